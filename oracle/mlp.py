"""TEST INFRASTRUCTURE ONLY (the checker): CPU restatement of the learned SGS
closure's network in numpy - forward, backward and Adam - with the same
rounding points as the GPU kernels (BF16 weights and activations, FP32
accumulation).  Used by tests/ only; the product never imports it.

Network: 10 inputs -> 64 -> 64 -> 64 -> 6, GELU (tanh form) after each hidden
layer, linear output; loss = mean over the batch and the 6 outputs of the
squared error.  Parameters are FP32 master copies; the forward pass uses
their BF16 rounding.  BF16 rounding is round-to-nearest-even of the FP32
value (as the kernels' conversion)."""
import numpy as np

SIZES = (10, 64, 64, 64, 6)
C0 = np.float32(0.7978845608028654)   # sqrt(2/pi)
C1 = np.float32(0.044715)


def bf16(x):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    x = np.asarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    lsb = (b >> 16) & 1
    r = ((b + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(z):
    z = z.astype(np.float32)
    t = np.tanh(C0 * (z + C1 * z * z * z))
    return (np.float32(0.5) * z * (np.float32(1) + t)).astype(np.float32)


def gelu_grad(z):
    z = z.astype(np.float32)
    u = C0 * (z + C1 * z * z * z)
    t = np.tanh(u)
    du = C0 * (np.float32(1) + np.float32(3) * C1 * z * z)
    return (np.float32(0.5) * (np.float32(1) + t) + np.float32(0.5) * z * (np.float32(1) - t * t) * du).astype(
        np.float32)


def init_params(seed=0):
    """He-style initialisation: W_l (out, in) ~ N(0, 2/in), b_l = 0 (FP32)."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(4):
        fan_in, fan_out = SIZES[i], SIZES[i + 1]
        W = (rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)).astype(np.float32)
        params.append((W, np.zeros(fan_out, np.float32)))
    return params


def _gelu64(z):
    return 0.5 * z * (1 + np.tanh(0.7978845608028654 * (z + 0.044715 * z ** 3)))


def _gelu_grad64(z):
    u = 0.7978845608028654 * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * 0.7978845608028654 * (1 + 3 * 0.044715 * z * z)


def forward(params, X, exact=False):
    """Returns (Y, cache): Y (B, 6) FP32; cache holds the BF16 layer inputs H_l
    and the pre-activations Z_l the backward pass reads.  exact=True: FP64
    without rounding (the mathematical network, for derivative checks)."""
    if exact:
        H = np.asarray(X, np.float64)
        cache = {"H": [H], "Z": []}
        for i, (W, b) in enumerate(params):
            Z = H @ np.asarray(W, np.float64).T + b
            if i < 3:
                cache["Z"].append(Z)
                H = _gelu64(Z)
                cache["H"].append(H)
            else:
                Y = Z
        return Y, cache
    H = bf16(X)
    cache = {"H": [H], "Z": []}
    for i, (W, b) in enumerate(params):
        Z = (H @ bf16(W).T).astype(np.float32) + b
        if i < 3:
            Zr = bf16(Z)
            cache["Z"].append(Zr)
            H = bf16(gelu(Zr))
            cache["H"].append(H)
        else:
            Y = Z.astype(np.float32)
    return Y, cache


def backward(params, cache, Y, T, exact=False, dY=None):
    """Gradients (dW_l, db_l) of the mean squared error, FP32, with the
    upstream gradients rounded to BF16 before each matrix product (as the
    MFMA kernels consume them).  exact=True: FP64, no rounding.  dY given:
    the gradients of any loss with dL/dY = dY (Y and T unused)."""
    B = (Y if dY is None else dY).shape[0]
    if exact:
        dZ = 2.0 / (B * SIZES[-1]) * (Y - T) if dY is None else np.asarray(dY, np.float64)
        grads = [None] * 4
        for i in (3, 2, 1, 0):
            grads[i] = (dZ.T @ cache["H"][i], dZ.sum(axis=0))
            if i > 0:
                dZ = (dZ @ np.asarray(params[i][0], np.float64)) * _gelu_grad64(cache["Z"][i - 1])
        return grads
    dZ = (np.float32(2.0 / (B * SIZES[-1])) * (Y - T)).astype(np.float32) if dY is None else \
        np.asarray(dY, np.float32)
    grads = [None] * 4
    for i in (3, 2, 1, 0):
        W, _ = params[i]
        dZb = bf16(dZ)
        Hin = cache["H"][i]
        grads[i] = ((dZb.T @ Hin).astype(np.float32), dZ.sum(axis=0).astype(np.float32))
        if i > 0:
            dH = (dZb @ bf16(W)).astype(np.float32)
            dZ = (dH * gelu_grad(cache["Z"][i - 1])).astype(np.float32)
    return grads


def loss(Y, T):
    return float(np.mean((Y.astype(np.float64) - T) ** 2))


class Adam:
    """Adam (Kingma and Ba, 2015) on FP32 master parameters."""

    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = [[np.zeros_like(W), np.zeros_like(b)] for W, b in params]
        self.v = [[np.zeros_like(W), np.zeros_like(b)] for W, b in params]
        self.t = 0

    def step(self, params, grads):
        self.t += 1
        c1 = 1.0 - self.b1 ** self.t
        c2 = 1.0 - self.b2 ** self.t
        out = []
        for i, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
            new = []
            for j, (p, g) in enumerate(((W, gW), (b, gb))):
                m = self.m[i][j] = (self.b1 * self.m[i][j] + (1 - self.b1) * g).astype(np.float32)
                v = self.v[i][j] = (self.b2 * self.v[i][j] + (1 - self.b2) * g * g).astype(np.float32)
                new.append((p - self.lr * (m / c1) / (np.sqrt(v / c2) + self.eps)).astype(np.float32))
            out.append(tuple(new))
        return out
