"""Learned SGS closure: training data, training and a priori judgement.

Data: a triply periodic velocity field on an n^3 grid of side 2 pi is
filtered with the Gaussian LES filter of width Delta, G(k) = exp(-k^2 Delta^2
/ 24) per direction, by this library's kernels: the 2D transforms of the
spectral object, its periodic-z transform-multiply-transform pass with G as
the (separable) factors, and the inverse transforms.  (A sharp spectral
cutoff leaves the exact stress of random-phase fields almost uncorrelated
with any local model; the smooth filter is the usual a priori setting.)  The
exact subgrid stress of the filtered field is

    tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j),

and the network's inputs are the filtered velocity gradient (cd06 compact
derivatives of the filtered field, this library's operators) scaled by the
filter width Delta, plus Delta itself:

    x = (Delta d bar(u_i) / d x_j (9 values, igrid's order) / U, Delta / Delta_0)
    t = deviatoric tau_ij (6 values: 11 12 13 22 23 33) / U^2

with U the rms of Delta |grad bar(u)| over the field and Delta_0 the finest
filter width of the dataset.

Baseline: the Smagorinsky stress -2 (C Delta)^2 |S| S_ij with its coefficient
fitted by least squares to the exact stress (the best any constant-coefficient
eddy viscosity, the dynamic procedure's planar constant included, can do a
priori in homogeneous turbulence), compared with the network on held-out
data by the correlation coefficient and the relative L2 error of the
deviatoric stress.
"""
import math

import numpy as np

import padeops_amd as P

IDX6 = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def _torch():
    import torch
    return torch


def gaussian_spectral(n, delta):
    """Spectral object of the n^3 periodic grid whose dealias pass applies the
    Gaussian filter exp(-k^2 delta^2 / 24) in each direction."""
    dx = 2 * math.pi / n
    sp = P.spectral()
    sp.init(n, n, n, dx, dx, dx, init_periodicInZ=True)
    k1, k2, k3 = sp.wavenumbers()[:3]
    g = [np.exp(-(k * k) * delta * delta / 24.0) for k in (k1, k2, k3)]
    sp.set_factors(*g)
    return sp


def apply_filter(sp, f, stream=None):
    """Apply the filter of the spectral object sp (its dealias factors, periodic
    z) to the real (n, n, n) field f (flat CUDA tensor, FP64)."""
    torch = _torch()
    n = sp.shape[0]
    fh = torch.empty(sp._nspec(), dtype=torch.complex128, device=f.device)
    out = torch.empty_like(f)
    sp.fft(f, fh, prec="f64", stream=stream)
    sp.dealias(fh, prec="f64", stream=stream)
    sp.ifft(fh, out, prec="f64", stream=stream)
    assert out.numel() == n ** 3
    return out


def filtered_dataset(u, v, w, n, delta, d0=None, stream=None):
    """(X, T, aux) for one filter width delta.  u, v, w: flat FP64 CUDA
    tensors of the n^3 field (Fortran order, x fastest).  X (n^3, 10),
    T (n^3, 6) FP32; aux holds the filtered strain rate S (n^3, 6), |S| (n^3)
    and the scale U."""
    torch = _torch()
    dx = 2 * math.pi / n
    sp = gaussian_spectral(n, delta)
    vel = [u, v, w]
    ub = [apply_filter(sp, c, stream) for c in vel]
    tau = []
    for i, j in IDX6:
        tau.append(apply_filter(sp, vel[i] * vel[j], stream) - ub[i] * ub[j])
    tr = (tau[0] + tau[3] + tau[5]) / 3.0
    for q in (0, 3, 5):
        tau[q] = tau[q] - tr
    d = P.cd06()
    d.init(n, dx, True)
    grads = []
    for c in ub:
        for dirn, meth in ((1, d.dd1), (2, d.dd2), (3, d.dd3)):
            g = torch.empty_like(c)
            meth(c, g, n, n, prec="f64", stream=stream)
            grads.append(g)
    d0 = d0 or delta
    G = torch.stack(grads, dim=1) * delta                      # (N, 9) Delta grad u
    U = float(torch.sqrt((G ** 2).sum(dim=1).mean()))
    X = torch.cat([G / U, torch.full((G.shape[0], 1), delta / d0, dtype=G.dtype, device=G.device)], dim=1)
    T = torch.stack(tau, dim=1) / (U * U)
    S = torch.stack([0.5 * (grads[3 * i + j] + grads[3 * j + i]) for i, j in IDX6], dim=1)
    Smag = torch.sqrt(2 * (S[:, 0] ** 2 + 2 * S[:, 1] ** 2 + 2 * S[:, 2] ** 2 + S[:, 3] ** 2 + 2 * S[:, 4] ** 2 +
                           S[:, 5] ** 2))
    return X.float().contiguous(), T.float().contiguous(), {"S": S, "Smag": Smag, "U": U, "delta": delta}


def smagorinsky_fit(aux, T):
    """Least-squares Smagorinsky stress -2 c |S| S_ij (deviatoric, scaled like T)."""
    torch = _torch()
    S, Smag, U = aux["S"], aux["Smag"], aux["U"]
    base = (-2.0 * aux["delta"] ** 2 * Smag[:, None] * S) / (U * U)
    tr = (base[:, 0] + base[:, 3] + base[:, 5]) / 3.0
    base = base.clone()
    for q in (0, 3, 5):
        base[:, q] -= tr
    w = torch.tensor([1, 2, 2, 1, 2, 1], dtype=base.dtype, device=base.device)
    Td = T.double()
    c = float((w * base * Td).sum() / (w * base * base).sum())
    return (c * base).float(), c


def stress_metrics(pred, T):
    """(mean correlation coefficient over the 6 components, relative L2 error
    of the deviatoric tensor, off-diagonals counted twice)."""
    torch = _torch()
    p, t = pred.double(), T.double()
    cors = []
    for q in range(6):
        a, b = p[:, q] - p[:, q].mean(), t[:, q] - t[:, q].mean()
        cors.append(float((a * b).sum() / torch.sqrt((a * a).sum() * (b * b).sum())))
    w = torch.tensor([1, 2, 2, 1, 2, 1], dtype=p.dtype, device=p.device)
    err = float(torch.sqrt((w * (p - t) ** 2).sum() / (w * t * t).sum()))
    return float(np.mean(cors)), err


def gradient_model(X):
    """The gradient (Clark) model tau_ij = Delta^2 / 12 d_k bar(u_i) d_k bar(u_j),
    the leading term of the Gaussian filter's exact stress, deviatoric, in the
    units of T: X_ik X_jk / 12.  A reference point for the a priori scores."""
    torch = _torch()
    g = X[:, :9].double().reshape(-1, 3, 3)
    gg = torch.einsum("nik,njk->nij", g, g) / 12.0
    tr = (gg[:, 0, 0] + gg[:, 1, 1] + gg[:, 2, 2]) / 3.0
    out = torch.stack([gg[:, i, j] for i, j in IDX6], dim=1)
    for q in (0, 3, 5):
        out[:, q] -= tr
    return out.float()


def train(net, X, T, steps, batch=65536, lr=2e-3, seed=0, stream=None, tscale=None):
    """Adam on random batches of (X, T); returns the loss of the last batch.

    The network learns T / tscale (tscale defaults to the rms of T): the
    subgrid stress in units of U^2 is O(1e-2), and an O(1) target keeps the
    output layer at its initial scale and the BF16 upstream gradients well
    resolved.  predict() multiplies by the same factor (kept on the net)."""
    torch = _torch()
    if tscale is None:
        tscale = float(torch.sqrt((T.double() ** 2).mean()))
    net.tscale = tscale
    Tn = (T / tscale).contiguous()
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    Y = torch.empty((batch, 6), dtype=torch.float32, device=X.device)
    for s in range(steps):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g, device=X.device)
        xb, tb = X[idx].contiguous(), Tn[idx].contiguous()
        lr_s = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
        net.train_step(xb, tb, lr=lr_s, Y=Y, stream=stream)
    return float(((Y - tb) ** 2).mean())


def predict(net, X, chunk=1 << 20, stream=None):
    """Network stress for the rows of X, in the units of T."""
    torch = _torch()
    out = torch.empty((X.shape[0], 6), dtype=torch.float32, device=X.device)
    for a in range(0, X.shape[0], chunk):
        net.forward(X[a:a + chunk].contiguous(), out[a:a + chunk], stream=stream)
    return out * getattr(net, "tscale", 1.0)
