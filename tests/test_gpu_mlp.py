"""GPU: the learned closure's BF16 MFMA kernels against the CPU oracle
(oracle/mlp.py, same rounding points).  Forward outputs and batch gradients
agree to the level of FP32 accumulation-order differences (relative L2
<= 2e-3, a BF16 rounding of a hidden activation can flip where the two
accumulations straddle a rounding boundary); Adam steps track the oracle's
parameters; training reduces the loss on a quadratic target."""
import numpy as np
import pytest

import oracle.mlp as M
import padeops_amd as P
from fields import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(cuda)


@pytest.mark.parametrize("B", [64, 1000, 65536])
def test_forward_and_gradients(cuda, B):
    import torch
    rng = np.random.default_rng(B)
    params = [(W, (0.1 * rng.standard_normal(b.shape)).astype(np.float32)) for W, b in M.init_params(3)]
    X = rng.standard_normal((B, 10)).astype(np.float32)
    T = rng.standard_normal((B, 6)).astype(np.float32)
    net = P.SGSNet(max(B, 64))
    net.set_params(P.SGSNet.pack(params))
    Xd, Td = _dev(X, cuda), _dev(T, cuda)
    Y = torch.empty((B, 6), dtype=torch.float32, device=cuda)
    assert net.forward(Xd, Y) == 0
    torch.cuda.synchronize()
    Yref, cache = M.forward(params, X)
    ey = rel_l2(Y.cpu().numpy(), Yref)
    assert net.gradients(Xd, Td, Y) == 0
    torch.cuda.synchronize()
    g = P.SGSNet.unpack(net.get_grads())
    gref = M.backward(params, cache, Yref, T)
    errs = [rel_l2(g[i][j], gref[i][j]) for i in range(4) for j in range(2)]
    print("B", B, "forward", ey, "grads", errs)
    assert ey <= 2e-3
    assert max(errs) <= 2e-3
    # padding entries of the parameter vector stay zero after an Adam step
    lib_before = net.get_params()
    from padeops_amd import _lib
    assert _lib.lib().pdo_mlp_adam_step(net._h, 1e-3, 0.9, 0.999, 1e-8, None) == 0
    after = net.get_params()
    pad = np.zeros(10448, bool)
    pad[0:1024] = (np.arange(1024) % 16) >= 10
    pad[9408 + 6 * 64:10432] = True
    pad[10432 + 6:10448] = True
    assert np.all(after[pad] == 0) and np.all(lib_before[pad] == 0)


@pytest.mark.parametrize("B", [5, 400003])
def test_inference_kernel(cuda, B):
    """The inference forward (persistent workgroups, several 16-sample tiles
    per wave at B = 400003, a partial last tile, and a single partial tile at
    B = 5) against the oracle and against the training forward, which keeps
    its activations and evaluates GELU through tanh."""
    import torch
    rng = np.random.default_rng(11)
    params = [(W, (0.1 * rng.standard_normal(b.shape)).astype(np.float32)) for W, b in M.init_params(4)]
    X = rng.standard_normal((B, 10)).astype(np.float32)
    net = P.SGSNet(max(B, 64))
    net.set_params(P.SGSNet.pack(params))
    Xd = _dev(X, cuda)
    Y = torch.full((B + 1, 6), 7.0, dtype=torch.float32, device=cuda)
    assert net.forward(Xd, Y[:B]) == 0
    Yt = torch.empty((B, 6), dtype=torch.float32, device=cuda)
    assert net.forward_train(Xd, Yt) == 0
    torch.cuda.synchronize()
    Yref, _ = M.forward(params, X)
    got = Y.cpu().numpy()
    e_oracle = rel_l2(got[:B], Yref)
    e_train = rel_l2(got[:B], Yt.cpu().numpy())
    print("B", B, "vs oracle", e_oracle, "vs training forward", e_train)
    assert e_oracle <= 2e-3 and e_train <= 2e-3
    # nothing written past the batch
    assert np.all(got[B] == 7.0)


def test_adam_tracks_oracle(cuda):
    import torch
    rng = np.random.default_rng(7)
    B = 4096
    params = M.init_params(5)
    X = rng.standard_normal((B, 10)).astype(np.float32)
    T = rng.standard_normal((B, 6)).astype(np.float32)
    net = P.SGSNet(B)
    net.set_params(P.SGSNet.pack(params))
    opt = M.Adam(params, lr=1e-3)
    Xd, Td = _dev(X, cuda), _dev(T, cuda)
    for _ in range(5):
        assert net.train_step(Xd, Td, lr=1e-3) == 0
        Y, cache = M.forward(params, X)
        params = opt.step(params, M.backward(params, cache, Y, T))
    torch.cuda.synchronize()
    got = P.SGSNet.unpack(net.get_params())
    for i in range(4):
        for j in range(2):
            # per-entry agreement: Adam's first steps move every entry by ~lr
            assert np.abs(got[i][j] - params[i][j]).max() <= 2e-4, (i, j)


def test_training_reduces_loss(cuda):
    import torch
    rng = np.random.default_rng(1)
    B = 65536
    X = rng.standard_normal((B, 10)).astype(np.float32)
    # a gradient-model-like quadratic target: T_c = sum_k A_ik A_jk / 12 of a 3x3 block of X
    A = X[:, :9].reshape(B, 3, 3)
    Q = np.einsum("bik,bjk->bij", A, A) / 12.0
    T = np.stack([Q[:, 0, 0], Q[:, 0, 1], Q[:, 0, 2], Q[:, 1, 1], Q[:, 1, 2], Q[:, 2, 2]], axis=1).astype(np.float32)
    net = P.SGSNet(B)
    net.set_params(P.SGSNet.pack(M.init_params(0)))
    Xd, Td = _dev(X, cuda), _dev(T, cuda)
    Y = torch.empty((B, 6), dtype=torch.float32, device=cuda)
    net.forward(Xd, Y)
    l0 = float(((Y - Td) ** 2).mean())
    for _ in range(400):
        assert net.train_step(Xd, Td, lr=2e-3) == 0
    net.forward(Xd, Y)
    torch.cuda.synchronize()
    l1 = float(((Y - Td) ** 2).mean())
    print("loss", l0, "->", l1)
    assert l1 < 0.1 * l0
