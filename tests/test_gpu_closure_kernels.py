"""GPU parity of the learned closure's pointwise kernels (pdo_closure_*) and of
the network's backward pass from a given dL/dY (pdo_mlp_backward) against the
oracle (oracle/closure.py, oracle/mlp.py).

  features: FP64 invariants, FP32 roots; the invariants (roots undone)
            to <= 2e-6 of each column's scale
  stress:   FP64 fields to 1e-12 relative; the igrid cell (11 12 22 33) and
            edge (13 23) layouts; the fallback keeps the prior values at
            exactly the points of negative dissipation
  loss:     the value to 1e-10, dL/dG to FP32 rounding
  backward: the MSE case equals pdo_mlp_gradients; a generic dY matches the
            oracle's BF16-rounded backward (relative L2 <= 2e-3, as the other
            MLP tests)."""
import numpy as np
import pytest

import oracle.closure as C
import oracle.mlp as M
import padeops_amd as P
from fields import rel_l2

pytestmark = pytest.mark.gpu


def _field(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, 3, 3)) * rng.uniform(0.2, 3.0, (n, 1, 1))
    A -= (np.trace(A, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    return A.reshape(n, 9)


def _planes(A, cuda, dtype):
    """(npts, 9) -> 9 component planes of npts (igrid's duidxj layout)"""
    import torch
    return torch.from_numpy(np.ascontiguousarray(A.T).ravel()).to(device=cuda, dtype=dtype)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_features(cuda, dt):
    import torch
    n = 5000
    A = _field(n, 1)
    dtype = torch.float64 if dt == "f64" else torch.float32
    if dt == "f32":
        A = A.astype(np.float32).astype(np.float64)
    cl = P.Closure(None, n)
    X = torch.empty((n, 10), dtype=torch.float32, device=cuda)
    A9 = torch.empty((n, 9), dtype=torch.float32, device=cuda)
    assert cl.features(_planes(A, cuda, dtype), n, 0.5, X, A9) == 0
    torch.cuda.synchronize()
    ref = C.features(A, 0.5)
    got = X.cpu().numpy().astype(np.float64)
    # compare the invariants themselves, sign(x) |x|^p with p the degree of
    # each root (a root near zero magnifies FP64 ordering differences)
    deg = np.array([2, 2, 3, 3, 4, 6, 1, 1, 1, 1], float)
    unroot = lambda x: np.sign(x) * np.abs(x) ** deg  # noqa: E731
    err = np.abs(unroot(got) - unroot(ref)).max(axis=0) / np.abs(unroot(ref)).max(axis=0)
    print("feature column errors", err)
    assert err.max() <= 2e-6
    g = C.scale(A)
    assert abs(cl.scale() - g) <= 1e-12 * g
    assert np.abs(A9.cpu().numpy() - A / g).max() <= 1e-6 * np.abs(A / g).max()


def test_stress_layouts_and_fallback(cuda):
    import torch
    n = 4096
    A = _field(n, 2)
    rng = np.random.default_rng(3)
    net = P.SGSNet(64)
    # the stress is checked against the oracle on the network's own outputs
    # for this field (forwarded again below)
    params = [(W, (0.1 * rng.standard_normal(b.shape)).astype(np.float32)) for W, b in M.init_params(5)]
    net.set_params(P.SGSNet.pack(params))
    cl = P.Closure(net, n)
    g = _planes(A, cuda, torch.float64)
    delta = 0.37
    tau = torch.empty(6 * n, dtype=torch.float64, device=cuda)
    assert cl.stress(g, n, delta, 0.5, tau, which=2) == 0
    X = torch.empty((n, 10), dtype=torch.float32, device=cuda)
    assert cl.features(g, n, 0.5, X) == 0
    Y = torch.empty((n, 6), dtype=torch.float32, device=cuda)
    assert net.forward(X, Y) == 0
    torch.cuda.synchronize()
    Gn = Y.cpu().numpy()
    ref = C.stress(A, Gn, delta)
    got = tau.cpu().numpy().reshape(6, n).T
    e = np.abs(got - ref).max() / np.abs(ref).max()
    print("stress", e)
    assert e <= 1e-12
    # igrid layouts
    tc = torch.empty(4 * n, dtype=torch.float64, device=cuda)
    te = torch.empty(2 * n, dtype=torch.float64, device=cuda)
    assert cl.stress(g, n, delta, 0.5, tc, which=0) == 0
    assert cl.stress(g, n, delta, 0.5, te, which=1, rescale=False) == 0
    torch.cuda.synchronize()
    assert np.abs(tc.cpu().numpy().reshape(4, n).T - ref[:, [0, 1, 3, 5]]).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(te.cpu().numpy().reshape(2, n).T - ref[:, [2, 4]]).max() <= 1e-12 * np.abs(ref).max()
    # fallback: marker values stay exactly where the dissipation is negative
    mark = torch.full((6 * n,), 1234.5, dtype=torch.float64, device=cuda)
    assert cl.stress(g, n, delta, 0.5, mark, which=2, fallback=True) == 0
    nf = cl.fallback_count()
    m = mark.cpu().numpy().reshape(6, n).T
    neg = C.dissipation_sign(A, Gn) < 0
    print("fallback points", nf, "of", n)
    assert nf == int(neg.sum()) and 0 < nf < n
    assert (m[neg] == 1234.5).all()
    assert np.abs(m[~neg] - ref[~neg]).max() <= 1e-12 * np.abs(ref).max()


def test_loss(cuda):
    import torch
    B = 3000
    rng = np.random.default_rng(4)
    A9 = (_field(B, 5) / 2.0).astype(np.float32)
    G = rng.standard_normal((B, 6)).astype(np.float32)
    T = rng.standard_normal((B, 6)).astype(np.float32)
    cl = P.Closure(None, 16)
    d = lambda a: torch.from_numpy(a).to(cuda)  # noqa: E731
    dG = torch.empty((B, 6), dtype=torch.float32, device=cuda)
    L = cl.loss(d(G), d(A9), d(T), dG)
    Lr, dGr = C.loss(G.astype(np.float64), A9.astype(np.float64), T.astype(np.float64))
    assert abs(L - Lr) <= 1e-10 * Lr
    assert np.abs(dG.cpu().numpy() - dGr).max() <= 1e-6 * np.abs(dGr).max()


def test_mlp_backward_from_dy(cuda):
    import torch
    B = 2048
    rng = np.random.default_rng(6)
    params = [(W, (0.1 * rng.standard_normal(b.shape)).astype(np.float32)) for W, b in M.init_params(7)]
    X = rng.standard_normal((B, 10)).astype(np.float32)
    T = rng.standard_normal((B, 6)).astype(np.float32)
    net = P.SGSNet(B)
    net.set_params(P.SGSNet.pack(params))
    Xd, Td = torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda)
    Y = torch.empty((B, 6), dtype=torch.float32, device=cuda)
    assert net.gradients(Xd, Td, Y) == 0
    torch.cuda.synchronize()
    g_mse = net.get_grads()
    # the same loss through dL/dY
    assert net.forward_train(Xd, Y) == 0
    dY = (Y - Td) * np.float32(2.0 / (B * 6))
    assert net.backward(dY.contiguous()) == 0
    torch.cuda.synchronize()
    assert rel_l2(net.get_grads(), g_mse) <= 1e-6
    # a generic upstream gradient against the oracle
    dYg = rng.standard_normal((B, 6)).astype(np.float32) * 1e-3
    assert net.forward_train(Xd, Y) == 0
    assert net.backward(torch.from_numpy(dYg).to(cuda)) == 0
    torch.cuda.synchronize()
    g = P.SGSNet.unpack(net.get_grads())
    _, cache = M.forward(params, X)
    gref = M.backward(params, cache, None, None, dY=dYg)
    for (dW, db), (rW, rb) in zip(g, gref):
        assert rel_l2(dW, rW) <= 2e-3
        assert rel_l2(db, rb) <= 2e-3
    # the batch of the last forward_train is enforced
    with pytest.raises(P.PdoError):
        net.backward(torch.from_numpy(dYg[:100]).to(cuda))
