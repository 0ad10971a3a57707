"""Pins of the MLP oracle (oracle/mlp.py): the backward formulas against
central finite differences of the exact (FP64) network, the BF16 rounding
against its definition, and one Adam step against the published update."""
import numpy as np

import oracle.mlp as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even
    assert r[2] == np.float32(1.0 + 2 ** -6)      # tie to even (up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.14159265) <= 2 ** -7 * 3.2
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = [(W.astype(np.float64), b.astype(np.float64) + 0.1 * rng.standard_normal(b.shape))
              for W, b in M.init_params(1)]
    X = rng.standard_normal((7, 10))
    T = rng.standard_normal((7, 6))
    Y, cache = M.forward(params, X, exact=True)
    grads = M.backward(params, cache, Y, T, exact=True)
    eps = 1e-6
    for layer in range(4):
        for which in (0, 1):
            p = params[layer][which]
            idx = tuple(rng.integers(0, s) for s in p.shape)
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in p.shape)
                orig = p[idx]
                p[idx] = orig + eps
                lp = M.loss(M.forward(params, X, exact=True)[0], T)
                p[idx] = orig - eps
                lm = M.loss(M.forward(params, X, exact=True)[0], T)
                p[idx] = orig
                fd = (lp - lm) / (2 * eps)
                assert abs(grads[layer][which][idx] - fd) <= 1e-7 * (1 + abs(fd)), (layer, which, idx)


def test_adam_first_step_is_signed_lr():
    params = [(np.ones((2, 2), np.float32), np.zeros(2, np.float32))]
    grads = [(np.array([[1.0, -2.0], [0.5, 0.0]], np.float32), np.array([3.0, -1.0], np.float32))]
    opt = M.Adam(params, lr=0.01)
    new = opt.step(params, grads)
    # first Adam step: p - lr * g / (|g| + eps') with bias-corrected moments
    assert np.allclose(new[0][0], 1.0 - 0.01 * np.sign(grads[0][0]), atol=1e-6)
    assert np.allclose(new[0][1], -0.01 * np.sign(grads[0][1]), atol=1e-6)
