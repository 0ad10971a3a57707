"""pdo_op_apply_noalloc against the ordinary entry points, through the C ABI.

The allocation-free entry point runs the same kernels as the ordinary calls,
with the general kernel's work area in a fixed device array instead of
stream-ordered memory, so its results must equal theirs bit for bit.  Each
case makes two calls in a row, the second with a smaller input; the fixed
array keeps the first call's contents, which the second must not depend on.
"""
import numpy as np
import pytest

import padeops_amd as P

pytestmark = pytest.mark.gpu


def _field(shape_words, dtype, cuda, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(shape_words, generator=g, dtype=torch.float64).to(device=cuda, dtype=dtype)


def _pair(obj, ordinary, kind, dir, n_in, n_out, na, nb, dtype, prec, cuda, seed, bc=()):
    import torch
    f = _field(n_in * na * nb, dtype, cuda, seed)
    ref = torch.full((n_out * na * nb,), float("nan"), dtype=dtype, device=cuda)
    got = torch.full((n_out * na * nb,), float("nan"), dtype=dtype, device=cuda)
    ordinary(f, ref, na, nb, *bc, prec=prec)
    obj.apply_noalloc(kind, dir, f, got, na, nb, *(bc or (0, 0)), prec=prec)
    torch.cuda.synchronize()
    return ref, got


# (pencil counts of the first call, of the second, smaller one)
SIZES = [((37, 11), (20, 5))]


@pytest.mark.parametrize("prec", ["f64", "f32_refine"])
@pytest.mark.parametrize("dir", [1, 2, 3])
def test_cd10_nonperiodic_general_kernel(cuda, prec, dir):
    # non-periodic x pencils run on the one-lane-per-pencil kernel, whose work
    # area is the fixed array; y and z (n = 64, two segments) on the register
    # path, which needs none
    import torch
    dt = torch.float64 if prec == "f64" else torch.float32
    n = 64
    d = P.cd10()
    assert d.init(n, 0.1, False, 0, 0) == 0
    fn = {1: d.dd1, 2: d.dd2, 3: d.dd3}[dir]
    for call, (na, nb) in enumerate(SIZES[0]):
        print("cd10 dd%d n=%d na=%d nb=%d prec=%s call %d" % (dir, n, na, nb, prec, call), flush=True)
        ref, got = _pair(d, fn, 0, dir, n, n, na, nb, dt, prec, cuda, 10 + call, bc=(1, -1))
        assert torch.isfinite(ref).all()
        assert torch.equal(ref, got)


@pytest.mark.parametrize("prec", ["f64", "f32_refine"])
def test_cd06stagg_edge_systems(cuda, prec):
    # 257-row edge systems of a 256-cell wall-normal grid, complex data
    import torch
    dt = torch.float64 if prec == "f64" else torch.float32
    n = 256
    s = P.cd06stagg()
    assert s.init(n, 1.0 / n, True, False, False, True) == 0
    for call, (n1, n2) in enumerate([(9, 6), (5, 3)]):
        print("cd06stagg ddz_C2E n=%d n1=%d n2=%d prec=%s call %d" % (n, n1, n2, prec, call), flush=True)
        f = _field(2 * n1 * n2 * n, dt, cuda, 20 + call)
        ref = torch.full((2 * n1 * n2 * (n + 1),), float("nan"), dtype=dt, device=cuda)
        got = torch.full_like(ref, float("nan"))
        s.ddz_C2E(f, ref, n1, n2, is_complex=True, prec=prec)
        s.apply_noalloc(11, 3, f, got, 2 * n1, n2, prec=prec)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(ref, got)


def test_periodic_register_path(cuda):
    # a periodic cf90 y pass (register path, no work area) in FP32+refine
    import torch
    n = 128
    c = P.cf90()
    assert c.init(n, True) == 0
    for call, (na, nb) in enumerate([(64, 3), (32, 2)]):
        print("cf90 filter2 n=%d na=%d nb=%d call %d" % (n, na, nb, call), flush=True)
        ref, got = _pair(c, c.filter2, 2, 2, n, n, na, nb, torch.float32, "f32_refine", cuda, 30 + call,
                         bc=(0, 0))
        assert torch.equal(ref, got)


def test_rejections(cuda):
    # checked on the host before any launch
    import torch
    n = 64
    d = P.cd10()
    assert d.init(n, 0.1, False, 0, 0) == 0
    f = torch.zeros(n * 70000, dtype=torch.float64, device=cuda)
    out = torch.zeros_like(f)
    # variant never built by an ordinary call
    with pytest.raises(P.PdoError):
        d.apply_noalloc(0, 1, f, out, 70000, 1, 1, 1, prec="f64")
    d.dd1(f[:n], out[:n], 1, 1, 1, 1)
    torch.cuda.synchronize()
    # 64 x 70000 FP64 words exceed the 32 MiB work area
    with pytest.raises(P.PdoError, match="too large"):
        d.apply_noalloc(0, 1, f, out, 70000, 1, 1, 1, prec="f64")
    # bad boundary flag: the reference's GracefulExit code
    with pytest.raises(P.GracefulExit):
        d.apply_noalloc(0, 1, f, out, 1, 1, 2, 0, prec="f64")
    assert np.isfinite(out[:n].cpu().numpy()).all()
