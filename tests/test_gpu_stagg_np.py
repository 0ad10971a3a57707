"""GPU parity of the non-periodic cd06stagg operators (the nine boundary
objects of Pade6stagg, real and complex fields) against the oracle's
restatement of cd06stagg.F90 / STAGG_CD06_files.  FP64 relative L2 <= 1e-12;
FP32 + refinement and FP32-storage / FP64-arithmetic <= 2e-7 against the
oracle on the same FP32-rounded input."""
import numpy as np
import pytest

import oracle as O
import padeops_amd as P
from fields import from_dev, rel_l2, to_dev

pytestmark = pytest.mark.gpu

METHODS = [("ddz_E2C", O.STG_D1_E2C), ("ddz_C2E", O.STG_D1_C2E), ("ddz_C2C", O.STG_D1_C2C),
           ("ddz_E2E", O.STG_D1_E2E), ("d2dz2_C2C", O.STG_D2_C2C), ("d2dz2_E2E", O.STG_D2_E2E),
           ("InterpZ_E2C", O.STG_IN_E2C), ("InterpZ_C2E", O.STG_IN_C2E)]
PADE6 = {"ES": (1, 1, 0, 1), "SE": (1, 1, 1, 0), "SS": (1, 1, 1, 1), "SO": (0, 0, 1, 0), "OS": (0, 0, 0, 1),
         "OO": (0, 0, 0, 0), "EE": (1, 1, 0, 0), "OE": (0, 1, 0, 0), "EO": (1, 0, 0, 0)}
DZ = 0.09


def _run(o, meth, op, n, flags, n1, n2, is_complex, dtype, prec, cuda, seed):
    import torch
    nin, nout = O.stagg_lengths(op, n)
    k = 2 if is_complex else 1
    f = np.asfortranarray(np.random.default_rng(seed).standard_normal((k * n1, n2, nin)))
    if dtype == torch.float32:
        f = f.astype(np.float32).astype(np.float64)
    fd = to_dev(f, dtype, cuda)
    out = torch.full((k * n1 * n2 * nout,), float("nan"), dtype=dtype, device=cuda)
    assert getattr(o, meth)(fd, out, n1, n2, is_complex=is_complex, prec=prec) == 0
    g = from_dev(out, (k * n1, n2, nout))
    ref = O.stagg_np(op, f, DZ, flags, n)
    return rel_l2(g, ref)


# n = 64 and 256: the edge systems (n + 1 rows) run on the register path with
# one extra row in the last segment; n = 63: edge systems of 64 rows on the
# register path, cell systems on the one-lane-per-pencil kernel; n = 40: both
# on the latter
@pytest.mark.parametrize("combo", list(PADE6))
@pytest.mark.parametrize("n", [256, 64, 63, 40])
def test_stagg_np_fp64(cuda, combo, n):
    import torch
    te, be, ts, bs = PADE6[combo]
    flags = te * O.TOP_EVEN + be * O.BOT_EVEN + ts * O.TOP_SIDED + bs * O.BOT_SIDED
    o = P.cd06stagg()
    assert o.init(n, DZ, bool(te), bool(be), bool(ts), bool(bs)) == 0
    for i, (meth, op) in enumerate(METHODS):
        for is_complex in (False, True):
            err = _run(o, meth, op, n, flags, 9, 4, is_complex, torch.float64, "f64", cuda, n + i)
            print(combo, n, meth, "complex" if is_complex else "real", err)
            assert err <= 1e-12


@pytest.mark.parametrize("prec", ["f32_refine", "f32_store"])
@pytest.mark.parametrize("combo", ["SS", "OE"])
def test_stagg_np_fp32(cuda, combo, prec):
    import torch
    n = 128
    te, be, ts, bs = PADE6[combo]
    flags = te * O.TOP_EVEN + be * O.BOT_EVEN + ts * O.TOP_SIDED + bs * O.BOT_SIDED
    o = P.cd06stagg()
    assert o.init(n, DZ, bool(te), bool(be), bool(ts), bool(bs)) == 0
    # every operator of a 128-cell grid is on the register path (4 segments)
    for kind in range(10, 18):
        assert o.partition(kind) == (32, 4)
    for i, (meth, op) in enumerate(METHODS):
        err = _run(o, meth, op, n, flags, 70, 3, False, torch.float32, prec, cuda, i)
        print(combo, prec, meth, err)
        assert err <= 2e-7
