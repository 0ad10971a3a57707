"""GPU: the Pade6stagg dispatch (PadeDerOps.F90) selects, for each (bot, top),
the cd06stagg object the reference selects, and writes zeros where the
reference does; periodic objects ignore (bot, top)."""
import numpy as np
import pytest

import oracle as O
import padeops_amd as P
from fields import from_dev, rel_l2, to_dev

pytestmark = pytest.mark.gpu

DZ = 0.1
# (bot, top) -> flags of the selected object; the letters of derXY are (top, bottom)
OBJ = {(-1, -1): (0, 0, 0, 0), (-1, 0): (0, 0, 1, 0), (-1, 1): (1, 0, 0, 0), (0, -1): (0, 0, 0, 1),
       (0, 0): (1, 1, 1, 1), (0, 1): (1, 1, 0, 1), (1, -1): (0, 1, 0, 0), (1, 0): (1, 1, 1, 0), (1, 1): (1, 1, 0, 0)}
OPS = [("ddz_E2C", O.STG_D1_E2C), ("ddz_C2E", O.STG_D1_C2E), ("d2dz2_C2C", O.STG_D2_C2C),
       ("d2dz2_E2E", O.STG_D2_E2E), ("interpz_E2C", O.STG_IN_E2C), ("interpz_C2E", O.STG_IN_C2E)]


def test_dispatch(cuda):
    import torch
    n, n1, n2 = 64, 5, 3
    d = P.Pade6stagg()
    assert d.init(n, DZ, False) == 0
    for meth, op in OPS:
        nin, nout = O.stagg_lengths(op, n)
        f = np.asfortranarray(np.random.default_rng(op).standard_normal((n1, n2, nin)))
        fd = to_dev(f, torch.float64, cuda)
        for (bot, top), (te, be, ts, bs) in OBJ.items():
            out = torch.full((n1 * n2 * nout,), float("nan"), dtype=torch.float64, device=cuda)
            assert getattr(d, meth)(fd, out, bot, top, n1, n2, prec="f64") == 0
            g = from_dev(out, (n1, n2, nout))
            if op in (O.STG_D2_C2C, O.STG_D2_E2E) and (bot == 0 or top == 0):
                assert np.all(g == 0.0), (meth, bot, top)
                continue
            flags = te * O.TOP_EVEN + be * O.BOT_EVEN + ts * O.TOP_SIDED + bs * O.BOT_SIDED
            assert rel_l2(g, O.stagg_np(op, f, DZ, flags, n)) <= 1e-12, (meth, bot, top)
        out = torch.full((n1 * n2 * nout,), 7.0, dtype=torch.float64, device=cuda)
        assert getattr(d, meth)(fd, out, 2, 1, n1, n2, prec="f64") == 0
        torch.cuda.synchronize()
        assert float(out.abs().max()) == 0.0


def test_periodic(cuda):
    import torch
    n, n1, n2 = 32, 4, 2
    d = P.Pade6stagg()
    assert d.init(n, DZ, True) == 0
    f = np.asfortranarray(np.random.default_rng(1).standard_normal((n1, n2, n + 1)))
    f[:, :, n] = f[:, :, 0]
    out = torch.empty(n1 * n2 * n, dtype=torch.float64, device=cuda)
    assert d.ddz_E2C(to_dev(f, torch.float64, cuda), out, 1, 1, n1, n2, prec="f64") == 0
    assert rel_l2(from_dev(out, (n1, n2, n)), O.apply(O.STG_D1_E2C, f, 3, DZ, True, n=n)) <= 1e-12
