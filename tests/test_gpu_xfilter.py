"""GPU: the explicit Gaussian and least-squares filters (gaussianstuff,
lstsqstuff) against the oracle, along x, y and z, periodic and
non-periodic with every boundary option, FP64 and FP32 storage, the
complex z variant, n == 1 and the argument errors."""
import numpy as np
import pytest

import oracle.explicit as X
import padeops_amd as P
from fields import from_dev, random_field, rel_l2, to_dev

pytestmark = pytest.mark.gpu

SHAPES = [(16, 8, 10), (8, 9, 12), (33, 12, 10), (64, 64, 64)]


def _obj(kind, n, periodic):
    o = P.gaussian() if kind == "gaussian" else P.lstsq()
    o.init(n, periodic)
    return o


@pytest.mark.parametrize("kind", ["gaussian", "lstsq"])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("d", [1, 2, 3])
def test_filter_fp64(cuda, kind, periodic, shape, d):
    import torch
    n = shape[d - 1]
    f = random_field(shape, seed=sum(shape) + d)
    bcs = [(0, 0)] if (periodic or kind == "lstsq") else [(0, 0), (1, 1), (-1, -1), (1, -1), (-1, 0)]
    o = _obj(kind, n, periodic)
    for bc1, bcn in bcs:
        ref = X.filter_axis(f, d - 1, kind, periodic, bc1, bcn)
        ft = to_dev(f, torch.float64, cuda)
        out = torch.full_like(ft, float("nan"))
        others = [s for i, s in enumerate(shape) if i != d - 1]
        getattr(o, "filter%d" % d)(ft, out, others[0], others[1], bc1, bcn, prec="f64")
        torch.cuda.synchronize()
        e = rel_l2(from_dev(out, shape), ref)
        print(kind, periodic, shape, d, bc1, bcn, e)
        assert e <= 1e-14


@pytest.mark.parametrize("prec", ["f32_refine", "f32_store", "f32"])
def test_filter_fp32(cuda, prec):
    import torch
    shape = (64, 32, 48)
    f = random_field(shape, seed=11)
    for kind in ("gaussian", "lstsq"):
        for d in (1, 2, 3):
            n = shape[d - 1]
            o = _obj(kind, n, False)
            ft = to_dev(f, torch.float32, cuda)
            out = torch.empty_like(ft)
            others = [s for i, s in enumerate(shape) if i != d - 1]
            getattr(o, "filter%d" % d)(ft, out, others[0], others[1], 1, -1, prec=prec)
            torch.cuda.synchronize()
            # the input rounded to FP32 is the oracle's input; FP64 arithmetic, one rounding out
            ref = X.filter_axis(from_dev(ft, shape), d - 1, kind, False, 1, -1)
            e = rel_l2(from_dev(out, shape), ref)
            assert e <= 1e-7, (kind, d, e)


def test_filter3_complex(cuda):
    import torch
    na, nb, n = 5, 7, 24
    rng = np.random.default_rng(2)
    fr = rng.standard_normal((na, nb, n))
    fi = rng.standard_normal((na, nb, n))
    z = np.stack([fr, fi], axis=0)  # interleaved (re, im) along a leading axis of 2
    o = _obj("gaussian", n, False)
    ft = to_dev(z, torch.float64, cuda)
    out = torch.empty_like(ft)
    o.filter3(ft, out, na, nb, -1, 1, prec="f64", complex_=True)
    torch.cuda.synchronize()
    got = from_dev(out, z.shape)
    for part in range(2):
        ref = X.filter_axis(z[part], 2, "gaussian", False, -1, 1)
        assert rel_l2(got[part], ref) <= 1e-14


def test_n1_and_errors(cuda):
    import torch
    o = _obj("gaussian", 1, False)
    f = torch.arange(12, dtype=torch.float64, device=cuda)
    out = torch.zeros_like(f)
    o.filter1(f, out, 3, 4, prec="f64")
    torch.cuda.synchronize()
    assert torch.equal(out, f)
    g = _obj("gaussian", 16, False)
    x = torch.zeros(16 * 4, dtype=torch.float64, device=cuda)
    with pytest.raises(P.GracefulExit):
        g.filter1(x, torch.empty_like(x), 2, 2, 2, 0, prec="f64")
    with pytest.raises(P.PdoError):
        P.gaussian().init(5, True)
