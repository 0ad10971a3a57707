"""GPU: the derivatives / filters dispatch types and the operators module
(derivatives.F90, filters.F90, ffts.F90, operators.F90) against oracle
compositions: cd10 / cd06 / Fourier derivatives per direction, cf90 /
Gaussian / least-squares filters per direction, gradient, curl, divergence
(with its reversed boundary signs) and filter3D."""
import numpy as np
import pytest

import oracle as O
import oracle.explicit as X
import padeops_amd as P
from fields import from_dev, random_field, rel_l2, to_dev

pytestmark = pytest.mark.gpu

L = 2 * np.pi


def _dev(a, dt, cuda):
    return to_dev(a, dt, cuda)


def _ref_der(method, f, d, dx, periodic, bc1, bcn, order):
    if method == "four":
        return X.fourier_deriv(f, d - 1, dx, order)
    if method == "cd10":
        return O.apply(O.CD10_D1 if order == 1 else O.CD10_D2, f, d, dx, periodic, bc1, bcn)
    return O.apply(O.CD06_D1, f, d, dx, periodic)


def _ref_fil(method, f, d, periodic, bc1, bcn):
    if method == "cf90":
        return O.apply(O.CF90, f, d, 1.0, periodic, bc1, bcn)
    return X.filter_axis(f, d - 1, method, periodic, bc1, bcn)


CASES = [
    # (shape, periodic xyz, methods xyz)
    ((32, 16, 32), (False, True, True), ("cd10", "cd06", "four")),
    ((16, 32, 16), (True, True, False), ("four", "four", "cd10")),
    ((24, 20, 32), (True, False, True), ("cd06", "cd10", "four")),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec", ["f64", "f32_refine"])
def test_derivatives_dispatch(cuda, case, prec):
    import torch
    shape, per, meth = CASES[case]
    dt = torch.float64 if prec == "f64" else torch.float32
    bar = 1e-12 if prec == "f64" else 2e-6
    dxs = [L / n for n in shape]
    der = P.derivatives()
    der.init(*shape, *dxs, *per, *meth)
    assert (der.getMethodx(), der.getMethody(), der.getMethodz()) == meth
    f = random_field(shape, seed=case)
    ft = _dev(f, dt, cuda)
    fin = from_dev(ft, shape)  # the input as stored
    for d, name in ((1, "ddx"), (2, "ddy"), (3, "ddz")):
        bc = (1, -1) if (meth[d - 1] == "cd10" and not per[d - 1]) else (0, 0)
        out = torch.full_like(ft, float("nan"))
        getattr(der, name)(ft, out, *bc, prec=prec)
        torch.cuda.synchronize()
        ref = _ref_der(meth[d - 1], fin, d, dxs[d - 1], per[d - 1], *bc, 1)
        e = rel_l2(from_dev(out, shape), ref)
        print(case, prec, name, meth[d - 1], e)
        assert e <= bar
    for d, name in ((1, "d2dx2"), (2, "d2dy2"), (3, "d2dz2")):
        out = torch.full_like(ft, float("nan"))
        if meth[d - 1] == "cd06":
            with pytest.raises(P.PdoError):
                getattr(der, name)(ft, out, prec=prec)
            continue
        bc = (1, 1) if (meth[d - 1] == "cd10" and not per[d - 1]) else (0, 0)
        getattr(der, name)(ft, out, *bc, prec=prec)
        torch.cuda.synchronize()
        ref = _ref_der(meth[d - 1], fin, d, dxs[d - 1], per[d - 1], *bc, 2)
        e = rel_l2(from_dev(out, shape), ref)
        print(case, prec, name, meth[d - 1], e)
        assert e <= bar * 10


def test_fourier_exact_and_sizes(cuda):
    import torch
    n = 32
    shape = (n, n, n)
    dx = L / n
    x = np.arange(n) * dx
    X3, Y3, Z3 = np.meshgrid(x, x, x, indexing="ij")
    f = np.sin(3 * X3) * np.cos(5 * Y3) + np.cos(7 * Z3)
    der = P.derivatives()
    der.init(n, n, n, dx, dx, dx, True, True, True, "four", "four", "four")
    ft = _dev(f, torch.float64, cuda)
    out = torch.empty_like(ft)
    der.ddx(ft, out, prec="f64")
    torch.cuda.synchronize()
    assert np.abs(from_dev(out, shape) - 3 * np.cos(3 * X3) * np.cos(5 * Y3)).max() < 1e-12
    der.d2dz2(ft, out, prec="f64")
    torch.cuda.synchronize()
    assert np.abs(from_dev(out, shape) + 49 * np.cos(7 * Z3)).max() < 1e-10
    with pytest.raises(P.PdoError):
        P.derivatives().init(24, 8, 8, dx, dx, dx, True, True, True, "four", "cd06", "cd06")
    with pytest.raises(P.PdoError):
        P.derivatives().init(16, 8, 8, dx, dx, dx, True, True, True, "cheb", "cd06", "cd06")


FCASES = [((32, 16, 24), (False, True, True), ("cf90", "gaussian", "lstsq")),
          ((16, 24, 32), (True, False, False), ("lstsq", "cf90", "gaussian"))]


@pytest.mark.parametrize("case", range(len(FCASES)))
def test_filters_dispatch(cuda, case):
    import torch
    shape, per, meth = FCASES[case]
    fil = P.filters()
    fil.init(*shape, *per, *meth)
    assert (fil.getMethodx(), fil.getMethody(), fil.getMethodz()) == meth
    f = random_field(shape, seed=10 + case)
    ft = _dev(f, torch.float64, cuda)
    for d, name in ((1, "filterx"), (2, "filtery"), (3, "filterz")):
        bc = (0, 0) if per[d - 1] else (1, -1)
        if meth[d - 1] == "cf90" and not per[d - 1]:
            bc = (0, 1)
        out = torch.full_like(ft, float("nan"))
        getattr(fil, name)(ft, out, *bc, prec="f64")
        torch.cuda.synchronize()
        ref = _ref_fil(meth[d - 1], f, d, per[d - 1], *bc)
        e = rel_l2(from_dev(out, shape), ref)
        print(case, name, meth[d - 1], e)
        assert e <= 1e-12


def test_operators(cuda):
    import torch
    shape = (32, 16, 32)
    per = (False, True, True)
    meth = ("cd10", "cd06", "four")
    dxs = [L / n for n in shape]
    der = P.derivatives()
    der.init(*shape, *dxs, *per, *meth)
    rng = np.random.default_rng(7)
    u, v, w = (rng.standard_normal(shape) for _ in range(3))
    U = [_dev(a, torch.float64, cuda) for a in (u, v, w)]
    N = u.size
    xbc = (1, -1)

    def D(f, d, sign=1):
        bc = (sign * xbc[0], sign * xbc[1]) if d == 1 else (0, 0)
        return _ref_der(meth[d - 1], f, d, dxs[d - 1], per[d - 1], *bc, 1)

    g = [torch.empty_like(U[0]) for _ in range(3)]
    P.gradient(der, U[0], *g, x_bc=xbc, prec="f64")
    curlu = torch.empty(3 * N, dtype=torch.float64, device=cuda)
    P.curl(der, *U, curlu, x_bc=xbc, prec="f64")
    div = torch.empty_like(U[0])
    P.divergence(der, *U, div, x_bc=xbc, prec="f64")
    torch.cuda.synchronize()
    for q in range(3):
        assert rel_l2(from_dev(g[q], shape), D(u, q + 1)) <= 1e-12
    c = [from_dev(curlu[q * N:(q + 1) * N], shape) for q in range(3)]
    assert rel_l2(c[0], D(w, 2) - D(v, 3)) <= 1e-12
    assert rel_l2(c[1], D(u, 3) - D(w, 1)) <= 1e-12
    assert rel_l2(c[2], D(v, 1) - D(u, 2)) <= 1e-12
    # divergence: reversed signs of the boundary options
    assert rel_l2(from_dev(div, shape), D(v, 2, -1) + D(u, 1, -1) + D(w, 3, -1)) <= 1e-12


@pytest.mark.parametrize("numtimes", [1, 2])
def test_filter3d(cuda, numtimes):
    import torch
    shape = (32, 16, 24)
    per = (False, True, True)
    meth = ("gaussian", "cf90", "lstsq")
    fil = P.filters()
    fil.init(*shape, *per, *meth)
    f = random_field(shape, seed=21)
    ft = _dev(f, torch.float64, cuda)
    P.filter3D(fil, ft, numtimes, x_bc=(1, 1), prec="f64")
    torch.cuda.synchronize()
    ref = f
    for d, bc in ((2, (0, 0)), (1, (1, 1)), (3, (0, 0))):
        for _ in range(numtimes):
            ref = _ref_fil(meth[d - 1], ref, d, per[d - 1], *bc)
    assert rel_l2(from_dev(ft, shape), ref) <= 1e-12
