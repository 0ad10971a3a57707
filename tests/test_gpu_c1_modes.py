"""The triply periodic 64^3 Fourier-mode workload on [0, 2 pi)^3 in FP64: a
scalar u = sum of 32 modes with integer wave vectors |k| <= 16 drawn from
seed 1234, amplitudes |k|^(-5/3) and phases uniform in [0, 2 pi).

cd06 and cd10 first and second derivatives and the cf90 filter along x, y
and z run on the GPU through the C ABI and are compared with
  1. the sum over the modes of each operator's closed-form Fourier symbol
     (a periodic compact operator is circulant, so a mode is an eigenvector):
     relative L2 <= 1e-12;
  2. the oracle on the same field: relative L2 <= 1e-12;
  3. the analytic derivatives (d/dx cos = -k sin, d2/dx2 cos = -k^2 cos;
     the filter against u itself): the GPU's error equals the schemes'
     truncation error that the symbols predict, to 1e-12.
"""
import numpy as np
import pytest

import oracle as O
import padeops_amd as P
from fields import fourier_modes, from_dev, rel_l2, to_dev
from test_oracle_pins import symbol

pytestmark = pytest.mark.gpu

N = 64
DX = 2 * np.pi / N

# (label, oracle operator, derivative order (0: filter), object factory, method stem)
OPS = [
    ("cd06_d1", O.CD06_D1, 1, lambda: P.cd06(), "dd"),
    ("cd06_d2", O.CD06_D2, 2, lambda: P.cd06(), "d2d"),
    ("cd10_d1", O.CD10_D1, 1, lambda: P.cd10(), "dd"),
    ("cd10_d2", O.CD10_D2, 2, lambda: P.cd10(), "d2d"),
    ("cf90", O.CF90, 0, lambda: P.cf90(), "filter"),
]

_FIELD = {}


def _field():
    if "u" not in _FIELD:
        _FIELD["u"], _FIELD["modes"] = fourier_modes(N, nmodes=32, kmax=16, seed=1234)
    return _FIELD["u"], _FIELD["modes"]


def _by_modes(modes, dir, fn):
    """sum over modes of a * Re(fn(k_dir) exp(i theta)), theta = k.x + phase"""
    x = np.arange(N) * DX
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((N, N, N))
    for kx, ky, kz, a, ph in modes:
        th = kx * X + ky * Y + kz * Z + ph
        k = (kx, ky, kz)[dir - 1]
        out += a * np.real(fn(k) * np.exp(1j * th))
    return out


@pytest.mark.parametrize("dir", [1, 2, 3])
@pytest.mark.parametrize("label,op,order,make,stem", OPS, ids=[o[0] for o in OPS])
def test_modes_fp64(cuda, label, op, order, make, stem, dir):
    import torch
    u, modes = _field()
    obj = make()
    assert (obj.init(N, True) if stem == "filter" else obj.init(N, DX, True)) == 0
    f = to_dev(u, torch.float64, cuda)
    out = torch.full_like(f, float("nan"))
    assert getattr(obj, "%s%d" % (stem, dir))(f, out, N, N) == 0
    torch.cuda.synchronize()
    got = from_dev(out, (N, N, N))

    by_symbol = _by_modes(modes, dir, lambda k: symbol(op, k * DX, DX))
    e_sym = rel_l2(got, by_symbol)
    e_orc = rel_l2(got, O.apply(op, u, dir, DX))
    if order == 0:
        exact = u
    else:
        exact = _by_modes(modes, dir, lambda k: (1j * k) ** order)
    e_exact = rel_l2(got, exact)
    e_trunc = rel_l2(by_symbol, exact)
    print("%s dir %d: vs symbols %.2e, vs oracle %.2e, vs analytic %.3e (scheme truncation %.3e)"
          % (label, dir, e_sym, e_orc, e_exact, e_trunc))
    assert e_sym <= 1e-12
    assert e_orc <= 1e-12
    assert abs(e_exact - e_trunc) <= 1e-12
    # the sixth- and tenth-order schemes resolve |k| <= 16 on 64 points to a
    # few percent at worst; the filter damps nothing below k dx = pi/2 by
    # more than a few percent either
    assert e_exact < 0.05
