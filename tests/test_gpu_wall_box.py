"""The wall-bounded box: 512 x 512 x 256 cells, periodic in x and y (side
2 pi), non-periodic in z (dz = dx, walls at z = 0 and z = pi), with the
log-law mean profile U(z) = ln(z / 1e-4) / 0.4 plus the seeded isotropic
turbulence (E(k) ~ k^4 exp(-2 (k/4)^2), rms 1); see
padeops_amd.synthetic.wall_bounded.

Operators, as the wall-normal direction of the igrid layout uses them:
  z: cd10 first and second derivatives with the one-sided closures
     (bc 0 at both walls), and the staggered cd06 operators ddz_C2E,
     InterpZ_C2E, ddz_E2C, InterpZ_E2C, d2dz2_E2E of two Pade6stagg boundary
     objects (even/sided ends for u, odd ends for w); the edge systems have
     257 rows;
  x, y: periodic cd10 first derivatives on 512-point lines.

Checks per operator:
  1. FP64 on the GPU against the oracle on 64 pencils of the box:
     relative L2 <= 1e-12;
  2. FP32 + refinement against FP64 on the same FP32-rounded input (the
     error of the FP32 solves): relative L2 <= 5e-7;
  3. FP32 + refinement against FP64 on the FP64 input (the whole mixed-precision
     error): reported, and no larger than the rounding floor of the FP32
     input (FP64 operator on the rounded input against FP64 on the exact
     one) plus twice the solve error of 2.  The log-law field has a mean of
     up to 26 and its wall-normal derivatives grow like 1/z, so the rounding
     of the stored input (about 1.6e-6 absolute) sets that floor, not the
     solves.
"""
import math

import numpy as np
import pytest

import oracle as O
import padeops_amd as P
from fields import rel_l2

pytestmark = pytest.mark.gpu

NX, NZ = 512, 256
DX = 2 * math.pi / NX
IDX = [0, 37, 100, 255, 256, 300, 450, 511]
# Pade6stagg boundary objects (isTopEven, isBotEven, isTopSided, isBotSided)
ES = (1, 1, 0, 1)
OO = (0, 0, 0, 0)


def _flags(te, be, ts, bs):
    return te * O.TOP_EVEN + be * O.BOT_EVEN + ts * O.TOP_SIDED + bs * O.BOT_SIDED


_F = {}


def _fields(cuda):
    if "uC" not in _F:
        import torch
        from padeops_amd.synthetic import wall_bounded
        _F["uC"], _F["uE"], _F["wE"] = wall_bounded(NX, NX, NZ, 4.0, 1234, torch.float64, cuda)
    return _F


def _zsub(t, nin):
    """the 8 x 8 pencils (IDX, IDX) of a z-layout field, as a Fortran (8, 8, nin) array"""
    a = t.view(nin, NX, NX)[:, IDX][:, :, IDX].cpu().numpy()  # (z, y, x)
    return np.asfortranarray(a.transpose(2, 1, 0))


def _xsub(t, nz):
    """x-lines at (y, z) = (IDX[i], i * nz / 8) as a Fortran (512, 8, 1) array"""
    v = t.view(nz, NX, NX)
    a = np.stack([v[(i * nz) // 8, IDX[i], :].cpu().numpy() for i in range(8)], axis=1)
    return np.asfortranarray(a.reshape(NX, 8, 1))


def _ysub(t, nz):
    """y-lines at (x, z) = (IDX[i], i * nz / 8) as a Fortran (8, 512, 1) array"""
    v = t.view(nz, NX, NX)
    a = np.stack([v[(i * nz) // 8, :, IDX[i]].cpu().numpy() for i in range(8)], axis=0)
    return np.asfortranarray(a.reshape(8, NX, 1))


def _cd10(periodic, n, dx):
    o = P.cd10()
    assert o.init(n, dx, periodic, 0, 0) == 0
    return o


def _stagg(flags):
    o = P.cd06stagg()
    assert o.init(NZ, DX, *(bool(b) for b in flags)) == 0
    return o


# (name, input field, input planes, output planes, apply(obj, f, out, prec), object, oracle on a subset,
#  subset extractor, second-derivative flag)
def _cases():
    c10z = lambda: _cd10(False, NZ, DX)  # noqa: E731
    c10x = lambda: _cd10(True, NX, DX)  # noqa: E731
    st = lambda fl: (lambda: _stagg(fl))  # noqa: E731
    return [
        ("cd10_dd3_walls", "uC", NZ, NZ, lambda o, f, g, p: o.dd3(f, g, NX, NX, 0, 0, prec=p), c10z,
         lambda a: O.apply(O.CD10_D1, a, 3, DX, False, 0, 0), "z", False),
        ("cd10_d2d3_walls", "uC", NZ, NZ, lambda o, f, g, p: o.d2d3(f, g, NX, NX, 0, 0, prec=p), c10z,
         lambda a: O.apply(O.CD10_D2, a, 3, DX, False, 0, 0), "z", True),
        ("stagg_ES_ddz_C2E", "uC", NZ, NZ + 1, lambda o, f, g, p: o.ddz_C2E(f, g, NX, NX, prec=p), st(ES),
         lambda a: O.stagg_np(O.STG_D1_C2E, a, DX, _flags(*ES), NZ), "z", False),
        ("stagg_ES_InterpZ_C2E", "uC", NZ, NZ + 1, lambda o, f, g, p: o.InterpZ_C2E(f, g, NX, NX, prec=p),
         st(ES), lambda a: O.stagg_np(O.STG_IN_C2E, a, DX, _flags(*ES), NZ), "z", False),
        ("stagg_ES_ddz_E2C", "uE", NZ + 1, NZ, lambda o, f, g, p: o.ddz_E2C(f, g, NX, NX, prec=p), st(ES),
         lambda a: O.stagg_np(O.STG_D1_E2C, a, DX, _flags(*ES), NZ), "z", False),
        ("stagg_ES_d2dz2_E2E", "uE", NZ + 1, NZ + 1, lambda o, f, g, p: o.d2dz2_E2E(f, g, NX, NX, prec=p),
         st(ES), lambda a: O.stagg_np(O.STG_D2_E2E, a, DX, _flags(*ES), NZ), "z", True),
        ("stagg_OO_ddz_E2C", "wE", NZ + 1, NZ, lambda o, f, g, p: o.ddz_E2C(f, g, NX, NX, prec=p), st(OO),
         lambda a: O.stagg_np(O.STG_D1_E2C, a, DX, _flags(*OO), NZ), "z", False),
        ("stagg_OO_InterpZ_E2C", "wE", NZ + 1, NZ, lambda o, f, g, p: o.InterpZ_E2C(f, g, NX, NX, prec=p),
         st(OO), lambda a: O.stagg_np(O.STG_IN_E2C, a, DX, _flags(*OO), NZ), "z", False),
        ("cd10_dd1_periodic", "uC", NZ, NZ, lambda o, f, g, p: o.dd1(f, g, NX, NZ, 0, 0, prec=p), c10x,
         lambda a: O.apply(O.CD10_D1, a, 1, DX, True), "x", False),
        ("cd10_dd2_periodic", "uC", NZ, NZ, lambda o, f, g, p: o.dd2(f, g, NX, NZ, 0, 0, prec=p), c10x,
         lambda a: O.apply(O.CD10_D1, a, 2, DX, True), "y", False),
    ]


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_wall_box(cuda, case):
    import torch
    name, key, nin, nout, apply, make, oracle, axis, second = case
    F = _fields(cuda)
    f64 = F[key]
    assert f64.numel() == NX * NX * nin
    obj = make()
    npts_out = NX * NX * nout
    g64 = torch.full((npts_out,), float("nan"), dtype=torch.float64, device=cuda)
    print("case", name, flush=True)
    assert apply(obj, f64, g64, "f64") == 0
    torch.cuda.synchronize()
    # 1. FP64 against the oracle on 64 pencils
    if axis == "z":
        e_orc = rel_l2(_zsub(g64, nout), oracle(_zsub(f64, nin)))
    elif axis == "x":
        e_orc = rel_l2(_xsub(g64, NZ), oracle(_xsub(f64, NZ)))
    else:
        e_orc = rel_l2(_ysub(g64, NZ), oracle(_ysub(f64, NZ)))
    # 2./3. FP32 + refinement
    f32 = f64.to(torch.float32)
    g32 = torch.full((npts_out,), float("nan"), dtype=torch.float32, device=cuda)
    assert apply(obj, f32, g32, "f32_refine") == 0
    gr = torch.full_like(g64, float("nan"))
    assert apply(obj, f32.double(), gr, "f64") == 0
    torch.cuda.synchronize()

    def rel(a, b):
        return float(torch.linalg.norm(a.double() - b) / torch.linalg.norm(b))

    e_solve = rel(g32, gr)
    e_total = rel(g32, g64)
    e_round = rel(gr, g64)
    print("%s: FP64 vs oracle %.2e; FP32+refine vs FP64 same input %.2e, vs FP64 %.2e (input rounding floor %.2e)"
          % (name, e_orc, e_solve, e_total, e_round))
    assert e_orc <= 1e-12
    assert e_solve <= 5e-7
    assert e_total <= e_round + 2 * e_solve
    del second
