"""CPU checks of the HIP library without a GPU: it loads, exports every entry
point include/padeops_hip.h declares, and its host-side operator tables agree
with the oracle (no device work is done here)."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import oracle as O
import padeops_amd as P
import spike_model as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "padeops_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdo_[a-z0-9_A-Z]+)\s*\(", text)))


def test_library_exports_header():
    L = P.lib()
    names = declared_symbols()
    assert len(names) > 40
    for name in names:
        assert hasattr(L, name), name
        assert name in P._lib.SIGNATURES, name


def test_no_device_needed_for_init_codes():
    # init returns the reference's codes without touching the device
    assert P.cd06().init(5, 0.1, True) == 3
    assert P.cd06().init(7, 0.1, False) == 2
    assert P.cd10().init(7, 0.1, True) == 2
    assert P.cf90().init(9, False) == 7
    assert P.cd06stagg().init(4, 0.1) == 21
    assert P.cd10().init(8, 0.1, True) == 0


def _variants():
    out = []
    for n in (32, 40, 64, 512):
        out.append(("cd06", n, True, 0, 0, 0, O.CD06_D1))
        out.append(("cd06", n, True, 0, 0, 1, O.CD06_D2))
        out.append(("cd06", n, False, 0, 0, 0, O.CD06_D1))
        out.append(("cd06", n, False, 1, 1, 0, O.CD06_D1))
        for sch, kind, oop in (("cd10", 0, O.CD10_D1), ("cd10", 1, O.CD10_D2), ("cf90", 2, O.CF90)):
            out.append((sch, n, True, 0, 0, kind, oop))
            for b1, bn in itertools.product([0, 1, -1], repeat=2):
                out.append((sch, n, False, b1, bn, kind, oop))
    return out


def _obj(sch, n, dx, periodic, b1, bn):
    if sch == "cd06":
        o = P.cd06()
        assert o.init(n, dx, periodic, b1, bn) == 0
    elif sch == "cd10":
        o = P.cd10()
        assert o.init(n, dx, periodic) == 0
    else:
        o = P.cf90()
        assert o.init(n, periodic) == 0
    return o


@pytest.mark.parametrize("sch,n,periodic,b1,bn,kind,oop", _variants())
def test_operator_matches_oracle(sch, n, periodic, b1, bn, kind, oop):
    """The library's system A y = B f is the oracle's operator."""
    dx = 0.37 if sch != "cf90" else 1.0
    o = _obj(sch, n, dx, periodic, b1, bn)
    A, B = o.dense(kind, b1, bn)
    Ao, Bo = O.dense(oop, n, dx, periodic, b1, bn)
    M, Mo = np.linalg.solve(A, B), np.linalg.solve(Ao, Bo)
    assert np.abs(M - Mo).max() <= 1e-12 * max(1, np.abs(Mo).max())


@pytest.mark.parametrize("sch,n,periodic,b1,bn,kind,oop", _variants())
def test_partitioned_tables_reproduce_oracle(sch, n, periodic, b1, bn, kind, oop):
    """The kernels' partitioned solve, run in numpy from the exported tables
    (both the register-resident and the single-pencil layouts), equals the
    oracle applied to the same pencil."""
    dx = 0.37 if sch != "cf90" else 1.0
    o = _obj(sch, n, dx, periodic, b1, bn)
    A, B = o.dense(kind, b1, bn)
    f = np.random.default_rng(n + kind).standard_normal(B.shape[1])
    ref = O.apply1d(oop, f, dx, periodic, b1, bn)
    for general in (False, True):
        t = S.tables(o, kind, b1, bn, general)
        if not general and n % 32 == 0:
            assert t["m"] == 32 and t["P"] == n // 32
        d = B @ f
        for r, v in S.closure_rows(t, f, B.shape[1]).items():
            assert abs(v - d[r]) <= 1e-10 * (1 + abs(d[r]))
        y = S.solve(t, d, periodic)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind,oop", [(10, O.STG_D1_E2C), (11, O.STG_D1_C2E), (12, O.STG_D2_C2C),
                                      (13, O.STG_D2_E2E), (14, O.STG_IN_E2C), (15, O.STG_IN_C2E)])
def test_staggered_operator_matches_oracle(kind, oop):
    n, dz = 64, 0.21
    o = P.cd06stagg()
    assert o.init(n, dz) == 0
    A, B = o.dense(kind)
    Ao, Bo = O.dense(oop, n, dz, True)
    assert np.abs(np.linalg.solve(A, B) - np.linalg.solve(Ao, Bo)).max() < 1e-12 * max(
        1, np.abs(np.linalg.solve(Ao, Bo)).max())
