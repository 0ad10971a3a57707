"""numpy model of the partitioned ("spike") solve the HIP kernels run, driven
by the tables the library exports (pdo_op_tables).  Test helper only: it lets
the CPU suite check the host-built tables without a GPU."""
import ctypes

import numpy as np

from padeops_amd import _lib

NBMAX, WBMAX = 4, 8


def tables(op, kind, bc1=0, bcn=0, general=False):
    L = _lib.lib()
    meta = (ctypes.c_int * 6)()
    rc = L.pdo_op_tables(op._handle(), kind, bc1, bcn, int(general), meta, None, None, None, None)
    if rc:
        raise RuntimeError("pdo_op_tables rc=%d" % rc)
    m, P, K, nbl, nbr, xr = list(meta)
    mt = m + xr
    fac = np.zeros(4 * mt * 5)
    spk = np.zeros(4 * mt * 4)
    G = np.zeros(P * 2 * K * 2 * K * P)
    sp = np.zeros(2 * NBMAX * WBMAX)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.pdo_op_tables(op._handle(), kind, bc1, bcn, int(general), meta, fac.ctypes.data_as(dp),
                         spk.ctypes.data_as(dp), G.ctypes.data_as(dp), sp.ctypes.data_as(dp))
    assert rc == 0
    return dict(m=m, P=P, K=K, nbl=nbl, nbr=nbr, xr=xr, fac=fac.reshape(4, mt, 5), spk=spk.reshape(4, mt, 4),
                G=G.reshape(P, 2 * K, 2 * K * P), sp=sp.reshape(2, NBMAX, WBMAX))


def seg_type(periodic, P, s):
    if periodic:
        return 1
    if P == 1:
        return 3
    return 0 if s == 0 else (2 if s == P - 1 else 1)


def solve(t, d, periodic):
    """y = A^-1 d for one pencil, the way the kernels do it (segments start
    every m rows; the last one holds m + xr)."""
    mb, P, K, xr = t["m"], t["P"], t["K"], t["xr"]
    z = []
    for s in range(P):
        ty = seg_type(periodic, P, s)
        m = mb + (xr if s == P - 1 else 0)
        F = t["fac"][ty]
        v = d[s * mb:s * mb + m].astype(float).copy()
        for j in range(1, m):
            v[j] -= F[j, 0] * v[j - 1]
            if K == 2 and j >= 2:
                v[j] -= F[j, 1] * v[j - 2]
        for j in range(m - 1, -1, -1):
            x = v[j] * F[j, 2]
            if j + 1 < m:
                x -= F[j, 3] * v[j + 1]
            if K == 2 and j + 2 < m:
                x -= F[j, 4] * v[j + 2]
            v[j] = x
        z.append(v)
    zh = np.concatenate([np.concatenate([z[s][:K], z[s][len(z[s]) - K:]]) for s in range(P)])
    y = np.zeros(P * mb + xr)
    for s in range(P):
        ty = seg_type(periodic, P, s)
        m = len(z[s])
        g = t["G"][s] @ zh
        yl, yr = g[:K], g[K:]
        S = t["spk"][ty][:m]
        y[s * mb:s * mb + m] = z[s] - S[:, :K] @ yl - S[:, 2:2 + K] @ yr
    return y


def closure_rows(t, f, n_in):
    """RHS rows of the closure tables, as the kernels evaluate them."""
    out = {}
    n = t["m"] * t["P"] + t["xr"]
    for j in range(t["nbl"]):
        out[j] = t["sp"][0, j] @ f[:WBMAX]
    for jr in range(t["nbr"]):
        out[n - t["nbr"] + jr] = t["sp"][1, jr] @ f[n_in - WBMAX:n_in]
    return out
