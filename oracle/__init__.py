"""ctypes front-end of the CPU FP64 oracle (oracle/padeops_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product library
(padeops_amd) never imports it.

Arrays are numpy float64 in Fortran order (first index fastest), as in the
reference.  Operator codes are the oracle's own numbering.
"""
import ctypes
import os
import subprocess

import numpy as np

CD06_D1, CD06_D2, CD10_D1, CD10_D2, CF90 = 1, 2, 3, 4, 5
STG_D1_E2C, STG_D1_C2E, STG_D2_C2C, STG_D2_E2E, STG_IN_E2C, STG_IN_C2E = 10, 11, 12, 13, 14, 15
STG_D1_C2C, STG_D1_E2E = 16, 17
# flag bits of the non-periodic staggered object (cd06stagg%init_nonperiodic)
TOP_EVEN, BOT_EVEN, TOP_SIDED, BOT_SIDED = 1, 2, 4, 8

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile oracle/liboracle.so with gcc (OpenMP)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _setup(L):
    dp = ctypes.POINTER(ctypes.c_double)
    L.orc_apply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    L.orc_dense.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, dp, dp]
    L.orc_fft_c2c_dir.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int]
    L.orc_poisson_periodic.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 2 + [dp, dp]
    L.orc_modk.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]
    L.orc_full_rhs.argtypes = [ctypes.c_int, ctypes.c_double] + [dp] * 8
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_stagg_np.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, dp, dp]
    L.orc_stagg_np_dense.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
    return L


def load(path):
    """A separately built copy of the oracle library (e.g. a deliberately
    perturbed build used to show that the pins detect a wrong constant)."""
    return _setup(ctypes.CDLL(path))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _setup(ctypes.CDLL(_LIB))
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _io_len(op, n):
    nin = nout = n
    if op in (STG_D1_E2C, STG_IN_E2C):
        nin = n + 1
    if op in (STG_D1_C2E, STG_IN_C2E):
        nout = n + 1
    if op == STG_D2_E2E:
        nin = nout = n + 1
    return nin, nout


class OracleError(RuntimeError):
    def __init__(self, ierr):
        super().__init__("oracle ierr=%d" % ierr)
        self.ierr = ierr


def apply(op, f, dir, dx, periodic=True, bc1=0, bcn=0, n=None):
    """Apply operator `op` along axis `dir` (1, 2 or 3) of the 3-D array f."""
    f = np.asfortranarray(f, dtype=np.float64)
    if f.ndim == 1:
        f = f.reshape(-1, 1, 1, order="F") if dir == 1 else None
    shape = list(f.shape)
    if n is None:
        n = shape[dir - 1]
        if op in (STG_D1_E2C, STG_IN_E2C, STG_D2_E2E):
            n -= 1
    nin, nout = _io_len(op, n)
    oshape = list(shape)
    oshape[dir - 1] = nout
    out = np.zeros(oshape, dtype=np.float64, order="F")
    ierr = lib().orc_apply(op, n, dx, int(periodic), bc1, bcn, dir, shape[0], shape[1], shape[2], _p(f), _p(out))
    if ierr:
        raise OracleError(ierr)
    return out


def apply1d(op, f, dx, periodic=True, bc1=0, bcn=0):
    f = np.asarray(f, dtype=np.float64)
    return apply(op, f.reshape(-1, 1, 1, order="F"), 1, dx, periodic, bc1, bcn).reshape(-1)


def dense(op, n, dx, periodic=True, bc1=0, bcn=0, library=None):
    """(A, B): the oracle's LHS matrix (n x n) and RHS stencil matrix (n x n_in)."""
    nin, _ = _io_len(op, n)
    A = np.zeros((n, n))
    B = np.zeros((n, nin))
    ierr = (library or lib()).orc_dense(op, n, dx, int(periodic), bc1, bcn, _p(A), _p(B))
    if ierr:
        raise OracleError(ierr)
    return A, B


def fft_c2c(z, dir, sign=-1):
    """Unnormalized in-place-equivalent c2c transform of a complex Fortran array along dir."""
    z = np.asfortranarray(z, dtype=np.complex128).copy(order="F")
    n1, n2, n3 = z.shape
    lib().orc_fft_c2c_dir(dir, n1, n2, n3, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), sign)
    return z


def modk(kind, n, dx):
    k = np.zeros(n)
    lib().orc_modk(kind, n, dx, _p(k))
    return k


def poisson_periodic(rhs, dx, dy, dz, modk_kind=0, zero_policy=0):
    rhs = np.asfortranarray(rhs, dtype=np.float64)
    out = np.zeros_like(rhs, order="F")
    nx, ny, nz = rhs.shape
    lib().orc_poisson_periodic(nx, ny, nz, dx, dy, dz, modk_kind, zero_policy, _p(rhs), _p(out))
    return out


def full_rhs(u, v, w, L=2 * np.pi):
    """Returns (grad[9], uo, vo, wo, p) of the full RHS evaluation (see padeops_oracle.c)."""
    n = u.shape[0]
    u, v, w = (np.asfortranarray(a, dtype=np.float64) for a in (u, v, w))
    grad = np.zeros((9,) + u.shape, dtype=np.float64)  # C order over the 9 fields, each Fortran-ordered
    grad_buf = np.zeros(9 * u.size, dtype=np.float64)
    uo, vo, wo, p = (np.zeros_like(u, order="F") for _ in range(4))
    e = lib().orc_full_rhs(n, L, _p(u), _p(v), _p(w), _p(grad_buf), _p(uo), _p(vo), _p(wo), _p(p))
    if e:
        raise OracleError(e)
    for q in range(9):
        grad[q] = grad_buf[q * u.size:(q + 1) * u.size].reshape(u.shape, order="F")
    return grad, uo, vo, wo, p


def num_threads():
    return lib().orc_num_threads()


def stagg_lengths(op, n):
    """(n_in, n_out) of a non-periodic staggered operator on n cells."""
    nin = n + 1 if op in (STG_D1_E2C, STG_IN_E2C, STG_D1_E2E, STG_D2_E2E) else n
    nout = n + 1 if op in (STG_D1_C2E, STG_IN_C2E, STG_D1_E2E, STG_D2_E2E) else n
    return nin, nout


def stagg_np(op, f, dz, flags, n):
    """Non-periodic cd06stagg operator along z of f(n1, n2, n_in), flags as TOP_EVEN | ..."""
    f = np.asfortranarray(f, dtype=np.float64)
    n1, n2, nin = f.shape
    assert nin == stagg_lengths(op, n)[0]
    out = np.zeros((n1, n2, stagg_lengths(op, n)[1]), order="F")
    ierr = lib().orc_stagg_np(op, n, dz, flags, n1, n2, _p(f), _p(out))
    if ierr:
        raise OracleError(ierr)
    return out


def stagg_np_dense(op, n, dz, flags, library=None):
    nin, nout = stagg_lengths(op, n)
    A = np.zeros((nout, nout))
    B = np.zeros((nout, nin))
    ierr = (library or lib()).orc_stagg_np_dense(op, n, dz, flags, _p(A), _p(B))
    if ierr:
        raise OracleError(ierr)
    return A, B
