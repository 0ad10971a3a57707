"""MI355X-native PadeOps hot path: compact derivatives, compact filter and the
periodic FFT Poisson projection, as hand-written gfx950 HIP kernels behind a C
ABI (include/padeops_hip.h).

This module mirrors the reference's Fortran type-bound API for the path:

    reference (Fortran)                          here (Python over the C ABI)
    type(cd06);    ierr = d%init(n,dx,per,b1,bn) d = cd06(); ierr = d.init(...)
    call d%dd1(f, df, na, nb)                    d.dd1(f, df, na, nb)
    type(cd10) dd1..3 / d2d1..3 (bc1, bcn)       cd10().dd1(f, df, na, nb, bc1, bcn)
    type(cf90) filter1..3 (bc1, bcn)             cf90().filter1(f, fil, na, nb, bc1, bcn)
    type(cd06stagg) ddz_E2C ... InterpZ_C2E      cd06stagg().ddz_E2C(fE, dfC, n1, n2)
    type(PoissonPeriodic) poisson_solve          PoissonPeriodic().poisson_solve(rhs, f)
    type(spectral) fft/ifft/mTimes_ik/dealias    spectral().fft(f, fhat) ...
    type(padepoisson) PressureProjection         padepoisson().PressureProjection(uh, vh, wh)

Field arguments are torch tensors resident on the GPU (PyTorch supplies the
device memory and stream only), holding the reference's column-major layout
in a flat, contiguous buffer: f(n, na, nb) for x, f(na, n, nb) for y,
f(na, nb, n) for z.  float64 tensors run the FP64 path; float32 tensors run
the precision chosen by ``prec`` ("f32_refine" by default).

Errors: the reference's init codes are returned (2, 3, 7, 21); a boundary
flag outside {-1,0,1} raises GracefulExit(code=324) like the reference's
GracefulExit(..., 324); library failures raise PdoError.
"""
import ctypes

from . import _lib
from ._lib import LibraryMissing, lib  # noqa: F401

PDO_F64, PDO_F32_REFINE, PDO_F32_STORE, PDO_F32 = 0, 1, 2, 3
PREC = {"f64": PDO_F64, "f32_refine": PDO_F32_REFINE, "f32_store": PDO_F32_STORE, "f32": PDO_F32}
MODK = {"fourier": 0, "cd06": 1, "cd06stagg": 2, "cd10": 3}


class PdoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("padeops_hip error %d: %s" % (code, msg))
        self.code = code


class GracefulExit(RuntimeError):
    """Raised where the reference calls GracefulExit (an MPI abort there)."""

    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


def _check(rc):
    if rc == 0:
        return 0
    msg = _lib.last_error()
    if rc == 324:
        raise GracefulExit(324, msg or "Incorrect boundary specification")
    if rc > 0:
        return rc
    raise PdoError(rc, msg)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("field buffers must be GPU tensors")
    if not t.is_contiguous():
        raise ValueError("field buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _prec(f, prec):
    import torch
    if f.dtype in (torch.float64, torch.complex128):
        return PDO_F64
    if f.dtype in (torch.float32, torch.complex64):
        return PREC[prec] if prec != "f64" else PDO_F32_REFINE
    raise TypeError("fields must be float64/complex128 or float32/complex64")


def _reals(t):
    """Number of real words a (real or complex) tensor holds."""
    return t.numel() * (2 if t.is_complex() else 1)


def _check_sizes(f, out, n_in, n_out, na, nb):
    if f.numel() < n_in * na * nb or out.numel() < n_out * na * nb:
        raise ValueError("buffer smaller than the declared shape")
    if f.dtype != out.dtype:
        raise TypeError("input and output dtypes differ")


class _Op:
    _init_fn = None

    def __init__(self):
        self._h = ctypes.c_void_p()

    def _handle(self):
        if not self._h:
            raise PdoError(-1, "object not initialised")
        return self._h

    def destroy(self):
        if self._h:
            lib().pdo_op_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def dense(self, kind, bc1=0, bcn=0):
        """(A, B) dense system of one variant (host-side description, no GPU)."""
        import numpy as np
        ns, ni = ctypes.c_int(), ctypes.c_int()
        _check(lib().pdo_op_dims(self._handle(), kind, ctypes.byref(ns), ctypes.byref(ni)))
        A = np.zeros((ns.value, ns.value))
        B = np.zeros((ns.value, ni.value))
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().pdo_op_dense(self._handle(), kind, bc1, bcn, A.ctypes.data_as(dp), B.ctypes.data_as(dp)))
        return A, B

    def apply_noalloc(self, kind, dir, f, out, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        """pdo_op_apply_noalloc: the operator `kind` (numbering of dense())
        along `dir` with no device allocation; the variant must have been
        built by one ordinary call first.  na counts real words (2 n1 for
        complex staggered data)."""
        ns, ni = ctypes.c_int(), ctypes.c_int()
        _check(lib().pdo_op_dims(self._handle(), kind, ctypes.byref(ns), ctypes.byref(ni)))
        # edge outputs of the staggered operators hold n + 1 points
        n_out = self.n + 1 if kind in (11, 13, 15, 17) else ns.value
        if _reals(f) < ni.value * na * nb or _reals(out) < n_out * na * nb:
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_op_apply_noalloc(self._handle(), kind, dir, _prec(f, prec), _ptr(f), _ptr(out), na,
                                                 nb, bc1, bcn, _stream(stream)))

    def partition(self, kind, bc1=0, bcn=0):
        m, P = ctypes.c_int(), ctypes.c_int()
        _check(lib().pdo_op_partition(self._handle(), kind, bc1, bcn, ctypes.byref(m), ctypes.byref(P)))
        return m.value, P.value


def _apply(fn, h, f, df, na, nb, n_in, n_out, prec, stream, *bc):
    _check_sizes(f, df, n_in, n_out, na, nb)
    return _check(fn(h, _prec(f, prec), _ptr(f), _ptr(df), na, nb, *bc, _stream(stream)))


class cd06(_Op):
    """type(cd06), src/derivatives/cd06.F90."""

    def init(self, n, dx, periodic, bc1=0, bcn=0):
        self.n = n
        return _check(lib().pdo_cd06_init(ctypes.byref(self._h), n, dx, int(periodic), bc1, bcn))

    def dd1(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_dd1, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)

    def dd2(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_dd2, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)

    def dd3(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_dd3, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)

    def d2d1(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_d2d1, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)

    def d2d2(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_d2d2, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)

    def d2d3(self, f, df, na, nb, prec="f32_refine", stream=None):
        return _apply(lib().pdo_cd06_d2d3, self._handle(), f, df, na, nb, self.n, self.n, prec, stream)


class cd10(_Op):
    """type(cd10), src/derivatives/cd10.F90."""

    def init(self, n, dx, periodic, bc1=0, bcn=0):
        self.n = n
        return _check(lib().pdo_cd10_init(ctypes.byref(self._h), n, dx, int(periodic), bc1, bcn))

    def _ap(self, name, f, df, na, nb, bc1, bcn, prec, stream):
        return _apply(getattr(lib(), name), self._handle(), f, df, na, nb, self.n, self.n, prec, stream, bc1, bcn)

    def dd1(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_dd1", f, df, na, nb, bc1, bcn, prec, stream)

    def dd2(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_dd2", f, df, na, nb, bc1, bcn, prec, stream)

    def dd3(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_dd3", f, df, na, nb, bc1, bcn, prec, stream)

    def d2d1(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_d2d1", f, df, na, nb, bc1, bcn, prec, stream)

    def d2d2(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_d2d2", f, df, na, nb, bc1, bcn, prec, stream)

    def d2d3(self, f, df, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cd10_d2d3", f, df, na, nb, bc1, bcn, prec, stream)


class cf90(_Op):
    """type(cf90), src/filters/cf90.F90."""

    def init(self, n, periodic):
        self.n = n
        return _check(lib().pdo_cf90_init(ctypes.byref(self._h), n, int(periodic)))

    def _ap(self, name, f, df, na, nb, bc1, bcn, prec, stream):
        return _apply(getattr(lib(), name), self._handle(), f, df, na, nb, self.n, self.n, prec, stream, bc1, bcn)

    def filter1(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cf90_filter1", f, fil, na, nb, bc1, bcn, prec, stream)

    def filter2(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_cf90_filter2", f, fil, na, nb, bc1, bcn, prec, stream)

    def filter3(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None, complex_=False):
        name = "pdo_cf90_filter3_cmplx" if complex_ else "pdo_cf90_filter3"
        k = 2 if complex_ else 1
        _check_sizes(f, fil, self.n * k, self.n * k, na, nb)
        return _check(getattr(lib(), name)(self._handle(), _prec(f, prec), _ptr(f), _ptr(fil), na, nb, bc1, bcn,
                                           _stream(stream)))


class _XFilter(_Op):
    """Explicit 9-point filters; filter1/2/3 as cf90's."""
    _init_name = None

    def init(self, n, periodic):
        self.n = n
        return _check(getattr(lib(), self._init_name)(ctypes.byref(self._h), n, int(periodic)))

    def destroy(self):
        if self._h:
            lib().pdo_xfilter_destroy(self._h)
            self._h = ctypes.c_void_p()

    def _ap(self, name, f, df, na, nb, bc1, bcn, prec, stream):
        return _apply(getattr(lib(), name), self._handle(), f, df, na, nb, self.n, self.n, prec, stream, bc1, bcn)

    def dense_matrix(self, bc1=0, bcn=0):
        """n x n matrix of the filter (host-side, no GPU)."""
        import numpy as np
        W = np.zeros((self.n, self.n))
        _check(lib().pdo_xfilter_dense(self._handle(), bc1, bcn, W.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return W

    def filter1(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_xfilter_filter1", f, fil, na, nb, bc1, bcn, prec, stream)

    def filter2(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_xfilter_filter2", f, fil, na, nb, bc1, bcn, prec, stream)

    def filter3(self, f, fil, na, nb, bc1=0, bcn=0, prec="f32_refine", stream=None, complex_=False):
        name = "pdo_xfilter_filter3_cmplx" if complex_ else "pdo_xfilter_filter3"
        k = 2 if complex_ else 1
        _check_sizes(f, fil, self.n * k, self.n * k, na, nb)
        return _check(getattr(lib(), name)(self._handle(), _prec(f, prec), _ptr(f), _ptr(fil), na, nb, bc1, bcn,
                                           _stream(stream)))


class gaussian(_XFilter):
    """type(gaussian), src/filters/gaussian.F90: the width-4dx Gaussian filter."""
    _init_name = "pdo_gaussian_init"


class lstsq(_XFilter):
    """type(lstsq), src/filters/lstsq.F90: the least-squares filter (no bc options)."""
    _init_name = "pdo_lstsq_init"


class cd06stagg(_Op):
    """type(cd06stagg), src/derivatives/cd06stagg.F90: init(n, dz) is the
    periodic object, init(n, dz, isTopEven, isBotEven[, isTopSided,
    isBotSided]) the non-periodic one (cd06stagg.F90:170-232)."""

    def init(self, n, dz, isTopEven=None, isBotEven=None, isTopSided=False, isBotSided=False):
        self.n = n
        if isTopEven is None and isBotEven is None:
            self.periodic = True
            return _check(lib().pdo_cd06stagg_init_periodic(ctypes.byref(self._h), n, dz))
        self.periodic = False
        return _check(lib().pdo_cd06stagg_init(ctypes.byref(self._h), n, dz, int(bool(isTopEven)),
                                               int(bool(isBotEven)), int(bool(isTopSided)), int(bool(isBotSided))))

    def _ap(self, name, fi, fo, n1, n2, n_in, n_out, is_complex, prec, stream):
        k = 2 if is_complex else 1
        if fi.numel() < k * n1 * n2 * n_in or fo.numel() < k * n1 * n2 * n_out:
            raise ValueError("buffer smaller than the declared shape")
        return _check(getattr(lib(), name)(self._handle(), _prec(fi, prec), _ptr(fi), _ptr(fo), n1, n2,
                                           int(is_complex), _stream(stream)))

    def ddz_E2C(self, fE, dfC, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_ddz_E2C", fE, dfC, n1, n2, self.n + 1, self.n, is_complex, prec, stream)

    def ddz_C2E(self, fC, dfE, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_ddz_C2E", fC, dfE, n1, n2, self.n, self.n + 1, is_complex, prec, stream)

    def d2dz2_C2C(self, fC, d2fC, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_d2dz2_C2C", fC, d2fC, n1, n2, self.n, self.n, is_complex, prec, stream)

    def d2dz2_E2E(self, fE, d2fE, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_d2dz2_E2E", fE, d2fE, n1, n2, self.n + 1, self.n + 1, is_complex, prec,
                        stream)

    def InterpZ_E2C(self, fE, ifC, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_interpz_E2C", fE, ifC, n1, n2, self.n + 1, self.n, is_complex, prec, stream)

    def InterpZ_C2E(self, fC, ifE, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("pdo_cd06stagg_interpz_C2E", fC, ifE, n1, n2, self.n, self.n + 1, is_complex, prec, stream)

    def ddz_C2C(self, fC, dfC, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        """non-periodic objects only"""
        return self._ap("pdo_cd06stagg_ddz_C2C", fC, dfC, n1, n2, self.n, self.n, is_complex, prec, stream)

    def ddz_E2E(self, fE, dfE, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        """non-periodic objects only"""
        return self._ap("pdo_cd06stagg_ddz_E2E", fE, dfE, n1, n2, self.n + 1, self.n + 1, is_complex, prec, stream)


def _wavenums(n, d):
    """GetWaveNums (PoissonPeriodic.F90:226-242): (-pi + i 2 pi / (n - n mod 2)) / d,
    i = 0 .. n-1, then ifftshift."""
    import numpy as np
    t = (-np.pi + np.arange(n) * 2 * np.pi / (n - n % 2)) / d
    return np.fft.ifftshift(t)


class PoissonPeriodic:
    """type(PoissonPeriodic), src/utilities/PoissonPeriodic.F90 (dir_id = 1)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, dx, dy, dz, nx, ny, nz, modk=("fourier", "fourier", "fourier"), zero_policy=0):
        """modk[d] names a built-in modified wavenumber ("fourier", "cd06",
        "cd06stagg", "cd10") or is a Get_ModK callback: a callable taking the
        wavenumbers times the spacing (numpy FP64) and returning their modified
        values, applied as PoissonPeriodic.F90:188-204 does."""
        import numpy as np
        self.shape = (nx, ny, nz)
        self.destroy()
        if any(callable(m) for m in modk):
            ks = []
            for n, d, m, cnt in ((nx, dx, modk[0], nx // 2 + 1), (ny, dy, modk[1], ny), (nz, dz, modk[2], nz)):
                k = _wavenums(n, d)[:cnt]
                if callable(m):
                    k = np.ascontiguousarray(np.asarray(m(k * d), dtype=np.float64).reshape(cnt)) / d
                elif (MODK[m] if isinstance(m, str) else int(m)) != 0:
                    raise ValueError("mix callbacks only with Fourier directions")
                ks.append(np.ascontiguousarray(k, dtype=np.float64))
            dp = ctypes.POINTER(ctypes.c_double)
            return _check(lib().pdo_poisson_periodic_init_k(ctypes.byref(self._h), nx, ny, nz,
                                                            *[k.ctypes.data_as(dp) for k in ks], zero_policy))
        mk = [MODK[m] if isinstance(m, str) else int(m) for m in modk]
        return _check(lib().pdo_poisson_periodic_init(ctypes.byref(self._h), nx, ny, nz, dx, dy, dz, mk[0], mk[1],
                                                      mk[2], zero_policy))

    def poisson_solve(self, rhs, f=None, prec="f32_refine", stream=None):
        """In place when f is None (poisson_solve_inplace), else out of place."""
        out = rhs if f is None else f
        n = self.shape[0] * self.shape[1] * self.shape[2]
        if rhs.numel() < n or out.numel() < n:
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_poisson_periodic_solve(self._h, _prec(rhs, prec), _ptr(rhs), _ptr(out),
                                                       _stream(stream)))

    def destroy(self):
        if self._h:
            lib().pdo_poisson_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


DER_METHOD = {"cd10": 1, "cd06": 2, "four": 3, "cheb": 4}
FIL_METHOD = {"cf90": 1, "gaussian": 2, "lstsq": 3}


def _ints(vals):
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def _field_check(obj, *ts):
    n = obj.nx * obj.ny * obj.nz
    for t in ts:
        if t is not None and t.numel() < n:
            raise ValueError("buffer smaller than the (nx, ny, nz) field")


class derivatives:
    """type(derivatives), src/derivatives/derivatives.F90 (serial init: the
    pencil sizes of every direction are (nx, ny, nz))."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, nx, ny, nz, dx, dy, dz, periodic_x, periodic_y, periodic_z, method_x, method_y, method_z):
        self.nx, self.ny, self.nz = nx, ny, nz
        meth = [DER_METHOD.get(m, 0) for m in (method_x, method_y, method_z)]
        return _check(lib().pdo_derivatives_init(ctypes.byref(self._h), nx, ny, nz, dx, dy, dz,
                                                 _ints([periodic_x, periodic_y, periodic_z]), _ints(meth)))

    def _method(self, d):
        m = ctypes.c_int()
        _check(lib().pdo_derivatives_method(self._h, d, ctypes.byref(m)))
        return {v: k for k, v in DER_METHOD.items()}[m.value]

    def getMethodx(self):
        return self._method(1)

    def getMethody(self):
        return self._method(2)

    def getMethodz(self):
        return self._method(3)

    def _ap(self, name, f, df, bc1, bcn, prec, stream):
        _field_check(self, f, df)
        return _check(getattr(lib(), name)(self._h, _prec(f, prec), _ptr(f), _ptr(df), bc1, bcn, _stream(stream)))

    def ddx(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_ddx", f, df, bc1, bcn, prec, stream)

    def ddy(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_ddy", f, df, bc1, bcn, prec, stream)

    def ddz(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_ddz", f, df, bc1, bcn, prec, stream)

    def d2dx2(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_d2dx2", f, df, bc1, bcn, prec, stream)

    def d2dy2(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_d2dy2", f, df, bc1, bcn, prec, stream)

    def d2dz2(self, f, df, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_derivatives_d2dz2", f, df, bc1, bcn, prec, stream)

    def destroy(self):
        if self._h:
            lib().pdo_derivatives_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class filters:
    """type(filters), src/filters/filters.F90 (serial init)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, nx, ny, nz, periodicx, periodicy, periodicz, methodx, methody, methodz):
        self.nx, self.ny, self.nz = nx, ny, nz
        meth = [FIL_METHOD.get(m, 0) for m in (methodx, methody, methodz)]
        return _check(lib().pdo_filters_init(ctypes.byref(self._h), nx, ny, nz,
                                             _ints([periodicx, periodicy, periodicz]), _ints(meth)))

    def _method(self, d):
        m = ctypes.c_int()
        _check(lib().pdo_filters_method(self._h, d, ctypes.byref(m)))
        return {v: k for k, v in FIL_METHOD.items()}[m.value]

    def getMethodx(self):
        return self._method(1)

    def getMethody(self):
        return self._method(2)

    def getMethodz(self):
        return self._method(3)

    def _ap(self, name, f, ff, bc1, bcn, prec, stream):
        _field_check(self, f, ff)
        return _check(getattr(lib(), name)(self._h, _prec(f, prec), _ptr(f), _ptr(ff), bc1, bcn, _stream(stream)))

    def filterx(self, f, ff, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_filters_filterx", f, ff, bc1, bcn, prec, stream)

    def filtery(self, f, ff, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_filters_filtery", f, ff, bc1, bcn, prec, stream)

    def filterz(self, f, ff, bc1=0, bcn=0, prec="f32_refine", stream=None):
        return self._ap("pdo_filters_filterz", f, ff, bc1, bcn, prec, stream)

    def destroy(self):
        if self._h:
            lib().pdo_filters_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _bcs(x_bc, y_bc, z_bc):
    return _ints(list(x_bc) + list(y_bc) + list(z_bc))


def gradient(der, f, dfdx, dfdy, dfdz, x_bc=(0, 0), y_bc=(0, 0), z_bc=(0, 0), prec="f32_refine", stream=None):
    """operators%gradient (src/utilities/operators.F90:18-53)."""
    _field_check(der, f, dfdx, dfdy, dfdz)
    return _check(lib().pdo_operators_gradient(der._h, _prec(f, prec), _ptr(f), _ptr(dfdx), _ptr(dfdy), _ptr(dfdz),
                                               _bcs(x_bc, y_bc, z_bc), _stream(stream)))


def curl(der, u, v, w, curlu, x_bc=(0, 0), y_bc=(0, 0), z_bc=(0, 0), prec="f32_refine", stream=None):
    """operators%curl (operators.F90:55-116); curlu holds 3 consecutive fields."""
    _field_check(der, u, v, w)
    if curlu.numel() < 3 * der.nx * der.ny * der.nz:
        raise ValueError("curl output must hold three fields")
    return _check(lib().pdo_operators_curl(der._h, _prec(u, prec), _ptr(u), _ptr(v), _ptr(w), _ptr(curlu),
                                           _bcs(x_bc, y_bc, z_bc), _stream(stream)))


def divergence(der, u, v, w, div, x_bc=(0, 0), y_bc=(0, 0), z_bc=(0, 0), prec="f32_refine", stream=None):
    """operators%divergence (operators.F90:118-156)."""
    _field_check(der, u, v, w, div)
    return _check(lib().pdo_operators_divergence(der._h, _prec(u, prec), _ptr(u), _ptr(v), _ptr(w), _ptr(div),
                                                 _bcs(x_bc, y_bc, z_bc), _stream(stream)))


def filter3D(fil, arr, numtimes=1, x_bc=(0, 0), y_bc=(0, 0), z_bc=(0, 0), prec="f32_refine", stream=None):
    """operators%filter3D (operators.F90:158-223): in place."""
    _field_check(fil, arr)
    return _check(lib().pdo_operators_filter3d(fil._h, _prec(arr, prec), _ptr(arr), numtimes,
                                               _bcs(x_bc, y_bc, z_bc), _stream(stream)))


class FullRHS:
    """The benchmarked RHS evaluation (see pdo_rhs_eval in include/padeops_hip.h)."""

    def __init__(self, n, L, prec="f64"):
        self._h = ctypes.c_void_p()
        self.n = n
        self.prec = PREC[prec]
        _check(lib().pdo_rhs_init(ctypes.byref(self._h), n, L, self.prec))

    def eval(self, u, v, w, grad, uo, vo, wo, stream=None):
        return _check(lib().pdo_rhs_eval(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(grad), _ptr(uo), _ptr(vo),
                                         _ptr(wo), _stream(stream)))

    def pressure_ptr(self):
        p = ctypes.c_void_p()
        _check(lib().pdo_rhs_pressure(self._h, ctypes.byref(p)))
        return p.value

    def max_divergence(self, u, v, w, stream=None):
        d = ctypes.c_double()
        _check(lib().pdo_rhs_max_divergence(self._h, _ptr(u), _ptr(v), _ptr(w), ctypes.byref(d), _stream(stream)))
        return d.value

    def destroy(self):
        if self._h:
            lib().pdo_rhs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class spectral:
    """type(spectral) of src/incompressible/spectral.F90 as igrid uses it:
    2D (x, y) transforms of (nx, ny, nz) fields into (nx/2+1, ny, nz) spectra,
    i k multiplication and dealiasing (see pdo_spectral_* in the C header)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, nx, ny, nz, dx, dy, dz, fixOddball=False, dealiasF=2.0 / 3.0, init_periodicInZ=False):
        self.shape = (nx, ny, nz)
        self.nxh = nx // 2 + 1
        return _check(lib().pdo_spectral_init(ctypes.byref(self._h), nx, ny, nz, dx, dy, dz, int(fixOddball),
                                              dealiasF, int(init_periodicInZ)))

    def _nspec(self, planes=None):
        return self.nxh * self.shape[1] * (self.shape[2] if planes is None else planes)

    def fft(self, f, fhat, prec="f32_refine", stream=None):
        if f.numel() < self.shape[0] * self.shape[1] * self.shape[2] or _reals(fhat) < 2 * self._nspec():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_spectral_fft(self._h, _prec(f, prec), _ptr(f), _ptr(fhat), _stream(stream)))

    def ifft(self, fhat, f, setOddball=False, prec="f32_refine", stream=None):
        if f.numel() < self.shape[0] * self.shape[1] * self.shape[2] or _reals(fhat) < 2 * self._nspec():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_spectral_ifft(self._h, _prec(f, prec), _ptr(fhat), _ptr(f), int(setOddball),
                                              _stream(stream)))

    def mTimes_ik(self, dir, fin, fout=None, prec="f32_refine", stream=None):
        """mTimes_ik1 (dir 1) / mTimes_ik2 (dir 2); in place when fout is None."""
        fout = fin if fout is None else fout
        if _reals(fin) < 2 * self._nspec() or _reals(fout) < 2 * self._nspec():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_spectral_mtimes_ik(self._h, _prec(fin, prec), dir, _ptr(fin), _ptr(fout),
                                                   _stream(stream)))

    def mTimes_ik1(self, fin, fout=None, **kw):
        return self.mTimes_ik(1, fin, fout, **kw)

    def mTimes_ik2(self, fin, fout=None, **kw):
        return self.mTimes_ik(2, fin, fout, **kw)

    def dealias(self, fhat, prec="f32_refine", stream=None):
        if _reals(fhat) < 2 * self._nspec():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_spectral_dealias(self._h, _prec(fhat, prec), _ptr(fhat), _stream(stream)))

    def dealias_edgeField(self, fhat, prec="f32_refine", stream=None):
        if _reals(fhat) < 2 * self._nspec(self.shape[2] + 1):
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_spectral_dealias_edge(self._h, _prec(fhat, prec), _ptr(fhat), _stream(stream)))

    def set_factors(self, g1=None, g2=None, g3=None):
        """Replace the dealias factors by any separable transfer function."""
        import numpy as np
        dp = ctypes.POINTER(ctypes.c_double)
        keep = [None if g is None else np.ascontiguousarray(g, np.float64) for g in (g1, g2, g3)]
        args = [None if g is None else g.ctypes.data_as(dp) for g in keep]
        return _check(lib().pdo_spectral_set_factors(self._h, *args))

    def wavenumbers(self):
        """(k1, k2, k3, g1, g2, g3) host arrays; the dealias mask is g1 g2 g3."""
        import numpy as np
        nx, ny, nz = self.shape
        arrs = [np.zeros(self.nxh), np.zeros(ny), np.zeros(nz), np.zeros(self.nxh), np.zeros(ny), np.zeros(nz)]
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().pdo_spectral_wavenumbers(self._h, *[a.ctypes.data_as(dp) for a in arrs]))
        return tuple(arrs)

    def destroy(self):
        if self._h:
            lib().pdo_spectral_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class padepoisson:
    """type(padepoisson) of src/incompressible/PadePoisson.F90 with cd06
    staggered z derivatives (Pade6stagg scheme cd06): periodic in z, or with
    walls (PeriodicInZ = False), then optionally with the Stokes pressure
    (computeStokesPressure = True and Lz; ignored when periodic, as in the
    reference)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, sp, dz, PeriodicInZ=True, computeStokesPressure=False, Lz=None):
        self.sp = sp
        if PeriodicInZ:
            return _check(lib().pdo_padepoisson_init_periodic(ctypes.byref(self._h), sp._h, dz))
        _check(lib().pdo_padepoisson_init_nonperiodic(ctypes.byref(self._h), sp._h, dz))
        if computeStokesPressure:
            if Lz is None:
                raise ValueError("the Stokes pressure needs Lz")
            _check(lib().pdo_padepoisson_set_stokes(self._h, float(Lz)))
        return 0

    def _sizes(self, uh, vh, wh):
        n = self.sp._nspec()
        ne = self.sp._nspec(self.sp.shape[2] + 1)
        if _reals(uh) < 2 * n or _reals(vh) < 2 * n or _reals(wh) < 2 * ne:
            raise ValueError("buffer smaller than the declared shape")

    def PressureProjection(self, uhat, vhat, what, prec="f32_refine", stream=None):
        self._sizes(uhat, vhat, what)
        return _check(lib().pdo_padepoisson_projection(self._h, _prec(uhat, prec), _ptr(uhat), _ptr(vhat),
                                                       _ptr(what), _stream(stream)))

    def _pressure(self, uhat, vhat, what, pressure, update, prec, stream):
        self._sizes(uhat, vhat, what)
        nx, ny, nz = self.sp.shape
        if pressure.numel() < nx * ny * nz:
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_padepoisson_pressure(self._h, _prec(uhat, prec), _ptr(uhat), _ptr(vhat), _ptr(what),
                                                     _ptr(pressure), update, _stream(stream)))

    def getPressure(self, uhat, vhat, what, pressure, prec="f32_refine", stream=None):
        """the real pressure (nx, ny, nz); the velocities stay as they are"""
        return self._pressure(uhat, vhat, what, pressure, 0, prec, stream)

    def getPressureAndUpdateRHS(self, uhat, vhat, what, pressure, prec="f32_refine", stream=None):
        """PressureProjection of (uhat, vhat, what) and its real pressure"""
        return self._pressure(uhat, vhat, what, pressure, 1, prec, stream)

    def DivergenceCheck(self, uhat, vhat, what, divergence, prec="f32_refine", stream=None):
        self._sizes(uhat, vhat, what)
        nx, ny, nz = self.sp.shape
        if divergence.numel() < nx * ny * nz:
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_padepoisson_divergence(self._h, _prec(uhat, prec), _ptr(uhat), _ptr(vhat),
                                                       _ptr(what), _ptr(divergence), _stream(stream)))

    def destroy(self):
        if self._h:
            lib().pdo_padepoisson_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Pade6stagg:
    """type(Pade6stagg) of src/incompressible/PadeDerOps.F90 with scheme cd06:
    ddz_C2E(input, output, bot, top) etc. on (n1, n2, nz) cell / (n1, n2, nz+1)
    edge fields, bot / top in {-1 odd, 0 one-sided, 1 even} (ignored when
    periodic)."""

    KINDS = {"ddz_E2C": (10, 1, 0), "ddz_C2E": (11, 0, 1), "d2dz2_C2C": (12, 0, 0), "d2dz2_E2E": (13, 1, 1),
             "interpz_E2C": (14, 1, 0), "interpz_C2E": (15, 0, 1)}

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, nz, dz, isPeriodic):
        self.nz, self.dz = nz, dz
        return _check(lib().pdo_pade6stagg_init(ctypes.byref(self._h), nz, dz, int(bool(isPeriodic))))

    def _ap(self, name, inp, out, bot, top, n1, n2, is_complex, prec, stream):
        kind, ein, eout = self.KINDS[name]
        k = 2 if is_complex else 1
        if _reals(inp) < k * n1 * n2 * (self.nz + ein) or _reals(out) < k * n1 * n2 * (self.nz + eout):
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_pade6stagg_apply(self._h, kind, bot, top, _prec(inp, prec), _ptr(inp), _ptr(out), n1,
                                                 n2, int(is_complex), _stream(stream)))

    def ddz_E2C(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("ddz_E2C", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    def ddz_C2E(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("ddz_C2E", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    def d2dz2_C2C(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("d2dz2_C2C", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    def d2dz2_E2E(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("d2dz2_E2E", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    def interpz_E2C(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("interpz_E2C", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    def interpz_C2E(self, inp, out, bot, top, n1, n2, is_complex=False, prec="f32_refine", stream=None):
        return self._ap("interpz_C2E", inp, out, bot, top, n1, n2, is_complex, prec, stream)

    @staticmethod
    def getModifiedWavenumbers(k, dz, scheme="cd06"):
        import numpy as np
        k = np.ascontiguousarray(k, dtype=np.float64)
        kp = np.zeros_like(k)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().pdo_pade6stagg_modk({"fd02": 0, "cd06": 1, "fourierColl": 2}[scheme], k.size, dz,
                                         k.ctypes.data_as(dp), kp.ctypes.data_as(dp)))
        return kp

    def destroy(self):
        if self._h:
            lib().pdo_pade6stagg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class duidxj:
    """igrid%compute_duidxj (src/incompressible/igrid.F90:2553-2683) over a cell
    and an edge spectral object and a Pade6stagg object."""

    NAMES = ("uhat", "vhat", "whatC", "what", "uEhat", "vEhat", "duidxjC", "duidxjE", "duidxjChat", "duidxjEhat",
             "d2wdz2hatE", "d2udz2hatC", "d2vdz2hatC")

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, spC, spE, dzop, bc=(0,) * 12, inviscid=False):
        self.spC, self.spE, self.dzop = spC, spE, dzop
        arr = (ctypes.c_int * 12)(*[int(b) for b in bc])
        return _check(lib().pdo_duidxj_init(ctypes.byref(self._h), spC._h, spE._h, dzop._h, arr, int(inviscid)))

    def compute(self, prec="f32_refine", stream=None, **fields):
        """fields: the tensors named in NAMES (the d2 outputs may be omitted when inviscid)."""
        F = _lib.DuidxjFields()
        first = None
        for name in self.NAMES:
            t = fields.get(name)
            if t is not None:
                setattr(F, name, _ptr(t).value)
                first = first if first is not None else t
        return _check(lib().pdo_duidxj_compute(self._h, _prec(first, prec), ctypes.cast(ctypes.pointer(F),
                                                                                       ctypes.c_void_p),
                                               _stream(stream)))

    def destroy(self):
        if self._h:
            lib().pdo_duidxj_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def interp_PrimitiveVars(op, uhat, vhat, what, whatC, wC, uEhat, uE, vEhat, vE, That=None, TEhat=None, TE=None,
                         TBC_bottom=1, TBC_top=1, botBC_Temp=1, Tsurf=0.0, prec="f32_refine", stream=None):
    """igrid%interp_PrimitiveVars (igrid.F90:1423-1463) on a duidxj object
    (its spectral objects, Pade6stagg and boundary kinds)."""
    F = _lib.InterpFields()
    for name, t in (("uhat", uhat), ("vhat", vhat), ("what", what), ("That", That), ("whatC", whatC), ("wC", wC),
                    ("uEhat", uEhat), ("uE", uE), ("vEhat", vEhat), ("vE", vE), ("TEhat", TEhat), ("TE", TE)):
        if t is not None:
            setattr(F, name, _ptr(t).value)
    F.TBC_bottom, F.TBC_top, F.botBC_Temp, F.Tsurf = int(TBC_bottom), int(TBC_top), int(botBC_Temp), float(Tsurf)
    return _check(lib().pdo_igrid_interp_primitive(op._h, _prec(uhat, prec),
                                                    ctypes.cast(ctypes.pointer(F), ctypes.c_void_p),
                                                    _stream(stream)))


class sgsRHS:
    """sgs_igrid%getRHS_SGS (sgsmod_igrid.F90:204-266) for the eddy-viscosity
    models: model "smagorinsky", "sigma" or "amd"; cmodelC / cmodelE:
    per-plane model constants (sequences of nz / nz+1) or None for cglobal;
    wall_model "moeng" with roughness z0 (walls in z; compute then needs the
    physical u, v)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, spC, spE, dzop, model="smagorinsky", explicit_edge=False, camd=(1.0, 1.0, 1.0), cglobal=1.0,
             cmodelC=None, cmodelE=None, wall_model=None, z0=0.0, prec="f32_refine"):
        self._keep = (spC, spE, dzop)
        p = _lib.SgsParams()
        p.model = SGS_MODELS[model]
        p.wall_model = {None: 0, "moeng": 1}[wall_model]
        p.z0 = z0
        p.explicit_edge = int(explicit_edge)
        p.camd_x, p.camd_y, p.camd_z = camd
        p.cglobal = cglobal
        arrs = []
        for name, c in (("cmodelC", cmodelC), ("cmodelE", cmodelE)):
            if c is not None:
                a = (ctypes.c_double * len(c))(*[float(x) for x in c])
                arrs.append(a)
                setattr(p, name, ctypes.cast(a, ctypes.POINTER(ctypes.c_double)))
        return _check(lib().pdo_sgsrhs_init(ctypes.byref(self._h), spC._h, spE._h, dzop._h,
                                            ctypes.cast(ctypes.pointer(p), ctypes.c_void_p), PREC[prec]))

    def compute(self, duidxjC, duidxjE, urhs, vrhs, wrhs, lambda_1d=None, u=None, v=None, prec="f32_refine",
                stream=None):
        """urhs, vrhs, wrhs -= the divergence of the SGS stress (in place);
        u, v: the physical cell velocities, for the wall model."""
        if u is not None or v is not None:
            return _check(lib().pdo_sgsrhs_compute_wall(self._h, _prec(duidxjC, prec), _ptr(duidxjC),
                                                        _ptr(duidxjE), _ptr(lambda_1d), _ptr(u), _ptr(v),
                                                        _ptr(urhs), _ptr(vrhs), _ptr(wrhs), _stream(stream)))
        return _check(lib().pdo_sgsrhs_compute(self._h, _prec(duidxjC, prec), _ptr(duidxjC), _ptr(duidxjE),
                                               _ptr(lambda_1d), _ptr(urhs), _ptr(vrhs), _ptr(wrhs),
                                               _stream(stream)))

    def wall_state(self):
        """(Uspmn, ustar, WallMFactor, Umn) of the last compute"""
        out = (ctypes.c_double * 4)()
        _check(lib().pdo_sgsrhs_wall_state(self._h, out))
        return tuple(out)

    def fields(self):
        """device pointers (tauC, tauE, nuC, nuE) of the last compute"""
        out = [ctypes.c_void_p() for _ in range(4)]
        _check(lib().pdo_sgsrhs_fields(self._h, *[ctypes.byref(o) for o in out]))
        return [o.value for o in out]

    def destroy(self):
        if self._h:
            lib().pdo_sgsrhs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class cgrid:
    """getRHS of type(cgrid) (src/compressible/cgrid.F90:822-894) for one
    species of ideal gas, with the artificial diffusivities of getLAD and the
    cgrid filter.  shape (nx, ny, nz), spacing (dx, dy, dz); periodic,
    bc ((x_bc1, x_bcn), (y_bc1, y_bcn), (z_bc1, z_bcn)); derivative and
    filter per direction ("cd10" / "cd06"; "cf90" / "gaussian" / "lstsq");
    inviscid (the reference's default) or the power-law viscosity mu_ref,
    T_ref, visc_n, Pr."""

    FIELDS = {"rho": 0, "u": 1, "v": 2, "w": 3, "e": 4, "p": 5, "T": 6, "mu": 7, "bulk": 8, "kap": 9,
              "duidxj": 10, "tau": 11, "q": 12}
    _DER = {"cd10": 1, "cd06": 2}
    _FIL = {"cf90": 1, "gaussian": 2, "lstsq": 3}

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, shape, spacing, periodic=(True, True, True), bc=((0, 0), (0, 0), (0, 0)),
             derivative=("cd10",) * 3, filter=("cf90",) * 3, gam=1.4, Rgas=1.0, Cmu=0.002, Cbeta=1.75,
             Ckap=0.01, inviscid=True, mu_ref=1.0, T_ref=1.0, visc_n=0.75, Pr=0.75, prec="f64"):
        p = _lib.CgridParams()
        p.nx, p.ny, p.nz = shape
        p.dx, p.dy, p.dz = spacing
        for i in range(3):
            p.periodic[i] = int(bool(periodic[i]))
            p.derivative[i] = self._DER[derivative[i]]
            p.filter[i] = self._FIL[filter[i]]
        p.x_bc[0], p.x_bc[1] = bc[0]
        p.y_bc[0], p.y_bc[1] = bc[1]
        p.z_bc[0], p.z_bc[1] = bc[2]
        p.gam, p.Rgas, p.Cmu, p.Cbeta, p.Ckap = gam, Rgas, Cmu, Cbeta, Ckap
        p.inviscid = int(bool(inviscid))
        p.mu_ref, p.T_ref, p.visc_n, p.Pr = mu_ref, T_ref, visc_n, Pr
        self.shape, self.prec = tuple(shape), prec
        return _check(lib().pdo_cgrid_init(ctypes.byref(self._h), ctypes.cast(ctypes.pointer(p), ctypes.c_void_p),
                                           PREC[prec]))

    def _n(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def getRHS(self, W, rhs, stream=None):
        """W, rhs: flat tensors of (nx, ny, nz, 5) (rho, rho u, rho v, rho w, TE)."""
        if _reals(W) < 5 * self._n() or _reals(rhs) < 5 * self._n():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_cgrid_rhs(self._h, _prec(W, self.prec), _ptr(W), _ptr(rhs), _stream(stream)))

    def filter(self, arr, numtimes=1, gaussian=False, stream=None):
        if _reals(arr) < self._n():
            raise ValueError("buffer smaller than the declared shape")
        return _check(lib().pdo_cgrid_filter(self._h, _prec(arr, self.prec), _ptr(arr), int(numtimes),
                                             int(bool(gaussian)), _stream(stream)))

    def field(self, name):
        """device pointer of a field of the last getRHS"""
        out = ctypes.c_void_p()
        _check(lib().pdo_cgrid_field(self._h, self.FIELDS[name], ctypes.byref(out)))
        return out.value

    def destroy(self):
        if self._h:
            lib().pdo_cgrid_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class igrid:
    """Time advancement of type(igrid) (src/incompressible/igrid.F90) on one
    GPU for domains periodic in x, y and z, or with walls in z: TVD-RK3
    (scheme 1) or SSP-RK45 (scheme 2) stages of populate_rhs (rotational
    nonlinear term, viscous term, optional SGS stress with the wall model,
    HIT forcing or a mean pressure gradient) and project_and_prep."""

    FIELDS = {"u": 0, "v": 1, "wC": 2, "w": 3, "uhat": 4, "vhat": 5, "what": 6, "duidxjC": 7, "duidxjE": 8,
              "urhs": 9, "vrhs": 10, "wrhs": 11, "uE": 12, "vE": 13, "whatC": 14}

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, nx, ny, nz, Lx, Ly, Lz, Re=0.0, scheme=2, CFL=0.0, dt=0.0, sgs=None, forcing=None,
             walls=None, prec="f64"):
        """sgs: dict(model, explicit_edge, camd, cglobal, wall_model, z0) or
        None; forcing: dict(kmin, kmax, nwaves, eps, seed) or None; walls:
        dict(botWall, topWall, dpFdx, stokes_pressure) for walls in z (botWall
        1 no-slip, 2 slip, 3 wall model with sgs wall_model "moeng";
        stokes_pressure defaults to True, igrid's ComputeStokesPressure
        default: without it the projection zeroes the wall w that each stage's
        update leaves there, and the wall cells keep that divergence), or None
        for periodic z."""
        p = _lib.IgridParams()
        p.nx, p.ny, p.nz, p.Lx, p.Ly, p.Lz = nx, ny, nz, Lx, Ly, Lz
        p.Re, p.scheme, p.CFL, p.dt = Re, scheme, CFL, dt
        if sgs is not None:
            p.use_sgs = 1
            p.sgs.model = SGS_MODELS[sgs.get("model", "smagorinsky")]
            p.sgs.explicit_edge = int(sgs.get("explicit_edge", False))
            p.sgs.camd_x, p.sgs.camd_y, p.sgs.camd_z = sgs.get("camd", (1.0, 1.0, 1.0))
            p.sgs.cglobal = sgs.get("cglobal", 1.0)
            p.sgs.wall_model = {None: 0, "moeng": 1}[sgs.get("wall_model")]
            p.sgs.z0 = sgs.get("z0", 0.0)
        if forcing is not None:
            p.use_forcing = 1
            p.kmin, p.kmax = forcing["kmin"], forcing["kmax"]
            p.nwaves, p.eps_amplitude, p.seed = forcing["nwaves"], forcing["eps"], forcing["seed"]
        if walls is not None:
            p.walls = 1
            p.botWall, p.topWall = int(walls.get("botWall", 1)), int(walls.get("topWall", 1))
            p.dpFdx = float(walls.get("dpFdx", 0.0))
            p.stokes_pressure = int(bool(walls.get("stokes_pressure", True)))
        self.shape = (nx, ny, nz)
        self.prec = prec
        return _check(lib().pdo_igrid_init(ctypes.byref(self._h), ctypes.cast(ctypes.pointer(p), ctypes.c_void_p),
                                           PREC[prec]))

    def set_velocity(self, u, v, w, stream=None):
        """physical u, v on cells (nx, ny, nz) and w on edges (nx, ny, nz+1)"""
        return _check(lib().pdo_igrid_set_velocity(self._h, _ptr(u), _ptr(v), _ptr(w), _stream(stream)))

    def step(self, nsteps=1, stream=None):
        return _check(lib().pdo_igrid_step(self._h, nsteps, _stream(stream)))

    def field(self, name):
        """a torch view of one of the object's fields (FIELDS)"""
        import torch
        ptr = ctypes.c_void_p()
        _check(lib().pdo_igrid_field(self._h, self.FIELDS[name], ctypes.byref(ptr)))
        nx, ny, nz = self.shape
        nxh = nx // 2 + 1
        sizes = {"u": nx * ny * nz, "v": nx * ny * nz, "wC": nx * ny * nz, "w": nx * ny * (nz + 1),
                 "uE": nx * ny * (nz + 1), "vE": nx * ny * (nz + 1), "duidxjC": 9 * nx * ny * nz,
                 "duidxjE": 9 * nx * ny * (nz + 1), "uhat": nxh * ny * nz, "vhat": nxh * ny * nz,
                 "what": nxh * ny * (nz + 1), "urhs": nxh * ny * nz, "vrhs": nxh * ny * nz,
                 "wrhs": nxh * ny * (nz + 1), "whatC": nxh * ny * nz}
        spectral_ = name in ("uhat", "vhat", "what", "urhs", "vrhs", "wrhs", "whatC")
        real = torch.float64 if self.prec == "f64" else torch.float32
        cplx = torch.complex128 if self.prec == "f64" else torch.complex64
        dt = cplx if spectral_ else real
        n = sizes[name]
        W = torch.empty((), dtype=dt).element_size()
        # wrap the device memory without copying (the object owns it)
        holder = _DeviceView(ptr.value, n * W)
        t = torch.as_tensor(holder, device="cuda").view(dt)
        return t[:n]

    def dumpRestartFile(self, outputdir, runID, stream=None):
        """RESTART_RunRR_{u,v,w,info}.TTTTTT of igrid.F90:2771-2800"""
        return _check(lib().pdo_igrid_dump_restart(self._h, str(outputdir).encode(), int(runID), _stream(stream)))

    def readRestartFile(self, inputdir, tid, rid, stream=None):
        """igrid.F90:2719-2769"""
        return _check(lib().pdo_igrid_read_restart(self._h, str(inputdir).encode(), int(rid), int(tid),
                                                   _stream(stream)))

    def set_closure(self, kind, closure=None, delta=0.0, dratio=1.0, deltaRat=2.0):
        """kind 0 eddy viscosity (init constants), 1 dynamic Smagorinsky, 2
        learned + dynamic fallback, 3 learned (pdo_igrid_set_closure)"""
        kinds = {"eddy": 0, "dynamic": 1, "learned+dynamic": 2, "learned": 3}
        k = kinds[kind] if isinstance(kind, str) else int(kind)
        self._closure = closure  # keep the closure object alive with the grid
        return _check(lib().pdo_igrid_set_closure(self._h, k, closure._h if closure is not None else None, delta,
                                                  dratio, deltaRat))

    def time(self):
        t, s, d = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
        _check(lib().pdo_igrid_time(self._h, ctypes.byref(t), ctypes.byref(s), ctypes.byref(d)))
        return t.value, s.value, d.value

    def destroy(self):
        if self._h:
            lib().pdo_igrid_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Closure:
    """The learned SGS closure's pointwise algebra on the device (pdo_closure_*):
    invariant features of a velocity-gradient field, the tensor-basis stress
    of the network's outputs, and the training loss."""

    def __init__(self, net, max_points):
        self._h = ctypes.c_void_p()
        self.net = net
        _check(lib().pdo_closure_create(ctypes.byref(self._h), net._h if net is not None else None, max_points))

    def features(self, duidxj, npts, dratio, X, A9=None, prec=None, stream=None):
        """X (npts, 10) and A9 (npts, 9) of duidxj (9 components npts apart)"""
        return _check(lib().pdo_closure_features(self._h, _prec(duidxj, prec or "f32_refine"), _ptr(duidxj), npts,
                                                 dratio, _ptr(X), _ptr(A9), _stream(stream)))

    def stress(self, duidxj, npts, delta, dratio, tau, which=2, fallback=False, rescale=True, prec=None,
               stream=None):
        return _check(lib().pdo_closure_stress(self._h, _prec(duidxj, prec or "f32_refine"), _ptr(duidxj), npts,
                                               delta, dratio, which, int(fallback), int(rescale), _ptr(tau),
                                               _stream(stream)))

    def fallback_count(self, stream=None):
        c = ctypes.c_long()
        _check(lib().pdo_closure_fallback_count(self._h, ctypes.byref(c), _stream(stream)))
        return c.value

    def scale(self, stream=None):
        v = ctypes.c_double()
        _check(lib().pdo_closure_scale(self._h, ctypes.byref(v), _stream(stream)))
        return v.value

    def loss(self, G, A9, T, dG, sync=True, stream=None):
        """training loss (a float when sync, else None) and dL/dG into dG"""
        v = ctypes.c_double()
        _check(lib().pdo_closure_loss(self._h, _ptr(G), _ptr(A9), _ptr(T), G.shape[0], _ptr(dG),
                                      ctypes.byref(v) if sync else None, _stream(stream)))
        return v.value if sync else None

    def destroy(self):
        if self._h:
            lib().pdo_closure_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class _DeviceView:
    """__cuda_array_interface__ of a byte range of device memory owned by a
    library object (torch.as_tensor wraps it without a copy)."""

    def __init__(self, ptr, nbytes):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False), "version": 3}


SGS_MODELS = {"smagorinsky": 0, "sigma": 1, "amd": 2}


def sgs_kernel(duidxj, Sij, nu_sgs, npts, model="smagorinsky", camd=(0.0, 0.0, 0.0), prec="f32_refine",
               stream=None):
    """get_Sij_from_duidxj + get_smagorinsky_kernel / get_sigma_kernel /
    get_amd_kernel (sgs_models)."""
    m = SGS_MODELS[model]
    for t, k in ((duidxj, 9), (Sij, 6), (nu_sgs, 1)):
        if t is not None and t.numel() < k * npts:
            raise ValueError("buffer smaller than the declared shape")
    return _check(lib().pdo_sgs_kernel(_prec(duidxj, prec), m, _ptr(duidxj), _ptr(Sij), _ptr(nu_sgs), npts,
                                       camd[0], camd[1], camd[2], _stream(stream)))


def sgs_scale(nu_sgs, nxy, nz, cplane=None, cglobal=1.0, prec="f32_refine", stream=None):
    """multiply_by_model_constant: cplane is a float64 CUDA tensor of nz values or None."""
    return _check(lib().pdo_sgs_scale(_prec(nu_sgs, prec), _ptr(nu_sgs), nxy, nz, _ptr(cplane), cglobal,
                                      _stream(stream)))


class DynamicProcedure:
    """DoStandardDynamicProcedure of sgs_igrid (standardDynamicProcedure.F90)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def init(self, spC, deltaRat=2.0):
        self.sp = spC
        return _check(lib().pdo_dynproc_init(ctypes.byref(self._h), spC._h, deltaRat))

    def compute(self, u, v, w, uhat, vhat, what, Sij, Dsgs, lambda_1d, prec="f32_refine", stream=None):
        return _check(lib().pdo_dynproc_compute(self._h, _prec(u, prec), _ptr(u), _ptr(v), _ptr(w), _ptr(uhat),
                                                _ptr(vhat), _ptr(what), _ptr(Sij), _ptr(Dsgs), _ptr(lambda_1d),
                                                _stream(stream)))

    def destroy(self):
        if self._h:
            lib().pdo_dynproc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class SGSNet:
    """The learned SGS closure network (pdo_mlp_* in the C header): BF16 MFMA
    forward / backward kernels with FP32 accumulation and Adam."""

    # offsets of the padded FP32 parameter vector
    LAYOUT = (("W1", 0, (64, 16)), ("b1", 1024, (64,)), ("W2", 1088, (64, 64)), ("b2", 5184, (64,)),
              ("W3", 5248, (64, 64)), ("b3", 9344, (64,)), ("W4", 9408, (16, 64)), ("b4", 10432, (16,)))

    def __init__(self, max_batch):
        self._h = ctypes.c_void_p()
        self.nparam = lib().pdo_mlp_nparam()
        _check(lib().pdo_mlp_create(ctypes.byref(self._h), max_batch))

    @classmethod
    def pack(cls, params):
        """[(W (out, in), b (out,))] x 4 of the 10-64-64-64-6 network -> padded FP32 vector."""
        import numpy as np
        v = np.zeros(10448, np.float32)
        for (name, off, shape), arr in zip(cls.LAYOUT, [a for wb in params for a in wb]):
            full = np.zeros(shape, np.float32)
            a = np.asarray(arr, np.float32)
            full[tuple(slice(0, s) for s in a.shape)] = a
            v[off:off + full.size] = full.ravel()
        return v

    @classmethod
    def unpack(cls, v, sizes=(10, 64, 64, 64, 6)):
        out = []
        for i in range(4):
            (_, ow, sw), (_, ob, sb) = cls.LAYOUT[2 * i], cls.LAYOUT[2 * i + 1]
            W = v[ow:ow + sw[0] * sw[1]].reshape(sw)[: sizes[i + 1], : sizes[i]]
            out.append((W.copy(), v[ob:ob + sb[0]][: sizes[i + 1]].copy()))
        return out

    def set_params(self, vec):
        import numpy as np
        vec = np.ascontiguousarray(vec, np.float32)
        _check(lib().pdo_mlp_set_params(self._h, vec.ctypes.data_as(ctypes.c_void_p)))

    def get_params(self):
        import numpy as np
        v = np.zeros(self.nparam, np.float32)
        _check(lib().pdo_mlp_get_params(self._h, v.ctypes.data_as(ctypes.c_void_p)))
        return v

    def get_grads(self):
        import numpy as np
        v = np.zeros(self.nparam, np.float32)
        _check(lib().pdo_mlp_get_grads(self._h, v.ctypes.data_as(ctypes.c_void_p)))
        return v

    def forward(self, X, Y, stream=None):
        return _check(lib().pdo_mlp_forward(self._h, _ptr(X), _ptr(Y), X.shape[0], _stream(stream)))

    def gradients(self, X, T, Y=None, stream=None):
        return _check(lib().pdo_mlp_gradients(self._h, _ptr(X), _ptr(T), X.shape[0], _ptr(Y), _stream(stream)))

    def train_step(self, X, T, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, Y=None, stream=None):
        return _check(lib().pdo_mlp_train_step(self._h, _ptr(X), _ptr(T), X.shape[0], lr, b1, b2, eps, _ptr(Y),
                                               _stream(stream)))

    def forward_train(self, X, Y, stream=None):
        """forward pass keeping the activations for backward()"""
        return _check(lib().pdo_mlp_forward_train(self._h, _ptr(X), _ptr(Y), X.shape[0], _stream(stream)))

    def backward(self, dY, stream=None):
        """parameter gradients of a loss with dL/dY = dY (B, 6) for the batch of
        the last forward_train"""
        return _check(lib().pdo_mlp_backward(self._h, _ptr(dY), dY.shape[0], _stream(stream)))

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        return _check(lib().pdo_mlp_adam_step(self._h, lr, b1, b2, eps, _stream(stream)))

    def destroy(self):
        if self._h:
            lib().pdo_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
