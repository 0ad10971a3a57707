"""TEST INFRASTRUCTURE ONLY (the checker): CPU restatement of the igrid spectral
operators and of the periodic PadePoisson projection, in numpy.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this module; the
product path never does.

Arrays are numpy arrays indexed like the reference's Fortran arrays: real
fields (nx, ny, nz), half spectra (nx/2+1, ny, nz) complex, edge fields with
nz+1 planes.  The discrete Fourier transforms (FFTW in the reference,
fft_3d.F90) are numpy.fft, whose unnormalized forward / 1/n inverse pair is
the DFT FFTW computes; tests/test_oracle_pins.py checks it against a
brute-force DFT.
"""
import numpy as np

from . import (BOT_EVEN, BOT_SIDED, STG_D1_C2E, STG_D1_E2C, STG_D2_C2C, STG_D2_E2E, STG_IN_C2E, STG_IN_E2C,
               TOP_EVEN, TOP_SIDED, apply, modk, stagg_np)


def getwavenums(n, dx):
    """GetWaveNums (spectral.F90:1748-1763): (-pi + (i-1) 2 pi/(n - n%2))/dx, then
    ifftshift: (0, 1, .., n/2-1, -n/2, .., -1) 2 pi / (n dx) for even n."""
    dummy = n - n % 2
    k = (-np.pi + np.arange(n) * 2.0 * np.pi / dummy) / dx
    return np.fft.ifftshift(k)


class Spectral:
    """type(spectral) with dimTransform = 2 and an x input pencil
    (spectral.F90:867-1260; the periodic-z dealiasing of :755-865)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, fix_oddball=False, dealias_fact=2.0 / 3.0, periodic_z=False):
        self.shape = (nx, ny, nz)
        self.nxh = nx // 2 + 1
        self.periodic_z = periodic_z
        k1 = getwavenums(nx, dx)
        k2 = getwavenums(ny, dy)
        k3 = getwavenums(nz, dz)
        # STEP 3 (:1027-1033): the oddball takes a + sign
        k1[nx // 2] = -k1[nx // 2]
        k2[ny // 2] = -k2[ny // 2]
        if nz > 1:
            k3[nz // 2] = -k3[nz // 2]
        self.k1 = k1[: self.nxh].copy()
        self.k2 = k2.copy()
        self.k3 = k3.copy()
        # STEP 7 (:1135-1148): 2/3 rule on (k1, k2), before the oddball fix
        kdx = ((2.0 / 3.0) * np.pi) / dx
        kdy = ((2.0 / 3.0) * np.pi) / dy
        self.g1 = (np.abs(self.k1) < kdx).astype(float)
        self.g2 = (np.abs(self.k2) < kdy).astype(float)
        self.g3 = np.ones(nz)
        # STEP 11 (:1201-1216, two periodic directions): zero x and y oddballs
        if fix_oddball:
            self.k2[ny // 2] = 0.0
            self.k1[nx // 2] = 0.0
        if periodic_z:
            # init_periodic_inZ_procedures (:783-813): mask of stored k1, k2, k3
            self.g1 = np.where(np.abs(self.k1) >= (dealias_fact * np.pi) / dx, 0.0, 1.0)
            self.g2 = np.where(np.abs(self.k2) >= (dealias_fact * np.pi) / dy, 0.0, 1.0)
            self.g3 = np.where(np.abs(self.k3) >= (dealias_fact * np.pi) / dz, 0.0, 1.0)

    def fft(self, f):
        """fft2_x2y (fft_3d.F90:645-663): r2c along x, c2c along y, unnormalized."""
        return np.fft.fft(np.fft.rfft(np.asarray(f, dtype=np.float64), axis=0), axis=1)

    def ifft(self, fhat, set_oddball=False):
        """ifft2_y2x (fft_3d.F90:616-643): inverse c2c along y, optionally zero
        the kx = nx/2 modes, c2r along x (real parts of kx = 0 and nx/2), times
        1/(nx ny)."""
        t = np.fft.ifft(np.asarray(fhat, dtype=np.complex128), axis=1)
        if set_oddball:
            t[self.shape[0] // 2] = 0.0
        return np.fft.irfft(t, n=self.shape[0], axis=0)

    def mtimes_ik(self, f, dir):
        """mTimes_ik1/ik2 (spectral.F90:235-312): i k f."""
        k = self.k1[:, None, None] if dir == 1 else self.k2[None, :, None]
        f = np.asarray(f, dtype=np.complex128)
        return (-k * f.imag) + 1j * (k * f.real)

    def dealias(self, fhat):
        """dealias (spectral.F90:314-341)."""
        g = self.g1[:, None, None] * self.g2[None, :, None]
        if not self.periodic_z:
            return fhat * g
        t = np.fft.fft(fhat, axis=2) * (g * self.g3[None, None, :])
        return np.fft.ifft(t, axis=2)

    def dealias_edge(self, fhat):
        """dealias_edgeField (spectral.F90:343-363): nz+1 planes, periodic z."""
        out = np.array(fhat, dtype=np.complex128)
        if not self.periodic_z:
            return out
        nz = self.shape[2]
        out[:, :, :nz] = self.dealias(out[:, :, :nz])
        out[:, :, nz] = out[:, :, 0]
        return out


def _stagg(op, f, dz, nz):
    """Periodic cd06stagg operator along z of a complex field (real operator
    applied to the real and imaginary parts)."""
    return apply(op, f.real, 3, dz, True, n=nz) + 1j * apply(op, f.imag, 3, dz, True, n=nz)


class PadePoissonPeriodic:
    """Periodic-in-z padepoisson with cd06 staggered z derivatives
    (PadePoisson.F90:76-128, 130-180, 386-432, 1165-1244)."""

    def __init__(self, sp, dx, dy, dz):
        nx, ny, nz = sp.shape
        self.sp, self.dz, self.nz = sp, dz, nz
        k1 = getwavenums(nx, dx)[: sp.nxh]
        k2 = getwavenums(ny, dy)
        # cd06stagg modified wavenumber of k3 (getmodCD06stagg via derivZ)
        k3m = modk(2, nz, dz)
        ks = k1[:, None, None] ** 2 + k2[None, :, None] ** 2 + k3m[None, None, :] ** 2
        self.kradsq_inv = np.where(ks <= 1e-14, 0.0, 1.0 / np.where(ks <= 1e-14, 1.0, ks))

    def _div(self, uhat, vhat, what):
        sp = self.sp
        k1 = sp.k1[:, None, None]
        k2 = sp.k2[None, :, None]
        dwdz = _stagg(STG_D1_E2C, np.asarray(what), self.dz, self.nz)
        return dwdz + 1j * (k1 * uhat + k2 * vhat)

    def projection(self, uhat, vhat, what):
        """PeriodicProjection: returns the projected (uhat, vhat, what)."""
        sp = self.sp
        f = self._div(uhat, vhat, what)
        p = np.fft.ifft(-self.kradsq_inv * np.fft.fft(f, axis=2), axis=2)
        dpdz = _stagg(STG_D1_C2E, p, self.dz, self.nz)
        k1 = sp.k1[:, None, None]
        k2 = sp.k2[None, :, None]
        return uhat - 1j * k1 * p, vhat - 1j * k2 * p, what - dpdz

    def divergence(self, uhat, vhat, what):
        """DivergenceCheck steps 1-3: the real divergence field."""
        return self.sp.ifft(self._div(uhat, vhat, what))


# Pade6stagg (PadeDerOps.F90:57-1008), scheme cd06: the boundary object of
# each (bot, top) in {-1 odd, 0 one-sided, 1 even}; the object derXY is
# initialised with X the top and Y the bottom kind, a one-sided side being
# marked even unless the other side is odd.
def pade6_flags(bot, top):
    te = 1 if top == 1 else (0 if top == -1 else (0 if bot == -1 else 1))
    be = 1 if bot == 1 else (0 if bot == -1 else (0 if top == -1 else 1))
    return te * TOP_EVEN + be * BOT_EVEN + (TOP_SIDED if top == 0 else 0) + (BOT_SIDED if bot == 0 else 0)


def pade6_apply(op, f, dz, nz, bot, top, periodic):
    """A Pade6stagg generic on a complex field f(n1, n2, n_in)."""
    f = np.asarray(f, dtype=np.complex128)
    if periodic:
        return _stagg(op, f, dz, nz)
    if op in (STG_D2_C2C, STG_D2_E2E) and (bot == 0 or top == 0) or bot not in (-1, 0, 1) or top not in (-1, 0, 1):
        nout = nz + 1 if op in (STG_D1_C2E, STG_IN_C2E, STG_D2_E2E) else nz
        return np.zeros(f.shape[:2] + (nout,), complex)
    fl = pade6_flags(bot, top)
    return stagg_np(op, np.asfortranarray(f.real), dz, fl, nz) + 1j * stagg_np(op, np.asfortranarray(f.imag), dz,
                                                                              fl, nz)


def duidxj(spC, spE, dz, periodic, bc, inviscid, uhat, vhat, whatC, what, uEhat, vEhat):
    """igrid%compute_duidxj (igrid.F90:2553-2683).  Returns a dict of the
    outputs: duidxjC (nx,ny,nz,9), duidxjE (nx,ny,nz+1,9), their spectra and
    d2wdz2hatE, d2udz2hatC, d2vdz2hatC (viscous)."""
    nx, ny, nz = spC.shape
    uB, uT, vB, vT, wB, wT, dUB, dUT, dVB, dVT, dWB, dWT = bc
    RC = np.zeros((nx, ny, nz, 9))
    RE = np.zeros((nx, ny, nz + 1, 9))
    HC = np.zeros((spC.nxh, ny, nz, 9), complex)
    HE = np.zeros((spC.nxh, ny, nz + 1, 9), complex)
    out = {}
    for k, c, e in ((1, uhat, uEhat), (2, uhat, uEhat), (4, vhat, vEhat), (5, vhat, vEhat), (7, whatC, what),
                    (8, whatC, what)):
        d = 1 if k in (1, 4, 7) else 2
        HC[..., k - 1] = spC.mtimes_ik(c, d)
        RC[..., k - 1] = spC.ifft(HC[..., k - 1])
        HE[..., k - 1] = spE.mtimes_ik(e, d)
        RE[..., k - 1] = spE.ifft(HE[..., k - 1])
    P6 = lambda op, f, b, t: pade6_apply(op, f, dz, nz, b, t, periodic)  # noqa: E731
    HC[..., 8] = P6(STG_D1_E2C, what, wB, wT)
    RC[..., 8] = spC.ifft(HC[..., 8])
    HE[..., 8] = P6(STG_IN_C2E, HC[..., 8], dWB, dWT)
    RE[..., 8] = spE.ifft(HE[..., 8])
    if not inviscid:
        out["d2wdz2hatE"] = P6(STG_D2_E2E, what, wB, wT)
    for k, c, b, t, db, dt, name in ((3, uhat, uB, uT, dUB, dUT, "d2udz2hatC"),
                                     (6, vhat, vB, vT, dVB, dVT, "d2vdz2hatC")):
        HE[..., k - 1] = P6(STG_D1_C2E, c, b, t)
        RE[..., k - 1] = spE.ifft(HE[..., k - 1])
        if not inviscid:
            if t == 0 or b == 0:
                out[name] = P6(STG_D1_E2C, P6(STG_D1_C2E, c, b, t), db, dt)
            else:
                out[name] = P6(STG_D2_C2C, c, b, t)
        HC[..., k - 1] = P6(STG_IN_E2C, HE[..., k - 1], db, dt)
        RC[..., k - 1] = spC.ifft(HC[..., k - 1])
    out.update(duidxjC=RC, duidxjE=RE, duidxjChat=HC, duidxjEhat=HE)
    return out


def interp_primitive(spC, spE, dz, periodic, bc, uhat, vhat, what, That=None, Tbc=(1, 1), botBC_Temp=1,
                     Tsurf=0.0):
    """igrid%interp_PrimitiveVars (igrid.F90:1423-1463): w to cells, u, v (and
    T) to edges with Pade6stagg, then the 2D inverse transforms.  bc = (uBC_bottom,
    uBC_top, vBC_bottom, vBC_top, wBC_bottom, wBC_top, ...)."""
    nx, ny, nz = spC.shape
    P6 = lambda op, f, b, t: pade6_apply(op, f, dz, nz, b, t, periodic)  # noqa: E731
    out = {}
    out["whatC"] = P6(STG_IN_E2C, what, bc[4], bc[5])
    out["wC"] = spC.ifft(out["whatC"])
    out["uEhat"] = P6(STG_IN_C2E, uhat, bc[0], bc[1])
    out["uE"] = spE.ifft(out["uEhat"])
    out["vEhat"] = P6(STG_IN_C2E, vhat, bc[2], bc[3])
    out["vE"] = spE.ifft(out["vEhat"])
    if That is not None:
        TE = P6(STG_IN_C2E, That, Tbc[0], Tbc[1])
        if botBC_Temp == 0:
            # :1449-1454: the bottom plane holds only the surface mean
            TE[:, :, 0] = 0.0
            TE[0, 0, 0] = Tsurf * nx * ny
        out["TEhat"] = TE
        out["TE"] = spE.ifft(TE)
    return out
