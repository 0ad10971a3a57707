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

    def _pressure_hat(self, uhat, vhat, what):
        f = self._div(uhat, vhat, what)
        return np.fft.ifft(-self.kradsq_inv * np.fft.fft(f, axis=2), axis=2)

    def get_pressure(self, uhat, vhat, what):
        """Periodic_getPressure (:716-750): the real pressure field."""
        return self.sp.ifft(self._pressure_hat(uhat, vhat, what))

    def get_pressure_and_update(self, uhat, vhat, what):
        """Periodic_getPressureAndUpdateRHS (:900-949): the projected
        (u, v, w) and the pressure."""
        return self.projection(uhat, vhat, what) + (self.get_pressure(uhat, vhat, what),)


class PadePoissonWalls:
    """padepoisson with walls in z (PeriodicInZ = .false.), literally as the
    reference computes it: init PadePoisson.F90:181-307, PressureProjection
    :434-624 (steps 0-9), getPressure :752-898, getPressureAndUpdateRHS
    :952-1163: the 2 nz-point even extension of f = i (k1 u + k2 v) and odd
    extension of w, their z transforms, the pressure
    -(F + i k3' e^{-i k3 dz/2} W) / |k|^2 (|k|^2 <= 1e-14 dropped),
    W -= i k3' e^{+i k3 dz/2} P, the inverse transforms (1 / (2 nz)), and
    w = 0 on both walls; u -= i k1 p, v -= i k2 p.

    Lz given: computeStokesPressure.  Step 0 (ProjectStokesPressure /
    GetStokesPressure, :320-384, :641-714) first removes, line by line, the
    wall-normal velocity on both walls with the potential
    phi = chat_bot cosh(lam (Lz - z)) + chat_top cosh(lam z), lam = |(k1, k2)|,
    and from there on k1, k2 are the plain GetWaveNums values (k1inZ, k2inZ),
    not the spectral object's (whose Nyquist modes carry the other sign).  The
    pressure adds phat_z1 + phat_z2 = i phi (:882-890, :1142-1150)."""

    def __init__(self, sp, dx, dy, dz, Lz=None):
        nx, ny, nz = sp.shape
        self.sp, self.dz, self.nz = sp, dz, nz
        k1 = getwavenums(nx, dx)[: sp.nxh]
        k2 = getwavenums(ny, dy)
        k3 = getwavenums(2 * nz, dz)
        k3m = modk(2, 2 * nz, dz)
        self.k3cm = k3m * np.exp(-0.5j * dz * k3)
        self.k3cp = k3m * np.exp(0.5j * dz * k3)
        ks = k1[:, None, None] ** 2 + k2[None, :, None] ** 2 + k3m[None, None, :] ** 2
        self.kradsq_inv = np.where(ks <= 1e-14, 0.0, 1.0 / np.where(ks <= 1e-14, 1.0, ks))
        self.k1in, self.k2in = k1, k2
        self.stokes = Lz is not None
        if self.stokes:
            self._stokes_tables(float(Lz))

    def _stokes_tables(self, Lz):
        """init :236-307 (denFact, cosh_bot/top on the cells, sinh_bot/top on
        the edges, with the reference's caps)."""
        nz = self.nz
        lam = np.sqrt(self.k1in[:, None] ** 2 + self.k2in[None, :] ** 2)
        t = lam * Lz
        den = np.where(t < 500.0, 1.0 / (lam * np.sinh(np.minimum(t, 500.0)) + 1e-13), 0.0)
        den = np.where(den < 1e-16, 0.0, den)
        den[0, 0] = 0.0
        # linspace (gridtools.F90:51-69): running sums of the step
        step = (Lz - 0.0) / (float(nz + 1) - 1.0)
        ze = np.zeros(nz + 1)
        for i in range(1, nz + 1):
            ze[i] = ze[i - 1] + step
        zc = 0.5 * (ze[:nz] + ze[1:])
        L3 = lam[:, :, None]

        def capped(t, f, cap):
            return np.where(t < 32.0, f(np.minimum(t, 32.0)), cap)

        self.den = den
        self.cosh_bot = capped(L3 * (Lz - zc), np.cosh, 4.0e13)
        self.cosh_top = capped(L3 * zc, np.cosh, 1.0e13)
        tb, tt = L3 * (Lz - ze), L3 * ze
        self.sinh_bot = np.where(tb < 32.0, -L3 * np.sinh(np.minimum(tb, 32.0)), -4.0e13)
        self.sinh_top = np.where(tt < 32.0, L3 * np.sinh(np.minimum(tt, 32.0)), 4.0e13)
        self.ze, self.zc, self.Lz = ze, zc, Lz

    def stokes_step(self, uhat, vhat, what):
        """ProjectStokesPressure / GetStokesPressure: the corrected copies
        (u, v, w) and phat_z1, phat_z2."""
        nz = self.nz
        k1 = self.k1in[:, None, None]
        k2 = self.k2in[None, :, None]
        u = np.array(uhat, dtype=np.complex128)
        v = np.array(vhat, dtype=np.complex128)
        w = np.array(what, dtype=np.complex128)
        cb = -w[:, :, 0] * self.den
        p1 = 1j * cb[:, :, None] * self.cosh_bot
        u = u - k1 * p1
        v = v - k2 * p1
        w[:, :, 0] = 0
        w[:, :, 1:] = w[:, :, 1:] - cb[:, :, None] * self.sinh_bot[:, :, 1:]
        ct = w[:, :, nz] * self.den
        p2 = 1j * ct[:, :, None] * self.cosh_top
        u = u - k1 * p2
        v = v - k2 * p2
        w[:, :, :nz] = w[:, :, :nz] - ct[:, :, None] * self.sinh_top[:, :, :nz]
        w[:, :, nz] = 0
        return u, v, w, p1, p2

    def _prep(self, uhat, vhat, what):
        sp = self.sp
        if self.stokes:
            u, v, w, p1, p2 = self.stokes_step(uhat, vhat, what)
            k1, k2, ps = self.k1in, self.k2in, p1 + p2
        else:
            u, v, w = np.asarray(uhat), np.asarray(vhat), np.asarray(what)
            k1, k2, ps = sp.k1, sp.k2, None
        return u, v, w, k1[:, None, None], k2[None, :, None], ps

    def _solve(self, f, w, update):
        """steps 3-7: the pressure on the cells and (update) the new w."""
        nz = self.nz
        fext = np.concatenate([f[:, :, ::-1], f], axis=2)
        wext = np.concatenate([-w[:, :, nz - 1:0:-1], w], axis=2)
        F = np.fft.fft(fext, axis=2)
        Wt = np.fft.fft(wext, axis=2)
        F = -(F + 1j * self.k3cm * Wt) * self.kradsq_inv
        p = np.fft.ifft(F, axis=2)[:, :, nz:]
        if not update:
            return p, None
        Wt = Wt - 1j * self.k3cp * F
        wn = np.fft.ifft(Wt, axis=2)[:, :, nz - 1:].copy()
        wn[:, :, 0] = 0
        wn[:, :, nz] = 0
        return p, wn

    def _project(self, uhat, vhat, what):
        u, v, w, k1, k2, ps = self._prep(uhat, vhat, what)
        p, wn = self._solve(1j * (k1 * u + k2 * v), w, True)
        return u - 1j * k1 * p, v - 1j * k2 * p, wn, p, ps

    def projection(self, uhat, vhat, what):
        return self._project(uhat, vhat, what)[:3]

    def get_pressure(self, uhat, vhat, what):
        """getPressure: the real pressure field; the inputs stay as they are."""
        u, v, w, k1, k2, ps = self._prep(uhat, vhat, what)
        p, _ = self._solve(1j * (k1 * u + k2 * v), w, False)
        if ps is not None:
            p = p + ps
        return self.sp.ifft(p)

    def get_pressure_and_update(self, uhat, vhat, what):
        """getPressureAndUpdateRHS: the projected (u, v, w) and the pressure.
        With the Stokes pressure, the i phi of this call is added (the
        reference adds phat_z1 + phat_z2 as its last getPressure left them:
        ProjectStokesPressure does not store them)."""
        u, v, w, p, ps = self._project(uhat, vhat, what)
        if ps is not None:
            p = p + ps
        return u, v, w, self.sp.ifft(p)


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
