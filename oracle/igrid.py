"""TEST INFRASTRUCTURE ONLY (the checker): numpy restatement of the time
advancement of igrid (src/incompressible/igrid.F90) for domains periodic in
z or with walls in z, composed from the oracle's spectral, Pade6stagg,
PadePoisson, duidxj and SGS restatements.  Only tests/ use it; the product
path never does.

  walls              PeriodicInZ = .false.: the boundary kinds of
                     get_boundary_conditions_stencil :5148-5220 for botWall /
                     topWall (1 no-slip, 2 slip, 3 wall model), the x-y
                     dealias, the PadePoisson of the 2 nz extensions, the SGS
                     wall model, addExtraForcingTerm :1759-1765 (dpFdx)

  project_and_prep   igrid.F90:1961-1992 (dealiasFields :1020-1037)
  populate_rhs       :1793-1912 with AddNonLinearTerm_Rot :1497-1566,
                     addViscousTerm :1914-1959, getRHS_SGS, HIT forcing
  TVD_RK3 / SSP_RK45 :1105-1170 / :1172-1299, compute_deltaT :1374-1400
  HIT_shell_forcing  forcingIsotropic.F90:123-314

The forcing's random wave vectors come from the generator the library uses
(splitmix64 of seed, draw counter, wave and variable): the reference draws
them with the compiler's random_number, which has no portable sequence.
"""
import numpy as np

from . import STG_IN_E2C
from .sgs import rhs_sgs
from .spectral import PadePoissonPeriodic, PadePoissonWalls, Spectral, duidxj, interp_primitive, pade6_apply

_M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def uniform01(seed, draw, ik, q):
    key = splitmix64((splitmix64((seed + draw) & _M64) + 3 * ik + q) & _M64)
    return (key >> 11) * 2.0 ** -53


# SSP-RK(5,4) coefficients, igrid.F90:1176-1183
B01, B12, B23, B34, B35, B45 = (0.39175222657189, 0.368410593050371, 0.25189177427169, 0.54497475022852,
                                0.06369246866629, 0.22600748323690)
A20, A21, A30, A32, A40, A43 = (0.444370493651235, 0.555629506348765, 0.620101851488403, 0.379898148511597,
                                0.17807995439313, 0.821920045606868)
A52, A53, A54 = 0.517231671970585, 0.096059710526147, 0.386708617503269


class Igrid:
    def __init__(self, nx, ny, nz, Lx, Ly, Lz, Re=0.0, scheme=2, CFL=0.0, dt=0.0, sgs=None, forcing=None,
                 walls=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.dy, self.dz = Lx / nx, Ly / ny, Lz / nz
        self.Re, self.scheme, self.CFL, self.dt_fixed = Re, scheme, CFL, dt
        self.sgs = sgs            # dict of rhs_sgs keyword arguments (z0 for the wall model), or None
        self.forcing = forcing    # dict(kmin, kmax, nwaves, eps, seed), or None
        # dict(botWall, topWall, dpFdx, stokes_pressure), or None; the Stokes
        # pressure is on unless stokes_pressure is False, as igrid's namelist
        # default ComputeStokesPressure = .true. (igrid.F90:330)
        self.walls = walls
        self.per = walls is None
        self.spC = Spectral(nx, ny, nz, self.dx, self.dy, self.dz, periodic_z=self.per)
        self.spE = Spectral(nx, ny, nz + 1, self.dx, self.dy, self.dz)
        if self.per:
            self.pp = PadePoissonPeriodic(self.spC, self.dx, self.dy, self.dz)
            # slip-wall boundary kinds (get_boundary_conditions_stencil,
            # igrid.F90:5148-5220, botWall = topWall = 2)
            self.bc = (1, 1, 1, 1, -1, -1, -1, -1, -1, -1, 0, 0)
        else:
            self.pp = PadePoissonWalls(self.spC, self.dx, self.dy, self.dz,
                                       Lz=Lz if walls.get("stokes_pressure", True) else None)
            bc = [1, 1, 1, 1, -1, -1, -1, -1, -1, -1, 0, 0]
            for side, wall in enumerate((walls["botWall"], walls["topWall"])):
                # (uBC = vBC, wBC, dUdzBC = dVdzBC, dWdzBC) of each wall kind
                uB, wB, dUB, dWB = {1: (-1, 1, 0, -1), 2: (1, -1, -1, 0), 3: (0, -1, 0, 0)}[wall]
                bc[0 + side], bc[2 + side], bc[4 + side] = uB, uB, wB
                bc[6 + side], bc[8 + side], bc[10 + side] = dUB, dUB, dWB
            self.bc = tuple(bc)
        self.tsim, self.step_count, self.dt, self.draws = 0.0, 0, 0.0, 0
        self.newTimeStep = True

    # project_and_prep
    def prep(self, uh, vh, wh):
        sp = self.spC
        uh, vh = sp.dealias(uh), sp.dealias(vh)
        wh = sp.dealias_edge(wh) if self.per else self.spE.dealias(wh)
        uh, vh, wh = self.pp.projection(uh, vh, wh)
        f = dict(uhat=uh, vhat=vh, what=wh, u=sp.ifft(uh), v=sp.ifft(vh), w=self.spE.ifft(wh))
        f.update(interp_primitive(sp, self.spE, self.dz, self.per, self.bc, uh, vh, wh))
        f.update(duidxj(sp, self.spE, self.dz, self.per, self.bc, self.Re <= 0, uh, vh, f["whatC"], wh, f["uEhat"],
                        f["vEhat"]))
        self.f = f
        return uh, vh, wh

    def set_velocity(self, u, v, w):
        self.state = self.prep(self.spC.fft(u), self.spC.fft(v), self.spE.fft(w))
        self.tsim, self.step_count, self.draws, self.newTimeStep = 0.0, 0, 0, True

    def forcing_rhs(self, ur, vr, wr):
        """getRHS_HITforcing (forcingIsotropic.F90:263-314) at the current stage."""
        F = self.forcing
        f = self.f
        nz = self.nz
        if self.newTimeStep:
            self.draws += 1
        Uz = np.fft.fft(f["uhat"], axis=2)
        Vz = np.fft.fft(f["vhat"], axis=2)
        k3 = 2 * np.pi * np.fft.fftfreq(nz, d=1.0 / nz) / (nz * self.dz)
        Wz = np.fft.fft(f["what"][:, :, :nz], axis=2) * np.exp(1j * k3 * self.dz / 2)[None, None, :]
        normfact = (float(self.nx) * self.ny * nz) ** 2
        fx = np.zeros_like(Uz)
        fy = np.zeros_like(Uz)
        fz = np.zeros_like(Uz)
        for ik in range(F["nwaves"]):
            kabs = F["kmin"] + (F["kmax"] - F["kmin"]) * uniform01(F["seed"], self.draws, ik, 0)
            zeta = -1.0 + 2.0 * uniform01(F["seed"], self.draws, ik, 1)
            theta = 2.0 * np.pi * uniform01(F["seed"], self.draws, ik, 2)
            s = np.sqrt(1 - zeta ** 2)
            wx = int(np.ceil(abs(kabs * s * np.cos(theta))))
            wy = int(np.ceil(abs(kabs * s * np.sin(theta))))
            wz = int(np.ceil(abs(kabs * zeta)))
            a, b, c = Uz[wx, wy, wz], Vz[wx, wy, wz], Wz[wx, wy, wz]
            den = abs(a) ** 2 + abs(b) ** 2 + abs(c) ** 2 + 1e-14
            fac = normfact * F["eps"] / den / F["nwaves"]
            fx[wx, wy, wz] += fac * np.conj(a)
            fy[wx, wy, wz] += fac * np.conj(b)
            fz[wx, wy, wz] += fac * np.conj(c)
        ur = ur + np.fft.ifft(fx, axis=2)
        vr = vr + np.fft.ifft(fy, axis=2)
        fzE = np.fft.ifft(fz * np.exp(-1j * k3 * self.dz / 2)[None, None, :], axis=2)
        wr = wr + np.concatenate([fzE, fzE[:, :, :1]], axis=2)
        return ur, vr, wr

    def rhs(self):
        """populate_rhs of the prepared stage."""
        f, spC, spE = self.f, self.spC, self.spE
        gC, gE = f["duidxjC"], f["duidxjE"]
        E2C = lambda x: pade6_apply(STG_IN_E2C, x, self.dz, self.nz, 0, 0, self.per)  # noqa: E731
        # AddNonLinearTerm_Rot
        ur = E2C(spE.fft((gE[..., 6] - gE[..., 2]) * f["w"])) + spC.fft((gC[..., 3] - gC[..., 1]) * f["v"])
        vr = E2C(spE.fft((gE[..., 7] - gE[..., 5]) * f["w"])) + spC.fft((gC[..., 1] - gC[..., 3]) * f["u"])
        wr = spE.fft((gE[..., 2] - gE[..., 6]) * f["uE"] + (gE[..., 5] - gE[..., 7]) * f["vE"])
        # addViscousTerm
        if self.Re > 0:
            ks = spC.k1[:, None, None] ** 2 + spC.k2[None, :, None] ** 2
            ur = ur + (1 / self.Re) * (-ks * f["uhat"] + f["d2udz2hatC"])
            vr = vr + (1 / self.Re) * (-ks * f["vhat"] + f["d2vdz2hatC"])
            wr = wr + (1 / self.Re) * (-ks * f["what"] + f["d2wdz2hatE"])
        if self.sgs is not None:
            kw = dict(self.sgs)
            z0 = kw.pop("z0", None)
            if z0 is not None:
                kw["wall"] = (f["u"], f["v"], f["uhat"], f["vhat"], z0)
            ur, vr, wr = rhs_sgs(spC, spE, self.dz, self.per, gC, gE, ur, vr, wr, **kw)[:3]
        if self.walls is not None and self.walls.get("dpFdx", 0.0) != 0.0:
            ur = ur.copy()
            ur[0, 0, :] += self.walls["dpFdx"] * self.nx * self.ny
        if self.forcing is not None:
            ur, vr, wr = self.forcing_rhs(ur, vr, wr)
        return ur, vr, wr

    def compute_dt(self):
        if not self.CFL > 0:
            return self.dt_fixed
        f = self.f
        ts = (np.abs(f["u"] / self.dx) + np.abs(f["v"] / self.dy) + np.abs(f["wC"] / self.dz)).max()
        dt = self.CFL / ts
        if self.Re > 0:
            dt = min(dt, self.CFL * self.Re * min(self.dx, self.dy, self.dz) ** 2)
        return dt

    def step(self):
        dt = self.dt = self.compute_dt()
        S0 = self.state
        if self.scheme == 1:
            R = self.rhs()
            self.newTimeStep = False
            S1 = self.prep(*[x + dt * r for x, r in zip(S0, R)])
            R = self.rhs()
            S1 = self.prep(*[0.75 * x + 0.25 * y + (0.25 * dt) * r for x, y, r in zip(S0, S1, R)])
            R = self.rhs()
            S0 = self.prep(*[(1 / 3) * x + (2 / 3) * y + ((2 / 3) * dt) * r for x, y, r in zip(S0, S1, R)])
        else:
            R = self.rhs()
            self.newTimeStep = False
            S1 = self.prep(*[x + (B01 * dt) * r for x, r in zip(S0, R)])
            R = self.rhs()
            S2 = self.prep(*[A20 * x + A21 * y + (B12 * dt) * r for x, y, r in zip(S0, S1, R)])
            R = self.rhs()
            S3 = self.prep(*[A30 * x + A32 * y + (B23 * dt) * r for x, y, r in zip(S0, S2, R)])
            R = self.rhs()
            S0 = self.prep(*[A40 * x + A43 * y + (B34 * dt) * r for x, y, r in zip(S0, S3, R)])
            R5 = self.rhs()
            S0 = self.prep(*[A52 * a + A53 * b + (B35 * dt) * r + A54 * x + (B45 * dt) * q
                             for a, b, r, x, q in zip(S2, S3, R, S0, R5)])
        self.state = S0
        self.step_count += 1
        self.tsim += dt
        self.newTimeStep = True
