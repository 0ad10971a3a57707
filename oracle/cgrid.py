"""TEST INFRASTRUCTURE ONLY (the checker): numpy restatement of the cgrid
compressible right-hand side for one species of ideal gas
(src/compressible/cgrid.F90): get_primitive :763-797, getRHS :822-894,
getRHS_x/y/z :896-1056, getLAD :1058-1257 (artificial shear and bulk
viscosity and conductivity), filter :1259-1336, get_tau :1360-1404, get_q
:1406-1452; the ideal gas (IdealGasEOS.F90:45-79) and the power-law
viscosity (PowerLawViscosity.F90:46-76).  Derivatives are the C oracle's
cd10 / cd06 operators (oracle.apply), the Gaussian filter oracle/explicit.py,
the cf90 filter the C oracle's.

Fields are (nx, ny, nz) float64 arrays; the conserved variables W
(nx, ny, nz, 5) in the order rho, rho u, rho v, rho w, TE (mass index 1,
mom_index 2, TE_index 5 of a single species).  Boundary codes (bc1, bcn)
per direction are the reference's x_bc, y_bc, z_bc."""
import numpy as np

from . import CD06_D1, CD06_D2, CD10_D1, CD10_D2, CF90, apply
from .explicit import filter_axis


class Cgrid:
    def __init__(self, shape, d, periodic=(True, True, True), bc=((0, 0), (0, 0), (0, 0)), method="cd10",
                 gam=1.4, Rgas=1.0, Cmu=0.002, Cbeta=1.75, Ckap=0.01, inviscid=True, mu_ref=1.0, T_ref=1.0,
                 visc_n=0.75, Pr=0.75):
        self.shape, self.d, self.periodic, self.bc = tuple(shape), tuple(d), tuple(periodic), bc
        self.D1, self.D2 = (CD10_D1, CD10_D2) if method == "cd10" else (CD06_D1, CD06_D2)
        self.gam, self.Rgas = gam, Rgas
        self.Cmu, self.Cbeta, self.Ckap = Cmu, Cbeta, Ckap
        self.inviscid = inviscid
        self.mu_ref, self.T_ref, self.visc_n, self.Pr = mu_ref, T_ref, visc_n, Pr

    # derivatives%ddx..d2dz2 along direction i (0, 1, 2) with (bc1, bcn)
    def dd(self, f, i, bc):
        return apply(self.D1, f, i + 1, self.d[i], self.periodic[i], bc[0], bc[1])

    def d2(self, f, i, bc):
        return apply(self.D2, f, i + 1, self.d[i], self.periodic[i], bc[0], bc[1])

    def gradient(self, f, bcs):
        """gradient (:446-480): (d/dx, d/dy, d/dz) with the given bc per direction."""
        return [self.dd(f, i, bcs[i]) for i in range(3)]

    def filter(self, arr, kind="gaussian", numtimes=1):
        """filter (:1259-1336): numtimes passes in y, then in x, then in z."""
        bc = self.bc
        out = np.asarray(arr, dtype=np.float64)
        for axis in (1, 0, 2):
            for _ in range(numtimes):
                if kind == "cf90":
                    out = apply(CF90, out, axis + 1, 1.0, self.periodic[axis], bc[axis][0], bc[axis][1])
                else:
                    out = filter_axis(out, axis, kind, self.periodic[axis], bc[axis][0], bc[axis][1])
        return out

    def primitive(self, W):
        """get_primitive with the ideal gas: rho, u, v, w, e, p, T."""
        rho = W[..., 0]
        onebyrho = 1.0 / rho
        u, v, w = W[..., 1] * onebyrho, W[..., 2] * onebyrho, W[..., 3] * onebyrho
        e = (W[..., 4] * onebyrho) - 0.5 * (u * u + v * v + w * w)
        p = (self.gam - 1.0) * rho * e
        T = (self.gam - 1.0) * e * (1.0 / self.Rgas)
        return dict(rho=rho, u=u, v=v, w=w, e=e, p=p, T=T)

    def transport(self, T):
        """get_transport_properties for one species (MixtureEOS.F90:175-203)."""
        if self.inviscid:
            z = np.zeros_like(T)
            return z, z.copy(), z.copy()
        Cp = self.gam * self.Rgas * (1.0 / (self.gam - 1.0))
        mu = self.mu_ref * (T / self.T_ref) ** self.visc_n
        return mu, np.zeros_like(T), Cp * mu / self.Pr

    def lad(self, P, g):
        """getLAD (:1058-1257) for one species: (mustar, bulkstar, kapstar)
        after their Gaussian filters (two passes per direction)."""
        bc = self.bc
        dx, dy, dz = self.d
        rho = P["rho"]
        dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz = g

        def d4(f, i):
            return self.d2(self.d2(f, i, bc[i]), i, bc[i])
        func = np.sqrt(dudx ** 2 + 0.5 * (dvdx + dudy) ** 2 + 0.5 * (dwdx + dudz) ** 2 + dvdy ** 2
                       + 0.5 * (dwdy + dvdz) ** 2 + dwdz ** 2)
        mustar = d4(func, 0) * dx ** 6
        mustar = mustar + d4(func, 2) * dz ** 6
        mustar = mustar + d4(func, 1) * dy ** 6
        mustar = self.filter(self.Cmu * rho * np.abs(mustar), numtimes=2)

        func = dudx + dvdy + dwdz
        r1, r2, r3 = [x * x for x in self.gradient(rho, bc)]
        s = r1 + r2 + r3 + 1.0e-32
        bulk = d4(func, 0) * dx ** 4 * (dx * r1 / s) ** 2
        bulk = bulk + d4(func, 2) * dz ** 4 * (dz * r3 / s) ** 2
        bulk = bulk + d4(func, 1) * dy ** 4 * (dy * r2 / s) ** 2
        c1, c2, c3 = dwdy - dvdz, dudz - dwdx, dvdx - dudy
        curl2 = c1 * c1 + c2 * c2 + c3 * c3
        dil2 = func * func
        fsw = np.where(func >= 0.0, 0.0, dil2 / (dil2 + curl2 + 1.0e-32))
        bulkstar = self.filter(self.Cbeta * rho * fsw * np.abs(bulk), numtimes=2)

        e = P["e"]
        e1, e2, e3 = [x * x for x in self.gradient(e, bc)]
        s = e1 + e2 + e3 + 1.0e-32
        kap = d4(e, 0) * dx ** 4 * (dx * e1 / s)
        kap = kap + d4(e, 2) * dz ** 4 * (dz * e3 / s)
        kap = kap + d4(e, 1) * dy ** 4 * (dy * e2 / s)
        sos = np.sqrt(self.gam * P["p"] / rho)
        kapstar = self.filter(self.Ckap * rho * sos * np.abs(kap) / P["T"], numtimes=2)
        return mustar, bulkstar, kapstar

    def rhs(self, W, parts=False):
        """getRHS (:822-894) without the problem source hook: the RHS
        (nx, ny, nz, 5); with parts also the primitives, gradients, transport
        properties after LAD, tau and q."""
        W = np.asarray(W, dtype=np.float64)
        P = self.primitive(W)
        bc = self.bc
        neg = lambda b: (-b[0], -b[1])  # noqa: E731
        gu = self.gradient(P["u"], [neg(bc[0]), bc[1], bc[2]])
        gv = self.gradient(P["v"], [bc[0], neg(bc[1]), bc[2]])
        gw = self.gradient(P["w"], [bc[0], bc[1], neg(bc[2])])
        g = gu + gv + gw
        mu, bulk, kap = self.transport(P["T"])
        ms, bs, ks = self.lad(P, g)
        mu, bulk, kap = mu + ms, bulk + bs, kap + ks
        dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz = g
        # get_tau (:1360-1404)
        bambda = (4.0 / 3.0) * mu + bulk
        lam = bulk - (2.0 / 3.0) * mu
        txy = mu * (dudy + dvdx)
        txz = mu * (dudz + dwdx)
        tyz = mu * (dvdz + dwdy)
        txx = bambda * dudx + lam * (dvdy + dwdz)
        tyy = bambda * dvdy + lam * (dudx + dwdz)
        tzz = bambda * dwdz + lam * (dudx + dvdy)
        # get_q (:1406-1452)
        T = P["T"]
        q = [-kap * dT for dT in self.gradient(T, bc)]
        rhou, rhov, rhow, TE = W[..., 1], W[..., 2], W[..., 3], W[..., 4]
        u, v, w, p = P["u"], P["v"], P["w"], P["p"]
        fluxes = [
            # getRHS_x (:896-952)
            [rhou, rhou * u + p - txx, rhou * v - txy, rhou * w - txz,
             (TE + p - txx) * u - v * txy - w * txz + q[0]],
            # getRHS_y (:954-998)
            [rhov, rhov * u - txy, rhov * v + p - tyy, rhov * w - tyz,
             (TE + p - tyy) * v - u * txy - w * tyz + q[1]],
            # getRHS_z (:1000-1056)
            [rhow, rhow * u - txz, rhow * v - tyz, rhow * w + p - tzz,
             (TE + p - tzz) * w - u * txz - v * tyz + q[2]],
        ]
        rhs = np.zeros(W.shape)
        for i in range(3):
            for c in range(5):
                # symmetric for the momentum along the direction, antisymmetric otherwise
                b = bc[i] if c == i + 1 else neg(bc[i])
                rhs[..., c] = rhs[..., c] - self.dd(fluxes[i][c], i, b)
        if not parts:
            return rhs
        return rhs, dict(P, grad=g, mu=mu, bulk=bulk, kap=kap, tau=(txx, txy, txz, tyy, tyz, tzz), q=q)
