"""TEST INFRASTRUCTURE ONLY (the checker): CPU restatement of igrid's SGS
eddy-viscosity kernels, the Moeng wall model and the standard dynamic
procedure, in numpy
(src/incompressible/sgs_models/*.F90).  Used by tests/ and by the learned
closure's comparison only; the product path never imports it.

Velocity gradients are (nx, ny, nz, 9) in igrid's order dudx dudy dudz dvdx
dvdy dvdz dwdx dwdy dwdz; strain rates (.., 6) in the order S11 S12 S13 S22
S23 S33."""
import numpy as np


def sij(g):
    """get_Sij_from_duidxj (eddyViscosity.F90:1-21)."""
    S = np.empty(g.shape[:-1] + (6,))
    S[..., 0], S[..., 3], S[..., 5] = g[..., 0], g[..., 4], g[..., 8]
    S[..., 1] = 0.5 * (g[..., 1] + g[..., 3])
    S[..., 2] = 0.5 * (g[..., 2] + g[..., 6])
    S[..., 4] = 0.5 * (g[..., 5] + g[..., 7])
    return S


def smagorinsky(S):
    """get_smagorinsky_kernel (smagorinsky.F90:45-68): sqrt(2 S_ij S_ij)."""
    q = S[..., 0] ** 2 + 2 * S[..., 1] ** 2 + 2 * S[..., 2] ** 2 + S[..., 3] ** 2 + 2 * S[..., 4] ** 2 + S[..., 5] ** 2
    return np.sqrt(2 * q)


def amd(g, S, cx, cy, cz):
    """get_amd_kernel (AMD.F90:23-73)."""
    c = np.array([cx, cy, cz])

    def row(i, j):
        return sum((g[..., 3 * i + q] * c[q]) * (g[..., 3 * j + q] * c[q]) for q in range(3))
    num = row(0, 0) * S[..., 0] + row(1, 1) * S[..., 3] + row(2, 2) * S[..., 5]
    num = num + 2 * row(0, 1) * S[..., 1] + 2 * row(0, 2) * S[..., 2] + 2 * row(1, 2) * S[..., 4]
    den = (g ** 2).sum(axis=-1)
    return np.maximum(-num / (den + 1e-32), 0.0)


def sigma(g):
    """get_sigma_kernel (sigma.F90:30-100): the singular values of the
    gradient from the invariants of G = A^T A (closed-form cubic roots, with
    the reference's guards), nu = s3 (s1 - s2) (s2 - s3) / s1^2."""
    d = [g[..., c] for c in range(9)]
    G11 = d[0] * d[0] + d[3] * d[3] + d[6] * d[6]
    G12 = d[0] * d[1] + d[3] * d[4] + d[6] * d[7]
    G13 = d[0] * d[2] + d[3] * d[5] + d[6] * d[8]
    G22 = d[1] * d[1] + d[4] * d[4] + d[7] * d[7]
    G23 = d[1] * d[2] + d[4] * d[5] + d[7] * d[8]
    G33 = d[2] * d[2] + d[5] * d[5] + d[8] * d[8]
    I1 = G11 + G22 + G33
    I1sq = I1 * I1
    I1cu = I1sq * I1
    I2 = 0.5 * (-G11 * G11 - G22 * G22 - G33 * G33 - 2 * G12 * G12 - 2 * G13 * G13 - 2 * G23 * G23 + I1sq)
    I3 = G11 * (G22 * G33 - G23 * G23) + G12 * (G13 * G23 - G12 * G33) + G13 * (G12 * G23 - G22 * G13)
    a1 = np.maximum(I1sq / 9.0 - I2 / 3.0, 0.0)
    a2 = I1cu / 27.0 - I1 * I2 / 6.0 + I3 / 2.0
    a1s = np.sqrt(a1)
    t = np.clip(a2 / (a1 * a1s + 1e-13), -1.0, 1.0)
    a3 = np.arccos(t) / 3.0
    s1sq = np.maximum(I1 / 3.0 + 2 * a1s * np.cos(a3), 0.0)
    s1 = np.sqrt(s1sq)
    s2 = np.sqrt(np.maximum(-2 * a1s * np.cos(np.pi / 3.0 + a3) + I1 / 3.0, 0.0))
    s3 = np.sqrt(np.maximum(-2 * a1s * np.cos(np.pi / 3.0 - a3) + I1 / 3.0, 0.0))
    return s3 * (s1 - s2) * (s2 - s3) / (s1sq + 1e-15)


KAPPA = 0.40   # constants.F90:34


def wall_model(spC, u, v, uhat, dz, z0):
    """computeWallStress with the standard Moeng wall model (wallmodel.F90:
    33-77 WallModel 1, compute_and_bcast_surface_Mn :106-134,
    getSurfaceQuantities :162-199 unstratified): the first-level means
    Umn = Re uhat(1,1,1) / (nx ny) and Uspmn of sqrt(u^2 + v^2) (through the
    same transform), ustar = Uspmn kappa / log(dz / 2 / z0),
    WallMFactor = -ustar^2 / (Uspmn + 1e-13).  Returns (Uspmn, ustar, M, Umn)."""
    nx, ny, _ = spC.shape
    mf = 1.0 / (nx * ny)
    Umn = np.real(uhat[0, 0, 0]) * mf
    Uspmn = np.real(spC.fft(np.sqrt(u * u + v * v))[0, 0, 0]) * mf
    ustar = Uspmn * KAPPA / np.log(dz / 2.0 / z0)
    return Uspmn, ustar, -ustar * ustar / (Uspmn + 1e-13), Umn


def dynamic_procedure(sp, dx, dy, deltaRat, u, v, w, uhat, vhat, what, S, Dsgs):
    """DoStandardDynamicProcedure (standardDynamicProcedure.F90:74-163) and
    planarAverage_and_TakeRatio (:222-250): returns lambda_1d (nz)."""
    kfx = (((2.0 / 3.0) * np.pi) / dx) / deltaRat
    kfy = (((2.0 / 3.0) * np.pi) / dy) / deltaRat
    G = ((np.abs(sp.k1) < kfx)[:, None, None] & (np.abs(sp.k2) < kfy)[None, :, None]).astype(float)
    tf = lambda f: sp.ifft(sp.fft(f) * G)  # noqa: E731
    Sf = np.stack([tf(S[..., c]) for c in range(6)], axis=-1)
    Df = smagorinsky(Sf)
    M = np.stack([tf(Dsgs * S[..., c]) - deltaRat ** 2 * Df * Sf[..., c] for c in range(6)], axis=-1)
    uf = [sp.ifft(h * G) for h in (uhat, vhat, what)]
    vel = (u, v, w)
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    Lij = np.stack([tf(vel[i] * vel[j]) - uf[i] * uf[j] for i, j in pairs], axis=-1)
    wgt = np.array([1, 2, 2, 1, 2, 1.0])
    num = (Lij * M * wgt).sum(axis=-1)
    den = 2 * (M * M * wgt).sum(axis=-1)
    nx, ny, nz = u.shape
    mf = 1.0 / (nx * ny)
    a = np.maximum(num.sum(axis=(0, 1)) * mf, 0.0)
    return a / (den.sum(axis=(0, 1)) * mf + 1e-18)


def rhs_sgs(spC, spE, dz, periodic, duidxjC, duidxjE, urhs, vrhs, wrhs, model=0, explicit_edge=False,
            cglobal=1.0, cmodelC=None, cmodelE=None, camd=(1.0, 1.0, 1.0), lam=None, wall=None):
    """sgs_igrid%getRHS_SGS (sgsmod_igrid.F90:204-266) with getTauSGS
    (:156-201), multiply_by_model_constant and interpolate_eddy_viscosity
    (eddyViscosity.F90:74-122) for the eddy-viscosity models (0 Smagorinsky,
    1 sigma, 2 AMD).  wall = (u, v, uhat, vhat, z0): the Moeng wall model
    (wall_model), whose stresses WallMFactor uhat / vhat of the first cell
    level, transformed back on the edge grid, replace tau13 and tau23 on the
    bottom edge plane (embed_WM_stress, wallmodel.F90:79-86).  Returns
    (urhs, vrhs, wrhs, tau dict, nuC, nuE)."""
    from .spectral import pade6_apply
    from . import STG_D1_C2E, STG_D1_E2C, STG_IN_C2E
    nx, ny, nz = spC.shape
    SC, SE = sij(duidxjC), sij(duidxjE)

    def kernel(g, S):
        return smagorinsky(S) if model == 0 else (sigma(g) if model == 1 else amd(g, S, *camd))
    nuC = kernel(duidxjC, SC)
    interp = lambda f: pade6_apply(STG_IN_C2E, f, dz, nz, 0, 0, periodic).real  # noqa: E731
    if lam is not None:
        nuC = nuC * np.asarray(lam)[None, None, :]
        nuE = interp(nuC)
    else:
        nuC = nuC * (np.asarray(cmodelC)[None, None, :] if cmodelC is not None else cglobal)
        if explicit_edge:
            nuE = kernel(duidxjE, SE) * (np.asarray(cmodelE)[None, None, :] if cmodelE is not None else cglobal)
        else:
            nuE = np.maximum(interp(nuC), 0.0)
    tau = {"11": -2 * nuC * SC[..., 0], "12": -2 * nuC * SC[..., 1], "13": -2 * nuE * SE[..., 2],
           "22": -2 * nuC * SC[..., 3], "23": -2 * nuE * SE[..., 4], "33": -2 * nuC * SC[..., 5]}
    if wall is not None:
        u, v, uhat, vhat, z0 = wall
        M = wall_model(spC, u, v, uhat, dz, z0)[2]
        for key, fh in (("13", uhat), ("23", vhat)):
            th = np.zeros((spC.nxh, ny, nz + 1), complex)
            th[:, :, 0] = M * fh[:, :, 0]
            tau[key] = tau[key].copy()
            tau[key][:, :, 0] = spE.ifft(th)[:, :, 0]
    ddz = lambda op, f: pade6_apply(op, f, dz, nz, 0, 0, periodic)  # noqa: E731
    urhs = urhs - spC.mtimes_ik(spC.fft(tau["11"]), 1)
    vrhs = vrhs - spC.mtimes_ik(spC.fft(tau["22"]), 2)
    wrhs = wrhs - ddz(STG_D1_C2E, spC.fft(tau["33"]))
    f12 = spC.fft(tau["12"])
    vrhs = vrhs - spC.mtimes_ik(f12, 1)
    urhs = urhs - spC.mtimes_ik(f12, 2)
    f13 = spE.fft(tau["13"])
    urhs = urhs - ddz(STG_D1_E2C, f13)
    wrhs = wrhs - spE.mtimes_ik(f13, 1)
    f23 = spE.fft(tau["23"])
    vrhs = vrhs - ddz(STG_D1_E2C, f23)
    wrhs = wrhs - spE.mtimes_ik(f23, 2)
    return urhs, vrhs, wrhs, tau, nuC, nuE
