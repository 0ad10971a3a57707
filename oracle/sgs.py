"""TEST INFRASTRUCTURE ONLY (the checker): CPU restatement of igrid's SGS
eddy-viscosity kernels and standard dynamic procedure, in numpy
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
