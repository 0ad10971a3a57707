"""Seeded test fields and layout helpers shared by the tests and the bench."""
import numpy as np


def to_dev(a, dtype, device):
    """numpy Fortran-ordered array -> flat contiguous GPU tensor (same memory order)."""
    import torch
    flat = np.ascontiguousarray(np.asarray(a).ravel(order="F"))
    return torch.from_numpy(flat).to(device=device, dtype=dtype)


def from_dev(t, shape):
    return t.detach().cpu().double().numpy().reshape(shape, order="F")


def rel_l2(a, b):
    """||a - b|| / ||b|| in FP64; complex arrays keep both parts."""
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    a = np.asarray(a, dtype=np.complex128 if cplx else np.float64)
    b = np.asarray(b, dtype=np.complex128 if cplx else np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def random_field(shape, seed=1234):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(shape))


def fourier_modes(n, nmodes=32, kmax=16, seed=1234, L=2 * np.pi):
    """Scalar field on [0,L)^3: sum of nmodes Fourier modes with integer wave
    vectors |k| <= kmax drawn from the seed, amplitude ~ |k|^(-5/3), uniform
    random phases.  Returns (u, modes) with modes = [(kx,ky,kz,amp,phase)]."""
    rng = np.random.default_rng(seed)
    modes = []
    while len(modes) < nmodes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        kk = np.sqrt((k ** 2).sum())
        if kk == 0 or kk > kmax:
            continue
        modes.append((int(k[0]), int(k[1]), int(k[2]), kk ** (-5.0 / 3.0), rng.uniform(0, 2 * np.pi)))
    x = np.arange(n) * (L / n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((n, n, n))
    s = 2 * np.pi / L
    for kx, ky, kz, a, ph in modes:
        u += a * np.cos(s * (kx * X + ky * Y + kz * Z) + ph)
    return np.asfortranarray(u), modes


def isotropic_velocity(n, k0=4.0, seed=1234, L=2 * np.pi):
    """Divergence-free random-phase velocity with E(k) ~ k^4 exp(-2 (k/k0)^2),
    rescaled to rms 1 per component on average.  Built in spectral space by
    projecting random complex amplitudes onto the plane normal to k."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    Kmag = np.sqrt(K2)
    Kmag[0, 0, 0] = 1.0
    E = Kmag ** 4 * np.exp(-2.0 * (Kmag / k0) ** 2)
    amp = np.sqrt(E / (4 * np.pi * Kmag ** 2))
    amp[0, 0, 0] = 0.0
    uh = [(rng.standard_normal(K2.shape) + 1j * rng.standard_normal(K2.shape)) * amp for _ in range(3)]
    kd = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / np.where(K2 == 0, 1.0, K2)
    uh = [uh[0] - KX * kd, uh[1] - KY * kd, uh[2] - KZ * kd]
    u = [np.real(np.fft.ifftn(c)) for c in uh]
    rms = np.sqrt(np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / 3.0)
    return [np.asfortranarray(c / rms) for c in u]
