"""Synthetic input fields generated on the device (torch supplies the random
numbers and the one-off spectral synthesis; nothing here is on the measured
path).

turbulence(): a divergence-free random-phase velocity field on the triply
periodic n^3 grid of side 2 pi with energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2),
scaled to rms 1 per component, seeded.  Returns the three components as
flat tensors in the reference's Fortran layout (x fastest)."""


def turbulence(n, k0, seed, dtype, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kr = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    KX = k1.view(n, 1, 1)
    KY = k1.view(1, n, 1)
    KZ = kr.view(1, 1, -1)
    K2 = KX * KX + KY * KY + KZ * KZ
    Km = torch.sqrt(K2).clamp_min(1.0)
    amp = torch.sqrt(Km ** 4 * torch.exp(-2.0 * (Km / k0) ** 2) / (4 * torch.pi * Km ** 2))
    amp[0, 0, 0] = 0.0
    uh = []
    for _ in range(3):
        re = torch.randn(K2.shape, generator=g, device=device, dtype=torch.float64)
        im = torch.randn(K2.shape, generator=g, device=device, dtype=torch.float64)
        uh.append(torch.complex(re, im) * amp)
    kd = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / torch.where(K2 == 0, torch.ones_like(K2), K2)
    uh = [uh[0] - KX * kd, uh[1] - KY * kd, uh[2] - KZ * kd]
    # torch axes (a0, a1, a2) with a2 fastest are the Fortran (z, y, x): the
    # x velocity is the component along a2
    u = [torch.fft.irfftn(c, s=(n, n, n)) for c in uh]
    del uh, kd
    rms = torch.sqrt((u[0] ** 2 + u[1] ** 2 + u[2] ** 2).mean() / 3.0)
    return [(c / rms).contiguous().view(-1).to(dtype) for c in (u[2], u[1], u[0])]


def wall_bounded(nx, ny, nz, k0, seed, dtype, device, z0=1e-4, kappa=0.4):
    """Fields of the wall-bounded box (nx, ny, nz) cells, periodic in x and y
    with side 2 pi, non-periodic in z over [0, Lz] with dz = 2 pi / nx (so
    Lz = nz dz): the log-law mean profile U(z) = ln(z / z0) / kappa plus the
    isotropic turbulence of turbulence(nx, k0, seed) restricted to its first
    planes.  Returns (uC, uE, wE): u at the nz cell centres z = (k + 1/2) dz,
    u at the nz + 1 edges z = k dz (U taken as ln(max(z, z0) / z0) / kappa,
    zero at the wall), and the w component at the edges; flat tensors in the
    Fortran layout (x fastest).  nx == ny and nz < nx are required."""
    import math

    import torch
    assert nx == ny and nz < nx
    n = nx
    u, _, w = turbulence(n, k0, seed, torch.float64, device)
    u = u.view(n, n, n)  # torch view (z, y, x) of the Fortran (x, y, z) layout
    w = w.view(n, n, n)
    dz = 2 * math.pi / n
    zc = (torch.arange(nz, device=device, dtype=torch.float64) + 0.5) * dz
    ze = torch.arange(nz + 1, device=device, dtype=torch.float64) * dz
    Uc = torch.log(zc / z0) / kappa
    Ue = torch.log(torch.clamp(ze, min=z0) / z0) / kappa
    # turbulence on cells: the average of the two neighbouring planes
    # (cell k lies between planes k and k + 1 of the periodic field)
    tc = 0.5 * (u[:nz] + u[1:nz + 1])
    uC = (tc + Uc.view(nz, 1, 1)).contiguous().view(-1).to(dtype)
    uE = (u[:nz + 1] + Ue.view(nz + 1, 1, 1)).contiguous().view(-1).to(dtype)
    wE = w[:nz + 1].contiguous().view(-1).to(dtype)
    del u, w
    return uC, uE, wE
