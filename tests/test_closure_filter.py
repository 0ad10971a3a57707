"""CPU checks of the LES filter of the closure data (padeops_amd/closure_dns.py
filter_snapshot, torch on the host): a DNS field band-limited to kc/2 passes
the cutoff unchanged, so its filtered samples must equal the analytic field
at the LES cell centres (u, v) and edges (w), which pins the half-cell z
shifts between the staggered grids, and its subgrid stress must vanish
(every product stays below the cutoff).  A mode above kc is removed."""
import math

import numpy as np
import torch

from padeops_amd import closure_dns as D


def _fields(N, modes):
    dz = 2 * math.pi / N
    x = torch.arange(N, dtype=torch.float64) * dz
    zc = (torch.arange(N, dtype=torch.float64) + 0.5) * dz
    ze = torch.arange(N + 1, dtype=torch.float64) * dz

    def f(z, ph):
        Z, Y, X = torch.meshgrid(z, x, x, indexing="ij")
        return sum(a * torch.cos(kx * X + ky * Y + kz * Z + ph) for a, kx, ky, kz in modes)
    return f(zc, 0.3), f(zc, 1.1), f(ze, -0.4), f


def test_band_limited_field_passes_and_has_no_stress():
    N, m = 32, 16
    kc = D.kc_of(m)
    assert kc == 5
    modes = [(1.0, 1, 2, -2), (0.5, -2, 1, 1), (0.3, 0, 0, 2)]
    u, v, w, f = _fields(N, modes)
    ub, vb, wb, tau = D.filter_snapshot(u, v, w, m)
    Dz = 2 * math.pi / m
    Zc = (torch.arange(m, dtype=torch.float64) + 0.5) * Dz
    Ze = torch.arange(m + 1, dtype=torch.float64) * Dz
    X = torch.arange(m, dtype=torch.float64) * Dz

    def at(z, ph):
        Zg, Y, Xg = torch.meshgrid(z, X, X, indexing="ij")
        return sum(a * torch.cos(kx * Xg + ky * Y + kz * Zg + ph) for a, kx, ky, kz in modes)
    assert torch.abs(ub - at(Zc, 0.3)).max() < 1e-12
    assert torch.abs(vb - at(Zc, 1.1)).max() < 1e-12
    assert torch.abs(wb - at(Ze, -0.4)).max() < 1e-12
    assert torch.abs(tau).max() < 1e-12


def test_cutoff_removes_high_modes_and_grids_share_transforms():
    N = 32
    modes = [(1.0, 1, 0, 0), (1.0, 0, 7, 0)]      # k = 7 > kc(16) = 5, < kc(32) = 10
    u, v, w, _ = _fields(N, modes)
    (u16, _, _, t16), (u32, _, _, t32) = D.filter_snapshot(u, v, w, (16, 32))
    X16 = torch.arange(16, dtype=torch.float64) * (2 * math.pi / 16)
    assert torch.abs(u16 - torch.cos(X16 + 0.3).view(1, 1, -1)).max() < 1e-12
    # the 32^3 grid keeps both modes; its stress is that of the products
    # above its cutoff (none of k = 1, 7 alone)
    assert torch.abs(u32.mean()) < 1e-12
    single = D.filter_snapshot(u, v, w, 16)
    assert torch.equal(single[0], u16) and torch.equal(single[3], t16)
    assert np.isfinite(t32.numpy()).all()
