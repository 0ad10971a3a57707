"""TEST INFRASTRUCTURE ONLY (the checker): numpy restatement of the explicit
filters gaussianstuff (src/filters/gaussian.F90) and lstsqstuff
(src/filters/lstsq.F90), and of the Fourier-collocation derivatives of
fftstuff (src/derivatives/ffts.F90).  Used by tests/ only; the product never
imports it.

Filters act along one axis of an array (axis 0 = the reference's first
index).  Row formulas follow the reference line by line:
  * periodic rows (gaussian.F90:137-181): c0 f_i + sum_q c_q (f_{i+q} + f_{i-q})
    with wrap-around;
  * non-periodic bc = 0 (gaussian.F90:186-199, 257-271; lstsq.F90:169-213):
    one-sided rows b1..b4 at each end, the interior stencil elsewhere;
  * non-periodic bc = +1 / -1 (gaussian.F90:200-251, 272-323): the interior
    stencil with the ghost values f(1-m) = bc f(1+m), f(n+m) = bc f(n-m).
Fourier derivatives (ffts.F90:185-308): r2c transform along the axis,
multiply by i k (or -k^2) with k from GetWaveNums (ffts.F90:370-388), the
last mode (index n/2) zeroed, c2r.
"""
from fractions import Fraction as Fr

import numpy as np

# gaussian.F90:15-19
GAUSS = [Fr(3565, 10368), Fr(3091, 12960), Fr(1997, 25920), Fr(149, 12960), Fr(107, 103680)]
# lstsq.F90:15-19 (default-kind literals: single precision values widened)
LSQ = [0.5, float(np.float32(0.6744132)) / 2, 0.0, float(np.float32(-0.1744132)) / 2, 0.0]
# gaussian.F90:24-43 / lstsq.F90:24-43: centre, then pairs at distance 1, 2, 3
BROWS = [[Fr(5, 6), Fr(1, 6)], [Fr(2, 3), Fr(1, 6)], [Fr(31, 64), Fr(7, 32), Fr(5, 128)],
         [Fr(17, 48), Fr(15, 64), Fr(7, 96), Fr(1, 64)]]


def coeffs(kind):
    return [float(c) for c in (GAUSS if kind == "gaussian" else LSQ)]


def _interior(f, c, i):
    n = f.shape[0]
    acc = c[0] * f[i]
    for q in range(1, 5):
        acc = acc + c[q] * (f[(i + q) % n] + f[(i - q) % n])
    return acc


def filter_axis0(f, kind, periodic, bc1=0, bcn=0):
    """Filter along axis 0 of f (float64 array, n = f.shape[0] >= 8 or 1)."""
    f = np.asarray(f, dtype=np.float64)
    n = f.shape[0]
    if n == 1:
        return f.copy()
    c = coeffs(kind)
    if kind == "lstsq":
        bc1 = bcn = 0
    out = np.empty_like(f)
    if periodic:
        for i in range(n):
            out[i] = _interior(f, c, i)
        return out
    for i in range(4, n - 4):
        out[i] = _interior(f, c, i)
    # left end
    for j in range(4):
        if bc1 == 0:
            # row 1 has no partner below the first point (gaussian.F90:188-189)
            b = [float(x) for x in BROWS[j]]
            acc = b[0] * f[j]
            for q in range(1, len(b)):
                acc = acc + b[q] * f[j + q]
                if j - q >= 0:
                    acc = acc + b[q] * f[j - q]
            out[j] = acc
        else:
            acc = c[0] * f[j]
            for q in range(1, 5):
                m = j - q
                lo = f[m] if m >= 0 else bc1 * f[-m]
                acc = acc + c[q] * (f[j + q] + lo)
            out[j] = acc
    # right end (mirror image)
    for j in range(4):
        i = n - 1 - j
        if bcn == 0:
            b = [float(x) for x in BROWS[j]]
            acc = b[0] * f[i]
            for q in range(1, len(b)):
                acc = acc + b[q] * f[i - q]
                if i + q <= n - 1:
                    acc = acc + b[q] * f[i + q]
            out[i] = acc
        else:
            acc = c[0] * f[i]
            for q in range(1, 5):
                m = i + q
                hi = f[m] if m <= n - 1 else bcn * f[2 * (n - 1) - m]
                acc = acc + c[q] * (hi + f[i - q])
            out[i] = acc
    return out


def filter_axis(f, axis, kind, periodic, bc1=0, bcn=0):
    g = np.moveaxis(np.asarray(f, dtype=np.float64), axis, 0)
    return np.moveaxis(filter_axis0(g, kind, periodic, bc1, bcn), 0, axis)


def wavenums(n, dx):
    """GetWaveNums (ffts.F90:370-388): the first n/2+1 entries of
    ifftshift((-pi + (i-1) 2 pi / (n - mod(n,2))) / dx)."""
    dummy = n - n % 2
    k = (-np.pi + np.arange(n) * 2.0 * np.pi / dummy) / dx
    if n % 2 == 0:
        k = np.concatenate([k[n // 2:], k[:n // 2]])
    else:
        k = np.concatenate([k[(n - 1) // 2:], k[:(n - 1) // 2]])
    return k[: n // 2 + 1]


def fourier_deriv(f, axis, dx, order=1):
    """ffts%dd1/dd2/dd3 (order 1) and d2d1/d2d2/d2d3 (order 2) along axis."""
    f = np.asarray(f, dtype=np.float64)
    n = f.shape[axis]
    if n == 1:
        return np.zeros_like(f)
    k = wavenums(n, dx)
    fh = np.fft.rfft(f, axis=axis)
    shape = [1] * f.ndim
    shape[axis] = k.size
    mult = (1j * k) if order == 1 else -(k * k)
    fh = fh * mult.reshape(shape)
    idx = [slice(None)] * f.ndim
    idx[axis] = n // 2
    fh[tuple(idx)] = 0
    return np.fft.irfft(fh, n=n, axis=axis)
