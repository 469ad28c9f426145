"""AMultilayer (include/AMultilayer.h, src/AMultilayer.cxx): thin-film stacks
evaluated with the coherent transfer-matrix method, and the (wavelength,
angle) tables the device reads.

The stack is a list of media from the incident (top) half-space through the
films to the exit (bottom) half-space; a medium's complex index is
n + i k from its ARefractiveIndex (ARefractiveIndex::GetComplexRefractiveIndex,
include/ARefractiveIndex.h:47-51). CoherentTMM follows the steps of
AMultilayer::CoherentTMM (src/AMultilayer.cxx:240-483): Snell's law with
complex angles, the forward-angle choice for the two half-spaces, per-layer
phase (clamped at an imaginary part of 35 for opaque films), the Fresnel
amplitudes of each interface, the product of the layer matrices, and the
power reflectance and transmittance. It is vectorised over angles and
wavelengths with numpy.

The device evaluates a multilayer through a TH2-style table of the
unpolarised reflectance and transmittance on bin centres, as
AMultilayer::PreCalculateCoherentTMM (include/AMultilayer.h:243-262) builds it
and CoherentTMMMixed then reads it with TH2::Interpolate. A stack that was not
precalculated, which the reference evaluates exactly at any wavelength, is
tabulated on 1 nm x 0.25 deg bins over 200-1000 nm widened to the traced
rays' wavelengths (default_grid): AOpticsManager re-flattens the scene when a
batch reaches outside the band its table covers.
"""
import math

import numpy as np

EPS = np.finfo(float).eps
kS, kP = 0, 1
NM = 1e-9 * (1e2 * 1.0)                       # AOpticsManager::nm() in cm
# default table: 1 nm steps over 200-1000 nm, 0.25 deg steps over 0-90 deg (bin centres)
DEFAULT_GRID = (801, 199.5 * NM, 1000.5 * NM, 361, math.radians(-0.125), math.radians(90.125))


def default_grid(lambda_range=None):
    """DEFAULT_GRID, its wavelength axis widened in whole 1 nm bins to hold
    lambda_range = (lo, hi) in cm with a bin to spare on either side: every
    wavelength in the range then lies between two bin centres."""
    nx, xlo, xhi, ny, ylo, yhi = DEFAULT_GRID
    if lambda_range is None:
        return DEFAULT_GRID
    lo, hi = float(lambda_range[0]), float(lambda_range[1])
    below = max(0, math.ceil((xlo + NM - lo) / NM)) if lo < xlo + NM else 0
    above = max(0, math.ceil((hi - (xhi - NM)) / NM)) if hi > xhi - NM else 0
    # never below 1 nm: the first bin centre stays positive
    below = min(below, int((xlo - 0.5 * NM) / NM))
    return (nx + below + above, xlo - below * NM, xhi + above * NM, ny, ylo, yhi)


def grid_covers(grid, lo, hi):
    """True if every wavelength in [lo, hi] (cm) lies between the first and
    the last bin centre of the table's wavelength axis."""
    nx, xlo, xhi = grid[0], grid[1], grid[2]
    w = (xhi - xlo) / nx
    return xlo + 0.5 * w <= lo and hi <= xhi - 0.5 * w


def _forward(n, th):
    """tmm's is_forward_angle (src/AMultilayer.cxx:120-176) without the
    diagnostics: the decaying wave, or the one with positive Poynting flux."""
    nc = n * np.cos(th)
    return np.where(np.abs(nc.imag) > 100 * EPS, nc.imag > 0, nc.real > 0)


def coherent_tmm(pol, n_list, d_list, th_0, lam_vac):
    """Power (R, T) of the stack for polarisation pol (kS / kP).
    n_list: complex indices [L, m] (or [L]); d_list: thicknesses [L] (inf at
    both ends); th_0, lam_vac: arrays broadcast to [m]."""
    th_0 = np.asarray(th_0, dtype=complex)
    lam_vac = np.asarray(lam_vac, dtype=float)
    m = np.broadcast(th_0, lam_vac).shape
    th_0 = np.broadcast_to(th_0, m).ravel()
    lam_vac = np.broadcast_to(lam_vac, m).ravel()
    n = np.asarray(n_list, dtype=complex)
    if n.ndim == 1:
        n = np.repeat(n[:, None], th_0.size, axis=1)
    L = n.shape[0]
    d = np.asarray(d_list, dtype=float)
    with np.errstate(all="ignore"):
        th = np.arcsin(n[0] * np.sin(th_0) / n)             # Snell in every layer
        th[0] = np.where(_forward(n[0], th[0]), th[0], np.pi - th[0])
        th[-1] = np.where(_forward(n[-1], th[-1]), th[-1], np.pi - th[-1])
        cos_th = np.cos(th)
        kz = 2 * np.pi * n * cos_th / lam_vac
        delta = np.zeros_like(kz)
        if L > 2:
            delta[1:-1] = kz[1:-1] * d[1:-1, None]
            opaque = delta[1:-1].imag > 35
            delta[1:-1] = np.where(opaque, delta[1:-1].real + 35j, delta[1:-1])
        ni, nf = n[:-1], n[1:]
        ci, cf = cos_th[:-1], cos_th[1:]
        ii = ni * ci
        if pol == kS:
            ff = nf * cf
            r = (ii - ff) / (ii + ff)
            t = 2 * ii / (ii + ff)
        else:
            fi, i_f = nf * ci, ni * cf
            r = (fi - i_f) / (fi + i_f)
            t = 2 * ii / (fi + i_f)
        # Mtilde = [[1, r0], [r0, 1]] / t0 * prod_i (1/t_i) diag(e^-jd, e^jd) [[1, r_i], [r_i, 1]]
        a, b, c, e = np.ones_like(th_0), np.zeros_like(th_0), np.zeros_like(th_0), np.ones_like(th_0)
        for k in range(1, L - 1):
            em, ep = np.exp(-1j * delta[k]), np.exp(1j * delta[k])
            m00, m01, m10, m11 = em / t[k], em * r[k] / t[k], ep * r[k] / t[k], ep / t[k]
            a, b, c, e = a * m00 + b * m10, a * m01 + b * m11, c * m00 + e * m10, c * m01 + e * m11
        a, b, c, e = (a + r[0] * c) / t[0], (b + r[0] * e) / t[0], (r[0] * a + c) / t[0], (r[0] * b + e) / t[0]
        rr = c / a
        tt = 1 / a
        R = np.abs(rr) ** 2
        n_i, n_f, c_i, c_f = n[0], n[-1], np.cos(th_0), cos_th[-1]
        if pol == kS:
            T = np.abs(tt * tt) * ((n_f * c_f).real / (n_i * c_i).real)
        else:
            T = np.abs(tt * tt) * ((n_f * np.conj(c_f)).real / (n_i * np.conj(c_i)).real)
    return R.reshape(m), T.reshape(m)


def th2_interpolate(values, nx, xlo, xhi, ny, ylo, yhi, x, y):
    """TH2::Interpolate(x, y) over a table of bin contents values[ny, nx]:
    bilinear between the four nearest bin centres, the neighbour index held
    inside the axis (so the outer half bins are flat along that axis), and 0
    outside the axis ranges. Restated from ROOT's documented behaviour."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    wx, wy = (xhi - xlo) / nx, (yhi - ylo) / ny

    def axis(v, n, lo, w, hi):
        inner = np.where((v >= lo) & (v < hi), n * (v - lo) / (hi - lo), 0.0)
        b = np.where(v < lo, 0, np.where(~(v < hi), n + 1, 1 + np.floor(inner).astype(np.int64)))
        b = np.clip(b, 0, n + 1)
        up = lo + b * w
        below = (up - v) > w / 2                         # in the lower half of the bin
        b1 = np.where(below, b - 1, b)
        c1 = lo + (b1 - 0.5) * w
        c2 = lo + (b1 + 0.5) * w
        return b, np.clip(b1, 1, n), np.clip(b1 + 1, 1, n), c1, c2

    bx, ix1, ix2, x1, x2 = axis(x, nx, xlo, wx, xhi)
    by, iy1, iy2, y1, y2 = axis(y, ny, ylo, wy, yhi)
    v = np.asarray(values, dtype=float)
    q11, q21 = v[iy1 - 1, ix1 - 1], v[iy1 - 1, ix2 - 1]
    q12, q22 = v[iy2 - 1, ix1 - 1], v[iy2 - 1, ix2 - 1]
    dd = (x2 - x1) * (y2 - y1)
    f = (q11 / dd * (x2 - x) * (y2 - y) + q21 / dd * (x - x1) * (y2 - y) + q12 / dd * (x2 - x) * (y - y1)
         + q22 / dd * (x - x1) * (y - y1))
    inside = (bx >= 1) & (bx <= nx) & (by >= 1) & (by <= ny)
    return np.where(inside, f, 0.0)


class AMultilayer:
    """include/AMultilayer.h:31-303 (coherent stacks)."""

    def __init__(self, top, bottom):
        self.media = [top, bottom]
        self.thickness = [math.inf, math.inf]
        self.table = None             # (nx, xlo, xhi, ny, ylo, yhi, R[ny, nx], T[ny, nx])

    def AddLayer(self, idx, thickness):
        """A film right below the top medium (src/AMultilayer.cxx:212-224)."""
        self.media.insert(1, idx)
        self.thickness.insert(1, float(thickness))

    def InsertLayer(self, idx, thickness):
        """A film right above the bottom medium (src/AMultilayer.cxx:226-238)."""
        self.media.insert(len(self.media) - 1, idx)
        self.thickness.insert(len(self.thickness) - 1, float(thickness))

    def ChangeThickness(self, i, thickness):
        if i < 1 or i > len(self.thickness) - 2:
            raise IndexError(f"cannot change the thickness of layer {i}")
        self.thickness[i] = float(thickness)
        self.table = None

    def _indices(self, lam, reverse=False):
        lam = np.atleast_1d(np.asarray(lam, dtype=float))
        media = self.media[::-1] if reverse else self.media
        n = np.array([[complex(m.GetRefractiveIndex(float(l)), m.GetExtinctionCoefficient(float(l))) for l in lam]
                      for m in media])
        d = self.thickness[::-1] if reverse else self.thickness
        return n, d

    def CoherentTMM(self, polarization, th_0, lam_vac, reverse=False):
        """(reflectance, transmittance) for one polarisation; th_0 and lam_vac
        may be arrays of one shape (or one a scalar)."""
        th = np.asarray(th_0, dtype=complex)
        lam = np.asarray(lam_vac, dtype=float)
        shape = np.broadcast(th, lam).shape
        lam_b = np.broadcast_to(lam, shape).ravel()
        n, d = self._indices(lam_b, reverse)
        R, T = coherent_tmm(polarization, n, d, np.broadcast_to(th, shape).ravel(), lam_b)
        if shape == ():
            return float(R[0]), float(T[0])
        return R.reshape(shape), T.reshape(shape)

    def CoherentTMMP(self, th_0, lam_vac):
        return self.CoherentTMM(kP, th_0, lam_vac)

    def CoherentTMMS(self, th_0, lam_vac):
        return self.CoherentTMM(kS, th_0, lam_vac)

    def CoherentTMMMixed(self, th_0, lam_vac):
        """Unpolarised (R, T): the mean of p and s, or the precalculated table
        read with TH2::Interpolate(lam_vac, Re th_0) once there is one."""
        if self.table is not None:
            nx, xlo, xhi, ny, ylo, yhi, R, T = self.table
            th = np.real(np.asarray(th_0, dtype=complex))
            r = th2_interpolate(R, nx, xlo, xhi, ny, ylo, yhi, lam_vac, th)
            t = th2_interpolate(T, nx, xlo, xhi, ny, ylo, yhi, lam_vac, th)
            return (float(r), float(t)) if np.ndim(r) == 0 else (r, t)
        rp, tp = self.CoherentTMM(kP, th_0, lam_vac)
        rs, ts = self.CoherentTMM(kS, th_0, lam_vac)
        return (rp + rs) / 2.0, (tp + ts) / 2.0

    def tabulate(self, lam_nbins, lam_min, lam_max, th_nbins, th_min, th_max):
        """Unpolarised R and T at the bin centres of the (lambda, theta) grid,
        arrays [th_nbins, lam_nbins]."""
        lam = lam_min + (np.arange(1, lam_nbins + 1) - 0.5) * ((lam_max - lam_min) / lam_nbins)
        th = th_min + (np.arange(1, th_nbins + 1) - 0.5) * ((th_max - th_min) / th_nbins)
        L, TH = np.meshgrid(lam, th)
        n, d = self._indices(lam)
        n = np.repeat(n[:, None, :], th_nbins, axis=1).reshape(n.shape[0], -1)
        rp, tp = coherent_tmm(kP, n, d, TH.ravel(), L.ravel())
        rs, ts = coherent_tmm(kS, n, d, TH.ravel(), L.ravel())
        return ((rp + rs) / 2.0).reshape(th_nbins, lam_nbins), ((tp + ts) / 2.0).reshape(th_nbins, lam_nbins)

    def PreCalculateCoherentTMM(self, lam_nbins, lam_min, lam_max, th_nbins, th_min, th_max):
        """include/AMultilayer.h:243-262: tabulate R and T on bin centres; later
        CoherentTMMMixed calls (and the device) interpolate the table."""
        R, T = self.tabulate(lam_nbins, lam_min, lam_max, th_nbins, th_min, th_max)
        self.table = (int(lam_nbins), float(lam_min), float(lam_max), int(th_nbins), float(th_min), float(th_max), R, T)

    def device_table(self, lambda_range=None):
        """The table the device reads: the precalculated one, or the default
        grid widened to lambda_range (default_grid)."""
        if self.table is None:
            nx, xlo, xhi, ny, ylo, yhi = default_grid(lambda_range)
            R, T = self.tabulate(nx, xlo, xhi, ny, ylo, yhi)
            return (nx, xlo, xhi, ny, ylo, yhi, R, T)
        return self.table
