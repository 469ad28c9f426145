"""Spot analysis after the trace, on the device: the host mirror of
AGeoUtil::ContainmentRadius (src/AGeoUtil.cxx:198-308) and of the TH2D it
reads (tutorials/MST.C:272-277, tutorials/unittest_robast.py:590-610).

    h2 = TH2D("", "", 1000, -3, 3, 1000, -3, 3)
    h2.FillRays(ctx, rays, n)                 # the rays' terminal (x, y), RB_FOCUS
    r, x, y = ContainmentRadius(h2, 0.80)     # D80

The fill and the search run in librbgpu (rbgpu_containment_radius); after the
search the histogram's counts and statistics are on the host for GetMean,
GetStdDev, Integral and GetBinContent.
"""
import math

import numpy as np

from . import _abi as A


class TH2D:
    """Fixed binning of ROOT's TH2D(name, title, nx, xlo, xhi, ny, ylo, yhi)."""

    def __init__(self, name, title, nx, xlo, xhi, ny, ylo, yhi):
        self.name, self.title = name, title
        self.nx, self.xlo, self.xhi = int(nx), float(xlo), float(xhi)
        self.ny, self.ylo, self.yhi = int(ny), float(ylo), float(yhi)
        self.counts = None   # [ny, nx] uint32 after ContainmentRadius
        self.stats = None    # {sum w, sum x, sum x^2, sum y, sum y^2}
        self.values = None   # [ny, nx] float64 bin contents set with SetBinContent
        self._source = None

    def SetBinContent(self, bx, by, v):
        """Set bin (bx, by) (1-based; under/overflow bins are not kept)."""
        if self.values is None:
            self.values = np.zeros((self.ny, self.nx))
        if 1 <= bx <= self.nx and 1 <= by <= self.ny:
            self.values[by - 1, bx - 1] = float(v)

    def contents(self):
        """Bin contents [ny, nx]: those set with SetBinContent, else the fill's counts."""
        if self.values is not None:
            return self.values
        self._filled()
        return self.counts.astype(np.float64)

    def Interpolate(self, x, y):
        """TH2::Interpolate: bilinear between the nearest bin centres, 0 outside the axes."""
        from .multilayer import th2_interpolate
        return float(th2_interpolate(self.contents(), self.nx, self.xlo, self.xhi, self.ny, self.ylo, self.yhi, x, y))

    def FillRays(self, ctx, rays, n, status=A.RB_FOCUS, device_ptrs=False):
        """Fill with the terminal (x, y) of the n rays (rb_rays) with `status`;
        the fill runs on the device, with the search, in ContainmentRadius."""
        self._source = (ctx, rays, int(n), int(status), bool(device_ptrs))

    def desc(self):
        d = A.rb_hist2d_desc()
        d.nx, d.ny, d.xlo, d.xhi, d.ylo, d.yhi = self.nx, self.ny, self.xlo, self.xhi, self.ylo, self.yhi
        d.status = self._source[3] if self._source else A.RB_FOCUS
        d.device_ptrs = 1 if self._source and self._source[4] else 0
        return d

    def _filled(self):
        if self.stats is None:
            raise RuntimeError("TH2D: not filled yet (FillRays, then ContainmentRadius)")

    def Integral(self):
        self._filled()
        return float(self.counts.sum(dtype="uint64"))

    def GetBinContent(self, bx, by):
        if 1 <= bx <= self.nx and 1 <= by <= self.ny:
            return float(self.contents()[by - 1, bx - 1])
        return 0.0

    def GetMean(self, axis=1):
        self._filled()
        w = self.stats[0]
        return 0.0 if w == 0 else self.stats[1 if axis == 1 else 3] / w

    def GetStdDev(self, axis=1):
        self._filled()
        w = self.stats[0]
        if w == 0:
            return 0.0
        m = self.GetMean(axis)
        return math.sqrt(max(self.stats[2 if axis == 1 else 4] / w - m * m, 0.0))


def ContainmentRadius(h2, fraction):
    """Radius and centre (r, x, y) of the circle holding `fraction` of the
    histogram's entries (AGeoUtil::ContainmentRadius), computed on the
    device; h2's counts and statistics are filled in on the way."""
    if h2._source is None:
        raise RuntimeError("ContainmentRadius: the TH2D has no ray source (FillRays)")
    ctx, rays, n = h2._source[:3]
    (r, x, y), h2.counts, h2.stats = ctx.containment_radius(rays, n, h2.desc(), fraction, with_histogram=True)
    return r, x, y
