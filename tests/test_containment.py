"""Spot analysis after the trace: the TH2D fill of the rays' terminal (x, y)
and AGeoUtil::ContainmentRadius (src/AGeoUtil.cxx:18-42, 198-308) over it,
as rbgpu_containment_radius runs them on the device.

CPU (pins the oracle): the reference's own known answer (testD80,
tutorials/unittest_robast.py:590-610: 1e7 points of a disk of radius 2 about
(1, -1), uniform in radius, 1000 x 1000 bins over [-3, 3]^2, D80 = 1.6 and the
centre within 1.5 %; numpy's generator stands in for TRandom3), the closed
form of an area-uniform disk (R80 = sqrt(0.8) R), the fill's bin assignment
by construction (points placed inside known bins, boundary and out-of-range
values), and the degenerate spots (no entries, one bin).

GPU: the device histogram equals the oracle's bin for bin; the statistics
agree to 1e-12; and given the device's counts and statistics, the oracle's
search returns the device's (r, x, y) bit for bit. Spots: the C2 fixture's
focal plane, the testD80 disk, empty and out-of-range input."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from robast_amd import _abi as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
D80_BINS = (1000, -3.0, 3.0, 1000, -3.0, 3.0)


def d80_points(n, seed):
    """testD80's source: gRandom->Circle(x, y, 2) scaled by U(0, 1), about (1, -1)."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    scale = rng.uniform(0.0, 1.0, n)
    return 2.0 * np.cos(phi) * scale + 1.0, 2.0 * np.sin(phi) * scale - 1.0


def oracle_d80(x, y, status, want, bins, fraction=0.8):
    counts, stats = O.hist2d(x, y, status, want, *bins)
    return O.containment(counts, stats, *bins, fraction), counts, stats


# ---------------------------------------------------------------- oracle (CPU)

def test_oracle_reference_d80_known_answer():
    x, y = d80_points(10_000_000, seed=1)
    (r, cx, cy), _, _ = oracle_d80(x, y, np.full(len(x), A.RB_FOCUS, np.int32), A.RB_FOCUS, D80_BINS)
    tol = 0.015  # unittest_robast.py:607
    assert abs(cx / 1.0 - 1.0) <= tol
    assert abs(cy / -1.0 - 1.0) <= tol
    assert abs(r / 2.0 - 0.8) / 0.8 <= tol


def test_oracle_uniform_disk_closed_form():
    """Area-uniform disk of radius R: 80 % lie within sqrt(0.8) R of its centre."""
    rng = np.random.default_rng(7)
    n, R, x0, y0 = 2_000_000, 1.2, 0.4, -0.3
    rad = R * np.sqrt(rng.uniform(0, 1, n))
    phi = rng.uniform(0, 2 * np.pi, n)
    x, y = x0 + rad * np.cos(phi), y0 + rad * np.sin(phi)
    (r, cx, cy), counts, stats = oracle_d80(x, y, np.full(n, 3, np.int32), 3, D80_BINS)
    assert abs(r / (math.sqrt(0.8) * R) - 1) < 0.01
    # a flat spot holds nearly the same count as the circle moves a little:
    # the centre is found more loosely than on a peaked one
    assert math.hypot(cx - x0, cy - y0) < 0.03 * R
    assert counts.sum() == n and stats[0] == n


def test_oracle_fill_bins_by_construction():
    nx, ny, lo, hi = 7, 5, -2.0, 3.0
    w = (hi - lo) / nx, (hi - lo) / ny
    rng = np.random.default_rng(3)
    bx = rng.integers(1, nx + 1, 500)
    by = rng.integers(1, ny + 1, 500)
    # inside bin (bx, by): centre plus an offset below 0.4 bin widths
    x = lo + (bx - 0.5) * w[0] + rng.uniform(-0.4, 0.4, 500) * w[0]
    y = lo + (by - 0.5) * w[1] + rng.uniform(-0.4, 0.4, 500) * w[1]
    want = np.zeros((ny, nx), np.uint32)
    np.add.at(want, (by - 1, bx - 1), 1)
    # out of range on either axis (hi itself is overflow), NaN, other status
    ox = np.array([lo - 1e-9, hi, hi + 1, np.nan, 0.0, 0.0, 0.0, 0.0])
    oy = np.array([0.0, 0.0, 0.0, 0.0, lo - 1, hi, np.inf, 0.0])
    st = np.full(508, 3, np.int32)
    st[-1] = 1
    X, Y = np.concatenate([x, ox]), np.concatenate([y, oy])
    counts, stats = O.hist2d(X, Y, st, 3, nx, lo, hi, ny, lo, hi)
    assert np.array_equal(counts, want)
    assert stats[0] == 500
    assert np.allclose(stats[1:], [x.sum(), (x * x).sum(), y.sum(), (y * y).sum()], rtol=1e-13)
    # the lowest edge is in bin 1
    c, _ = O.hist2d(np.array([lo]), np.array([lo]), np.array([3], np.int32), 3, nx, lo, hi, ny, lo, hi)
    assert c[0, 0] == 1


def test_oracle_degenerate_spots():
    bins = (10, -1.0, 1.0, 10, -1.0, 1.0)
    # no entries: centre at the origin, radius 0
    r = O.containment(np.zeros((10, 10), np.uint32), np.zeros(5), *bins, 0.8)
    assert r == (0.0, 0.0, 0.0)
    # every entry in one bin: zero spread, radius 0, centre at the mean
    x = np.full(100, 0.25)
    y = np.full(100, -0.35)
    (rr, cx, cy), _, stats = oracle_d80(x, y, np.full(100, 3, np.int32), 3, bins)
    assert rr == 0.0 and (cx, cy) == (stats[1] / stats[0], stats[3] / stats[0])
    assert math.isclose(cx, 0.25, rel_tol=1e-14) and math.isclose(cy, -0.35, rel_tol=1e-14)


# ---------------------------------------------------------------- device (GPU)

def _device_vs_oracle(ctx, x, y, status, want, bins, fraction=0.8, device=False):
    from robast_amd._lib import make_rays
    n = len(x)
    h = A.rb_hist2d_desc()
    h.nx, h.xlo, h.xhi, h.ny, h.ylo, h.yhi = bins
    h.status = want
    if device:
        import torch
        t = {f: torch.zeros(max(n, 1), dtype=torch.float64, device="cuda") for f in
             ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
        t.update({f: torch.zeros(max(n, 1), dtype=torch.int32, device="cuda") for f in
                  ("status", "npoints", "last_node")})
        t["x"][:n] = torch.from_numpy(x)
        t["y"][:n] = torch.from_numpy(y)
        t["status"][:n] = torch.from_numpy(status)
        from robast_amd._lib import make_rays_device
        rays = make_rays_device(t)
        h.device_ptrs = 1
    else:
        soa = O.new_soa(n)
        soa["x"][:], soa["y"][:], soa["status"][:] = x, y, status
        rays = make_rays(soa)
    (r, cx, cy), counts, stats = ctx.containment_radius(rays, n, h, fraction, with_histogram=True)
    oc, ost = O.hist2d(x, y, status, want, *bins)
    assert np.array_equal(counts, oc)
    assert np.allclose(stats, ost, rtol=1e-12, atol=0)
    # the search itself, on the device's own histogram: bit for bit
    assert O.containment(counts, stats, *bins, fraction) == (r, cx, cy)
    return r, cx, cy


@pytest.mark.gpu
def test_gpu_c2_focal_spot_matches_oracle(ctx):
    print("containment: C2 fixture focal plane", flush=True)
    z = np.load(os.path.join(GOLDEN, "c2_sample.npz"))
    x, y, st = z["out_x"].copy(), z["out_y"].copy(), z["out_status"].copy()
    assert (st == A.RB_FOCUS).sum() > 2000
    bins = (400, -125.0, 125.0, 400, -125.0, 125.0)
    for device in (False, True):
        r, cx, cy = _device_vs_oracle(ctx, x, y, st, A.RB_FOCUS, bins, device=device)
        assert 0 < r < 125


@pytest.mark.gpu
def test_gpu_reference_d80_known_answer(ctx):
    print("containment: testD80 disk, 1e7 points", flush=True)
    x, y = d80_points(10_000_000, seed=1)
    r, cx, cy = _device_vs_oracle(ctx, x, y, np.full(len(x), A.RB_FOCUS, np.int32), A.RB_FOCUS, D80_BINS)
    tol = 0.015
    assert abs(cx - 1.0) <= tol and abs(cy + 1.0) <= tol and abs(r / 2.0 - 0.8) / 0.8 <= tol


@pytest.mark.gpu
def test_gpu_containment_edge_cases(ctx):
    from robast_amd._lib import RbError, make_rays
    print("containment: empty, out of range, bad arguments", flush=True)
    bins = (10, -1.0, 1.0, 10, -1.0, 1.0)
    # empty batch: arrays may be null
    h = A.rb_hist2d_desc()
    h.nx, h.xlo, h.xhi, h.ny, h.ylo, h.yhi = bins
    h.status = A.RB_FOCUS
    assert ctx.containment_radius(A.rb_rays(), 0, h, 0.8) == (0.0, 0.0, 0.0)
    # nothing in range, or nothing with the status
    x = np.array([5.0, -5.0, np.nan, 0.5])
    y = np.array([0.0, 0.0, 0.0, 0.5])
    st = np.array([3, 3, 3, 1], np.int32)
    assert _device_vs_oracle(ctx, x, y, st, 3, bins) == (0.0, 0.0, 0.0)
    # one populated bin
    x, y = np.full(64, 0.25), np.full(64, -0.35)
    r, cx, cy = _device_vs_oracle(ctx, x, y, np.full(64, 3, np.int32), 3, bins)
    assert r == 0.0 and math.isclose(cx, 0.25, rel_tol=1e-14) and math.isclose(cy, -0.35, rel_tol=1e-14)
    soa = O.new_soa(4)
    for bad in (dict(nx=0), dict(xhi=-2.0), dict(fraction=0.0), dict(fraction=1.5)):
        hb = A.rb_hist2d_desc()
        hb.nx, hb.xlo, hb.xhi, hb.ny, hb.ylo, hb.yhi = bins
        frac = bad.pop("fraction", 0.8)
        for k, v in bad.items():
            setattr(hb, k, v)
        with pytest.raises(RbError):
            ctx.containment_radius(make_rays(soa), 4, hb, frac)


@pytest.mark.gpu
def test_gpu_geoutil_mirror(ctx):
    from robast_amd import geoutil
    from robast_amd._lib import make_rays
    print("containment: TH2D / ContainmentRadius mirror", flush=True)
    x, y = d80_points(200_000, seed=2)
    soa = O.new_soa(len(x))
    soa["x"][:], soa["y"][:] = x, y
    soa["status"][:] = A.RB_FOCUS
    h2 = geoutil.TH2D("", "", *D80_BINS)
    h2.FillRays(ctx, make_rays(soa), len(x))
    r, cx, cy = geoutil.ContainmentRadius(h2, 0.80)
    (orr, ocx, ocy), counts, stats = oracle_d80(x, y, soa["status"], A.RB_FOCUS, D80_BINS)
    assert h2.Integral() == counts.sum() == len(x)
    assert math.isclose(h2.GetMean(1), stats[1] / stats[0], rel_tol=1e-12)
    assert math.isclose(h2.GetStdDev(2), math.sqrt(stats[4] / stats[0] - (stats[3] / stats[0]) ** 2), rel_tol=1e-9)
    assert O.containment(h2.counts, h2.stats, *D80_BINS, 0.80) == (r, cx, cy)
    assert abs(r / 1.6 - 1) < 0.02
