"""Pins for the oracle functions that the marching and closed-form tests of
test_oracle_geometry.py leave open:

- AGeoWinstonConePoly: the profile r(z) against the compound-parabola
  definition (focus at the opposite exit edge, axis tilted by theta) solved
  in mpmath, containment against the apothem polygon test, and the exit
  distance against a 40-digit scan plus bisection of that exact profile;
- the normal of every shape against the tangent plane found by bisecting
  the containment test along four neighbouring parallel rays (central
  differences), so a sign, sector-index or transposed-operand error in any
  *_normal fails.
"""
import math
import zlib

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O
from robast_amd import geometry as G
from test_oracle_geometry import SHAPES, _asph_shape, _march, _shape_scene


def _winston(r1=2.5, r2=1.0, n=6):
    G._reset_registry()
    return G.AGeoWinstonConePoly("wc", r1, r2, n)


# ---------------------------------------------------------------- Winston cone


def _cpc_profile(w, dps=40):
    """r(z) of the compound parabolic concentrator, solved in mpmath from the
    parabola's definition: |P - F| = (P - F).a + 2f with the focus F at the
    opposite exit edge (-r2, -dz), axis a = (-sin theta, cos theta) and focal
    length f = r2 (1 + sin theta), theta = asin(r2 / r1). Nothing here uses
    the reference's quadratic in the tilted frame (AGeoWinstonCone2D::CalcR)."""
    mp.mp.dps = dps
    r1, r2 = mp.mpf(w.r1), mp.mpf(w.r2)
    th = mp.asin(r2 / r1)
    f = r2 * (1 + mp.sin(th))
    dz = (r1 + r2) / mp.tan(th) / 2
    ax, az = -mp.sin(th), mp.cos(th)

    def g(x, z):
        px, pz = x + r2, z + dz  # P - F
        return mp.sqrt(px * px + pz * pz) - (px * ax + pz * az) - 2 * f

    def r_of(z):
        z = mp.mpf(z)
        # g decreases through zero between the exit radius and beyond the entrance
        return mp.findroot(lambda x: g(x, z), (r2 - mp.mpf("0.5") * r2, r1 + r1), solver="anderson")

    return r_of, dz, th, f


def test_winston_dimensions_are_the_cpc_edge_rays():
    """The parabola through both edges: r(-dz) = r2 and r(+dz) = r1, with the
    flattened theta, dz and f the CPC's (sin theta = r2/r1)."""
    for r1, r2 in ((2.5, 1.0), (1.25, 0.5), (3.0, 2.9)):
        w = _winston(r1, r2)
        r_of, dz, th, f = _cpc_profile(w)
        assert abs(float(r_of(-dz)) - r2) < 1e-13 * r1
        assert abs(float(r_of(dz)) - r1) < 1e-13 * r1
        assert w.theta == pytest.approx(float(th), rel=1e-15)
        assert w.dz == pytest.approx(float(dz), rel=1e-14)
        assert w.f == pytest.approx(float(f), rel=1e-15)


def test_winston_contains_matches_cpc_polygon():
    """Containment: |z| <= dz and the largest apothem projection
    max_k (x cos phi_k + y sin phi_k), phi_k = 2 pi k / n, within r(z). The
    oracle folds the azimuth into one sector with atan2 instead."""
    for n_sides in (6, 4, 3, 8):
        w = _winston(2.5, 1.0, n_sides)
        flat, sid = _shape_scene(w)
        r_of, _, _, _ = _cpc_profile(w, dps=20)
        zs = np.linspace(-w.dz, w.dz, 161)
        rz = np.array([float(r_of(z)) for z in zs])
        rng = np.random.default_rng(n_sides)
        q = np.column_stack([rng.uniform(-3.2, 3.2, 40000), rng.uniform(-3.2, 3.2, 40000),
                             rng.uniform(-1.1 * w.dz, 1.1 * w.dz, 40000)])
        phi = 2 * np.pi * np.arange(n_sides) / n_sides
        apo = (q[:, :1] * np.cos(phi) + q[:, 1:2] * np.sin(phi)).max(axis=1)
        # piecewise-linear profile between 161 exact samples: exact to ~1e-4,
        # so points within 1e-3 of the wall are left out
        r_lin = np.interp(q[:, 2], zs, rz)
        clear = (np.abs(apo - r_lin) > 1e-3) & (np.abs(np.abs(q[:, 2]) - w.dz) > 1e-9)
        want = (np.abs(q[:, 2]) <= w.dz) & (apo <= r_lin)
        got = O.contains_batch(flat, sid, q)
        assert clear.sum() > 39000
        np.testing.assert_array_equal(got[clear], want[clear])


def test_winston2d_contains_matches_cpc_trough():
    """AGeoWinstonCone2D (src/AGeoWinstonCone2D.cxx:175-190): |y| <= dy,
    |z| <= dz and |x| within the exact CPC profile r(z)."""
    G._reset_registry()
    w = G.AGeoWinstonCone2D("w2", 2.5, 1.0, 3.0)
    flat, sid = _shape_scene(w)
    r_of, dz, _, _ = _cpc_profile(w, dps=20)
    assert w.dz == pytest.approx(float(dz), rel=1e-14)
    zs = np.linspace(-w.dz, w.dz, 161)
    rz = np.array([float(r_of(z)) for z in zs])
    rng = np.random.default_rng(21)
    q = np.column_stack([rng.uniform(-3, 3, 40000), rng.uniform(-3.5, 3.5, 40000),
                         rng.uniform(-1.1 * w.dz, 1.1 * w.dz, 40000)])
    r_lin = np.interp(q[:, 2], zs, rz)
    clear = (np.abs(np.abs(q[:, 0]) - r_lin) > 1e-3) & (np.abs(np.abs(q[:, 2]) - w.dz) > 1e-9) & \
        (np.abs(np.abs(q[:, 1]) - 3.0) > 1e-9)
    want = (np.abs(q[:, 2]) <= w.dz) & (np.abs(q[:, 0]) <= r_lin) & (np.abs(q[:, 1]) <= 3.0)
    got = O.contains_batch(flat, sid, q)
    assert clear.sum() > 39000 and want[clear].sum() > 5000
    np.testing.assert_array_equal(got[clear], want[clear])


def _exact_exit_distance(w, r_of, p, d, tmax):
    """Distance from an inside point to the first wall or end cap, in mpmath:
    a 2000-step scan of h(t) = max_k apothem_k(p + t d) - r(z) and of the
    caps, then bisection of h to 1e-30."""
    mp.mp.dps = 40
    phi = [2 * mp.pi * k / w.n for k in range(w.n)]
    cs = [(mp.cos(a), mp.sin(a)) for a in phi]
    P = [mp.mpf(float(v)) for v in p]
    D = [mp.mpf(float(v)) for v in d]
    dz = mp.mpf(w.dz)

    def h(t):
        x, y, z = (P[k] + t * D[k] for k in range(3))
        if abs(z) > dz:
            return mp.mpf(1)
        return max(x * c + y * s for c, s in cs) - r_of(z)

    t_cap = mp.inf
    if D[2] > 0:
        t_cap = (dz - P[2]) / D[2]
    elif D[2] < 0:
        t_cap = (-dz - P[2]) / D[2]
    steps = 2000
    lo = mp.mpf(0)
    for i in range(1, steps + 1):
        t = mp.mpf(tmax) * i / steps
        if t >= t_cap:
            break
        if h(t) > 0:
            hi = t
            for _ in range(110):
                mid = (lo + hi) / 2
                if h(mid) > 0:
                    hi = mid
                else:
                    lo = mid
            return (lo + hi) / 2
        lo = t
    return t_cap


def test_winston_exit_distance_matches_mpmath_scan():
    """wins_dist_in (AGeoWinstonConePoly::DistFromInside over
    AGeoWinstonCone2D::DistToParabola) against the exact CPC wall and caps,
    for rays from inside, steep and shallow, including light-guide rays that
    run from the entrance towards the exit."""
    w = _winston(2.5, 1.0, 6)
    flat, sid = _shape_scene(w)
    r_of, _, _, _ = _cpc_profile(w, dps=40)
    rng = np.random.default_rng(77)
    checked = walls = 0
    while checked < 60:
        z0 = rng.uniform(-0.95 * w.dz, 0.95 * w.dz)
        rr = float(r_of(z0))
        ang = rng.uniform(0, 2 * np.pi)
        rad = rng.uniform(0, 0.95) * rr
        p = np.array([rad * math.cos(ang), rad * math.sin(ang), z0])
        d = rng.normal(size=3)
        if checked % 2:
            d[2] = -abs(d[2]) - 1.0  # towards the exit, as light in a guide
        d /= np.linalg.norm(d)
        inside, _, din, _ = O.shape_query(flat, sid, p, d)
        if not inside:
            continue
        t = _exact_exit_distance(w, r_of, p, d, tmax=4 * w.dz)
        z_end = p[2] + float(t) * d[2]
        walls += abs(abs(z_end) - w.dz) > 1e-9
        assert abs(din - float(t)) < 1e-9, (p, d, din, float(t))
        checked += 1
    assert walls > 30


def test_winston_marching_from_outside_and_inside():
    """Distances from anywhere, by marching the containment test (which the
    CPC test above pins), as for the other shapes."""
    w = _winston(2.5, 1.0, 6)
    flat, sid = _shape_scene(w)
    rng = np.random.default_rng(zlib.crc32(b"winston"))
    n_in = 0
    for _ in range(300):
        p = rng.uniform(-4, 4, 3)
        p[2] = rng.uniform(-1.2 * w.dz, 1.2 * w.dz)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        inside, dout, din, _ = O.shape_query(flat, sid, p, d)
        t = _march(flat, sid, p, d, inside, tmax=30.0, n=6000)
        if inside:
            n_in += 1
            assert t is not None and din == pytest.approx(t, abs=2e-7), (p, d, din, t)
        elif t is not None and abs(dout - t) <= 2e-7:
            pass
        elif dout < 1e29 and (t is None or dout < t):
            assert O.contains_batch(flat, sid, p + (dout + 1e-7) * d)[0], (p, d, dout, t)
        else:
            assert t is None and dout >= 1e29, (p, d, dout, t)
    assert n_in > 40


# ---------------------------------------------------------------- normals


def _crossing(flat, sid, p, d, t0, inside_start):
    """Bisection of the containment flip along p + t d near t0."""
    lo, hi = t0 - 1e-2, t0 + 1e-2
    c = O.contains_batch(flat, sid, np.array([p + lo * d, p + hi * d]))
    if c[0] != inside_start or c[1] == inside_start:
        return None
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if O.contains_batch(flat, sid, p + mid * d)[0] == inside_start:
            lo = mid
        else:
            hi = mid
    return p + 0.5 * (lo + hi) * d


def _fd_normal(flat, sid, p, d, t, inside, h=1e-3):
    """Unit normal of the surface the ray p + t d crosses, from the crossings
    of four parallel rays offset by +-h along two directions perpendicular to
    d (central differences); None where the neighbourhood is not one smooth
    surface (an edge, a seam)."""
    e1 = np.cross(d, [1.0, 0, 0] if abs(d[0]) < 0.9 else [0, 1.0, 0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(d, e1)
    c0 = _crossing(flat, sid, p, d, t, inside)
    if c0 is None:
        return None
    pts = []
    for e in (e1, -e1, e2, -e2):
        c = _crossing(flat, sid, p + h * e, d, t, inside)
        if c is None:
            return None
        pts.append(c)
    # a kink between the samples shows as a large second difference
    for a, b in ((pts[0], pts[1]), (pts[2], pts[3])):
        if np.linalg.norm(a + b - 2 * c0) > 2e-5:
            return None
    n = np.cross(pts[0] - pts[1], pts[2] - pts[3])
    return n / np.linalg.norm(n)


_NORMAL_SHAPES = dict(SHAPES)
_NORMAL_SHAPES["winston"] = _winston(2.5, 1.0, 6)
_NORMAL_SHAPES["aspheric"] = _asph_shape()


@pytest.mark.parametrize("name", sorted(_NORMAL_SHAPES))
def test_normals_match_surface_tangent_plane(name):
    """The oracle's normal where a ray enters the shape (from outside) or
    leaves it (from inside), oriented along the ray (d.n > 0, as
    FindNormal's caller orients it), against the tangent plane of the
    containment boundary."""
    shape = _NORMAL_SHAPES[name]
    flat, sid = _shape_scene(shape)
    rng = np.random.default_rng(zlib.crc32(("n" + name).encode()))
    span = 14.0 if name not in ("winston",) else 4.0
    checked = tries = 0
    while checked < 60 and tries < 3000:
        tries += 1
        p = rng.uniform(-span, span, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        inside, dout, din, _ = O.shape_query(flat, sid, p, d)
        t = din if inside else dout
        n = O.shape_query(flat, sid, p, d, exit_normal=inside)[3]
        if not (1e-3 < t < 1e29):
            continue
        fd = _fd_normal(flat, sid, p, d, t, inside)
        if fd is None:
            continue
        if np.dot(fd, d) < 0:
            fd = -fd
        assert np.dot(n, d) >= 0, (name, p, d, n)
        assert np.linalg.norm(n - fd) < 2e-5, (name, p, d, n, fd)
        checked += 1
    assert checked >= 40, (name, checked, tries)
