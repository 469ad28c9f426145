"""Device ray sources: the ARayShooter patterns (src/ARayShooter.cxx:122-459)
and the CORSIKA photon-bunch conversion of ACorsikaIACTFile::GetRayArray
(src/ACorsikaIACTFile.cxx:71-131).

CPU (pins the oracle): the grid and ring patterns against their closed forms
(counts, spacings, the n = 1 half step, the centre ray translated but not
rotated, a 90 degree rotation applied to positions and direction before the
translation); the random patterns against the distributions the reference
draws from (uniform disk, rectangle, sphere, cosine band of the spherical
cone, the cone's aim points) by Kolmogorov-Smirnov tests; the reference's
empty-array cases; the bunch conversion against the geometry it encodes (each
ray, followed to the observation level, arrives at the bunch's (x, y) at the
bunch's time), its photon count (the float loop `j < photons`), the telescope
and array filter, the order of the output and the wavelength draw. The host
ARayShooter mirror agrees with the oracle, and the library's host-side count
function with the oracle's.

GPU: every pattern and the bunch conversion on the device against the oracle
ray for ray (integers and the +-*/sqrt-only paths bit for bit; paths through
sin/cos/acos within 1e-12 relative), count-only and capacity behaviour, the
error cases, and a device-shot batch traced in place against the oracle."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from robast_amd import ARayShooter, TGeoRotation, TGeoTranslation
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd import sources as S

nm, cm, m = G.nm, G.cm, G.m
LAM = 400 * nm
ROT90Z = TGeoRotation(90, 0, 0)  # (x, y, z) -> (-y, x, z)
TR = TGeoTranslation(10.0, -20.0, 30.0)
KINDS = {
    "rectangle": (A.RB_SHOOT_RECTANGLE, 12.0, 7.0, 9, 5, 0),
    "circle": (A.RB_SHOOT_CIRCLE, 15.0, 0.0, 6, 5, 0),
    "random_circle": (A.RB_SHOOT_RANDOM_CIRCLE, 15.0, 0.0, 0, 0, 20000),
    "random_cone": (A.RB_SHOOT_RANDOM_CONE, 7.5 * m, 16 * m, 0, 0, 20000),
    "random_rectangle": (A.RB_SHOOT_RANDOM_RECTANGLE, 4.0, 9.0, 0, 0, 20000),
    "random_sphere": (A.RB_SHOOT_RANDOM_SPHERE, 0.0, 0.0, 0, 0, 20000),
    "random_spherical_cone": (A.RB_SHOOT_RANDOM_SPHERICAL_CONE, 12.0, 0.0, 0, 0, 20000),
}


def desc(name, rot=None, tr=None, v=None, seed=11, **over):
    kind, a, b, nx, ny, n = KINDS[name]
    g = S.shooter_desc(kind, LAM, a, b, nx, ny, rot, tr, v, seed)
    for k, val in over.items():
        setattr(g, k, val)
    return g, n


def shot(name, **kw):
    g, n = desc(name, **kw)
    return O.shoot(g, n), g


def ks_uniform(x, lo, hi):
    return stats.kstest((np.asarray(x) - lo) / (hi - lo), "uniform").pvalue


def assert_fresh(s):
    assert np.all(s["t"] == 0) and np.all(s["status"] == A.RB_RUN)
    assert np.all(s["npoints"] == 1) and np.all(s["last_node"] == -1)
    assert np.all(s["lambda_"] == LAM)


# ---------------------------------------------------------------- oracle (CPU)

def test_rectangle_closed_form():
    s, _ = shot("rectangle")
    assert len(s["x"]) == 9 * 5
    i, j = np.divmod(np.arange(45), 5)   # i over nx outer, j over ny inner (:426-431)
    assert np.allclose(s["x"], i * 12.0 / 8 - 6.0, rtol=0, atol=1e-15)
    assert np.allclose(s["y"], j * 7.0 / 4 - 3.5, rtol=0, atol=1e-15)
    assert s["x"][0] == -6.0 and s["x"][-1] == 6.0 and s["y"][-1] == 3.5
    assert np.all(s["z"] == 0) and np.all(s["dz"] == 1)
    assert_fresh(s)
    # one ray per axis: a half step (dx / 2) that is never taken: the ray sits at -dx / 2
    one, _ = shot("rectangle", nx=1, ny=1)
    assert (one["x"][0], one["y"][0]) == (-6.0, -3.5)


def test_circle_closed_form():
    s, g = shot("circle")
    nr, nphi, rmax = 6, 5, 15.0
    assert len(s["x"]) == 1 + nphi * nr * (nr + 1) // 2
    r = np.hypot(s["x"], s["y"])
    k = 1
    for i in range(nr):
        ring = slice(k, k + nphi * (i + 1))
        assert np.allclose(r[ring], rmax * (i + 1) / nr, rtol=1e-15)
        phi = np.unwrap(np.arctan2(s["y"][ring], s["x"][ring]))
        assert np.allclose(np.diff(phi), 2 * math.pi / (nphi * (i + 1)), rtol=1e-12)
        assert abs(phi[0]) < 1e-15
        k += nphi * (i + 1)
    assert (s["x"][0], s["y"][0], s["z"][0]) == (0.0, 0.0, 0.0)
    assert_fresh(s)


def test_rotation_then_translation():
    """rot turns positions and the v direction, tr moves positions only; the
    Circle's centre ray is translated without rotation (:150-157)."""
    v = (1.0, 2.0, 2.0)
    base, _ = shot("circle", v=v)
    s, _ = shot("circle", rot=ROT90Z, tr=TR, v=v)
    assert np.allclose(s["x"], -base["y"] + 10.0, rtol=0, atol=1e-13)
    assert np.allclose(s["y"], base["x"] - 20.0, rtol=0, atol=1e-13)
    assert np.allclose(s["z"], 30.0, rtol=0, atol=0)
    assert (s["x"][0], s["y"][0], s["z"][0]) == (10.0, -20.0, 30.0)
    assert np.allclose(np.stack([s["dx"], s["dy"], s["dz"]], 1), [-2 / 3, 1 / 3, 2 / 3], rtol=0, atol=1e-15)
    rect, _ = shot("rectangle", rot=ROT90Z, tr=TR)
    plain, _ = shot("rectangle")
    assert np.allclose(rect["x"], -plain["y"] + 10.0, atol=1e-13) and np.allclose(rect["y"], plain["x"] - 20.0,
                                                                                   atol=1e-13)


def test_random_circle_distribution():
    rot = TGeoRotation(0, 90, 0)  # the disk's normal turns from +z to rot . ez
    s, g = shot("random_circle", rot=rot, tr=TR)
    n = np.array(rot.rot)[:, 2]
    p = np.stack([s["x"], s["y"], s["z"]], 1) - np.array(TR.tr)
    assert np.abs(p @ n).max() < 1e-12                                  # in the rotated plane
    rr = np.sqrt((p * p).sum(1)) / 15.0
    assert rr.max() <= 1.0 + 1e-15
    assert ks_uniform(rr ** 2, 0, 1) > 1e-3                             # area-uniform
    assert ks_uniform(np.arctan2(p @ np.array(rot.rot)[:, 1], p @ np.array(rot.rot)[:, 0]), -math.pi, math.pi) > 1e-3
    assert np.allclose(np.stack([s["dx"], s["dy"], s["dz"]], 1), n, atol=1e-15)   # v = +z, rotated


def test_random_rectangle_distribution():
    s, _ = shot("random_rectangle", tr=TR)
    x, y = s["x"] - 10.0, s["y"] + 20.0
    assert x.min() >= -2.0 and x.max() <= 2.0 and y.min() >= -4.5 and y.max() <= 4.5
    assert ks_uniform(x, -2.0, 2.0) > 1e-3 and ks_uniform(y, -4.5, 4.5) > 1e-3
    assert np.all(s["z"] == 30.0)


def test_random_cone_aims_into_the_disk():
    s, _ = shot("random_cone", tr=TR)
    assert np.all(s["x"] == 10.0) and np.all(s["y"] == -20.0) and np.all(s["z"] == 30.0)
    d = np.stack([s["dx"], s["dy"], s["dz"]], 1)
    assert np.allclose((d * d).sum(1), 1, rtol=0, atol=2e-15)
    t = 16 * m / d[:, 2]                                  # reach the plane z = d above the start
    hx, hy = t * d[:, 0] / (7.5 * m), t * d[:, 1] / (7.5 * m)
    r2 = hx * hx + hy * hy
    assert r2.max() <= 1 + 1e-12
    assert ks_uniform(r2, 0, 1) > 1e-3


def test_random_sphere_isotropic():
    s, _ = shot("random_sphere", tr=TR, rot=ROT90Z, v=(1, 0, 0))  # rot and v are not used
    assert np.all(s["x"] == 10.0) and np.all(s["z"] == 30.0)
    d = np.stack([s["dx"], s["dy"], s["dz"]], 1)
    assert np.allclose((d * d).sum(1), 1, rtol=0, atol=2e-15)
    assert ks_uniform(d[:, 2], -1, 1) > 1e-3
    assert ks_uniform(np.arctan2(d[:, 1], d[:, 0]), -math.pi, math.pi) > 1e-3


def test_random_spherical_cone_band():
    rot = TGeoRotation(30, 40, 0)
    s, _ = shot("random_spherical_cone", rot=rot, tr=TR)
    axis = np.array(rot.rot)[:, 2]
    d = np.stack([s["dx"], s["dy"], s["dz"]], 1)
    c = d @ axis
    lo = math.cos(12 * math.pi / 180)
    assert c.min() >= lo - 1e-15 and c.max() <= 1 + 1e-15
    assert ks_uniform(c, lo, 1) > 1e-3               # uniform in solid angle inside the cone
    assert np.all(s["x"] == 10.0)


def test_zero_direction_is_left_zero():
    s, _ = shot("rectangle", v=(0.0, 0.0, 0.0))
    assert np.all(s["dx"] == 0) and np.all(s["dy"] == 0) and np.all(s["dz"] == 0)


def test_empty_patterns_follow_the_reference():
    cases = [("rectangle", dict(a=-1.0)), ("rectangle", dict(nx=0)), ("circle", dict(a=-1.0)),
             ("circle", dict(ny=0)), ("random_circle", dict(a=-1.0)), ("random_rectangle", dict(b=-1.0))]
    for name, over in cases:
        g, n = desc(name, **over)
        assert O.shooter_count(g, n) == 0
    g, _ = desc("random_rectangle")
    assert O.shooter_count(g, 0) == 0
    g, _ = desc("random_cone")
    assert O.shooter_count(g, 0) == 0 and O.shooter_count(g, 5) == 5


def test_library_count_matches_oracle():
    from robast_amd._lib import load
    L = load()
    for name in KINDS:
        for over in ({}, dict(a=-1.0), dict(nx=0), dict(ny=3), dict(b=-2.0)):
            g, n = desc(name, **over)
            for nn in (0, 1, n):
                assert L.rbgpu_shooter_count(g, nn) == O.shooter_count(g, nn), (name, over, nn)
    g, _ = desc("circle", nx=1 << 20, ny=1 << 20)
    assert L.rbgpu_shooter_count(g, 0) == -1
    g.kind = 99
    assert L.rbgpu_shooter_count(g, 0) == -1


def test_host_mirror_matches_oracle():
    cases = [(ARayShooter.Rectangle(LAM, 12.0, 7.0, 9, 5, ROT90Z, TR, (1, 2, 2)), "rectangle", 0),
             (ARayShooter.Square(LAM, 12.0, 9, ROT90Z, TR), "rectangle", 0),
             (ARayShooter.Circle(LAM, 15.0, 6, 5, TGeoRotation(30, 20, 10), TR, (0, 1, 0)), "circle", 1e-13)]
    for arr, name, tol in cases:
        if arr.n == 81:
            g, n = desc(name, rot=ROT90Z, tr=TR, b=12.0, ny=9)
        elif name == "rectangle":
            g, n = desc(name, rot=ROT90Z, tr=TR, v=(1, 2, 2))
        else:
            g, n = desc(name, rot=TGeoRotation(30, 20, 10), tr=TR, v=(0, 1, 0))
        s = O.shoot(g, n)
        for f in ("x", "y", "z", "dx", "dy", "dz"):
            assert np.abs(getattr(arr, f) - s[f]).max() <= tol, (name, f)


def test_host_random_patterns():
    a = ARayShooter.RandomSquare(LAM, 4.0, 20000, None, TR, seed=5)
    assert ks_uniform(a.x - 10.0, -2.0, 2.0) > 1e-3 and np.all(a.dz == 1)
    c = ARayShooter.RandomCircle(LAM, 3.0, 20000, seed=5)
    assert ks_uniform((c.x ** 2 + c.y ** 2) / 9.0, 0, 1) > 1e-3
    k = ARayShooter.RandomSphericalCone(LAM, 20000, 20.0, seed=5)
    assert ks_uniform(k.dz, math.cos(20 * math.pi / 180), 1) > 1e-3
    o = ARayShooter.RandomCone(LAM, 2.0, 10.0, 5000, seed=5)
    assert np.all(o.x == 0) and np.all(np.hypot(o.dx / o.dz, o.dy / o.dz) * 10.0 <= 2.0 + 1e-12)
    assert ARayShooter.RandomCircle(LAM, -1.0, 10).n == 0 and ARayShooter.Circle(LAM, 1.0, 0, 3).n == 0


# ---------------------------------------------------------------- CORSIKA bunches

def bunch_table(n, seed, lambda_zero=0.3, tels=(0, 1), arrays=(0, 2)):
    rng = np.random.default_rng(seed)
    cx = rng.uniform(-0.2, 0.2, n).astype(np.float32)
    cy = rng.uniform(-0.2, 0.2, n).astype(np.float32)
    lam = rng.uniform(300, 550, n).astype(np.float32)
    lam[rng.uniform(0, 1, n) < lambda_zero] = 0
    return S.CorsikaBunches(rng.choice(tels, n), rng.choice(arrays, n), rng.uniform(-800, 800, n),
                            rng.uniform(-800, 800, n), rng.uniform(5e5, 2e6, n), rng.uniform(0, 60, n), cx, cy, lam,
                            rng.uniform(0.05, 4.5, n))


def opts(**kw):
    base = dict(tel_no=1, array_no=2, z=3000.0, tel_z=120.0, refractive_index=1.0003, lambda_min=250.0,
                lambda_max=700.0, seed=0xC0, id_base=0)
    base.update(kw)
    return S.bunch_opts(**base)


def test_bunch_counts_follow_the_float_loop():
    photons = np.array([0.0, -1.0, 0.3, 1.0, 2.5, np.nan, 7.0, 1e-30, 3.0000002], np.float32)
    n = len(photons)
    b = S.CorsikaBunches(np.ones(n), np.full(n, 2), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n),
                         np.zeros(n), np.zeros(n), np.full(n, 400.0), photons)
    s = O.bunches_to_rays(b, opts())
    want = [0, 0, 1, 1, 3, 0, 7, 1, 4]
    assert len(s["x"]) == sum(want)
    assert np.all(s["status"] == A.RB_RUN) and np.all(s["npoints"] == 1)


def test_bunch_geometry_and_time():
    """Followed along its direction to the observation level, each ray meets
    the bunch's (x, y) at the bunch's arrival time."""
    b = bunch_table(3000, seed=1)
    o = opts()
    s = O.bunches_to_rays(b, o)
    sel = np.nonzero((b.tel_no == 1) & (b.array_no == 2))[0]
    reps = np.ceil(b.photons[sel]).astype(int)
    src = np.repeat(sel, reps)                                    # bunch order, as the reference's loops
    assert len(s["x"]) == len(src) > 1000
    d = np.stack([s["dx"], s["dy"], s["dz"]], 1)
    assert np.allclose((d * d).sum(1), 1, rtol=0, atol=2e-15)
    step = (o.tel_z - o.z) / d[:, 2]                              # path length down to z = tel_z
    assert np.all(step > 0)
    assert np.allclose(s["x"] + step * d[:, 0], b.x[src], rtol=0, atol=1e-9)
    assert np.allclose(s["y"] + step * d[:, 1], b.y[src], rtol=0, atol=1e-9)
    # the path is timed as if (cx, cy, cz) were unit; in float it is not quite
    c = np.stack([b.cx[src], b.cy[src], b.cz[src]], 1).astype(float)
    arrival = s["t"] + step / np.linalg.norm(c, axis=1) / (2.99792458e10 / o.refractive_index)
    assert np.allclose(arrival, b.time[src].astype(float) * 1e-9, rtol=1e-12, atol=0)
    assert np.all(s["z"] == o.z)
    # the direction is the bunch's, normalised
    assert np.allclose(d, c / np.linalg.norm(c, axis=1)[:, None], rtol=0, atol=1e-15)


def test_bunch_wavelengths():
    b = bunch_table(20000, seed=2, lambda_zero=0.5)
    o = opts(lambda_min=300.0, lambda_max=600.0)
    s = O.bunches_to_rays(b, o)
    sel = np.nonzero((b.tel_no == 1) & (b.array_no == 2))[0]
    src = np.repeat(sel, np.ceil(b.photons[sel]).astype(int))
    lam = s["lambda_"] / nm
    fixed = b.lambda_[src] != 0
    assert np.array_equal(s["lambda_"][fixed], b.lambda_[src][fixed].astype(float) * (1e-9 * 100.0))
    drawn = lam[~fixed]
    assert len(drawn) > 5000
    assert drawn.min() >= 300.0 * (1 - 1e-12) and drawn.max() <= 600.0 * (1 + 1e-12)
    assert ks_uniform(1 / drawn, 1 / 600.0, 1 / 300.0) > 1e-3    # 1/lambda uniform (a 1/lambda^2 spectrum)
    # the reference's upper bound is fMaxPhotonBunches (100000): most draws stay near lambda_min
    wide = O.bunches_to_rays(b, opts(lambda_min=300.0, lambda_max=100000.0))["lambda_"][~fixed] / nm
    assert wide.max() > 5000 and np.median(wide) < 700


def test_bunch_filters_and_limits():
    b = bunch_table(500, seed=3)
    for tel, arr in ((0, 0), (1, 0), (0, 2), (5, 2)):
        s = O.bunches_to_rays(b, opts(tel_no=tel, array_no=arr))
        sel = (b.tel_no == tel) & (b.array_no == arr)
        assert len(s["x"]) == int(np.ceil(b.photons[sel]).sum())
    big = S.CorsikaBunches([1, 0], [2, 2], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [400, 400],
                           [1.0, 3e7])
    assert len(O.bunches_to_rays(big, opts())["x"]) == 1          # the 3e7 bunch is another telescope's
    with pytest.raises(ValueError):
        O.bunches_to_rays(big, opts(tel_no=0))


# ---------------------------------------------------------------- device (GPU)

def _device_soa(dr):
    return {f: dr.t[f][:dr.n].cpu().numpy() for f in dr.t}


def _match(g, o, exact):
    assert len(g["x"]) == len(o["x"])
    for f in ("status", "npoints", "last_node"):
        assert np.array_equal(g[f], o[f]), f
    for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_"):
        if exact:
            assert np.array_equal(g[f], o[f]), f
        else:
            scale = max(np.abs(o[f]).max(initial=0), 1e-300)
            assert np.abs(g[f] - o[f]).max(initial=0) <= 1e-12 * scale, f


EXACT = {"rectangle", "random_rectangle"}   # no transcendental on their path


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(KINDS))
def test_gpu_shooter_matches_oracle(ctx, name):
    print(f"shooter {name}", flush=True)
    for kw in (dict(), dict(rot=TGeoRotation(30, 20, 10), tr=TR, v=(0.2, -0.4, 1.0))):
        g, n = desc(name, **kw)
        dr = S.shoot(ctx, g, n)
        _match(_device_soa(dr), O.shoot(g, n), name in EXACT)


@pytest.mark.gpu
def test_gpu_shooter_large_and_offset(ctx):
    print("shooter: 4e6-ray random circle with id_base, 2000-ring circle", flush=True)
    g, _ = desc("random_circle", tr=TR, id_base=1 << 33)
    dr = S.shoot(ctx, g, 4_000_000)
    sub = np.arange(0, 4_000_000, 997)
    gs = _device_soa(dr)
    g2, _ = desc("random_circle", tr=TR, id_base=(1 << 33) + 0)
    o = O.shoot(g2, 4_000_000)
    for f in ("x", "y", "z"):
        assert np.abs(gs[f][sub] - o[f][sub]).max() <= 1e-12 * 30
    g, _ = desc("circle", nx=2000, ny=3)            # 6,003,001 rays: ring lookup far from the start
    dr = S.shoot(ctx, g, 0)
    assert dr.n == 1 + 3 * 2000 * 2001 // 2
    gs, o = _device_soa(dr), O.shoot(g, 0)
    assert np.abs(gs["x"] - o["x"]).max() <= 1e-12 * 15 and np.abs(gs["y"] - o["y"]).max() <= 1e-12 * 15


@pytest.mark.gpu
def test_gpu_shooter_errors_and_empty(ctx):
    from robast_amd._lib import RbError
    print("shooter: errors and empty patterns", flush=True)
    g, _ = desc("rectangle")
    dr = S.DeviceRays(ctx, 45)
    with pytest.raises(RbError, match="pattern has 45"):
        ctx.shoot(g, dr.rays(), 44)
    g.kind = 42
    with pytest.raises(RbError, match="bad kind"):
        ctx.shoot(g, dr.rays(), 0)
    g, _ = desc("circle", a=-1.0)
    assert S.shoot(ctx, g, 0).n == 0
    host = O.new_soa(45)
    from robast_amd._lib import make_rays
    g, _ = desc("rectangle")
    with pytest.raises(RbError, match="not GPU memory"):
        ctx.shoot(g, make_rays(host), 45)


@pytest.mark.gpu
@pytest.mark.parametrize("device_ptrs", [False, True])
def test_gpu_bunches_match_oracle(ctx, device_ptrs):
    import torch
    print(f"bunches -> rays, device_ptrs={device_ptrs}", flush=True)
    b = bunch_table(200_000, seed=4)
    o = opts(device_ptrs=device_ptrs, id_base=77)
    if device_ptrs:
        keep = {f: torch.from_numpy(getattr(b, f)).cuda() for f in A.BUNCH_FIELDS_I + A.BUNCH_FIELDS_F}
        d = A.rb_bunches()
        for f, t in keep.items():
            setattr(d, f, t.data_ptr())
    else:
        d = b.desc()
    n = ctx.bunches_to_rays(d, b.n, o)
    ref = O.bunches_to_rays(b, o)
    assert n == len(ref["x"]) > 100_000
    dr = S.DeviceRays(ctx, n)
    assert ctx.bunches_to_rays(d, b.n, o, dr.rays(), n) == n
    _match(_device_soa(dr), ref, exact=True)


@pytest.mark.gpu
def test_gpu_bunch_edge_cases(ctx):
    from robast_amd._lib import RbError
    print("bunches: capacity, limits, empty", flush=True)
    b = bunch_table(1000, seed=5)
    o = opts()
    n = ctx.bunches_to_rays(b.desc(), b.n, o)
    dr = S.DeviceRays(ctx, n)
    with pytest.raises(RbError, match="exceed the capacity"):
        ctx.bunches_to_rays(b.desc(), b.n, o, dr.rays(), n - 1)
    big = S.CorsikaBunches([1, 0], [2, 2], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [0, 0], [400, 400], [1.0, 3e7])
    assert ctx.bunches_to_rays(big.desc(), big.n, o) == 1
    with pytest.raises(RbError, match="2\\^24"):
        ctx.bunches_to_rays(big.desc(), big.n, opts(tel_no=0))
    assert ctx.bunches_to_rays(A.rb_bunches(), 0, o) == 0
    assert ctx.bunches_to_rays(b.desc(), b.n, opts(tel_no=9)) == 0
    with pytest.raises(RbError, match="positive"):
        ctx.bunches_to_rays(b.desc(), b.n, opts(refractive_index=0.0))
    # a sparse selection: tiles of output rays span more bunches than LDS stages
    sparse = bunch_table(300_000, seed=6, tels=tuple(range(100)), arrays=(2,))
    o = opts()
    n = ctx.bunches_to_rays(sparse.desc(), sparse.n, o)
    ref = O.bunches_to_rays(sparse, o)
    assert n == len(ref["x"]) > 5000
    dr = S.DeviceRays(ctx, n)
    ctx.bunches_to_rays(sparse.desc(), sparse.n, o, dr.rays(), n)
    _match(_device_soa(dr), ref, exact=True)
    # one bunch spanning many tiles
    one = S.CorsikaBunches([1, 1, 1], [2, 2, 2], [1, 2, 3], [0, 0, 0], [0, 0, 0], [5, 6, 7], [0, 0.1, 0], [0, 0, 0.1],
                           [0, 400, 0], [3.5, 5000.2, 700])
    n = ctx.bunches_to_rays(one.desc(), one.n, o)
    assert n == 4 + 5001 + 700
    dr = S.DeviceRays(ctx, n)
    ctx.bunches_to_rays(one.desc(), one.n, o, dr.rays(), n)
    _match(_device_soa(dr), O.bunches_to_rays(one, o), exact=True)
    rays = S.GetRayArray(ctx, b, 1, 2, 3000.0, 1.0003, 120.0, 250.0, seed=0xC0)
    ref = O.bunches_to_rays(b, opts(lambda_max=100000.0))
    _match(_device_soa(rays), ref, exact=True)
    assert S.GetRayArray(ctx, b, -1, 2, 3000.0, 1.0, 120.0, 250.0) is None


@pytest.mark.gpu
def test_gpu_shot_rays_traced_in_place(ctx):
    """A device-shot batch traced where it lies equals the oracle's trace of
    the oracle's rays (HESS1.C's source: RandomCircle toward a paraboloid)."""
    from helpers import compare
    from robast_amd import configs
    print("device-shot RandomCircle traced in place (C1 telescope)", flush=True)
    mgr = configs.c1_scene()
    v = (math.sin(0.5 * G.deg), 0.0, -math.cos(0.5 * G.deg))
    tr = TGeoTranslation(0.0, 0.0, 2 * configs.C1_F)
    rays = ARayShooter.RandomCircle(LAM, 140.0, 300_000, None, tr, v, seed=9, ctx=mgr.context())
    g = S.shooter_desc(A.RB_SHOOT_RANDOM_CIRCLE, LAM, 140.0, 0.0, 0, 0, None, tr, v, 9)
    ora = O.shoot(g, 300_000)
    mgr.TraceNonSequential(rays)
    O.trace(mgr.flat(), ora, mgr.trace_opts(), nthreads=8)
    compare(_device_soa(rays), ora)
    assert (ora["status"] == A.RB_FOCUS).sum() > 200_000
    arr = rays.to_array()
    assert arr.GetFocused().GetEntries() == (ora["status"] == A.RB_FOCUS).sum()
