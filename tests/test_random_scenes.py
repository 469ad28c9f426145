"""Seeded random scenes: every supported shape (box with an offset origin,
tube with a bore, theta-limited sphere with and without a phi range,
spherical cap, paraboloid, Pgon and
Pcon -- full, hollow, over a wedge of at most 180 deg and over a wider one --,
aspheric disk, Winston cone (polygonal and the 2D trough), and union / subtraction / intersection
composites up to two levels) under random rotations, as mirrors, lenses,
obscurations, focal surfaces and plain components, hit by random rays with
Fresnel reflection on. Objects occupy separate cells of a 3 x 3 x 3 grid
(no overlapping nodes, whose resolution is ROOT-internal). A second set of
the same scenes puts border conditions on about half of the mirrors and
lenses: Gaussian roughness, Lambertian reflection, or a one- or two-film
coating (read from the default table or a coarse precalculated one), on
either side of the surface.

CPU: the scenes build, flatten and trace in the oracle with every ray
terminated. GPU: librbgpu reproduces the oracle per ray, with bars set by
each ray's own conditioning rather than by observed differences. The oracle
re-traces the rays from start points moved by 1e-10 cm; a ray's gain is its
largest terminal move over 1e-10 cm. Paths that bounce between curved lens
and mirror surfaces amplify a difference at each one (gains up to 1e11 over
40-80 points), flat-walled and short paths do not (gains of 1-100). The
test allows a rounding-level difference of 1e-11 cm at each recorded point,
amplified by the gain: |dpos| <= 1e-9 + 1e-11 * gain * npoints cm; rays
with gain <= 1e3 are also held to 1e-7 cm outright. Statuses, point counts
and last nodes must be identical except on rays that the 1e-10 cm move
itself changes."""
import math

import numpy as np
import pytest

from helpers import conditioned_compare, copy_soa, sensitivity
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.manager import AOpticsManager
from robast_amd.materials import ARefractiveIndex, ASchottFormula, ASellmeierFormula, TGraph

SEEDS = [11, 23, 37, 41, 59]
# the per-point rounding-level difference the parity bars allow, in cm
# (helpers.conditioned_compare: bar = 1e-9 + EPS * gain * npoints)
EPS = 1e-11
CELL = 40.0     # cm, grid pitch; every object fits in a 15 cm radius


def _shape(rng, k, tag):
    kinds = ["box", "tube", "sphere", "cap", "paraboloid", "pgon", "aspheric", "winston", "box_minus_tube",
             "sphere_and_box", "union_rot", "nested"]
    kind = kinds[k % len(kinds)]
    u = lambda a, b: float(rng.uniform(a, b))  # noqa: E731
    n = f"{tag}{k}"
    if kind == "box":
        return kind, G.TGeoBBox(n, u(4, 10), u(4, 10), u(3, 8), (u(-2, 2), u(-2, 2), u(-2, 2)))
    if kind == "tube":
        return kind, G.TGeoTube(n, u(1, 4), u(6, 11), u(3, 9))
    if kind == "sphere":
        if rng.uniform() < 0.4:  # a phi range, narrower or wider than 180 deg
            f1 = u(-180, 180)
            return kind, G.TGeoSphere(n, u(3, 6), u(8, 12), u(10, 60), u(100, 170), f1, f1 + u(60, 320))
        return kind, G.TGeoSphere(n, u(3, 6), u(8, 12), u(10, 60), u(100, 170))
    if kind == "cap":
        return kind, G.TGeoSphere(n, u(8, 10), u(11, 13), u(120, 150), 180)
    if kind == "paraboloid":
        return kind, G.TGeoParaboloid(n, u(0, 2), u(6, 11), u(3, 8))
    if kind == "pgon":
        # a full polygon, or (TGeoPgon / TGeoPcon) hollow, over a wedge of at
        # most 180 deg, or over a wider one
        variant = int(rng.integers(0, 7))
        pcon = variant >= 4
        dphi = (360.0, 360.0, u(60, 180), u(200, 330))[variant % 4 if not pcon else (variant - 3)]
        hollow = variant in (1, 4, 6)
        rin = (lambda a, b: u(a, b)) if hollow else (lambda a, b: 0.0)
        if pcon:
            pg = G.TGeoPcon(n, u(0, 360), dphi, 3)
        else:
            pg = G.TGeoPgon(n, u(0, 60), dphi, int(rng.integers(3, 9)), 3)
        pg.DefineSection(0, -u(4, 8), rin(1, 3), u(4, 8))
        pg.DefineSection(1, 0, rin(1, 4), u(6, 10))
        pg.DefineSection(2, u(4, 8), rin(0, 2), u(3, 7))
        return kind, pg
    if kind == "aspheric":
        a = G.AGeoAsphericDisk(n, -u(1, 3), 1.0 / u(30, 60), u(1, 3), 1.0 / u(40, 80), u(8, 12), u(0, 3))
        a.SetConicConstants(u(-1, 0), u(-1, 0))
        a.SetPolynomials(2, [u(-1e-4, 1e-4), u(-1e-7, 1e-7)], 1, [u(-1e-4, 1e-4)])
        return kind, a
    if kind == "winston":
        if rng.uniform() < 0.35:  # the trough, AGeoWinstonCone2D
            return kind, G.AGeoWinstonCone2D(n, u(4, 7), u(1.5, 3), u(3, 8))
        return kind, G.AGeoWinstonConePoly(n, u(4, 7), u(1.5, 3), int(rng.integers(4, 9)))
    if kind == "box_minus_tube":
        G.TGeoBBox(n + "b", u(5, 9), u(5, 9), u(5, 9))
        G.TGeoTube(n + "t", 0, u(2, 4), 12)
        G.TGeoTranslation(n + "tr", u(-2, 2), u(-2, 2), 0).RegisterYourself()
        return kind, G.TGeoCompositeShape(n, f"{n}b - {n}t:{n}tr")
    if kind == "sphere_and_box":
        G.TGeoSphere(n + "s", 0, u(7, 11))
        G.TGeoBBox(n + "b", u(4, 8), u(4, 8), u(4, 8))
        return kind, G.TGeoCompositeShape(n, f"{n}s * {n}b")
    if kind == "union_rot":
        G.TGeoBBox(n + "a", u(5, 8), u(2, 4), u(2, 4))
        G.TGeoBBox(n + "b", u(2, 4), u(2, 4), u(5, 9))
        G.TGeoRotation(n + "r", u(0, 90), u(0, 60), u(0, 90)).RegisterYourself()
        return kind, G.TGeoCompositeShape(n, f"{n}a + {n}b:{n}r")
    G.TGeoBBox(n + "a", u(6, 9), u(6, 9), u(6, 9))
    G.TGeoTube(n + "t", 0, u(2, 4), 11)
    G.TGeoSphere(n + "s", 0, u(8, 11))
    return kind, G.TGeoCompositeShape(n, f"({n}a - {n}t) * {n}s")


def _surface(rng, world, vol, k):
    """A random border condition between the world and vol (either way):
    Gaussian roughness, Lambertian reflection or a one- or two-film coating."""
    from robast_amd.multilayer import AMultilayer
    first, second = (world, vol) if rng.uniform() < 0.7 else (vol, world)
    cond = G.ABorderSurfaceCondition(first, second)
    pick = int(rng.integers(0, 3))
    if pick == 0:
        cond.SetGaussianRoughness(float(rng.uniform(0.2, 3.0)) * G.deg)
    elif pick == 1:
        cond.EnableLambertian(True)
    else:
        n_sub = float(rng.uniform(1.3, 1.8))
        ml = AMultilayer(ARefractiveIndex(1.0), ARefractiveIndex(n_sub, float(rng.choice([0.0, 2.0]))))
        for _ in range(int(rng.integers(1, 3))):
            ml.InsertLayer(ARefractiveIndex(float(rng.uniform(1.3, 2.4)), float(rng.choice([0.0, 0.05]))),
                           float(rng.uniform(40, 200)) * G.nm)
        if rng.uniform() < 0.5:   # a coarse table: most rays fall between its bin centres
            ml.PreCalculateCoherentTMM(25, 280e-7, 620e-7, 16, -3 * G.deg, 93 * G.deg)
        cond.SetMultilayer(ml)
        if rng.uniform() < 0.5:
            cond.SetGaussianRoughness(float(rng.uniform(0.2, 2.0)) * G.deg)


def random_scene(seed, surfaces=False):
    """World of 100 cm half size; up to 27 objects, one per grid cell. With
    surfaces, about half of the mirrors and lenses get a border condition."""
    G._reset_registry()
    rng = np.random.default_rng(seed)
    mgr = AOpticsManager("manager", f"random {seed}")
    world = G.AOpticalComponent("world", G.TGeoBBox("world", 100, 100, 100))
    mgr.SetTopVolume(world)
    graph = TGraph()
    graph.SetPoint(0, 300e-7, 0.7)
    graph.SetPoint(1, 600e-7, 0.95)
    k = 0
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            for iz in (-1, 0, 1):
                if rng.uniform() < 0.25:
                    continue
                kind, sh = _shape(rng, int(rng.integers(0, 12)), f"s{k}_")
                role = int(rng.integers(0, 6))
                name = f"v{k}"
                if role == 0 or role == 1:
                    vol = G.AMirror(name, sh)
                    if role == 1:
                        vol.SetReflectance(graph)
                elif role == 2:
                    vol = G.ALens(name, sh)
                    u_idx = rng.uniform()
                    if u_idx < 0.4:
                        vol.SetRefractiveIndex(ARefractiveIndex(float(rng.uniform(1.3, 1.8))))
                    elif u_idx < 0.7:
                        vol.SetRefractiveIndex(ASellmeierFormula(1.03961212, 0.231792344, 1.01046945,
                                                                 0.00600069867, 0.0200179144, 103.560653))
                    else:
                        # N-BK7 in Schott's form (src/ASchottFormula.cxx:43-55)
                        vol.SetRefractiveIndex(ASchottFormula(2.2718929, -1.0108077e-2, 1.0592509e-2,
                                                              2.0816965e-4, -7.6472538e-6, 4.9240991e-7))
                elif role == 3:
                    vol = G.AObscuration(name, sh)
                elif role == 4:
                    vol = G.AFocalSurface(name, sh)
                else:
                    vol = G.AOpticalComponent(name, sh)
                rot = G.TGeoRotation(f"r{k}", float(rng.uniform(0, 360)), float(rng.uniform(0, 180)),
                                     float(rng.uniform(0, 360)))
                mat = G.TGeoCombiTrans(ix * CELL, iy * CELL, iz * CELL, rot)
                world.AddNode(vol, 1, mat)
                if surfaces and role <= 2 and rng.uniform() < 0.6:
                    _surface(rng, world, vol, k)
                k += 1
    mgr.CloseGeometry()
    return mgr


def random_rays(seed, n):
    """Rays from a 90 cm sphere towards random points of the occupied region."""
    rng = np.random.default_rng(seed + 1000)
    soa = O.new_soa(n)
    u = rng.normal(size=(n, 3))
    o = 90.0 * u / np.linalg.norm(u, axis=1)[:, None]
    target = rng.uniform(-60, 60, size=(n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    soa["x"], soa["y"], soa["z"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
    soa["dx"], soa["dy"], soa["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    soa["lambda_"][:] = rng.uniform(300e-7, 600e-7, n)
    soa["status"][:] = A.RB_RUN
    return soa


@pytest.mark.parametrize("surfaces", [False, True])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_traces_in_oracle(seed, surfaces):
    mgr = random_scene(seed, surfaces)
    soa = random_rays(seed, 4000)
    O.trace(mgr.flat(), soa, mgr.trace_opts(), nthreads=8)
    c = np.bincount(soa["status"], minlength=6)
    assert c[A.RB_RUN] == 0
    assert c[A.RB_EXIT] < len(soa["x"])


@pytest.mark.gpu
@pytest.mark.parametrize("surfaces", [False, True])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_gpu_matches_oracle(ctx, seed, surfaces):
    from robast_amd._lib import make_rays
    print(f"random scene {seed}, border conditions {surfaces}", flush=True)
    mgr = random_scene(seed, surfaces)
    if surfaces:
        assert len(mgr.flat().borders) > 0
    soa = random_rays(seed, 20000)
    g = copy_soa(soa)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(g), len(g["x"]), mgr.trace_opts())
    o = copy_soa(soa)
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    gain, unstable = sensitivity(mgr.flat(), soa, mgr.trace_opts())
    mism, worst = conditioned_compare(g, o, gain, unstable, eps=EPS)
    dp = np.sqrt((g["x"] - o["x"]) ** 2 + (g["y"] - o["y"]) ** 2 + (g["z"] - o["z"]) ** 2)
    calm = gain <= 1e3
    print(f"  {mism} integer mismatches ({int(unstable.sum())} unstable rays), worst |dpos| / bar {worst:.2e}, "
          f"well-conditioned max |dpos| {dp[calm].max():.2e} cm", flush=True)
    assert dp[calm].max() <= 1e-7


@pytest.mark.gpu
def test_random_scene_sweep_gpu(ctx):
    """40 more scenes (seeds 100-119, plain and with border conditions), 2e4
    rays each, held to the same per-ray conditioning bars."""
    from robast_amd._lib import make_rays
    rays = mism = 0
    for seed in range(100, 120):
        for surfaces in (False, True):
            mgr = random_scene(seed, surfaces)
            soa = random_rays(seed, 20000)
            g, o = copy_soa(soa), copy_soa(soa)
            ctx.trace(ctx.upload(mgr.flat()), make_rays(g), len(g["x"]), mgr.trace_opts())
            O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
            gain, unstable = sensitivity(mgr.flat(), soa, mgr.trace_opts())
            m, _ = conditioned_compare(g, o, gain, unstable, eps=EPS)
            dp = np.sqrt((g["x"] - o["x"]) ** 2 + (g["y"] - o["y"]) ** 2 + (g["z"] - o["z"]) ** 2)
            assert dp[gain <= 1e3].max() <= 1e-7, (seed, surfaces)
            rays += len(g["x"])
            mism += m
        print(f"sweep seed {seed}: {mism} integer mismatches in {rays} rays so far", flush=True)
    assert mism <= 1e-5 * rays
