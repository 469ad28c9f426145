"""TGeoPcon (any rmin, any phi range) and TGeoPgon with rmin > 0 or a phi
range, in scenes: the reference tutorial's hollow Pcon obscuration
(tutorials/AshraOptics.C:370-386, sections given in decreasing z), a
polygonal mirror over a 150 deg wedge and a hollow conical lens over a 270
deg wedge, none of them in a composite, so the scene is otherwise one that
the shallow kernel form would take: the upload must route it to the form
that holds the polycone code. The shapes' distances, containment and
normals are pinned in tests/test_oracle_geometry.py and
tests/test_oracle_pins.py; here the oracle traces the scene, and the GPU
matches it per ray with the bars of each ray's own conditioning."""
import numpy as np
import pytest

from helpers import conditioned_compare, copy_soa, sensitivity
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.manager import AOpticsManager
from robast_amd.materials import ARefractiveIndex

mm = G.mm

# tutorials/AshraOptics.C:370-379
HAUBE_Z = [0, 4, 4, 194, 194, 202, 202, 298, 613, 717, 717, 720, 720, 730]
HAUBE_RMAX = [320, 320, 315, 315, 315, 315, 291, 291, 119, 119, 119, 119, 93, 93]
HAUBE_RMIN = [310, 310, 310, 310, 287, 287, 287, 287, 115, 115, 0, 0, 0, 0]


def haube(offset=400.0):
    pc = G.TGeoPcon("pip_pcon1", 0, 360, len(HAUBE_Z))
    for i in range(len(HAUBE_Z)):
        pc.DefineSection(i, offset * mm - HAUBE_Z[i] * mm, HAUBE_RMIN[i] * mm, HAUBE_RMAX[i] * mm)
    return pc


def test_sections_in_decreasing_z_are_reordered():
    G._reset_registry()
    pc = haube()
    assert pc.z == sorted(pc.z)
    assert pc.z[0] == pytest.approx((400 - 730) * 0.1) and pc.rmax[0] == pytest.approx(9.3)
    assert pc.rmin[-1] == pytest.approx(31.0) and pc.rmax[-1] == pytest.approx(32.0)


def scene():
    G._reset_registry()
    mgr = AOpticsManager("manager", "polycones")
    world = G.AOpticalComponent("world", G.TGeoBBox("world", 200, 200, 200))
    mgr.SetTopVolume(world)
    world.AddNode(G.AObscuration("haube", haube()), 1)
    pg = G.TGeoPgon("wedgeMirror", 40, 150, 3, 2)
    pg.DefineSection(0, -5, 20, 28)
    pg.DefineSection(1, 4, 24, 35)
    world.AddNode(G.AMirror("mirror", pg), 1, G.TGeoTranslation(0, 0, 80))
    pc = G.TGeoPcon("lensPcon", -30, 270, 3)
    pc.DefineSection(0, -6, 5, 20)
    pc.DefineSection(1, 0, 8, 26)
    pc.DefineSection(2, 6, 4, 15)
    lens = G.ALens("lens", pc)
    lens.SetRefractiveIndex(ARefractiveIndex(1.5))
    world.AddNode(lens, 1, G.TGeoTranslation(0, 0, -90))
    focal = G.TGeoPgon("focalPgon", 0, 360, 8, 2)
    focal.DefineSection(0, -1, 10, 60)
    focal.DefineSection(1, 1, 10, 60)
    world.AddNode(G.AFocalSurface("focal", focal), 1, G.TGeoTranslation(0, 0, -150))
    mgr.CloseGeometry()
    return mgr


def rays(n, seed=3):
    rng = np.random.default_rng(seed)
    soa = O.new_soa(n)
    u = rng.normal(size=(n, 3))
    o = 190.0 * u / np.linalg.norm(u, axis=1)[:, None]
    target = rng.uniform(-60, 60, size=(n, 3))
    target[:, 2] = rng.uniform(-160, 100, n)
    d = target - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    soa["x"], soa["y"], soa["z"] = o[:, 0].copy(), o[:, 1].copy(), o[:, 2].copy()
    soa["dx"], soa["dy"], soa["dz"] = d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy()
    soa["lambda_"][:] = 400e-7
    soa["status"][:] = A.RB_RUN
    return soa


def test_polycone_scene_traces_in_oracle():
    mgr = scene()
    t = mgr.flat().shapes["type"]
    assert (t == A.RB_SHAPE_PCON).sum() == 2 and (t == A.RB_SHAPE_PGON).sum() == 2
    soa = rays(20000)
    soa["path_xyzt"] = np.zeros((len(soa["x"]), 12, 4))
    soa["path_node"] = np.full((len(soa["x"]), 12), -1, dtype=np.int32)
    O.trace(mgr.flat(), soa, mgr.trace_opts(), nthreads=8)
    c = np.bincount(soa["status"], minlength=6)
    assert c[A.RB_RUN] == 0 and c[A.RB_STOP] > 1000 and c[A.RB_FOCUS] > 100 and c[A.RB_EXIT] > 1000
    # every node is met by some ray: stopped on the obscuration, reflected on
    # the mirror, refracted into and out of the lens, focused
    met, counts = np.unique(soa["path_node"], return_counts=True)
    assert set(met) >= {1, 2, 3, 4} and counts.min() > 50


@pytest.mark.gpu
def test_polycone_scene_gpu_matches_oracle(ctx):
    from robast_amd._lib import make_rays
    print("polycone scene: Pcon obscuration, wedge Pgon mirror, reflex Pcon lens", flush=True)
    mgr = scene()
    soa = rays(50000)
    g, o = copy_soa(soa), copy_soa(soa)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(g), len(g["x"]), mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    gain, unstable = sensitivity(mgr.flat(), soa, mgr.trace_opts())
    mism, worst = conditioned_compare(g, o, gain, unstable, eps=1e-11)
    dp = np.sqrt((g["x"] - o["x"]) ** 2 + (g["y"] - o["y"]) ** 2 + (g["z"] - o["z"]) ** 2)
    print(f"  {mism} integer mismatches, worst |dpos| / bar {worst:.2e}, max |dpos| {dp.max():.2e} cm", flush=True)
    assert mism == 0
    assert dp[gain <= 1e3].max() <= 1e-7
