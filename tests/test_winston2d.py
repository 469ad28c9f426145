"""AGeoWinstonCone2D in the reference tutorial's hexagonal light guide,
tutorials/HexWinstonCone.C:29-75 mode 1: the intersection of three troughs
rotated by 0, 60 and 120 deg ("coneV*(coneV:rot60)*(coneV:rot120)", two
boolean levels), cut out of a hexagonal Pgon ("pgon:rot30 - coneComp1",
three levels), as a mirror over a hexagonal focal surface. Mode 0 (the
polygonal cone, "pgon:rot30 - hexV") is traced beside it. The shape itself
is pinned in tests/test_oracle_geometry.py and tests/test_oracle_pins.py.

CPU: every ray that enters the entrance aperture within the cutoff angle
reaches the exit in both modes; and since three troughs at 60 deg bound the
hexagon of the polygonal cone, modes 0 and 1 are one solid reached through
different shape code: the oracle traces both to the same fate and points.
GPU: per-ray parity with the oracle, with each ray's own conditioning
bars."""
import math

import numpy as np
import pytest

from helpers import conditioned_compare, copy_soa, sensitivity
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.manager import AOpticsManager

mm = G.mm
K_RIN, K_ROUT = 20 * mm, 10 * mm


def hex_winston_scene(mode):
    """tutorials/HexWinstonCone.C:29-75 (mode 0: polygonal cone; mode 1: the
    intersection of three 2D cones)."""
    G._reset_registry()
    mgr = AOpticsManager("manager", "HexWinstonCone")
    world = G.AOpticalComponent("world", G.TGeoBBox("worldbox", 30 * G.m, 30 * G.m, 30 * G.m))
    mgr.SetTopVolume(world)
    G.TGeoRotation("rot30", 30, 0, 0).RegisterYourself()
    G.TGeoRotation("rot60", 60, 0, 0).RegisterYourself()
    G.TGeoRotation("rot120", 120, 0, 0).RegisterYourself()
    cone_v = G.AGeoWinstonCone2D("coneV", K_RIN, K_ROUT, K_RIN * 1.733)
    dz = cone_v.GetDZ()
    pgon = G.TGeoPgon("pgon", 0, 360, 6, 4)
    pgon.DefineSection(0, -dz * 0.999, 0, K_ROUT * 1.1)
    pgon.DefineSection(1, -dz * 0.5, 0, K_RIN * 0.9)
    pgon.DefineSection(2, -dz * 0.0, 0, K_RIN * 0.99)
    pgon.DefineSection(3, dz * 0.999, 0, K_RIN * 1.001)
    if mode == 0:
        G.AGeoWinstonConePoly("hexV", K_RIN, K_ROUT, 6)
        comp = G.TGeoCompositeShape("coneComp1", "pgon:rot30 - hexV")
    else:
        G.TGeoCompositeShape("coneComp1", "coneV*(coneV:rot60)*(coneV:rot120)")
        comp = G.TGeoCompositeShape("coneComp2", "pgon:rot30 - coneComp1")
    world.AddNode(G.AMirror("coneMirror", comp), 1)
    pmt = G.TGeoPgon("pgonPMT", 0, 360, 6, 2)
    pmt.DefineSection(0, -dz - 0.01 * mm, 0, K_ROUT * 1.01)
    pmt.DefineSection(1, -dz, 0, K_ROUT * 1.01)
    world.AddNode(G.AFocalSurface("pmt", pmt), 1, G.TGeoRotation("r30", 30, 0, 0))
    mgr.CloseGeometry()
    mgr.cone_dz = dz
    return mgr


def beams(dz, n, theta_deg, seed=1):
    """Parallel rays over the entrance aperture (a disk of radius 0.95 r1
    just above it) at theta_deg off the axis in random azimuths, as the
    tutorial's effective-area scan sends them."""
    rng = np.random.default_rng(seed)
    soa = O.new_soa(n)
    r = 0.95 * K_RIN * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    th = np.radians(theta_deg)
    ph = rng.uniform(0, 2 * np.pi, n)
    d = np.column_stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.full(n, np.cos(th))])
    start = np.column_stack([r * np.cos(a), r * np.sin(a), np.full(n, dz + 0.5 * mm)])
    for k, c in enumerate("xyz"):
        soa[c][:] = start[:, k]
        soa["d" + c][:] = d[:, k]
    soa["lambda_"][:] = 400e-7
    soa["status"][:] = A.RB_RUN
    return soa


def test_composite_depth_three_builds():
    mgr = hex_winston_scene(1)
    flat = mgr.flat()
    assert (flat.shapes["type"] == A.RB_SHAPE_WINSTON2D).sum() == 1
    comps = flat.shapes["type"] == A.RB_SHAPE_COMPOSITE
    assert comps.sum() == 3


@pytest.mark.parametrize("mode", [0, 1])
def test_on_axis_rays_reach_the_exit(mode):
    """Within the cutoff angle asin(r2/r1) = 30 deg every ray that enters is
    guided to the exit; well outside it most are turned back."""
    mgr = hex_winston_scene(mode)
    soa = beams(mgr.cone_dz, 4000, 5.0)
    O.trace(mgr.flat(), soa, mgr.trace_opts(), nthreads=8)
    assert np.mean(soa["status"] == A.RB_FOCUS) > 0.97
    out = beams(mgr.cone_dz, 4000, 50.0)
    O.trace(mgr.flat(), out, mgr.trace_opts(), nthreads=8)
    assert np.mean(out["status"] == A.RB_FOCUS) < 0.5


@pytest.mark.parametrize("theta", [0.0, 10.0, 25.0, 35.0, 60.0])
def test_three_troughs_are_the_hexagonal_cone(theta):
    """Three troughs at 0, 60 and 120 deg bound |x . n_k| <= R(z) for the six
    side normals of a hexagon with apothem R(z): mode 1 is mode 0's solid,
    reached through AGeoWinstonCone2D's slab and parabola tests and a
    two-level intersection instead of AGeoWinstonConePoly's sector tests.
    Both trace every ray to the same fate and the same points."""
    res = []
    for mode in (0, 1):
        mgr = hex_winston_scene(mode)
        soa = beams(mgr.cone_dz, 10000, theta, seed=int(theta))
        O.trace(mgr.flat(), soa, mgr.trace_opts(), nthreads=8)
        res.append(soa)
    a, b = res
    assert np.array_equal(a["status"], b["status"]) and np.array_equal(a["npoints"], b["npoints"])
    assert a["npoints"].max() >= 4  # some rays reflect more than once
    dp = np.sqrt((a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2)
    assert dp.max() < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_hex_winston_gpu_matches_oracle(ctx, mode):
    from robast_amd._lib import make_rays
    print(f"HexWinstonCone mode {mode}", flush=True)
    mgr = hex_winston_scene(mode)
    parts = [beams(mgr.cone_dz, 10000, th, seed=int(th)) for th in (0.0, 10.0, 25.0, 35.0, 60.0)]
    soa = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    g, o = copy_soa(soa), copy_soa(soa)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(g), len(g["x"]), mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    gain, unstable = sensitivity(mgr.flat(), soa, mgr.trace_opts())
    mism, worst = conditioned_compare(g, o, gain, unstable, eps=1e-11)
    dp = np.sqrt((g["x"] - o["x"]) ** 2 + (g["y"] - o["y"]) ** 2 + (g["z"] - o["z"]) ** 2)
    print(f"  {mism} integer mismatches, worst |dpos| / bar {worst:.2e}, max |dpos| {dp.max():.2e} cm, "
          f"focused {np.mean(g['status'] == A.RB_FOCUS):.3f}", flush=True)
    assert dp[gain <= 1e3].max() <= 1e-7
