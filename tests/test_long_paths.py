"""Deterministic long paths (unit reflectance, no absorption, Fresnel
reflection off: no ray draws a random number), compared point by point
through the path record:

- a mirror sphere shell (the reference's testLimit geometry,
  unittest_robast.py:390-413): rays from inside the hollow bounce on the
  inner wall until the 200-point limit suspends them;
- a hexagonal light pipe, a 2 m mirror tube "pgon - pgon" (a composite)
  with a focal surface at its far end: rays enter at up to 20 deg and
  reflect up to 57 times on its flat walls;
- the C4 light-guide camera made deterministic: BK7 window without
  absorption or Fresnel reflection above 1,027 Winston cones with unit wall
  reflectance.

CPU: the oracle's paths obey the geometry (sphere: every bounce against a
40-digit construction of the chain; pipe: every bounce point on a wall,
the axial direction cosine conserved). GPU: every recorded point within 1e-7
cm of the oracle's, status, point count and node history identical."""
import mpmath as mp
import numpy as np
import pytest

from helpers import copy_soa
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import configs
from robast_amd import geometry as G
from robast_amd.manager import AOpticsManager

SPHERE_LIMIT = 200
PIPE_LEN = 200.0     # cm
PIPE_APOTHEM = 1.0   # cm, inner


def sphere_scene():
    G._reset_registry()
    mgr = AOpticsManager("manager", "sphere")
    world = G.AOpticalComponent("world", G.TGeoBBox("worldbox", 2000.0, 2000.0, 2000.0))
    mgr.SetTopVolume(world)
    world.AddNode(G.AMirror("mirror", G.TGeoSphere("mirrorsphere", 10.0, 20.0)), 1)
    mgr.SetLimit(SPHERE_LIMIT)
    mgr.DisableFresnelReflection(True)
    mgr.CloseGeometry()
    return mgr


def sphere_rays(n, seed=5):
    rng = np.random.default_rng(seed)
    soa = O.new_soa(n)
    p = rng.uniform(-4, 4, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    for k, a in enumerate("xyz"):
        soa[a][:] = p[:, k]
        soa["d" + a][:] = d[:, k]
    soa["lambda_"][:] = 400e-7
    soa["status"][:] = A.RB_RUN
    return soa


def pipe_scene():
    G._reset_registry()
    mgr = AOpticsManager("manager", "light pipe")
    world = G.AOpticalComponent("world", G.TGeoBBox("worldbox", 300.0, 300.0, 300.0))
    mgr.SetTopVolume(world)
    h = PIPE_LEN / 2
    outer = G.TGeoPgon("pipeOut", 0, 360, 6, 2)
    outer.DefineSection(0, -h, 0, PIPE_APOTHEM + 0.3)
    outer.DefineSection(1, h, 0, PIPE_APOTHEM + 0.3)
    inner = G.TGeoPgon("pipeIn", 0, 360, 6, 2)
    inner.DefineSection(0, -h - 1.0, 0, PIPE_APOTHEM)
    inner.DefineSection(1, h + 1.0, 0, PIPE_APOTHEM)
    world.AddNode(G.AMirror("pipe", G.TGeoCompositeShape("pipeComp", "pipeOut - pipeIn")), 1)
    end = G.TGeoPgon("endPgon", 0, 360, 6, 2)
    end.DefineSection(0, -h - 0.6, 0, PIPE_APOTHEM + 0.3)
    end.DefineSection(1, -h - 0.5, 0, PIPE_APOTHEM + 0.3)
    world.AddNode(G.AFocalSurface("end", end), 1)
    mgr.SetLimit(200)
    mgr.DisableFresnelReflection(True)
    mgr.CloseGeometry()
    return mgr


def pipe_rays(n, seed=7):
    """Rays entering the pipe's top opening at 3-20 deg to its axis."""
    rng = np.random.default_rng(seed)
    soa = O.new_soa(n)
    th = np.radians(rng.uniform(3, 20, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    r = 0.8 * PIPE_APOTHEM * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    soa["x"][:], soa["y"][:] = r * np.cos(a), r * np.sin(a)
    soa["z"][:] = PIPE_LEN / 2 + 0.5
    soa["dx"][:] = np.sin(th) * np.cos(ph)
    soa["dy"][:] = np.sin(th) * np.sin(ph)
    soa["dz"][:] = -np.cos(th)
    soa["lambda_"][:] = 400e-7
    soa["status"][:] = A.RB_RUN
    return soa


def c4_det():
    mgr = configs.c4_scene(deterministic=True)
    return mgr, O.generate(configs.c4_raygen(mgr.c4_dz), 20000)


def with_path(soa, cap):
    s = copy_soa(soa)
    s["path_xyzt"] = np.zeros((len(s["x"]), cap, 4))
    s["path_node"] = np.full((len(s["x"]), cap), -1, dtype=np.int32)
    return s


def test_sphere_paths_obey_the_geometry():
    mgr = sphere_scene()
    soa = with_path(sphere_rays(300), SPHERE_LIMIT)
    O.trace(mgr.flat(), soa, mgr.trace_opts())
    assert np.all(soa["status"] == A.RB_SUSPEND) and np.all(soa["npoints"] == SPHERE_LIMIT)
    # the 199 points the trace added: every one a bounce on the inner wall
    # r = 10 cm, stepped back 2e-6 cm along the incoming direction
    pts = soa["path_xyzt"][:, : SPHERE_LIMIT - 1, :3]
    r = np.linalg.norm(pts, axis=2)
    assert np.all(r < 10.0) and np.all(r > 10.0 - 2.1e-6)
    # the whole chain of 199 bounces of a few rays, in 40-digit arithmetic:
    # the hit on r = 10, the specular reflection about the radial normal and
    # the 2e-6 cm step back along the incoming direction
    mp.mp.dps = 40
    src = sphere_rays(300)
    for i in (0, 77, 151, 299):
        p = [mp.mpf(float(src[k][i])) for k in "xyz"]
        d = [mp.mpf(float(src["d" + k][i])) for k in "xyz"]
        for j in range(SPHERE_LIMIT - 1):
            b = sum(p[k] * d[k] for k in range(3))
            c = sum(v * v for v in p) - 100
            t = -b + mp.sqrt(b * b - c)
            h = [p[k] + t * d[k] for k in range(3)]
            nrm = [v / 10 for v in h]
            cs = sum(d[k] * nrm[k] for k in range(3))
            d2 = [d[k] - 2 * cs * nrm[k] for k in range(3)]
            m = mp.sqrt(sum(v * v for v in d2))
            p = [h[k] - mp.mpf("2e-6") * d[k] for k in range(3)]
            d = [v / m for v in d2]
            got = soa["path_xyzt"][i, j, :3]
            assert max(abs(float(p[k]) - got[k]) for k in range(3)) < 1e-11, (i, j)


def test_pipe_paths_obey_the_geometry():
    mgr = pipe_scene()
    soa = with_path(pipe_rays(2000), 200)
    O.trace(mgr.flat(), soa, mgr.trace_opts())
    assert np.all(soa["status"] == A.RB_FOCUS)
    assert soa["npoints"].min() >= 4 and soa["npoints"].max() >= 50
    # flat walls conserve the axial direction cosine
    dz = soa["dz"]
    th = np.cos(np.arcsin(np.hypot(pipe_rays(2000)["dx"], pipe_rays(2000)["dy"])))
    assert np.abs(dz + th).max() < 1e-12
    # every bounce (all points but the first and the last) sits on a wall:
    # the largest apothem projection equals the inner apothem, less the step back
    for i in range(0, 2000, 97):
        k = soa["npoints"][i]
        p = soa["path_xyzt"][i, : k - 2, :2]
        phi = np.radians(30.0 + np.arange(6) * 60.0)  # TGeoPgon sector middles from phi1 = 0
        apo = (p[:, :1] * np.cos(phi) + p[:, 1:2] * np.sin(phi)).max(axis=1)
        assert np.abs(apo - PIPE_APOTHEM).max() < 3e-6


def test_deterministic_c4_draws_nothing():
    """Two oracle traces with different seeds give identical rays."""
    mgr, soa = c4_det()
    a, b = copy_soa(soa), copy_soa(soa)
    oa = mgr.trace_opts()
    ob = mgr.trace_opts()
    ob.seed = oa.seed + 12345
    O.trace(mgr.flat(), a, oa, nthreads=8)
    O.trace(mgr.flat(), b, ob, nthreads=8)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert a["npoints"].max() >= 8


def _gpu_vs_oracle(ctx, mgr, soa, cap):
    from robast_amd._lib import make_rays
    g, o = with_path(soa, cap), with_path(soa, cap)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(g, g["path_xyzt"], g["path_node"]), len(g["x"]), mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    for k in ("status", "npoints", "last_node", "path_node"):
        assert np.array_equal(g[k], o[k]), k
    n = o["npoints"]
    valid = np.arange(cap)[None, :] < np.minimum(n, cap)[:, None]
    dp = np.linalg.norm(g["path_xyzt"][..., :3] - o["path_xyzt"][..., :3], axis=2)[valid]
    dt = np.abs(g["path_xyzt"][..., 3] - o["path_xyzt"][..., 3])[valid]
    return g, o, dp, dt


@pytest.mark.gpu
def test_sphere_gpu_path_matches_oracle_point_by_point(ctx):
    print("long paths: sphere shell, 200 points per ray", flush=True)
    mgr = sphere_scene()
    g, o, dp, dt = _gpu_vs_oracle(ctx, mgr, sphere_rays(2000), SPHERE_LIMIT)
    print(f"  max |dpoint| {dp.max():.2e} cm over {dp.size} points, max |dt| {dt.max():.2e} s", flush=True)
    assert np.all(o["npoints"] == SPHERE_LIMIT)
    assert dp.max() <= 1e-7 and dt.max() <= 1e-15


@pytest.mark.gpu
def test_pipe_gpu_path_matches_oracle_point_by_point(ctx):
    print("long paths: hexagonal light pipe, up to 57 bounces", flush=True)
    mgr = pipe_scene()
    g, o, dp, dt = _gpu_vs_oracle(ctx, mgr, pipe_rays(20000), 200)
    print(f"  max |dpoint| {dp.max():.2e} cm over {dp.size} points, max |dt| {dt.max():.2e} s", flush=True)
    assert o["npoints"].max() >= 50
    assert dp.max() <= 1e-7 and dt.max() <= 1e-15


@pytest.mark.gpu
def test_deterministic_c4_gpu_path_matches_oracle_point_by_point(ctx):
    print("long paths: deterministic C4", flush=True)
    mgr, soa = c4_det()
    g, o, dp, dt = _gpu_vs_oracle(ctx, mgr, soa, 100)
    print(f"  max |dpoint| {dp.max():.2e} cm over {dp.size} points, longest path {o['npoints'].max()}", flush=True)
    assert dp.max() <= 1e-7
