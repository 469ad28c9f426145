"""Border-surface conditions (ABorderSurfaceCondition, include/ABorderSurfaceCondition.h:24-45):
Gaussian roughness of the facet normal (AOpticsManager::GetFacetNormal,
src/AOpticsManager.cxx:250-301) and Lambertian reflection (DoReflection,
:209-224), selected per pair of volumes (AOpticalComponent::FindBorderSurfaceCondition,
src/AOpticalComponent.cxx:51-65).

The reference's own tests, restated with seeded Philox streams on the oracle
and the GPU: testMirrorScattaring (unittest_robast.py:333-381: 1 deg roughness
spreads the reflected direction by 2 deg per axis) and testLambertian
(:782-840: the cosine law, checked here on the emitted directions directly).
The condition lookup (pair, direction, first match) is checked exactly, and
the GPU reproduces the oracle per ray on a scene with rough and Lambertian
mirrors, a rough lens and a rough focal surface with an angular QE."""
import math

import numpy as np
import pytest
from scipy import stats

from robast_amd import (ABorderSurfaceCondition, AFocalSurface, ALens, AMirror, ARayArray, ARefractiveIndex,
                        TGeoBBox, TGeoTranslation, TGraph)
from robast_amd import _abi as A
from robast_amd import geometry as G
from test_reference_units import BACKENDS, make_the_world, rays_down, run

um, mm, cm, m, deg = G.um, G.mm, G.cm, G.m, G.deg


@pytest.mark.parametrize("backend", BACKENDS)
def test_mirror_scattering(backend):
    # unittest_robast.py:333-381
    mgr = make_the_world()
    mirror = AMirror("mirror", TGeoBBox("mirrorbox", 0.5 * m, 0.5 * m, 0.5 * m))
    cond = ABorderSurfaceCondition(mgr.GetTopVolume(), mirror)
    sigma = 1.0
    cond.SetGaussianRoughness(sigma * deg)
    mgr.GetTopVolume().AddNode(mirror, 1)
    mgr.CloseGeometry()
    N = 40000
    arr = rays_down(N, 0.8 * m)
    run(mgr, arr, backend)
    assert np.all(arr.status == A.RB_EXIT)
    # a facet tilted by alpha turns the reflection by 2 alpha: each direction
    # component is normal with width 2 sigma (the reference fits exactly this)
    for comp in (arr.dx / deg, arr.dy / deg):
        assert abs(comp.mean()) < 3 * 2 * sigma / math.sqrt(N)
        assert abs(comp.std() - 2 * sigma) < 3 * 2 * sigma / math.sqrt(2 * N)
    assert np.all(arr.dz > 0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_lambertian_cosine_law(backend):
    # unittest_robast.py:782-840: diffuse reflection, radiance independent of angle
    mgr = make_the_world()
    mirror = AMirror("mirror", TGeoBBox("mirrorbox", 5 * cm, 5 * cm, 1 * um))
    ABorderSurfaceCondition(mgr.GetTopVolume(), mirror).EnableLambertian(True)
    mgr.GetTopVolume().AddNode(mirror, 1)
    mgr.CloseGeometry()
    N = 100000
    th = 30 * deg   # oblique incidence: the lobe is about the normal, not the specular direction
    arr = rays_down(N, 1 * cm, d=(math.sin(th), 0, -math.cos(th)))
    arr.x[:] = -1 * cm * math.tan(th)
    run(mgr, arr, backend)
    assert np.all(arr.status == A.RB_EXIT)
    assert np.all(arr.dz > 0)
    # cos-weighted hemisphere: sin^2(theta) and the azimuth are uniform
    s2 = 1 - arr.dz ** 2
    assert stats.kstest(s2, "uniform").pvalue > 1e-3
    phi = (np.arctan2(arr.dy, arr.dx) + math.pi) / (2 * math.pi)
    assert stats.kstest(phi, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("backend", BACKENDS)
def test_condition_lookup(backend):
    """FindBorderSurfaceCondition: conditions belong to component1 and match
    component2; the first registered match applies."""
    def scene(setup):
        G._reset_registry()
        mgr = make_the_world()
        mirror = AMirror("mirror", TGeoBBox("mirrorbox", 0.5 * m, 0.5 * m, 0.5 * m))
        other = AMirror("other", TGeoBBox("otherbox", 0.1 * m, 0.1 * m, 0.1 * m))
        setup(mgr.GetTopVolume(), mirror, other)
        mgr.GetTopVolume().AddNode(mirror, 1)
        mgr.GetTopVolume().AddNode(other, 1, TGeoTranslation(0, 0, -5 * m))
        mgr.CloseGeometry()
        arr = rays_down(2000, 0.8 * m)
        run(mgr, arr, backend)
        assert np.all(arr.status == A.RB_EXIT)
        return arr

    def specular(arr):
        return np.all(arr.dz == 1.0) and np.all(arr.dx == 0.0)

    # a condition for another pair, or registered on the mirror (mirror -> world), does not apply
    assert specular(scene(lambda w, mi, o: ABorderSurfaceCondition(w, o).EnableLambertian(True)))
    assert specular(scene(lambda w, mi, o: ABorderSurfaceCondition(mi, w).EnableLambertian(True)))
    # world -> mirror applies; the first of two matching conditions wins
    def two(w, mi, o):
        ABorderSurfaceCondition(w, mi).SetGaussianRoughness(0.5 * deg)
        ABorderSurfaceCondition(w, mi).EnableLambertian(True)
    a = scene(two)
    assert not specular(a) and np.abs(a.dx).max() < 0.2   # rough (narrow), not Lambertian (wide)


def _mixed_scene():
    G._reset_registry()
    mgr = make_the_world()
    world = mgr.GetTopVolume()
    rough = AMirror("rough", TGeoBBox("roughbox", 20 * cm, 20 * cm, 1 * cm))
    ABorderSurfaceCondition(world, rough).SetGaussianRoughness(2 * deg)
    world.AddNode(rough, 1, TGeoTranslation(-30 * cm, 0, 0))
    diffuse = AMirror("diffuse", TGeoBBox("diffusebox", 20 * cm, 20 * cm, 1 * cm))
    ABorderSurfaceCondition(world, diffuse).EnableLambertian(True)
    world.AddNode(diffuse, 1, TGeoTranslation(30 * cm, 0, 0))
    lens = ALens("lens", TGeoBBox("lensbox", 40 * cm, 40 * cm, 2 * cm))
    lens.SetRefractiveIndex(ARefractiveIndex(1.5))
    ABorderSurfaceCondition(world, lens).SetGaussianRoughness(1 * deg)
    ABorderSurfaceCondition(lens, world).SetGaussianRoughness(3 * deg)
    world.AddNode(lens, 1, TGeoTranslation(0, 0, 40 * cm))
    focal = AFocalSurface("focal", TGeoBBox("focalbox", 80 * cm, 80 * cm, 1 * cm))
    qe = TGraph()
    qe.SetPoint(0, 0.0, 1.0)
    qe.SetPoint(1, 90 * deg, 0.2)
    focal.SetQuantumEfficiencyAngle(qe)
    ABorderSurfaceCondition(world, focal).SetGaussianRoughness(5 * deg)
    world.AddNode(focal, 1, TGeoTranslation(0, 0, 90 * cm))
    mgr.CloseGeometry()
    rng = np.random.default_rng(17)
    n = 50000
    x, y = rng.uniform(-50, 50, n) * cm, rng.uniform(-15, 15, n) * cm
    d = np.stack([rng.normal(0, 0.05, n), rng.normal(0, 0.05, n), np.full(n, -1.0)], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    arr = ARayArray.from_arrays(x, y, np.full(n, 70 * cm), d[:, 0].copy(), d[:, 1].copy(), d[:, 2].copy(),
                                rng.uniform(300, 600, n) * G.nm)
    return mgr, arr


def test_oracle_mixed_scene_runs():
    from oracle import oracle as O
    mgr, arr = _mixed_scene()
    O.trace(mgr.flat(), arr.soa(), mgr.trace_opts(), nthreads=8)
    c = np.bincount(arr.status, minlength=6)
    assert c[A.RB_RUN] == 0 and c[A.RB_FOCUS] > 1000 and c[A.RB_EXIT] > 1000


@pytest.mark.gpu
def test_gpu_border_surfaces_per_ray(ctx):
    from helpers import compare, copy_soa
    from oracle import oracle as O
    from robast_amd._lib import make_rays
    print("border surfaces: rough/Lambertian mirrors, rough lens and focal surface", flush=True)
    mgr, arr = _mixed_scene()
    soa = arr.soa()
    g, o = copy_soa(soa), copy_soa(soa)
    scene = ctx.upload(mgr.flat())
    ctx.trace(scene, make_rays(g), len(g["x"]), mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    compare(g, o)
    assert (g["status"] == A.RB_FOCUS).sum() > 1000


@pytest.mark.gpu
def test_gpu_adjoint_rejects_border_surfaces(ctx):
    import torch
    from robast_amd._lib import RbError, make_rays_device
    mgr, _ = _mixed_scene()
    scene = ctx.upload(mgr.flat())
    t = {f: torch.zeros(64, dtype=torch.float64, device="cuda") for f in
         ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
    t.update({f: torch.zeros(64, dtype=torch.int32, device="cuda") for f in ("status", "npoints", "last_node")})
    adj = A.rb_adjoint_desc()
    node_param = torch.full((mgr.flat().desc.n_nodes, 3), -1, dtype=torch.int32, device="cuda")
    vals = torch.zeros(1, dtype=torch.float64, device="cuda")
    grad = torch.zeros(1, dtype=torch.float64, device="cuda")
    adj.node_param, adj.param_values, adj.n_params, adj.z_param, adj.focal_node = (
        node_param.data_ptr(), vals.data_ptr(), 1, -1, 1)
    with pytest.raises(RbError, match="not differentiated"):
        ctx.adjoint(scene, make_rays_device(t), 64, mgr.trace_opts(device_ptrs=1), adj, grad.data_ptr())
