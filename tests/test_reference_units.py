"""The reference's own hot-path unit tests (tutorials/unittest_robast.py, the
rows that touch TraceNonSequential), restated with seeded Philox streams and
run on both the CPU oracle and the GPU (through librbgpu's C ABI)."""
import math

import numpy as np
import pytest

from robast_amd import (AFocalSurface, ALens, AMirror, AOpticalComponent, AOpticsManager, ARay, ARayArray,
                        ARayShooter, ARefractiveIndex, TGeoBBox, TGeoSphere, TGeoTranslation, TGraph)
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.materials import ASchottFormula

nm, um, mm, m, deg = G.nm, G.um, G.mm, G.m, G.deg

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def run(mgr, arr, backend):
    if backend == "oracle":
        # the scene TraceNonSequential would upload for this batch's wavelengths
        from oracle import oracle as O
        flat = mgr.flat_for(float(arr.lambda_.min()), float(arr.lambda_.max())) if arr.n else mgr.flat()
        O.trace(flat, arr.soa(), mgr.trace_opts(), nthreads=8)
    else:
        mgr.TraceNonSequential(arr)


def make_the_world():
    # unittest_robast.py:47-55
    mgr = AOpticsManager("manager", "manager")
    world = AOpticalComponent("world", TGeoBBox("worldbox", 20 * m, 20 * m, 20 * m))
    mgr.SetTopVolume(world)
    return mgr


def rays_down(n, z, lam=400 * nm, d=(0, 0, -1)):
    return ARayArray.from_arrays(np.zeros(n), np.zeros(n), np.full(n, z), np.full(n, d[0] * 1.0),
                                 np.full(n, d[1] * 1.0), np.full(n, d[2] * 1.0), np.full(n, lam))


def within_3sigma(p, k, n):
    return abs(k - n * p) < 3 * math.sqrt(n * p * (1 - p))


@pytest.mark.parametrize("backend", BACKENDS)
def test_snells_law(backend):
    # unittest_robast.py:428-468
    mgr = make_the_world()
    mgr.DisableFresnelReflection(True)
    lens = ALens("lens", TGeoBBox("lensbox", 0.5 * m, 0.5 * m, 1 * mm))
    idx = 1.5
    lens.SetRefractiveIndex(ARefractiveIndex(idx))
    mgr.GetTopVolume().AddNode(lens, 1)
    lens.AddNode(AFocalSurface("focal", TGeoBBox("focalbox", 0.5 * m, 0.5 * m, 0.1 * mm)), 1)
    mgr.CloseGeometry()
    th = 30 * deg
    arr = ARayArray()
    arr.Add(ARay(0, 400 * nm, 0, 0, 2 * mm, 0, math.sin(th), 0, -math.cos(th)))
    run(mgr, arr, backend)
    assert arr.status[0] == A.RB_FOCUS
    assert round(arr.dx[0] - math.sin(th) / idx, 7) == 0
    assert round(arr.dy[0], 7) == 0


@pytest.mark.parametrize("backend", BACKENDS)
def test_snells_law_schott_dispersion(backend):
    """test_snells_law through a lens whose index follows Schott's formula
    (src/ASchottFormula.cxx:43-55; an N-BK7 fit): at each wavelength the
    refracted direction is sin(30 deg) / n(lambda), n from the closed form."""
    a = (2.2718929, -1.0108077e-2, 1.0592509e-2, 2.0816965e-4, -7.6472538e-6, 4.9240991e-7)
    mgr = make_the_world()
    mgr.DisableFresnelReflection(True)
    lens = ALens("lens", TGeoBBox("lensbox", 0.5 * m, 0.5 * m, 1 * mm))
    lens.SetRefractiveIndex(ASchottFormula(*a))
    mgr.GetTopVolume().AddNode(lens, 1)
    lens.AddNode(AFocalSurface("focal", TGeoBBox("focalbox", 0.5 * m, 0.5 * m, 0.1 * mm)), 1)
    mgr.CloseGeometry()
    th = 30 * deg
    lams = (350 * nm, 450 * nm, 587.6 * nm, 800 * nm)
    arr = ARayArray()
    for k, lam in enumerate(lams):
        arr.Add(ARay(k, lam, 0, 0, 2 * mm, 0, math.sin(th), 0, -math.cos(th)))
    run(mgr, arr, backend)
    for k, lam in enumerate(lams):
        l = lam / um
        n = math.sqrt(a[0] + a[1] * l ** 2 + a[2] * l ** -2 + a[3] * l ** -4 + a[4] * l ** -6 + a[5] * l ** -8)
        assert arr.status[k] == A.RB_FOCUS
        assert abs(arr.dx[k] - math.sin(th) / n) < 1e-12, (lam, arr.dx[k], math.sin(th) / n)
    # dispersion: the blue ray bends most
    assert arr.dx[0] < arr.dx[1] < arr.dx[2] < arr.dx[3]


@pytest.mark.parametrize("backend", BACKENDS)
def test_quantum_efficiency(backend):
    # unittest_robast.py:470-522
    mgr = make_the_world()
    focal = AFocalSurface("focal", TGeoBBox("focalbox", 0.5 * m, 0.5 * m, 1 * mm))
    mgr.GetTopVolume().AddNode(focal, 1)
    mgr.CloseGeometry()
    qe_lambda = TGraph()
    qe_lambda.SetPoint(0, 300 * nm, 0.0)
    qe_lambda.SetPoint(1, 500 * nm, 1.0)
    qe_angle = TGraph()
    qe_angle.SetPoint(0, 0 * deg, 1.0)
    qe_angle.SetPoint(1, 90 * deg, 0.0)
    N = 1000 ** 2
    for i in range(3):
        if i == 1:
            focal.SetQuantumEfficiency(qe_lambda)
        elif i == 2:
            focal.SetQuantumEfficiencyAngle(qe_angle)
        mgr.CloseGeometry()
        raytr = TGeoTranslation("raytr", 0, 0, 2 * mm)
        arr = ARayShooter.Square(400 * nm, 1 * mm, 1000, None, raytr, (math.cos(45 * deg), 0, -math.sin(45 * deg)))
        run(mgr, arr, backend)
        nf, ns = arr.GetFocused().GetEntries(), arr.GetStopped().GetEntries()
        assert nf + ns == N
        if i == 0:
            assert nf == N
        elif i == 1:
            assert within_3sigma(0.5, nf, N)
        else:
            assert within_3sigma(0.25, nf, N)


@pytest.mark.parametrize("backend", BACKENDS)
def test_limit_for_suspended(backend):
    # unittest_robast.py:390-413: exactly 1000 points, no leak in 999 reflections
    mgr = make_the_world()
    mgr.SetLimit(1000)
    mgr.GetTopVolume().AddNode(AMirror("mirror", TGeoSphere("mirrorsphere", 0.1 * m, 0.2 * m)), 1)
    mgr.CloseGeometry()
    if backend == "gpu":
        ray = ARay(0, 400 * nm, 0, 0, 0, 0, 0, 0, -1)
        mgr.TraceNonSequential(ray)
        assert ray.GetNpoints() == 1000
        assert ray.IsSuspended()
        # the path bounces along the z axis between the inner wall at +-10 cm
        zs = np.array([p[2] for p in ray.points[1:]])
        assert np.all(np.abs(np.abs(zs) - 10.0) < 1e-5)
    else:
        arr = ARayArray()
        arr.Add(ARay(0, 400 * nm, 0, 0, 0, 0, 0, 0, -1))
        run(mgr, arr, backend)
        assert arr.npoints[0] == 1000
        assert arr.status[0] == A.RB_SUSPEND


@pytest.mark.parametrize("backend", BACKENDS)
def test_fresnel_reflection_normal_incidence(backend):
    # unittest_robast.py:122-160 (the multilayer half is out of scope)
    mgr = make_the_world()
    mgr.DisableFresnelReflection(False)
    lens = ALens("lens", TGeoBBox("lensbox", 0.5 * m, 0.5 * m, 0.5 * m))
    idx, wl, absl = 3.0, 400 * nm, 1 * um
    lens.SetRefractiveIndex(ARefractiveIndex(idx, ARefractiveIndex.AbsorptionLengthToExtinctionCoefficient(absl, wl)))
    mgr.GetTopVolume().AddNode(lens, 1)
    mgr.CloseGeometry()
    N = 100000
    arr = rays_down(N, 0.8 * m, wl)
    run(mgr, arr, backend)
    n = arr.GetExited().GetEntries()
    ref = (idx - 1) ** 2 / (idx + 1) ** 2
    assert (n - n ** 0.5 * 3) / N < ref < (n + n ** 0.5 * 3) / N
    # everything that entered is absorbed within a few um
    assert arr.GetExited().GetEntries() + arr.GetAbsorbed().GetEntries() == N


@pytest.mark.parametrize("backend", BACKENDS)
def test_mirror_reflectance_graph(backend):
    # unittest_robast.py:186-218 (the TGraph2D half is out of scope)
    mgr = make_the_world()
    mirror = AMirror("mirror", TGeoBBox("mirrorbox", 0.5 * m, 0.5 * m, 0.5 * m))
    mgr.GetTopVolume().AddNode(mirror, 1)
    g = TGraph()
    g.SetPoint(0, 300 * nm, 0.0)
    g.SetPoint(1, 500 * nm, 0.5)
    mirror.SetReflectance(g)
    assert round(mirror.GetReflectance(400 * nm, 0) - 0.25, 6) == 0
    mgr.CloseGeometry()
    N = 10000
    arr = rays_down(N, 0.8 * m)
    run(mgr, arr, backend)
    n = arr.GetExited().GetEntries()
    assert (n - n ** 0.5 * 3) / N < 0.25 < (n + n ** 0.5 * 3) / N
    assert n + arr.GetAbsorbed().GetEntries() == N


@pytest.mark.parametrize("backend", BACKENDS)
def test_absorption_length(backend):
    # unittest_robast.py:67-120: absorbed depth is exponential with mean 1 mm
    mgr = make_the_world()
    lens = ALens("lens", TGeoBBox("lensbox", 0.5 * m, 0.5 * m, 0.5 * m))
    wl, absl = 400 * nm, 1 * mm
    idx = ARefractiveIndex()
    gn = TGraph()
    gn.SetPoint(0, wl, 1)
    idx.SetRefractiveIndex(gn)
    gk = TGraph()
    gk.SetPoint(0, wl, ARefractiveIndex.AbsorptionLengthToExtinctionCoefficient(absl, wl))
    idx.SetExtinctionCoefficient(gk)
    lens.SetRefractiveIndex(idx)
    mgr.GetTopVolume().AddNode(lens, 1)
    mgr.CloseGeometry()
    N = 10000
    arr = ARayShooter.RandomSphere(400 * nm, N, seed=3)
    run(mgr, arr, backend)
    ab = arr.GetAbsorbed()
    assert ab.GetEntries() == N
    p = ab.last_points()
    dist = np.sqrt((p[:, :3] ** 2).sum(1))
    mean = dist.mean()
    assert abs(mean - absl) < 3 * absl / math.sqrt(N)
    # the time stamp is distance / c at n = 1
    assert np.allclose(p[:, 3], dist / 2.99792458e10, rtol=1e-12)
