"""Mirror reflectance from a TH2 R(lambda, angle) (AMirror::SetReflectance(TH2),
read first by AMirror::GetReflectance, src/AMirror.cxx:39-60, with
TH2::Interpolate and clamped to [0, 1]).

The reference's testMirrorReflection (tutorials/unittest_robast.py:186-248)
sets a 2D reflectance through the four corners (300 nm, 0), (300 nm, 90 deg),
(500 nm, 0), (500 nm, 90 deg) = 0, 0.3, 0.7, 1 and expects 0.5 at
(400 nm, 45 deg) and an exited fraction of 0.5 within 3 sigma. It does so
with a TGraph2D; the same corners as the centres of a 2 x 2 TH2 give the same
value by bilinear interpolation, and that is restated here on the oracle and
the GPU. Also: the TH2 takes priority over a TGraph, values outside [0, 1]
are clamped, and the GPU matches the oracle per ray over random angles and
wavelengths."""
import math

import numpy as np
import pytest

from robast_amd import AMirror, ARayArray, TGeoBBox
from robast_amd import _abi as A
from robast_amd import geometry as G
from robast_amd.geoutil import TH2D
from robast_amd.materials import TGraph
from test_reference_units import BACKENDS, make_the_world, run

nm, m, deg = G.nm, G.m, G.deg


def corners_th2(values=(0.0, 0.3, 0.7, 1.0)):
    """Bin centres at lambda = 300, 500 nm and angle = 0, 90 deg."""
    h = TH2D("", "", 2, 200 * nm, 600 * nm, 2, -45 * deg, 135 * deg)
    (a, b, c, d) = values
    h.SetBinContent(1, 1, a)   # (300 nm, 0)
    h.SetBinContent(1, 2, b)   # (300 nm, 90 deg)
    h.SetBinContent(2, 1, c)   # (500 nm, 0)
    h.SetBinContent(2, 2, d)   # (500 nm, 90 deg)
    return h


def mirror_world(refl):
    G._reset_registry()
    mgr = make_the_world()
    mirror = AMirror("mirror", TGeoBBox("mirrorbox", 0.5 * m, 0.5 * m, 0.5 * m))
    for r in refl:
        mirror.SetReflectance(r)
    mgr.GetTopVolume().AddNode(mirror, 1)
    mgr.CloseGeometry()
    return mgr, mirror


def rays_45(n, lam=400 * nm):
    s = math.sqrt(2) / 2
    return ARayArray.from_arrays(np.zeros(n), np.zeros(n), np.full(n, 0.51 * m), np.full(n, s), np.zeros(n),
                                 np.full(n, -s), np.full(n, lam))


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_2d_reflectance(backend):
    mgr, mirror = mirror_world([corners_th2()])
    assert abs(mirror.GetReflectance(400 * nm, 45 * deg) - 0.5) < 1e-12
    N = 10000
    arr = rays_45(N)
    run(mgr, arr, backend)
    n = (arr.status == A.RB_EXIT).sum()
    assert abs(n - 0.5 * N) < 3 * math.sqrt(N * 0.25)
    assert n + (arr.status == A.RB_ABSORB).sum() == N


@pytest.mark.parametrize("backend", BACKENDS)
def test_th2_priority_and_clamping(backend):
    g = TGraph()
    g.SetPoint(0, 300 * nm, 0.0)
    g.SetPoint(1, 500 * nm, 0.0)
    for vals, want in (((1.7, 1.7, 1.7, 1.7), A.RB_EXIT), ((-0.5, -0.5, -0.5, -0.5), A.RB_ABSORB)):
        mgr, mirror = mirror_world([g, corners_th2(vals)])   # the TGraph alone would absorb every ray
        assert mirror.GetReflectance(400 * nm, 45 * deg) == (1.0 if want == A.RB_EXIT else 0.0)
        arr = rays_45(2000)
        run(mgr, arr, backend)
        assert np.all(arr.status == want)


@pytest.mark.gpu
def test_gpu_th2_mirror_per_ray(ctx):
    from helpers import compare, copy_soa
    from oracle import oracle as O
    from robast_amd._lib import make_rays
    print("TH2 mirror reflectance per ray", flush=True)
    h = TH2D("", "", 12, 250 * nm, 650 * nm, 9, 0.0, 90 * deg)
    rng = np.random.default_rng(4)
    for bx in range(1, 13):
        for by in range(1, 10):
            h.SetBinContent(bx, by, float(rng.uniform(-0.1, 1.1)))
    mgr, _ = mirror_world([h])
    n = 50000
    th = rng.uniform(0, 85, n) * deg
    ph = rng.uniform(0, 2 * math.pi, n)
    arr = ARayArray.from_arrays(np.zeros(n), np.zeros(n), np.full(n, 0.51 * m), np.sin(th) * np.cos(ph),
                                np.sin(th) * np.sin(ph), -np.cos(th), rng.uniform(230, 680, n) * nm)
    soa = arr.soa()
    g, o = copy_soa(soa), copy_soa(soa)
    ctx.trace(ctx.upload(mgr.flat()), make_rays(g), n, mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    compare(g, o)
    c = np.bincount(g["status"], minlength=6)
    assert c[A.RB_EXIT] > 5000 and c[A.RB_ABSORB] > 5000
