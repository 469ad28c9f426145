"""The bounded path record when a ray enters the trace with a history (an
ARay that already holds several points, as a retraced ray does): slot k
holds the (k+1)-th point the trace adds, whatever the entry point count,
and ARayArray / the ARay overload keep the earlier points in front of the
new ones, as ARay::AddPoint appends to its TGeoTrack (src/ARay.cxx:42-63)."""
import numpy as np
import pytest

from helpers import copy_soa
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import configs
from robast_amd.rays import ARay

CAP = 8


def _with_history(soa, extra):
    s = copy_soa(soa)
    s["npoints"][:] += extra
    s["path_xyzt"] = np.full((len(s["x"]), CAP, 4), -7.0)
    s["path_node"] = np.full((len(s["x"]), CAP), -5, dtype=np.int32)
    return s


def _check(fresh, hist, extra):
    assert np.array_equal(hist["npoints"], fresh["npoints"] + extra)
    for k in ("x", "y", "z", "t", "status", "last_node"):
        assert np.array_equal(hist[k], fresh[k]), k
    assert np.array_equal(hist["path_node"], fresh["path_node"])
    assert np.array_equal(hist["path_xyzt"], fresh["path_xyzt"])
    # slots past the points added stay as they were
    added = fresh["npoints"] - 1
    untouched = np.arange(CAP)[None, :] >= added[:, None]
    assert np.all(hist["path_node"][untouched] == -5)


def _sample():
    mgr = configs.c4_scene()
    return mgr, O.generate(configs.c4_raygen(mgr.c4_dz), 3000)


def test_oracle_record_starts_at_the_first_added_point():
    mgr, soa = _sample()
    fresh, hist = _with_history(soa, 0), _with_history(soa, 3)
    O.trace(mgr.flat(), fresh, mgr.trace_opts(), nthreads=8)
    O.trace(mgr.flat(), hist, mgr.trace_opts(), nthreads=8)
    _check(fresh, hist, 3)
    assert fresh["npoints"].max() > CAP  # some paths overflow the record


@pytest.mark.gpu
def test_gpu_record_starts_at_the_first_added_point(ctx):
    from robast_amd._lib import make_rays
    print("path record with a ray history, C4 sample", flush=True)
    mgr, soa = _sample()
    scene = ctx.upload(mgr.flat())
    fresh, hist, ora = _with_history(soa, 0), _with_history(soa, 3), _with_history(soa, 3)
    for s in (fresh, hist):
        ctx.trace(scene, make_rays(s, s["path_xyzt"], s["path_node"]), len(s["x"]), mgr.trace_opts())
    _check(fresh, hist, 3)
    O.trace(mgr.flat(), ora, mgr.trace_opts(), nthreads=8)
    assert np.array_equal(hist["path_node"], ora["path_node"])
    assert np.abs(hist["path_xyzt"] - ora["path_xyzt"]).max() < 1e-6


@pytest.mark.gpu
def test_retraced_aray_keeps_its_history():
    """An ARay holding three points is traced through the single-ray
    overload: its points are the three, then the new ones; GetNpoints counts
    them all."""
    from test_reference_units import make_the_world
    from robast_amd.geometry import AFocalSurface, AMirror, TGeoBBox, TGeoTranslation
    mgr = make_the_world()
    world = mgr.GetTopVolume()
    world.AddNode(AMirror("mirror", TGeoBBox("mbox", 10.0, 10.0, 1.0)), 1, TGeoTranslation(0, 0, -10.0))
    world.AddNode(AFocalSurface("focal", TGeoBBox("fbox", 10.0, 10.0, 1.0)), 1, TGeoTranslation(0, 0, 20.0))
    mgr.CloseGeometry()
    ray = ARay(0, 400e-7, 1.0, 2.0, 50.0, 0.0, 0, 0, -1)
    ray.points = [(9.0, 9.0, 9.0, -3.0), (8.0, 8.0, 8.0, -2.0), (1.0, 2.0, 5.0, -1.0)]
    ray.nodes = ["a", "b"]
    mgr.TraceNonSequential(ray)
    assert ray.IsFocused()
    assert ray.points[:3] == [(9.0, 9.0, 9.0, -3.0), (8.0, 8.0, 8.0, -2.0), (1.0, 2.0, 5.0, -1.0)]
    assert ray.nodes[:2] == ["a", "b"]
    # from z = 5 down to the mirror top (z = -9), back up to the focal bottom (z = 19)
    assert ray.GetNpoints() == 5 and len(ray.nodes) == 4
    assert ray.points[3][2] == pytest.approx(-9.0 + 2e-6, abs=1e-12)
    assert ray.points[4][2] == pytest.approx(19.0, abs=1e-12)
