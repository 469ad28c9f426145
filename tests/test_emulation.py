"""The device tracer's own source (robast_amd/csrc/rb_trace.h, rb_device.h,
rb_scene_build.h) built for the CPU through rb_host_emulation.h, with
AddressSanitizer and UndefinedBehaviorSanitizer, and compared with the
oracle ray for ray: the configurations' samples, the plain and the mixed
forms, and seeded random scenes with every shape and border condition. An
out-of-bounds access, a use of uninitialised or freed memory, or undefined
arithmetic in the device algorithms aborts the program and fails the test.
The same program traces with the oracle (oracle/rb_oracle.cpp, linked in)
under the same sanitizers, which must reproduce the oracle library's
results bit for bit.

File layout read by tests/emulation/emul_trace.cpp: for each array, an
int64 element count and an int64 byte count, then the bytes. In order: the
nine scene arrays of rb_scene_desc, rb_trace_opts (one), the eight FP64 ray
arrays, the three int32 ray arrays, then the path record (xyzt, node; empty
arrays when there is none). Output: the int64 count of the mixed form's
FP64 reruns, then the ray arrays and the path record in the same order,
raw."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import copy_soa
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emulation", "emul_trace.cpp")
ORACLE_SRC = os.path.join(os.path.dirname(HERE), "oracle", "rb_oracle.cpp")
F64 = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
I32 = ("status", "npoints", "last_node")


@pytest.fixture(scope="session")
def emul(tmp_path_factory):
    """The sanitized CPU build of the device tracer (a program)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emul") / "emul_trace")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-omit-frame-pointer",
                           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-o", exe, SRC, ORACLE_SRC,
                           "-lpthread"])
    return exe


def _put(fh, a, count=None):
    a = np.ascontiguousarray(a)
    if count is not None:
        a = a[:count]
    fh.write(np.array([a.shape[0] if a.ndim else 0, a.nbytes], dtype=np.int64).tobytes())
    fh.write(a.tobytes())


def emul_trace(exe, flat, soa, opts, tmpdir, oracle=False, lane=False):
    """Traces soa (host SoA dict, with optional path_xyzt / path_node) in
    place with the sanitized device code (with oracle: the sanitized
    oracle; with lane: the lane kernel's form even for scenes with a lens);
    returns the mixed form's count of FP64 reruns."""
    n = len(soa["x"])
    fin, fout = os.path.join(tmpdir, "in.bin"), os.path.join(tmpdir, "out.bin")
    with open(fin, "wb") as fh:
        d = flat.desc
        for a, count in ((flat.nodes, d.n_nodes), (flat.children, d.n_children), (flat.volumes, d.n_volumes),
                         (flat.shapes, d.n_shapes), (flat.params, d.n_params), (flat.tables, d.n_tables),
                         (flat.table_xy, d.n_table_xy), (flat.borders, d.n_borders), (flat.grids, d.n_grids)):
            _put(fh, a, count)
        fh.write(np.array([1, A.C.sizeof(opts)], dtype=np.int64).tobytes())
        fh.write(bytes(opts))
        for k in F64:
            _put(fh, soa[k].astype(np.float64))
        for k in I32:
            _put(fh, soa[k].astype(np.int32))
        path = soa.get("path_xyzt")
        _put(fh, path.reshape(-1) if path is not None else np.zeros(0))
        _put(fh, soa["path_node"].reshape(-1) if path is not None else np.zeros(0, dtype=np.int32))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + (["--oracle"] if oracle else ["--lane"] if lane else []) + [fin, fout],
                       capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, f"emul_trace exited {r.returncode}:\n{r.stderr[-4000:]}"
    raw = open(fout, "rb").read()
    fallbacks = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    off = 8
    for k in F64:
        soa[k][:] = np.frombuffer(raw[off:off + 8 * n], dtype=np.float64)
        off += 8 * n
    for k in I32:
        soa[k][:] = np.frombuffer(raw[off:off + 4 * n], dtype=np.int32)
        off += 4 * n
    if "path_xyzt" in soa:
        m = soa["path_xyzt"].size
        soa["path_xyzt"].reshape(-1)[:] = np.frombuffer(raw[off:off + 8 * m], dtype=np.float64)
        off += 8 * m
        soa["path_node"].reshape(-1)[:] = np.frombuffer(raw[off:off + 4 * soa["path_node"].size], dtype=np.int32)
    return fallbacks


def _inputs(name, mgr, n):
    if name == "c1":
        return configs.c1_rays(1.0, nside=64).soa()
    gen = {"c2": configs.c2_raygen, "c3": configs.c3_raygen}.get(name)
    return O.generate(gen() if gen else configs.c4_raygen(mgr.c4_dz), n)


def _agree(e, o, pos_tol, long_tol=None, long_from=12):
    assert np.array_equal(e["status"], o["status"])
    assert np.array_equal(e["npoints"], o["npoints"])
    assert np.array_equal(e["last_node"], o["last_node"])
    dp = np.sqrt((e["x"] - o["x"]) ** 2 + (e["y"] - o["y"]) ** 2 + (e["z"] - o["z"]) ** 2)
    # a navigation step that degraded to TGeoShape::Big leaves the point near
    # 1e31 cm: compared relative to its size
    pn = np.sqrt(o["x"] ** 2 + o["y"] ** 2 + o["z"] ** 2)
    big = pn > 1e4
    assert np.all(dp[big] <= 1e-12 * pn[big])
    dp = np.where(big, 0.0, dp)
    short = o["npoints"] <= long_from
    assert dp[short].max(initial=0) <= pos_tol, dp[short].max()
    if long_tol is not None:
        assert dp[~short].max(initial=0) <= long_tol, dp[~short].max()
    return dp


@pytest.mark.parametrize("name, n", [("c1", 0), ("c2", 3000), ("c3", 1500), ("c4", 600)])
@pytest.mark.parametrize("precision", [A.RB_PRECISION_FP64, A.RB_PRECISION_MIXED])
def test_device_code_under_sanitizers_matches_oracle(emul, tmp_path, name, n, precision):
    mgr = {"c1": configs.c1_scene, "c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene}[name]()
    soa = _inputs(name, mgr, n)
    e, o = copy_soa(soa), copy_soa(soa)
    opts = mgr.trace_opts()
    opts.precision = precision
    fallbacks = emul_trace(emul, mgr.flat(), e, opts, str(tmp_path))
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    if precision == A.RB_PRECISION_FP64:
        assert fallbacks == 0
        _agree(e, o, 1e-7, long_tol=1e-3 if name == "c4" else None)
    else:
        # the mixed form against the FP64 oracle: the bench's error bar
        flips = int((e["status"] != o["status"]).sum())
        assert flips <= max(1, len(e["x"]) // 1000), flips
        f = (e["status"] == A.RB_FOCUS) & (o["status"] == A.RB_FOCUS)
        dp = np.sqrt((e["x"] - o["x"]) ** 2 + (e["y"] - o["y"]) ** 2 + (e["z"] - o["z"]) ** 2)[f]
        assert np.sqrt(np.mean(dp ** 2)) < 1e-4  # 1 um


@pytest.mark.parametrize("surfaces", [False, True])
@pytest.mark.parametrize("seed", [11, 41])
def test_random_scenes_under_sanitizers_match_oracle(emul, tmp_path, seed, surfaces):
    from test_random_scenes import random_rays, random_scene
    mgr = random_scene(seed, surfaces)
    soa = random_rays(seed, 1500)
    e, o = copy_soa(soa), copy_soa(soa)
    emul_trace(emul, mgr.flat(), e, mgr.trace_opts(), str(tmp_path))
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    _agree(e, o, 1e-7)


def test_path_record_under_sanitizers(emul, tmp_path):
    """The bounded path record (capacity 6, shorter than some paths) written
    by the device code equals the oracle's, slot for slot."""
    mgr = configs.c4_scene()
    soa = _inputs("c4", mgr, 400)
    cap = 6
    e, o = copy_soa(soa), copy_soa(soa)
    for d in (e, o):
        d["path_xyzt"] = np.zeros((len(d["x"]), cap, 4))
        d["path_node"] = np.full((len(d["x"]), cap), -1, dtype=np.int32)
    emul_trace(emul, mgr.flat(), e, mgr.trace_opts(), str(tmp_path))
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    assert np.array_equal(e["npoints"], o["npoints"]) and e["npoints"].max() > cap
    assert np.array_equal(e["path_node"], o["path_node"])
    assert np.abs(e["path_xyzt"] - o["path_xyzt"]).max() < 1e-6


@pytest.mark.parametrize("name, n", [("c1", 0), ("c2", 3000), ("c3", 1500), ("c4", 600), ("r41", 1500)])
def test_oracle_under_sanitizers_is_the_oracle(emul, tmp_path, name, n):
    """The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
    runs clean and gives the oracle library's results bit for bit (C1-C4 and
    a random scene with border conditions, path record on)."""
    if name == "r41":
        from test_random_scenes import random_rays, random_scene
        mgr, soa = random_scene(41, True), random_rays(41, n)
    else:
        mgr = {"c1": configs.c1_scene, "c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene}[name]()
        soa = _inputs(name, mgr, n)
    e, o = copy_soa(soa), copy_soa(soa)
    for d in (e, o):
        d["path_xyzt"] = np.zeros((len(d["x"]), 5, 4))
        d["path_node"] = np.full((len(d["x"]), 5), -1, dtype=np.int32)
    emul_trace(emul, mgr.flat(), e, mgr.trace_opts(), str(tmp_path), oracle=True)
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    for k in o:
        assert np.array_equal(e[k], o[k]), k


@pytest.mark.parametrize("name", ["c4", "c4_det", "r41"])
def test_work_queue_form_equals_lane_form_under_sanitizers(emul, tmp_path, name):
    """The work-queue kernels' form of the device code -- the last-daughter
    hint, the searches and relocations within neighbour cells, the boolean
    nodes' child-box tests -- gives the lane kernel's results bit for bit
    (path record included), on C4 (2,055 daughters: cells built), its
    deterministic variant, and a random scene with a lens and composites."""
    if name == "r41":
        from test_random_scenes import random_rays, random_scene
        mgr, soa = random_scene(41, True), random_rays(41, 1500)
    else:
        mgr = configs.c4_scene(deterministic=name == "c4_det")
        soa = _inputs("c4", mgr, 1500)
    a, b = copy_soa(soa), copy_soa(soa)
    for d in (a, b):
        d["path_xyzt"] = np.zeros((len(d["x"]), 8, 4))
        d["path_node"] = np.full((len(d["x"]), 8), -1, dtype=np.int32)
    emul_trace(emul, mgr.flat(), a, mgr.trace_opts(), str(tmp_path))
    emul_trace(emul, mgr.flat(), b, mgr.trace_opts(), str(tmp_path), lane=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), (name, k)
    assert a["npoints"].max() >= 4
