"""Per-ray parity of the HIP tracer (through librbgpu's C ABI) with the CPU
FP64 oracle on identical inputs: status, point count and last node
identical; terminal point within 1e-7 cm (1e-9 m); direction within 1e-9."""
import numpy as np
import pytest

from helpers import compare, copy_soa, counts
from oracle import oracle as O
from robast_amd import _abi as A
from robast_amd import configs
from robast_amd._lib import RbError, make_rays, make_rays_device

pytestmark = pytest.mark.gpu


def _trace_both(ctx, mgr, soa, nthreads=8, path_cap=0):
    g = copy_soa(soa)
    o = copy_soa(soa)
    if path_cap:
        for d in (g, o):
            d["path_xyzt"] = np.zeros((len(d["x"]), path_cap, 4))
            d["path_node"] = np.full((len(d["x"]), path_cap), -1, dtype=np.int32)
    scene = ctx.upload(mgr.flat())
    rays = make_rays(g, g.get("path_xyzt"), g.get("path_node"))
    ctx.trace(scene, rays, len(g["x"]), mgr.trace_opts())
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=nthreads)
    return g, o


@pytest.mark.parametrize("deg", configs.C1_ANGLES_DEG)
def test_c1_parabolic_per_ray(ctx, deg):
    print(f"C1 {deg} deg, 316x316 grid", flush=True)
    mgr = configs.c1_scene()
    g, o = _trace_both(ctx, mgr, configs.c1_rays(deg).soa())
    compare(g, o)
    c = counts(g["status"])
    assert c[A.RB_FOCUS] > 60000 and c[A.RB_RUN] == 0


def test_c1_path_record_matches(ctx):
    mgr = configs.c1_scene()
    soa = configs.c1_rays(1.0, nside=64).soa()
    g, o = _trace_both(ctx, mgr, soa, path_cap=4)
    compare(g, o)
    assert np.array_equal(g["path_node"], o["path_node"])
    assert np.abs(g["path_xyzt"][..., :3] - o["path_xyzt"][..., :3]).max() < 1e-7


@pytest.mark.parametrize("deterministic", [False, True])
def test_c2_davies_cotton_sample_per_ray(ctx, deterministic):
    print(f"C2 sample deterministic={deterministic}", flush=True)
    mgr = configs.c2_scene(deterministic=deterministic)
    soa = O.generate(configs.c2_raygen(), 200000)
    g, o = _trace_both(ctx, mgr, soa)
    compare(g, o)
    c = counts(g["status"])
    if deterministic:
        assert c[A.RB_ABSORB] == 0
    else:
        assert c[A.RB_ABSORB] > 0


def test_generator_matches_oracle(ctx):
    import torch
    n = 4096
    dev = torch.device("cuda")
    t = {f: torch.zeros(n, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
    for f in ("status", "npoints", "last_node"):
        t[f] = torch.zeros(n, dtype=torch.int32, device=dev)
    ctx.generate(configs.c2_raygen(), make_rays_device(t), n)
    got = {k: v.cpu().numpy() for k, v in t.items()}
    want = O.generate(configs.c2_raygen(), n)
    for f in ("x", "y", "z", "dx", "dy", "dz", "lambda_"):
        assert np.abs(got[f] - want[f]).max() < 1e-9, f
    assert np.all(got["status"] == A.RB_RUN) and np.all(got["npoints"] == 1)


def test_c2_full_size_properties_and_prefix_parity(ctx):
    """1e7 rays resident in HBM: terminal-state invariants at full size, and
    the first 2e5 rays identical to the oracle (Philox streams are keyed by
    ray id, so a prefix of the big batch is a self-contained parity case)."""
    import torch
    print("C2 full size 1e7", flush=True)
    N, P = 10_000_000, 200_000
    mgr = configs.c2_scene()
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
    for f in ("status", "npoints", "last_node"):
        t[f] = torch.empty(N, dtype=torch.int32, device=dev)
    rays = make_rays_device(t)
    ctx.generate(configs.c2_raygen(), rays, N)
    src = {k: v[:P].cpu().numpy().copy() for k, v in t.items()}
    ctx.trace(scene, rays, N, mgr.trace_opts(device_ptrs=1))
    st = t["status"]
    c = torch.bincount(st, minlength=6).cpu().numpy()
    assert c.sum() == N and c[A.RB_RUN] == 0 and c[A.RB_SUSPEND] == 0
    norm = torch.sqrt(t["dx"] ** 2 + t["dy"] ** 2 + t["dz"] ** 2)
    assert float((norm - 1).abs().max()) < 1e-12
    foc = st == A.RB_FOCUS
    # focused rays end on the camera's lower face z = f (the tube spans f .. f + 2 mm)
    assert float((t["z"][foc] - configs.C2_F).abs().max()) < 1e-9
    assert float(torch.sqrt(t["x"][foc] ** 2 + t["y"][foc] ** 2).max()) <= configs.C2_CAMERA_R + 1e-9
    o = copy_soa(src)
    O.trace(mgr.flat(), o, mgr.trace_opts(), nthreads=8)
    g = {k: v[:P].cpu().numpy() for k, v in t.items()}
    compare(g, o)
    # the sample's focused fraction is the full run's within binomial error
    p_full = c[A.RB_FOCUS] / N
    p_samp = (o["status"] == A.RB_FOCUS).mean()
    assert abs(p_samp - p_full) < 4 * np.sqrt(p_full * (1 - p_full) / P)


def test_device_pointer_path_equals_host_path(ctx):
    import torch
    mgr = configs.c1_scene()
    soa = configs.c1_rays(0.5, nside=128).soa()
    h = copy_soa(soa)
    scene = ctx.upload(mgr.flat())
    ctx.trace(scene, make_rays(h), len(h["x"]), mgr.trace_opts())
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in soa.items()}
    ctx.trace(scene, make_rays_device(t), len(h["x"]), mgr.trace_opts(device_ptrs=1))
    for k in h:
        assert np.array_equal(t[k].cpu().numpy(), h[k]), k
    assert ctx.last_trace_ms() > 0 and ctx.last_trace_launches() >= 1


def test_non_running_rays_untouched(ctx):
    mgr = configs.c1_scene()
    soa = configs.c1_rays(0.0, nside=32).soa()
    soa["status"][::3] = A.RB_STOP
    before = copy_soa(soa)
    scene = ctx.upload(mgr.flat())
    ctx.trace(scene, make_rays(soa), len(soa["x"]), mgr.trace_opts())
    for k in soa:
        assert np.array_equal(soa[k][::3], before[k][::3]), k


def test_unsupported_scene_is_reported(ctx):
    mgr = configs.c1_scene()
    flat = mgr.flat()
    flat.shapes[0]["type"] = 42
    with pytest.raises(RbError, match="unsupported shape type"):
        ctx.upload(flat)
    flat.shapes[0]["type"] = A.RB_SHAPE_BOX
    flat.nodes[1]["volume"] = 99
    with pytest.raises(RbError, match="node volume out of range"):
        ctx.upload(flat)


def test_c3_aspheric_sample_per_ray(ctx):
    print("C3 Schwarzschild-Couder sample", flush=True)
    mgr = configs.c3_scene()
    soa = O.generate(configs.c3_raygen(), 200000)
    g, o = _trace_both(ctx, mgr, soa)
    compare(g, o)
    assert counts(g["status"])[A.RB_FOCUS] > 10000


@pytest.mark.parametrize("config", ["c1", "c2", "c3", "c4"])
def test_mixed_precision_against_fp64(ctx, config):
    """Mixed form (FP32 search in rebased frames + FP64 polish) against the
    plain FP64 kernel at 1e6 rays: status flips rare, focused-hit RMS < 1 um."""
    import torch
    print(f"mixed vs fp64 {config}", flush=True)
    N = 1_000_000
    mgr = {"c1": configs.c1_scene, "c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene}[config]()
    gen = {"c1": configs.c1_raygen(1.0, 1000), "c2": configs.c2_raygen(), "c3": configs.c3_raygen(),
           "c4": configs.c4_raygen(getattr(mgr, "c4_dz", 0.0))}[config]
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    out = {}
    for mixed in (False, True):
        t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
        for f in ("status", "npoints", "last_node"):
            t[f] = torch.empty(N, dtype=torch.int32, device=dev)
        rays = make_rays_device(t)
        ctx.generate(gen, rays, N)
        mgr.SetPrecision(mixed)
        ctx.trace(scene, rays, N, mgr.trace_opts(device_ptrs=1))
        out[mixed] = t
    a, b = out[False], out[True]
    flips = int((a["status"] != b["status"]).sum())
    both = (a["status"] == A.RB_FOCUS) & (b["status"] == A.RB_FOCUS)
    d2 = (a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2
    rms_um = float(torch.sqrt(d2[both].mean())) * 1e4
    fa = float((a["status"] == A.RB_FOCUS).double().mean())
    fb = float((b["status"] == A.RB_FOCUS).double().mean())
    print(f"  flips {flips} / {N}, focused hit RMS {rms_um:.3e} um, focused {fa:.6f} vs {fb:.6f}", flush=True)
    # the focused fraction agrees within binomial error
    assert abs(fa - fb) < 3 * np.sqrt(fa * (1 - fa) / N)
    # composites are measured in FP64, primitives ranked in FP32 and polished:
    # rays pair one to one, multi-bounce light guides included
    assert flips <= 10
    assert rms_um < 1.0


def test_c4_light_guides_sample_per_ray(ctx):
    print("C4 BK7 window + 1027 Winston cones sample", flush=True)
    mgr = configs.c4_scene()
    soa = O.generate(configs.c4_raygen(mgr.c4_dz), 100000)
    g, o = _trace_both(ctx, mgr, soa)
    compare(g, o)
    c = counts(g["status"])
    assert c[A.RB_FOCUS] > 50000 and c[A.RB_ABSORB] > 0
    assert g["npoints"].max() > 10


@pytest.mark.parametrize("config", ["c2", "c4"])
def test_coherence_order_gives_identical_results(ctx, config):
    """Tracing in coherence order (sorted by first daughter-box entry, traced
    through a permutation) and in input order gives bit-identical rays, one
    lane per ray, from the persistent work queue, or with the block's rays
    regrouped by phase before every step. The schedules are
    separately compiled kernels; with no FMA contraction every operation
    rounds as written in both, so they agree bit for bit too."""
    import torch
    print(f"order input vs coherent, lane vs persistent {config}", flush=True)
    N = 1 << 20
    mgr = {"c2": configs.c2_scene, "c4": configs.c4_scene}[config]()
    gen = configs.c2_raygen() if config == "c2" else configs.c4_raygen(mgr.c4_dz)
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    out = {}
    for order, schedule in ((A.RB_ORDER_INPUT, A.RB_SCHEDULE_LANE), (A.RB_ORDER_COHERENT, A.RB_SCHEDULE_LANE),
                            (A.RB_ORDER_COHERENT, A.RB_SCHEDULE_PERSISTENT),
                            (A.RB_ORDER_INPUT, A.RB_SCHEDULE_PERSISTENT),
                            (A.RB_ORDER_COHERENT, A.RB_SCHEDULE_REGROUP), (A.RB_ORDER_INPUT, A.RB_SCHEDULE_REGROUP)):
        t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
        for f in ("status", "npoints", "last_node"):
            t[f] = torch.empty(N, dtype=torch.int32, device=dev)
        ctx.generate(gen, make_rays_device(t), N)
        t["status"][::7] = A.RB_STOP   # some rays not running: left untouched in every order
        ctx.trace(scene, make_rays_device(t), N, mgr.trace_opts(device_ptrs=1, order=order, schedule=schedule))
        assert ctx.last_trace_launches() == (4 if order == A.RB_ORDER_COHERENT else 1)  # key, sort, trace, copy-back
        out[order, schedule] = t
    for schedule in (A.RB_SCHEDULE_LANE, A.RB_SCHEDULE_PERSISTENT, A.RB_SCHEDULE_REGROUP):
        a, b = out[A.RB_ORDER_INPUT, schedule], out[A.RB_ORDER_COHERENT, schedule]
        for k in a:
            assert torch.equal(a[k], b[k]), (schedule, k)
    for schedule in (A.RB_SCHEDULE_PERSISTENT, A.RB_SCHEDULE_REGROUP):
        a, b = out[A.RB_ORDER_INPUT, A.RB_SCHEDULE_LANE], out[A.RB_ORDER_INPUT, schedule]
        for k in a:
            assert torch.equal(a[k], b[k]), (schedule, k)


def test_auto_order_work_queue_threshold(ctx):
    """The Winston cone array (a lens scene, traced with the work-queue
    schedule) is sorted by the auto order from 2^18 rays: just above that
    the trace runs the key kernel, sort, trace and copy-back, just below it
    the trace alone, and both equal input order."""
    import torch
    mgr = configs.c4_scene()
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    for N, launches in (((1 << 18) + 5, 4), ((1 << 18) - 5, 1)):
        print(f"C4 auto order n={N}", flush=True)
        out = {}
        for order in (A.RB_ORDER_AUTO, A.RB_ORDER_INPUT):
            t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
            for f in ("status", "npoints", "last_node"):
                t[f] = torch.empty(N, dtype=torch.int32, device=dev)
            ctx.generate(configs.c4_raygen(mgr.c4_dz), make_rays_device(t), N)
            ctx.trace(scene, make_rays_device(t), N, mgr.trace_opts(device_ptrs=1, order=order))
            assert ctx.last_trace_launches() == (launches if order == A.RB_ORDER_AUTO else 1)
            out[order] = t
        for k in out[A.RB_ORDER_AUTO]:
            assert torch.equal(out[A.RB_ORDER_AUTO][k], out[A.RB_ORDER_INPUT][k]), (N, k)


def test_coherent_order_with_path_record(ctx):
    """A coherent-order trace (rays read and written through their staged
    lines) with the path record on equals the input-order trace: terminal
    states, every recorded point and node, and the record's indexing from the
    point count a ray enters with (some rays enter with 3 points)."""
    import torch
    N, cap = 100_000, 4
    print(f"C2 path record, coherent vs input order, n={N}", flush=True)
    mgr = configs.c2_scene()
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    out = {}
    for order in (A.RB_ORDER_INPUT, A.RB_ORDER_COHERENT):
        t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
        for f in ("status", "npoints", "last_node"):
            t[f] = torch.empty(N, dtype=torch.int32, device=dev)
        t["path_xyzt"] = torch.zeros((N, cap, 4), dtype=torch.float64, device=dev)
        t["path_node"] = torch.full((N, cap), -1, dtype=torch.int32, device=dev)
        ctx.generate(configs.c2_raygen(), make_rays_device(t), N)
        t["npoints"][::5] = 3
        ctx.trace(scene, make_rays_device(t), N, mgr.trace_opts(device_ptrs=1, order=order))
        out[order] = t
    a, b = out[A.RB_ORDER_INPUT], out[A.RB_ORDER_COHERENT]
    for k in a:
        assert torch.equal(a[k], b[k]), k
    assert int((a["path_node"][:, 1] >= 0).sum()) > N // 2


def test_host_memory_passed_as_device_pointers_is_an_error(ctx):
    """device_ptrs = 1 with pageable host arrays returns an error (checked on
    the host before any launch) instead of faulting in a kernel."""
    import torch
    mgr = configs.c1_scene()
    scene = ctx.upload(mgr.flat())
    soa = configs.c1_rays(0.0, nside=16).soa()
    with pytest.raises(RbError, match="not GPU memory"):
        ctx.trace(scene, make_rays(soa), len(soa["x"]), mgr.trace_opts(device_ptrs=1))
    with pytest.raises(RbError, match="not GPU memory"):
        ctx.generate(configs.c2_raygen(), make_rays(soa), len(soa["x"]))
    # the same arrays in HBM are accepted
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in soa.items()}
    ctx.trace(scene, make_rays_device(t), len(soa["x"]), mgr.trace_opts(device_ptrs=1))


@pytest.mark.parametrize("n", [0, 1, 257])
def test_empty_and_ragged_batches(ctx, n):
    """n = 0 is a no-op; batches that do not fill a workgroup trace exactly
    as the oracle does, through host and device pointers."""
    import torch
    mgr = configs.c2_scene()
    scene = ctx.upload(mgr.flat())
    soa = O.generate(configs.c2_raygen(), max(n, 1))
    soa = {k: v[:n].copy() for k, v in soa.items()}
    g = copy_soa(soa)
    ctx.trace(scene, make_rays(g), n, mgr.trace_opts())
    o = copy_soa(soa)
    O.trace(mgr.flat(), o, mgr.trace_opts())
    if n:
        compare(g, o)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in soa.items()}
    if n:
        ctx.trace(scene, make_rays_device(t), n, mgr.trace_opts(device_ptrs=1))
        for k in g:
            assert np.array_equal(t[k].cpu().numpy(), g[k]), k
    assert ctx.last_trace_launches() == (1 if n else 0)


def test_auto_order_just_above_threshold(ctx):
    """A batch of 2^20 + 3 random Davies-Cotton rays (auto order sorts it:
    key kernel, sort, trace, copy-back of the staged lines) equals the same
    batch traced in input order."""
    import torch
    N = (1 << 20) + 3
    mgr = configs.c2_scene()
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    out = {}
    for order in (A.RB_ORDER_AUTO, A.RB_ORDER_INPUT):
        t = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
        for f in ("status", "npoints", "last_node"):
            t[f] = torch.empty(N, dtype=torch.int32, device=dev)
        ctx.generate(configs.c2_raygen(), make_rays_device(t), N)
        ctx.trace(scene, make_rays_device(t), N, mgr.trace_opts(device_ptrs=1, order=order))
        assert ctx.last_trace_launches() == (4 if order == A.RB_ORDER_AUTO else 1)
        out[order] = t
    for k in out[A.RB_ORDER_AUTO]:
        assert torch.equal(out[A.RB_ORDER_AUTO][k], out[A.RB_ORDER_INPUT][k]), k

