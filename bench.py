"""Benchmark: Mrays/s to termination for the 85-facet Davies-Cotton
reflector at 1e7 rays on one MI355X (rays resident in HBM), with the trace
kernel's roofline position and the CPU oracle timed beside it.

A step = one rbgpu_trace of the whole 1e7-ray batch (restored from a
pristine HBM copy first). `value` is the FP64 tracer (the drop-in); the same
run also times the mixed FP32/FP64 tracer on the same batch (with its error
against the FP64 result: status flips and focused-hit RMS) and the adjoint
alignment gradient (loss forward + rbgpu_adjoint over 8 angles x 1.25e6 rays,
per-facet tip/tilt/R + camera z; one adjoint launch over all angles gives
loss and gradient). Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s to termination, FP64 vs mixed FP32; adjoint grad evals/s at 1e7 rays"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
BYTES_PER_RAY = 132  # in: 8 doubles + status; out: status, point(4), dir(3), npoints, node


def _profile_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summaries, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _algorithmic_flops(key):
    """FP64 FLOPs per ray from profiles/algorithmic_flops.md (instrumented
    oracle counts x nominal per-operation costs), or None."""
    path = os.path.join(ROOT, "profiles", "algorithmic_flops.md")
    if not os.path.exists(path):
        return None
    m = re.search(key + r"\s*:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def vs_gpu(soa, gpu):
    """Error of the GPU's FP64 results against the oracle on the rays both
    traced (the same seeded rays, keyed by the same ids): status flips and
    the focused-hit RMS and maximum distance."""
    import numpy as np
    m = min(len(soa["x"]), len(gpu["x"]))
    so, sg = soa["status"][:m], gpu["status"][:m]
    both = (so == 3) & (sg == 3)
    d2 = sum((soa[k][:m] - gpu[k][:m]) ** 2 for k in ("x", "y", "z"))[both]
    return {"rays": int(m), "status_flips": int((so != sg).sum()),
            "focused_hit_rms_um": round(float(np.sqrt(d2.mean())) * 1e4 if d2.size else 0.0, 9),
            "focused_hit_max_um": round(float(np.sqrt(d2.max())) * 1e4 if d2.size else 0.0, 9)}


def cpu_baseline(mgr, gen, target_s=10.0, gpu_main=None, gpu_configs=None):
    """The CPU FP64 oracle (a restatement, not the ROOT-built reference) on a
    bounded sample of the same workload, all usable host cores (max 16); the
    GPU's FP64 results on the same rays are checked against it (vs_gpu)."""
    from oracle import oracle as O
    cores = min(len(os.sched_getaffinity(0)), 16)
    flat = mgr.flat()
    probe = O.generate(gen, 500_000)
    t0 = time.perf_counter()
    O.trace(flat, probe, mgr.trace_opts(), nthreads=cores)
    rate = 500_000 / max(time.perf_counter() - t0, 1e-6)
    n = int(min(max(rate * target_s, 500_000), 100_000_000))
    soa = O.generate(gen, n)
    t0 = time.perf_counter()
    O.trace(flat, soa, mgr.trace_opts(), nthreads=cores)
    dt = time.perf_counter() - t0
    out = {"value": round(n / dt / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
           "sample": f"{n} rays of the same seeded Davies-Cotton ray stream (the timed GPU batch is its first "
                     f"1e7), oracle/rb_oracle.cpp, std::thread x {cores}, {dt:.1f} s"}
    if gpu_main is not None:
        out["vs_gpu"] = vs_gpu(soa, gpu_main)
    # the other configurations, on samples of about 2 s each
    from robast_amd import configs
    per = {}
    for name, m, g, k in (("c1", configs.c1_scene(), configs.c1_raygen(1.0, 1000), 1_000_000),
                          ("c3", configs.c3_scene(), configs.c3_raygen(), 400_000),
                          ("c4", None, None, 200_000)):
        if name == "c4":
            m = configs.c4_scene()
            g = configs.c4_raygen(m.c4_dz)
        soa = O.generate(g, k)
        t0 = time.perf_counter()
        O.trace(m.flat(), soa, m.trace_opts(), nthreads=cores)
        per[name] = {"value": round(k / (time.perf_counter() - t0) / 1e6, 4), "rays": k}
        if gpu_configs and name in gpu_configs:
            per[name]["vs_gpu"] = vs_gpu(soa, gpu_configs[name])
    out["configs"] = per
    return out


# rays per configuration in the sweep: C2-C4 at the 1e7 rays their
# workloads are specified at, C1 as a 1000 x 1000 grid (its own workload is
# 4 angles x 99,856 rays)
SWEEP_RAYS = {"c1": 1_000_000, "c2": 10_000_000, "c3": 10_000_000, "c4": 10_000_000, "c2_surfaces": 10_000_000}


def config_sweep(ctx, torch, A, configs, make_rays_device, sizes=None, steps=3, hosts=None):
    """C1-C4, and C2 with rough, coated facets (border-surface kernel form), at
    SWEEP_RAYS rays: FP64 and mixed device Mrays/s, and the mixed form's error
    against the FP64 result (status flips, focused-hit RMS)."""
    import math
    dev = torch.device("cuda")
    out = {}
    for name in ("c1", "c2", "c3", "c4", "c2_surfaces"):
        n = (sizes or SWEEP_RAYS)[name]
        mgr = {"c1": configs.c1_scene, "c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene,
               "c2_surfaces": lambda: configs.c2_scene(surfaces=True)}[name]()
        scene = ctx.upload(mgr.flat())
        src = {f: torch.empty(n, dtype=torch.float64, device=dev) for f in ("x", "y", "z", "t", "dx", "dy", "dz",
                                                                            "lambda_")}
        src.update({f: torch.empty(n, dtype=torch.int32, device=dev) for f in ("status", "npoints", "last_node")})
        if name == "c1":
            side = int(math.isqrt(n))
            g = configs.c1_raygen(1.0, side)
            n_c = side * side
        else:
            g = {"c2": configs.c2_raygen, "c2_surfaces": configs.c2_raygen, "c3": configs.c3_raygen}.get(name)
            g = g() if g else configs.c4_raygen(mgr.c4_dz)
            n_c = n
        ctx.generate(g, make_rays_device(src), n_c)
        res = {}
        for prec in ("fp64", "mixed"):
            mgr.SetPrecision(prec == "mixed")
            opts = mgr.trace_opts(device_ptrs=1)
            best = None
            for _ in range(steps):
                work = {f: v[:n_c].clone() for f, v in src.items()}
                torch.cuda.synchronize()
                ctx.trace(scene, make_rays_device(work), n_c, opts)
                best = ctx.last_trace_ms() if best is None else min(best, ctx.last_trace_ms())
            res[prec] = (best, work)
        a, b = res["fp64"][1], res["mixed"][1]
        if hosts is not None:
            hosts[name] = {k: a[k].cpu().numpy() for k in ("status", "x", "y", "z")}
        both = (a["status"] == A.RB_FOCUS) & (b["status"] == A.RB_FOCUS)
        d2 = (a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2
        out[name] = {"rays": n_c, "fp64_mrays_s": round(n_c / res["fp64"][0] / 1e3, 2),
                     "mixed_mrays_s": round(n_c / res["mixed"][0] / 1e3, 2),
                     "mixed_status_flips": int((a["status"] != b["status"]).sum()),
                     "mixed_focused_hit_rms_um": round(float(torch.sqrt(d2[both].mean())) * 1e4, 6),
                     "focused_fraction": round(float((a["status"] == A.RB_FOCUS).double().mean()), 6)}
        print(f"bench: config {name} {out[name]['fp64_mrays_s']} Mrays/s", file=sys.stderr, flush=True)
    return out


def sources_bench(ctx, torch, A, rays, n, reps=3):
    """Device ray sources writing n rays into resident buffers: ARayShooter's
    RandomCircle (HESS1.C's 7.5 m source plane) and the CORSIKA bunch
    conversion over a synthetic bunch table of about n photons."""
    import numpy as np

    from robast_amd import sources as S
    out_bytes = 8 * 8 + 3 * 4                # x y z t dx dy dz lambda, status npoints last_node
    g = S.shooter_desc(A.RB_SHOOT_RANDOM_CIRCLE, 400e-7, 750.0, 0.0, 0, 0, None, None, (0.0, 0.0, -1.0), seed=7)
    shoot_ms = []
    for _ in range(reps):
        ctx.shoot(g, rays, n)
        shoot_ms.append(ctx.last_kernel_ms())
    rng = np.random.default_rng(5)
    nb = n // 5                              # photons ~ U(0.5, 7.5): 4.5 rays per bunch on average
    cx = rng.uniform(-0.05, 0.05, nb).astype(np.float32)
    cy = rng.uniform(-0.05, 0.05, nb).astype(np.float32)
    host = S.CorsikaBunches(np.zeros(nb), np.zeros(nb), rng.uniform(-1000, 1000, nb), rng.uniform(-1000, 1000, nb),
                            np.full(nb, 1e6), rng.uniform(0, 50, nb), cx, cy, np.zeros(nb),
                            rng.uniform(0.5, 7.5, nb))
    keep = {f: torch.from_numpy(getattr(host, f)).cuda() for f in A.BUNCH_FIELDS_I + A.BUNCH_FIELDS_F}
    d = A.rb_bunches()
    for f, t in keep.items():
        setattr(d, f, t.data_ptr())
    o = S.bunch_opts(0, 0, 3000.0, 0.0, 1.0, 300.0, 700.0, seed=9, device_ptrs=True)
    total = ctx.bunches_to_rays(d, nb, o)
    bunch_rays = min(total, n)
    sel = total if total <= n else None
    conv_ms, expand_ms = [], []
    if sel is not None:
        for _ in range(reps):
            ctx.bunches_to_rays(d, nb, o, rays, n)
            conv_ms.append(ctx.last_trace_ms())
            expand_ms.append(ctx.last_kernel_ms())
    ks = min(shoot_ms)
    res = {"shoot_random_circle": {"rays": n, "kernel_ms": round(ks, 3),
                                   "mrays_s": round(n / (ks * 1e-3) / 1e6, 1),
                                   "write_gb_s": round(n * out_bytes / (ks * 1e-3) / 1e9, 1),
                                   "bytes_per_ray": out_bytes}}
    if conv_ms:
        ke = min(expand_ms)
        res["corsika_bunches"] = {"bunches": nb, "rays": bunch_rays, "device_ms": round(min(conv_ms), 3),
                                  "expand_kernel_ms": round(ke, 3),
                                  "mrays_s": round(bunch_rays / (min(conv_ms) * 1e-3) / 1e6, 1),
                                  "expand_write_gb_s": round(bunch_rays * out_bytes / (ke * 1e-3) / 1e9, 1),
                                  "workload": "photons ~ U(0.5, 7.5) per bunch, wavelengths drawn (lambda = 0)"}
    del keep
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=10_000_000)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--order", choices=("auto", "input", "coherent"), default="auto",
                    help="trace order of the headline batch (auto: coherent for >= 2^20 rays)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(rays would shard by contiguous ranges with the scene replicated per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch

    from robast_amd import _abi as A
    from robast_amd import configs
    from robast_amd._lib import Context, make_rays_device

    N = args.rays
    mgr = configs.c2_scene()
    ctx = Context(0)
    scene = ctx.upload(mgr.flat())
    dev = torch.device("cuda")
    fields_f = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
    fields_i = ("status", "npoints", "last_node")
    src = {f: torch.empty(N, dtype=torch.float64, device=dev) for f in fields_f}
    src.update({f: torch.empty(N, dtype=torch.int32, device=dev) for f in fields_i})
    work = {f: torch.empty_like(v) for f, v in src.items()}
    gen = configs.c2_raygen()
    ctx.generate(gen, make_rays_device(src), N)
    rays = make_rays_device(work)
    # one resident copy of the batch per timed step (the trace updates rays in
    # place): the timed region holds the traces only, not the resets
    batches = [{f: torch.empty_like(v) for f, v in src.items()} for _ in range(args.steps)]
    batch_rays = [make_rays_device(b) for b in batches]

    def run(precision, steps, warmup):
        mgr.SetPrecision(precision == "mixed")
        opts = mgr.trace_opts(device_ptrs=1, order={"auto": A.RB_ORDER_AUTO, "input": A.RB_ORDER_INPUT,
                                                    "coherent": A.RB_ORDER_COHERENT}[args.order])

        for _ in range(warmup):
            for f in src:
                work[f].copy_(src[f])
            torch.cuda.current_stream().synchronize()
            ctx.trace(scene, rays, N, opts)
        for b in batches:
            for f in src:
                b[f].copy_(src[f])
        torch.cuda.synchronize()
        device_ms, kernel_ms = [], []
        t0 = time.perf_counter()
        for k in range(steps):
            ctx.trace(scene, batch_rays[k], N, opts)
            device_ms.append(ctx.last_trace_ms())
            kernel_ms.append(ctx.last_kernel_ms())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for f in src:
            work[f].copy_(batches[-1][f])
        return wall * 1e3 / steps, sum(kernel_ms) / len(kernel_ms), sum(device_ms) / len(device_ms)

    other = "mixed" if args.precision == "fp64" else "fp64"
    ms_o, kms_o, _ = run(other, args.steps, args.warmup)
    res_o = {f: v.clone() for f, v in work.items()}
    ms_per_step, kms, dms = run(args.precision, args.steps, args.warmup)
    value = N / (ms_per_step * 1e-3) / 1e6
    print(f"bench: headline {value:.1f} Mrays/s, trace kernel {kms:.3f} ms", file=sys.stderr, flush=True)
    status = torch.bincount(work["status"], minlength=6).cpu().tolist()
    a, b = (work, res_o) if args.precision == "fp64" else (res_o, work)
    flips = int((a["status"] != b["status"]).sum())
    both = (a["status"] == A.RB_FOCUS) & (b["status"] == A.RB_FOCUS)
    d2 = (a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2
    rms_um = float(torch.sqrt(d2[both].mean())) * 1e4
    ms_fp64, ms_mixed = (ms_per_step, ms_o) if args.precision == "fp64" else (ms_o, ms_per_step)
    kms_fp64, kms_mixed = (kms, kms_o) if args.precision == "fp64" else (kms_o, kms)
    # spot analysis after the trace: containment radius (D80) of the batch's
    # focal-plane hits, histogram filled and searched on the device
    hist = A.rb_hist2d_desc()
    hist.nx = hist.ny = 1000
    hist.xlo, hist.xhi, hist.ylo, hist.yhi = -120.0, 120.0, -120.0, 120.0
    hist.status, hist.device_ptrs = A.RB_FOCUS, 1
    spot_ms, search_ms = [], []
    for _ in range(3):
        d80 = ctx.containment_radius(rays, N, hist, 0.8)
        spot_ms.append(ctx.last_trace_ms())
        search_ms.append(ctx.last_kernel_ms())
    spot = {"d80_radius_cm": round(d80[0], 4), "centre_cm": [round(d80[1], 4), round(d80[2], 4)],
            "device_ms": round(min(spot_ms), 3), "search_ms": round(min(search_ms), 3),
            "fill_and_prefix_ms": round(min(s - k for s, k in zip(spot_ms, search_ms)), 3),
            "workload": "AGeoUtil::ContainmentRadius(0.80) of the focused rays of the batch (all field angles), "
                        "TH2D 1000 x 1000 bins over +-120 cm"}
    gpu_main = None if args.no_cpu_baseline else {k: a[k].cpu().numpy() for k in ("status", "x", "y", "z")}
    sources = sources_bench(ctx, torch, A, rays, N)
    del src, work, res_o, rays, a, b
    torch.cuda.empty_cache()

    adjoint = None
    if not args.no_adjoint:
        from robast_amd import alignment as AL
        model = AL.DaviesCottonAlignment(ctx, rays_per_angle=N // len(AL.C5_ANGLES_DEG))
        p0 = AL.c5_initial_params()
        model.loss_and_gradient(p0)          # warm-up evaluation (sets the reference centroids)
        torch.cuda.synchronize()
        adj_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            model.loss_and_gradient(p0)
            adj_ms.append(model.last_adjoint_ms)
        torch.cuda.synchronize()
        ms_eval = (time.perf_counter() - t0) * 1e3 / args.steps
        print(f"bench: adjoint {ms_eval:.2f} ms per evaluation", file=sys.stderr, flush=True)
        adjoint = {"grad_evals_per_s": round(1e3 / ms_eval, 3), "ms_per_eval": round(ms_eval, 3),
                   "adjoint_device_ms": round(sum(adj_ms) / len(adj_ms), 3),
                   "form": "one adjoint launch over the 8 angles returning loss and gradient (moments form)",
                   "rays": model.n * len(model.angles), "params": model.n_params,
                   "workload": "Davies-Cotton alignment: 170 tip/tilt + 85 radii + camera z, loss = sum of "
                               "spot RMS^2 over 8 angles 0-3.5 deg, unit reflectance, FP64 forward, "
                               "adjoint FP32 per ray with FP64 accumulation"}
        del model
        torch.cuda.empty_cache()
        # the aspheric figure fit: forward traces (spot RMS^2 per angle) and
        # the adjoint through both AGeoAsphericDisk mirrors and the curved
        # focal surface
        from robast_amd import aspheric_fit as AF
        fig = AF.AsphericFigureFit(ctx, rays_per_angle=N // len(AF.FIT_ANGLES_DEG))
        q0 = AF.figure_params()
        fig.loss_and_gradient(q0)            # warm-up evaluation (sets the reference centroids)
        torch.cuda.synchronize()
        fig_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fig.loss_and_gradient(q0)
            fig_ms.append(fig.last_adjoint_ms)
        torch.cuda.synchronize()
        ms_fig = (time.perf_counter() - t0) * 1e3 / args.steps
        print(f"bench: figure adjoint {ms_fig:.2f} ms per evaluation", file=sys.stderr, flush=True)
        adjoint["figure"] = {"grad_evals_per_s": round(1e3 / ms_fig, 3), "ms_per_eval": round(ms_fig, 3),
                             "adjoint_device_ms": round(sum(fig_ms) / len(fig_ms), 3),
                             "rays": fig.n * len(fig.angles), "params": AF.NPARAM,
                             "workload": "Schwarzschild-Couder figure: curvature, kappa and r^2..r^16 "
                                         "coefficients of both aspheric mirrors + focal offset, loss = sum "
                                         "of spot RMS^2 over 4 angles 0-4.5 deg; one moments-form adjoint "
                                         "launch over the 4 angles returning loss and gradient (FP64 chain)"}
        del fig
        torch.cuda.empty_cache()

    gpu_configs = {}
    per_config = None if args.no_configs else config_sweep(ctx, torch, A, configs, make_rays_device,
                                                           hosts=gpu_configs)

    achieved = BYTES_PER_RAY * N / (kms * 1e-3) / 1e9
    traffic = _profile_traffic(f"k_trace_c2_{args.precision}_pmc.txt")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic: seeded Philox4x32-10 rays, uniform 12.5 m disk, 0-4 deg off-axis, 300-600 nm",
        "config": {"workload": "Davies-Cotton 85 hexagonal facets f=16 m, 1e7 rays, stochastic reflectance "
                               "0.80-0.90, camera r=1.2 m + 2.5 m square obscuration",
                   "rays": N, "precision": args.precision, "limit": mgr.limit,
                   "order": {"auto": "coherent (auto for >= 2^20 rays): sorted by first daughter-box entry",
                             "input": "input order", "coherent": "coherent: sorted by first daughter-box entry"}[
                       args.order]},
        "kernel_ms": round(kms, 3),
        "kernel_mrays_s": round(N / (kms * 1e-3) / 1e6, 3),
        "device_ms": round(dms, 3),
        "order_ms": round(dms - kms, 3),
        "status_counts": dict(zip(A.STATUS_NAMES, status)),
        "fp64_mrays_s": round(N / (ms_fp64 * 1e-3) / 1e6, 3),
        "mixed_mrays_s": round(N / (ms_mixed * 1e-3) / 1e6, 3),
        "mixed_speedup": round(ms_fp64 / ms_mixed, 3),
        "mixed_kernel_speedup": round(kms_fp64 / kms_mixed, 3),
        "mixed_vs_fp64": {"status_flips": flips, "focused_hit_rms_um": round(rms_um, 6)},
        "adjoint": adjoint,
        "spot": spot,
        "sources": sources,
        "configs": per_config,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "note": "algorithmic ray-state bytes (132 B/ray) over the k_trace launch time (ordering "
                             "kernels excluded; they are in device_ms)"},
    }
    flops = _algorithmic_flops("c2_fp64_flops_per_ray")
    if flops is not None:
        tf = flops * N / (kms * 1e-3) / 1e12
        out["valu_roofline"] = {"bound": "valu_fp64", "achieved": round(tf, 4), "peak": FP64_PEAK_TFS,
                                "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFS, 6), "flops_per_ray": flops,
                                "note": "algorithmic FP64 FLOPs per ray (profiles/algorithmic_flops.md) over the "
                                        "k_trace launch time; MI355X vector FP64 peak"}
    if not args.no_cpu_baseline:
        mgr.SetPrecision(False)
        out["cpu_baseline"] = cpu_baseline(mgr, gen, gpu_main=gpu_main, gpu_configs=gpu_configs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
