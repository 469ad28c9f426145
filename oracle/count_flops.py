"""Algorithmic FLOPs per ray for the tracing configurations (test
infrastructure; writes profiles/algorithmic_flops.md). Run from the
repository root: `python oracle/count_flops.py`.

The instrumented oracle counts, per configuration, every operation a
bounding-box-culled trace performs (the oracle tests a daughter's shape only
when the ray crosses the daughter's box, as the device BVH does). Each
operation has a nominal FP64 cost counted from the written formulas of the
device code (robast_amd/csrc/rb_device.h, rb_trace.h): add/sub/mul/div/sqrt
= 1, fma = 2 (i.e. every add and multiply counted once), comparisons and
fabs free. Bounding-box slab tests run in FP32 and are counted apart.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from robast_amd import configs  # noqa: E402

# nominal FP64 FLOPs per operation (the Davies-Cotton / Winston / aspheric
# parameters: 6-sided Pgon with 2 sections, 8 polynomial terms)
COST_FP64 = {
    "xform_point": 18,      # (p - t): 3, R^T q: 9 mul + 6 add
    "xform_dir": 15,        # R^T d
    "dist_box": 15,         # per-axis offsets, safeties, one to three face solves
    "dist_tube": 25,        # cylinder quadratic (18) + z planes and checks
    "dist_sphere": 90,      # two sphere quadratics (32), theta cone (20), candidate checks (~40)
    "dist_paraboloid": 30,  # quadratic in s (20) + range checks
    "dist_pgon": 85,        # 6 side planes x (2 dots + sub + div) + z planes + checks
    "dist_winston_cap": 10,
    "asph_seed": 36,        # Matsui's spherical seed
    "asph_newton": 72,      # 40 + 4 x 8 polynomial terms (Horner for sag and slope)
    "asph_wall": 73,        # cylinder roots (25) + two sags (2 x 24)
    "winston_side": 50,     # tilted-parabola quadratic of one side + checks
    "in_box": 6,
    "in_tube": 6,
    "in_sphere": 14,
    "in_paraboloid": 6,
    "in_pgon": 20,          # 6 side projections + z
    "in_aspheric": 52,      # radius + two sags
    "in_winston": 40,
    "normal": 15,
    "reflect": 21,          # d.n, d - 2(d.n) n, renormalise
    "refract": 60,          # Snell + unpolarised Fresnel
    "box_test": 0,          # FP32, below
    "step": 6,              # boundary push
}
COST_FP32 = {"box_test": 12}  # slab test: 6 sub + 6 mul


def sample(name, n):
    if name == "c1":
        mgr = configs.c1_scene()
        return mgr, configs.c1_rays(1.0, nside=int(n ** 0.5)).soa()
    mgr = {"c2": configs.c2_scene, "c3": configs.c3_scene, "c4": configs.c4_scene}[name]()
    gen = {"c2": configs.c2_raygen, "c3": configs.c3_raygen}.get(name)
    return mgr, O.generate(gen() if gen else configs.c4_raygen(mgr.c4_dz), n)


def flops_per_ray(name, n=100_000):
    mgr, soa = sample(name, n)
    counts = O.trace_counts(mgr.flat(), soa, mgr.trace_opts())
    n = len(soa["x"])
    f64 = sum(COST_FP64[k] * v for k, v in counts.items()) / n
    f32 = sum(COST_FP32.get(k, 0) * v for k, v in counts.items()) / n
    return f64, f32, {k: v / n for k, v in counts.items()}


def main():
    lines = ["# Algorithmic FLOPs per ray (instrumented oracle, nominal costs)", "",
             "Command: `python oracle/count_flops.py` (writes this file).", "",
             "Counts are per ray, averaged over the sample (C1: 1.0 deg 316x316 grid; C2-C4: the first 1e5",
             "rays of the seeded sources). FP64 FLOPs = sum(count x nominal cost). The oracle tests every",
             "daughter's box (a brute-force prefilter) where the device walks a BVH, so the FP32 box-test",
             "count below is an upper bound of the device's and is reported apart from the FP64 figure.", ""]
    rows = {}
    for name in ("c1", "c2", "c3", "c4"):
        f64, f32, per = flops_per_ray(name, 99856 if name == "c1" else 100_000)
        rows[name] = (f64, f32, per)
        print(name, round(f64, 1), round(f32, 1), flush=True)
    lines.append("| op | FP64 cost | " + " | ".join(f"{c} count/ray" for c in rows) + " |")
    lines.append("|---|---|" + "---|" * len(rows))
    for k in O.OP_NAMES:
        lines.append(f"| {k} | {COST_FP64[k]}{' (FP32 12)' if k == 'box_test' else ''} | "
                     + " | ".join(f"{rows[c][2][k]:.3f}" for c in rows) + " |")
    lines.append("")
    for c, (f64, f32, _) in rows.items():
        lines.append(f"{c}_fp64_flops_per_ray: {f64:.1f}")
        lines.append(f"{c}_fp32_flops_per_ray: {f32:.1f}")
    out = os.path.join(ROOT, "profiles", "algorithmic_flops.md")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
