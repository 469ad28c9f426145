"""Shared helpers for the parity tests."""
import mpmath as mp
import numpy as np

from robast_amd import _abi as A

FIELDS = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node")


def copy_soa(soa):
    return {k: np.array(v, copy=True) for k, v in soa.items()}


def compare(gpu, ora, pos_tol=1e-7, dir_tol=1e-9, t_tol=1e-15, check_nodes=True):
    """Per-ray parity: identical status / npoints / last node, terminal point
    within pos_tol cm (1e-7 cm = 1e-9 m), direction within dir_tol."""
    st = gpu["status"] == ora["status"]
    bad = np.nonzero(~st)[0]
    assert len(bad) == 0, f"{len(bad)} status mismatches, first {bad[:5]}: gpu {gpu['status'][bad[:5]]} oracle {ora['status'][bad[:5]]}"
    assert np.array_equal(gpu["npoints"], ora["npoints"])
    if check_nodes:
        assert np.array_equal(gpu["last_node"], ora["last_node"])
    dp = np.sqrt((gpu["x"] - ora["x"]) ** 2 + (gpu["y"] - ora["y"]) ** 2 + (gpu["z"] - ora["z"]) ** 2)
    dd = np.sqrt((gpu["dx"] - ora["dx"]) ** 2 + (gpu["dy"] - ora["dy"]) ** 2 + (gpu["dz"] - ora["dz"]) ** 2)
    assert dp.max() <= pos_tol, f"max |dpos| {dp.max():.3e} cm"
    assert dd.max() <= dir_tol, f"max |ddir| {dd.max():.3e}"
    fin = np.isfinite(ora["t"]) & (np.abs(ora["t"]) < 1e10)
    assert np.abs(gpu["t"] - ora["t"])[fin].max(initial=0) <= t_tol
    return dp.max()


def counts(status):
    return np.bincount(np.asarray(status), minlength=6)


def c1_exact_chain(flat, src, i, step_back):
    """50-digit construction of ray i of a C1 (paraboloid) batch: the hit on
    the inner surface z = a r^2 + b + sag/2, the normal, the specular
    reflection, the step back along -d1 by `step_back` cm (the navigator's
    move after a reflection, an input here) and the focal-plane (z = f)
    crossing. Returns (end point [x, y, z], arrival time)."""
    from robast_amd import configs
    mp.mp.dps = 50
    para = [k for k in range(len(flat.shapes)) if flat.shapes[k]["type"] == A.RB_SHAPE_PARABOLOID][0]
    off = flat.shapes[para]["param_offset"]
    a = mp.mpf(float(flat.params[off + 3]))
    b = mp.mpf(float(flat.params[off + 4]))
    shift = mp.mpf(configs.C1_R ** 2 / 4.0 / configs.C1_F / 2.0)
    f = mp.mpf(configs.C1_F)
    c = mp.mpf(2.99792458e10)
    o = [mp.mpf(float(src[k][i])) for k in ("x", "y", "z")]
    d = [mp.mpf(float(src[k][i])) for k in ("dx", "dy", "dz")]
    A2 = a * (d[0] ** 2 + d[1] ** 2)
    B2 = 2 * a * (o[0] * d[0] + o[1] * d[1]) - d[2]
    C2 = a * (o[0] ** 2 + o[1] ** 2) + b + shift - o[2]
    sq = mp.sqrt(B2 ** 2 - 4 * A2 * C2)
    qq = -(B2 + (sq if B2 >= 0 else -sq)) / 2
    roots = [C2 / qq] + ([qq / A2] if A2 != 0 else [])
    t = min(r for r in roots if r > 0)
    h = [o[k] + t * d[k] for k in range(3)]
    n = [2 * a * h[0], 2 * a * h[1], mp.mpf(-1)]
    nn = mp.sqrt(sum(v * v for v in n))
    n = [v / nn for v in n]
    cs = sum(d[k] * n[k] for k in range(3))
    d2 = [d[k] - 2 * cs * n[k] for k in range(3)]
    m = mp.sqrt(sum(v * v for v in d2))
    d2 = [v / m for v in d2]
    hb = [h[k] - mp.mpf(step_back) * d[k] for k in range(3)]
    s2 = (f - hb[2]) / d2[2]
    end = [hb[k] + s2 * d2[k] for k in range(3)]
    return end, t / c + (s2 + mp.mpf("1e-10")) / c


def sensitivity(flat, soa, opts, delta=1e-10, trials=2, seed=0):
    """Per-ray conditioning of the trace, measured with the oracle: the rays
    are re-traced `trials` times with their start points moved by `delta` cm
    in random directions. Returns (gain, unstable): gain[i] is the largest
    terminal-point move over delta (how much ray i amplifies a small
    difference along its path; the terminal point of a path that bounces
    between curved surfaces can move by 1e9 delta), unstable[i] is true when
    a move changed the ray's status, point count or last node."""
    from oracle import oracle as O
    base = copy_soa(soa)
    O.trace(flat, base, opts, nthreads=8)
    rng = np.random.default_rng(seed)
    n = len(soa["x"])
    gain = np.zeros(n)
    unstable = np.zeros(n, dtype=bool)
    for _ in range(trials):
        s = copy_soa(soa)
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        s["x"] += delta * u[:, 0]
        s["y"] += delta * u[:, 1]
        s["z"] += delta * u[:, 2]
        O.trace(flat, s, opts, nthreads=8)
        unstable |= (s["status"] != base["status"]) | (s["npoints"] != base["npoints"]) | \
            (s["last_node"] != base["last_node"])
        mv = np.sqrt((s["x"] - base["x"]) ** 2 + (s["y"] - base["y"]) ** 2 + (s["z"] - base["z"]) ** 2)
        gain = np.maximum(gain, mv / delta)
    return gain, unstable


def conditioned_compare(g, o, gain, unstable, eps=1e-12, floor=1e-9):
    """Per-ray parity with bars from each ray's own conditioning: statuses,
    point counts and last nodes identical except on rays that a 1e-10 cm
    input move also changes; terminal points within floor + eps * gain *
    npoints cm, i.e. a rounding-level difference of eps cm at each recorded
    point, amplified by the ray's measured gain. Returns (integer
    mismatches, largest |dpos| / bar over the matched rays)."""
    bad = (g["status"] != o["status"]) | (g["npoints"] != o["npoints"]) | (g["last_node"] != o["last_node"])
    assert np.all(unstable[bad]), f"integer mismatch on well-conditioned rays {np.nonzero(bad & ~unstable)[0][:5]}"
    dp = np.sqrt((g["x"] - o["x"]) ** 2 + (g["y"] - o["y"]) ** 2 + (g["z"] - o["z"]) ** 2)
    # a navigation step that degraded to TGeoShape::Big leaves the point near
    # 1e31 cm: compared relative to its size
    pn = np.sqrt(o["x"] ** 2 + o["y"] ** 2 + o["z"] ** 2)
    bar = floor + eps * np.maximum(gain, 1.0) * o["npoints"] + 1e-12 * np.where(pn > 1e4, pn, 0.0)
    ratio = (dp / bar)[~bad]
    worst = int(np.argmax(dp / bar * ~bad))
    assert ratio.max(initial=0) <= 1.0, \
        f"ray {worst}: |dpos| {dp[worst]:.3e} > bar {bar[worst]:.3e} (gain {gain[worst]:.3e}, {o['npoints'][worst]} points)"
    return int(bad.sum()), float(ratio.max(initial=0))
