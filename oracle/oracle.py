"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU FP64 oracle
(oracle/rb_oracle.cpp). Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from robast_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call([os.path.join(_HERE, "build.sh")])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_trace.argtypes = [C.POINTER(A.rb_scene_desc), C.POINTER(A.rb_trace_opts), C.POINTER(A.rb_rays),
                                   C.c_int64, C.c_int]
        L.oracle_trace.restype = C.c_int
        L.oracle_set_navigator.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p]
        L.oracle_set_navigator.restype = None
        L.oracle_shape_query.argtypes = [C.POINTER(A.rb_scene_desc), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_shape_query.restype = C.c_int
        L.oracle_table_eval.argtypes = [C.POINTER(A.rb_scene_desc), C.c_int, C.c_double]
        L.oracle_table_eval.restype = C.c_double
        L.oracle_lens_index.argtypes = [C.POINTER(A.rb_scene_desc), C.c_int, C.c_double]
        L.oracle_lens_index.restype = C.c_double
        L.oracle_philox.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_philox.restype = None
        L.oracle_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.oracle_uniform.restype = C.c_double
        L.oracle_contains_batch.argtypes = [C.POINTER(A.rb_scene_desc), C.c_int, C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_contains_batch.restype = C.c_int
        L.oracle_generate_rays.argtypes = [C.POINTER(A.rb_raygen_desc), C.POINTER(A.rb_rays), C.c_int64]
        L.oracle_generate_rays.restype = C.c_int
        L.oracle_th2_interpolate.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double]
        L.oracle_th2_interpolate.restype = C.c_double
        L.oracle_shooter_count.argtypes = [C.POINTER(A.rb_shooter_desc), C.c_int64]
        L.oracle_shooter_count.restype = C.c_int64
        L.oracle_shoot_rays.argtypes = [C.POINTER(A.rb_shooter_desc), C.POINTER(A.rb_rays), C.c_int64]
        L.oracle_shoot_rays.restype = C.c_int
        L.oracle_bunches_to_rays.argtypes = [C.POINTER(A.rb_bunches), C.c_int64, C.POINTER(A.rb_bunch_opts),
                                             C.POINTER(A.rb_rays), C.c_int64]
        L.oracle_bunches_to_rays.restype = C.c_int64
        L.oracle_adjoint_dual.argtypes = [C.POINTER(A.rb_scene_desc), C.POINTER(A.rb_trace_opts), C.POINTER(A.rb_rays),
                                          C.c_int64, C.POINTER(A.rb_adjoint_desc), C.c_void_p, C.c_int]
        L.oracle_adjoint_dual.restype = C.c_int64
        L.oracle_trace_counts.argtypes = [C.POINTER(A.rb_scene_desc), C.POINTER(A.rb_trace_opts),
                                          C.POINTER(A.rb_rays), C.c_int64, C.c_void_p, C.c_int]
        L.oracle_trace_counts.restype = C.c_int
        L.oracle_hist2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int, C.c_double,
                                    C.c_double, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        L.oracle_hist2d.restype = C.c_int
        L.oracle_containment.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_void_p]
        L.oracle_containment.restype = None
        _lib = L
    return _lib


def new_soa(n, path_capacity=0):
    soa = {f: np.zeros(n) for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")}
    soa["status"] = np.zeros(n, dtype=np.int32)
    soa["npoints"] = np.ones(n, dtype=np.int32)
    soa["last_node"] = np.full(n, -1, dtype=np.int32)
    if path_capacity:
        soa["path_xyzt"] = np.zeros((n, path_capacity, 4))
        soa["path_node"] = np.full((n, path_capacity), -1, dtype=np.int32)
    return soa


def rays_struct(soa):
    r = A.rb_rays()
    for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node"):
        setattr(r, f, A.ptr(soa[f]))
    if "path_xyzt" in soa:
        r.path_capacity = soa["path_xyzt"].shape[1]
        r.path_xyzt = A.ptr(soa["path_xyzt"])
        r.path_node = A.ptr(soa["path_node"])
    return r


NAVIGATOR_DEFAULTS = (2e-6, 100.0, 1e-8)


def set_navigator(step_back=0.0, push_scale=0.0, bool_push=0.0):
    """Sets the oracle's navigator constants (0 keeps one): the step back
    after a reflection (cm), the scale of the locate push after a crossing
    (push = scale (1 + |T| + step) 1e-10 cm) and the boolean iteration's
    inside-test offset (cm). Returns the constants in force."""
    out = np.zeros(3)
    lib().oracle_set_navigator(step_back, push_scale, bool_push, out.ctypes.data)
    return tuple(float(v) for v in out)


class navigator:
    """Context manager: the oracle traces with other navigator constants
    inside the block and with the defaults after it."""

    def __init__(self, step_back=0.0, push_scale=0.0, bool_push=0.0):
        self.args = (step_back, push_scale, bool_push)

    def __enter__(self):
        return set_navigator(*self.args)

    def __exit__(self, *exc):
        set_navigator(*NAVIGATOR_DEFAULTS)
        return False


def trace(flat, soa, opts, nthreads=1):
    """Trace in place (soa: dict of host arrays)."""
    n = len(soa["x"])
    r = rays_struct(soa)
    rc = lib().oracle_trace(C.byref(flat.desc), C.byref(opts), C.byref(r), n, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_trace failed ({rc})")
    return soa


def adjoint_dual(flat, soa, opts, adj, n_params, nthreads=8):
    """Dual-number loss gradient (forward mode) of the rays in `soa` (not
    modified); adj is an rb_adjoint_desc. Returns (grad[n_params], rays used)."""
    n = len(soa["x"])
    grad = np.zeros(n_params)
    r = rays_struct(soa)
    used = lib().oracle_adjoint_dual(C.byref(flat.desc), C.byref(opts), C.byref(r), n, C.byref(adj),
                                     grad.ctypes.data, nthreads)
    if used < 0:
        raise RuntimeError("oracle_adjoint_dual failed")
    return grad, int(used)


OP_NAMES = ("xform_point", "xform_dir", "dist_box", "dist_tube", "dist_sphere", "dist_paraboloid", "dist_pgon",
            "dist_winston_cap", "asph_seed", "asph_newton", "asph_wall", "winston_side", "in_box", "in_tube",
            "in_sphere", "in_paraboloid", "in_pgon", "in_aspheric", "in_winston", "normal", "reflect", "refract",
            "box_test", "step")


def trace_counts(flat, soa, opts):
    """Single-threaded trace in place that also returns the operation counts
    (dict OP_NAMES -> count) the FLOP figure is built from."""
    n = len(soa["x"])
    c = np.zeros(64, dtype=np.uint64)
    r = rays_struct(soa)
    k = lib().oracle_trace_counts(C.byref(flat.desc), C.byref(opts), C.byref(r), n, c.ctypes.data, 64)
    if k != len(OP_NAMES):
        raise RuntimeError(f"oracle_trace_counts: {k} counters, expected {len(OP_NAMES)}")
    return dict(zip(OP_NAMES, (int(v) for v in c[:k])))


def generate(gen, n):
    soa = new_soa(n)
    r = rays_struct(soa)
    lib().oracle_generate_rays(C.byref(gen), C.byref(r), n)
    return soa


def th2_interpolate(grid, values, x, y):
    """The oracle's TH2::Interpolate over values[ny, nx] (grid: a GRID_DTYPE record)."""
    g = np.zeros(1, dtype=A.GRID_DTYPE)
    g[0] = grid
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib().oracle_th2_interpolate(g.ctypes.data, v.ctypes.data, float(x), float(y))


def shooter_count(desc, n=0):
    return int(lib().oracle_shooter_count(C.byref(desc), n))


def shoot(desc, n=0):
    """oracle_shoot_rays: the ARayShooter restatement; returns the SoA."""
    count = shooter_count(desc, n)
    soa = new_soa(count)
    rc = lib().oracle_shoot_rays(C.byref(desc), C.byref(rays_struct(soa)), count)
    assert rc == 0, rc
    return soa


def bunches_to_rays(bunches, opts):
    """oracle_bunches_to_rays over a CorsikaBunches table; returns the SoA
    (or raises for a bunch above 2^24 photons)."""
    d = bunches.desc()
    total = lib().oracle_bunches_to_rays(C.byref(d), bunches.n, C.byref(opts), None, 0)
    if total < 0:
        raise ValueError("a selected bunch holds more than 2^24 photons")
    soa = new_soa(total)
    got = lib().oracle_bunches_to_rays(C.byref(d), bunches.n, C.byref(opts), C.byref(rays_struct(soa)), total)
    assert got == total
    return soa


def shape_query(flat, shape, p, d, exit_normal=False):
    """(inside, distance in from outside, distance out from inside, normal):
    the normal at the entry crossing, or with exit_normal at the exit one."""
    out = np.zeros(9)
    p = np.ascontiguousarray(p, dtype=float)
    d = np.ascontiguousarray(d, dtype=float)
    lib().oracle_shape_query(C.byref(flat.desc), shape, p.ctypes.data, d.ctypes.data, out.ctypes.data)
    return bool(out[0]), out[1], out[2], (out[6:9] if exit_normal else out[3:6])


def contains_batch(flat, shape, pts):
    pts = np.ascontiguousarray(pts, dtype=float).reshape(-1, 3)
    out = np.zeros(len(pts), dtype=np.uint8)
    lib().oracle_contains_batch(C.byref(flat.desc), shape, len(pts), pts.ctypes.data, out.ctypes.data)
    return out.astype(bool)


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def uniform(seed, ray_id, draw, stream=0):
    return lib().oracle_uniform(seed, ray_id, draw, stream)


def table_eval(flat, table, x):
    return lib().oracle_table_eval(C.byref(flat.desc), table, x)


def lens_index(flat, volume, lam):
    return lib().oracle_lens_index(C.byref(flat.desc), volume, lam)


def hist2d(x, y, status, want, nx, xlo, xhi, ny, ylo, yhi):
    """TH2D(nx, xlo, xhi, ny, ylo, yhi) filled with the (x, y) of the rays with
    status `want`: counts [ny, nx] (uint32) and stats {sum w, x, x^2, y, y^2}."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    status = np.ascontiguousarray(status, dtype=np.int32)
    counts = np.zeros((ny, nx), dtype=np.uint32)
    stats = np.zeros(5)
    rc = lib().oracle_hist2d(x.ctypes.data, y.ctypes.data, status.ctypes.data, len(x), want, nx, xlo, xhi, ny, ylo,
                             yhi, counts.ctypes.data, stats.ctypes.data)
    assert rc == 0, rc
    return counts, stats


def containment(counts, stats, nx, xlo, xhi, ny, ylo, yhi, fraction):
    """AGeoUtil::ContainmentRadius over a filled histogram: (r, x, y)."""
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    out = np.zeros(3)
    lib().oracle_containment(counts.ctypes.data, stats.ctypes.data, nx, xlo, xhi, ny, ylo, yhi, fraction,
                             out.ctypes.data)
    return float(out[0]), float(out[1]), float(out[2])
