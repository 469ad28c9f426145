"""Loader and thin ctypes wrapper for librbgpu.so (the HIP tracer).

There is no CPU fallback: if the library is missing or no GPU is visible,
every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbgpu.so")

_lib = None

EXPORTS = (
    "rbgpu_ctx_create", "rbgpu_ctx_destroy", "rbgpu_last_error", "rbgpu_scene_upload",
    "rbgpu_scene_free", "rbgpu_trace", "rbgpu_generate_rays", "rbgpu_last_trace_ms",
    "rbgpu_last_trace_launches", "rbgpu_abi_sizes", "rbgpu_adjoint", "rbgpu_last_kernel_ms",
    "rbgpu_containment_radius", "rbgpu_shooter_count", "rbgpu_shoot_rays", "rbgpu_bunches_to_rays",
    "rbgpu_trace_noalloc",
)

# rbgpu_trace_noalloc's largest batch (RB_NOALLOC_MAX_RAYS in include/rbgpu.h)
NOALLOC_MAX_RAYS = 4194304


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"librbgpu.so not built ({LIB_PATH}); run __graft_entry__.build()")
    # torch ships its own libamdhip64 with the same SONAME; loading torch
    # first makes librbgpu bind to that copy, so one HIP runtime serves both
    # (two runtimes in one process leave the second without devices)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    lib.rbgpu_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    lib.rbgpu_ctx_create.restype = C.c_int
    lib.rbgpu_ctx_destroy.argtypes = [C.c_void_p]
    lib.rbgpu_ctx_destroy.restype = None
    lib.rbgpu_last_error.argtypes = [C.c_void_p]
    lib.rbgpu_last_error.restype = C.c_char_p
    lib.rbgpu_scene_upload.argtypes = [C.c_void_p, C.POINTER(A.rb_scene_desc), C.POINTER(C.c_void_p)]
    lib.rbgpu_scene_upload.restype = C.c_int
    lib.rbgpu_scene_free.argtypes = [C.c_void_p]
    lib.rbgpu_scene_free.restype = None
    lib.rbgpu_trace.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(A.rb_trace_opts), C.POINTER(A.rb_rays),
                                C.c_int64]
    lib.rbgpu_trace.restype = C.c_int
    lib.rbgpu_trace_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(A.rb_trace_opts), C.POINTER(A.rb_rays),
                                        C.c_int64]
    lib.rbgpu_trace_noalloc.restype = C.c_int
    lib.rbgpu_generate_rays.argtypes = [C.c_void_p, C.POINTER(A.rb_raygen_desc), C.POINTER(A.rb_rays), C.c_int64]
    lib.rbgpu_generate_rays.restype = C.c_int
    lib.rbgpu_last_trace_ms.argtypes = [C.c_void_p]
    lib.rbgpu_last_trace_ms.restype = C.c_double
    lib.rbgpu_last_kernel_ms.argtypes = [C.c_void_p]
    lib.rbgpu_last_kernel_ms.restype = C.c_double
    lib.rbgpu_last_trace_launches.argtypes = [C.c_void_p]
    lib.rbgpu_last_trace_launches.restype = C.c_int
    lib.rbgpu_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(A.rb_trace_opts), C.POINTER(A.rb_rays), C.c_int64,
                                  C.POINTER(A.rb_adjoint_desc), C.c_void_p]
    lib.rbgpu_adjoint.restype = C.c_int
    lib.rbgpu_containment_radius.argtypes = [C.c_void_p, C.POINTER(A.rb_rays), C.c_int64,
                                             C.POINTER(A.rb_hist2d_desc), C.c_double, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
    lib.rbgpu_containment_radius.restype = C.c_int
    lib.rbgpu_shooter_count.argtypes = [C.POINTER(A.rb_shooter_desc), C.c_int64]
    lib.rbgpu_shooter_count.restype = C.c_int64
    lib.rbgpu_shoot_rays.argtypes = [C.c_void_p, C.POINTER(A.rb_shooter_desc), C.POINTER(A.rb_rays), C.c_int64]
    lib.rbgpu_shoot_rays.restype = C.c_int
    lib.rbgpu_bunches_to_rays.argtypes = [C.c_void_p, C.POINTER(A.rb_bunches), C.c_int64, C.POINTER(A.rb_bunch_opts),
                                          C.POINTER(A.rb_rays), C.c_int64, C.POINTER(C.c_int64)]
    lib.rbgpu_bunches_to_rays.restype = C.c_int
    lib.rbgpu_abi_sizes.argtypes = [C.POINTER(C.c_int32)]
    lib.rbgpu_abi_sizes.restype = C.c_int
    _lib = lib
    return lib


class RbError(RuntimeError):
    pass


class Context:
    """One librbgpu context (a HIP device + stream)."""

    def __init__(self, device=0):
        self.lib = load()
        self.device = int(device)
        h = C.c_void_p()
        rc = self.lib.rbgpu_ctx_create(device, C.byref(h))
        if rc != 0:
            raise RbError(f"rbgpu_ctx_create failed ({rc}); a HIP device is required")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.rbgpu_last_error(self.h)
            raise RbError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def upload(self, flat):
        h = C.c_void_p()
        self._check(self.lib.rbgpu_scene_upload(self.h, C.byref(flat.desc), C.byref(h)), "rbgpu_scene_upload")
        return Scene(self, h)

    def trace(self, scene, rays, n, opts):
        self._check(self.lib.rbgpu_trace(self.h, scene.h, C.byref(opts), C.byref(rays), C.c_int64(n)),
                    "rbgpu_trace")

    def trace_noalloc(self, scene, rays, n, opts):
        """rbgpu_trace_noalloc: rbgpu_trace on device-resident rays with no
        allocation (opts.device_ptrs must be set; n <= NOALLOC_MAX_RAYS)."""
        self._check(self.lib.rbgpu_trace_noalloc(self.h, scene.h, C.byref(opts), C.byref(rays), C.c_int64(n)),
                    "rbgpu_trace_noalloc")

    def generate(self, gen, rays, n):
        self._check(self.lib.rbgpu_generate_rays(self.h, C.byref(gen), C.byref(rays), C.c_int64(n)),
                    "rbgpu_generate_rays")

    def shoot(self, desc, rays, n):
        """rbgpu_shoot_rays: the ARayShooter pattern `desc` into device rays."""
        self._check(self.lib.rbgpu_shoot_rays(self.h, C.byref(desc), C.byref(rays), C.c_int64(n)),
                    "rbgpu_shoot_rays")

    def bunches_to_rays(self, bunches, n_bunches, opts, rays=None, capacity=0):
        """rbgpu_bunches_to_rays: returns the ray count (rays=None: count only)."""
        n = C.c_int64(0)
        self._check(self.lib.rbgpu_bunches_to_rays(self.h, C.byref(bunches), C.c_int64(n_bunches), C.byref(opts),
                                                   C.byref(rays) if rays is not None else None,
                                                   C.c_int64(capacity), C.byref(n)), "rbgpu_bunches_to_rays")
        return int(n.value)

    def adjoint(self, scene, rays, n, opts, adj, grad_ptr):
        self._check(self.lib.rbgpu_adjoint(self.h, scene.h, C.byref(opts), C.byref(rays), C.c_int64(n), C.byref(adj),
                                           C.c_void_p(grad_ptr)), "rbgpu_adjoint")

    def containment_radius(self, rays, n, hist, fraction, with_histogram=False):
        """rbgpu_containment_radius: (r, x, y), plus (counts [ny, nx] uint32,
        stats {sum w, x, x^2, y, y^2}) with_histogram."""
        out = np.zeros(3)
        counts = np.zeros((hist.ny, hist.nx), dtype=np.uint32) if with_histogram else None
        stats = np.zeros(5) if with_histogram else None
        self._check(self.lib.rbgpu_containment_radius(self.h, C.byref(rays), C.c_int64(n), C.byref(hist),
                                                      C.c_double(fraction), A.ptr(out), A.ptr(counts), A.ptr(stats)),
                    "rbgpu_containment_radius")
        r = (float(out[0]), float(out[1]), float(out[2]))
        return (r, counts, stats) if with_histogram else r

    def last_trace_ms(self):
        return self.lib.rbgpu_last_trace_ms(self.h)

    def last_kernel_ms(self):
        return self.lib.rbgpu_last_kernel_ms(self.h)

    def last_trace_launches(self):
        return self.lib.rbgpu_last_trace_launches(self.h)

    def close(self):
        if self.h:
            self.lib.rbgpu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.ctx.lib.rbgpu_scene_free(self.h)
                self.h = None
        except Exception:
            pass


def make_rays(soa, path_xyzt=None, path_node=None):
    """rb_rays over host numpy SoA arrays (dict of field -> array)."""
    r = A.rb_rays()
    for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node"):
        a = soa[f]
        setattr(r, f, A.ptr(a))
    if path_xyzt is not None:
        r.path_capacity = path_xyzt.shape[1]
        r.path_xyzt = A.ptr(path_xyzt)
        r.path_node = A.ptr(path_node)
    return r


def make_rays_device(t):
    """rb_rays over torch device tensors (dict of field -> tensor)."""
    r = A.rb_rays()
    for f in ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_", "status", "npoints", "last_node"):
        setattr(r, f, t[f].data_ptr())
    if "path_xyzt" in t:
        r.path_capacity = t["path_xyzt"].shape[1]
        r.path_xyzt = t["path_xyzt"].data_ptr()
        r.path_node = t["path_node"].data_ptr()
    return r
