"""AOpticsManager: the drop-in entry point. Holds the geometry tree and the
tracer options, flattens the closed geometry once, and runs
TraceNonSequential on the GPU through librbgpu.

Mirrors include/AOpticsManager.h:28-103: SetTopVolume / CloseGeometry,
SetLimit (ignores n <= 0), DisableFresnelReflection, and the
TraceNonSequential overloads for ARay and ARayArray.
"""
import numpy as np

from . import _abi as A
from . import geometry as G
from .flatten import flatten
from .rays import ARay, ARayArray


class AOpticsManager:
    def __init__(self, name="", title=""):
        G._reset_registry()
        self.name, self.title = name, title
        self.top = None
        self.limit = 100                    # src/AOpticsManager.cxx:29
        self.disable_fresnel = False        # include/AOpticsManager.h:73-75
        self.precision = A.RB_PRECISION_FP64
        self.seed = 0
        self.max_threads = 1
        self.record_path = False
        self.device = 0
        self._flat = None
        self._ctx = None
        self._scene = None

    # ---- construction (TGeoManager API subset)
    def SetTopVolume(self, vol):
        self.top = vol
        self._flat = None

    def GetTopVolume(self):
        return self.top

    def CloseGeometry(self):
        if self.top is None:
            raise RuntimeError("CloseGeometry: no top volume")
        self._flat = flatten(self.top)
        self._scene = None
        return self._flat

    def flat_for(self, lam_min, lam_max):
        """The flattened scene for rays with wavelengths in [lam_min, lam_max]
        (cm). Coatings that were not precalculated, which the reference
        evaluates exactly at any wavelength, are read on the device from a
        table; when the batch reaches outside the band that table covers,
        the scene is flattened again with the band widened to it."""
        from .multilayer import grid_covers
        flat = self.flat()
        g = flat.coating_grid
        if g is not None and not grid_covers(g, lam_min, lam_max):
            lo = min(lam_min, g[1] + 0.5 * (g[2] - g[1]) / g[0])
            hi = max(lam_max, g[2] - 0.5 * (g[2] - g[1]) / g[0])
            self._flat = flatten(self.top, lambda_range=(lo, hi))
            self._scene = None
        return self._flat

    def SetNsegments(self, n):
        pass

    def SetMaxThreads(self, n):
        self.max_threads = int(n)

    def GetMaxThreads(self):
        return self.max_threads

    def SetMultiThread(self, flag=True):
        pass

    # ---- tracer options
    def SetLimit(self, n):
        if n > 0:
            self.limit = int(n)

    def DisableFresnelReflection(self, disable):
        self.disable_fresnel = bool(disable)

    def SetPrecision(self, mixed):
        self.precision = A.RB_PRECISION_MIXED if mixed else A.RB_PRECISION_FP64

    def SetSeed(self, seed):
        self.seed = int(seed)

    def SetRecordPath(self, flag=True):
        self.record_path = bool(flag)

    def flat(self):
        if self._flat is None:
            self.CloseGeometry()
        return self._flat

    def trace_opts(self, device_ptrs=0, ray_id_base=0, order=A.RB_ORDER_AUTO, schedule=A.RB_SCHEDULE_AUTO):
        o = A.rb_trace_opts()
        o.order = order
        o.schedule = schedule
        o.limit = self.limit
        o.fresnel = 0 if self.disable_fresnel else 1
        o.precision = self.precision
        o.device_ptrs = device_ptrs
        o.seed = self.seed
        o.ray_id_base = ray_id_base
        return o

    # ---- GPU plumbing
    def context(self):
        if self._ctx is None:
            from ._lib import Context
            self._ctx = Context(self.device)
        return self._ctx

    def gpu_scene(self):
        if self._scene is None:
            self._scene = self.context().upload(self.flat())
        return self._scene

    # ---- TraceNonSequential overloads
    def TraceNonSequential(self, obj):
        if obj is None:
            return
        if isinstance(obj, ARay):
            arr = ARayArray()
            arr.Add(obj)
            self._trace_array(arr, record=True)
            r = arr.ray(0)
            obj.points = r.points
            obj.nodes = r.nodes
            obj.direction = r.direction
            obj.status = r.status
            return
        if isinstance(obj, ARayArray):
            self._trace_array(obj, record=self.record_path)
            return
        from .sources import DeviceRays
        if isinstance(obj, DeviceRays):  # rays already in HBM: traced in place
            if obj.n:
                if self.flat().coating_grid is not None:
                    lam = obj.t["lambda_"][:obj.n]
                    self.flat_for(float(lam.min()), float(lam.max()))
                self.context().trace(self.gpu_scene(), obj.rays(), obj.n, self.trace_opts(device_ptrs=1))
            return
        raise TypeError(f"TraceNonSequential: unsupported argument {type(obj).__name__}")

    def _trace_array(self, arr, record):
        from ._lib import make_rays
        n = arr.n
        if n == 0:
            return
        if arr.first is None:
            arr.first = np.stack([arr.x, arr.y, arr.z, arr.t], axis=1).copy()
        flat = self.flat_for(float(arr.lambda_.min()), float(arr.lambda_.max()))
        path_xyzt = path_node = None
        if record:
            cap = self.limit
            path_xyzt = np.zeros((n, cap, 4))
            path_node = np.full((n, cap), -1, dtype=np.int32)
        npoints_in = arr.npoints.copy()
        rays = make_rays(arr.soa(), path_xyzt, path_node)
        ctx = self.context()
        ctx.trace(self.gpu_scene(), rays, n, self.trace_opts())
        if record:
            arr.path_xyzt, arr.path_node, arr.npoints_in = path_xyzt, path_node, npoints_in
        arr.node_names = flat.node_names
