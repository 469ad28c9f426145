"""Ray sources that write straight into HBM: ARayShooter's patterns
(src/ARayShooter.cxx:122-459) and the CORSIKA photon-bunch conversion of
ACorsikaIACTFile::GetRayArray (src/ACorsikaIACTFile.cxx:71-131), run by
librbgpu (rbgpu_shoot_rays, rbgpu_bunches_to_rays). The rays stay on the
device as a DeviceRays batch, which AOpticsManager.TraceNonSequential traces
in place; DeviceRays.to_array() brings them back as an ARayArray.
"""
import numpy as np

from . import _abi as A

_F64 = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
_I32 = ("status", "npoints", "last_node")


class DeviceRays:
    """n rays in HBM (torch tensors on the context's device), laid out as
    rb_rays: the device-resident counterpart of ARayArray."""

    def __init__(self, ctx, n):
        import torch
        dev = torch.device("cuda", ctx.device)
        self.ctx = ctx
        self.n = int(n)
        m = max(self.n, 1)
        self.t = {f: torch.empty(m, dtype=torch.float64, device=dev) for f in _F64}
        self.t.update({f: torch.empty(m, dtype=torch.int32, device=dev) for f in _I32})

    def rays(self):
        from ._lib import make_rays_device
        return make_rays_device(self.t)

    def __len__(self):
        return self.n

    def to_array(self):
        """Copy to the host as an ARayArray (status bins in ray order)."""
        from .rays import ARayArray
        arr = ARayArray(self.n)
        for f in _F64 + _I32:
            getattr(arr, f)[:] = self.t[f][:self.n].cpu().numpy()
        return arr


def _matrix(m, what):
    """The 3x3 rotation of a TGeoRotation-like object (None or 0: identity)."""
    if m is None or (isinstance(m, (int, float)) and m == 0):
        return np.eye(3)
    r = np.asarray(m.rot, dtype=float)
    if r.shape != (3, 3):
        raise ValueError(f"{what}: expected a 3x3 rotation")
    return r


def _translation(t):
    if t is None or (isinstance(t, (int, float)) and t == 0):
        return np.zeros(3)
    return np.asarray(t.tr, dtype=float).reshape(3)


def _vector(v):
    """ARayShooter's `v` argument (a TVector3 or a 3-sequence; None or 0: +z)."""
    if v is None or (isinstance(v, (int, float)) and v == 0):
        return np.array([0.0, 0.0, 1.0])
    if hasattr(v, "X"):
        return np.array([v.X(), v.Y(), v.Z()], dtype=float)
    return np.asarray(v, dtype=float).reshape(3)


def shooter_desc(kind, lam, a=0.0, b=0.0, nx=0, ny=0, rot=None, tr=None, v=None, seed=0, id_base=0):
    """rb_shooter_desc for one ARayShooter call (null rot/tr/v as the reference's 0)."""
    g = A.rb_shooter_desc()
    g.kind, g.nx, g.ny = int(kind), int(nx), int(ny)
    g.lambda_, g.a, g.b = float(lam), float(a), float(b)
    g.rot[:] = list(_matrix(rot, "rot").reshape(9))
    g.tr[:] = list(_translation(tr))
    g.dir[:] = list(_vector(v))
    g.seed, g.id_base = int(seed), int(id_base)
    return g


def shooter_count(desc, n=0):
    """Rays the reference's function makes for `desc` (rbgpu_shooter_count)."""
    from ._lib import load
    c = load().rbgpu_shooter_count(desc, int(n))
    if c < 0:
        raise ValueError("bad shooter kind or a pattern above 2^40 rays")
    return int(c)


def shoot(ctx, desc, n=0):
    """Run one shooter descriptor on the device; returns DeviceRays."""
    count = shooter_count(desc, n)
    out = DeviceRays(ctx, count)
    if count:
        ctx.shoot(desc, out.rays(), count)
    return out


class CorsikaBunches:
    """A photon-bunch table as ACorsikaIACTFile's reader fills it (one entry
    per bunch; the TTree branches telNo, arrayNo, x, y, zem, time, cx, cy, cz,
    lambda, photons of src/ACorsikaIACTFile.cxx:330-342, FP32 as there)."""

    def __init__(self, tel_no, array_no, x, y, zem, time, cx, cy, lambda_, photons, cz=None):
        self.tel_no = np.ascontiguousarray(tel_no, dtype=np.int32)
        self.array_no = np.ascontiguousarray(array_no, dtype=np.int32)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        self.x, self.y, self.zem, self.time = f32(x), f32(y), f32(zem), f32(time)
        self.cx, self.cy = f32(cx), f32(cy)
        if cz is None:  # the reader's cz = -1. * sqrt(1. - cx * cx - cy * cy) (:452): float squares, double rest
            cx2, cy2 = (self.cx * self.cx).astype(np.float64), (self.cy * self.cy).astype(np.float64)
            cz = -1.0 * np.sqrt(1.0 - cx2 - cy2)
        self.cz = f32(cz)
        self.lambda_, self.photons = f32(lambda_), f32(photons)
        n = len(self.tel_no)
        for f in A.BUNCH_FIELDS_I + A.BUNCH_FIELDS_F:
            if len(getattr(self, f)) != n:
                raise ValueError(f"CorsikaBunches: {f} has {len(getattr(self, f))} entries, expected {n}")
        self.n = n

    def desc(self):
        d = A.rb_bunches()
        for f in A.BUNCH_FIELDS_I + A.BUNCH_FIELDS_F:
            setattr(d, f, A.ptr(getattr(self, f)))
        return d


def bunch_opts(tel_no, array_no, z, tel_z, refractive_index, lambda_min, lambda_max, seed=0, id_base=0,
               device_ptrs=False):
    o = A.rb_bunch_opts()
    o.tel_no, o.array_no = int(tel_no), int(array_no)
    o.z, o.tel_z, o.refractive_index = float(z), float(tel_z), float(refractive_index)
    o.lambda_min, o.lambda_max = float(lambda_min), float(lambda_max)
    o.seed, o.id_base, o.device_ptrs = int(seed), int(id_base), int(bool(device_ptrs))
    return o


def GetRayArray(ctx, bunches, telNo, arrayNo, z, refractiveIndex, tel_z, min_wavelength, max_photon_bunches=100000,
                seed=0):
    """ACorsikaIACTFile::GetRayArray(telNo, arrayNo, z, refractiveIndex)
    (src/ACorsikaIACTFile.cxx:71-131) on the device. tel_z is
    GetTelescopeZ(telNo) (cm), min_wavelength fMinWavelength (nm). Bunches
    without a CORSIKA wavelength draw 1/lambda uniformly between
    1/min_wavelength and 1/max_photon_bunches, the reference's bound
    (:123-125; fMaxPhotonBunches defaults to 100000). Returns DeviceRays,
    or None for a telescope or array index the reference rejects (:79-82)."""
    if telNo < 0 or arrayNo < 0 or arrayNo >= 100:
        return None
    o = bunch_opts(telNo, arrayNo, z, tel_z, refractiveIndex, min_wavelength, max_photon_bunches, seed)
    d = bunches.desc()
    n = ctx.bunches_to_rays(d, bunches.n, o)
    out = DeviceRays(ctx, n)
    if n:
        ctx.bunches_to_rays(d, bunches.n, o, out.rays(), n)
    return out
