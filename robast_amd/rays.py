"""Ray containers mirroring ROBAST's ARay / ARayArray / ARayShooter, stored
as structure-of-arrays so 1e7-ray batches move to and from HBM in bulk.

ARay (include/ARay.h:24-68): points (x, y, z, t), unit direction, lambda,
status, node history. ARayArray (src/ARayArray.cxx:41-73): rays binned by
status in input order. ARayShooter grids restate src/ARayShooter.cxx:401-459.
"""
import math

import numpy as np

from . import _abi as A

_FIELDS_F = ("x", "y", "z", "t", "dx", "dy", "dz", "lambda_")
_FIELDS_I = ("status", "npoints", "last_node")


class ARay:
    """One ray. Constructor order as ARay(id, lambda, x, y, z, t, dx, dy, dz)
    (src/ARay.cxx:28-36); the direction is normalised like SetDirection."""

    def __init__(self, id=0, lam=0.0, x=0.0, y=0.0, z=0.0, t=0.0, dx=1.0, dy=0.0, dz=0.0):
        self.id = int(id)
        self.lam = float(lam)
        self.points = [(float(x), float(y), float(z), float(t))]
        self.nodes = []
        self.direction = (1.0, 0.0, 0.0)
        self.SetDirection(dx, dy, dz)
        self.status = A.RB_RUN

    def SetDirection(self, dx, dy=None, dz=None):
        if dy is None:
            dx, dy, dz = dx
        mag = math.sqrt(dx * dx + dy * dy + dz * dz)
        if mag > 0:
            self.direction = (dx / mag, dy / mag, dz / mag)

    def GetDirection(self, out=None):
        if out is not None:
            out[0], out[1], out[2] = self.direction
        return self.direction

    def GetLambda(self):
        return self.lam

    def GetNpoints(self):
        return len(self.points)

    def GetLastPoint(self, out=None):
        p = self.points[-1]
        if out is not None:
            for i in range(4):
                out[i] = p[i]
        return p

    def GetFirstPoint(self):
        return self.points[0]

    def GetPoint(self, i):
        return self.points[i]

    def GetNodeHistory(self):
        return list(self.nodes)

    def FindNodeNumberStartWith(self, name):
        for i, n in enumerate(self.nodes):
            if n is not None and n.startswith(name):
                return i
        return -1

    def FindNodeStartWith(self, name):
        i = self.FindNodeNumberStartWith(name)
        return self.nodes[i] if i >= 0 else None

    def IsRunning(self):
        return self.status == A.RB_RUN

    def IsStopped(self):
        return self.status == A.RB_STOP

    def IsExited(self):
        return self.status == A.RB_EXIT

    def IsFocused(self):
        return self.status == A.RB_FOCUS

    def IsSuspended(self):
        return self.status == A.RB_SUSPEND

    def IsAbsorbed(self):
        return self.status == A.RB_ABSORB


class RayBin:
    """One status bin of an ARayArray (a TObjArray of rays in input order)."""

    def __init__(self, array, idx):
        self.array = array
        self.indices = idx

    def GetLast(self):
        return len(self.indices) - 1

    def GetEntries(self):
        return len(self.indices)

    def __len__(self):
        return len(self.indices)

    def At(self, i):
        return self.array.ray(int(self.indices[i]))

    __getitem__ = At

    def last_points(self):
        a = self.array
        i = self.indices
        return np.stack([a.x[i], a.y[i], a.z[i], a.t[i]], axis=1)


class ARayArray:
    """Rays in SoA form with ARayArray's status bins."""

    def __init__(self, n=0):
        self.n = n
        for f in _FIELDS_F:
            setattr(self, f, np.zeros(n))
        self.status = np.zeros(n, dtype=np.int32)
        self.npoints = np.ones(n, dtype=np.int32)
        self.last_node = np.full(n, -1, dtype=np.int32)
        self.ids = np.arange(n, dtype=np.int64)
        self.first = None          # first points, kept for GetFirstPoint
        self.path_xyzt = None      # optional [n, cap, 4]
        self.path_node = None      # optional [n, cap]
        self.node_names = None

    @classmethod
    def from_arrays(cls, x, y, z, dx, dy, dz, lam, t=None):
        a = cls(len(x))
        a.x[:], a.y[:], a.z[:] = x, y, z
        if t is not None:
            a.t[:] = t
        d = np.stack([dx, dy, dz], axis=1).astype(float)
        mag = np.sqrt((d * d).sum(1))
        ok = mag > 0
        d[ok] /= mag[ok, None]
        d[~ok] = (1.0, 0.0, 0.0)
        a.dx[:], a.dy[:], a.dz[:] = d[:, 0], d[:, 1], d[:, 2]
        a.lambda_[:] = lam
        return a

    def Add(self, ray):
        """ARayArray::Add (src/ARayArray.cxx:41-56): append one ray."""
        if ray is None:
            return
        i = self.n
        self._grow(1)
        p = ray.points[-1]
        self.x[i], self.y[i], self.z[i], self.t[i] = p
        self.dx[i], self.dy[i], self.dz[i] = ray.direction
        self.lambda_[i] = ray.lam
        self.status[i] = ray.status
        self.npoints[i] = len(ray.points)
        self.ids[i] = ray.id
        self._objs = getattr(self, "_objs", {})
        self._objs[i] = ray

    def _grow(self, k):
        n = self.n + k
        for f in _FIELDS_F:
            setattr(self, f, np.concatenate([getattr(self, f), np.zeros(k)]))
        self.status = np.concatenate([self.status, np.zeros(k, np.int32)])
        self.npoints = np.concatenate([self.npoints, np.ones(k, np.int32)])
        self.last_node = np.concatenate([self.last_node, np.full(k, -1, np.int32)])
        self.ids = np.concatenate([self.ids, np.zeros(k, np.int64)])
        self.n = n

    def _bin(self, code):
        return RayBin(self, np.nonzero(self.status == code)[0])

    def GetRunning(self):
        return self._bin(A.RB_RUN)

    def GetStopped(self):
        return self._bin(A.RB_STOP)

    def GetExited(self):
        return self._bin(A.RB_EXIT)

    def GetFocused(self):
        return self._bin(A.RB_FOCUS)

    def GetSuspended(self):
        return self._bin(A.RB_SUSPEND)

    def GetAbsorbed(self):
        return self._bin(A.RB_ABSORB)

    def ray(self, i):
        """Materialise ray i as an ARay (points from the path record when kept)."""
        r = ARay(int(self.ids[i]), self.lambda_[i])
        if self.first is not None:
            r.points = [tuple(self.first[i])]
        else:
            r.points = []
        if self.path_xyzt is not None:
            cap = self.path_xyzt.shape[1]
            k = min(int(self.npoints[i]) - 1, cap)
            r.points += [tuple(p) for p in self.path_xyzt[i, :k]]
            names = self.node_names or []
            r.nodes = [names[j] if 0 <= j < len(names) else None for j in self.path_node[i, :k]]
        if not r.points or r.points[-1] != (self.x[i], self.y[i], self.z[i], self.t[i]):
            r.points.append((self.x[i], self.y[i], self.z[i], self.t[i]))
        r.direction = (self.dx[i], self.dy[i], self.dz[i])
        r.status = int(self.status[i])
        return r

    def soa(self):
        return {f: getattr(self, f) for f in _FIELDS_F + _FIELDS_I}


class ARayShooter:
    """Grid sources of src/ARayShooter.cxx:401-459."""

    @staticmethod
    def Rectangle(lam, dx, dy, nx, ny, rot=None, tr=None, v=None):
        if dx < 0 or dy < 0 or nx < 1 or ny < 1:
            return ARayArray(0)
        d = np.array([0.0, 0.0, 1.0]) if v is None else np.asarray(v, dtype=float)
        nd = rot.LocalToMasterVect(d) if rot is not None else d
        deltax = dx / 2 if nx == 1 else dx / (nx - 1)
        deltay = dy / 2 if ny == 1 else dy / (ny - 1)
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        pts = np.stack([i.ravel() * deltax - dx / 2, j.ravel() * deltay - dy / 2, np.zeros(nx * ny)], axis=1)
        if rot is not None:
            pts = pts @ rot.rot.T
        if tr is not None:
            pts = pts @ tr.rot.T + tr.tr
        n = len(pts)
        return ARayArray.from_arrays(pts[:, 0], pts[:, 1], pts[:, 2], np.full(n, nd[0]), np.full(n, nd[1]),
                                     np.full(n, nd[2]), np.full(n, float(lam)))

    @staticmethod
    def Square(lam, d, n, rot=None, tr=None, v=None):
        return ARayShooter.Rectangle(lam, d, d, n, n, rot, tr, v)

    @staticmethod
    def RandomSphere(lam, n, tr=None, seed=0):
        """n rays from the origin (or tr) with isotropic directions
        (src/ARayShooter.cxx RandomSphere; numpy's seeded generator replaces gRandom)."""
        rng = np.random.default_rng(seed)
        cos_t = rng.uniform(-1, 1, n)
        phi = rng.uniform(0, 2 * math.pi, n)
        sin_t = np.sqrt(1 - cos_t * cos_t)
        o = np.zeros((n, 3)) if tr is None else np.tile(tr.tr, (n, 1))
        return ARayArray.from_arrays(o[:, 0], o[:, 1], o[:, 2], sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t,
                                     np.full(n, float(lam)))
