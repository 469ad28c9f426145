"""Ray containers mirroring ROBAST's ARay / ARayArray / ARayShooter, stored
as structure-of-arrays so 1e7-ray batches move to and from HBM in bulk.

ARay (include/ARay.h:24-68): points (x, y, z, t), unit direction, lambda,
status, node history. ARayArray (src/ARayArray.cxx:41-73): rays binned by
status in input order. ARayShooter restates src/ARayShooter.cxx:122-459.
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
        self.path_xyzt = None      # optional [n, cap, 4]: the points the last trace added
        self.path_node = None      # optional [n, cap]
        self.npoints_in = None     # point counts when the last recorded trace began
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
        """Materialise ray i as an ARay: the points it held when added (an ARay
        passed to Add keeps its whole history; otherwise its first point),
        then the points the last recorded trace added, then its last point."""
        r = ARay(int(self.ids[i]), self.lambda_[i])
        obj = getattr(self, "_objs", {}).get(i)
        if obj is not None:
            r.points, r.nodes = list(obj.points), list(obj.nodes)
        elif self.first is not None:
            r.points = [tuple(self.first[i])]
        else:
            r.points = []
        if self.path_xyzt is not None:
            cap = self.path_xyzt.shape[1]
            base = int(self.npoints_in[i]) if self.npoints_in is not None else 1
            k = min(int(self.npoints[i]) - base, cap)
            r.points += [tuple(p) for p in self.path_xyzt[i, :k]]
            names = self.node_names or []
            r.nodes += [names[j] if 0 <= j < len(names) else None for j in self.path_node[i, :k]]
        if not r.points or r.points[-1] != (self.x[i], self.y[i], self.z[i], self.t[i]):
            r.points.append((self.x[i], self.y[i], self.z[i], self.t[i]))
        r.direction = (self.dx[i], self.dy[i], self.dz[i])
        r.status = int(self.status[i])
        return r

    def soa(self):
        return {f: getattr(self, f) for f in _FIELDS_F + _FIELDS_I}


class ARayShooter:
    """The ray sources of src/ARayShooter.cxx:122-459, with the reference's
    signatures (rot, tr and v may be None or 0 for the reference's null
    pointers). Without `ctx` each returns an ARayArray built on the host;
    with `ctx` (a librbgpu Context) the rays are written into HBM by
    rbgpu_shoot_rays and a DeviceRays batch is returned, ready for
    AOpticsManager.TraceNonSequential. Random patterns draw from numpy's
    generator seeded with `seed` on the host and from Philox streams keyed by
    (seed, ray index) on the device; the reference draws from gRandom, so the
    distributions agree and the individual rays do not."""

    # ---- deterministic patterns
    @staticmethod
    def Rectangle(lam, dx, dy, nx, ny, rot=None, tr=None, v=None, ctx=None):
        """:401-452: nx x ny grid, x_i = i dx / (nx - 1) - dx / 2 (dx / 2 steps for n = 1)."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RECTANGLE, lam, dx, dy, nx, ny, rot, tr, v)
        if dx < 0 or dy < 0 or nx < 1 or ny < 1:
            return ARayArray(0)
        deltax = dx / 2 if nx == 1 else dx / (nx - 1)
        deltay = dy / 2 if ny == 1 else dy / (ny - 1)
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        pts = np.stack([i.ravel() * deltax - dx / 2, j.ravel() * deltay - dy / 2, np.zeros(nx * ny)], axis=1)
        return _host(lam, _place(pts, rot, tr), _turn(_vector(v), rot))

    @staticmethod
    def Square(lam, d, n, rot=None, tr=None, v=None, ctx=None):
        """:455-459."""
        return ARayShooter.Rectangle(lam, d, d, n, n, rot, tr, v, ctx=ctx)

    @staticmethod
    def Circle(lam, rmax, nr, nphi, rot=None, tr=None, v=None, ctx=None):
        """:122-183: a ray at the centre, then rings r = rmax (i + 1) / nr of
        nphi (i + 1) rays at phi = 2 pi / nphi / (i + 1) j."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_CIRCLE, lam, rmax, 0.0, nr, nphi, rot, tr, v)
        if rmax < 0 or nr < 1 or nphi < 1:
            return ARayArray(0)
        xs, ys = [0.0], [0.0]
        for i in range(nr):
            r = rmax * (i + 1) / nr
            j = np.arange(nphi * (i + 1), dtype=float)
            phi = 2 * math.pi / nphi / (i + 1) * j
            xs.append(r * np.cos(phi))
            ys.append(r * np.sin(phi))
        x, y = np.hstack(xs), np.hstack(ys)
        pts = _place(np.stack([x, y, np.zeros(len(x))], axis=1), rot, tr)
        pts[0] = _translation(tr)  # the centre is translated, not rotated (:150-157)
        return _host(lam, pts, _turn(_vector(v), rot))

    # ---- random patterns
    @staticmethod
    def RandomCircle(lam, rmax, n, rot=None, tr=None, v=None, seed=0, ctx=None):
        """:186-237: uniform in the disk of radius rmax."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RANDOM_CIRCLE, lam, rmax, 0.0, 0, 0, rot, tr, v, seed, n)
        if rmax < 0 or n < 1:
            return ARayArray(0)
        rng = np.random.default_rng(seed)
        rho, phi = rmax * np.sqrt(rng.uniform(0, 1, n)), rng.uniform(0, 2 * math.pi, n)
        pts = np.stack([rho * np.cos(phi), rho * np.sin(phi), np.zeros(n)], axis=1)
        return _host(lam, _place(pts, rot, tr), _turn(_vector(v), rot))

    @staticmethod
    def RandomCone(lam, r, d, n, rot=None, tr=None, seed=0, ctx=None):
        """:240-285: from the (translated) origin toward a point uniform in the
        disk of radius r at z = d (rotated, translated)."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RANDOM_CONE, lam, r, d, 0, 0, rot, tr, None, seed, n)
        if n < 1:
            return ARayArray(0)
        rng = np.random.default_rng(seed)
        rho, phi = r * np.sqrt(rng.uniform(0, 1, n)), rng.uniform(0, 2 * math.pi, n)
        goal = _place(np.stack([rho * np.cos(phi), rho * np.sin(phi), np.full(n, float(d))], axis=1), rot, tr)
        start = np.tile(_translation(tr), (n, 1))
        return _host(lam, start, goal - start)

    @staticmethod
    def RandomRectangle(lam, dx, dy, n, rot=None, tr=None, v=None, seed=0, ctx=None):
        """:288-335: uniform in [-dx/2, dx/2] x [-dy/2, dy/2]."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RANDOM_RECTANGLE, lam, dx, dy, 0, 0, rot, tr, v, seed, n)
        if dx < 0 or dy < 0 or n < 1:
            return ARayArray(0)
        rng = np.random.default_rng(seed)
        pts = np.stack([rng.uniform(-dx / 2, dx / 2, n), rng.uniform(-dy / 2, dy / 2, n), np.zeros(n)], axis=1)
        return _host(lam, _place(pts, rot, tr), _turn(_vector(v), rot))

    @staticmethod
    def RandomSquare(lam, d, n, rot=None, tr=None, v=None, seed=0, ctx=None):
        """:394-398."""
        return ARayShooter.RandomRectangle(lam, d, d, n, rot, tr, v, seed=seed, ctx=ctx)

    @staticmethod
    def RandomSphere(lam, n, tr=None, seed=0, ctx=None):
        """:338-356: isotropic directions from the origin (or tr)."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RANDOM_SPHERE, lam, 0.0, 0.0, 0, 0, None, tr, None, seed, n)
        if n < 1:
            return ARayArray(0)
        rng = np.random.default_rng(seed)
        cos_t = rng.uniform(-1, 1, n)
        phi = rng.uniform(0, 2 * math.pi, n)
        sin_t = np.sqrt(1 - cos_t * cos_t)
        o = np.tile(_translation(tr), (n, 1))
        return _host(lam, o, np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1))

    @staticmethod
    def RandomSphericalCone(lam, n, theta, rot=None, tr=None, seed=0, ctx=None):
        """:359-391: cos(angle to +z) uniform in [cos theta, 1] (theta in
        degrees), then rotated; all from the origin (or tr)."""
        if ctx is not None:
            return _device(ctx, A.RB_SHOOT_RANDOM_SPHERICAL_CONE, lam, theta, 0.0, 0, 0, rot, tr, None, seed, n)
        if n < 1:
            return ARayArray(0)
        rng = np.random.default_rng(seed)
        th = np.arccos(rng.uniform(math.cos(theta * math.pi / 180.0), 1, n))
        phi = rng.uniform(0, 2 * math.pi, n)
        d = np.stack([np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), np.cos(th)], axis=1)
        o = np.tile(_translation(tr), (n, 1))
        return _host(lam, o, _turn(d, rot))


def _rotation(rot):
    if rot is None or (isinstance(rot, (int, float)) and rot == 0):
        return None
    return np.asarray(rot.rot, dtype=float)


def _translation(tr):
    if tr is None or (isinstance(tr, (int, float)) and tr == 0):
        return np.zeros(3)
    return np.asarray(tr.tr, dtype=float).reshape(3)


def _vector(v):
    if v is None or (isinstance(v, (int, float)) and v == 0):
        return np.array([0.0, 0.0, 1.0])
    if hasattr(v, "X"):
        return np.array([v.X(), v.Y(), v.Z()], dtype=float)
    return np.asarray(v, dtype=float).reshape(3)


def _rot_rows(R, p):
    """TGeoRotation::LocalToMaster row by row: m_i = l0 R_i0 + l1 R_i1 + l2 R_i2."""
    return np.stack([p[..., 0] * R[i, 0] + p[..., 1] * R[i, 1] + p[..., 2] * R[i, 2] for i in range(3)], axis=-1)


def _place(pts, rot, tr):
    R = _rotation(rot)
    if R is not None:
        pts = _rot_rows(R, pts)
    return pts + _translation(tr)


def _turn(d, rot):
    R = _rotation(rot)
    return d if R is None else _rot_rows(R, d)


def _host(lam, pts, d):
    """new ARay(0, lam, x, y, z, 0, d) per point: SetDirection normalises a
    non-zero vector and leaves a zero one (0, 0, 0) (src/ARay.cxx:28-36, 217-223)."""
    n = len(pts)
    d = np.broadcast_to(np.asarray(d, dtype=float), (n, 3))
    mag = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    u = np.zeros((n, 3))
    ok = mag > 0
    u[ok] = d[ok] / mag[ok, None]
    a = ARayArray(n)
    a.x[:], a.y[:], a.z[:] = pts[:, 0], pts[:, 1], pts[:, 2]
    a.dx[:], a.dy[:], a.dz[:] = u[:, 0], u[:, 1], u[:, 2]
    a.lambda_[:] = lam
    return a


def _device(ctx, kind, lam, a, b, nx, ny, rot, tr, v, seed=0, n=0):
    from . import sources
    return sources.shoot(ctx, sources.shooter_desc(kind, lam, a, b, nx, ny, rot, tr, v, seed), n)
