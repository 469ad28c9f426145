"""Host-side mirror of the ROOT/ROBAST geometry-construction API used to
describe a scene for the tracer: units, TGeo matrices, primitive and
composite shapes, and ROBAST's optical volume classes.

Only construction and the per-shape parameters the tracer needs live here;
ray tracing itself runs in librbgpu (robast_amd/csrc). Names, argument
order and units follow the reference so macros such as
tutorials/SimpleParabolicTelescope.C read the same in Python.
"""
import math
import re

import numpy as np

from . import _abi as A

# --------------------------------------------------------------------- units
# include/AOpticsManager.h:59-71 (computed the same way, so rounding matches)
cm = 1.0
m = 1e2 * cm
km = 1e3 * m
mm = 1e-3 * m
um = 1e-6 * m
nm = 1e-9 * m
inch = 2.54 * cm
s = 1.0
ms = 1e-3 * s
us = 1e-6 * s
ns = 1e-9 * s
deg = math.pi / 180.0
rad = 1.0


class _Registry:
    """Named shapes and registered matrices, as gGeoManager holds them for
    TGeoCompositeShape expressions."""

    def __init__(self):
        self.shapes = {}
        self.matrices = {}


_registry = _Registry()


def _reset_registry():
    global _registry
    _registry = _Registry()


# ------------------------------------------------------------------ matrices
class TGeoMatrix:
    """Affine transform master = R local + T (TGeoMatrix::LocalToMaster)."""

    def __init__(self, name=""):
        self.name = name
        self.rot = np.eye(3)
        self.tr = np.zeros(3)

    def RegisterYourself(self):
        if self.name:
            _registry.matrices[self.name] = self
        return self

    def LocalToMaster(self, local, master=None):
        out = self.rot @ np.asarray(local, dtype=float) + self.tr
        if master is not None:
            master[:] = out
        return out

    def LocalToMasterVect(self, local, master=None):
        out = self.rot @ np.asarray(local, dtype=float)
        if master is not None:
            master[:] = out
        return out

    def MasterToLocal(self, master):
        return self.rot.T @ (np.asarray(master, dtype=float) - self.tr)

    def GetTranslation(self):
        return self.tr.copy()

    def GetRotationMatrix(self):
        return self.rot.reshape(9).copy()

    def compose(self, other):
        """self * other: apply `other` first (local of other -> local of self)."""
        out = TGeoHMatrix()
        out.rot = self.rot @ other.rot
        out.tr = self.rot @ other.tr + self.tr
        return out

    def as_record(self):
        return np.concatenate([self.rot.reshape(9), self.tr])


class TGeoHMatrix(TGeoMatrix):
    pass


class TGeoTranslation(TGeoMatrix):
    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        if len(args) == 3:
            self.tr = np.array(args, dtype=float)


def _euler(phi, theta, psi):
    """TGeoRotation::SetAngles (Euler angles in degrees, z-x-z)."""
    d = math.pi / 180.0
    sph, cph = math.sin(d * phi), math.cos(d * phi)
    sth, cth = math.sin(d * theta), math.cos(d * theta)
    sps, cps = math.sin(d * psi), math.cos(d * psi)
    return np.array([
        [cps * cph - cth * sph * sps, -sps * cph - cth * sph * cps, sth * sph],
        [cps * sph + cth * cph * sps, -sps * sph + cth * cph * cps, -sth * cph],
        [sps * sth, cps * sth, cth],
    ])


class TGeoRotation(TGeoMatrix):
    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        if len(args) == 3:
            self.SetAngles(*args)

    def SetAngles(self, phi, theta, psi):
        self.rot = _euler(phi, theta, psi)

    def RotateZ(self, angle):
        self.rot = _euler(angle, 0, 0) @ self.rot

    def MultiplyBy(self, rot, after=True):
        """after=True: this = this * rot; after=False: this = rot * this."""
        if after:
            self.rot = self.rot @ rot.rot
        else:
            self.rot = rot.rot @ self.rot


class TGeoCombiTrans(TGeoMatrix):
    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        if len(args) == 2 and isinstance(args[0], TGeoMatrix):
            self.tr = args[0].tr.copy()
            self.rot = args[1].rot.copy() if args[1] is not None else np.eye(3)
        elif len(args) == 4:
            self.tr = np.array(args[:3], dtype=float)
            self.rot = args[3].rot.copy() if args[3] is not None else np.eye(3)


IDENTITY = TGeoHMatrix("identity")


# -------------------------------------------------------------------- shapes
class TGeoShape:
    shape_type = 0

    def __init__(self, name=""):
        self.name = name
        if name:
            _registry.shapes[name] = self

    def params(self):
        raise NotImplementedError

    def local_bbox(self):
        """(origin[3], half[3]) of the shape's local bounding box."""
        raise NotImplementedError

    def GetName(self):
        return self.name


class TGeoBBox(TGeoShape):
    shape_type = A.RB_SHAPE_BOX

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        self.dx, self.dy, self.dz = (float(v) for v in args[:3])
        self.origin = np.array(args[3], dtype=float) if len(args) > 3 else np.zeros(3)

    def params(self):
        return [self.dx, self.dy, self.dz, *self.origin]

    def local_bbox(self):
        return self.origin.copy(), np.array([self.dx, self.dy, self.dz])


class TGeoTube(TGeoShape):
    shape_type = A.RB_SHAPE_TUBE

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        self.rmin, self.rmax, self.dz = (float(v) for v in args[:3])

    def params(self):
        return [self.rmin, self.rmax, self.dz]

    def local_bbox(self):
        return np.zeros(3), np.array([self.rmax, self.rmax, self.dz])


class TGeoSphere(TGeoShape):
    shape_type = A.RB_SHAPE_SPHERE

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        vals = [float(v) for v in args] + [0.0, 180.0, 0.0, 360.0][len(args) - 2:]
        self.rmin, self.rmax, self.theta1, self.theta2, self.phi1, self.phi2 = vals[:6]
        # the phi range [phi1, phi2], going round once when phi2 <= phi1
        self.dphi = self.phi2 - self.phi1
        if self.dphi <= 0:
            self.dphi += 360.0

    def params(self):
        t1, t2 = math.radians(self.theta1), math.radians(self.theta2)
        f1, f2 = math.radians(self.phi1), math.radians(self.phi1 + self.dphi)
        return [self.rmin, self.rmax, self.theta1, self.theta2,
                math.cos(t1), math.cos(t2), math.sin(t1), math.sin(t2),
                self.phi1, self.dphi, math.cos(f1), math.sin(f1), math.cos(f2), math.sin(f2)]

    def local_bbox(self):
        r = self.rmax
        # z range of the theta band
        c1, c2 = math.cos(math.radians(self.theta1)), math.cos(math.radians(self.theta2))
        zs = [r * c1, r * c2, self.rmin * c1, self.rmin * c2]
        zmax = r if self.theta1 <= 0 else max(zs)
        zmin = -r if self.theta2 >= 180 else min(zs)
        # radial extent
        if self.theta1 <= 90 <= self.theta2:
            rho = r
        else:
            rho = r * max(math.sin(math.radians(self.theta1)), math.sin(math.radians(self.theta2)))
        return (np.array([0, 0, 0.5 * (zmax + zmin)]),
                np.array([rho, rho, 0.5 * (zmax - zmin)]))


class TGeoParaboloid(TGeoShape):
    shape_type = A.RB_SHAPE_PARABOLOID

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        self.rlo, self.rhi, self.dz = (float(v) for v in args[:3])
        d = self.rhi * self.rhi - self.rlo * self.rlo
        # TGeoParaboloid::SetParaboloidDimensions: -dz = a rlo^2 + b, dz = a rhi^2 + b
        self.a = 2.0 * self.dz / d
        self.b = -self.dz * (self.rlo * self.rlo + self.rhi * self.rhi) / d

    def params(self):
        return [self.rlo, self.rhi, self.dz, self.a, self.b]

    def local_bbox(self):
        r = max(self.rlo, self.rhi)
        return np.zeros(3), np.array([r, r, self.dz])


class TGeoPcon(TGeoShape):
    shape_type = A.RB_SHAPE_PCON

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        self.phi1, self.dphi, self.nz = float(args[0]), float(args[1]), int(args[2])
        self.z = [0.0] * self.nz
        self.rmin = [0.0] * self.nz
        self.rmax = [0.0] * self.nz

    def DefineSection(self, i, z, rmin, rmax):
        """TGeoPcon::DefineSection: once the last section is defined, sections
        given in decreasing z are put in increasing order."""
        self.z[i], self.rmin[i], self.rmax[i] = float(z), float(rmin), float(rmax)
        if i == self.nz - 1 and self.z[0] > self.z[i]:
            self.z.reverse()
            self.rmin.reverse()
            self.rmax.reverse()

    def _check(self):
        for i in range(1, self.nz):
            if self.z[i] < self.z[i - 1]:
                raise ValueError(f"{type(self).__name__} {self.GetName()}: sections must not decrease in z")

    def _wedge(self):
        """{cos sin} of phi1 and of phi1 + dphi (the phi half-planes)."""
        a, b = math.radians(self.phi1), math.radians(self.phi1 + self.dphi)
        return [math.cos(a), math.sin(a), math.cos(b), math.sin(b)]

    def params(self):
        self._check()
        p = [self.phi1, self.dphi, self.nz]
        for i in range(self.nz):
            p += [self.z[i], self.rmin[i], self.rmax[i]]
        return p + self._wedge()

    def local_bbox(self):
        r = max(self.rmax)
        z0, z1 = min(self.z), max(self.z)
        return np.array([0, 0, 0.5 * (z0 + z1)]), np.array([r, r, 0.5 * (z1 - z0)])


class TGeoPgon(TGeoPcon):
    shape_type = A.RB_SHAPE_PGON

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        TGeoShape.__init__(self, name)
        self.phi1, self.dphi = float(args[0]), float(args[1])
        self.nedges, self.nz = int(args[2]), int(args[3])
        self.z = [0.0] * self.nz
        self.rmin = [0.0] * self.nz
        self.rmax = [0.0] * self.nz

    def params(self):
        self._check()
        p = [self.phi1, self.dphi, float(self.nedges), float(self.nz)]
        for i in range(self.nz):
            p += [self.z[i], self.rmin[i], self.rmax[i]]
        # outward normal (cos, sin) of each side: the middle of its phi sector
        div = self.dphi / self.nedges
        for k in range(self.nedges):
            ph = math.radians(self.phi1 + div * (k + 0.5))
            p += [math.cos(ph), math.sin(ph)]
        return p + self._wedge()

    def local_bbox(self):
        # the polygon's corners lie at rmax / cos(half a sector) from the axis
        r = max(self.rmax) / math.cos(math.radians(min(self.dphi, 360.0) / self.nedges / 2))
        z0, z1 = min(self.z), max(self.z)
        return np.array([0, 0, 0.5 * (z0 + z1)]), np.array([r, r, 0.5 * (z1 - z0)])


class AGeoAsphericDisk(TGeoShape):
    """Disk bounded by two aspheric surfaces z = z_i + c r^2 / (1 + sqrt(1 -
    (1+K) c^2 r^2)) + sum_k K_k r^(2k) and cylinders at rmin / rmax
    (src/AGeoAsphericDisk.cxx; SetAsphDimensions orders z1 < z2 and rmin < rmax)."""
    shape_type = A.RB_SHAPE_ASPHERIC

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        z1, c1, z2, c2, rmax = (float(v) for v in args[:5])
        rmin = float(args[5]) if len(args) > 5 else 0.0
        if z1 < z2:
            self.z1, self.z2, self.c1, self.c2 = z1, z2, c1, c2
        else:
            self.z1, self.z2, self.c1, self.c2 = z2, z1, c2, c1
        rmax, rmin = abs(rmax), abs(rmin)
        self.rmax, self.rmin = (rmax, rmin) if rmax > rmin else (rmin, rmax)
        self.conic1 = self.conic2 = 0.0
        self.k1, self.k2 = [], []

    def SetConicConstants(self, conic1, conic2):
        self.conic1, self.conic2 = float(conic1), float(conic2)

    def SetPolynomials(self, n1, k1, n2, k2):
        self.k1 = [float(k1[i]) for i in range(n1)] if n1 > 0 else []
        self.k2 = [float(k2[i]) for i in range(n2)] if n2 > 0 else []

    def _sag(self, which, r):
        z, c, kap, ks = ((self.z1, self.c1, self.conic1 + 1, self.k1) if which == 1 else
                         (self.z2, self.c2, self.conic2 + 1, self.k2))
        p = r * r * c * c * kap
        if 1 - p < 0:
            return None
        v = z + r * r * c / (1 + math.sqrt(1 - p))
        for i, k in enumerate(ks):
            v += k * r ** (2 * (i + 1))
        return v

    def CalcF1(self, r):
        return self._sag(1, r)

    def CalcF2(self, r):
        return self._sag(2, r)

    def params(self):
        # the K_k r^2k terms again as Kn_k (r / rmax)^2k, Kn_k = K_k rmax^2k, so
        # single precision never forms r^16 (1e42 cm^16 overflows FP32)
        R = self.rmax if self.rmax > 0 else 1.0
        kn1 = [k * R ** (2 * (i + 1)) for i, k in enumerate(self.k1)]
        kn2 = [k * R ** (2 * (i + 1)) for i, k in enumerate(self.k2)]
        return [self.z1, self.c1, self.z2, self.c2, self.rmax, self.rmin, self.conic1 + 1, self.conic2 + 1,
                float(len(self.k1)), float(len(self.k2)), *self.k1, *self.k2, *kn1, *kn2]

    def _sag_samples(self, which, rs):
        """_sag over an array of radii, the points where it is undefined
        dropped."""
        z, c, kap, ks = ((self.z1, self.c1, self.conic1 + 1, self.k1) if which == 1 else
                         (self.z2, self.c2, self.conic2 + 1, self.k2))
        rs = rs[1 - rs * rs * c * c * kap >= 0]
        v = z + rs * rs * c / (1 + np.sqrt(1 - rs * rs * c * c * kap))
        for i, k in enumerate(ks):
            v = v + k * rs ** (2 * (i + 1))
        return v

    def local_bbox(self):
        rs = np.linspace(self.rmin, self.rmax, 4001)
        f1 = self._sag_samples(1, rs)
        f2 = self._sag_samples(2, rs)
        zmin = float(f1.min()) if f1.size else self.z1
        zmax = float(f2.max()) if f2.size else self.z2
        pad = 1e-3 * (zmax - zmin) + 0.1
        zmin, zmax = zmin - pad, zmax + pad
        return np.array([0, 0, 0.5 * (zmin + zmax)]), np.array([self.rmax, self.rmax, 0.5 * (zmax - zmin)])


class AGeoWinstonConePoly(TGeoShape):
    """Polygonal Winston cone (src/AGeoWinstonConePoly.cxx:356-386): r1, r2
    are the entrance/exit apothems, n >= 3 sides; theta = asin(r2/r1),
    half length dz = (r1 + r2) / tan(theta) / 2, parabola focal length
    f = r2 (1 + sin theta). The wide end is at z = +dz."""
    shape_type = A.RB_SHAPE_WINSTON

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        r1, r2, n = abs(float(args[0])), abs(float(args[1])), int(args[2])
        self.r1, self.r2 = (r1, r2) if r1 > r2 else (r2, r1)
        self.n = n if n >= 3 else 3
        self.theta = math.asin(self.r2 / self.r1)
        self.dz = (self.r1 + self.r2) / math.tan(self.theta) / 2.0
        self.f = self.r2 * (1 + math.sin(self.theta))

    def GetDZ(self):
        return self.dz

    def params(self):
        p = [self.r1, self.r2, float(self.n), self.theta, self.dz, self.f,
             math.cos(self.theta), math.sin(self.theta), math.tan(math.pi / self.n)]
        for i in range(self.n):
            a = i * 2 * math.pi / self.n
            p += [math.cos(a), math.sin(a)]
        return p

    def local_bbox(self):
        r = self.r1 / math.cos(math.pi / self.n)
        return np.zeros(3), np.array([r, r, self.dz])


class AGeoWinstonCone2D(TGeoShape):
    """Two-dimensional Winston cone, a trough (src/AGeoWinstonCone2D.cxx:551-567):
    r1, r2 are the half widths of the entrance and exit apertures in x, and
    y the half length of the trough; theta = asin(r2/r1), half length
    dz = (r1 + r2) / tan(theta) / 2, focal length f = r2 (1 + sin theta).
    The wide end is at z = +dz."""
    shape_type = A.RB_SHAPE_WINSTON2D

    def __init__(self, *args):
        name = ""
        if args and isinstance(args[0], str):
            name, args = args[0], args[1:]
        super().__init__(name)
        r1, r2, y = abs(float(args[0])), abs(float(args[1])), abs(float(args[2]))
        self.r1, self.r2 = (r1, r2) if r1 > r2 else (r2, r1)
        self.dy = y
        self.theta = math.asin(self.r2 / self.r1)
        self.dz = (self.r1 + self.r2) / math.tan(self.theta) / 2.0
        self.f = self.r2 * (1 + math.sin(self.theta))

    def GetDZ(self):
        return self.dz

    def GetDY(self):
        return self.dy

    def params(self):
        # the polygonal cone's layout with two sides, at phi = 0 and pi, then dy
        return [self.r1, self.r2, 2.0, self.theta, self.dz, self.f, math.cos(self.theta), math.sin(self.theta), 0.0,
                1.0, 0.0, -1.0, 0.0, self.dy]

    def local_bbox(self):
        return np.zeros(3), np.array([self.r1, self.dy, self.dz])


class TGeoBoolNode:
    def __init__(self, op, left, right, lmat, rmat):
        self.op, self.left, self.right = op, left, right
        self.lmat = lmat if lmat is not None else IDENTITY
        self.rmat = rmat if rmat is not None else IDENTITY


def _box_corners(origin, half, mat):
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=float)
    pts = origin + corners * half
    return (mat.rot @ pts.T).T + mat.tr


class _BoolShape(TGeoShape):
    """One boolean node of a composite expression."""
    shape_type = A.RB_SHAPE_COMPOSITE

    def __init__(self, node, name=""):
        super().__init__(name)
        self.node = node

    def params(self):
        n = self.node
        return [float(n.op), None, None, *n.lmat.as_record(), *n.rmat.as_record()]

    def children(self):
        return self.node.left, self.node.right

    def local_bbox(self):
        n = self.node
        la = _box_corners(*n.left.local_bbox(), n.lmat)
        ra = _box_corners(*n.right.local_bbox(), n.rmat)
        lo_l, hi_l = la.min(0), la.max(0)
        lo_r, hi_r = ra.min(0), ra.max(0)
        if n.op == A.RB_OP_UNION:
            lo, hi = np.minimum(lo_l, lo_r), np.maximum(hi_l, hi_r)
        elif n.op == A.RB_OP_INTERSECTION:
            lo, hi = np.maximum(lo_l, lo_r), np.minimum(hi_l, hi_r)
        else:
            lo, hi = lo_l, hi_l
        return 0.5 * (lo + hi), 0.5 * (hi - lo)


_TOKEN = re.compile(r"\s*(?:([A-Za-z_][A-Za-z0-9_]*)(?::([A-Za-z_][A-Za-z0-9_]*))?|([-+*()]))")


def _parse_expr(expr):
    """Parse a TGeoCompositeShape expression ("a:t1 - b:t2", "+", "*",
    parentheses; * binds tighter than + and -, all left-associative)."""
    toks = []
    pos = 0
    expr = expr.strip()
    while pos < len(expr):
        mt = _TOKEN.match(expr, pos)
        if not mt or mt.end() == pos:
            raise ValueError(f"cannot parse composite expression at: {expr[pos:]!r}")
        pos = mt.end()
        if mt.group(3):
            toks.append(("op", mt.group(3)))
        else:
            toks.append(("id", mt.group(1), mt.group(2)))
    it = [0]

    def peek():
        return toks[it[0]] if it[0] < len(toks) else None

    def take():
        t = toks[it[0]]
        it[0] += 1
        return t

    def primary():
        t = take()
        if t[0] == "op" and t[1] == "(":
            v = additive()
            if take() != ("op", ")"):
                raise ValueError("missing ')' in composite expression")
            return v
        if t[0] != "id":
            raise ValueError(f"unexpected token {t}")
        shape = _registry.shapes.get(t[1])
        if shape is None:
            raise ValueError(f"unknown shape {t[1]!r} in composite expression")
        mat = IDENTITY
        if t[2]:
            mat = _registry.matrices.get(t[2])
            if mat is None:
                raise ValueError(f"unknown matrix {t[2]!r} (call RegisterYourself())")
        return (shape, mat)

    def term():
        left = primary()
        while peek() == ("op", "*"):
            take()
            right = primary()
            left = (_BoolShape(TGeoBoolNode(A.RB_OP_INTERSECTION, left[0], right[0], left[1], right[1])), IDENTITY)
        return left

    def additive():
        left = term()
        while peek() in (("op", "+"), ("op", "-")):
            op = A.RB_OP_UNION if take()[1] == "+" else A.RB_OP_SUBTRACTION
            right = term()
            left = (_BoolShape(TGeoBoolNode(op, left[0], right[0], left[1], right[1])), IDENTITY)
        return left

    result = additive()
    if peek() is not None:
        raise ValueError(f"trailing tokens in composite expression: {toks[it[0]:]}")
    return result


class TGeoCompositeShape(TGeoShape):
    shape_type = A.RB_SHAPE_COMPOSITE

    def __init__(self, name, expr):
        shape, mat = _parse_expr(expr)
        if not isinstance(shape, _BoolShape):
            raise ValueError("composite expression needs a boolean operator")
        super().__init__(name)
        n = shape.node
        if mat is not IDENTITY:
            raise ValueError("a bare matrix on the whole expression is not allowed")
        self.node = n
        self.expr = expr

    params = _BoolShape.params
    children = _BoolShape.children
    local_bbox = _BoolShape.local_bbox


# ------------------------------------------------------------------- volumes
class TGeoVolume:
    """A plain TGeoVolume: optical kind kOther (include/AOpticsManager.h:45-53)."""
    kind = A.RB_KOTHER

    def __init__(self, name, shape, medium=None):
        self.name = name
        self.shape = shape
        self.nodes = []  # (volume, copy, matrix)

    def AddNode(self, volume, copy_no=1, matrix=None):
        self.nodes.append((volume, int(copy_no), matrix if matrix is not None else IDENTITY))

    AddNodeOverlap = AddNode

    def GetShape(self):
        return self.shape

    def GetName(self):
        return self.name

    def GetNdaughters(self):
        return len(self.nodes)


class AOpticalComponent(TGeoVolume):
    kind = A.RB_KOPT

    def AddBorderSurfaceCondition(self, condition):
        """src/AOpticalComponent.cxx:40-48."""
        if not hasattr(self, "border_conditions"):
            self.border_conditions = []
        self.border_conditions.append(condition)

    def FindBorderSurfaceCondition(self, component2):
        """src/AOpticalComponent.cxx:51-65: the first condition registered
        with this component whose second component is component2."""
        for c in getattr(self, "border_conditions", []):
            if c.component2 is component2:
                return c
        return None


class ABorderSurfaceCondition:
    """include/ABorderSurfaceCondition.h:24-45 (src/ABorderSurfaceCondition.cxx):
    the surface between component1 and component2 (None = leaving the top
    volume), registered with component1. Gaussian roughness (rad) perturbs
    the facet normal; Lambertian makes reflections diffuse; a multilayer
    (robast_amd.multilayer.AMultilayer) sets the reflectance and
    transmittance of the surface."""

    def __init__(self, component1, component2):
        self.component1 = component1
        self.component2 = component2
        self.sigma = 0.0
        self.lambertian = False
        self.multilayer = None
        if component1 is not None:
            component1.AddBorderSurfaceCondition(self)

    def GetComponent1(self):
        return self.component1

    def GetComponent2(self):
        return self.component2

    def SetGaussianRoughness(self, sigma):
        self.sigma = abs(float(sigma))

    def GetGaussianRoughness(self):
        return self.sigma

    def EnableLambertian(self, enable):
        self.lambertian = bool(enable)

    def IsLambertian(self):
        return self.lambertian

    def SetMultilayer(self, layer):
        self.multilayer = layer

    def GetMultilayer(self):
        return self.multilayer


class AObscuration(AOpticalComponent):
    kind = A.RB_KOBS


class AMirror(AOpticalComponent):
    """src/AMirror.cxx: reflectance is constant (default 1.0), or a TGraph in
    lambda; GetReflectance clamps to [0, 1] (src/AMirror.cxx:39-60)."""
    kind = A.RB_KMIRROR

    def __init__(self, name, shape, medium=None):
        super().__init__(name, shape, medium)
        self.reflectance = 1.0
        self.reflectance_graph = None

    def SetReflectance(self, r):
        """A constant, a TGraph R(lambda), or a TH2 R(lambda, angle [rad])
        (robast_amd.geoutil.TH2D with SetBinContent); src/AMirror.cxx:39-60
        reads the TH2 first, then the TGraph, then the constant."""
        from .geoutil import TH2D
        from .materials import TGraph
        if isinstance(r, TGraph):
            self.reflectance_graph = r
        elif isinstance(r, TH2D):
            self.reflectance_th2 = r
        else:
            self.reflectance = float(r)
            self.reflectance_graph = None

    def GetReflectance(self, lam, angle=0.0):
        h = getattr(self, "reflectance_th2", None)
        if h is not None:
            r = h.Interpolate(lam, angle)
        else:
            r = self.reflectance_graph.Eval(lam) if self.reflectance_graph is not None else self.reflectance
        return min(1.0, max(0.0, r))


class ALens(AOpticalComponent):
    """src/ALens.cxx:36-60: n (1 without index), k (0 without index), and
    absorption length lambda / (4 pi k) (infinite for k <= 0)."""
    kind = A.RB_KLENS

    def __init__(self, name, shape, medium=None):
        super().__init__(name, shape, medium)
        self.index = None

    def SetRefractiveIndex(self, index):
        self.index = index

    def GetRefractiveIndex(self, lam):
        return self.index.GetRefractiveIndex(lam) if self.index is not None else 1.0

    def GetExtinctionCoefficient(self, lam):
        return self.index.GetExtinctionCoefficient(lam) if self.index is not None else 0.0

    def GetAbsorptionLength(self, lam):
        return self.index.GetAbsorptionLength(lam) if self.index is not None else math.inf


class AFocalSurface(AOpticalComponent):
    """src/AFocalSurface.cxx:35-52: QE(lambda) * QE(angle)."""
    kind = A.RB_KFOCUS

    def __init__(self, name, shape, medium=None):
        super().__init__(name, shape, medium)
        self.qe_lambda = None
        self.qe_angle = None

    def SetQuantumEfficiency(self, g):
        self.qe_lambda = g

    def SetQuantumEfficiencyAngle(self, g):
        self.qe_angle = g

    def HasQEAngle(self):
        return self.qe_angle is not None

    def GetQuantumEfficiency(self, lam, angle=None):
        qe = self.qe_lambda.Eval(lam) if self.qe_lambda is not None else 1.0
        if angle is not None and self.qe_angle is not None:
            qe *= self.qe_angle.Eval(angle)
        return qe
