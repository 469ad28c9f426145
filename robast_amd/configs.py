"""Scene and ray-source builders for the benchmark/parity workloads.

C1 parabolic telescope (after tutorials/SimpleParabolicTelescope.C:20-88)
C2 Davies-Cotton 85-facet reflector (after tutorials/DaviesCotton.C:58-157)
"""
import math

import numpy as np

from . import _abi as A
from .geometry import (AFocalSurface, AMirror, AObscuration, AOpticalComponent, TGeoBBox, TGeoCombiTrans,
                       TGeoCompositeShape, TGeoParaboloid, TGeoPgon, TGeoRotation, TGeoSphere, TGeoTranslation,
                       TGeoTube, cm, m, mm, nm, um)
from .manager import AOpticsManager
from .materials import TGraph
from .rays import ARayArray, ARayShooter

# ------------------------------------------------------------------ C1
C1_F = 8 * m
C1_R = 1.5 * m
C1_ANGLES_DEG = (0.0, 0.5, 1.0, 1.5)
C1_NSIDE = 316


def c1_scene():
    """f = 8 m paraboloid of 3 m aperture made as a 1 um composite shell
    (para:tr2 - para:tr1), R = 1. Focal disk r = 50 cm at z = f, obscuration
    disk of radius 0.6 m 10 cm behind it, shadowing the whole focal disk from
    sky-side rays as the tutorial's obscuration does. World: 40 m cube."""
    sag = C1_R * C1_R / 4.0 / C1_F
    mgr = AOpticsManager("manager", "C1 parabolic")
    world = AOpticalComponent("world", TGeoBBox("worldbox", 20 * m, 20 * m, 20 * m))
    mgr.SetTopVolume(world)
    TGeoParaboloid("mirror_para", 0, C1_R, sag / 2.0)
    TGeoTranslation("mirror_tr1", 0, 0, sag / 2.0).RegisterYourself()
    TGeoTranslation("mirror_tr2", 0, 0, sag / 2.0 - 1 * um).RegisterYourself()
    comp = TGeoCompositeShape("mirror_comp", "mirror_para:mirror_tr2 - mirror_para:mirror_tr1")
    world.AddNode(AMirror("mirror", comp), 1)
    focal = AFocalSurface("focal", TGeoTube("focal_tube", 0, 50 * cm, 10 * um))
    world.AddNode(focal, 1, TGeoTranslation("focal_tr", 0, 0, C1_F + 10 * um))
    obs = AObscuration("obs", TGeoTube("obs_tube", 0, 60 * cm, 10 * um))
    world.AddNode(obs, 1, TGeoTranslation("obs_tr", 0, 0, C1_F + 10 * cm + 10 * um))
    mgr.CloseGeometry()
    return mgr


def c1_rays(deg, nside=C1_NSIDE, lam=400 * nm):
    """ARayShooter::Square(lambda, 3 m, nside) from (-2f sin th, 0, 2f cos th),
    direction SetMagThetaPhi(1, pi - th, 0) (SimpleParabolicTelescope.C:104-111)."""
    th = deg * math.pi / 180.0
    tr = TGeoTranslation("raytr", -C1_F * 2 * math.sin(th), 0, C1_F * 2 * math.cos(th))
    t2 = math.pi - th
    v = (math.sin(t2) * math.cos(0.0), math.sin(t2) * math.sin(0.0), math.cos(t2))
    return ARayShooter.Square(lam, 2 * C1_R, nside, None, tr, v)


# ------------------------------------------------------------------ C2
C2_F = 16 * m
C2_MIRROR_R = 2 * C2_F
C2_MIRROR_D = 1.2 * m
C2_MIRROR_T = 0.1 * mm
C2_CAMERA_R = 1.2 * m
C2_OBS_SIDE = 2.5 * m
C2_NFACETS = 85
C2_DISK_D = 12.5 * m


def dc_facet_positions(n=C2_NFACETS, pitch=C2_MIRROR_D):
    """Centres of the n hexagonal facets nearest the axis on the lattice of
    tutorials/DaviesCotton.C:75-105 (x = 1.5 i dx, y = j dy, i + j even).
    85 = every lattice point with r^2 <= 21 pitch^2 (complete shells)."""
    dx = pitch / math.sqrt(3)
    dy = pitch / 2.0
    pts = []
    for i in range(-12, 13):
        for j in range(-24, 25):
            if (i + j) % 2:
                continue
            x, y = 1.5 * dx * i, dy * j
            pts.append((x * x + y * y, math.atan2(y, x), x, y))
    pts.sort(key=lambda p: (round(p[0] / (pitch * pitch), 6), p[1]))
    r2n = round(pts[n - 1][0] / (pitch * pitch), 6)
    assert round(pts[n][0] / (pitch * pitch), 6) > r2n, "facet count does not close a shell"
    return [(p[2], p[3]) for p in pts[:n]]


def _dc_facet_matrix(x, y, f=C2_F):
    """DaviesCotton.C:107-124: centre on the sphere of radius f, aimed at (0,0,2f)."""
    r2d = 180.0 / math.pi
    r2 = x * x + y * y
    z = f - math.sqrt(f * f - r2)
    trans = TGeoTranslation(x, y, z)
    phi = math.atan2(y, x) * r2d
    rot = TGeoRotation("", -phi + 90.0, 0, 0)
    theta = math.atan2(math.sqrt(r2), 2 * f - z) * r2d
    rot2 = TGeoRotation("", phi - 90.0, theta, 0)
    rot.MultiplyBy(rot2, False)
    return TGeoCombiTrans(trans, rot)


def dc_facet_shape(name, radius=C2_MIRROR_R):
    """Spherical cap (thickness 0.1 mm) cut by a hexagonal prism
    (DaviesCotton.C:58-72): "sphere:transZ * hexCut"."""
    cut = TGeoPgon(name + "_cut", 0.0, 360.0, 6, 2)
    cut.DefineSection(0, -100 * mm, 0, C2_MIRROR_D / 2.0)
    cut.DefineSection(1, 100 * mm, 0, C2_MIRROR_D / 2.0)
    theta = math.asin(C2_MIRROR_D / math.sqrt(3) / radius) * 180.0 / math.pi
    TGeoSphere(name + "_sph", radius, radius + C2_MIRROR_T, 180.0 - theta, 180.0)
    TGeoTranslation(name + "_tz", 0, 0, radius).RegisterYourself()
    return TGeoCompositeShape(name, f"{name}_sph:{name}_tz*{name}_cut")


def c2_reflectance():
    g = TGraph()
    g.SetPoint(0, 300 * nm, 0.80)
    g.SetPoint(1, 600 * nm, 0.90)
    return g


def c2_scene(deterministic=False, per_facet=False, facet_radius=None, facet_tilt=None, camera_dz=0.0):
    """85-facet Davies-Cotton, f = 16 m. Facet reflectance linear in lambda
    0.80 -> 0.90 over 300-600 nm (R = 1 when deterministic). Camera disk
    r = 1.2 m at z = f; square 2.5 m obscuration plate 1 cm behind it.
    per_facet: one volume per facet (needed for per-facet radius / tilt)."""
    mgr = AOpticsManager("manager", "Davies-Cotton")
    mgr.DisableFresnelReflection(True)
    world = AOpticalComponent("world", TGeoBBox("boxWorld", 20 * m, 20 * m, 20 * m))
    mgr.SetTopVolume(world)
    refl = None if deterministic else c2_reflectance()
    pos = dc_facet_positions()
    shared = None
    for i, (x, y) in enumerate(pos):
        if per_facet or shared is None:
            rad = facet_radius[i] if facet_radius is not None else C2_MIRROR_R
            mirror = AMirror(f"mirror{i}" if per_facet else "mirror", dc_facet_shape(f"mirComposite{i}", rad))
            if refl is not None:
                mirror.SetReflectance(refl)
            if not per_facet:
                shared = mirror
        else:
            mirror = shared
        mat = _dc_facet_matrix(x, y)
        if facet_tilt is not None:
            tip, tilt = facet_tilt[i]
            extra = TGeoRotation()
            extra.rot = _small_rotation(tip, tilt)
            mat.rot = mat.rot @ extra.rot
        world.AddNode(mirror, i + 1, mat)
    zc = C2_F + camera_dz
    cam = AFocalSurface("focalPlane", TGeoTube("tubeCamera", 0, C2_CAMERA_R, 1 * mm))
    world.AddNode(cam, 1, TGeoTranslation(0, 0, zc + 1 * mm))
    obs = AObscuration("cameraBox", TGeoBBox("boxCamera", C2_OBS_SIDE / 2, C2_OBS_SIDE / 2, 5 * cm))
    world.AddNode(obs, 1, TGeoTranslation(0, 0, zc + 2 * mm + 1 * cm + 5 * cm))
    mgr.CloseGeometry()
    return mgr


def _small_rotation(tip, tilt):
    """Rotation by `tip` about local x then `tilt` about local y (radians)."""
    cx, sx, cy, sy = math.cos(tip), math.sin(tip), math.cos(tilt), math.sin(tilt)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    return rx @ ry


def c2_raygen(seed=0xC2, theta_max_deg=4.0):
    g = A.rb_raygen_desc()
    g.kind = A.RB_RAYGEN_DISK
    g.n_side = 0
    g.size = C2_DISK_D
    g.distance = 2 * C2_F
    g.theta_min = 0.0
    g.theta_max = theta_max_deg * math.pi / 180.0
    g.lambda_min = 300 * nm
    g.lambda_max = 600 * nm
    g.seed = seed
    return g
