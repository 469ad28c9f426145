"""Scene and ray-source builders for the benchmark/parity workloads.

C1 parabolic telescope (after tutorials/SimpleParabolicTelescope.C:20-88)
C2 Davies-Cotton 85-facet reflector (after tutorials/DaviesCotton.C:58-157)
C3 Schwarzschild-Couder aspheric two-mirror system (after tutorials/SchwarzschildCouder.C:33-90)
C4 BK7 window + 1,027 hexagonal Winston-cone light guides (after tutorials/HexWinstonCone.C:29-75)
"""
import math

import numpy as np

from . import _abi as A
from .geometry import (AFocalSurface, AGeoAsphericDisk, AGeoWinstonConePoly, ALens, AMirror, AObscuration, AOpticalComponent, TGeoBBox, TGeoCombiTrans,
                       TGeoCompositeShape, TGeoParaboloid, TGeoPgon, TGeoRotation, TGeoSphere, TGeoTranslation,
                       TGeoTube, cm, m, mm, nm, um)
from .manager import AOpticsManager
from .materials import ASellmeierFormula, ARefractiveIndex, TGraph
from .rays import ARayShooter

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


def c1_raygen(deg, nside=C1_NSIDE, lam=400 * nm):
    """The device-side source of c1_rays: a GRID of nside x nside rays over
    the 3 m square at distance 2f."""
    g = A.rb_raygen_desc()
    g.kind = A.RB_RAYGEN_GRID
    g.n_side = nside
    g.size = 2 * C1_R
    g.distance = 2 * C1_F
    g.theta_min = g.theta_max = deg * math.pi / 180.0
    g.lambda_min = g.lambda_max = lam
    return g


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


def c2_scene(deterministic=False, per_facet=False, facet_radius=None, facet_tilt=None, camera_dz=0.0, surfaces=False):
    """85-facet Davies-Cotton, f = 16 m. Facet reflectance linear in lambda
    0.80 -> 0.90 over 300-600 nm (R = 1 when deterministic). Camera disk
    r = 1.2 m at z = f; square 2.5 m obscuration plate 1 cm behind it.
    per_facet: one volume per facet (needed for per-facet radius / tilt).
    surfaces: the facets' front surface gets a border condition, 0.5 deg
    Gaussian roughness and an aluminium coating under 25 nm of silica
    (constant indices n = 0.49 + 4.86 i and 1.46; the default table)."""
    mgr = AOpticsManager("manager", "Davies-Cotton")
    mgr.DisableFresnelReflection(True)
    world = AOpticalComponent("world", TGeoBBox("boxWorld", 20 * m, 20 * m, 20 * m))
    mgr.SetTopVolume(world)
    refl = None if deterministic else c2_reflectance()
    pos = dc_facet_positions()
    shared = None
    rots = None
    if facet_tilt is not None:
        rots = facet_rotations([_dc_facet_matrix(x, y).rot for x, y in pos], facet_tilt)
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
        if rots is not None:
            mat.rot = rots[i]
        world.AddNode(mirror, i + 1, mat)
    if surfaces:
        from .geometry import ABorderSurfaceCondition
        from .multilayer import AMultilayer
        if per_facet:
            raise ValueError("c2_scene: surfaces apply to the shared facet volume")
        coat = AMultilayer(ARefractiveIndex(1.0), ARefractiveIndex(0.49, 4.86))
        coat.InsertLayer(ARefractiveIndex(1.46), 25 * nm)
        cond = ABorderSurfaceCondition(world, shared)
        cond.SetGaussianRoughness(math.radians(0.5))
        cond.SetMultilayer(coat)
    zc = C2_F + camera_dz
    cam = AFocalSurface("focalPlane", TGeoTube("tubeCamera", 0, C2_CAMERA_R, 1 * mm))
    world.AddNode(cam, 1, TGeoTranslation(0, 0, zc + 1 * mm))
    obs = AObscuration("cameraBox", TGeoBBox("boxCamera", C2_OBS_SIDE / 2, C2_OBS_SIDE / 2, 5 * cm))
    world.AddNode(obs, 1, TGeoTranslation(0, 0, zc + 2 * mm + 1 * cm + 5 * cm))
    mgr.CloseGeometry()
    return mgr


def facet_rotations(base, tilts):
    """base @ Rx(tip) @ Ry(tilt) for every facet: base (N, 3, 3) placement
    rotations, tilts (N, 2) (tip, tilt) in radians, a rotation by `tip` about
    the facet's local x then `tilt` about its local y. One vectorised
    evaluation, shared by the scene builder and the alignment's scene updater
    so that both give the same bits."""
    t = np.asarray(tilts, dtype=float).reshape(-1, 2)
    cx, sx, cy, sy = np.cos(t[:, 0]), np.sin(t[:, 0]), np.cos(t[:, 1]), np.sin(t[:, 1])
    small = np.zeros((len(t), 3, 3))
    small[:, 0, 0], small[:, 0, 2] = cy, sy
    small[:, 1, 0], small[:, 1, 1], small[:, 1, 2] = sx * sy, cx, -(sx * cy)
    small[:, 2, 0], small[:, 2, 1], small[:, 2, 2] = -(cx * sy), sx, cx * cy
    return np.matmul(np.asarray(base, dtype=float).reshape(-1, 3, 3), small)


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


# ------------------------------------------------------------------ C3
C3_FP = 16.915 * m      # primary focal length (SchwarzschildCouder.C:33-45)
C3_FS = -3.553 * m      # secondary focal length
C3_ZS = 9.980 * m       # secondary position
C3_ZF = 7.631 * m       # focal surface position
C3_RP_MAX, C3_RP_MIN = 4.8 * m, 2.2 * m
C3_RS_MAX = 2.7 * m
C3_FOCAL_RC = 1.0 * m   # focal surface: spherical cap, radius of curvature 1 m
C3_FOCAL_R = 0.2 * m    # 0.4 m across
C3_NPOL = 8             # r^2 .. r^16


def c3_coefficients(seed=0xC3):
    """Seeded draws: conic ~ U[-1, 0]; polynomial coefficient k (of r^2k) ~
    N(0, 1e-3 / R^(2k-1)) with R the surface's outer radius in cm."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = {}
    for name, rmax in (("primary", C3_RP_MAX), ("secondary", C3_RS_MAX)):
        conic = rng.uniform(-1.0, 0.0)
        ks = [rng.normal(0.0, 1e-3 / rmax ** (2 * k - 1)) for k in range(1, C3_NPOL + 1)]
        out[name] = (conic, ks)
    return out


def c3_figure(seed=0xC3):
    """The mirrors' figures as ((c, kappa, K[]) primary, (c, kappa, K[])
    secondary): base curvature 1/(2F), kappa = conic + 1, r^2k coefficients."""
    co = c3_coefficients(seed)
    (kp, polp), (ks, pols) = co["primary"], co["secondary"]
    return ((0.5 / C3_FP, kp + 1.0, list(polp)), (0.5 / C3_FS, ks + 1.0, list(pols)))


def c3_scene(seed=0xC3, figure=None, focal_dz=0.0):
    """Primary (R 4.8 m, hole 2.2 m) and secondary (R 2.7 m) 1 um thick
    AGeoAsphericDisk mirrors with base curvature 1/(2F) from the tutorial,
    seeded conics and r^2..r^16 terms; spherical-cap focal surface (R_c = 1 m,
    0.4 m across) with an obscuration disk behind it. World -> top -> optics,
    as the tutorial nests them. figure: the mirrors' (c, kappa, K[]) in place
    of the seeded ones (both surfaces of each disk; c3_figure's layout);
    focal_dz moves the focal surface along z."""
    if figure is None:  # the seeded conics as drawn (kappa - 1 would round)
        co = c3_coefficients(seed)
        (conic_p, polp), (conic_s, pols) = co["primary"], co["secondary"]
        cp, cs = 0.5 / C3_FP, 0.5 / C3_FS
    else:
        (cp, kap_p, polp), (cs, kap_s, pols) = figure
        conic_p, conic_s = kap_p - 1.0, kap_s - 1.0
    mgr = AOpticsManager("manager", "SC")
    world = AOpticalComponent("world", TGeoBBox("worldbox", 30 * m, 30 * m, 30 * m))
    mgr.SetTopVolume(world)
    top = AOpticalComponent("top", TGeoBBox("topbox", 25 * m, 25 * m, 25 * m))
    prim = AGeoAsphericDisk("primaryV", -1 * um, cp, 0.0, cp, C3_RP_MAX, C3_RP_MIN)
    prim.SetConicConstants(conic_p, conic_p)
    prim.SetPolynomials(len(polp), polp, len(polp), polp)
    top.AddNode(AMirror("primaryMirror", prim), 1)
    sec = AGeoAsphericDisk("secondaryV", C3_ZS, cs, C3_ZS + 1 * um, cs, C3_RS_MAX, 0.0)
    sec.SetConicConstants(conic_s, conic_s)
    sec.SetPolynomials(len(pols), pols, len(pols), pols)
    top.AddNode(AMirror("secondaryMirror", sec), 1)
    cf = -1.0 / C3_FOCAL_RC
    foc = AGeoAsphericDisk("focalV", C3_ZF - 1 * mm, cf, C3_ZF, cf, C3_FOCAL_R, 0.0)
    if focal_dz:
        top.AddNode(AFocalSurface("focalPlane", foc), 1, TGeoTranslation(0.0, 0.0, float(focal_dz)))
    else:
        top.AddNode(AFocalSurface("focalPlane", foc), 1)
    zlow = foc.CalcF1(C3_FOCAL_R)
    obs = AObscuration("focalObs", TGeoTube("focalObsV", 0, C3_FOCAL_R * 1.25, 0.5 * cm))
    top.AddNode(obs, 1, TGeoTranslation(0, 0, zlow - 1 * cm))
    world.AddNode(top, 1)
    mgr.CloseGeometry()
    return mgr


def c3_raygen(seed=0xC3):
    """Parallel beams over a 10 m disk from 1.2 z_s up the axis, 0-5 deg."""
    g = A.rb_raygen_desc()
    g.kind = A.RB_RAYGEN_DISK
    g.size = 10 * m
    g.distance = 1.2 * C3_ZS
    g.theta_min = 0.0
    g.theta_max = 5.0 * math.pi / 180.0
    g.lambda_min = g.lambda_max = 400 * nm
    g.seed = seed
    return g


# ------------------------------------------------------------------ C4
C4_R1, C4_R2 = 12.5 * mm, 5.0 * mm     # entrance / exit apothems (25 mm, 10 mm flat-to-flat)
C4_RINGS = 18                           # 1 + 3 * 18 * 19 = 1027 cones
C4_WALL_R = 0.9
C4_BK7 = (1.0396, 0.23179, 1.01047, 6.0007e-3, 2.0018e-2, 103.56)
C4_ABS_400 = 1.0 * m                    # absorption length at 400 nm
C4_FIELD_DEG = 6.0
C4_FNUMBER = 1.2


def c4_positions(rings=C4_RINGS, pitch=None):
    """Centres of a hexagonal array of hexagonal cones whose flats face +-x
    (the cone's side normals are at 0, 60, ... deg): all lattice points
    i a1 + j a2 with a1 = pitch (1, 0), a2 = pitch (1/2, sqrt(3)/2) and hex
    distance max(|i|, |j|, |i + j|) <= rings."""
    if pitch is None:
        pitch = 2 * C4_R1 * 1.002
    pts = []
    for i in range(-rings, rings + 1):
        for j in range(-rings, rings + 1):
            if max(abs(i), abs(j), abs(i + j)) <= rings:
                pts.append((pitch * (i + 0.5 * j), pitch * (math.sqrt(3) / 2) * j))
    return pts


def c4_scene(rings=C4_RINGS, deterministic=False):
    """BK7 window (Sellmeier; k from a 1 m absorption length at 400 nm, so the
    length scales with lambda) above a hexagonal array of hexagonal Winston
    cones, each a mirror block "pgon:rot30 - hexV" with wall reflectance 0.9
    and a thin hexagonal focal surface at its exit (HexWinstonCone.C mode 0).
    The cone length follows from r1, r2: 2 dz = 40.1 mm. deterministic: unit
    wall reflectance, no absorption in the window and Fresnel reflection off,
    so no ray draws a random number."""
    mgr = AOpticsManager("manager", "light guides")
    world = AOpticalComponent("world", TGeoBBox("worldbox", 1 * m, 1 * m, 1 * m))
    mgr.SetTopVolume(world)
    hexv = AGeoWinstonConePoly("hexV", C4_R1, C4_R2, 6)
    dz = hexv.GetDZ()
    TGeoRotation("rot30", 30, 0, 0).RegisterYourself()
    pgon = TGeoPgon("pgon", 0, 360, 6, 4)
    pgon.DefineSection(0, -dz * 0.999, 0, C4_R2 * 1.1)
    pgon.DefineSection(1, -dz * 0.5, 0, C4_R1 * 0.9)
    pgon.DefineSection(2, -dz * 0.0, 0, C4_R1 * 0.99)
    pgon.DefineSection(3, dz * 0.999, 0, C4_R1 * 1.001)
    mirror = AMirror("coneMirror", TGeoCompositeShape("coneComp1", "pgon:rot30 - hexV"))
    mirror.SetReflectance(1.0 if deterministic else C4_WALL_R)
    pmt_shape = TGeoPgon("pgonPMT", 0, 360, 6, 2)
    pmt_shape.DefineSection(0, -dz - 0.01 * mm, 0, C4_R2 * 1.01)
    pmt_shape.DefineSection(1, -dz, 0, C4_R2 * 1.01)
    pmt = AFocalSurface("pmt", pmt_shape)
    rot30 = TGeoRotation("", 30, 0, 0)
    pos = c4_positions(rings)
    for k, (x, y) in enumerate(pos):
        world.AddNode(mirror, k + 1, TGeoTranslation(x, y, 0.0))
        world.AddNode(pmt, k + 1, TGeoCombiTrans(x, y, 0.0, rot30))
    rmax = max(math.hypot(x, y) for x, y in pos) + 2 * C4_R1
    bk7 = ASellmeierFormula(*C4_BK7)
    kg = TGraph()
    kg.SetPoint(0, 400 * nm, ARefractiveIndex.AbsorptionLengthToExtinctionCoefficient(C4_ABS_400, 400 * nm))
    if not deterministic:
        bk7.SetExtinctionCoefficient(kg)
    lens = ALens("lens", TGeoTube("lensV", 0, rmax, 0.5 * cm))
    lens.SetRefractiveIndex(bk7)
    world.AddNode(lens, 1, TGeoTranslation(0, 0, dz + 2 * cm))
    if deterministic:
        mgr.DisableFresnelReflection(True)
    mgr.CloseGeometry()
    mgr.c4_dz = dz
    mgr.c4_rmax = rmax
    return mgr


def c4_raygen(dz, seed=0xC4, rings=C4_RINGS):
    """f/1.2 beams converging on the cone entrance plane z = dz, field angle
    0-6 deg (the 6 deg image at the array's edge), lambda 300-700 nm."""
    g = A.rb_raygen_desc()
    g.kind = A.RB_RAYGEN_CONE
    r_edge = rings * 2 * C4_R1 * 1.002
    F = r_edge / math.tan(math.radians(C4_FIELD_DEG))
    g.distance = F
    g.size = F / C4_FNUMBER
    g.theta_min = 0.0
    g.theta_max = math.radians(C4_FIELD_DEG)
    g.lambda_min = 300 * nm
    g.lambda_max = 700 * nm
    g.seed = seed
    g.z0 = dz
    g.start = 5 * cm
    return g
