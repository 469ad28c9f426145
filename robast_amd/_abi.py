"""ctypes/numpy mirror of include/rbgpu.h (the librbgpu C ABI).

Struct layouts here must match rbgpu.h byte for byte; tests/test_abi.py
checks the sizes against the compiled library.
"""
import ctypes as C

import numpy as np

# ray status (include/ARay.h:26)
RB_RUN, RB_STOP, RB_EXIT, RB_FOCUS, RB_SUSPEND, RB_ABSORB = range(6)
RB_SCHEDULE_AUTO, RB_SCHEDULE_LANE, RB_SCHEDULE_PERSISTENT, RB_SCHEDULE_REGROUP = range(4)
STATUS_NAMES = ("running", "stopped", "exited", "focused", "suspended", "absorbed")

# volume kinds (include/AOpticsManager.h:45-53)
RB_KLENS, RB_KOBS, RB_KMIRROR, RB_KFOCUS, RB_KOPT, RB_KOTHER, RB_KNULL = range(7)

RB_SHAPE_BOX = 1
RB_SHAPE_TUBE = 2
RB_SHAPE_SPHERE = 3
RB_SHAPE_PARABOLOID = 4
RB_SHAPE_PGON = 5
RB_SHAPE_PCON = 6
RB_SHAPE_COMPOSITE = 7
RB_SHAPE_ASPHERIC = 8
RB_SHAPE_WINSTON = 9
RB_SHAPE_WINSTON2D = 10

RB_OP_UNION, RB_OP_SUBTRACTION, RB_OP_INTERSECTION = 0, 1, 2

RB_INDEX_NONE, RB_INDEX_TABLE, RB_INDEX_SELLMEIER, RB_INDEX_SCHOTT = 0, 1, 2, 3

RB_PRECISION_FP64, RB_PRECISION_MIXED = 0, 1

RB_ORDER_AUTO, RB_ORDER_INPUT, RB_ORDER_COHERENT = 0, 1, 2

RB_ADJ_CENTROID, RB_ADJ_MOMENTS, RB_ADJ_ENCIRCLED = 0, 1, 2
RB_ADJ_MAX_POLY = 10

RB_RAYGEN_GRID, RB_RAYGEN_DISK, RB_RAYGEN_CONE = 0, 1, 2
(RB_SHOOT_RECTANGLE, RB_SHOOT_CIRCLE, RB_SHOOT_RANDOM_CIRCLE, RB_SHOOT_RANDOM_CONE, RB_SHOOT_RANDOM_RECTANGLE,
 RB_SHOOT_RANDOM_SPHERE, RB_SHOOT_RANDOM_SPHERICAL_CONE) = range(7)

SHAPE_DTYPE = np.dtype(
    [("type", "<i4"), ("param_offset", "<i4"), ("n_params", "<i4"), ("pad", "<i4"),
     ("bbox_origin", "<f8", 3), ("bbox_half", "<f8", 3)], align=True)
VOLUME_DTYPE = np.dtype(
    [("refl_const", "<f8"), ("shape", "<i4"), ("kind", "<i4"), ("refl_table", "<i4"),
     ("index_model", "<i4"), ("index_ref", "<i4"), ("k_table", "<i4"),
     ("qe_lambda_table", "<i4"), ("qe_angle_table", "<i4"), ("border_first", "<i4"), ("border_count", "<i4"),
     ("refl_grid", "<i4"), ("pad", "<i4")],
    align=True)
BORDER_DTYPE = np.dtype([("volume2", "<i4"), ("lambertian", "<i4"), ("sigma", "<f8"), ("multilayer", "<i4"),
                         ("pad", "<i4")], align=True)
GRID_DTYPE = np.dtype([("nx", "<i4"), ("ny", "<i4"), ("xlo", "<f8"), ("xhi", "<f8"), ("ylo", "<f8"), ("yhi", "<f8"),
                       ("r_offset", "<i4"), ("t_offset", "<i4")], align=True)
NODE_DTYPE = np.dtype(
    [("rot", "<f8", 9), ("tr", "<f8", 3), ("volume", "<i4"), ("parent", "<i4"),
     ("first_child", "<i4"), ("n_children", "<i4")], align=True)
TABLE_DTYPE = np.dtype([("offset", "<i4"), ("n", "<i4")], align=True)

assert SHAPE_DTYPE.itemsize == 64
assert VOLUME_DTYPE.itemsize == 56
assert BORDER_DTYPE.itemsize == 24
assert GRID_DTYPE.itemsize == 48
assert NODE_DTYPE.itemsize == 112
assert TABLE_DTYPE.itemsize == 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class rb_scene_desc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("nodes", C.c_void_p),
        ("n_children", C.c_int32), ("children", C.c_void_p),
        ("n_volumes", C.c_int32), ("volumes", C.c_void_p),
        ("n_shapes", C.c_int32), ("shapes", C.c_void_p),
        ("n_params", C.c_int32), ("params", C.c_void_p),
        ("n_tables", C.c_int32), ("tables", C.c_void_p),
        ("n_table_xy", C.c_int32), ("table_xy", C.c_void_p),
        ("n_borders", C.c_int32), ("borders", C.c_void_p),
        ("n_grids", C.c_int32), ("grids", C.c_void_p),
    ]


class rb_rays(C.Structure):
    _fields_ = [
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("t", C.c_void_p),
        ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p),
        ("lambda_", C.c_void_p), ("status", C.c_void_p), ("npoints", C.c_void_p),
        ("last_node", C.c_void_p),
        ("path_capacity", C.c_int32), ("pad", C.c_int32),
        ("path_xyzt", C.c_void_p), ("path_node", C.c_void_p),
    ]


class rb_trace_opts(C.Structure):
    _fields_ = [
        ("limit", C.c_int32), ("fresnel", C.c_int32), ("precision", C.c_int32),
        ("device_ptrs", C.c_int32), ("seed", C.c_uint64), ("ray_id_base", C.c_uint64),
        ("order", C.c_int32), ("schedule", C.c_int32),
    ]


class rb_raygen_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_side", C.c_int32), ("size", C.c_double),
        ("distance", C.c_double), ("theta_min", C.c_double), ("theta_max", C.c_double),
        ("lambda_min", C.c_double), ("lambda_max", C.c_double), ("seed", C.c_uint64),
        ("z0", C.c_double), ("start", C.c_double),
    ]


class rb_adjoint_desc(C.Structure):
    _fields_ = [
        ("node_param", C.c_void_p), ("param_values", C.c_void_p), ("n_params", C.c_int32),
        ("z_param", C.c_int32), ("focal_node", C.c_int32), ("mode", C.c_int32),
        ("centroid_x", C.c_double), ("centroid_y", C.c_double), ("weight", C.c_double),
        ("ref_xy", C.c_void_p), ("group_size", C.c_int64), ("n_groups", C.c_int32), ("pad", C.c_int32),
        ("surface_param", C.c_void_p), ("ee_radius", C.c_double), ("ee_width", C.c_double),
    ]


class rb_hist2d_desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("xlo", C.c_double), ("xhi", C.c_double),
        ("ylo", C.c_double), ("yhi", C.c_double), ("status", C.c_int32), ("device_ptrs", C.c_int32),
    ]


class rb_shooter_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("pad", C.c_int32),
        ("lambda_", C.c_double), ("a", C.c_double), ("b", C.c_double),
        ("rot", C.c_double * 9), ("tr", C.c_double * 3), ("dir", C.c_double * 3),
        ("seed", C.c_uint64), ("id_base", C.c_uint64),
    ]


BUNCH_FIELDS_I = ("tel_no", "array_no")
BUNCH_FIELDS_F = ("x", "y", "zem", "time", "cx", "cy", "cz", "lambda_", "photons")


class rb_bunches(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in BUNCH_FIELDS_I + BUNCH_FIELDS_F]


class rb_bunch_opts(C.Structure):
    _fields_ = [
        ("tel_no", C.c_int32), ("array_no", C.c_int32), ("z", C.c_double), ("tel_z", C.c_double),
        ("refractive_index", C.c_double), ("lambda_min", C.c_double), ("lambda_max", C.c_double),
        ("seed", C.c_uint64), ("id_base", C.c_uint64), ("device_ptrs", C.c_int32), ("pad", C.c_int32),
    ]


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
