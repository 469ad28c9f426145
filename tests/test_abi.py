"""The C-ABI library loads and exports every entry point include/rbgpu.h
declares; struct layouts agree with the Python binding. No GPU needed."""
import ctypes as C
import os
import re

import numpy as np

from robast_amd import _abi as A
from robast_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rbgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rbgpu_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), f"librbgpu.so does not export {s}"
    assert set(syms) == set(_lib.EXPORTS)


def test_struct_sizes_match_binding():
    lib = _lib.load()
    out = (C.c_int32 * 16)()
    assert lib.rbgpu_abi_sizes(out) == 16
    sizes = list(out)[:16]
    assert sizes[0] == 96                              # rb_matrix
    assert sizes[1] == A.SHAPE_DTYPE.itemsize
    assert sizes[2] == A.VOLUME_DTYPE.itemsize == 56
    assert sizes[3] == A.NODE_DTYPE.itemsize
    assert sizes[4] == A.TABLE_DTYPE.itemsize
    assert sizes[5] == C.sizeof(A.rb_scene_desc)
    assert sizes[6] == C.sizeof(A.rb_rays)
    assert sizes[7] == C.sizeof(A.rb_trace_opts)
    assert sizes[8] == C.sizeof(A.rb_raygen_desc)
    assert sizes[9] == C.sizeof(A.rb_adjoint_desc)
    assert sizes[10] == C.sizeof(A.rb_hist2d_desc) == 48
    assert sizes[11] == A.BORDER_DTYPE.itemsize == 24
    assert C.sizeof(A.rb_trace_opts) == 40
    assert sizes[12] == C.sizeof(A.rb_shooter_desc) == 176
    assert sizes[13] == C.sizeof(A.rb_bunches) == 11 * 8
    assert sizes[14] == C.sizeof(A.rb_bunch_opts) == 72
    assert sizes[15] == A.GRID_DTYPE.itemsize == 48


def test_null_context_is_an_error_not_a_crash():
    lib = _lib.load()
    assert lib.rbgpu_trace(None, None, None, None, 0) == -1
    assert lib.rbgpu_trace_noalloc(None, None, None, None, 0) == -1
    assert lib.rbgpu_scene_upload(None, None, None) == -1
    assert lib.rbgpu_last_error(None) == b"null context"
    assert lib.rbgpu_containment_radius(None, None, 0, None, 0.8, None, None, None) == -1
    assert lib.rbgpu_shoot_rays(None, None, None, 0) == -1
    assert lib.rbgpu_bunches_to_rays(None, None, 0, None, None, 0, None) == -1
    assert lib.rbgpu_shooter_count(None, 0) == -1
