"""The C-ABI library loads and exports every entry point include/*.h declares (CPU only)."""
import ctypes
import glob
import os
import re

import pytest

from difftactile_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(dt_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "dt_mpm_p2g" in names and "dt_last_error" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert set(_lib.declared_symbols()) == set(declared())


def test_error_path_without_gpu_work():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.dt_version() >= 1
    # a NULL parameter struct is rejected on the host before any launch
    rc = L.dt_mpm_grid_occupy(None, None, None, None)
    assert rc == -1
    assert b"NULL" in L.dt_last_error()


def test_call_rejects_host_tensors_before_any_launch():
    import torch
    with pytest.raises(ValueError):
        _lib._arg(torch.zeros(3))
    assert _lib._arg(None) is None and _lib._arg(7) == 7


def test_binding_constants_match_header():
    src = open(os.path.join(ROOT, "include", "difftactile.h")).read()
    m = re.search(r"#define\s+DT_CONTACT_TRI_RUN\s+(\d+)", src)
    assert m and int(m.group(1)) == _lib.TRI_RUN
    m = re.search(r"#define\s+DT_CONTACT_TRI_GROUP\s+(\d+)", src)
    assert m and int(m.group(1)) == _lib.TRI_GROUP
    # run boxes then group boxes, two float4 each
    assert [_lib.tri_bounds_count(n) for n in (0, 1, 8, 9, 64, 65, 3042)] == \
        [0, 4, 4, 6, 18, 22, 2 * (381 + 48)]


def test_jaw_frame_size_matches_header():
    src = open(os.path.join(ROOT, "include", "difftactile.h")).read()
    m = re.search(r"#define\s+DT_JAW_FRAME_FLOATS\s+(\d+)", src)
    assert m and int(m.group(1)) == _lib.JAW_FRAME_FLOATS == 6 * 16
    assert ctypes.sizeof(_lib.DtJawParams) == 4 * (16 + 32 + 1)


def test_dome_pose_layout_matches_header():
    src = open(os.path.join(ROOT, "include", "difftactile.h")).read()
    names = dict(trans_world="TRANS_WORLD", trans_local="TRANS_LOCAL", trans_h="TRANS_H",
                 inv_trans_h="INV_TRANS_H", dtrans_h="DTRANS_H", rot_world="ROT_WORLD",
                 rot_local="ROT_LOCAL", rot_h="ROT_H", inv_rot="INV_ROT")
    end = 0
    for field, (off, shape) in _lib.DOME_POSE.items():
        m = re.search(r"#define\s+DT_DOME_%s\s+(\d+)" % names[field], src)
        assert m and int(m.group(1)) == off == end, field
        end = off + shape[0] * shape[1]
    m = re.search(r"#define\s+DT_DOME_POSE_FLOATS\s+(\d+)", src)
    assert m and int(m.group(1)) == _lib.DOME_POSE_FLOATS == end
