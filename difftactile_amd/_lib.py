"""ctypes binding of libdifftactile_hip.so (the C-ABI in include/difftactile.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises.  There is no CPU or PyTorch fallback.
"""
import ctypes
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdifftactile_hip.so")

_c_int = ctypes.c_int32
_c_float = ctypes.c_float
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


# centroids per run and runs per group of the accelerated closest-triangle
# search (DT_CONTACT_TRI_RUN, DT_CONTACT_TRI_GROUP)
TRI_RUN = 8
TRI_GROUP = 8


def tri_bounds_count(n_tris):
    """float4 dt_contact_tri_batches writes for n_tris triangles: two per
    run box and per group box (DT_CONTACT_TRI_BOUNDS)."""
    return 2 * (-(-n_tris // TRI_RUN) + -(-n_tris // (TRI_RUN * TRI_GROUP)))


class DtMpmParams(ctypes.Structure):
    _fields_ = [
        ("n_grid", ctypes.c_int32), ("bound", ctypes.c_int32),
        ("sticky_floor", ctypes.c_int32), ("clamp_grad", ctypes.c_int32),
        ("dx", ctypes.c_float), ("inv_dx", ctypes.c_float), ("dt", ctypes.c_float),
        ("p_vol", ctypes.c_float), ("p_mass", ctypes.c_float), ("eps", ctypes.c_float),
        ("gravity", ctypes.c_float * 3), ("grad_clamp", ctypes.c_float),
        ("soa", ctypes.c_int32), ("svd_exact", ctypes.c_int32), ("n_env", ctypes.c_int32),
        ("env_id", ctypes.c_void_p), ("active_bricks", ctypes.c_void_p),
        ("n_active_bricks", ctypes.c_void_p), ("env_ngrid", ctypes.c_void_p),
        ("env_scale", ctypes.c_void_p),
    ]


class DtJawParams(ctypes.Structure):
    _fields_ = [("ee_world", ctypes.c_float * 16), ("finger_t", (ctypes.c_float * 16) * 2),
                ("dt", ctypes.c_float)]


class DtLaunchConfig(ctypes.Structure):
    _fields_ = [("brick_blocks", ctypes.c_int32), ("node_blocks", ctypes.c_int32),
                ("sort_blocks", ctypes.c_int32)]


# floats per frame of the jaw chain (DT_JAW_FRAME_FLOATS)
JAW_FRAME_FLOATS = 96
# dome pad pose state (DT_DOME_*): field offsets and size
DOME_POSE = dict(trans_world=(0, (4, 4)), trans_local=(16, (4, 4)), trans_h=(32, (4, 4)),
                 inv_trans_h=(48, (4, 4)), dtrans_h=(64, (4, 4)), rot_world=(80, (3, 3)),
                 rot_local=(89, (3, 3)), rot_h=(98, (3, 3)), inv_rot=(107, (3, 3)))
DOME_POSE_FLOATS = 116
# blocks of the fixed-order reductions (DT_REDUCE_BLOCKS)
REDUCE_BLOCKS = 256
# floats of the Kabsch-angle state (DT_KABSCH_STATE); the angle is entry 47
KABSCH_STATE = 64


class DtRopeParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("p_mass", ctypes.c_float),
                ("table_height", ctypes.c_float), ("floor_friction", ctypes.c_float),
                ("gravity", ctypes.c_float * 3), ("rest_length", ctypes.c_float),
                ("bending_length", ctypes.c_float), ("stretch_compliance", ctypes.c_float),
                ("bending_compliance", ctypes.c_float), ("stretch_relaxation", ctypes.c_float),
                ("bending_relaxation", ctypes.c_float)]


class DtMaterial(ctypes.Structure):
    _fields_ = [
        ("mu", ctypes.c_float), ("lam", ctypes.c_float), ("yield_stress", ctypes.c_float),
        ("dp_alpha", ctypes.c_float), ("kind", ctypes.c_int32), ("pad", ctypes.c_int32 * 3),
    ]


class DtFemParams(ctypes.Structure):
    _fields_ = [
        ("mu", ctypes.c_float), ("lam", ctypes.c_float), ("dt", ctypes.c_float),
        ("mass", ctypes.c_float),
    ]


class DtContactParams(ctypes.Structure):
    _fields_ = [
        ("kn", ctypes.c_float), ("kd", ctypes.c_float), ("kt", ctypes.c_float),
        ("mu_f", ctypes.c_float), ("norm_eps", ctypes.c_float), ("sdf_offset", ctypes.c_float),
        ("unclamped_vn", ctypes.c_int32),
    ]


# name -> (restype, argtypes)
_P = ctypes.POINTER
_SIGS = {
    "dt_last_error": (ctypes.c_char_p, []),
    "dt_version": (_c_int, []),
    "dt_timing_enable": (_c_int, [_c_int]),
    "dt_timing_read": (_c_int, [_c_int, _P(ctypes.c_double), _P(ctypes.c_int64)]),
    "dt_graphs_enable": (_c_int, [_c_int]),
    "dt_graph_stats": (_c_int, [_P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "dt_mpm_init_object": (_c_int, [_c_int, _vp, _P(ctypes.c_float), _P(ctypes.c_float), _vp,
                                    _vp, _vp, _vp]),
    "dt_mpm_compute_new_F": (_c_int, [_c_int, _vp, _vp, _vp, _c_float, _vp]),
    "dt_mpm_compute_new_F_adj": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_float, _vp]),
    "dt_mpm_svd": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_mpm_svd_adj": (_c_int, [_c_int] + [_vp] * 7 + [_vp]),
    "dt_mpm_p2g": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 10 + [_vp]),
    "dt_mpm_p2g_adj": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 7
                       + [_vp] * 3 + [_vp] * 7 + [_vp, _vp]),
    "dt_mpm_grid_occupy": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp]),
    "dt_mpm_grid_op": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_mpm_grid_op_adj": (_c_int, [_P(DtMpmParams)] + [_vp] * 8 + [_vp]),
    "dt_mpm_g2p": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_mpm_g2p_adj": (_c_int, [_P(DtMpmParams), _c_int] + [_vp] * 7 + [_vp]),
    "dt_mpm_substep_fwd": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 10 + [_vp]),
    "dt_mpm_substep_adj": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 16
                           + [_vp, _vp]),
    "dt_mpm_run_fwd": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                                _c_int, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
    "dt_mpm_run_adj": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                                _c_int] + [_vp] * 12 + [_vp, _sz, _vp, _vp]),
    "dt_mpm_advance": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 8
                       + [_c_int, _vp, _vp, _vp, _sz, _vp]),
    "dt_mpm_advance_static": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 4
                              + [_c_int, _vp]),
    "dt_mpm_workspace_bytes": (_c_int, [_P(DtMpmParams), _c_int, _P(_sz)]),
    "dt_mpm_workspace_history_bytes": (_c_int, [_P(DtMpmParams), _c_int, _c_int, _P(_sz)]),
    "dt_mpm_bin": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _sz, _vp]),
    "dt_mpm_bin_perm": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _P(_vp), _P(_vp)]),
    "dt_mpm_bin_slot_bytes": (_c_int, [_P(DtMpmParams), _c_int, _P(_sz)]),
    "dt_mpm_bin_copy": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _vp, _c_int, _vp]),
    "dt_mpm_active_bricks": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _P(_vp), _P(_vp)]),
    "dt_mpm_zero_nodes": (_c_int, [_P(DtMpmParams), _vp, _c_int, _vp]),
    "dt_mpm_brick_nodes_copy": (_c_int, [_P(DtMpmParams), _c_int, _P(_vp), _P(_vp), _P(_c_int),
                                         _c_int, _c_int, _vp, _vp]),
    "dt_mpm_p2g_binned": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 5
                          + [_vp, _sz, _vp]),
    "dt_mpm_grid_op4_tiles": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _vp, _vp, _vp]),
    "dt_mpm_tiles_sum": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _vp, _vp]),
    "dt_mpm_tiles_sum_zero": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _vp, _vp, _c_int, _vp]),
    "dt_mpm_g2p4_adj_binned": (_c_int, [_P(DtMpmParams), _c_int] + [_vp] * 6 + [_vp, _sz, _vp]),
    "dt_mpm_grid_op4_adj_tiles": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz, _vp, _vp, _vp,
                                           _vp]),
    "dt_mpm_grid_op4_adj_tiles_f": (_c_int, [_P(DtMpmParams), _c_int, _vp, _sz] + [_vp] * 5 + [_vp]),
    "dt_mpm_brick_count": (ctypes.c_int64, [_P(DtMpmParams)]),
    "dt_mpm_mark_bricks": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp]),
    "dt_mpm_zero4_bricks": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp]),
    "dt_mpm_grid_op4_bricks": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp, _vp]),
    "dt_mpm_grid_op4_adj_bricks": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp, _vp, _vp]),
    "dt_mpm_zero4": (_c_int, [ctypes.c_int64, _vp, _vp]),
    "dt_mpm_p2g_fused": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 6 + [_vp]),
    "dt_mpm_grid_op4": (_c_int, [_P(DtMpmParams), _vp, _vp, _vp, _vp]),
    "dt_mpm_g2p4": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_mpm_g2p4_adj": (_c_int, [_P(DtMpmParams), _c_int] + [_vp] * 7 + [_vp]),
    "dt_mpm_grid_op4_adj": (_c_int, [_P(DtMpmParams)] + [_vp] * 5 + [_vp]),
    "dt_mpm_p2g_fused_adj": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp] + [_vp] * 10
                             + [_vp, _vp]),
    "dt_mpm_cell_keys": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp]),
    "dt_fem_init_B": (_c_int, [_c_int, _vp, _vp, _vp, _vp]),
    "dt_fem_update": (_c_int, [_P(DtFemParams), _c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_fem_update_csr": (_c_int, [_P(DtFemParams), _c_int, _c_int] + [_vp] * 7 + [_vp]),
    "dt_fem_update_adj_csr": (_c_int, [_P(DtFemParams), _c_int, _c_int] + [_vp] * 8 + [_vp]),
    "dt_fem_update_adj": (_c_int, [_P(DtFemParams), _c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_fem_update2": (_c_int, [_P(DtFemParams), _c_int] + [_vp] * 10 + [_vp]),
    "dt_fem_update2_adj": (_c_int, [_P(DtFemParams), _c_int] + [_vp] * 10 + [_vp]),
    "dt_clamp": (_c_int, [ctypes.c_int64, _vp, _c_float, _vp]),
    "dt_contact_centroids": (_c_int, [_c_int, _vp, _vp, _vp, _vp]),
    "dt_contact_closest": (_c_int, [_P(DtMpmParams), _vp, _c_int, _c_int, _vp, _vp, _c_int,
                                    _c_int, _vp]),
    "dt_jaw_chain": (_c_int, [_P(DtJawParams), _c_int, _vp, _vp, _vp]),
    "dt_jaw_chain_adj": (_c_int, [_P(DtJawParams), _c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_jaw_control_vel": (_c_int, [_c_int, _c_int, _c_int, _vp, _vp, _c_float, _vp, _vp]),
    "dt_jaw_control_vel_adj": (_c_int, [_c_int, _c_int, _c_int, _vp, _c_float, _vp, _vp, _vp]),
    "dt_graph_replay": (_c_int, [_vp, _vp, _c_int, _P(ctypes.c_int32)]),
    "dt_graph_begin": (_c_int, [_vp, _P(ctypes.c_int32)]),
    "dt_graph_end": (_c_int, [_vp, _vp, _vp, _c_int]),
    "dt_graph_abort": (_c_int, [_vp]),
    "dt_set_launch_config": (_c_int, [_P(DtLaunchConfig)]),
    "dt_get_launch_config": (_c_int, [_P(DtLaunchConfig)]),
    "dt_dome_pose_control": (_c_int, [_c_int, _vp, _c_float, _vp, _vp, _vp]),
    "dt_dome_pose_control_adj": (_c_int, [_c_int, _vp, _c_float, _vp, _vp, _vp, _vp, _vp]),
    "dt_dome_control_vel": (_c_int, [_c_int, _vp, _vp, _c_float, _vp, _vp]),
    "dt_dome_control_vel_adj": (_c_int, [_c_int, _vp, _vp, _vp, _c_float, _vp, _vp, _vp, _vp]),
    "dt_mpm_grid_mass": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp]),
    "dt_mpm_grid_mass_adj": (_c_int, [_P(DtMpmParams), _c_int, _vp, _vp, _vp, _vp]),
    "dt_deform_target_sdf": (_c_int, [_c_int, _c_float, _vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_deform_loss": (_c_int, [ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_deform_loss_adj": (_c_int, [ctypes.c_int64, _vp, _vp, _vp, _c_float, _c_float, _vp, _vp,
                                    _vp]),
    "dt_grid_iou_terms": (_c_int, [ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]),
    "dt_kabsch_angle": (_c_int, [_c_int, _vp, _vp, _c_float, _vp, _vp, _vp, _vp]),
    "dt_kabsch_angle_adj": (_c_int, [_c_int, _vp, _vp, _c_float, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_rope_step": (_c_int, [_P(DtRopeParams), _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_rope_step_adj": (_c_int, [_P(DtRopeParams), _c_int] + [_vp] * 10 + [_vp]),
    "dt_rope_closest": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _c_int, _c_int, _vp]),
    "dt_rope_closest_batched": (_c_int, [_c_int, _vp, _c_int, _vp, _vp, _vp, _vp, _c_int, _c_int,
                                         _vp]),
    "dt_rope_collision": (_c_int, [_P(DtContactParams), _c_float, _c_int, _vp, _vp, _vp, _c_int,
                                   _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_rope_collision_adj": (_c_int, [_P(DtContactParams), _c_float, _c_int, _vp, _vp, _vp,
                                       _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp, _vp]),
    "dt_contact_tri_batches": (_c_int, [_c_int, _vp, _vp, _vp]),
    "dt_contact_closest_batched": (_c_int, [_P(DtMpmParams), _vp, _c_int, _vp, _c_int, _vp, _vp,
                                            _vp, _vp, _c_int, _c_int, _vp]),
    "dt_contact_closest_env": (_c_int, [_P(DtMpmParams), _vp, _c_int, _vp, _c_int, _vp, _vp,
                                        _c_int, _c_int, _vp]),
    "dt_contact_collision": (_c_int, [_P(DtMpmParams), _P(DtContactParams), _vp, _c_int, _vp,
                                      _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                      _P(ctypes.c_float), _vp, _vp, _vp]),
    "dt_contact_collision_adj": (_c_int, [_P(DtMpmParams), _P(DtContactParams), _vp, _c_int, _vp,
                                          _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                          _P(ctypes.c_float), _vp, _vp, _vp, _c_int, _vp, _c_int,
                                          _vp, _vp, _vp]),
    "dt_contact_collision_fx": (_c_int, [_P(DtMpmParams), _P(DtContactParams), _vp, _c_int, _vp,
                                         _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                         _P(ctypes.c_float), _vp, _c_int, _vp, _vp, _vp, _vp]),
    "dt_contact_collision_adj_fx": (_c_int, [_P(DtMpmParams), _P(DtContactParams), _vp, _c_int,
                                             _vp, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp,
                                             _vp, _P(ctypes.c_float), _vp, _vp, _vp, _c_int, _vp,
                                             _c_int, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_rope_collision_fx": (_c_int, [_P(DtContactParams), _c_float, _c_int, _vp, _vp, _vp,
                                      _c_int, _c_int, _c_int, _vp, _vp, _vp, _P(ctypes.c_float),
                                      _vp, _c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_rope_collision_adj_fx": (_c_int, [_P(DtContactParams), _c_float, _c_int, _vp, _vp, _vp,
                                          _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                          _P(ctypes.c_float), _vp, _vp, _vp, _vp, _c_int, _vp,
                                          _vp, _vp, _vp, _vp]),
    "dt_loss_com": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "dt_loss_com_adj": (_c_int, [_c_int, _c_int, _vp, _vp, _vp]),
    "dt_contact_force_sum": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_contact_force_sum_adj": (_c_int, [_c_int, _vp, _vp, _vp, _vp]),
    "dt_force_loss": (_c_int, [_c_int, _vp, _P(ctypes.c_float), _c_float, _c_float, _vp, _vp]),
    "dt_marker_offset": (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    "dt_marker_offset_adj": (_c_int, [_c_int, _vp, _vp, _vp, _vp]),
    "dt_markers_fwd": (_c_int, [_c_int, _vp, _vp, _vp, _P(ctypes.c_float), _P(ctypes.c_float),
                                _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_markers_adj": (_c_int, [_c_int, _vp, _vp, _P(ctypes.c_float), _P(ctypes.c_float), _c_int,
                                _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dt_sort_tmp_bytes": (ctypes.c_int64, [_c_int]),
    "dt_sort_by_key": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp]),
    "dt_gather_rows": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "dt_scatter_rows": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "dt_gather_rows_soa": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "dt_scatter_rows_soa": (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    "dt_permute_soa": (_c_int, [_c_int, _c_int, _vp, _c_int, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """Load the HIP library once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m difftactile_amd.csrc.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Entry points this binding expects the library to export."""
    return sorted(_SIGS)


def check(rc):
    if rc != 0:
        msg = lib().dt_last_error().decode(errors="replace")
        raise RuntimeError(f"difftactile HIP call failed ({rc}): {msg}")


def ptr(t, dtype=torch.float32):
    """Device pointer of a contiguous CUDA/HIP tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream


_ARG_DTYPES = (torch.float32, torch.int32, torch.uint8)


def _arg(a):
    if isinstance(a, torch.Tensor):
        if not a.is_cuda:
            raise ValueError("tensor arguments must live on the GPU")
        if not a.is_contiguous():
            raise ValueError("tensor arguments must be contiguous")
        if a.dtype not in _ARG_DTYPES:
            raise TypeError(f"unsupported tensor dtype {a.dtype}")
        return a.data_ptr()
    return a


def call(name, *args):
    """Call entry point `name`.  Tensor arguments are passed as device pointers;
    the `args` tuple keeps every tensor (temporaries included) alive until the
    launch has been enqueued, after which stream order protects the memory: a
    block freed here is only reused by work enqueued later on the same stream.
    Never pass `ptr()` of a temporary: the temporary is freed before the call
    and its block can be reused by the next argument's copy."""
    conv = [_arg(a) for a in args]
    check(getattr(lib(), name)(*conv))
