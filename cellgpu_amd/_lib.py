"""Loader of the native library libcellgpu_amd.so (HIP kernels + C++ host layer).

There is no fallback: if the library is missing or no GPU is visible, calls
raise. The C ABI is declared in include/cellgpu_amd.h.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcellgpu_amd.so")
_LIB = None


class CellGPUError(RuntimeError):
    pass


STATUS = {
    0: "ok", -1: "bad arguments", -2: "HIP error", -3: "Delaunay degree exceeds ring capacity",
    -4: "cell-list bin overflow", -5: "topology inconsistent", -6: "non-finite position", -7: "no device",
    -8: "input exceeds the static scratch",
}


class SysParams(C.Structure):
    _fields_ = [("off", C.c_int), ("n", C.c_int), ("bin_off", C.c_int), ("xsize", C.c_int),
                ("L", C.c_double), ("boxsize", C.c_double), ("KA", C.c_double), ("KP", C.c_double),
                ("dt", C.c_double), ("mu", C.c_double), ("T", C.c_double), ("seed", C.c_uint64),
                ("ncells", C.c_int), ("tension", C.c_int), ("gamma", C.c_double)]


class SysState(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "any_fail", "nf_count", "mode", "need_global", "nn_delta", "tri_delta", "max_deg", "neigh_max", "step",
        "active", "steps_left", "error", "nflag", "pad_s")] + [
        (n, C.c_longlong) for n in ("global_fixes", "repair_events", "local_repairs", "t1_events")] + [
        ("tri_hash_all", C.c_ulonglong), ("tri_hash_own", C.c_ulonglong), ("tri_area", C.c_longlong),
        ("skip_off", C.c_int), ("pad_k", C.c_int)]


class VorWorkspace(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("num_systems", "num_cells", "total_bins", "num_chunks", "num_bin_chunks",
                                       "precision")] + [("bytes", C.c_longlong)]


class FireState(C.Structure):
    """Mirror of cgpu_fire_state (include/cellgpu_amd.h)."""

    _fields_ = [(n, C.c_double) for n in (
        "alpha", "alpha_start", "alpha_dec", "dt", "dt_max", "dt_inc", "dt_dec", "force_cutoff", "power",
        "force_max", "scaling", "mix_alpha", "max_f2", "pad_d")] + [(n, C.c_int) for n in (
            "nmin", "nsince", "iterations", "max_iterations", "converged", "zero_v", "pending", "pad_i")]


class VertexBatchView(C.Structure):
    """Mirror of cgpu_vertex_batch (include/cellgpu_amd.h)."""

    _fields_ = [("num_systems", C.c_int), ("num_vertices", C.c_int), ("num_cells", C.c_int), ("precision", C.c_int)] + [
        (name, C.c_void_p) for name in (
            "sys", "st", "vert_sys", "cell_sys", "pos", "pos_next", "vn", "vcn", "cvn", "cv", "AP", "dAP", "APpref",
            "force", "flag_count", "flag_list", "flag_cells", "flag_touched", "cell_min", "noise", "theta",
            "motility", "chunk_sys", "chunk_begin", "chunk_end")] + [("num_chunks", C.c_int), ("vertex_cap", C.c_int)]


PROFILE_SLOTS = 14  # CGPU_PROFILE_SLOTS


class LaunchShape(C.Structure):
    """Mirror of cgpu_launch_shape (include/cellgpu_amd.h)."""

    _fields_ = [(n, C.c_int) for n in (
        "force_block", "test_block", "cell_block", "star_blocks", "repair_blocks", "queue_lanes", "force_waves",
        "pad_shape")]


class VoronoiBatchView(C.Structure):
    _fields_ = [("num_systems", C.c_int), ("num_cells", C.c_int), ("total_bins", C.c_int), ("precision", C.c_int)] + [
        (name, C.c_void_p) for name in (
            "sys", "st", "cell_sys", "pos", "pos_next", "nn", "nbr", "oth", "AP", "dAP", "APpref", "mot", "vel",
            "theta", "force", "excl", "ext_force", "bin_count", "bin_start", "cell_bin", "cell_rank", "sorted_id",
            "sorted_pos", "repair", "mark", "flagged", "nfq", "gq",
            "counters", "noise", "chunk_sys", "chunk_begin", "chunk_end", "bchunk_sys", "bchunk_begin",
            "bchunk_end", "sys_bchunk", "bchunk_sum")] + [("num_chunks", C.c_int), ("num_bin_chunks", C.c_int)] + [
        ("cell_type", C.c_void_p), ("tension", C.c_void_p), ("num_types", C.c_int), ("ring_cap", C.c_int),
        ("ld_tri", C.c_void_p), ("bin_mark_wide", C.c_void_p), ("cell_rmax", C.c_void_p),
        ("ld_regions", C.c_void_p), ("ld_region_cap", C.c_int), ("ld_tri_cap", C.c_int),
        ("ld_part", C.c_void_p),
        ("shape", LaunchShape)]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise CellGPUError("libcellgpu_amd.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    P = C.POINTER
    sig = {
        "cgpu_device_count": (i, []),
        "cgpu_last_error": (C.c_char_p, []),
        "cgpu_set_block_size": (i, [i]),
        "cgpu_vor_create": (vp, [i, vp, vp, vp, vp, vp, vp, i, i]),
        "cgpu_vor_destroy": (None, [vp]),
        "cgpu_vor_set_params": (i, [vp, i, d, d, d, d, d]),
        "cgpu_vor_set_prefs": (i, [vp, vp, vp]),
        "cgpu_vor_set_motility": (i, [vp, vp, vp]),
        "cgpu_vor_set_exclusions": (i, [vp, vp]),
        "cgpu_vor_set_cell_types": (i, [vp, vp]),
        "cgpu_vor_set_tension": (i, [vp, i, i, d]),
        "cgpu_vor_set_tension_matrix": (i, [vp, i, vp]),
        "cgpu_vor_set_positions": (i, [vp, vp, vp, i]),
        "cgpu_vor_set_noise_mode": (i, [vp, i]),
        "cgpu_vor_upload_noise": (i, [vp, vp]),
        "cgpu_vor_compute_forces": (i, [vp]),
        "cgpu_vor_enforce_topology": (i, [vp]),
        "cgpu_vm_enforce_topology": (i, [vp]),
        "cgpu_vor_spv_steps": (i, [vp, i, i]),
        "cgpu_vor_brownian_steps": (i, [vp, i]),
        "cgpu_vor_sync": (i, [vp]),
        "cgpu_vor_status": (i, [vp]),
        "cgpu_vor_get": (i, [vp] + [vp] * 9),
        "cgpu_vor_get_topology": (i, [vp, vp, vp, i, vp]),
        "cgpu_vor_get_counters": (i, [vp, i, vp]),
        "cgpu_vor_num_cells": (i, [vp]),
        "cgpu_vor_energy": (d, [vp, i]),
        "cgpu_vor_set_active": (i, [vp, vp, vp]),
        "cgpu_vor_get_active": (i, [vp, vp, vp]),
        "cgpu_vor_device_bytes": (C.c_longlong, [vp]),
        "cgpu_vor_merge": (vp, [i, vp, vp]),
        "cgpu_vor_get_external_forces": (i, [vp, vp, vp]),
        "cgpu_vor_autotune": (i, [vp, i, P(LaunchShape), P(i)]),
        "cgpu_vor_set_launch_shape": (i, [vp, P(LaunchShape)]),
        "cgpu_vor_get_launch_shape": (i, [vp, P(LaunchShape)]),
        "cgpu_vor_spv_steps_profiled": (i, [vp, i, vp, vp]),
        "cgpu_profile_slot_name": (C.c_char_p, [i]),
        "cgpu_vor_view": (i, [vp, P(VoronoiBatchView)]),
        "cgpu_vor_default_params": (i, [i, vp, vp, vp, P(SysParams)]),
        "cgpu_vor_workspace_query": (i, [i, P(SysParams), i, P(VorWorkspace)]),
        "cgpu_vor_workspace_bind": (i, [P(VorWorkspace), P(SysParams), vp, C.c_size_t, P(VoronoiBatchView), vp]),
        "cgpu_vm_view": (i, [vp, P(VertexBatchView)]),
        "cgpu_vm_stream": (vp, [vp]),
        "cgpu_vor_stream": (vp, [vp]),
        "cgpu_compute_cell_list": (i, [P(VoronoiBatchView), i, vp]),
        "cgpu_test_and_geometry": (i, [P(VoronoiBatchView), i, vp]),
        "cgpu_repair_topology": (i, [P(VoronoiBatchView), vp]),
        "cgpu_global_triangulation": (i, [P(VoronoiBatchView), vp]),
        "cgpu_voronoi_forces": (i, [P(VoronoiBatchView), vp]),
        "cgpu_spv_force_eom_move": (i, [P(VoronoiBatchView), vp]),
        "cgpu_spv_step": (i, [P(VoronoiBatchView), vp]),
        "cgpu_brownian_step": (i, [P(VoronoiBatchView), vp]),
        "cgpu_brownian_force_eom_move": (i, [P(VoronoiBatchView), vp]),
        "cgpu_fill_normal": (i, [P(VoronoiBatchView), i, i, vp, vp]),
        "cgpu_step_begin": (i, [P(VoronoiBatchView), vp]),
        "cgpu_step_end": (i, [P(VoronoiBatchView), vp]),
        "cgpu_device_synchronize": (i, []),
        "cgpu_vor_spv_steps_timed": (i, [vp, i, vp]),
        "cgpu_philox4x32_10": (None, [vp, vp, vp]),
        "cgpu_normal_pair": (None, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, vp]),
        "cgpu_put_in_box": (None, [d, P(d), P(d)]),
        "cgpu_vor_spatial_sort": (i, [vp, vp]),
        "cgpu_fire_create": (vp, [vp]),
        "cgpu_fire_destroy": (None, [vp]),
        "cgpu_fire_set": (i, [vp, i, d, d, d, d, d, d, d, i]),
        "cgpu_fire_minimize": (i, [vp]),
        "cgpu_fire_get": (i, [vp, i, vp]),
        "cgpu_vm_create": (vp, [i, vp, vp, vp, vp, vp, i, i, i]),
        "cgpu_vm_destroy": (None, [vp]),
        "cgpu_vm_num_vertices": (i, [vp]),
        "cgpu_vm_set_params": (i, [vp, i, d, d, d, d, d, d]),
        "cgpu_vm_set_prefs": (i, [vp, vp, vp]),
        "cgpu_vm_set_positions": (i, [vp, vp, vp]),
        "cgpu_vm_set_noise_mode": (i, [vp, i]),
        "cgpu_vm_upload_noise": (i, [vp, vp]),
        "cgpu_vm_compute_forces": (i, [vp]),
        "cgpu_vm_brownian_steps": (i, [vp, i, i]),
        "cgpu_vm_get": (i, [vp] + [vp] * 6),
        "cgpu_vm_get_topology": (i, [vp, vp, vp, vp, vp, i]),
        "cgpu_vm_get_counters": (i, [vp, i, vp]),
        "cgpu_vm_status": (i, [vp]),
        "cgpu_vm_upload_cell_noise": (i, [vp, vp]),
        "cgpu_vm_set_motility": (i, [vp, vp, vp]),
        "cgpu_vm_set_directors": (i, [vp, vp]),
        "cgpu_vm_get_directors": (i, [vp, vp]),
        "cgpu_vm_spp_steps": (i, [vp, i, i]),
        "cgpu_vm_fire_create": (vp, [vp]),
        "cgpu_vm_set_topology": (i, [vp, vp, vp, vp, vp, i]),
        "cgpu_vm_cell_events": (i, [vp, i, vp, vp, vp, vp, vp]),
        "cgpu_vm_spatial_sort": (i, [vp, vp, vp]),
        "cgpu_vor_ring_capacity": (i, [vp]),
        "cgpu_vm_vertex_capacity": (i, [vp]),
        "cgpu_global_triangulation_active": (i, [P(VoronoiBatchView), vp]),
        "cgpu_vor_sort_workspace": (C.c_size_t, [i]),
        "cgpu_vor_hilbert_order": (i, [P(VoronoiBatchView), vp, vp, C.c_size_t, vp]),
        "cgpu_vor_permute_cells": (i, [P(VoronoiBatchView), vp, vp, vp]),
        "cgpu_vm_set_active": (i, [vp, vp]),
        "cgpu_vm_get_active": (i, [vp, vp]),
        "cgpu_fire_set_param": (i, [vp, i, d]),
        "cgpu_vm_get_sizes": (i, [vp, vp, vp]),
        "cgpu_vor_set_directors": (i, [vp, vp]),
        "cgpu_vor_set_velocities": (i, [vp, vp, vp]),
        "cgpu_vor_set_topology": (i, [vp, vp, vp, i, vp]),
        "cgpu_vor_get_del_other": (i, [vp, vp, i]),
        "cgpu_derive_topology": (i, [P(VoronoiBatchView), vp]),
        "cgpu_vor_set_step_counters": (i, [vp, vp]),
        "cgpu_vor_get_tags": (i, [vp, vp]),
        "cgpu_vor_cell_events": (i, [vp, i, vp, vp, vp, vp, vp]),
        "cgpu_vor_get_sizes": (i, [vp, vp]),
        "cgpu_db_open": (vp, [C.c_char_p, i, i, i, i]),
        "cgpu_db_close": (None, [vp]),
        "cgpu_db_last_error": (C.c_char_p, []),
        "cgpu_db_records": (i, [vp]),
        "cgpu_db_spv_write": (i, [vp, i, d, d, d] + [vp] * 6),
        "cgpu_db_spv_read": (i, [vp, i] + [vp] * 8),
        "cgpu_db_avm_write": (i, [vp, i, d, d, d] + [vp] * 10),
        "cgpu_db_avm_read": (i, [vp, i] + [vp] * 9),
        "cgpu_db_spv_write_batch": (i, [vp, vp, i, d, i]),
        "cgpu_db_spv_read_batch": (i, [vp, vp, i, i]),
        "cgpu_db_avm_write_batch": (i, [vp, vp, i, d, i]),
        "cgpu_db_avm_read_batch": (i, [vp, vp, i, i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)  # a missing export is a build error: AttributeError
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(status, what="call"):
    if status != 0:
        msg = STATUS.get(status, "status %d" % status)
        extra = lib().cgpu_last_error()
        raise CellGPUError("%s failed: %s %s" % (what, msg, extra.decode() if extra else ""))
    return status


def require_device():
    n = lib().cgpu_device_count()
    if n < 1:
        raise CellGPUError("no HIP device visible: cellgpu_amd runs on the GPU only")
    return n
