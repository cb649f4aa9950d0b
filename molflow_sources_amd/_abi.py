"""ctypes mirror of include/molflow_gpu.h (the C ABI of the tracer library).

Only plain structs and the library loader live here. The field order must
match the header exactly; tests check sizes against the library's own
struct_size fields.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmolflow_gpu.so")

MFG_ABI_VERSION = 1
MFG_OK = 0
MFG_ERR_LIMIT = 6
STATIC_MAX_RAYS = 65536       # MFG_STATIC_MAX_RAYS

DES_NONE, DES_UNIFORM, DES_COSINE, DES_COSINE_N, DES_ANGLEMAP = range(5)
PROFILE_NONE, PROFILE_U, PROFILE_V, PROFILE_ANGULAR, PROFILE_VELOCITY, PROFILE_ORT_VELOCITY, PROFILE_TAN_VELOCITY = range(7)
PROFILE_SIZE = 100

# event kinds of the per-particle trace
EV_DESORB, EV_BOUNCE, EV_ABSORB, EV_TRANS, EV_TELEPORT, EV_LINK, EV_LEAK, EV_VOLUME, EV_SCATTER = range(9)

_d, _i32, _u32, _u64, _u8 = C.c_double, C.c_int32, C.c_uint32, C.c_uint64, C.c_uint8


class FacetDesc(C.Structure):
    _fields_ = [
        ("sticking", _d), ("opacity", _d), ("temperature", _d), ("accomodation", _d),
        ("outgassing", _d), ("desorb_exponent", _d), ("diffuse_part", _d), ("specular_part", _d),
        ("cosine_exponent", _d), ("sojourn_freq", _d), ("sojourn_e", _d),
        ("tex_width_precise", _d), ("tex_height_precise", _d), ("am_theta_limit", _d),
        ("fh_distance_binsize", _d), ("fh_distance_max", _d), ("fh_time_binsize", _d), ("fh_time_max", _d),
        ("og_ratio_u", _d), ("og_ratio_v", _d), ("total_outgassing", _d),
        ("sticking_param", _i32), ("opacity_param", _i32), ("outgassing_param", _i32), ("temperature_param", _i32),
        ("desorb_type", _i32), ("is_2sided", _i32), ("super_idx", _i32), ("super_dest", _i32),
        ("teleport_dest", _i32), ("profile_type", _i32), ("is_moving", _i32), ("enable_sojourn", _i32),
        ("is_textured", _i32), ("count_des", _i32), ("count_abs", _i32), ("count_refl", _i32),
        ("count_trans", _i32), ("count_dir", _i32), ("am_record", _i32),
        ("am_phi_width", _u32), ("am_theta_lower_res", _u32), ("am_theta_higher_res", _u32),
        ("fh_record_bounce", _i32), ("fh_record_distance", _i32), ("fh_record_time", _i32),
        ("fh_bounce_binsize", _u32), ("fh_bounce_max", _u32),
        ("use_outgassing_map", _i32), ("og_map_width", _u32), ("og_map_height", _u32), ("reserved0", _i32),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("gas_mass", _d), ("half_life", _d), ("time_window", _d), ("low_flux_cutoff", _d),
        ("motion_vector1", _d * 3), ("motion_vector2", _d * 3), ("torque_ref", _d * 3),
        ("scatter_mass_ratio", _d), ("scatter_mean_free_path_cm", _d), ("scatter_cutoff_speed", _d),
        ("scatter_rho", _d),
        ("gh_distance_binsize", _d), ("gh_distance_max", _d), ("gh_time_binsize", _d), ("gh_time_max", _d),
        ("enable_decay", _i32), ("use_maxwell", _i32), ("calc_constant_flow", _i32), ("low_flux_mode", _i32),
        ("motion_type", _i32), ("enable_force", _i32), ("scatter_enabled", _i32), ("scatter_enable_cutoff", _i32),
        ("gh_record_bounce", _i32), ("gh_record_distance", _i32), ("gh_record_time", _i32),
        ("gh_bounce_binsize", _u32), ("gh_bounce_max", _u32),
        ("log_enabled", _i32), ("log_facet_id", _i32), ("log_limit", _u32),
    ]


class ModelView(C.Structure):
    _fields_ = [
        ("struct_size", _u32), ("abi_version", _u32), ("n_vertices", _u32), ("n_facets", _u32),
        ("n_structures", _u32), ("n_params", _u32), ("n_moments", _u32), ("reserved0", _u32),
        ("vertices", C.POINTER(_d)), ("facet_vertex_offset", C.POINTER(_u32)),
        ("facet_vertex_index", C.POINTER(_u32)), ("facets", C.POINTER(FacetDesc)),
        ("settings", C.POINTER(Settings)), ("param_offset", C.POINTER(_u32)),
        ("param_x", C.POINTER(_d)), ("param_y", C.POINTER(_d)),
        ("param_log_x", C.POINTER(_u8)), ("param_log_y", C.POINTER(_u8)),
        ("moment_time", C.POINTER(_d)), ("moment_window", C.POINTER(_d)),
        ("og_map_offset", C.POINTER(_u64)), ("og_map_values", C.POINTER(_d)),
        ("am_pdf_offset", C.POINTER(_u64)), ("am_pdf", C.POINTER(_u64)),
    ]


class GlobalCounters(C.Structure):
    _fields_ = [("nb_mc_hit", _u64), ("nb_hit_equiv", _d), ("nb_desorbed", _u64), ("nb_abs_equiv", _d),
                ("nb_leak_total", _u64), ("dist_traveled_total", _d), ("dist_traveled_full_hits_only", _d),
                ("nb_decayed", _u64)]


class TallyLayout(C.Structure):
    _fields_ = [("struct_size", _u32), ("n_facets", _u32), ("n_slots", _u32), ("n_structures", _u32),
                ("n_tex_cells", _u64), ("n_dir_cells", _u64), ("n_profile_slices", _u64),
                ("n_angle_cells", _u64), ("n_global_hist_bins", _u64), ("n_facet_hist_bins", _u64),
                ("log_capacity", _u64), ("leak_capacity", _u64)]


class TallyView(C.Structure):
    _fields_ = [("struct_size", _u32), ("reserved0", _u32),
                ("global_", C.c_void_p), ("facets", C.c_void_p), ("forces", C.c_void_p),
                ("textures", C.c_void_p), ("directions", C.c_void_p), ("profiles", C.c_void_p),
                ("angle_maps", C.c_void_p), ("global_hist", C.c_void_p), ("facet_hist", C.c_void_p),
                ("log", C.c_void_p), ("log_count", C.c_void_p), ("leak_pos_dir", C.c_void_p),
                ("leak_count", C.c_void_p)]


class TuneResult(C.Structure):
    _fields_ = [("waves", _u32), ("block", _u32), ("blocks_per_cu", _u32), ("n_hot", _u32), ("hits_per_s", _d),
                ("bvh_width", _u32), ("leaf_size", _u32)]


class StepResult(C.Structure):
    _fields_ = [("status", _i32), ("stopped", _i32), ("desorbed", _u64), ("hits", _u64), ("kernel_ms", _d)]


# numpy dtypes matching the tally structs
def _np_dtypes():
    import numpy as np
    facet = np.dtype([("nb_mc_hit", "<u8"), ("nb_hit_equiv", "<f8"), ("nb_desorbed", "<u8"),
                      ("nb_abs_equiv", "<f8"), ("sum_1_per_ort_velocity", "<f8"), ("sum_v_ort", "<f8"),
                      ("sum_1_per_velocity", "<f8"), ("reserved", "<f8")])
    tex = np.dtype([("count_equiv", "<f8"), ("sum_v_ort_per_area", "<f8"), ("sum_1_per_ort_velocity", "<f8")])
    prof = np.dtype([("count_equiv", "<f8"), ("sum_v_ort", "<f8"), ("sum_1_per_ort_velocity", "<f8")])
    dirc = np.dtype([("dir", "<f8", (3,)), ("count", "<u8")])
    log = np.dtype([("u", "<f8"), ("v", "<f8"), ("hit_theta", "<f8"), ("hit_phi", "<f8"), ("ori_ratio", "<f8"),
                    ("particle_decay_moment", "<f8"), ("time", "<f8"), ("velocity", "<f8")])
    glob = np.dtype([(n, "<u8" if t is _u64 else "<f8") for n, t in GlobalCounters._fields_])
    return facet, tex, prof, dirc, log, glob


def load_library(path: str | None = None) -> C.CDLL:
    """Load the HIP tracer library; raises if it has not been built."""
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"molflow GPU library not built: {p} is missing (run __graft_entry__.build())")
    lib = C.CDLL(p)
    vp, i32, u32, u64, dp = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.POINTER(C.c_double)
    sigs = {
        "mfg_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "mfg_destroy": (None, [vp]),
        "mfg_last_error": (C.c_char_p, [vp]),
        "mfg_abi_version": (C.c_int, []),
        "mfg_load_model": (C.c_int, [vp, C.POINTER(ModelView)]),
        "mfg_build_accel": (C.c_int, [vp, C.c_int, C.c_int]),
        "mfg_reset": (C.c_int, [vp, u64, u64]),
        "mfg_run": (C.c_int, [vp, u64, u64, C.POINTER(StepResult)]),
        "mfg_request_stop": (C.c_int, [vp]),
        "mfg_set_time_limit": (C.c_int, [vp, C.c_double]),
        "mfg_get_layout": (C.c_int, [vp, C.POINTER(TallyLayout)]),
        "mfg_get_facet_offsets": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64), C.c_size_t]),
        "mfg_read_tallies": (C.c_int, [vp, C.POINTER(TallyView)]),
        "mfg_take_tallies": (C.c_int, [vp, C.POINTER(TallyView)]),
        "mfg_clear_tallies": (C.c_int, [vp]),
        "mfg_reinit_particle_log": (C.c_int, [vp]),
        "mfg_sanity_check": (C.c_int, [C.POINTER(ModelView), C.c_char_p, C.c_size_t, C.POINTER(u32)]),
        "mfg_model_totals": (C.c_int, [C.POINTER(ModelView), dp]),
        "mfg_hits_size": (C.c_int, [vp, C.POINTER(u64)]),
        "mfg_intersect_batch": (C.c_int, [vp, u64, dp, dp, C.POINTER(i32), i32, C.POINTER(i32), dp, dp, dp]),
        "mfg_intersect_batch_static": (C.c_int, [vp, u64, dp, dp, C.POINTER(i32), i32, C.POINTER(i32), dp, dp, dp]),
        "mfg_maxwell_cdf": (C.c_int, [C.c_double, dp, dp, C.c_size_t]),
        "mfg_parse_user_moment": (C.c_int, [C.c_char_p, C.c_double, dp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "mfg_integrated_desorption": (C.c_int, [dp, dp, C.c_size_t, C.c_int, C.c_int, C.c_double, dp, dp,
                                                C.c_size_t, C.POINTER(C.c_size_t)]),
        "mfg_set_trace": (C.c_int, [vp, u32, u32]),
        "mfg_get_trace": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.c_size_t]),
        "mfg_load_variants": (C.c_int, [vp, C.POINTER(C.POINTER(ModelView)), u32]),
        "mfg_run_variants": (C.c_int, [vp, C.POINTER(u64), C.POINTER(u64), u64, C.POINTER(StepResult)]),
        "mfg_get_variant_layout": (C.c_int, [vp, u32, C.POINTER(TallyLayout)]),
        "mfg_read_variant_tallies": (C.c_int, [vp, u32, C.POINTER(TallyView)]),
        "mfg_set_launch": (C.c_int, [vp, u32, u32]),
        "mfg_enable_graph": (C.c_int, [vp, C.c_int]),
        "mfg_set_kernel_variant": (C.c_int, [vp, u32]),
        "mfg_set_carry": (C.c_int, [vp, C.c_int]),
        "mfg_in_flight": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "mfg_autotune": (C.c_int, [vp, u64, C.POINTER(TuneResult)]),
        "mfg_adjoint_setup": (C.c_int, [vp, C.POINTER(i32), u32, C.POINTER(i32), u32]),
        "mfg_adjoint_set_sticking": (C.c_int, [vp, dp, u32]),
        "mfg_adjoint_read": (C.c_int, [vp, dp, dp]),
        "mfg_adjoint_backward": (C.c_int, [vp, dp, dp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


EXPORTED_SYMBOLS = (
    "mfg_create", "mfg_destroy", "mfg_last_error", "mfg_abi_version", "mfg_load_model", "mfg_build_accel",
    "mfg_reset", "mfg_run", "mfg_request_stop", "mfg_set_time_limit", "mfg_get_layout", "mfg_get_facet_offsets", "mfg_read_tallies",
    "mfg_take_tallies", "mfg_clear_tallies", "mfg_reinit_particle_log", "mfg_sanity_check", "mfg_model_totals",
    "mfg_hits_size", "mfg_intersect_batch", "mfg_intersect_batch_static", "mfg_maxwell_cdf", "mfg_parse_user_moment",
    "mfg_integrated_desorption", "mfg_set_trace", "mfg_get_trace", "mfg_load_variants", "mfg_run_variants",
    "mfg_get_variant_layout", "mfg_read_variant_tallies", "mfg_set_launch", "mfg_enable_graph",
    "mfg_set_kernel_variant", "mfg_autotune", "mfg_set_carry", "mfg_in_flight",
    "mfg_adjoint_setup", "mfg_adjoint_set_sticking", "mfg_adjoint_read", "mfg_adjoint_backward",
)
