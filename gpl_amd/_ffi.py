"""ctypes declarations of include/gpl_hip.h (the C ABI of libgpl_hip.so).

The library is loaded from this package's directory only; a missing library is an error, never a
fallback: there is no CPU path in this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpl_hip.so")

GPL_OK = 0
GPL_ERR_ARG = -1
GPL_ERR_HIP = -2
GPL_ERR_STATE = -3
GPL_ERR_REFUSED = -4

GPL_CELL_INSTANCE = 1
GPL_CELL_MACRO = 2
GPL_CELL_LOCKED = 4

GPL_SLOT_CUR_SLP = 0
GPL_SLOT_NEXT_SLP = 1
GPL_SLOT_PREV_SLP = 2
GPL_SLOT_CUR = 3
GPL_SLOT_NEXT = 4

i32, i64, u32, f32, f64, u8 = C.c_int32, C.c_int64, C.c_uint32, C.c_float, C.c_double, C.c_uint8
P = C.POINTER


class Opts(C.Structure):
    _fields_ = [
        ("density", f32), ("uniform_density", i32), ("bin_cnt", i32), ("init_density_penalty", f32),
        ("init_wirelength_coef", f32), ("min_phi_coef", f32), ("max_phi_coef", f32), ("overflow", f32),
        ("reference_hpwl", f32), ("max_iter", i32), ("seed", i32), ("perturb_dist", f32),
        ("disable_revert_if_diverge", i32), ("pad_left", i32), ("pad_right", i32), ("skip_io", i32),
        ("disable_pin_density_adjust", i32), ("initial_place_max_iter", i32), ("initial_place_max_fanout", i32),
        ("initial_place_min_diff_length", i32), ("initial_place_max_solver_iter", i32),
        ("initial_place_net_weight_scale", f32), ("skip_initial_place", i32),
        ("routability", i32), ("routability_end_overflow", f32), ("routability_snapshot_overflow", f32),
        ("routability_target_rc", f32), ("routability_inflation_ratio_coef", f32), ("routability_max_inflation_ratio", f32),
        ("routability_max_density", f32), ("routability_min_congestion", f32), ("routability_max_inflation_total", f32),
        ("routability_net_weight_max", f32), ("routability_congested_nets_percentage", f32), ("routability_rc_k", f32 * 4),
        ("routability_max_passes", i32), ("route_gcells", i32), ("route_tracks", i32),
    ]


class GenConfig(C.Structure):
    _fields_ = [
        ("num_cells", i32), ("num_nets", i32), ("degree_p", f64), ("degree_cap", i32), ("num_hf_nets", i32),
        ("hf_min_pins", i32), ("hf_max_pins", i32), ("num_io", i32), ("utilization", f64), ("macro_area_frac", f64),
        ("macro_rows_min", i32), ("macro_rows_max", i32), ("row_height", i32), ("site_width", i32),
        ("width_min_sites", i32), ("width_max_sites", i32), ("core_rows", i32), ("core_sites", i32), ("dbu", i32),
        ("locality_window", i32), ("seed", u32), ("num_groups", i32), ("group_cells", P(i32)),
    ]


class DesignInfo(C.Structure):
    _fields_ = [
        ("dbu", i32), ("die", i32 * 4), ("core", i32 * 4), ("num_insts", i32), ("num_movable", i32), ("num_fixed", i32),
        ("num_nets", i32), ("num_pins", i32), ("num_regions", i32), ("movable_area", i64), ("fixed_area", i64),
        ("initial_hpwl", i64), ("ip_rounds", i32), ("ip_error", f32),
    ]


class Region(C.Structure):
    _fields_ = [
        ("lx", i32), ("ly", i32), ("ux", i32), ("uy", i32), ("cnt_x", i32), ("cnt_y", i32), ("size_x", f64),
        ("size_y", f64), ("target_density", f32), ("uniform_target_density", f32), ("cell_begin", i32),
        ("cell_end", i32), ("num_fillers", i32), ("filler_dx", i32), ("filler_dy", i32), ("std_insts_area", i64),
        ("macro_insts_area", i64), ("non_place_insts_area", i64), ("white_space_area", i64), ("movable_area", i64),
        ("total_filler_area", i64), ("non_place", P(i64)), ("non_place_unscaled", P(i64)), ("bin_x", P(i32)),
        ("bin_y", P(i32)), ("wx", P(f32)), ("wy", P(f32)), ("group", i32), ("non_place_raw", P(i64)),
        ("non_place_union", P(i64)),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("num_cells", i32), ("cell_box", P(i32)), ("cell_dbox", P(i32)), ("density_scale", P(f32)),
        ("cell_flags", P(u8)), ("cell_inst", P(i32)), ("cell_pin_start", P(i32)), ("cell_pins", P(i32)),
        ("num_pins", i32), ("pin_cell", P(i32)), ("pin_offset", P(i32)), ("pin_xy", P(i32)), ("pin_net", P(i32)),
        ("num_nets", i32), ("net_pin_start", P(i32)), ("net_pins", P(i32)), ("net_weight", P(f32)),
        ("num_regions", i32), ("regions", P(Region)), ("opts", Opts), ("dbu", i32), ("site_x", i32), ("site_y", i32),
        ("inst_box", P(i32)), ("num_insts", i32), ("die", i32 * 4),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iter", i32), ("converged", i32), ("status", i32), ("total_trials", i32), ("hpwl", i64), ("overflow", f32),
        ("overflow_scaled", f32), ("density_penalty", f32), ("step_length", f32), ("wl_coef", f32),
        ("base_wl_coef", f32), ("reference_hpwl", f32), ("wl_grad_sum", f32), ("density_grad_sum", f32),
        ("num_diverge_reverts", i32), ("init_seconds", f64), ("loop_seconds", f64),
        ("first_below_target_seconds", f64),
    ]


class IterRow(C.Structure):
    _fields_ = [
        ("iter", i32), ("trials", i32), ("overflow", f32), ("density_penalty", f32), ("step_length", f32),
        ("wl_coef", f32), ("hpwl", i64),
    ]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("launches", i32), ("ms", f64)]


class RudyGrid(C.Structure):
    _fields_ = [("lx", i32), ("ly", i32), ("ux", i32), ("uy", i32), ("cnt_x", i32), ("cnt_y", i32), ("tile_size", i32),
                ("wire_width", i32)]


class RouteVars(C.Structure):
    _fields_ = [("min_congestion_for_inflation", f32), ("inflation_ratio_coef", f32), ("max_inflation_ratio", f32),
                ("ignore_edge_ratio", f32), ("rc_k", f32 * 4)]


class RudyStats(C.Structure):
    _fields_ = [("total_overflow", f64), ("overflowed_tiles", i32), ("actionable_tiles", i32), ("top", f64 * 4), ("rc", f32)]


class RoutePass(C.Structure):
    _fields_ = [("iter", i32), ("rc", f32), ("overflowed_tiles", i32), ("actionable_tiles", i32), ("inflation", i64),
                ("target_density", f32), ("fillers", i32), ("weighted_nets", i32), ("need", i32), ("revert", i32)]


class RouteStats(C.Structure):
    _fields_ = [("passes", i32), ("route_iters", i32), ("final_rc", f32), ("min_rc", f32), ("total_inflation", i64),
                ("original_area", i64)]


# (name, restype, argtypes) of every exported function declared in include/gpl_hip.h
class DensityIO(C.Structure):
    """gpl_density_io: device pointers (as integers) of gpl_density_scatter_noalloc's arrays."""
    _fields_ = [
        ("num_cells", i32), ("dbox", C.c_void_p), ("density_scale", C.c_void_p), ("cell_flags", C.c_void_p),
        ("lx", i32), ("ly", i32), ("ux", i32), ("uy", i32), ("cnt_x", i32), ("cnt_y", i32),
        ("size_x", f64), ("size_y", f64), ("target_density", f32), ("bin_x", C.c_void_p), ("bin_y", C.c_void_p),
        ("non_place_f", C.c_void_p), ("inst_area", C.c_void_p), ("filler_area", C.c_void_p), ("density", C.c_void_p),
        ("overflow", C.c_void_p), ("stream", C.c_void_p),
    ]


SIGNATURES = [
    ("gpl_opts_default", None, [P(Opts)]),
    ("gpl_design_generate", i32, [P(GenConfig), P(Opts), P(C.c_void_p)]),
    ("gpl_design_destroy", None, [C.c_void_p]),
    ("gpl_design_initial_place", i32, [C.c_void_p]),
    ("gpl_design_initial_place_gpu", i32, [C.c_void_p]),
    ("gpl_design_get_info", i32, [C.c_void_p, P(DesignInfo)]),
    ("gpl_init_nesterov_place", i32, [C.c_void_p, P(Opts)]),
    ("gpl_design_problem", i32, [C.c_void_p, P(Problem)]),
    ("gpl_ctx_create", i32, [P(Problem), P(C.c_void_p)]),
    ("gpl_ctx_destroy", None, [C.c_void_p]),
    ("gpl_last_error", C.c_char_p, [C.c_void_p]),
    ("gpl_init_nesterov_loop", i32, [C.c_void_p]),
    ("gpl_run_nesterov_loop", i32, [C.c_void_p, i32, P(i32)]),
    ("gpl_do_nesterov_place", i32, [C.c_void_p, P(i32)]),
    ("gpl_place_batch", i32, [P(C.c_void_p), i32, P(i32)]),
    ("gpl_continue_nesterov_place", i32, [C.c_void_p, f32, i32, i32, P(i32)]),
    ("gpl_get_stats", i32, [C.c_void_p, P(Stats)]),
    ("gpl_loop_begin", i32, [C.c_void_p, i32]),
    ("gpl_loop_step", i32, [C.c_void_p, i32, P(i32), P(i32)]),
    ("gpl_loop_end", i32, [C.c_void_p, P(i32)]),
    ("gpl_get_history", i32, [C.c_void_p, P(IterRow), i32]),
    ("gpl_get_log", i32, [C.c_void_p, C.c_char_p, i32]),
    ("gpl_placed_locations", i32, [C.c_void_p, P(i32), i32]),
    ("gpl_set_coordinates", i32, [C.c_void_p, i32, i32, P(f32)]),
    ("gpl_get_coordinates", i32, [C.c_void_p, i32, i32, P(f32)]),
    ("gpl_density_scatter", i32, [C.c_void_p, i32, i32, P(i64), P(i64), P(f32), P(f64)]),
    ("gpl_density_solve", i32, [C.c_void_p, i32, P(f32), P(f32), P(f32)]),
    ("gpl_set_density", i32, [C.c_void_p, i32, P(f32)]),
    ("gpl_wa_forward", i32, [C.c_void_p, f32, f32, P(f32), P(i32), P(i64)]),
    ("gpl_wa_backward", i32, [C.c_void_p, i32, P(f32)]),
    ("gpl_density_gather", i32, [C.c_void_p, i32, P(f32)]),
    ("gpl_update_gradients", i32, [C.c_void_p, i32, i32, f32, P(f32), P(f32), P(f32), P(f64)]),
    ("gpl_nesterov_step", i32, [C.c_void_p, i32, f32, f32, P(f32), P(f32)]),
    ("gpl_objective_forward", i32, [C.c_void_p, P(f32), f32, f32, P(f64), P(f64), P(f64)]),
    ("gpl_objective_backward", i32, [C.c_void_p, P(f32), f32, f32, P(f32)]),
    ("gpl_rudy_wire_width", i32, [P(i32), i32]),
    ("gpl_rudy", i32, [C.c_void_p, P(RudyGrid), P(C.c_uint8), P(f32), P(f32)]),
    ("gpl_rudy_tiles", i32, [C.c_void_p, P(RudyGrid), P(RouteVars), P(f32), P(C.c_uint8), P(RudyStats)]),
    ("gpl_route_inflate", i32, [C.c_void_p, i32, P(RudyGrid), P(f32), P(i32), P(C.c_int64)]),
    ("gpl_update_density_size", i32, [C.c_void_p, i32, P(i32)]),
    ("gpl_get_density_boxes", i32, [C.c_void_p, i32, P(i32), P(f32)]),
    ("gpl_get_route_stats", i32, [C.c_void_p, P(RouteStats), P(RoutePass), i32]),
    ("gpl_debug_inject", i32, [C.c_void_p, i32, i32]),
    ("gpl_kernel_times", i32, [C.c_void_p, P(KernelTime), i32]),
    ("gpl_set_profiling", i32, [C.c_void_p, i32]),
    ("gpl_util_mt19937", None, [u32, i32, P(u32)]),
    ("gpl_util_uniform_int", None, [u32, i32, P(i32), P(i32), P(i32)]),
    ("gpl_util_normal", u32, [u32, C.c_double, C.c_double, i32, P(f64)]),
    ("gpl_util_wavenumber_x", f32, [i32, i32]),
    ("gpl_util_wavenumber_y", f32, [i32, i32, f32, f32]),
    ("gpl_util_union_area", i64, [P(i32), i32]),
    ("gpl_util_round_down_pow2", u32, [u32]),
    ("gpl_util_struct_sizes", i32, [P(i64), i32]),
    ("gpl_set_launch_mode", i32, [C.c_void_p, i32]),
    ("gpl_set_resort_policy", i32, [C.c_void_p, i32, f32]),
    ("gpl_set_tuning", i32, [C.c_void_p, i32, i32, i32, i32]),
    ("gpl_get_tuning", i32, [C.c_void_p, P(i32), i32]),
    ("gpl_density_scatter_noalloc", i32, [P(DensityIO)]),
    ("gpl_design_write_locations", i32, [C.c_void_p, P(i32), i32, P(i64)]),
    ("gpl_design_bases", i32, [C.c_void_p, P(i32), i32]),
    ("gpl_design_write_def", i32, [C.c_void_p, C.c_char_p]),
]

_lib = None


def lib():
    """The loaded libgpl_hip.so, with every signature declared.  Raises when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m gpl_amd.build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return [name for name, _, _ in SIGNATURES]
