"""TEST INFRASTRUCTURE -- ctypes access to the oracle (liboracle.so, built from oracle.c and
dct.c by the Makefile here).  Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, as the checker; the product (gpl_amd/) never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = [
            ("orc_fast_exp", C.c_float, [C.c_float]),
            ("orc_distance", C.c_float, [f32p, f32p, C.c_int32]),
            ("orc_derived_reference_hpwl", C.c_float, [C.c_float, C.c_int64]),
            ("orc_wirelength_coef_factor", C.c_float, [C.c_float]),
            ("orc_phi_coef", C.c_float, [C.c_float, C.c_float, C.c_float]),
            ("orc_diverge_fallback", C.c_int32, [C.c_int32, C.c_int32, C.c_int32]),
            ("orc_max_iter_takes_snapshot", C.c_int32,
             [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_float, C.c_float, C.c_int64]),
            ("orc_dct2_f32", None, [C.c_int, f32p]),
            ("orc_dct3_f32", None, [C.c_int, f32p]),
            ("orc_dst3_f32", None, [C.c_int, f32p]),
            ("orc_dct2_f64", None, [C.c_int, f64p]),
            ("orc_dct3_f64", None, [C.c_int, f64p]),
            ("orc_dst3_f64", None, [C.c_int, f64p]),
            ("orc_direct_dct2_f64", None, [C.c_int, f64p]),
            ("orc_direct_dct3_f64", None, [C.c_int, f64p]),
            ("orc_direct_dst3_f64", None, [C.c_int, f64p]),
            ("orc_set_batch_isa", None, [C.c_int]),
            ("orc_ddct2d_fwd_f32", None, [C.c_int, C.c_int, f32p]),
            ("orc_ddct2d_inv_f32", None, [C.c_int, C.c_int, f32p]),
            ("orc_ddsct2d_inv_f32", None, [C.c_int, C.c_int, f32p]),
            ("orc_ddcst2d_inv_f32", None, [C.c_int, C.c_int, f32p]),
            ("orc_ddct2d_fwd_f64", None, [C.c_int, C.c_int, f64p]),
            ("orc_ddct2d_inv_f64", None, [C.c_int, C.c_int, f64p]),
            ("orc_ddsct2d_inv_f64", None, [C.c_int, C.c_int, f64p]),
            ("orc_ddcst2d_inv_f64", None, [C.c_int, C.c_int, f64p]),
            ("orc_fft_solve_f32", None, [C.c_int32, C.c_int32, f32p, f32p, f32p, f32p, f32p, f32p]),
            ("orc_fft_solve_f64", None, [C.c_int32, C.c_int32, f32p, f32p, f32p, f64p, f64p, f64p]),
            ("orc_create", C.c_void_p, [C.c_void_p]),
            ("orc_destroy", None, [C.c_void_p]),
            ("orc_init", C.c_int32, [C.c_void_p]),
            ("orc_run", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, i32p]),
            ("orc_stats", None, [C.c_void_p, C.c_void_p]),
            ("orc_history", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32]),
            ("orc_get_slot", None, [C.c_void_p, C.c_int32, C.c_int32, f32p]),
            ("orc_get_grad", None, [C.c_void_p, C.c_int32, C.c_int32, f32p]),
            ("orc_get_bins", None, [C.c_void_p, C.c_int32, f32p, f32p, f32p, i64p]),
            ("orc_get_scalars", None, [C.c_void_p, C.c_int32, f32p]),
            ("orc_set_inject", None, [C.c_void_p, C.c_int32]),
            ("orc_np_hpwl", C.c_int64, [C.c_void_p]),
            ("orc_scatter", None, [C.c_void_p, C.c_int32, f32p, i64p, i64p, f32p, i64p, f64p]),
            ("orc_place_pins", None, [C.c_void_p, C.c_int32, f32p]),
            ("orc_wa_forward", None, [C.c_void_p, C.c_float, C.c_float, f32p, i32p, f64p, i64p]),
            ("orc_wa_backward", None, [C.c_void_p, C.c_int32, f32p, f64p, f64p]),
            ("orc_gather", None, [C.c_void_p, C.c_int32, f32p, f32p, f32p, f64p, f64p, f64p]),
            ("orc_get_dbox", None, [C.c_void_p, i32p]),
            ("orc_get_pins", None, [C.c_void_p, i32p]),
            ("orc_wa_net_f64", C.c_double, [C.c_int32, f64p, C.c_double, C.c_int32, f64p]),
            ("orc_objective_value", None, [C.c_void_p, f32p, C.c_float, f64p, f64p]),
            ("orc_objective_fd", None, [C.c_void_p, f32p, C.c_float, C.c_float, C.c_int32, i32p, C.c_double, f64p]),
            ("orc_solve_phi_f64", None, [C.c_int32, C.c_int32, f32p, f32p, f64p, f64p]),
            ("orc_inside", C.c_float, [C.c_float, C.c_int32, C.c_int32, C.c_int32]),
            ("orc_nesterov_points", None, [C.c_int32, f32p, f32p, f32p, u8p, i32p, i32p, C.c_float, C.c_float, f32p, f32p]),
            ("orc_step_length", C.c_int32, [C.c_int32, f32p, f32p, f32p, f32p, f32p, f32p, i32p]),
            ("orc_precondition", None, [C.c_int32, f32p, f32p, i32p, i32p, C.c_float, f32p, f32p, f32p]),
            ("orc_init_penalty", C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
            ("orc_init_step", C.c_float, [C.c_int32, f32p, f32p, f32p, f32p, f32p]),
            ("orc_check_divergence", C.c_int32, [C.c_float, C.c_float, C.c_float, C.c_int64, C.c_float, C.c_int64]),
            ("orc_momentum", C.c_float, [C.c_float, f32p]),
            ("orc_set_inject_control", None, [C.c_void_p, C.c_int32, C.c_int32]),
            ("orc_region_rows", C.c_int32, [C.c_void_p, f32p, C.c_int32]),
            ("orc_set_field_noise", None, [C.c_void_p, C.c_float, C.c_uint64]),
            ("orc_set_wide_sums", None, [C.c_int32]),
            ("orc_set_threads", None, [C.c_int32]),
            ("orc_rudy_wire_width", C.c_int32, [i32p, C.c_int32]),
            ("orc_rudy_rect", None, [i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, i32p]),
            ("orc_rudy", None, [i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, i32p, u8p, f32p, f32p]),
            ("orc_rudy_congestion", C.c_float, [C.c_float, C.c_float]),
            ("orc_top_averages", None, [C.c_int32, f64p, f64p]),
            ("orc_rudy_tiles", None, [C.c_int32, C.c_int32, C.c_int32, f32p, f32p, C.c_int32, i32p, f32p, f32p, u8p, f64p]),
            ("orc_set_size", None, [i32p, C.c_int32, C.c_int32, i32p]),
            ("orc_route_inflate", C.c_int64, [C.c_int32, i32p, i32p, u8p, C.c_int32, C.c_int32, C.c_int32, f32p]),
            ("orc_density_size", None, [C.c_int32, i32p, C.c_double, C.c_double, i32p, f32p]),
            ("orc_route_stats", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
            ("orc_non_place_at", None, [C.c_int32, i64p, i64p, C.c_float, i64p, i64p]),
            ("orc_cut_fillers_plan", C.c_int32, [C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_int64, f32p]),
            ("orc_budget_trim", C.c_int32, [C.c_double, C.c_double, C.c_double, C.c_int32, f32p]),
            ("orc_congested_net_weights", C.c_int32, [C.c_int32, f32p, C.c_float, C.c_float, f32p]),
        ]
        for name, res, args in sig:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def f32(a):
    return ptr(a, f32p)


def f64(a):
    return ptr(a, f64p)


def i64(a):
    return ptr(a, i64p)


def i32(a):
    return ptr(a, i32p)


class Oracle:
    """The faithful CPU placer on a problem view (a gpl_amd Problem ctypes struct)."""

    def __init__(self, problem):
        self.problem = problem
        self._h = lib().orc_create(C.addressof(problem))
        self.regions = [problem.regions[b] for b in range(problem.num_regions)]

    def n(self, b=0):
        R = self.regions[b]
        return R.cell_end - R.cell_begin

    def m(self, b=0):
        R = self.regions[b]
        return R.cnt_x * R.cnt_y

    def init(self):
        return lib().orc_init(self._h)

    def run(self, start_iter=0, iter_budget=-1, time_budget_s=0.0):
        it = C.c_int32(0)
        rc = lib().orc_run(self._h, start_iter, iter_budget, time_budget_s, C.byref(it))
        return rc, it.value

    def stats(self):
        from gpl_amd._ffi import Stats
        s = Stats()
        lib().orc_stats(self._h, C.byref(s))
        return s

    def route_stats(self):
        """The routability record: (totals dict, one dict per routability pass), as
        NesterovGPU.route_stats."""
        from gpl_amd._ffi import RoutePass, RouteStats
        st = RouteStats()
        n = lib().orc_route_stats(self._h, C.byref(st), None, 0)
        rows = (RoutePass * max(n, 1))()
        lib().orc_route_stats(self._h, None, rows, n)
        passes = [{k: getattr(r, k) for k, _ in RoutePass._fields_} for r in rows[:n]]
        return {k: getattr(st, k) for k, _ in RouteStats._fields_}, passes

    def history(self):
        from gpl_amd._ffi import IterRow
        n = lib().orc_history(self._h, None, 0)
        rows = (IterRow * max(n, 1))()
        lib().orc_history(self._h, rows, n)
        return [dict(iter=r.iter, trials=r.trials, overflow=r.overflow, density_penalty=r.density_penalty,
                     step_length=r.step_length, wl_coef=r.wl_coef, hpwl=r.hpwl) for r in rows[:n]]

    def slot(self, b, slot):
        a = np.zeros(2 * self.n(b), np.float32)
        lib().orc_get_slot(self._h, b, slot, f32(a))
        return a.reshape(-1, 2)

    def grad(self, b, which):
        a = np.zeros(2 * self.n(b), np.float32)
        lib().orc_get_grad(self._h, b, which, f32(a))
        return a.reshape(-1, 2)

    def bins(self, b=0):
        M = self.m(b)
        d, fx, fy = (np.zeros(M, np.float32) for _ in range(3))
        ov = np.zeros(2, np.int64)
        lib().orc_get_bins(self._h, b, f32(d), f32(fx), f32(fy), i64(ov))
        return d, fx, fy, ov

    def scalars(self, b=0):
        a = np.zeros(10, np.float32)
        lib().orc_get_scalars(self._h, b, f32(a))
        keys = ["step_length", "density_penalty", "wl_grad_sum", "density_grad_sum", "sum_overflow",
                "sum_overflow_unscaled", "wl_coef", "base_wl_coef", "reference_hpwl", "coordi_distance"]
        return dict(zip(keys, a.tolist()))

    def hpwl(self):
        return lib().orc_np_hpwl(self._h)

    def scatter(self, b, xy):
        M = self.m(b)
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1)
        inst, fill = np.zeros(M, np.int64), np.zeros(M, np.int64)
        dens = np.zeros(M, np.float32)
        ovi, ovf = np.zeros(2, np.int64), np.zeros(2, np.float64)
        lib().orc_scatter(self._h, b, f32(xy), i64(inst), i64(fill), f32(dens), i64(ovi), f64(ovf))
        return inst, fill, dens, ovi, ovf

    def place_pins(self, b, xy):
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1)
        lib().orc_place_pins(self._h, b, f32(xy))

    def wa_forward(self, cx, cy):
        E = self.problem.num_nets
        sums = np.zeros(8 * E, np.float32)
        box = np.zeros(4 * E, np.int32)
        s64 = np.zeros(8 * E, np.float64)
        h = C.c_int64(0)
        lib().orc_wa_forward(self._h, cx, cy, f32(sums), i32(box), f64(s64), C.byref(h))
        return sums.reshape(-1, 8), box.reshape(-1, 4), s64.reshape(-1, 8), h.value

    def wa_backward(self, b, sums_f64):
        n = self.n(b)
        wl = np.zeros(2 * n, np.float32)
        w64 = np.zeros(2 * n, np.float64)
        s = np.ascontiguousarray(sums_f64, np.float64).reshape(-1)
        lib().orc_wa_backward(self._h, b, f32(wl), f64(s), f64(w64))
        return wl.reshape(-1, 2), w64.reshape(-1, 2)

    def gather(self, b, fx, fy, fx64=None, fy64=None):
        n = self.n(b)
        den = np.zeros(2 * n, np.float32)
        d64 = np.zeros(2 * n, np.float64)
        fx = np.ascontiguousarray(fx, np.float32)
        fy = np.ascontiguousarray(fy, np.float32)
        fx64 = None if fx64 is None else np.ascontiguousarray(fx64, np.float64)
        fy64 = None if fy64 is None else np.ascontiguousarray(fy64, np.float64)
        lib().orc_gather(self._h, b, f32(fx), f32(fy), f32(den), f64(fx64), f64(fy64), f64(d64) if fx64 is not None else None)
        return den.reshape(-1, 2), d64.reshape(-1, 2)

    def dbox(self):
        a = np.zeros(4 * self.problem.num_cells, np.int32)
        lib().orc_get_dbox(self._h, i32(a))
        return a.reshape(-1, 4)

    def pins(self):
        a = np.zeros(2 * self.problem.num_pins, np.int32)
        lib().orc_get_pins(self._h, i32(a))
        return a.reshape(-1, 2)

    def set_inject(self, after_iter):
        lib().orc_set_inject(self._h, after_iter)

    def region_rows(self):
        """Every region's state after each iteration: rows (iter, region, overflow, penalty, step,
        converged, WA gradient sum, density gradient sum)."""
        n = lib().orc_region_rows(self._h, None, 0)
        a = np.zeros(8 * max(n, 1), np.float32)
        lib().orc_region_rows(self._h, f32(a), n)
        return a[:8 * n].reshape(-1, 8)

    def set_field_noise(self, eps, seed=1):
        """The sensitivity probe: every field value of every solve scaled by (1 + eps u), u uniform
        in [-1, 1); eps = 0 turns it off."""
        lib().orc_set_field_noise(self._h, eps, seed)

    def set_inject_control(self, kind, at):
        """Kind 5: the WA gradient sum read as zero once at iteration `at`; kind 6: a divergence
        reported at iteration `at` (twice for two iterations); kind 0 clears them."""
        lib().orc_set_inject_control(self._h, kind, at)

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_threads(n):
    """Host threads for the loop's independent per-cell, per-net, per-row and per-bin work
    (process-wide; 1, the default, is the reference's single thread).  Any count gives the same
    bits: the sums whose order the reference fixes stay sequential."""
    lib().orc_set_threads(int(n))


def set_wide_sums(on):
    """The oracle's global float sums (step-length distances, |wl| and |den|, overflow) in double
    (on) or as the reference's float sums (off, the default); process-wide."""
    lib().orc_set_wide_sums(1 if on else 0)


def solve_f64(nx, ny, wx, wy, density_ym):
    M = nx * ny
    phi, fx, fy = (np.zeros(M, np.float64) for _ in range(3))
    d = np.ascontiguousarray(density_ym, np.float32)
    lib().orc_fft_solve_f64(nx, ny, f32(np.ascontiguousarray(wx, np.float32)), f32(np.ascontiguousarray(wy, np.float32)),
                            f32(d), f64(phi), f64(fx), f64(fy))
    return phi, fx, fy


def solve_f32(nx, ny, wx, wy, density_ym):
    """Fft::solve in float on y-major bins: returns phi, field_x, field_y (y-major)."""
    d = np.ascontiguousarray(np.asarray(density_ym, np.float32).reshape(ny, nx).T.reshape(-1))  # x-major
    phi, fx, fy = (np.zeros(nx * ny, np.float32) for _ in range(3))
    lib().orc_fft_solve_f32(nx, ny, f32(np.ascontiguousarray(wx, np.float32)), f32(np.ascontiguousarray(wy, np.float32)),
                            f32(d), f32(phi), f32(fx), f32(fy))
    back = lambda a: a.reshape(nx, ny).T.reshape(-1).copy()
    return back(phi), back(fx), back(fy)


def objective_value(problem, xy, wl_coef):
    """FP64 value of the objective's two parts (wirelength, 1/2 sum rho phi) at region 0's cell
    centres xy (problem order)."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1)
    w, d = C.c_double(0.0), C.c_double(0.0)
    lib().orc_objective_value(C.addressof(problem), f32(xy), wl_coef, C.byref(w), C.byref(d))
    return w.value, d.value


def objective_fd(problem, xy, wl_coef, lam, sample, h):
    """FP64 central differences of the objective for the sampled cells: rows (d/dx, d/dy,
    kink_x, kink_y)."""
    xy = np.ascontiguousarray(xy, np.float32).reshape(-1)
    sample = np.ascontiguousarray(sample, np.int32)
    out = np.zeros(4 * len(sample), np.float64)
    lib().orc_objective_fd(C.addressof(problem), f32(xy), wl_coef, lam, len(sample), i32(sample), h, f64(out))
    return out.reshape(-1, 4)


def solve_phi_f64(nx, ny, wx, wy, density):
    d = np.ascontiguousarray(density, np.float64)
    phi = np.zeros(nx * ny, np.float64)
    lib().orc_solve_phi_f64(nx, ny, f32(np.ascontiguousarray(wx, np.float32)), f32(np.ascontiguousarray(wy, np.float32)), f64(d), f64(phi))
    return phi


# ---- the loop control's arithmetic on plain arrays (pinned by tests/test_oracle_control.py) ----

def _f2(a):
    return np.ascontiguousarray(a, np.float32).reshape(-1)


def inside(c, dd, lo, hi):
    return lib().orc_inside(np.float32(c), int(dd), int(lo), int(hi))


def nesterov_points(cur_slp, sum_, cur, locked, dd, box, step, coeff):
    """next, next_slp of nesterovUpdateCoordinates (n x 2 float32 each)."""
    cs, sm, cu = _f2(cur_slp), _f2(sum_), _f2(cur)
    n = len(cs) // 2
    lk = np.ascontiguousarray(locked, np.uint8)
    d = np.ascontiguousarray(dd, np.int32).reshape(-1)
    b = np.ascontiguousarray(box, np.int32)
    nx, ns = np.zeros(2 * n, np.float32), np.zeros(2 * n, np.float32)
    lib().orc_nesterov_points(n, f32(cs), f32(sm), f32(cu), lk.ctypes.data_as(u8p), i32(d), i32(b), np.float32(step),
                              np.float32(coeff), f32(nx), f32(ns))
    return nx.reshape(-1, 2), ns.reshape(-1, 2)


def step_length(cur_slp, next_slp, cur_sum, next_sum, old_step):
    """(accepted, new step, coordinate distance, diverged) of nesterovUpdateStepLength."""
    a, b, c, d = _f2(cur_slp), _f2(next_slp), _f2(cur_sum), _f2(next_sum)
    st, cd, dv = C.c_float(old_step), C.c_float(0.0), C.c_int32(0)
    ok = lib().orc_step_length(len(a) // 2, f32(a), f32(b), f32(c), f32(d), C.byref(st), C.byref(cd), C.byref(dv))
    return bool(ok), st.value, cd.value, bool(dv.value)


def precondition(wl, den, npins, dxy, penalty):
    """(sum n x 2, |wl| sum, |den| sum) of updateGradients' arithmetic."""
    w, d = _f2(wl), _f2(den)
    n = len(w) // 2
    s = np.zeros(2 * n, np.float32)
    wa, da = C.c_float(0.0), C.c_float(0.0)
    lib().orc_precondition(n, f32(w), f32(d), i32(np.ascontiguousarray(npins, np.int32)),
                           i32(np.ascontiguousarray(dxy, np.int32).reshape(-1)), np.float32(penalty), f32(s), C.byref(wa), C.byref(da))
    return s.reshape(-1, 2), wa.value, da.value


def init_penalty(wl_sum, den_sum, cur, init):
    return lib().orc_init_penalty(np.float32(wl_sum), np.float32(den_sum), np.float32(cur), np.float32(init))


def init_step(prev_slp, cur_slp, prev_sum, cur_sum):
    a, b, c, d = _f2(prev_slp), _f2(cur_slp), _f2(prev_sum), _f2(cur_sum)
    cd = C.c_float(0.0)
    s = lib().orc_init_step(len(a) // 2, f32(a), f32(b), f32(c), f32(d), C.byref(cd))
    return s, cd.value


def check_divergence(ovf, min_ovf, hpwl_at_min, prev_hpwl, prev_reported_ovf, prev_reported_hpwl):
    return lib().orc_check_divergence(np.float32(ovf), np.float32(min_ovf), np.float32(hpwl_at_min), int(prev_hpwl),
                                      np.float32(prev_reported_ovf), int(prev_reported_hpwl))


def momentum(prev_a):
    """(a', coeff) of the loop's momentum step."""
    c = C.c_float(0.0)
    a = lib().orc_momentum(np.float32(prev_a), C.byref(c))
    return a, c.value


# ---- routability's congestion measurement (route.c) --------------------------------------------
def rudy_wire_width(pitches):
    p = np.ascontiguousarray(pitches, np.int32)
    return int(lib().orc_rudy_wire_width(i32(p), len(p)))


def rudy(die, cnt_x, cnt_y, tile, wire_width, boxes, cap_red=None):
    """Rudy::calculate: (demand, blockage), x-major (index x * cnt_y + y), from net boxes (n x 4,
    xl yl xh yh, in the nets' order) and optional (capacity, reduction) pairs per tile."""
    n = cnt_x * cnt_y
    demand, blockage = np.zeros(n, np.float32), np.zeros(n, np.float32)
    b = np.ascontiguousarray(boxes, np.int32).reshape(-1)
    cr = None if cap_red is None else np.ascontiguousarray(cap_red, np.uint8).reshape(-1)
    lib().orc_rudy(i32(np.ascontiguousarray(die, np.int32)), cnt_x, cnt_y, tile, wire_width, len(b) // 4, i32(b),
                   None if cr is None else cr.ctypes.data_as(C.POINTER(C.c_uint8)), f32(demand), f32(blockage))
    return demand, blockage


def rudy_rect(die, cnt_x, cnt_y, tile, x, y):
    r = np.zeros(4, np.int32)
    lib().orc_rudy_rect(i32(np.ascontiguousarray(die, np.int32)), cnt_x, cnt_y, tile, x, y, i32(r))
    return tuple(int(v) for v in r)


def rudy_congestion(demand, blockage):
    return float(lib().orc_rudy_congestion(C.c_float(demand), C.c_float(blockage)))


def top_averages(desc_list):
    a = np.ascontiguousarray(desc_list, np.float64)
    out = np.zeros(4, np.float64)
    lib().orc_top_averages(len(a), f64(a), f64(out))
    return out


def rudy_tiles(cnt_x, cnt_y, tile, demand, blockage, centres, min_cong=0.95, coef=2.0, max_ratio=3.0, ignore_edge=0.8,
               rc_k=(1.0, 1.0, 0.0, 0.0)):
    """calculateRudyTiles' tiles and updateRudyAverage: (inflation ratio, actionable) per tile
    (y-major), and dict(total_overflow, overflowed, actionable, top (4), rc)."""
    n = cnt_x * cnt_y
    ratio, act = np.zeros(n, np.float32), np.zeros(n, np.uint8)
    c = np.ascontiguousarray(centres, np.int32).reshape(-1)
    v = np.array([min_cong, coef, max_ratio, ignore_edge, *rc_k], np.float32)
    st = np.zeros(8, np.float64)
    lib().orc_rudy_tiles(cnt_x, cnt_y, tile, f32(np.ascontiguousarray(demand, np.float32)), f32(np.ascontiguousarray(blockage, np.float32)),
                         len(c) // 2, i32(c), f32(v), f32(ratio), act.ctypes.data_as(C.POINTER(C.c_uint8)), f64(st))
    return ratio, act, dict(total_overflow=st[0], overflowed=int(st[1]), actionable=int(st[2]), top=st[3:7].copy(), rc=float(st[7]))


def set_size(box, dx, dy):
    out = np.zeros(4, np.int32)
    lib().orc_set_size(i32(np.ascontiguousarray(box, np.int32)), dx, dy, i32(out))
    return tuple(int(v) for v in out)


def route_inflate(boxes, centres, std_cell, cnt_x, cnt_y, tile, ratio):
    """The per-cell inflation: (new boxes (n x 4), area added)."""
    b = np.ascontiguousarray(boxes, np.int32).reshape(-1).copy()
    c = np.ascontiguousarray(centres, np.int32).reshape(-1)
    s = np.ascontiguousarray(std_cell, np.uint8)
    d = lib().orc_route_inflate(len(s), i32(b), i32(c), s.ctypes.data_as(C.POINTER(C.c_uint8)), cnt_x, cnt_y, tile,
                                f32(np.ascontiguousarray(ratio, np.float32)))
    return b.reshape(-1, 4), int(d)


def density_size(sizes, bin_x, bin_y):
    """updateDensitySize: (density sizes (n x 2), scales)."""
    s = np.ascontiguousarray(sizes, np.int32).reshape(-1)
    n = len(s) // 2
    ds, sc = np.zeros(2 * n, np.int32), np.zeros(n, np.float32)
    lib().orc_density_size(n, i32(s), bin_x, bin_y, i32(ds), f32(sc))
    return ds.reshape(-1, 2), sc


# ---- routability passes' host arithmetic (oracle.c) ----------------------------------------------
def non_place_at(raw, uni, td):
    """updateBinsNonPlaceArea at target density td from the raw overlap sums and union areas:
    (non_place, non_place_unscaled)."""
    r = np.ascontiguousarray(raw, np.int64)
    u = np.ascontiguousarray(uni, np.int64)
    a, b = np.zeros(len(r), np.int64), np.zeros(len(r), np.int64)
    lib().orc_non_place_at(len(r), i64(r), i64(u), C.c_float(td), i64(a), i64(b))
    return a, b


def cut_fillers_plan(area, single, fillers, insts_area, white_space, td):
    """cutFillerCells' arithmetic: (fillers taken out, target density afterwards)."""
    t = np.array([td], np.float32)
    k = lib().orc_cut_fillers_plan(area, single, fillers, insts_area, white_space, f32(t))
    return int(k), float(t[0])


def budget_trim(budget, spent, wanted, ratios):
    """scaleInflationToBudget's trim: (budget spent, the trimmed ratios)."""
    r = np.ascontiguousarray(ratios, np.float32).copy()
    spent_flag = lib().orc_budget_trim(budget, spent, wanted, len(r), f32(r))
    return bool(spent_flag), r


def congested_net_weights(demand, pct, wmax):
    """weightCongestedNets' weights: (count of nets with demand or -1, weight per net, 0 = none)."""
    d = np.ascontiguousarray(demand, np.float32)
    w = np.zeros(len(d), np.float32)
    n = lib().orc_congested_net_weights(len(d), f32(d), C.c_float(pct), C.c_float(wmax), f32(w))
    return int(n), w

