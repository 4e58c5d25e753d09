"""gpl_amd: the ePlace Nesterov global-placement loop of vyges-gpl on MI355X (gfx950).

Host API over libgpl_hip.so (C ABI in include/gpl_hip.h):

* :class:`Design` -- a seeded synthetic design (the ODB reader's stand-in), the placer built from
  it, its initial placement and the Nesterov placer's construction;
* :class:`NesterovGPU` -- the placer on the GPU: ``init_nesterov_loop`` / ``run_nesterov_loop`` /
  ``do_nesterov_place`` / ``continue_nesterov_place`` / ``placed_locations`` as the reference names
  them (nesterov.rs:4099-4247), plus the per-kernel entries the parity tests drive;
* :func:`global_placement` -- the command (run.rs:115): option parsing with the reference's
  prefix rules, initial placement, Nesterov placement, the outcome.

Everything computes on the GPU through the HIP library; when it is not built the first call
raises -- there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import (GPL_CELL_INSTANCE, GPL_CELL_LOCKED, GPL_CELL_MACRO, GPL_ERR_ARG, GPL_ERR_HIP, GPL_ERR_REFUSED,
                   GPL_ERR_STATE, GPL_OK, GPL_SLOT_CUR, GPL_SLOT_CUR_SLP, GPL_SLOT_NEXT, GPL_SLOT_NEXT_SLP,
                   GPL_SLOT_PREV_SLP, DesignInfo, GenConfig, IterRow, KernelTime, Opts, Problem, Stats)

__all__ = [
    "Design", "NesterovGPU", "GplError", "global_placement", "default_opts", "gen_config", "CONFIGS",
    "GPL_SLOT_CUR_SLP", "GPL_SLOT_NEXT_SLP", "GPL_SLOT_PREV_SLP", "GPL_SLOT_CUR", "GPL_SLOT_NEXT",
]


class GplError(RuntimeError):
    """A non-zero status: positive values are the reference's GPL message ids (304..307)."""

    def __init__(self, code, msg=""):
        super().__init__(f"status {code}: {msg}" if msg else f"status {code}")
        self.code = code


def _ok(rc, what, ctx=None):
    if rc != GPL_OK:
        msg = what
        if ctx is not None:
            err = _ffi.lib().gpl_last_error(ctx)
            if err:
                msg += ": " + err.decode(errors="replace")
        raise GplError(rc, msg)
    return rc


def default_opts(**kw):
    """`Opts` with the reference's defaults (nesterov.rs:59-82, initial_place.rs:27-31), then kw."""
    o = Opts()
    _ffi.lib().gpl_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


# The workloads: seeded synthetic designs of the sizes the headline configurations name.
#   c1: 2,000 one-row cells (1-8 sites), 2,200 nets of degree 2+Geom(0.55) capped at 32, 96 IO
#       ports; core near-square for 70 % utilisation (36 rows x 357 sites); 64 x 64 bins;
#       5 initial-placement rounds.
#   c2: 200k cells, 210k nets 2+Geom(0.5) capped at 64, 20 nets of 500-2,000 pins; square core at
#       65 % utilisation, target density 0.70; 512 x 512 bins.
#   c3: 1M cells, 1.05M nets capped at 128, 50 nets of 1k-5k pins, 5 % of the core in fixed
#       macros of 50-400 rows; 1024 x 1024 bins.
#   c4r: 8 region groups of 5k-150k cells (485k) plus 20k top-level cells (a region needs
#       placeable cells), each region with the bin grid its size derives (32^2 .. 512^2).
#   c4s: the 50k-cell, 52k-net design placed for many seeds (256 x 256 bins).
#   c4r_small: c4r at 1/30 scale for the oracle comparison.
CONFIGS = {
    "c1": dict(gen=dict(num_cells=2000, num_nets=2200, degree_p=0.55, degree_cap=32, num_io=96, utilization=0.70,
                        core_rows=36, core_sites=357),
               opts=dict(bin_cnt=64, density=0.70, initial_place_max_iter=5)),
    "c2": dict(gen=dict(num_cells=200_000, num_nets=210_000, degree_p=0.5, degree_cap=64, num_hf_nets=20,
                        hf_min_pins=500, hf_max_pins=2000, num_io=512, utilization=0.65),
               opts=dict(bin_cnt=512, density=0.70, initial_place_max_iter=5)),
    "c3": dict(gen=dict(num_cells=1_000_000, num_nets=1_050_000, degree_p=0.5, degree_cap=128, num_hf_nets=50,
                        hf_min_pins=1000, hf_max_pins=5000, num_io=1024, utilization=0.65, macro_area_frac=0.05,
                        macro_rows_min=50, macro_rows_max=400),
               opts=dict(bin_cnt=1024, density=0.70, initial_place_max_iter=5)),
    "c4r": dict(gen=dict(num_cells=20_000, num_nets=530_000, degree_p=0.5, degree_cap=64, num_io=256, utilization=0.65,
                         group_cells=[5_000, 10_000, 20_000, 40_000, 60_000, 80_000, 120_000, 150_000]),
                opts=dict(bin_cnt=0, density=0.70, initial_place_max_iter=5)),
    "c4s": dict(gen=dict(num_cells=50_000, num_nets=52_000, degree_p=0.5, degree_cap=64, num_io=128, utilization=0.65),
                opts=dict(bin_cnt=256, density=0.70, initial_place_max_iter=5)),
    "c5": dict(gen=dict(num_cells=500_000, num_nets=520_000, degree_p=0.5, degree_cap=64, num_io=512, utilization=0.65),
               opts=dict(bin_cnt=768, density=0.70, initial_place_max_iter=5, routability=1, route_gcells=1024, route_tracks=8,
                         routability_target_rc=1.0, routability_max_inflation_ratio=2.5, routability_max_passes=5,
                         routability_net_weight_max=2.0)),
    "c4r_small": dict(gen=dict(num_cells=1_000, num_nets=18_000, degree_p=0.5, degree_cap=64, num_io=96, utilization=0.65,
                               group_cells=[250, 500, 1_000, 1_500, 2_000, 3_000, 4_000, 5_000]),
                      opts=dict(bin_cnt=0, density=0.70, initial_place_max_iter=5)),
}


def gen_config(name_or_dict, seed=1, **over):
    """A GenConfig from a CONFIGS entry (or a dict of GenConfig fields), with overrides."""
    d = dict(CONFIGS[name_or_dict]["gen"]) if isinstance(name_or_dict, str) else dict(name_or_dict)
    d.update(over)
    groups = d.pop("group_cells", None)
    g = GenConfig()
    for k, v in d.items():
        setattr(g, k, v)
    g.seed = seed
    if groups:
        arr = (C.c_int32 * len(groups))(*groups)
        g.num_groups = len(groups)
        g.group_cells = C.cast(arr, C.POINTER(C.c_int32))
        g._keep = arr
    return g


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Design:
    """A synthetic design and its placer (gpl_design_*)."""

    def __init__(self, cfg, opts=None):
        L = _ffi.lib()
        self.opts = opts if opts is not None else default_opts()
        self._cfg = cfg
        h = C.c_void_p()
        _ok(L.gpl_design_generate(C.byref(cfg), C.byref(self.opts), C.byref(h)), "gpl_design_generate")
        self._h = h
        self._built = False

    @classmethod
    def from_config(cls, name, seed=1, gen_over=None, **opt_over):
        o = dict(CONFIGS[name]["opts"])
        o.update(opt_over)
        return cls(gen_config(name, seed=seed, **(gen_over or {})), default_opts(**o))

    def initial_place(self, device=False):
        """B2B + BiCGSTAB initial placement; device=True runs its rounds on the GPU (the same
        placement bit for bit)."""
        if device:
            _ok(_ffi.lib().gpl_design_initial_place_gpu(self._h), "gpl_design_initial_place_gpu")
        else:
            _ok(_ffi.lib().gpl_design_initial_place(self._h), "gpl_design_initial_place")

    def init_nesterov_place(self, opts=None):
        if opts is not None:
            self.opts = opts
        _ok(_ffi.lib().gpl_init_nesterov_place(self._h, C.byref(self.opts)), "gpl_init_nesterov_place")
        self._built = True

    def info(self):
        i = DesignInfo()
        _ok(_ffi.lib().gpl_design_get_info(self._h, C.byref(i)), "gpl_design_get_info")
        return i

    def problem(self):
        """The ctypes view of the Nesterov problem (valid while this design lives)."""
        p = Problem()
        _ok(_ffi.lib().gpl_design_problem(self._h, C.byref(p)), "gpl_design_problem")
        return p

    def arrays(self):
        """Copies of the problem's arrays as numpy arrays, regions as dicts."""
        p = self.problem()
        N, Pn, E = p.num_cells, p.num_pins, p.num_nets
        out = dict(
            num_cells=N, num_pins=Pn, num_nets=E,
            cell_box=_arr(p.cell_box, 4 * N, np.int32).reshape(-1, 4),
            cell_dbox=_arr(p.cell_dbox, 4 * N, np.int32).reshape(-1, 4),
            density_scale=_arr(p.density_scale, N, np.float32),
            cell_flags=_arr(p.cell_flags, N, np.uint8),
            cell_inst=_arr(p.cell_inst, N, np.int32),
            cell_pin_start=_arr(p.cell_pin_start, N + 1, np.int32),
            pin_cell=_arr(p.pin_cell, Pn, np.int32),
            pin_offset=_arr(p.pin_offset, 2 * Pn, np.int32).reshape(-1, 2),
            pin_xy=_arr(p.pin_xy, 2 * Pn, np.int32).reshape(-1, 2),
            pin_net=_arr(p.pin_net, Pn, np.int32),
            net_pin_start=_arr(p.net_pin_start, E + 1, np.int32),
            net_weight=_arr(p.net_weight, E, np.float32),
        )
        out["cell_pins"] = _arr(p.cell_pins, int(out["cell_pin_start"][-1]), np.int32)
        out["net_pins"] = _arr(p.net_pins, int(out["net_pin_start"][-1]), np.int32)
        regs = []
        for b in range(p.num_regions):
            R = p.regions[b]
            M = R.cnt_x * R.cnt_y
            regs.append(dict(
                lx=R.lx, ly=R.ly, ux=R.ux, uy=R.uy, cnt_x=R.cnt_x, cnt_y=R.cnt_y, size_x=R.size_x, size_y=R.size_y,
                target_density=R.target_density, cell_begin=R.cell_begin, cell_end=R.cell_end, num_fillers=R.num_fillers,
                filler_dx=R.filler_dx, filler_dy=R.filler_dy, std_insts_area=R.std_insts_area,
                macro_insts_area=R.macro_insts_area, non_place_insts_area=R.non_place_insts_area,
                white_space_area=R.white_space_area, non_place=_arr(R.non_place, M, np.int64),
                non_place_unscaled=_arr(R.non_place_unscaled, M, np.int64), bin_x=_arr(R.bin_x, R.cnt_x + 1, np.int32),
                bin_y=_arr(R.bin_y, R.cnt_y + 1, np.int32), wx=_arr(R.wx, R.cnt_x, np.float32),
                wy=_arr(R.wy, R.cnt_y, np.float32), group=R.group))
        out["regions"] = regs
        return out

    def problem_sha256(self):
        """A fingerprint of the Nesterov problem: sha256 over every array of arrays() (cells, pins,
        nets, the regions' bins), so that two builds of a design can be told apart."""
        import hashlib
        a = self.arrays()
        h = hashlib.sha256()

        def feed(k, v):
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes() if isinstance(v, np.ndarray) else repr(v).encode())

        for k in sorted(k for k in a if k != "regions"):
            v = a[k]
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes() if isinstance(v, np.ndarray) else str(v).encode())
        for r in a["regions"]:
            for k in sorted(r):
                feed(k, r[k])
        return h.hexdigest()

    def close(self):
        if getattr(self, "_h", None):
            _ffi.lib().gpl_design_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def place_concurrently(contexts, max_workers=8, launch_mode=2):
    """do_nesterov_place on each of several independent contexts from a pool of host threads (the
    library releases the interpreter lock in every call), so that the GPU interleaves the
    placements' kernels; returns the iteration counts in order.  The results are each context's
    own (the same as running them one after another).  Each context launches on one stream
    (launch_mode 2): many placements' graphs and side streams would contend for the device's few
    hardware queues, one stream each interleaves best (16 seeds of the 50k-cell design: 3.3x the
    sequential graph form with 8 threads, 1.8x with graphs and 16)."""
    from concurrent.futures import ThreadPoolExecutor
    if not contexts:
        return []
    if launch_mode is not None:
        for g in contexts:
            g.set_launch_mode(launch_mode)
    with ThreadPoolExecutor(max_workers=min(max_workers or len(contexts), len(contexts))) as ex:
        return list(ex.map(lambda g: g.do_nesterov_place(), contexts))


def place_batched(contexts):
    """do_nesterov_place on each of several independent contexts in lockstep (gpl_place_batch):
    each iteration's first trial of all of them as one launch per kernel; returns the iteration
    counts in order.  Each context's result is its own do_nesterov_place's."""
    if not contexts:
        return []
    n = len(contexts)
    hs = (C.c_void_p * n)(*[g._h for g in contexts])
    its = (C.c_int32 * n)()
    rc = _ffi.lib().gpl_place_batch(hs, n, its)
    if rc != GPL_OK:
        for g in contexts:
            msg = _ffi.lib().gpl_last_error(g._h)
            if msg:
                raise GplError(rc, f"gpl_place_batch: {msg.decode(errors='replace')}")
        raise GplError(rc, "gpl_place_batch")
    return list(its)


def rudy_wire_width(pitches):
    """Rudy::new's wire width (route_base.rs:143-152) from the routing layers' pitches."""
    p = np.ascontiguousarray(pitches, np.int32)
    return int(_ffi.lib().gpl_rudy_wire_width(p.ctypes.data_as(C.POINTER(C.c_int32)), len(p)))


class NesterovGPU:
    """The Nesterov placer of one design on the GPU (gpl_ctx_*)."""

    def __init__(self, design):
        if not design._built:
            raise GplError(GPL_ERR_STATE, "init_nesterov_place first")
        self.design = design
        self.problem = design.problem()
        h = C.c_void_p()
        _ok(_ffi.lib().gpl_ctx_create(C.byref(self.problem), C.byref(h)), "gpl_ctx_create")
        self._h = h
        self.regions = [self.problem.regions[b] for b in range(self.problem.num_regions)]

    # ---- the reference's entry points ----
    def init_nesterov_loop(self):
        return _ok(_ffi.lib().gpl_init_nesterov_loop(self._h), "init_nesterov_loop", self._h)

    def run_nesterov_loop(self, start_iter=0, check=True):
        it = C.c_int32(0)
        rc = _ffi.lib().gpl_run_nesterov_loop(self._h, start_iter, C.byref(it))
        if check:
            _ok(rc, "run_nesterov_loop", self._h)
        return it.value if check else (rc, it.value)

    def do_nesterov_place(self):
        it = C.c_int32(0)
        _ok(_ffi.lib().gpl_do_nesterov_place(self._h, C.byref(it)), "do_nesterov_place", self._h)
        return it.value

    def continue_nesterov_place(self, overflow, max_iter, start_iter):
        it = C.c_int32(0)
        _ok(_ffi.lib().gpl_continue_nesterov_place(self._h, overflow, max_iter, start_iter, C.byref(it)),
            "continue_nesterov_place", self._h)
        return it.value

    def loop_begin(self, start_iter=0):
        return _ok(_ffi.lib().gpl_loop_begin(self._h, start_iter), "gpl_loop_begin", self._h)

    def loop_step(self, n):
        """Runs at most n iterations; returns (iteration, done)."""
        it, done = C.c_int32(0), C.c_int32(0)
        _ok(_ffi.lib().gpl_loop_step(self._h, n, C.byref(it), C.byref(done)), "gpl_loop_step", self._h)
        return it.value, bool(done.value)

    def loop_end(self, check=True):
        it = C.c_int32(0)
        rc = _ffi.lib().gpl_loop_end(self._h, C.byref(it))
        if check:
            _ok(rc, "gpl_loop_end", self._h)
        return it.value if check else (rc, it.value)

    def stats(self):
        s = Stats()
        _ok(_ffi.lib().gpl_get_stats(self._h, C.byref(s)), "gpl_get_stats")
        return s

    def history(self):
        n = _ffi.lib().gpl_get_history(self._h, None, 0)
        rows = (IterRow * max(n, 1))()
        _ffi.lib().gpl_get_history(self._h, rows, n)
        return [dict(iter=r.iter, trials=r.trials, overflow=r.overflow, density_penalty=r.density_penalty,
                     step_length=r.step_length, wl_coef=r.wl_coef, hpwl=r.hpwl) for r in rows[:n]]

    def log(self):
        n = _ffi.lib().gpl_get_log(self._h, None, 0)
        buf = C.create_string_buffer(n + 1)
        _ffi.lib().gpl_get_log(self._h, buf, n + 1)
        return buf.value.decode()

    def placed_locations(self):
        n = _ffi.lib().gpl_placed_locations(self._h, None, 0)
        out = np.zeros(3 * max(n, 1), dtype=np.int32)
        _ffi.lib().gpl_placed_locations(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return out[: 3 * n].reshape(-1, 3)

    def debug_inject(self, kind, after_iter):
        _ok(_ffi.lib().gpl_debug_inject(self._h, kind, after_iter), "gpl_debug_inject")

    def set_profiling(self, on):
        _ok(_ffi.lib().gpl_set_profiling(self._h, 1 if on else 0), "gpl_set_profiling")

    def set_resort_policy(self, min_gap, growth):
        _ok(_ffi.lib().gpl_set_resort_policy(self._h, min_gap, growth), "gpl_set_resort_policy", self._h)

    def set_launch_mode(self, mode):
        """0: one HIP graph per back-tracking trial (default); 1: kernels one by one on the side
        streams; 2: one by one on one stream; 3: one graph captured on one stream; 4-6:
        device-driven trials (the loop's decisions taken by the trial's last kernel) as a graph,
        one by one on the side streams, on one stream; 7: grouped trials (one launch per kernel
        variant and step over the regions).  The results do not depend on it."""
        _ok(_ffi.lib().gpl_set_launch_mode(self._h, mode), "gpl_set_launch_mode", self._h)

    def set_tuning(self, region, chunk, table_bits, stage):
        """Forces region's kernel variants (scatter chunk 256/512, LDS sum width 32/64, gradient
        staging 0/1) and turns the autotuner off; region -1: the tuner on (chunk 1) or off (0)."""
        _ok(_ffi.lib().gpl_set_tuning(self._h, region, chunk, table_bits, stage), "gpl_set_tuning", self._h)

    def tuning(self):
        """Per region: dict(chunk, table_bits, stage, source, scatter_us, gradient_us)."""
        n = _ffi.lib().gpl_get_tuning(self._h, None, 0)
        out = np.zeros(6 * max(n, 1), np.int32)
        _ffi.lib().gpl_get_tuning(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
        src = {0: "default", 1: "measured", 2: "cached"}
        return [dict(chunk=int(a), table_bits=int(b), stage=int(c), source=src.get(int(d), str(d)), scatter_us=e / 100.0,
                     gradient_us=f / 100.0) for a, b, c, d, e, f in out[:6 * n].reshape(-1, 6)]

    def kernel_times(self):
        arr = (KernelTime * 16)()
        n = _ffi.lib().gpl_kernel_times(self._h, arr, 16)
        return {a.name.decode(): (a.launches, a.ms) for a in arr[:n]}

    # ---- routability's congestion measurement ----
    def rudy(self, grid, cap_red=None):
        """gpl_rudy: (demand, blockage) per tile, x-major (x * cnt_y + y); grid a dict of the
        gpl_rudy_grid fields; cap_red optional (capacity, reduction) uint8 pairs per tile."""
        g = _ffi.RudyGrid(**grid)
        n = g.cnt_x * g.cnt_y
        demand, blockage = np.zeros(n, np.float32), np.zeros(n, np.float32)
        cr = None if cap_red is None else np.ascontiguousarray(cap_red, np.uint8).reshape(-1)
        _ok(_ffi.lib().gpl_rudy(self._h, C.byref(g), None if cr is None else cr.ctypes.data_as(C.POINTER(C.c_uint8)),
                                demand.ctypes.data_as(C.POINTER(C.c_float)), blockage.ctypes.data_as(C.POINTER(C.c_float))),
            "gpl_rudy", self._h)
        return demand, blockage

    def rudy_tiles(self, grid, min_cong=0.95, coef=2.0, max_ratio=3.0, ignore_edge=0.8, rc_k=(1.0, 1.0, 0.0, 0.0)):
        """gpl_rudy_tiles on the last rudy(): (inflation ratio, actionable) per tile (y-major) and
        dict(total_overflow, overflowed, actionable, top (4), rc)."""
        g = _ffi.RudyGrid(**grid)
        v = _ffi.RouteVars(min_cong, coef, max_ratio, ignore_edge, (C.c_float * 4)(*rc_k))
        n = g.cnt_x * g.cnt_y
        ratio, act = np.zeros(n, np.float32), np.zeros(n, np.uint8)
        st = _ffi.RudyStats()
        _ok(_ffi.lib().gpl_rudy_tiles(self._h, C.byref(g), C.byref(v), ratio.ctypes.data_as(C.POINTER(C.c_float)),
                                      act.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(st)), "gpl_rudy_tiles", self._h)
        return ratio, act, dict(total_overflow=st.total_overflow, overflowed=st.overflowed_tiles, actionable=st.actionable_tiles,
                                top=np.array(st.top[:], np.float64), rc=float(st.rc))

    def route_inflate(self, region, grid, inflated_ratio, sizes):
        """gpl_route_inflate: (new sizes (n x 2), area added)."""
        g = _ffi.RudyGrid(**grid)
        r = np.ascontiguousarray(inflated_ratio, np.float32)
        s = np.ascontiguousarray(sizes, np.int32).copy()
        d = C.c_int64(0)
        _ok(_ffi.lib().gpl_route_inflate(self._h, region, C.byref(g), r.ctypes.data_as(C.POINTER(C.c_float)),
                                         s.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(d)), "gpl_route_inflate", self._h)
        return s.reshape(-1, 2), d.value

    def update_density_size(self, region, sizes):
        s = np.ascontiguousarray(sizes, np.int32)
        _ok(_ffi.lib().gpl_update_density_size(self._h, region, s.ctypes.data_as(C.POINTER(C.c_int32))), "gpl_update_density_size",
            self._h)

    def density_boxes(self, region):
        """(density boxes (n x 4), density scales) of the region, the problem's order."""
        n = self._n(region)
        db, sc = np.zeros(4 * max(n, 1), np.int32), np.zeros(max(n, 1), np.float32)
        _ok(_ffi.lib().gpl_get_density_boxes(self._h, region, db.ctypes.data_as(C.POINTER(C.c_int32)),
                                             sc.ctypes.data_as(C.POINTER(C.c_float))), "gpl_get_density_boxes", self._h)
        return db[:4 * n].reshape(-1, 4), sc[:n]

    def route_stats(self):
        """gpl_get_route_stats: (totals dict, one dict per routability pass)."""
        st = _ffi.RouteStats()
        n = _ffi.lib().gpl_get_route_stats(self._h, C.byref(st), None, 0)
        rows = (_ffi.RoutePass * max(n, 1))()
        _ffi.lib().gpl_get_route_stats(self._h, None, rows, n)
        passes = [{k: getattr(r, k) for k, _ in _ffi.RoutePass._fields_} for r in rows[:n]]
        return {k: getattr(st, k) for k, _ in _ffi.RouteStats._fields_}, passes

    # ---- kernel-level entries ----
    def _n(self, region):
        R = self.regions[region]
        return R.cell_end - R.cell_begin

    def _m(self, region):
        R = self.regions[region]
        return R.cnt_x * R.cnt_y

    def set_coordinates(self, region, slot, xy):
        a = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        assert a.size == 2 * self._n(region)
        _ok(_ffi.lib().gpl_set_coordinates(self._h, region, slot, _fp(a)), "gpl_set_coordinates", self._h)

    def get_coordinates(self, region, slot):
        a = np.zeros(2 * self._n(region), dtype=np.float32)
        _ok(_ffi.lib().gpl_get_coordinates(self._h, region, slot, _fp(a)), "gpl_get_coordinates", self._h)
        return a.reshape(-1, 2)

    def density_scatter(self, region, slot):
        M = self._m(region)
        inst = np.zeros(M, np.int64)
        fill = np.zeros(M, np.int64)
        dens = np.zeros(M, np.float32)
        ovf = np.zeros(2, np.float64)
        _ok(_ffi.lib().gpl_density_scatter(self._h, region, slot, inst.ctypes.data_as(C.POINTER(C.c_int64)),
                                           fill.ctypes.data_as(C.POINTER(C.c_int64)), _fp(dens),
                                           ovf.ctypes.data_as(C.POINTER(C.c_double))), "gpl_density_scatter", self._h)
        return inst, fill, dens, ovf

    def set_density(self, region, density):
        a = np.ascontiguousarray(density, dtype=np.float32).reshape(-1)
        assert a.size == self._m(region)
        _ok(_ffi.lib().gpl_set_density(self._h, region, _fp(a)), "gpl_set_density", self._h)

    def density_solve(self, region, want_phi=False):
        M = self._m(region)
        phi = np.zeros(M, np.float32) if want_phi else None
        fx = np.zeros(M, np.float32)
        fy = np.zeros(M, np.float32)
        _ok(_ffi.lib().gpl_density_solve(self._h, region, _fp(phi) if want_phi else None, _fp(fx), _fp(fy)),
            "gpl_density_solve", self._h)
        return phi, fx, fy

    def objective_forward(self, xy, wl_coef, lam):
        """The objective WA + lam * 1/2 sum rho phi at region 0's centres xy: (value, wa, density)."""
        a = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        assert a.size == 2 * self._n(0)
        v, w, d = C.c_double(), C.c_double(), C.c_double()
        _ok(_ffi.lib().gpl_objective_forward(self._h, _fp(a), wl_coef, lam, C.byref(v), C.byref(w), C.byref(d)),
            "gpl_objective_forward", self._h)
        return v.value, w.value, d.value

    def objective_backward(self, xy, wl_coef, lam):
        """d/dxy of the objective per cell of region 0 (problem order)."""
        a = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1)
        assert a.size == 2 * self._n(0)
        g = np.zeros_like(a)
        _ok(_ffi.lib().gpl_objective_backward(self._h, _fp(a), wl_coef, lam, _fp(g)), "gpl_objective_backward", self._h)
        return g.reshape(-1, 2)

    def wa_forward(self, coef_x, coef_y):
        E = self.problem.num_nets
        sums = np.zeros(8 * E, np.float32)
        box = np.zeros(4 * E, np.int32)
        h = C.c_int64(0)
        _ok(_ffi.lib().gpl_wa_forward(self._h, coef_x, coef_y, _fp(sums), box.ctypes.data_as(C.POINTER(C.c_int32)),
                                      C.byref(h)), "gpl_wa_forward", self._h)
        return sums.reshape(-1, 8), box.reshape(-1, 4), h.value

    def update_gradients(self, region, slot, penalty):
        n = self._n(region)
        wl = np.zeros(2 * n, np.float32)
        den = np.zeros(2 * n, np.float32)
        sm = np.zeros(2 * n, np.float32)
        ab = np.zeros(2, np.float64)
        _ok(_ffi.lib().gpl_update_gradients(self._h, region, slot, penalty, _fp(wl), _fp(den), _fp(sm),
                                            ab.ctypes.data_as(C.POINTER(C.c_double))), "gpl_update_gradients", self._h)
        return wl.reshape(-1, 2), den.reshape(-1, 2), sm.reshape(-1, 2), ab

    def nesterov_step(self, region, step, coeff):
        n = self._n(region)
        nx = np.zeros(2 * n, np.float32)
        ns = np.zeros(2 * n, np.float32)
        _ok(_ffi.lib().gpl_nesterov_step(self._h, region, step, coeff, _fp(nx), _fp(ns)), "gpl_nesterov_step", self._h)
        return nx.reshape(-1, 2), ns.reshape(-1, 2)

    def close(self):
        if getattr(self, "_h", None):
            _ffi.lib().gpl_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- the command (run.rs) -------------------------------------------------------------------

_KEYS = [
    "-bin_grid_count", "-density", "-init_density_penalty", "-init_wirelength_coef", "-min_phi_coef", "-max_phi_coef",
    "-overflow", "-reference_hpwl", "-initial_place_max_iter", "-initial_place_max_fanout",
    "-routability_check_overflow", "-routability_snapshot_overflow", "-routability_max_density",
    "-routability_target_rc_metric", "-routability_inflation_ratio_coef", "-routability_max_inflation_ratio",
    "-routability_min_congestion_for_inflation", "-routability_max_inflation_total", "-routability_net_weight_max",
    "-routability_congested_nets_percentage", "-routability_rc_coefficients", "-timing_driven_net_reweight_overflow",
    "-timing_driven_net_weight_max", "-timing_driven_nets_percentage", "-timing_driven_repair_tns_end_percent",
    "-keep_resize_below_overflow", "-virtual_cts_max_skew_fraction", "-random_seed", "-perturb_dist", "-pad_left",
    "-pad_right",
]
_FLAGS = [
    "-skip_initial_place", "-force_center_initial_place", "-skip_nesterov_place", "-timing_driven",
    "-timing_driven_repair_timing", "-routability_driven", "-virtual_cts", "-routability_use_grt", "-skip_io",
    "-place_ios", "-incremental", "-disable_revert_if_diverge", "-disable_pin_density_adjust",
    "-enable_routing_congestion",
]
_MODELLED_FLAGS = {"-skip_initial_place", "-skip_nesterov_place", "-skip_io", "-disable_revert_if_diverge",
                   "-disable_pin_density_adjust", "-force_center_initial_place", "-routability_driven"}


def parse_key_args(args, keys=_KEYS, flags=_FLAGS):
    """sta::parse_key_args + check_argc_eq0 (run.rs:56-92): an exact key or flag, else the first
    key, then the first flag, the argument is a prefix of.  Raises GplError with the STA id."""
    k, f = {}, []
    positional = False
    i = 0
    while i < len(args):
        a = args[i]
        if not (len(a) >= 2 and a[0] == "-" and a[1].isalpha()):
            positional = True
            i += 1
            continue
        key = a if a in keys else (None if a in flags else next((x for x in keys if x.startswith(a)), None))
        flag = None
        if key is None:
            flag = a if a in flags else next((x for x in flags if x.startswith(a)), None)
        if key is not None:
            if i + 1 >= len(args):
                raise GplError(560 if a in keys else 561, f"global_placement {key} missing value.")
            k[key] = args[i + 1]
            i += 2
        elif flag is not None:
            f.append(flag)
            i += 1
        else:
            raise GplError(562, f"global_placement {a} is not a known keyword or flag.")
    if positional:
        raise GplError(564, "global_placement positional arguments not supported.")
    return k, f


def opts_from_args(args, base=None):
    """The options `global_placement` reads from its arguments (run.rs:225-272)."""
    keys, flags = parse_key_args(args)
    o = base if base is not None else default_opts()
    for fl in flags:
        if fl not in _MODELLED_FLAGS:
            raise GplError(GPL_ERR_REFUSED, f"{fl} is not modelled by this engine")
    conv = {
        "-bin_grid_count": ("bin_cnt", int), "-init_density_penalty": ("init_density_penalty", float),
        "-init_wirelength_coef": ("init_wirelength_coef", float), "-min_phi_coef": ("min_phi_coef", float),
        "-max_phi_coef": ("max_phi_coef", float), "-overflow": ("overflow", float),
        "-reference_hpwl": ("reference_hpwl", float), "-initial_place_max_iter": ("initial_place_max_iter", int),
        "-initial_place_max_fanout": ("initial_place_max_fanout", int), "-random_seed": ("seed", int),
        "-perturb_dist": ("perturb_dist", float), "-pad_left": ("pad_left", int), "-pad_right": ("pad_right", int),
    }
    route_conv = {
        "-routability_check_overflow": "routability_end_overflow", "-routability_snapshot_overflow": "routability_snapshot_overflow",
        "-routability_target_rc_metric": "routability_target_rc", "-routability_max_density": "routability_max_density",
        "-routability_inflation_ratio_coef": "routability_inflation_ratio_coef",
        "-routability_max_inflation_ratio": "routability_max_inflation_ratio",
        "-routability_min_congestion_for_inflation": "routability_min_congestion",
        "-routability_max_inflation_total": "routability_max_inflation_total",
        "-routability_net_weight_max": "routability_net_weight_max",
        "-routability_congested_nets_percentage": "routability_congested_nets_percentage",
    }
    for key, val in keys.items():
        if key in route_conv:
            setattr(o, route_conv[key], float(val))
        elif key == "-routability_rc_coefficients":
            try:
                k = [float(x) for x in val.split()]
            except ValueError:
                k = []
            if len(k) != 4:
                raise GplError(GPL_ERR_REFUSED, f"-routability_rc_coefficients {val}: not four numbers")
            o.routability_rc_k[:] = k
        elif key == "-density":
            if val == "uniform":
                o.uniform_density = 1
            else:
                o.density = float(val)
        elif key in conv:
            name, ty = conv[key]
            setattr(o, name, ty(val))
        else:
            raise GplError(GPL_ERR_REFUSED, f"{key} is not modelled by this engine")
    o.skip_initial_place = int("-skip_initial_place" in flags)
    o.skip_io = int("-skip_io" in flags)
    o.disable_revert_if_diverge = int("-disable_revert_if_diverge" in flags)
    o.disable_pin_density_adjust = int("-disable_pin_density_adjust" in flags)
    # -routability_driven, unless -skip_io turned it off (run.rs:253-269)
    o.routability = int(o.routability or ("-routability_driven" in flags and not o.skip_io))
    if o.routability and "-routability_use_grt" in flags:
        raise GplError(GPL_ERR_REFUSED, "-routability_use_grt is not modelled (no global router)")
    return o, flags


def global_placement(cfg, args=(), seed=1):
    """The `global_placement` command on a synthetic design (run.rs:115-201): returns a dict with
    the final iteration, the placed locations (placer instance, lx, ly), the log and stats."""
    if isinstance(cfg, str):
        base = default_opts(**CONFIGS[cfg]["opts"])
        gcfg = gen_config(cfg, seed=seed)
    else:
        base, gcfg = default_opts(), cfg
    o, flags = opts_from_args(list(args), base)
    d = Design(gcfg, o)
    try:
        if not o.skip_initial_place:
            d.initial_place()
        if "-skip_nesterov_place" in flags:
            return dict(outcome="InitiallyPlaced", iter=None, info=d.info())
        d.init_nesterov_place(o)
        g = NesterovGPU(d)
        try:
            it = g.do_nesterov_place()
            return dict(outcome="Placed", iter=it, locations=g.placed_locations(), log=g.log(), stats=g.stats(),
                        history=g.history(), info=d.info())
        finally:
            g.close()
    finally:
        d.close()
