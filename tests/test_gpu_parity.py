"""GPU parity: every kernel of the HIP path against the oracle on the same inputs, through the C
ABI, and the whole loop against the oracle's run.

Bars: integer work (bin areas, net boxes, HPWL) bit-exact; the density map and the per-net sums
of nets up to 128 pins bit-exact (same float expression, same order); the WA gradient within
1e-5 (normwise, relative to its largest magnitude) of the FP64 oracle and bit-exact with the
faithful float one; the field within 1e-4 of the FP64 solve; the final HPWL within 0.5 % of the
oracle's run on the same seed, both reaching overflow 0.10."""
import numpy as np
import pytest

import gpl_amd as G
import oracle as O
from tests.tinyproblem import Tiny, two_by_two

pytestmark = pytest.mark.gpu


def rel(x, y):
    return float(np.max(np.abs(np.asarray(x, np.float64) - y)) / max(np.max(np.abs(y)), 1e-300))


def centres(a):
    db = a["cell_dbox"]
    return np.stack([(db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2], 1).astype(np.float32)


@pytest.fixture(scope="module")
def c1():
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


def test_scatter_is_bit_exact_on_the_small_workload(c1):
    a = c1.arrays()
    g = G.NesterovGPU(c1)
    o = O.Oracle(c1.problem())
    rng = np.random.default_rng(1)
    xy = centres(a)
    for k in range(3):  # the init centres, then two random moves
        g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
        inst, fill, dens, ovf = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
        oi, of_, od, ovi, ovf64 = o.scatter(0, xy)
        assert np.array_equal(inst, oi) and np.array_equal(fill, of_)
        assert np.array_equal(dens, od)
        assert abs(ovf[0] - ovf64[0]) <= 1e-12 * max(1.0, ovf64[0]) and abs(ovf[1] - ovf64[1]) <= 1e-12 * max(1.0, ovf64[1])
        xy = (xy + rng.normal(0, 3000, xy.shape)).astype(np.float32)
    g.close()


def test_scatter_hand_computed_on_the_gpu():
    cells = [dict(box=(50, 50, 150, 150)), dict(box=(0, 0, 100, 100), scale=0.5, flags=0),
             dict(box=(120, 20, 180, 80), scale=0.3333, flags=1 | 2)]
    t = two_by_two(cells, target_density=0.8, non_place=np.array([0, 1000, 0, 0]))
    g = _ctx(t)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, np.array([[100, 100], [50, 50], [150, 50]], np.float32))
    inst, fill, dens, ovf = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    a_macro = int(np.float32(np.float32(60 * 60) * np.float32(0.3333)) * np.float32(0.8))
    assert inst.tolist() == [2500, 2500 + a_macro, 2500, 2500]
    assert fill.tolist() == [5000, 0, 0, 0]
    g.close()


@pytest.mark.parametrize("bin_size", [1000, 100_000])
def test_scatter_tables_of_both_widths_match_the_oracle(bin_size):
    """3,000 random cells (several chunks, overlapping, fillers and macros) on 8 x 8 bins: bins of
    1e3 DBU take the 32-bit LDS sums, bins of 1e5 DBU (sums up to ~1e13) the 64-bit ones."""
    rng = np.random.default_rng(bin_size)
    n, side = 3000, 8 * bin_size
    cells = []
    for i in range(n):
        w, h = (rng.uniform(0.2, 2.0, 2) * bin_size).astype(int) + 1
        flags = 1 | (2 if i % 97 == 0 else 0) if i % 5 else 0
        cells.append(dict(box=(0, 0, int(w), int(h)), scale=float(rng.uniform(0.3, 1.0)), flags=flags))
    t = Tiny(cells, [], regions=[dict(lx=0, ly=0, ux=side, uy=side, cnt_x=8, cnt_y=8, target_density=0.9)])
    xy = rng.uniform(0.1 * side, 0.9 * side, (n, 2)).astype(np.float32)
    g = _ctx(t)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
    inst, fill, dens, _ = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    g.close()
    o = O.Oracle(t.problem)
    oi, of_, od, _, _ = o.scatter(0, xy)
    o.close()
    assert np.array_equal(inst, oi) and np.array_equal(fill, of_) and np.array_equal(dens, od)
    if bin_size == 100_000:
        assert inst.max() > 2 ** 32  # the sums need 64 bits


def _ctx(tiny):
    """A GPU context on a hand-built problem (no design object)."""
    import ctypes as C
    from gpl_amd import _ffi as F
    h = C.c_void_p()
    rc = F.lib().gpl_ctx_create(C.byref(tiny.problem), C.byref(h))
    assert rc == 0
    g = G.NesterovGPU.__new__(G.NesterovGPU)
    g._h = h
    g.problem = tiny.problem
    g.regions = [tiny.problem.regions[b] for b in range(tiny.problem.num_regions)]
    g._tiny = tiny
    return g


@pytest.mark.parametrize("nx,ny,bx,by", [(64, 64, 1116, 1125), (512, 256, 100, 200), (1024, 1024, 100, 100), (768, 768, 100, 100),
                                         (96, 384, 300, 100), (24, 48, 1000, 700)])
def test_solve_matches_fp64_within_1e4(nx, ny, bx, by):
    """Powers of two and 3 * 2^k sides (the rows' half-length FFT with a radix-3 stage) against the
    oracle's FP64 solve."""
    t = Tiny([dict(box=(0, 0, 10, 10))], [], regions=[dict(lx=0, ly=0, ux=nx * bx, uy=ny * by, cnt_x=nx, cnt_y=ny)])
    R = t.problem.regions[0]
    wx = np.ctypeslib.as_array(R.wx, (nx,)).copy()
    wy = np.ctypeslib.as_array(R.wy, (ny,)).copy()
    g = _ctx(t)
    rng = np.random.default_rng(nx + ny)
    # a density map of the kind the loop makes: smooth blobs plus per-bin noise
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny))
    dens = (0.8 * np.exp(-((X - nx / 3) ** 2 + (Y - ny / 2) ** 2) / (0.02 * nx * ny)) + 0.3 * rng.random((ny, nx))).astype(np.float32).reshape(-1)
    g.set_density(0, dens)
    phi, fx, fy = g.density_solve(0, want_phi=True)
    p64, fx64, fy64 = O.solve_f64(nx, ny, wx, wy, dens)
    assert rel(fx, fx64) < 1e-4 and rel(fy, fy64) < 1e-4 and rel(phi, p64) < 1e-4
    # a single cosine mode is an eigenfunction: field_x = wx_u / (wx_u^2 + wy_v^2) sin cos
    u, v = 3, 5
    cx, sx = np.cos(np.pi * u * (X + 0.5) / nx), np.sin(np.pi * u * (X + 0.5) / nx)
    cy = np.cos(np.pi * v * (Y + 0.5) / ny)
    g.set_density(0, (cx * cy).astype(np.float32).reshape(-1))
    _, fx, _ = g.density_solve(0)
    w2 = float(wx[u]) ** 2 + float(wy[v]) ** 2
    want = (float(wx[u]) / w2 * sx * cy).reshape(-1)
    assert rel(fx, want) < 1e-4
    g.close()


def test_wa_forward_and_gradients_match_the_oracle(c1):
    a = c1.arrays()
    g = G.NesterovGPU(c1)
    o = O.Oracle(c1.problem())
    xy = centres(a) + np.float32(0.0)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
    _, _, dens, _ = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    o.scatter(0, xy)
    for coef in (np.float32(2e-5), np.float32(2e-4), np.float32(1e-3)):
        sums, box, h = g.wa_forward(coef, coef)
        osums, obox, os64, oh = o.wa_forward(coef, coef)
        assert np.array_equal(box, obox) and h == oh
        assert np.array_equal(sums, osums)          # every net here has <= 128 pins
        assert rel(sums, os64) < 1e-5
        _, fx, fy = g.density_solve(0)
        wl, den, sm, ab = g.update_gradients(0, G.GPL_SLOT_PREV_SLP, np.float32(1e-8))
        owl, ow64 = o.wa_backward(0, os64)
        assert np.array_equal(wl, owl)
        assert rel(wl, ow64) < 1e-5
        oden, _ = o.gather(0, fx, fy)
        assert np.array_equal(den, oden)


def test_gather_matches_fp64_field(c1):
    a = c1.arrays()
    R = a["regions"][0]
    g = G.NesterovGPU(c1)
    o = O.Oracle(c1.problem())
    xy = centres(a)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
    _, _, dens, _ = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    o.scatter(0, xy)
    _, fx, fy = g.density_solve(0)
    p64, fx64, fy64 = O.solve_f64(R["cnt_x"], R["cnt_y"], R["wx"], R["wy"], dens)
    g.wa_forward(np.float32(1e-4), np.float32(1e-4))
    wl, den, sm, ab = g.update_gradients(0, G.GPL_SLOT_PREV_SLP, np.float32(0.0))
    _, d64 = o.gather(0, fx, fy, fx64, fy64)
    assert rel(den, d64) < 1e-4
    g.close()


def test_nesterov_step_is_the_reference_arithmetic():
    """nesterovUpdateCoordinates on a hand-built region from a nonzero preconditioned gradient (the
    density part of updateGradients, penalty 1): next and next_slp bit-exact with the oracle's
    update, a locked cell holding still and a cell pushed against the grid's edge."""
    cells = [dict(box=(0, 0, 100, 100)), dict(box=(0, 0, 300, 100)), dict(box=(0, 0, 100, 100), flags=1 | 4),
             dict(box=(0, 0, 700, 500)), dict(box=(0, 0, 200, 200), flags=0)]
    t = Tiny(cells, [], regions=[dict(lx=0, ly=0, ux=10000, uy=10000, cnt_x=8, cnt_y=8)])
    g = _ctx(t)
    cur_slp = np.array([[500.5, 600.25], [9900.0, 50.0], [700.0, 700.0], [800.0, 900.0], [650.0, 580.0]], np.float32)
    cur = np.array([[400.0, 650.0], [9800.0, 100.0], [650.0, 690.0], [700.0, 950.0], [600.0, 600.0]], np.float32)
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, cur_slp)
    g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    g.density_solve(0)
    _, _, sm, _ = g.update_gradients(0, G.GPL_SLOT_CUR_SLP, np.float32(1.0))  # fills the cur_slp sum slot
    assert np.abs(sm).max() > 0
    g.set_coordinates(0, G.GPL_SLOT_CUR, cur)
    dd = np.array([[c["box"][2] - c["box"][0], c["box"][3] - c["box"][1]] for c in cells], np.int32)
    locked = np.array([bool(c.get("flags", 1) & 4) for c in cells])
    for s in (np.float32(3.5), np.float32(5e4 / np.abs(sm).max())):
        coeff = np.float32(0.25)
        nx, ns = g.nesterov_step(0, s, coeff)
        ox, osl = O.nesterov_points(cur_slp, sm, cur, locked, dd, (0, 0, 10000, 10000), s, coeff)
        assert np.array_equal(nx, ox) and np.array_equal(ns, osl)
        assert nx[2].tolist() == cur[2].tolist() and ns[2].tolist() == cur_slp[2].tolist()
    assert np.any(ns - dd / 2 <= 0) or np.any(ns + dd / 2 >= 10000)  # the large step reached an edge
    g.close()


@pytest.mark.parametrize("bin_size", [1000, 100_000])
@pytest.mark.parametrize("layout", ["dense", "middle", "hash"])
def test_scatter_chunk_paths_match_the_oracle(layout, bin_size):
    """Two 512-cell chunks on 128 x 128 bins, laid out for each path of the scatter: every cell in
    one 20 x 20-bin patch (the dense tile over the chunk's box); the middle half of each chunk in
    the patch and its first and last quarters spread over the grid (the tile over the middle half,
    the rest sent straight to the bins); every cell spread (the LDS hash table).  Both LDS table
    widths (bins of 1e3 and 1e5 DBU).  The cells start at one point, so the device keeps the
    problem's order and the chunks are the problem's cells 0-511 and 512-1023."""
    rng = np.random.default_rng(11)
    n, nb = 1024, 128
    side = nb * bin_size
    cells = []
    for i in range(n):
        w, h = (rng.uniform(0.3, 1.4, 2) * bin_size).astype(int) + 1
        flags = 0 if i % 7 == 0 else (1 | (2 if i % 53 == 0 else 0))
        cells.append(dict(box=(0, 0, int(w), int(h)), scale=float(rng.uniform(0.4, 1.0)), flags=flags))
    t = Tiny(cells, [], regions=[dict(lx=0, ly=0, ux=side, uy=side, cnt_x=nb, cnt_y=nb, target_density=0.8)])
    patch = rng.uniform(50 * bin_size, 70 * bin_size, (n, 2))
    spread = rng.uniform(0.02 * side, 0.98 * side, (n, 2))
    k = np.arange(n) % 512
    middle = (k >= 128) & (k < 384)
    xy = {"dense": patch, "middle": np.where(middle[:, None], patch, spread), "hash": spread}[layout].astype(np.float32)
    g = _ctx(t)
    before = g.kernel_times()["scatter_global_adds"][1]
    g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
    inst, fill, dens, _ = g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
    adds = g.kernel_times()["scatter_global_adds"][1] - before
    g.close()
    o = O.Oracle(t.problem)
    oi, of_, od, _, _ = o.scatter(0, xy)
    o.close()
    assert np.array_equal(inst, oi) and np.array_equal(fill, of_) and np.array_equal(dens, od)
    touched = int(np.count_nonzero(oi) + np.count_nonzero(of_))
    print(f"{layout} {bin_size}: {adds:.0f} global adds, {touched} nonzero bin accumulators")
    assert adds >= 1


def test_loop_matches_the_oracle_on_the_small_workload(c1):
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    g = G.NesterovGPU(d)
    it = g.do_nesterov_place()
    s = g.stats()
    o = O.Oracle(d.problem())
    assert o.init() == 0
    rc, oit = o.run()
    os_ = o.stats()
    assert s.converged == 1 and os_.converged == 1
    assert s.overflow <= 0.1 and os_.overflow <= 0.1
    assert abs(s.hpwl - os_.hpwl) / os_.hpwl < 0.005, (s.hpwl, os_.hpwl)
    assert abs(it - oit) <= max(10, 0.05 * oit), (it, oit)
    # the first rows agree closely (the trajectories separate only by FP64-vs-float sums and the
    # transform's rounding)
    gh, oh = g.history(), o.history()
    assert abs(gh[0]["hpwl"] - oh[0]["hpwl"]) / oh[0]["hpwl"] < 1e-4
    assert abs(gh[0]["overflow"] - oh[0]["overflow"]) < 1e-4
    log = g.log()
    assert "[INFO GPL-1001] Global placement finished at iteration" in log
    assert "Iteration | Overflow |     HPWL (um) |  HPWL(%) |   Penalty | Group" in log
    g.close()
    d.close()


def test_loop_is_deterministic(c1):
    runs = []
    for _ in range(2):
        d = G.Design.from_config("c1")
        d.initial_place()
        d.init_nesterov_place()
        g = G.NesterovGPU(d)
        g.do_nesterov_place()
        runs.append((g.history(), g.placed_locations()))
        g.close()
        d.close()
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])


@pytest.mark.parametrize("cfg", ["c1", "c2", "c4r_small"])
def test_launch_modes_give_the_same_placement(cfg):
    """A trial as one HIP graph (replayed every trial, with the loop's eager work -- re-sorts,
    snapshots, statistics -- between replays), as kernels on the side streams, as kernels on one
    stream, recorded and launched grouped over the regions (mode 7; the region groups: nine regions
    of six grid sizes), and profiled (the bins pass apart from the solve's first row pass, where it
    runs fused otherwise): the same iterations, rows and placement."""
    d = G.Design.from_config(cfg)
    d.initial_place()
    d.init_nesterov_place()
    out = []
    for mode in (0, 1, 2, 7, "profiled"):
        g = G.NesterovGPU(d)
        if mode == "profiled":
            g.set_profiling(True)
        else:
            g.set_launch_mode(mode)
        it = g.do_nesterov_place()
        out.append((it, g.history(), g.placed_locations(), g.stats().hpwl))
        g.close()
    d.close()
    for o in out[1:]:
        assert o[0] == out[0][0] and o[3] == out[0][3]
        assert o[1] == out[0][1]
        assert np.array_equal(o[2], out[0][2])


def test_general_kernel_paths_give_the_same_placement():
    """The gradient's per-cell gathers of the pin terms (the path for blocks whose cells hold more
    pins than the LDS stage) and the bin edges read from the region's tables (the path for bin
    grids not built by init_bins), forced by debug kind 4, against the fast paths: the same
    iterations, rows and placement, bit for bit."""
    d = G.Design.from_config("c2")
    d.initial_place()
    d.init_nesterov_place()
    out = []
    for general in (False, True):
        g = G.NesterovGPU(d)
        if general:
            g.debug_inject(4, 0)
        it = g.do_nesterov_place()
        out.append((it, g.history(), g.placed_locations(), g.stats().hpwl))
        g.close()
    d.close()
    assert out[1][0] == out[0][0] and out[1][3] == out[0][3]
    assert out[1][1] == out[0][1]
    assert np.array_equal(out[1][2], out[0][2])


def test_region_groups_match_the_oracle():
    """Nine regions (eight region groups and the top level) with their own bin grids, 8^2 .. 128^2:
    the GPU loop against the oracle's, and the trial graph (nine scatter and gradient nodes)
    against launches one by one on one stream."""
    d = G.Design.from_config("c4r_small")
    d.initial_place()
    d.init_nesterov_place()
    p = d.problem()
    assert p.num_regions == 9
    runs = []
    for mode in (0, 2):
        g = G.NesterovGPU(d)
        g.set_launch_mode(mode)
        it = g.do_nesterov_place()
        runs.append((it, g.history(), g.stats()))
        g.close()
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    o = O.Oracle(p)
    o.init()
    o.run()
    so = o.stats()
    o.close()
    s = runs[0][2]
    assert s.converged and so.converged and s.overflow <= 0.1
    assert abs(s.hpwl - so.hpwl) <= 0.005 * so.hpwl, (s.hpwl, so.hpwl)
    assert abs(runs[0][0] - so.iter) <= 10, (runs[0][0], so.iter)
    d.close()


def test_independent_seeds_run_concurrently_with_the_same_results():
    """Six placements of one design (seeds 1..6: their own perturbations) in six contexts run from
    six host threads at once give each seed the result it gets alone."""
    d = G.Design.from_config("c1")
    d.initial_place()
    sets = []
    for _ in range(2):
        ctxs = []
        for seed in range(1, 7):
            d.init_nesterov_place(G.default_opts(**dict(G.CONFIGS["c1"]["opts"], seed=seed)))
            ctxs.append(G.NesterovGPU(d))
        sets.append(ctxs)
    seq = [g.do_nesterov_place() for g in sets[0]]
    par = G.place_concurrently(sets[1])
    assert seq == par
    for a, b in zip(sets[0], sets[1]):
        assert a.history() == b.history()
        assert np.array_equal(a.placed_locations(), b.placed_locations())
    assert len({a.stats().hpwl for a in sets[0]}) > 1  # the seeds differ
    for g in sets[0] + sets[1]:
        g.close()
    d.close()


@pytest.mark.parametrize("kind,code", [(1, 306), (2, 305)])
def test_divergence_is_reported(kind, code):
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    g = G.NesterovGPU(d)
    g.init_nesterov_loop()
    g.debug_inject(kind, 5)
    rc, it = g.run_nesterov_loop(check=False)
    assert rc == code
    assert f"[ERROR GPL-0{code}]" in g.log()
    g.close()
    d.close()


def test_cell_resorts_do_not_change_the_placement(c1):
    """Re-sorting the device cells (and with them the nets) every iteration (the storage orders the
    scatter and WA tiles rely on) leaves the trajectory as it was: the same iterations, overflow
    and HPWL to FP64-sum rounding, the placed locations in the problem's order, and every net's WA
    box and sums in the problem's order."""
    runs, ctxs = [], []
    for kind in (0, 3):
        d = G.Design.from_config("c1")
        d.initial_place()
        d.init_nesterov_place()
        g = G.NesterovGPU(d)
        g.init_nesterov_loop()
        if kind:
            g.debug_inject(3, 0)
        g.run_nesterov_loop()
        runs.append((g.stats(), g.history(), g.placed_locations(), g.get_coordinates(0, G.GPL_SLOT_CUR)))
        ctxs.append((g, d))
    (s0, h0, p0, c0), (s1, h1, p1, c1_) = runs
    try:
        assert s0.iter == s1.iter and s1.converged
        assert abs(s0.hpwl - s1.hpwl) / s0.hpwl < 1e-4
        assert np.array_equal(p0[:, 0], p1[:, 0])           # same instances, same order
        assert np.max(np.abs(p0[:, 1:] - p1[:, 1:])) <= 64  # positions within rounding drift
        assert np.max(np.abs(c0 - c1_)) < 64.0
        # the nets were re-sorted with the cells: at the same points, every net's box and WA sums
        # come back in the problem's order, bit for bit
        outs = []
        for g, _ in ctxs:
            g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, c0)
            g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
            outs.append(g.wa_forward(np.float32(1e-3), np.float32(1e-3)))
        (a0, b0, q0), (a1, b1, q1) = outs
        assert q0 == q1 and np.array_equal(b0, b1) and np.array_equal(a0, a1)
    finally:
        for g, d in ctxs:
            g.close()
            d.close()
