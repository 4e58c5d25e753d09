"""The oracle's routability congestion measurement (route.c), pinned by the reference's own unit-test
vectors (route_base.rs:682-757) and by what the definitions fix: the demand a box leaves on a grid of
uniform tiles sums to hpwl * wire_width / tile_area * 100 whatever tiles it covers, and the tile
inflation rule's threshold, power and cap on hand-built tiles."""
import numpy as np
import pytest

import oracle as O

DIE = (0, 0, 250, 200)  # route_base.rs:684: 2 x 2 tiles of 100, the last column and row stretched


def test_wire_width_is_the_truncated_harmonic_pitch():
    assert O.rudy_wire_width([100, 100]) == 62
    assert O.rudy_wire_width([]) == 100


def test_last_column_takes_the_die_remainder():
    rects = [O.rudy_rect(DIE, 2, 2, 100, x, y) for x in range(2) for y in range(2)]  # x-major
    assert rects == [(0, 0, 100, 100), (0, 100, 100, 200), (100, 0, 250, 100), (100, 100, 250, 200)]


def test_rudy_deposits_on_strict_overlap_only():
    imax, imin = 2 ** 31 - 1, -2 ** 31
    boxes = [(0, 0, 100, 50), (10, 10, 10, 90), (imax, imax, imin, imin)]
    demand, _ = O.rudy(DIE, 2, 2, 100, 62, boxes, cap_red=[(10, 0)] * 4)
    # hpwl 150 * 62 / 5000 = 1.86, half the tile, times 100 (all in float)
    assert demand[0] == np.float32(np.float32(1.86) * np.float32(0.5) * np.float32(100.0))
    assert demand[1:].tolist() == [0.0, 0.0, 0.0]


def test_blockage_derates_the_tile():
    _, blockage = O.rudy(DIE, 2, 2, 100, 62, [], cap_red=[(10, 5), (0, 0), (4, 9), (10, 0)])
    assert blockage.tolist() == [0.5, 1.0, 1.0, 0.0]
    assert O.rudy_congestion(50.0, 0.5) == 1.0
    assert np.isinf(O.rudy_congestion(50.0, 1.0))


def test_top_shares_round_up():
    assert O.top_averages(np.arange(10, 0, -1, dtype=np.float64)).tolist() == [10.0, 10.0, 10.0, 10.0]
    assert O.top_averages(np.arange(100, 0, -1, dtype=np.float64)).tolist() == [100.0, 100.0, 99.5, 98.0]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_total_demand_is_the_boxes_wire_area(seed):
    """On a grid of uniform tiles (the die a whole number of tiles) every box inside the die leaves
    sum over tiles of (overlap / tile area) = box area / tile area, so the grid's total is
    sum(hpwl * wire_width / tile_area * 100) -- independent of how the boxes cut the tiles."""
    rng = np.random.default_rng(seed)
    tile, nx, ny, ww = 50, 40, 30, 7
    die = (0, 0, nx * tile, ny * tile)
    lo = rng.integers(0, nx * tile - 1, (200, 1)), rng.integers(0, ny * tile - 1, (200, 1))
    w, h = rng.integers(1, 400, (200, 1)), rng.integers(1, 300, (200, 1))
    xl, yl = lo[0], lo[1]
    boxes = np.hstack([xl, yl, np.minimum(xl + w, nx * tile), np.minimum(yl + h, ny * tile)])
    demand, blockage = O.rudy(die, nx, ny, tile, ww, boxes)
    assert not blockage.any()
    want = sum(((b[2] - b[0]) + (b[3] - b[1])) * ww / (tile * tile) * 100.0 for b in boxes)
    assert abs(float(demand.astype(np.float64).sum()) - want) <= 1e-5 * want


def test_tile_inflation_threshold_power_and_cap():
    """Four tiles of 10 (y-major index y * 2 + x; the demand x-major): congestion 0.9 (below the
    0.95 threshold: ratio 1), 1.14 (0.95 * 1.2: (1.2)^2 = 1.44), 3.0 ((3/0.95)^2 > 3: the cap), and a
    blocked tile (blockage 0.8: not actionable, not averaged); only tiles holding a centre count."""
    cnt, tile = 2, 10
    # x-major: index x * 2 + y
    demand = np.zeros(4, np.float32)
    blockage = np.zeros(4, np.float32)
    demand[0 * 2 + 0] = 90.0    # tile (0, 0)
    demand[1 * 2 + 0] = 114.0   # tile (1, 0)
    demand[0 * 2 + 1] = 300.0   # tile (0, 1)
    blockage[1 * 2 + 1] = 0.8   # tile (1, 1)
    demand[1 * 2 + 1] = 500.0
    centres = [(5, 5), (15, 5), (5, 15), (15, 15)]
    ratio, act, st = O.rudy_tiles(cnt, cnt, tile, demand, blockage, centres)
    assert act.tolist() == [1, 1, 1, 0]
    assert ratio[0] == 1.0
    assert abs(ratio[1] - 1.44) < 1e-6
    assert ratio[2] == 3.0
    assert ratio[3] == 1.0
    assert st["actionable"] == 3 and st["overflowed"] == 2
    assert abs(st["total_overflow"] - (0.14 + 2.0)) < 1e-6
    # three tiles: every share is the top one (3.0), the weighted congestion too
    assert st["top"].tolist() == [3.0, 3.0, 3.0, 3.0] and st["rc"] == 3.0
    # a tile without a centre is not actionable whatever its congestion
    _, act2, st2 = O.rudy_tiles(cnt, cnt, tile, demand, blockage, centres[:1])
    assert act2.tolist() == [1, 0, 0, 0] and st2["actionable"] == 1


def test_set_size_rebuilds_from_halves():
    """nesterov.rs:4841-4845: box (0, 0, 10, 4) set to 7 x 5 becomes (2, 0, 8, 4)."""
    assert O.set_size((0, 0, 10, 4), 7, 5) == (2, 0, 8, 4)


def test_inflation_grows_by_the_root_of_the_ratio():
    """A standard cell of 10 x 4 in a tile of ratio 2.25 grows by 1.5 per side: 15 x 6, rebuilt
    from halves: 14 x 6 (area 84, 44 added); a cell in a tile of ratio 1 and a macro stay."""
    boxes = [(0, 0, 10, 4), (20, 0, 30, 4), (0, 20, 10, 24)]
    centres = [(5, 2), (25, 2), (5, 22)]
    std = [1, 1, 0]
    ratio = np.array([2.25, 1.0, 2.25, 2.25], np.float32)  # 2 x 2 tiles of 16, y-major
    out, d = O.route_inflate(boxes, centres, std, 2, 2, 16, ratio)
    assert out.tolist() == [[-2, -1, 12, 5], [20, 0, 30, 4], [0, 20, 10, 24]]
    assert d == 14 * 6 - 40


def test_density_size_threshold_and_scale():
    """Bins of 10: the threshold is sqrt(2) * 10 = 14.142...; a side of 14 spreads to
    int(float(sqrt 2 * 10)) = 14 with the scale 14 / 14.142..., a side of 15 keeps its size at scale 1."""
    ds, sc = O.density_size([(14, 15), (15, 14), (3, 40)], 10.0, 10.0)
    thr = np.float32(np.sqrt(2.0) * 10.0)
    assert ds.tolist() == [[14, 15], [15, 14], [14, 40]]
    assert sc[0] == np.float32(np.float32(14) / thr) and sc[1] == sc[0]
    assert sc[2] == np.float32(np.float32(3) / thr)


def test_density_size_is_what_init_built():
    """The 2k-cell design's density boxes and scales, as the host's init left them, are
    updateDensitySize's at the cells' own sizes (each box 2 (d / 2) wide, rebuilt from halves)."""
    import gpl_amd as G
    d = G.Design.from_config("c1")
    d.init_nesterov_place()
    try:
        a = d.arrays()
        R = d.problem().regions[0]
        box, db = a["cell_box"], a["cell_dbox"]
        sizes = np.stack([box[:, 2] - box[:, 0], box[:, 3] - box[:, 1]], 1)
        ds, sc = O.density_size(sizes, R.size_x, R.size_y)
        assert np.array_equal(np.stack([db[:, 2] - db[:, 0], db[:, 3] - db[:, 1]], 1), 2 * (ds // 2))
        assert np.array_equal(sc, a["density_scale"])
        assert (sc < 1.0).any()  # some cells spread
    finally:
        d.close()


# ---- the routability pass's host arithmetic (RouteBase::routability and its helpers) -------------
def test_non_place_at_rebuilds_the_construction():
    """updateBinsNonPlaceArea (nesterov.rs:1956-2018) from its target-independent halves at the
    initial target density gives the construction's fixed-area map bit for bit (a 20k-cell design
    with 5 % of its core in macros: the bivariate-normal overlaps and the unions; the 1.10 cap is
    pinned by the hand-computed bins)."""
    import gpl_amd as G
    d = G.Design.from_config("c2", gen_over=dict(num_cells=20_000, num_nets=21_000, num_hf_nets=0, macro_area_frac=0.05,
                                                  macro_rows_min=5, macro_rows_max=20), bin_cnt=128)
    d.init_nesterov_place()
    try:
        R = d.problem().regions[0]
        m = R.cnt_x * R.cnt_y
        raw = np.ctypeslib.as_array(R.non_place_raw, (m,)).copy()
        uni = np.ctypeslib.as_array(R.non_place_union, (m,)).copy()
        np_, npu = O.non_place_at(raw, uni, R.target_density)
        assert np.array_equal(np_, np.ctypeslib.as_array(R.non_place, (m,)))
        assert np.array_equal(npu, np.ctypeslib.as_array(R.non_place_unscaled, (m,)))
        assert (uni >= 0).sum() > 100 and (np_ > 0).sum() > 100
        assert (raw != np.maximum(uni, 0)).any()  # the macros' bivariate-normal overlaps differ from the union
    finally:
        d.close()


def test_non_place_at_hand_computed():
    td = np.float32(0.8)
    raw, uni = [1000, 0, 500, 1000], [900, 400, -1, 1000]
    a, b = O.non_place_at(raw, uni, float(td))
    f = lambda v: int(np.float32(v) * td)  # noqa: E731
    cap = int(np.float32(np.float32(900) * td) * np.float32(1.10))
    assert a.tolist() == [min(f(1000), cap), 0, f(500), f(1000)]
    assert b.tolist() == [f(900), f(400), 0, f(1000)]


def test_cut_fillers_plan_hand_computed():
    """cutFillerCells (nesterov.rs:3826-3866): whole fillers for the area, at most the fillers there
    are; the target density raised only when more than a filler's area is left over."""
    assert O.cut_fillers_plan(350, 100, 10, 50_000, 100_000, 0.7) == (3, pytest.approx(0.7))
    assert O.cut_fillers_plan(1100, 100, 10, 50_000, 100_000, 0.7) == (10, pytest.approx(0.7))  # 100 left: not more
    k, td = O.cut_fillers_plan(1250, 100, 10, 50_000, 100_000, 0.7)
    assert k == 10 and td == np.float32(np.float32(50_000 + 1000 + 0 + 250) / np.float32(100_000))
    assert O.cut_fillers_plan(-5, 100, 10, 50_000, 100_000, 0.7) == (0, pytest.approx(0.7))
    assert O.cut_fillers_plan(500, 100, 0, 50_000, 100_000, 0.7)[0] == 0


def test_budget_trim_hand_computed():
    """scaleInflationToBudget (nesterov.rs:3946-3978): a pass may add min(budget / 6, what is left);
    over it every excess above 1 is scaled by allowed / wanted."""
    spent, r = O.budget_trim(600.0, 0.0, 50.0, [3.0, 1.0, 0.8])
    assert not spent and r.tolist() == [3.0, 1.0, pytest.approx(0.8)]
    spent, r = O.budget_trim(600.0, 0.0, 200.0, [3.0, 1.0, 1.5])
    assert not spent and r.tolist() == [2.0, 1.0, 1.25]
    spent, r = O.budget_trim(600.0, 550.0, 100.0, [3.0])
    assert spent and r.tolist() == [2.0]
    spent, _ = O.budget_trim(600.0, 590.0, 5.0, [1.2])
    assert not spent
    spent, _ = O.budget_trim(600.0, 590.0, 10.0, [1.2])
    assert spent


def test_congested_net_weights_hand_computed():
    """weightCongestedNets (route_base.rs:381-426): the top share of the depositing nets, weighted
    linearly from the maximum at the worst to 1 at the cutoff."""
    n, w = O.congested_net_weights([0, 5, 3, 1, 5, 2], 50.0, 2.0)
    assert n == 5 and w.tolist() == [0.0, 2.0, 1.0, 0.0, 2.0, 0.0]  # ceil(5 * 0.5) = 3 nets: 5, 5, 3
    n, w = O.congested_net_weights([0, 5, 3, 1, 5, 2], 1.0, 1.5)
    assert n == 5 and w.tolist() == [0.0, 1.5, 0.0, 0.0, 0.0, 0.0]  # one net, worst == cutoff: the maximum
    n, w = O.congested_net_weights([0, 4, 2, 1], 100.0, 3.0)
    assert w.tolist() == [0.0, 3.0, pytest.approx(1.0 + 2.0 * 1.0 / 3.0), 1.0]
    assert O.congested_net_weights([1, 2], 10.0, 1.0)[0] == -1
    assert O.congested_net_weights([0, 0], 10.0, 2.0)[0] == -1
