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
