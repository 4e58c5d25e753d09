"""Routability's congestion measurement on the GPU (gpl_rudy, gpl_rudy_tiles) against the oracle's
restatement of route_base.rs on the same net boxes and cell centres.

Bars: the tile demand and blockage bit-exact (the reference's float sums in the nets' order); the
actionable tiles, the counts, the total overflow, the four top averages and the weighted congestion
bit-exact; the inflation ratios within one float ulp (the device squares the ratio, x * x, where the
oracle calls the C library's powf(x, 2), which is not correctly rounded everywhere)."""
import numpy as np
import pytest

import gpl_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


def spread_points(design, iters):
    g = G.NesterovGPU(design)
    try:
        g.init_nesterov_loop()
        g.loop_begin(0)
        g.loop_step(iters)
        return np.round(g.get_coordinates(0, G.GPL_SLOT_CUR)).astype(np.float32)
    finally:
        g.close()


def grid_for(design, cnt, tracks=8):
    R = design.problem().regions[0]
    die = (R.lx, R.ly, R.ux, R.uy)
    tile = max(1, (R.ux - R.lx) // cnt)
    # two routing layers of pitch tile / tracks (8 horizontal and 8 vertical tracks per tile)
    ww = G.rudy_wire_width([tile // tracks, tile // tracks])
    assert ww == O.rudy_wire_width([tile // tracks, tile // tracks])
    return dict(lx=die[0], ly=die[1], ux=die[2], uy=die[3], cnt_x=cnt, cnt_y=cnt, tile_size=tile, wire_width=ww), die


def measure(design, xy, cnt, cap_red=None):
    """The GPU's demand, blockage, tiles and statistics, and the oracle's on the GPU's net boxes and
    the oracle's own density centres at the same points."""
    a = design.arrays()
    grid, die = grid_for(design, cnt)
    g = G.NesterovGPU(design)
    o = O.Oracle(design.problem())
    try:
        g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
        g.density_scatter(0, G.GPL_SLOT_CUR_SLP)  # the cells centred at xy
        _, boxes, _ = g.wa_forward(np.float32(1e-4), np.float32(1e-4))
        demand, blockage = g.rudy(grid, cap_red)
        ratio, act, st = g.rudy_tiles(grid)
        o.scatter(0, xy)
        db = o.dbox()
    finally:
        g.close()
        o.close()
    od, ob = O.rudy(die, cnt, cnt, grid["tile_size"], grid["wire_width"], boxes, cap_red)
    flags = a["cell_flags"]
    std = (flags & 1).astype(bool) & ~(flags & 2).astype(bool)
    centres = np.stack([(db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2], 1)[std]
    oratio, oact, ost = O.rudy_tiles(cnt, cnt, grid["tile_size"], demand, blockage, centres)
    return (demand, blockage, ratio, act, st), (od, ob, oratio, oact, ost)


@pytest.fixture(scope="module")
def c1():
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


def check(gpu, orc):
    demand, blockage, ratio, act, st = gpu
    od, ob, oratio, oact, ost = orc
    assert np.array_equal(demand, od) and np.array_equal(blockage, ob)
    assert demand.max() > 0
    assert np.array_equal(act, oact)
    ulp = np.spacing(np.maximum(np.abs(oratio), 1.0).astype(np.float32))
    assert np.all(np.abs(ratio - oratio) <= ulp)
    assert st["actionable"] == ost["actionable"] and st["overflowed"] == ost["overflowed"]
    assert st["total_overflow"] == ost["total_overflow"]
    assert np.array_equal(st["top"], ost["top"]) and st["rc"] == np.float32(ost["rc"])


@pytest.mark.parametrize("cnt", [16, 64])
def test_rudy_and_tiles_match_the_oracle(c1, cnt):
    """The 2k-cell design spread by 150 iterations of the loop, RUDY on 16 x 16 and 64 x 64 tiles
    (a stretched last column and row where the die is not a whole number of tiles)."""
    xy = spread_points(c1, 150)
    gpu, orc = measure(c1, xy, cnt)
    check(gpu, orc)
    assert gpu[4]["actionable"] > 0


def test_blockage_and_unactionable_tiles_match_the_oracle(c1):
    """Random (capacity, reduction) pairs: zero capacities (blockage 1, congestion infinite),
    reductions above the capacity (blockage capped at 1) and tiles blocked past the 0.8 edge ratio
    (not actionable, not averaged)."""
    cnt = 32
    rng = np.random.default_rng(4)
    cap = rng.integers(0, 20, cnt * cnt).astype(np.uint8)
    red = rng.integers(0, 25, cnt * cnt).astype(np.uint8)
    cap_red = np.stack([cap, red], 1)
    xy = spread_points(c1, 150)
    gpu, orc = measure(c1, xy, cnt, cap_red)
    check(gpu, orc)
    assert (gpu[1] == 1.0).any() and (gpu[1] < 0.8).any()


def test_rudy_at_200k_cells():
    """The 200k-cell design (20 nets of 500-2,000 pins spanning many tiles: the block-per-net pairs)
    on 512 x 512 tiles: bit-exact with the oracle."""
    d = G.Design.from_config("c2")
    d.initial_place()
    d.init_nesterov_place()
    try:
        xy = spread_points(d, 100)
        gpu, orc = measure(d, xy, 512)
    finally:
        d.close()
    check(gpu, orc)


def test_inflation_and_density_size_match_the_oracle(c1):
    """The tiles' inflated ratios (max(r, 1)) of a measured RUDY grid: the standard cells in
    inflated tiles grown on the GPU as the oracle grows them (sizes and area added, exact), then the
    density boxes and scales of every cell at the new sizes (updateDensitySize) as the oracle
    computes them (a box of density size d spans 2 (d / 2), rebuilt from halves)."""
    cnt = 32
    a = c1.arrays()
    R = c1.problem().regions[0]
    xy = spread_points(c1, 150)
    grid, die = grid_for(c1, cnt)
    box = a["cell_box"]
    sizes = np.stack([box[:, 2] - box[:, 0], box[:, 3] - box[:, 1]], 1).astype(np.int32)
    g = G.NesterovGPU(c1)
    try:
        g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
        g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
        g.rudy(grid)
        ratio, _, _ = g.rudy_tiles(grid, min_cong=0.3)  # a low threshold: many inflated tiles
        infl = np.maximum(ratio, np.float32(1.0))
        assert (infl > 1.0).sum() > 10
        new, delta = g.route_inflate(0, grid, infl, sizes)
        db, _ = g.density_boxes(0)
        g.update_density_size(0, new)
        db2, sc2 = g.density_boxes(0)
    finally:
        g.close()
    flags = a["cell_flags"]
    std = ((flags & 1) != 0) & ((flags & 2) == 0)
    cx, cy = (db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2
    boxes = np.stack([cx - sizes[:, 0] // 2, cy - sizes[:, 1] // 2, cx - sizes[:, 0] // 2 + sizes[:, 0],
                      cy - sizes[:, 1] // 2 + sizes[:, 1]], 1)
    obox, odelta = O.route_inflate(boxes, np.stack([cx, cy], 1), std.astype(np.uint8), cnt, cnt, grid["tile_size"], infl)
    osz = np.stack([obox[:, 2] - obox[:, 0], obox[:, 3] - obox[:, 1]], 1)
    assert np.array_equal(new, osz) and delta == odelta and delta > 0
    ds, sc = O.density_size(new, R.size_x, R.size_y)
    assert np.array_equal(np.stack([db2[:, 2] - db2[:, 0], db2[:, 3] - db2[:, 1]], 1), 2 * (ds // 2))
    assert np.array_equal(sc2, sc)
    # the centres stay
    assert np.array_equal((db2[:, 0] + db2[:, 2]) // 2, cx) and np.array_equal((db2[:, 1] + db2[:, 3]) // 2, cy)
