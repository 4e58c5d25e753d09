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
