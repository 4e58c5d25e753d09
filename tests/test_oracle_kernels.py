"""The oracle's per-kernel arithmetic on hand-built problems with hand-computed answers, and on
the small workload through invariants the mathematics fixes."""
import math

import numpy as np

import gpl_amd as G
import oracle as O
from tests.tinyproblem import Tiny, two_by_two


def test_scatter_hand_computed_2x2():
    cells = [
        dict(box=(50, 50, 150, 150)),                                   # 50x50 in each of the four bins
        dict(box=(0, 0, 100, 100), scale=0.5, flags=0),                 # a filler in bin (0,0) at half scale
        dict(box=(120, 20, 180, 80), scale=0.3333, flags=1 | 2),        # a macro: x target density
    ]
    t = two_by_two(cells, target_density=0.8, non_place=np.array([0, 1000, 0, 0]))
    o = O.Oracle(t.problem)
    centres = np.array([[100, 100], [50, 50], [150, 50]], np.float32)
    inst, fill, dens, ovi, ovf = o.scatter(0, centres)
    a_macro = int(np.float32(np.float32(60 * 60) * np.float32(0.3333)) * np.float32(0.8))
    assert inst.tolist() == [2500, 2500 + a_macro, 2500, 2500]
    assert fill.tolist() == [5000, 0, 0, 0]
    scaled = np.float32(100 * 100) * np.float32(0.8)
    want = [(np.float32(inst[k]) + np.float32(fill[k]) + np.float32([0, 1000, 0, 0][k])) / scaled for k in range(4)]
    assert dens.tolist() == [float(w) for w in want]
    # overflow: max(0, inst + non_place - area * target) per bin, summed
    terms = [max(0.0, float(np.float32(inst[k]) + np.float32([0, 1000, 0, 0][k]) - scaled)) for k in range(4)]
    assert abs(ovf[0] - sum(terms)) < 1e-6


def test_scatter_truncates_each_addend():
    # scale 0.3333 on a 50 x 50 overlap: 833.25 truncated per bin, summed over the bins
    t = two_by_two([dict(box=(50, 50, 150, 150), scale=0.3333)])
    o = O.Oracle(t.problem)
    inst, *_ = o.scatter(0, np.array([[100, 100]], np.float32))
    assert inst.tolist() == [833] * 4


def test_gather_hand_computed_2x2():
    t = two_by_two([dict(box=(50, 50, 150, 150), scale=0.5)])
    o = O.Oracle(t.problem)
    o.scatter(0, np.array([[100, 100]], np.float32))
    fx = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    fy = np.array([-1.0, 0.5, 0.25, 8.0], np.float32)
    den, d64 = o.gather(0, fx, fy, fx.astype(np.float64), fy.astype(np.float64))
    a = 2500 * 0.5
    assert den[0].tolist() == [a * 10.0, a * 7.75]
    assert np.allclose(d64[0], [a * 10.0, a * 7.75])


def test_wa_gradient_expression_is_minus_the_derivative():
    """With the exact exponential the reference's gmin - gmax is -dW/dx (central differences in
    FP64), for random nets of 2..12 pins and several coefficients."""
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(2, 13))
        xs = rng.uniform(0, 5000, n)
        coef = float(rng.choice([1e-4, 5e-4, 2e-3]))
        g = np.zeros(n)
        O.lib().orc_wa_net_f64(n, O.f64(xs), coef, 0, O.f64(g))
        h = 1e-3
        for i in range(n):
            xp, xm = xs.copy(), xs.copy()
            xp[i] += h
            xm[i] -= h
            fd = (O.lib().orc_wa_net_f64(n, O.f64(xp), coef, 0, None) - O.lib().orc_wa_net_f64(n, O.f64(xm), coef, 0, None)) / (2 * h)
            assert abs(g[i] + fd) < 1e-6 * max(1.0, abs(fd)), (trial, i, g[i], fd)
        # translation invariance: the gradients of a net sum to zero
        assert abs(g.sum()) < 1e-9


def test_wa_two_pin_net_faithful_matches_expression():
    """A two-pin net on two one-pin cells: the oracle's float gradient (fast exponential, double
    quotient) agrees with the one-net expression evaluated in double with the fast exponential."""
    cells = [dict(box=(0, 0, 100, 100), pins=[(0, 0)]), dict(box=(0, 0, 100, 100), pins=[(0, 0)])]
    t = Tiny(cells, [[0, 1]], regions=[dict(lx=0, ly=0, ux=20000, uy=20000, cnt_x=4, cnt_y=4)])
    o = O.Oracle(t.problem)
    xy = np.array([[1000, 3000], [9000, 4000]], np.float32)
    o.place_pins(0, xy)
    coef = np.float32(2.5e-4)
    sums, box, s64, h = o.wa_forward(coef, coef)
    assert box[0].tolist() == [1000, 3000, 9000, 4000] and h == 8000 + 1000
    wl, w64 = o.wa_backward(0, s64)
    for axis in range(2):
        xs = xy[:, axis].astype(np.float64)
        g = np.zeros(2)
        O.lib().orc_wa_net_f64(2, O.f64(xs), float(coef), 1, O.f64(g))
        assert np.allclose(wl[:, axis], g, rtol=1e-5), (axis, wl[:, axis], g)
    # two pins: equal and opposite
    assert np.allclose(wl[0], -wl[1], rtol=1e-5)


def test_scatter_conserves_area_on_the_small_workload(c1_design):
    """Sum over bins of the integer areas equals the cells' scaled areas inside the region, less
    under one unit per touched bin (each addend is truncated)."""
    a = c1_design.arrays()
    R = a["regions"][0]
    o = O.Oracle(c1_design.problem())
    db = a["cell_dbox"]
    xy = np.stack([(db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2], 1).astype(np.float32)
    inst, fill, dens, ovi, ovf = o.scatter(0, xy)
    boxes = o.dbox()
    total = 0.0
    touched = 0
    for g in range(a["num_cells"]):
        lx, ly, ux, uy = boxes[g]
        cx0, cy0 = max(lx, R["lx"]), max(ly, R["ly"])
        cx1, cy1 = min(ux, R["ux"]), min(uy, R["uy"])
        area = max(0, cx1 - cx0) * max(0, cy1 - cy0)
        total += area * float(a["density_scale"][g])
        touched += (math.ceil((ux - R["lx"]) / R["size_x"]) - int((lx - R["lx"]) / R["size_x"])) * \
                   (math.ceil((uy - R["ly"]) / R["size_y"]) - int((ly - R["ly"]) / R["size_y"]))
    got = int(inst.sum() + fill.sum())
    assert total - touched - 1 <= got <= total + 1e-3 * total / 1e6 + 1
    # the faithful overflow (float through an integer) is close to the exact sum
    assert abs(ovi[0] - ovf[0]) <= 1e-4 * max(ovf[0], 1.0) + 64


def test_faithful_loop_converges_on_the_small_workload(c1_design):
    o = O.Oracle(c1_design.problem())
    assert o.init() == 0
    rc, it = o.run()
    s = o.stats()
    assert rc == 0 and s.converged == 1
    assert s.overflow <= 0.1
    h = o.history()
    assert h[-1]["overflow"] <= 0.1 < h[-2]["overflow"]
    # the density penalty grows on average (phi >= 0.95 each step, mostly above 1)
    assert h[-1]["density_penalty"] > h[0]["density_penalty"]
