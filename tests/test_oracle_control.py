"""Pins of the oracle's loop control (CPU only): the Nesterov update and its clamping, the step
length's ratio / acceptance / clamp / divergence rules, the preconditioned gradient, the first
penalty and step, both divergence rules, the momentum sequence, and a whole run of those
functions on a convex quadratic.

Every expected value here is worked out by hand from the reference's rule (nesterov.rs lines in
each test), or follows from mathematics the functions must satisfy (Nesterov's sequence identity,
the step length of an isotropic quadratic, convergence to a quadratic's minimiser); none restates
the oracle's expression."""
import math

import numpy as np
import pytest

import oracle as O

BOX = (0, 0, 1000, 1000)


def f32(v):
    return float(np.float32(v))


# ---- getDensityCoordiLayoutInside{X,Y} (nesterov.rs:2188-2213) ------------------------------

@pytest.mark.parametrize("c, dd, expect", [
    (500.0, 100, 500.0),   # inside: unchanged
    (20.0, 100, 50.0),     # 20 - 50 < 0: moved to 0 + 50
    (990.0, 100, 950.0),   # 990 + 50 > 1000: moved to 1000 - 50
    (50.0, 100, 50.0),     # touching the lower edge is inside
    (950.0, 100, 950.0),   # touching the upper edge is inside
    (10.0, 101, 50.5),     # half of an odd width is exact in float
    (500.0, 1200, 400.0),  # wider than the grid: both rules fire, the upper one last
])
def test_inside_keeps_a_box_in_the_grid(c, dd, expect):
    assert O.inside(c, dd, 0, 1000) == expect


# ---- nesterovUpdateCoordinates (nesterov.rs:2718-2746) ---------------------------------------

def test_nesterov_points_hand_computed():
    # cell 0: free; next = (100, 200) + 2 (10, -5) = (120, 190);
    #         next_slp = next + 0.5 (next - cur) with cur = (90, 210) -> (135, 180)
    # cell 1: locked: next = cur, next_slp = cur_slp
    # cell 2: next.x = 30 + 1 (-15) = 15, clamped by half-width 20 to 20; the look-ahead starts from
    #         the unclamped 15: 15 + 1 (15 - 10) = 20 (inside); from the clamped point it would be 30.
    #         y: 970 + 15 = 985 -> 980; look-ahead 985 + (985 - 990) = 980 (inside).
    cur_slp = [[100, 200], [300, 400], [30, 970]]
    cur = [[90, 210], [310, 390], [10, 990]]
    s = [[10, -5], [7, 7], [-15, 15]]
    dd = [[10, 10], [10, 10], [40, 40]]
    nx, ns = O.nesterov_points(cur_slp[:2], s[:2], cur[:2], [0, 1], dd[:2], BOX, 2.0, 0.5)
    assert nx.tolist() == [[120, 190], [310, 390]]
    assert ns.tolist() == [[135, 180], [300, 400]]
    nx, ns = O.nesterov_points(cur_slp[2:], s[2:], cur[2:], [0], dd[2:], BOX, 1.0, 1.0)
    assert nx.tolist() == [[20, 980]]
    assert ns.tolist() == [[20, 980]]


def test_nesterov_points_zero_step_is_the_look_ahead_of_the_momentum_alone():
    # s = 0: next = slp, next_slp = slp + coeff (slp - cur)
    rng = np.random.default_rng(3)
    slp = rng.uniform(200, 800, (50, 2)).astype(np.float32)
    cur = (slp + rng.uniform(-50, 50, slp.shape)).astype(np.float32)
    g = rng.normal(0, 10, slp.shape).astype(np.float32)
    nx, ns = O.nesterov_points(slp, g, cur, np.zeros(50), np.full((50, 2), 8), BOX, 0.0, 0.25)
    assert np.array_equal(nx, slp)
    np.testing.assert_allclose(ns, slp + 0.25 * (slp - cur), rtol=0, atol=1e-4)


# ---- nesterovUpdateStepLength (nesterov.rs:2750-2772), getDistance (319) -------------------------

def test_step_length_ratio_and_acceptance():
    # coordinates moved by (3, 4): RMS = sqrt(25 / 2); gradients by (6, 8): sqrt(100 / 2) -> 0.5
    a, b = [[0, 0]], [[3, 4]]
    ga, gb = [[0, 0]], [[6, 8]]
    ok, st, cd, dv = O.step_length(a, b, ga, gb, 1.0)
    assert (ok, st, dv) == (True, 0.5, False)
    assert cd == f32(math.sqrt(12.5))
    # 0.5 above 0.95 x 0.52 = 0.494: taken, not accepted
    ok, st, _, dv = O.step_length(a, b, ga, gb, 0.52)
    assert (ok, st, dv) == (False, 0.5, False)
    # 0.5 at most 0.95 x 0.5264 = 0.50008: accepted
    assert O.step_length(a, b, ga, gb, 0.5264)[:2] == (True, 0.5)


def test_step_length_floor_of_one_hundredth():
    # moved by (0.03, 0.04) against (6, 8): 0.005 -> set to 0.01, not accepted
    ok, st, _, dv = O.step_length([[0, 0]], [[0.03, 0.04]], [[0, 0]], [[6, 8]], 1.0)
    assert (ok, st, dv) == (False, f32(0.01), False)


def test_step_length_nan_and_inf_diverge():
    same = [[1, 2], [3, 4]]
    ok, st, _, dv = O.step_length(same, same, same, same, 0.7)  # 0 / 0
    assert (ok, dv) == (False, True) and st == f32(0.7)  # the step is left as it was
    ok, _, _, dv = O.step_length(same, [[2, 2], [3, 4]], same, same, 0.7)  # x / 0
    assert (ok, dv) == (False, True)


@pytest.mark.parametrize("curv, pre", [(2.0, 1.0), (0.25, 3.0), (8.0, 5.0)])
def test_step_length_of_an_isotropic_quadratic_is_pre_over_curvature(curv, pre):
    # f = curv/2 |x - c|^2 preconditioned by pre: sum = -curv (x - c) / pre, so the gradient moves
    # by curv / pre times the points and the ratio is pre / curv for any two point sets
    rng = np.random.default_rng(int(curv * 10))
    c = rng.uniform(0, 1000, (200, 2))
    x1, x2 = rng.uniform(0, 1000, (2, 200, 2))
    g1, g2 = (-curv * (x - c) / pre for x in (x1, x2))
    _, st, _, _ = O.step_length(x1, x2, g1, g2, 1e30)
    assert st == pytest.approx(pre / curv, rel=2e-5)


# ---- updateGradients' preconditioned sum (nesterov.rs:2469-2489) ------------------------------

def test_precondition_hand_computed():
    # cell 0: (wl + 0.5 den) / (2 pins + 0.5 x 2 x 2) = ((3 + 0.5, 4 + 1) / 4) = (0.875, 1.25)
    # cell 1: preconditioner 0 + 0.5 x 1 x 1 = 0.5 < 1 -> 1: (-1 - 1.5, -2 + 2) = (-2.5, 0)
    # cell 2: preconditioner exactly 1 (one pin, no area): wl + 0.5 den = (0.5, 0.5)
    wl = [[3, 4], [-1, -2], [0, 0]]
    den = [[1, 2], [-3, 4], [1, 1]]
    s, wa, da = O.precondition(wl, den, [2, 0, 1], [[2, 2], [1, 1], [0, 0]], 0.5)
    assert s.tolist() == [[0.875, 1.25], [-2.5, 0.0], [0.5, 0.5]]
    assert (wa, da) == (10.0, 12.0)


def test_precondition_abs_sums_are_float_sums_in_cell_order():
    # 2^24 + 1 + 1 ... in float: each +1 is lost after 2^24, so the order is visible
    wl = [[2.0 ** 24, 0]] + [[1.0, 0]] * 4
    _, wa, _ = O.precondition(wl, np.zeros((5, 2)), [1] * 5, [[1, 1]] * 5, 0.0)
    assert wa == 2.0 ** 24
    _, wa, _ = O.precondition(wl[::-1], np.zeros((5, 2)), [1] * 5, [[1, 1]] * 5, 0.0)
    assert wa == 2.0 ** 24 + 4


def test_wide_sums_switch_sums_in_double():
    # the same lost-ones case with the switch on: every +1 counts in either order; the distances
    # likewise (2^24 + four 1s of squared difference)
    wl = [[2.0 ** 24, 0]] + [[1.0, 0]] * 4
    a = np.zeros((5, 2))
    b = np.array([[2.0 ** 12, 0]] + [[1.0, 0]] * 4)
    try:
        O.set_wide_sums(True)
        _, wa, _ = O.precondition(wl, np.zeros((5, 2)), [1] * 5, [[1, 1]] * 5, 0.0)
        assert wa == 2.0 ** 24 + 4
        st, cd = O.init_step(a, b, a, a)
        assert cd == f32(np.sqrt((2.0 ** 24 + 4) / 10.0))
    finally:
        O.set_wide_sums(False)
    _, wa, _ = O.precondition(wl, np.zeros((5, 2)), [1] * 5, [[1, 1]] * 5, 0.0)
    assert wa == 2.0 ** 24
    st, cd = O.init_step(a, b, a, a)
    assert cd == f32(np.sqrt(2.0 ** 24 / 10.0))


# ---- initDensity2's first penalty and step (nesterov.rs:2600-2617) -----------------------------

def test_initial_penalty_and_step():
    assert O.init_penalty(2.0, 8.0, 7.0, 8e-5) == f32(2e-5)  # (2 / 8) x 8e-5
    assert O.init_penalty(0.0, 8.0, 7.0, 8e-5) == 7.0       # zero WA sum: left as it is
    st, cd = O.init_step([[0, 0]], [[3, 4]], [[0, 0]], [[6, 8]])
    assert st == 0.5 and cd == f32(math.sqrt(12.5))


# ---- checkDivergence (nesterov.rs:2976-2991) ---------------------------------------------------

@pytest.mark.parametrize("args, bits", [
    # rule 1: overflow < 0.2, 0.02 above its minimum, HPWL > 1.2 x the HPWL at the minimum (120)
    ((0.125, 0.0625, 100.0, 121, 0.0, 0), 1),
    ((0.125, 0.0625, 100.0, 120, 0.0, 0), 0),       # 1.2 x 100 is not below 120
    ((0.25, 0.0625, 100.0, 121, 0.0, 0), 0),        # overflow not below 0.2
    ((0.125, 0.109375, 100.0, 121, 0.0, 0), 0),     # only 0.015625 above the minimum
    # rule 2: overflow up >= 0.05 and HPWL up >= 25 % since the last progress row
    ((0.3125, 1e30, 1e30, 1250, 0.25, 1000), 2),
    ((0.3125, 1e30, 1e30, 1249, 0.25, 1000), 0),    # HPWL up 24.9 %
    ((0.28125, 1e30, 1e30, 1250, 0.25, 1000), 0),   # overflow up 0.03125
    ((0.3125, 1e30, 1e30, 1250, 0.0, 1000), 0),     # no overflow reported yet
    ((0.3125, 1e30, 1e30, 1250, 0.25, 0), 0),       # no HPWL reported yet
    # both
    ((0.125, 0.0625, 100.0, 1250, 0.0625, 1000), 3),
])
def test_divergence_rules(args, bits):
    assert O.check_divergence(*args) == bits


# ---- the momentum (nesterov.rs:4484-4486) -------------------------------------------------------

def test_momentum_is_nesterovs_sequence():
    a, c = O.momentum(1.0)
    assert a == f32((1 + math.sqrt(5)) / 2) and c == 0.0
    a2, c2 = O.momentum(a)
    assert a2 == pytest.approx(2.1935271, abs=1e-6) and c2 == pytest.approx(0.2817535, abs=1e-6)
    # a_k^2 - a_k = a_{k-1}^2, and the coefficient (a_{k-1} - 1) / a_k rises towards 1
    prev, last_c = 1.0, -1.0
    for _ in range(200):
        a, c = O.momentum(prev)
        assert a * a - a == pytest.approx(prev * prev, rel=1e-5)
        assert last_c < c < 1.0
        prev, last_c = a, c
    assert prev == pytest.approx(101.0, rel=0.02)  # a_k ~ (k + 2) / 2


# ---- a whole run of the control functions on a convex quadratic (SURVEY: "convergence on a convex
# quadratic") ----------------------------------------------------------------------------------

def quadratic_run(curv, target, x0, npins, locked, iters, sign=1.0):
    """The reference's loop structure (nesterov.rs:4484-4486, 2776-2808, 2837-2859) driven by the
    oracle's functions on f = sum_i curv_i / 2 |x_i - target_i|^2, whose negative gradient plays the
    WA part and whose density part is zero."""
    n = len(x0)
    dd = np.full((n, 2), 10)
    dxy = np.ones((n, 2), np.int32)
    zero = np.zeros((n, 2), np.float32)

    def grad(x):
        s, _, _ = O.precondition((-sign * curv[:, None] * (x - target)).astype(np.float32), zero, npins, dxy, 0.0)
        return s

    cur_slp = x0.astype(np.float32)
    cur = cur_slp.copy()
    cur_sum = grad(cur_slp)
    # np_init: a step back along the gradient, the first step from the two points
    prev_slp, _ = O.nesterov_points(cur_slp, cur_sum, cur, locked, dd, BOX, -100.0, 0.0)
    step, _ = O.init_step(prev_slp, cur_slp, grad(prev_slp), cur_sum)
    a = 1.0
    trials = 0
    for _ in range(iters):
        a, coeff = O.momentum(a)
        for _ in range(10):
            trials += 1
            nx, ns = O.nesterov_points(cur_slp, cur_sum, cur, locked, dd, BOX, step, coeff)
            nsum = grad(ns)
            ok, step, _, dv = O.step_length(cur_slp, ns, cur_sum, nsum, step)
            if dv:
                return None, trials
            if not ok:  # a step that did not shrink by 5 % ends the back-tracking (2804-2806)
                break
        cur_slp, cur_sum, cur = ns, nsum, nx
    return cur, trials


def test_the_control_functions_converge_on_a_convex_quadratic():
    rng = np.random.default_rng(7)
    n = 400
    curv = rng.uniform(0.05, 1.0, n)            # condition number 20
    target = rng.uniform(100, 900, (n, 2))
    x0 = rng.uniform(100, 900, (n, 2))
    npins = rng.integers(1, 5, n)                # preconditioners 1..4
    locked = np.zeros(n, np.uint8)
    locked[::37] = 1                             # some cells hold still
    x, trials = quadratic_run(curv, target, x0, npins, locked, 300)
    free = locked == 0
    err0 = np.linalg.norm(x0[free] - target[free])
    assert np.linalg.norm(x[free] - target[free]) < 5e-5 * err0
    assert np.array_equal(x[~free], x0[~free].astype(np.float32))
    assert trials < 2 * 300


def test_the_quadratic_run_fails_with_the_gradient_sign_flipped():
    """The same run with the gradient's sign planted wrong (ascent) must not converge: the test
    above can tell a sign mistake in the update."""
    rng = np.random.default_rng(7)
    n = 400
    curv = rng.uniform(0.05, 1.0, n)
    target = rng.uniform(100, 900, (n, 2))
    x0 = rng.uniform(100, 900, (n, 2))
    x, _ = quadratic_run(curv, target, x0, rng.integers(1, 5, n), np.zeros(n, np.uint8), 150, sign=-1.0)
    assert x is None or np.linalg.norm(x - target) > 0.1 * np.linalg.norm(x0 - target)
