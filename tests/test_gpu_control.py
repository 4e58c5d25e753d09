"""The loop's control on the GPU against the oracle, through the C ABI: the Nesterov step from a
real preconditioned gradient, the preconditioned gradient and its FP64 sums, the zero-WA-sum
recursion (nesterov.rs:2512-2528) and the min-HPWL snapshot revert with resume, stop and the
GPL-0307 error (nesterov.rs:3009-3051), each forced from the host (gpl_debug_inject kinds 5 and 6)
on both sides; and the objective op pair against the loop's own analytic gradient.

Bars: the step and the preconditioned sum bit-exact (the same float expressions); the FP64 sums
within 1e-12 of numpy's; loops with a forced case reach the same outcome as the oracle's under
the same forcing (status, reverts, converged), the iteration within 10 (the GPU's FP64 reductions
and FFT rounding move the trajectory by an iteration or two) and the HPWL within 0.5 %."""
import numpy as np
import pytest

import gpl_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


def centres(a):
    db = a["cell_dbox"]
    return np.stack([(db[:, 0] + db[:, 2]) // 2, (db[:, 1] + db[:, 3]) // 2], 1).astype(np.float32)


def loop_points(design, iters):
    """Region 0's current points after `iters` iterations of the GPU loop (spread-out cells)."""
    g = G.NesterovGPU(design)
    try:
        g.init_nesterov_loop()
        g.loop_begin(0)
        g.loop_step(iters)
        return g.get_coordinates(0, G.GPL_SLOT_CUR)
    finally:
        g.close()


def test_preconditioned_gradient_and_step_match_the_oracle(c1):
    """updateGradients into the current slot at a loop point (every cell's WA and density parts
    bit-exact with the oracle), its preconditioned sum bit-exact with the oracle's arithmetic on
    the same parts, the FP64 |wl| and |den| sums; then nesterovUpdateCoordinates from that sum,
    bit-exact with the oracle's update (some cells pushed against the grid's edges)."""
    a = c1.arrays()
    R = a["regions"][0]
    xy = np.round(loop_points(c1, 60)).astype(np.float32)
    g = G.NesterovGPU(c1)
    o = O.Oracle(c1.problem())
    try:
        g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
        g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
        o.scatter(0, xy)
        coef = np.float32(1.0 / R["size_x"])
        g.wa_forward(coef, coef)
        _, _, os64, _ = o.wa_forward(coef, coef)
        _, fx, fy = g.density_solve(0)
        pen = np.float32(3e-4)
        wl, den, sm, ab = g.update_gradients(0, G.GPL_SLOT_CUR_SLP, pen)
        owl, _ = o.wa_backward(0, os64)
        oden, _ = o.gather(0, fx, fy)
        assert np.array_equal(wl, owl) and np.array_equal(den, oden)
        box = a["cell_box"]
        npins = np.diff(a["cell_pin_start"])
        dxy = np.stack([box[:, 2] - box[:, 0], box[:, 3] - box[:, 1]], 1)
        osm, _, _ = O.precondition(owl, oden, npins, dxy, pen)
        assert np.array_equal(sm, osm)
        assert np.abs(sm).max() > 0
        w64, d64 = np.abs(wl.astype(np.float64)).sum(), np.abs(den.astype(np.float64)).sum()
        assert abs(ab[0] - w64) <= 1e-12 * w64 and abs(ab[1] - d64) <= 1e-12 * d64
        # the step: next = slp + s sum, next_slp = next + coeff (next - cur), kept inside
        rng = np.random.default_rng(5)
        cur = (xy + rng.normal(0, 0.3 * R["size_x"], xy.shape)).astype(np.float32)
        g.set_coordinates(0, G.GPL_SLOT_CUR, cur)
        db = a["cell_dbox"]
        dd = np.stack([db[:, 2] - db[:, 0], db[:, 3] - db[:, 1]], 1)
        locked = (a["cell_flags"] & G.GPL_CELL_LOCKED) != 0
        gbox = (R["lx"], R["ly"], R["ux"], R["uy"])
        for s in (np.float32(2.0 * R["size_x"] / np.abs(sm).max()), np.float32(60.0 * R["size_x"] / np.abs(sm).max())):
            coeff = np.float32(0.37)
            nx, ns = g.nesterov_step(0, s, coeff)
            ox, osl = O.nesterov_points(xy, sm, cur, locked, dd, gbox, s, coeff)
            assert np.array_equal(nx, ox) and np.array_equal(ns, osl)
        # the large step pushed cells against the grid's edges: the clamping ran
        at_edge = ((ns[:, 0] - dd[:, 0] / 2 <= R["lx"] + 1e-3) | (ns[:, 0] + dd[:, 0] / 2 >= R["ux"] - 1e-3) |
                   (ns[:, 1] - dd[:, 1] / 2 <= R["ly"] + 1e-3) | (ns[:, 1] + dd[:, 1] / 2 >= R["uy"] - 1e-3))
        assert at_edge.sum() > 0
    finally:
        g.close()
        o.close()


def run_both(design, injections):
    """The GPU loop and the oracle's, each with the same forced cases (kind, iteration)."""
    g = G.NesterovGPU(design)
    g.init_nesterov_loop()
    for kind, at in injections:
        g.debug_inject(kind, at)
    rc, it = g.run_nesterov_loop(check=False)
    res = dict(rc=rc, it=it, stats=g.stats(), log=g.log(), hist=g.history())
    g.close()
    o = O.Oracle(design.problem())
    assert o.init() == 0
    for kind, at in injections:
        o.set_inject_control(kind, at)
    orc, oit = o.run()
    ores = dict(rc=orc, it=oit, stats=o.stats(), hist=o.history())
    o.close()
    return res, ores


def test_zero_wirelength_sum_halves_the_coefficients_and_retries(c1):
    """A WA gradient sum of zero in iteration 100 (forced): the coefficients and the base are halved,
    the forces recomputed and the gradient evaluated again (nesterov.rs:2516-2522); the run goes on
    with half the base coefficient and converges as the oracle's does under the same forcing."""
    ref = G.NesterovGPU(c1)
    ref.do_nesterov_place()
    base0 = ref.stats().base_wl_coef
    ref.close()
    r, o = run_both(c1, [(5, 100)])
    s, so = r["stats"], o["stats"]
    assert r["rc"] == 0 and o["rc"] == 0 and s.converged and so.converged
    assert s.base_wl_coef == np.float32(base0 * 0.5) and so.base_wl_coef == s.base_wl_coef
    assert abs(r["it"] - o["it"]) <= 10, (r["it"], o["it"])
    assert abs(s.hpwl - so.hpwl) <= 0.005 * so.hpwl, (s.hpwl, so.hpwl)
    # the row of iteration 100 carries the halved coefficient (the WA coefficient is base x factor)
    row = {h["iter"]: h for h in r["hist"]}
    assert row[101]["wl_coef"] < 0.6 * row[99]["wl_coef"]


def test_divergence_reverts_to_the_min_hpwl_snapshot_and_resumes(c1):
    """Divergences forced in iterations 400 and 420, both after the overflow fell below 0.25 (so a
    min-HPWL snapshot exists): each reverts to the snapshot and resumes with the momentum reset
    (GPL-0998, GPL-0999), the second after a new snapshot; the run converges as the oracle's."""
    r, o = run_both(c1, [(6, 400), (6, 420)])
    s, so = r["stats"], o["stats"]
    assert r["rc"] == 0 and o["rc"] == 0
    assert s.num_diverge_reverts == so.num_diverge_reverts == 2
    assert s.converged and so.converged
    assert r["log"].count("[WARNING GPL-0998]") == 2 and "[WARNING GPL-0999] Revert to iter:" in r["log"]
    assert abs(r["it"] - o["it"]) <= 10, (r["it"], o["it"])
    assert abs(s.hpwl - so.hpwl) <= 0.005 * so.hpwl


def test_divergence_again_without_a_new_snapshot_reverts_and_stops(c1):
    """Divergences in iterations 400 and 401: the first reverts and resumes; at the second no new
    snapshot was taken, so the loop reverts once more and stops (GPL-0997), not converged, at the
    same iteration as the oracle, with the snapshot's HPWL."""
    r, o = run_both(c1, [(6, 400), (6, 401)])
    s, so = r["stats"], o["stats"]
    assert r["rc"] == 0 and o["rc"] == 0
    assert s.num_diverge_reverts == so.num_diverge_reverts == 1
    assert not s.converged and not so.converged
    assert r["it"] == o["it"] == 401
    assert "[WARNING GPL-0997] Divergence detected again" in r["log"]
    assert abs(s.hpwl - so.hpwl) <= 0.005 * so.hpwl, (s.hpwl, so.hpwl)


def test_divergence_without_a_snapshot_is_gpl_0307(c1):
    """A divergence in iteration 300, while the overflow is still above 0.25 (no min-HPWL snapshot
    yet): the loop stops with GPL-0307 at that iteration, as the oracle's does."""
    r, o = run_both(c1, [(6, 300)])
    assert r["rc"] == o["rc"] == 307
    assert r["it"] == o["it"] == 300
    assert "[ERROR GPL-0307]" in r["log"]


def test_objective_backward_against_the_loops_analytic_gradient(c1):
    """gpl_objective_backward against the loop's own gradient at the same integer points (the
    density boxes then sit where the loop's truncation puts them): the wirelength part is minus
    the loop's WA gradient (nesterov.rs:2392-2430; the objective's exact exponential against the
    loop's fast_exp: bar 2 % normwise), and the density part is minus the loop's density gradient
    (2434-2449) times 1/(size_x^2 size_y t) in x and 1/(size_y^3 t) in y -- the field is the
    potential's derivative per bin, the adjoint per DBU, and the density is normalised by the bin
    area and the target.  The two are different discretisations of one derivative: the loop sums
    the spectral field at the centres of the bins a box covers, the adjoint takes the potential's
    jump across the box's edge bins, so per cell they differ by the potential's content near the
    grid's Nyquist frequency (large after 150 iterations of a 64 x 64 grid).  Bars: the factor
    fitted over all cells within 1 % of the analytic one per axis (measured 0.4 %), and the
    residual after scaling by the analytic factor below 50 % normwise (measured 44 %)."""
    a = c1.arrays()
    R = a["regions"][0]
    xy = np.round(loop_points(c1, 150)).astype(np.float32)
    coef = np.float32(1.0 / R["size_x"])
    g = G.NesterovGPU(c1)
    try:
        ob0 = g.objective_backward(xy, coef, 0.0).astype(np.float64)
        ob1 = g.objective_backward(xy, coef, 1.0).astype(np.float64)
        g.set_coordinates(0, G.GPL_SLOT_CUR_SLP, xy)
        g.density_scatter(0, G.GPL_SLOT_CUR_SLP)
        g.wa_forward(coef, coef)
        g.density_solve(0)
        wl, den, _, _ = g.update_gradients(0, G.GPL_SLOT_CUR_SLP, np.float32(0.0))
    finally:
        g.close()
    wl, den = wl.astype(np.float64), den.astype(np.float64)
    e_wl = np.linalg.norm(ob0 + wl) / np.linalg.norm(wl)
    dobj = ob1 - ob0
    t = float(R["target_density"])
    bx, by = float(R["size_x"]), float(R["size_y"])
    kx, ky = 1.0 / (bx * bx * by * t), 1.0 / (by ** 3 * t)
    fit = [float(-(dobj[:, i] @ den[:, i]) / (den[:, i] @ den[:, i])) for i in range(2)]
    res = np.linalg.norm(dobj + den * np.array([kx, ky])) / np.linalg.norm(dobj)
    print(f"wirelength part {e_wl:.2e}; density factor fitted {fit[0]:.4e} {fit[1]:.4e} vs {kx:.4e} {ky:.4e}, residual {res:.3f}")
    assert e_wl < 0.02
    assert abs(fit[0] / kx - 1) < 0.01 and abs(fit[1] / ky - 1) < 0.01
    assert res < 0.5
