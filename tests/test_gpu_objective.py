"""The differentiable objective on the GPU (gpl_objective_forward / gpl_objective_backward)
against the oracle: the value against the FP64 objective, the adjoint against FP64 central
differences (the wirelength part with lambda = 0, the density part with a lambda that makes it
dominate), on the small workload and on 2,000 seeded cells of the 1M-cell workload (h = 1e-3
bin).

Tolerances: the GPU keeps the reference's float state (float pin positions and WA sums, the
quotients in double), so a pin's WA term carries the float error of x*S0 - S1 at coordinates of
~1e6 DBU: about 1e-7 * c * x relative, 1e-4..1e-3 of a term at these coefficients; the density
adjoint is FP64 arithmetic on the FP32 potential (1e-6 relative to the FP64 solve).  Cells with a
box edge within h of a bin edge (a kink of the density term) are left out, as the oracle flags
them."""
import numpy as np
import pytest

import gpl_amd as G
import oracle as O
from tests.test_oracle_objective import spread_points

pytestmark = pytest.mark.gpu


def _check(design, xy, n_sample, seed, wl_tol, den_tol):
    p = design.problem()
    R = p.regions[0]
    n = R.cell_end - R.cell_begin
    wl, h = 4.0 / R.size_x, 1e-3 * R.size_x
    rng = np.random.default_rng(seed)
    sample = np.sort(rng.choice(n, min(n_sample, n), replace=False)).astype(np.int32)
    g = G.NesterovGPU(design)
    try:
        g0 = g.objective_backward(xy, wl, 0.0).astype(np.float64)
        g1 = g.objective_backward(xy, wl, 1.0).astype(np.float64)
        den1 = g1 - g0
        # a lambda at which the density part is ~100x the wirelength part
        lam = float(100.0 * np.median(np.abs(g0[sample]) + 1e-3) / max(np.median(np.abs(den1[sample])), 1e-30))
        gl = g.objective_backward(xy, wl, lam).astype(np.float64)
    finally:
        g.close()
    fd0 = O.objective_fd(p, xy, wl, 0.0, sample, h)
    fdl = O.objective_fd(p, xy, wl, lam, sample, h)
    ok = fdl[:, 2:] == 0
    assert ok.mean() > 0.95
    # the wirelength part: every sampled cell, both axes
    e0 = np.abs(g0[sample] - fd0[:, :2])
    s0 = np.abs(fd0[:, :2]) + 1.0
    assert np.max(e0 / s0) <= wl_tol, float(np.max(e0 / s0))
    # the whole gradient with the density part dominating, off the kinks
    el = np.abs(gl[sample] - fdl[:, :2])[ok]
    sl = (np.abs(fdl[:, :2]) + lam * np.median(np.abs(den1[sample])))[ok]
    assert np.max(el / sl) <= den_tol, float(np.max(el / sl))
    return lam, float(np.max(e0 / s0)), float(np.max(el / sl))


@pytest.fixture(scope="module")
def c1():
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


def test_objective_value_matches_fp64(c1):
    p = c1.problem()
    R = p.regions[0]
    xy = spread_points(c1)
    wl, lam = 4.0 / R.size_x, 30.0
    g = G.NesterovGPU(c1)
    v, w, d = g.objective_forward(xy, wl, lam)
    g.close()
    w64, d64 = O.objective_value(p, xy, wl)
    assert abs(w - w64) <= 1e-5 * abs(w64)
    assert abs(d - d64) <= 1e-4 * abs(d64)
    assert v == pytest.approx(w + lam * d, rel=1e-12)


def test_objective_gradient_matches_central_differences(c1):
    _check(c1, spread_points(c1), 600, 4, wl_tol=2e-3, den_tol=1e-4)


def test_objective_gradient_at_1m_cells(c3_design):
    """2,000 seeded cells of the 1M-cell workload at the loop's points after 100 iterations."""
    d = c3_design
    g = G.NesterovGPU(d)
    g.init_nesterov_loop()
    g.loop_begin(0)
    g.loop_step(100)
    xy = g.get_coordinates(0, G.GPL_SLOT_CUR)
    g.close()
    lam, e_wl, e_den = _check(d, xy, 2000, 7, wl_tol=5e-3, den_tol=1e-4)
    print(f"1M cells: lambda {lam:.3g}, worst wirelength error {e_wl:.2e}, worst total error {e_den:.2e}")
