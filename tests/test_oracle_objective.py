"""The differentiable objective's oracle side (CPU): the identities its central differences rest on.

f = sum_e w_e WA_e + lambda * 1/2 rho^T K rho.  Its density gradient is phi^T d rho/dx only if the
Poisson operator K is symmetric, and the oracle's central difference takes phi^T (rho(x+h) -
rho(x-h)) instead of re-solving; both shortcuts are checked here against their definitions."""
import numpy as np
import pytest

import oracle as O
from tests.test_oracle_dct import _wavenumbers


@pytest.mark.parametrize("nx,ny,bx,by", [(64, 32, 1000.0, 1500.0), (16, 128, 3.0, 1.0)])
def test_poisson_operator_is_symmetric(nx, ny, bx, by):
    wx, wy = _wavenumbers(nx, ny, bx, by)
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal(nx * ny), rng.standard_normal(nx * ny)
    ka = O.solve_phi_f64(nx, ny, wx, wy, a)
    kb = O.solve_phi_f64(nx, ny, wx, wy, b)
    lhs, rhs = float(a @ kb), float(b @ ka)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(a) * np.linalg.norm(kb)
    # and not trivially: K is far from a multiple of the identity
    assert abs(float(a @ ka)) < 0.9 * np.linalg.norm(a) * np.linalg.norm(ka)


def spread_points(design, seed=5, spread_bins=3.0):
    """Region 0's cells at their density centres moved by a seeded normal of `spread_bins` bins,
    kept inside the region (problem order)."""
    p = design.problem()
    R = p.regions[0]
    a = design.arrays()
    db = a["cell_dbox"][R.cell_begin:R.cell_end].astype(np.float64)
    xy = np.stack([(db[:, 0] + db[:, 2]) / 2, (db[:, 1] + db[:, 3]) / 2], 1)
    rng = np.random.default_rng(seed)
    xy += rng.normal(0.0, spread_bins, xy.shape) * np.array([R.size_x, R.size_y])
    xy[:, 0] = np.clip(xy[:, 0], R.lx, R.ux)
    xy[:, 1] = np.clip(xy[:, 1], R.ly, R.uy)
    return xy.astype(np.float32)


def test_central_difference_shortcut_equals_re_evaluation(c1_design):
    """The oracle's difference (nets on the cell re-evaluated, phi^T delta-rho) against the full
    objective evaluated from scratch at x +- h (a new density and a new FP64 solve each time)."""
    p = c1_design.problem()
    R = p.regions[0]
    xy = spread_points(c1_design)
    wl, h = 4.0 / R.size_x, 1e-3 * R.size_x
    rng = np.random.default_rng(2)
    sample = rng.choice(R.cell_end - R.cell_begin, 12, replace=False).astype(np.int32)
    checked = 0
    # lambda 0: the wirelength alone; 30: the density part a few times the wirelength's
    for lam, i, k in [(lam, i, k) for lam in (0.0, 30.0) for i, k in enumerate(sample)]:
        fd = O.objective_fd(p, xy, wl, lam, sample, h)
        for axis in range(2):
            if fd[i, 2 + axis]:
                continue
            vals = []
            for sgn in (1.0, -1.0):
                x2 = xy.astype(np.float64).copy()
                x2[k, axis] += sgn * h
                w, d = O.objective_value(p, x2.astype(np.float32), wl)
                vals.append(w + lam * d)
            # the moved coordinate is rounded to float: use the step actually taken
            step = float(np.float32(xy[k, axis] + h)) - float(np.float32(xy[k, axis] - h))
            brute = (vals[0] - vals[1]) / step
            assert abs(brute - fd[i, axis]) <= 1e-6 * max(abs(fd[i, axis]), 1.0), (lam, k, axis, brute, fd[i, axis])
            checked += 1
    assert checked >= 32
