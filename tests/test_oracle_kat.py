"""The oracle's scalar rules against the reference's own known-answer tests (tests/golden)."""
import math

import numpy as np

import oracle as O


def test_fast_exp_is_ten_squarings(kats):
    L = O.lib()
    assert L.orc_fast_exp(0.0) == 1.0
    e = np.float32(1.0) - np.float32(1.0) / np.float32(1024.0)
    for _ in range(10):
        e = np.float32(e * e)
    assert L.orc_fast_exp(-1.0) == e
    assert L.orc_fast_exp(-1.0) != np.float32(math.exp(-1.0))


def test_distance_is_rms_per_coordinate(kats):
    k = kats["distance"]
    a = np.array(k["a"], np.float32).reshape(-1)
    b = np.array(k["b"], np.float32).reshape(-1)
    got = O.lib().orc_distance(O.f32(a), O.f32(b), 2)
    assert got == np.float32(math.sqrt(k["expected_f32_of_sqrt"]))


def test_derived_reference_hpwl(kats):
    L = O.lib()
    for user, hpwl, want in kats["derived_reference_hpwl"]["cases"]:
        got = L.orc_derived_reference_hpwl(user, hpwl)
        if want == "0.01f32 * 1e11f32":
            want = np.float32(np.float32(0.01) * np.float32(1e11))
            assert got != np.float32(1e11)
        assert got == np.float32(want)


def test_diverge_fallback_table(kats):
    for saved, reverts, disable, want in kats["diverge_fallback"]["cases"]:
        assert O.lib().orc_diverge_fallback(saved, reverts, disable) == want


def test_max_iter_snapshot_rule(kats):
    k = kats["max_iter_takes_min_hpwl_snapshot"]
    for it, dis, saved, rev, shpwl, sov, ov, hpwl, want in k["cases"]:
        assert O.lib().orc_max_iter_takes_snapshot(it, k["max_iter"], dis, saved, rev, shpwl, sov, ov, hpwl) == want


def test_wirelength_coef_schedule_closed_form():
    # 1 / 10^((ovf - 0.1) * 20/9 - 1): 10 at 0.1, 1 at 0.55, 0.1 at 1.0; clamps outside
    L = O.lib()
    assert L.orc_wirelength_coef_factor(0.05) == 10.0
    assert L.orc_wirelength_coef_factor(1.5) == np.float32(0.1)
    assert abs(L.orc_wirelength_coef_factor(0.55) - 1.0) < 1e-6
    assert abs(L.orc_wirelength_coef_factor(1.0) - 0.1) < 1e-7
    assert abs(L.orc_wirelength_coef_factor(0.1) - 10.0) < 1e-5
    # monotone decreasing in between
    v = [L.orc_wirelength_coef_factor(x) for x in np.linspace(0.1, 1.0, 50)]
    assert all(a > b for a, b in zip(v, v[1:]))


def test_phi_coef_rule():
    L = O.lib()
    assert L.orc_phi_coef(1.05, 0.95, -0.5) == np.float32(1.05)
    assert abs(L.orc_phi_coef(1.05, 0.95, 1.0) - 1.0) < 1e-6  # m * m^-1
    assert L.orc_phi_coef(1.05, 0.95, 100.0) == np.float32(0.95)  # floored at min
