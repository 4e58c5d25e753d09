"""Host side of the product on the CPU: the C ABI library loads and exports every function the
header declares; the construction's helpers against the reference's known answers; the option
parser's rules; the synthetic design and the Nesterov placer's construction rules."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import gpl_amd as G
from gpl_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function():
    hdr = open(os.path.join(ROOT, "include", "gpl_hip.h")).read()
    declared = set(re.findall(r"^\s*[\w\s\*]+?\b(gpl_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 35
    lib = C.CDLL(F.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(F.exported_symbols())


def test_ctypes_structs_match_the_c_layout():
    L = F.lib()
    n = L.gpl_util_struct_sizes(None, 0)
    out = (C.c_int64 * n)()
    assert L.gpl_util_struct_sizes(out, n) == n == 8
    mirror = [F.Opts, F.GenConfig, F.DesignInfo, F.Region, F.Problem, F.Stats, F.IterRow, F.KernelTime]
    assert [C.sizeof(m) for m in mirror] == list(out)


def test_mt19937_reference_output(kats):
    k = kats["mt19937"]
    out = (C.c_uint32 * k["nth"])()
    F.lib().gpl_util_mt19937(k["seed"], k["nth"], out)
    assert out[k["nth"] - 1] == k["value"]


def test_uniform_int_draws_as_boost_does(kats):
    k = kats["uniform_int"]
    lo, hi, want = [], [], []
    for r, draws in k["draws"]:
        lo += [-r] * len(draws)
        hi += [r] * len(draws)
        want += draws
    n = len(want)
    out = (C.c_int32 * n)()
    F.lib().gpl_util_uniform_int(k["seed"], n, (C.c_int32 * n)(*lo), (C.c_int32 * n)(*hi), out)
    assert list(out) == want


def test_normal_draws_match_boost_bit_for_bit(kats):
    """The perturbation's normal: the reference's 16 draws of Boost 1.89 (one through a wedge) as
    bit patterns, and the engine's next word after 1000 draws (any draw taking a wrong count of
    engine words, the tail and the density test included, shifts it)."""
    import struct
    k = kats["normal_seed1_bits"]
    sigma = 0.3071963263271184 * 100.0
    n = len(k["bits"])
    out = (C.c_double * n)()
    F.lib().gpl_util_normal(k["seed"], 0.0, sigma, n, out)
    got = [struct.pack(">d", v).hex() for v in out]
    assert got == k["bits"]
    many = (C.c_double * 1000)()
    assert F.lib().gpl_util_normal(k["seed"], 0.0, sigma, 1000, many) == k["engine_word_after_1000_draws"]


def test_normal_draws_follow_the_unit_normal():
    """Moments and tail mass of 400k draws (the ziggurat's tail and wedges included): mean 0,
    variance sigma^2, P(|z| > 3.4426) (beyond the first layer's edge) as the normal's."""
    import math
    import numpy as np
    n = 400_000
    out = (C.c_double * n)()
    F.lib().gpl_util_normal(7, 2.0, 3.0, n, out)
    z = (np.frombuffer(out, np.float64) - 2.0) / 3.0
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 0.01
    tail = np.mean(np.abs(z) > 3.442619855896652)
    want = math.erfc(3.442619855896652 / math.sqrt(2.0))
    assert abs(tail - want) < 5 * math.sqrt(want / n)
    assert np.mean(z > 0) == pytest.approx(0.5, abs=0.005)


def test_wavenumbers_in_extended_precision(kats):
    L = F.lib()
    named = {"0": 0.0, "FRAC_PI_2": np.float32(math.pi / 2), "PI": np.float32(math.pi), "FRAC_PI_4": np.float32(math.pi / 4)}
    for i, n, want in kats["wavenumbers"]["x"]:
        assert L.gpl_util_wavenumber_x(i, n) == named[want]
    for i, n, bx, by, want in kats["wavenumbers"]["y"]:
        v = L.gpl_util_wavenumber_x(32, 64) if want == "x(32,64)" else named[want]
        assert L.gpl_util_wavenumber_y(i, n, bx, by) == v
    # i = n gives the extended pi rounded once to float
    assert L.gpl_util_wavenumber_x(64, 64) == np.float32(math.pi)


def test_round_down_to_power_of_two(kats):
    for x, want in kats["round_down_to_power_of_two"]["cases"]:
        assert F.lib().gpl_util_round_down_pow2(x) == want


def test_union_area(kats):
    for rects, want in kats["union_area"]["cases"]:
        flat = [v for r in rects for v in r]
        arr = (C.c_int32 * max(1, len(flat)))(*flat)
        assert F.lib().gpl_util_union_area(arr, len(rects)) == want
    # brute force on random small rectangles
    rng = np.random.default_rng(5)
    for _ in range(20):
        rs = []
        for _ in range(rng.integers(1, 6)):
            x0, y0 = rng.integers(0, 20, 2)
            rs.append([int(x0), int(y0), int(x0 + rng.integers(1, 10)), int(y0 + rng.integers(1, 10))])
        grid = np.zeros((40, 40), bool)
        for a, b, c, d in rs:
            grid[a:c, b:d] = True
        flat = [v for r in rs for v in r]
        assert F.lib().gpl_util_union_area((C.c_int32 * len(flat))(*flat), len(rs)) == int(grid.sum())


def test_keys_match_by_prefix(kats):
    k = kats["parse_key_args"]
    keys, flags = G.parse_key_args(k["args"])
    assert keys == k["keys"]
    assert flags == k["flags"]
    for bad in ("unknown", "positional"):
        arg, code, _line = k[bad]
        with pytest.raises(G.GplError) as e:
            G.parse_key_args([arg])
        assert e.value.code == code


def test_unmodelled_modes_are_refused():
    for a in (["-timing_driven"], ["-incremental"], ["-place_ios"], ["-routability_driven", "-routability_use_grt"],
              ["-routability_driven", "-routability_rc_coefficients", "1 1 0"]):
        with pytest.raises(G.GplError) as e:
            G.opts_from_args(a)
        assert e.value.code == G.GPL_ERR_REFUSED
    o, _ = G.opts_from_args(["-bin", "128", "-den", "0.8", "-overflow", "0.2", "-random_seed", "7"])
    assert (o.bin_cnt, round(o.density, 6), round(o.overflow, 6), o.seed) == (128, 0.8, 0.2, 7)
    assert o.routability == 0


def test_routability_options_as_run_reads_them():
    """run.rs:251-269: -routability_driven turns the mode on (not with -skip_io); its keys by
    prefix; the defaults of RouteVars (route_base.rs:29-45) and Opts (nesterov.rs:77-78)."""
    o, _ = G.opts_from_args(["-routability_driven"])
    assert o.routability == 1
    assert [round(v, 6) for v in (o.routability_end_overflow, o.routability_snapshot_overflow, o.routability_target_rc,
                                  o.routability_inflation_ratio_coef, o.routability_max_inflation_ratio,
                                  o.routability_max_density, o.routability_min_congestion, o.routability_max_inflation_total,
                                  o.routability_net_weight_max, o.routability_congested_nets_percentage)] == \
        [0.3, 0.6, 1.01, 2.0, 3.0, 0.99, 0.95, 0.3, 1.0, 1.0]
    assert list(o.routability_rc_k) == [1.0, 1.0, 0.0, 0.0] and o.routability_max_passes == 50
    o, _ = G.opts_from_args(["-routability_driven", "-routability_check", "0.25", "-routability_target_rc", "1.2",
                             "-routability_rc_coefficients", "1 2 3 4"])
    assert (round(o.routability_end_overflow, 6), round(o.routability_target_rc, 6)) == (0.25, 1.2)
    assert list(o.routability_rc_k) == [1.0, 2.0, 3.0, 4.0]
    o, _ = G.opts_from_args(["-routability_driven", "-skip_io"])
    assert o.routability == 0


def test_generator_is_deterministic_and_sized():
    a = G.Design.from_config("c1")
    b = G.Design.from_config("c1")
    ia, ib = a.info(), b.info()
    assert (ia.num_movable, ia.num_nets, ia.num_pins) == (ib.num_movable, ib.num_nets, ib.num_pins)
    assert ia.num_movable == 2000 and ia.num_nets == 2200
    # 2,200 nets of mean degree ~2.82 plus 96 ports
    assert 5700 < ia.num_pins < 6800
    assert (ia.core[2] - ia.core[0], ia.core[3] - ia.core[1]) == (357 * 200, 36 * 2000)
    c = G.Design.from_config("c1", seed=2)
    assert c.info().num_pins != ia.num_pins or c.info().movable_area != ia.movable_area


def test_nesterov_construction_rules(c1_design):
    a = c1_design.arrays()
    R = a["regions"][0]
    n = R["cell_end"] - R["cell_begin"]
    assert n == a["num_cells"]
    flags = a["cell_flags"]
    nf = R["num_fillers"]
    # instances first, then the fillers
    assert np.all(flags[: n - nf] & G.GPL_CELL_INSTANCE) and not np.any(flags[n - nf:] & G.GPL_CELL_INSTANCE)
    # bins: the requested count, edges rounded from the f64 size
    assert (R["cnt_x"], R["cnt_y"]) == (64, 64)
    assert R["bin_x"][0] == R["lx"] and R["bin_x"][-1] == R["ux"]
    # f64::round: halves away from zero (size_x = 1115.625 makes real ties)
    assert np.all(R["bin_x"] == R["lx"] + np.floor(np.arange(65) * R["size_x"] + 0.5).astype(np.int64))
    # density sizes: a side under sqrt(2) bins is spread to sqrt(2) bins, scale = size / threshold
    box, dbox, sc = a["cell_box"], a["cell_dbox"], a["density_scale"]
    thr_x, thr_y = math.sqrt(2) * R["size_x"], math.sqrt(2) * R["size_y"]
    dx, dy = box[:, 2] - box[:, 0], box[:, 3] - box[:, 1]
    ddx, ddy = dbox[:, 2] - dbox[:, 0], dbox[:, 3] - dbox[:, 1]
    sx = np.where(dx < thr_x, dx / np.float32(thr_x), 1.0)
    sy = np.where(dy < thr_y, dy / np.float32(thr_y), 1.0)
    assert np.allclose(sc, (sx * sy).astype(np.float32), rtol=2e-7)
    wide = dx >= thr_x
    assert np.all(np.abs(ddx[wide] - dx[wide]) <= 1)
    assert np.all(np.abs(ddx[~wide] - int(np.float32(thr_x))) <= 1)
    # fillers all of the filler size, inside the region
    fb = box[n - nf:]
    assert np.all(fb[:, 2] - fb[:, 0] == 2 * (R["filler_dx"] // 2))
    # the filler area fills the target share: total_filler ~ count * filler area
    assert nf == (R["white_space_area"] * 0 + nf)
    # pins: each cell's pins point back at it; nets cover every pin once
    cps, pins = a["cell_pin_start"], a["cell_pins"]
    for g in range(0, n, 97):
        for q in pins[cps[g]:cps[g + 1]]:
            assert a["pin_cell"][q] == g
    assert sorted(a["net_pins"].tolist()) == list(range(a["num_pins"]))
    # wavenumbers as the extended-precision helpers give them
    L = F.lib()
    assert R["wx"][5] == L.gpl_util_wavenumber_x(5, 64)
    assert R["wy"][5] == L.gpl_util_wavenumber_y(5, 64, np.float32(R["size_x"]), np.float32(R["size_y"]))


def test_target_density_falls_back_to_uniform_when_too_low():
    # 70 % utilisation at target 0.70 after the pin-density adjustment leaves no filler room:
    # GPL-0302 raises the target to the uniform density (nesterov.rs:1719-1728)
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    R = d.arrays()["regions"][0]
    assert R["target_density"] > 0.70
    assert abs(R["target_density"] * 100 - round(R["target_density"] * 100)) < 1e-4
