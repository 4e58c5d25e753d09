"""The autotuner (gpl_set_tuning / gpl_get_tuning): every kernel variant it chooses among -- the
scatter's chunk of 256 or 512 cells and its 32- or 64-bit LDS sums, the gradient's LDS staging on or
off -- gives the same placement bit for bit (rows, iterations, placed locations); the tuner's own
pick is measured once per shape and then taken from its cache."""
import itertools

import numpy as np
import pytest

import gpl_amd as G

pytestmark = pytest.mark.gpu


def _run(d, forced=None):
    g = G.NesterovGPU(d)
    try:
        if forced is not None:
            for b in range(d.problem().num_regions):
                g.set_tuning(b, *forced)
        it = g.do_nesterov_place()
        return it, g.history(), g.placed_locations(), g.tuning()
    finally:
        g.close()


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_every_variant_gives_the_same_placement(cfg):
    d = G.Design.from_config(cfg)
    d.initial_place()
    d.init_nesterov_place()
    try:
        it0, h0, p0, t0 = _run(d)
        assert t0[0]["source"] in ("measured", "cached")
        assert t0[0]["chunk"] in (256, 512) and t0[0]["table_bits"] in (32, 64) and t0[0]["stage"] in (0, 1)
        combos = list(itertools.product((512, 256), (32, 64), (1, 0)))
        if cfg == "c2":
            combos = [(512, 64, 1), (256, 32, 0), (256, 64, 1)]
        for forced in combos:
            it, h, p, t = _run(d, forced)
            assert (t[0]["chunk"], t[0]["stage"]) == (forced[0], forced[2]) and t[0]["source"] == "default"
            assert it == it0 and h == h0, forced
            assert np.array_equal(p, p0), forced
        # a second context of the same shape takes the cached pick at its init
        g = G.NesterovGPU(d)
        g.init_nesterov_loop()
        t2 = g.tuning()
        g.close()
        assert t2[0]["source"] == "cached"
    finally:
        d.close()
