"""Independent placements in lockstep (gpl_place_batch): each iteration's first trial of every
context as one launch per kernel, whose blocks take each context's arguments in turn, the rest of
each iteration by its own context.  Every context's iterations, history rows, log, statistics and
placed locations are those of its own do_nesterov_place, bit for bit: seeds of one design (one
group), designs of different shapes in one batch (one group per shape), and a batch whose
contexts cannot be grouped (profiling on: they run one after another)."""
import numpy as np
import pytest

import gpl_amd as G

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iter", "converged", "status", "total_trials", "hpwl", "overflow", "overflow_scaled", "density_penalty",
             "step_length", "wl_coef", "base_wl_coef", "wl_grad_sum", "density_grad_sum", "num_diverge_reverts")


def _result(g, it):
    s = g.stats()
    # (a NaN sum, as a diverged run leaves it, compares as the string "nan")
    stats = {k: (v if v == v else "nan") for k, v in ((k, getattr(s, k)) for k in STAT_KEYS)}
    return dict(it=it, hist=g.history(), log=g.log(), stats=stats, loc=g.placed_locations())


def _contexts(cfg, seeds):
    d = G.Design.from_config(cfg)
    d.initial_place()
    out = []
    for seed in seeds:
        d.init_nesterov_place(G.default_opts(**dict(G.CONFIGS[cfg]["opts"], seed=seed)))
        out.append(G.NesterovGPU(d))
    return d, out


def _same(a, b):
    assert a["it"] == b["it"]
    assert a["stats"] == b["stats"]
    assert a["hist"] == b["hist"]
    assert a["log"] == b["log"]
    assert np.array_equal(a["loc"], b["loc"])


def _check(specs, profiling=False, inject=None):
    """specs: (config, seeds) pairs; the contexts placed one by one, then as one batch.  inject:
    {context index: [(debug kind, after iteration)]} forced on that context in both."""
    inject = inject or {}
    designs, alone, batch = [], [], []
    try:
        k = 0
        for cfg, seeds in specs:
            d, cs = _contexts(cfg, seeds)
            designs.append(d)
            for g in cs:
                for kind, at in inject.get(k, []):
                    g.debug_inject(kind, at)
                it = g.run_nesterov_loop(check=False)[1] if g.init_nesterov_loop() == 0 else None
                alone.append(_result(g, it))
                g.close()
                k += 1
        for (cfg, seeds), d in zip(specs, designs):
            for seed in seeds:
                d.init_nesterov_place(G.default_opts(**dict(G.CONFIGS[cfg]["opts"], seed=seed)))
                g = G.NesterovGPU(d)
                if profiling:
                    g.set_profiling(True)
                for kind, at in inject.get(len(batch), []):
                    g.debug_inject(kind, at)
                batch.append(g)
        its = G.place_batched(batch)
        assert len(its) == len(alone)
        for a, g, it in zip(alone, batch, its):
            _same(a, _result(g, it))
    finally:
        for g in batch:
            g.close()
        for d in designs:
            d.close()


def test_seeds_of_one_design_in_one_batch():
    _check([("c1", (1, 2, 3, 4, 5, 6))])


def test_nine_seeds_in_one_batch():
    _check([("c1", tuple(range(1, 10)))])


def test_seeds_of_the_50k_cell_design_in_one_batch():
    _check([("c4s", (1, 2, 3, 4))])


def test_designs_of_two_shapes_in_one_batch():
    _check([("c1", (1, 2)), ("c2", (1,)), ("c1", (3,))])


def test_a_batch_that_cannot_be_grouped_runs_one_by_one():
    _check([("c1", (1, 2))], profiling=True)


def test_contexts_that_diverge_or_resort_every_iteration_keep_their_own_outcome():
    """In one batch: a context whose gradient sum turns NaN at iteration 5 (GPL-0306: it leaves the
    batch there), one whose step length does at 30 (GPL-0305), one re-sorting its cells every
    iteration (its arrays swapped between the grouped launches), and two plain ones."""
    _check([("c1", (1, 2, 3, 4, 5))], inject={1: [(1, 5)], 2: [(3, 0)], 3: [(2, 30)]})
