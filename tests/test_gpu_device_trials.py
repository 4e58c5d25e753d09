"""Device-driven trials (launch mode 4: the trial graph ends in a control kernel that takes the
loop's decisions on the GPU, the host stepping in only on the events it handles) against the
host-driven loop (mode 0) and the eager one (mode 2): the same iterations, history rows, log,
statistics and placed locations, bit for bit -- the control kernel evaluates the host's float and
double expressions in the same order, and every event the host handles (a zero or non-finite sum,
a divergence, a new min-HPWL snapshot, a convergence, a re-sort, the caller's iteration limit)
halts the device before any of its effects."""
import numpy as np
import pytest

import gpl_amd as G

pytestmark = pytest.mark.gpu

STAT_KEYS = ("iter", "converged", "status", "total_trials", "hpwl", "overflow", "overflow_scaled", "density_penalty",
             "step_length", "wl_coef", "base_wl_coef", "wl_grad_sum", "density_grad_sum", "num_diverge_reverts")


def _stats(g):
    s = g.stats()
    # (a NaN sum, as an error run leaves it, compares as the string "nan")
    return {k: (v if v == v else "nan") for k, v in ((k, getattr(s, k)) for k in STAT_KEYS)}


def run(design, mode, injections=(), slices=None):
    g = G.NesterovGPU(design)
    try:
        g.set_launch_mode(mode)
        g.init_nesterov_loop()
        for kind, at in injections:
            g.debug_inject(kind, at)
        if slices:
            g.loop_begin(0)
            done = False
            while not done:
                _, done = g.loop_step(slices)
            rc, it = g.loop_end(check=False)
        else:
            rc, it = g.run_nesterov_loop(check=False)
        return dict(rc=rc, it=it, hist=g.history(), log=g.log(), stats=_stats(g), loc=g.placed_locations(),
                    times=g.kernel_times())
    finally:
        g.close()


def same(a, b):
    assert a["rc"] == b["rc"] and a["it"] == b["it"], (a["rc"], a["it"], b["rc"], b["it"])
    assert a["hist"] == b["hist"]
    assert a["stats"] == b["stats"]
    assert a["log"] == b["log"]
    assert np.array_equal(a["loc"], b["loc"])


@pytest.fixture(scope="module")
def c1():
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    yield d
    d.close()


@pytest.mark.parametrize("cfg", ["c1", "c2", "c4r_small"])
def test_device_trials_give_the_host_loops_placement(cfg):
    """Whole runs (one region; nine regions with their own grids and convergence points): mode 4
    against mode 0, and the device halted only on rare events (far fewer halts than iterations)."""
    d = G.Design.from_config(cfg)
    d.initial_place()
    d.init_nesterov_place()
    try:
        host = run(d, 0)
        dev = run(d, 4)
    finally:
        d.close()
    same(host, dev)
    halts, _ = dev["times"]["device_halts"]
    print(f"{cfg}: {dev['it']} iterations, {halts} halts, {dev['times']['device_launched'][0]} trials launched")
    assert halts < dev["it"] / 4, (halts, dev["it"])


def test_device_trials_in_slices(c1):
    """gpl_loop_step in slices of 7 iterations (the device stops at each slice's end and the next
    call takes it on): the same run as one call and as the host loop in slices."""
    whole = run(c1, 4)
    host = run(c1, 0, slices=7)
    dev = run(c1, 4, slices=7)
    same(host, dev)
    same(whole, dev)


@pytest.mark.parametrize("inj", [[(5, 100)], [(6, 400), (6, 420)], [(6, 400), (6, 401)], [(6, 300)], [(3, 200)], [(1, 5)],
                                 [(2, 5)]],
                         ids=["zero_wl", "diverge_resume", "diverge_stop", "gpl_0307", "resort_every_iter", "nan_sum",
                              "nan_step"])
def test_forced_events_on_the_device(c1, inj):
    """Each host-handled event forced from the host (the WA sum read as zero, divergences with
    resume, stop and GPL-0307, a re-sort at every iteration from 200, a NaN gradient sum GPL-0306,
    a NaN step GPL-0305): the device halts where the host loop would act, and the run is the host
    loop's."""
    host = run(c1, 0, inj)
    dev = run(c1, 4, inj)
    same(host, dev)


@pytest.mark.parametrize("mode", [5, 6])
def test_eager_device_trials_agree(c1, mode):
    """Device-driven trials launched kernel by kernel (mode 5 on the side streams, 6 on one
    stream) against host-driven launches one by one on one stream (mode 2)."""
    same(run(c1, 2), run(c1, mode))
