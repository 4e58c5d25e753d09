"""The oracle on several host threads (the all-cores CPU baseline): the same bits as on one -- the
2-D transforms row by row, the bins' integer areas by atomic adds, the WA nets, the gradients and
the centres cell by cell, while every float sum whose order the reference fixes stays sequential.
A whole run of the 2k-cell placement (history, statistics, placed positions) and a 256 x 256 solve
are compared."""
import numpy as np
import pytest

import gpl_amd as G
import oracle as O


def _run(design, threads):
    O.set_threads(threads)
    try:
        o = O.Oracle(design.problem())
        assert o.init() == 0
        rc, it = o.run()
        s = o.stats()
        out = (rc, it, o.history(), s.hpwl, s.overflow, s.total_trials)
        o.close()
        return out
    finally:
        O.set_threads(1)


@pytest.mark.parametrize("threads", [2, 5])
def test_threads_give_the_single_threads_run(threads):
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    try:
        a = _run(d, 1)
        b = _run(d, threads)
    finally:
        d.close()
    assert a[:2] == b[:2]
    assert a[2] == b[2]
    assert a[3:] == b[3:]


def test_threads_give_the_same_solve():
    rng = np.random.default_rng(3)
    n = 256
    dens = rng.random((n, n)).astype(np.float32)
    wx = (np.pi * np.arange(n) / n).astype(np.float32)
    wy = wx.copy()
    outs = []
    for t in (1, 4):
        O.set_threads(t)
        try:
            outs.append(O.solve_f32(n, n, wx, wy, dens.reshape(-1)))
        finally:
            O.set_threads(1)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
