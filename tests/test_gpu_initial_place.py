"""The B2B + BiCGSTAB initial placement on the GPU (gpl_design_initial_place_gpu) against the host's
rounds: the same placement bit for bit -- the Nesterov problem built after it has the same
fingerprint (every cell box, pin position, bin array) -- and the same solver residuals, on the
2k-cell and 200k-cell workloads and on the region groups (locked cells, many regions)."""
import time

import pytest

import gpl_amd as G

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", ["c1", "c2", "c4r_small"])
def test_device_rounds_place_as_the_host_does(cfg):
    out = []
    for device in (False, True):
        d = G.Design.from_config(cfg)
        t = time.perf_counter()
        d.initial_place(device=device)
        dt = time.perf_counter() - t
        info = d.info()
        d.init_nesterov_place()
        out.append((d.problem_sha256(), info.ip_rounds, info.ip_error, info.initial_hpwl, dt))
        d.close()
    (h0, r0, e0, p0, t0), (h1, r1, e1, p1, t1) = out
    print(f"{cfg}: host {t0:.2f}s, device {t1:.2f}s, rounds {r0}, residual {e0:.3e}")
    assert (r1, e1, p1) == (r0, e0, p0)
    assert h1 == h0
