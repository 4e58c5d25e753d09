"""The job front end through the GPU loop (main.rs:355-371): the report of a placed run, the placed
locations written into the design, and the DEF's components at exactly those locations."""
import numpy as np
import pytest

import gpl_amd as G
from gpl_amd.job import run_job
from tests.test_job import parse_def

pytestmark = pytest.mark.gpu


def test_placed_job_report_and_def(tmp_path):
    path = str(tmp_path / "c1_placed.def")
    rep = run_job({"steps": [{"cmd": "read_synthetic", "args": ["c1"]}, {"cmd": "global_placement", "args": ["-density", "0.7"]}],
                   "write_def": path})
    assert rep["status"] == "placed", rep.get("reason")
    assert rep["nesterov_iterations"] > 100
    assert any("[INFO GPL-1001] Global placement finished at iteration" in l for l in rep["lines"])
    assert rep["lines"][-1].startswith("[INFO GPL-1018] Final HPWL (um): ")
    # the same run through the API: the DEF's components sit at its placed locations
    d = G.Design.from_config("c1")
    d.initial_place()
    d.init_nesterov_place()
    g = G.NesterovGPU(d)
    g.do_nesterov_place()
    loc = g.placed_locations()
    g.close()
    d.close()
    dd = parse_def(path)
    pos = {int(n[1:]): (int(x), int(y)) for n, _, _, x, y in dd["comps"]}
    assert len(pos) == len(loc) == 2000
    # placer instance k is design instance k for this design (no fixed instances, no dummies)
    assert all(pos[int(k)] == (int(x), int(y)) for k, x, y in loc)
