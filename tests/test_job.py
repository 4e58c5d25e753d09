"""The job front end (main.rs:240-400) on the CPU: the report's fields and statuses, refusals, and the
DEF written after initial placement (-skip_nesterov_place: no GPU needed).  The placed run through
the Nesterov loop is checked on the GPU (test_gpu_job.py)."""
import os
import re

import gpl_amd as G
from gpl_amd.job import run_job


def parse_def(path):
    """Components (name, status, x, y), pin count and net count of a DEF the writer produced."""
    with open(path) as f:
        text = f.read()
    comps = re.findall(r"^- (i\d+) (\S+) \+ (PLACED|FIXED) \( (-?\d+) (-?\d+) \) N ;$", text, re.M)
    die = tuple(map(int, re.search(r"DIEAREA \( (-?\d+) (-?\d+) \) \( (-?\d+) (-?\d+) \) ;", text).groups()))
    n_comp = int(re.search(r"^COMPONENTS (\d+) ;", text, re.M).group(1))
    n_pins = int(re.search(r"^PINS (\d+) ;", text, re.M).group(1))
    n_nets = int(re.search(r"^NETS (\d+) ;", text, re.M).group(1))
    nets = re.findall(r"^- n\d+( \( .*?)? ;$", text, re.M)
    return dict(comps=comps, die=die, n_comp=n_comp, n_pins=n_pins, n_nets=n_nets, nets=nets, text=text)


def test_initial_placement_job_writes_the_report_and_def(tmp_path):
    path = str(tmp_path / "c1.def")
    rep = run_job({"steps": [{"cmd": "read_synthetic", "args": ["c1"]},
                             {"cmd": "global_placement", "args": ["-skip_nesterov_place"]}], "write_def": path})
    assert rep["tool"] == "vyges-gpl" and rep["status"] == "placed" and rep["reason"] is None
    assert rep["insts"] == 2000 and rep["nets"] == 2200 and rep["pins"] > 6000
    assert rep["bases"] == [dict(group=-1, place=2000, fixed=0, dummies=0)]
    d = parse_def(path)
    assert d["n_comp"] == len(d["comps"]) == 2000
    assert d["n_pins"] == 96 and d["n_nets"] == 2200 and len(d["nets"]) == 2200
    lx, ly, ux, uy = d["die"]
    assert all(lx <= int(x) <= ux and ly <= int(y) <= uy for _, _, _, x, y in d["comps"])
    assert all(st == "PLACED" for _, _, st, _, _ in d["comps"])


def test_job_statuses():
    assert run_job({"steps": []})["status"] == "vacuous"
    r = run_job({"steps": [{"cmd": "global_placement", "args": []}]})
    assert r["status"] == "refused" and "read a design" in r["reason"]
    r = run_job({"steps": [{"cmd": "read_db", "args": ["x.odb"]}]})
    assert r["status"] == "refused"
    r = run_job({"steps": [{"cmd": "read_synthetic", "args": ["c1"]}, {"cmd": "global_placement", "args": ["-bogus", "1"]}]})
    assert r["status"] == "error" and r["code"] == "STA-0562"
    r = run_job({"steps": [{"cmd": "read_synthetic", "args": ["c1"]}, {"cmd": "global_placement", "args": ["-timing_driven"]}]})
    assert r["status"] == "refused" and "not modelled" in r["reason"]
