"""The job front end (main.rs:240-400): a JSON job ``{"steps": [{"cmd", "args"}], "write_def"}`` run
step by step, and the report the reference returns: ``{"tool": "vyges-gpl", "status", "reason",
"lines", "trace", "insts", "nets", "pins", "nesterov_iterations", "bases"}`` (``"code"`` on an
error), with the placed design written as DEF when the job names a path.

The reference's first step reads an OpenDB database (``read_db``); no database library exists here,
so the step ``read_synthetic <config> [-seed N]`` builds one of the seeded synthetic designs in its
place.  ``global_placement`` takes the reference's arguments (the prefix rules and refusals of
:func:`gpl_amd.opts_from_args`).  Statuses follow main.rs's ``finish``: ``placed``; ``error`` with
``code`` ``GPL-NNNN``; ``refused`` with the reason; ``vacuous`` when no step places anything.

    python -m gpl_amd.job job.json [-o report.json]
"""
import ctypes as C
import json
import sys

import numpy as np

from . import (CONFIGS, Design, GplError, NesterovGPU, _ffi, default_opts, gen_config, opts_from_args)


def _bases(design):
    L = _ffi.lib()
    n = L.gpl_design_bases(design._h, None, 0)
    out = np.zeros(4 * max(n, 1), np.int32)
    L.gpl_design_bases(design._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
    return [dict(group=int(g), place=int(p), fixed=int(f), dummies=int(dm)) for g, p, f, dm in out[:4 * n].reshape(-1, 4)]


def _counts(design):
    i = design.info()
    return dict(insts=int(i.num_insts), nets=int(i.num_nets), pins=int(i.num_pins))


def _write_def(design, path):
    rc = _ffi.lib().gpl_design_write_def(design._h, path.encode())
    if rc != 0:
        raise OSError(f"{path}: cannot write")


def run_job(job):
    """Runs a job (a dict, or a path to its JSON) and returns the report dict."""
    if isinstance(job, str):
        with open(job) as f:
            job = json.load(f)
    lines, trace = [], []
    design, cfg_name = None, None

    def finish(stop=None, extra=None):
        status, reason = "placed", None
        if stop is not None:
            status, _, reason = stop.partition(": ")
            reason = reason or None
        v = {"tool": "vyges-gpl", "status": status, "reason": reason, "lines": list(lines), "trace": trace}
        v.update(extra or {})
        return v

    try:
        for step in job.get("steps", []):
            cmd, args = step.get("cmd"), [str(a) for a in step.get("args", [])]
            if cmd == "read_synthetic":
                if not args or args[-1] not in CONFIGS:
                    return finish(f"refused: read_synthetic needs one of {sorted(CONFIGS)}")
                seed = int(args[args.index("-seed") + 1]) if "-seed" in args else 1
                cfg_name = args[-1]
                if design is not None:
                    design.close()
                design = Design(gen_config(cfg_name, seed=seed), default_opts(**CONFIGS[cfg_name]["opts"]))
            elif cmd == "global_placement":
                if design is None:
                    return finish("refused: read a design before global_placement")
                try:
                    o, flags = opts_from_args(args, default_opts(**CONFIGS[cfg_name]["opts"]))
                except GplError as e:
                    if e.code > 0:
                        return dict(finish(f"error: STA-{e.code:04d}"), code=f"STA-{e.code:04d}")
                    return finish(f"refused: {e}")
                if not o.skip_initial_place:
                    design.initial_place()
                if "-skip_nesterov_place" in flags:
                    # initial placement is the whole command: each instance it moved is PLACED where it is
                    extra = dict(_counts(design), bases=_bases(design))
                    if job.get("write_def"):
                        _write_def(design, job["write_def"])
                    return finish(None, extra)
                design.init_nesterov_place(o)
                g = NesterovGPU(design)
                try:
                    try:
                        g.init_nesterov_loop()
                        rc, it = g.run_nesterov_loop(check=False)
                    finally:
                        lines.extend(g.log().splitlines())
                    if rc > 0:
                        return dict(finish(f"error: GPL-{rc:04d}"), code=f"GPL-{rc:04d}")
                    if rc < 0:
                        return finish(f"refused: status {rc}")
                    loc = np.ascontiguousarray(g.placed_locations(), np.int32)
                finally:
                    g.close()
                hp = C.c_int64(0)
                _ffi.lib().gpl_design_write_locations(design._h, loc.ctypes.data_as(C.POINTER(C.c_int32)), len(loc), C.byref(hp))
                info = design.info()
                lines.append(f"[INFO GPL-1018] Final HPWL (um): {hp.value / info.dbu:.2f}")  # db_router.rs:112
                extra = dict(_counts(design), nesterov_iterations=int(it), bases=_bases(design))
                if job.get("write_def"):
                    _write_def(design, job["write_def"])
                return finish(None, extra)
            else:
                return finish(f"refused: {cmd}: not a step this engine runs")
        return finish("vacuous: the job has no global_placement step")
    except GplError as e:
        if e.code > 0:
            return dict(finish(f"error: GPL-{e.code:04d}"), code=f"GPL-{e.code:04d}")
        return finish(f"refused: {e}")
    finally:
        if design is not None:
            design.close()


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    out = None
    if "-o" in argv:
        k = argv.index("-o")
        out = argv[k + 1]
        del argv[k:k + 2]
    if len(argv) != 1:
        print("usage: python -m gpl_amd.job <job.json> [-o FILE]", file=sys.stderr)
        return 2
    rep = run_job(argv[0])
    text = json.dumps(rep, indent=1)
    if out:
        with open(out, "w") as f:
            f.write(text + "\n")
    else:
        print(text)
    return 0 if rep["status"] == "placed" else 1


if __name__ == "__main__":
    sys.exit(main())
