"""Writes the end-to-end loop fixtures: the oracle's full run to overflow 0.10 on the 200k-cell
(512 x 512 bins) and 1M-cell (1024 x 1024 bins) synthetic placements, seed 1, and the other
workloads (c4r, c4s; c5 with its routability passes).

The oracle is the single-thread C restatement of the reference's loop (oracle/); a 1M-cell run
takes about eleven minutes on one core, too long for a test, so its outcome is kept here as data:
final HPWL, iteration, overflow, the first progress rows, the CPU time, and a hash of the
generated problem (the GPU test checks that it places the same problem).

    python tests/golden/make_loop_fixtures.py c2 c3
    python tests/golden/make_loop_fixtures.py --perturb 3 c4r   # adds the sensitivity runs
    python tests/golden/make_loop_fixtures.py --wide-rows c2 c3 c4r c4s c5   # adds the wide-sum first rows

--perturb N adds N more oracle runs of the same problem with every field value scaled by
(1 + 2e-7 u), u uniform in [-1, 1) (the oracle's sensitivity probe: the size by which the GPU's
transforms differ from the oracle's), recording how far the loop's outcome moves under
rounding-level perturbations.

--wide-rows adds each fixture's first rows from the oracle with its global float sums in double
(the step length's distances, |wl| and |den|, the overflow), the precision of the GPU path's FP64
reductions: the GPU's first rows equal these, while the reference's float sums move the plain
rows by up to ~1e-4 at 500k-1M cells.
"""
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402


def problem_hash(design):
    """sha256 over every array of the Nesterov problem (cells, pins, nets, regions' bins)."""
    return design.problem_sha256()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def one_perturbed(name, eps, k, seed=1):
    """One sensitivity run: the oracle's loop with noise seed k (a process of its own)."""
    import gpl_amd as G
    import oracle as O
    d = G.Design.from_config(name, seed=seed)
    d.initial_place()
    d.init_nesterov_place()
    o = O.Oracle(d.problem())
    o.set_field_noise(eps, k)
    o.init()
    rc, it = o.run()
    s = o.stats()
    out = dict(field_noise=eps, noise_seed=k, status=int(rc), converged=bool(s.converged), final_iter=int(it),
               final_hpwl=int(s.hpwl), final_overflow=float(s.overflow))
    if d.problem().opts.routability:
        _, passes = o.route_stats()
        out["route_passes"] = [{k2: (float(v) if isinstance(v, float) else int(v)) for k2, v in q.items()} for q in passes]
    print(f"{name} perturbed {k}: iter {it} hpwl {s.hpwl}", flush=True)
    o.close()
    d.close()
    return out


def one_wide_rows(name, seed=1, iters=3):
    """The oracle's first rows with its global float sums in double (the GPU path's FP64
    reductions): what the GPU's first rows must equal, the reference's float sums aside."""
    import gpl_amd as G
    import oracle as O
    O.set_wide_sums(True)
    d = G.Design.from_config(name, seed=seed)
    d.initial_place()
    d.init_nesterov_place()
    o = O.Oracle(d.problem())
    o.init()
    o.run(iter_budget=iters)
    rows = [dict(iter=r["iter"], trials=r["trials"], overflow=float(r["overflow"]), hpwl=int(r["hpwl"]),
                 density_penalty=float(r["density_penalty"]), step_length=float(r["step_length"])) for r in o.history()[:iters]]
    o.close()
    d.close()
    print(f"{name} wide rows: {[(r['iter'], r['hpwl']) for r in rows]}", flush=True)
    return name, rows


def wide_rows(names):
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=max(1, min(len(names), os.cpu_count() or 1))) as ex:
        for name, rows in ex.map(one_wide_rows, names):
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"loop_{name}.json")
            with open(path) as f:
                fx = json.load(f)
            fx["rows_wide"] = rows
            with open(path, "w") as f:
                json.dump(fx, f, indent=1)


def perturbed(name, n, eps=2e-7, seed=1):
    """n sensitivity runs, one process each (at most the host's cores at once)."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=max(1, min(n, os.cpu_count() or 1))) as ex:
        out = list(ex.map(one_perturbed, [name] * n, [eps] * n, range(1, n + 1), [seed] * n))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"loop_{name}.json")
    with open(path) as f:
        fx = json.load(f)
    fx["perturbed"] = out
    with open(path, "w") as f:
        json.dump(fx, f, indent=1)


def run(name, seed=1):
    import gpl_amd as G
    import oracle as O
    t0 = time.time()
    d = G.Design.from_config(name, seed=seed)
    d.initial_place()
    d.init_nesterov_place()
    setup = time.time() - t0
    ph = problem_hash(d)
    o = O.Oracle(d.problem())
    t1 = time.time()
    rc_init = o.init()
    t_init = time.time() - t1
    rc, it = o.run()
    t_run = time.time() - t1
    s = o.stats()
    hist = o.history()
    out = dict(
        config=name, seed=seed, problem_sha256=ph,
        cells=int(d.problem().num_cells), pins=int(d.problem().num_pins), nets=int(d.problem().num_nets),
        status=int(rc), init_status=int(rc_init), converged=bool(s.converged), final_iter=int(it),
        final_hpwl=int(s.hpwl), final_overflow=float(s.overflow), total_trials=int(s.total_trials),
        first_below_target_seconds=float(s.first_below_target_seconds), init_seconds=float(t_init),
        cpu_seconds_total=float(t_run), setup_seconds=float(setup),
        cpu=cpu_model(), threads=1,
        rows=[dict(iter=r["iter"], trials=r["trials"], overflow=float(r["overflow"]), hpwl=int(r["hpwl"]),
                   density_penalty=float(r["density_penalty"]), step_length=float(r["step_length"]))
              for r in hist[:10]],
        every_50th=[dict(iter=r["iter"], overflow=float(r["overflow"]), hpwl=int(r["hpwl"])) for r in hist[::50]],
    )
    if d.problem().opts.routability:
        # the routability-driven loop's record: totals and one row per pass
        tot, passes = o.route_stats()
        out["route"] = dict(stats={k: (float(v) if isinstance(v, float) else int(v)) for k, v in tot.items()},
                            passes=[{k: (float(v) if isinstance(v, float) else int(v)) for k, v in p.items()} for p in passes])
    o.close()
    d.close()
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"loop_{name}.json")
    if os.path.exists(path):
        # a run that reproduces the fixture's results (a re-timing) keeps its sensitivity runs and
        # wide-sum rows; a run with other results drops them (they belong to the old trajectory)
        with open(path) as f:
            old = json.load(f)
        same = all(old.get(k) == out.get(k) for k in ("problem_sha256", "final_iter", "final_hpwl", "final_overflow",
                                                        "total_trials", "rows", "every_50th", "route"))
        print(f"{name}: {'the same results as' if same else 'results differ from'} the previous fixture", flush=True)
        if same:
            for k in ("perturbed", "rows_wide"):
                if k in old:
                    out[k] = old[k]
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"{name}: iter {it} hpwl {s.hpwl} overflow {s.overflow:.4f} cpu {t_run:.1f}s -> {path}", flush=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args[:1] == ["--wide-rows"]:
        wide_rows(args[1:])
    elif args[:1] == ["--perturb"]:
        k = int(args[1])
        for n in args[2:]:
            perturbed(n, k)
    else:
        for n in args or ["c2", "c3"]:
            run(n)
