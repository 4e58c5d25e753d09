#!/usr/bin/env python3
"""Throughput of the ePlace Nesterov placement loop on one MI355X.

One step is one outer Nesterov iteration (its back-tracking trials included: position update,
density scatter, Poisson solve, WA wirelength forward and gradient, density gather, step length)
of the 1M-cell synthetic placement (1,000,000 one-row cells, 1,050,000 nets, 50 nets of 1k-5k pins,
5 % fixed macros, 1024 x 1024 bins).  The design is generated and initially placed on the host
before anything is timed; np_init and W warmup iterations run untimed, then exactly K iterations
are timed between two stream synchronisations, then the loop runs on to overflow 0.10 for the
time-to-overflow figure.  One JSON line on stdout.

Besides the throughput the line carries:
  roofline      the dominant kernel family's algorithmic bytes per launch over its average device
                duration (HIP events on the library's stream), against MI355X's 8 TB/s;
  cpu_baseline  the oracle (the C restatement of the reference's loop, one host thread) on a bounded
                sample of the same placement: np_init and the first iterations within --cpu-budget
                seconds; and (all_cores) the same on every host thread the box gives (OMP_NUM_THREADS)
                within half that budget;
  steady_state  iterations/s from iteration 50 to convergence of the same run (the whole run's
                rate, not only the timed window), and the time to overflow 0.10;
  hpwl_delta    GPU vs CPU final HPWL on the 1M-cell placement: the GPU run above against the
                oracle's full run on the same problem (tests/golden/loop_c3.json: one core, about
                twelve minutes, too long to repeat here; the problems' fingerprints must match),
                and a fresh GPU-vs-oracle run to convergence on the 2k-cell placement;
  batched       16 seeds of the 50k-cell design placed one after another, then as one lockstep
                batch (each iteration's first trial of all 16 as one launch per kernel, each
                context its own control) and from concurrent contexts, with the results compared;
  routability   the routability-driven loop on the 500k-cell design (768 x 768 bins, RUDY on
                1024 x 1024 tiles), np_init to convergence, against the oracle's full run;
  region_groups the region-group placement (nine regions, grids of 64 to 512 bins a side), np_init
                to convergence with its trials grouped over the regions and with one graph per
                trial, the results compared.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Nesterov iters/s + time to overflow 0.10 at 1M cells vs CPU; HPWL delta"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c1": "2k-cell synthetic placement: 2,000 cells, 2,200 nets, 96 IO ports, 64x64 bins",
    "c2": "200k-cell synthetic placement: 200,000 cells, 210,000 nets (+20 of 500-2,000 pins), 512x512 bins",
    "c3": "1M-cell synthetic placement: 1,000,000 cells, 1,050,000 nets (+50 of 1k-5k pins), 5% fixed macros, 1024x1024 bins",
    "c4s": "50k-cell synthetic placement: 50,000 cells, 52,000 nets, 256x256 bins",
    "c4r": "8 region groups (5k-150k cells, 485k) + 20k top-level cells, per-region bin grids",
    "c5": "500k-cell routability-driven placement: 500,000 cells, 520,000 nets, 768x768 bins, RUDY on 1024x1024 tiles",
}


# kernels of each family, the first one launched once per family launch
FAMILY_KERNELS = {
    "update": ["k_place_update"],
    "scatter": ["k_place_scatter<2, unsigned int>", "k_place_scatter<2, unsigned long long>"],
    "bin_density": ["k_bin_density"],
    "dct_solve": ["k_rows<0, true, 10>", "k_rows<6, true, 10>", "k_rows<3, true, 10>"],
    "wa_forward": ["k_wa_tiles<true>", "k_wa_big<true>"],
    "gradient": ["k_gradient"],
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_c3.txt")


def algorithmic_bytes(fam, N, P, E, M):
    """Compulsory bytes per launch of each kernel family (every array touched once; N cells,
    P pins, E nets, M bins):
    update          81 N          per cell: read cur_slp, cur_sum, cur (24), box (16), flags (1);
                                  write next, next_slp (16), box (16), centre (8)
    scatter         21 N + 16 M   per cell: read box (16), flags (1), scale (4); the two int64 bin
                                  accumulators written once (16 per bin)
    bin_density     44 M          read inst, filler (16), the non-placeable areas as a float pair
                                  (8); write the density (4); clear the accumulators (16)
    dct_solve       12 M          read the density (4); write the field pairs (8)
    wa_forward      32 P + 8 E    per pin: cell (4), offset (8), the cell's centre (8), cell-pin
                                  index (4), its gradient term written (8); per net: start (4),
                                  weight (4)
    gradient        64 N + 8 P + 8 M   per cell: box, pin start, preconditioner, scale (32), the 3
                                  slots the step length reads (24), the preconditioned sum written (8);
                                  per cell-pin: its gradient term (8); the two fields once (8 per bin)
    """
    return {
        "update": 81 * N,
        "scatter": 21 * N + 16 * M,
        "bin_density": 44 * M,
        "dct_solve": 12 * M,
        "wa_forward": 32 * P + 8 * E,
        "gradient": 64 * N + 8 * P + 8 * M,
    }.get(fam)


def read_traffic():
    """HBM bytes per launch of each kernel family, from the committed rocprofv3 --pmc summary
    (lines 'traffic <family> <bytes>'); {} when absent."""
    out = {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                t = line.split()
                if len(t) == 3 and t[0] == "traffic":
                    out[t[1]] = float(t[2])
    except OSError:
        pass
    return out


def traffic_table(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_mib):
    """Per kernel family: HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by what the counter reads for a plain copy of copy_mib MiB."""
    import csv
    from collections import defaultdict

    def per_kernel(path):
        tot, disp = defaultdict(float), defaultdict(set)
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                k = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("gplk::", "").strip()
                tot[k] += float(r["Counter_Value"])
                disp[k].add(r["Dispatch_Id"])
        return tot, {k: len(v) for k, v in disp.items()}

    known = copy_mib * 1024.0  # KiB read and written by the copy
    cf, _ = per_kernel(copy_fetch_csv)
    cw, cn = per_kernel(copy_write_csv)
    probe = max(cw, key=lambda k: cw[k] / cn[k] if "neg" in k or "copy" in k.lower() else 0)
    f_scale = known / (cf[probe] / cn[probe])
    w_scale = known / (cw[probe] / cn[probe])
    ft, fn = per_kernel(fetch_csv)
    wt, wn = per_kernel(write_csv)
    print(f"# copy probe {probe}: FETCH_SIZE x{f_scale:.3f}, WRITE_SIZE x{w_scale:.3f} (known {copy_mib} MiB each way)")
    print(f"# {'family':16s} {'launches':>8s} {'fetch_MB':>10s} {'write_MB':>10s}")
    for fam, ks in FAMILY_KERNELS.items():
        anchor = next((k for k in ks if fn.get(k, 0)), None)  # the family's kernel launched once per launch
        if anchor is None:
            continue
        n = fn[anchor]
        fetch = sum(ft.get(k, 0.0) for k in ks) / n * f_scale * 1024.0
        write = sum(wt.get(k, 0.0) for k in ks) / max(wn.get(anchor, n), 1) * w_scale * 1024.0
        print(f"# {fam:16s} {n:8d} {fetch / 1e6:10.2f} {write / 1e6:10.2f}")
        print(f"traffic {fam} {fetch + write:.0f}")


def copy_probe(mib):
    """A plain copy of `mib` MiB (read mib, write mib) on cuda:0, three times: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch against the known size give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": 4 * n, "bytes_written": 4 * n}), flush=True)


def summarize(paths):
    """Text summary of rocprofv3 CSV files: per-kernel call count and durations from a
    *_kernel_stats.csv, per-kernel average counter values per dispatch from a *_counter_collection.csv."""
    import csv
    from collections import defaultdict

    def short(name):
        name = name.replace("(anonymous namespace)::", "").split("(")[0]
        return name.replace("void ", "").strip()

    for path in paths:
        with open(path, newline="") as f:
            rows = list(csv.DictReader(f))
        if not rows:
            continue
        if "AverageNs" in rows[0]:
            print(f"# {os.path.basename(path)}")
            print(f"{'kernel':60s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}")
            for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
                print(f"{short(r['Name'])[:60]:60s} {int(r['Calls']):7d} {float(r['TotalDurationNs']) / 1e6:10.3f} "
                      f"{float(r['AverageNs']) / 1e3:9.2f} {float(r['MinNs']) / 1e3:9.2f} {float(r['MaxNs']) / 1e3:9.2f} "
                      f"{float(r['Percentage']):6.2f}")
        else:
            acc = defaultdict(lambda: [0, 0.0])
            disp = defaultdict(set)
            for r in rows:
                key = (short(r["Kernel_Name"]), r["Counter_Name"])
                acc[key][1] += float(r["Counter_Value"])
                disp[key].add(r.get("Dispatch_Id", r.get("Correlation_Id", "")))
            print(f"# {os.path.basename(path)}")
            print(f"{'kernel':60s} {'counter':>12s} {'dispatches':>10s} {'avg_per_dispatch':>18s}")
            for (k, cn), (_, tot) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
                nd = max(len(disp[(k, cn)]), 1)
                print(f"{k[:60]:60s} {cn:>12s} {nd:10d} {tot / nd:18.1f}")
        print()


def timeline(path):
    """Median offsets (us from the trial's first kernel) of each kernel of a back-tracking trial,
    from a rocprofv3 --kernel-trace CSV of the timed (graph) path: a trial runs from
    k_place_update to k_finish; the gap to the next trial is the host's turn (statistics copy,
    decisions, graph update, launch).  The last 35 complete trials but 5 are used."""
    import csv
    import statistics as st
    from collections import defaultdict

    def short(name):
        return name.split("(")[0].replace("void ", "").replace("gplk::", "").strip()

    with open(path, newline="") as f:
        ev = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"]), r["Queue_Id"])
                    for r in csv.DictReader(f))
    trials, cur = [], None
    for e in ev:
        if e[2] == "k_place_update":
            if cur:
                trials.append(cur)
            cur = [e]
        elif cur is not None:
            cur.append(e)
    trials = [t for t in trials if any(e[2] == "k_finish" for e in t)][-40:-5]
    span, gaps, at = [], [], defaultdict(list)
    for i, t in enumerate(trials):
        t0 = t[0][0]
        t1 = next(e[1] for e in t if e[2] == "k_finish")
        span.append(t1 - t0)
        if i + 1 < len(trials):
            gaps.append(trials[i + 1][0][0] - t1)
        seen = defaultdict(int)
        for e in t:
            if e[1] <= t1:
                at[(e[2], seen[e[2]])].append((e[0] - t0, e[1] - t0, e[3]))
                seen[e[2]] += 1
    busy = [sum(e[1] - e[0] for e in t if e[1] <= next(x[1] for x in t if x[2] == "k_finish")) for t in trials]
    print(f"# {os.path.basename(path)}: {len(trials)} trials")
    print(f"trial span (k_place_update start to k_finish end): median {st.median(span) / 1e3:.1f} us; "
          f"kernel-busy {st.median(busy) / 1e3:.1f} us; host turn between trials: median {st.median(gaps) / 1e3:.1f} us")
    print(f"{'kernel':44s} {'start_us':>9s} {'end_us':>9s} {'dur_us':>9s} {'queue':>6s}")
    for (k, n), v in sorted(at.items(), key=lambda kv: st.median(x[0] for x in kv[1])):
        print(f"{(k + (f' #{n + 1}' if n else ''))[:44]:44s} {st.median(x[0] for x in v) / 1e3:9.1f} "
              f"{st.median(x[1] for x in v) / 1e3:9.1f} {st.median(x[1] - x[0] for x in v) / 1e3:9.1f} {v[0][2]:>6s}")


def load_fixture(cfg):
    """The oracle's full run of a workload (tests/golden/loop_<cfg>.json), or None."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", f"loop_{cfg}.json")) as f:
            return json.load(f)
    except OSError:
        return None


def batched(cfg, k, emit=True, batch_only=False):
    """K independent placements (seeds 1..K: their own perturbations) of one design: run one
    after another, then as one lockstep batch (gpl_place_batch: each iteration's first trial of
    all K as one launch per kernel), and for comparison at once from host threads (one context
    each); one JSON line when `emit`, and the record returned."""
    import gpl_amd as G

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    t = time.perf_counter()
    d = G.Design.from_config(cfg)
    d.initial_place()

    def contexts():
        out = []
        for seed in range(1, k + 1):
            d.init_nesterov_place(G.default_opts(**dict(G.CONFIGS[cfg]["opts"], seed=seed)))
            out.append(G.NesterovGPU(d))
        return out

    a = contexts()
    p = d.problem()
    log(f"[bench] {cfg} x {k} seeds: cells {p.num_cells} nets {p.num_nets}; setup {time.perf_counter() - t:.1f}s")
    if batch_only:
        t0 = time.perf_counter()
        it_bat = G.place_batched(a)
        t_bat = time.perf_counter() - t0
        for g in a:
            g.close()
        d.close()
        print(json.dumps({"batch_only": True, "config": cfg, "seeds": k, "seconds": round(t_bat, 4), "iterations": it_bat}), flush=True)
        return None
    t0 = time.perf_counter()
    it_seq = [g.do_nesterov_place() for g in a]
    t_seq = time.perf_counter() - t0
    b = contexts()
    t0 = time.perf_counter()
    it_bat = G.place_batched(b)
    t_bat = time.perf_counter() - t0
    same = it_seq == it_bat and all(x.stats().hpwl == y.stats().hpwl for x, y in zip(a, b))
    for g in b:
        g.close()
    c = contexts()
    t0 = time.perf_counter()
    it_par = G.place_concurrently(c)
    t_par = time.perf_counter() - t0
    same_par = it_seq == it_par and all(x.stats().hpwl == y.stats().hpwl for x, y in zip(a, c))
    for g in a + c:
        g.close()
    d.close()
    log(f"[bench] sequential {t_seq:.3f}s, batch {t_bat:.3f}s, concurrent contexts {t_par:.3f}s, same results: {same} {same_par}")
    line = {
        "metric": f"placements/s, {k} seeds of one design, each to overflow 0.10, one lockstep batch",
        "value": round(k / t_bat, 3), "unit": "placements/s", "n_gpus": 1, "higher_is_better": True,
        "sequential_value": round(k / t_seq, 3), "speedup_vs_sequential": round(t_seq / t_bat, 2),
        "concurrent_contexts_value": round(k / t_par, 3), "concurrent_speedup_vs_sequential": round(t_seq / t_par, 2),
        "iterations": it_bat, "same_results_as_sequential": bool(same and same_par), "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS.get(cfg, cfg), "name": cfg, "seeds": k}}
    if emit:
        print(json.dumps(line), flush=True)
    return line


def region_groups_leg(cfg="c4r"):
    """The region-group workload (nine regions of 5k-209k cells on grids of 64 to 512 bins a side,
    one placement): np_init to convergence with its trials grouped over the regions (the default
    for region groups: one launch per kernel variant and step) and with one HIP graph per trial."""
    import gpl_amd as G

    t = time.perf_counter()
    d = G.Design.from_config(cfg)
    d.initial_place(device=True)
    d.init_nesterov_place()
    setup = time.perf_counter() - t
    p = d.problem()
    out = {"config": {"workload": WORKLOADS.get(cfg, cfg), "name": cfg, "regions": p.num_regions,
                      "grids": [int(p.regions[b].cnt_x) for b in range(p.num_regions)]}, "setup_s": round(setup, 2)}
    runs = {}
    for name, mode in (("grouped", 7), ("graph", 0)):
        g = G.NesterovGPU(d)
        g.set_launch_mode(mode)
        g.init_nesterov_loop()
        t0 = time.perf_counter()
        it = g.run_nesterov_loop()
        el = time.perf_counter() - t0
        runs[name] = (it, el, g.stats().hpwl, g.history())
        g.close()
    d.close()
    it, el, hpwl, hist = runs["grouped"]
    out.update({"iterations": it, "seconds": round(el, 4), "us_per_iteration": round(1e6 * el / max(it, 1), 1),
                "final_hpwl": hpwl, "graph_seconds": round(runs["graph"][1], 4),
                "speedup_vs_graph": round(runs["graph"][1] / el, 2),
                "same_results_as_graph": runs["graph"][0] == it and runs["graph"][2] == hpwl and runs["graph"][3] == hist})
    return out


def routability_leg(cfg="c5"):
    """The routability-driven loop on the 500k-cell design (768 x 768 bins, RUDY on 1024 x 1024
    tiles of 8 + 8 tracks, at most 5 passes): the whole placement timed from np_init to convergence
    on the GPU, against the oracle's full run of the same problem (tests/golden/loop_c5.json)."""
    import gpl_amd as G

    t = time.perf_counter()
    d = G.Design.from_config(cfg)
    d.initial_place(device=True)
    d.init_nesterov_place()
    setup = time.perf_counter() - t
    g = G.NesterovGPU(d)
    t = time.perf_counter()
    it = g.do_nesterov_place()
    el = time.perf_counter() - t
    s = g.stats()
    tot, passes = g.route_stats()
    g.close()
    out = {"config": {"workload": "500k-cell synthetic placement: 500,000 cells, 520,000 nets, 768x768 bins, RUDY on "
                                  "1024x1024 tiles (8 H + 8 V tracks), target congestion 1.0, at most 5 inflation passes",
                      "name": cfg},
           "iterations": it, "seconds": round(el, 3), "iters_per_s": round(it / el, 1) if el > 0 else None,
           "converged": bool(s.converged), "final_overflow": round(s.overflow, 4), "final_hpwl": s.hpwl,
           "passes": tot["passes"], "routability_iterations": tot["route_iters"], "final_congestion": round(tot["final_rc"], 4),
           "min_congestion": round(tot["min_rc"], 4),
           "inflation_pct": round(100.0 * tot["total_inflation"] / tot["original_area"], 2) if tot["original_area"] else None,
           "pass_congestion": [round(p["rc"], 4) for p in passes], "setup_s": round(setup, 2)}
    fx = load_fixture(cfg)
    if fx and fx.get("route"):
        out["cpu_run"] = {"iterations": fx["final_iter"], "seconds": round(fx["cpu_seconds_total"], 1), "final_hpwl": fx["final_hpwl"],
                          "passes": fx["route"]["stats"]["passes"], "same_problem": d.problem_sha256() == fx["problem_sha256"],
                          "hpwl_delta": round((s.hpwl - fx["final_hpwl"]) / fx["final_hpwl"], 6),
                          "speedup": round(fx["cpu_seconds_total"] / el, 1) if el > 0 else None,
                          "source": f"tests/golden/loop_{cfg}.json (the oracle's full run, one core)"}
    d.close()
    print(f"[bench] {cfg}: {it} iterations, {tot['passes']} routability passes in {el:.3f}s", file=sys.stderr, flush=True)
    return out


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400, help="timed iterations (K)")
    ap.add_argument("--warmup", type=int, default=50, help="untimed iterations after np_init (W)")
    ap.add_argument("--config", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-baseline sampling")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the HPWL check")
    ap.add_argument("--hpwl-config", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-batched", action="store_true", help="skip the 16-seed batched comparison")
    ap.add_argument("--no-routability", action="store_true", help="skip the routability-driven loop on the 500k-cell design")
    ap.add_argument("--no-region-groups", action="store_true", help="skip the region-group placement (nine regions)")
    ap.add_argument("--launch-mode", type=int, default=0, choices=(0, 1, 2, 3, 4, 5, 6, 7),
                    help="0: each trial one HIP graph (branches on side streams); 1: kernels launched one by "
                         "one on the side streams; 2: one by one on one stream; 3: one graph captured on one stream; "
                         "4: device-driven trials (the graph's control kernel takes the loop's decisions); "
                         "7: grouped trials (one launch per kernel variant and step over the regions)")
    ap.add_argument("--profile-only", action="store_true",
                    help="np_init + W + K iterations of the profiled (serial) path only, the run rocprofv3 wraps; "
                         "no convergence, no CPU leg")
    ap.add_argument("--loop-only", action="store_true",
                    help="np_init + W + K iterations of the timed (graph) path only, the run rocprofv3 wraps for the "
                         "production kernels; no convergence, no CPU leg")
    ap.add_argument("--batch-only", action="store_true",
                    help="with --seeds: the lockstep batch alone (no sequential or concurrent runs), for rocprofv3")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a plain device copy of this many MiB three times and exit (the HBM counter "
                         "calibration under rocprofv3 --pmc)")
    ap.add_argument("--seeds", type=int, default=0,
                    help="the batched form instead of the headline: this many seeds of --config (c4s: the "
                         "50k-cell design) placed one after another, then as one lockstep batch and from concurrent contexts")
    ap.add_argument("--summarize", nargs="+", metavar="CSV",
                    help="print a text summary of rocprofv3 CSV output (kernel_stats.csv or counter_collection.csv) and exit")
    ap.add_argument("--timeline", metavar="KERNEL_TRACE_CSV",
                    help="print the median kernel offsets within a back-tracking trial from a rocprofv3 "
                         "--kernel-trace CSV of the timed (graph) path, and exit")
    ap.add_argument("--traffic-table", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="print HBM bytes per launch of each kernel family from rocprofv3 --pmc runs of the loop "
                         "and of --copy-probe-mib 256, and exit")
    args = ap.parse_args()

    if args.summarize:
        summarize(args.summarize)
        return
    if args.timeline:
        timeline(args.timeline)
        return
    if args.traffic_table:
        traffic_table(*args.traffic_table, copy_mib=256)
        return
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return

    if args.seeds:
        batched(args.config, args.seeds, batch_only=args.batch_only)
        return
    if args.config != "c3":
        args.no_batched = True

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: one placement runs on one MI355X; only N=1 is measured "
              "(placements spread across GPUs one per device; sharding one placement is not built)", file=sys.stderr)
        sys.exit(2)

    import gpl_amd as G

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    # ---- the workload, built on the host (untimed) ----
    t = time.perf_counter()
    d = G.Design.from_config(args.config)
    t_gen = time.perf_counter()
    d.initial_place(device=True)  # the B2B rounds on the GPU: the host's placement bit for bit
    t_ip = time.perf_counter()
    d.init_nesterov_place()
    setup_s = time.perf_counter() - t
    setup_split = {"generate_s": round(t_gen - t, 2), "initial_place_gpu_s": round(t_ip - t_gen, 2),
                   "init_nesterov_s": round(time.perf_counter() - t_ip, 2)}
    p = d.problem()
    N, P, E = p.num_cells, p.num_pins, p.num_nets
    M = sum(p.regions[b].cnt_x * p.regions[b].cnt_y for b in range(p.num_regions))
    log(f"[bench] {args.config}: cells {N} pins {P} nets {E} bins {M}; setup {setup_s:.1f}s")

    # ---- the timed run ----
    g = G.NesterovGPU(d)
    g.set_launch_mode(args.launch_mode)
    if args.profile_only:
        g.set_profiling(True)  # the serial path the roofline leg measures, for rocprofv3 to wrap
    g.init_nesterov_loop()
    g.loop_begin(0)
    it, done = g.loop_step(args.warmup)
    it0 = it
    t0 = time.perf_counter()
    it, done = g.loop_step(args.steps)
    t1 = time.perf_counter()
    steps = it - it0 if not done else it - it0 + 1
    log(f"[bench] timed {steps} iterations in {t1 - t0:.3f}s")
    if args.profile_only or args.loop_only:
        kt = g.kernel_times()
        g.close()
        fam = {k: {"launches": v[0], "avg_us": round(1000.0 * v[1] / max(v[0], 1), 2)} for k, v in kt.items()
               if k in FAMILY_KERNELS}
        print(json.dumps({"profile_only" if args.profile_only else "loop_only": True, "config": args.config, "steps": steps,
                          "iters_per_s": round(steps / (t1 - t0), 2), "families": fam}), flush=True)
        d.close()
        return
    # the steady-state window: iteration 50 (or the timed window's start, when later) to convergence
    t50, it50 = (t0, it0) if it0 >= 50 else (None, None)
    if t50 is None and not done and it < 50:
        it, done = g.loop_step(50 - it)
    if t50 is None and not done:
        t50, it50 = time.perf_counter(), it
    while not done:
        it, done = g.loop_step(1 << 20)
    t_end = time.perf_counter()
    rc, final_iter = g.loop_end(check=False)
    s = g.stats()
    hist = g.history()
    # the host's share of each graph-launched trial: node argument updates, the launch call, the
    # wait for the graph, and the host's own work between two trials
    host = {k: round(1000.0 * v[1] / v[0], 2) for k, v in g.kernel_times().items()
            if k.startswith("host_") and v[0] > 0}
    tuning = g.tuning()
    tune_runs, tune_ms = g.kernel_times().get("autotune", (0, 0.0))
    g.close()
    timed_s = t1 - t0
    value = steps / timed_s if timed_s > 0 else 0.0
    trials = sum(h["trials"] for h in hist) / max(len(hist), 1)
    log(f"[bench] converged={s.converged} at iter {final_iter}; time to overflow {s.first_below_target_seconds:.3f}s; "
        f"hpwl {s.hpwl}; overflow {s.overflow:.4f}")

    # ---- a profiled run for the roofline: device times per kernel family (HIP events) ----
    gp = G.NesterovGPU(d)
    gp.set_profiling(True)
    gp.init_nesterov_loop()
    gp.loop_begin(0)
    # the whole run, np_init to convergence: the families' averages over every phase of the run
    pdone = False
    while not pdone:
        _, pdone = gp.loop_step(1 << 20)
    kt = gp.kernel_times()
    gp.close()
    resorts, global_adds = kt.pop("scatter_global_adds", (0, 0.0))
    fams = {k: v for k, v in kt.items() if algorithmic_bytes(k, N, P, E, M) is not None and v[0] > 0}
    dom = max(fams, key=lambda k: fams[k][1])
    launches, ms = fams[dom]
    avg_s = ms / 1000.0 / launches
    bytes_launch = algorithmic_bytes(dom, N, P, E, M)
    achieved = bytes_launch / avg_s / 1e9
    families = {k: {"launches": v[0], "avg_us": round(1000.0 * v[1] / max(v[0], 1), 2),
                    "share": round(v[1] / max(sum(x[1] for k2, x in kt.items() if k2 != "autotune"), 1e-12), 4),
                    "algorithmic_bytes": algorithmic_bytes(k, N, P, E, M),
                    "achieved_GBs": (round(algorithmic_bytes(k, N, P, E, M) / (v[1] / 1000.0 / v[0]) / 1e9, 1)
                                     if algorithmic_bytes(k, N, P, E, M) and v[0] else None)}
                for k, v in kt.items() if k != "autotune"}
    traffic = read_traffic().get(dom) if args.config == "c3" else None
    roofline = {"bound": "hbm", "kernel": dom, "kernels": FAMILY_KERNELS[dom], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None, "avg_launch_us": round(avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": bytes_launch, "families": families,
                "resorts": resorts, "scatter_global_adds": global_adds}

    # ---- CPU baseline (the oracle, one host thread, bounded sample of the same placement) ----
    cpu = None
    hpwl_delta = None
    if not args.no_cpu:
        import oracle as O
        o = O.Oracle(d.problem())
        o.init()
        orc, oit = o.run(time_budget_s=args.cpu_budget)
        os_ = o.stats()
        n_it = len(o.history())
        model, ncpu = cpu_info()
        cpu_val = n_it / os_.loop_seconds if os_.loop_seconds > 0 else 0.0
        cpu = {"value": round(cpu_val, 4), "unit": "iters/s", "cores": 1, "kind": "port",
               "sample": f"np_init ({os_.init_seconds:.2f}s) + the first {n_it} iterations of the same {args.config} "
                         f"placement in {os_.loop_seconds:.1f}s on one host thread ({model}, {ncpu} CPUs)",
               "init_seconds": round(os_.init_seconds, 3)}
        o.close()
        log(f"[bench] cpu baseline {cpu_val:.3f} it/s over {n_it} iterations")
        # the secondary, stronger baseline: the same sample on the host's cores (the oracle's
        # independent per-cell, per-net, per-row work on every thread, the same bits)
        threads = max(1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))
        O.set_threads(threads)
        try:
            o = O.Oracle(d.problem())
            o.init()
            o.run(time_budget_s=args.cpu_budget / 2)
            os2 = o.stats()
            n2 = len(o.history())
            o.close()
        finally:
            O.set_threads(1)
        all_val = n2 / os2.loop_seconds if os2.loop_seconds > 0 else 0.0
        cpu["all_cores"] = {"value": round(all_val, 4), "unit": "iters/s", "cores": threads, "kind": "port",
                            "sample": f"np_init ({os2.init_seconds:.2f}s) + the first {n2} iterations in {os2.loop_seconds:.1f}s "
                                      f"on {threads} host threads"}
        log(f"[bench] cpu baseline on {threads} threads {all_val:.3f} it/s over {n2} iterations")
        fx = load_fixture(args.config)
        if fx:
            cpu["full_run"] = {"time_to_overflow_s": round(fx["first_below_target_seconds"], 1), "iterations": fx["final_iter"],
                               "iters_per_s": round(fx["final_iter"] / fx["cpu_seconds_total"], 4), "cores": 1,
                               "host": fx["cpu"], "source": f"tests/golden/loop_{args.config}.json (the oracle's full run, "
                                                            "one core of the build container)"}
            same_problem = d.problem_sha256() == fx["problem_sha256"]
            hpwl_delta = {"config": args.config, "gpu_hpwl": s.hpwl, "cpu_hpwl": fx["final_hpwl"],
                          "delta": round((s.hpwl - fx["final_hpwl"]) / fx["final_hpwl"], 6), "gpu_iter": final_iter,
                          "cpu_iter": fx["final_iter"], "gpu_overflow": round(s.overflow, 4),
                          "cpu_overflow": round(fx["final_overflow"], 4), "same_problem": same_problem,
                          "cpu_run": f"tests/golden/loop_{args.config}.json"}
            log(f"[bench] HPWL vs the oracle's full run: {100 * hpwl_delta['delta']:+.4f} % (same problem: {same_problem})")
        # full-run HPWL parity on the small placement
        dh = G.Design.from_config(args.hpwl_config)
        dh.initial_place()
        dh.init_nesterov_place()
        gh = G.NesterovGPU(dh)
        gh.do_nesterov_place()
        sg = gh.stats()
        gh.close()
        oh = O.Oracle(dh.problem())
        oh.init()
        oh.run()
        so = oh.stats()
        oh.close()
        small = {"config": args.hpwl_config, "gpu_hpwl": sg.hpwl, "cpu_hpwl": so.hpwl,
                 "delta": round((sg.hpwl - so.hpwl) / so.hpwl, 6), "gpu_iter": sg.iter, "cpu_iter": so.iter,
                 "gpu_overflow": round(sg.overflow, 4), "cpu_overflow": round(so.overflow, 4)}
        if hpwl_delta is None:
            hpwl_delta = small
        else:
            hpwl_delta["small"] = small
        dh.close()

    batched_line = None
    if not args.no_batched:
        batched_line = batched("c4s", 16, emit=False)
    route_line = None if args.no_routability or args.config != "c3" else routability_leg("c5")
    groups_line = None if args.no_region_groups or args.config != "c3" else region_groups_leg("c4r")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * timed_s / max(steps, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "name": args.config, "cells": N, "pins": P, "nets": E, "bins": M,
                   "seed": 1},
        "time_to_overflow_s": round(s.first_below_target_seconds, 4),
        "iterations_to_overflow": final_iter,
        "converged": bool(s.converged),
        "final_overflow": round(s.overflow, 4),
        "final_hpwl": s.hpwl,
        "trials_per_iteration": round(trials, 3),
        "np_init_s": round(s.init_seconds, 4),
        "host_us_per_trial": host,
        "launch_mode": args.launch_mode,
        "autotune": {"regions": tuning[:4], "measurements": tune_runs, "host_ms": round(tune_ms, 2)},
        "setup_s": round(setup_s, 2),
        "setup": setup_split,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu and cpu["value"] > 0 else None,
        "speedup_vs_cpu_all_cores": (round(value / cpu["all_cores"]["value"], 1)
                                     if cpu and cpu.get("all_cores", {}).get("value") else None),
        "hpwl_delta": hpwl_delta,
        "steady_state": ({"iters_per_s": round((final_iter - it50) / (t_end - t50), 2), "window": [it50, final_iter],
                          "note": "iteration 50 to convergence of the timed run"}
                         if t50 is not None and t_end > t50 and final_iter > it50 else None),
        "batched": batched_line,
        "routability": route_line,
        "region_groups": groups_line,
    }
    print(json.dumps(out), flush=True)
    d.close()


if __name__ == "__main__":
    main()
