"""Benchmark of the incremental HIT-Leiden update on one MI355X.

A step is one edge-delta batch applied by hl_apply_batch_device to the
device-resident graph and hierarchy (apply dG, local moving, refinement,
aggregation per level, def-update), with the batch already in HBM when the
timed region starts.  Default workload: R-MAT scale 24 (16,777,216 vertices,
edge factor 16, a,b,c,d = 0.57,0.19,0.19,0.05, symmetrised, deduplicated,
unit weights), 4 levels, throughput mode, batches of 95,000 R-MAT inserts +
5,000 deletes (one batch per step, warmup + steps batches generated).

Prints one JSON line: edges/s over the timed steps, p50/p99 batch latency,
the roofline of the dominant kernel (HIP events on the engine stream), the
CPU oracle's rate on a bounded sample of the same workload on this host, and
(in "extras") the batched and served forms against their plain
one-batch-per-call paths, each measured by this script in a child process.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "incremental edges/s + p99 update latency vs plain GPU path; HBM GB/s"

# profiler slots that bracket exactly one kernel (both template instances for
# the proposal kernels: movement and refinement), with the name rocprofv3
# reports; the dominant kernel of the roofline is the one of these with the
# most time
SLOT_KERNEL = {
    "propose_small": "k_prop2_first", "try_half": "k_prop2_tab<hl::(Move|Refine)PolicyT, 32u?, 64u?>",
    "try_wave": "k_prop2_tab<hl::(Move|Refine)PolicyT, 64u?, 128u?>",
    "try_lds": "k_prop2_hub<hl::(Move|Refine)PolicyT, 256u?, 2048u?>", "propose_lds": "k_prop2_lds",
    "try_hub": "k_prop2_hub<hl::(Move|Refine)PolicyT, 512u?, 4096u?>",
    "try_giant": "k_prop2_hub<hl::(Move|Refine)PolicyT, 1024u?, 8192u?>", "propose_big": "k_prop2_big",
    "decouple": "k_dec_small", "apply_moves": "k_apply4", "mark_moved": "k_marks_flat",
    "row_merge": "k_row_merge", "resolve_runs": "k_resolve_runs", "move_rounds": "k_move_rounds",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, nbatches, seed):
    from hit_leiden_amd import gen
    if name == "mid":
        w = gen.mid_config(batches=nbatches)
        desc = {"workload": "planted-partition n=1e6, 1000 log-normal blocks (sigma=1), mean degree 32, mu=0.3, "
                            "w~U[0.5,1.5]; batches of 9000 new-edge inserts + 1000 deletes",
                "n": w.n, "m": w.m, "batch_edges": 10000}
    elif name == "small":
        w = gen.small_config()
        while len(w.batches) < nbatches:
            w.make_batches(nbatches, 1000, 0, mode=0, seed=seed)
        desc = {"workload": "LFR-style n=9877, m=21778, tau1=2.5, tau2=1.5 in [5,200], mu=0.2; batches of 1000 "
                            "degree-proportional new edges", "n": w.n, "m": w.m, "batch_edges": 1000}
    elif name == "large":
        w = gen.large_config(batches=nbatches)
        desc = {"workload": "R-MAT scale 24, edge factor 16, (0.57,0.19,0.19,0.05), symmetrised, deduplicated; "
                            "batches of 95000 R-MAT inserts + 5000 deletes", "n": w.n, "m": w.m,
                "batch_edges": 100000}
    else:
        raise SystemExit(f"unknown config {name}")
    return w, desc


def cpu_baseline(w, levels, gamma, budget_s, threads):
    """The CPU oracle (a C++ port of the reference path) on this host: load the
    same graph (untimed), then apply the workload's batches until budget_s of
    batch time is spent."""
    from oracle import hl_oracle as O
    st = O.OracleState(levels=levels, gamma=gamma, mode=O.PAR, threads=threads)
    t0 = time.time()
    st.load(w.n, w.u, w.v, w.w)
    load_s = time.time() - t0
    log(f"[bench] cpu baseline: load {load_s:.1f}s")
    spent, edges, done = 0.0, 0, 0
    for (u, v, a) in w.batches:
        t = time.perf_counter()
        st.apply(u, v, a)
        spent += time.perf_counter() - t
        edges += len(u)
        done += 1
        if spent >= budget_s:
            break
    log(f"[bench] cpu baseline: {done} batches in {spent:.1f}s")
    q = st.modularity(st.hierarchy()[-1])
    return {"value": edges / spent if spent > 0 else 0.0, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"oracle throughput mode ({levels} levels) on the same graph: load {load_s:.1f}s untimed, "
                      f"then the first {done} batches ({edges} edges) timed",
            "load_seconds": load_s, "modularity": q}


CPU_VARIANTS = {
    # name: (oracle mode, threads (0: the box's), description)
    "deterministic_1_thread": ("DET", 1, "oracle deterministic mode (RunMode::Deterministic), one thread"),
    "throughput_as_written": ("PAR|AS_WRITTEN", 0, "oracle throughput mode with the reference's own "
                              "recomputations (CSR rebuilt through a hash map per batch and level, O(n) scans "
                              "per refined vertex and candidate subcommunity)"),
    "deterministic_as_written_1_thread": ("DET|AS_WRITTEN", 1, "oracle deterministic mode with the reference's "
                                          "own recomputations, one thread"),
}


def cpu_variant_child(config, variant, levels, gamma, budget_s, nbatches, threads):
    """One CPU_VARIANTS entry on the workload's graph and batches; prints one
    JSON line.  Runs in a child process that never touches the GPU."""
    from oracle import hl_oracle as O
    mode_s, t, desc = CPU_VARIANTS[variant]
    mode = 0
    for part in mode_s.split("|"):
        mode |= getattr(O, part)
    threads = t or threads
    if config == "large":
        from hit_leiden_amd import gen
        w = gen.large_config(scale=20, batches=nbatches)
    else:
        w, _ = make_workload(config, nbatches, 1)
    st = O.OracleState(levels=levels, gamma=gamma, mode=mode, threads=threads)
    t0 = time.time()
    st.load(w.n, w.u, w.v, w.w)
    load_s = time.time() - t0
    log(f"[bench] cpu variant {variant}: load {load_s:.1f}s")
    spent, edges, done = 0.0, 0, 0
    for (u, v, a) in w.batches:
        t = time.perf_counter()
        st.apply(u, v, a)
        spent += time.perf_counter() - t
        edges += len(u)
        done += 1
        print(json.dumps({"value": edges / spent if spent > 0 else 0.0, "unit": "edges/s", "cores": threads,
                          "kind": "port", "sample": f"{desc} ({levels} levels)"
                          + (" on an R-MAT scale-20 sample" if config == "large" else "")
                          + f": load {load_s:.1f}s untimed, then the first {done} batches ({edges} edges) timed",
                          "load_seconds": round(load_s, 3), "batches": done, "seconds_per_batch":
                          round(spent / done, 4)}), flush=True)
        if spent >= budget_s:
            break


def cpu_variant(args, nbatches, variant, threads):
    """cpu_variant_child in a child process, bounded by --det-timeout; the last
    line it printed (the batches done by then) when the bound cuts it short."""
    import subprocess
    if args.config == "large" and variant.startswith("deterministic"):
        return {"skipped": "the deterministic sweep is superlinear on R-MAT hubs: one thread loads an R-MAT "
                           "scale-14 sample (16,384 vertices) in 31 s and a scale-18 one in more than 10 minutes"}
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", args.config, "--cpu-variant", variant,
           "--levels", str(args.levels), "--gamma", str(args.gamma), "--cpu-budget", str(args.det_budget),
           "--steps", str(nbatches), "--warmup", "0", "--cpu-threads", str(threads)]
    out = ""
    log(f"[bench] cpu variant {variant}: at most {args.det_timeout:.0f}s")
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=sys.stderr, timeout=args.det_timeout, text=True)
        out = r.stdout
    except subprocess.TimeoutExpired as e:
        out = e.stdout.decode() if isinstance(e.stdout, bytes) else (e.stdout or "")
        if not any(ln.startswith("{") for ln in out.splitlines()):
            return {"skipped": f"the load and first batch did not finish within {args.det_timeout:.0f} s"}
    lines = [ln for ln in out.splitlines() if ln.startswith("{")]
    if not lines:
        return {"error": "no result"}
    return json.loads(lines[-1])


def bench_batched(args):
    """The batched form: many independent graphs in one engine, one pass per set
    of batches, against the plain path (one engine per graph, one call per batch)
    on a sample of the same graphs."""
    import hit_leiden_amd as hl
    from hit_leiden_amd import gen, hip
    nb = args.warmup + args.steps
    t0 = time.time()
    ws = gen.batched_config(count=args.graphs, steps=nb)
    graphs = [hl.GraphInput(w.n, w.u, w.v) for w in ws]
    sum_n, sum_m = sum(w.n for w in ws), sum(w.m for w in ws)
    log(f"[bench] batched: {len(ws)} graphs, {sum_n} vertices, {sum_m} edges, gen {time.time() - t0:.1f}s")
    cfg = hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=args.levels, resolution=args.gamma)
    shape = {}
    if args.autotune:
        # launch shapes picked on a sample of the graphs (their first step's deltas), then used for all
        from hit_leiden_amd import autotune
        t0 = time.time()
        pick = np.random.default_rng(5).choice(len(ws), size=min(args.tune_sample, len(ws)), replace=False)
        tuned = autotune.tune_graphs([graphs[i] for i in pick],
                                     {j: hl.GraphInput(ws[i].n, *ws[i].batches[0]) for j, i in enumerate(pick)},
                                     levels=args.levels, gamma=args.gamma, log=log)
        shape = tuned._asdict()
        cfg = tuned.config(mode=hl.RunMode.THROUGHPUT, levels=args.levels, resolution=args.gamma)
        log(f"[bench] autotune picked {shape} on {len(pick)} graphs in {time.time() - t0:.1f}s")
    u = hl.HitLeiden(cfg)
    t0 = time.time()
    u.load_graphs(graphs)
    load_s = time.time() - t0
    dev = []
    for t in range(nb):
        src, dst, a = u.graphs_delta({i: hl.GraphInput(w.n, *w.batches[t]) for i, w in enumerate(ws)})
        dev.append((len(src), hip.DeviceArray(src), hip.DeviceArray(dst), hip.DeviceArray(a)))
    hip.synchronize()
    for t in range(args.warmup):
        k, du, dv, da = dev[t]
        u.apply_batch_device(k, du.address, dv.address, da.address)
    lat, edges = [], 0
    hip.synchronize()
    for t in range(args.warmup, nb):
        k, du, dv, da = dev[t]
        t1 = time.perf_counter()
        u.apply_batch_device(k, du.address, dv.address, da.address)
        lat.append(time.perf_counter() - t1)
        edges += k
    hip.synchronize()
    batched = edges / sum(lat)
    q = u.modularity_graphs()
    u.close()
    for _, a, b, c in dev:
        a.free(), b.free(), c.free()
    log(f"[bench] batched pass: {np.mean(lat) * 1e3:.2f} ms for {edges // args.steps} records; load {load_s:.1f}s")

    # plain path: one engine and one call per graph and batch, on a sample
    rng = np.random.default_rng(3)
    sample = rng.choice(len(ws), size=min(args.plain_sample, len(ws)), replace=False)
    plain_t, plain_e, per_graph = 0.0, 0, []
    for i in sample:
        w = ws[i]
        e = hl.HitLeiden(cfg)
        e.load(graphs[i])
        bd = [(len(b[0]), hip.DeviceArray(b[0]), hip.DeviceArray(b[1]), hip.DeviceArray(b[2])) for b in w.batches]
        hip.synchronize()
        for t in range(args.warmup):
            k, du, dv, da = bd[t]
            e.apply_batch_device(k, du.address, dv.address, da.address)
        for t in range(args.warmup, nb):
            k, du, dv, da = bd[t]
            t1 = time.perf_counter()
            e.apply_batch_device(k, du.address, dv.address, da.address)
            dt = time.perf_counter() - t1
            plain_t += dt
            plain_e += k
            per_graph.append(dt)
        e.close()
        for _, a, b, c in bd:
            a.free(), b.free(), c.free()
    plain = plain_e / plain_t
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC, "value": round(batched, 1), "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(float(lat_ms.mean()), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(ws)} SBM graphs, n log-uniform in [200, 50000], 4-64 blocks, mean degree 8, "
                               "mu=0.25, unit weights; per graph and step one batch of k new edges, k log-uniform "
                               "in [16, 2000]; a step is one pass over every graph's batch",
                   "graphs": len(ws), "vertices": sum_n, "edges": sum_m, "records_per_step": edges // args.steps,
                   "levels": args.levels, "mode": "throughput", "gamma": args.gamma,
                   "launch_shape": shape or "defaults"},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 4), "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "plain_edges_per_s": round(plain, 1), "plain_sample_graphs": len(sample),
        "plain_p50_ms_per_graph": round(float(np.percentile(per_graph, 50)) * 1e3, 4),
        "plain_p99_ms_per_graph": round(float(np.percentile(per_graph, 99)) * 1e3, 4),
        "speedup_vs_plain": round(batched / plain, 2), "load_seconds": round(load_s, 3),
        "modularity_mean": round(float(np.mean(q)), 6), "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_served(args):
    """The served form: resident graphs under Poisson request streams, one open-loop
    run per offered rate with continuous batching, then the plain path (one
    request per pass) at the highest rate."""
    import hit_leiden_amd as hl
    from hit_leiden_amd import gen
    t0 = time.time()
    ws = gen.served_config(count=args.graphs if args.graphs != 4096 else 256)
    sizes = [w.n for w in ws]
    log(f"[bench] served: {len(ws)} graphs, {sum(sizes)} vertices, {sum(w.m for w in ws)} edges, "
        f"gen {time.time() - t0:.1f}s")
    cfg = hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=args.levels, resolution=args.gamma)
    u = hl.HitLeiden(cfg)
    t0 = time.time()
    u.load_graphs([hl.GraphInput(w.n, w.u, w.v) for w in ws])
    load_s = time.time() - t0
    log(f"[bench] load {load_s:.1f}s")
    del ws
    rates = [float(x) for x in args.rates.split(",")]
    runs = []

    def run(rate, count, budget, seed, coalesce):
        rq = gen.served_requests(sizes, rate, count, seed=seed)
        reqs = [(g, hl.GraphInput(sizes[g], a, b), t) for g, a, b, t in rq]
        done, pass_of, st = u.serve(reqs, max_pass_records=budget, coalesce=coalesce)
        arr = np.array([t for _, _, _, t in rq])
        lat = (done - arr) * 1e3
        first = (u.last_emitted - arr) * 1e3  # communities final (the top frontier empty)
        edges = sum(len(a) for _, a, _, _ in rq)
        r = {"offered_req_per_s": rate, "requests": count, "completed_req_per_s": round(count / st.seconds, 2),
             "edges_per_s": round(edges / st.seconds, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
             "p99_ms": round(float(np.percentile(lat, 99)), 3),
             "first_community_p50_ms": round(float(np.percentile(first, 50)), 3),
             "first_community_p99_ms": round(float(np.percentile(first, 99)), 3),
             "max_pass_bytes": int(st.max_pass_bytes), "footprint_cut_passes": int(st.footprint_cut_passes),
             "passes": int(st.passes), "requests_per_pass": round(count / max(st.passes, 1), 2),
             "busy_frac": round(st.busy_seconds / st.seconds, 3), "max_pass_records": budget,
             "coalesce": coalesce}
        log(f"[bench] rate {rate:g}/s budget {budget}: {r}")
        return r

    for i, rate in enumerate(rates):
        runs.append(run(rate, int(min(max(rate * args.served_seconds, 200), 8000)), args.pass_records, 100 + i,
                        not args.no_coalesce))
    u.close()
    del u
    # plain path: one engine per graph, one hit_leiden call per request, first come first served
    top = rates[-1]
    ws = gen.served_config(count=len(sizes))
    engines = []
    t0 = time.time()
    for w in ws:
        e = hl.HitLeiden(cfg)
        e.load(hl.GraphInput(w.n, w.u, w.v))
        engines.append(e)
    del ws
    log(f"[bench] plain: {len(engines)} one-graph engines loaded in {time.time() - t0:.1f}s")
    rq = gen.served_requests(sizes, top, args.plain_requests, seed=999)
    done = np.zeros(len(rq))
    t0 = time.perf_counter()
    for i, (g, a, b, t) in enumerate(rq):
        now = time.perf_counter() - t0
        if now < t:
            time.sleep(t - now)
        engines[g].hit_leiden(hl.GraphInput(sizes[g], a, b))
        done[i] = time.perf_counter() - t0
    span = done[-1]
    lat = (done - np.array([t for _, _, _, t in rq])) * 1e3
    edges = sum(len(a) for _, a, _, _ in rq)
    plain = {"offered_req_per_s": top, "requests": len(rq), "completed_req_per_s": round(len(rq) / span, 2),
             "edges_per_s": round(edges / span, 1), "p50_ms": round(float(np.percentile(lat, 50)), 3),
             "p99_ms": round(float(np.percentile(lat, 99)), 3), "engines": len(engines)}
    log(f"[bench] plain at {top:g}/s: {plain}")
    for e in engines:
        e.close()
    best = runs[-1]
    out = {
        "metric": METRIC, "value": best["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "steps": len(runs),
        "warmup": 0, "ms_per_step": best["p50_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(sizes)} resident LFR-style graphs, n log-uniform in [1e4, 2e6], degree "
                               "exponent 2.5, mean degree 10, mu=0.2; Poisson requests of 10-5000 insertions "
                               "(log-uniform) to a Zipf(1.1) graph", "vertices": int(sum(sizes)),
                   "rates": rates, "levels": args.levels, "mode": "throughput", "gamma": args.gamma},
        "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "runs": runs, "plain": plain,
        "speedup_vs_plain": round(best["completed_req_per_s"] / plain["completed_req_per_s"], 2),
        "load_seconds": round(load_s, 3), "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def batch_summary(stats):
    """Per-batch detail of the timed steps: movement rounds per level, priority
    rounds of the multi-kernel decoupling and the candidates its in-order
    settle took, and the device allocations the batch made (count, GiB, host
    seconds), so that a slow batch can be attributed."""
    return [{"seconds": round(s["seconds"], 5), "rounds": s["rounds"], "decide_rounds": s["decide_rounds"],
             "settled": s["settled"], "evaluated": s["evaluated"], "moved": s["moved"],
             "level_n": s["level_n"], "allocs": s["alloc_count"], "alloc_gib": round(s["alloc_bytes"] / 2**30, 3),
             "alloc_seconds": round(s["alloc_seconds"], 4), "host_syncs": s["host_syncs"]} for s in stats]


def run_extra(args, config, extra):
    """The batched or served form, measured by this script in a child process
    (a fresh device context, run before this process uses the device), as a summary
    of its JSON line with the speed-up over its plain path."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--config", config, "--no-cpu-baseline", "--no-extras",
           "--levels", str(args.levels), "--gamma", str(args.gamma)] + extra
    log(f"[bench] extra: {' '.join(cmd[2:])}")
    t0 = time.time()
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=sys.stderr, timeout=args.extra_timeout, text=True)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {args.extra_timeout}s"}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}"}
    j = json.loads(lines[-1])
    keep = ("value", "unit", "ms_per_step", "p50_ms", "p99_ms", "speedup_vs_plain", "plain_edges_per_s",
            "plain_sample_graphs", "plain", "runs", "steps", "warmup")
    out = {k: j[k] for k in keep if k in j}
    out["workload"] = j["config"]["workload"]
    if "launch_shape" in j["config"]:
        out["launch_shape"] = j["config"]["launch_shape"]
    out["command"] = " ".join(["python", "bench.py"] + cmd[2:])
    out["wall_seconds"] = round(time.time() - t0, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="large", choices=["small", "mid", "large", "batched", "served"])
    ap.add_argument("--rates", default="100,500,2000", help="offered request rates of the served configuration")
    ap.add_argument("--served-seconds", type=float, default=3.0, help="arrival span per offered rate")
    ap.add_argument("--pass-records", type=int, default=1 << 20, help="record budget of a served pass")
    ap.add_argument("--plain-requests", type=int, default=600, help="requests of the one-per-pass run")
    ap.add_argument("--no-coalesce", action="store_true",
                    help="served passes take one request per graph instead of all its waiting ones")
    ap.add_argument("--graphs", type=int, default=4096, help="graphs of the batched configuration")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="graphs timed one by one for the plain path of the batched configuration")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--mode", default="throughput", choices=["throughput", "deterministic"])
    ap.add_argument("--gamma", type=float, default=1.0)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle batch time")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--det-budget", type=float, default=10.0, help="seconds of batch time of each CPU variant "
                    "(deterministic one-thread, as-written)")
    ap.add_argument("--det-timeout", type=float, default=60.0, help="wall-time bound of each CPU variant")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-variant", default="deterministic_1_thread", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched and served forms that the plain configurations report in 'extras'")
    ap.add_argument("--extra-timeout", type=float, default=400.0, help="seconds per extra child process")
    ap.add_argument("--traffic", default=None,
                    help="JSON {slot: HBM bytes per launch} from the FETCH_SIZE / WRITE_SIZE counter runs "
                         "(default: profiles/traffic_<config>.json when present)")
    ap.add_argument("--dump-profile", default=None)
    ap.add_argument("--key-cache", action="store_true",
                    help="cache each row entry's neighbour community for the throughput movement")
    ap.add_argument("--tune-sample", type=int, default=128, help="graphs the batched form's autotuning times")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the degree-bucket thresholds on the warmup batches first")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run device-to-device copies of this many bytes (counter calibration)")
    ap.add_argument("--gather-probe", action="store_true",
                    help="only run the random-gather ceiling probe (hit_leiden_amd/hl_gather_probe)")
    ap.add_argument("--rocprof-summary", default=None, metavar="DIR",
                    help="only print a text summary of a rocprofv3 csv output directory")
    ap.add_argument("--kernel", default=None, help="kernel regex for --rocprof-summary")
    ap.add_argument("--command", default="", help="command line recorded in --rocprof-summary")
    args = ap.parse_args()

    if args.cpu_child:
        cpu_variant_child(args.cpu_child, args.cpu_variant, args.levels, args.gamma, args.cpu_budget, args.steps,
                          args.cpu_threads)
        return
    if args.rocprof_summary:
        from hit_leiden_amd import profiling
        print(profiling.summary_text(args.rocprof_summary, args.command, pattern=args.kernel), end="")
        return
    if args.gather_probe:
        import subprocess
        exe = os.path.join(ROOT, "hit_leiden_amd", "hl_gather_probe")
        if not os.path.exists(exe):
            raise SystemExit("hl_gather_probe is not built (python -m hit_leiden_amd.build)")
        sys.exit(subprocess.run([exe]).returncode)
    if args.copy_probe:
        from hit_leiden_amd import profiling
        profiling.copy_probe(args.copy_probe)
        log(f"[bench] copied {args.copy_probe} bytes device-to-device 4 times")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              f"(the path runs on a single MI355X; sharding over several GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    if args.config == "batched":
        bench_batched(args)
        return
    if args.config == "served":
        bench_served(args)
        return

    extras = None
    if not args.no_extras and args.config in ("large", "mid"):
        # the batched and served forms against their plain one-batch-per-call
        # GPU paths, each in a child process, run before this process uses the
        # device: the runtime grants cooperative launches (the persistent
        # movement rounds) to one process at a time
        extras = {
            "batched": run_extra(args, "batched", ["--steps", "4", "--warmup", "1", "--plain-sample", "32",
                                                   "--autotune", "--tune-sample", "96"]),
            "served": run_extra(args, "served", ["--rates", "2000", "--served-seconds", "2",
                                                 "--plain-requests", "300"]),
        }

    import hit_leiden_amd as hl
    from hit_leiden_amd import hip

    nb = args.warmup + args.steps
    t0 = time.time()
    w, desc = make_workload(args.config, max(nb, 1), 1)
    log(f"[bench] workload {args.config}: n={w.n} m={w.m} batches={len(w.batches)} gen {time.time() - t0:.1f}s")
    if len(w.batches) < nb:
        raise SystemExit("not enough batches generated")

    mode = hl.RunMode.THROUGHPUT if args.mode == "throughput" else hl.RunMode.DETERMINISTIC
    shape = {}
    if args.autotune and mode == hl.RunMode.THROUGHPUT:
        from hit_leiden_amd import autotune
        t0 = time.time()
        tuned = autotune.tune(hl.GraphInput(w.n, w.u, w.v, w.w), w.batches[:args.warmup], levels=args.levels,
                              gamma=args.gamma, log=log)
        shape = tuned._asdict()
        log(f"[bench] autotune picked {shape} in {time.time() - t0:.1f}s")
    g = hl.HitLeiden(hl.RunConfig(resolution=args.gamma, mode=mode, levels=args.levels, key_cache=int(args.key_cache),
                                  **shape))
    t0 = time.time()
    ls = g.load(hl.GraphInput(w.n, w.u, w.v, w.w))
    load_s = time.time() - t0
    q0 = g.modularity(hl.PART_COMMUNITY)
    free, total = hip.mem_info()
    log(f"[bench] load {load_s:.2f}s Q={q0:.4f} levels={ls.as_dict()['level_n']} "
        f"device memory in use {(total - free) / 2**30:.1f} GiB")

    dev = []
    for (u, v, a) in w.batches[:nb]:
        dev.append((len(u), hip.DeviceArray(u), hip.DeviceArray(v), hip.DeviceArray(a)))
    hip.synchronize()

    # warmup with every kernel family timed: picks the dominant kernel
    g.profile(True)
    for i in range(args.warmup):
        k, du, dv, da = dev[i]
        g.apply_batch_device(k, du.address, dv.address, da.address)
    free, total = hip.mem_info()
    mem_gib = (total - free) / 2**30
    log(f"[bench] after warmup: device memory in use {mem_gib:.1f} GiB")
    prof_warm = g.profile_read()
    single = {k: v for k, v in prof_warm.items() if k in SLOT_KERNEL}
    dominant = max(single.items(), key=lambda kv: kv[1][1])[0] if single else None
    log(f"[bench] warmup profile (ms): " + ", ".join(f"{k}={v[1]:.2f}" for k, v in
                                                   sorted(prof_warm.items(), key=lambda kv: -kv[1][1])[:8]))

    # timed region: only the dominant kernel is bracketed by events
    g.profile(True, only=dominant) if dominant else g.profile(False)
    hip.synchronize()
    lat, edges = [], 0
    stats = []
    t_start = time.perf_counter()
    for i in range(args.warmup, nb):
        k, du, dv, da = dev[i]
        t = time.perf_counter()
        s = g.apply_batch_device(k, du.address, dv.address, da.address)
        lat.append(time.perf_counter() - t)
        edges += k
        stats.append(s.as_dict())
    hip.synchronize()
    total = time.perf_counter() - t_start
    prof = g.profile_read()
    g.profile(False)
    q1 = g.modularity(hl.PART_COMMUNITY)
    lat_ms = np.array(lat) * 1000.0
    value = edges / total

    roof = None
    if dominant and dominant in prof:
        launches, ms, byts = prof[dominant]
        per_launch_bytes = byts / launches
        avg_ms = ms / launches
        achieved = (per_launch_bytes / (avg_ms * 1e-3)) / 1e9 if avg_ms > 0 else 0.0
        traffic = None
        tpath = args.traffic or os.path.join(ROOT, "profiles", f"traffic_{args.config}.json")
        if os.path.exists(tpath):
            tj = json.load(open(tpath))
            traffic = tj.get(dominant, {}).get("bytes_per_launch") if isinstance(tj.get(dominant), dict) \
                else tj.get(dominant)
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": SLOT_KERNEL[dominant].replace("hl::(Move|Refine)PolicyT, ", "").replace("u?", ""),
                "kernel_regex": SLOT_KERNEL[dominant],
                "slot": dominant,
                "launches": launches, "avg_launch_us": round(avg_ms * 1000.0, 3),
                "alg_bytes_per_launch": round(per_launch_bytes, 1)}
    bytes_alg = sum(s["bytes_alg"] for s in stats)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(float(lat_ms.mean()), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": dict(desc, levels=args.levels, mode=args.mode, gamma=args.gamma,
                       small_degree=shape.get("small_degree", 16), wave_degree=shape.get("wave_degree", 512),
                       marks_block=shape.get("marks_block", 256), marks_grid=shape.get("marks_grid", 1024),
                       autotuned=bool(shape), key_cache=int(args.key_cache)),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 4), "p99_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "path_gbs": round(bytes_alg / total / 1e9, 3), "path_hbm_frac": round(bytes_alg / total / 1e9 / HBM_PEAK_GBS, 6),
        "modularity_after_load": round(q0, 6), "modularity_after_steps": round(q1, 6), "load_seconds": round(load_s, 3),
        "device_memory_gib": round(mem_gib, 1),
        "roofline": roof,
        "warmup_profile_ms": {k: round(v[1], 3) for k, v in sorted(prof_warm.items(), key=lambda kv: -kv[1][1])[:10]},
        "batches": batch_summary(stats),
    }
    if args.dump_profile:
        json.dump({"warmup": prof_warm, "timed": prof, "stats": stats}, open(args.dump_profile, "w"), indent=1)
    g.close()
    for _, a, b, c in dev:
        a.free(), b.free(), c.free()

    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        log(f"[bench] cpu baseline: oracle with {threads} threads")
        if args.config == "large":
            # the oracle needs minutes for the scale-24 load: time it on a scale-20
            # R-MAT of the same parameters and batch shape instead
            from hit_leiden_amd import gen
            del w
            ws = gen.large_config(scale=20, batches=nb)
            cb = cpu_baseline(ws, args.levels, args.gamma, args.cpu_budget, threads)
            cb["sample"] = "R-MAT scale 20 (same a,b,c, edge factor, batch shape) instead of scale 24: " + cb["sample"]
            out["cpu_baseline"] = cb
            del ws
        else:
            out["cpu_baseline"] = cpu_baseline(w, args.levels, args.gamma, args.cpu_budget, threads)
        out["cpu_baseline"]["cpu_model"] = cpu_model()
        out["cpu_baseline"]["variants"] = {v: cpu_variant(args, nb, v, threads) for v in CPU_VARIANTS}
    else:
        out["cpu_baseline"] = None

    if extras is not None:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
