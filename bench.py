#!/usr/bin/env python3
"""Benchmark: traversed edges/s per LPA iteration, MG8 asynchronous, on R-MAT scale 24.

Workload (one GPU): R-MAT scale 24 (16,777,216 vertices), edge factor 16, (a, b, c, d) =
(0.57, 0.19, 0.19, 0.05), seeded, symmetrised, deduplicated, self-loops dropped, float32 weights
~ U[0.5, 1.5) per unordered pair; MG8 sketch, asynchronous mode, tolerance 0.05, at most 20
iterations (RakOptions defaults). The graph is uploaded once (resident in HBM) before timing.

A step is one complete label-propagation run on that graph: label initialisation, marking, and
every iteration until convergence (the span the reference times, inc/rakLowmemCuda.hxx:588-595).
`value` = traversed edges of the timed steps / their wall time. Traversed edges of an iteration =
sum of the degrees of the vertices it processes.

Roofline: the move launches of each iteration are bracketed by HIP events on the graph's stream;
`achieved` = algorithmic bytes of the move phase (per iteration: N flag bytes + A (8+4+1) +
12 E + 4 C + 5 X, where A, E, C, X are the active vertices, their edges, the moved vertices and
their edges; an iteration that flags every vertex at its end writes N flag bytes instead of 5 X,
one that marks at its end adds 3 N) / measured move time, against the 8 TB/s HBM peak.

cpu_baseline: the oracle's restatement of rakLowmemStaticOmp<true,true,8> (pinned bit-exact to
the compiled reference single-threaded) on the same graph, with all host cores; beside it the same
method's first iteration on one thread, and GVE-LPA (rakStaticOmp restated) with all host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "rak-lowmem-communities-cuda_amd"))

HBM_PEAK_GBS = 8000.0
METHOD = {8: "MG8", 1: "BM", 0: "exact hashtable vote (rakStaticCuda)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(nbytes):
    """One hipMemcpy device-to-device of nbytes: the copy kernel whose counters read a known size."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes device to device")


def load_traffic(args):
    """HBM bytes per move-phase iteration measured by rocprofv3 counters on this workload (profiles/)."""
    path = os.path.join(ROOT, "profiles", "rmat24_mg8_async_traffic.json")
    if (args.workload != "rmat" or args.scale != 24 or args.unit or args.slots != 8 or args.group_degree != "0,0,0,0"
            or args.sweeps != 0 or args.blanket):
        return None
    try:
        with open(path) as f:
            t = json.load(f)
        return {"bytes_per_iteration": t["traffic_bytes_per_iteration"], "source": os.path.relpath(path, ROOT)}
    except (OSError, ValueError, KeyError):
        return None


def bench_batch(args):
    """Batched workload: seeded SBM graphs (vertex counts log-uniform in [1k, 1M], 4..1024 blocks,
    degree U[4, 32], 70-90% intra-block edges, unit weights) packed into one device arena and run
    in the deterministic synchronous mode (MG8), all graphs at once vs one graph per launch."""
    import hashlib

    import numpy as np

    import rlc

    P = rlc.sbm_mix_params(args.graphs, seed=args.batch_seed)
    nsum = int(P[0].sum())
    ecap = int(1.15 * float((P[0].astype(np.float64) * P[2]).sum())) + (1 << 20)
    ctx = rlc.Context(0)
    b = ctx.batch(nsum + args.graphs, ecap, args.graphs)
    t0 = time.time()
    chunk = 256
    m_total = 0
    sample = {}  # a few graphs kept on the host for the check against rlc_run
    pick = set(np.random.default_rng(7).choice(args.graphs, size=min(8, args.graphs), replace=False).tolist())
    pick |= {int(np.argmin(P[0])), int(np.argmax(P[0]))}
    for c0 in range(0, args.graphs, chunk):
        sl = slice(c0, min(c0 + chunk, args.graphs))
        S = rlc.gen_sbm_set(*(x[sl] for x in P))
        for i in range(len(S)):
            g = S.graph(i)
            gid = b.add(g)
            assert gid == c0 + i
            m_total += g.m
            if gid in pick:
                sample[gid] = rlc.Graph(g.offsets.copy(), g.edges.copy(), None)
        log(f"batch: {min(c0 + chunk, args.graphs)}/{args.graphs} graphs generated and added "
            f"({time.time() - t0:.0f}s, {m_total} entries)")
    opts = rlc.options(mode=rlc.MODE_SYNC, slots=args.slots)
    t1 = time.time()
    tuned = b.tune(opts)
    log(f"autotuner: {tuned.as_dict()} ({time.time() - t1:.1f}s)")
    for _ in range(args.warmup):
        b.run(opts)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        st = b.run(opts)
    wall_b = time.perf_counter() - t_start
    log(f"batched: {args.steps} runs, {wall_b:.3f}s, steps={st.steps}, device {st.time_ms:.1f} ms/run")
    digest = []
    for gid in range(args.graphs):
        lab, it = b.result(gid)
        digest.append((hashlib.blake2b(lab.tobytes(), digest_size=16).hexdigest(), it))
    if args.batched_only:
        b.close()
        ctx.close()
        print(json.dumps({"metric": "batched graphs/s (batched leg only)", "value": args.graphs * args.steps / wall_b,
                          "unit": "graphs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": wall_b * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": f"{args.graphs} seeded SBM graphs", "batch_steps": st.steps,
                                     "traversed_edges_per_run": st.traversed_edges}}), flush=True)
        return
    # reference: each sampled graph alone through rlc_run (its own upload, deterministic mode)
    for gid, g in sample.items():
        with ctx.upload(g) as dg:
            r = dg.run(opts)
        lab, it = b.result(gid)
        assert it == r.iterations and np.array_equal(lab, r.membership), f"graph {gid} differs from rlc_run"
    # one launch per graph, with a tuning of its own (autotuned on the largest graph alone)
    big = int(np.argmax(P[0]))
    tuned_1 = b.tune(opts, ids=[big])
    log(f"autotuner, one graph per launch: {tuned_1.as_dict()}")
    t_start = time.perf_counter()
    for gid in range(args.graphs):
        b.run(opts, ids=[gid], tune=tuned_1)
    wall_1 = time.perf_counter() - t_start
    # and the single-graph entry point on a random sample: rlc_run on each graph uploaded alone
    # (upload and label copy-back outside the timed call)
    t_run = 0.0
    rsample = sorted(np.random.default_rng(11).choice(args.graphs, size=min(args.run_sample, args.graphs),
                                                      replace=False).tolist())
    for c0 in range(0, len(rsample), 64):
        part = rsample[c0:c0 + 64]
        S = rlc.gen_sbm_set(*(x[part] for x in P))
        for i, gid in enumerate(part):
            with ctx.upload(S.graph(i)) as dg:
                dg.run(opts, labels=False)  # warm: graph capture
                t = time.perf_counter()
                dg.run(opts, labels=False)
                t_run += time.perf_counter() - t
    run_rate = len(rsample) / t_run if t_run > 0 else 0.0
    log(f"rlc_run on {len(rsample)} sampled graphs: {t_run:.3f}s ({run_rate:.1f} graphs/s)")
    same = 0
    for gid in range(args.graphs):
        lab, it = b.result(gid)
        same += (hashlib.blake2b(lab.tobytes(), digest_size=16).hexdigest(), it) == digest[gid]
    log(f"one-by-one: {wall_1:.3f}s; identical results {same}/{args.graphs}")
    value = args.graphs * args.steps / wall_b
    one = args.graphs / wall_1
    line = {
        "metric": "batched vs 1-by-1 graphs/s",
        "value": value,
        "unit": "graphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_b * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.graphs} seeded SBM graphs, n log-uniform [1e3, 1e6], 4-1024 blocks, degree U[4,32], "
                        f"intra U[0.7,0.9], unit weights, one CSR arena, {'MG8' if args.slots == 8 else 'BM'} "
                        f"deterministic synchronous mode, tol 0.05, max 20 iterations",
            "vertices": nsum, "entries": m_total, "tuning": tuned.as_dict(),
            "one_by_one_graphs_per_s": one, "one_by_one_tuning": tuned_1.as_dict(),
            "rlc_run_graphs_per_s": run_rate, "rlc_run_sample": len(rsample),
            "speedup_vs_one_launch_per_graph": value / one if one > 0 else None,
            "speedup_vs_rlc_run": value / run_rate if run_rate > 0 else None,
            "identical_labels": f"{same}/{args.graphs}", "batch_steps": st.steps,
            "traversed_edges_per_run": st.traversed_edges,
            "rlc_run_checked_graphs": sorted(sample),
        },
    }
    b.close()
    ctx.close()
    if not args.no_cpu_baseline:
        import oracle

        # the oracle's deterministic mode, graph after graph, until about 10 s of CPU work
        cores = len(os.sched_getaffinity(0))
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
        t_cpu, k = 0.0, 0
        while k < args.graphs and t_cpu < 10.0:
            g = rlc.gen_sbm(int(P[0][k]), int(P[1][k]), float(P[2][k]), float(P[3][k]), int(P[4][k]))
            t = time.perf_counter()
            oracle.lpa_sync(g, slots=args.slots, threads=cores)
            t_cpu += time.perf_counter() - t
            k += 1
        line["cpu_baseline"] = {
            "value": k / t_cpu, "unit": "graphs/s", "cores": cores, "kind": "port",
            "sample": f"the first {k} graphs of the mix (random order of the mix) through the oracle's "
                      f"deterministic synchronous mode, {cores} threads, {t_cpu:.1f} s",
        }
    print(json.dumps(line), flush=True)


def serve_stream(ctx, rlc, graphs, arrivals, opts, cfg):
    """Submit graphs[i] at arrivals[i] (seconds from the start) to a fresh server; returns the
    latencies (ms, submit to labels on the host), the makespan (s) and the server's stats."""
    with rlc.Server(ctx, cfg, opts) as srv:
        t0 = time.perf_counter()
        ids = []
        for g, t in zip(graphs, arrivals):
            while True:
                dt = t - (time.perf_counter() - t0)
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.002))
            ids.append(srv.submit(g))
        srv.drain()
        infos = [srv.result(r)[1] for r in ids]
        st = srv.stats()
    lat = [i.done_ms - i.submit_ms for i in infos]
    span = (max(i.done_ms for i in infos) - min(i.submit_ms for i in infos)) / 1e3
    return lat, span, st, infos


def bench_serve(args):
    """Serving workload: a Poisson stream of requests, graphs drawn from the batched workload's
    SBM distribution with 1% R-MAT scale 22 graphs, MG8 asynchronous; continuous batching vs
    one graph resident at a time, on the same arrival times."""
    import numpy as np

    import rlc

    N = args.requests
    P = rlc.sbm_mix_params(N, seed=args.batch_seed + 100)
    t0 = time.time()
    S = rlc.gen_sbm_set(*P)
    graphs = [S.graph(i) for i in range(N)]
    rmat = [rlc.gen_rmat(22, 16, 0.57, 0.19, 0.19, 1.0, 1.0, seed=s) for s in (1, 2)]
    n_rmat = 0
    for i in range(0, N, 100):  # 1% of the requests
        graphs[i] = rmat[(i // 100) % 2]
        n_rmat += 1
    log(f"serve: {N} requests generated ({time.time() - t0:.0f}s), {n_rmat} R-MAT 22 "
        f"(n={rmat[0].n}, m={rmat[0].m}), SBM entries {int(S.estart[-1])}")
    opts = rlc.options(slots=args.slots)  # asynchronous
    budget = int(args.serve_gb * (1 << 30))
    # arena split: ~20 entries per vertex entry (the mix has degree 4-32, R-MAT 22 about 30)
    V = budget // (13 + 20 * 4)
    E = 20 * V
    cfg = rlc.server_config(vertex_capacity=V, edge_capacity=E, max_graphs=4096, steps_per_round=args.round_steps)
    ctx = rlc.Context(0)
    # capacity: a closed burst of the first requests
    burst = graphs[: min(N, 300)]
    lat_b, span_b, _, _ = serve_stream(ctx, rlc, burst, [0.0] * len(burst), opts, cfg)
    cap = len(burst) / span_b
    lam = args.load * cap
    arrivals = np.cumsum(np.random.default_rng(args.batch_seed).exponential(1.0 / lam, N))
    log(f"serve: burst capacity {cap:.1f} graphs/s; Poisson rate {lam:.1f}/s ({args.load:.0%} of it)")
    lat, span, st, infos = serve_stream(ctx, rlc, graphs, arrivals, opts, cfg)
    log(f"continuous batching: makespan {span:.2f}s, rounds {st.rounds}, resident max {st.resident_max}")
    cfg1 = rlc.server_config(vertex_capacity=V, edge_capacity=E, max_graphs=1, steps_per_round=args.round_steps)
    lat1, span1, st1, _ = serve_stream(ctx, rlc, graphs, arrivals, opts, cfg1)
    log(f"one at a time: makespan {span1:.2f}s")
    ctx.close()
    p = lambda x, q: float(np.percentile(np.array(x), q))
    line = {
        "metric": "p99 latency (serving)",
        "value": p(lat, 99),
        "unit": "ms",
        "n_gpus": 1,
        "steps": N,
        "warmup": len(burst),
        "ms_per_step": span * 1e3 / N,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{N} requests, Poisson arrivals at {lam:.1f}/s ({args.load:.0%} of the burst capacity), "
                        f"SBM graphs n log-uniform [1e3, 1e6] + 1% R-MAT scale 22, MG8 asynchronous, "
                        f"{args.serve_gb} GiB arena, {args.round_steps} iterations per round",
            "p50_ms": p(lat, 50), "p99_ms": p(lat, 99), "graphs_per_s": N / span,
            "one_at_a_time": {"p50_ms": p(lat1, 50), "p99_ms": p(lat1, 99), "graphs_per_s": N / span1},
            "burst_capacity_graphs_per_s": cap, "rate_per_s": lam, "resident_max": st.resident_max,
            "rounds": st.rounds, "rmat22_requests": n_rmat,
            "rmat22_p50_ms": p([l for l, i in zip(lat, infos) if i.n == rmat[0].n], 50),
        },
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--group-degree", type=str, default="0,0,0,0",
                    help="degrees at which 4, 16, 64, 256 lanes scan a vertex (0,0,0,0 = library defaults)")
    ap.add_argument("--sweeps", type=int, default=0)
    ap.add_argument("--default-only", action="store_true",
                    help="skip the opt-in blanket leg beside the default (profiling runs)")
    ap.add_argument("--blanket", action="store_true",
                    help="asynchronous MG8 with blanket iterations (rlc_options::blanket = 1) for the timed steps")
    ap.add_argument("--workload", choices=["rmat", "grid", "sbm", "batch", "serve"], default="rmat",
                    help="rmat: R-MAT 24 async (default); grid: 10^8-vertex road grid; sbm: the 65,536-vertex SBM "
                         "in the deterministic synchronous mode; batch: 4096-graph arena; serve: continuous batching")
    ap.add_argument("--unit", action="store_true", help="rmat workload: unit weights (the scale-27 configuration)")
    ap.add_argument("--grid-side", type=int, default=10000, help="grid workload: lattice side (10000: 1e8 vertices)")
    ap.add_argument("--requests", type=int, default=1000, help="serve workload: requests")
    ap.add_argument("--load", type=float, default=0.8, help="serve workload: arrival rate / burst capacity")
    ap.add_argument("--serve-gb", type=float, default=24.0, help="serve workload: arena size (GiB)")
    ap.add_argument("--round-steps", type=int, default=2, help="serve workload: iterations per round")
    ap.add_argument("--graphs", type=int, default=4096, help="batch workload: number of graphs")
    ap.add_argument("--batch-seed", type=int, default=1)
    ap.add_argument("--run-sample", type=int, default=256,
                    help="batch workload: graphs timed through rlc_run one at a time")
    ap.add_argument("--batched-only", action="store_true",
                    help="batch workload: time the batched leg only (no one-by-one legs, no checks; for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after timing, run one device-to-device copy of this many bytes (the known-size kernel "
                         "that scales FETCH_SIZE / WRITE_SIZE counters in a rocprofv3 --pmc run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return

    import numpy as np

    import rlc

    t0 = time.time()
    sync = args.workload == "sbm"
    if args.workload == "sbm":
        g = rlc.gen_sbm(65536, 256, 16.0, 0.8, 1)
        name = ("SBM 65,536 vertices in 256 equal blocks, expected degree 16, 80% intra-block, unit weights, "
                "deterministic synchronous mode")
    elif args.workload == "grid":
        side = args.grid_side
        g = rlc.gen_grid(side, side, 0.10, 0.02, seed=1)
        name = (f"road-like grid {side}x{side} (4-neighbour lattice, 10% of lattice edges dropped, shortcuts "
                f"for 2% of the kept edges), unit weights")
    else:
        wmin, wmax = (1.0, 1.0) if args.unit else (0.5, 1.5)
        g = rlc.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, wmin, wmax, seed=1)
        name = (f"R-MAT scale {args.scale}, edge factor 16, a=.57 b=.19 c=.19, symmetrised+deduplicated, "
                f"{'unit weights' if args.unit else 'w~U[0.5,1.5)'}")
    log(f"graph: {name}: n={g.n} m={g.m} average degree {g.m / g.n:.2f} max_degree={int(g.degrees().max())} "
        f"({time.time() - t0:.1f}s to generate)")
    opts = rlc.options(mode=rlc.MODE_SYNC if sync else rlc.MODE_ASYNC, slots=args.slots, timing=rlc.TIMING_PER_ITERATION,
                       group_degree=[int(x) for x in args.group_degree.split(",")], sweeps=args.sweeps,
                       blanket=1 if args.blanket else 0)
    ctx = rlc.Context(0)
    t0 = time.time()
    dg = ctx.upload(g)
    log(f"upload (device layout): {time.time() - t0:.2f}s")
    for _ in range(args.warmup):
        dg.run(opts, labels=False)
    rlc.lib()  # loaded
    total_edges = 0
    total_bytes = 0
    move_ms = 0.0
    iters = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = dg.run(opts, labels=False)
        st = r.stats
        total_edges += st.traversed_edges
        total_bytes += st.algorithmic_bytes
        move_ms += sum(st.iter_ms[i] for i in range(min(st.iterations, rlc.MAX_ITER_STATS)))
        iters.append(st.iterations)
    wall = time.perf_counter() - t_start
    q, ncom = dg.modularity()
    per_iter = [
        {"l": i, "active": int(st.active[i]), "edges": int(st.edges[i]), "changed": int(st.changed[i]),
         "changed_edges": int(st.changed_edges[i]), "blanket": bool(st.blanket_iterations >> i & 1),
         "move_ms": round(st.iter_ms[i], 4)}
        for i in range(min(st.iterations, rlc.MAX_ITER_STATS))
    ]
    log("last step per iteration:", json.dumps(per_iter))
    achieved = total_bytes / (move_ms * 1e-3) / 1e9 if move_ms > 0 else 0.0
    value = total_edges / wall
    line = {
        "metric": f"traversed edges/s per LPA iter ({'MG8' if args.slots == 8 else METHOD[args.slots]})",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{name}, {METHOD[args.slots]} {'sync' if sync else 'async'}, tol 0.05, max 20 iterations",
            "n": g.n, "m": g.m, "iterations": iters, "modularity": q, "communities": ncom,
            "pruning": "blanket iterations (opt-in)" if args.blanket else "the reference's neighbour marks",
            # with blanket iterations every vertex is processed after an iteration that flagged all of
            # them: the time per step is the work-independent figure
            "traversed_edges_per_step": total_edges / args.steps,
            "fill_iterations_last_step": [i for i in range(min(st.iterations, rlc.MAX_ITER_STATS))
                                          if st.blanket_iterations >> i & 1],
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "move phase (the lane/wave/workgroup move launches of each iteration and its end-of-iteration flag passes)",
            "algorithmic_bytes_per_step": total_bytes / args.steps, "move_ms_per_step": move_ms / args.steps,
        },
    }
    # ceiling of the access pattern itself: one pass of the targets (and weights) of every edge
    # position with the label gathers, no sketch arithmetic (rlc_probe_gather), against the
    # first iteration (every vertex active: every position scanned once)
    probe_ms = dg.probe_gather(5)
    ceil = g.m / (probe_ms * 1e-3) if probe_ms > 0 else 0.0
    it0 = st.edges[0] / (st.iter_ms[0] * 1e-3) if st.iter_ms[0] > 0 else 0.0
    line["roofline"]["gather_probe"] = {
        "ms_per_pass": probe_ms, "edges_per_s": ceil, "first_iteration_edges_per_s": it0,
        "first_iteration_fraction": it0 / ceil if ceil > 0 else None,
    }
    log(f"gather probe: {probe_ms:.3f} ms per pass ({ceil:.3g} edges/s); first iteration {it0:.3g} edges/s")
    if not sync and args.slots == 8 and not args.blanket and not args.default_only:
        # beside the default: the same steps with the opt-in blanket iterations
        ob = rlc.options(mode=rlc.MODE_ASYNC, slots=8, timing=rlc.TIMING_PER_ITERATION,
                         group_degree=[int(x) for x in args.group_degree.split(",")], sweeps=args.sweeps, blanket=1)
        for _ in range(args.warmup):
            dg.run(ob, labels=False)
        e_b, it_b = 0, []
        t_b = time.perf_counter()
        for _ in range(args.steps):
            rb = dg.run(ob, labels=False)
            e_b += rb.stats.traversed_edges
            it_b.append(rb.iterations)
        w_b = time.perf_counter() - t_b
        q_b, _ = dg.modularity()
        line["config"]["blanket_opt_in"] = {"ms_per_step": w_b * 1e3 / args.steps, "edges_per_s": e_b / w_b,
                                            "iterations": it_b, "modularity": q_b}
        log(f"blanket iterations (opt-in): {w_b * 1e3 / args.steps:.2f} ms per step, Q {q_b:.6g}")
    dg.close()
    ctx.close()
    if args.calibrate_copy:
        copy_probe(args.calibrate_copy)
    traffic = load_traffic(args)
    if traffic is not None:
        line["roofline"]["traffic"] = traffic["bytes_per_iteration"]
        line["roofline"]["traffic_source"] = traffic["source"]
    if not args.no_cpu_baseline and sync:
        import oracle

        cores = len(os.sched_getaffinity(0))
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
        # the oracle's deterministic synchronous mode on the same graph, repeated for about 10 s
        t_cpu, k, trav = 0.0, 0, 0
        while t_cpu < 10.0:
            t = time.perf_counter()
            _, it, ost = oracle.lpa_sync(g, slots=args.slots, threads=cores)
            t_cpu += time.perf_counter() - t
            trav += int(np.asarray(ost)[:, 1].sum())
            k += 1
        line["cpu_baseline"] = {
            "value": trav / t_cpu, "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": f"{k} runs of the oracle's deterministic synchronous mode (MG8) on the same graph, "
                      f"{it} iterations each, {t_cpu:.1f} s",
        }
    elif not args.no_cpu_baseline:
        import oracle

        cores = len(os.sched_getaffinity(0))
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
        # a bounded sample: the whole run when the graph is small enough, else its first iteration
        # (every vertex active: the iteration with the most traversed edges)
        full = g.m <= (1 << 30)
        lab, it, tms, trav = oracle.lowmem_omp(g, args.slots == 8, threads=cores,
                                               max_iterations=20 if full else 1)
        what = "one full run" if full else "the first iteration (all vertices active)"
        line["cpu_baseline"] = {
            "value": trav / (tms * 1e-3), "unit": "edges/s", "cores": cores, "kind": "port",
            "sample": f"{what} of {'MG8' if args.slots == 8 else 'BM'} (rakLowmemStaticOmp restated) on the same "
                      f"graph: {it} iterations, {trav} traversed edges, {tms / 1e3:.2f} s",
        }
        # the same method on one thread (its first iteration), and GVE-LPA (rakStaticOmp restated,
        # exact weighted vote) with all cores
        _, it1, tms1, trav1 = oracle.lowmem_omp(g, args.slots == 8, threads=1, max_iterations=1)
        line["cpu_baseline"]["single_thread"] = {
            "value": trav1 / (tms1 * 1e-3), "cores": 1,
            "sample": f"the first iteration on one thread: {trav1} traversed edges, {tms1 / 1e3:.2f} s"}
        _, itg, tmsg, travg = oracle.rak_omp(g, threads=cores, max_iterations=20 if full else 1)
        line["cpu_baseline"]["gve_lpa"] = {
            "value": travg / (tmsg * 1e-3), "cores": cores,
            "sample": f"{what} of GVE-LPA (rakStaticOmp restated): {itg} iterations, {travg} traversed edges, "
                      f"{tmsg / 1e3:.2f} s"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
