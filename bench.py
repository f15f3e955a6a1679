#!/usr/bin/env python3
"""Benchmark of the energy-game solver on one MI355X.

Default workload (N=1): one standalone energy game of 100,000 positions
(50% attacker, geometric out-degree with mean 4, uniform targets, updates
identity 0.5 / one Decrement 0.35 / one Min 0.15), seed 2, solved by the plain
one-game path with the graph already resident in HBM.  A step is one
EnergyGame::run: a full solve (energies reset on the device, fixpoint to
convergence) and the download of every position's energies to the host; the
solve alone is reported beside it.  The same line carries the rest of the
metric: the batched-vs-plain summary of 2,048 games (`batch`) and the serving
summary at 200 and 2,000 jobs/s and at saturation (`serve`, completed and
failed jobs counted apart).

The launch shape is picked on the game by the autotuner before the timed
steps (--no-tune keeps the defaults).  Prints one JSON line: edge updates/s as
`value`, the k_eval roofline (kernel time from start/stop events carried by
its dispatches on the solver's stream; traffic from the rocprofv3 summary under
profiles/) and the oracle's single-core time on the same game as
`cpu_baseline`.

Other workloads (--workload): batch (2,048 games in one device-resident solve
against one call per game), serve (Poisson arrivals into the continuous-batching
server against the plain path), spectroscopy (the full 13-equivalence pipeline
of a 300-state LTS, GPU against the oracle), gamebuild (a 50,000-state
power-law LTS with 4,096 pairs).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, AMD spec


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measured_traffic(profile: str) -> float | None:
    """HBM bytes per launch of a kernel, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/ (null when absent)."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_cores() -> int:
    """Cores this process may use, at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(graph, reach, edge_updates: int, budget_s: float = 8.0) -> dict:
    """The oracle on the same game: single-threaded, then with the round's
    evaluations spread over every host core (OpenMP); each about budget_s."""
    from oracle import oracle as orc

    def timed(threads: int):
        orc.set_threads(threads)
        times = []
        t_end = time.perf_counter() + budget_s
        while True:
            t0 = time.perf_counter()
            orc.solve(6, 3, graph.row_offsets, graph.column_indices, graph.weights.array.reshape(-1),
                      graph.attacker_pos, reach)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() > t_end or len(times) >= 200:
                break
        orc.set_threads(1)
        return statistics.median(times), len(times)

    med1, n1 = timed(1)
    cores = host_cores()
    medn, nn = timed(cores) if cores > 1 else (med1, n1)
    return {"value": edge_updates / medn, "unit": "edge updates/s", "cores": cores, "kind": "port",
            "sample": f"the same 100k-position game solved by the oracle {nn} times with {cores} threads "
                      f"(median {medn*1e3:.1f} ms) and {n1} times on one thread (median {med1*1e3:.1f} ms), "
                      f"{cpu_model()}",
            "single_thread_value": edge_updates / med1}


def bench_batch(args) -> dict:
    """2,048 random games (position counts log-uniform in [1,000, 200,000]):
    one batched device-resident solve vs the same games solved one per launch
    sequence (each game uploaded beforehand, so both paths start from HBM)."""
    from gpuequiv_amd import _lib, workloads
    from gpuequiv_amd.batch import GameBatch

    t0 = time.perf_counter()
    graphs = workloads.game_batch(args.games, args.seed, 1000, 200000)
    gen_s = time.perf_counter() - t0
    V = sum(g.n_vertices() for g in graphs)
    E = sum(g.n_edges() for g in graphs)
    print(f"bench: {len(graphs)} games, {V} positions, {E} edges generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    b = GameBatch(graphs)
    for _ in range(args.warmup):
        b.solve()
    tune = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tune = b.autotune(repeats=1)
        tune["tuning_s"] = time.perf_counter() - t0
        print(f"bench: autotune {tune}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = b.solve()
    batched_s = (time.perf_counter() - t0) / args.steps
    prof = b.solve(profile=True)
    print(f"bench: batched solve {batched_s*1e3:.2f} ms", file=sys.stderr, flush=True)
    # the other solve mode on the same batch: replayed round graphs (1) or one
    # persistent launch with a block per game (3); same lists and counts
    mode = b.solve_mode()
    want = b.raw_energies()
    other = 3 if mode == 1 else 1
    b.set_solve_mode(other)
    b.solve()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        so = b.solve()
    other_s = (time.perf_counter() - t0) / args.steps
    same = all(np.array_equal(o0, o1) and np.array_equal(d0, d1) for (o0, d0), (o1, d1) in zip(want, b.raw_energies()))
    same = same and (so["edge_updates"], so["node_evals"], so["rounds"]) == (s["edge_updates"], s["node_evals"], s["rounds"])
    b.set_solve_mode(mode)
    modes = {"used": mode, f"mode_{mode}_ms": batched_s * 1e3, f"mode_{other}_ms": other_s * 1e3,
             "same_lists_and_counts": bool(same),
             "modes": "1: replayed round graphs over the whole batch; 3: one persistent launch, a block per game"}
    print(f"bench: solve modes {modes}", file=sys.stderr, flush=True)

    if args.no_plain:          # profiling runs: the batched solve alone
        kms = prof["eval_kernel_ms"]
        return {"metric": "batched solve only (profiling run)", "ms_per_step": batched_s * 1e3,
                "edge_updates_per_step": int(s["edge_updates"]), "node_evals_per_step": int(s["node_evals"]),
                "rounds": int(s["rounds"]), "k_eval_solves": args.warmup + args.steps + 1,
                "k_eval_ms": kms, "alg_bytes": int(prof["alg_bytes"])}
    lib = _lib.lib()
    ctx = _lib.Context.get()
    handles, keep = [], []
    for g in graphs:
        gs = g._c_struct(keep)
        h = C.c_void_p()
        _lib.check(lib.geq_game_upload(ctx.handle, C.byref(gs), C.byref(h)))
        handles.append(h)
    st = _lib.geq_stats()
    # every game solved twice first (a game's first solve launches its rounds
    # directly, the second captures them): its round graph is captured and
    # its standard reach found on the device, as for the default plain line
    for h in handles:
        for _ in range(2):
            _lib.check(lib.geq_game_solve_standard(h, 0, C.byref(st)))
    t0 = time.perf_counter()
    plain_edges = 0
    lat = []
    for h in handles:
        t1 = time.perf_counter()
        _lib.check(lib.geq_game_solve_standard(h, 0, C.byref(st)))
        lat.append(time.perf_counter() - t1)
        plain_edges += st.edge_updates
    plain_s = time.perf_counter() - t0
    print(f"bench: plain path {plain_s*1e3:.1f} ms", file=sys.stderr, flush=True)

    # parity of the two paths on a sample of games
    rng = np.random.default_rng(0)
    sample = sorted(rng.choice(len(graphs), size=min(32, len(graphs)), replace=False).tolist())
    batched_out = b.raw_energies()
    mismatches = 0
    for i in sample:
        out = _lib.geq_energies()
        _lib.check(lib.geq_game_energies(handles[i], C.byref(out)))
        try:
            o, d = _lib.energies_to_numpy(out)
        finally:
            lib.geq_energies_free(C.byref(out))
        bo, bd = batched_out[i]
        if not (np.array_equal(o, bo) and np.array_equal(d, bd)):
            mismatches += 1
    for h in handles:
        lib.geq_game_free(h)
    b.close()

    edge_updates = int(s["edge_updates"])
    assert edge_updates == plain_edges, "batched and plain paths relaxed different edge counts"
    kms = prof["eval_kernel_ms"]
    achieved = prof["alg_bytes"] / (kms / 1e3) / 1e9 if kms else 0.0
    avg_us = kms * 1e3 / max(1, int(prof["eval_launches"]))
    # measured HBM traffic per launch (rocprofv3 summary of this workload) over this run's launch time
    traffic = measured_traffic("batch_2048_k_eval_pmc.md") if len(graphs) == 2048 else None
    traffic_gbs = traffic / (avg_us * 1e-6) / 1e9 if traffic and avg_us else None
    lat_ms = np.array(lat) * 1e3
    return {
        "metric": "edge updates/sec + games/sec (batched vs plain), p50/p99 solve latency, 1 GPU",
        "value": edge_updates / batched_s,
        "unit": "edge updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": batched_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"{len(graphs)} independent random energy games, positions log-uniform in "
                               f"[1000, 200000], seed {args.seed}; one batched device-resident solve per step",
                   "games": len(graphs), "positions": V, "edges": E, "rounds": int(s["rounds"]),
                   "edge_updates_per_step": edge_updates},
        "games_per_s": {"batched": len(graphs) / batched_s, "plain": len(graphs) / plain_s,
                        "batched_over_plain": plain_s / batched_s},
        "plain_path": {"total_ms": plain_s * 1e3, "edge_updates_per_s": plain_edges / plain_s,
                       "latency_ms_p50": float(np.percentile(lat_ms, 50)),
                       "latency_ms_p99": float(np.percentile(lat_ms, 99))},
        "parity_sample": {"games_compared": len(sample), "mismatches": mismatches},
        "solve_modes": modes,
        "autotune": tune,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_eval",
                     "traffic_source": "profiles/batch_2048_k_eval_pmc.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
                     "traffic_gbs": traffic_gbs,
                     "launches": int(prof["eval_launches"]),
                     "avg_launch_us": avg_us,
                     "alg_bytes_per_launch": prof["alg_bytes"] / max(1, int(prof["eval_launches"]))},
    }


def bench_spectroscopy(args) -> dict:
    """Full spectroscopy of a random LTS (300 states, 1,200 transitions with
    uniform sources and targets, 8 labels): bisimulation minimisation, the
    game of every pair in the same enabledness class built on the GPU, the
    solve, and the 13 verdicts of every ordered state pair."""
    from gpuequiv_amd import GameBuild, TransitionSystem, workloads
    from gpuequiv_amd.energy import std_equivalences
    from gpuequiv_amd.equivalence import Equivalence, preorder_matrix_of

    n, src, dst, lbl = workloads.random_lts(args.states, 4 * args.states, 8, args.seed)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    eqs = [std_equivalences.get(nm) for nm in std_equivalences.NAMES]

    def pipeline(check: bool = False):
        t0 = time.perf_counter()
        reduced, part = lts.bisimilar_minimize()
        t1 = time.perf_counter()
        b, info = GameBuild.compare_all(reduced)
        t2 = time.perf_counter()
        st = b.solve()
        t3 = time.perf_counter()
        # the starting points' 13 verdicts on the device (only they come back)
        holds = b.preorder_bits(eqs)
        t4 = time.perf_counter()
        mat = preorder_matrix_of(info, part, holds)   # all 13 preorders of all ordered state pairs
        t5 = time.perf_counter()
        if check:   # the same matrix as from the downloaded budgets
            eq = Equivalence(info, b.energies(0, b.n_starting_points))
            eq.set_minimization(part)
            if not np.array_equal(mat, eq.preorder_matrix(eqs)):
                raise RuntimeError("device verdicts differ from the downloaded budgets' verdicts")
        sizes = (b.n_positions(), int(b.build_stats["edges"]), len(info.starting_points))
        b.close()
        return [t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4], st, sizes

    pipeline(check=True)

    for _ in range(args.warmup):
        pipeline()
    runs = [pipeline() for _ in range(args.steps)]
    phases = np.median(np.array([r[0] for r in runs]), axis=0)
    st, (V, E, n_pairs) = runs[-1][1], runs[-1][2]
    total = float(phases.sum())
    line = {
        "metric": "edge updates/sec + games/sec (batched vs plain), p50/p99 solve latency, 1 GPU",
        "value": n * n / total, "unit": "ordered state pairs decided/s (13 verdicts each)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": total * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"full spectroscopy (13 equivalences, all state pairs) of a random LTS: {n} states, "
                               f"{len(src)} transitions (uniform sources/targets), 8 labels, seed {args.seed}",
                   "positions": V, "edges": E, "start_pairs": n_pairs, "rounds": int(st["rounds"]),
                   "edge_updates": int(st["edge_updates"])},
        "phases_ms": {"minimise_gpu": phases[0] * 1e3, "gamebuild_gpu": phases[1] * 1e3,
                      "solve_gpu": phases[2] * 1e3, "verdicts_gpu": phases[3] * 1e3, "matrix_host": phases[4] * 1e3},
        "solve_edge_updates_per_s": int(st["edge_updates"]) / phases[2],
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        t0 = time.perf_counter()
        part, cnt, orow, odst, olbl = orc.bisimilar_minimize(n, lts.row_offsets, lts.targets, lts.labels)
        red = TransitionSystem(orow, odst, olbl)
        pairs = [(p, q) for grp in red.enabledness_partitions() for p in grp for q in grp if p != q]
        t1 = time.perf_counter()
        g = orc.gamebuild(red.n_vertices(), red.row_offsets, red.targets, red.labels, pairs)
        t2 = time.perf_counter()
        reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
        orc.solve(6, 3, g["row"], g["col"], g["wt"], g["atk"], reach)
        t3 = time.perf_counter()
        cpu_total = t3 - t0
        line["cpu_baseline"] = {"value": n * n / cpu_total, "unit": line["unit"], "cores": 1, "kind": "port",
                                "sample": f"the same pipeline on the oracle once: minimise {1e3*(t1-t0):.1f} ms, "
                                          f"gamebuild {1e3*(t2-t1):.0f} ms, solve {1e3*(t3-t2):.0f} ms "
                                          f"(verdicts not timed), single thread, {cpu_model()}"}
        line["gpu_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    return line


def bench_gamebuild(args) -> dict:
    """A power-law LTS (50,000 states, 250,000 transitions, out-degree
    alpha 2.1 capped at 200, 60 labels with Zipf(1.2) frequencies, 10% tau)
    and 4,096 uniformly sampled ordered pairs p != q: the game is built on the
    GPU, then solved.  The build stops after the level that passes
    --max-positions (the game of this workload can outgrow any budget)."""
    from gpuequiv_amd import GameBuild, TransitionSystem, workloads

    n, src, dst, lbl = workloads.powerlaw_lts(args.states, 5 * args.states, 60, args.seed)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = workloads.sample_pairs(n, args.pairs, args.seed)
    for _ in range(args.warmup):
        GameBuild(lts, pairs, max_positions=args.max_positions).close()
    builds, solves = [], []
    for i in range(args.steps):
        b = GameBuild(lts, pairs, max_positions=args.max_positions)
        bs = b.build_stats
        builds.append(bs["build_ms"])
        if not bs["truncated"] or i == args.steps - 1:
            # a capped game is solved as a game of its own (its unexpanded
            # frontier is never a position to reach): timed once, and reported
            # as the capped game's solve, not the reference's answer
            b.solve(capped=True)
            solves.append(b.solve(capped=True))
        b.close()
    V, E = int(bs["positions"]), int(bs["edges"])
    bms = float(np.median(builds))
    line = {
        "metric": "edge updates/sec + games/sec (batched vs plain), p50/p99 solve latency, 1 GPU",
        "value": V / (bms / 1e3), "unit": "positions built/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": bms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"gamebuild + solve: power-law LTS {n} states / {len(src)} transitions, 60 labels "
                               f"Zipf(1.2), 10% tau, {len(pairs)} sampled pairs, seed {args.seed}, "
                               f"build cap {args.max_positions} positions",
                   "positions": V, "edges": E, "levels": int(bs["levels"]), "truncated": bool(bs["truncated"])},
        "gamebuild_edges_per_s": E / (bms / 1e3),
    }
    if solves:
        s = solves[-1]
        line["solve"] = {"of": "the capped game" if bs["truncated"] else "the game",
                         "rounds": int(s["rounds"]), "edge_updates": int(s["edge_updates"]), "solve_ms": s["solve_ms"],
                         "edge_updates_per_s": int(s["edge_updates"]) / (s["solve_ms"] / 1e3),
                         "alg_bytes_per_s": int(s["alg_bytes"]) / (s["solve_ms"] / 1e3)}
    if not args.no_cpu:
        # the oracle's breadth-first build of the same LTS and pairs, capped lower (bounded sample)
        from oracle import oracle as orc
        cap = min(args.max_positions, 2000000)
        t0 = time.perf_counter()
        og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=cap,
                           positions_wanted=False)
        cpu_s = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": og["V"] / cpu_s, "unit": "positions built/s", "cores": 1, "kind": "port",
                                "sample": f"the oracle's build of the same LTS and pairs capped at {cap} positions: "
                                          f"{og['V']} positions, {og['E']} edges in {cpu_s:.1f} s, single thread, "
                                          f"{cpu_model()}"}
        line["gpu_over_cpu"] = line["value"] / line["cpu_baseline"]["value"]
    return line


def _serve_jobs(n_jobs: int, seed: int):
    from gpuequiv_amd import TransitionSystem, workloads
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n_jobs):
        n, src, dst, lbl, pairs = workloads.serve_job(rng)
        jobs.append((TransitionSystem.from_arrays(n, src, dst, lbl), pairs))
    return jobs


def _fifo_latency(arrivals, service):
    """Latencies of one server taking jobs in arrival order (the plain path)."""
    done, lat = 0.0, []
    for a, sv in zip(arrivals, service):
        done = max(done, a) + sv
        lat.append(done - a)
    return np.array(lat)


def bench_serve(args) -> dict:
    """Serving: Poisson arrivals of spectroscopy jobs (random LTS, states
    log-uniform in [100, 20000], 4 transitions per state, 4..64 labels, 5% tau,
    1..256 pairs) into the continuous-batching server, at several arrival
    rates; against the plain path (each job built and solved alone, one after
    the other) on the same jobs."""
    from gpuequiv_amd import GameBuild, _lib
    from gpuequiv_amd.serve import Server

    jobs = _serve_jobs(args.jobs, args.seed)
    free, _total = _lib.Context.get().memory()
    budget = int(min(args.budget_gb * 1e9, 0.6 * free))
    max_pos = args.max_positions
    print(f"bench: {len(jobs)} jobs, budget {budget/1e9:.0f} GB, max {max_pos} positions per job", file=sys.stderr,
          flush=True)

    # plain path: each job alone (build, solve, read the starting budgets back)
    service, sizes, plain_res, plain_edges = [], [], [], 0
    def build_alone(lts, pairs):
        # None: the job fails as the server fails it (a level past the 32-bit
        # totals, or device memory)
        try:
            return GameBuild(lts, pairs, max_positions=max_pos)
        except _lib.Error:
            return None

    for lts, pairs in jobs[:8]:                       # warm-up
        b = build_alone(lts, pairs)
        if b is not None and not b.build_stats["truncated"]:
            b.solve()
        if b is not None:
            b.close()
    for lts, pairs in jobs:
        t0 = time.perf_counter()
        b = build_alone(lts, pairs)
        if b is None:
            service.append(time.perf_counter() - t0)
            sizes.append((0, 0))
            plain_res.append(None)
            continue
        if b.build_stats["truncated"]:
            res = None
        else:
            st = b.solve()
            res = b.energies_raw(0, len(pairs))
            plain_edges += int(st["edge_updates"])
        service.append(time.perf_counter() - t0)
        sizes.append((b.n_positions(), int(b.build_stats["edges"])))
        plain_res.append(res)
        b.close()
    service = np.array(service)
    ok = [r is not None for r in plain_res]
    plain_s = float(service.sum())
    print(f"bench: plain path {plain_s:.2f} s for {len(jobs)} jobs ({sum(ok)} within the size limit)", file=sys.stderr,
          flush=True)

    # one long-lived server for every rate (creating and freeing a pool of
    # this size per rate would time the release and re-clearing of device memory)
    srv = Server(budget_bytes=budget, rounds_per_step=args.rounds_per_step, max_positions=max_pos,
                 builders=args.builders)
    srv.run_until_done([srv.submit(lts, pairs) for lts, pairs in jobs[:16]])     # warm-up
    rates = [float(r) for r in args.rates.split(",")]
    sweep = []
    for rate in rates:
        rng = np.random.default_rng(args.seed + int(rate))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
        st0 = srv.stats()
        ids, submit_at, delay = [], {}, {}
        out = {}
        nxt = 0
        t0 = time.perf_counter()
        while len(out) < len(jobs):
            now = time.perf_counter() - t0
            while nxt < len(jobs) and arrivals[nxt] <= now:
                lts, pairs = jobs[nxt]
                j = srv.submit(lts, pairs)
                ids.append(j)
                delay[j] = (time.perf_counter() - t0) - arrivals[nxt]
                nxt += 1
            resident, waiting = srv.step()
            for j in srv.finished():
                out[j] = srv.poll(j)
            if len(out) == nxt and nxt < len(jobs):        # idle until the next arrival
                time.sleep(max(0.0, min(1e-3, arrivals[nxt] - (time.perf_counter() - t0))))
            elif resident == 0:                              # games still being built
                time.sleep(1e-4)
        wall = time.perf_counter() - t0
        st1 = srv.stats()
        st = {k: st1[k] - st0[k] for k in ("steps", "rounds", "compactions", "builds", "jobs_done")}
        st.update({k: st1[k] for k in ("node_slots", "edge_slots", "arena_entries")})
        for j in ids:
            srv.release(j)
        lat, mism, failed = [], 0, 0
        for k, j in enumerate(ids):
            info, res = out[j]
            if info["state_name"] != "done":
                failed += 1
                continue
            lat.append(info["t_done_ms"] - info["t_submit_ms"] + delay[j] * 1e3)
            pr = plain_res[k]
            if pr is None or not (np.array_equal(pr[0], res[0]) and np.array_equal(pr[1], res[1])):
                mism += 1
        lat = np.array(lat)
        plain_lat = _fifo_latency(arrivals, service)[np.array(ok)] * 1e3
        span = wall - arrivals[0]
        done = len(jobs) - failed
        sweep.append({
            "rate_jobs_per_s": rate, "saturation": len(jobs) / rate < 0.01,   # every job within 10 ms
            "jobs": len(jobs), "completed": done, "failed_over_size_limit": failed,
            "completed_jobs_per_s": done / span, "finished_jobs_per_s": len(jobs) / span, "served_wall_s": wall,
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "plain_fifo_latency_ms_p50": float(np.percentile(plain_lat, 50)),
            "plain_fifo_latency_ms_p99": float(np.percentile(plain_lat, 99)),
            "mismatches_vs_alone": mism, "server": st,
        })
        print(f"bench: rate {rate:g}/s: {done/span:.1f} completed jobs/s ({failed} failed), p50 "
              f"{sweep[-1]['latency_ms_p50']:.1f} ms, p99 {sweep[-1]['latency_ms_p99']:.1f} ms, mismatches {mism}",
              file=sys.stderr, flush=True)

    srv.close()
    best = max(sweep, key=lambda r: r["completed_jobs_per_s"])
    V = sum(v for (v, _), k in zip(sizes, ok) if k)
    E = sum(e for (_, e), k in zip(sizes, ok) if k)
    plain_done = sum(ok) / plain_s
    return {
        "metric": "edge updates/sec + games/sec (batched vs plain), p50/p99 solve latency, 1 GPU",
        "value": best["completed_jobs_per_s"], "unit": "completed jobs/s",
        "n_gpus": 1, "steps": len(jobs), "warmup": 8, "ms_per_step": 1e3 / best["completed_jobs_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"serving: {len(jobs)} spectroscopy jobs, Poisson arrivals at {args.rates} jobs/s, "
                               "LTS states log-uniform in [100, 20000], 4 transitions per state, 4..64 labels, "
                               f"5% tau, 1..256 pairs; jobs over {max_pos} positions rejected; seed {args.seed}",
                   "positions": V, "edges": E, "budget_bytes": budget},
        "plain_path": {"completed_jobs_per_s": plain_done, "completed": int(sum(ok)),
                       "failed_over_size_limit": len(jobs) - int(sum(ok)), "total_s": plain_s, "edge_updates": plain_edges,
                       # the jobs over the size limit: standalone builds check it at level ends
                       "over_limit_s": float(service[~np.array(ok)].sum()),
                       "over_limit_positions": int(sum(v for (v, _), k in zip(sizes, ok) if not k)),
                       "service_ms_p50": float(np.percentile(service * 1e3, 50)),
                       "service_ms_p99": float(np.percentile(service * 1e3, 99))},
        "completed_over_plain": best["completed_jobs_per_s"] / plain_done,
        "sweep": sweep,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["plain", "batch", "serve", "spectroscopy", "gamebuild"], default="plain")
    ap.add_argument("--states", type=int, default=None, help="spectroscopy / gamebuild: LTS states")
    ap.add_argument("--pairs", type=int, default=4096, help="gamebuild: sampled pairs")
    ap.add_argument("--jobs", type=int, default=400, help="serve: jobs per arrival rate")
    ap.add_argument("--rates", default="50,200,1000,2000,100000",
                    help="serve: arrival rates (jobs/s); 100000 submits all jobs at once, after the other rates have "
                         "warmed the server (its saturation throughput)")
    ap.add_argument("--max-positions", type=int, default=None,
                    help="serve: per-job game size limit (2M); gamebuild: build cap (200M)")
    ap.add_argument("--rounds-per-step", type=int, default=4, help="serve: rounds per server step")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM admission budget (GB)")
    ap.add_argument("--builders", type=int, default=3, help="serve: game-building threads")
    ap.add_argument("--positions", type=int, default=100000)
    ap.add_argument("--games", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="plain / batch: keep the default launch shape")
    ap.add_argument("--no-plain", action="store_true", help="batch: skip the one-game-per-call comparison")
    ap.add_argument("--no-batch", action="store_true", help="default line: skip the batched-vs-plain summary")
    ap.add_argument("--no-serve", action="store_true", help="default line: skip the serving summary")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB copied once by a plain device copy first (a known-size kernel to scale PMC byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"bench: copy probe of {args.copy_probe} MiB done", file=sys.stderr, flush=True)
        del x, y
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a launched path", file=sys.stderr)
        return 2
    if args.workload in ("spectroscopy", "gamebuild"):
        if args.seed is None:
            args.seed = 1
        if args.workload == "spectroscopy":
            args.states = args.states or 300
            print(json.dumps(bench_spectroscopy(args)), flush=True)
        else:
            args.states = args.states or 50000
            args.max_positions = args.max_positions or 20000000
            print(json.dumps(bench_gamebuild(args)), flush=True)
        return 0
    if args.workload == "serve":
        if args.seed is None:
            args.seed = 5
        args.max_positions = args.max_positions or 2000000
        print(json.dumps(bench_serve(args)), flush=True)
        return 0
    if args.workload == "batch":
        if args.seed is None:
            args.seed = 4
        print(json.dumps(bench_batch(args)), flush=True)
        return 0
    if args.seed is None:
        args.seed = 2

    from gpuequiv_amd import _lib, workloads

    lib = _lib.lib()
    ctx = _lib.Context.get()
    graph = workloads.random_game(args.positions, args.seed)
    deg = np.diff(graph.row_offsets.astype(np.int64))
    reach = np.ascontiguousarray(np.nonzero((graph.attacker_pos == 0) & (deg == 0))[0], dtype=np.uint32)
    keep: list = []
    gs = graph._c_struct(keep)
    handle = C.c_void_p()
    _lib.check(lib.geq_game_upload(ctx.handle, C.byref(gs), C.byref(handle)))
    st = _lib.geq_stats()

    def solve(profile: int = 0):
        # EnergyGame::standard_reach: the defender dead ends (the `reach` list
        # above), found on the device at the first solve and kept with the game
        _lib.check(lib.geq_game_solve_standard(handle, profile, C.byref(st)))
        return st.as_dict()

    for _ in range(args.warmup):
        solve()
    tune = None
    if not args.no_tune:
        # launch shape picked on this game by the autotuner, outside the timed region
        t0 = time.perf_counter()
        rep = _lib.geq_tune_report()
        _lib.check(lib.geq_game_autotune(handle, _lib.u32p(reach), len(reach), 5, C.byref(rep)))
        tune = rep.as_dict()
        tune["tuning_s"] = time.perf_counter() - t0
        print(f"bench: autotune {tune}", file=sys.stderr, flush=True)
    out = _lib.geq_energies()

    def download():
        # the energies of every position back on the host, as EnergyGame::run
        # returns them (src/energygame.rs:379-383)
        _lib.check(lib.geq_game_energies(handle, C.byref(out)))
        lib.geq_energies_free(C.byref(out))

    # a step: EnergyGame::run on the resident game = solve + energies download
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = solve()
        download()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    edge_updates = int(s["edge_updates"])
    value = edge_updates / (ms_per_step / 1e3)
    # the solve alone (no download), as the same number of steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
    solve_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # dominant kernel: k_eval.  The timed steps replay captured round graphs,
    # so its launches are timed in extra solves of the same game on the same
    # stream that launch each round's kernels directly, with HIP events
    # around every k_eval / k_expand.
    n_prof = max(3, min(args.steps, 20))
    eval_ms = expand_ms = alg = 0.0
    launches = 0
    for _ in range(n_prof):
        prof = solve(profile=1)
        eval_ms += prof["eval_kernel_ms"]
        expand_ms += prof["expand_kernel_ms"]
        alg += prof["alg_bytes"]
        launches += int(prof["eval_launches"])
    assert int(prof["edge_updates"]) == edge_updates and int(prof["rounds"]) == int(s["rounds"])
    achieved = alg / (eval_ms / 1e3) / 1e9 if eval_ms > 0 else 0.0
    avg_us = eval_ms * 1e3 / max(1, launches)
    traffic = measured_traffic("plain_100k_k_eval_pmc.md") if (args.positions, args.seed) == (100000, 2) else None
    traffic_gbs = traffic / (avg_us * 1e-6) / 1e9 if traffic and avg_us else None
    phase = {"profiled_solves": n_prof, "k_eval_ms_per_solve": eval_ms / n_prof,
             "k_expand_ms_per_solve": expand_ms / n_prof, "rounds": int(prof["rounds"]),
             "profiled_solve_ms": prof["solve_ms"]}
    _lib.lib().geq_game_free(handle)

    line = {
        "metric": "edge updates/sec + games/sec (batched vs plain), p50/p99 solve latency, 1 GPU",
        "value": value,
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": f"single standalone energy game, {args.positions} positions (50% attacker), "
                               f"geometric out-degree mean 4, uniform targets, seed {args.seed}; plain one-game path",
                   "positions": graph.n_vertices(), "edges": graph.n_edges(), "rounds": int(s["rounds"]),
                   "edge_updates_per_solve": edge_updates, "node_evals_per_solve": int(s["node_evals"])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/plain_100k_k_eval_pmc.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
                     "traffic_gbs": traffic_gbs,
                     "kernel": "k_eval", "launches": launches, "avg_launch_us": avg_us,
                     "alg_bytes_per_launch": alg / max(1, launches)},
        "per_kernel_profile": phase,
        "step": "solve + download of every position's energies (EnergyGame::run on the resident game)",
        "solve_only": {"ms_per_step": solve_ms, "edge_updates_per_s": edge_updates / (solve_ms / 1e3)},
        "device_solve_ms": s["solve_ms"],
        "autotune": tune,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(graph, reach, edge_updates)
        line["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    # the rest of the metric on the same run: batched vs plain games/s, and
    # serving latency and throughput against the plain path
    if not args.no_batch:
        b = bench_batch(argparse.Namespace(games=2048, seed=4, steps=3, warmup=1, no_tune=True, no_plain=False))
        line["batch"] = {k: b[k] for k in ("config", "games_per_s", "plain_path", "parity_sample", "ms_per_step", "solve_modes")}
        line["batch"]["edge_updates_per_s"] = b["value"]
        line["batch"]["roofline"] = b["roofline"]
        print(f"bench: batch {b['games_per_s']}", file=sys.stderr, flush=True)
    if not args.no_serve:
        sv = bench_serve(argparse.Namespace(jobs=400, seed=5, rates="200,2000,100000", max_positions=2000000,
                                            rounds_per_step=4, budget_gb=200.0, builders=3))
        line["serve"] = {k: sv[k] for k in ("config", "plain_path", "completed_over_plain", "sweep")}
        line["serve"]["completed_jobs_per_s"] = sv["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
