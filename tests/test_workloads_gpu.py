"""GPU parity on the generators of the power-law gamebuild workload (50,000
states / 250,000 transitions, out-degree power law alpha 2.1 capped at 200,
60 labels with Zipf(1.2) frequencies, 10% tau, sampled pairs) and of the
serving workload (states log-uniform in [100, 20000], 4 transitions per
state, 4..64 labels, 5% tau, 1..256 pairs), against the oracle.

The power-law games are bimodal: a pair whose processes reach the LTS's
high-degree core grows a game of many millions of positions (2,000 states
and 64 uniform pairs pass 3M positions), the rest close within a few
thousand.  The oracle-sized cases are therefore (a) 64 pairs of a 2,000-state
instance drawn uniformly and kept when their own game closes within 20,000
positions, (b) a 200-state instance whose 16 uniform pairs close at 557,749
positions, and (c) the full 50,000-state instance with 4,096 pairs, capped,
against the oracle's capped breadth-first walk."""
import numpy as np
import pytest

from gpuequiv_amd import GameBuild, TransitionSystem, std_equivalences, workloads
from gpuequiv_amd.serve import Server
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _oracle_game_and_budgets(lts, pairs, positions_wanted=False, threads=8):
    og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs,
                       positions_wanted=positions_wanted)
    reach = [v for v in range(og["V"]) if not og["atk"][v] and og["row"][v] == og["row"][v + 1]]
    orc.set_threads(threads)
    try:
        offs, data, st = orc.solve(6, 3, og["row"], og["col"], og["wt"], og["atk"], reach)
    finally:
        orc.set_threads(1)
    return og, offs, data, st


def _verdicts(offs, data, n_pairs):
    """preorder bits of the 13 equivalences for every start pair (src/energy.rs:307-318)."""
    eqs = np.array([std_equivalences.get(nm).to_vec() for nm in H.EQ_NAMES], dtype=np.int64)
    out = np.zeros((n_pairs, 13), dtype=bool)
    for i in range(n_pairs):
        es = np.array([orc.unpack_energy(6, 3, [int(w)]) for w in data[offs[i]:offs[i + 1]]], dtype=np.int64)
        out[i] = True if es.size == 0 else ~np.all(es[:, None, :] <= eqs[None, :, :], axis=2).any(axis=0)
    return out


def _check_build(b, og, offs, data, st, positions=False):
    h = b._download()
    assert not b.truncated() and b.n_positions() == og["V"]
    for k in ("row", "col", "wt", "atk"):
        assert np.array_equal(h[k], og[k]), k
    if positions:
        from tests.test_gamebuild_gpu import _oracle_pos
        assert b.nodes == [_oracle_pos(p) for p in og["positions"]]
    s = b.solve()
    o, d = b.energies_raw(0, og["V"])
    assert H.canon_lists(o, d) == H.canon_lists(offs, data)
    assert s["rounds"] == st["rounds"] and s["edge_updates"] == st["edge_updates"]
    return o, d


def test_powerlaw_lts_sampled_pairs_game_energies_verdicts():
    print("power-law LTS 2000/10000, 60 Zipf labels, 10% tau; 64 closing pairs", flush=True)
    n, src, dst, lbl = workloads.powerlaw_lts(2000, 10000, 60, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    rng = np.random.default_rng(1)
    pairs = []
    while len(pairs) < 64:
        p, q = (int(x) for x in rng.integers(0, n, 2))
        if p == q or (p, q) in pairs:
            continue
        one = orc.gamebuild(n, lts.row_offsets, lts.targets, lts.labels, [(p, q)], max_positions=20000,
                            positions_wanted=False)
        if not one["truncated"]:
            pairs.append((p, q))
    og, offs, data, st = _oracle_game_and_budgets(lts, pairs, positions_wanted=True)
    b = GameBuild(lts, pairs)
    o, d = _check_build(b, og, offs, data, st, positions=True)
    assert np.array_equal(_verdicts(o, d, len(pairs)), _verdicts(offs, data, len(pairs)))
    b.close()


def test_powerlaw_lts_closed_half_million_position_game():
    print("power-law LTS 200/1000, 16 uniform pairs: 557,749 positions", flush=True)
    n, src, dst, lbl = workloads.powerlaw_lts(200, 1000, 60, 3)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = workloads.sample_pairs(n, 16, 3)
    og, offs, data, st = _oracle_game_and_budgets(lts, pairs)
    assert og["V"] == 557749
    b = GameBuild(lts, pairs)
    o, d = _check_build(b, og, offs, data, st)
    assert np.array_equal(_verdicts(o, d, len(pairs)), _verdicts(offs, data, len(pairs)))
    b.close()


def test_powerlaw_50k_4096_pairs_capped_cut_matches_oracle():
    print("power-law LTS 50000/250000, 4096 pairs, capped at 2M positions", flush=True)
    n, src, dst, lbl = workloads.powerlaw_lts(50000, 250000, 60, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = workloads.sample_pairs(n, 4096, 1)
    cap = 2000000
    b = GameBuild(lts, pairs, max_positions=cap)
    og = orc.gamebuild(n, lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=cap, positions_wanted=False)
    h = b._download()
    assert og["truncated"] and b.truncated() and b.n_positions() == og["V"]
    for k in ("row", "col", "wt", "atk"):
        assert np.array_equal(h[k], og[k]), k
    b.close()


def _serve_jobs(seed, draws, cap):
    rng = np.random.default_rng(seed)
    jobs, skipped = [], 0
    for _ in range(draws):
        n, src, dst, lbl, pairs = workloads.serve_job(rng)
        lts = TransitionSystem.from_arrays(n, src, dst, lbl)
        og = orc.gamebuild(n, lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=cap,
                           positions_wanted=False)
        if og["truncated"]:
            skipped += 1
            continue
        reach = [v for v in range(og["V"]) if not og["atk"][v] and og["row"][v] == og["row"][v + 1]]
        offs, data, _ = orc.solve(6, 3, og["row"], og["col"], og["wt"], og["atk"], reach)
        k = len(pairs)
        jobs.append((lts, pairs, H.canon_lists(offs[:k + 1], data[:offs[k]]), (og["V"], og["E"])))
    return jobs, skipped


def test_serving_workload_jobs_same_alone_in_the_mix_and_oracle():
    # 32 jobs of the serving generator; those whose game the oracle builds
    # within 250,000 positions (about five in six) are submitted as one burst
    print("serve: 32 draws of the serving workload, burst", flush=True)
    jobs, skipped = _serve_jobs(21, 32, 250000)
    assert len(jobs) >= 24, (len(jobs), skipped)
    alone = []
    for lts, pairs, want, size in jobs:
        b = GameBuild(lts, pairs)
        assert (b.n_positions(), int(b.build_stats["edges"])) == size
        b.solve()
        o, d = b.energies_raw(0, len(pairs))
        alone.append(H.canon_lists(o, d))
        b.close()
    srv = Server(budget_bytes=4 << 30, rounds_per_step=4)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    for k, ((lts, pairs, want, size), j) in enumerate(zip(jobs, ids)):
        info, r = res[j]
        assert info["state_name"] == "done", (k, info)
        assert (info["positions"], info["edges"]) == size, k
        got = H.canon_lists(*r)
        assert got == alone[k] == want, k
    st = srv.stats()
    assert st["node_slots_used"] == 0 and st["edge_slots_used"] == 0
    srv.close()
