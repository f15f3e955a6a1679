"""Serving: every job's budgets are the same alone and in the mix (and equal
the oracle's), also when the HBM budget forces jobs to wait for others and
when the energy arena is packed between steps."""
import numpy as np
import pytest

from gpuequiv_amd import GameBuild, TransitionSystem, workloads
from gpuequiv_amd.serve import Server
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

NODE_BYTES, EDGE_BYTES = 87, 12          # per-slot bytes of the position / edge pools


def _budget(positions, edges_per_position):
    """Bytes for pools of `positions` slots (the pools get 7/8 of the budget)."""
    return int(positions * (NODE_BYTES + edges_per_position * EDGE_BYTES) * 8 / 7) + 64


def _alone(lts, pairs, max_v=0):
    """Budgets of the starting positions solved on their own (None if the game
    has more than max_v positions) and the game's size."""
    b = GameBuild(lts, pairs, max_positions=max_v)
    if b.build_stats["truncated"]:
        b.close()
        return None, None
    b.solve()
    o, d = b.energies_raw(0, len(pairs))
    size = (b.n_positions(), int(b.build_stats["edges"]))
    b.close()
    return H.canon_lists(o, d), size


def _jobs(n_jobs, seed, n_lo=10, n_hi=50, l_lo=4, l_hi=7, max_v=60000):
    """Random LTS (3 transitions per state) with 1-11 distinct pairs; jobs
    whose game exceeds max_v positions are skipped.  Returns (lts, pairs,
    budgets alone, (positions, edges))."""
    rng = np.random.default_rng(seed)
    jobs = []
    k = 0
    while len(jobs) < n_jobs:
        k += 1
        n = int(rng.integers(n_lo, n_hi))
        n, src, dst, lbl = workloads.random_lts(n, 3 * n, int(rng.integers(l_lo, l_hi)), 1000 * seed + k)
        lts = TransitionSystem.from_arrays(n, src, dst, lbl)
        want = int(rng.integers(1, 12))
        pairs = []
        for _ in range(100):
            p, q = (int(x) for x in rng.integers(0, n, 2))
            if p != q and (p, q) not in pairs:
                pairs.append((p, q))
            if len(pairs) == want:
                break
        alone, size = _alone(lts, pairs, max_v)
        if alone is not None:
            jobs.append((lts, pairs, alone, size))
    return jobs


def _oracle(lts, pairs):
    g = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs)
    reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
    offs, data, _ = orc.solve(6, 3, g["row"], g["col"], g["wt"], g["atk"], reach)
    return H.canon_lists(offs[:len(pairs) + 1], data[:offs[len(pairs)]])


def test_jobs_same_alone_and_in_the_mix():
    print("serve: 24 jobs, one pool", flush=True)
    jobs = _jobs(24, 1)
    srv = Server(budget_bytes=256 << 20, rounds_per_step=2)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    for k, ((lts, pairs, alone, size), j) in enumerate(zip(jobs, ids)):
        info, r = res[j]
        assert info["state_name"] == "done", info
        got = H.canon_lists(*r)
        assert (info["positions"], info["edges"]) == size
        assert got == alone, k
        if size[0] <= 20000:
            assert got == _oracle(lts, pairs), k
    st = srv.stats()
    assert st["jobs_done"] == len(jobs) and st["node_slots_used"] == 0 and st["edge_slots_used"] == 0
    srv.close()


def test_placement_short_of_memory_waits(monkeypatch):
    # GEQ_SERVER_DEFER_PLACEMENTS=3: the first three placements find no device
    # memory for their scratch (forced on the host); their jobs go back to wait
    # and are placed by later steps, with the same results.
    print("serve: placements deferred three times", flush=True)
    monkeypatch.setenv("GEQ_SERVER_DEFER_PLACEMENTS", "3")
    jobs = _jobs(8, 5)
    srv = Server(budget_bytes=256 << 20, rounds_per_step=2)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    for k, j in enumerate(ids):
        assert res[j][0]["state_name"] == "done", res[j][0]
        assert H.canon_lists(*res[j][1]) == jobs[k][2], k
    st = srv.stats()
    assert st["deferred_admissions"] >= 3
    assert st["jobs_done"] == len(jobs) and st["node_slots_used"] == 0 and st["edge_slots_used"] == 0
    srv.close()


def test_tiny_budget_defers_admission():
    print("serve: budget for about one game", flush=True)
    jobs = _jobs(10, 2)
    big_v = max(size[0] for *_, size in jobs)
    ratio = max(4.0, 1.1 * max(e / max(v, 1) for *_, (v, e) in jobs))
    budget = _budget((big_v + 1) * 1.3, ratio)
    srv = Server(budget_bytes=budget, rounds_per_step=1, edges_per_position=ratio)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    for k, j in enumerate(ids):
        assert res[j][0]["state_name"] == "done", res[j][0]
        assert H.canon_lists(*res[j][1]) == jobs[k][2], k
    first_done = min(res[j][0]["t_done_ms"] for j in ids)
    assert any(res[j][0]["t_admit_ms"] >= first_done for j in ids)
    srv.close()


def test_job_larger_than_budget_fails_alone():
    print("serve: one job larger than the budget", flush=True)
    jobs = sorted(_jobs(4, 3), key=lambda j: j[3][0])
    big_v = jobs[-1][3][0]
    if big_v < 2 * jobs[-2][3][0] + 200:
        pytest.skip("games too alike in size")
    ratio = max(4.0, 1.1 * max(e / max(v, 1) for *_, (v, e) in jobs))
    srv = Server(budget_bytes=_budget(big_v // 2, ratio), edges_per_position=ratio)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    assert res[ids[-1]][0]["state_name"] == "failed"
    for k, j in enumerate(ids[:-1]):
        assert H.canon_lists(*res[j][1]) == jobs[k][2], k
    srv.close()


def test_arena_compaction_keeps_results(monkeypatch):
    print("serve: compaction after every retirement", flush=True)
    monkeypatch.setenv("GEQ_ARENA_ENTRIES", "4096")
    jobs = _jobs(16, 4, 20, 45, 3, 5)
    srv = Server(budget_bytes=64 << 20, rounds_per_step=1, compact_entries=1)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    st = srv.stats()
    for k, j in enumerate(ids):
        assert H.canon_lists(*res[j][1]) == jobs[k][2], k
    print("serve stats", st, flush=True)
    assert st["compactions"] > 0
    srv.close()


def test_invalid_and_empty_jobs_finish_alone():
    print("serve: an invalid job, a job without pairs, and valid jobs in one batch", flush=True)
    jobs = _jobs(3, 5)
    bad = TransitionSystem(np.array([0, 2, 1], np.uint32), np.array([0, 1], np.uint32), np.array([0, 0], np.int32))
    srv = Server(budget_bytes=64 << 20)
    ids = [srv.submit(jobs[0][0], jobs[0][1]), srv.submit(bad, [(0, 1)]), srv.submit(jobs[1][0], []),
           srv.submit(jobs[2][0], jobs[2][1])]
    res = srv.run_until_done(ids)
    assert res[ids[1]][0]["state_name"] == "failed" and res[ids[1]][0]["status"] == 4     # INVALID_ARG
    assert res[ids[2]][0]["state_name"] == "done" and res[ids[2]][1][0].tolist() == [0]
    assert H.canon_lists(*res[ids[0]][1]) == jobs[0][2]
    assert H.canon_lists(*res[ids[3]][1]) == jobs[2][2]
    assert srv.stats()["builds"] >= 1
    srv.close()


@pytest.mark.parametrize("gen", ["sub", "wave", "block"])
def test_chunked_levels_and_the_position_cap(monkeypatch, gen):
    # Tiny chunks: every level of the merged builds runs in pieces of 37
    # frontier positions, with the per-job caps checked after each piece.  The
    # jobs within the cap give exactly their standalone results; a job past it
    # fails alone.
    # With gen wave / block every attack position with a non-empty Q is
    # generated by a wave / block of its own (k_generate_wave /
    # k_generate_heavy), in chunks and ahead of the waits too.
    print("serve: chunked levels, one job over the position cap, generator", gen, flush=True)
    monkeypatch.setenv("GEQ_BUILD_CHUNK", "37")
    monkeypatch.setenv("GEQ_BUILD_GEN", gen)
    jobs = _jobs(10, 8)
    cap = max(size[0] for *_, size in jobs)
    rng = np.random.default_rng(99)
    big = None
    for k in range(40):
        n, src, dst, lbl = workloads.random_lts(120, 360, 4, 9000 + k)
        lts = TransitionSystem.from_arrays(n, src, dst, lbl)
        pairs = [(int(p), int(q)) for p, q in rng.integers(0, n, (8, 2)) if p != q]
        if pairs and _alone(lts, pairs, cap)[0] is None:
            big = (lts, pairs)
            break
    if big is None:
        pytest.skip("no game over the cap found")
    srv = Server(budget_bytes=256 << 20, rounds_per_step=2, max_positions=cap)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    big_id = srv.submit(*big)
    res = srv.run_until_done(ids + [big_id])
    assert res[big_id][0]["state_name"] == "failed"
    for k, j in enumerate(ids):
        info, r = res[j]
        assert info["state_name"] == "done", info
        assert (info["positions"], info["edges"]) == jobs[k][3], k
        assert H.canon_lists(*r) == jobs[k][2], k
    srv.close()


def test_one_build_of_more_than_256_jobs(monkeypatch):
    """A merged build of 300 jobs: the job extraction sorts positions by a
    9-bit job id (two radix passes).  256 jobs are built first and left
    unadmitted, so the builder waits (the ready queue is full) until the
    first step, and then takes all 300 queued jobs at once."""
    import time
    print("serve: one merged build of 300 jobs", flush=True)
    monkeypatch.setenv("GEQ_SERVER_MAX_BATCH", "1024")
    fill = TransitionSystem.from_edges([(0, 1, 0), (1, 0, 0)])
    jobs = _jobs(300, 11, 5, 14, 2, 4)
    srv = Server(budget_bytes=512 << 20, rounds_per_step=4)
    first = [srv.submit(fill, [(0, 1)]) for _ in range(256)]
    deadline = time.time() + 60
    while any(srv.poll(j, with_result=False)[0]["state_name"] != "built" for j in first):
        assert time.time() < deadline, "the first 256 builds did not finish"
        time.sleep(0.01)
    builds0 = srv.stats()["builds"]
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(first + ids)
    assert srv.stats()["builds"] == builds0 + 1
    for j in first:
        assert res[j][0]["state_name"] == "done"
    for k, ((lts, pairs, alone, size), j) in enumerate(zip(jobs, ids)):
        info, r = res[j]
        assert info["state_name"] == "done", info
        assert (info["positions"], info["edges"]) == size, k
        assert H.canon_lists(*r) == alone, k
    srv.close()


def test_failed_merged_build_is_split_per_job(monkeypatch):
    """A merged build that fails (here: a level past a lowered per-level total,
    as a level past 2^32 candidates would) is rebuilt one job at a time, so
    only the jobs that fail by themselves fail; the others equal their
    standalone results."""
    print("serve: a failed merged build is split into one-job builds", flush=True)
    jobs = _jobs(12, 23, 20, 60, 3, 6)
    # the smallest limit at which standalone builds of a third or fewer of the jobs fail
    limit, alone_fails = None, None
    for lim in (1000, 2000, 5000, 10000, 20000, 50000, 100000, 400000):
        monkeypatch.setenv("GEQ_BUILD_LEVEL_LIMIT", str(lim))
        fails = []
        for lts, pairs, _, _ in jobs:
            try:
                GameBuild(lts, pairs).close()
                fails.append(False)
            except Exception as e:   # the refusal: GEQ_ERR_OUT_OF_MEMORY
                assert "2^32 or more" in str(e), e
                fails.append(True)
        if sum(fails) <= len(jobs) // 3:
            limit, alone_fails = lim, fails
            break
    assert limit is not None
    print("level limit", limit, "jobs failing alone", sum(alone_fails), flush=True)
    srv = Server(budget_bytes=256 << 20, rounds_per_step=2)
    ids = [srv.submit(lts, pairs) for lts, pairs, _, _ in jobs]
    res = srv.run_until_done(ids)
    for k, ((lts, pairs, alone, size), j, fails) in enumerate(zip(jobs, ids, alone_fails)):
        info, r = res[j]
        if fails:
            assert info["state_name"] == "failed", (k, info)
        else:
            assert info["state_name"] == "done", (k, info)
            assert H.canon_lists(*r) == alone, k
    st = srv.stats()
    assert st["split_builds"] >= 1, st
    srv.close()
