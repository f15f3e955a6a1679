"""Solve mode 3 (one persistent launch, a block per game) against mode 1
(replayed round graphs over the whole batch) and the oracle: the same
energies, node evaluations, edge updates and round count, also when the
energy arena overflows mid-solve and when games re-tag their slots."""
import numpy as np
import pytest

from gpuequiv_amd import EnergyConf, workloads
from gpuequiv_amd.batch import GameBatch
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _graphs(n, seed, lo=50, hi=5000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(int)
    return [workloads.random_game(int(v), 10 * seed + i) for i, v in enumerate(sizes)]


def _solve(graphs, mode, **tune):
    b = GameBatch(graphs)
    if tune:
        b.set_tune(**tune)
    b.set_solve_mode(mode)
    assert b.solve_mode() == mode
    st = b.solve()
    res = [H.canon_lists(o, d) for o, d in b.raw_energies()]
    b.close()
    return res, st


def _counts(st):
    return st["rounds"], st["node_evals"], st["edge_updates"], st["alg_bytes"]


def test_games_mode_matches_graph_mode_and_oracle():
    graphs = _graphs(48, 3)
    print("solve mode 3:", len(graphs), "games", flush=True)
    want, s1 = _solve(graphs, 1)
    got, s3 = _solve(graphs, 3)
    assert s3["mode"] == 3 and s1["mode"] == 1
    assert got == want
    assert _counts(s3) == _counts(s1)
    for g, lists in list(zip(graphs, got))[:12]:
        o, d, _ = H.oracle_solve_game(g, H.standard_reach(g))
        assert lists == H.canon_lists(o, d)


@pytest.mark.parametrize("waves", [1, 4, 8])
def test_games_mode_block_sizes(waves):
    graphs = [workloads.random_game(v, 500 + v) for v in (20000, 3, 1500, 60000, 1, 700)]
    print("solve mode 3, waves per block", waves, flush=True)
    want, s1 = _solve(graphs, 1)
    got, s3 = _solve(graphs, 3, waves_per_block=waves)
    assert got == want and _counts(s3) == _counts(s1)


def test_games_mode_renumbered_batch(monkeypatch):
    # games renumbered in reverse breadth-first order at upload keep their
    # contiguous ranges, which are what a block of mode 3 works on
    monkeypatch.setenv("GEQ_RENUMBER", "1")
    graphs = _graphs(20, 13)
    want, s1 = _solve(graphs, 1)
    got, s3 = _solve(graphs, 3)
    assert got == want and _counts(s3) == _counts(s1)


def test_games_mode_single_game():
    g = workloads.random_game(100000, 42)
    print("solve mode 3, one game of 100000 positions", flush=True)
    want, s1 = _solve([g], 1)
    got, s3 = _solve([g], 3)
    assert got == want and _counts(s3) == _counts(s1)
    # the device time covers the persistent launch (one block runs the game)
    assert s3["solve_ms"] > 0 and s1["solve_ms"] > 0


def test_autotune_keeps_replayed_rounds_for_one_large_game():
    # one block per game cannot beat the whole GPU on a single large game
    g = workloads.random_game(100000, 43)
    b = GameBatch([g])
    b.autotune(repeats=2)
    assert b.solve_mode() == 1
    b.close()


def test_games_mode_arena_overflow_relaunches(monkeypatch):
    # a tiny arena: launches stop at round boundaries, the arena grows, the
    # games go on; re-run rounds count twice, so only values and rounds compare
    graphs = _graphs(24, 7, 200, 20000)
    print("solve mode 3 with a 64-entry arena", flush=True)
    want, s1 = _solve(graphs, 1)
    monkeypatch.setenv("GEQ_ARENA_ENTRIES", "64")
    got, s3 = _solve(graphs, 3)
    assert got == want
    assert s3["rounds"] == s1["rounds"]
    assert s3["eval_launches"] > 1


@pytest.mark.parametrize("interval", [1, 3, 7])
def test_games_mode_retags(monkeypatch, interval):
    # games re-tag their slots every `interval` rounds, and games longer than
    # the interval re-tag to round 0 when they finish
    graphs = _graphs(16, 11, 100, 30000)
    print("solve mode 3, re-tag every", interval, "rounds", flush=True)
    want, s1 = _solve(graphs, 1)
    monkeypatch.setenv("GEQ_RETAG_ROUNDS", str(interval))
    got, s3 = _solve(graphs, 3)
    assert got == want and _counts(s3) == _counts(s1)


def test_games_mode_resolve_and_empty_games():
    graphs = [workloads.random_game(v, 900 + v) for v in (1, 2, 5000, 1, 300)]
    b = GameBatch(graphs)
    b.set_solve_mode(3)
    s_a = b.solve()
    a = [H.canon_lists(o, d) for o, d in b.raw_energies()]
    s_b = b.solve()
    assert [H.canon_lists(o, d) for o, d in b.raw_energies()] == a
    assert _counts(s_a) == _counts(s_b)
    b.close()
    want, _ = _solve(graphs, 1)
    assert a == want


def test_games_mode_rejects_list_confs_and_bad_modes():
    conf = EnergyConf(8, 2)   # past the bitmap layout: the list kernels
    g = workloads.random_game_conf(300, 5, conf)
    b = GameBatch([g])
    with pytest.raises(RuntimeError, match="bitmap energy layout"):
        b.set_solve_mode(3)
    with pytest.raises(RuntimeError, match="solve mode"):
        b.set_solve_mode(0)
    b.close()
