"""GPU gamebuild (geq_gamebuild) and the LTS pipeline against the oracle and
the reference's own tests (tests/complete.rs, src/lib.rs doctests)."""
import os

import numpy as np
import pytest

from gpuequiv_amd import (EnergyArray, EnergyConf, GameBuild, Position, TransitionSystem, std_equivalences,
                          workloads)
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu
STD = EnergyConf.STANDARD
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fixture_pos(p):
    if p[0] == "A":
        return Position.attack(p[1], p[2])
    if p[0] == "C":
        return Position.clause(p[1], p[2])
    return Position.defend(p[1], p[2], p[3])


def _oracle_pos(p):
    return _fixture_pos(p)


def _assert_same_game(b: GameBuild, lts, pairs):
    og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs)
    h = b._download()
    assert b.n_positions() == og["V"]
    assert b.nodes == [_oracle_pos(p) for p in og["positions"]]
    assert np.array_equal(h["row"], og["row"])
    assert np.array_equal(h["col"], og["col"])
    assert np.array_equal(h["wt"], og["wt"])
    assert np.array_equal(h["atk"], og["atk"])
    return og


def _assert_same_energies(b: GameBuild, og):
    reach = [v for v in range(og["V"]) if not og["atk"][v] and og["row"][v] == og["row"][v + 1]]
    offs, data, st = orc.solve(6, 3, og["row"], og["col"], og["wt"], og["atk"], reach)
    s = b.solve()
    o, d = b.energies_raw(0, og["V"])
    assert H.canon_lists(o, d) == H.canon_lists(offs, data)
    assert s["rounds"] == st["rounds"] and s["edge_updates"] == st["edge_updates"]


# ---------------------------------------------------------------- tests/complete.rs
def test_bisping_fig4_game_and_energies():
    d = H.load("complete_lts.json")["bisping_fig4"]
    lts = TransitionSystem.from_edges(d["transitions"])
    b = GameBuild.compare(lts, *d["compare"])
    assert b.nodes == [_fixture_pos(p) for p in d["positions"]]
    g = b.game
    for v, adj in enumerate(d["adjacency_sorted"]):
        assert sorted(int(x) for x in g.adj(v)) == adj
    b.solve()
    got = b.energies(0, b.n_positions())
    for v, exp in enumerate(d["energies"]):
        assert got[v] == EnergyArray.from_conf(exp, STD), v


def test_bisimilar_compare_is_empty():
    d = H.load("complete_lts.json")["bisimilar"]
    lts = TransitionSystem.from_edges(d["transitions"])
    e = lts.compare(*d["compare"])
    assert e == EnergyArray.empty(STD)
    assert e.test_equivalence(std_equivalences.bisimulation())


def test_failure_trace_verdicts():
    d = H.load("complete_lts.json")["failure_trace"]
    lts = TransitionSystem.from_edges(d["transitions"])
    budgets = lts.compare(*d["compare"])
    for name in d["holds"]:
        assert budgets.test_equivalence(getattr(std_equivalences, name)()), name
    for name in d["fails"]:
        assert not budgets.test_equivalence(getattr(std_equivalences, name)()), name


def test_doctests():
    d = H.load("complete_lts.json")["doctest"]
    lts = TransitionSystem.from_edges(d["transitions"])
    e = lts.compare(0, 4)
    assert not e.test_equivalence(std_equivalences.simulation())
    assert e.test_equivalence(std_equivalences.traces())
    eq = lts.equivalences()
    for p, q, name in d["preorder_true"]:
        assert eq.preorder(p, q, getattr(std_equivalences, name)())
    for p, q, name in d["preorder_false"]:
        assert not eq.preorder(p, q, getattr(std_equivalences, name)())
    for p, q, name in d["equiv_true"]:
        assert eq.equiv(p, q, getattr(std_equivalences, name)())
    for p, q, name in d["equiv_false"]:
        assert not eq.equiv(p, q, getattr(std_equivalences, name)())
    traces = eq.relation(std_equivalences.traces())
    assert traces.count_classes() == d["traces_count_classes"]
    assert list(traces.class_of(0)) == d["traces_class_of_0"]


# ---------------------------------------------------------------- parity with the oracle gamebuild
@pytest.mark.parametrize("n,m,labels,seed", [(12, 30, 2, 1), (30, 90, 3, 2), (60, 200, 4, 3), (300, 1200, 8, 1)])
def test_compare_all_matches_oracle(n, m, labels, seed):
    print("random lts", n, m, labels, seed, flush=True)
    n, src, dst, lbl = workloads.random_lts(n, m, labels, seed)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    if n >= 300:
        lts, _ = lts.bisimilar_minimize()
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    b = GameBuild(lts, pairs)
    og = _assert_same_game(b, lts, pairs)
    _assert_same_energies(b, og)


@pytest.mark.parametrize("heavy_min,label_mask", [("1", "1"), ("40", "1"), ("0", "1"), ("1", "0"), ("0", "0")])
@pytest.mark.parametrize("n,m,labels,seed", [(60, 200, 4, 3), (300, 1200, 8, 1)])
def test_block_generated_positions_match_oracle(monkeypatch, heavy_min, label_mask, n, m, labels, seed):
    # With one thread per position (GEQ_BUILD_GEN=thread), attack positions
    # weighing at least GEQ_BUILD_HEAVY_MIN are generated by a block each (1:
    # every one with a non-empty Q; 0: none).  label_mask 0 turns off the
    # per-state label bit sets (the label walks search the rows instead).  The
    # game and its energies are the oracle's either way, capped builds
    # included.
    print("block-generated positions, heavy_min", heavy_min, "label_mask", label_mask, "lts", n, m, labels, seed,
          flush=True)
    monkeypatch.setenv("GEQ_BUILD_GEN", "thread")
    monkeypatch.setenv("GEQ_BUILD_HEAVY_MIN", heavy_min)
    monkeypatch.setenv("GEQ_BUILD_LABEL_MASK", label_mask)
    n, src, dst, lbl = workloads.random_lts(n, m, labels, seed)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    if n >= 300:
        lts, _ = lts.bisimilar_minimize()
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    b = GameBuild(lts, pairs)
    og = _assert_same_game(b, lts, pairs)
    _assert_same_energies(b, og)
    b2 = GameBuild(lts, pairs[:40], max_positions=3000)
    og2 = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs[:40], max_positions=3000,
                        positions_wanted=False)
    h = b2._download()
    assert bool(b2.build_stats["truncated"]) == og2["truncated"] and b2.n_positions() == og2["V"]
    for k in ("row", "col", "wt"):
        assert np.array_equal(h[k], og2[k]), k


GEN_CASES = [   # (states, transitions, labels, seed, pairs: "part" = enabledness partitions or a random count, cap)
    (60, 200, 4, 3, "part", 0),
    (300, 1200, 8, 1, "part", 0),
    (40, 480, 3, 5, 10, 0),             # 12 transitions per state, 3 labels: 238,657 positions
    (100, 1300, 70, 4, 40, 0),          # labels past 63: no label bit sets
    (150, 2000, 20, 9, 30, 200000),     # capped at 200,000 positions
]


@pytest.mark.parametrize("gen,heavy_min,label_mask", [("thread", "1024", "1"), ("sub", "1024", "1"), ("sub", "0", "1"),
                                                      ("sub", "1024", "0"), ("wave", "1024", "1"), ("wave", "1024", "0"),
                                                      ("block", "1024", "1")])
@pytest.mark.parametrize("case", range(len(GEN_CASES)))
def test_generators_match_oracle(monkeypatch, gen, heavy_min, label_mask, case):
    # Positions are generated by one thread each (GEQ_BUILD_GEN=thread) or by
    # eight lanes each (the default), positions past the eight-lane limits by
    # a wave or a block each (heavy_min 0: by the eight-lane kernel's first
    # lane); wave / block send every attack position that fits to the wave /
    # block kernel.  The dense LTSs give states with more transitions than the
    # lanes hold in registers, sets past the eight-lane limits and label groups
    # large enough for the wave's bitonic sort.  Same game and energies every
    # way; capped builds equal the oracle's capped walk.
    n, m, labels, seed, pmode, cap = GEN_CASES[case]
    print("generator", gen, "heavy_min", heavy_min, "label_mask", label_mask, "lts", n, m, labels, seed, pmode, cap,
          flush=True)
    monkeypatch.setenv("GEQ_BUILD_GEN", gen)
    monkeypatch.setenv("GEQ_BUILD_HEAVY_MIN", heavy_min)
    monkeypatch.setenv("GEQ_BUILD_LABEL_MASK", label_mask)
    n, src, dst, lbl = workloads.random_lts(n, m, labels, seed)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    if pmode == "part":
        if n >= 300:
            lts, _ = lts.bisimilar_minimize()
        pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    else:
        rng = np.random.default_rng(seed)
        pairs = [(int(p), int(q)) for p, q in rng.integers(0, n, (4 * pmode, 2)) if p != q][:pmode]
    if cap == 0:
        b = GameBuild(lts, pairs)
        og = _assert_same_game(b, lts, pairs)
        _assert_same_energies(b, og)
        pairs, cap = pairs[:40], 3000
    b2 = GameBuild(lts, pairs, max_positions=cap)
    og2 = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=cap,
                        positions_wanted=False)
    h = b2._download()
    assert bool(b2.build_stats["truncated"]) == og2["truncated"] and b2.n_positions() == og2["V"]
    for k in ("row", "col", "wt", "atk"):
        assert np.array_equal(h[k], og2[k]), k


@pytest.mark.parametrize("tail_max,tail_gen", [("0", "128"), ("100", "512"), ("4096", "0"), ("4096", "16"), ("4096", "128"),
                                               ("4096", "512")])
def test_one_block_chunks_match_oracle(monkeypatch, tail_max, tail_gen):
    # Chunks whose frontier, candidates and next bounds fit GEQ_BUILD_TAIL_MAX
    # items finish in one block (0: none; 100: the small levels of this build;
    # 4096, the default: most of them), which then builds the following levels
    # of at most GEQ_BUILD_TAIL_GEN positions itself (0: none; 128 the
    # default; 512 the most).  Same game and energies every way, a capped
    # build (whose cap stops the block's levels) included.
    print("one-block chunks up to", tail_max, "levels generated in the block up to", tail_gen, flush=True)
    monkeypatch.setenv("GEQ_BUILD_TAIL_MAX", tail_max)
    monkeypatch.setenv("GEQ_BUILD_TAIL_GEN", tail_gen)
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    lts, _ = lts.bisimilar_minimize()
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    b = GameBuild(lts, pairs)
    og = _assert_same_game(b, lts, pairs)
    _assert_same_energies(b, og)
    b2 = GameBuild(lts, pairs[:40], max_positions=3000)
    og2 = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs[:40], max_positions=3000,
                        positions_wanted=False)
    h = b2._download()
    assert bool(b2.build_stats["truncated"]) == og2["truncated"] and b2.n_positions() == og2["V"]
    for k in ("row", "col", "wt"):
        assert np.array_equal(h[k], og2[k]), k


@pytest.mark.parametrize("gen,heavy_min,label_mask", [("sub", "1024", "1"), ("sub", "1024", "0"), ("thread", "1024", "1"),
                                                      ("thread", "1", "1"), ("wave", "1024", "1")])
def test_large_lts_capped_build_matches_oracle(monkeypatch, gen, heavy_min, label_mask):
    # An LTS too large for k_generate's LDS copy: its walks read global memory,
    # where the per-state label bit sets apply (label_mask 0 turns them off).
    # A capped build of 300 pairs equals the oracle's capped breadth-first walk,
    # whichever kernels generate the positions.
    print("large lts, capped build, generator", gen, "heavy_min", heavy_min, "label_mask", label_mask, flush=True)
    monkeypatch.setenv("GEQ_BUILD_GEN", gen)
    monkeypatch.setenv("GEQ_BUILD_HEAVY_MIN", heavy_min)
    monkeypatch.setenv("GEQ_BUILD_LABEL_MASK", label_mask)
    n, src, dst, lbl = workloads.random_lts(3000, 12000, 4, 7)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    rng = np.random.default_rng(7)
    pairs = [(int(p), int(q)) for p, q in rng.integers(0, n, (600, 2)) if p != q][:300]
    cap = 200000
    b = GameBuild(lts, pairs, max_positions=cap)
    og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=cap,
                       positions_wanted=False)
    h = b._download()
    assert og["truncated"] and bool(b.build_stats["truncated"]) and b.n_positions() == og["V"]
    for k in ("row", "col", "wt"):
        assert np.array_equal(h[k], og[k]), k


def test_vlts_peterson_matches_oracle():
    lts = TransitionSystem.from_csv_file(os.path.join(GOLDEN, "peterson_mutex_weak.csv"))
    lts, _ = lts.bisimilar_minimize()
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    b = GameBuild(lts, pairs)
    og = _assert_same_game(b, lts, pairs)
    _assert_same_energies(b, og)


def test_duplicate_start_pairs_keep_the_last_id():
    # GameBuild::new_node maps a repeated start position to its last id
    lts = TransitionSystem.from_edges([(0, 1, 0), (1, 0, 0), (0, 2, 1), (2, 2, 0), (1, 2, 1)])
    pairs = [(0, 1), (1, 0), (0, 1), (2, 1)]
    b = GameBuild(lts, pairs)
    _assert_same_game(b, lts, pairs)


def test_repeated_start_pair_survives_table_growth():
    # The table of a game with thousands of positions grows (and is rebuilt)
    # several times; the earlier copy of a repeated start pair stays out of it,
    # so its successors' edges lead to the last copy's id, as in the oracle.
    n, src, dst, lbl = workloads.random_lts(60, 200, 4, 3)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
    pairs = [pairs[1], pairs[0]] + pairs
    print("repeated start pairs, table growth:", len(pairs), "pairs", flush=True)
    b = GameBuild(lts, pairs)
    assert b.n_positions() > 2048
    _assert_same_game(b, lts, pairs)


@pytest.mark.parametrize("mode", ["1", "3"])
def test_built_games_reuse_freed_buffers(monkeypatch, mode):
    # A freed built game's device buffers serve the next built game that fits
    # in them: bigger, smaller and equal games in a row each give the oracle's
    # game and energies, also when the reused buffers still hold a larger game
    # (and, in solve mode 3, the previous game's bounds and seed grouping).
    print("built games in a row (buffer reuse), solve mode", mode, flush=True)
    monkeypatch.setenv("GEQ_SOLVE_MODE", mode)
    shapes = [(60, 200, 4, 3), (12, 30, 2, 1), (60, 200, 4, 3), (30, 90, 3, 2), (60, 200, 4, 3), (12, 30, 2, 1)]
    for n, m, labels, seed in shapes:
        n2, src, dst, lbl = workloads.random_lts(n, m, labels, seed)
        lts = TransitionSystem.from_arrays(n2, src, dst, lbl)
        pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q]
        b = GameBuild(lts, pairs)
        og = _assert_same_game(b, lts, pairs)
        _assert_same_energies(b, og)
        b.close()


def test_truncated_build_is_well_formed():
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q][:50]
    b = GameBuild(lts, pairs, max_positions=2000)
    assert b.build_stats["truncated"] == 1
    h = b._download()
    V = b.n_positions()
    assert V > 2000 and h["row"][-1] == len(h["col"]) and np.all(np.diff(h["row"].astype(np.int64)) >= 0)
    assert h["col"].max() < V
    # the same cut as the oracle's capped breadth-first walk: identical game
    og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs, max_positions=2000,
                       positions_wanted=False)
    assert og["truncated"] and og["V"] == V
    for k in ("row", "col", "wt"):
        assert np.array_equal(h[k], og[k]), k


def test_truncated_build_refuses_a_complete_solve():
    # A capped build is not the complete game: its last level's positions were
    # never expanded.  geq_built_solve refuses it; the capped solve leaves that
    # frontier out of the positions to reach and equals the oracle's solve of
    # the cut game with the same seeds.
    from gpuequiv_amd._lib import Error
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q][:50]
    b = GameBuild(lts, pairs, max_positions=2000)
    assert b.truncated()
    front = b.frontier()
    assert 0 < front < b.n_positions()
    with pytest.raises(Error) as ei:
        b.solve()
    assert ei.value.code == 4
    h = b._download()
    dead = [v for v in range(b.n_positions()) if not h["atk"][v] and h["row"][v] == h["row"][v + 1]]
    reach = [v for v in dead if v < front]
    assert len(reach) < len(dead)          # the frontier holds "dead ends" that are only unexpanded
    offs, data, st = orc.solve(6, 3, h["row"], h["col"], h["wt"], h["atk"], reach)
    s = b.solve(capped=True)
    o, d = b.energies_raw(0, b.n_positions())
    assert H.canon_lists(o, d) == H.canon_lists(offs, data)
    assert s["rounds"] == st["rounds"] and s["edge_updates"] == st["edge_updates"]
    b.close()
    # a complete build: the capped solve is the ordinary one
    b = GameBuild(lts, pairs[:3])
    assert not b.truncated() and b.frontier() == b.n_positions()
    og = orc.gamebuild(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels, pairs[:3], positions_wanted=False)
    reach = [v for v in range(og["V"]) if not og["atk"][v] and og["row"][v] == og["row"][v + 1]]
    offs, data, _ = orc.solve(6, 3, og["row"], og["col"], og["wt"], og["atk"], reach)
    b.solve(capped=True)
    o, d = b.energies_raw(0, og["V"])
    assert H.canon_lists(o, d) == H.canon_lists(offs, data)
    b.close()


def test_full_spectroscopy_verdict_matrix():
    """300 states, 1,200 transitions, 8 labels: the 300 x 300 x 13 preorder
    matrix of equivalences() equals the oracle's."""
    print("spectroscopy 300/1200/8 seed 1", flush=True)
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    want = H.oracle_verdicts(n, lts.row_offsets, lts.targets, lts.labels)
    eq = lts.equivalences()
    eqs = [getattr(std_equivalences, nm)() for nm in H.EQ_NAMES]
    got = np.zeros_like(want)
    part = eq.minimization
    seeded = set(eq.pos_to_idx)
    for p in range(n):
        for q in range(n):
            if part[p] == part[q] or (part[p], part[q]) in seeded:
                got[p, q] = [eq.preorder(p, q, e) for e in eqs]
    assert np.array_equal(got, want)
    assert np.array_equal(eq.preorder_matrix(eqs), want)
    assert want.sum() > n * 13          # verdicts beyond the diagonal


def test_device_verdicts_equal_the_oracle_matrix():
    """The starting points' 13 verdicts computed on the device
    (geq_built_preorders), assembled into the 300 x 300 x 13 matrix, equal the
    oracle's; per point they equal test_equivalence on the downloaded budgets,
    also over a sub-range; energies not packed for the conf are refused."""
    from gpuequiv_amd._lib import Error
    from gpuequiv_amd.equivalence import preorder_matrix_of
    print("device verdicts, spectroscopy 300/1200/8 seed 1", flush=True)
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    want = H.oracle_verdicts(n, lts.row_offsets, lts.targets, lts.labels)
    red, part = lts.bisimilar_minimize()
    b, info = GameBuild.compare_all(red)
    b.solve()
    eqs = [getattr(std_equivalences, nm)() for nm in H.EQ_NAMES]
    bits = b.preorder_bits(eqs)
    assert bits.shape == (b.n_starting_points, len(eqs))
    assert np.array_equal(preorder_matrix_of(info, part, bits), want)
    ens = b.energies(0, b.n_starting_points)
    for i, a in enumerate(ens):
        assert [a.test_equivalence(e) for e in eqs] == bits[i].tolist(), i
    assert np.array_equal(b.preorder_bits(eqs, first=3, count=5), bits[3:8])
    assert 0 < bits.sum() < bits.size      # both verdicts occur

    class Bad:
        def raw_data(self):
            return (0xFFFFFFFF,)
    with pytest.raises(Error):
        b.preorder_bits([Bad()])


def test_level_past_the_scan_limit_is_refused(monkeypatch):
    # A level whose candidate or set-pool-word bounds add up past 32 bits is
    # refused with OUT_OF_MEMORY before any candidate is written (the bound
    # scan saturates instead of wrapping).  GEQ_BUILD_LEVEL_LIMIT lowers that
    # limit so a small game reaches it; the context builds normally afterwards.
    from gpuequiv_amd._lib import Error
    print("level past the bound-scan limit", flush=True)
    n, src, dst, lbl = workloads.random_lts(300, 1200, 8, 1)
    lts = TransitionSystem.from_arrays(n, src, dst, lbl)
    pairs = [(p, q) for part in lts.enabledness_partitions() for p in part for q in part if p != q][:50]
    for limit in ("40", "600"):          # refused at the first level / at a later one
        monkeypatch.setenv("GEQ_BUILD_LEVEL_LIMIT", limit)
        with pytest.raises(Error) as ei:
            GameBuild(lts, pairs)
        assert ei.value.code == 5 and "2^32" in str(ei.value)
    monkeypatch.delenv("GEQ_BUILD_LEVEL_LIMIT")
    b = GameBuild(lts, pairs[:5])
    og = _assert_same_game(b, lts, pairs[:5])
    _assert_same_energies(b, og)
    b.close()
