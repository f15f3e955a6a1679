"""Host-side LTS handling: CSV ingest, label-sorted rows, bisimulation
minimisation (native code, checked against the oracle).  CPU only."""
import os

import numpy as np
import pytest

from gpuequiv_amd import TransitionSystem, Transition, workloads
from oracle import oracle as orc
from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_from_edges_sorts_rows_stably_by_label():
    lts = TransitionSystem.from_edges([(0, 3, 2), (0, 1, 1), (0, 2, 1), (1, 0, 0)])
    assert lts.adj(0) == [Transition(1, 1), Transition(2, 1), Transition(3, 2)]
    assert lts.adj(1) == [Transition(0, 0)]
    assert lts.n_vertices() == 4 and lts.adj(3) == []


def test_csv_label_numbering():
    # src/lib.rs:372-406: "i" is label 0, then first appearance order, quotes stripped
    lts = TransitionSystem.from_csv_lines(['0,1,"a"', "1,2,i", "2,0,b", '0,2,"i"'])
    assert lts.adj(0) == [Transition(2, 0), Transition(1, 1)]
    assert lts.adj(1) == [Transition(2, 0)]
    assert lts.adj(2) == [Transition(0, 2)]
    with pytest.raises(ValueError):
        TransitionSystem.from_csv_lines(["0,1"])


def test_vlts_csv_matches_oracle_order():
    lts = TransitionSystem.from_csv_file(os.path.join(GOLDEN, "peterson_mutex_weak.csv"))
    assert lts.n_vertices() == 20 and lts.n_transitions() == 138
    for p in range(lts.n_vertices()):
        labels = [t.label for t in lts.adj(p)]
        assert labels == sorted(labels)


def _lts_random(n, m, labels, seed):
    n, src, dst, lbl = workloads.random_lts(n, m, labels, seed)
    return TransitionSystem.from_arrays(n, src, dst, lbl)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bisimulation_matches_oracle(seed):
    lts = _lts_random(200, 500, 3, seed)
    part, cnt = lts.signature_refinement(on_device=False)
    opart, ocnt = orc.signature_refinement(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels)
    assert cnt == ocnt and part == [int(x) for x in opart]
    red, part2 = lts.bisimilar_minimize(on_device=False)
    op, oc, orow, odst, olbl = orc.bisimilar_minimize(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels)
    assert part2 == [int(x) for x in op]
    assert np.array_equal(red.row_offsets, orow) and np.array_equal(red.targets, odst) and np.array_equal(red.labels, olbl)


def test_bisimilar_golden():
    d = H.load("complete_lts.json")["bisimilar"]
    lts = TransitionSystem.from_edges(d["transitions"])
    part, _ = lts.signature_refinement(on_device=False)
    assert part[0] == part[5]


def test_bisimulation_on_vlts():
    lts = TransitionSystem.from_csv_file(os.path.join(GOLDEN, "vasy_0_1.csv"))
    part, cnt = lts.signature_refinement(on_device=False)
    opart, ocnt = orc.signature_refinement(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels)
    assert cnt == ocnt and part == [int(x) for x in opart]


def test_preorder_matrix_matches_preorder_calls():
    # host-side only: an Equivalence over hand-made budgets
    from gpuequiv_amd import EnergyArray, EnergyConf, std_equivalences
    from gpuequiv_amd.equivalence import Equivalence, StartInfo
    from gpuequiv_amd.gamebuild import AttackPosition
    STD = EnergyConf.STANDARD
    pairs = [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)]
    budgets = [[[1, 1]], [[3, 1]], [], [[2, 2, 0, 0, 1, 1], [0, 0, 0, 0, 3]], [[3, 3, 3, 3, 3, 3]], [[1]]]
    info = StartInfo([AttackPosition(p, (q,)) for p, q in pairs], 4)
    eq = Equivalence(info, [EnergyArray.from_conf(b, STD) for b in budgets])
    eq.set_minimization([0, 1, 2, 2])
    eqs = [getattr(std_equivalences, nm)() for nm in std_equivalences.NAMES]
    mat = eq.preorder_matrix(eqs)
    for p in range(4):
        for q in range(4):
            for k, e in enumerate(eqs):
                assert mat[p, q, k] == eq.preorder(p, q, e), (p, q, k)


def test_preorder_matrix_random_budgets_and_repeated_pairs():
    # the vectorised matrix against per-pair preorder() calls: random budgets
    # (some pairs with none), a minimisation that merges states, and a pair
    # seeded twice (its last entry counts, as in the pair lookup)
    import numpy as np
    from gpuequiv_amd import EnergyArray, EnergyConf, std_equivalences
    from gpuequiv_amd.equivalence import Equivalence, StartInfo
    from gpuequiv_amd.gamebuild import AttackPosition
    STD = EnergyConf.STANDARD
    rng = np.random.default_rng(11)
    m = 40
    pairs = [(int(p), int(q)) for p, q in rng.integers(0, m, (300, 2)) if p != q]
    pairs.append(pairs[5])
    budgets = [rng.integers(0, 4, (int(rng.integers(0, 6)), 6)).tolist() for _ in pairs]
    info = StartInfo([AttackPosition(p, (q,)) for p, q in pairs], 50)
    eq = Equivalence(info, [EnergyArray.from_conf(b, STD) for b in budgets])
    # every reduced state is the image of some state (as for a real partition)
    eq.set_minimization(list(range(m)) + [int(x) for x in rng.integers(0, m, 50 - m)])
    eqs = [getattr(std_equivalences, nm)() for nm in std_equivalences.NAMES]
    mat = eq.preorder_matrix(eqs)
    assert mat.shape == (50, 50, len(eqs))
    for p in range(50):
        for q in range(50):
            for k, e in enumerate(eqs):
                assert mat[p, q, k] == eq.preorder(p, q, e), (p, q, k)


def test_preorder_matrix_from_verdicts():
    # preorder_matrix_of (the assembly the device verdicts take) given each
    # starting point's test_equivalence results equals preorder_matrix, with
    # empty budgets, a merging minimisation and a pair seeded twice
    import numpy as np
    from gpuequiv_amd import EnergyArray, EnergyConf, std_equivalences
    from gpuequiv_amd.equivalence import Equivalence, StartInfo, preorder_matrix_of
    from gpuequiv_amd.gamebuild import AttackPosition
    STD = EnergyConf.STANDARD
    rng = np.random.default_rng(12)
    m = 30
    pairs = [(int(p), int(q)) for p, q in rng.integers(0, m, (200, 2)) if p != q]
    pairs.append(pairs[3])
    budgets = [rng.integers(0, 4, (int(rng.integers(0, 5)), 6)).tolist() for _ in pairs]
    info = StartInfo([AttackPosition(p, (q,)) for p, q in pairs], 36)
    arrays = [EnergyArray.from_conf(b, STD) for b in budgets]
    eq = Equivalence(info, arrays)
    part = list(range(m)) + [int(x) for x in rng.integers(0, m, 36 - m)]
    eq.set_minimization(part)
    eqs = [getattr(std_equivalences, nm)() for nm in std_equivalences.NAMES]
    holds = np.array([[a.test_equivalence(e) for e in eqs] for a in arrays], dtype=bool)
    assert np.array_equal(preorder_matrix_of(info, part, holds), eq.preorder_matrix(eqs))
    with pytest.raises(ValueError):
        preorder_matrix_of(info, part, holds[:-1])
