"""Device signature refinement (geq_signature_refinement_gpu /
geq_bisimilar_minimize_gpu) against the host routine and the oracle: the same
partition (classes numbered by first occurrence, src/lib.rs:444-478) and the
same quotient LTS."""
import os

import numpy as np
import pytest

from gpuequiv_amd import TransitionSystem, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _hubs():
    """States whose transition lists take each sorting path of a refinement
    round: up to 32 keys (one thread), up to 4096 (a block in LDS) and more (a
    block in global memory), with repeated (label, target) pairs to drop."""
    rng = np.random.default_rng(7)
    n = 3000
    fan = {0: 40, 1: 700, 2: 4096, 3: 4097, 4: 6000, 5: 30000, 6: 33}
    src, dst, lbl = [], [], []
    for s in range(n):
        k = fan.get(s, int(rng.integers(0, 5)))
        src += [s] * k
        dst += [int(x) for x in rng.integers(0, n if s > 6 else 200, k)]
        lbl += [int(x) for x in rng.integers(0, 3, k)]
    return TransitionSystem.from_arrays(n, src, dst, lbl)


def _cases():
    yield "empty transitions", TransitionSystem.from_arrays(5, [], [], [])
    yield "one state loop", TransitionSystem.from_edges([(0, 0, 0)])
    yield "aceto exercise", TransitionSystem.from_edges([(0, 1, 0), (0, 2, 0), (1, 3, 0), (1, 4, 1), (2, 4, 0), (3, 0, 0),
                                                         (4, 0, 0), (5, 6, 0), (5, 8, 0), (6, 7, 0), (6, 7, 1), (7, 5, 0),
                                                         (8, 9, 0), (9, 0, 0)])
    for seed, (n, m, l) in enumerate([(40, 80, 2), (300, 1200, 8), (2000, 4000, 3), (20000, 60000, 5)]):
        yield f"random {n}/{m}/{l}", TransitionSystem.from_arrays(*workloads.random_lts(n, m, l, 100 + seed))
    yield "power-law 50000", TransitionSystem.from_arrays(*workloads.powerlaw_lts(50000, 250000, 60, 1))
    yield "hub states", _hubs()
    for f in ("peterson_mutex_weak.csv", "vasy_0_1.csv"):
        yield f, TransitionSystem.from_csv_file(os.path.join(GOLDEN, f))


@pytest.mark.parametrize("name,lts", list(_cases()))
def test_device_refinement_matches_host_and_oracle(name, lts):
    print("bisimulation", name, lts.n_vertices(), lts.n_transitions(), flush=True)
    part_d, cnt_d = lts.signature_refinement()
    part_h, cnt_h = lts.signature_refinement(on_device=False)
    assert (part_d, cnt_d) == (part_h, cnt_h)
    if lts.n_vertices() <= 20000:
        opart, ocnt = orc.signature_refinement(lts.n_vertices(), lts.row_offsets, lts.targets, lts.labels)
        assert part_d == [int(x) for x in opart] and cnt_d == ocnt
    red_d, p_d = lts.bisimilar_minimize()
    red_h, p_h = lts.bisimilar_minimize(on_device=False)
    assert p_d == p_h
    assert np.array_equal(red_d.row_offsets, red_h.row_offsets)
    assert np.array_equal(red_d.targets, red_h.targets) and np.array_equal(red_d.labels, red_h.labels)
