"""Batched solving: every game's result equals its standalone (oracle) result."""
import numpy as np
import pytest

from gpuequiv_amd import EnergyGame, workloads
from gpuequiv_amd.batch import GameBatch
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_batch_matches_oracle_per_game():
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(50), np.log(5000), 40)).astype(int)
    graphs = [workloads.random_game(int(v), 1000 + i) for i, v in enumerate(sizes)]
    print("batch of", len(graphs), "games", flush=True)
    b = GameBatch(graphs)
    st = b.solve()
    res = b.raw_energies()
    rounds = []
    for g, (offs, data) in zip(graphs, res):
        o, d, s = H.oracle_solve_game(g, H.standard_reach(g))
        assert H.canon_lists(offs, data) == H.canon_lists(o, d)
        rounds.append(s["rounds"])
    # the union runs as many rounds as its slowest game
    assert st["rounds"] == max(rounds)
    b.close()


def test_batch_equals_plain_path():
    graphs = [workloads.random_game(v, 77 + v) for v in (3000, 100, 20000, 1, 700)]
    b = GameBatch(graphs)
    b.solve()
    batched = b.raw_energies()
    for g, (offs, data) in zip(graphs, batched):
        game = EnergyGame.standard_reach(g)
        out = game.run()
        assert [tuple(sorted(int(x) for x in a.array.reshape(-1))) for a in out] == H.canon_lists(offs, data)
    # solving again from scratch on the resident batch gives the same result
    b.solve()
    assert [H.canon_lists(o, d) for o, d in b.raw_energies()] == [H.canon_lists(o, d) for o, d in batched]
    b.close()
