"""GPU parity of the list kernels: EnergyGame.run on energy configurations
outside the bitmap layout (elements > 6 or max > 3), as the reference's
EnergyGame::run accepts any EnergyConf (src/energy.rs:36-78, shadergen.rs:159-189),
against the oracle."""
import numpy as np
import pytest

from gpuequiv_amd import EnergyArray, EnergyConf, EnergyGame, GameGraph, workloads
from gpuequiv_amd._lib import Error
from gpuequiv_amd.batch import GameBatch
from oracle import oracle as orc
from tests import helpers as H

pytestmark = pytest.mark.gpu

CONFS = [EnergyConf.SILENT_STEP, EnergyConf.ONE_DIMENSION, EnergyConf(20, 3), EnergyConf(6, 100), EnergyConf(12, 15),
         EnergyConf(4, 0xFFFFFFFF), EnergyConf(10, 70000), EnergyConf(30, 3)]


def _rows(arr: EnergyArray) -> list:
    return sorted(tuple(int(x) for x in r) for r in arr.array)


def _oracle_rows(offs, data, esz, v):
    return sorted(tuple(int(x) for x in data[(int(offs[v]) + i) * esz:(int(offs[v]) + i + 1) * esz])
                  for i in range(int(offs[v + 1] - offs[v])))


def _check(g, game, out, offs, data, st):
    esz = g.conf.energy_size()
    for v in range(g.n_vertices()):
        assert _rows(out[v]) == _oracle_rows(offs, data, esz, v), v
    s = game.last_stats
    assert s["rounds"] == st["rounds"]
    assert s["node_evals"] == st["node_evals"]
    assert s["edge_updates"] == st["edge_updates"]
    assert s["alg_bytes"] == st["alg_bytes"]


@pytest.mark.parametrize("conf", CONFS, ids=lambda c: f"{c.elements}x{c.max}")
@pytest.mark.parametrize("V,seed,anyu", [(60, 1, False), (3000, 2, False), (2000, 3, True)])
def test_random_games_match_oracle(conf, V, seed, anyu):
    print("list kernels", conf, V, seed, anyu, flush=True)
    g = workloads.random_game_conf(V, seed, conf, any_updates=anyu)
    offs, data, st = H.oracle_solve_game(g, H.standard_reach(g))
    game = EnergyGame.standard_reach(g)
    out = game.run()
    _check(g, game, out, offs, data, st)


@pytest.mark.parametrize("conf", [EnergyConf.SILENT_STEP, EnergyConf(6, 100)], ids=lambda c: f"{c.elements}x{c.max}")
def test_long_antichains_and_scratch_regrowth(conf, monkeypatch):
    # dense, decrement-heavy games: long antichains, defender folds whose
    # pairwise suprema outgrow a wave's first scratch region, and (with a tiny
    # arena) arena regrowth; both are recovered by running the round again
    print("long antichains", conf, flush=True)
    monkeypatch.setenv("GEQ_ARENA_ENTRIES", "256")
    g = workloads.random_game_conf(2500, 9, conf, mean_degree=8.0, p_ident=0.1, p_dec=0.7)
    offs, data, st = H.oracle_solve_game(g, H.standard_reach(g))
    lens = np.diff(offs.astype(np.int64))
    assert lens.max() >= 12
    game = EnergyGame.standard_reach(g)
    out = game.run()
    _check(g, game, out, offs, data, st)


def test_defender_fold_needs_more_scratch():
    # one defender over two attackers holding 150 incomparable SILENT_STEP
    # energies each: 22,500 suprema in one fold step (more than the 16,384
    # words a wave starts with)
    conf = EnergyConf.SILENT_STEP
    rng = np.random.default_rng(4)
    lay = set()
    while len(lay) < 150:                 # points of the layer with component sum 8: an antichain
        v = rng.integers(0, 3, 8)
        if v.sum() == 8:
            lay.add(tuple(int(x) for x in v))
    lay = sorted(lay)
    a = EnergyArray.from_conf(lay, conf)
    b = EnergyArray.from_conf([tuple(reversed(x)) for x in lay], conf)
    graph = GameGraph.new(3, [(0, 1, [0]), (0, 2, [-1])], [False, True, True], conf)
    game = EnergyGame.with_reach(graph, [1, 2])
    game.energies = [EnergyArray.empty(conf), a, b]
    out = game.run()
    io, idt = [0, 0, len(lay), 2 * len(lay)], np.concatenate([a.array.reshape(-1), b.array.reshape(-1)])
    offs, data, st = H.oracle_solve_game(graph, [1, 2], init=(np.array(io, np.uint32), idt))
    assert _rows(out[0]) == _oracle_rows(offs, data, 1, 0)
    assert out[0].n_energies() > 1


def test_explicit_initial_energies_and_one_dimension_wrap():
    # ONE_DIMENSION: a position at u32::MAX behind a Decrement wraps to 0
    # before the min with max (u32 arithmetic, shadergen.rs:103-112)
    conf = EnergyConf.ONE_DIMENSION
    graph = GameGraph.new(4, [(0, 1, [-1]), (0, 2, [-1]), (3, 1, [0])], [True, True, True, True], conf)
    game = EnergyGame.with_reach(graph, [1, 2])
    game.energies = [EnergyArray.empty(conf), EnergyArray.from_conf([[0xFFFFFFFF]], conf),
                     EnergyArray.from_conf([[5]], conf), EnergyArray.empty(conf)]
    out = game.run()
    assert _rows(out[0]) == [(0,)] and _rows(out[3]) == [(0xFFFFFFFF,)]
    io = np.array([0, 0, 1, 2, 2], np.uint32)
    offs, data, _ = H.oracle_solve_game(graph, [1, 2], init=(io, np.array([0xFFFFFFFF, 5], np.uint32)))
    for v in range(4):
        assert _rows(out[v]) == _oracle_rows(offs, data, 1, v)


def test_embedded_golden_games():
    # the reference's golden games (tests/energygame.rs) with their components
    # renamed into EnergyConf{20, 3} across its word boundary
    from tests.test_oracle import _embed_energy, _embed_game
    conf = EnergyConf(20, 3)
    comps = [13, 14, 15, 16, 17, 4]
    for d in H.load("energygame_games.json")["games"]:
        g = H.fixture_game(d)
        out = EnergyGame.standard_reach(_embed_game(g, conf, comps)).run()
        for v, exp in enumerate(d["expected"]):
            want = sorted(_embed_energy(orc.pack_energy(6, 3, e)[0], conf, comps) for e in exp)
            assert _rows(out[v]) == want, (d["name"], v)


def test_batch_of_generic_games():
    conf = EnergyConf.SILENT_STEP
    games = [workloads.random_game_conf(int(v), 100 + i, conf) for i, v in enumerate([40, 700, 1500, 90, 2600])]
    b = GameBatch(games)
    b.solve()
    outs = b.energies()
    for g, out in zip(games, outs):
        offs, data, _ = H.oracle_solve_game(g, H.standard_reach(g))
        for v in range(g.n_vertices()):
            assert _rows(out[v]) == _oracle_rows(offs, data, 1, v)
    b.close()


def test_unsupported_and_invalid_confs():
    # energies of more than 32 words are refused; so is an update naming a
    # component past `elements`
    g = GameGraph.new(2, [(0, 1, [-1])], [True, False], EnergyConf(40, 0xFFFFFFFF))
    with pytest.raises(Error) as ei:
        EnergyGame.standard_reach(g).run()
    assert ei.value.code == 6
    conf = EnergyConf.SILENT_STEP
    bad = GameGraph(conf, np.array([0, 1, 1], np.uint32), np.array([1], np.uint32),
                    np.array([[11 << 0]], np.uint32), np.array([1, 0], np.uint8))   # Min(10) with 8 elements
    with pytest.raises(Error) as ei:
        EnergyGame.standard_reach(bad).run()
    assert ei.value.code == 4
