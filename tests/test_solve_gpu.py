"""GPU parity of the fixpoint solver (EnergyGame.run through the C ABI) against
the reference's golden games, the kernel fixtures and the oracle."""
import numpy as np
import pytest

from gpuequiv_amd import EnergyArray, EnergyConf, EnergyGame, GameGraph, Update, Upd, update
from gpuequiv_amd import workloads
from gpuequiv_amd.energygame import energies_from_flat
from tests import helpers as H

pytestmark = pytest.mark.gpu
STD = EnergyConf.STANDARD


def _run(graph, reach=None, energies=None):
    game = EnergyGame.standard_reach(graph) if reach is None else EnergyGame.with_reach(graph, reach)
    if energies is not None:
        game.energies = energies
    out = game.run()
    return game, out


def _canon(arrays):
    return [tuple(sorted(int(x) for x in a.array.reshape(-1))) for a in arrays]


# ---------------------------------------------------------------- tests/energygame.rs
@pytest.mark.parametrize("name", ["simple", "simple_one_dimension", "fork", "combinations"])
def test_golden_games(name):
    d = next(g for g in H.load("energygame_games.json")["games"] if g["name"] == name)
    print("golden game", name, flush=True)
    _, out = _run(H.fixture_game(d))
    assert out == H.expected_arrays(d["expected"])


def test_wide_game():
    d = H.load("energygame_games.json")["wide_game"]
    graph, expected = H.wide(d["single"], d["double"])
    _, out = _run(graph)
    assert out == H.expected_arrays(expected)


def test_simple_one_dimension_by_set_weights():
    # tests/energygame.rs:73-104 rewrites the weights of the built graph in place
    d = next(g for g in H.load("energygame_games.json")["games"] if g["name"] == "simple")
    exp = next(g for g in H.load("energygame_games.json")["games"] if g["name"] == "simple_one_dimension")
    graph = H.fixture_game(d)
    for i in range(graph.n_edges()):
        graph.weights.set(i, Update.new([Upd.Decrement], STD))
    _, out = _run(graph)
    assert out == H.expected_arrays(exp["expected"])


# ---------------------------------------------------------------- src/energygame/tests.rs
def _defender_over(lists, weights, conf=STD):
    """Defender 0 with one attacker successor per list; the successors hold the
    lists as their energies and are the nodes to reach, so node 0 is visited."""
    k = len(lists)
    edges = [(0, i + 1, w) for i, w in enumerate(weights)]
    graph = GameGraph.new(k + 1, edges, [False] + [True] * k, conf)
    energies = [EnergyArray.empty(conf)] + [EnergyArray.from_conf(l, conf) for l in lists]
    game = EnergyGame.with_reach(graph, list(range(1, k + 1)))
    game.energies = energies
    return game.run()[0]


def test_defend_four_successors():
    fx = H.load("kernel_fixtures.json")["four_successors"]
    got = _defender_over(fx["successors"], [update(*w) for w in fx["weights"]])
    assert got == EnergyArray.from_conf(fx["expected"], STD)


def test_defend_antichain_141():
    fx = H.load("kernel_fixtures.json")["antichain"]["energies"]
    got = _defender_over([fx] * 4, [update(0)] * 4)
    assert got.n_energies() == 141
    assert got == EnergyArray.from_conf(fx, STD)


def test_defend_mixed_nodes():
    # src/energygame/tests.rs:345-483: three defenders solved in the same launch
    fx = H.load("kernel_fixtures.json")
    ac = fx["antichain"]["energies"]
    ac2 = [[v + 1 for v in e] for e in ac]
    four = fx["four_successors"]
    conf = STD
    lists = [ac] * 4 + [ac2] * 4 + four["successors"]
    n = 3 + len(lists)
    edges = []
    heads = [0, 1, 2]
    ws = [[update(0)] * 4, [update(0)] * 4, [update(*w) for w in four["weights"]]]
    s = 3
    for h, wl in zip(heads, ws):
        for w in wl:
            edges.append((h, s, w))
            s += 1
    atk = [False, False, False] + [True] * len(lists)
    graph = GameGraph.new(n, edges, atk, conf)
    game = EnergyGame.with_reach(graph, list(range(3, n)))
    game.energies = [EnergyArray.empty(conf)] * 3 + [EnergyArray.from_conf(l, conf) for l in lists]
    out = game.run()
    assert out[0] == EnergyArray.from_conf(ac, conf)
    assert out[1] == EnergyArray.from_conf(ac2, conf)
    assert out[2] == EnergyArray.from_conf(four["expected"], conf)


# ---------------------------------------------------------------- parity with the oracle
@pytest.mark.parametrize("V,seed", [(50, 1), (200, 2), (1000, 3), (5000, 4), (20000, 5), (100000, 42)])
def test_random_games_match_oracle(V, seed):
    g = workloads.random_game(V, seed)
    print("random game", V, seed, flush=True)
    reach = H.standard_reach(g)
    offs, data, st = H.oracle_solve_game(g, reach)
    game, out = _run(g)
    assert _canon(out) == H.canon_lists(offs, data)
    s = game.last_stats
    # the same Jacobi rounds: identical schedule counters
    assert s["rounds"] == st["rounds"]
    assert s["node_evals"] == st["node_evals"]
    assert s["edge_updates"] == st["edge_updates"]
    assert s["alg_bytes"] == st["alg_bytes"]


@pytest.mark.parametrize("seed", [7, 8])
def test_random_games_varied_shapes(seed):
    # dense, decrement-heavy games build long antichains (arena lists > 7)
    g = workloads.random_game(3000, seed, mean_degree=8.0, p_ident=0.2, p_dec=0.6)
    reach = H.standard_reach(g)
    offs, data, _ = H.oracle_solve_game(g, reach)
    _, out = _run(g)
    assert _canon(out) == H.canon_lists(offs, data)


@pytest.mark.parametrize("V,seed", [(300, 11), (3000, 12), (20000, 13)])
def test_arbitrary_updates_match_oracle(V, seed):
    # every field of every update random: several Decrements, chained and
    # self-referencing Mins in one update (the inverse update's field order matters)
    g = workloads.random_game(V, seed, any_updates=True)
    print("any-update game", V, seed, flush=True)
    reach = H.standard_reach(g)
    offs, data, st = H.oracle_solve_game(g, reach)
    game, out = _run(g)
    assert _canon(out) == H.canon_lists(offs, data)
    assert game.last_stats["rounds"] == st["rounds"] and game.last_stats["edge_updates"] == st["edge_updates"]


def test_initial_energies_and_empty_reach():
    d = next(g for g in H.load("energygame_games.json")["games"] if g["name"] == "combinations")
    graph = H.fixture_game(d)
    game = EnergyGame.with_reach(graph, [])
    out = game.run()
    assert all(a.is_empty() for a in out)


def test_small_confs_match_oracle():
    # EnergyConf{elements: 4, max: 2}: 2-bit fields, saturation at 2
    conf = EnergyConf(4, 2)
    rng = np.random.default_rng(3)
    V = 400
    edges = []
    for v in range(V):
        for _ in range(rng.integers(0, 5)):
            i = int(rng.integers(0, 4))
            u = [0] * 4
            r = rng.random()
            if r < 0.4:
                u[i] = -1
            elif r < 0.6:
                u[i] = int(rng.integers(1, 5))
            edges.append((v, int(rng.integers(0, V)), update(*u)))
    atk = (rng.random(V) < 0.5).tolist()
    g = GameGraph.new(V, edges, atk, conf)
    reach = H.standard_reach(g)
    offs, data, _ = H.oracle_solve_game(g, reach)
    _, out = _run(g)
    assert _canon(out) == H.canon_lists(offs, data)


def _solve_raw(graph, profile, **tune):
    import ctypes as C
    from gpuequiv_amd import _lib
    lib = _lib.lib()
    ctx = _lib.Context.get()
    keep = []
    gs = graph._c_struct(keep)
    h = C.c_void_p()
    _lib.check(lib.geq_game_upload(ctx.handle, C.byref(gs), C.byref(h)))
    try:
        if tune:
            p = _lib.geq_tune_params()
            _lib.check(lib.geq_game_get_tune(h, C.byref(p)))
            for k, v in tune.items():
                setattr(p, k, int(v))
            _lib.check(lib.geq_game_set_tune(h, C.byref(p)))
        reach = np.ascontiguousarray(H.standard_reach(graph), dtype=np.uint32)
        st = _lib.geq_stats()
        _lib.check(lib.geq_game_solve(h, _lib.u32p(reach), len(reach), None, profile, C.byref(st)))
        out = _lib.geq_energies()
        _lib.check(lib.geq_game_energies(h, C.byref(out)))
        try:
            offs, data = _lib.energies_to_numpy(out)
        finally:
            lib.geq_energies_free(C.byref(out))
        return H.canon_lists(offs, data), st.as_dict()
    finally:
        lib.geq_game_free(h)


def test_standard_reach_solve_matches_explicit_reach():
    # geq_game_solve_standard seeds with the defender dead ends found on the
    # device (once per game); the result equals the explicit standard_reach
    # list's, on repeated solves of the same resident game and with profiling.
    import ctypes as C
    from gpuequiv_amd import _lib
    lib = _lib.lib()
    ctx = _lib.Context.get()
    for V, seed in ((3000, 4), (50000, 9)):
        graph = workloads.random_game(V, seed)
        print("standard reach solve", V, seed, flush=True)
        want, sw = _solve_raw(graph, 0)
        keep = []
        gs = graph._c_struct(keep)
        h = C.c_void_p()
        _lib.check(lib.geq_game_upload(ctx.handle, C.byref(gs), C.byref(h)))
        try:
            for profile in (0, 0, 1):
                st = _lib.geq_stats()
                _lib.check(lib.geq_game_solve_standard(h, profile, C.byref(st)))
                out = _lib.geq_energies()
                _lib.check(lib.geq_game_energies(h, C.byref(out)))
                try:
                    offs, data = _lib.energies_to_numpy(out)
                finally:
                    lib.geq_energies_free(C.byref(out))
                assert H.canon_lists(offs, data) == want
                s = st.as_dict()
                assert (s["rounds"], s["edge_updates"], s["node_evals"]) == (sw["rounds"], sw["edge_updates"],
                                                                            sw["node_evals"])
        finally:
            lib.geq_game_free(h)


def test_profiled_launches_match_graph_replay():
    g = workloads.random_game(20000, 17)
    a, sa = _solve_raw(g, 0)
    b, sb = _solve_raw(g, 1)
    assert a == b
    assert sa["rounds"] == sb["rounds"] and sa["edge_updates"] == sb["edge_updates"]
    assert sb["eval_kernel_ms"] > 0


def test_arena_regrow_gives_identical_results(monkeypatch):
    # a tiny energy arena forces k_eval to overflow and the round to be re-run
    g = workloads.random_game(3000, 8, mean_degree=8.0, p_ident=0.2, p_dec=0.6)
    ref, sref = _solve_raw(g, 0)
    monkeypatch.setenv("GEQ_ARENA_ENTRIES", "64")
    small, ssmall = _solve_raw(g, 0)
    assert small == ref
    assert ssmall["rounds"] == sref["rounds"]


def test_all_launch_modes_agree(monkeypatch):
    # hipGraph of captured rounds (default) and per-kernel launches
    g = workloads.random_game(30000, 19)
    a, sa = _solve_raw(g, 0)
    assert sa["mode"] == 1
    c, sc = _solve_raw(g, 1)
    assert sc["mode"] == 2
    assert a == c
    assert sa["rounds"] == sc["rounds"] and sa["node_evals"] == sc["node_evals"] and sa["alg_bytes"] == sc["alg_bytes"]


@pytest.mark.parametrize("rounds", [1, 5])
def test_fused_expansion_matches_separate_kernel(rounds):
    # expand_blocks_per_cu = 0: k_eval queues the predecessors itself
    g = workloads.random_game(20000, 23)
    a, sa = _solve_raw(g, 0, expand_blocks_per_cu=4, rounds_per_launch=rounds)
    b, sb = _solve_raw(g, 0, expand_blocks_per_cu=0, rounds_per_launch=rounds)
    assert a == b
    assert (sa["rounds"], sa["node_evals"], sa["edge_updates"], sa["alg_bytes"]) == \
        (sb["rounds"], sb["node_evals"], sb["edge_updates"], sb["alg_bytes"])
    c, sc = _solve_raw(g, 1, expand_blocks_per_cu=0)
    assert c == a and sc["rounds"] == sa["rounds"] and sc["eval_kernel_ms"] > 0
    o, d, so = H.oracle_solve_game(g, H.standard_reach(g))
    assert a == H.canon_lists(o, d) and sa["rounds"] == so["rounds"]


@pytest.mark.parametrize("expand_bpc", [4, 0])
def test_arena_regrow_fused_and_separate(monkeypatch, expand_bpc):
    # an overflow in the middle of a batch of rounds: the interrupted round runs
    # again from its start (its half-written slots and queued predecessors are dropped)
    g = workloads.random_game(3000, 8, mean_degree=8.0, p_ident=0.2, p_dec=0.6)
    ref, sref = _solve_raw(g, 0)
    monkeypatch.setenv("GEQ_ARENA_ENTRIES", "64")
    small, ssmall = _solve_raw(g, 0, expand_blocks_per_cu=expand_bpc, rounds_per_launch=3)
    assert small == ref
    # the interrupted round's evaluations ran and are counted too
    assert ssmall["rounds"] == sref["rounds"] and ssmall["node_evals"] >= sref["node_evals"]


@pytest.mark.parametrize("expand_bpc,rounds", [(4, 1), (0, 1), (0, 2)])
def test_slot_retag_keeps_results(monkeypatch, expand_bpc, rounds):
    # re-tag every node's two list slots after every batch of rounds (normally
    # every 16384 rounds): which slot is current must survive it
    g = workloads.random_game(8000, 31, mean_degree=6.0)
    ref, sref = _solve_raw(g, 0)
    monkeypatch.setenv("GEQ_RETAG_ROUNDS", "1")
    got, sgot = _solve_raw(g, 0, expand_blocks_per_cu=expand_bpc, rounds_per_launch=rounds)
    assert got == ref and sgot["rounds"] == sref["rounds"]


def test_concurrent_host_threads_share_the_context():
    # the crate's types are Send + Sync (Cargo.toml:18-21): games solved from
    # several host threads at once through one context give their own results
    import threading
    graphs = [workloads.random_game(v, 500 + i) for i, v in enumerate((8000, 300, 15000, 2000, 9000, 50))]
    want = [_canon(_run(g)[1]) for g in graphs]
    got = [None] * len(graphs)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                got[i] = _canon(_run(graphs[i])[1])
        except Exception as e:   # surfaced below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(graphs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not errors, errors
    assert got == want


def test_json_example_game_matches_oracle():
    # the game of examples/json.rs, written as SerdeGameGraph JSON and read back
    from tests.test_json import json_example_graph
    g = GameGraph.from_json(json_example_graph().to_json())
    offs, data, st = H.oracle_solve_game(g, H.standard_reach(g))
    game, out = _run(g)
    assert _canon(out) == H.canon_lists(offs, data)
    assert game.last_stats["rounds"] == st["rounds"]


@pytest.mark.parametrize("order,renumber", [("1", "1"), ("0", "0"), ("1", "0")])
def test_worklist_order_and_renumbering_keep_results(monkeypatch, order, renumber):
    # Rounds listed in id order from the queued bitmap (GEQ_WORKLIST_ORDER=1,
    # the default from 2^20 positions) or in append order (0), positions
    # renumbered in reverse breadth-first order at upload (GEQ_RENUMBER=1) or
    # not (the default): the same sets, and the same Jacobi schedule counters.
    monkeypatch.setenv("GEQ_WORKLIST_ORDER", order)
    monkeypatch.setenv("GEQ_RENUMBER", renumber)
    for V, seed in [(5000, 21), (30000, 22)]:
        print("order", order, "renumber", renumber, "game", V, seed, flush=True)
        g = workloads.random_game(V, seed)
        offs, data, st = H.oracle_solve_game(g, H.standard_reach(g))
        game, out = _run(g)
        assert _canon(out) == H.canon_lists(offs, data)
        for k in ("rounds", "node_evals", "edge_updates", "alg_bytes"):
            assert game.last_stats[k] == st[k], k
