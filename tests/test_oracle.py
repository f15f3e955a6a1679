"""Pins the CPU oracle (and the host energy model) against the reference's own
known answers, golden games and an independent up-set fixpoint.  CPU only."""
import numpy as np
import pytest

from gpuequiv_amd import Energy, EnergyArray, EnergyConf, GameGraph, Upd, Update, UpdateArray, std_equivalences, update
from gpuequiv_amd import workloads
from oracle import oracle as orc
from tests import helpers as H
from tests import upsets as U

STD = EnergyConf.STANDARD


# ---------------------------------------------------------------- src/energy.rs:805-937
def test_energy_conf_known_answers():
    # src/energy.rs:810-842
    for (el, mx), exp in [((4, 3), (1, 2, 0x3, 1, 4, 0xF)), ((6, 100), (2, 8, 0xFF, 1, 4, 0xF)),
                          ((30, 3), (2, 2, 0x3, 8, 8, 0xFF)), ((4, 0xFFFFFFFF), (4, 32, 0xFFFFFFFF, 1, 4, 0xF))]:
        assert tuple(orc.conf_info(el, mx)) == exp
        c = EnergyConf(el, mx)
        assert (c.energy_size(), c.energy_bits(), c.energy_mask(), c.update_size(), c.update_bits(),
                c.update_mask()) == exp


def test_energy_packing_known_answers():
    # :846-856
    vals = [0, 1, 2, 3, 4, 5, 6, 7, 8, 13, 14, 15]
    assert orc.pack_energy(12, 15, vals) == [0x76543210, 0xFED8]
    e = Energy.new(vals, EnergyConf(12, 15))
    assert list(e.data) == [0x76543210, 0xFED8] and e.to_vec() == vals
    assert orc.unpack_energy(12, 15, [0x76543210, 0xFED8]) == vals
    # :859-871
    assert orc.pack_energy(6, 3, [0, 1, 2, 0, 3]) == [0x324]
    assert Energy.new([0, 1, 2, 0, 3], STD).to_vec() == [0, 1, 2, 0, 3, 0]
    assert orc.pack_energy(6, 3, [4]) is None and Energy.new([4], STD) is None


def test_energy_partial_order():
    # :874-886
    e0 = Energy.zero(STD)
    e1 = Energy.new([1] * 6, STD)
    e2 = Energy.new([0, 0, 0, 0, 0, 1], STD)
    e3 = Energy.new([1, 0, 0, 0, 0, 0], STD)
    assert e0 < e1 and e0 < e2 and e2 < e1
    assert e2.partial_cmp(e3) is None and e3.partial_cmp(e2) is None
    assert e2.partial_cmp(e2) == 0


def test_update_known_answers():
    # :898-936
    assert Upd.from_int(0) == Upd.Zero and Upd.from_int(-1) == Upd.Decrement
    assert Upd.from_int(8) == Upd.Min(8) and Upd.Min(8).to_int() == 8
    vals = [Upd.Zero, Upd.Decrement] + [Upd.Min(k) for k in (1, 2, 3, 4, 5, 6, 9, 10, 11, 12)]
    u = Update.new(vals, EnergyConf(12, 3))
    assert list(u.data) == [0x76543210, 0xDCBA]
    assert u.to_vec() == vals
    assert u.to_ints() == [0, -1, 1, 2, 3, 4, 5, 6, 9, 10, 11, 12]
    assert orc.pack_update(12, 3, [0, -1, 1, 2, 3, 4, 5, 6, 9, 10, 11, 12]) == [0x76543210, 0xDCBA]
    assert UpdateArray.from_conf([[Upd.Zero, Upd.Decrement, Upd.Min(2)]], STD) == UpdateArray.from_conf([[0, -1, 2]], STD)
    assert Update.new([Upd.Min(7)], STD) is None and orc.pack_update(6, 3, [7]) is None


def test_std_equivalences_table():
    fx = H.load("complete_lts.json")["std_equivalences"]
    for name in std_equivalences.NAMES:
        assert getattr(std_equivalences, name)().to_vec() == fx[name]
    assert std_equivalences.enabledness() == Energy.new([1, 1], STD)


def test_inv_update_closed_form_cases():
    # shadergen.rs:103-128: decrement first (saturating at max), then Min fields in index order
    def inv(vals, upd):
        e = orc.pack_energy(6, 3, vals)
        u = orc.pack_update(6, 3, upd)
        return orc.unpack_energy(6, 3, orc.inv_update(6, 3, e, u))
    assert inv([0, 0, 0, 0, 0, 0], [-1]) == [1, 0, 0, 0, 0, 0]
    assert inv([3, 0, 0, 0, 0, 0], [-1]) == [3, 0, 0, 0, 0, 0]          # saturation at max
    assert inv([0, 2, 0, 0, 0, 0], [0, 1]) == [2, 2, 0, 0, 0, 0]        # Min(1) at 1: e0 = max(e0, e1)
    assert inv([0, 2, 0, 0, 0, 0], [0, 1, 2]) == [2, 2, 0, 0, 0, 0]
    # sequential: Min(2) at 0 then Min(3) at 1 reads the updated e1
    assert inv([3, 0, 1, 0, 0, 0], [2, 3]) == [3, 3, 3, 0, 0, 0]
    # decrement before Min: Min(1) at 1 sees e1 + 1
    assert inv([0, 1, 0, 0, 0, 0], [0, -1]) == [0, 2, 0, 0, 0, 0]
    # Decrement at 0, Min(1) at 1: e0 = max(e0 + 1, e1); Min-first would give e0 = 2
    assert inv([0, 1, 0, 0, 0, 0], [-1, 1]) == [1, 1, 0, 0, 0, 0]
    assert inv([0, 1, 0, 0, 0, 0], [1, -1]) == [0, 2, 0, 0, 0, 0]      # Min(1) at 0 refers to itself


# ---------------------------------------------------------------- tests/energygame.rs
@pytest.mark.parametrize("name", ["simple", "simple_one_dimension", "fork", "combinations"])
def test_oracle_golden_games(name):
    d = next(g for g in H.load("energygame_games.json")["games"] if g["name"] == name)
    graph = H.fixture_game(d)
    offs, data, _ = H.oracle_solve_game(graph, H.standard_reach(graph))
    got = [EnergyArray(data[offs[v]:offs[v + 1]].reshape(-1, 1), STD) for v in range(graph.n_vertices())]
    assert got == H.expected_arrays(d["expected"])


def test_oracle_wide_game():
    d = H.load("energygame_games.json")["wide_game"]
    graph, expected = H.wide(d["single"], d["double"])
    offs, data, _ = H.oracle_solve_game(graph, H.standard_reach(graph))
    got = [EnergyArray(data[offs[v]:offs[v + 1]].reshape(-1, 1), STD) for v in range(graph.n_vertices())]
    assert got == H.expected_arrays(expected)


# ---------------------------------------------------------------- src/energygame/tests.rs
def _packed_lists(lists):
    return [[orc.pack_energy(6, 3, e) for e in lst] for lst in lists]


def test_oracle_defend_four_successors():
    fx = H.load("kernel_fixtures.json")["four_successors"]
    w = [orc.pack_update(6, 3, u) for u in fx["weights"]]
    st, out = orc.defend_fold(6, 3, _packed_lists(fx["successors"]), w, mem=64)
    assert st == 2
    assert sorted(x[0] for x in out) == sorted(orc.pack_energy(6, 3, e)[0] for e in fx["expected"])


def test_oracle_defend_antichain_overflow_statuses():
    fx = H.load("kernel_fixtures.json")["antichain"]
    ac = _packed_lists([fx["energies"]])[0]
    assert len(ac) == 141
    zero = [orc.pack_update(6, 3, [0])] * 4
    exp = fx["overflow_status"]
    assert orc.defend_fold(6, 3, [ac] * 4, zero, mem=64)[0] == exp["mem64"]
    assert orc.defend_fold(6, 3, [ac] * 4, zero, mem=256)[0] == exp["mem256"]
    st, out = orc.defend_fold(6, 3, [ac] * 4, zero, mem=512)
    assert st == exp["mem512"]
    assert sorted(x[0] for x in out) == sorted(x[0] for x in ac)
    # shifted antichain (+1 everywhere, :220-228) folds to itself as well
    ac2 = [[x[0] + 0x555] for x in ac]
    st, out = orc.defend_fold(6, 3, [ac2] * 4, zero, mem=512)
    assert st == 141 and sorted(x[0] for x in out) == sorted(x[0] for x in ac2)


def test_antichain_fixture_is_middle_layer():
    # the 141-element fixture is exactly the rank-6 layer of {0,1,2}^6
    fx = H.load("kernel_fixtures.json")["antichain"]["energies"]
    import itertools
    layer = sorted(list(t) for t in itertools.product(range(3), repeat=6) if sum(t) == 6)
    assert sorted(fx) == layer


# ---------------------------------------------------------------- tests/complete.rs + doctests
def _fig4():
    return H.load("complete_lts.json")["bisping_fig4"]


def test_oracle_gamebuild_bisping_fig4():
    d = _fig4()
    n = H.n_states_of(d["transitions"])
    row, dst, lbl = H.lts_csr(n, d["transitions"])
    g = orc.gamebuild(n, row, dst, lbl, [tuple(d["compare"])])
    exp_pos = [tuple(x if not isinstance(x, list) else list(x) for x in p) for p in d["positions"]]
    assert [tuple(p) for p in g["positions"]] == exp_pos
    for v, adj in enumerate(d["adjacency_sorted"]):
        assert sorted(int(x) for x in g["col"][g["row"][v]:g["row"][v + 1]]) == adj
    reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
    offs, data, _ = orc.solve(6, 3, g["row"], g["col"], g["wt"], g["atk"], reach)
    for v, exp in enumerate(d["energies"]):
        got = EnergyArray(data[offs[v]:offs[v + 1]].reshape(-1, 1), STD)
        assert got == EnergyArray.from_conf(exp, STD), v


def test_oracle_bisimilar():
    d = H.load("complete_lts.json")["bisimilar"]
    n = H.n_states_of(d["transitions"])
    row, dst, lbl = H.lts_csr(n, d["transitions"])
    part, _ = orc.signature_refinement(n, row, dst, lbl)
    assert part[0] == part[5]
    assert H.oracle_compare(d["transitions"], *d["compare"]) == []


def test_oracle_failure_trace_verdicts():
    d = H.load("complete_lts.json")["failure_trace"]
    fx = H.load("complete_lts.json")["std_equivalences"]
    e = H.oracle_compare(d["transitions"], *d["compare"])
    for name in d["holds"]:
        assert H.holds_equivalence(e, fx[name]), name
    for name in d["fails"]:
        assert not H.holds_equivalence(e, fx[name]), name


def test_oracle_doctest_compare():
    d = H.load("complete_lts.json")["doctest"]
    fx = H.load("complete_lts.json")["std_equivalences"]
    e = H.oracle_compare(d["transitions"], 0, 4)
    for name, holds in d["compare_0_4"].items():
        assert H.holds_equivalence(e, fx[name]) == holds


# ---------------------------------------------------------------- independent fixpoint
@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_oracle_matches_upset_fixpoint(seed):
    V = [12, 25, 40, 60, 80, 120][seed - 1]
    g = workloads.random_game(V, seed)
    reach = H.standard_reach(g)
    offs, data, _ = H.oracle_solve_game(g, reach)
    W = U.kleene(g.row_offsets, g.column_indices, g.weights.array.reshape(-1), g.attacker_pos, reach)
    lists = H.canon_lists(offs, data)
    for v in range(V):
        deg = int(g.row_offsets[v + 1] - g.row_offsets[v])
        mask = np.zeros(U.N, dtype=bool)
        mask[list(lists[v])] = True
        if not g.attacker_pos[v] and deg == 1:
            # copied, not minimised: same up-set, and exactly the successor's images
            assert np.array_equal(U.up(mask), W[v])
            s = int(g.column_indices[g.row_offsets[v]])
            t = U.inv_table(int(g.weights.array[g.row_offsets[v], 0]))
            assert lists[v] == tuple(sorted(int(t[x]) for x in lists[s])) or not W[s].any()
        else:
            assert lists[v] == U.as_points(U.minimal(W[v])), v
            assert len(set(lists[v])) == len(lists[v])


def test_upset_fixpoint_pins_golden_games():
    # the independent fixpoint reproduces the reference's expected budgets ...
    for d in H.load("energygame_games.json")["games"]:
        g = H.fixture_game(d)
        W = U.kleene(g.row_offsets, g.column_indices, g.weights.array.reshape(-1), g.attacker_pos, H.standard_reach(g))
        for v, exp in enumerate(d["expected"]):
            pts = tuple(sorted(orc.pack_energy(6, 3, e)[0] for e in exp))
            assert U.as_points(U.minimal(W[v])) == pts, (d["name"], v)


def _golden_upsets_match(variant: str) -> bool:
    games = H.load("energygame_games.json")["games"]
    for d in games:
        g = H.fixture_game(d)
        W = U.kleene(g.row_offsets, g.column_indices, g.weights.array.reshape(-1), g.attacker_pos,
                     H.standard_reach(g), variant=variant)
        for v, exp in enumerate(d["expected"]):
            if U.as_points(U.minimal(W[v])) != tuple(sorted(orc.pack_energy(6, 3, e)[0] for e in exp)):
                return False
    d = _fig4()
    n = H.n_states_of(d["transitions"])
    row, dst, lbl = H.lts_csr(n, d["transitions"])
    g = orc.gamebuild(n, row, dst, lbl, [tuple(d["compare"])])
    reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
    W = U.kleene(g["row"], g["col"], g["wt"], g["atk"], reach, variant=variant)
    for v, exp in enumerate(d["energies"]):
        if U.as_points(U.minimal(W[v])) != tuple(sorted(orc.pack_energy(6, 3, e)[0] for e in exp)):
            return False
    return True


@pytest.mark.parametrize("variant", ["transposed", "no_min"])
def test_planted_mistake_is_caught_by_goldens(variant):
    # the reference's golden games and Bisping's Fig. 8 pin the Min fields of
    # inv_update: each of these slips changes at least one expected budget
    assert _golden_upsets_match("ok")
    assert not _golden_upsets_match(variant)


def test_decrement_before_min_is_invisible_to_goldens():
    # No golden (nor any gamebuild / graphgen update shape) has a Decrement and
    # a Min touching the same component, so the step order of inv_update is
    # pinned only by the WGSL source order (attack_update.wgsl:50-58) through
    # the closed-form cases of test_inv_update_closed_form_cases.
    assert _golden_upsets_match("min_first")
    t_ok = U.inv_table(orc.pack_update(6, 3, [-1, 1])[0])
    t_bad = U.inv_table(orc.pack_update(6, 3, [-1, 1])[0], variant="min_first")
    p = orc.pack_energy(6, 3, [0, 1])[0]
    assert t_ok[p] != t_bad[p]


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_oracle_memory_retry_does_not_change_results(seed):
    # the defender retry protocol (src/energygame.rs:919-923) only defers work
    g = workloads.random_game(3000, seed)
    reach = H.standard_reach(g)
    a = H.oracle_solve_game(g, reach, emulate_mem=False)
    b = H.oracle_solve_game(g, reach, emulate_mem=True)
    assert H.canon_lists(a[0], a[1]) == H.canon_lists(b[0], b[1])


def test_bruteforce_minimal_and_fold_identity():
    # up(min{sup(a,b)}) == up(A) & up(B), and min() of the oracle fold, on random lists
    rng = np.random.default_rng(7)
    zero = [orc.pack_update(6, 3, [0])]
    for _ in range(20):
        A = rng.integers(0, 4096, rng.integers(1, 12))
        B = rng.integers(0, 4096, rng.integers(1, 12))
        st, out = orc.defend_fold(6, 3, [[[int(a)] for a in A], [[int(b)] for b in B]], zero * 2)
        ma = np.zeros(U.N, bool); ma[A] = True
        mb = np.zeros(U.N, bool); mb[B] = True
        inter = U.up(ma) & U.up(mb)
        assert tuple(sorted(x[0] for x in out)) == U.as_points(U.minimal(inter))


def test_oracle_threads_give_identical_results():
    # the oracle's round evaluations spread over threads (the bench's
    # all-cores CPU baseline) compute exactly the sequential results
    from gpuequiv_amd import workloads
    g = workloads.random_game(6000, 77)
    reach = H.standard_reach(g)
    ref = H.oracle_solve_game(g, reach)
    orc.set_threads(4)
    try:
        got = H.oracle_solve_game(g, reach)
    finally:
        orc.set_threads(1)
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    assert ref[2] == got[2]


# ---------------------------------------------------------------- energy confs beyond STANDARD
def _edge_fields(g):
    """Per edge: its update's field values (0 Zero, 1 Decrement, v >= 2 Min(v - 1))."""
    c = g.conf
    return [[0 if u.kind == "zero" else 1 if u.kind == "dec" else u.idx + 1 for u in Update(g.weights.array[k], c).to_vec()]
            for k in range(g.n_edges())]


def _rows(offs, data, esz, v):
    return sorted(tuple(int(x) for x in data[(int(offs[v]) + i) * esz:(int(offs[v]) + i + 1) * esz])
                  for i in range(int(offs[v + 1] - offs[v])))


@pytest.mark.parametrize("elements,mx,V,seed,anyu", [(8, 2, 14, 1, False), (8, 2, 30, 2, True), (2, 7, 40, 3, True),
                                                    (3, 5, 40, 4, False), (3, 5, 30, 5, True)])
def test_oracle_matches_grid_fixpoint_small_confs(elements, mx, V, seed, anyu):
    # SILENT_STEP (src/energy.rs:45) and other small confs: the oracle's lists
    # against a Kleene iteration over every point of {0..max}^elements
    conf = EnergyConf(elements, mx)
    g = workloads.random_game_conf(V, seed, conf, any_updates=anyu)
    reach = H.standard_reach(g)
    offs, data, _ = H.oracle_solve_game(g, reach)
    grid = U.Grid(elements, mx)
    fields = _edge_fields(g)
    W = U.kleene_grid(grid, g.row_offsets, g.column_indices, fields, g.attacker_pos, reach)
    esz = conf.energy_size()
    for v in range(V):
        pts = [grid.index(Energy(r, conf).to_vec()) for r in _rows(offs, data, esz, v)]
        mask = np.zeros(grid.n, dtype=bool)
        mask[pts] = True
        deg = int(g.row_offsets[v + 1] - g.row_offsets[v])
        assert np.array_equal(grid.up(mask), W[v]), v
        if g.attacker_pos[v] or deg != 1:
            assert sorted(pts) == sorted(np.nonzero(grid.minimal(W[v]))[0].tolist()), v


@pytest.mark.parametrize("seed,anyu", [(1, False), (2, True), (3, False)])
def test_oracle_one_dimension_closed_form(seed, anyu):
    # ONE_DIMENSION (src/energy.rs:46): 32-bit fields, no saturation in reach
    conf = EnergyConf.ONE_DIMENSION
    g = workloads.random_game_conf(400, seed, conf, mean_degree=2.0, p_ident=0.1, p_dec=0.9, p_attacker=0.7,
                                   any_updates=anyu)
    reach = H.standard_reach(g)
    offs, data, _ = H.oracle_solve_game(g, reach)
    val = U.one_dimension_values(g.row_offsets, g.column_indices, _edge_fields(g), g.attacker_pos, reach)
    for v in range(g.n_vertices()):
        got = [int(x) for x in data[offs[v]:offs[v + 1]]]
        assert got == ([] if val[v] is None else [val[v]]), v
    if not anyu:
        assert max(x for x in val if x is not None) > 3    # budgets past STANDARD's saturation


def test_oracle_one_dimension_wraps_like_wgsl():
    # inv_update adds the Decrement in u32 arithmetic before the min with max
    # (shadergen.rs:103-112): a field at u32::MAX wraps to 0
    c = EnergyConf.ONE_DIMENSION
    assert orc.inv_update(1, 0xFFFFFFFF, [0xFFFFFFFF], orc.pack_update(1, 0xFFFFFFFF, [-1])) == [0]
    assert orc.inv_update(1, 0xFFFFFFFF, [41], orc.pack_update(1, 0xFFFFFFFF, [-1])) == [42]
    assert orc.inv_update(4, 0xFFFFFFFF, [7, 0xFFFFFFFF, 3, 0], orc.pack_update(4, 0xFFFFFFFF, [0, -1, 1, -1])) == [7, 0, 3, 1]
    assert c.energy_size() == 1 and c.update_bits() == 2


def _embed_game(g, conf, comps):
    """The STANDARD game g with its 6 components placed at `comps` of conf
    (Min indices renamed alike)."""
    fields = np.zeros((g.n_edges(), conf.elements), dtype=np.int64)
    for k, f in enumerate(_edge_fields(g)):
        for i, u in enumerate(f):
            fields[k, comps[i]] = u if u < 2 else comps[u - 2] + 2
    w = workloads.pack_fields(fields, conf.update_bits(), conf.update_size())
    return GameGraph(conf, g.row_offsets, g.column_indices, w, g.attacker_pos)


def _embed_energy(e_std, conf, comps):
    vals = [0] * conf.elements
    for i, x in enumerate(Energy([e_std], EnergyConf.STANDARD).to_vec()):
        vals[comps[i]] = x
    return tuple(Energy.new(vals, conf).data)


@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_multiword_embedding(seed):
    # Components are interchangeable: a STANDARD game (one Min field per
    # update at most, so the field order of inv_update does not matter)
    # renamed into 6 of the 20 components of EnergyConf{20, 3}, straddling its
    # two words, gives the STANDARD lists renamed alike.  This pins the
    # multi-word compares, suprema and updates of the oracle against the
    # single-word STANDARD path and its golden vectors.
    conf = EnergyConf(20, 3)
    comps = [13, 14, 15, 16, 17, 4]
    games = [workloads.random_game(300, seed)] + [H.fixture_game(d) for d in H.load("energygame_games.json")["games"]]
    for g in games:
        reach = H.standard_reach(g)
        o1, d1, s1 = H.oracle_solve_game(g, reach)
        ge = _embed_game(g, conf, comps)
        o2, d2, s2 = H.oracle_solve_game(ge, reach)
        for v in range(g.n_vertices()):
            want = sorted(_embed_energy(int(x), conf, comps) for x in d1[o1[v]:o1[v + 1]])
            assert _rows(o2, d2, 2, v) == want, v
        assert s1["rounds"] == s2["rounds"] and s1["edge_updates"] == s2["edge_updates"]
