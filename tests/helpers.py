"""Shared test helpers: fixtures, canonical forms, oracle pipelines."""
from __future__ import annotations

import json
import os

import numpy as np

from gpuequiv_amd import EnergyArray, EnergyConf, EnergyGame, GameGraph, update
from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STD = EnergyConf.STANDARD


def load(name: str) -> dict:
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def fixture_game(d: dict) -> GameGraph:
    conf = EnergyConf(*d["conf"])
    atk = [i in set(d["attacker"]) for i in range(d["n"])]
    edges = [(a, b, update(*w)) for a, b, w in d["edges"]]
    return GameGraph.new(d["n"], edges, atk, conf)


def wide(single: int, double: int):
    """tests/energygame.rs:209-237: the game and its expected per-node lists."""
    n = 2 * single + 3 * double + 1
    atk = [False] * n
    expected = [[[0]] for _ in range(n)]
    expected[0] = [[1, 1, 2], [1, 1, 1, 1]]
    edges = []
    for i in range(1, single + 1):
        edges.append((0, i, update(2)))
        edges.append((i, i + single, update(-1)))
        atk[i] = True
        expected[i] = [[1]]
    for i in range(1, double + 1):
        edges.append((0, 2 * single + i, update(0, 0, -1)))
        edges.append((2 * single + i, 2 * single + double + i, update(0, 0, -1)))
        edges.append((2 * single + i, 2 * single + double + i + 1, update(0, 0, 0, -1)))
        atk[2 * single + i] = True
        expected[2 * single + i] = [[0, 0, 1], [0, 0, 0, 1]]
    return GameGraph.new(n, edges, atk, STD), expected


def expected_arrays(expected: list, conf: EnergyConf = STD) -> list:
    return [EnergyArray.from_conf(rows, conf) for rows in expected]


def canon_lists(offs, data) -> list:
    """Per node: sorted tuple of packed words (multiset)."""
    return [tuple(sorted(int(x) for x in data[offs[v]:offs[v + 1]])) for v in range(len(offs) - 1)]


def oracle_solve_game(graph: GameGraph, reach, init=None, emulate_mem=False):
    c = graph.conf
    if init is None:
        offs, data, st = orc.solve(c.elements, c.max, graph.row_offsets, graph.column_indices,
                                   graph.weights.array.reshape(-1), graph.attacker_pos, reach, emulate_mem=emulate_mem)
    else:
        io, idt = init
        offs, data, st = orc.solve(c.elements, c.max, graph.row_offsets, graph.column_indices,
                                   graph.weights.array.reshape(-1), graph.attacker_pos, reach, io, idt,
                                   emulate_mem=emulate_mem)
    return offs, data, st


def standard_reach(graph: GameGraph) -> list:
    deg = np.diff(graph.row_offsets.astype(np.int64))
    return np.nonzero((graph.attacker_pos == 0) & (deg == 0))[0].tolist()


def lts_csr(n: int, transitions):
    """Edge list -> label-sorted CSR (TransitionSystem::from, src/lib.rs:508-535)."""
    adj = [[] for _ in range(n)]
    for a, b, l in transitions:
        adj[int(a)].append((int(b), int(l)))
    row = [0]
    dst, lbl = [], []
    for r in adj:
        r = sorted(r, key=lambda t: t[1])    # stable
        for b, l in r:
            dst.append(b)
            lbl.append(l)
        row.append(len(dst))
    return np.array(row, np.uint32), np.array(dst, np.uint32), np.array(lbl, np.int32)


def n_states_of(transitions) -> int:
    return max(max(a, b) for a, b, _ in transitions) + 1


def oracle_compare(transitions, p: int, q: int, minimize: bool = True):
    """TransitionSystem::compare (src/lib.rs:275-296) on the oracle."""
    n = n_states_of(transitions)
    row, dst, lbl = lts_csr(n, transitions)
    if minimize:
        part, cnt, row, dst, lbl = orc.bisimilar_minimize(n, row, dst, lbl)
        p, q = int(part[p]), int(part[q])
        if p == q:
            return []
        n = cnt
    g = orc.gamebuild(n, row, dst, lbl, [(p, q)])
    reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
    offs, data, _ = orc.solve(6, 3, g["row"], g["col"], g["wt"], g["atk"], reach)
    return [int(x) for x in data[offs[0]:offs[1]]]


def energies_le(e: int, eq: list) -> bool:
    return all(((e >> (2 * i)) & 3) <= eq[i] for i in range(6))


def holds_equivalence(energies: list, eq: list) -> bool:
    """EnergyArray::test_equivalence on packed STANDARD words (src/energy.rs:307-318)."""
    return not any(energies_le(e, eq) for e in energies)


EQ_NAMES = ("enabledness", "traces", "failures", "failure_traces", "readiness", "readiness_traces", "revivals",
            "impossible_futures", "possible_futures", "simulation", "ready_simulation", "nested_2simulation",
            "bisimulation")


def oracle_verdicts(n_states, lrow, ldst, llbl):
    """The full spectroscopy of an LTS on the oracle: minimise, build the game
    of every pair in the same enabledness class, solve, and decide the 13
    preorders for every ordered pair (src/lib.rs:329-343,
    src/equivalence.rs:125-132).  Returns a bool array [n, n, 13]."""
    from gpuequiv_amd.lts import TransitionSystem
    from gpuequiv_amd.energy import std_equivalences
    part, cnt, orow, odst, olbl = orc.bisimilar_minimize(n_states, lrow, ldst, llbl)
    red = TransitionSystem(orow, odst, olbl)
    pairs = [(p, q) for grp in red.enabledness_partitions() for p in grp for q in grp if p != q]
    g = orc.gamebuild(red.n_vertices(), red.row_offsets, red.targets, red.labels, pairs)
    reach = [v for v in range(g["V"]) if not g["atk"][v] and g["row"][v] == g["row"][v + 1]]
    offs, data, _ = orc.solve(6, 3, g["row"], g["col"], g["wt"], g["atk"], reach)
    eqs = np.array([std_equivalences.get(nm).to_vec() for nm in EQ_NAMES], dtype=np.int64)   # [13, 6]
    red_v = np.zeros((cnt, cnt, 13), dtype=bool)
    for i, (p, q) in enumerate(pairs):
        es = np.array([orc.unpack_energy(6, 3, [int(w)]) for w in data[offs[i]:offs[i + 1]]], dtype=np.int64)
        if es.size == 0:
            red_v[p, q] = True
        else:
            below = np.all(es[:, None, :] <= eqs[None, :, :], axis=2)      # [k, 13]
            red_v[p, q] = ~below.any(axis=0)
    part = np.asarray(part, dtype=np.int64)
    out = red_v[part[:, None], part[None, :]]
    same = part[:, None] == part[None, :]
    out[same] = True
    return out
