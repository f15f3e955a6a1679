"""GameGraph and EnergyGame: the public entry to the solver.

Host-side mirror of ``src/energygame.rs`` (Gobbel2000/gpuequiv v1.0.0).  The
CSR layout and the constructors follow the crate; ``EnergyGame.run`` calls the
HIP fixpoint through the C ABI (``geq_solve``) in place of the crate's wgpu
``GPURunner`` (src/energygame.rs:1430-1524).
"""
from __future__ import annotations

import ctypes as C
import json
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from .energy import EnergyArray, EnergyConf, Update, UpdateArray, update


class GameGraph:
    """CSR game graph (src/energygame.rs:112-250)."""

    def __init__(self, conf: EnergyConf, row_offsets, column_indices, weights, attacker_pos,
                 rev_row_offsets=None, rev_column_indices=None):
        self.conf = conf
        self.row_offsets = np.ascontiguousarray(row_offsets, dtype=np.uint32)
        self.column_indices = np.ascontiguousarray(column_indices, dtype=np.uint32)
        w = np.ascontiguousarray(weights, dtype=np.uint32)
        self.weights = UpdateArray(w.reshape(-1, conf.update_size()), conf)
        self.attacker_pos = np.ascontiguousarray(attacker_pos, dtype=np.uint8)
        self._rev_row = None if rev_row_offsets is None else np.ascontiguousarray(rev_row_offsets, dtype=np.uint32)
        self._rev_col = None if rev_column_indices is None else np.ascontiguousarray(rev_column_indices, dtype=np.uint32)

    # the reverse CSR is derived on first use (the C ABI derives it itself when absent)
    @property
    def rev_row_offsets(self) -> np.ndarray:
        if self._rev_row is None:
            self.make_reverse()
        return self._rev_row

    @property
    def rev_column_indices(self) -> np.ndarray:
        if self._rev_col is None:
            self.make_reverse()
        return self._rev_col

    @classmethod
    def new(cls, n_vertices: int, edges: Iterable, attacker_pos: Sequence[bool], conf: EnergyConf) -> "GameGraph":
        """``GameGraph::new(n, edges, attacker_pos, conf)`` (:141-194): edges are
        (from, to, weight) with weight an update list (ints / Upd) or an Update."""
        edges = list(edges)
        adj = [[] for _ in range(n_vertices)]
        for (a, b, w) in edges:
            adj[int(a)].append((int(b), w))
        row = np.zeros(n_vertices + 1, dtype=np.uint32)
        col, wts = [], []
        for v in range(n_vertices):
            for (b, w) in adj[v]:
                col.append(b)
                if isinstance(w, Update):
                    wts.append(list(w.data))
                else:
                    u = Update.new(list(w), conf)
                    if u is None:
                        raise ValueError("Updates don't match configuration")
                    wts.append(list(u.data))
            row[v + 1] = len(col)
        warr = np.array(wts, dtype=np.uint32).reshape(len(col), conf.update_size())
        return cls(conf, row, np.array(col, dtype=np.uint32), warr, np.array(attacker_pos, dtype=np.uint8))

    @classmethod
    def from_csr(cls, conf, row_offsets, column_indices, weights, attacker_pos) -> "GameGraph":
        return cls(conf, row_offsets, column_indices, weights, attacker_pos)

    @classmethod
    def empty(cls, conf: EnergyConf) -> "GameGraph":
        return cls(conf, np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros((0, conf.update_size()), np.uint32),
                   np.zeros(0, np.uint8))

    def n_vertices(self) -> int:
        return int(self.attacker_pos.shape[0])

    def n_edges(self) -> int:
        return int(self.column_indices.shape[0])

    def get_conf(self) -> EnergyConf:
        return self.conf

    def adj(self, v: int) -> np.ndarray:
        return self.column_indices[self.row_offsets[v]:self.row_offsets[v + 1]]

    def reverse(self, v: int) -> np.ndarray:
        return self.rev_column_indices[self.rev_row_offsets[v]:self.rev_row_offsets[v + 1]]

    def make_reverse(self) -> None:
        """Predecessor CSR in increasing source order (:235-250)."""
        V = self.n_vertices()
        src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(self.row_offsets.astype(np.int64)))
        order = np.argsort(self.column_indices, kind="stable")
        self._rev_col = src[order].astype(np.uint32)
        counts = np.bincount(self.column_indices, minlength=V) if V else np.zeros(0, np.int64)
        self._rev_row = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)

    # serde form (SerdeGameGraph, :34-109) as JSON
    def to_json(self) -> str:
        adj, weights = [], []
        for v in range(self.n_vertices()):
            a, b = int(self.row_offsets[v]), int(self.row_offsets[v + 1])
            adj.append([int(x) for x in self.column_indices[a:b]])
            weights.append([Update(self.weights.array[k], self.conf).to_ints() for k in range(a, b)])
        return json.dumps({"conf": {"elements": self.conf.elements, "max": self.conf.max}, "adj": adj,
                           "weights": weights, "attacker_pos": [bool(x) for x in self.attacker_pos]})

    @classmethod
    def from_json(cls, text: str) -> "GameGraph":
        d = json.loads(text)
        conf = EnergyConf(int(d["conf"]["elements"]), int(d["conf"]["max"]))
        n = len(d["adj"])
        if len(d["weights"]) != n:
            raise ValueError("Weight list size doesn't match adjacecy list size")
        if len(d["attacker_pos"]) != n:
            raise ValueError("attacker_pos list has wrong length")
        edges = []
        for v, (succ, ws) in enumerate(zip(d["adj"], d["weights"])):
            if len(succ) != len(ws):
                raise ValueError("Out-degree in adjacency list doesn't match out-degree in weight list")
            if not all(0 <= s < n for s in succ):
                raise ValueError("Node index too high")
            edges.extend((v, s, update(*w)) for s, w in zip(succ, ws))
        return cls.new(n, edges, d["attacker_pos"], conf)

    def _c_struct(self, keep: list) -> _lib.geq_graph:
        g = _lib.geq_graph()
        g.conf = _lib.geq_conf(self.conf.elements, self.conf.max)
        g.n_vertices = self.n_vertices()
        g.n_edges = self.n_edges()
        w = np.ascontiguousarray(self.weights.array.reshape(-1))
        keep.extend([w])
        g.row_offsets = _lib.u32p(self.row_offsets)
        g.column_indices = _lib.u32p(self.column_indices)
        g.weights = _lib.u32p(w)
        g.attacker_pos = _lib.u8p(self.attacker_pos)
        if self._rev_row is not None and self._rev_col is not None:
            g.rev_row_offsets = _lib.u32p(self._rev_row)
            g.rev_column_indices = _lib.u32p(self._rev_col)
        return g


def energies_from_flat(offs: np.ndarray, data: np.ndarray, conf: EnergyConf) -> list:
    esz = conf.energy_size()
    arr = data.reshape(-1, esz) if esz else data.reshape(-1, 1)
    return [EnergyArray(arr[offs[v]:offs[v + 1]], conf) for v in range(len(offs) - 1)]


def flat_from_energies(energies: Sequence[EnergyArray], conf: EnergyConf):
    offs = np.zeros(len(energies) + 1, dtype=np.uint32)
    parts = []
    for v, e in enumerate(energies):
        offs[v + 1] = offs[v] + e.n_energies()
        parts.append(e.array.reshape(-1))
    data = np.concatenate(parts).astype(np.uint32) if parts else np.zeros(0, np.uint32)
    return offs, data


class EnergyGame:
    """Energy game: graph + per-node winning budgets + nodes to reach (:326-384)."""

    def __init__(self, graph: GameGraph, energies: list, to_reach: Sequence[int]):
        self.graph = graph
        self.energies = energies
        self.to_reach = [int(v) for v in to_reach]
        self._initial_default = True
        self.last_stats: dict | None = None

    @classmethod
    def standard_reach(cls, graph: GameGraph) -> "EnergyGame":
        """Attacker wants to reach the defender dead ends (:347-352)."""
        deg = np.diff(graph.row_offsets.astype(np.int64))
        to_reach = np.nonzero((graph.attacker_pos == 0) & (deg == 0))[0]
        return cls.with_reach(graph, to_reach.tolist())

    @classmethod
    def with_reach(cls, graph: GameGraph, to_reach: Sequence[int]) -> "EnergyGame":
        conf = graph.get_conf()
        energies = [EnergyArray.empty(conf) for _ in range(graph.n_vertices())]
        for v in to_reach:
            energies[int(v)] = EnergyArray.zero(1, conf)
        return cls(graph, energies, to_reach)

    def run(self) -> list:
        """Solve on the GPU; fills and returns ``self.energies`` (:379-383)."""
        ctx = _lib.Context.get()
        keep: list = []
        g = self.graph._c_struct(keep)
        reach = np.ascontiguousarray(self.to_reach, dtype=np.uint32)
        offs, data = flat_from_energies(self.energies, self.graph.conf)
        init = _lib.geq_energies(len(self.energies), self.graph.conf.energy_size(), _lib.u32p(offs),
                                 _lib.u32p(data if data.size else np.zeros(1, np.uint32)))
        out = _lib.geq_energies()
        st = _lib.geq_stats()
        _lib.check(_lib.lib().geq_solve(ctx.handle, C.byref(g), _lib.u32p(reach), len(reach), C.byref(init),
                                        C.byref(out), C.byref(st)))
        try:
            o, d = _lib.energies_to_numpy(out)
        finally:
            _lib.lib().geq_energies_free(C.byref(out))
        self.energies = energies_from_flat(o, d, self.graph.conf)
        self.last_stats = st.as_dict()
        return self.energies
