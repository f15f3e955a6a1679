"""TransitionSystem: the crate's top-level API (src/lib.rs:169-536).

The LTS is kept as a label-sorted CSR (rows stably sorted by label, as
TransitionSystem::from does, src/lib.rs:489-493, :508-535).  compare /
equivalences build the spectroscopy game on the GPU (geq_gamebuild), solve it
there (geq_built_solve) and read the starting positions' budgets back.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from .energy import EnergyArray, EnergyConf

STD = EnergyConf.STANDARD


@dataclass(frozen=True)
class Transition:
    """A labelled transition to `process` (src/lib.rs:169-179)."""

    process: int
    label: int


class TransitionSystem:
    def __init__(self, row_offsets, targets, labels):
        self.row_offsets = np.ascontiguousarray(row_offsets, dtype=np.uint32)
        self.targets = np.ascontiguousarray(targets, dtype=np.uint32)
        self.labels = np.ascontiguousarray(labels, dtype=np.int32)

    # -------------------------------------------------------------- construction
    @classmethod
    def from_adj_unchecked(cls, adj: Sequence[Sequence]) -> "TransitionSystem":
        """Adjacency lists already sorted by label (src/lib.rs:480-487)."""
        row = [0]
        dst, lbl = [], []
        for r in adj:
            for t in r:
                p, l = (t.process, t.label) if isinstance(t, Transition) else t
                dst.append(int(p))
                lbl.append(int(l))
            row.append(len(dst))
        return cls(row, dst, lbl)

    @classmethod
    def from_adj(cls, adj: Sequence[Sequence]) -> "TransitionSystem":
        """From adjacency lists, stably sorted by label (From<Vec<Vec<Transition>>>)."""
        rows = []
        for r in adj:
            pairs = [(t.process, t.label) if isinstance(t, Transition) else tuple(t) for t in r]
            rows.append(sorted(pairs, key=lambda x: x[1]))
        return cls.from_adj_unchecked(rows)

    @classmethod
    def from_edges(cls, edges: Iterable) -> "TransitionSystem":
        """From (from, to, label) triples (From<Vec<Edge>>, src/lib.rs:517-535)."""
        e = np.asarray(list(edges), dtype=np.int64).reshape(-1, 3)
        n = int(e[:, :2].max()) + 1 if len(e) else 0
        return cls.from_arrays(n, e[:, 0], e[:, 1], e[:, 2])

    @classmethod
    def from_arrays(cls, n_states: int, src, dst, lbl) -> "TransitionSystem":
        """Vectorised form of from_edges: rows keep the input order, then a
        stable sort by label."""
        src = np.asarray(src, dtype=np.int64)
        dst = np.asarray(dst, dtype=np.uint32)
        lbl = np.asarray(lbl, dtype=np.int32)
        order = np.lexsort((lbl, src))     # by source, then label; stable for ties
        counts = np.bincount(src, minlength=n_states) if len(src) else np.zeros(n_states, np.int64)
        row = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        return cls(row, dst[order], lbl[order])

    @classmethod
    def from_csv_lines(cls, lines: Iterable[str]) -> "TransitionSystem":
        """`from,to,label` lines; label "i" is 0, other labels get 1, 2, ... in
        order of first appearance; quotes are stripped (src/lib.rs:372-406)."""
        labels = {"i": 0}
        src, dst, lbl = [], [], []
        for line in lines:
            line = line.rstrip("\n").rstrip("\r")
            parts = line.split(",", 2)
            if len(parts) < 3:
                raise ValueError("insufficient number of values in line")
            a, b, name = int(parts[0]), int(parts[1]), parts[2]
            if len(name) >= 2 and name.startswith('"') and name.endswith('"'):
                name = name[1:-1]
            if name not in labels:
                labels[name] = len(labels)
            src.append(a)
            dst.append(b)
            lbl.append(labels[name])
        n = max(max(src, default=-1), max(dst, default=-1)) + 1
        return cls.from_arrays(n, src, dst, lbl)

    @classmethod
    def from_csv_file(cls, path: str) -> "TransitionSystem":
        with open(path) as f:
            return cls.from_csv_lines(f)

    # -------------------------------------------------------------- access
    def n_vertices(self) -> int:
        return int(self.row_offsets.shape[0] - 1)

    def n_transitions(self) -> int:
        return int(self.targets.shape[0])

    def adj(self, p: int) -> list:
        a, b = int(self.row_offsets[p]), int(self.row_offsets[p + 1])
        return [Transition(int(self.targets[k]), int(self.labels[k])) for k in range(a, b)]

    def _c_struct(self) -> _lib.geq_lts:
        s = _lib.geq_lts()
        s.n_states = self.n_vertices()
        s.n_transitions = self.n_transitions()
        s.row_offsets = _lib.u32p(self.row_offsets)
        self._t = self.targets if self.targets.size else np.zeros(1, np.uint32)
        self._l = self.labels if self.labels.size else np.zeros(1, np.int32)
        s.targets = _lib.u32p(self._t)
        s.labels = _lib.i32p(self._l)
        return s

    # -------------------------------------------------------------- bisimulation
    def signature_refinement(self, on_device: bool = True):
        """(partition, class count) of strong bisimilarity (src/lib.rs:444-478),
        refined on the GPU (geq_signature_refinement_gpu) or, with
        on_device=False, by the library's host routine."""
        n = self.n_vertices()
        part = np.zeros(max(n, 1), np.uint32)
        cnt = C.c_uint32()
        s = self._c_struct()
        if on_device:
            _lib.check(_lib.lib().geq_signature_refinement_gpu(_lib.Context.get().handle, C.byref(s), _lib.u32p(part),
                                                               C.byref(cnt), None))
        else:
            _lib.check(_lib.lib().geq_signature_refinement(C.byref(s), _lib.u32p(part), C.byref(cnt)))
        return part[:n].tolist(), int(cnt.value)

    def bisimilar_minimize(self, on_device: bool = True):
        """(quotient LTS, partition) (src/lib.rs:413-432); the partition is
        refined on the GPU unless on_device=False."""
        n, m = self.n_vertices(), self.n_transitions()
        part = np.zeros(max(n, 1), np.uint32)
        cnt = C.c_uint32()
        orow = np.zeros(n + 1, np.uint32)
        odst = np.zeros(max(m, 1), np.uint32)
        olbl = np.zeros(max(m, 1), np.int32)
        om = C.c_uint32()
        s = self._c_struct()
        if on_device:
            _lib.check(_lib.lib().geq_bisimilar_minimize_gpu(_lib.Context.get().handle, C.byref(s), _lib.u32p(part),
                                                             C.byref(cnt), _lib.u32p(orow), _lib.u32p(odst),
                                                             _lib.i32p(olbl), C.byref(om)))
        else:
            _lib.check(_lib.lib().geq_bisimilar_minimize(C.byref(s), _lib.u32p(part), C.byref(cnt), _lib.u32p(orow),
                                                         _lib.u32p(odst), _lib.i32p(olbl), C.byref(om)))
        c, k = int(cnt.value), int(om.value)
        return TransitionSystem(orow[:c + 1], odst[:k], olbl[:k]), part[:n].tolist()

    # -------------------------------------------------------------- spectroscopy
    def enabledness_partitions(self, processes: Sequence[int] | None = None) -> list:
        """Processes grouped by their set of enabled labels (src/gamebuild.rs:104-133).
        Groups are ordered by their first member (the reference iterates a hash map)."""
        members, sizes = self._enabledness_groups(processes)
        out, k = [], 0
        ml = members.tolist()
        for sz in sizes.tolist():
            out.append(ml[k:k + sz])
            k += sz
        return out

    def _enabledness_groups(self, processes: Sequence[int] | None = None):
        """enabledness_partitions as arrays: the members, group after group, and
        the group sizes."""
        import numpy as np
        n = self.n_vertices()
        procs = np.arange(n, dtype=np.int64) if processes is None else np.asarray(list(processes), dtype=np.int64)
        lab = np.asarray(self.labels)
        if procs.size and (lab.size == 0 or (lab.min() >= 0 and lab.max() < 64)):
            # the label set of a state as a bit mask; groups in order of their
            # first member, members in process order (as the dict below)
            row = np.asarray(self.row_offsets, dtype=np.int64)
            mask = np.zeros(n, dtype=np.uint64)
            if lab.size:
                deg = np.diff(row)
                full = deg > 0   # rows with transitions: their segments of `bits` in order
                bits = np.left_shift(np.uint64(1), lab.astype(np.uint64))
                mask[full] = np.bitwise_or.reduceat(bits, row[:-1][full])
            _, first, inv = np.unique(mask[procs], return_index=True, return_inverse=True)
            rank = np.empty(len(first), dtype=np.int64)
            rank[np.argsort(first, kind="stable")] = np.arange(len(first))
            g = rank[inv.reshape(-1)]
            return procs[np.argsort(g, kind="stable")], np.bincount(g, minlength=len(first))
        groups: dict = {}
        for p in procs.tolist():
            a, b = int(self.row_offsets[p]), int(self.row_offsets[p + 1])
            key = tuple(sorted(set(self.labels[a:b].tolist())))
            groups.setdefault(key, []).append(int(p))
        parts = list(groups.values())
        members = np.array([p for part in parts for p in part], dtype=np.int64)
        return members, np.array([len(part) for part in parts], dtype=np.int64)

    def compare(self, p: int, q: int) -> EnergyArray:
        """Budgets of the attack position (p, {q}) after bisimulation
        minimisation; empty if p and q are bisimilar (src/lib.rs:275-283)."""
        reduced, part = self.bisimilar_minimize()
        pm, qm = part[p], part[q]
        if pm == qm:
            return EnergyArray.empty(STD)
        return reduced.compare_unminimized(pm, qm)

    def compare_unminimized(self, p: int, q: int) -> EnergyArray:
        from .gamebuild import GameBuild
        b = GameBuild.compare(self, p, q)
        b.solve()
        return b.energies(0, 1)[0]

    def equivalences(self):
        """All pairs, after bisimulation minimisation (src/lib.rs:329-334)."""
        reduced, part = self.bisimilar_minimize()
        eq = reduced.equivalences_unminimized()
        eq.set_minimization(part)
        return eq

    def equivalences_unminimized(self):
        from .gamebuild import GameBuild
        b, info = GameBuild.compare_all(self)
        b.solve()
        return info.equivalence(b.energies(0, b.n_starting_points))
