"""GameBuild and the game positions (src/gamebuild.rs, src/gamebuild/gamepos.rs).

The game is built on the GPU by geq_gamebuild (positions get the ids of the
reference's breadth-first construction) and stays in HBM; ``solve`` runs the
fixpoint there, ``energies`` reads budgets back, ``game`` / ``nodes`` download
the graph and positions on demand.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _lib
from .energy import EnergyConf, UpdateArray
from .energygame import GameGraph, energies_from_flat

STD = EnergyConf.STANDARD


@dataclass(frozen=True)
class AttackPosition:
    p: int
    q: tuple


@dataclass(frozen=True)
class SingletonPosition:
    p: int
    q: int


@dataclass(frozen=True)
class DefendPosition:
    p: int
    q: tuple
    qx: tuple


class Position:
    """Constructors mirroring Position::{attack, clause, defend} (gamepos.rs:489-513)."""

    @staticmethod
    def attack(p: int, q: Sequence[int]) -> AttackPosition:
        return AttackPosition(int(p), tuple(int(x) for x in q))

    @staticmethod
    def clause(p: int, q: int) -> SingletonPosition:
        return SingletonPosition(int(p), int(q))

    @staticmethod
    def defend(p: int, q: Sequence[int], qx: Sequence[int]) -> DefendPosition:
        return DefendPosition(int(p), tuple(int(x) for x in q), tuple(int(x) for x in qx))

    @staticmethod
    def is_attack(pos) -> bool:
        return not isinstance(pos, DefendPosition)


class GameBuild:
    """Device-resident spectroscopy game seeded with attack positions (p, {q})."""

    ENERGY_CONF = STD

    def __init__(self, lts, pairs: Sequence, max_positions: int = 0):
        self.ctx = _lib.Context.get()
        pr = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        if pr.size and (pr.min() < 0 or pr.max() > 0xFFFFFFFF):
            raise ValueError("process ids out of range")
        self.n_starting_points = pr.shape[0]
        self._pairs = pr                      # [k, 2] of (p, q)
        arr = np.ascontiguousarray(pr.astype(np.uint32).reshape(-1))
        if arr.size == 0:
            arr = np.zeros(2, np.uint32)
        s = lts._c_struct()
        h = C.c_void_p()
        st = _lib.geq_build_stats()
        _lib.check(_lib.lib().geq_gamebuild(self.ctx.handle, C.byref(s), _lib.u32p(arr), len(self._pairs),
                                            int(max_positions), C.byref(h), C.byref(st)))
        self.handle = h
        self.build_stats = st.as_dict()
        self._host = None
        self.last_stats = None

    # reference constructors (src/gamebuild.rs:165-215)
    @classmethod
    def compare(cls, lts, p: int, q: int) -> "GameBuild":
        return cls(lts, [(p, q)])

    @staticmethod
    def _class_pairs(members, sizes):
        """Every ordered pair (p, q), p != q, within each group (members group
        after group, `sizes` of them), in the order of `for p in group for q in
        group` (src/gamebuild.rs:182-194)."""
        members = np.asarray(members, dtype=np.int64)
        sizes = np.asarray(sizes, dtype=np.int64)
        sq = sizes * sizes
        total = int(sq.sum())
        if total == 0:
            return np.zeros((0, 2), np.int64)
        start = np.concatenate(([0], np.cumsum(sizes)[:-1]))          # first member of each group
        pstart = np.concatenate(([0], np.cumsum(sq)[:-1]))           # first pair of each group
        sz = np.repeat(sizes, sq)
        t = np.arange(total, dtype=np.int64) - np.repeat(pstart, sq)  # pair index within its group
        a, b = t // sz, t % sz
        base = np.repeat(start, sq)
        keep = a != b
        return np.stack([members[base + a][keep], members[base + b][keep]], axis=1)

    @classmethod
    def compare_multiple(cls, lts, processes: Sequence[int]):
        from .equivalence import StartInfo
        pairs = cls._class_pairs(*lts._enabledness_groups(processes))
        b = cls(lts, pairs)
        return b, StartInfo(None, lts.n_vertices(), pairs=pairs)

    @classmethod
    def compare_all(cls, lts):
        from .equivalence import StartInfo
        pairs = cls._class_pairs(*lts._enabledness_groups())
        b = cls(lts, pairs)
        return b, StartInfo(None, lts.n_vertices(), pairs=pairs)

    def starting_points(self) -> list:
        return [AttackPosition(p, (q,)) for p, q in self._pairs.tolist()]

    def n_positions(self) -> int:
        return int(self.build_stats["positions"])

    def truncated(self) -> bool:
        """max_positions stopped the build: the last level was never expanded."""
        return bool(self.build_stats["truncated"])

    def frontier(self) -> int:
        """First unexpanded position of a capped build (the position count otherwise)."""
        t, f = C.c_uint32(), C.c_uint32()
        _lib.check(_lib.lib().geq_built_frontier(self.handle, C.byref(t), C.byref(f)))
        return int(f.value)

    # -------------------------------------------------------------- device solve
    def solve(self, profile: bool = False, capped: bool = False) -> dict:
        """EnergyGame::standard_reach(game).run() on the device.  A capped build
        raises (it is not the complete game) unless ``capped`` asks for the
        solve of the capped game as a game of its own: its unexpanded frontier
        is then never a position to reach, and the budgets are not the
        reference's."""
        st = _lib.geq_stats()
        fn = _lib.lib().geq_built_solve_capped if capped else _lib.lib().geq_built_solve
        _lib.check(fn(self.handle, 1 if profile else 0, C.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def energies_raw(self, first: int, count: int):
        out = _lib.geq_energies()
        _lib.check(_lib.lib().geq_built_energies(self.handle, int(first), int(count), C.byref(out)))
        try:
            return _lib.energies_to_numpy(out)
        finally:
            _lib.lib().geq_energies_free(C.byref(out))

    def energies(self, first: int, count: int) -> list:
        o, d = self.energies_raw(first, count)
        return energies_from_flat(o, d, STD)

    def preorder_bits(self, equivalences, first: int = 0, count: int = None):
        """After solve(): bool array [count, len(equivalences)], True where no
        budget of position first + i lies below equivalence k
        (EnergyArray::test_equivalence, src/energy.rs:307-318), computed on the
        device; with first 0 and count n_starting_points these are the
        starting points' verdicts (equivalence.preorder_matrix_of)."""
        import numpy as np
        if count is None:
            count = self.n_starting_points - first
        eqs = np.array([e.raw_data()[0] for e in equivalences], dtype=np.uint32)
        for e in equivalences:
            if len(e.raw_data()) != 1:
                raise ValueError("one word per energy")
        out = np.zeros(max(int(count) * len(eqs), 1), dtype=np.uint8)
        _lib.check(_lib.lib().geq_built_preorders(self.handle, int(first), int(count), _lib.u32p(eqs), len(eqs),
                                                  out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out[:int(count) * len(eqs)].reshape(int(count), len(eqs)).astype(bool)

    # -------------------------------------------------------------- host copies
    def _download(self):
        if self._host is None:
            g = _lib.geq_built_game()
            _lib.check(_lib.lib().geq_built_download(self.handle, C.byref(g)))
            try:
                V, E = g.n_positions, g.n_edges

                def arr(p, n, dt=np.uint32):
                    return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dt, copy=True)
                self._host = {
                    "row": arr(g.row_offsets, V + 1), "col": arr(g.column_indices, E), "wt": arr(g.weights, E),
                    "atk": arr(g.attacker_pos, V, np.uint8), "kind": arr(g.kind, V, np.uint8), "p": arr(g.p, V),
                    "q_off": arr(g.q_off, V), "q_len": arr(g.q_len, V), "x_off": arr(g.x_off, V),
                    "x_len": arr(g.x_len, V), "pool": arr(g.pool, int(g.pool_len)),
                }
            finally:
                _lib.lib().geq_built_game_free(C.byref(g))
        return self._host

    @property
    def game(self) -> GameGraph:
        h = self._download()
        return GameGraph(STD, h["row"], h["col"], h["wt"].reshape(-1, 1), h["atk"])

    @property
    def nodes(self) -> list:
        h = self._download()
        out = []
        pool = h["pool"]
        for i in range(len(h["kind"])):
            k, p = int(h["kind"][i]), int(h["p"][i])
            q = tuple(int(x) for x in pool[h["q_off"][i]:h["q_off"][i] + h["q_len"][i]])
            if k == 0:
                out.append(AttackPosition(p, q))
            elif k == 1:
                out.append(SingletonPosition(p, q[0]))
            else:
                qx = tuple(int(x) for x in pool[h["x_off"][i]:h["x_off"][i] + h["x_len"][i]])
                out.append(DefendPosition(p, q, qx))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().geq_built_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
