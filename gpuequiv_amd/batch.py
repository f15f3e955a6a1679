"""Batched solving: many independent energy games in one device-resident solve.

Games are laid out back to back in HBM and share one device worklist, so every
round evaluates the frontier of all games together (geq_batch_* in
include/gpuequiv.h).  Each game's result equals its standalone result.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from .energygame import GameGraph, energies_from_flat


def standard_reach(graph: GameGraph) -> np.ndarray:
    deg = np.diff(graph.row_offsets.astype(np.int64))
    return np.ascontiguousarray(np.nonzero((graph.attacker_pos == 0) & (deg == 0))[0], dtype=np.uint32)


class GameBatch:
    """Device-resident batch of games (standard_reach seeding per game)."""

    def __init__(self, graphs: Sequence[GameGraph], reach: Sequence[np.ndarray] | None = None):
        self.graphs = list(graphs)
        self.ctx = _lib.Context.get()
        self._keep: list = []
        arr = (_lib.geq_graph * len(self.graphs))()
        for i, g in enumerate(self.graphs):
            arr[i] = g._c_struct(self._keep)
        self._structs = arr
        h = C.c_void_p()
        _lib.check(_lib.lib().geq_batch_upload(self.ctx.handle, len(self.graphs), arr, C.byref(h)))
        self.handle = h
        if reach is None:
            # every game's standard_reach, found on the device (NULL lists)
            self.reach = None
            self._reach_ptrs = None
            self._reach_n = None
        else:
            self.reach = [np.ascontiguousarray(r, dtype=np.uint32) for r in reach]
            self._reach_ptrs = (C.POINTER(C.c_uint32) * len(self.reach))(*[_lib.u32p(r) for r in self.reach])
            self._reach_n = np.array([len(r) for r in self.reach], dtype=np.uint32)
        self.last_stats: dict | None = None

    def _n_ptr(self):
        return None if self._reach_n is None else _lib.u32p(self._reach_n)

    def solve(self, profile: bool = False) -> dict:
        st = _lib.geq_stats()
        _lib.check(_lib.lib().geq_batch_solve(self.handle, self._reach_ptrs, self._n_ptr(),
                                              1 if profile else 0, C.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def autotune(self, repeats: int = 3) -> dict:
        """Pick workgroup size, grid sizes, rounds per graph replay and arena
        capacity by timing candidates on this batch (geq_batch_autotune)."""
        rep = _lib.geq_tune_report()
        _lib.check(_lib.lib().geq_batch_autotune(self.handle, self._reach_ptrs, self._n_ptr(), int(repeats),
                                                 C.byref(rep)))
        return rep.as_dict()

    def set_solve_mode(self, mode: int) -> None:
        """1: replayed round graphs over the whole batch (default); 3: one
        persistent launch with a block per game (geq_game_set_solve_mode)."""
        _lib.check(_lib.lib().geq_game_set_solve_mode(self.handle, int(mode)))

    def solve_mode(self) -> int:
        m = C.c_uint32()
        _lib.check(_lib.lib().geq_game_get_solve_mode(self.handle, C.byref(m)))
        return int(m.value)

    def tune(self) -> dict:
        p = _lib.geq_tune_params()
        _lib.check(_lib.lib().geq_game_get_tune(self.handle, C.byref(p)))
        return p.as_dict()

    def set_tune(self, **kw) -> None:
        p = _lib.geq_tune_params()
        _lib.check(_lib.lib().geq_game_get_tune(self.handle, C.byref(p)))
        for k, v in kw.items():
            setattr(p, k, int(v))
        _lib.check(_lib.lib().geq_game_set_tune(self.handle, C.byref(p)))

    def energies(self) -> list:
        n = len(self.graphs)
        outs = (_lib.geq_energies * n)()
        _lib.check(_lib.lib().geq_batch_energies(self.handle, outs))
        res = []
        try:
            for i in range(n):
                o, d = _lib.energies_to_numpy(outs[i])
                res.append(energies_from_flat(o, d, self.graphs[i].conf))
        finally:
            for i in range(n):
                _lib.lib().geq_energies_free(C.byref(outs[i]))
        return res

    def raw_energies(self) -> list:
        """Per game (offsets, data) numpy arrays, canonical order."""
        n = len(self.graphs)
        outs = (_lib.geq_energies * n)()
        _lib.check(_lib.lib().geq_batch_energies(self.handle, outs))
        res = []
        try:
            for i in range(n):
                res.append(_lib.energies_to_numpy(outs[i]))
        finally:
            for i in range(n):
                _lib.lib().geq_energies_free(C.byref(outs[i]))
        return res

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().geq_game_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
