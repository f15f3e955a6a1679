"""Serving with continuous batching (geq_server_* in include/gpuequiv.h).

Jobs are (TransitionSystem, [(p, q), ...]) pairs: the spectroscopy game of
compare_unminimized seeded with every pair.  Jobs that arrive while others
are on the GPU are built on a second stream and join the shared device
worklist at the next round boundary; each finished job yields the budgets of
its starting positions.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import Sequence

import numpy as np

from . import _lib
from .energy import EnergyConf
from .energygame import energies_from_flat

STATES = {0: "queued", 1: "building", 2: "built", 3: "resident", 4: "done", 5: "failed"}


class Server:
    def __init__(self, budget_bytes: int = 4 << 30, max_jobs: int = 4096, rounds_per_step: int = 4,
                 edges_per_position: float = 4.0, compact_entries: int = 0, max_positions: int = 0,
                 builders: int = 3):
        self.ctx = _lib.Context.get()
        o = _lib.geq_server_opts(int(budget_bytes), float(edges_per_position), int(max_jobs), int(rounds_per_step),
                                 int(compact_entries), int(max_positions), int(builders), 0)
        h = C.c_void_p()
        _lib.check(_lib.lib().geq_server_create(self.ctx.handle, C.byref(o), C.byref(h)))
        self.handle = h

    def submit(self, lts, pairs: Sequence) -> int:
        arr = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1))
        if arr.size == 0:
            arr = np.zeros(2, np.uint32)
        s = lts._c_struct()
        jid = C.c_uint64()
        _lib.check(_lib.lib().geq_server_submit(self.handle, C.byref(s), _lib.u32p(arr), len(pairs), C.byref(jid)))
        return int(jid.value)

    def step(self):
        r, w = C.c_uint32(), C.c_uint32()
        _lib.check(_lib.lib().geq_server_step(self.handle, C.byref(r), C.byref(w)))
        return int(r.value), int(w.value)

    def poll(self, job_id: int, with_result: bool = True):
        info = _lib.geq_job_info()
        out = _lib.geq_energies()
        _lib.check(_lib.lib().geq_server_poll(self.handle, int(job_id), C.byref(info),
                                              C.byref(out) if with_result else None))
        d = info.as_dict()
        d["state_name"] = STATES.get(d["state"], "?")
        res = None
        if with_result and info.state == 4:
            try:
                res = _lib.energies_to_numpy(out)
            finally:
                _lib.lib().geq_energies_free(C.byref(out))
        return d, res

    def stats(self) -> dict:
        st = _lib.geq_server_stats()
        _lib.check(_lib.lib().geq_server_get_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def release(self, job_id: int) -> None:
        _lib.check(_lib.lib().geq_server_release(self.handle, int(job_id)))

    def finished(self, cap: int = 4096) -> list:
        """Ids of the jobs that finished since the last call."""
        buf = (C.c_uint64 * cap)()
        n = C.c_uint32()
        _lib.check(_lib.lib().geq_server_finished(self.handle, buf, cap, C.byref(n)))
        return list(buf[:n.value])

    def run_until_done(self, job_ids, max_steps: int = 1 << 30):
        """Step until every given job is done or failed; returns {id: (info, (offs, data))}."""
        pending = set(job_ids)
        out = {}
        for _ in range(max_steps):
            resident, _waiting = self.step()
            if resident == 0:
                time.sleep(2e-4)          # games still being built
            for j in self.finished():
                if j in pending:
                    out[j] = self.poll(j)
                    pending.discard(j)
            if not pending:
                return out
        raise RuntimeError("jobs did not finish")

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib.lib().geq_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def energies_of(res, conf: EnergyConf = EnergyConf.STANDARD) -> list:
    offs, data = res
    return energies_from_flat(offs, data, conf)
