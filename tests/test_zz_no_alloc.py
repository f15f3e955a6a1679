"""geq_solve_noalloc (the entry point that allocates no device memory: its
buffers live in a static device array) against geq_solve, through the C ABI,
on two calls in a row, the second with a smaller game."""
import ctypes as C

import numpy as np
import pytest

from gpuequiv_amd import EnergyConf, EnergyGame, _lib, workloads
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _solve_ref(g, reach):
    game = EnergyGame.with_reach(g, reach)
    out = game.run()
    return [tuple(sorted(int(x) for x in a.array.reshape(-1))) for a in out], game.last_stats


def _solve_noalloc(g, reach, cap=None):
    ctx = _lib.Context.get()
    keep = []
    gs = g._c_struct(keep)
    r = np.ascontiguousarray(reach, dtype=np.uint32)
    V = g.n_vertices()
    offs = np.zeros(V + 1, np.uint32)
    cap = 8 * V if cap is None else cap
    data = np.zeros(max(cap, 1), np.uint32)
    n = C.c_uint64()
    st = _lib.geq_stats()
    rc = _lib.lib().geq_solve_noalloc(ctx.handle, C.byref(gs), _lib.u32p(r) if len(r) else None, len(r), _lib.u32p(offs),
                                      _lib.u32p(data), cap, C.byref(n), C.byref(st))
    return rc, offs, data, int(n.value), st.as_dict()


def _lists(offs, data):
    return [tuple(int(x) for x in data[offs[v]:offs[v + 1]]) for v in range(len(offs) - 1)]


@pytest.mark.parametrize("first,second", [((100000, 2, None), (3000, 5, None)),
                                          ((20000, 7, EnergyConf(4, 1)), (500, 8, EnergyConf(4, 1)))])
def test_noalloc_matches_geq_solve_twice(first, second):
    for V, seed, conf in (first, second):
        print("noalloc vs geq_solve:", V, seed, conf, flush=True)
        g = workloads.random_game(V, seed) if conf is None else workloads.random_game_conf(V, seed, conf)
        reach = H.standard_reach(g)
        want, ws = _solve_ref(g, reach)
        rc, offs, data, n, st = _solve_noalloc(g, reach)
        assert rc == 0, _lib.lib().geq_last_error()
        got = _lists(offs, data)
        assert got == want
        assert n == int(offs[-1])
        for k in ("rounds", "node_evals", "edge_updates", "alg_bytes", "changed"):
            assert st[k] == ws[k], k


def test_noalloc_limits_and_small_output():
    # a game beyond the static scratch is refused before any device work; a
    # result larger than the caller's array reports its size
    big = workloads.random_game(2000000, 3)
    rc, *_ = _solve_noalloc(big, H.standard_reach(big), cap=16)
    assert rc == 5 and b"static scratch" in _lib.lib().geq_last_error()
    g = workloads.random_game(4000, 11)
    reach = H.standard_reach(g)
    rc, offs, data, n, _ = _solve_noalloc(g, reach, cap=3)
    assert rc == 4 and n > 3
    rc, offs, data, n2, _ = _solve_noalloc(g, reach, cap=n)
    assert rc == 0 and n2 == n
    assert _lists(offs, data) == _solve_ref(g, reach)[0]
