"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The shipped library never uses
it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
P = C.POINTER


class orc_lists(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("esz", C.c_uint32), ("offsets", P(C.c_uint32)), ("data", P(C.c_uint32))]


class orc_game(C.Structure):
    _fields_ = [("V", C.c_uint32), ("E", C.c_uint32), ("n_start", C.c_uint32),
                ("row", P(C.c_uint32)), ("col", P(C.c_uint32)), ("wt", P(C.c_uint32)), ("atk", P(C.c_uint8)),
                ("kind", P(C.c_uint8)), ("p", P(C.c_uint32)), ("q1", P(C.c_uint32)),
                ("q_off", P(C.c_uint32)), ("qx_off", P(C.c_uint32)), ("pool", P(C.c_uint32))]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        u32, i32, u8 = P(C.c_uint32), P(C.c_int32), P(C.c_uint8)
        h.orc_conf_info.argtypes = [C.c_uint32, C.c_uint32, u32]
        h.orc_pack_energy.argtypes = [C.c_uint32, C.c_uint32, u32, C.c_uint32, u32]
        h.orc_pack_energy.restype = C.c_int
        h.orc_unpack_energy.argtypes = [C.c_uint32, C.c_uint32, u32, u32]
        h.orc_pack_update.argtypes = [C.c_uint32, C.c_uint32, i32, C.c_uint32, u32]
        h.orc_pack_update.restype = C.c_int
        h.orc_inv_update.argtypes = [C.c_uint32, C.c_uint32, u32, u32, u32]
        h.orc_defend_fold.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32, u32, u32, C.c_uint64, u32, C.c_uint32,
                                      P(C.c_int32)]
        h.orc_defend_fold.restype = C.c_int
        h.orc_solve.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32, u32, u32, u8, u32, u32, u32, C.c_uint32,
                                C.c_int, P(C.c_uint64)]
        h.orc_solve.restype = P(orc_lists)
        h.orc_lists_free.argtypes = [P(orc_lists)]
        h.orc_gamebuild.argtypes = [C.c_uint32, u32, u32, i32, C.c_uint32, u32]
        h.orc_gamebuild.restype = P(orc_game)
        h.orc_game_free.argtypes = [P(orc_game)]
        h.orc_gamebuild_capped.argtypes = [C.c_uint32, u32, u32, i32, C.c_uint32, u32, C.c_uint64, P(C.c_int)]
        h.orc_gamebuild_capped.restype = P(orc_game)
        h.orc_signature_refinement.argtypes = [C.c_uint32, u32, u32, i32, u32]
        h.orc_signature_refinement.restype = C.c_uint32
        h.orc_bisimilar_minimize.argtypes = [C.c_uint32, u32, u32, i32, u32, u32, u32, i32]
        h.orc_bisimilar_minimize.restype = C.c_uint32
        h.orc_set_threads.argtypes = [C.c_int]
        _lib = h
    return _lib


def set_threads(n: int) -> None:
    """Threads for the evaluations of one fixpoint round (1: sequential)."""
    lib().orc_set_threads(int(n))


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(P(C.c_uint32))


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(P(C.c_int32))


def conf_info(elements: int, mx: int) -> list:
    out = np.zeros(6, np.uint32)
    lib().orc_conf_info(elements, mx, out.ctypes.data_as(P(C.c_uint32)))
    return [int(x) for x in out]


def pack_energy(elements: int, mx: int, vals) -> list | None:
    esz = conf_info(elements, mx)[0]
    v, vp = _u32(vals)
    out, op = _u32(np.zeros(max(esz, 1)))
    if lib().orc_pack_energy(elements, mx, vp, len(v), op) != 0:
        return None
    return [int(x) for x in out[:esz]]


def unpack_energy(elements: int, mx: int, words) -> list:
    w, wp = _u32(list(words) + [0] * 32)
    out, op = _u32(np.zeros(elements))
    lib().orc_unpack_energy(elements, mx, wp, op)
    return [int(x) for x in out]


def pack_update(elements: int, mx: int, ints) -> list | None:
    usz = conf_info(elements, mx)[3]
    v, vp = _i32(ints)
    out, op = _u32(np.zeros(max(usz, 1)))
    if lib().orc_pack_update(elements, mx, vp, len(v), op) != 0:
        return None
    return [int(x) for x in out[:usz]]


def inv_update(elements: int, mx: int, e_words, u_words) -> list:
    esz = conf_info(elements, mx)[0]
    e, ep = _u32(list(e_words) + [0] * 32)
    u, up = _u32(list(u_words) + [0] * 1024)
    out, op = _u32(np.zeros(32))
    lib().orc_inv_update(elements, mx, ep, up, op)
    return [int(x) for x in out[:esz]]


def defend_fold(elements: int, mx: int, succ_lists, weights, mem: int = 1 << 40):
    """Defender combination of src/energygame/tests.rs form: returns (status, list of packed words)."""
    esz = conf_info(elements, mx)[0]
    offs = [0]
    flat = []
    for lst in succ_lists:
        for e in lst:
            flat.extend(e)
        offs.append(offs[-1] + len(lst))
    o, op = _u32(offs)
    d, dp = _u32(flat if flat else [0])
    w, wp = _u32([x for u in weights for x in u])
    cap = 1 << 16
    out, outp = _u32(np.zeros(cap * esz))
    st = C.c_int32()
    rc = lib().orc_defend_fold(elements, mx, len(succ_lists), op, dp, wp, mem, outp, cap, C.byref(st))
    assert rc == 0
    s = st.value
    if s < 0:
        return s, []
    return s, [list(map(int, out[i * esz:(i + 1) * esz])) for i in range(s)]


def solve(elements: int, mx: int, row, col, wt, atk, reach, init_off=None, init_data=None, emulate_mem: bool = False):
    """Run the oracle fixpoint.  Returns (offsets, data, stats dict); lists keep
    the reference's element order."""
    row, rowp = _u32(row)
    V = len(row) - 1
    col, colp = _u32(col)
    wt, wtp = _u32(wt if len(wt) else [0])
    atk = np.ascontiguousarray(atk, dtype=np.uint8)
    nreach = len(reach)
    reach, reachp = _u32(reach if nreach else [0])
    if init_off is None:
        esz = conf_info(elements, mx)[0]
        counts = np.zeros(V, np.int64)
        rl = [int(x) for x in reach[:nreach]]
        for v in rl:
            counts[v] = 1
        init_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        init_data = np.zeros(max(int(init_off[-1]) * esz, 1), np.uint32)
    io, iop = _u32(init_off)
    idt, idp = _u32(init_data if len(init_data) else [0])
    stats = (C.c_uint64 * 7)()
    r = lib().orc_solve(elements, mx, V, rowp, colp, wtp, atk.ctypes.data_as(P(C.c_uint8)), iop, idp, reachp, nreach,
                        1 if emulate_mem else 0, stats)
    if not r:
        raise ValueError("oracle: unsupported configuration")
    try:
        esz = r.contents.esz
        offs = np.ctypeslib.as_array(r.contents.offsets, shape=(V + 1,)).copy()
        tot = int(offs[-1]) * esz
        data = np.ctypeslib.as_array(r.contents.data, shape=(max(tot, 1),))[:tot].copy()
    finally:
        lib().orc_lists_free(r)
    keys = ["rounds", "node_evals", "atk_evals", "def_evals", "edge_updates", "alg_bytes", "def_retries"]
    return offs, data, dict(zip(keys, [int(x) for x in stats]))


def gamebuild(n_states: int, lrow, ldst, llbl, pairs, max_positions: int = 0, positions_wanted: bool = True):
    """Build the spectroscopy game seeded with the given (p, q) attack positions.
    max_positions > 0: stop after the breadth-first level that passes it (the
    nodes that level created keep empty rows; "truncated" in the result).
    positions_wanted=False skips the Python list of position tuples."""
    lrow, lrp = _u32(lrow)
    ldst, ldp = _u32(ldst if len(ldst) else [0])
    llbl, llp = _i32(llbl if len(llbl) else [0])
    pr, prp = _u32(np.asarray(pairs, dtype=np.uint32).reshape(-1) if len(pairs) else [0])
    trunc = C.c_int(0)
    g = lib().orc_gamebuild_capped(n_states, lrp, ldp, llp, len(pairs), prp, int(max_positions), C.byref(trunc))
    try:
        c = g.contents
        V, E = c.V, c.E
        arr = lambda p, n: np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy()
        qo = arr(c.q_off, V + 1)
        xo = arr(c.qx_off, V + 1)
        pool = arr(c.pool, int(qo[-1]) if V else 0)
        kinds = arr(c.kind, V)
        ps = arr(c.p, V)
        q1 = arr(c.q1, V)
        positions = []
        for i in (range(V) if positions_wanted else ()):
            k = int(kinds[i])
            if k == 0:
                positions.append(("A", int(ps[i]), [int(x) for x in pool[qo[i]:xo[i]]]))
            elif k == 1:
                positions.append(("C", int(ps[i]), int(q1[i])))
            else:
                positions.append(("D", int(ps[i]), [int(x) for x in pool[qo[i]:xo[i]]],
                                  [int(x) for x in pool[xo[i]:qo[i + 1]]]))
        return {
            "V": V, "E": E, "n_start": c.n_start,
            "row": arr(c.row, V + 1), "col": arr(c.col, E), "wt": arr(c.wt, E), "atk": arr(c.atk, V),
            "positions": positions, "truncated": bool(trunc.value),
        }
    finally:
        lib().orc_game_free(g)


def signature_refinement(n: int, lrow, ldst, llbl):
    lrow, lrp = _u32(lrow)
    ldst, ldp = _u32(ldst if len(ldst) else [0])
    llbl, llp = _i32(llbl if len(llbl) else [0])
    part = np.zeros(max(n, 1), np.uint32)
    cnt = lib().orc_signature_refinement(n, lrp, ldp, llp, part.ctypes.data_as(P(C.c_uint32)))
    return part[:n], int(cnt)


def bisimilar_minimize(n: int, lrow, ldst, llbl):
    lrow, lrp = _u32(lrow)
    E = int(lrow[-1]) if len(lrow) else 0
    ldst, ldp = _u32(ldst if len(ldst) else [0])
    llbl, llp = _i32(llbl if len(llbl) else [0])
    part = np.zeros(max(n, 1), np.uint32)
    orow = np.zeros(n + 1, np.uint32)
    odst = np.zeros(max(E, 1), np.uint32)
    olbl = np.zeros(max(E, 1), np.int32)
    cnt = lib().orc_bisimilar_minimize(n, lrp, ldp, llp, part.ctypes.data_as(P(C.c_uint32)),
                                       orow.ctypes.data_as(P(C.c_uint32)), odst.ctypes.data_as(P(C.c_uint32)),
                                       olbl.ctypes.data_as(P(C.c_int32)))
    k = int(orow[cnt])
    return part[:n], int(cnt), orow[:cnt + 1].copy(), odst[:k].copy(), olbl[:k].copy()
