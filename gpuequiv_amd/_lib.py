"""ctypes binding of the HIP library (include/gpuequiv.h).

The library is built in-tree (``gpuequiv_amd/libgpuequiv_hip.so``) by
``gpuequiv_amd/build.py``.  There is no fallback: when the library is missing
every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

LIB_NAME = "libgpuequiv_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

GEQ_OK = 0
_KINDS = {1: "NoAdapter", 2: "NoDevice", 3: "BufferMap", 4: "InvalidArg", 5: "OutOfMemory", 6: "Unsupported", 7: "Internal"}


class Error(RuntimeError):
    """gpuequiv::Error (src/error.rs:3-40) plus the C ABI's extra codes."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.kind = _KINDS.get(code, "Internal")
        super().__init__(f"{self.kind}: {message}")


class geq_conf(C.Structure):
    _fields_ = [("elements", C.c_uint32), ("max", C.c_uint32)]


class geq_graph(C.Structure):
    _fields_ = [
        ("conf", geq_conf),
        ("n_vertices", C.c_uint32),
        ("n_edges", C.c_uint32),
        ("row_offsets", C.POINTER(C.c_uint32)),
        ("column_indices", C.POINTER(C.c_uint32)),
        ("weights", C.POINTER(C.c_uint32)),
        ("attacker_pos", C.POINTER(C.c_uint8)),
        ("rev_row_offsets", C.POINTER(C.c_uint32)),
        ("rev_column_indices", C.POINTER(C.c_uint32)),
    ]


class geq_energies(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("energy_size", C.c_uint32),
        ("offsets", C.POINTER(C.c_uint32)),
        ("data", C.POINTER(C.c_uint32)),
    ]


class geq_stats(C.Structure):
    _fields_ = [
        ("rounds", C.c_uint64),
        ("node_evals", C.c_uint64),
        ("edge_updates", C.c_uint64),
        ("alg_bytes", C.c_uint64),
        ("changed", C.c_uint64),
        ("solve_ms", C.c_double),
        ("eval_kernel_ms", C.c_double),
        ("expand_kernel_ms", C.c_double),
        ("eval_launches", C.c_uint64),
        ("solve_kernel_ms", C.c_double),
        ("mode", C.c_uint32),
        ("pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class geq_lts(C.Structure):
    _fields_ = [
        ("n_states", C.c_uint32),
        ("n_transitions", C.c_uint32),
        ("row_offsets", C.POINTER(C.c_uint32)),
        ("targets", C.POINTER(C.c_uint32)),
        ("labels", C.POINTER(C.c_int32)),
    ]


class geq_build_stats(C.Structure):
    _fields_ = [
        ("positions", C.c_uint64),
        ("edges", C.c_uint64),
        ("levels", C.c_uint64),
        ("set_words", C.c_uint64),
        ("build_ms", C.c_double),
        ("truncated", C.c_uint32),
        ("pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class geq_built_game(C.Structure):
    _fields_ = [
        ("n_positions", C.c_uint32),
        ("n_edges", C.c_uint32),
        ("n_start", C.c_uint32),
        ("row_offsets", C.POINTER(C.c_uint32)),
        ("column_indices", C.POINTER(C.c_uint32)),
        ("weights", C.POINTER(C.c_uint32)),
        ("attacker_pos", C.POINTER(C.c_uint8)),
        ("kind", C.POINTER(C.c_uint8)),
        ("p", C.POINTER(C.c_uint32)),
        ("q_off", C.POINTER(C.c_uint32)),
        ("q_len", C.POINTER(C.c_uint32)),
        ("x_off", C.POINTER(C.c_uint32)),
        ("x_len", C.POINTER(C.c_uint32)),
        ("pool", C.POINTER(C.c_uint32)),
        ("pool_len", C.c_uint64),
    ]


class geq_server_opts(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("edges_per_position", C.c_double), ("max_jobs", C.c_uint32),
                ("rounds_per_step", C.c_uint32), ("compact_entries", C.c_uint64), ("max_positions", C.c_uint64),
                ("builders", C.c_uint32), ("pad", C.c_uint32)]


class geq_job_info(C.Structure):
    _fields_ = [("state", C.c_uint32), ("status", C.c_uint32), ("positions", C.c_uint32), ("edges", C.c_uint32),
                ("t_submit_ms", C.c_double), ("t_built_ms", C.c_double), ("t_admit_ms", C.c_double),
                ("t_done_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class geq_server_stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("steps", "rounds", "compactions", "node_slots", "node_slots_used", "edge_slots",
                                          "edge_slots_used", "arena_entries", "arena_top", "jobs_resident",
                                          "jobs_waiting", "jobs_done", "builds", "split_builds",
                                          "deferred_admissions")]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class geq_tune_params(C.Structure):
    _fields_ = [("waves_per_block", C.c_uint32), ("eval_blocks_per_cu", C.c_uint32), ("expand_blocks_per_cu", C.c_uint32),
                ("rounds_per_launch", C.c_uint32), ("arena_entries", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class geq_tune_report(C.Structure):
    _fields_ = [("best", geq_tune_params), ("default_ms", C.c_double), ("best_ms", C.c_double),
                ("candidates", C.c_uint32), ("repeats", C.c_uint32), ("arena_peak", C.c_uint64)]

    def as_dict(self) -> dict:
        return {"best": self.best.as_dict(), "default_ms": self.default_ms, "best_ms": self.best_ms,
                "candidates": int(self.candidates), "repeats": int(self.repeats), "arena_peak": int(self.arena_peak)}


P = C.POINTER
_SIGS = {
    "geq_last_error": (C.c_char_p, []),
    "geq_ctx_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
    "geq_ctx_destroy": (None, [C.c_void_p]),
    "geq_ctx_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, P(C.c_int)]),
    "geq_ctx_memory": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint64)]),
    "geq_solve": (C.c_int, [C.c_void_p, P(geq_graph), P(C.c_uint32), C.c_uint32, P(geq_energies), P(geq_energies), P(geq_stats)]),
    "geq_energies_free": (None, [P(geq_energies)]),
    "geq_solve_noalloc": (C.c_int, [C.c_void_p, P(geq_graph), P(C.c_uint32), C.c_uint32, P(C.c_uint32), P(C.c_uint32),
                                    C.c_uint64, P(C.c_uint64), P(geq_stats)]),
    "geq_game_upload": (C.c_int, [C.c_void_p, P(geq_graph), P(C.c_void_p)]),
    "geq_game_solve": (C.c_int, [C.c_void_p, P(C.c_uint32), C.c_uint32, P(geq_energies), C.c_int, P(geq_stats)]),
    "geq_game_solve_standard": (C.c_int, [C.c_void_p, C.c_int, P(geq_stats)]),
    "geq_game_energies": (C.c_int, [C.c_void_p, P(geq_energies)]),
    "geq_game_free": (None, [C.c_void_p]),
    "geq_standard_reach": (C.c_int, [P(geq_graph), P(C.c_uint32), P(C.c_uint32)]),
    "geq_batch_upload": (C.c_int, [C.c_void_p, C.c_uint32, P(geq_graph), P(C.c_void_p)]),
    "geq_batch_solve": (C.c_int, [C.c_void_p, P(P(C.c_uint32)), P(C.c_uint32), C.c_int, P(geq_stats)]),
    "geq_batch_energies": (C.c_int, [C.c_void_p, P(geq_energies)]),
    "geq_gamebuild": (C.c_int, [C.c_void_p, P(geq_lts), P(C.c_uint32), C.c_uint32, C.c_uint64, P(C.c_void_p),
                                P(geq_build_stats)]),
    "geq_built_download": (C.c_int, [C.c_void_p, P(geq_built_game)]),
    "geq_built_game_free": (None, [P(geq_built_game)]),
    "geq_built_solve": (C.c_int, [C.c_void_p, C.c_int, P(geq_stats)]),
    "geq_built_solve_capped": (C.c_int, [C.c_void_p, C.c_int, P(geq_stats)]),
    "geq_built_frontier": (C.c_int, [C.c_void_p, P(C.c_uint32), P(C.c_uint32)]),
    "geq_built_energies": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(geq_energies)]),
    "geq_built_preorders": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_uint32), C.c_uint32, P(C.c_uint8)]),
    "geq_built_free": (None, [C.c_void_p]),
    "geq_signature_refinement": (C.c_int, [P(geq_lts), P(C.c_uint32), P(C.c_uint32)]),
    "geq_server_create": (C.c_int, [C.c_void_p, P(geq_server_opts), P(C.c_void_p)]),
    "geq_server_destroy": (None, [C.c_void_p]),
    "geq_server_submit": (C.c_int, [C.c_void_p, P(geq_lts), P(C.c_uint32), C.c_uint32, P(C.c_uint64)]),
    "geq_server_step": (C.c_int, [C.c_void_p, P(C.c_uint32), P(C.c_uint32)]),
    "geq_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, P(geq_job_info), P(geq_energies)]),
    "geq_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "geq_server_get_stats": (C.c_int, [C.c_void_p, P(geq_server_stats)]),
    "geq_signature_refinement_gpu": (C.c_int, [C.c_void_p, P(geq_lts), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
    "geq_bisimilar_minimize_gpu": (C.c_int, [C.c_void_p, P(geq_lts), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                             P(C.c_uint32), P(C.c_int32), P(C.c_uint32)]),
    "geq_game_set_solve_mode": (C.c_int, [C.c_void_p, C.c_uint32]),
    "geq_game_get_solve_mode": (C.c_int, [C.c_void_p, P(C.c_uint32)]),
    "geq_game_get_tune": (C.c_int, [C.c_void_p, P(geq_tune_params)]),
    "geq_game_set_tune": (C.c_int, [C.c_void_p, P(geq_tune_params)]),
    "geq_game_autotune": (C.c_int, [C.c_void_p, P(C.c_uint32), C.c_uint32, C.c_uint32, P(geq_tune_report)]),
    "geq_batch_autotune": (C.c_int, [C.c_void_p, P(P(C.c_uint32)), P(C.c_uint32), C.c_uint32, P(geq_tune_report)]),
    "geq_server_finished": (C.c_int, [C.c_void_p, P(C.c_uint64), C.c_uint32, P(C.c_uint32)]),
    "geq_bisimilar_minimize": (C.c_int, [P(geq_lts), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                         P(C.c_int32), P(C.c_uint32)]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: run gpuequiv_amd/build.py (or __graft_entry__.build())")
            h = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            _lib = h
        return _lib


def check(code: int) -> None:
    if code != GEQ_OK:
        msg = lib().geq_last_error()
        raise Error(code, msg.decode() if msg else "")


def u32p(a: np.ndarray):
    return a.ctypes.data_as(P(C.c_uint32))


def u8p(a: np.ndarray):
    return a.ctypes.data_as(P(C.c_uint8))


def i32p(a: np.ndarray):
    return a.ctypes.data_as(P(C.c_int32))


class Context:
    """Process-wide device context, like GPUCommon::get_gpu (src/energygame/gpuutils.rs:15-31)."""

    _instance = None
    _ilock = threading.Lock()

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().geq_ctx_create(device, C.byref(h)))
        self.handle = h

    @classmethod
    def get(cls) -> "Context":
        with cls._ilock:
            if cls._instance is None:
                cls._instance = Context(int(os.environ.get("GEQ_DEVICE", "0")))
            return cls._instance

    def memory(self) -> tuple:
        """(free, total) bytes of device memory."""
        f, t = C.c_uint64(), C.c_uint64()
        check(lib().geq_ctx_memory(self.handle, C.byref(f), C.byref(t)))
        return int(f.value), int(t.value)

    def info(self) -> tuple:
        buf = C.create_string_buffer(128)
        ncu = C.c_int()
        check(lib().geq_ctx_info(self.handle, buf, 128, C.byref(ncu)))
        return buf.value.decode(), ncu.value

    def __del__(self):
        try:
            if getattr(self, "handle", None) and _lib is not None:
                _lib.geq_ctx_destroy(self.handle)
        except Exception:
            pass


def energies_to_numpy(e: geq_energies):
    n = e.n_nodes
    offs = np.ctypeslib.as_array(e.offsets, shape=(n + 1,)).copy()
    tot = int(offs[-1]) * e.energy_size
    data = np.ctypeslib.as_array(e.data, shape=(max(tot, 1),))[:tot].copy() if tot else np.zeros(0, np.uint32)
    return offs, data
