"""The C ABI library loads, exports every function include/*.h declares, and
reports errors without a device (CPU only: no kernel is launched here)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from gpuequiv_amd import _lib, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(geq_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_functions():
    assert "geq_solve" in declared_functions()
    assert "geq_ctx_create" in declared_functions()


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = C.CDLL(_lib.LIB_PATH)
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_no_device_maps_to_no_adapter():
    h = C.c_void_p()
    rc = _lib.lib().geq_ctx_create(0, C.byref(h))
    if rc == 0:
        _lib.lib().geq_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == 1   # GEQ_ERR_NO_ADAPTER (gpuequiv::Error::NoAdapter)
    assert b"device" in _lib.lib().geq_last_error()


def test_invalid_ordinal_is_no_adapter():
    h = C.c_void_p()
    rc = _lib.lib().geq_ctx_create(1 << 20, C.byref(h))
    assert rc == 1


def test_standard_reach_host_helper():
    g = workloads.random_game(500, 9)
    keep = []
    gs = g._c_struct(keep)
    out = np.zeros(g.n_vertices(), np.uint32)
    n = C.c_uint32()
    _lib.check(_lib.lib().geq_standard_reach(C.byref(gs), _lib.u32p(out), C.byref(n)))
    deg = np.diff(g.row_offsets.astype(np.int64))
    exp = np.nonzero((g.attacker_pos == 0) & (deg == 0))[0]
    assert list(out[:n.value]) == list(exp)


def test_null_arguments_are_invalid_arg():
    rc = _lib.lib().geq_solve(None, None, None, 0, None, None, None)
    assert rc == 4
