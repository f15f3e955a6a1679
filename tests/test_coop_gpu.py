"""The cooperative one-launch fixpoint (k_solve) against the round-per-launch
path and the oracle."""
import pytest

from gpuequiv_amd import workloads
from tests import helpers as H
from tests.test_solve_gpu import _solve_raw

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="cooperative k_solve is switched off: the 20000-position case hung twice; the grid "
                            "(min(occupancy API, 8) blocks/CU) can exceed the 7 waves/SIMD its 106 SGPRs allow"),
]


@pytest.mark.parametrize("V,seed", [(2000, 3), (20000, 5), (100000, 42)])
def test_cooperative_solve_matches(V, seed, monkeypatch):
    g = workloads.random_game(V, seed)
    print("cooperative k_solve", V, seed, flush=True)
    ref, sref = _solve_raw(g, 0)
    monkeypatch.setenv("GEQ_SOLVE_MODE", "0")
    got, s = _solve_raw(g, 0)
    assert s["mode"] == 0
    assert got == ref
    assert s["rounds"] == sref["rounds"] and s["node_evals"] == sref["node_evals"]
    assert s["alg_bytes"] == sref["alg_bytes"]
    o, d, _ = H.oracle_solve_game(g, H.standard_reach(g))
    assert got == H.canon_lists(o, d)
