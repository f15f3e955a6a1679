"""The C++ host API (include/gpuequiv.hpp) against the reference's own test
vectors: packing known answers (src/energy.rs:846-936), the games of
tests/energygame.rs, the LTS cases of tests/complete.rs and the src/lib.rs
doctests.  The test program is tests/cpp/test_api.cpp; the case file is
written here from the committed fixtures."""
import os
import subprocess

import pytest

from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_api")


def _binary():
    if not os.path.exists(BIN):
        from gpuequiv_amd import build
        build.build_cpp_test()
    return BIN


def _energies(rows):
    out = [str(len(rows))]
    for e in rows:
        out.append(" ".join([str(len(e))] + [str(int(x)) for x in e]))
    return " ".join(out)


def _position(p):
    if p[0] == "A":
        return f"A {p[1]} {len(p[2])} " + " ".join(map(str, p[2]))
    if p[0] == "C":
        return f"C {p[1]} {p[2]}"
    return (f"D {p[1]} {len(p[2])} " + " ".join(map(str, p[2])) + f" {len(p[3])} " + " ".join(map(str, p[3])))


def write_cases(path, with_device_error=False):
    L = []
    # src/energy.rs:846-871, :898-936
    L.append("pack_energy 12 15 12 0 1 2 3 4 5 6 7 8 13 14 15 2 %d %d" % (0x76543210, 0xFED8))
    L.append("pack_energy 6 3 5 0 1 2 0 3 1 %d" % 0x324)
    L.append("pack_update 12 3 12 0 -1 1 2 3 4 5 6 9 10 11 12 2 %d %d" % (0x76543210, 0xDCBA))
    if with_device_error:
        L.append("no_device")
    for g in H.load("energygame_games.json")["games"]:
        el, mx = g["conf"]
        L.append(f"game {g['name']} {el} {mx} {g['n']}")
        L.append(f"{len(g['attacker'])} " + " ".join(map(str, g["attacker"])))
        L.append(str(len(g["edges"])))
        for a, b, u in g["edges"]:
            L.append(f"{a} {b} {len(u)} " + " ".join(map(str, u)))
        for exp in g["expected"]:
            L.append(_energies(exp))
    c = H.load("complete_lts.json")

    def lts(name, d):
        L.append(f"lts {name} {len(d['transitions'])}")
        for a, b, l in d["transitions"]:
            L.append(f"{a} {b} {l}")

    d = c["bisping_fig4"]
    lts("bisping_fig4", d)
    p, q = d["compare"]
    L.append(f"game {p} {q} {len(d['positions'])}")
    for pos in d["positions"]:
        L.append(_position(pos))
    for adj in d["adjacency_sorted"]:
        L.append(f"{len(adj)} " + " ".join(map(str, adj)))
    for e in d["energies"]:
        L.append(_energies(e))
    L.append("end")
    d = c["bisimilar"]
    lts("bisimilar", d)
    L.append(f"compare {d['compare'][0]} {d['compare'][1]} 1 bisimulation 0 1 0")
    L.append("end")
    d = c["failure_trace"]
    lts("failure_trace", d)
    L.append(f"compare {d['compare'][0]} {d['compare'][1]} {len(d['holds'])} " + " ".join(d["holds"]) +
             f" {len(d['fails'])} " + " ".join(d["fails"]) + " 0")
    L.append("end")
    d = c["doctest"]
    lts("doctest", d)
    L.append("compare 0 4 1 traces 1 simulation 0")
    for kind in ("preorder", "equiv"):
        for val, key in ((1, f"{kind}_true"), (0, f"{kind}_false")):
            for p, q, name in d[key]:
                L.append(f"{kind} {p} {q} {name} {val}")
    L.append(f"classes traces {d['traces_count_classes']} {len(d['traces_class_of_0'])} " +
             " ".join(map(str, d["traces_class_of_0"])))
    L.append("end")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def _run(args, timeout=120):
    print("run", args, flush=True)
    r = subprocess.run([_binary()] + args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-2000:])
    return r


def test_cpp_api_cpu(tmp_path):
    import torch
    cases = tmp_path / "cases.txt"
    write_cases(cases, with_device_error=not torch.cuda.is_available())
    r = _run(["--cpu", str(cases)])
    assert r.returncode == 0 and "ALL OK" in r.stdout


@pytest.mark.gpu
def test_cpp_api_gpu(tmp_path):
    cases = tmp_path / "cases.txt"
    write_cases(cases)
    r = _run([str(cases)])
    assert r.returncode == 0 and "ALL OK" in r.stdout
    assert "lts doctest ok" in r.stdout and "lts bisping_fig4 ok" in r.stdout
