"""Build the HIP library in-tree: gpuequiv_amd/libgpuequiv_hip.so (gfx950)."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgpuequiv_hip.so")
ARCH = os.environ.get("GEQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def _flags() -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            # packed FP32 instructions stay off (no kernel here needs them)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-parameter", "-Wno-unused-variable"]


def _run(cmd: list, verbose: bool) -> subprocess.CompletedProcess:
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False) -> str:
    """Compile every source to an object in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor

    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.basename(src) + ".o") for src in sources]
    jobs = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda so: _run([hipcc()] + _flags() + ["-c", so[0], "-o", so[1]], verbose),
                                zip(sources, objs)))
    msgs = []
    failed = False
    for r in results:
        msgs += [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln]
        failed |= r.returncode != 0
    if not failed:
        r = _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs, verbose)
        msgs += [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln]
        failed = r.returncode != 0
    if failed:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("hipcc failed building libgpuequiv_hip.so")
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(OUT + ".tmp", OUT)
    return OUT


def build_cpp_test(verbose: bool = False) -> str:
    """The C++ host-API test program (tests/cpp/test_api.cpp), linked against
    the library built above."""
    root = os.path.dirname(HERE)
    out_dir = os.path.join(root, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "test_api")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include"),
           os.path.join(root, "tests", "cpp", "test_api.cpp"), "-L", HERE, "-lgpuequiv_hip",
           "-Wl,-rpath,$ORIGIN/../../../gpuequiv_amd", "-Wl,--allow-shlib-undefined", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the C++ API test")
    os.replace(out + ".tmp", out)
    return out


def build_sanitized(verbose: bool = False) -> tuple:
    """Host code under ASan + UBSan, for CPU-only runs: the library built with
    the sanitizers on its host code only (-Xarch_host; device code unchanged),
    the C++ API test program and tests/cpp/test_host_sanitize.cpp (the C ABI's
    no-device paths and the CPU oracle) linked against it.  Returns the two
    program paths."""
    from concurrent.futures import ThreadPoolExecutor

    root = os.path.dirname(HERE)
    out_dir = os.path.join(root, "tests", "cpp", "_build", "sanitize")
    os.makedirs(out_dir, exist_ok=True)
    san = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
    host_san = [x for f in san for x in ("-Xarch_host", f)]
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    objs = [os.path.join(out_dir, os.path.basename(src) + ".o") for src in sources]
    flags = [f"--offload-arch={ARCH}", "-O1", "-g", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-w"] + host_san
    jobs = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as pool:
        results = list(pool.map(lambda so: _run([hipcc()] + flags + ["-c", so[0], "-o", so[1]], verbose), zip(sources, objs)))
    for r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed building the sanitized library")
    lib = os.path.join(out_dir, "libgpuequiv_hip_san.so")
    r = _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib] + objs, verbose)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking the sanitized library failed")
    clang = os.environ.get("GEQ_CLANGXX", "/opt/rocm/llvm/bin/clang++")
    progs = []
    for name, srcs in (("test_api_san", [os.path.join(root, "tests", "cpp", "test_api.cpp")]),
                       ("test_host_sanitize", [os.path.join(root, "tests", "cpp", "test_host_sanitize.cpp"),
                                               os.path.join(root, "oracle", "oracle.cpp")])):
        out = os.path.join(out_dir, name)
        cmd = [clang, "-std=c++17", "-O1", "-g", "-I", os.path.join(root, "include")] + san + srcs + \
              [lib, f"-Wl,-rpath,{out_dir}", "-Wl,-rpath,/opt/rocm/lib", "-o", out]
        r = _run(cmd, verbose)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building {name} failed")
        progs.append(out)
    return tuple(progs)


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_test(verbose=True))
