"""Build recipes of the backend's native libraries (in-tree, no JIT cache).

* ``capaz_amd/libcapaz_amd.so`` -- the product: HIP kernels for gfx950 plus the
  C ABI of ``include/capaz.h``, built by hipcc.
* ``tests/emul/libcapaz_emul.so`` -- the same per-cell kernel bodies compiled
  for the host by g++, so tests can check the kernels' arithmetic against the
  oracle without a GPU (test infrastructure; never loaded by the product).
* ``tests/hip/libcapaz_selftest.so`` -- device checks of the kernels' exact
  arithmetic rewrites against IEEE operations (test infrastructure).

All code is compiled without floating-point contraction: the kernels must
reproduce the reference's IEEE operation sequence, and the only FMAs are the
explicit ones of the exact division.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcapaz_amd.so")
EMUL_SRC = os.path.join(ROOT, "tests", "emul", "capaz_emul.cpp")
EMUL_LIB = os.path.join(ROOT, "tests", "emul", "libcapaz_emul.so")
SELFTEST_SRC = os.path.join(ROOT, "tests", "hip", "capaz_selftest.hip")
SELFTEST_LIB = os.path.join(ROOT, "tests", "hip", "libcapaz_selftest.so")

ARCH = os.environ.get("CAPAZ_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    # packed FP32 instructions are kept out of every HIP build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _hipcc() -> str:
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stdout.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {cmd[0]}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature; keep other warnings visible
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out)


def build_library() -> str:
    srcs = [os.path.join(CSRC, f) for f in
            ("capaz_kernels.hip", "capaz_batch.hip", "capaz_adjoint.hip", "capaz_api.cpp", "capaz_batch_api.cpp",
             "capaz_adjoint_api.cpp", "capaz_state.cpp", "capaz_plot.cpp", "capaz_noalloc.hip", "capaz_energy2.hip")]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), *HIP_FLAGS, "-x", "hip", *srcs, "-o", tmp]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_emulator() -> str:
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-march=x86-64-v3", "-fno-builtin-pow", "-I", CSRC, EMUL_SRC, "-o", EMUL_LIB, "-lm"]
    _run(cmd)
    return EMUL_LIB


def build_selftest() -> str:
    tmp = SELFTEST_LIB + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-I", CSRC, "-x", "hip", SELFTEST_SRC, "-o", tmp])
    os.replace(tmp, SELFTEST_LIB)
    return SELFTEST_LIB


def build_all() -> None:
    build_library()
    if os.path.exists(EMUL_SRC):
        build_emulator()
    if os.path.exists(SELFTEST_SRC):
        build_selftest()


if __name__ == "__main__":
    build_all()
