"""AddressSanitizer and UndefinedBehaviorSanitizer build of the CPU code
paths: the oracle's C restatement (oracle/capaz_oracle.c) and the host build
of the kernels' per-cell bodies (capaz_amd/csrc/capaz_bodies.h), driven side
by side by tests/sanitize/sanitize_driver.cpp on small, thin and
one-cell-wide grids of every model.  Any out-of-bounds access, use of
uninitialised heap memory or undefined arithmetic (signed overflow in the
index math, invalid shifts) aborts the run; the two implementations must
also agree bit for bit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
IEEE = ["-ffp-contract=off", "-fno-fast-math", "-fno-builtin-pow"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_oracle_and_kernel_bodies_under_asan_ubsan(tmp_path):
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "sanitize_driver")
    subprocess.run(["gcc", "-std=c11", *SAN, *IEEE, "-c", os.path.join(ROOT, "oracle", "capaz_oracle.c"), "-o", obj],
                   check=True)
    subprocess.run(["g++", "-std=c++17", *SAN, *IEEE, "-I", os.path.join(ROOT, "capaz_amd", "csrc"),
                    "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "sanitize", "sanitize_driver.cpp"),
                    obj, "-o", exe, "-lm"], check=True)
    # verify_asan_link_order=0: the sanitizer runtime need not be the first
    # library of the process (a preloaded library may come before it)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "sanitize: ok" in r.stdout
