"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host library's
sources and of the FP64 oracle (CPU only).

tests/sanitize/driver.c exercises every host entry point that needs no GPU,
the -verify_stf program path and an oracle run with every optional feature;
it is compiled here with gcc -fsanitize=address,undefined together with
se2dr_fork_amd/csrc/host/*.c and oracle/ref_cpu.c (linked against the built
libse2w_hip.so for the GPU symbols the solver references, none of which the
driver calls) and run.  Any sanitizer report fails the test.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HOST = os.path.join(ROOT, "se2dr_fork_amd", "csrc", "host")
LIB = os.path.join(ROOT, "se2dr_fork_amd", "lib")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_host_library_and_oracle_under_asan_ubsan(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libse2w_hip.so")):
        pytest.skip("libse2w_hip.so is not built")
    exe = str(tmp_path / "driver")
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".c"))
    cmd = ["gcc", "-std=gnu99", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off", "-fopenmp", "-pthread",
           "-I", os.path.join(ROOT, "include"), "-I", HOST,
           os.path.join(ROOT, "tests", "sanitize", "driver.c"), *srcs,
           "-L", LIB, "-lse2w_hip", "-lm", f"-Wl,-rpath,{LIB}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["LSAN_OPTIONS"] = f"suppressions={os.path.join(ROOT, 'tests', 'sanitize', 'lsan.supp')}"
    env["OMP_NUM_THREADS"] = "2"
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=300, env=env, cwd=str(tmp_path))
    report = r.stdout[-3000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "sanitized driver ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
