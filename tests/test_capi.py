"""The C-ABI libraries load and export every function their headers declare.

No GPU call is made here; the HIP entry points are exercised by the `gpu`
tests.
"""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADERS = {"se2w.h": "libse2w_hip.so", "se2w_mlp.h": "libse2w_hip.so", "se2wave.h": "libse2w_host.so"}


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(se2w_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(se2w, header):
    lib = os.path.join(ROOT, "se2dr_fork_amd", "lib", HEADERS[header])
    names = declared_functions(header)
    assert len(names) > 10
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, f"{header}: not exported by {HEADERS[header]}: {missing}"
    handle = ctypes.CDLL(lib)
    for n in names:
        getattr(handle, n)


def test_error_strings(se2w):
    hip, _ = se2w.libraries()
    for code, text in [(0, b"success"), (-1, b"invalid argument"), (-5, b"unsupported configuration")]:
        assert hip.se2w_strerror(code) == text
    assert hip.se2w_version() >= 1
    assert hip.se2w_kernel_count() >= 2


def test_c_headers_compile_as_c99():
    """The ABI headers are plain C (no torch or C++ types in the signatures)."""
    src = '#include "se2w.h"\n#include "se2w_mlp.h"\n#include "se2wave.h"\nint main(void){return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_missing_library_raises(tmp_path):
    """Without the built library the package raises; it has no CPU fallback."""
    code = ("import se2dr_fork_amd as S\n"
            f"S.LIB_DIR = {str(tmp_path)!r}\n"
            "try:\n"
            "    S.libraries()\n"
            "except S.Se2wError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    raise SystemExit('no error')\n")
    r = subprocess.run(["python", "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "is missing" in r.stdout


def test_context_without_gpu_raises(se2w):
    """On a host without a GPU, creating a step context fails with the HIP
    error instead of running anywhere else."""
    try:
        ctx = se2w.Context(mx=4, my=4, p=2, precision=se2w.FP32)
    except se2w.Se2wError as e:
        assert "HIP" in str(e) or "device" in str(e), e
        return
    ctx.close()
    pytest.skip("a GPU is present")
