"""Shared test setup: repository root on sys.path and the `gpu` marker.

Tests marked `gpu` need an MI355X; everything else runs on the CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs an AMD GPU (MI355X) and the built HIP library")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def se2w():
    import se2dr_fork_amd as S

    S.libraries()
    return S


def read_vts(path):
    """Arrays of a VTK XML StructuredGrid file with raw appended data (UInt64
    block headers, little-endian Float64): {name: flat array}."""
    import re

    data = open(path, "rb").read()
    head, sep, raw = data.partition(b'<AppendedData encoding="raw">\n_')
    assert sep, "no raw appended data"
    out = {}
    for m in re.finditer(rb'Name="([^"]+)" NumberOfComponents="(\d+)" format="appended" offset="(\d+)"', head):
        off = int(m.group(3))
        nbytes = int(np.frombuffer(raw, dtype="<u8", count=1, offset=off)[0])
        out[m.group(1).decode()] = np.frombuffer(raw, dtype="<f8", count=nbytes // 8, offset=off + 8)
    return out
