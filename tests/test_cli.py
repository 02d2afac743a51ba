"""The se2wave executable's CPU-only paths: -verify_stf (se2wave.c:4809-4914)
and option handling."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "se2dr_fork_amd", "bin", "se2wave")


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_verify_stf_ricker(tmp_path, oracle):
    r = run(["-verify_stf", "-stf_ricker", "-tmax", "0.3", "-dt", "1e-3", "-stf.ricker.freq", "5"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "[se2wave] Requested time period: 3.0000e-01" in r.stdout
    text = (tmp_path / "SeismicSTF_Ricker.log").read_text().splitlines()
    header = [l for l in text if l.startswith("#")]
    assert '#     type  "Ricker"' in header
    assert "#     freq  +5.000000000000e+00 (Hz)" in header  # the option is printed ...
    rows = np.array([[float(x) for x in l.split()] for l in text if not l.startswith("#")])
    # ... but the evaluated STF is Ricker(0.15, 12, 1), as in the reference (se2wave.c:4837)
    t = 0.0
    for row in rows:
        assert row[0] == pytest.approx(t, abs=5e-7)
        assert float("%+1.12e" % oracle.stf("ricker", [0.15, 12.0, 1.0], t)) == row[1]
        t = t + 1e-3
    assert len(rows) == 300


def test_verify_stf_yoffe(tmp_path, oracle):
    r = run(["-verify_stf", "-stf_yoffe", "-tmax", "1.4", "-dt", "1e-2"], tmp_path)
    assert r.returncode == 0, r.stderr
    text = (tmp_path / "SeismicSTF_Yoffe.log").read_text().splitlines()
    rows = np.array([[float(x) for x in l.split()] for l in text if not l.startswith("#")])
    t = 0.0
    for row in rows:
        assert float("%+1.12e" % oracle.stf("yoffe", [0.15, 1.0, 1.1], t)) == row[1]
        t = t + 1e-2


def test_verify_stf_needs_a_choice(tmp_path):
    r = run(["-verify_stf"], tmp_path)
    assert r.returncode != 0
    assert "No STF implementation specified" in r.stderr


def test_bad_options_are_rejected(tmp_path):
    r = run(["-verify_stf", "-stf_ricker", "-tmax", "abc"], tmp_path)
    assert r.returncode != 0 and "not a real number" in r.stderr
