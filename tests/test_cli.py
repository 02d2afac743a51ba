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


def test_source_impl_option_errors(tmp_path):
    """-source_impl takes the reference's numbering 0..3 (se2wave.c:2202-2229);
    the cubic-spline and P0 forms assemble moment sources only."""
    r = run(["-mx", "4", "-source_impl", "7"], tmp_path)
    assert r.returncode != 0 and "valid choices are 0,1,2,3" in r.stderr
    r = run(["-mx", "4", "-source_impl", "3", "-src_xy", "10,10", "-src_type", "force", "-src_values", "1,0,0,0",
             "-src_ricker", "0.1,10,1"], tmp_path)
    assert r.returncode != 0 and "assemble moment sources" in r.stderr
    assert "[source implementation]: P0" in r.stdout
