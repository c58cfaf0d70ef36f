"""The Fortran driver's input checks (no GPU: they run before the engine is
created).  The reference STOPs on a particle-count mismatch
(main.f90:415-421); the driver also STOPs with a message, instead of a
runtime error, on a short or malformed input.xyz and on a malformed
control.prm."""
import os
import subprocess

import pytest

import tree_amd as ta
from tree_amd import plasma_io

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tree_amd", "bin", "octree_gpu")
pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built")


def _run(d):
    r = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=60)
    return r.stdout + r.stderr


def _case(tmp_path, ne=20, ni=20, header=None, drop=0, rest="OFF"):
    qel, qion = ta.gen_plasma(3, ne, ni, 10.0)
    plasma_io.write_control(tmp_path / "control.prm", ne, ni, rest=rest)
    plasma_io.write_xyz(tmp_path / "input.xyz", qel, qion, velocities=rest != "OFF")
    lines = (tmp_path / "input.xyz").read_text().splitlines()
    if header is not None:
        lines[0] = str(header)
    if drop:
        lines = lines[:-drop]
    (tmp_path / "input.xyz").write_text("\n".join(lines) + "\n")


def test_count_mismatch_stops(tmp_path):
    _case(tmp_path, header=41)
    out = _run(tmp_path)
    assert "particle count differs from control.prm" in out and "parsing error" in out, out
    assert not (tmp_path / "res001.xyz").exists()


def test_short_input_stops(tmp_path):
    _case(tmp_path, drop=3)
    out = _run(tmp_path)
    assert "input.xyz: short or malformed" in out and "38 of" in out, out


def test_restart_input_without_velocities_stops(tmp_path):
    _case(tmp_path)
    plasma_io.write_control(tmp_path / "control.prm", 20, 20, rest="ON")
    out = _run(tmp_path)
    assert "input.xyz: short or malformed" in out, out


def test_malformed_control_stops(tmp_path):
    _case(tmp_path)
    txt = (tmp_path / "control.prm").read_text().splitlines()
    k = txt.index("#temper")
    txt[k + 1] = "ten"
    (tmp_path / "control.prm").write_text("\n".join(txt) + "\n")
    out = _run(tmp_path)
    assert "cannot read temperatures" in out and "control.prm: malformed" in out, out


def test_missing_control_stops(tmp_path):
    out = _run(tmp_path)
    assert "control.prm: cannot open" in out, out
