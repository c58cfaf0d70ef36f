"""The output records of the driver against the reference's formats.

tests/golden/formats.txt holds records written by amdflang with the
reference's FORMAT statements (TEMPORAL sample.f90:90-93, Restart
print.f90:73-106, RDF sample.f90:36-43) for fixed values; the formatters in
tree_amd.plasma_io must reproduce them character for character (the GPU
driver test then checks the driver's files against those formatters)."""
import os

import pytest

from tree_amd import plasma_io

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture():
    lines = open(os.path.join(HERE, "golden", "formats.txt")).read().splitlines()
    out = {}
    for k in range(0, len(lines), 2):
        head = lines[k].split()
        assert head[0] == "#V"
        kind = head[1]
        nums = head[2:]
        out[kind] = (nums, lines[k + 1])
    return out


@pytest.mark.parametrize("kind", ["ener", "res_P", "res_O", "rdf"])
def test_formatters_match_fortran_records(kind):
    nums, rec = _fixture()[kind]
    if kind == "ener":
        got = plasma_io.format_ener([float(v) for v in nums])
    elif kind == "rdf":
        got = plasma_io.format_rdf([float(v) for v in nums])
    else:
        v = [float(x) for x in nums[:6]]
        got = plasma_io.format_res(kind[-1], v[:3], v[3:], int(nums[6]))
    assert got == rec, (got, rec)


def test_readers_parse_fortran_records(tmp_path):
    fx = _fixture()
    (tmp_path / "e.dat").write_text("# header\n" + fx["ener"][1] + "\n")
    e = plasma_io.read_ener(tmp_path / "e.dat")
    assert e.shape == (1, 9) and e[0, 1] == -2.5e-3 and e[0, 5] == 1e-120
    (tmp_path / "r.xyz").write_text("2\n frame = 1 energy = 0\n" + fx["res_P"][1] + "\n" + fx["res_O"][1] + "\n")
    qe, qi = plasma_io.read_res(tmp_path / "r.xyz")
    assert len(qe) == 1 and len(qi) == 1 and qe["id"][0] == 12 and qi["id"][0] == 12345678
    assert qe["xx"][0][1] == -2.5e-3 and qi["xv"][0][2] == 0.33
