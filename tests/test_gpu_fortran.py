"""End-to-end run of the Fortran driver (tree_amd/bin/octree_gpu) through
the ISO_C_BINDING interface: control.prm + input.xyz in, ener###.dat,
rdf###.dat and res###.xyz out, checked against the same steps driven from
Python through the C-ABI (ES14.6 output: 1e-6 relative)."""
import os
import subprocess

import numpy as np
import pytest

import tree_amd as ta
from tree_amd import plasma_io
from tests.conftest import announce

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tree_amd", "bin", "octree_gpu")
KB = 8.617343e-5


@pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built")
def test_fortran_driver_matches_capi(tmp_path):
    announce("fortran_driver")
    ne = ni = 1500
    rs = 6.6493e-3
    qel, qion = ta.gen_plasma(8, ne, ni, 100.0)
    plasma_io.write_control(tmp_path / "control.prm", ne, ni, te=10.0, ti=1e-3, tmax=0.05, trest=0.025,
                            tdump=0.025, tsamp=0.005, dt=0.005, rs=rs, theta=0.5)
    plasma_io.write_xyz(tmp_path / "input.xyz", qel, qion)
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    e0 = plasma_io.read_ener(tmp_path / "ener000.dat")
    e1 = plasma_io.read_ener(tmp_path / "ener001.dat")
    assert e0.shape == (5, 9) and e1.shape == (5, 9)
    f_e, f_i = plasma_io.read_res(tmp_path / "res002.xyz")
    assert len(f_e) == ne and len(f_i) == ni
    rdf_files = [plasma_io.read_rdf(tmp_path / ("rdf%03d.dat" % k)) for k in (1, 2)]
    assert all(r.shape == (ta.NDIST + 1, 4) for r in rdf_files)

    # the same run through the C-ABI from Python
    v_e, v_i = ta.gen_plasma(1, ne, ni, 1.0, te=10.0, ti=1e-3)
    qel["xv"], qion["xv"] = v_e["xv"], v_i["xv"]
    dt = 0.005 / ta.TPS
    kbm_e, kbm_i = KB * (3 * ne - 6), KB * (3 * ni - 6)
    kbm_er, kbm_ir, kbm_ec, kbm_ic = KB * (ne - 6), KB * (ni - 6), KB * (2 * ne - 6), KB * (2 * ni - 6)
    rows = []
    rdfs = []
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        eng.rdf_enable(0.01, 0.01)
        for k in range(10):
            eng.vv1(dt)
            eng.build()
            ep = eng.force()
            s = eng.vv2(dt, True)
            if (k + 1) % 5 == 0:
                rdfs.append(eng.rdf_get(reset=True))
            tel, tion, trad_e, trad_i = s
            rows.append([(k + 1) * dt * ta.TPS, tel / kbm_e, tion / kbm_i, trad_e / kbm_er, trad_i / kbm_ir,
                         (tel - trad_e) / kbm_ec, (tion - trad_i) / kbm_ic, ep[0] + ep[1],
                         0.5 * (tel + tion) + ep[0] + ep[1]])
        e, i = eng.get_particles()
    ref = np.array(rows)
    got = np.r_[e0, e1]
    assert np.allclose(got, ref, rtol=2e-6, atol=1e-12), np.abs(got - ref).max(axis=0)
    assert np.allclose(f_e["xx"], e["xx"], rtol=1e-6, atol=1e-9)
    r = np.arange(1, ta.NDIST + 2) * 0.01
    for got_rdf, (hei, hii) in zip(rdf_files, rdfs):
        assert hei.sum() > 0 and hii.sum() > 0
        assert np.allclose(got_rdf[:, 0], r, rtol=1e-6) and np.allclose(got_rdf[:, 2], r, rtol=1e-6)
        assert np.allclose(got_rdf[:, 1], hei / 5.0, rtol=1e-6, atol=0)
        assert np.allclose(got_rdf[:, 3], hii / 5.0, rtol=1e-6, atol=0)
    assert np.array_equal(f_i["id"], np.arange(1, ni + 1))
    _check_records(tmp_path)


def _check_records(d):
    """Every record of the driver's ener / res / rdf files is laid out as the
    reference's FORMAT statements write it (plasma_io's formatters, pinned to
    amdflang's output by tests/test_formats.py)."""
    for path in sorted(d.glob("ener*.dat")):
        for ln in path.read_text().splitlines()[1:]:
            vals = [plasma_io.fortran_float(v) for v in ln.split()]
            assert ln == plasma_io.format_ener(vals), (path.name, ln)
    for path in sorted(d.glob("rdf*.dat")):
        for ln in path.read_text().splitlines()[1:50]:
            vals = [plasma_io.fortran_float(v) for v in ln.split()]
            assert ln == plasma_io.format_rdf(vals), (path.name, ln)
    for path in sorted(d.glob("res*.xyz")):
        for ln in path.read_text().splitlines()[2:]:
            t = ln.split()
            v = [plasma_io.fortran_float(x) for x in t[1:7]]
            assert ln == plasma_io.format_res(t[0], v[:3], v[3:], int(t[7])), (path.name, ln)


@pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built")
def test_fortran_driver_resume(tmp_path):
    """Resume (rest /= 'OFF', main.f90:333-380 reading, Vinit_REST main.f90:796-834):
    10 steps in one run against 5 steps, res001.xyz, and 5 more steps resumed
    from it.  The resumed run starts from the 7-digit positions and velocities
    of the restart file, so it matches the uninterrupted run to that
    precision: positions within 1e-6 of the largest coordinate, velocities
    and energies (T_e, T_i, Epot, Etot) within 1e-4 relative."""
    announce("fortran_driver_resume")
    ne = ni = 1200
    rs = 6.6493e-3
    qel, qion = ta.gen_plasma(9, ne, ni, 100.0)
    full, first, second = tmp_path / "full", tmp_path / "first", tmp_path / "second"
    for d, tmax in ((full, 0.05), (first, 0.025)):
        d.mkdir()
        plasma_io.write_control(d / "control.prm", ne, ni, tmax=tmax, trest=0.025, tdump=0.025, tsamp=0.005,
                                dt=0.005, rs=rs, theta=0.5)
        plasma_io.write_xyz(d / "input.xyz", qel, qion)
        r = subprocess.run([BIN], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        _check_records(d)
    second.mkdir()
    plasma_io.write_control(second / "control.prm", ne, ni, tmax=0.025, trest=0.025, tdump=0.025, tsamp=0.005,
                            dt=0.005, rs=rs, theta=0.5, rest="ON")
    (second / "input.xyz").write_text((first / "res001.xyz").read_text())
    r = subprocess.run([BIN], cwd=second, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _check_records(second)
    a_e, a_i = plasma_io.read_res(full / "res002.xyz")
    b_e, b_i = plasma_io.read_res(second / "res001.xyz")
    assert np.array_equal(a_e["id"], b_e["id"]) and np.array_equal(a_i["id"], b_i["id"])
    for a, b in ((a_e, b_e), (a_i, b_i)):
        assert np.abs(a["xx"] - b["xx"]).max() <= 1e-6 * np.abs(a["xx"]).max()
        assert np.abs(a["xv"] - b["xv"]).max() <= 1e-4 * np.abs(a["xv"]).max()
    ea = plasma_io.read_ener(full / "ener001.dat")    # steps 6..10
    eb = plasma_io.read_ener(second / "ener000.dat")  # the resumed run's steps 1..5
    assert ea.shape == eb.shape == (5, 9)
    for c in (1, 2, 7, 8):
        assert np.allclose(ea[:, c], eb[:, c], rtol=1e-4, atol=0), (c, ea[:, c], eb[:, c])


@pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built")
def test_fortran_driver_mixed_walk_writes_rdf(tmp_path):
    """TREEMI_WALK=mixed: the driver steps with the mixed walk and still writes
    rdf###.dat (every pair in range, see rdf32.hip), in the reference's record
    format, with the histograms of the same steps driven through the C-ABI."""
    announce("fortran_driver_mixed")
    ne = ni = 1000
    rs = 6.6493e-3
    qel, qion = ta.gen_plasma(10, ne, ni, 100.0)
    plasma_io.write_control(tmp_path / "control.prm", ne, ni, tmax=0.025, trest=0.025, tdump=0.025, tsamp=0.005,
                            dt=0.005, rs=rs, theta=0.4)
    plasma_io.write_xyz(tmp_path / "input.xyz", qel, qion)
    env = dict(os.environ, TREEMI_WALK="mixed")
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    _check_records(tmp_path)
    got = plasma_io.read_rdf(tmp_path / "rdf001.dat")
    v_e, v_i = ta.gen_plasma(1, ne, ni, 1.0, te=10.0, ti=1e-3)
    qel["xv"], qion["xv"] = v_e["xv"], v_i["xv"]
    dt = 0.005 / ta.TPS
    with ta.Engine(rs=rs, theta=0.4, theta_mixed=0.4, walk="mixed") as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        eng.rdf_enable(0.01, 0.01)
        for k in range(5):
            eng.vv1(dt)
            eng.build()
            eng.force()
            eng.vv2(dt, True)
        hei, hii = eng.rdf_get()
    assert hei.sum() > 0 and hii.sum() > 0
    assert np.allclose(got[:, 1], hei / 5.0, rtol=1e-6, atol=0) and np.allclose(got[:, 3], hii / 5.0, rtol=1e-6, atol=0)
