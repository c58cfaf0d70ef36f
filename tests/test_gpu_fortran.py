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
