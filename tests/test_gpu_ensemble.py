"""Batched ensemble: realisations of uneven size stepped in one persistent
walk launch give the same forces and energies as each realisation run alone
on its own handle (same mixed walk, same group size)."""
import numpy as np
import pytest

import tree_amd as ta
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def test_ensemble_matches_single_runs():
    announce("ensemble_vs_single")
    rng = np.random.default_rng(4)
    sizes = [3000, 17000, 800, 9000, 25000]
    tes = [1.0, 10.0, 100.0, 3.0, 30.0]
    cfgs = [ta.make_config(rs=ta.kelbg_length(t), theta=0.5, theta_mixed=0.5, walk="mixed", group=128)
            for t in tes]
    data = [ta.gen_plasma(100 + r, n // 2, n // 2, 120.0 + 10 * r, te=tes[r]) for r, n in enumerate(sizes)]
    dt = 0.005 / ta.TPS
    with ta.Ensemble(cfgs) as ens:
        for r, (qe, qi) in enumerate(data):
            ens.set_particles(r, qe, qi)
        ep = ens.force()
        got = [ens.get_particles(r) for r in range(len(sizes))]
        out = ens.step(dt, 3)
        after = [ens.get_particles(r) for r in range(len(sizes))]
    for r, (qe, qi) in enumerate(data):
        with ta.Engine(rs=ta.kelbg_length(tes[r]), theta=0.5, theta_mixed=0.5, walk="mixed", group=128) as eng:
            eng.set_particles(qe, qi)
            eng.build()
            ep1 = eng.force()
            e1, i1 = eng.get_particles()
            out1 = eng.step(dt, 3)
            e2, i2 = eng.get_particles()
        # same kernel, same groups: forces identical; energies differ only by summation order
        assert np.array_equal(got[r][0]["ff"], e1["ff"]) and np.array_equal(got[r][1]["ff"], i1["ff"])
        assert np.allclose(ep[r], ep1, rtol=1e-12)
        assert np.array_equal(after[r][0]["xx"], e2["xx"]) and np.array_equal(after[r][1]["xv"], i2["xv"])
        assert np.allclose(out[r], out1, rtol=1e-12)
    # the autotuner (last: it sets the list lengths later walks of these sizes use)
    with ta.Ensemble(cfgs) as ens:
        for r, (qe, qi) in enumerate(data):
            ens.set_particles(r, qe, qi)
        ens.force()
        g, ms = ens.autotune()
        assert g in (64, 128, 256) and np.all(ms > 0)
        assert ens.walk_lists() in ((11, 24), (9, 16), (7, 12))
