"""GPU parity of the plain FP64 path against the CPU oracle.

Integer and structural results (membership, order, node table, next
pointers) must be bit-exact; FP64 moments bit-exact (contraction off on both
sides); forces within 1e-12 of the largest force component (exp() in the
Kelbg leaves may round differently on the GPU), Epot within 1e-12 relative.
"""
import numpy as np
import pytest

import tree_amd as ta
from oracle import pyoracle
from tests.conftest import announce

pytestmark = pytest.mark.gpu

RS10 = 6.6493e-3


def _engine(**kw):
    kw.setdefault("rs", RS10)
    kw.setdefault("theta", 0.5)
    return ta.Engine(**kw)


def _plasma(seed, ne, ni, sigma):
    return ta.gen_plasma(seed, ne, ni, sigma)


def _cmp_tree(g, o):
    n = len(o["depth"])
    assert len(g["depth"]) == n
    for k in ("depth", "leaf", "count", "next"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(np.where(g["leaf"] == 1, g["pt"], -1), np.where(o["leaf"] == 1, o["pt"], -1))
    for k in ("xx", "lx", "qeff", "dx", "D", "Q", "QQ"):
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("seed,n,sigma", [(1, 3000, 100.0), (2, 1, 100.0), (3, 2, 100.0), (4, 20000, 400.0)])
def test_tree_bitexact(seed, n, sigma):
    announce("tree_bitexact seed=%d n=%d" % (seed, n))
    qel, qion = _plasma(seed, n, n + 3, sigma)
    with _engine() as eng:
        eng.set_particles(qel, qion)
        eng.build()
        for sp, q, sign in ((0, qel, -1), (1, qion, 1)):
            _cmp_tree(eng.export_tree(sp), pyoracle.export_tree(q["xx"], sign))


def test_tree_ties_and_lattice():
    """Particles on cell centres (ties go high) and an exact lattice."""
    announce("tree_ties_and_lattice")
    g = np.arange(-3, 4, dtype=np.float64) * 10.0
    lat = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T.copy()
    qel = np.zeros(len(lat), ta.PT_DTYPE)
    qel["xx"] = lat
    qion = np.zeros(len(lat), ta.PT_DTYPE)
    qion["xx"] = lat * 0.5 + 1.25
    with _engine() as eng:
        eng.set_particles(qel, qion)
        eng.build()
        _cmp_tree(eng.export_tree(0), pyoracle.export_tree(lat, -1))
        _cmp_tree(eng.export_tree(1), pyoracle.export_tree(lat * 0.5 + 1.25, 1))


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_force_vs_oracle(theta):
    announce("force_vs_oracle theta=%g" % theta)
    qel, qion = _plasma(11, 4096, 4096, 100.0)
    fe_o, fi_o, ep_o, cnt_o = pyoracle.local_coulomb_sum(qel["xx"], qion["xx"], theta, RS10)
    with _engine(theta=theta) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        ep = eng.force()
        e, i = eng.get_particles()
        st = eng.count_interactions()
    for f, fo in ((e["ff"], fe_o), (i["ff"], fi_o)):
        assert np.abs(f - fo).max() <= 1e-12 * np.abs(fo).max()
    assert np.allclose(ep, ep_o, rtol=1e-12, atol=0)
    # interaction counts equal the oracle's (same MAC decisions): ee, ie, ei, ii
    order = [(0, 1, 2), (3, 4, 5), (6, 7, 8), (9, 10, 11)]
    for p, (a, b, c) in enumerate(order):
        assert st["visits"][p] == cnt_o[a] and st["m2p"][p] == cnt_o[b] and st["p2p"][p] == cnt_o[c]


def test_theta_tiny_equals_direct():
    """theta -> 0 opens every twig: the walk is the direct sum."""
    announce("theta_tiny_equals_direct")
    qel, qion = _plasma(12, 700, 650, 50.0)
    with _engine(theta=1e-12) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        e, i = eng.get_particles()
        tsp = np.r_[np.zeros(700, np.int32), np.ones(650, np.int32)]
        tix = np.r_[np.arange(700), np.arange(650)].astype(np.int32)
        fd = eng.direct_sample(tsp, tix)
    f = np.r_[e["ff"], i["ff"]]
    assert np.abs(f - fd).max() <= 1e-11 * np.abs(fd).max()


def test_direct_vs_oracle():
    announce("direct_vs_oracle")
    qel, qion = _plasma(13, 3000, 2900, 100.0)
    rng = np.random.default_rng(99)
    tsp = rng.integers(0, 2, 256).astype(np.int32)
    tix = np.where(tsp == 0, rng.integers(0, 3000, 256), rng.integers(0, 2900, 256)).astype(np.int32)
    fo, _ = pyoracle.direct(qel["xx"], qion["xx"], RS10, tsp, tix)
    with _engine() as eng:
        eng.set_particles(qel, qion)
        fd = eng.direct_sample(tsp, tix)
    assert np.abs(fd - fo).max() <= 1e-12 * np.abs(fo).max()


def test_verlet_vs_oracle():
    """VV1, build, force, VV2 for 3 steps against the oracle doing the same."""
    announce("verlet_vs_oracle")
    dt = 0.005 / ta.TPS  # 5 ps
    qel, qion = _plasma(14, 1500, 1500, 100.0)
    xe, ve, xi, vi = qel["xx"].copy(), qel["xv"].copy(), qion["xx"].copy(), qion["xv"].copy()
    fe, fi, _, _ = pyoracle.local_coulomb_sum(xe, xi, 0.5, RS10)
    with _engine() as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        for step in range(3):
            xe, ve = pyoracle.vv1(xe, ve, fe, ta.XM_E, dt)
            xi, vi = pyoracle.vv1(xi, vi, fi, ta.XM_I, dt)
            fe, fi, ep_o, _ = pyoracle.local_coulomb_sum(xe, xi, 0.5, RS10)
            ve, se = pyoracle.vv2(xe, ve, fe, ta.XM_E, dt)
            vi, si = pyoracle.vv2(xi, vi, fi, ta.XM_I, dt)
            eng.vv1(dt)
            eng.build()
            ep = eng.force()
            sums = eng.vv2(dt, True)
            assert np.allclose(ep, ep_o, rtol=1e-11, atol=0)
            assert np.allclose(sums, [se[0], si[0], se[1], si[1]], rtol=1e-12, atol=0)
        e, i = eng.get_particles()
    for a, b in ((e["xx"], xe), (i["xx"], xi), (e["xv"], ve), (i["xv"], vi)):
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()


def test_step_fused_equals_separate():
    """tb_step (fused KDK between steps) equals the separate VV1/build/force/VV2 calls."""
    announce("step_fused_equals_separate")
    dt = 0.005 / ta.TPS
    qel, qion = _plasma(15, 2000, 2000, 100.0)
    with _engine() as a, _engine() as b:
        for eng in (a, b):
            eng.set_particles(qel, qion)
            eng.build()
            eng.force()
        out = a.step(dt, 4)
        for k in range(4):
            b.vv1(dt)
            b.build()
            ep = b.force()
            sums = b.vv2(dt, k == 3)
        ea, ia = a.get_particles()
        eb, ib = b.get_particles()
    assert np.array_equal(ea["xx"], eb["xx"]) and np.array_equal(ia["xv"], ib["xv"])
    assert np.allclose(out[:2], ep, rtol=1e-14) and np.allclose(out[2:], [sums[0], sums[1], sums[2], sums[3]],
                                                                  rtol=1e-14)


def test_rdf_matches_reference_and_oracle():
    """The FP64 walk's leaf-pair RDF histograms (tree.f90:916-919, 1021-1024,
    1125-1129) equal the reference's own at C1 (4096 + 4096, theta 0.5, dx 0.01 um)
    bin for bin; two forces add twice; the counting pass adds nothing; reset
    clears.  With dx = 0.001 um most pairs land beyond the block-local bins
    (global atomics) and the result still equals the oracle's."""
    import os
    announce("rdf C1")
    gold = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_c1.npz"), allow_pickle=False))
    qel, qion = ta.gen_plasma(1, 4096, 4096, 100.0)
    dx = float(gold["rdf_dx"])
    with _engine(theta=0.5) as eng:
        eng.set_particles(qel, qion)
        eng.rdf_enable(dx, dx)
        eng.build()
        eng.force()
        ei, ii = eng.rdf_get()
        assert np.array_equal(ei, gold["rdf_ei"]) and np.array_equal(ii, gold["rdf_ii"])
        eng.force()
        eng.count_interactions()
        ei2, ii2 = eng.rdf_get(reset=True)
        assert np.array_equal(ei2, 2 * gold["rdf_ei"]) and np.array_equal(ii2, 2 * gold["rdf_ii"])
        ei3, ii3 = eng.rdf_get()
        assert not ei3.any() and not ii3.any()
        eng.rdf_enable(0.001, 0.0015)
        eng.force()
        ei, ii = eng.rdf_get()
    *_, oei, oii = pyoracle.local_coulomb_sum_rdf(qel["xx"], qion["xx"], 0.5, RS10, 0.001, 0.0015, nthreads=4)
    assert np.array_equal(ei, oei) and np.array_equal(ii, oii)
    assert ei[2048:].sum() > ei[:2048].sum() > 0
    with _engine(theta=0.5, walk="mixed", theta_mixed=0.5) as eng:
        with pytest.raises(ta.TreeError):  # the mixed walk's histograms live in LDS: at most 10240 bins
            eng.rdf_enable(dx, dx, nbin=10241)


def test_errors():
    announce("errors")
    qel, qion = _plasma(16, 100, 100, 100.0)
    with _engine() as eng:
        with pytest.raises(ta.TreeError) as ex:
            eng.force()
        assert ex.value.code == 5  # STATE: force before build
        q2 = qel.copy()
        q2["xx"][5] = q2["xx"][17]
        eng.set_particles(q2, qion)
        with pytest.raises(ta.TreeError) as ex:
            eng.build()
        assert ex.value.code == 3  # DEPTH: coincident
        q3 = qel.copy()
        q3["xx"][3, 1] = np.nan
        eng.set_particles(q3, qion)
        with pytest.raises(ta.TreeError) as ex:
            eng.build()
        assert ex.value.code == 6  # NONFINITE
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
    with pytest.raises(ta.TreeError) as ex:
        ta.Engine(rs=RS10, theta=0.0)
    assert ex.value.code == 1
    with pytest.raises(ta.TreeError) as ex:
        e = ta.Engine(rs=RS10, theta=0.5, mem_cap=1 << 16)
        e.set_particles(*_plasma(1, 100000, 100000, 400.0))
    assert ex.value.code == 4


def test_force_targets_and_sampled_timing():
    """The target-list walk gives exactly tb_force's forces for those targets."""
    announce("force_targets")
    qel, qion = _plasma(17, 5000, 4000, 100.0)
    rng = np.random.default_rng(3)
    tsp = rng.integers(0, 2, 300).astype(np.int32)
    tix = np.where(tsp == 0, rng.integers(0, 5000, 300), rng.integers(0, 4000, 300)).astype(np.int32)
    with _engine(theta=0.4) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        e, i = eng.get_particles()
        f = eng.force_targets(tsp, tix)
        ms, nw = eng.time_walk(8)
        # blocks 0, 8, 16, ... of 64 targets: 10 of 79 electron blocks, 8 of 63 ion blocks (all whole)
        assert nw == 10 * 64 + 8 * 64, nw
    fr = np.where(tsp[:, None] == 0, e["ff"][np.where(tsp == 0, tix, 0)], i["ff"][np.where(tsp == 1, tix, 0)])
    assert np.array_equal(f, fr)
    assert ms > 0
