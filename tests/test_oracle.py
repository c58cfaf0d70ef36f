"""Pins the CPU oracle (oracle/refcpu.c).

1. Against the reference's own tree code: golden vectors in tests/golden/
   written by oracle/_ref (tree.f90 compiled from the reference sources):
   node tables, forces, energies and a 10-step trajectory, bit for bit.
2. Where the reference harness is built (this container), against a live run
   on fresh inputs.
3. Against closed forms that a plausible transcription slip would break:
   theta -> 0 equals the direct sum; tiny trees by hand; centroid = mean;
   D == 0; the pair forces are minus the gradients of the pair energies;
   velocity Verlet on a constant force is the exact quadratic.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle import pyoracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RS10 = 6.6493e-3
EPS = 0.0014399835661145824


def _load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _oracle_table(x, sign):
    """Oracle node table in the reference harness's record layout."""
    t = pyoracle.export_tree(x, sign)
    n = len(t["depth"])
    ileaf = np.where(t["leaf"] == 1, sign, 0)
    pt = np.where(t["leaf"] == 1, t["pt"] + 1, 0)
    ints = np.stack([t["depth"], ileaf, pt, t["next"]], axis=1).astype(np.int32)
    dbl = np.concatenate([t["xx"], t["lx"], t["qeff"][:, None], t["dx"][:, None], t["D"], t["Q"], t["QQ"]], axis=1)
    return ints.reshape(n, 4), dbl


@pytest.mark.parametrize("name", ["ref_n96.npz", "ref_lattice.npz"])
def test_tree_tables_match_reference(name):
    g = _load(name)
    for sp, (x, sign) in enumerate(((g["xe"], -1), (g["xi"], 1))):
        ints, dbl = _oracle_table(x, sign)
        ref_i, ref_d = g["tree%d_int" % sp], g["tree%d_dbl" % sp]
        # the reference leaves pt of a twig as 0 and its root qeff starts from 0 here and there
        assert ints.shape == ref_i.shape
        assert np.array_equal(ints[:, [0, 1, 3]], ref_i[:, [0, 1, 3]])
        assert np.array_equal(np.where(ints[:, 1] != 0, ints[:, 2], 0), np.where(ref_i[:, 1] != 0, ref_i[:, 2], 0))
        assert np.array_equal(dbl, ref_d)


@pytest.mark.parametrize("name", ["ref_n96.npz", "ref_lattice.npz"])
def test_forces_match_reference(name):
    g = _load(name)
    fe, fi, ep, _ = pyoracle.local_coulomb_sum(g["xe"], g["xi"], float(g["theta"]), float(g["rs"]), nthreads=1)
    assert np.array_equal(fe, g["fe"]) and np.array_equal(fi, g["fi"])
    assert np.array_equal(ep, g["epot"])


def test_c1_matches_reference():
    """C1 (4096 + 4096, sigma 100 um): forces, Epot and node-table digests."""
    import tree_amd as ta
    g = _load("ref_c1.npz")
    qel, qion = ta.gen_plasma(1, 4096, 4096, 100.0)
    xe, xi = qel["xx"].copy(), qion["xx"].copy()
    assert _digest(np.r_[xe, xi]) == str(g["pos_digest"])
    fe, fi, ep, _ = pyoracle.local_coulomb_sum(xe, xi, 0.5, RS10, nthreads=1)
    assert np.array_equal(fe, g["fe"]) and np.array_equal(fi, g["fi"]) and np.array_equal(ep, g["epot"])
    for sp, (x, sign) in enumerate(((xe, -1), (xi, 1))):
        ints, dbl = _oracle_table(x, sign)
        assert len(ints) == g["nnodes"][sp]
        ints[:, 2] = np.where(ints[:, 1] != 0, ints[:, 2], 0)
        assert _digest(ints) == str(g["tree%d_int_digest" % sp])
        assert _digest(dbl) == str(g["tree%d_dbl_digest" % sp])


def test_trajectory_matches_reference():
    """10 steps of VV1 / build + force / VV2 equal the reference's own integrator bit for bit."""
    g = _load("ref_traj.npz")
    xe, ve, xi, vi = g["xe"], g["ve"], g["xi"], g["vi"]
    th, rs, dt = float(g["theta"]), float(g["rs"]), float(g["dt"])
    xme, xmi = 5.4858e-4, 85.0
    fe, fi, _, _ = pyoracle.local_coulomb_sum(xe, xi, th, rs, nthreads=1)
    for k in range(10):
        xe, ve = pyoracle.vv1(xe, ve, fe, xme, dt)
        xi, vi = pyoracle.vv1(xi, vi, fi, xmi, dt)
        fe, fi, ep, _ = pyoracle.local_coulomb_sum(xe, xi, th, rs, nthreads=1)
        ve, se = pyoracle.vv2(xe, ve, fe, xme, dt)
        vi, si = pyoracle.vv2(xi, vi, fi, xmi, dt)
        assert np.array_equal(np.r_[ep, se[0], si[0], se[1], si[1]], g["steps"][k]), k
    assert np.array_equal(np.c_[xe, ve, fe], g["e"]) and np.array_equal(np.c_[xi, vi, fi], g["i"])


@pytest.mark.parametrize("name", ["ref_n96.npz", "ref_lattice.npz"])
def test_rdf_matches_reference(name):
    """Leaf-pair RDF histograms (tree.f90:916-919, 1021-1024, 1125-1129) equal the reference's, bin for bin."""
    g = _load(name)
    dx = float(g["rdf_dx"])
    *_, hei, hii = pyoracle.local_coulomb_sum_rdf(g["xe"], g["xi"], float(g["theta"]), float(g["rs"]), dx, dx,
                                                  nthreads=1)
    assert np.array_equal(hei, g["rdf_ei"]) and np.array_equal(hii, g["rdf_ii"])


def test_rdf_c1_matches_reference():
    import tree_amd as ta
    g = _load("ref_c1.npz")
    qel, qion = ta.gen_plasma(1, 4096, 4096, 100.0)
    dx = float(g["rdf_dx"])
    *_, hei, hii = pyoracle.local_coulomb_sum_rdf(qel["xx"], qion["xx"], 0.5, RS10, dx, dx, nthreads=4)
    assert np.array_equal(hei, g["rdf_ei"]) and np.array_equal(hii, g["rdf_ii"])


def test_rdf_totals_are_the_leaf_counts():
    """With bins covering every distance the histograms hold exactly the walk's
    leaf interactions: ie + ei for RDF_ei, ii (beyond the cut) for RDF_ii."""
    rng = np.random.default_rng(8)
    xe, xi = rng.normal(size=(600, 3)) * 50.0, rng.normal(size=(550, 3)) * 50.0
    *_, c, hei, hii = pyoracle.local_coulomb_sum_rdf(xe, xi, 0.5, RS10, 0.1, 0.1, nthreads=2)
    assert hei.sum() == c[5] + c[8] and hii.sum() == c[11]
    assert hei.sum() > 0 and hii.sum() > 0


def test_rdf_theta_zero_is_the_pair_histogram():
    """theta -> 0 makes every pair a leaf pair: the histograms become the histograms
    of all ordered e-i pairs (seen from both sides) and i-i pairs beyond 0.1 um,
    binned as INT(r / dx) + 1 up to Ndist + 1."""
    rng = np.random.default_rng(9)
    xe, xi = rng.normal(size=(70, 3)) * 3.0, rng.normal(size=(60, 3)) * 3.0
    dx = 0.0013
    *_, hei, hii = pyoracle.local_coulomb_sum_rdf(xe, xi, 1e-12, RS10, dx, dx, nthreads=1)
    nb = pyoracle.NDIST + 1

    def hist(r):
        idx = np.floor(r / dx).astype(np.int64) + 1
        idx = idx[(idx >= 1) & (idx <= nb)]
        return np.bincount(idx - 1, minlength=nb)

    rei = np.sqrt(((xe[:, None, :] - xi[None, :, :]) ** 2).sum(-1)).ravel()
    dii = xi[:, None, :] - xi[None, :, :]
    r2 = (dii ** 2).sum(-1)
    rii = np.sqrt(r2[r2 > 0.01])
    assert np.array_equal(hei, 2 * hist(rei)) and np.array_equal(hii, hist(rii))
    assert (rei / dx < nb).sum() > 100 and (rei / dx >= nb).sum() > 100  # both sides of the last bin


@pytest.mark.skipif(not pyoracle.ref_available(), reason="reference harness not built here")
def test_live_reference_random():
    rng = np.random.default_rng(77)
    xe, xi = rng.normal(size=(700, 3)) * 80.0, rng.normal(size=(650, 3)) * 80.0
    z = np.zeros_like
    buf = pyoracle.run_ref(1, xe, z(xe), xi, z(xi), 0.6, RS10, (5.4858e-4, 85.0), 0.05)
    trees, fe_r, fi_r, ep_r = pyoracle.parse_ref_mode1(buf, 700, 650)
    fe, fi, ep, _ = pyoracle.local_coulomb_sum(xe, xi, 0.6, RS10, nthreads=1)
    assert np.array_equal(fe, fe_r) and np.array_equal(fi, fi_r) and np.array_equal(ep, ep_r)
    for sp, (x, sign) in enumerate(((xe, -1), (xi, 1))):
        _, dbl = _oracle_table(x, sign)
        assert np.array_equal(dbl, trees[sp]["d"])


# ------------------------------------------------------------ closed forms

def test_theta_zero_is_direct_sum():
    rng = np.random.default_rng(5)
    xe, xi = rng.normal(size=(300, 3)) * 20.0, rng.normal(size=(280, 3)) * 20.0
    fe, fi, ep, cnt = pyoracle.local_coulomb_sum(xe, xi, 1e-300, RS10)
    assert cnt[1] == cnt[4] == cnt[7] == cnt[10] == 0  # nothing accepted
    tsp = np.r_[np.zeros(300, np.int32), np.ones(280, np.int32)]
    tix = np.r_[np.arange(300), np.arange(280)].astype(np.int32)
    fd, pot = pyoracle.direct(xe, xi, RS10, tsp, tix)
    f = np.r_[fe, fi]
    assert np.abs(f - fd).max() <= 1e-12 * np.abs(fd).max()
    # Epot = 1/2 sum over targets of their pair energies (e: ee+ei+ie halves, i: ii)
    assert np.isclose(ep[0] + ep[1], 0.5 * pot.sum(), rtol=1e-12)


def test_two_particle_chain():
    """Two particles: a chain down to the first level where they separate."""
    x = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    t = pyoracle.export_tree(x, -1)
    # root box [0,1]^3 centre 0.5: they separate at once -> root + 2 leaves
    assert list(t["depth"]) == [0, 1, 1] and list(t["leaf"]) == [0, 1, 1]
    x = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.01, 0.0, 0.0]])
    t = pyoracle.export_tree(x, -1)
    # particle 0 and 2 share the low octant down to the level where cells are < 0.01 wide
    d = t["depth"][t["leaf"] == 1]
    assert d.max() == 1 + int(np.ceil(np.log2(1.0 / 0.01))) - 1 or d.max() == int(np.ceil(np.log2(1 / 0.01)))
    assert t["count"][0] == 3 and t["qeff"][0] == -3.0


def test_one_per_octant_depth_one():
    c = np.array([[sx, sy, sz] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)], dtype=np.float64)
    t = pyoracle.export_tree(c, 1)
    assert list(t["depth"]) == [0] + [1] * 8
    # children in the reference's order 1..8: x fastest, then y, then z
    assert np.array_equal(t["pt"][1:], np.arange(8))
    assert np.array_equal(t["xx"][0], [0, 0, 0]) and t["dx"][0] == 0.0
    assert np.allclose(t["Q"][0], 4.0) and np.allclose(t["QQ"][0], 0.0)  # sum q x^2/2 = 8*1/2


def test_centroid_and_dipole():
    rng = np.random.default_rng(9)
    x = rng.normal(size=(500, 3)) * 50.0
    t = pyoracle.export_tree(x, -1)
    assert np.allclose(t["xx"][0], x.mean(axis=0), rtol=0, atol=1e-12 * 50)
    tw = t["leaf"] == 0
    assert np.abs(t["D"][tw]).max() <= 1e-9 * 50 * 500  # D == 0 up to rounding
    assert np.all(t["qeff"][tw] == -t["count"][tw])


def test_pair_forces_are_energy_gradients():
    """Finite differences of the direct pair energies give the direct forces."""
    rng = np.random.default_rng(3)
    xe, xi = rng.normal(size=(40, 3)) * 0.2, rng.normal(size=(40, 3)) * 0.2
    h = 1e-6
    for tsp, arr in ((0, xe), (1, xi)):
        f, _ = pyoracle.direct(xe, xi, RS10, [tsp], [3])
        g = np.zeros(3)
        for k in range(3):
            p, m = arr.copy(), arr.copy()
            p[3, k] += h
            m[3, k] -= h
            xe_p, xi_p = (p, xi) if tsp == 0 else (xe, p)
            xe_m, xi_m = (m, xi) if tsp == 0 else (xe, m)
            _, ep = pyoracle.direct(xe_p, xi_p, RS10, [tsp], [3])
            _, em = pyoracle.direct(xe_m, xi_m, RS10, [tsp], [3])
            g[k] = -(ep[0] - em[0]) / (2 * h)
        assert np.allclose(f[0], g, rtol=1e-5, atol=1e-9 * np.abs(f).max())


def test_kelbg_limits():
    """phi -> 1/rs - r/(2 rs^2) for r << rs and -> 1/r for r >> rs (one unlike pair)."""
    for r, expect in ((1e-5, 1 / RS10 - 1e-5 / (2 * RS10 ** 2)), (10.0, 1 / 10.0)):
        xe = np.array([[0.0, 0.0, 0.0]])
        xi = np.array([[r, 0.0, 0.0]])
        _, pot = pyoracle.direct(xe, xi, RS10, [0], [0])
        assert np.isclose(pot[0], -EPS * expect, rtol=1e-6)


def test_like_cutoff_excludes_close_pairs():
    xe = np.array([[0.0, 0.0, 0.0], [0.09, 0.0, 0.0], [5.0, 0.0, 0.0]])
    xi = np.array([[100.0, 0.0, 0.0]])
    f, _ = pyoracle.direct(xe, xi, RS10, [0], [0])
    fd, _ = pyoracle.direct(xe[[0, 2]], xi, RS10, [0], [0])
    assert np.array_equal(f, fd)


def test_verlet_constant_force_exact():
    x = np.array([[1.0, -2.0, 3.0]])
    v = np.array([[0.5, 0.25, -1.0]])
    f = np.array([[0.125, -0.5, 2.0]])
    m, dt = 2.0, 0.25
    x1, v1 = pyoracle.vv1(x, v, f, m, dt)
    v2, s = pyoracle.vv2(x1, v1, f, m, dt)
    a = f / m
    assert np.allclose(x1, x + v * dt + 0.5 * a * dt * dt, rtol=0, atol=1e-15)
    assert np.allclose(v2, v + a * dt, rtol=0, atol=1e-15)
    assert np.isclose(s[0], m * (v2 ** 2).sum())
    assert np.isclose(s[1], m * ((x1 * v2).sum()) ** 2 / (x1 ** 2).sum())
