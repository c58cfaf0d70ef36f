"""The mixed walk's far-field local expansion (tree_amd/csrc/local_exp.h) on
the GPU against the FP64 oracle (oracle/local_ref.py), and the walk with its
far-cell path against the walk without it."""
import numpy as np
import pytest

import tree_amd as ta
from oracle import local_ref as lr
from tests.conftest import announce

pytestmark = pytest.mark.gpu

RS10 = 6.6493e-3


def _cells(rng, n, dmin, dmax):
    u = rng.normal(size=(n, 3))
    R = u / np.linalg.norm(u, axis=1)[:, None] * rng.uniform(dmin, dmax, size=n)[:, None]
    q = rng.normal(size=n)
    D = rng.normal(size=(n, 3)) * 0.3
    A = rng.normal(size=(n, 3, 3)) * 0.1
    Qp = A + A.transpose(0, 2, 1)
    Qp -= np.trace(Qp, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    # a traceless symmetric third moment from random point-charge offsets
    O = np.stack([lr.traceless3(lr.third_moment(rng.normal(size=(5, 3)) * 0.3, rng.normal(size=5)))
                  for _ in range(n)])
    # and a traceless symmetric fourth moment
    H = np.stack([lr.traceless4(lr.fourth_moment(rng.normal(size=(5, 3)) * 0.3, rng.normal(size=5)))
                  for _ in range(n)])
    cells = np.c_[R, q, D, Qp[:, 0, 0], Qp[:, 1, 1], Qp[:, 0, 1], Qp[:, 1, 2], Qp[:, 2, 0],
                  O[:, 0, 0, 0], O[:, 0, 0, 1], O[:, 0, 0, 2], O[:, 0, 1, 1], O[:, 0, 1, 2], O[:, 1, 1, 1], O[:, 1, 1, 2],
                  H[:, 0, 0, 0, 0], H[:, 0, 0, 0, 1], H[:, 0, 0, 0, 2], H[:, 0, 0, 1, 1], H[:, 0, 0, 1, 2],
                  H[:, 0, 1, 1, 1], H[:, 0, 1, 1, 2], H[:, 1, 1, 1, 1], H[:, 1, 1, 1, 2]]
    return cells, R, q, D, Qp, O, H


@pytest.mark.parametrize("order", [3, 4, 5, 6])
def test_local_expansion_vs_oracle(order):
    """FP32 coefficients summed over 150 cells (a partial last pass of the
    wavefront) and evaluated at 70 offsets: within 2e-5 of the FP64 oracle,
    relative to the field's scale.  Orders 4 to 6 (the harmonic forms; 6 is
    the walk's, with the terms to total order 7) include the cells' third
    moments, and order 6 their fourth moments."""
    announce("local_eval order %d" % order)
    rng = np.random.default_rng(order)
    cells, R, q, D, Qp, O, H = _cells(rng, 150, 5.0, 20.0)
    y = rng.uniform(-1.0, 1.0, size=(70, 3))
    with ta.Engine(rs=RS10) as eng:
        out = eng.local_eval(order, cells, y)
    L = {}
    for k in range(len(cells)):
        Lk = lr.m2l(-R[k], q[k], D[k], Qp[k], np.zeros(3), order, O=O[k] if order >= 4 else None,
                    tmax=7 if order == 6 else None, H=H[k] if order == 6 else None)
        for b, v in Lk.items():
            L[b] = L.get(b, 0.0) + v
    ref = np.array([[*lr.l2p(L, yy)[:1], *lr.l2p(L, yy)[1]] for yy in y])
    escale = np.abs(ref[:, 1:]).max()
    pscale = np.abs(ref[:, 0]).max()
    assert np.abs(out[:, 1:] - ref[:, 1:]).max() <= 2e-5 * escale
    assert np.abs(out[:, 0] - ref[:, 0]).max() <= 2e-5 * pscale
    if order == 3:
        # and against the cells' own fields (M2P, FP64): the local truncation, ~(|y|/R)^order
        direct = np.zeros((len(y), 4))
        for k in range(len(cells)):
            for j, yy in enumerate(y):
                ph, E = lr.m2p(-R[k], q[k], D[k], Qp[k], yy)
                direct[j, 0] += ph
                direct[j, 1:] += E
        rho = np.sqrt(3.0) / 5.0
        assert np.abs(out[:, 1:] - direct[:, 1:]).max() <= 30 * rho ** order * np.abs(direct[:, 1:]).max()


def _forces(eng, ne, ni):
    e, i = eng.get_particles()
    return np.r_[e["ff"], i["ff"]]


def test_walk_far_cells_through_local_expansion():
    """The same accepted cells with and without the far path: the M2P cells
    with it plus its far cells equal the M2P cells without it (the route
    changes, the acceptance does not); forces differ by less than the walk's
    own error; both stay close to the FP64 direct sum; the total potential
    energy stays within the local expansion's truncation (1e-4 relative at
    K = 8, 1e-3 at K = 4) of an accurate one (the walk without the far path at
    theta 0.1)."""
    announce("walk_far_path")
    n = 1 << 16
    qel, qion = ta.gen_plasma(31, n, n, 300.0)
    rng = np.random.default_rng(4)
    tsp = np.r_[np.zeros(512, np.int32), np.ones(512, np.int32)]
    tix = np.r_[rng.choice(n, 512, replace=False), rng.choice(n, 512, replace=False)].astype(np.int32)
    out = {}
    for far in (-1.0, 8.0, 4.0):
        with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.3, walk="mixed", far_ratio=far) as eng:
            eng.set_particles(qel, qion)
            if far < 0:
                fd = eng.direct_sample(tsp, tix)
            eng.build()
            ep = eng.force()
            f = _forces(eng, n, n)
            st = eng.count_interactions()
            out[far] = (f, ep, st)
    with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.1, walk="mixed", far_ratio=-1.0) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        ep_ref = sum(eng.force())
    f0, ep0, st0 = out[-1.0]
    assert sum(st0["m2l"]) == 0
    for far in (8.0, 4.0):
        f1, ep1, st1 = out[far]
        assert sum(st1["m2l"]) > 0
        assert sum(st1["m2p"]) + sum(st1["m2l"]) == sum(st0["m2p"]), (far, st0, st1)
        assert st1["p2p"] == st0["p2p"]
        d = np.linalg.norm(f1 - f0, axis=1) / np.linalg.norm(f0, axis=1)
        bound = 2e-3 if far == 8.0 else 1e-2
        assert np.sqrt(np.mean(d ** 2)) < bound, (far, np.sqrt(np.mean(d ** 2)))
        ebound = 1e-4 if far == 8.0 else 1e-3
        assert abs(sum(ep1) - ep_ref) <= ebound * abs(ep_ref), (far, ep0, ep1, ep_ref)
    idx = np.where(tsp == 0, tix, n + tix)
    for far in (-1.0, 8.0):
        f = out[far][0][idx]
        rel = np.linalg.norm(f - fd, axis=1) / np.linalg.norm(fd, axis=1)
        out[far] = out[far] + (np.sqrt(np.mean(rel ** 2)),)
    assert out[8.0][3] < 1.3 * out[-1.0][3], (out[-1.0][3], out[8.0][3])


def test_far_buffer_regrows(monkeypatch):
    """A far-cell chunk buffer too small for the walk (one chunk) is detected
    after the walk, grown, and the walk re-run: forces and energies are
    bit-identical to a run whose first buffer was large enough."""
    announce("far_buffer_regrow")
    n = 1 << 14
    qel, qion = ta.gen_plasma(33, n, n, 200.0)
    out = []
    for chunks in (None, "1"):
        if chunks:
            monkeypatch.setenv("TREEMI_FAR_CHUNKS", chunks)
        else:
            monkeypatch.delenv("TREEMI_FAR_CHUNKS", raising=False)
        with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.4, walk="mixed") as eng:
            eng.set_particles(qel, qion)
            eng.build()
            ep = eng.force()
            e, i = eng.get_particles()
            st = eng.count_interactions()
            out.append((np.r_[e["ff"], i["ff"]], ep, st))
    assert sum(out[0][2]["m2l"]) > 0
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_far_buffer_regrows_learned(monkeypatch):
    """The learned walk's re-run after a far-buffer regrow decides its band
    cells as the first attempt would: the first attempt has overwritten the
    forces its E_min feature reads, so the re-run takes E_min saved before
    it.  A first force at theta 0.6 sizes the buffer; the second, at 0.35,
    needs more far chunks (smaller accepted cells) and is re-run, with the
    first force's field as its previous force.  Forces and energies are
    bit-identical to the same two forces with a buffer large from the start."""
    announce("far_buffer_regrow learned")
    n = 1 << 14
    qel, qion = ta.gen_plasma(34, n, n, 200.0)
    with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.37, theta_lo=0.33, walk="mixed") as eng:
        eng.set_particles(qel, qion)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 18)
        assert len(X) > 1000
        eng.mlp_init(0)
        eng.mlp_train(X, y, steps=300, batch=4096, lr=2e-3)
        params = eng.mlp_get()
    out = []
    for chunks in (None, "1"):
        if chunks:
            monkeypatch.setenv("TREEMI_FAR_CHUNKS", chunks)
        else:
            monkeypatch.delenv("TREEMI_FAR_CHUNKS", raising=False)
        with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.6, walk="mixed") as eng:
            eng.set_particles(qel, qion)
            eng.mlp_set(params)
            eng.build()
            eng.force()
            eng.configure(theta_mixed=0.37, theta_lo=0.33, mac="learned", err_target=1e-5)
            eng.build()
            ep = eng.force()
            e, i = eng.get_particles()
            st = eng.count_interactions()
            out.append((np.r_[e["ff"], i["ff"]], ep, st))
    assert sum(out[0][2]["m2l"]) > 0
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("mac", ["fixed", "learned"])
def test_two_level_far_field(mac):
    """far_levels 2 (segments of up to 8 groups walk the top of the tree once;
    cells far from the whole segment go through the segment's local
    expansion) against the group level alone: the same M2P, P2P and per-target
    far counts (acceptance and route per target are unchanged), fewer cells
    visited, forces within the local expansions' truncation of each other, and
    no further from the FP64 direct sum.  The learned criterion runs a
    network from tb_mlp_init (any network: both runs take the same decisions)."""
    announce("two_level_far %s" % mac)
    n = 1 << 16
    qel, qion = ta.gen_plasma(35, n, n, 300.0)
    rng = np.random.default_rng(6)
    tsp = np.r_[np.zeros(512, np.int32), np.ones(512, np.int32)]
    tix = np.r_[rng.choice(n, 512, replace=False), rng.choice(n, 512, replace=False)].astype(np.int32)
    idx = np.where(tsp == 0, tix, n + tix)
    out = {}
    for lv in (1, 2):
        with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.3, walk="mixed", mac=mac, group=128,
                       far_levels=lv) as eng:
            if mac == "learned":
                eng.mlp_init(3)
            eng.set_particles(qel, qion)
            if lv == 1:
                fd = eng.direct_sample(tsp, tix)
            eng.build()
            ep = eng.force()
            out[lv] = (_forces(eng, n, n), ep, eng.count_interactions())
    (f1, ep1, s1), (f2, ep2, s2) = out[1], out[2]
    assert sum(s1["seg_pairs"]) == 0 and sum(s2["seg_pairs"]) > 0
    assert s2["m2p"] == s1["m2p"] and s2["p2p"] == s1["p2p"] and s2["m2l"] == s1["m2l"], (s1, s2)
    assert sum(s2["far_pairs"]) < sum(s1["far_pairs"])
    assert sum(s2["visits"]) < sum(s1["visits"])
    d = np.linalg.norm(f2 - f1, axis=1) / np.linalg.norm(f1, axis=1)
    assert np.sqrt(np.mean(d ** 2)) < 2e-4, np.sqrt(np.mean(d ** 2))
    assert abs(sum(ep2) - sum(ep1)) <= 1e-5 * abs(sum(ep1)), (ep1, ep2)
    e = []
    for f in (f1, f2):
        rel = np.linalg.norm(f[idx] - fd, axis=1) / np.linalg.norm(fd, axis=1)
        e.append(np.sqrt(np.mean(rel ** 2)))
    assert e[1] < 1.1 * e[0], e


def test_two_level_far_buffer_regrows(monkeypatch):
    """With two levels the segments' frontier and far cells share the far-cell
    chunk buffer: one chunk to start with is detected, grown and re-run,
    bit-identical to a run whose buffer was large enough."""
    announce("two_level_far_regrow")
    n = 1 << 14
    qel, qion = ta.gen_plasma(37, n, n, 200.0)
    out = []
    for chunks in (None, "1"):
        if chunks:
            monkeypatch.setenv("TREEMI_FAR_CHUNKS", chunks)
        else:
            monkeypatch.delenv("TREEMI_FAR_CHUNKS", raising=False)
        with ta.Engine(rs=RS10, theta=0.5, theta_mixed=0.4, walk="mixed", far_levels=2) as eng:
            eng.set_particles(qel, qion)
            eng.build()
            ep = eng.force()
            out.append((_forces(eng, n, n), ep, eng.count_interactions()))
    assert sum(out[0][2]["seg_pairs"]) > 0
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]
