"""Every configuration of the workload family, run on the GPU at its full size.

C1  4096 e- + 4096 ions, sigma 100 um, theta 0.5, dt 5 ps: 100 velocity-Verlet
    steps of the FP64 walk against the oracle stepping the same plasma
    (main.f90:123-167 time loop, vverlet.f90:9-112, tree.f90:39-1133).
C2  2^19 + 2^19, sigma 400 um: the FP64 walk against the oracle's walk on 4096
    seeded targets; the mixed walk against the FP64 direct sum; the sampled
    walk timing (every stride-th block of 64 targets) against a full force.
C3  2^21 + 2^21, sigma 600 um: the mixed walk at theta 0.3 / 0.5 / 0.7 for
    T_e 1 / 10 / 100 K, and at the headline's theta (rms <= 1e-4).
C4  64 realisations, N_total log-uniform in [16k, 256k], in one ensemble:
    forces equal to each realisation run alone, and within the walk's error
    of the FP64 direct sum.

Error sample: 4096 targets (2048 e- + 2048 ions, seed 99), rms and max of
|F - F_dir| / |F_dir|.  At these sizes the direct sum on the GPU
(tb_direct_sample, FP64) is the reference; it is itself checked here against
the CPU oracle's direct sum on 64 of the targets at the same full size.
"""
import time

import numpy as np
import pytest

import tree_amd as ta
from oracle import pyoracle
from tests.conftest import announce

pytestmark = pytest.mark.gpu

DT = 0.005 / ta.TPS  # 5 ps


def sample_targets(ne, ni, seed=99, n=4096):
    rng = np.random.default_rng(seed)
    te = rng.choice(ne, n // 2, replace=False)
    ti = rng.choice(ni, n // 2, replace=False)
    tsp = np.r_[np.zeros(n // 2, np.int32), np.ones(n // 2, np.int32)]
    return tsp, np.r_[te, ti].astype(np.int32)


def rel_err(f, fd):
    rel = np.linalg.norm(f - fd, axis=1) / np.linalg.norm(fd, axis=1)
    return float(np.sqrt(np.mean(rel ** 2))), float(rel.max())


def pick(e, i, tsp, tix):
    return np.where(tsp[:, None] == 0, e["ff"][np.where(tsp == 0, tix, 0)], i["ff"][np.where(tsp == 1, tix, 0)])


def direct_reference(qel, qion, rs, tsp, tix):
    """FP64 direct sum on the GPU, its first 32 e- and 32 ion targets checked
    against the CPU oracle's direct sum (1e-12 of the largest component)."""
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        fd = eng.direct_sample(tsp, tix)
    n2 = len(tsp) // 2
    sub = np.r_[np.arange(32), n2 + np.arange(32)]
    fo, _ = pyoracle.direct(qel["xx"], qion["xx"], rs, tsp[sub], tix[sub])
    assert np.abs(fd[sub] - fo).max() <= 1e-12 * np.abs(fo).max()
    return fd


def test_c1_hundred_steps_vs_oracle():
    """C1 as configured: 100 steps.  The oracle's own trajectories separate
    from a 1-ulp change of the initial positions to 2e-14 (positions) and
    2e-13 (velocities) of their largest values over these 100 steps, so the
    bounds here (1e-10, 1e-9, energies 1e-10) leave room only for rounding
    (exp() in the Kelbg leaves), not for a different algorithm."""
    announce("c1_hundred_steps")
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(1, 4096, 4096, 100.0)
    xe, ve, xi, vi = qel["xx"].copy(), qel["xv"].copy(), qion["xx"].copy(), qion["xv"].copy()
    fe, fi, _, _ = pyoracle.local_coulomb_sum(xe, xi, 0.5, rs)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        for step in range(100):
            xe, ve = pyoracle.vv1(xe, ve, fe, ta.XM_E, DT)
            xi, vi = pyoracle.vv1(xi, vi, fi, ta.XM_I, DT)
            fe, fi, ep_o, _ = pyoracle.local_coulomb_sum(xe, xi, 0.5, rs)
            ve, se = pyoracle.vv2(xe, ve, fe, ta.XM_E, DT)
            vi, si = pyoracle.vv2(xi, vi, fi, ta.XM_I, DT)
            eng.vv1(DT)
            eng.build()
            ep = eng.force()
            sums = eng.vv2(DT, True)
            assert np.allclose(ep, ep_o, rtol=1e-10, atol=0), (step, ep, ep_o)
            assert np.allclose(sums, [se[0], si[0], se[1], si[1]], rtol=1e-10, atol=0), step
            if step % 10 == 9 or step == 99:
                e, i = eng.get_particles()
                for a, b, tol in ((e["xx"], xe, 1e-10), (i["xx"], xi, 1e-10), (e["xv"], ve, 1e-9),
                                  (i["xv"], vi, 1e-9)):
                    assert np.abs(a - b).max() <= tol * np.abs(b).max(), step


def test_c2_fp64_walk_vs_oracle_full_size():
    """C2 at full size (2^19 + 2^19): the FP64 walk's forces on 4096 sampled
    targets equal the oracle's walk of the same trees (1e-12 of the largest
    component), and a full tb_force gives those targets the same forces."""
    announce("c2_fp64_vs_oracle")
    n = 1 << 19
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(2, n, n, 400.0)
    tsp, tix = sample_targets(n, n)
    fo = pyoracle.force_targets(qel["xx"], qion["xx"], 0.5, rs, tsp, tix)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        fg = eng.force_targets(tsp, tix)
        eng.force()
        e, i = eng.get_particles()
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    assert np.array_equal(pick(e, i, tsp, tix), fg)


def test_c2_mixed_walk_error():
    """C2, mixed walk at theta 0.5 against the FP64 direct sum: rms < 1.3e-3,
    max < 0.05 (measured 1.04e-3 / 0.024); the FP64 walk at the same theta:
    rms < 3e-2 (measured 2.0e-2; per-species expansions of a quasi-neutral
    plasma)."""
    announce("c2_mixed_error")
    n = 1 << 19
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(2, n, n, 400.0)
    tsp, tix = sample_targets(n, n)
    fd = direct_reference(qel, qion, rs, tsp, tix)
    with ta.Engine(rs=rs, theta=0.5, theta_mixed=0.5, walk="mixed") as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        e, i = eng.get_particles()
    rms, mx = rel_err(pick(e, i, tsp, tix), fd)
    print("C2 mixed theta 0.5: rms %.3e max %.3e" % (rms, mx))
    assert rms < 1.3e-3 and mx < 0.05, (rms, mx)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        rms64, mx64 = rel_err(eng.force_targets(tsp, tix), fd)
    print("C2 FP64 theta 0.5: rms %.3e max %.3e" % (rms64, mx64))
    assert rms64 < 3e-2, (rms64, mx64)


def test_c2_sampled_walk_timing_matches_full_walk():
    """The plain walk's sampled timing (every stride-th block of 64
    consecutive targets of each species, scaled to all targets) against a
    full force at C2, theta 0.15: within 10% for strides 4 and 16 (per-target
    walks are independent; whole wavefront blocks keep the full walk's
    coherence between neighbouring lanes)."""
    announce("c2_sampled_timing")
    n = 1 << 19
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(2, n, n, 400.0)
    with ta.Engine(rs=rs, theta=0.15) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        full = []
        for _ in range(2):
            eng.force()
            full.append(eng.stats()["ms_force"])
        t_full = min(full)
        for stride in (4, 16):
            runs = [eng.time_walk(stride) for _ in range(2)]
            ms, nsamp = min(r[0] for r in runs), runs[0][1]
            scaled = ms * 2 * n / nsamp
            print("C2 theta 0.15: full walk %.1f ms, stride %d scaled %.1f ms" % (t_full, stride, scaled))
            assert abs(scaled / t_full - 1.0) < 0.10, (stride, scaled, t_full)


@pytest.mark.parametrize("te", [1.0, 10.0, 100.0])
def test_c3_mixed_walk_theta_sweep(te):
    """C3 at full size (2^21 + 2^21, sigma 600 um): the mixed walk at theta
    0.3 / 0.5 / 0.7 against the FP64 direct sum, rms below 3.8e-5 / 3.9e-4 /
    3.2e-3 (measured 3.22e-5 / 3.25e-4 / 2.51e-3 with far cells beyond 3.5
    group radii through the local expansion and cells beyond 1.5 opening radii
    to the third moment, alike for every T_e: the Kelbg
    length only changes close unlike pairs), max below 0.1 (measured 0.067 at
    theta 0.7: one target of weak net force); the error rises with theta.  The target groups follow the species' key order over the joint
    box."""
    announce("c3_sweep te=%g" % te)
    n = 1 << 21
    rs = ta.kelbg_length(te)
    qel, qion = ta.gen_plasma(3, n, n, 600.0, te=te)
    tsp, tix = sample_targets(n, n)
    fd = direct_reference(qel, qion, rs, tsp, tix)
    errs = []
    with ta.Engine(rs=rs, theta=0.5, theta_mixed=0.3, walk="mixed") as eng:
        eng.set_particles(qel, qion)
        rows = []
        for th, bound in ((0.3, 3.8e-5), (0.5, 3.9e-4), (0.7, 3.2e-3)):
            eng.configure(theta_mixed=th)
            eng.build()
            eng.force()
            e, i = eng.get_particles()
            rms, mx = rel_err(pick(e, i, tsp, tix), fd)
            print("C3 T_e %g theta %.1f: rms %.3e max %.3e" % (te, th, rms, mx))
            rows.append((th, bound, rms, mx))
            errs.append(rms)
    for th, bound, rms, mx in rows:
        assert np.isfinite(rms) and rms < bound and mx < 0.1, rows
    assert errs[0] < errs[1] < errs[2], errs


def test_c3_headline_theta_reaches_1e4():
    """C3 at the headline's accuracy: the mixed walk at theta 0.38 (the
    bench's tuned theta is ~0.395) is within rms 1e-4 of the FP64 direct sum
    (max < 3e-3), at T_e 10 K."""
    announce("c3_headline_theta")
    n = 1 << 21
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(3, n, n, 600.0)
    tsp, tix = sample_targets(n, n)
    fd = direct_reference(qel, qion, rs, tsp, tix)
    with ta.Engine(rs=rs, theta=0.5, theta_mixed=0.38, walk="mixed") as eng:
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        e, i = eng.get_particles()
    rms, mx = rel_err(pick(e, i, tsp, tix), fd)
    print("C3 theta 0.38: rms %.3e max %.3e" % (rms, mx))
    assert rms <= 1e-4 and mx < 3e-3, (rms, mx)


def c4_realisations():
    """64 realisations: N_total log-uniform in [16k, 256k], T_e log-uniform in
    [1, 100] K, sigma from a total peak density of 1e15 m^-3."""
    rng = np.random.default_rng(1000)
    R = 64
    ntot = np.exp(rng.uniform(np.log(16384), np.log(262144), R)).astype(int) // 2 * 2
    tes = np.exp(rng.uniform(0.0, np.log(100.0), R))
    out = []
    for r in range(R):
        sigma = (ntot[r] / ((2 * np.pi) ** 1.5 * 1e15)) ** (1 / 3) * 1e6
        qe, qi = ta.gen_plasma(1000 + r, int(ntot[r]) // 2, int(ntot[r]) // 2, sigma, te=float(tes[r]))
        out.append((float(tes[r]), qe, qi))
    return out


def test_c4_ensemble_all_64_realisations():
    """C4 as configured: all 64 realisations in one ensemble (one persistent
    walk launch).  Every realisation's forces equal those of the same
    realisation run alone on its own handle (same kernel and groups), and on
    256 sampled targets each is within the mixed walk's error at theta 0.5 of
    the oracle's FP64 direct sum (rms < 3e-3 per realisation)."""
    announce("c4_ensemble_64")
    reals = c4_realisations()
    cfgs = [ta.make_config(rs=ta.kelbg_length(te), theta=0.5, theta_mixed=0.5, walk="mixed", group=128)
            for te, _, _ in reals]
    with ta.Ensemble(cfgs) as ens:
        for r, (_, qe, qi) in enumerate(reals):
            ens.set_particles(r, qe, qi)
        ens.force()
        got = [ens.get_particles(r) for r in range(len(reals))]
    worst = 0.0
    for r, (te, qe, qi) in enumerate(reals):
        with ta.Engine(rs=ta.kelbg_length(te), theta=0.5, theta_mixed=0.5, walk="mixed", group=128) as eng:
            eng.set_particles(qe, qi)
            eng.build()
            eng.force()
            e1, i1 = eng.get_particles()
        assert np.array_equal(got[r][0]["ff"], e1["ff"]) and np.array_equal(got[r][1]["ff"], i1["ff"]), r
        tsp, tix = sample_targets(len(qe), len(qi), seed=99 + r, n=256)
        fd, _ = pyoracle.direct(qe["xx"], qi["xx"], ta.kelbg_length(te), tsp, tix)
        rms, mx = rel_err(pick(got[r][0], got[r][1], tsp, tix), fd)
        worst = max(worst, rms)
        assert rms < 3e-3, (r, len(qe) + len(qi), te, rms, mx)
    print("C4: 64 realisations, worst rms %.3e" % worst)


def test_c5_learned_vs_fixed_16M():
    """C5 as configured (2^23 + 2^23, sigma 1 mm): the learned criterion,
    trained on labels the method generates at 2^19 + 2^19 (sigma 400 um),
    band 0.97 .. 1.03 about theta 0.5, its err_target bisected to the fixed
    theta-0.5 walk's rms error (within 1%).  It reaches that error with at
    most 1.03x the fixed walk's work (measured 0.97-1.00x: in a neutral
    plasma the network has no systematic gain), and both stay within rms
    1.5e-3 (measured 1.26e-3) of the FP64 direct sum."""
    announce("c5_learned_vs_fixed")
    n = 1 << 23
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(5, n, n, 1000.0)
    tsp, tix = sample_targets(n, n)
    fd = direct_reference(qel, qion, rs, tsp, tix)
    q2e, q2i = ta.gen_plasma(4, 1 << 19, 1 << 19, 400.0)

    def work(st):
        return sum(st["m2p"]) + sum(st["p2p"]) + 8 * sum(st["far_pairs"])

    with ta.Engine(rs=rs, theta=0.5, walk="mixed", theta_mixed=0.515, theta_lo=0.485, err_target=1e-7,
                   mac="learned") as eng:
        eng.set_particles(q2e, q2i)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 20)
        eng.mlp_init(0)
        eng.mlp_train(X, y, steps=3000, batch=8192, lr=2e-3)
        eng.configure(mac="fixed", theta_mixed=0.5)
        eng.set_particles(qel, qion)
        eng.build()
        eng.force()
        e, i = eng.get_particles()
        rms_f, mx_f = rel_err(pick(e, i, tsp, tix), fd)
        w_f = work(eng.count_interactions())
        eng.configure(mac="learned", theta_mixed=0.515)
        lo, hi, best = -10.0, -1.0, None
        for _ in range(10):
            mid = 0.5 * (lo + hi)
            eng.configure(err_target=10.0 ** mid)
            eng.build()
            eng.force()
            e, i = eng.get_particles()
            rms_l, mx_l = rel_err(pick(e, i, tsp, tix), fd)
            if rms_l <= 1.01 * rms_f:
                best, lo = (rms_l, work(eng.count_interactions())), mid
            else:
                hi = mid
    print("C5: fixed rms %.3e work %.3g | learned rms %s work x%s" %
          (rms_f, w_f, best and "%.3e" % best[0], best and "%.3f" % (best[1] / w_f)))
    assert rms_f < 1.5e-3, rms_f
    assert best is not None and best[1] <= 1.03 * w_f, (rms_f, w_f, best)
