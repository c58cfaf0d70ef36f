"""The GPU MLP kernels against the FP64 oracle (oracle/mlp_ref.py), the
GPU trainer, and the learned opening criterion against the fixed-theta walk
(its fallback and its yardstick)."""
import numpy as np
import pytest

import tree_amd as ta
from oracle import mlp_ref
from tests.conftest import announce

pytestmark = pytest.mark.gpu

RS10 = 6.6493e-3


def _eng(**kw):
    kw.setdefault("rs", RS10)
    return ta.Engine(**kw)


def test_forward_and_gradient_vs_oracle():
    """fp32 forward within 1e-5 of FP64; the bf16 forward (the walk's matrix-core
    path) within 2e-3 of the oracle's bf16 arithmetic (bf16 rounding of the first
    hidden layer may flip on ties that FP32 and FP64 sums round differently) and
    within 3e-2 of FP64; gradient within 1e-4 relative.  5003 samples: the last
    wavefront is partial."""
    announce("mlp_grad")
    rng = np.random.default_rng(1)
    p = mlp_ref.init_params(5).astype(np.float32)
    X = rng.normal(size=(5003, 8)).astype(np.float32)
    t = rng.normal(size=5003).astype(np.float32)
    y_ref, _ = mlp_ref.forward(p.astype(np.float64), X.astype(np.float64))
    loss_ref, g_ref = mlp_ref.loss_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64))
    with _eng() as eng:
        eng.mlp_set(p)
        assert np.array_equal(eng.mlp_get(), p)
        y32 = eng.mlp_forward(X, bf16=False)
        y16 = eng.mlp_forward(X, bf16=True)
        loss, g = eng.mlp_loss_grad(X, t)
    assert np.abs(y32 - y_ref).max() <= 1e-5 * max(1.0, np.abs(y_ref).max())
    y_bf = mlp_ref.forward_bf16(p.astype(np.float64), X.astype(np.float64))
    scale = max(1.0, np.abs(y_ref).max())
    assert np.median(np.abs(y16 - y_bf)) <= 1e-5 * scale
    assert np.abs(y16 - y_bf).max() <= 2e-3 * scale
    assert np.abs(y16 - y_ref).max() <= 3e-2 * scale
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_training_fits_a_known_function():
    """Adam on the GPU fits y = sum(sin(x)) well below its variance."""
    announce("mlp_train")
    rng = np.random.default_rng(2)
    X = rng.uniform(-2, 2, size=(20000, 8)).astype(np.float32)
    y = np.sin(X).sum(axis=1).astype(np.float32)
    with _eng() as eng:
        eng.mlp_set(mlp_ref.init_params(0).astype(np.float32))
        hist = eng.mlp_train(X, y, steps=3000, batch=2048, lr=3e-3)
        pred = eng.mlp_forward(X, bf16=False)
    mse = float(np.mean((pred - y) ** 2))
    assert hist[-50:].mean() < 0.05 * hist[:10].mean()
    assert mse < 0.05 * float(np.var(y)), mse


def _direct_err(eng, qel, qion, n):
    tsp = np.r_[np.zeros(n, np.int32), np.ones(n, np.int32)]
    tix = np.r_[np.arange(n), np.arange(n)].astype(np.int32)
    fd = eng.direct_sample(tsp, tix)
    eng.build()
    eng.force()
    e, i = eng.get_particles()
    f = np.r_[e["ff"][:n], i["ff"][:n]]
    rel = np.linalg.norm(f - fd, axis=1) / np.linalg.norm(fd, axis=1)
    return float(np.sqrt(np.mean(rel ** 2)))


@pytest.mark.parametrize("theta", [0.26, 0.5])
def test_learned_mac_matches_fixed(theta):
    """Trained on labels the method generates, the learned criterion is tuned
    (err_target bisected in log10) to the fixed-theta walk's rms error within
    1%, as in the other learned-criterion tests (opening every band cell does
    not lower the error monotonically, so an exact match is not always
    reachable), sampled against the FP64 direct sum, and its work (M2P + P2P
    interactions) is compared there: 2^17 + 2^17 particles; cells beyond
    s/(0.97 theta) always accepted, cells within s/(1.03 theta) always opened,
    the network deciding the near band cells in between (far ones are decided
    at the band's geometric mean, as the fixed walk decides them).  The
    learned walk must reach the fixed walk's error within 3% of its work.
    Measured at this seed: 1.014x (theta 0.26, where every band cell opened
    gives rms 8.15e-5 against the fixed walk's 8.10e-5) and 0.986x (theta
    0.5); over two seeds and three bands 0.96-1.05x.  With quadrupole cells the network
    saved 4-13%; the cells' third moments and the far path through order-6
    local expansions leave it no systematic gain."""
    announce("learned_mac_vs_fixed theta=%g" % theta)
    n = 1 << 17
    qel, qion = ta.gen_plasma(31, n, n, 300.0)
    with _eng(theta=theta, theta_mixed=1.03 * theta, theta_lo=0.97 * theta, walk="mixed", err_target=1e-7,
              mac="learned") as eng:
        eng.set_particles(qel, qion)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 20)
        assert len(X) > 5000
        eng.mlp_set(mlp_ref.init_params(0).astype(np.float32))
        eng.mlp_train(X, y, steps=3000, batch=8192, lr=2e-3)
        eng.configure(mac="fixed", theta_mixed=theta)
        err_fix = _direct_err(eng, qel, qion, 2048)
        st = eng.count_interactions()
        work_fix = sum(st["m2p"]) + sum(st["p2p"])
        eng.configure(mac="learned", theta_mixed=1.03 * theta)
        lo, hi, best, rows = -10.0, -2.0, None, []
        for _ in range(12):
            mid = 0.5 * (lo + hi)
            eng.configure(err_target=10.0 ** mid)
            err_l = _direct_err(eng, qel, qion, 2048)
            st = eng.count_interactions()
            work_l = sum(st["m2p"]) + sum(st["p2p"])
            rows.append((mid, err_l, work_l))
            print("err_target 1e%.2f: learned rms %.3e work %d | fixed rms %.3e work %d"
                  % (mid, err_l, work_l, err_fix, work_fix))
            if err_l <= 1.01 * err_fix:
                best, lo = (mid, err_l, work_l), mid
            else:
                hi = mid
    assert best is not None, rows
    assert best[2] < 1.03 * work_fix, (best, work_fix, rows)


def _learned_vs_fixed(qel, qion, theta, blo, bhi, m=2048, tol=1.01):
    """Fixed-theta walk at theta against the learned criterion in the band
    blo theta .. bhi theta, trained on labels of this plasma, its err_target
    bisected (log10) to the fixed walk's rms error (within tol): returns
    (fixed rms, fixed work, learned rms, learned work, fixed ms, learned ms),
    work = M2P + P2P interactions + 8 x far (cell, group) pairs."""
    tsp = np.r_[np.zeros(m, np.int32), np.ones(m, np.int32)]
    tix = np.r_[np.arange(m), np.arange(m)].astype(np.int32)

    def work(st):
        return sum(st["m2p"]) + sum(st["p2p"]) + 8 * sum(st["far_pairs"])

    def ms(eng):
        eng.force()
        return float(np.mean([(eng.force(), eng.stats()["ms_walk"])[1] for _ in range(3)]))

    with _eng(theta=theta, theta_mixed=bhi * theta, theta_lo=blo * theta, walk="mixed", err_target=1e-7,
              mac="learned", group=128) as eng:
        eng.set_particles(qel, qion)
        fd = eng.direct_sample(tsp, tix)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 21)
        eng.mlp_init(0)
        eng.mlp_train(X, y, steps=3000, batch=8192, lr=2e-3)
        eng.configure(mac="fixed", theta_mixed=theta)
        err_fix = _direct_err_fd(eng, fd, m)
        work_fix, ms_fix = work(eng.count_interactions()), ms(eng)
        eng.configure(mac="learned", theta_mixed=bhi * theta)
        lo, hi, best = -10.0, -1.0, None
        for _ in range(12):
            mid = 0.5 * (lo + hi)
            eng.configure(err_target=10.0 ** mid)
            e = _direct_err_fd(eng, fd, m)
            if e <= tol * err_fix:
                best, lo = (e, work(eng.count_interactions()), ms(eng)), mid
            else:
                hi = mid
    assert best is not None, "the learned walk never reaches the fixed walk's error %g" % err_fix
    return err_fix, work_fix, best[0], best[1], ms_fix, best[2]


def _direct_err_fd(eng, fd, m):
    eng.build()
    eng.force()
    e, i = eng.get_particles()
    f = np.r_[e["ff"][:m], i["ff"][:m]]
    rel = np.linalg.norm(f - fd, axis=1) / np.linalg.norm(fd, axis=1)
    return float(np.sqrt(np.mean(rel ** 2)))


@pytest.mark.parametrize("theta", [0.4, 0.5])
def test_learned_mac_gains_in_non_neutral_plasma(theta):
    """Where the learned walk pays: a non-neutral plasma, the electron cloud
    contracted to 0.7 of the ion cloud's width (2^17 + 2^17, sigma_i 300 um),
    so the net field is strong in the shell and vanishes at the centre.  At
    equal rms error (within 1%), band 0.8 .. 1.25 theta, the learned walk
    does at most 0.9 of the fixed walk's work (M2P + P2P + 8 x far pairs;
    measured 0.74 at theta 0.4 and 0.81-0.82 at 0.5 over two seeds, walk time
    0.77-0.80 and 0.90-0.97).  Most of the saving is the band's upper edge:
    a fixed walk taking cells within the far-path radius at 1.25 theta does
    the same work, but with 13% (theta 0.4) and 3% (theta 0.5) more error;
    the network's per-cell decisions remove that excess.  In the neutral
    plasmas of the benchmark configurations there is no such gain
    (test_learned_mac_matches_fixed)."""
    announce("learned_mac_non_neutral theta=%g" % theta)
    n = 1 << 17
    qel, qion = ta.gen_plasma(32, n, n, 300.0)
    qel["xx"] *= 0.7
    # at theta 0.4 this plasma's rms is set by a floor the opening angle does
    # not move (the fixed walk: 2.31e-4 at theta 0.32, 2.40e-4 at 0.36, 2.36e-4
    # at 0.40; the far cells' local expansions), and the learned walk's lowest
    # error, every band cell opened, is 2.42e-4: matched within 3%
    ef, wf, el, wl, tf, tl = _learned_vs_fixed(qel, qion, theta, 0.8, 1.25, tol=1.03 if theta < 0.45 else 1.01)
    print("non-neutral theta %.2f: fixed rms %.3e work %.3g %.2f ms | learned rms %.3e work x%.3f time x%.3f"
          % (theta, ef, wf, tf, el, wl / wf, tl / tf))
    assert wl <= 0.9 * wf, (ef, wf, el, wl)


# VALU instructions per interaction in the compiled walk kernels (gfx950):
# M2P per (cell, target) pair and P2P per (leaf, target) pair from the flush
# loops of k_walk32, a far cell's contribution to a group's local expansion
# per (cell, group) pair from k_far_eval
C_M2P, C_M2P3, C_P2P, C_FAR = 126.5, 77.0, 15.1, 1632.0


def _cost(st):
    m3 = sum(st["m2p3"])
    return (C_M2P * (sum(st["m2p"]) - m3) + C_M2P3 * m3 + C_P2P * sum(st["p2p"]) +
            C_FAR * sum(st["far_pairs"]))


def test_cost_weight_spends_the_error_where_it_saves_work():
    """The cost-aware threshold (cost_weight w): a band cell is accepted when
    the predicted log10 error is below log10(err_target) + w log10 S, S the
    work its acceptance saves in M2Ps (its children's P2P or M2P against its
    own M2P).  Neutral plasma at the C2 configuration's size and width (2^19 +
    2^19, sigma 400 um), theta 0.4 (near the benchmark's tuned 0.396), band
    0.97 .. 1.03 theta, err_target bisected to the fixed walk's rms error
    (within 1%) for w = 0 and w = 0.5; work in VALU instructions per force
    (M2P 126.5 to the fourth moment, 77 to the third, P2P 15.1 per pair, 1632
    per far (cell, group) pair).  Without
    the cost term the network accepts cells whose children are leaves (one
    M2P for a few P2Ps) and does more work than the fixed walk; with it, less.
    Measured: 1.053x and 0.959x the fixed walk's instructions (at N = 4M,
    theta 0.396: 1.07x and 0.96x; at 2^17 + 2^17, sigma 300 um: 0.977x and
    0.961x)."""
    announce("cost_weight theta=0.4")
    n, m, theta = 1 << 19, 2048, 0.4
    qel, qion = ta.gen_plasma(33, n, n, 400.0)
    tsp = np.r_[np.zeros(m, np.int32), np.ones(m, np.int32)]
    tix = np.r_[np.arange(m), np.arange(m)].astype(np.int32)
    with _eng(theta=theta, theta_mixed=1.03 * theta, theta_lo=0.97 * theta, walk="mixed", err_target=1e-7,
              mac="learned", group=128) as eng:
        eng.set_particles(qel, qion)
        fd = eng.direct_sample(tsp, tix)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 21)
        eng.mlp_init(0)
        eng.mlp_train(X, y, steps=3000, batch=8192, lr=2e-3)
        eng.configure(mac="fixed", theta_mixed=theta)
        err_fix = _direct_err_fd(eng, fd, m)
        cost_fix = _cost(eng.count_interactions())
        cost = {}
        for w in (0.0, 0.5):
            eng.configure(mac="learned", theta_mixed=1.03 * theta, cost_weight=w)
            lo, hi, best = -10.0, -1.0, None
            for _ in range(12):
                mid = 0.5 * (lo + hi)
                eng.configure(err_target=10.0 ** mid)
                e = _direct_err_fd(eng, fd, m)
                if e <= 1.01 * err_fix:
                    best, lo = (e, _cost(eng.count_interactions())), mid
                else:
                    hi = mid
            assert best is not None, "w %g: the learned walk never reaches the fixed walk's error %g" % (w, err_fix)
            cost[w] = best[1]
            print("w %.1f: rms %.3e (fixed %.3e), VALU work x%.3f of the fixed walk's" % (w, best[0], err_fix,
                                                                                      best[1] / cost_fix))
    assert cost[0.5] <= 0.97 * cost[0.0], cost
    assert cost[0.5] <= cost_fix, (cost, cost_fix)
