"""Headline benchmark: steps/s of the Barnes-Hut step at N = 4M on one MI355X.

Workload (default, the 4M configuration): 2^21 electrons + 2^21 singly
charged ions (m_i = 85 amu), seeded spherical Gaussian positions (sigma =
600 um), T_e = 10 K, T_i = 1 mK, Kelbg length at T_e, dt = 5 ps.  A step is
VV1 + tree build + force + VV2 with all particle state resident in HBM.

Every walk is first tuned (untimed) to an rms relative force error of 1e-4
against an FP64 direct sum on 4096 sampled targets (2048 e- + 2048 ions, seed
99):
  * plain FP64: the reference's algorithm (per-target MAC, child-point
    moments, two species trees).  Its theta is bisected on the sampled targets
    only, and its step time is measured with the walk timed on every S-th
    target (at least 2^17 targets per species, enough to fill the GPU) and
    scaled to all targets (per-target walks are independent), plus the
    measured build and Verlet times.  At 1e-4 its theta is small and a full
    step takes minutes, which is why it is sampled.
  * mixed: the wavefront-group walk on the joint tree, theta bisected on full
    force evaluations; K full steps timed.
  * mixed+learned (headline): the MLP opening criterion, trained on the GPU
    on labels collected by the method itself on a 2^20-particle plasma; its
    error threshold is bisected to 1e-4; K full steps timed.

Prints one JSON line.  `roofline` covers the headline's walk kernel (counted
interactions x flop weights / event-timed walk duration); `cpu_baseline`
times the reference's own tree code (oracle/_ref, compiled from its sources)
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import struct
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tree_amd as ta  # noqa: E402

# FP64 walk flop weights per event (FMA = 2): MAC test, M2P on accept, P2P Coulomb / Kelbg
W_VISIT, W_M2P, W_P2P_C, W_P2P_K = 11, 144, 20, 28
# mixed walk, per target-entry interaction, counted from walk32.hip's compiled
# FP32 code (monopole + dipole + traceless quadrupole + traceless third and
# fourth moments M2P 202: 404 flops in 253 VALU instructions per listed cell
# for a lane's 2 targets; Coulomb P2P 18, Kelbg P2P 22); MAC tests are per
# cell and group and are not counted
W32_M2P, W32_P2P_C, W32_P2P_K = 202, 18, 22
# cells far from the group for their size take the M2P to the third moment:
# 122 flops in 77 VALU instructions per (cell, target)
W32_M2P3 = 122
# far cells: flops per (cell, target group) pair of the order-6 local
# expansion with the cell's third and fourth moments, terms to total order 7
# (FMA = 2), counted from the compiled loop of k_far_eval (m2l_acc_p6hxt7;
# 1632 VALU instructions with the cell-record loads)
W32_M2L_PAIR = 2776
# TFLOP/s, MI355X published vector peaks.  The FP32 figure counts packed FP32
# (two lanes per instruction); this build keeps packed FP32 off, so one
# instruction per lane per cycle (78.6) is its own ceiling, stated beside it.
PEAK = {"fp64": 78.6, "fp32": 157.3, "fp32_nonpacked": 78.6}
TARGET = 1e-4


def log(s):
    print(s, file=sys.stderr, flush=True)


def sample_targets(ne, ni, seed=99, n=4096):
    rng = np.random.default_rng(seed)
    te = rng.choice(ne, n // 2, replace=False)
    ti = rng.choice(ni, n // 2, replace=False)
    tsp = np.r_[np.zeros(n // 2, np.int32), np.ones(n // 2, np.int32)]
    return tsp, np.r_[te, ti].astype(np.int32)


def rel_err(f, fdir):
    rel = np.linalg.norm(f - fdir, axis=1) / np.linalg.norm(fdir, axis=1)
    return float(np.sqrt(np.mean(rel ** 2))), float(rel.max())


def err_full(eng, tsp, tix, fdir):
    eng.build()
    eng.force()
    e, i = eng.get_particles()
    f = np.where(tsp[:, None] == 0, e["ff"][np.where(tsp == 0, tix, 0)], i["ff"][np.where(tsp == 1, tix, 0)])
    return rel_err(f, fdir)


def err_sampled64(eng, tsp, tix, fdir):
    eng.build()
    return rel_err(eng.force_targets(tsp, tix), fdir)


def bisect(set_param, measure, lo, hi, iters, geometric=True):
    """Largest parameter in [lo, hi] with rms <= TARGET (error increases with it)."""
    best = None
    for _ in range(iters):
        mid = np.sqrt(lo * hi) if geometric else 0.5 * (lo + hi)
        set_param(mid)
        rms, mx = measure()
        log("    %.5g -> rms %.3e max %.3e" % (mid, rms, mx))
        if rms <= TARGET:
            best = (mid, rms, mx)
            lo = mid
        else:
            hi = mid
    if best is None:
        set_param(lo)
        rms, mx = measure()
        best = (lo, rms, mx)
    set_param(best[0])
    return best


def walk_flops64(st):
    fl = 0
    for p in range(4):
        wp = W_P2P_C if p in (0, 3) else W_P2P_K
        fl += W_VISIT * st["visits"][p] + W_M2P * st["m2p"][p] + wp * st["p2p"][p]
    return fl


def walk_flops32(st):
    fl = 0
    for p in range(4):
        wp = W32_P2P_C if p in (0, 3) else W32_P2P_K
        fl += W32_M2P * st["m2p"][p] + wp * st["p2p"][p]
    # of the M2P pairs, those to the third moment only
    fl -= (W32_M2P - W32_M2P3) * sum(st.get("m2p3", [0, 0]))
    # far cells: (cell, group) pairs, and (cell, segment) pairs of the two-level far field
    return fl + W32_M2L_PAIR * (sum(st.get("far_pairs", [0, 0])) + sum(st.get("seg_pairs", [0, 0])))


def profiled_traffic(kernels, path=os.path.join(ROOT, "profiles", "hbm_traffic_4M.md")):
    """HBM bytes per launch of the walk (the sum over `kernels`) at N = 4M, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (its table gives calibrated MB per dispatch); None without a row for each."""
    rows = {}
    try:
        with open(path) as f:
            for ln in f:
                c = [x.strip() for x in ln.strip().strip("|").split("|")]
                if len(c) == 4 and c[0].startswith("`"):
                    try:
                        rows[c[0].strip("`")] = (float(c[2]) + float(c[3])) * 1e6
                    except ValueError:
                        pass
    except OSError:
        return None
    if not all(k in rows for k in kernels):
        return None
    return sum(rows[k] for k in kernels)


def interactions(st):
    """(target, cell or particle) interactions of a force: M2P, far cells taken
    through the group's local expansion (per target), and P2P."""
    return sum(st["m2p"]) + sum(st.get("m2l", [0])) + sum(st["p2p"])


def timed_steps(eng, dt, steps, warmup):
    eng.build()
    ep = eng.force()
    k0 = eng.vv2(0.0, True)  # kinetic sums of the current state (zero kick)
    e0 = 0.5 * (k0[0] + k0[1]) + ep[0] + ep[1]
    if warmup:
        eng.step(dt, warmup)
    ms_force, ms_build, ms_walk = [], [], []
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = eng.step(dt, 1)  # returns after the stream has synchronised
        st = eng.stats()
        ms_force.append(st["ms_force"])
        ms_build.append(st["ms_build"])
        ms_walk.append(st["ms_walk"])
    t1 = time.perf_counter()
    e1 = 0.5 * (out[2] + out[3]) + out[0] + out[1]
    return {"s": (t1 - t0) / steps, "ms_force": float(np.mean(ms_force)), "ms_build": float(np.mean(ms_build)),
            "ms_walk": float(np.mean(ms_walk)), "drift": (e1 - e0) / abs(e0)}


def cpu_baseline(qel, qion, theta, rs, threads):
    """Reference tree code on the host: serial build + OpenMP walk of every stride-th target."""
    try:
        from oracle import pyoracle
    except Exception as ex:  # pragma: no cover
        return {"value": None, "error": str(ex)}
    n = len(qel) + len(qion)
    kind = "reference" if pyoracle.ref_available() else "port"

    def run(stride):
        if kind == "reference":
            buf = pyoracle.run_ref(3, qel["xx"], qel["xv"], qion["xx"], qion["xv"], theta, rs, (ta.XM_E, ta.XM_I),
                                   0.0, stride=stride, threads=threads)
            return struct.unpack_from("ddi", buf, 0)
        return pyoracle.time_sample(qel["xx"], qion["xx"], theta, rs, stride, threads)

    # ~128 targets first; if their walk is short, widen the sample to ~15 s of walking
    stride = max(1, len(qel) // 64)
    tb, tw, nw = run(stride)
    if tw < 5.0 and stride > 1:
        stride = max(1, int(stride * max(tw, 1e-3) / 15.0))
        tb, tw, nw = run(stride)
    t_step = tb + tw * n / nw
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": kind,
            "cpu_model": model, "host_cpus": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)),
            "sample": "one serial tree build of both species (%.1f s) + walk of every %d-th target (%d targets, "
                      "%.1f s) at the plain walk's tuned theta=%.4f, scaled to all %d targets"
                      % (tb, stride, nw, tw, theta, n)}


def run_plain(qel, qion, rs, tsp, tix, fdir, a):
    eng = ta.Engine(rs=rs, theta=0.3, walk="ref64")
    eng.set_particles(qel, qion)
    if a.theta64 > 0:
        eng.configure(theta=a.theta64)
        th = a.theta64
        rms, mx = err_sampled64(eng, tsp, tix, fdir)
    else:
        log("tuning the plain FP64 walk (sampled targets)")
        th, rms, mx = bisect(lambda v: eng.configure(theta=v), lambda: err_sampled64(eng, tsp, tix, fdir),
                             0.01, 0.3, 7)
    eng.build()
    ms_build = eng.stats()["ms_build"]
    # sampled walk timing: every stride-th block of 64 consecutive targets of
    # each species, at least 2^17 targets per species so that the sample
    # fills the GPU (the scaled time is a throughput, not one wave's latency)
    n = len(qel) + len(qion)
    stride = max(1, min(len(qel), len(qion)) // (1 << 17))
    ms, nsamp = eng.time_walk(stride)
    ms_walk = ms * n / nsamp
    t0 = time.perf_counter()
    eng.vv2(0.0, False)
    eng.vv2(0.0, True)
    ms_vv = 1e3 * (time.perf_counter() - t0)
    s = (ms_build + ms_walk + ms_vv) * 1e-3
    eng.close()
    log("plain FP64: theta=%.4f rms=%.2e -> walk %.1f ms (every %d-th target timed), %.4f steps/s"
        % (th, rms, ms_walk, stride, 1 / s))
    return {"variant": "plain-fp64", "theta": th, "rms": rms, "max": mx, "s": s, "ms_force": ms_walk,
            "ms_build": ms_build, "stride": stride, "prec": "fp64"}


def run_mixed(qel, qion, rs, dt, tsp, tix, fdir, a):
    """The fixed-theta mixed walk, theta bisected to the error target.  With
    the group size left to the autotuner, theta is bisected at each group
    size (the theta that meets the target depends on the groups' extent) and
    the size whose walk is fastest at the target is kept, with the list
    lengths the autotuner picks there."""
    eng = ta.Engine(rs=rs, theta=0.5, walk="mixed", theta_mixed=0.5, group=a.group)
    eng.set_particles(qel, qion)

    def tune():
        if a.theta32 > 0:
            eng.configure(theta_mixed=a.theta32)
            return (a.theta32,) + tuple(err_full(eng, tsp, tix, fdir))
        return bisect(lambda v: eng.configure(theta_mixed=v), lambda: err_full(eng, tsp, tix, fdir), 0.05, 1.0, 10)

    if a.group == 0:
        log("tuning the mixed walk at each group size")
        best = None
        for k, gs in enumerate((64, 128, 256)):
            eng.configure(group=gs)
            th, rms, mx = tune()
            eng.build()
            _, ms = eng.autotune()  # per group size, at this theta and the best list lengths
            log("  groups of %d: theta %.4f rms %.2e, walk %.2f ms" % (gs, th, rms, ms[k]))
            if best is None or ms[k] < best[0]:
                best = (ms[k], gs, th, rms, mx)
        _, g, th, rms, mx = best
        eng.configure(group=g, theta_mixed=th)
        eng.build()
        eng.autotune()  # the list lengths for this group size and theta
        log("  chosen: groups of %d, lists %s" % (g, eng.walk_lists()))
    else:
        log("tuning the mixed walk")
        g = a.group
        th, rms, mx = tune()
    lists = eng.walk_lists()
    eng.build()
    eng.force()
    st = eng.count_interactions()
    eng.set_particles(qel, qion)
    r = timed_steps(eng, dt, a.steps, a.warmup)
    eng.close()
    log("mixed: theta=%.4f rms=%.2e %.3f steps/s walk %.2f ms" % (th, rms, 1 / r["s"], r["ms_walk"]))
    return {"variant": "mixed", "theta": th, "rms": rms, "max": mx, "st": st, "prec": "fp32", "group": g,
            "lists": lists, **r}


def train_mlp(a, th_lo, th_hi, group=0):
    """Labels from the method itself on a 2^20-particle plasma (the band the
    learned walk will decide), then GPU training.  The labels' target groups
    have the size the learned walk will use (`group`, 0: autotuned): two of
    the features (group extent, targets in the group) depend on it."""
    n = 1 << 19
    qel, qion = ta.gen_plasma(a.seed + 1, n, n, 400.0, te=a.te, ti=1e-3)
    rs = ta.kelbg_length(a.te)
    with ta.Engine(rs=rs, theta=0.5, walk="mixed", theta_mixed=th_hi, theta_lo=th_lo, group=group) as eng:
        eng.set_particles(qel, qion)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 20)
        eng.mlp_init(0)
        hist = eng.mlp_train(X, y, steps=4000, batch=8192, lr=2e-3, seed=1)
        p = eng.mlp_get()
    log("  MLP: %d labelled band pairs, loss %.3f -> %.3f" % (len(X), hist[:20].mean(), hist[-50:].mean()))
    return p, {"samples": int(len(X)), "loss_first": float(hist[:20].mean()), "loss_last": float(hist[-50:].mean()),
               "band_theta": [th_lo, th_hi], "train_n_total": 2 * n}


def run_learned(qel, qion, rs, dt, tsp, tix, fdir, a, params, th_lo, th_hi):
    """The learned criterion in the band theta_lo .. theta_hi, its error
    threshold bisected to the target.  The walk runs with the target groups
    the network was trained on (a.group: two features depend on the group's
    size); list lengths as autotuned for this size."""
    assert a.group in (64, 128, 256), "the learned walk runs at the group size its labels were collected with"
    eng = ta.Engine(rs=rs, theta=0.5, walk="mixed", theta_mixed=th_hi, theta_lo=th_lo, mac="learned",
                    err_target=1e-9, group=a.group, cost_weight=a.cost_weight)
    eng.set_particles(qel, qion)
    eng.mlp_set(params)
    eng.build()
    eng.force()  # the fixed-theta fallback: gives the first previous forces
    g = a.group
    lists = eng.walk_lists()
    if a.err_target > 0:
        et = a.err_target
        eng.configure(err_target=et)
        rms_l, mx_l = err_full(eng, tsp, tix, fdir)
    else:
        log("tuning the learned criterion's error threshold (band theta %.3f .. %.3f)" % (th_lo, th_hi))
        et, rms_l, mx_l = bisect(lambda v: eng.configure(err_target=v), lambda: err_full(eng, tsp, tix, fdir),
                                 1e-9, 1e-1, 9)
    eng.build()
    eng.force()
    st_l = eng.count_interactions()
    eng.set_particles(qel, qion)
    r = timed_steps(eng, dt, a.steps, a.warmup)
    eng.close()
    log("mixed+learned: err_target=%.2e rms=%.2e %.3f steps/s walk %.2f ms"
        % (et, rms_l, 1 / r["s"], r["ms_walk"]))
    return {"variant": "mixed+learned", "theta": th_hi, "theta_lo": th_lo, "err_target": et,
            "cost_weight": a.cost_weight, "rms": rms_l,
            "max": mx_l, "st": st_l, "prec": "fp32", "group": g, "lists": lists, **r}


# ------------------------------------------------------------------ other configurations
#
# `--config c2|c3|c4|c5` measures the other configurations of the workload
# family (one JSON line each; the default run is the headline above).


def _steps_per_s(eng, dt, steps):
    eng.build()
    eng.force()
    t0 = time.perf_counter()
    eng.step(dt, steps)
    return steps / (time.perf_counter() - t0), eng.stats()


def run_c2(a):
    """2^19 + 2^19, sigma 400 um, theta 0.5: the plain FP64 walk and the mixed walk at the same theta,
    each stepped a.steps times (steps/s and energy drift); the FP64 walk's counted flops against its
    kernel time (FP64 VALU roofline); the sampled walk timing that the headline's plain-walk figure uses,
    checked against a full force."""
    n, dt = 1 << 19, 0.005 / ta.TPS
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(2, n, n, 400.0, te=10.0, ti=1e-3)
    tsp, tix = sample_targets(n, n)
    out = {"config": "C2: 2^19 e- + 2^19 ions, sigma 400 um, T_e 10 K, theta 0.5", "steps": a.steps}
    for walk in ("ref64", "mixed"):
        with ta.Engine(rs=rs, theta=0.5, theta_mixed=0.5, walk=walk) as eng:
            eng.set_particles(qel, qion)
            if walk == "ref64":
                fdir = eng.direct_sample(tsp, tix)
            rms, mx = err_full(eng, tsp, tix, fdir)
            st = eng.count_interactions()
            eng.set_particles(qel, qion)
            r = timed_steps(eng, dt, a.steps, 0)
            sps = 1.0 / r["s"]
            row = {"steps_per_s": sps, "rms": rms, "max": mx, "ms_build": r["ms_build"], "ms_walk": r["ms_walk"],
                   "energy_drift": r["drift"], "interactions_per_s": interactions(st) * sps}
            if walk == "ref64":
                fl = walk_flops64(st)
                row["walk_tflops"] = fl / (r["ms_walk"] * 1e-3) / 1e12
                row["walk_frac_fp64_peak"] = row["walk_tflops"] / PEAK["fp64"]
                row["flops_per_force"] = fl
                # the sampled timing of the headline's plain walk, against the full force just timed
                eng.set_particles(qel, qion)
                eng.build()
                stride = max(1, n // (1 << 17))
                runs = [eng.time_walk(stride) for _ in range(3)]
                ms, nsamp = min(x[0] for x in runs), runs[0][1]
                row["sampled_walk_check"] = {"stride": stride, "scaled_ms": ms * 2 * n / nsamp,
                                             "full_ms": r["ms_walk"]}
                # and at the headline's small theta, where the full walk takes tens of seconds
                if a.theta64 > 0:
                    eng.configure(theta=a.theta64)
                    eng.build()
                    t0 = time.perf_counter()
                    eng.force()
                    full = eng.stats()["ms_force"]
                    log("C2 FP64 full walk at theta %.4f: %.1f s" % (a.theta64, time.perf_counter() - t0))
                    ms, nsamp = eng.time_walk(stride)
                    row["sampled_walk_check_small_theta"] = {"theta": a.theta64, "stride": stride,
                                                             "scaled_ms": ms * 2 * n / nsamp, "full_ms": full}
            out[walk] = row
            log("C2 %s: %.2f steps/s rms %.2e drift %.2e" % (walk, sps, rms, r["drift"]))
    print(json.dumps(out))


def run_c3(a):
    """4M, sigma 600 um: the mixed walk over theta in {0.3, 0.5, 0.7} and T_e in {1, 10, 100} K."""
    n, dt = 1 << 21, 0.005 / ta.TPS
    rows = []
    for te in (1.0, 10.0, 100.0):
        rs = ta.kelbg_length(te)
        qel, qion = ta.gen_plasma(3, n, n, 600.0, te=te, ti=1e-3)
        tsp, tix = sample_targets(n, n)
        with ta.Engine(rs=rs, theta=0.5) as eng:
            eng.set_particles(qel, qion)
            fdir = eng.direct_sample(tsp, tix)
        for th in (0.3, 0.5, 0.7):
            with ta.Engine(rs=rs, theta=0.5, theta_mixed=th, walk="mixed") as eng:
                eng.set_particles(qel, qion)
                rms, mx = err_full(eng, tsp, tix, fdir)
                st = eng.count_interactions()
                eng.set_particles(qel, qion)
                sps, stt = _steps_per_s(eng, dt, a.steps)
                fl = walk_flops32(st)
                rows.append({"T_e": te, "theta": th, "steps_per_s": sps, "rms": rms, "max": mx,
                             "ms_walk": stt["ms_walk"], "ms_build": stt["ms_build"],
                             "walk_tflops": fl / (stt["ms_walk"] * 1e-3) / 1e12,
                             "walk_frac_nonpacked_fp32": fl / (stt["ms_walk"] * 1e-3) / 1e12 / PEAK["fp32_nonpacked"],
                             "build_Mparticles_per_s": 2 * n / (stt["ms_build"] * 1e-3) / 1e6})
                log("C3 T_e %g theta %.1f: %.2f steps/s rms %.2e walk %.1f ms" % (te, th, sps, rms, stt["ms_walk"]))
    print(json.dumps({"config": "C3: 2^21 e- + 2^21 ions, sigma 600 um, mixed walk", "steps": a.steps, "rows": rows}))


def run_c4(a):
    """64 realisations, N_total log-uniform in [16k, 256k], T_e log-uniform in [1, 100] K,
    sigma from a total peak density of 1e15 m^-3: one persistent walk launch per step,
    against the same realisations stepped one engine at a time."""
    rng = np.random.default_rng(1000)
    R, dt, steps = 64, 0.005 / ta.TPS, a.steps
    ntot = np.exp(rng.uniform(np.log(16384), np.log(262144), R)).astype(int) // 2 * 2
    tes = np.exp(rng.uniform(0.0, np.log(100.0), R))
    cfgs, data = [], []
    for r in range(R):
        sigma = (ntot[r] / ((2 * np.pi) ** 1.5 * 1e15)) ** (1 / 3) * 1e6
        cfgs.append(ta.make_config(rs=ta.kelbg_length(tes[r]), theta=0.5, theta_mixed=0.5, walk="mixed"))
        data.append(ta.gen_plasma(1000 + r, int(ntot[r]) // 2, int(ntot[r]) // 2, sigma, te=float(tes[r]), ti=1e-3))
    with ta.Ensemble(cfgs) as ens:
        for r, (qe, qi) in enumerate(data):
            ens.set_particles(r, qe, qi)
        ens.force()
        g, ms = ens.autotune()
        lists = ens.walk_lists()
        ens.step(dt, 2)
        t0 = time.perf_counter()
        ens.step(dt, steps)
        t_ens = time.perf_counter() - t0
    t_seq = 0.0  # the same steps, one engine at a time (set-up untimed, as for the ensemble)
    for r, (qe, qi) in enumerate(data):
        with ta.Engine(rs=ta.kelbg_length(tes[r]), theta=0.5, theta_mixed=0.5, walk="mixed", group=g) as eng:
            eng.set_particles(qe, qi)
            eng.build()
            eng.force()
            eng.step(dt, 2)
            t0 = time.perf_counter()
            eng.step(dt, steps)
            t_seq += time.perf_counter() - t0
    out = {"config": "C4: 64 realisations, N_total log-uniform 16k..256k (sum %d), T_e 1..100 K, theta 0.5, "
                     "mixed walk" % int(ntot.sum()), "steps": steps, "group": int(g), "lists": list(lists),
           "ensemble_steps_per_s": steps / t_ens, "particle_steps_per_s": steps * float(ntot.sum()) / t_ens,
           "one_engine_at_a_time_s": t_seq, "ensemble_s": t_ens, "batching_speedup": t_seq / t_ens}
    log("C4: %.2f ensemble steps/s, %.3g particle-steps/s, batching x%.2f" %
        (out["ensemble_steps_per_s"], out["particle_steps_per_s"], out["batching_speedup"]))
    print(json.dumps(out))


def run_c5(a):
    """2^23 + 2^23, sigma 1 mm: the learned criterion (trained at 2^20) against the fixed-theta walk at
    theta 0.5, its fallback and yardstick; err_target bisected to the fixed walk's rms error."""
    global TARGET
    n, dt = 1 << 23, 0.005 / ta.TPS
    rs = ta.kelbg_length(10.0)
    qel, qion = ta.gen_plasma(5, n, n, 1000.0, te=10.0, ti=1e-3)
    tsp, tix = sample_targets(n, n)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        fdir = eng.direct_sample(tsp, tix)
    g = a.group if a.group else 128  # both walks on groups of one size (the network's features depend on it)
    with ta.Engine(rs=rs, theta=0.5, theta_mixed=0.5, walk="mixed", group=g) as eng:
        eng.set_particles(qel, qion)
        rms_f, mx_f = err_full(eng, tsp, tix, fdir)
        st_f = eng.count_interactions()
        eng.set_particles(qel, qion)
        sps_f, stt_f = _steps_per_s(eng, dt, a.steps)
    th_lo, th_hi = a.band[0] * 0.5, min(1.0, a.band[1] * 0.5)
    params, info = train_mlp(a, th_lo, th_hi, group=g)
    TARGET = rms_f
    r = run_learned(qel, qion, rs, dt, tsp, tix, fdir, argparse.Namespace(**{**vars(a), "err_target": 0.0,
                                                                            "group": g}),
                    params, th_lo, th_hi)
    out = {"config": "C5: 2^23 e- + 2^23 ions, sigma 1 mm, fixed theta 0.5 vs learned", "steps": a.steps,
           "fixed": {"steps_per_s": sps_f, "rms": rms_f, "max": mx_f, "ms_walk": stt_f["ms_walk"],
                     "interactions": interactions(st_f)},
           "learned": {"steps_per_s": 1.0 / r["s"], "rms": r["rms"], "max": r["max"], "ms_walk": r["ms_walk"],
                       "err_target": r["err_target"], "interactions": interactions(r["st"])},
           "mlp": info}
    log("C5: fixed %.3f steps/s rms %.2e | learned %.3f steps/s rms %.2e" % (sps_f, rms_f, 1 / r["s"], r["rms"]))
    print(json.dumps(out))


def run_nonneutral(a):
    """Where the learned criterion pays: 2^21 + 2^21 with the electron cloud
    contracted to 0.7 of the ions' sigma (600 um), theta 0.5.  The network is
    trained on labels from a 2^19 + 2^19 plasma of the same kind; its
    err_target is bisected to the fixed walk's rms error (within 1%); both
    walks are then stepped and timed."""
    n, dt = 1 << 21, 0.005 / ta.TPS
    rs = ta.kelbg_length(10.0)
    theta, blo, bhi = 0.5, 0.8, 1.25
    qel, qion = ta.gen_plasma(6, n, n, 600.0)
    qel["xx"] *= 0.7
    tsp, tix = sample_targets(n, n)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        fdir = eng.direct_sample(tsp, tix)
    q2e, q2i = ta.gen_plasma(7, 1 << 19, 1 << 19, 400.0)
    q2e["xx"] *= 0.7
    with ta.Engine(rs=rs, theta=theta, walk="mixed", theta_mixed=bhi * theta, theta_lo=blo * theta,
                   group=128) as eng:
        eng.set_particles(q2e, q2i)
        eng.build()
        X, y = eng.mlp_collect(cap=1 << 21)
        eng.mlp_init(0)
        eng.mlp_train(X, y, steps=4000, batch=8192, lr=2e-3, seed=1)
        params = eng.mlp_get()
    out = {"config": "non-neutral: 2^21 e- + 2^21 ions, sigma_i 600 um, sigma_e 420 um, theta 0.5, "
                     "learned band 0.8 .. 1.25 theta", "steps": a.steps, "labels": int(len(X))}
    with ta.Engine(rs=rs, theta=theta, walk="mixed", theta_mixed=theta, group=128) as eng:
        eng.set_particles(qel, qion)
        rms, mx = err_full(eng, tsp, tix, fdir)
        st = eng.count_interactions()
        eng.set_particles(qel, qion)
        r = timed_steps(eng, dt, a.steps, a.warmup)
        out["fixed"] = {"rms": rms, "max": mx, "steps_per_s": 1.0 / r["s"], "ms_walk": r["ms_walk"],
                        "interactions": interactions(st)}
    target = rms
    with ta.Engine(rs=rs, theta=theta, walk="mixed", theta_mixed=bhi * theta, theta_lo=blo * theta,
                   mac="learned", err_target=1e-6, group=128) as eng:
        eng.set_particles(qel, qion)
        eng.mlp_set(params)
        eng.build()
        eng.force()
        lo, hi, best = -10.0, -1.0, None
        for _ in range(12):
            mid = 0.5 * (lo + hi)
            eng.configure(err_target=10.0 ** mid)
            rl, ml = err_full(eng, tsp, tix, fdir)
            log("    err_target 1e%.2f -> rms %.3e" % (mid, rl))
            if rl <= 1.01 * target:
                best, lo = (mid, rl, ml), mid
            else:
                hi = mid
        if best is None:
            out["learned"] = None
        else:
            eng.configure(err_target=10.0 ** best[0])
            eng.build()
            eng.force()
            st = eng.count_interactions()
            eng.set_particles(qel, qion)
            r = timed_steps(eng, dt, a.steps, a.warmup)
            out["learned"] = {"rms": best[1], "max": best[2], "err_target": 10.0 ** best[0],
                              "steps_per_s": 1.0 / r["s"], "ms_walk": r["ms_walk"],
                              "interactions": interactions(st)}
            out["learned_over_fixed_steps_per_s"] = out["learned"]["steps_per_s"] / out["fixed"]["steps_per_s"]
            out["learned_over_fixed_interactions"] = out["learned"]["interactions"] / out["fixed"]["interactions"]
    log("non-neutral: %s" % json.dumps(out))
    print(json.dumps(out))


def run_build(a):
    """The tree build alone at N = 4M: the FP64 walk's two species trees, and the
    mixed walk's species sort + target groups + joint tree + cell records."""
    n = 1 << 21
    qel, qion = ta.gen_plasma(3, n, n, 600.0)
    out = {"config": "build: 2^21 e- + 2^21 ions, sigma 600 um", "builds": a.steps}
    for walk in ("ref64", "mixed"):
        with ta.Engine(rs=ta.kelbg_length(10.0), theta=0.5, theta_mixed=0.152, walk=walk) as eng:
            eng.set_particles(qel, qion)
            eng.build()
            ms = []
            for _ in range(a.steps):
                eng.build()
                ms.append(eng.stats()["ms_build"])
            out[walk] = {"ms_build": float(np.mean(ms)), "Mparticles_per_s": 2 * n / (np.mean(ms) * 1e-3) / 1e6}
            log("build %s: %.3f ms" % (walk, np.mean(ms)))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 21, help="particles per species")
    ap.add_argument("--sigma", type=float, default=600.0)
    ap.add_argument("--te", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--theta64", type=float, default=0.0, help="fixed theta for the FP64 walk (skip tuning)")
    ap.add_argument("--theta32", type=float, default=0.0, help="fixed theta for the mixed walk (skip tuning)")
    ap.add_argument("--group", type=int, default=0, help="mixed-walk targets per group (0 = autotuned)")
    ap.add_argument("--err-target", type=float, default=0.0,
                    help="fixed error threshold for the learned criterion (skip tuning)")
    # the band and the cost weight were swept together (below): with the
    # cost term the band's width matters little; without it wider bands cost
    # more (0.8 .. 1.25: walk 44.0 ms against 42.6 at 0.97 .. 1.03)
    ap.add_argument("--band", type=float, nargs=2, default=(0.97, 1.03), metavar=("LO", "HI"),
                    help="learned band: theta_lo = LO * theta, theta_hi = HI * theta (theta: the tuned fixed walk's)")
    # measured at N = 4M, 1e-4 rms (walk ms; fixed theta 36.5): band 0.97 ..
    # 1.03, w 0 / 0.25 / 0.5 / 0.75: 42.6 / 38.5 / 38.1 / 38.3; band 0.95 ..
    # 1.05: 38.6 / 38.4 / 37.9 (w 0.25 / 0.5 / 0.75); 0.9 .. 1.12: 38.9 / 38.2
    # / 38.2
    ap.add_argument("--cost-weight", type=float, default=0.5,
                    help="learned criterion: weight of the work a band cell's acceptance saves in its threshold")
    ap.add_argument("--config", choices=("c2", "c3", "c4", "c5", "build", "nonneutral"), default=None,
                    help="measure another configuration instead of the headline (one JSON line)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the engine holds one whole domain per GPU and this build has no "
              "multi-GPU path (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.config:
        {"c2": run_c2, "c3": run_c3, "c4": run_c4, "c5": run_c5, "build": run_build,
         "nonneutral": run_nonneutral}[a.config](a)
        return

    n = a.n
    rs = ta.kelbg_length(a.te)
    dt = 0.005 / ta.TPS
    qel, qion = ta.gen_plasma(a.seed, n, n, a.sigma, te=a.te, ti=1e-3)
    tsp, tix = sample_targets(n, n)
    with ta.Engine(rs=rs, theta=0.5) as eng:
        eng.set_particles(qel, qion)
        fdir = eng.direct_sample(tsp, tix)

    plain = None if a.no_plain else run_plain(qel, qion, rs, tsp, tix, fdir, a)
    fixed = run_mixed(qel, qion, rs, dt, tsp, tix, fdir, a)
    learned, mlp_info = None, None
    if not a.no_learned:
        th_lo, th_hi = a.band[0] * fixed["theta"], min(1.0, a.band[1] * fixed["theta"])
        # the network is trained on, and decides for, the fixed walk's autotuned groups
        al = argparse.Namespace(**{**vars(a), "group": fixed["group"]})
        log("collecting labels and training the opening-criterion MLP (groups of %d)" % al.group)
        params, mlp_info = train_mlp(al, th_lo, th_hi, group=al.group)
        learned = run_learned(qel, qion, rs, dt, tsp, tix, fdir, al, params, th_lo, th_hi)
    head = learned if learned is not None else fixed

    flops = walk_flops32(head["st"])
    mode = 1 if head["variant"] == "mixed+learned" else 0
    traffic = None
    if (n, a.sigma, a.te) == (1 << 21, 600.0, 10.0):
        traffic = profiled_traffic(["k_walk32<%d, %d, false>" % (head["group"] // 64, mode), "k_far_eval"])
    achieved = flops / (head["ms_walk"] * 1e-3) / 1e12
    peak = PEAK["fp32"]

    def summary(v):
        if v is None:
            return None
        out = {"steps_per_s": 1.0 / v["s"], "theta": v["theta"], "rms": v["rms"], "max": v["max"],
               "ms_force": v["ms_force"], "ms_build": v["ms_build"]}
        if "ms_walk" in v:
            out["ms_walk_kernel"] = v["ms_walk"]
        for k in ("drift", "stride", "err_target", "cost_weight", "theta_lo", "group", "lists"):
            if k in v:
                out[k] = v[k]
        if "st" in v:
            out["interactions_per_s"] = interactions(v["st"]) / v["s"]
            st = v["st"]
            out["counts_per_force"] = {"m2p": int(sum(st["m2p"])), "m2p_third_order": int(sum(st.get("m2p3", [0]))),
                                       "p2p": int(sum(st["p2p"])),
                                       "far_m2l": int(sum(st.get("m2l", [0]))),
                                       "far_pairs": int(sum(st.get("far_pairs", [0])))}
        return out

    res = {
        "metric": "steps/s at N=4M, mixed+learned vs plain FP64, at 1e-4 rms force err",
        "value": 1.0 / head["s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Philox Gaussian plasma, Maxwell-Boltzmann velocities)",
        "config": {"workload": "barnes-hut plasma step, %d e- + %d ions, sigma %g um, T_e %g K, dt 5 ps, "
                               "each walk tuned to 1e-4 rms force error" % (n, n, a.sigma, a.te),
                   "n_total": 2 * n, "variant": head["variant"]},
        "accuracy": {"rms_rel_force_err": head["rms"], "max_rel_force_err": head["max"],
                     "energy_drift": head.get("drift"), "sample": "4096 targets vs FP64 direct sum"},
        "variants": {"mixed+learned": summary(learned), "mixed": summary(fixed), "plain_fp64": summary(plain)},
        "speedup_vs_plain_fp64": None if plain is None else plain["s"] / head["s"],
        "mlp": mlp_info,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "frac_of_nonpacked_fp32": achieved / PEAK["fp32_nonpacked"],
                     "kernel": "k_walk32 + k_far_eval (one persistent walk launch over both target species, then the far cells)",
                     "ms_per_launch": head["ms_walk"], "flops_per_launch": flops},
    }
    if not a.no_cpu:
        log("timing the reference CPU path on a bounded sample")
        # every CPU this job may use: the OpenMP thread count the host sets
        # for it (the CPU share of one GPU), else the CPUs it is bound to
        avail = len(os.sched_getaffinity(0))
        threads = min(avail, int(os.environ.get("OMP_NUM_THREADS") or avail))
        th_cpu = plain["theta"] if plain is not None else 0.5
        res["cpu_baseline"] = cpu_baseline(qel, qion, th_cpu, rs, threads)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
