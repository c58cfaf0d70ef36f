"""Headline benchmark: steps/s of the Barnes-Hut step at N = 4M on one MI355X.

Workload (default): 2^21 electrons + 2^21 singly charged ions (m_i = 85 amu),
seeded spherical Gaussian positions (sigma = 600 um), T_e = 10 K, T_i = 1 mK,
Kelbg length at T_e, dt = 5 ps.  A step is VV1 + tree build + force + VV2,
with all particle state resident in HBM.  Each walk's opening angle is first
tuned (untimed bisection) so that its rms relative force error against an FP64
direct sum on 4096 sampled targets (2048 e- + 2048 ions, seed 99) is 1e-4.

Prints one JSON line (rank 0).  `value` is steps/s of the headline variant;
the plain FP64 walk is reported beside it.  `roofline` covers the dominant
kernel (the force walk) from device-event timings and interaction counts;
`cpu_baseline` times the reference's own tree code (compiled from its sources
into oracle/_ref) on a bounded sample of the same workload.
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

# algorithmic flop weights per event (FMA = 2): MAC test, M2P on accept, P2P Coulomb / Kelbg
W_VISIT, W_M2P, W_P2P_C, W_P2P_K = 11, 144, 20, 28
# mixed walk, per target-entry interaction, counted from walk32.hip's FP32 code
# (monopole + dipole + traceless-quadrupole M2P 72, Coulomb P2P 18, Kelbg P2P
# 30); MAC tests are per cell and group and are not counted
W32_M2P, W32_P2P_C, W32_P2P_K = 72, 18, 30
PEAK = {"fp64": 78.6, "fp32": 157.3}  # TFLOP/s, MI355X vector


def sample_targets(ne, ni, seed=99, n=4096):
    rng = np.random.default_rng(seed)
    te = rng.choice(ne, n // 2, replace=False)
    ti = rng.choice(ni, n // 2, replace=False)
    tsp = np.r_[np.zeros(n // 2, np.int32), np.ones(n // 2, np.int32)]
    return tsp, np.r_[te, ti].astype(np.int32)


def force_error(eng, tsp, tix, fdir):
    eng.build()
    eng.force()
    e, i = eng.get_particles()
    f = np.where(tsp[:, None] == 0, e["ff"][np.where(tsp == 0, tix, 0)], i["ff"][np.where(tsp == 1, tix, 0)])
    rel = np.linalg.norm(f - fdir, axis=1) / np.linalg.norm(fdir, axis=1)
    return float(np.sqrt(np.mean(rel ** 2))), float(rel.max())


def tune_theta(eng, key, tsp, tix, fdir, target=1e-4, lo=0.1, hi=1.0, iters=7, log=None):
    """Largest theta with rms error <= target (bisection on a monotone curve)."""
    best = None
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        eng.configure(**{key: mid})
        rms, mx = force_error(eng, tsp, tix, fdir)
        if log:
            log("  %s=%.4f rms=%.3e max=%.3e" % (key, mid, rms, mx))
        if rms <= target:
            best = (mid, rms, mx)
            lo = mid
        else:
            hi = mid
    if best is None:
        eng.configure(**{key: lo})
        rms, mx = force_error(eng, tsp, tix, fdir)
        best = (lo, rms, mx)
    eng.configure(**{key: best[0]})
    return best


def walk_flops(st, like_pairs=(0, 3)):
    fl = 0
    for p in range(4):
        wp = W_P2P_C if p in like_pairs else W_P2P_K
        fl += W_VISIT * st["visits"][p] + W_M2P * st["m2p"][p] + wp * st["p2p"][p]
    return fl


def walk32_flops(st, like_pairs=(0, 3)):
    fl = 0
    for p in range(4):
        wp = W32_P2P_C if p in like_pairs else W32_P2P_K
        fl += W32_M2P * st["m2p"][p] + wp * st["p2p"][p]
    return fl


def timed_run(eng, dt, steps, warmup):
    eng.build()
    ep = eng.force()
    v0 = eng.vv2(0.0, True)  # kinetic sums of the current state (zero kick)
    e0 = 0.5 * (v0[0] + v0[1]) + ep[0] + ep[1]
    if warmup:
        eng.step(dt, warmup)
    ms_force = []
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = eng.step(dt, 1)
        ms_force.append(eng.stats()["ms_force"])
    t1 = time.perf_counter()
    e1 = 0.5 * (out[2] + out[3]) + out[0] + out[1]
    return (t1 - t0) / steps, float(np.mean(ms_force)), (e1 - e0) / abs(e0)


def cpu_baseline(qel, qion, theta, rs, threads):
    """Reference tree code on the host: serial build + OpenMP walk of every stride-th target."""
    try:
        from oracle import pyoracle
    except Exception as ex:  # pragma: no cover
        return {"value": None, "error": str(ex)}
    stride = 256
    n = len(qel) + len(qion)
    if pyoracle.ref_available():
        buf = pyoracle.run_ref(3, qel["xx"], qel["xv"], qion["xx"], qion["xv"], theta, rs, (ta.XM_E, ta.XM_I),
                               0.0, stride=stride, threads=threads)
        tb, tw, nw = struct.unpack_from("ddi", buf, 0)
        kind = "reference"
    else:
        tb, tw, nw = pyoracle.time_sample(qel["xx"], qion["xx"], theta, rs, stride, threads)
        kind = "port"
    t_step = tb + tw * n / nw
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": "one serial tree build of both species (%.1f s) + walk of every %d-th target (%d targets, "
                      "%.1f s), scaled to all %d targets; theta=%.3f" % (tb, stride, nw, tw, n, theta)}


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
    ap.add_argument("--theta64", type=float, default=0.0, help="fixed theta for the FP64 walk (skip tuning)")
    ap.add_argument("--theta32", type=float, default=0.0, help="fixed theta for the mixed walk (skip tuning)")
    ap.add_argument("--group", type=int, default=0, help="mixed-walk targets per group (0 = autotuned)")
    ap.add_argument("--only", choices=["all", "mixed", "ref64"], default="all")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the engine holds one whole domain per GPU and this build has no "
              "multi-GPU path (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    log = (lambda s: None) if a.quiet else (lambda s: print(s, file=sys.stderr, flush=True))

    n = a.n
    rs = ta.kelbg_length(a.te)
    dt = 0.005 / ta.TPS
    qel, qion = ta.gen_plasma(a.seed, n, n, a.sigma, te=a.te, ti=1e-3)
    tsp, tix = sample_targets(n, n)

    eng = ta.Engine(rs=rs, theta=0.5, walk="ref64")
    eng.set_particles(qel, qion)
    fdir = eng.direct_sample(tsp, tix)
    plain = None
    if a.only in ("all", "ref64"):
        if a.theta64 > 0:
            eng.configure(theta=a.theta64)
            th64 = a.theta64
            rms64, mx64 = force_error(eng, tsp, tix, fdir)
        else:
            log("tuning plain FP64 walk")
            th64, rms64, mx64 = tune_theta(eng, "theta", tsp, tix, fdir, log=log)
        eng.build()
        st64 = eng.count_interactions()
        eng.set_particles(qel, qion)
        s64, msf64, drift64 = timed_run(eng, dt, a.steps, a.warmup)
        log("plain FP64: theta=%.4f rms=%.2e %.2f steps/s walk %.1f ms" % (th64, rms64, 1 / s64, msf64))
        plain = {"variant": "plain-fp64", "theta": th64, "rms": rms64, "max": mx64, "s": s64, "ms_force": msf64,
                 "drift": drift64, "st": st64, "prec": "fp64"}
    mixed = None
    if a.only in ("all", "mixed"):
        eng.set_particles(qel, qion)
        eng.configure(walk="mixed", theta_mixed=0.5, group=a.group)
        if a.theta32 > 0:
            eng.configure(theta_mixed=a.theta32)
            th32 = a.theta32
            rms32, mx32 = force_error(eng, tsp, tix, fdir)
        else:
            log("tuning mixed walk")
            th32, rms32, mx32 = tune_theta(eng, "theta_mixed", tsp, tix, fdir, log=log)
        eng.build()
        st32 = eng.count_interactions()
        eng.set_particles(qel, qion)
        s32, msf32, drift32 = timed_run(eng, dt, a.steps, a.warmup)
        log("mixed: theta=%.4f rms=%.2e %.2f steps/s walk %.1f ms" % (th32, rms32, 1 / s32, msf32))
        mixed = {"variant": "mixed", "theta": th32, "rms": rms32, "max": mx32, "s": s32, "ms_force": msf32,
                 "drift": drift32, "st": st32, "prec": "fp32"}
    eng.close()

    head = mixed if mixed is not None else plain
    flops = walk32_flops(head["st"]) if head["prec"] == "fp32" else walk_flops(head["st"])
    achieved = flops / (head["ms_force"] * 1e-3) / 1e12
    peak = PEAK[head["prec"]]
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
        "dtype": head["prec"],
        "data": "synthetic (seeded Philox Gaussian plasma, Maxwell-Boltzmann velocities)",
        "config": {"workload": "barnes-hut plasma step, %d e- + %d ions, sigma %g um, T_e %g K, dt 5 ps, "
                               "theta tuned to 1e-4 rms force error" % (n, n, a.sigma, a.te),
                   "n_total": 2 * n, "variant": head["variant"], "theta": head["theta"]},
        "accuracy": {"rms_rel_force_err": head["rms"], "max_rel_force_err": head["max"],
                     "energy_drift": head["drift"], "sample": "4096 targets vs FP64 direct sum"},
        "plain_fp64": None if plain is None else {
            "steps_per_s": 1.0 / plain["s"], "theta": plain["theta"], "rms": plain["rms"], "max": plain["max"],
            "ms_walk": plain["ms_force"], "energy_drift": plain["drift"],
            "walk_tflops": walk_flops(plain["st"]) / (plain["ms_force"] * 1e-3) / 1e12},
        "speedup_vs_plain_fp64": None if plain is None or mixed is None else plain["s"] / mixed["s"],
        "interactions_per_s": (sum(head["st"]["m2p"]) + sum(head["st"]["p2p"])) / head["s"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": "k_walk32 (persistent, both species)" if head["prec"] == "fp32"
                     else "k_walk64 (both target species)",
                     "flops_per_step": flops},
    }
    if not a.no_cpu:
        log("timing the reference CPU path on a bounded sample")
        threads = min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(qel, qion, head["theta"], rs, threads)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
