"""ctypes wrappers of the CPU oracle (liboracle.so) and the reference harness.

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker or the timed CPU baseline; the
product (tree_amd) never imports it.
"""
import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "ref_harness")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_local_coulomb_sum.argtypes = [I, P, I, P, D, D, P, P, P, P, I]
        L.orc_local_coulomb_sum_rdf.argtypes = [I, P, I, P, D, D, P, P, P, P, I, D, D, I, P, P]
        L.orc_export_tree.argtypes = [I, P, I, I] + [P] * 12
        L.orc_direct.argtypes = [I, P, I, P, D, I, P, P, P, P, I]
        L.orc_direct.restype = None
        L.orc_vv1.argtypes = [I, P, P, P, D, D]
        L.orc_vv1.restype = None
        L.orc_vv2.argtypes = [I, P, P, P, D, D, I, P]
        L.orc_vv2.restype = None
        L.orc_time_sample.argtypes = [I, P, I, P, D, D, I, I, P, P, P]
        L.orc_force_targets.argtypes = [I, P, I, P, D, D, I, P, P, P, I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _aos(x):
    return np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)


NDIST = 10000  # data.f90:38: the histograms hold NDIST + 1 bins


def local_coulomb_sum(xe, xi, theta, rs, nthreads=0):
    """Forces (fe, fi), (Epot_e, Epot_i) and 12 interaction counters."""
    xe, xi = _aos(xe), _aos(xi)
    fe, fi = np.zeros_like(xe), np.zeros_like(xi)
    ep = np.zeros(2)
    cnt = np.zeros(12, np.int64)
    rc = lib().orc_local_coulomb_sum(len(xe), _p(xe), len(xi), _p(xi), theta, rs, _p(fe), _p(fi), _p(ep),
                                     _p(cnt), nthreads)
    if rc:
        raise RuntimeError("oracle: coincident particles")
    return fe, fi, ep, cnt


def export_tree(x, sign):
    x = _aos(x)
    cap = 64 * len(x) + 8
    a = {k: np.zeros(cap, np.int32) for k in ("depth", "leaf", "pt", "count", "next")}
    for k in ("xx", "lx", "D", "Q", "QQ"):
        a[k] = np.zeros((cap, 3))
    for k in ("qeff", "dx"):
        a[k] = np.zeros(cap)
    n = lib().orc_export_tree(len(x), _p(x), sign, cap, _p(a["depth"]), _p(a["leaf"]), _p(a["pt"]),
                              _p(a["count"]), _p(a["next"]), _p(a["xx"]), _p(a["lx"]), _p(a["qeff"]),
                              _p(a["dx"]), _p(a["D"]), _p(a["Q"]), _p(a["QQ"]))
    if n < 0:
        raise RuntimeError("oracle export_tree failed (%d)" % n)
    return {k: v[:n] for k, v in a.items()}


def direct(xe, xi, rs, tspecies, tidx, nthreads=0):
    xe, xi = _aos(xe), _aos(xi)
    tspecies = np.ascontiguousarray(tspecies, np.int32)
    tidx = np.ascontiguousarray(tidx, np.int32)
    f = np.zeros((len(tidx), 3))
    pot = np.zeros(len(tidx))
    lib().orc_direct(len(xe), _p(xe), len(xi), _p(xi), rs, len(tidx), _p(tspecies), _p(tidx), _p(f), _p(pot),
                     nthreads)
    return f, pot


def force_targets(xe, xi, theta, rs, tspecies, tidx, nthreads=0):
    """The tree walk's forces on listed targets only (same trees, same order as local_coulomb_sum)."""
    xe, xi = _aos(xe), _aos(xi)
    tspecies = np.ascontiguousarray(tspecies, np.int32)
    tidx = np.ascontiguousarray(tidx, np.int32)
    f = np.zeros((len(tidx), 3))
    rc = lib().orc_force_targets(len(xe), _p(xe), len(xi), _p(xi), theta, rs, len(tidx), _p(tspecies), _p(tidx),
                                 _p(f), nthreads)
    if rc:
        raise RuntimeError("oracle: coincident particles")
    return f


def vv1(x, v, f, xm, dt):
    x, v, f = _aos(x).copy(), _aos(v).copy(), _aos(f)
    lib().orc_vv1(len(x), _p(x), _p(v), _p(f), xm, dt)
    return x, v


def vv2(x, v, f, xm, dt, sample=True):
    x, v, f = _aos(x), _aos(v).copy(), _aos(f)
    out = np.zeros(2)
    lib().orc_vv2(len(x), _p(x), _p(v), _p(f), xm, dt, 1 if sample else 0, _p(out))
    return v, out


def time_sample(xe, xi, theta, rs, stride, nthreads):
    """Serial build + OpenMP walk of every stride-th target: (t_build, t_walk, n_walked)."""
    xe, xi = _aos(xe), _aos(xi)
    tb, tw = ctypes.c_double(), ctypes.c_double()
    nw = ctypes.c_int64()
    rc = lib().orc_time_sample(len(xe), _p(xe), len(xi), _p(xi), theta, rs, stride, nthreads, ctypes.byref(tb),
                               ctypes.byref(tw), ctypes.byref(nw))
    if rc < 0:
        raise RuntimeError("oracle: coincident particles")
    return tb.value, tw.value, nw.value


# ------------------------------------------------------ reference harness

def ref_available():
    return os.path.exists(REF)


def run_ref(mode, xe, ve, xi, vi, theta, rs, xm, dt, nsteps=0, stride=1, threads=1):
    """Run the reference's own BINTREE (oracle/_ref/ref_harness) and return raw output bytes."""
    xe, ve, xi, vi = _aos(xe), _aos(ve), _aos(xi), _aos(vi)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("5i", mode, len(xe), len(xi), nsteps, stride))
            f.write(struct.pack("5d", theta, rs, xm[0], xm[1], dt))
            for a in (xe, ve, xi, vi):
                f.write(a.tobytes())
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run([REF, fin, fout], check=True, env=env)
        with open(fout, "rb") as f:
            return f.read()


_NODE = np.dtype([("i", "<i4", (4,)), ("d", "<f8", (17,))])


def parse_ref_mode1(buf, ne, ni):
    """Node tables (depth, ileaf, pt (1-based), next, 17 doubles) and forces of a mode-1 run."""
    off = 0
    trees = []
    for _ in range(2):
        n = struct.unpack_from("i", buf, off)[0]
        off += 4
        trees.append(np.frombuffer(buf, dtype=_NODE, count=n, offset=off).copy())
        off += n * _NODE.itemsize
    f = np.frombuffer(buf, dtype="<f8", offset=off, count=3 * ne + 3 * ni + 2)
    fe = f[: 3 * ne].reshape(ne, 3).copy()
    fi = f[3 * ne: 3 * ne + 3 * ni].reshape(ni, 3).copy()
    ep = f[3 * ne + 3 * ni: 3 * ne + 3 * ni + 2].copy()
    return trees, fe, fi, ep


def parse_ref_rdf(buf, ne, ni):
    """The RDF histograms (RDF_ei, RDF_ii: NDIST + 1 int32 each) at the end of a mode-1 run."""
    h = np.frombuffer(buf[len(buf) - 8 * (NDIST + 1):], dtype="<i4")
    return h[: NDIST + 1].astype(np.int64), h[NDIST + 1:].astype(np.int64)


def local_coulomb_sum_rdf(xe, xi, theta, rs, dx_ei, dx_ii, nbin=None, nthreads=0):
    """local_coulomb_sum plus the leaf-pair RDF histograms (rdf_ei, rdf_ii)."""
    nbin = NDIST + 1 if nbin is None else nbin
    xe, xi = _aos(xe), _aos(xi)
    fe, fi = np.zeros_like(xe), np.zeros_like(xi)
    ep = np.zeros(2)
    cnt = np.zeros(12, np.int64)
    hei, hii = np.zeros(nbin, np.int64), np.zeros(nbin, np.int64)
    rc = lib().orc_local_coulomb_sum_rdf(len(xe), _p(xe), len(xi), _p(xi), theta, rs, _p(fe), _p(fi), _p(ep),
                                         _p(cnt), nthreads, dx_ei, dx_ii, nbin, _p(hei), _p(hii))
    if rc:
        raise RuntimeError("oracle: coincident particles")
    return fe, fi, ep, cnt, hei, hii


def parse_ref_mode2(buf, ne, ni, nsteps):
    f = np.frombuffer(buf, dtype="<f8")
    steps = f[: 6 * nsteps].reshape(nsteps, 6).copy()
    rest = f[6 * nsteps:]
    e = rest[: 9 * ne].reshape(ne, 9).copy()
    i = rest[9 * ne: 9 * ne + 9 * ni].reshape(ni, 9).copy()
    return steps, e, i
