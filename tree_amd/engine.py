"""Host-side mirror of the reference's force / time-step interface over the
C-ABI of libtreemi.so (include/treemi.h).

The reference's subroutines and the calls that replace them:

    BuildLocalTree(NS, gel, gion, Roote, Rooti, ttime)   tree.f90:39     -> BuildLocalTree(eng)
    LocalCoulombSum(NS, Roote, Rooti, qel, qion, sys, ...) tree.f90:662  -> LocalCoulombSum(eng)
    Force(NS, qel, qion, param, ttime)                   main.f90:506    -> Force(eng)
    VVerletNotemp1(NS, qel, qion, param, dt, ttime)      vverlet.f90:9   -> VVerletNotemp1(eng, dt)
    VVerletNotemp2(NS, qel, qion, param, sys, dt, ttime) vverlet.f90:49  -> VVerletNotemp2(eng, dt, sample)

Particle arrays use PT_DTYPE, the 80-byte mirror of TYPE PT (data.f90:47-50).
Errors raise TreeError (the reference STOPs).  There is no CPU fallback: if
libtreemi.so is missing or no GPU is present the calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtreemi.so")

# TYPE PT: xx(3), xv(3), ff(3) REAL*8, id INTEGER (+4 bytes of padding)
PT_DTYPE = np.dtype([("xx", "<f8", (3,)), ("xv", "<f8", (3,)), ("ff", "<f8", (3,)), ("id", "<i4"), ("pad_", "<i4")])
assert PT_DTYPE.itemsize == 80

EPS = 0.0014399835661145824      # eV*um, data.f90:40
TPS = 0.10180505305371923        # ns per code time unit, data.f90:39
XM_E = 5.4858e-4                 # amu
XM_I = 85.0                      # amu (Rb-85)

NOALLOC_MAX_TARGETS = 65536  # TB_NOALLOC_MAX_TARGETS, include/treemi.h
NDIST = 10000  # histogram bins of the reference's RDF, less one (data.f90:38)
WALK = {"ref64": 0, "mixed": 1}
MLP_NPARAMS = 1377
MAC = {"fixed": 0, "learned": 1}

ERRORS = {1: "ARG", 2: "HIP", 3: "DEPTH", 4: "NOMEM", 5: "STATE", 6: "NONFINITE", 7: "NODEV"}


class TreeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tb error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("xm", ctypes.c_double * 2),
        ("rs", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("theta_mixed", ctypes.c_double),
        ("theta_lo", ctypes.c_double),
        ("err_target", ctypes.c_double),
        ("walk", ctypes.c_int32),
        ("mac", ctypes.c_int32),
        ("group", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("mem_cap", ctypes.c_int64),
        ("far_ratio", ctypes.c_double),
        ("far_levels", ctypes.c_int32),
        ("cost_weight", ctypes.c_float),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int64 * 2),
        ("max_depth", ctypes.c_int32 * 2),
        ("pad_", ctypes.c_int32),
        ("visits", ctypes.c_int64 * 4),
        ("m2p", ctypes.c_int64 * 4),
        ("p2p", ctypes.c_int64 * 4),
        ("ms_build", ctypes.c_double),
        ("ms_force", ctypes.c_double),
        ("ms_vv", ctypes.c_double),
        ("ms_walk", ctypes.c_double),
        ("m2l", ctypes.c_int64 * 4),
        ("far_pairs", ctypes.c_int64 * 2),
        ("seg_pairs", ctypes.c_int64 * 2),
        ("m2p3", ctypes.c_int64 * 2),
    ]


_lib = None

# every entry point declared in include/treemi.h: (name, restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
SIGNATURES = [
    ("tb_create", ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(Config)]),
    ("tb_destroy", ctypes.c_int, [_P]),
    ("tb_set_config", ctypes.c_int, [_P, ctypes.POINTER(Config)]),
    ("tb_set_particles", ctypes.c_int, [_P, _I, _P, _I, _P]),
    ("tb_get_particles", ctypes.c_int, [_P, _P, _P]),
    ("tb_build", ctypes.c_int, [_P]),
    ("tb_force", ctypes.c_int, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    ("tb_vv1", ctypes.c_int, [_P, _D]),
    ("tb_vv2", ctypes.c_int, [_P, _D, _I, _P]),
    ("tb_step", ctypes.c_int, [_P, _D, _I, _P]),
    ("tb_export_tree", ctypes.c_int, [_P, _I, _I, _P] + [_P] * 12),
    ("tb_direct_sample", ctypes.c_int, [_P, _I, _P, _P, _P]),
    ("tb_direct_sample_noalloc", ctypes.c_int, [_P, _I, _P, _P, _P]),
    ("tb_local_eval", ctypes.c_int, [_P, _I, _I, _P, _I, _P, _P]),
    ("tb_force_targets", ctypes.c_int, [_P, _I, _P, _P, _P]),
    ("tb_time_walk", ctypes.c_int, [_P, _I, _P, _P]),
    ("tb_count_interactions", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("tb_get_stats", ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    ("tb_rdf_enable", ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, _I]),
    ("tb_rdf_get", ctypes.c_int, [_P, _I, _P, _P, _I]),
    ("tb_mlp_set", ctypes.c_int, [_P, _P, _I]),
    ("tb_mlp_get", ctypes.c_int, [_P, _P, _I]),
    ("tb_mlp_init", ctypes.c_int, [_P, ctypes.c_uint64]),
    ("tb_mlp_forward", ctypes.c_int, [_P, _I, _P, _I, _P]),
    ("tb_mlp_loss_grad", ctypes.c_int, [_P, _I, _P, _P, _P, _P]),
    ("tb_mlp_train", ctypes.c_int, [_P, _I, _P, _P, _I, _I, ctypes.c_float, ctypes.c_uint64, _P]),
    ("tb_mlp_collect", ctypes.c_int, [_P, _I, _P, _P, _P]),
    ("tb_autotune", ctypes.c_int, [_P, _P, _P]),
    ("tb_ensemble_create", ctypes.c_int, [ctypes.POINTER(_P), _I, _P]),
    ("tb_ensemble_destroy", ctypes.c_int, [_P]),
    ("tb_ensemble_set_particles", ctypes.c_int, [_P, _I, _I, _P, _I, _P]),
    ("tb_ensemble_get_particles", ctypes.c_int, [_P, _I, _P, _P]),
    ("tb_ensemble_force", ctypes.c_int, [_P, _P]),
    ("tb_ensemble_step", ctypes.c_int, [_P, _D, _I, _P]),
    ("tb_ensemble_autotune", ctypes.c_int, [_P, _P, _P]),
    ("tb_ensemble_walk_lists", ctypes.c_int, [_P, _P, _P]),
    ("tb_walk_lists", ctypes.c_int, [_P, _P, _P]),
    ("tb_last_error", ctypes.c_char_p, []),
    ("tb_abi_sizes", ctypes.c_int, [_P, _P, _P]),
    ("tb_kelbg_length", ctypes.c_double, [_D, _D, _D]),
    ("tb_gen_plasma", ctypes.c_int, [ctypes.c_uint64, _I, _I, _D, _D, _D, _D, _D, _P, _P]),
]


def load_library(path=LIB_PATH):
    """Load libtreemi.so (built by tree_amd/build.py); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libtreemi.so not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise TreeError(rc, _lib.tb_last_error().decode())


def kelbg_length(te, xm_e=XM_E, xm_i=XM_I):
    """Thermal de Broglie length hbar/sqrt(2 mu k_B T_e) in um."""
    return load_library().tb_kelbg_length(te, xm_e, xm_i)


def gen_plasma(seed, ne, ni, sigma, te=10.0, ti=1e-3, xm_e=XM_E, xm_i=XM_I):
    """Seeded Gaussian plasma (positions sigma um, Maxwell-Boltzmann velocities)."""
    lib = load_library()
    qel = np.zeros(ne, PT_DTYPE)
    qion = np.zeros(ni, PT_DTYPE)
    _check(lib.tb_gen_plasma(seed, ne, ni, sigma, te, ti, xm_e, xm_i, _ptr(qel), _ptr(qion)))
    return qel, qion


class Engine:
    """One resident domain on one GPU (a tb_handle)."""

    def __init__(self, rs, theta=0.5, xm=(XM_E, XM_I), walk="ref64", theta_mixed=0.0, mac="fixed",
                 theta_lo=0.0, err_target=0.0, group=0, device=0, mem_cap=0, far_ratio=0.0, far_levels=0,
                 cost_weight=0.0):
        self.lib = load_library()
        self.cfg = Config()
        self.cfg.xm[0], self.cfg.xm[1] = xm
        self.cfg.rs = rs
        self.cfg.theta = theta
        self.cfg.theta_mixed = theta_mixed
        self.cfg.theta_lo = theta_lo
        self.cfg.err_target = err_target
        self.cfg.walk = WALK[walk]
        self.cfg.mac = MAC[mac]
        self.cfg.group = group
        self.cfg.device = device
        self.cfg.mem_cap = mem_cap
        self.cfg.far_ratio = far_ratio
        self.cfg.far_levels = far_levels
        self.cfg.cost_weight = cost_weight
        h = ctypes.c_void_p()
        _check(self.lib.tb_create(ctypes.byref(h), ctypes.byref(self.cfg)))
        self.h = h
        self.ne = self.ni = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.tb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def configure(self, **kw):
        for k, v in kw.items():
            if k == "walk":
                v = WALK[v]
            elif k == "mac":
                v = MAC[v]
            setattr(self.cfg, k, v)
        _check(self.lib.tb_set_config(self.h, ctypes.byref(self.cfg)))

    def set_particles(self, qel, qion):
        qel = np.ascontiguousarray(qel, dtype=PT_DTYPE)
        qion = np.ascontiguousarray(qion, dtype=PT_DTYPE)
        _check(self.lib.tb_set_particles(self.h, len(qel), _ptr(qel), len(qion), _ptr(qion)))
        self.ne, self.ni = len(qel), len(qion)

    def get_particles(self):
        qel = np.zeros(self.ne, PT_DTYPE)
        qion = np.zeros(self.ni, PT_DTYPE)
        _check(self.lib.tb_get_particles(self.h, _ptr(qel), _ptr(qion)))
        return qel, qion

    def build(self):
        _check(self.lib.tb_build(self.h))

    def force(self):
        ee, ei = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.tb_force(self.h, ctypes.byref(ee), ctypes.byref(ei)))
        return ee.value, ei.value

    def vv1(self, dt):
        _check(self.lib.tb_vv1(self.h, dt))

    def vv2(self, dt, sample=True):
        out = np.zeros(4)
        _check(self.lib.tb_vv2(self.h, dt, 1 if sample else 0, _ptr(out)))
        return out

    def step(self, dt, nsteps=1):
        out = np.zeros(6)
        _check(self.lib.tb_step(self.h, dt, nsteps, _ptr(out)))
        return out

    def export_tree(self, species):
        n = self.ne if species == 0 else self.ni
        cap = 22 * n + 8
        a = {k: np.zeros(cap, np.int32) for k in ("depth", "leaf", "pt", "count", "next")}
        for k in ("xx", "lx", "D", "Q", "QQ"):
            a[k] = np.zeros((cap, 3))
        for k in ("qeff", "dx"):
            a[k] = np.zeros(cap)
        nn = ctypes.c_int32()
        _check(self.lib.tb_export_tree(self.h, species, cap, ctypes.byref(nn), _ptr(a["depth"]), _ptr(a["leaf"]),
                                       _ptr(a["pt"]), _ptr(a["count"]), _ptr(a["next"]), _ptr(a["xx"]),
                                       _ptr(a["lx"]), _ptr(a["qeff"]), _ptr(a["dx"]), _ptr(a["D"]), _ptr(a["Q"]),
                                       _ptr(a["QQ"])))
        return {k: v[: nn.value] for k, v in a.items()}

    def direct_sample(self, tspecies, tidx):
        tspecies = np.ascontiguousarray(tspecies, np.int32)
        tidx = np.ascontiguousarray(tidx, np.int32)
        f = np.zeros((len(tidx), 3))
        _check(self.lib.tb_direct_sample(self.h, len(tidx), _ptr(tspecies), _ptr(tidx), _ptr(f)))
        return f

    def direct_sample_noalloc(self, tspecies, tidx):
        """tb_direct_sample with no device allocation (at most NOALLOC_MAX_TARGETS targets)."""
        tspecies = np.ascontiguousarray(tspecies, np.int32)
        tidx = np.ascontiguousarray(tidx, np.int32)
        f = np.zeros((len(tidx), 3))
        _check(self.lib.tb_direct_sample_noalloc(self.h, len(tidx), _ptr(tspecies), _ptr(tidx), _ptr(f)))
        return f

    def local_eval(self, order, cells, y):
        """The far-field local expansion check: cells [n][28], offsets y [m][3] ->
        [m][4] potential and field (FP32)."""
        cells = np.ascontiguousarray(cells, np.float32)
        if cells.ndim != 2 or cells.shape[1] != 28:
            raise ValueError("local_eval: cells must be [n][28]")
        y = np.ascontiguousarray(y, np.float32)
        out = np.zeros((len(y), 4), np.float32)
        _check(self.lib.tb_local_eval(self.h, order, len(cells), _ptr(cells), len(y), _ptr(y), _ptr(out)))
        return out

    def force_targets(self, tspecies, tidx):
        tspecies = np.ascontiguousarray(tspecies, np.int32)
        tidx = np.ascontiguousarray(tidx, np.int32)
        f = np.zeros((len(tidx), 3))
        _check(self.lib.tb_force_targets(self.h, len(tidx), _ptr(tspecies), _ptr(tidx), _ptr(f)))
        return f

    def time_walk(self, stride):
        """(ms, targets walked) of the FP64 walk over every stride-th block of 64 targets."""
        ms = ctypes.c_float()
        nw = ctypes.c_int64()
        _check(self.lib.tb_time_walk(self.h, stride, ctypes.byref(ms), ctypes.byref(nw)))
        return ms.value, nw.value

    def count_interactions(self):
        s = Stats()
        _check(self.lib.tb_count_interactions(self.h, ctypes.byref(s)))
        return _stats_dict(s)

    # ------------------------------------------------ learned opening criterion
    def mlp_set(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(self.lib.tb_mlp_set(self.h, _ptr(p), len(p)))

    def mlp_init(self, seed=0):
        """Seeded initial network (He-uniform weights, zero biases)."""
        _check(self.lib.tb_mlp_init(self.h, seed))

    def mlp_get(self):
        p = np.zeros(MLP_NPARAMS, np.float32)
        _check(self.lib.tb_mlp_get(self.h, _ptr(p), len(p)))
        return p

    def mlp_forward(self, X, bf16=True):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(len(X), np.float32)
        _check(self.lib.tb_mlp_forward(self.h, len(X), _ptr(X), 1 if bf16 else 0, _ptr(out)))
        return out

    def mlp_loss_grad(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        _check(self.lib.tb_mlp_loss_grad(self.h, len(X), _ptr(X), _ptr(y), _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def mlp_train(self, X, y, steps=2000, batch=4096, lr=3e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        hist = np.zeros(steps, np.float32)
        _check(self.lib.tb_mlp_train(self.h, len(X), _ptr(X), _ptr(y), steps, min(batch, len(X)), lr, seed,
                                     _ptr(hist)))
        return hist

    def mlp_collect(self, cap=200000):
        X = np.zeros((cap, 8), np.float32)
        y = np.zeros(cap, np.float32)
        n = ctypes.c_int32()
        _check(self.lib.tb_mlp_collect(self.h, cap, _ptr(X), _ptr(y), ctypes.byref(n)))
        return X[: n.value].copy(), y[: n.value].copy()

    def autotune(self):
        g = ctypes.c_int32()
        ms = np.zeros(3, np.float32)
        _check(self.lib.tb_autotune(self.h, ctypes.byref(g), _ptr(ms)))
        return g.value, ms

    def walk_lists(self):
        """(cells, leaves): the list lengths at which the mixed walk evaluates its interaction lists."""
        m, p = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.tb_walk_lists(self.h, ctypes.byref(m), ctypes.byref(p)))
        return m.value, p.value

    def rdf_enable(self, dx_ei, dx_ii, nbin=NDIST + 1):
        """Accumulate the leaf-pair RDF histograms in every force (FP64 walk); nbin = 0 disables."""
        _check(self.lib.tb_rdf_enable(self.h, dx_ei, dx_ii, nbin))
        self.rdf_nbin = nbin

    def rdf_get(self, reset=False):
        """(rdf_ei, rdf_ii) int64 histograms accumulated since enable or the last reset."""
        ei = np.zeros(self.rdf_nbin, np.int64)
        ii = np.zeros(self.rdf_nbin, np.int64)
        _check(self.lib.tb_rdf_get(self.h, self.rdf_nbin, _ptr(ei), _ptr(ii), 1 if reset else 0))
        return ei, ii

    def stats(self):
        s = Stats()
        _check(self.lib.tb_get_stats(self.h, ctypes.byref(s)))
        return _stats_dict(s)


def make_config(rs, theta=0.5, xm=(XM_E, XM_I), walk="ref64", theta_mixed=0.0, mac="fixed", theta_lo=0.0,
                err_target=0.0, group=0, device=0, mem_cap=0, far_ratio=0.0, far_levels=0, cost_weight=0.0):
    c = Config()
    c.xm[0], c.xm[1] = xm
    c.rs, c.theta, c.theta_mixed, c.theta_lo, c.err_target = rs, theta, theta_mixed, theta_lo, err_target
    c.walk, c.mac, c.group, c.device, c.mem_cap = WALK[walk], MAC[mac], group, device, mem_cap
    c.far_ratio = far_ratio
    c.far_levels = far_levels
    c.cost_weight = cost_weight
    return c


class Ensemble:
    """Independent realisations stepped together: one build read-back and one
    persistent mixed-walk launch per step (tb_ensemble_*)."""

    def __init__(self, configs):
        self.lib = load_library()
        arr = (Config * len(configs))(*configs)
        h = ctypes.c_void_p()
        _check(self.lib.tb_ensemble_create(ctypes.byref(h), len(configs), ctypes.cast(arr, ctypes.c_void_p)))
        self.h = h
        self.n = len(configs)
        self.sizes = [(0, 0)] * self.n

    def close(self):
        if getattr(self, "h", None):
            self.lib.tb_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_particles(self, r, qel, qion):
        qel = np.ascontiguousarray(qel, dtype=PT_DTYPE)
        qion = np.ascontiguousarray(qion, dtype=PT_DTYPE)
        _check(self.lib.tb_ensemble_set_particles(self.h, r, len(qel), _ptr(qel), len(qion), _ptr(qion)))
        self.sizes[r] = (len(qel), len(qion))

    def get_particles(self, r):
        qel = np.zeros(self.sizes[r][0], PT_DTYPE)
        qion = np.zeros(self.sizes[r][1], PT_DTYPE)
        _check(self.lib.tb_ensemble_get_particles(self.h, r, _ptr(qel), _ptr(qion)))
        return qel, qion

    def force(self):
        ep = np.zeros((self.n, 2))
        _check(self.lib.tb_ensemble_force(self.h, _ptr(ep)))
        return ep

    def step(self, dt, nsteps=1):
        out = np.zeros((self.n, 6))
        _check(self.lib.tb_ensemble_step(self.h, dt, nsteps, _ptr(out)))
        return out

    def autotune(self):
        g = ctypes.c_int32()
        ms = np.zeros(3, np.float32)
        _check(self.lib.tb_ensemble_autotune(self.h, ctypes.byref(g), _ptr(ms)))
        return g.value, ms

    def walk_lists(self):
        m, p = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.tb_ensemble_walk_lists(self.h, ctypes.byref(m), ctypes.byref(p)))
        return m.value, p.value


def _stats_dict(s):
    return {
        "nodes": list(s.nodes), "max_depth": list(s.max_depth),
        "visits": list(s.visits), "m2p": list(s.m2p), "p2p": list(s.p2p),
        "ms_build": s.ms_build, "ms_force": s.ms_force, "ms_vv": s.ms_vv, "ms_walk": s.ms_walk,
        "m2l": list(s.m2l), "far_pairs": list(s.far_pairs), "seg_pairs": list(s.seg_pairs), "m2p3": list(s.m2p3),
    }


# ---------------------------------------------------------------- reference names

def BuildLocalTree(eng):
    """tree.f90:39 -- both species' trees from the current positions."""
    eng.build()


def LocalCoulombSum(eng):
    """tree.f90:662 -- forces into qel/qion%ff; returns (Epot_e, Epot_i)."""
    return eng.force()


def Force(eng):
    """main.f90:506 (local part :531-541) -- build then LocalCoulombSum."""
    eng.build()
    return eng.force()


def VVerletNotemp1(eng, dt):
    """vverlet.f90:9 -- half kick and drift."""
    eng.vv1(dt)


def VVerletNotemp2(eng, dt, sample=True):
    """vverlet.f90:49 -- half kick; (mv2el, mv2ion, Te_rad, Ti_rad) when sampling."""
    return eng.vv2(dt, sample)
