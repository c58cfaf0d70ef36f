"""The C-ABI library without a GPU: it loads, exports every entry point that
include/treemi.h declares, and its host-only helpers (generator, Kelbg
length) behave as stated."""
import os
import re
import ctypes

import numpy as np

import tree_amd as ta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_every_declared_symbol():
    lib = ta.load_library()
    hdr = open(os.path.join(ROOT, "include", "treemi.h")).read()
    names = set(re.findall(r"^\s*(?:int|double|const char \*|void)\s*\**\s*(tb_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 17
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_struct_layouts():
    """The Python mirrors have the sizes the compiled library uses."""
    lib = ta.load_library()
    s = np.zeros(3, np.int32)
    assert lib.tb_abi_sizes(s[0:].ctypes.data, s[1:].ctypes.data, s[2:].ctypes.data) == 0
    assert s[0] == ta.PT_DTYPE.itemsize == 80  # TYPE PT + padding
    assert s[1] == ctypes.sizeof(ta.engine.Config)
    assert s[2] == ctypes.sizeof(ta.engine.Stats)


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors (Salmon et al., Random123 KAT)."""
    lib = ta.load_library()
    f = lib.tb_philox4x32_10
    f.argtypes = [ctypes.c_void_p] * 3
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        o = np.zeros(4, np.uint32)
        f(c.ctypes.data, k.ctypes.data, o.ctypes.data)
        assert tuple(int(v) for v in o) == want


def test_kelbg_length_table():
    for te, lam in ((1.0, 2.1027e-2), (10.0, 6.6493e-3), (100.0, 2.1027e-3)):
        assert abs(ta.kelbg_length(te) / lam - 1) < 2e-4


def test_generator_statistics_and_temperature():
    n = 20000
    qel, qion = ta.gen_plasma(42, n, n, sigma=100.0, te=10.0, ti=1e-3)
    for q in (qel, qion):
        x = q["xx"]
        assert abs(x.mean()) < 3 * 100.0 / np.sqrt(3 * n)
        assert abs(x.std() / 100.0 - 1) < 0.02
        assert np.array_equal(q["id"], np.arange(1, n + 1))
    kb = 8.617343e-5
    for q, m, t in ((qel, ta.XM_E, 10.0), (qion, ta.XM_I, 1e-3)):
        mv2 = m * (q["xv"] ** 2).sum()
        assert np.isclose(mv2 / kb / (3 * n - 6), t, rtol=1e-12)  # exact-T rescale, main.f90:731-737
    a, b = ta.gen_plasma(42, 100, 100, 100.0)
    c, d = ta.gen_plasma(42, 100, 100, 100.0)
    assert np.array_equal(a, c) and np.array_equal(b, d)
    e, _ = ta.gen_plasma(43, 100, 100, 100.0)
    assert not np.array_equal(a["xx"], e["xx"])


def test_no_gpu_fails_loudly():
    """Without a GPU the engine refuses to start (no CPU fallback)."""
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        return
    try:
        ta.Engine(rs=6.6e-3)
    except ta.TreeError as e:
        assert e.code in (7, 2)
    else:
        raise AssertionError("engine started without a GPU")


def test_config_is_checked_before_any_device():
    """tb_create rejects a bad configuration with TB_ERR_ARG and a message
    before it looks for a device (the same on a GPU box as without one):
    theta, group, far_ratio, far_levels and the learned criterion's cost
    weight outside their ranges."""
    bad = [dict(theta=0.0), dict(group=100), dict(far_ratio=1.5), dict(far_levels=3),
           dict(cost_weight=-0.1), dict(cost_weight=5.0), dict(cost_weight=float("nan"))]
    for kw in bad:
        try:
            ta.Engine(**{"rs": 6.6e-3, "walk": "mixed", **kw})
        except ta.TreeError as e:
            assert e.code == 1, (kw, e)
            assert str(e), kw
        else:
            raise AssertionError("accepted %r" % kw)
