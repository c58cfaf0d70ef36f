"""The far-field local expansion oracle (oracle/local_ref.py), pinned by
closed forms, finite differences, harmonicity, the Taylor series of 1/|R + y|
and direct sums over point charges."""
import numpy as np
import pytest

from oracle import local_ref as lr

R0 = np.array([0.7, -1.3, 0.9])


def test_taylor_closed_forms():
    T = lr.taylor_coeffs(R0, 3)
    x, y, z = R0
    r = np.sqrt(R0 @ R0)
    assert T[(0, 0, 0)] == pytest.approx(1 / r, rel=1e-15)
    assert T[(1, 0, 0)] == pytest.approx(-x / r ** 3, rel=1e-14)
    assert T[(0, 0, 1)] == pytest.approx(-z / r ** 3, rel=1e-14)
    assert T[(2, 0, 0)] == pytest.approx((3 * x * x - r * r) / (2 * r ** 5), rel=1e-13)
    assert T[(0, 1, 1)] == pytest.approx(3 * y * z / r ** 5, rel=1e-13)
    assert T[(3, 0, 0)] == pytest.approx((-15 * x ** 3 / r ** 7 + 9 * x / r ** 5) / 6, rel=1e-12)
    assert T[(2, 1, 0)] == pytest.approx((-15 * x * x * y / r ** 7 + 3 * y / r ** 5) / 2, rel=1e-12)
    assert T[(1, 1, 1)] == pytest.approx(-15 * x * y * z / r ** 7, rel=1e-12)


@pytest.mark.parametrize("a", [(4, 0, 0), (1, 2, 1), (0, 3, 1), (2, 2, 0)])
def test_taylor_vs_finite_differences(a):
    T = lr.taylor_coeffs(R0, 4)
    fd = lr.inv_r_derivative_fd(R0, a, h=2e-3)
    assert T[a] == pytest.approx(fd, rel=2e-4, abs=1e-6)


def test_taylor_harmonic():
    # sum_i (a_i + 2)(a_i + 1) T_{a + 2 e_i} = 0: 1/r is harmonic
    T = lr.taylor_coeffs(R0, 6)
    for a in lr.all_indices(4):
        s = 0.0
        scale = 0.0
        for i in range(3):
            b = list(a)
            b[i] += 2
            v = (a[i] + 2) * (a[i] + 1) * T[tuple(b)]
            s += v
            scale += abs(v)
        assert abs(s) <= 1e-13 * scale


@pytest.mark.parametrize("N", [3, 4, 6])
def test_taylor_series_converges(N):
    rng = np.random.default_rng(N)
    T = lr.taylor_coeffs(R0, N)
    r = np.sqrt(R0 @ R0)
    errs = []
    for rho in (0.1, 0.05):
        e = 0.0
        for _ in range(20):
            u = rng.normal(size=3)
            yv = rho * r * u / np.linalg.norm(u)
            s = sum(t * yv[0] ** a[0] * yv[1] ** a[1] * yv[2] ** a[2] for a, t in T.items())
            e = max(e, abs(s * np.linalg.norm(R0 + yv) - 1.0))
        errs.append(e)
    assert errs[0] < 3 * 0.1 ** (N + 1)
    # order of convergence: halving rho divides the error by about 2^(N+1)
    assert errs[0] / errs[1] > 0.5 * 2 ** (N + 1)


def _cell(rng, n=40, b=0.3):
    s = rng.uniform(-b, b, size=(n, 3))
    q = rng.choice([-1.0, 1.0], size=n)
    X = np.array([0.2, 0.1, -0.4])
    M = np.einsum("j,ji,jk->ik", q, s, s)
    Qp = 3 * M - np.trace(M) * np.eye(3)
    return X, X + s, q, q.sum(), q @ s, Qp


@pytest.mark.parametrize("P", [3, 4, 5, 6])
def test_local_expansion_converges_to_cell_field(P):
    rng = np.random.default_rng(10 + P)
    X, _, _, q, D, Qp = _cell(rng)
    g = X + np.array([6.0, -3.0, 4.0])
    L = lr.m2l(X, q, D, Qp, g, P)
    d = np.linalg.norm(g - X)
    errs = []
    for rho in (0.08, 0.04):
        ep = ee = 0.0
        for _ in range(20):
            u = rng.normal(size=3)
            yv = rho * d * u / np.linalg.norm(u)
            phi, E = lr.l2p(L, yv)
            phi0, E0 = lr.m2p(X, q, D, Qp, g + yv)
            ep = max(ep, abs(phi - phi0) / abs(phi0))
            ee = max(ee, np.linalg.norm(E - E0) / np.linalg.norm(E0))
        errs.append((ep, ee))
    # field through order P-1 in y: error ~ rho^P; potential ~ rho^(P+1)
    assert errs[0][1] < 60 * 0.08 ** P
    assert errs[0][1] / errs[1][1] > 0.4 * 2 ** P
    assert errs[0][0] / errs[1][0] > 0.4 * 2 ** (P + 1)


def test_local_expansion_vs_point_charges():
    # cell truncation (b/d)^3 and local truncation (|y|/d)^4 against a direct sum
    rng = np.random.default_rng(3)
    X, xs, qs, q, D, Qp = _cell(rng, n=60, b=0.25)
    g = X + np.array([-8.0, 5.0, 3.0])
    L = lr.m2l(X, q, D, Qp, g, 4)
    d = np.linalg.norm(g - X)
    for _ in range(10):
        u = rng.normal(size=3)
        yv = 0.1 * d * u / np.linalg.norm(u)
        x = g + yv
        dr = x - xs
        r = np.linalg.norm(dr, axis=1)
        E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
        phi_dir = (qs / r).sum()
        phi, E = lr.l2p(L, yv)
        scale = np.abs(qs).sum() / d ** 2
        assert np.linalg.norm(E - E_dir) < 20 * scale * ((0.25 * np.sqrt(3) / d) ** 3 + 0.1 ** 4)
        assert abs(phi - phi_dir) < 20 * scale * d * ((0.25 * np.sqrt(3) / d) ** 3 + 0.1 ** 5)


def test_moment_trace_drops_out():
    # adding c I to the second moment (not traceless) leaves L unchanged
    rng = np.random.default_rng(5)
    X, _, _, q, D, Qp = _cell(rng)
    g = X + np.array([3.0, 4.0, -5.0])
    L1 = lr.m2l(X, q, D, Qp, g, 4)
    T = lr.taylor_coeffs(g - X, 6)
    M = lr.cell_moments(q, D, Qp)
    for b in lr.all_indices(4):
        extra = 0.0
        for c in lr.multi_indices(2):
            if max(c) == 2:  # diagonal: add 1.7 to M_ii
                a = (b[0] + c[0], b[1] + c[1], b[2] + c[2])
                extra += lr._c(b, c) * T[a] * 1.7
        assert abs(extra) <= 1e-12 * max(1.0, abs(L1[b]))


@pytest.mark.parametrize("P", [4, 6])
def test_local_coefficients_harmonic(P):
    # each homogeneous part of the local expansion is harmonic:
    # sum_i (b_i + 2)(b_i + 1) L_{b + 2 e_i} = 0 for |b| <= P - 2.  The walk
    # forms only the coefficients with b_z <= 1 and recovers the others from this.
    rng = np.random.default_rng(8)
    X, xs, qs, q, D, Qp = _cell(rng)
    L = lr.m2l(X, q, D, Qp, X + np.array([4.0, -2.0, 3.0]), P, O=lr.third_moment(xs - X, qs))
    for b in lr.all_indices(P - 2):
        terms = []
        for i in range(3):
            c = list(b)
            c[i] += 2
            terms.append((b[i] + 2) * (b[i] + 1) * L[tuple(c)])
        assert abs(sum(terms)) <= 1e-12 * sum(abs(t) for t in terms)


def test_octupole_local_expansion_vs_point_charges():
    # with the third moment the cell's truncation falls to (b/d)^4
    rng = np.random.default_rng(11)
    X, xs, qs, q, D, Qp = _cell(rng, n=60, b=0.25)
    O = lr.third_moment(xs - X, qs)
    for d_scale in (1.0, 2.0):
        g = X + d_scale * np.array([-4.0, 2.5, 1.5])
        d = np.linalg.norm(g - X)
        L2 = lr.m2l(X, q, D, Qp, g, 4)
        L3 = lr.m2l(X, q, D, Qp, g, 4, O=O)
        e2 = e3 = 0.0
        for _ in range(10):
            u = rng.normal(size=3)
            x = g + 0.02 * d * u / np.linalg.norm(u)
            dr = x - xs
            r = np.linalg.norm(dr, axis=1)
            E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
            e2 = max(e2, np.linalg.norm(lr.l2p(L2, x - g)[1] - E_dir))
            e3 = max(e3, np.linalg.norm(lr.l2p(L3, x - g)[1] - E_dir))
        scale = np.abs(qs).sum() / d ** 2
        b = 0.25 * np.sqrt(3)
        assert e3 < 0.5 * e2                      # the third moment helps
        assert e3 < 20 * scale * (b / d) ** 4     # and leaves an order-4 remainder
    # the trace of O drops out: its traceless part gives the same coefficients
    L3t = lr.m2l(X, q, D, Qp, g, 4, O=lr.traceless3(O))
    for k in L3:
        assert abs(L3t[k] - L3[k]) <= 1e-12 * max(1.0, abs(L3[k]))


def test_octupole_m2p_vs_point_charges_and_gradient():
    # the walk's M2P with the traceless third moment: E = -grad phi (finite
    # differences), and against point charges the error falls as (b/d)^4
    # instead of (b/d)^3
    rng = np.random.default_rng(12)
    X, xs, qs, q, D, Qp = _cell(rng, n=80, b=0.25)
    Op = lr.traceless3(lr.third_moment(xs - X, qs))
    x0 = X + np.array([1.3, -0.9, 1.1])
    h = 1e-5
    g = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        g[i] = (lr.m2p(X, q, D, Qp, x0 + e, Op)[0] - lr.m2p(X, q, D, Qp, x0 - e, Op)[0]) / (2 * h)
    E = lr.m2p(X, q, D, Qp, x0, Op)[1]
    assert np.allclose(E, -g, rtol=1e-6, atol=1e-9 * np.abs(E).max())
    errs = []
    for d in (2.0, 4.0):
        u = np.array([0.6, -0.48, 0.64])
        x = X + d * u
        dr = x - xs
        r = np.linalg.norm(dr, axis=1)
        E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
        e2 = np.linalg.norm(lr.m2p(X, q, D, Qp, x)[1] - E_dir)
        e3 = np.linalg.norm(lr.m2p(X, q, D, Qp, x, Op)[1] - E_dir)
        assert e3 < e2
        errs.append(e3)
    # order: doubling d divides the octupole-truncated error by about 2^6 (field ~ d^-2 * (b/d)^4)
    assert errs[0] / errs[1] > 0.5 * 2 ** 6


def test_total_order_truncation():
    # The walk's order-6 expansion keeps only the terms with |b| + |c| <= 7.
    # tmax = P + 3 keeps every term; and for a cell at the walk's opening
    # angle (extent / distance ~ 0.27) seen from a third of the distance,
    # what the truncation drops stays below the order-6 truncation itself.
    rng = np.random.default_rng(12)
    X, xs, qs, q, D, Qp = _cell(rng, n=60, b=0.25)
    O = lr.third_moment(xs - X, qs)
    g = X + np.array([1.1, -0.9, 0.8])
    d = np.linalg.norm(g - X)
    assert 0.25 * np.sqrt(3) / d < 0.35
    full = lr.m2l(X, q, D, Qp, g, 6, O=O)
    same = lr.m2l(X, q, D, Qp, g, 6, O=O, tmax=9)
    for k in full:
        assert same[k] == full[k]
    t7 = lr.m2l(X, q, D, Qp, g, 6, O=O, tmax=7)
    t6 = lr.m2l(X, q, D, Qp, g, 6, O=O, tmax=6)
    worst_cut = worst_t6 = worst_trunc = 0.0
    for _ in range(20):
        u = rng.normal(size=3)
        y = d / 3.0 * u / np.linalg.norm(u)
        dr = g + y - xs
        r = np.linalg.norm(dr, axis=1)
        E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
        E_full = lr.l2p(full, y)[1]
        worst_trunc = max(worst_trunc, np.linalg.norm(E_full - E_dir))
        worst_cut = max(worst_cut, np.linalg.norm(lr.l2p(t7, y)[1] - E_full))
        worst_t6 = max(worst_t6, np.linalg.norm(lr.l2p(t6, y)[1] - E_full))
    assert worst_cut < worst_trunc, (worst_cut, worst_trunc)
    assert worst_cut < worst_t6  # each further order dropped costs more


def test_traceless4_is_traceless_and_idempotent():
    # every pair trace of the projection vanishes, the projection of a
    # traceless tensor is itself, and what it removes is a pure trace part
    rng = np.random.default_rng(21)
    s = rng.normal(size=(30, 3))
    qs = rng.choice([-1.0, 1.0], size=30)
    H = lr.fourth_moment(s, qs)
    Hp = lr.traceless4(H)
    scale = np.abs(H).max()
    for sub in ("iikl->kl", "ikil->kl", "ikli->kl", "kiil->kl", "kili->kl", "klii->kl"):
        assert np.abs(np.einsum(sub, Hp)).max() <= 1e-13 * scale
    assert np.allclose(lr.traceless4(Hp), Hp, rtol=0, atol=1e-13 * scale)
    # symmetric under every index swap
    for perm in ((1, 0, 2, 3), (0, 2, 1, 3), (0, 1, 3, 2), (3, 1, 2, 0)):
        assert np.allclose(Hp.transpose(perm), Hp, rtol=0, atol=1e-14 * scale)
    # r.r.H.r.r of a harmonic quartic: the Laplacian of H'rrrr vanishes at any point
    x0 = rng.normal(size=3)
    h = 1e-3
    f = lambda x: np.einsum("ijkl,i,j,k,l->", Hp, x, x, x, x)
    lap = sum((f(x0 + h * e) - 2 * f(x0) + f(x0 - h * e)) / h ** 2 for e in np.eye(3))
    assert abs(lap) <= 1e-6 * scale * (1 + x0 @ x0)


def test_hexadecapole_local_expansion_vs_point_charges():
    # with the fourth moment the cell's truncation falls to (b/d)^5
    rng = np.random.default_rng(13)
    X, xs, qs, q, D, Qp = _cell(rng, n=60, b=0.25)
    O = lr.third_moment(xs - X, qs)
    H = lr.fourth_moment(xs - X, qs)
    for d_scale in (1.0, 2.0):
        g = X + d_scale * np.array([-4.0, 2.5, 1.5])
        d = np.linalg.norm(g - X)
        L3 = lr.m2l(X, q, D, Qp, g, 5, O=O)
        L4 = lr.m2l(X, q, D, Qp, g, 5, O=O, H=H)
        e3 = e4 = 0.0
        for _ in range(10):
            u = rng.normal(size=3)
            x = g + 0.02 * d * u / np.linalg.norm(u)
            dr = x - xs
            r = np.linalg.norm(dr, axis=1)
            E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
            e3 = max(e3, np.linalg.norm(lr.l2p(L3, x - g)[1] - E_dir))
            e4 = max(e4, np.linalg.norm(lr.l2p(L4, x - g)[1] - E_dir))
        scale = np.abs(qs).sum() / d ** 2
        b = 0.25 * np.sqrt(3)
        assert e4 < 0.5 * e3                      # the fourth moment helps
        assert e4 < 20 * scale * (b / d) ** 5     # and leaves an order-5 remainder
    # the traces of H drop out: its traceless part gives the same coefficients
    L4t = lr.m2l(X, q, D, Qp, g, 5, O=O, H=lr.traceless4(H))
    for k in L4:
        assert abs(L4t[k] - L4[k]) <= 1e-12 * max(1.0, abs(L4[k]))


def test_hexadecapole_m2p_vs_point_charges_and_gradient():
    # the walk's M2P with the traceless third and fourth moments: E = -grad
    # phi (finite differences), and against point charges the error falls as
    # (b/d)^5 instead of (b/d)^4
    rng = np.random.default_rng(14)
    X, xs, qs, q, D, Qp = _cell(rng, n=80, b=0.25)
    Op = lr.traceless3(lr.third_moment(xs - X, qs))
    Hp = lr.traceless4(lr.fourth_moment(xs - X, qs))
    x0 = X + np.array([1.3, -0.9, 1.1])
    h = 1e-5
    g = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        g[i] = (lr.m2p(X, q, D, Qp, x0 + e, Op, Hp)[0] - lr.m2p(X, q, D, Qp, x0 - e, Op, Hp)[0]) / (2 * h)
    E = lr.m2p(X, q, D, Qp, x0, Op, Hp)[1]
    assert np.allclose(E, -g, rtol=1e-6, atol=1e-9 * np.abs(E).max())
    errs = []
    for d in (2.0, 4.0):
        u = np.array([0.6, -0.48, 0.64])
        x = X + d * u
        dr = x - xs
        r = np.linalg.norm(dr, axis=1)
        E_dir = (qs[:, None] * dr / r[:, None] ** 3).sum(0)
        e3 = np.linalg.norm(lr.m2p(X, q, D, Qp, x, Op)[1] - E_dir)
        e4 = np.linalg.norm(lr.m2p(X, q, D, Qp, x, Op, Hp)[1] - E_dir)
        assert e4 < e3
        errs.append(e4)
    # order: doubling d divides the error by about 2^7 (field ~ d^-2 * (b/d)^5)
    assert errs[0] / errs[1] > 0.5 * 2 ** 7


def test_hexadecapole_m2l_matches_m2p_near_the_centre():
    # the local expansion of a cell with all four moments tends to the
    # cell's own M2P field (same moments) as the offset shrinks
    rng = np.random.default_rng(15)
    X, xs, qs, q, D, Qp = _cell(rng)
    Op = lr.traceless3(lr.third_moment(xs - X, qs))
    Hp = lr.traceless4(lr.fourth_moment(xs - X, qs))
    g = X + np.array([5.0, -4.0, 3.0])
    d = np.linalg.norm(g - X)
    L = lr.m2l(X, q, D, Qp, g, 6, O=Op, H=Hp)
    errs = []
    for rho in (0.08, 0.04):
        ee = 0.0
        for _ in range(20):
            u = rng.normal(size=3)
            yv = rho * d * u / np.linalg.norm(u)
            E = lr.l2p(L, yv)[1]
            E0 = lr.m2p(X, q, D, Qp, g + yv, Op, Hp)[1]
            ee = max(ee, np.linalg.norm(E - E0) / np.linalg.norm(E0))
        errs.append(ee)
    assert errs[0] < 60 * 0.08 ** 6
    assert errs[0] / errs[1] > 0.4 * 2 ** 6
