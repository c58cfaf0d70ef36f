"""FP64 reference of the far-field local expansion used by the mixed walk.

TEST INFRASTRUCTURE ONLY: the checker for the walk's far-cell path
(tree_amd/csrc/local_exp.h).  The reference code has no local expansion;
this is the project's own method, so the oracle is pinned by mathematics
(tests/test_local_oracle.py): closed-form derivatives of 1/r, finite
differences, harmonicity, the Taylor series of 1/|R + y|, direct sums over
point charges, and the walk's own monopole + dipole + traceless-quadrupole
cell field (m2p below, the expression of the walk's M2P loop) as the limit
for targets close to the expansion centre.

Notation.  For a multi-index a = (a1, a2, a3), |a| = a1 + a2 + a3 and
a! = a1! a2! a3!.  T_a(R) = (1/a!) d^a (1/|R|) / dR^a, so that
1/|R + y| = sum_a T_a(R) y^a.  For |a| = n >= 1 (all terms with a negative
index dropped):

    n |R|^2 T_a = -(2n - 1) sum_i R_i T_{a - e_i} - (n - 1) sum_i T_{a - 2 e_i}

A cell with centre X and moments q (net charge), D (dipole), traceless
quadrupole Q' = 3 M - tr(M) I (M the second moment about X) and, optionally,
the third moment O (O_ijk = sum q s_i s_j s_k, s the offset from X) and the
fourth moment H (H_ijkl = sum q s_i s_j s_k s_l) produces at a point g + y,
with R = g - X:

    phi(y) = sum_b L_b y^b,
    L_b = sum_c C(b + c, b) (-1)^|c| T_{b+c}(R) M_c,   |c| <= 2 (3, or 4),

with M_0 = q, M_{e_i} = D_i, M_c = Q'_c / 3 for |c| = 2, M_c = O_c for
|c| = 3 and M_c = H_c for |c| = 4.  Trace parts of the moment tensors drop
out because 1/r is harmonic, so Q'/3 may stand for M and the traceless
parts of O and H for O and H.  The
field is E = -grad phi.  Keeping |b| <= P gives the field through order
P - 1 in y.  With tmax, only the terms of total order |b| + |c| <= tmax are
kept (the walk's order-6 expansion keeps tmax = 7: the dropped terms, a
higher moment times a high power of y, are below the order-P truncation for
cells accepted by the walk).
"""
from math import comb, factorial

import numpy as np


def multi_indices(n):
    """All (a1, a2, a3) with a1 + a2 + a3 = n, a1 descending, then a2."""
    return [(a, b, n - a - b) for a in range(n, -1, -1) for b in range(n - a, -1, -1)]


def all_indices(nmax):
    out = []
    for n in range(nmax + 1):
        out += multi_indices(n)
    return out


def taylor_coeffs(R, nmax):
    """T_a(R) for |a| <= nmax by the recurrence above (dict keyed by a)."""
    R = np.asarray(R, dtype=np.float64)
    r2 = float(R @ R)
    T = {(0, 0, 0): 1.0 / np.sqrt(r2)}
    for n in range(1, nmax + 1):
        for a in multi_indices(n):
            s1 = 0.0
            s2 = 0.0
            for i in range(3):
                if a[i] >= 1:
                    b = list(a)
                    b[i] -= 1
                    s1 += R[i] * T[tuple(b)]
                if a[i] >= 2:
                    b = list(a)
                    b[i] -= 2
                    s2 += T[tuple(b)]
            T[a] = -((2 * n - 1) * s1 + (n - 1) * s2) / (n * r2)
    return T


def _c(b, c):
    return comb(b[0] + c[0], b[0]) * comb(b[1] + c[1], b[1]) * comb(b[2] + c[2], b[2])


def cell_moments(q, D, Qp, O=None, H=None):
    """M_c for |c| <= 2 (3 with O, 4 with H) from q, D[3], the traceless Q'
    (3x3, symmetric), the third moment O (3x3x3, symmetric) and the fourth
    moment H (3x3x3x3, symmetric)."""
    Qp = np.asarray(Qp, dtype=np.float64)
    M = {(0, 0, 0): float(q)}
    for i in range(3):
        e = [0, 0, 0]
        e[i] = 1
        M[tuple(e)] = float(D[i])
    for c in multi_indices(2):
        idx = [i for i in range(3) for _ in range(c[i])]
        M[c] = Qp[idx[0], idx[1]] / 3.0
    if O is not None:
        O = np.asarray(O, dtype=np.float64)
        for c in multi_indices(3):
            idx = [i for i in range(3) for _ in range(c[i])]
            M[c] = O[idx[0], idx[1], idx[2]]
    if H is not None:
        H = np.asarray(H, dtype=np.float64)
        for c in multi_indices(4):
            idx = [i for i in range(3) for _ in range(c[i])]
            M[c] = H[idx[0], idx[1], idx[2], idx[3]]
    return M


def third_moment(s, q):
    """O_ijk = sum_j q_j s_ji s_jj s_jk of point charges at offsets s."""
    return np.einsum("j,ji,jk,jl->ikl", q, s, s, s)


def fourth_moment(s, q):
    """H_ijkl = sum_j q_j s_ji s_jj s_jk s_jl of point charges at offsets s."""
    return np.einsum("j,ji,jk,jl,jm->iklm", q, s, s, s, s)


def traceless4(H):
    """The traceless part of a symmetric rank-4 tensor: H - (d_ij T_kl + 5
    more pairings) / 7 + (d_ij d_kl + d_ik d_jl + d_il d_jk) tr T / 35, with
    T_kl = H_iikl."""
    T = np.einsum("iikl->kl", H)
    t0 = np.trace(T)
    d = np.eye(3)
    pair = (np.einsum("ij,kl->ijkl", d, T) + np.einsum("ik,jl->ijkl", d, T) + np.einsum("il,jk->ijkl", d, T) +
            np.einsum("jk,il->ijkl", d, T) + np.einsum("jl,ik->ijkl", d, T) + np.einsum("kl,ij->ijkl", d, T))
    dd = np.einsum("ij,kl->ijkl", d, d) + np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)
    return H - pair / 7.0 + dd * t0 / 35.0


def traceless3(O):
    """The traceless part of a symmetric rank-3 tensor."""
    t = np.einsum("iik->k", O)
    d = np.eye(3)
    return O - (np.einsum("ij,k->ijk", d, t) + np.einsum("ik,j->ijk", d, t) + np.einsum("jk,i->ijk", d, t)) / 5.0


def m2l(X, q, D, Qp, g, P, O=None, tmax=None, H=None):
    """Local coefficients L_b (|b| <= P) about g of a cell at X with moments
    q, D, Q' (and the third moment O, the fourth moment H); with tmax, only
    the terms with |b| + |c| <= tmax."""
    R = np.asarray(g, dtype=np.float64) - np.asarray(X, dtype=np.float64)
    T = taylor_coeffs(R, P + (4 if H is not None else 3 if O is not None else 2))
    M = cell_moments(q, D, Qp, O, H)
    L = {}
    for b in all_indices(P):
        s = 0.0
        for c, mc in M.items():
            if tmax is not None and sum(b) + sum(c) > tmax:
                continue
            a = (b[0] + c[0], b[1] + c[1], b[2] + c[2])
            s += _c(b, c) * (-1.0) ** sum(c) * T[a] * mc
        L[b] = s
    return L


def l2p(L, y):
    """Potential and field at offset y from the expansion centre."""
    y = np.asarray(y, dtype=np.float64)
    phi = 0.0
    E = np.zeros(3)
    for b, lb in L.items():
        phi += lb * y[0] ** b[0] * y[1] ** b[1] * y[2] ** b[2]
        for i in range(3):
            if b[i] == 0:
                continue
            e = list(b)
            e[i] -= 1
            E[i] -= lb * b[i] * y[0] ** e[0] * y[1] ** e[1] * y[2] ** e[2]
    return phi, E


def m2p(X, q, D, Qp, x, Op=None, Hp=None):
    """The walk's cell expansion evaluated directly at x (FP64):
    phi = q/r + D.r/r^3 + r.Q'.r/(2 r^5) [+ 5/2 O'rrr / r^7] [+ 35/8 H'rrrr / r^9],
    E = -grad phi, r = x - X, O' and H' the traceless third and fourth
    moments (the walk's M2P keeps both)."""
    r = np.asarray(x, dtype=np.float64) - np.asarray(X, dtype=np.float64)
    Qp = np.asarray(Qp, dtype=np.float64)
    r2 = float(r @ r)
    ri = 1.0 / np.sqrt(r2)
    ri3, ri5, ri7, ri9 = ri ** 3, ri ** 5, ri ** 7, ri ** 9
    dp = float(np.dot(D, r))
    Qr = Qp @ r
    rQr = float(r @ Qr)
    phi = q * ri + dp * ri3 + 0.5 * rQr * ri5
    E = r * (q * ri3 + 3.0 * dp * ri5 + 2.5 * rQr * ri7) - np.asarray(D) * ri3 - Qr * ri5
    if Op is not None:
        v = np.einsum("ijk,j,k->i", np.asarray(Op, dtype=np.float64), r, r)
        s3 = float(r @ v)
        phi += 2.5 * s3 * ri7
        E += -7.5 * v * ri7 + 17.5 * s3 * r * ri9
    if Hp is not None:
        w = np.einsum("ijkl,j,k,l->i", np.asarray(Hp, dtype=np.float64), r, r, r)
        s4 = float(r @ w)
        phi += 4.375 * s4 * ri9
        E += -17.5 * w * ri9 + 39.375 * s4 * r * ri9 / r2
    return phi, E


def inv_r_derivative_fd(R, a, h=1e-3):
    """(1/a!) d^a (1/|R|) by nested central differences (independent check)."""
    R = np.asarray(R, dtype=np.float64)

    def f(x):
        return 1.0 / np.sqrt(x @ x)

    def d(fun, i):
        def g(x):
            e = np.zeros(3)
            e[i] = h
            return (fun(x + e) - fun(x - e)) / (2 * h)
        return g

    fun = f
    for i in range(3):
        for _ in range(a[i]):
            fun = d(fun, i)
    return fun(R) / (factorial(a[0]) * factorial(a[1]) * factorial(a[2]))
