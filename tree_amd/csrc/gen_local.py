"""Writes local_exp.h: the far-field local expansion of the mixed walk as
straight-line FP32 device code (run by hand after a change; the header is
committed).

For a cell with centre X and moments q, D, traceless Q' (and the traceless
third and fourth moments O', H') and an expansion centre g, with R = g - X:

  T_a   Taylor coefficients of 1/|R + y| (a multi-index, |a| <= P + |c|max), by
        n r^2 T_a = -(2n - 1) sum_i R_i T_{a-e_i} - (n - 1) sum_i T_{a-2e_i};
  L_b  += sum_c C(b + c, b) m_c T_{b+c}   (|b| <= P, |c| <= 2, 3 or 4), with
        m_0 = q, m_{e_i} = -D_i, m_c = Q'_c / 3 for |c| = 2 and
        m_c = -O'_c for |c| = 3 (O' the traceless third moment) and
        m_c = H'_c for |c| = 4 (H' the traceless fourth moment);
  phi(y) = sum_b L_b y^b,  E = -grad phi.

The same mathematics in FP64, with loops instead of unrolled code, is the
checker (oracle/local_ref.py).
"""
import os
from math import comb

HERE = os.path.dirname(os.path.abspath(__file__))


def multi_indices(n):
    return [(a, b, n - a - b) for a in range(n, -1, -1) for b in range(n - a, -1, -1)]


def all_indices(nmax):
    out = []
    for n in range(nmax + 1):
        out += multi_indices(n)
    return out


def name(a):
    return "t%d%d%d" % a


def fl(v):
    s = repr(float(v))
    return s + "f"


def harmonic_indices(P):
    """The independent coefficients of a harmonic expansion: b_z <= 1."""
    return [b for b in all_indices(P) if b[2] <= 1]


def emit_m2l(P, harmonic=False, octupole=False, tmax=None, hexadecapole=False):
    """Keeps the terms m_c T_{b+c} with |b| <= P and total order |b| + |c| <=
    tmax (Taylor coefficients to order tmax); tmax None: every moment with
    every |b| <= P (to order P + 3).  hexadecapole adds the traceless fourth
    moment (m_c = H'_c for |c| = 4)."""
    L = harmonic_indices(P) if harmonic else all_indices(P)
    lidx = {b: i for i, b in enumerate(L)}
    sfx = ("h" if harmonic else "") + ("x" if hexadecapole else "o" if octupole else "") + \
        ("t%d" % tmax if tmax else "")
    cmax = 4 if hexadecapole else 3 if octupole else 2
    tmax = tmax or P + cmax
    # moments m_c: name and expression
    mom = {(0, 0, 0): "q", (1, 0, 0): "mdx", (0, 1, 0): "mdy", (0, 0, 1): "mdz",
           (2, 0, 0): "mxx", (0, 2, 0): "myy", (0, 0, 2): "mzz",
           (1, 1, 0): "mxy", (0, 1, 1): "myz", (1, 0, 1): "mzx"}
    out = []
    out.append("// adds one cell's local coefficients about the expansion centre to L[%d]" % len(L))
    out.append("// (order P = %d; R = centre - cell centre; Q' traceless, Q'zz = -Q'xx - Q'yy)" % P)
    out.append("__device__ inline void m2l_acc_p%d%s(float rx, float ry, float rz, float q, float dx, float dy, "
               "float dz," % (P, sfx))
    if octupole:
        out.append("                                  float qxx, float qyy, float qxy, float qyz, float qzx,")
        if hexadecapole:
            out.append("                                  float oxxx, float oxxy, float oxxz, float oxyy, float oxyz, "
                       "float oyyy, float oyyz,")
            out.append("                                  float hxxxx, float hxxxy, float hxxxz, float hxxyy, "
                       "float hxxyz, float hxyyy, float hxyyz,")
            out.append("                                  float hyyyy, float hyyyz, float (&L)[%d])" % len(L))
        else:
            out.append("                                  float oxxx, float oxxy, float oxxz, float oxyy, float oxyz, "
                       "float oyyy, float oyyz, float (&L)[%d])" % len(L))
    else:
        out.append("                                  float qxx, float qyy, float qxy, float qyz, float qzx, "
                   "float (&L)[%d])" % len(L))
    out.append("{")
    out.append("    const float r2 = rx * rx + ry * ry + rz * rz;")
    out.append("    const float mdx = -dx, mdy = -dy, mdz = -dz;")
    out.append("    const float mxx = qxx * (1.0f / 3.0f), myy = qyy * (1.0f / 3.0f), "
               "mzz = -(qxx + qyy) * (1.0f / 3.0f);")
    out.append("    const float mxy = qxy * (1.0f / 3.0f), myz = qyz * (1.0f / 3.0f), mzx = qzx * (1.0f / 3.0f);")
    if octupole:
        # m_c = -O'_c; the traceless third moment's z-heavy components follow from the others
        out.append("    const float m300 = -oxxx, m210 = -oxxy, m201 = -oxxz, m120 = -oxyy, m111 = -oxyz;")
        out.append("    const float m030 = -oyyy, m021 = -oyyz, m102 = oxxx + oxyy, m012 = oxxy + oyyy, "
                   "m003 = oxxz + oyyz;")
        for c in multi_indices(3):
            mom[c] = "m%d%d%d" % c
    if hexadecapole:
        # m_c = H'_c; the traceless fourth moment's components with two or more z from the others
        out.append("    const float m400 = hxxxx, m310 = hxxxy, m301 = hxxxz, m220 = hxxyy, m211 = hxxyz;")
        out.append("    const float m130 = hxyyy, m121 = hxyyz, m040 = hyyyy, m031 = hyyyz;")
        out.append("    const float m202 = -(hxxxx + hxxyy), m112 = -(hxxxy + hxyyy), m022 = -(hxxyy + hyyyy);")
        out.append("    const float m103 = -(hxxxz + hxyyz), m013 = -(hxxyz + hyyyz), "
                   "m004 = hxxxx + 2.0f * hxxyy + hyyyy;")
        for c in multi_indices(4):
            mom[c] = "m%d%d%d" % c
    out.append("    const float t000 = __builtin_amdgcn_rsqf(r2);")
    out.append("    const float ir2 = t000 * t000;")
    R = ["rx", "ry", "rz"]

    def contrib(a):
        # L_b += C(b + c, b) m_c T_a for every c <= a with |c| <= 2 and |a - c| <= P
        lines = []
        for c, m in mom.items():
            b = tuple(a[i] - c[i] for i in range(3))
            if min(b) < 0 or sum(b) > P or b not in lidx:
                continue
            k = comb(b[0] + c[0], b[0]) * comb(b[1] + c[1], b[1]) * comb(b[2] + c[2], b[2])
            if k == 1:
                lines.append("    L[%d] = fmaf(%s, %s, L[%d]);" % (lidx[b], m, name(a), lidx[b]))
            else:
                lines.append("    L[%d] = fmaf(%s * %s, %s, L[%d]);" % (lidx[b], fl(k), m, name(a), lidx[b]))
        return lines

    out += contrib((0, 0, 0))
    for n in range(1, tmax + 1):
        c1 = "%s * ir2" % fl((2 * n - 1) / n)
        c2 = "%s * ir2" % fl((n - 1) / n)
        out.append("    const float c1_%d = %s, c2_%d = %s;" % (n, c1, n, c2))
        for a in multi_indices(n):
            if harmonic and a[2] > 1 + cmax:
                continue  # only b_z <= 1 coefficients are formed: they need T with a_z <= 1 + |c|max
            s1 = []
            s2 = []
            for i in range(3):
                if a[i] >= 1:
                    b = list(a)
                    b[i] -= 1
                    s1.append("%s * %s" % (R[i], name(tuple(b))))
                if a[i] >= 2:
                    b = list(a)
                    b[i] -= 2
                    s2.append(name(tuple(b)))
            e1 = " + ".join(s1)
            if s2:
                out.append("    const float %s = -fmaf(c1_%d, %s, c2_%d * (%s));" % (name(a), n, e1, n, " + ".join(s2)))
            else:
                out.append("    const float %s = -c1_%d * (%s);" % (name(a), n, e1))
            out += contrib(a)
    out.append("}")
    return out


def emit_l2p(P, harmonic=False):
    L = all_indices(P)
    out = []
    if harmonic:
        H = harmonic_indices(P)
        hidx = {b: i for i, b in enumerate(H)}
        out.append("// potential and field at offset y from the expansion centre (order P = %d, from the" % P)
        out.append("// independent coefficients b_z <= 1; each homogeneous part of a harmonic")
        out.append("// expansion is harmonic, which fixes the others)")
        out.append("__device__ inline void l2p_p%dh(const float (&Lh)[%d], float x, float y, float z, float &phi, "
                   "float &ex, float &ey, float &ez)" % (P, len(H)))
        out.append("{")
        out.append("    float L[%d];" % len(L))
        full = {b: i for i, b in enumerate(L)}
        for b in L:
            if b[2] <= 1:
                out.append("    L[%d] = Lh[%d];" % (full[b], hidx[b]))
        # b_z >= 2 in increasing b_z: (b_z)(b_z - 1) L_b = -(b_x+2)(b_x+1) L_{b-2e_z+2e_x} - (b_y+2)(b_y+1) L_{b-2e_z+2e_y}
        for bz in range(2, P + 1):
            for b in L:
                if b[2] != bz:
                    continue
                c = (b[0], b[1], b[2] - 2)
                kx = (c[0] + 2) * (c[0] + 1)
                ky = (c[1] + 2) * (c[1] + 1)
                kz = (c[2] + 2) * (c[2] + 1)
                ix = full[(c[0] + 2, c[1], c[2])]
                iy = full[(c[0], c[1] + 2, c[2])]
                out.append("    L[%d] = -(%s * L[%d] + %s * L[%d]);" % (full[b], fl(kx / kz), ix, fl(ky / kz), iy))
    else:
        out.append("// potential and field at offset y from the expansion centre (order P = %d)" % P)
        out.append("__device__ inline void l2p_p%d(const float (&L)[%d], float x, float y, float z, float &phi, "
                   "float &ex, float &ey, float &ez)" % (P, len(L)))
        out.append("{")
    # powers
    for v in "xyz":
        out.append("    const float %s0 = 1.0f, %s1 = %s;" % (v, v, v))
        for k in range(2, P + 1):
            out.append("    const float %s%d = %s%d * %s;" % (v, k, v, k - 1, v))
    ph, gx, gy, gz = [], [], [], []
    for i, b in enumerate(L):
        mono = "x%d * y%d * z%d" % b
        ph.append("L[%d] * (%s)" % (i, mono))
        if b[0]:
            gx.append("%s * L[%d] * (x%d * y%d * z%d)" % (fl(b[0]), i, b[0] - 1, b[1], b[2]))
        if b[1]:
            gy.append("%s * L[%d] * (x%d * y%d * z%d)" % (fl(b[1]), i, b[0], b[1] - 1, b[2]))
        if b[2]:
            gz.append("%s * L[%d] * (x%d * y%d * z%d)" % (fl(b[2]), i, b[0], b[1], b[2] - 1))
    out.append("    phi = " + "\n        + ".join(ph) + ";")
    out.append("    ex = -(" + "\n        + ".join(gx) + ");")
    out.append("    ey = -(" + "\n        + ".join(gy) + ");")
    out.append("    ez = -(" + "\n        + ".join(gz) + ");")
    out.append("}")
    return out


def main():
    lines = ["// local_exp.h -- written by gen_local.py; do not edit by hand.",
             "//",
             "// Far-field local expansion of the mixed walk: Taylor coefficients of",
             "// 1/|R + y| by their three-term recurrence, contracted with a cell's",
             "// net charge, dipole, traceless quadrupole and (orders 4 to 6, the",
             "// harmonic forms) traceless third moment (order 6, the walk's: and",
             "// traceless fourth moment) into the local coefficients",
             "// L_b of the potential about a target group's centre (|b| <= P,",
             "// multi-indices in the order of all_indices; the harmonic forms keep",
             "// b_z <= 1; the order-6 form keeps the terms of total order |b| + |c|",
             "// <= 7 only), and the evaluation of potential and field at a target.",
             "#pragma once",
             "#include <hip/hip_runtime.h>",
             ""]
    lines.append("constexpr int TB_LOCAL_N3 = %d;   // order 3, all coefficients" % len(all_indices(3)))
    lines.append("constexpr int TB_LOCAL_N4H = %d;  // order 4, independent coefficients only" % len(harmonic_indices(4)))
    lines.append("constexpr int TB_LOCAL_N5H = %d;  // order 5, independent coefficients only" % len(harmonic_indices(5)))
    lines.append("constexpr int TB_LOCAL_N6H = %d;  // order 6 (terms to total order 7), independent coefficients only"
                 % len(harmonic_indices(6)))
    lines.append("")
    lines += emit_m2l(3)
    lines.append("")
    lines += emit_l2p(3)
    lines.append("")
    for P in (4, 5):
        lines += emit_m2l(P, harmonic=True, octupole=True)
        lines.append("")
        lines += emit_l2p(P, harmonic=True)
        lines.append("")
    # the walk's: order 6, terms to total order 7, cells to the fourth moment
    lines += emit_m2l(6, harmonic=True, octupole=True, tmax=7, hexadecapole=True)
    lines.append("")
    lines += emit_l2p(6, harmonic=True)
    lines.append("")
    with open(os.path.join(HERE, "local_exp.h"), "w") as f:
        f.write("\n".join(lines))


if __name__ == "__main__":
    main()
