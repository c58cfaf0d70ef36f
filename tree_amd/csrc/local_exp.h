// local_exp.h -- written by gen_local.py; do not edit by hand.
//
// Far-field local expansion of the mixed walk: Taylor coefficients of
// 1/|R + y| by their three-term recurrence, contracted with a cell's
// net charge, dipole, traceless quadrupole and (orders 4 to 6, the
// harmonic forms) traceless third moment (order 6, the walk's: and
// traceless fourth moment) into the local coefficients
// L_b of the potential about a target group's centre (|b| <= P,
// multi-indices in the order of all_indices; the harmonic forms keep
// b_z <= 1; the order-6 form keeps the terms of total order |b| + |c|
// <= 7 only), and the evaluation of potential and field at a target.
#pragma once
#include <hip/hip_runtime.h>

constexpr int TB_LOCAL_N3 = 20;   // order 3, all coefficients
constexpr int TB_LOCAL_N4H = 25;  // order 4, independent coefficients only
constexpr int TB_LOCAL_N5H = 36;  // order 5, independent coefficients only
constexpr int TB_LOCAL_N6H = 49;  // order 6 (terms to total order 7), independent coefficients only

// adds one cell's local coefficients about the expansion centre to L[20]
// (order P = 3; R = centre - cell centre; Q' traceless, Q'zz = -Q'xx - Q'yy)
__device__ inline void m2l_acc_p3(float rx, float ry, float rz, float q, float dx, float dy, float dz,
                                  float qxx, float qyy, float qxy, float qyz, float qzx, float (&L)[20])
{
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float mdx = -dx, mdy = -dy, mdz = -dz;
    const float mxx = qxx * (1.0f / 3.0f), myy = qyy * (1.0f / 3.0f), mzz = -(qxx + qyy) * (1.0f / 3.0f);
    const float mxy = qxy * (1.0f / 3.0f), myz = qyz * (1.0f / 3.0f), mzx = qzx * (1.0f / 3.0f);
    const float t000 = __builtin_amdgcn_rsqf(r2);
    const float ir2 = t000 * t000;
    L[0] = fmaf(q, t000, L[0]);
    const float c1_1 = 1.0f * ir2, c2_1 = 0.0f * ir2;
    const float t100 = -c1_1 * (rx * t000);
    L[1] = fmaf(q, t100, L[1]);
    L[0] = fmaf(mdx, t100, L[0]);
    const float t010 = -c1_1 * (ry * t000);
    L[2] = fmaf(q, t010, L[2]);
    L[0] = fmaf(mdy, t010, L[0]);
    const float t001 = -c1_1 * (rz * t000);
    L[3] = fmaf(q, t001, L[3]);
    L[0] = fmaf(mdz, t001, L[0]);
    const float c1_2 = 1.5f * ir2, c2_2 = 0.5f * ir2;
    const float t200 = -fmaf(c1_2, rx * t100, c2_2 * (t000));
    L[4] = fmaf(q, t200, L[4]);
    L[1] = fmaf(2.0f * mdx, t200, L[1]);
    L[0] = fmaf(mxx, t200, L[0]);
    const float t110 = -c1_2 * (rx * t010 + ry * t100);
    L[5] = fmaf(q, t110, L[5]);
    L[2] = fmaf(mdx, t110, L[2]);
    L[1] = fmaf(mdy, t110, L[1]);
    L[0] = fmaf(mxy, t110, L[0]);
    const float t101 = -c1_2 * (rx * t001 + rz * t100);
    L[6] = fmaf(q, t101, L[6]);
    L[3] = fmaf(mdx, t101, L[3]);
    L[1] = fmaf(mdz, t101, L[1]);
    L[0] = fmaf(mzx, t101, L[0]);
    const float t020 = -fmaf(c1_2, ry * t010, c2_2 * (t000));
    L[7] = fmaf(q, t020, L[7]);
    L[2] = fmaf(2.0f * mdy, t020, L[2]);
    L[0] = fmaf(myy, t020, L[0]);
    const float t011 = -c1_2 * (ry * t001 + rz * t010);
    L[8] = fmaf(q, t011, L[8]);
    L[3] = fmaf(mdy, t011, L[3]);
    L[2] = fmaf(mdz, t011, L[2]);
    L[0] = fmaf(myz, t011, L[0]);
    const float t002 = -fmaf(c1_2, rz * t001, c2_2 * (t000));
    L[9] = fmaf(q, t002, L[9]);
    L[3] = fmaf(2.0f * mdz, t002, L[3]);
    L[0] = fmaf(mzz, t002, L[0]);
    const float c1_3 = 1.6666666666666667f * ir2, c2_3 = 0.6666666666666666f * ir2;
    const float t300 = -fmaf(c1_3, rx * t200, c2_3 * (t100));
    L[10] = fmaf(q, t300, L[10]);
    L[4] = fmaf(3.0f * mdx, t300, L[4]);
    L[1] = fmaf(3.0f * mxx, t300, L[1]);
    const float t210 = -fmaf(c1_3, rx * t110 + ry * t200, c2_3 * (t010));
    L[11] = fmaf(q, t210, L[11]);
    L[5] = fmaf(2.0f * mdx, t210, L[5]);
    L[4] = fmaf(mdy, t210, L[4]);
    L[2] = fmaf(mxx, t210, L[2]);
    L[1] = fmaf(2.0f * mxy, t210, L[1]);
    const float t201 = -fmaf(c1_3, rx * t101 + rz * t200, c2_3 * (t001));
    L[12] = fmaf(q, t201, L[12]);
    L[6] = fmaf(2.0f * mdx, t201, L[6]);
    L[4] = fmaf(mdz, t201, L[4]);
    L[3] = fmaf(mxx, t201, L[3]);
    L[1] = fmaf(2.0f * mzx, t201, L[1]);
    const float t120 = -fmaf(c1_3, rx * t020 + ry * t110, c2_3 * (t100));
    L[13] = fmaf(q, t120, L[13]);
    L[7] = fmaf(mdx, t120, L[7]);
    L[5] = fmaf(2.0f * mdy, t120, L[5]);
    L[1] = fmaf(myy, t120, L[1]);
    L[2] = fmaf(2.0f * mxy, t120, L[2]);
    const float t111 = -c1_3 * (rx * t011 + ry * t101 + rz * t110);
    L[14] = fmaf(q, t111, L[14]);
    L[8] = fmaf(mdx, t111, L[8]);
    L[6] = fmaf(mdy, t111, L[6]);
    L[5] = fmaf(mdz, t111, L[5]);
    L[3] = fmaf(mxy, t111, L[3]);
    L[1] = fmaf(myz, t111, L[1]);
    L[2] = fmaf(mzx, t111, L[2]);
    const float t102 = -fmaf(c1_3, rx * t002 + rz * t101, c2_3 * (t100));
    L[15] = fmaf(q, t102, L[15]);
    L[9] = fmaf(mdx, t102, L[9]);
    L[6] = fmaf(2.0f * mdz, t102, L[6]);
    L[1] = fmaf(mzz, t102, L[1]);
    L[3] = fmaf(2.0f * mzx, t102, L[3]);
    const float t030 = -fmaf(c1_3, ry * t020, c2_3 * (t010));
    L[16] = fmaf(q, t030, L[16]);
    L[7] = fmaf(3.0f * mdy, t030, L[7]);
    L[2] = fmaf(3.0f * myy, t030, L[2]);
    const float t021 = -fmaf(c1_3, ry * t011 + rz * t020, c2_3 * (t001));
    L[17] = fmaf(q, t021, L[17]);
    L[8] = fmaf(2.0f * mdy, t021, L[8]);
    L[7] = fmaf(mdz, t021, L[7]);
    L[3] = fmaf(myy, t021, L[3]);
    L[2] = fmaf(2.0f * myz, t021, L[2]);
    const float t012 = -fmaf(c1_3, ry * t002 + rz * t011, c2_3 * (t010));
    L[18] = fmaf(q, t012, L[18]);
    L[9] = fmaf(mdy, t012, L[9]);
    L[8] = fmaf(2.0f * mdz, t012, L[8]);
    L[2] = fmaf(mzz, t012, L[2]);
    L[3] = fmaf(2.0f * myz, t012, L[3]);
    const float t003 = -fmaf(c1_3, rz * t002, c2_3 * (t001));
    L[19] = fmaf(q, t003, L[19]);
    L[9] = fmaf(3.0f * mdz, t003, L[9]);
    L[3] = fmaf(3.0f * mzz, t003, L[3]);
    const float c1_4 = 1.75f * ir2, c2_4 = 0.75f * ir2;
    const float t400 = -fmaf(c1_4, rx * t300, c2_4 * (t200));
    L[10] = fmaf(4.0f * mdx, t400, L[10]);
    L[4] = fmaf(6.0f * mxx, t400, L[4]);
    const float t310 = -fmaf(c1_4, rx * t210 + ry * t300, c2_4 * (t110));
    L[11] = fmaf(3.0f * mdx, t310, L[11]);
    L[10] = fmaf(mdy, t310, L[10]);
    L[5] = fmaf(3.0f * mxx, t310, L[5]);
    L[4] = fmaf(3.0f * mxy, t310, L[4]);
    const float t301 = -fmaf(c1_4, rx * t201 + rz * t300, c2_4 * (t101));
    L[12] = fmaf(3.0f * mdx, t301, L[12]);
    L[10] = fmaf(mdz, t301, L[10]);
    L[6] = fmaf(3.0f * mxx, t301, L[6]);
    L[4] = fmaf(3.0f * mzx, t301, L[4]);
    const float t220 = -fmaf(c1_4, rx * t120 + ry * t210, c2_4 * (t020 + t200));
    L[13] = fmaf(2.0f * mdx, t220, L[13]);
    L[11] = fmaf(2.0f * mdy, t220, L[11]);
    L[7] = fmaf(mxx, t220, L[7]);
    L[4] = fmaf(myy, t220, L[4]);
    L[5] = fmaf(4.0f * mxy, t220, L[5]);
    const float t211 = -fmaf(c1_4, rx * t111 + ry * t201 + rz * t210, c2_4 * (t011));
    L[14] = fmaf(2.0f * mdx, t211, L[14]);
    L[12] = fmaf(mdy, t211, L[12]);
    L[11] = fmaf(mdz, t211, L[11]);
    L[8] = fmaf(mxx, t211, L[8]);
    L[6] = fmaf(2.0f * mxy, t211, L[6]);
    L[4] = fmaf(myz, t211, L[4]);
    L[5] = fmaf(2.0f * mzx, t211, L[5]);
    const float t202 = -fmaf(c1_4, rx * t102 + rz * t201, c2_4 * (t002 + t200));
    L[15] = fmaf(2.0f * mdx, t202, L[15]);
    L[12] = fmaf(2.0f * mdz, t202, L[12]);
    L[9] = fmaf(mxx, t202, L[9]);
    L[4] = fmaf(mzz, t202, L[4]);
    L[6] = fmaf(4.0f * mzx, t202, L[6]);
    const float t130 = -fmaf(c1_4, rx * t030 + ry * t120, c2_4 * (t110));
    L[16] = fmaf(mdx, t130, L[16]);
    L[13] = fmaf(3.0f * mdy, t130, L[13]);
    L[5] = fmaf(3.0f * myy, t130, L[5]);
    L[7] = fmaf(3.0f * mxy, t130, L[7]);
    const float t121 = -fmaf(c1_4, rx * t021 + ry * t111 + rz * t120, c2_4 * (t101));
    L[17] = fmaf(mdx, t121, L[17]);
    L[14] = fmaf(2.0f * mdy, t121, L[14]);
    L[13] = fmaf(mdz, t121, L[13]);
    L[6] = fmaf(myy, t121, L[6]);
    L[8] = fmaf(2.0f * mxy, t121, L[8]);
    L[5] = fmaf(2.0f * myz, t121, L[5]);
    L[7] = fmaf(mzx, t121, L[7]);
    const float t112 = -fmaf(c1_4, rx * t012 + ry * t102 + rz * t111, c2_4 * (t110));
    L[18] = fmaf(mdx, t112, L[18]);
    L[15] = fmaf(mdy, t112, L[15]);
    L[14] = fmaf(2.0f * mdz, t112, L[14]);
    L[5] = fmaf(mzz, t112, L[5]);
    L[9] = fmaf(mxy, t112, L[9]);
    L[6] = fmaf(2.0f * myz, t112, L[6]);
    L[8] = fmaf(2.0f * mzx, t112, L[8]);
    const float t103 = -fmaf(c1_4, rx * t003 + rz * t102, c2_4 * (t101));
    L[19] = fmaf(mdx, t103, L[19]);
    L[15] = fmaf(3.0f * mdz, t103, L[15]);
    L[6] = fmaf(3.0f * mzz, t103, L[6]);
    L[9] = fmaf(3.0f * mzx, t103, L[9]);
    const float t040 = -fmaf(c1_4, ry * t030, c2_4 * (t020));
    L[16] = fmaf(4.0f * mdy, t040, L[16]);
    L[7] = fmaf(6.0f * myy, t040, L[7]);
    const float t031 = -fmaf(c1_4, ry * t021 + rz * t030, c2_4 * (t011));
    L[17] = fmaf(3.0f * mdy, t031, L[17]);
    L[16] = fmaf(mdz, t031, L[16]);
    L[8] = fmaf(3.0f * myy, t031, L[8]);
    L[7] = fmaf(3.0f * myz, t031, L[7]);
    const float t022 = -fmaf(c1_4, ry * t012 + rz * t021, c2_4 * (t002 + t020));
    L[18] = fmaf(2.0f * mdy, t022, L[18]);
    L[17] = fmaf(2.0f * mdz, t022, L[17]);
    L[9] = fmaf(myy, t022, L[9]);
    L[7] = fmaf(mzz, t022, L[7]);
    L[8] = fmaf(4.0f * myz, t022, L[8]);
    const float t013 = -fmaf(c1_4, ry * t003 + rz * t012, c2_4 * (t011));
    L[19] = fmaf(mdy, t013, L[19]);
    L[18] = fmaf(3.0f * mdz, t013, L[18]);
    L[8] = fmaf(3.0f * mzz, t013, L[8]);
    L[9] = fmaf(3.0f * myz, t013, L[9]);
    const float t004 = -fmaf(c1_4, rz * t003, c2_4 * (t002));
    L[19] = fmaf(4.0f * mdz, t004, L[19]);
    L[9] = fmaf(6.0f * mzz, t004, L[9]);
    const float c1_5 = 1.8f * ir2, c2_5 = 0.8f * ir2;
    const float t500 = -fmaf(c1_5, rx * t400, c2_5 * (t300));
    L[10] = fmaf(10.0f * mxx, t500, L[10]);
    const float t410 = -fmaf(c1_5, rx * t310 + ry * t400, c2_5 * (t210));
    L[11] = fmaf(6.0f * mxx, t410, L[11]);
    L[10] = fmaf(4.0f * mxy, t410, L[10]);
    const float t401 = -fmaf(c1_5, rx * t301 + rz * t400, c2_5 * (t201));
    L[12] = fmaf(6.0f * mxx, t401, L[12]);
    L[10] = fmaf(4.0f * mzx, t401, L[10]);
    const float t320 = -fmaf(c1_5, rx * t220 + ry * t310, c2_5 * (t120 + t300));
    L[13] = fmaf(3.0f * mxx, t320, L[13]);
    L[10] = fmaf(myy, t320, L[10]);
    L[11] = fmaf(6.0f * mxy, t320, L[11]);
    const float t311 = -fmaf(c1_5, rx * t211 + ry * t301 + rz * t310, c2_5 * (t111));
    L[14] = fmaf(3.0f * mxx, t311, L[14]);
    L[12] = fmaf(3.0f * mxy, t311, L[12]);
    L[10] = fmaf(myz, t311, L[10]);
    L[11] = fmaf(3.0f * mzx, t311, L[11]);
    const float t302 = -fmaf(c1_5, rx * t202 + rz * t301, c2_5 * (t102 + t300));
    L[15] = fmaf(3.0f * mxx, t302, L[15]);
    L[10] = fmaf(mzz, t302, L[10]);
    L[12] = fmaf(6.0f * mzx, t302, L[12]);
    const float t230 = -fmaf(c1_5, rx * t130 + ry * t220, c2_5 * (t030 + t210));
    L[16] = fmaf(mxx, t230, L[16]);
    L[11] = fmaf(3.0f * myy, t230, L[11]);
    L[13] = fmaf(6.0f * mxy, t230, L[13]);
    const float t221 = -fmaf(c1_5, rx * t121 + ry * t211 + rz * t220, c2_5 * (t021 + t201));
    L[17] = fmaf(mxx, t221, L[17]);
    L[12] = fmaf(myy, t221, L[12]);
    L[14] = fmaf(4.0f * mxy, t221, L[14]);
    L[11] = fmaf(2.0f * myz, t221, L[11]);
    L[13] = fmaf(2.0f * mzx, t221, L[13]);
    const float t212 = -fmaf(c1_5, rx * t112 + ry * t202 + rz * t211, c2_5 * (t012 + t210));
    L[18] = fmaf(mxx, t212, L[18]);
    L[11] = fmaf(mzz, t212, L[11]);
    L[15] = fmaf(2.0f * mxy, t212, L[15]);
    L[12] = fmaf(2.0f * myz, t212, L[12]);
    L[14] = fmaf(4.0f * mzx, t212, L[14]);
    const float t203 = -fmaf(c1_5, rx * t103 + rz * t202, c2_5 * (t003 + t201));
    L[19] = fmaf(mxx, t203, L[19]);
    L[12] = fmaf(3.0f * mzz, t203, L[12]);
    L[15] = fmaf(6.0f * mzx, t203, L[15]);
    const float t140 = -fmaf(c1_5, rx * t040 + ry * t130, c2_5 * (t120));
    L[13] = fmaf(6.0f * myy, t140, L[13]);
    L[16] = fmaf(4.0f * mxy, t140, L[16]);
    const float t131 = -fmaf(c1_5, rx * t031 + ry * t121 + rz * t130, c2_5 * (t111));
    L[14] = fmaf(3.0f * myy, t131, L[14]);
    L[17] = fmaf(3.0f * mxy, t131, L[17]);
    L[13] = fmaf(3.0f * myz, t131, L[13]);
    L[16] = fmaf(mzx, t131, L[16]);
    const float t122 = -fmaf(c1_5, rx * t022 + ry * t112 + rz * t121, c2_5 * (t102 + t120));
    L[15] = fmaf(myy, t122, L[15]);
    L[13] = fmaf(mzz, t122, L[13]);
    L[18] = fmaf(2.0f * mxy, t122, L[18]);
    L[14] = fmaf(4.0f * myz, t122, L[14]);
    L[17] = fmaf(2.0f * mzx, t122, L[17]);
    const float t113 = -fmaf(c1_5, rx * t013 + ry * t103 + rz * t112, c2_5 * (t111));
    L[14] = fmaf(3.0f * mzz, t113, L[14]);
    L[19] = fmaf(mxy, t113, L[19]);
    L[15] = fmaf(3.0f * myz, t113, L[15]);
    L[18] = fmaf(3.0f * mzx, t113, L[18]);
    const float t104 = -fmaf(c1_5, rx * t004 + rz * t103, c2_5 * (t102));
    L[15] = fmaf(6.0f * mzz, t104, L[15]);
    L[19] = fmaf(4.0f * mzx, t104, L[19]);
    const float t050 = -fmaf(c1_5, ry * t040, c2_5 * (t030));
    L[16] = fmaf(10.0f * myy, t050, L[16]);
    const float t041 = -fmaf(c1_5, ry * t031 + rz * t040, c2_5 * (t021));
    L[17] = fmaf(6.0f * myy, t041, L[17]);
    L[16] = fmaf(4.0f * myz, t041, L[16]);
    const float t032 = -fmaf(c1_5, ry * t022 + rz * t031, c2_5 * (t012 + t030));
    L[18] = fmaf(3.0f * myy, t032, L[18]);
    L[16] = fmaf(mzz, t032, L[16]);
    L[17] = fmaf(6.0f * myz, t032, L[17]);
    const float t023 = -fmaf(c1_5, ry * t013 + rz * t022, c2_5 * (t003 + t021));
    L[19] = fmaf(myy, t023, L[19]);
    L[17] = fmaf(3.0f * mzz, t023, L[17]);
    L[18] = fmaf(6.0f * myz, t023, L[18]);
    const float t014 = -fmaf(c1_5, ry * t004 + rz * t013, c2_5 * (t012));
    L[18] = fmaf(6.0f * mzz, t014, L[18]);
    L[19] = fmaf(4.0f * myz, t014, L[19]);
    const float t005 = -fmaf(c1_5, rz * t004, c2_5 * (t003));
    L[19] = fmaf(10.0f * mzz, t005, L[19]);
}

// potential and field at offset y from the expansion centre (order P = 3)
__device__ inline void l2p_p3(const float (&L)[20], float x, float y, float z, float &phi, float &ex, float &ey, float &ez)
{
    const float x0 = 1.0f, x1 = x;
    const float x2 = x1 * x;
    const float x3 = x2 * x;
    const float y0 = 1.0f, y1 = y;
    const float y2 = y1 * y;
    const float y3 = y2 * y;
    const float z0 = 1.0f, z1 = z;
    const float z2 = z1 * z;
    const float z3 = z2 * z;
    phi = L[0] * (x0 * y0 * z0)
        + L[1] * (x1 * y0 * z0)
        + L[2] * (x0 * y1 * z0)
        + L[3] * (x0 * y0 * z1)
        + L[4] * (x2 * y0 * z0)
        + L[5] * (x1 * y1 * z0)
        + L[6] * (x1 * y0 * z1)
        + L[7] * (x0 * y2 * z0)
        + L[8] * (x0 * y1 * z1)
        + L[9] * (x0 * y0 * z2)
        + L[10] * (x3 * y0 * z0)
        + L[11] * (x2 * y1 * z0)
        + L[12] * (x2 * y0 * z1)
        + L[13] * (x1 * y2 * z0)
        + L[14] * (x1 * y1 * z1)
        + L[15] * (x1 * y0 * z2)
        + L[16] * (x0 * y3 * z0)
        + L[17] * (x0 * y2 * z1)
        + L[18] * (x0 * y1 * z2)
        + L[19] * (x0 * y0 * z3);
    ex = -(1.0f * L[1] * (x0 * y0 * z0)
        + 2.0f * L[4] * (x1 * y0 * z0)
        + 1.0f * L[5] * (x0 * y1 * z0)
        + 1.0f * L[6] * (x0 * y0 * z1)
        + 3.0f * L[10] * (x2 * y0 * z0)
        + 2.0f * L[11] * (x1 * y1 * z0)
        + 2.0f * L[12] * (x1 * y0 * z1)
        + 1.0f * L[13] * (x0 * y2 * z0)
        + 1.0f * L[14] * (x0 * y1 * z1)
        + 1.0f * L[15] * (x0 * y0 * z2));
    ey = -(1.0f * L[2] * (x0 * y0 * z0)
        + 1.0f * L[5] * (x1 * y0 * z0)
        + 2.0f * L[7] * (x0 * y1 * z0)
        + 1.0f * L[8] * (x0 * y0 * z1)
        + 1.0f * L[11] * (x2 * y0 * z0)
        + 2.0f * L[13] * (x1 * y1 * z0)
        + 1.0f * L[14] * (x1 * y0 * z1)
        + 3.0f * L[16] * (x0 * y2 * z0)
        + 2.0f * L[17] * (x0 * y1 * z1)
        + 1.0f * L[18] * (x0 * y0 * z2));
    ez = -(1.0f * L[3] * (x0 * y0 * z0)
        + 1.0f * L[6] * (x1 * y0 * z0)
        + 1.0f * L[8] * (x0 * y1 * z0)
        + 2.0f * L[9] * (x0 * y0 * z1)
        + 1.0f * L[12] * (x2 * y0 * z0)
        + 1.0f * L[14] * (x1 * y1 * z0)
        + 2.0f * L[15] * (x1 * y0 * z1)
        + 1.0f * L[17] * (x0 * y2 * z0)
        + 2.0f * L[18] * (x0 * y1 * z1)
        + 3.0f * L[19] * (x0 * y0 * z2));
}

// adds one cell's local coefficients about the expansion centre to L[25]
// (order P = 4; R = centre - cell centre; Q' traceless, Q'zz = -Q'xx - Q'yy)
__device__ inline void m2l_acc_p4ho(float rx, float ry, float rz, float q, float dx, float dy, float dz,
                                  float qxx, float qyy, float qxy, float qyz, float qzx,
                                  float oxxx, float oxxy, float oxxz, float oxyy, float oxyz, float oyyy, float oyyz, float (&L)[25])
{
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float mdx = -dx, mdy = -dy, mdz = -dz;
    const float mxx = qxx * (1.0f / 3.0f), myy = qyy * (1.0f / 3.0f), mzz = -(qxx + qyy) * (1.0f / 3.0f);
    const float mxy = qxy * (1.0f / 3.0f), myz = qyz * (1.0f / 3.0f), mzx = qzx * (1.0f / 3.0f);
    const float m300 = -oxxx, m210 = -oxxy, m201 = -oxxz, m120 = -oxyy, m111 = -oxyz;
    const float m030 = -oyyy, m021 = -oyyz, m102 = oxxx + oxyy, m012 = oxxy + oyyy, m003 = oxxz + oyyz;
    const float t000 = __builtin_amdgcn_rsqf(r2);
    const float ir2 = t000 * t000;
    L[0] = fmaf(q, t000, L[0]);
    const float c1_1 = 1.0f * ir2, c2_1 = 0.0f * ir2;
    const float t100 = -c1_1 * (rx * t000);
    L[1] = fmaf(q, t100, L[1]);
    L[0] = fmaf(mdx, t100, L[0]);
    const float t010 = -c1_1 * (ry * t000);
    L[2] = fmaf(q, t010, L[2]);
    L[0] = fmaf(mdy, t010, L[0]);
    const float t001 = -c1_1 * (rz * t000);
    L[3] = fmaf(q, t001, L[3]);
    L[0] = fmaf(mdz, t001, L[0]);
    const float c1_2 = 1.5f * ir2, c2_2 = 0.5f * ir2;
    const float t200 = -fmaf(c1_2, rx * t100, c2_2 * (t000));
    L[4] = fmaf(q, t200, L[4]);
    L[1] = fmaf(2.0f * mdx, t200, L[1]);
    L[0] = fmaf(mxx, t200, L[0]);
    const float t110 = -c1_2 * (rx * t010 + ry * t100);
    L[5] = fmaf(q, t110, L[5]);
    L[2] = fmaf(mdx, t110, L[2]);
    L[1] = fmaf(mdy, t110, L[1]);
    L[0] = fmaf(mxy, t110, L[0]);
    const float t101 = -c1_2 * (rx * t001 + rz * t100);
    L[6] = fmaf(q, t101, L[6]);
    L[3] = fmaf(mdx, t101, L[3]);
    L[1] = fmaf(mdz, t101, L[1]);
    L[0] = fmaf(mzx, t101, L[0]);
    const float t020 = -fmaf(c1_2, ry * t010, c2_2 * (t000));
    L[7] = fmaf(q, t020, L[7]);
    L[2] = fmaf(2.0f * mdy, t020, L[2]);
    L[0] = fmaf(myy, t020, L[0]);
    const float t011 = -c1_2 * (ry * t001 + rz * t010);
    L[8] = fmaf(q, t011, L[8]);
    L[3] = fmaf(mdy, t011, L[3]);
    L[2] = fmaf(mdz, t011, L[2]);
    L[0] = fmaf(myz, t011, L[0]);
    const float t002 = -fmaf(c1_2, rz * t001, c2_2 * (t000));
    L[3] = fmaf(2.0f * mdz, t002, L[3]);
    L[0] = fmaf(mzz, t002, L[0]);
    const float c1_3 = 1.6666666666666667f * ir2, c2_3 = 0.6666666666666666f * ir2;
    const float t300 = -fmaf(c1_3, rx * t200, c2_3 * (t100));
    L[9] = fmaf(q, t300, L[9]);
    L[4] = fmaf(3.0f * mdx, t300, L[4]);
    L[1] = fmaf(3.0f * mxx, t300, L[1]);
    L[0] = fmaf(m300, t300, L[0]);
    const float t210 = -fmaf(c1_3, rx * t110 + ry * t200, c2_3 * (t010));
    L[10] = fmaf(q, t210, L[10]);
    L[5] = fmaf(2.0f * mdx, t210, L[5]);
    L[4] = fmaf(mdy, t210, L[4]);
    L[2] = fmaf(mxx, t210, L[2]);
    L[1] = fmaf(2.0f * mxy, t210, L[1]);
    L[0] = fmaf(m210, t210, L[0]);
    const float t201 = -fmaf(c1_3, rx * t101 + rz * t200, c2_3 * (t001));
    L[11] = fmaf(q, t201, L[11]);
    L[6] = fmaf(2.0f * mdx, t201, L[6]);
    L[4] = fmaf(mdz, t201, L[4]);
    L[3] = fmaf(mxx, t201, L[3]);
    L[1] = fmaf(2.0f * mzx, t201, L[1]);
    L[0] = fmaf(m201, t201, L[0]);
    const float t120 = -fmaf(c1_3, rx * t020 + ry * t110, c2_3 * (t100));
    L[12] = fmaf(q, t120, L[12]);
    L[7] = fmaf(mdx, t120, L[7]);
    L[5] = fmaf(2.0f * mdy, t120, L[5]);
    L[1] = fmaf(myy, t120, L[1]);
    L[2] = fmaf(2.0f * mxy, t120, L[2]);
    L[0] = fmaf(m120, t120, L[0]);
    const float t111 = -c1_3 * (rx * t011 + ry * t101 + rz * t110);
    L[13] = fmaf(q, t111, L[13]);
    L[8] = fmaf(mdx, t111, L[8]);
    L[6] = fmaf(mdy, t111, L[6]);
    L[5] = fmaf(mdz, t111, L[5]);
    L[3] = fmaf(mxy, t111, L[3]);
    L[1] = fmaf(myz, t111, L[1]);
    L[2] = fmaf(mzx, t111, L[2]);
    L[0] = fmaf(m111, t111, L[0]);
    const float t102 = -fmaf(c1_3, rx * t002 + rz * t101, c2_3 * (t100));
    L[6] = fmaf(2.0f * mdz, t102, L[6]);
    L[1] = fmaf(mzz, t102, L[1]);
    L[3] = fmaf(2.0f * mzx, t102, L[3]);
    L[0] = fmaf(m102, t102, L[0]);
    const float t030 = -fmaf(c1_3, ry * t020, c2_3 * (t010));
    L[14] = fmaf(q, t030, L[14]);
    L[7] = fmaf(3.0f * mdy, t030, L[7]);
    L[2] = fmaf(3.0f * myy, t030, L[2]);
    L[0] = fmaf(m030, t030, L[0]);
    const float t021 = -fmaf(c1_3, ry * t011 + rz * t020, c2_3 * (t001));
    L[15] = fmaf(q, t021, L[15]);
    L[8] = fmaf(2.0f * mdy, t021, L[8]);
    L[7] = fmaf(mdz, t021, L[7]);
    L[3] = fmaf(myy, t021, L[3]);
    L[2] = fmaf(2.0f * myz, t021, L[2]);
    L[0] = fmaf(m021, t021, L[0]);
    const float t012 = -fmaf(c1_3, ry * t002 + rz * t011, c2_3 * (t010));
    L[8] = fmaf(2.0f * mdz, t012, L[8]);
    L[2] = fmaf(mzz, t012, L[2]);
    L[3] = fmaf(2.0f * myz, t012, L[3]);
    L[0] = fmaf(m012, t012, L[0]);
    const float t003 = -fmaf(c1_3, rz * t002, c2_3 * (t001));
    L[3] = fmaf(3.0f * mzz, t003, L[3]);
    L[0] = fmaf(m003, t003, L[0]);
    const float c1_4 = 1.75f * ir2, c2_4 = 0.75f * ir2;
    const float t400 = -fmaf(c1_4, rx * t300, c2_4 * (t200));
    L[16] = fmaf(q, t400, L[16]);
    L[9] = fmaf(4.0f * mdx, t400, L[9]);
    L[4] = fmaf(6.0f * mxx, t400, L[4]);
    L[1] = fmaf(4.0f * m300, t400, L[1]);
    const float t310 = -fmaf(c1_4, rx * t210 + ry * t300, c2_4 * (t110));
    L[17] = fmaf(q, t310, L[17]);
    L[10] = fmaf(3.0f * mdx, t310, L[10]);
    L[9] = fmaf(mdy, t310, L[9]);
    L[5] = fmaf(3.0f * mxx, t310, L[5]);
    L[4] = fmaf(3.0f * mxy, t310, L[4]);
    L[2] = fmaf(m300, t310, L[2]);
    L[1] = fmaf(3.0f * m210, t310, L[1]);
    const float t301 = -fmaf(c1_4, rx * t201 + rz * t300, c2_4 * (t101));
    L[18] = fmaf(q, t301, L[18]);
    L[11] = fmaf(3.0f * mdx, t301, L[11]);
    L[9] = fmaf(mdz, t301, L[9]);
    L[6] = fmaf(3.0f * mxx, t301, L[6]);
    L[4] = fmaf(3.0f * mzx, t301, L[4]);
    L[3] = fmaf(m300, t301, L[3]);
    L[1] = fmaf(3.0f * m201, t301, L[1]);
    const float t220 = -fmaf(c1_4, rx * t120 + ry * t210, c2_4 * (t020 + t200));
    L[19] = fmaf(q, t220, L[19]);
    L[12] = fmaf(2.0f * mdx, t220, L[12]);
    L[10] = fmaf(2.0f * mdy, t220, L[10]);
    L[7] = fmaf(mxx, t220, L[7]);
    L[4] = fmaf(myy, t220, L[4]);
    L[5] = fmaf(4.0f * mxy, t220, L[5]);
    L[2] = fmaf(2.0f * m210, t220, L[2]);
    L[1] = fmaf(2.0f * m120, t220, L[1]);
    const float t211 = -fmaf(c1_4, rx * t111 + ry * t201 + rz * t210, c2_4 * (t011));
    L[20] = fmaf(q, t211, L[20]);
    L[13] = fmaf(2.0f * mdx, t211, L[13]);
    L[11] = fmaf(mdy, t211, L[11]);
    L[10] = fmaf(mdz, t211, L[10]);
    L[8] = fmaf(mxx, t211, L[8]);
    L[6] = fmaf(2.0f * mxy, t211, L[6]);
    L[4] = fmaf(myz, t211, L[4]);
    L[5] = fmaf(2.0f * mzx, t211, L[5]);
    L[3] = fmaf(m210, t211, L[3]);
    L[2] = fmaf(m201, t211, L[2]);
    L[1] = fmaf(2.0f * m111, t211, L[1]);
    const float t202 = -fmaf(c1_4, rx * t102 + rz * t201, c2_4 * (t002 + t200));
    L[11] = fmaf(2.0f * mdz, t202, L[11]);
    L[4] = fmaf(mzz, t202, L[4]);
    L[6] = fmaf(4.0f * mzx, t202, L[6]);
    L[3] = fmaf(2.0f * m201, t202, L[3]);
    L[1] = fmaf(2.0f * m102, t202, L[1]);
    const float t130 = -fmaf(c1_4, rx * t030 + ry * t120, c2_4 * (t110));
    L[21] = fmaf(q, t130, L[21]);
    L[14] = fmaf(mdx, t130, L[14]);
    L[12] = fmaf(3.0f * mdy, t130, L[12]);
    L[5] = fmaf(3.0f * myy, t130, L[5]);
    L[7] = fmaf(3.0f * mxy, t130, L[7]);
    L[2] = fmaf(3.0f * m120, t130, L[2]);
    L[1] = fmaf(m030, t130, L[1]);
    const float t121 = -fmaf(c1_4, rx * t021 + ry * t111 + rz * t120, c2_4 * (t101));
    L[22] = fmaf(q, t121, L[22]);
    L[15] = fmaf(mdx, t121, L[15]);
    L[13] = fmaf(2.0f * mdy, t121, L[13]);
    L[12] = fmaf(mdz, t121, L[12]);
    L[6] = fmaf(myy, t121, L[6]);
    L[8] = fmaf(2.0f * mxy, t121, L[8]);
    L[5] = fmaf(2.0f * myz, t121, L[5]);
    L[7] = fmaf(mzx, t121, L[7]);
    L[3] = fmaf(m120, t121, L[3]);
    L[2] = fmaf(2.0f * m111, t121, L[2]);
    L[1] = fmaf(m021, t121, L[1]);
    const float t112 = -fmaf(c1_4, rx * t012 + ry * t102 + rz * t111, c2_4 * (t110));
    L[13] = fmaf(2.0f * mdz, t112, L[13]);
    L[5] = fmaf(mzz, t112, L[5]);
    L[6] = fmaf(2.0f * myz, t112, L[6]);
    L[8] = fmaf(2.0f * mzx, t112, L[8]);
    L[3] = fmaf(2.0f * m111, t112, L[3]);
    L[2] = fmaf(m102, t112, L[2]);
    L[1] = fmaf(m012, t112, L[1]);
    const float t103 = -fmaf(c1_4, rx * t003 + rz * t102, c2_4 * (t101));
    L[6] = fmaf(3.0f * mzz, t103, L[6]);
    L[3] = fmaf(3.0f * m102, t103, L[3]);
    L[1] = fmaf(m003, t103, L[1]);
    const float t040 = -fmaf(c1_4, ry * t030, c2_4 * (t020));
    L[23] = fmaf(q, t040, L[23]);
    L[14] = fmaf(4.0f * mdy, t040, L[14]);
    L[7] = fmaf(6.0f * myy, t040, L[7]);
    L[2] = fmaf(4.0f * m030, t040, L[2]);
    const float t031 = -fmaf(c1_4, ry * t021 + rz * t030, c2_4 * (t011));
    L[24] = fmaf(q, t031, L[24]);
    L[15] = fmaf(3.0f * mdy, t031, L[15]);
    L[14] = fmaf(mdz, t031, L[14]);
    L[8] = fmaf(3.0f * myy, t031, L[8]);
    L[7] = fmaf(3.0f * myz, t031, L[7]);
    L[3] = fmaf(m030, t031, L[3]);
    L[2] = fmaf(3.0f * m021, t031, L[2]);
    const float t022 = -fmaf(c1_4, ry * t012 + rz * t021, c2_4 * (t002 + t020));
    L[15] = fmaf(2.0f * mdz, t022, L[15]);
    L[7] = fmaf(mzz, t022, L[7]);
    L[8] = fmaf(4.0f * myz, t022, L[8]);
    L[3] = fmaf(2.0f * m021, t022, L[3]);
    L[2] = fmaf(2.0f * m012, t022, L[2]);
    const float t013 = -fmaf(c1_4, ry * t003 + rz * t012, c2_4 * (t011));
    L[8] = fmaf(3.0f * mzz, t013, L[8]);
    L[3] = fmaf(3.0f * m012, t013, L[3]);
    L[2] = fmaf(m003, t013, L[2]);
    const float t004 = -fmaf(c1_4, rz * t003, c2_4 * (t002));
    L[3] = fmaf(4.0f * m003, t004, L[3]);
    const float c1_5 = 1.8f * ir2, c2_5 = 0.8f * ir2;
    const float t500 = -fmaf(c1_5, rx * t400, c2_5 * (t300));
    L[16] = fmaf(5.0f * mdx, t500, L[16]);
    L[9] = fmaf(10.0f * mxx, t500, L[9]);
    L[4] = fmaf(10.0f * m300, t500, L[4]);
    const float t410 = -fmaf(c1_5, rx * t310 + ry * t400, c2_5 * (t210));
    L[17] = fmaf(4.0f * mdx, t410, L[17]);
    L[16] = fmaf(mdy, t410, L[16]);
    L[10] = fmaf(6.0f * mxx, t410, L[10]);
    L[9] = fmaf(4.0f * mxy, t410, L[9]);
    L[5] = fmaf(4.0f * m300, t410, L[5]);
    L[4] = fmaf(6.0f * m210, t410, L[4]);
    const float t401 = -fmaf(c1_5, rx * t301 + rz * t400, c2_5 * (t201));
    L[18] = fmaf(4.0f * mdx, t401, L[18]);
    L[16] = fmaf(mdz, t401, L[16]);
    L[11] = fmaf(6.0f * mxx, t401, L[11]);
    L[9] = fmaf(4.0f * mzx, t401, L[9]);
    L[6] = fmaf(4.0f * m300, t401, L[6]);
    L[4] = fmaf(6.0f * m201, t401, L[4]);
    const float t320 = -fmaf(c1_5, rx * t220 + ry * t310, c2_5 * (t120 + t300));
    L[19] = fmaf(3.0f * mdx, t320, L[19]);
    L[17] = fmaf(2.0f * mdy, t320, L[17]);
    L[12] = fmaf(3.0f * mxx, t320, L[12]);
    L[9] = fmaf(myy, t320, L[9]);
    L[10] = fmaf(6.0f * mxy, t320, L[10]);
    L[7] = fmaf(m300, t320, L[7]);
    L[5] = fmaf(6.0f * m210, t320, L[5]);
    L[4] = fmaf(3.0f * m120, t320, L[4]);
    const float t311 = -fmaf(c1_5, rx * t211 + ry * t301 + rz * t310, c2_5 * (t111));
    L[20] = fmaf(3.0f * mdx, t311, L[20]);
    L[18] = fmaf(mdy, t311, L[18]);
    L[17] = fmaf(mdz, t311, L[17]);
    L[13] = fmaf(3.0f * mxx, t311, L[13]);
    L[11] = fmaf(3.0f * mxy, t311, L[11]);
    L[9] = fmaf(myz, t311, L[9]);
    L[10] = fmaf(3.0f * mzx, t311, L[10]);
    L[8] = fmaf(m300, t311, L[8]);
    L[6] = fmaf(3.0f * m210, t311, L[6]);
    L[5] = fmaf(3.0f * m201, t311, L[5]);
    L[4] = fmaf(3.0f * m111, t311, L[4]);
    const float t302 = -fmaf(c1_5, rx * t202 + rz * t301, c2_5 * (t102 + t300));
    L[18] = fmaf(2.0f * mdz, t302, L[18]);
    L[9] = fmaf(mzz, t302, L[9]);
    L[11] = fmaf(6.0f * mzx, t302, L[11]);
    L[6] = fmaf(6.0f * m201, t302, L[6]);
    L[4] = fmaf(3.0f * m102, t302, L[4]);
    const float t230 = -fmaf(c1_5, rx * t130 + ry * t220, c2_5 * (t030 + t210));
    L[21] = fmaf(2.0f * mdx, t230, L[21]);
    L[19] = fmaf(3.0f * mdy, t230, L[19]);
    L[14] = fmaf(mxx, t230, L[14]);
    L[10] = fmaf(3.0f * myy, t230, L[10]);
    L[12] = fmaf(6.0f * mxy, t230, L[12]);
    L[7] = fmaf(3.0f * m210, t230, L[7]);
    L[5] = fmaf(6.0f * m120, t230, L[5]);
    L[4] = fmaf(m030, t230, L[4]);
    const float t221 = -fmaf(c1_5, rx * t121 + ry * t211 + rz * t220, c2_5 * (t021 + t201));
    L[22] = fmaf(2.0f * mdx, t221, L[22]);
    L[20] = fmaf(2.0f * mdy, t221, L[20]);
    L[19] = fmaf(mdz, t221, L[19]);
    L[15] = fmaf(mxx, t221, L[15]);
    L[11] = fmaf(myy, t221, L[11]);
    L[13] = fmaf(4.0f * mxy, t221, L[13]);
    L[10] = fmaf(2.0f * myz, t221, L[10]);
    L[12] = fmaf(2.0f * mzx, t221, L[12]);
    L[8] = fmaf(2.0f * m210, t221, L[8]);
    L[7] = fmaf(m201, t221, L[7]);
    L[6] = fmaf(2.0f * m120, t221, L[6]);
    L[5] = fmaf(4.0f * m111, t221, L[5]);
    L[4] = fmaf(m021, t221, L[4]);
    const float t212 = -fmaf(c1_5, rx * t112 + ry * t202 + rz * t211, c2_5 * (t012 + t210));
    L[20] = fmaf(2.0f * mdz, t212, L[20]);
    L[10] = fmaf(mzz, t212, L[10]);
    L[11] = fmaf(2.0f * myz, t212, L[11]);
    L[13] = fmaf(4.0f * mzx, t212, L[13]);
    L[8] = fmaf(2.0f * m201, t212, L[8]);
    L[6] = fmaf(4.0f * m111, t212, L[6]);
    L[5] = fmaf(2.0f * m102, t212, L[5]);
    L[4] = fmaf(m012, t212, L[4]);
    const float t203 = -fmaf(c1_5, rx * t103 + rz * t202, c2_5 * (t003 + t201));
    L[11] = fmaf(3.0f * mzz, t203, L[11]);
    L[6] = fmaf(6.0f * m102, t203, L[6]);
    L[4] = fmaf(m003, t203, L[4]);
    const float t140 = -fmaf(c1_5, rx * t040 + ry * t130, c2_5 * (t120));
    L[23] = fmaf(mdx, t140, L[23]);
    L[21] = fmaf(4.0f * mdy, t140, L[21]);
    L[12] = fmaf(6.0f * myy, t140, L[12]);
    L[14] = fmaf(4.0f * mxy, t140, L[14]);
    L[7] = fmaf(6.0f * m120, t140, L[7]);
    L[5] = fmaf(4.0f * m030, t140, L[5]);
    const float t131 = -fmaf(c1_5, rx * t031 + ry * t121 + rz * t130, c2_5 * (t111));
    L[24] = fmaf(mdx, t131, L[24]);
    L[22] = fmaf(3.0f * mdy, t131, L[22]);
    L[21] = fmaf(mdz, t131, L[21]);
    L[13] = fmaf(3.0f * myy, t131, L[13]);
    L[15] = fmaf(3.0f * mxy, t131, L[15]);
    L[12] = fmaf(3.0f * myz, t131, L[12]);
    L[14] = fmaf(mzx, t131, L[14]);
    L[8] = fmaf(3.0f * m120, t131, L[8]);
    L[7] = fmaf(3.0f * m111, t131, L[7]);
    L[6] = fmaf(m030, t131, L[6]);
    L[5] = fmaf(3.0f * m021, t131, L[5]);
    const float t122 = -fmaf(c1_5, rx * t022 + ry * t112 + rz * t121, c2_5 * (t102 + t120));
    L[22] = fmaf(2.0f * mdz, t122, L[22]);
    L[12] = fmaf(mzz, t122, L[12]);
    L[13] = fmaf(4.0f * myz, t122, L[13]);
    L[15] = fmaf(2.0f * mzx, t122, L[15]);
    L[8] = fmaf(4.0f * m111, t122, L[8]);
    L[7] = fmaf(m102, t122, L[7]);
    L[6] = fmaf(2.0f * m021, t122, L[6]);
    L[5] = fmaf(2.0f * m012, t122, L[5]);
    const float t113 = -fmaf(c1_5, rx * t013 + ry * t103 + rz * t112, c2_5 * (t111));
    L[13] = fmaf(3.0f * mzz, t113, L[13]);
    L[8] = fmaf(3.0f * m102, t113, L[8]);
    L[6] = fmaf(3.0f * m012, t113, L[6]);
    L[5] = fmaf(m003, t113, L[5]);
    const float t104 = -fmaf(c1_5, rx * t004 + rz * t103, c2_5 * (t102));
    L[6] = fmaf(4.0f * m003, t104, L[6]);
    const float t050 = -fmaf(c1_5, ry * t040, c2_5 * (t030));
    L[23] = fmaf(5.0f * mdy, t050, L[23]);
    L[14] = fmaf(10.0f * myy, t050, L[14]);
    L[7] = fmaf(10.0f * m030, t050, L[7]);
    const float t041 = -fmaf(c1_5, ry * t031 + rz * t040, c2_5 * (t021));
    L[24] = fmaf(4.0f * mdy, t041, L[24]);
    L[23] = fmaf(mdz, t041, L[23]);
    L[15] = fmaf(6.0f * myy, t041, L[15]);
    L[14] = fmaf(4.0f * myz, t041, L[14]);
    L[8] = fmaf(4.0f * m030, t041, L[8]);
    L[7] = fmaf(6.0f * m021, t041, L[7]);
    const float t032 = -fmaf(c1_5, ry * t022 + rz * t031, c2_5 * (t012 + t030));
    L[24] = fmaf(2.0f * mdz, t032, L[24]);
    L[14] = fmaf(mzz, t032, L[14]);
    L[15] = fmaf(6.0f * myz, t032, L[15]);
    L[8] = fmaf(6.0f * m021, t032, L[8]);
    L[7] = fmaf(3.0f * m012, t032, L[7]);
    const float t023 = -fmaf(c1_5, ry * t013 + rz * t022, c2_5 * (t003 + t021));
    L[15] = fmaf(3.0f * mzz, t023, L[15]);
    L[8] = fmaf(6.0f * m012, t023, L[8]);
    L[7] = fmaf(m003, t023, L[7]);
    const float t014 = -fmaf(c1_5, ry * t004 + rz * t013, c2_5 * (t012));
    L[8] = fmaf(4.0f * m003, t014, L[8]);
    const float c1_6 = 1.8333333333333333f * ir2, c2_6 = 0.8333333333333334f * ir2;
    const float t600 = -fmaf(c1_6, rx * t500, c2_6 * (t400));
    L[16] = fmaf(15.0f * mxx, t600, L[16]);
    L[9] = fmaf(20.0f * m300, t600, L[9]);
    const float t510 = -fmaf(c1_6, rx * t410 + ry * t500, c2_6 * (t310));
    L[17] = fmaf(10.0f * mxx, t510, L[17]);
    L[16] = fmaf(5.0f * mxy, t510, L[16]);
    L[10] = fmaf(10.0f * m300, t510, L[10]);
    L[9] = fmaf(10.0f * m210, t510, L[9]);
    const float t501 = -fmaf(c1_6, rx * t401 + rz * t500, c2_6 * (t301));
    L[18] = fmaf(10.0f * mxx, t501, L[18]);
    L[16] = fmaf(5.0f * mzx, t501, L[16]);
    L[11] = fmaf(10.0f * m300, t501, L[11]);
    L[9] = fmaf(10.0f * m201, t501, L[9]);
    const float t420 = -fmaf(c1_6, rx * t320 + ry * t410, c2_6 * (t220 + t400));
    L[19] = fmaf(6.0f * mxx, t420, L[19]);
    L[16] = fmaf(myy, t420, L[16]);
    L[17] = fmaf(8.0f * mxy, t420, L[17]);
    L[12] = fmaf(4.0f * m300, t420, L[12]);
    L[10] = fmaf(12.0f * m210, t420, L[10]);
    L[9] = fmaf(4.0f * m120, t420, L[9]);
    const float t411 = -fmaf(c1_6, rx * t311 + ry * t401 + rz * t410, c2_6 * (t211));
    L[20] = fmaf(6.0f * mxx, t411, L[20]);
    L[18] = fmaf(4.0f * mxy, t411, L[18]);
    L[16] = fmaf(myz, t411, L[16]);
    L[17] = fmaf(4.0f * mzx, t411, L[17]);
    L[13] = fmaf(4.0f * m300, t411, L[13]);
    L[11] = fmaf(6.0f * m210, t411, L[11]);
    L[10] = fmaf(6.0f * m201, t411, L[10]);
    L[9] = fmaf(4.0f * m111, t411, L[9]);
    const float t402 = -fmaf(c1_6, rx * t302 + rz * t401, c2_6 * (t202 + t400));
    L[16] = fmaf(mzz, t402, L[16]);
    L[18] = fmaf(8.0f * mzx, t402, L[18]);
    L[11] = fmaf(12.0f * m201, t402, L[11]);
    L[9] = fmaf(4.0f * m102, t402, L[9]);
    const float t330 = -fmaf(c1_6, rx * t230 + ry * t320, c2_6 * (t130 + t310));
    L[21] = fmaf(3.0f * mxx, t330, L[21]);
    L[17] = fmaf(3.0f * myy, t330, L[17]);
    L[19] = fmaf(9.0f * mxy, t330, L[19]);
    L[14] = fmaf(m300, t330, L[14]);
    L[12] = fmaf(9.0f * m210, t330, L[12]);
    L[10] = fmaf(9.0f * m120, t330, L[10]);
    L[9] = fmaf(m030, t330, L[9]);
    const float t321 = -fmaf(c1_6, rx * t221 + ry * t311 + rz * t320, c2_6 * (t121 + t301));
    L[22] = fmaf(3.0f * mxx, t321, L[22]);
    L[18] = fmaf(myy, t321, L[18]);
    L[20] = fmaf(6.0f * mxy, t321, L[20]);
    L[17] = fmaf(2.0f * myz, t321, L[17]);
    L[19] = fmaf(3.0f * mzx, t321, L[19]);
    L[15] = fmaf(m300, t321, L[15]);
    L[13] = fmaf(6.0f * m210, t321, L[13]);
    L[12] = fmaf(3.0f * m201, t321, L[12]);
    L[11] = fmaf(3.0f * m120, t321, L[11]);
    L[10] = fmaf(6.0f * m111, t321, L[10]);
    L[9] = fmaf(m021, t321, L[9]);
    const float t312 = -fmaf(c1_6, rx * t212 + ry * t302 + rz * t311, c2_6 * (t112 + t310));
    L[17] = fmaf(mzz, t312, L[17]);
    L[18] = fmaf(2.0f * myz, t312, L[18]);
    L[20] = fmaf(6.0f * mzx, t312, L[20]);
    L[13] = fmaf(6.0f * m201, t312, L[13]);
    L[11] = fmaf(6.0f * m111, t312, L[11]);
    L[10] = fmaf(3.0f * m102, t312, L[10]);
    L[9] = fmaf(m012, t312, L[9]);
    const float t303 = -fmaf(c1_6, rx * t203 + rz * t302, c2_6 * (t103 + t301));
    L[18] = fmaf(3.0f * mzz, t303, L[18]);
    L[11] = fmaf(9.0f * m102, t303, L[11]);
    L[9] = fmaf(m003, t303, L[9]);
    const float t240 = -fmaf(c1_6, rx * t140 + ry * t230, c2_6 * (t040 + t220));
    L[23] = fmaf(mxx, t240, L[23]);
    L[19] = fmaf(6.0f * myy, t240, L[19]);
    L[21] = fmaf(8.0f * mxy, t240, L[21]);
    L[14] = fmaf(4.0f * m210, t240, L[14]);
    L[12] = fmaf(12.0f * m120, t240, L[12]);
    L[10] = fmaf(4.0f * m030, t240, L[10]);
    const float t231 = -fmaf(c1_6, rx * t131 + ry * t221 + rz * t230, c2_6 * (t031 + t211));
    L[24] = fmaf(mxx, t231, L[24]);
    L[20] = fmaf(3.0f * myy, t231, L[20]);
    L[22] = fmaf(6.0f * mxy, t231, L[22]);
    L[19] = fmaf(3.0f * myz, t231, L[19]);
    L[21] = fmaf(2.0f * mzx, t231, L[21]);
    L[15] = fmaf(3.0f * m210, t231, L[15]);
    L[14] = fmaf(m201, t231, L[14]);
    L[13] = fmaf(6.0f * m120, t231, L[13]);
    L[12] = fmaf(6.0f * m111, t231, L[12]);
    L[11] = fmaf(m030, t231, L[11]);
    L[10] = fmaf(3.0f * m021, t231, L[10]);
    const float t222 = -fmaf(c1_6, rx * t122 + ry * t212 + rz * t221, c2_6 * (t022 + t202 + t220));
    L[19] = fmaf(mzz, t222, L[19]);
    L[20] = fmaf(4.0f * myz, t222, L[20]);
    L[22] = fmaf(4.0f * mzx, t222, L[22]);
    L[15] = fmaf(2.0f * m201, t222, L[15]);
    L[13] = fmaf(8.0f * m111, t222, L[13]);
    L[12] = fmaf(2.0f * m102, t222, L[12]);
    L[11] = fmaf(2.0f * m021, t222, L[11]);
    L[10] = fmaf(2.0f * m012, t222, L[10]);
    const float t213 = -fmaf(c1_6, rx * t113 + ry * t203 + rz * t212, c2_6 * (t013 + t211));
    L[20] = fmaf(3.0f * mzz, t213, L[20]);
    L[13] = fmaf(6.0f * m102, t213, L[13]);
    L[11] = fmaf(3.0f * m012, t213, L[11]);
    L[10] = fmaf(m003, t213, L[10]);
    const float t204 = -fmaf(c1_6, rx * t104 + rz * t203, c2_6 * (t004 + t202));
    L[11] = fmaf(4.0f * m003, t204, L[11]);
    const float t150 = -fmaf(c1_6, rx * t050 + ry * t140, c2_6 * (t130));
    L[21] = fmaf(10.0f * myy, t150, L[21]);
    L[23] = fmaf(5.0f * mxy, t150, L[23]);
    L[14] = fmaf(10.0f * m120, t150, L[14]);
    L[12] = fmaf(10.0f * m030, t150, L[12]);
    const float t141 = -fmaf(c1_6, rx * t041 + ry * t131 + rz * t140, c2_6 * (t121));
    L[22] = fmaf(6.0f * myy, t141, L[22]);
    L[24] = fmaf(4.0f * mxy, t141, L[24]);
    L[21] = fmaf(4.0f * myz, t141, L[21]);
    L[23] = fmaf(mzx, t141, L[23]);
    L[15] = fmaf(6.0f * m120, t141, L[15]);
    L[14] = fmaf(4.0f * m111, t141, L[14]);
    L[13] = fmaf(4.0f * m030, t141, L[13]);
    L[12] = fmaf(6.0f * m021, t141, L[12]);
    const float t132 = -fmaf(c1_6, rx * t032 + ry * t122 + rz * t131, c2_6 * (t112 + t130));
    L[21] = fmaf(mzz, t132, L[21]);
    L[22] = fmaf(6.0f * myz, t132, L[22]);
    L[24] = fmaf(2.0f * mzx, t132, L[24]);
    L[15] = fmaf(6.0f * m111, t132, L[15]);
    L[14] = fmaf(m102, t132, L[14]);
    L[13] = fmaf(6.0f * m021, t132, L[13]);
    L[12] = fmaf(3.0f * m012, t132, L[12]);
    const float t123 = -fmaf(c1_6, rx * t023 + ry * t113 + rz * t122, c2_6 * (t103 + t121));
    L[22] = fmaf(3.0f * mzz, t123, L[22]);
    L[15] = fmaf(3.0f * m102, t123, L[15]);
    L[13] = fmaf(6.0f * m012, t123, L[13]);
    L[12] = fmaf(m003, t123, L[12]);
    const float t114 = -fmaf(c1_6, rx * t014 + ry * t104 + rz * t113, c2_6 * (t112));
    L[13] = fmaf(4.0f * m003, t114, L[13]);
    const float t060 = -fmaf(c1_6, ry * t050, c2_6 * (t040));
    L[23] = fmaf(15.0f * myy, t060, L[23]);
    L[14] = fmaf(20.0f * m030, t060, L[14]);
    const float t051 = -fmaf(c1_6, ry * t041 + rz * t050, c2_6 * (t031));
    L[24] = fmaf(10.0f * myy, t051, L[24]);
    L[23] = fmaf(5.0f * myz, t051, L[23]);
    L[15] = fmaf(10.0f * m030, t051, L[15]);
    L[14] = fmaf(10.0f * m021, t051, L[14]);
    const float t042 = -fmaf(c1_6, ry * t032 + rz * t041, c2_6 * (t022 + t040));
    L[23] = fmaf(mzz, t042, L[23]);
    L[24] = fmaf(8.0f * myz, t042, L[24]);
    L[15] = fmaf(12.0f * m021, t042, L[15]);
    L[14] = fmaf(4.0f * m012, t042, L[14]);
    const float t033 = -fmaf(c1_6, ry * t023 + rz * t032, c2_6 * (t013 + t031));
    L[24] = fmaf(3.0f * mzz, t033, L[24]);
    L[15] = fmaf(9.0f * m012, t033, L[15]);
    L[14] = fmaf(m003, t033, L[14]);
    const float t024 = -fmaf(c1_6, ry * t014 + rz * t023, c2_6 * (t004 + t022));
    L[15] = fmaf(4.0f * m003, t024, L[15]);
    const float c1_7 = 1.8571428571428572f * ir2, c2_7 = 0.8571428571428571f * ir2;
    const float t700 = -fmaf(c1_7, rx * t600, c2_7 * (t500));
    L[16] = fmaf(35.0f * m300, t700, L[16]);
    const float t610 = -fmaf(c1_7, rx * t510 + ry * t600, c2_7 * (t410));
    L[17] = fmaf(20.0f * m300, t610, L[17]);
    L[16] = fmaf(15.0f * m210, t610, L[16]);
    const float t601 = -fmaf(c1_7, rx * t501 + rz * t600, c2_7 * (t401));
    L[18] = fmaf(20.0f * m300, t601, L[18]);
    L[16] = fmaf(15.0f * m201, t601, L[16]);
    const float t520 = -fmaf(c1_7, rx * t420 + ry * t510, c2_7 * (t320 + t500));
    L[19] = fmaf(10.0f * m300, t520, L[19]);
    L[17] = fmaf(20.0f * m210, t520, L[17]);
    L[16] = fmaf(5.0f * m120, t520, L[16]);
    const float t511 = -fmaf(c1_7, rx * t411 + ry * t501 + rz * t510, c2_7 * (t311));
    L[20] = fmaf(10.0f * m300, t511, L[20]);
    L[18] = fmaf(10.0f * m210, t511, L[18]);
    L[17] = fmaf(10.0f * m201, t511, L[17]);
    L[16] = fmaf(5.0f * m111, t511, L[16]);
    const float t502 = -fmaf(c1_7, rx * t402 + rz * t501, c2_7 * (t302 + t500));
    L[18] = fmaf(20.0f * m201, t502, L[18]);
    L[16] = fmaf(5.0f * m102, t502, L[16]);
    const float t430 = -fmaf(c1_7, rx * t330 + ry * t420, c2_7 * (t230 + t410));
    L[21] = fmaf(4.0f * m300, t430, L[21]);
    L[19] = fmaf(18.0f * m210, t430, L[19]);
    L[17] = fmaf(12.0f * m120, t430, L[17]);
    L[16] = fmaf(m030, t430, L[16]);
    const float t421 = -fmaf(c1_7, rx * t321 + ry * t411 + rz * t420, c2_7 * (t221 + t401));
    L[22] = fmaf(4.0f * m300, t421, L[22]);
    L[20] = fmaf(12.0f * m210, t421, L[20]);
    L[19] = fmaf(6.0f * m201, t421, L[19]);
    L[18] = fmaf(4.0f * m120, t421, L[18]);
    L[17] = fmaf(8.0f * m111, t421, L[17]);
    L[16] = fmaf(m021, t421, L[16]);
    const float t412 = -fmaf(c1_7, rx * t312 + ry * t402 + rz * t411, c2_7 * (t212 + t410));
    L[20] = fmaf(12.0f * m201, t412, L[20]);
    L[18] = fmaf(8.0f * m111, t412, L[18]);
    L[17] = fmaf(4.0f * m102, t412, L[17]);
    L[16] = fmaf(m012, t412, L[16]);
    const float t403 = -fmaf(c1_7, rx * t303 + rz * t402, c2_7 * (t203 + t401));
    L[18] = fmaf(12.0f * m102, t403, L[18]);
    L[16] = fmaf(m003, t403, L[16]);
    const float t340 = -fmaf(c1_7, rx * t240 + ry * t330, c2_7 * (t140 + t320));
    L[23] = fmaf(m300, t340, L[23]);
    L[21] = fmaf(12.0f * m210, t340, L[21]);
    L[19] = fmaf(18.0f * m120, t340, L[19]);
    L[17] = fmaf(4.0f * m030, t340, L[17]);
    const float t331 = -fmaf(c1_7, rx * t231 + ry * t321 + rz * t330, c2_7 * (t131 + t311));
    L[24] = fmaf(m300, t331, L[24]);
    L[22] = fmaf(9.0f * m210, t331, L[22]);
    L[21] = fmaf(3.0f * m201, t331, L[21]);
    L[20] = fmaf(9.0f * m120, t331, L[20]);
    L[19] = fmaf(9.0f * m111, t331, L[19]);
    L[18] = fmaf(m030, t331, L[18]);
    L[17] = fmaf(3.0f * m021, t331, L[17]);
    const float t322 = -fmaf(c1_7, rx * t222 + ry * t312 + rz * t321, c2_7 * (t122 + t302 + t320));
    L[22] = fmaf(6.0f * m201, t322, L[22]);
    L[20] = fmaf(12.0f * m111, t322, L[20]);
    L[19] = fmaf(3.0f * m102, t322, L[19]);
    L[18] = fmaf(2.0f * m021, t322, L[18]);
    L[17] = fmaf(2.0f * m012, t322, L[17]);
    const float t313 = -fmaf(c1_7, rx * t213 + ry * t303 + rz * t312, c2_7 * (t113 + t311));
    L[20] = fmaf(9.0f * m102, t313, L[20]);
    L[18] = fmaf(3.0f * m012, t313, L[18]);
    L[17] = fmaf(m003, t313, L[17]);
    const float t304 = -fmaf(c1_7, rx * t204 + rz * t303, c2_7 * (t104 + t302));
    L[18] = fmaf(4.0f * m003, t304, L[18]);
    const float t250 = -fmaf(c1_7, rx * t150 + ry * t240, c2_7 * (t050 + t230));
    L[23] = fmaf(5.0f * m210, t250, L[23]);
    L[21] = fmaf(20.0f * m120, t250, L[21]);
    L[19] = fmaf(10.0f * m030, t250, L[19]);
    const float t241 = -fmaf(c1_7, rx * t141 + ry * t231 + rz * t240, c2_7 * (t041 + t221));
    L[24] = fmaf(4.0f * m210, t241, L[24]);
    L[23] = fmaf(m201, t241, L[23]);
    L[22] = fmaf(12.0f * m120, t241, L[22]);
    L[21] = fmaf(8.0f * m111, t241, L[21]);
    L[20] = fmaf(4.0f * m030, t241, L[20]);
    L[19] = fmaf(6.0f * m021, t241, L[19]);
    const float t232 = -fmaf(c1_7, rx * t132 + ry * t222 + rz * t231, c2_7 * (t032 + t212 + t230));
    L[24] = fmaf(2.0f * m201, t232, L[24]);
    L[22] = fmaf(12.0f * m111, t232, L[22]);
    L[21] = fmaf(2.0f * m102, t232, L[21]);
    L[20] = fmaf(6.0f * m021, t232, L[20]);
    L[19] = fmaf(3.0f * m012, t232, L[19]);
    const float t223 = -fmaf(c1_7, rx * t123 + ry * t213 + rz * t222, c2_7 * (t023 + t203 + t221));
    L[22] = fmaf(6.0f * m102, t223, L[22]);
    L[20] = fmaf(6.0f * m012, t223, L[20]);
    L[19] = fmaf(m003, t223, L[19]);
    const float t214 = -fmaf(c1_7, rx * t114 + ry * t204 + rz * t213, c2_7 * (t014 + t212));
    L[20] = fmaf(4.0f * m003, t214, L[20]);
    const float t160 = -fmaf(c1_7, rx * t060 + ry * t150, c2_7 * (t140));
    L[23] = fmaf(15.0f * m120, t160, L[23]);
    L[21] = fmaf(20.0f * m030, t160, L[21]);
    const float t151 = -fmaf(c1_7, rx * t051 + ry * t141 + rz * t150, c2_7 * (t131));
    L[24] = fmaf(10.0f * m120, t151, L[24]);
    L[23] = fmaf(5.0f * m111, t151, L[23]);
    L[22] = fmaf(10.0f * m030, t151, L[22]);
    L[21] = fmaf(10.0f * m021, t151, L[21]);
    const float t142 = -fmaf(c1_7, rx * t042 + ry * t132 + rz * t141, c2_7 * (t122 + t140));
    L[24] = fmaf(8.0f * m111, t142, L[24]);
    L[23] = fmaf(m102, t142, L[23]);
    L[22] = fmaf(12.0f * m021, t142, L[22]);
    L[21] = fmaf(4.0f * m012, t142, L[21]);
    const float t133 = -fmaf(c1_7, rx * t033 + ry * t123 + rz * t132, c2_7 * (t113 + t131));
    L[24] = fmaf(3.0f * m102, t133, L[24]);
    L[22] = fmaf(9.0f * m012, t133, L[22]);
    L[21] = fmaf(m003, t133, L[21]);
    const float t124 = -fmaf(c1_7, rx * t024 + ry * t114 + rz * t123, c2_7 * (t104 + t122));
    L[22] = fmaf(4.0f * m003, t124, L[22]);
    const float t070 = -fmaf(c1_7, ry * t060, c2_7 * (t050));
    L[23] = fmaf(35.0f * m030, t070, L[23]);
    const float t061 = -fmaf(c1_7, ry * t051 + rz * t060, c2_7 * (t041));
    L[24] = fmaf(20.0f * m030, t061, L[24]);
    L[23] = fmaf(15.0f * m021, t061, L[23]);
    const float t052 = -fmaf(c1_7, ry * t042 + rz * t051, c2_7 * (t032 + t050));
    L[24] = fmaf(20.0f * m021, t052, L[24]);
    L[23] = fmaf(5.0f * m012, t052, L[23]);
    const float t043 = -fmaf(c1_7, ry * t033 + rz * t042, c2_7 * (t023 + t041));
    L[24] = fmaf(12.0f * m012, t043, L[24]);
    L[23] = fmaf(m003, t043, L[23]);
    const float t034 = -fmaf(c1_7, ry * t024 + rz * t033, c2_7 * (t014 + t032));
    L[24] = fmaf(4.0f * m003, t034, L[24]);
}

// potential and field at offset y from the expansion centre (order P = 4, from the
// independent coefficients b_z <= 1; each homogeneous part of a harmonic
// expansion is harmonic, which fixes the others)
__device__ inline void l2p_p4h(const float (&Lh)[25], float x, float y, float z, float &phi, float &ex, float &ey, float &ez)
{
    float L[35];
    L[0] = Lh[0];
    L[1] = Lh[1];
    L[2] = Lh[2];
    L[3] = Lh[3];
    L[4] = Lh[4];
    L[5] = Lh[5];
    L[6] = Lh[6];
    L[7] = Lh[7];
    L[8] = Lh[8];
    L[10] = Lh[9];
    L[11] = Lh[10];
    L[12] = Lh[11];
    L[13] = Lh[12];
    L[14] = Lh[13];
    L[16] = Lh[14];
    L[17] = Lh[15];
    L[20] = Lh[16];
    L[21] = Lh[17];
    L[22] = Lh[18];
    L[23] = Lh[19];
    L[24] = Lh[20];
    L[26] = Lh[21];
    L[27] = Lh[22];
    L[30] = Lh[23];
    L[31] = Lh[24];
    L[9] = -(1.0f * L[4] + 1.0f * L[7]);
    L[15] = -(3.0f * L[10] + 1.0f * L[13]);
    L[18] = -(1.0f * L[11] + 3.0f * L[16]);
    L[25] = -(6.0f * L[20] + 1.0f * L[23]);
    L[28] = -(3.0f * L[21] + 3.0f * L[26]);
    L[32] = -(1.0f * L[23] + 6.0f * L[30]);
    L[19] = -(0.3333333333333333f * L[12] + 0.3333333333333333f * L[17]);
    L[29] = -(1.0f * L[22] + 0.3333333333333333f * L[27]);
    L[33] = -(0.3333333333333333f * L[24] + 1.0f * L[31]);
    L[34] = -(0.16666666666666666f * L[25] + 0.16666666666666666f * L[32]);
    const float x0 = 1.0f, x1 = x;
    const float x2 = x1 * x;
    const float x3 = x2 * x;
    const float x4 = x3 * x;
    const float y0 = 1.0f, y1 = y;
    const float y2 = y1 * y;
    const float y3 = y2 * y;
    const float y4 = y3 * y;
    const float z0 = 1.0f, z1 = z;
    const float z2 = z1 * z;
    const float z3 = z2 * z;
    const float z4 = z3 * z;
    phi = L[0] * (x0 * y0 * z0)
        + L[1] * (x1 * y0 * z0)
        + L[2] * (x0 * y1 * z0)
        + L[3] * (x0 * y0 * z1)
        + L[4] * (x2 * y0 * z0)
        + L[5] * (x1 * y1 * z0)
        + L[6] * (x1 * y0 * z1)
        + L[7] * (x0 * y2 * z0)
        + L[8] * (x0 * y1 * z1)
        + L[9] * (x0 * y0 * z2)
        + L[10] * (x3 * y0 * z0)
        + L[11] * (x2 * y1 * z0)
        + L[12] * (x2 * y0 * z1)
        + L[13] * (x1 * y2 * z0)
        + L[14] * (x1 * y1 * z1)
        + L[15] * (x1 * y0 * z2)
        + L[16] * (x0 * y3 * z0)
        + L[17] * (x0 * y2 * z1)
        + L[18] * (x0 * y1 * z2)
        + L[19] * (x0 * y0 * z3)
        + L[20] * (x4 * y0 * z0)
        + L[21] * (x3 * y1 * z0)
        + L[22] * (x3 * y0 * z1)
        + L[23] * (x2 * y2 * z0)
        + L[24] * (x2 * y1 * z1)
        + L[25] * (x2 * y0 * z2)
        + L[26] * (x1 * y3 * z0)
        + L[27] * (x1 * y2 * z1)
        + L[28] * (x1 * y1 * z2)
        + L[29] * (x1 * y0 * z3)
        + L[30] * (x0 * y4 * z0)
        + L[31] * (x0 * y3 * z1)
        + L[32] * (x0 * y2 * z2)
        + L[33] * (x0 * y1 * z3)
        + L[34] * (x0 * y0 * z4);
    ex = -(1.0f * L[1] * (x0 * y0 * z0)
        + 2.0f * L[4] * (x1 * y0 * z0)
        + 1.0f * L[5] * (x0 * y1 * z0)
        + 1.0f * L[6] * (x0 * y0 * z1)
        + 3.0f * L[10] * (x2 * y0 * z0)
        + 2.0f * L[11] * (x1 * y1 * z0)
        + 2.0f * L[12] * (x1 * y0 * z1)
        + 1.0f * L[13] * (x0 * y2 * z0)
        + 1.0f * L[14] * (x0 * y1 * z1)
        + 1.0f * L[15] * (x0 * y0 * z2)
        + 4.0f * L[20] * (x3 * y0 * z0)
        + 3.0f * L[21] * (x2 * y1 * z0)
        + 3.0f * L[22] * (x2 * y0 * z1)
        + 2.0f * L[23] * (x1 * y2 * z0)
        + 2.0f * L[24] * (x1 * y1 * z1)
        + 2.0f * L[25] * (x1 * y0 * z2)
        + 1.0f * L[26] * (x0 * y3 * z0)
        + 1.0f * L[27] * (x0 * y2 * z1)
        + 1.0f * L[28] * (x0 * y1 * z2)
        + 1.0f * L[29] * (x0 * y0 * z3));
    ey = -(1.0f * L[2] * (x0 * y0 * z0)
        + 1.0f * L[5] * (x1 * y0 * z0)
        + 2.0f * L[7] * (x0 * y1 * z0)
        + 1.0f * L[8] * (x0 * y0 * z1)
        + 1.0f * L[11] * (x2 * y0 * z0)
        + 2.0f * L[13] * (x1 * y1 * z0)
        + 1.0f * L[14] * (x1 * y0 * z1)
        + 3.0f * L[16] * (x0 * y2 * z0)
        + 2.0f * L[17] * (x0 * y1 * z1)
        + 1.0f * L[18] * (x0 * y0 * z2)
        + 1.0f * L[21] * (x3 * y0 * z0)
        + 2.0f * L[23] * (x2 * y1 * z0)
        + 1.0f * L[24] * (x2 * y0 * z1)
        + 3.0f * L[26] * (x1 * y2 * z0)
        + 2.0f * L[27] * (x1 * y1 * z1)
        + 1.0f * L[28] * (x1 * y0 * z2)
        + 4.0f * L[30] * (x0 * y3 * z0)
        + 3.0f * L[31] * (x0 * y2 * z1)
        + 2.0f * L[32] * (x0 * y1 * z2)
        + 1.0f * L[33] * (x0 * y0 * z3));
    ez = -(1.0f * L[3] * (x0 * y0 * z0)
        + 1.0f * L[6] * (x1 * y0 * z0)
        + 1.0f * L[8] * (x0 * y1 * z0)
        + 2.0f * L[9] * (x0 * y0 * z1)
        + 1.0f * L[12] * (x2 * y0 * z0)
        + 1.0f * L[14] * (x1 * y1 * z0)
        + 2.0f * L[15] * (x1 * y0 * z1)
        + 1.0f * L[17] * (x0 * y2 * z0)
        + 2.0f * L[18] * (x0 * y1 * z1)
        + 3.0f * L[19] * (x0 * y0 * z2)
        + 1.0f * L[22] * (x3 * y0 * z0)
        + 1.0f * L[24] * (x2 * y1 * z0)
        + 2.0f * L[25] * (x2 * y0 * z1)
        + 1.0f * L[27] * (x1 * y2 * z0)
        + 2.0f * L[28] * (x1 * y1 * z1)
        + 3.0f * L[29] * (x1 * y0 * z2)
        + 1.0f * L[31] * (x0 * y3 * z0)
        + 2.0f * L[32] * (x0 * y2 * z1)
        + 3.0f * L[33] * (x0 * y1 * z2)
        + 4.0f * L[34] * (x0 * y0 * z3));
}

// adds one cell's local coefficients about the expansion centre to L[36]
// (order P = 5; R = centre - cell centre; Q' traceless, Q'zz = -Q'xx - Q'yy)
__device__ inline void m2l_acc_p5ho(float rx, float ry, float rz, float q, float dx, float dy, float dz,
                                  float qxx, float qyy, float qxy, float qyz, float qzx,
                                  float oxxx, float oxxy, float oxxz, float oxyy, float oxyz, float oyyy, float oyyz, float (&L)[36])
{
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float mdx = -dx, mdy = -dy, mdz = -dz;
    const float mxx = qxx * (1.0f / 3.0f), myy = qyy * (1.0f / 3.0f), mzz = -(qxx + qyy) * (1.0f / 3.0f);
    const float mxy = qxy * (1.0f / 3.0f), myz = qyz * (1.0f / 3.0f), mzx = qzx * (1.0f / 3.0f);
    const float m300 = -oxxx, m210 = -oxxy, m201 = -oxxz, m120 = -oxyy, m111 = -oxyz;
    const float m030 = -oyyy, m021 = -oyyz, m102 = oxxx + oxyy, m012 = oxxy + oyyy, m003 = oxxz + oyyz;
    const float t000 = __builtin_amdgcn_rsqf(r2);
    const float ir2 = t000 * t000;
    L[0] = fmaf(q, t000, L[0]);
    const float c1_1 = 1.0f * ir2, c2_1 = 0.0f * ir2;
    const float t100 = -c1_1 * (rx * t000);
    L[1] = fmaf(q, t100, L[1]);
    L[0] = fmaf(mdx, t100, L[0]);
    const float t010 = -c1_1 * (ry * t000);
    L[2] = fmaf(q, t010, L[2]);
    L[0] = fmaf(mdy, t010, L[0]);
    const float t001 = -c1_1 * (rz * t000);
    L[3] = fmaf(q, t001, L[3]);
    L[0] = fmaf(mdz, t001, L[0]);
    const float c1_2 = 1.5f * ir2, c2_2 = 0.5f * ir2;
    const float t200 = -fmaf(c1_2, rx * t100, c2_2 * (t000));
    L[4] = fmaf(q, t200, L[4]);
    L[1] = fmaf(2.0f * mdx, t200, L[1]);
    L[0] = fmaf(mxx, t200, L[0]);
    const float t110 = -c1_2 * (rx * t010 + ry * t100);
    L[5] = fmaf(q, t110, L[5]);
    L[2] = fmaf(mdx, t110, L[2]);
    L[1] = fmaf(mdy, t110, L[1]);
    L[0] = fmaf(mxy, t110, L[0]);
    const float t101 = -c1_2 * (rx * t001 + rz * t100);
    L[6] = fmaf(q, t101, L[6]);
    L[3] = fmaf(mdx, t101, L[3]);
    L[1] = fmaf(mdz, t101, L[1]);
    L[0] = fmaf(mzx, t101, L[0]);
    const float t020 = -fmaf(c1_2, ry * t010, c2_2 * (t000));
    L[7] = fmaf(q, t020, L[7]);
    L[2] = fmaf(2.0f * mdy, t020, L[2]);
    L[0] = fmaf(myy, t020, L[0]);
    const float t011 = -c1_2 * (ry * t001 + rz * t010);
    L[8] = fmaf(q, t011, L[8]);
    L[3] = fmaf(mdy, t011, L[3]);
    L[2] = fmaf(mdz, t011, L[2]);
    L[0] = fmaf(myz, t011, L[0]);
    const float t002 = -fmaf(c1_2, rz * t001, c2_2 * (t000));
    L[3] = fmaf(2.0f * mdz, t002, L[3]);
    L[0] = fmaf(mzz, t002, L[0]);
    const float c1_3 = 1.6666666666666667f * ir2, c2_3 = 0.6666666666666666f * ir2;
    const float t300 = -fmaf(c1_3, rx * t200, c2_3 * (t100));
    L[9] = fmaf(q, t300, L[9]);
    L[4] = fmaf(3.0f * mdx, t300, L[4]);
    L[1] = fmaf(3.0f * mxx, t300, L[1]);
    L[0] = fmaf(m300, t300, L[0]);
    const float t210 = -fmaf(c1_3, rx * t110 + ry * t200, c2_3 * (t010));
    L[10] = fmaf(q, t210, L[10]);
    L[5] = fmaf(2.0f * mdx, t210, L[5]);
    L[4] = fmaf(mdy, t210, L[4]);
    L[2] = fmaf(mxx, t210, L[2]);
    L[1] = fmaf(2.0f * mxy, t210, L[1]);
    L[0] = fmaf(m210, t210, L[0]);
    const float t201 = -fmaf(c1_3, rx * t101 + rz * t200, c2_3 * (t001));
    L[11] = fmaf(q, t201, L[11]);
    L[6] = fmaf(2.0f * mdx, t201, L[6]);
    L[4] = fmaf(mdz, t201, L[4]);
    L[3] = fmaf(mxx, t201, L[3]);
    L[1] = fmaf(2.0f * mzx, t201, L[1]);
    L[0] = fmaf(m201, t201, L[0]);
    const float t120 = -fmaf(c1_3, rx * t020 + ry * t110, c2_3 * (t100));
    L[12] = fmaf(q, t120, L[12]);
    L[7] = fmaf(mdx, t120, L[7]);
    L[5] = fmaf(2.0f * mdy, t120, L[5]);
    L[1] = fmaf(myy, t120, L[1]);
    L[2] = fmaf(2.0f * mxy, t120, L[2]);
    L[0] = fmaf(m120, t120, L[0]);
    const float t111 = -c1_3 * (rx * t011 + ry * t101 + rz * t110);
    L[13] = fmaf(q, t111, L[13]);
    L[8] = fmaf(mdx, t111, L[8]);
    L[6] = fmaf(mdy, t111, L[6]);
    L[5] = fmaf(mdz, t111, L[5]);
    L[3] = fmaf(mxy, t111, L[3]);
    L[1] = fmaf(myz, t111, L[1]);
    L[2] = fmaf(mzx, t111, L[2]);
    L[0] = fmaf(m111, t111, L[0]);
    const float t102 = -fmaf(c1_3, rx * t002 + rz * t101, c2_3 * (t100));
    L[6] = fmaf(2.0f * mdz, t102, L[6]);
    L[1] = fmaf(mzz, t102, L[1]);
    L[3] = fmaf(2.0f * mzx, t102, L[3]);
    L[0] = fmaf(m102, t102, L[0]);
    const float t030 = -fmaf(c1_3, ry * t020, c2_3 * (t010));
    L[14] = fmaf(q, t030, L[14]);
    L[7] = fmaf(3.0f * mdy, t030, L[7]);
    L[2] = fmaf(3.0f * myy, t030, L[2]);
    L[0] = fmaf(m030, t030, L[0]);
    const float t021 = -fmaf(c1_3, ry * t011 + rz * t020, c2_3 * (t001));
    L[15] = fmaf(q, t021, L[15]);
    L[8] = fmaf(2.0f * mdy, t021, L[8]);
    L[7] = fmaf(mdz, t021, L[7]);
    L[3] = fmaf(myy, t021, L[3]);
    L[2] = fmaf(2.0f * myz, t021, L[2]);
    L[0] = fmaf(m021, t021, L[0]);
    const float t012 = -fmaf(c1_3, ry * t002 + rz * t011, c2_3 * (t010));
    L[8] = fmaf(2.0f * mdz, t012, L[8]);
    L[2] = fmaf(mzz, t012, L[2]);
    L[3] = fmaf(2.0f * myz, t012, L[3]);
    L[0] = fmaf(m012, t012, L[0]);
    const float t003 = -fmaf(c1_3, rz * t002, c2_3 * (t001));
    L[3] = fmaf(3.0f * mzz, t003, L[3]);
    L[0] = fmaf(m003, t003, L[0]);
    const float c1_4 = 1.75f * ir2, c2_4 = 0.75f * ir2;
    const float t400 = -fmaf(c1_4, rx * t300, c2_4 * (t200));
    L[16] = fmaf(q, t400, L[16]);
    L[9] = fmaf(4.0f * mdx, t400, L[9]);
    L[4] = fmaf(6.0f * mxx, t400, L[4]);
    L[1] = fmaf(4.0f * m300, t400, L[1]);
    const float t310 = -fmaf(c1_4, rx * t210 + ry * t300, c2_4 * (t110));
    L[17] = fmaf(q, t310, L[17]);
    L[10] = fmaf(3.0f * mdx, t310, L[10]);
    L[9] = fmaf(mdy, t310, L[9]);
    L[5] = fmaf(3.0f * mxx, t310, L[5]);
    L[4] = fmaf(3.0f * mxy, t310, L[4]);
    L[2] = fmaf(m300, t310, L[2]);
    L[1] = fmaf(3.0f * m210, t310, L[1]);
    const float t301 = -fmaf(c1_4, rx * t201 + rz * t300, c2_4 * (t101));
    L[18] = fmaf(q, t301, L[18]);
    L[11] = fmaf(3.0f * mdx, t301, L[11]);
    L[9] = fmaf(mdz, t301, L[9]);
    L[6] = fmaf(3.0f * mxx, t301, L[6]);
    L[4] = fmaf(3.0f * mzx, t301, L[4]);
    L[3] = fmaf(m300, t301, L[3]);
    L[1] = fmaf(3.0f * m201, t301, L[1]);
    const float t220 = -fmaf(c1_4, rx * t120 + ry * t210, c2_4 * (t020 + t200));
    L[19] = fmaf(q, t220, L[19]);
    L[12] = fmaf(2.0f * mdx, t220, L[12]);
    L[10] = fmaf(2.0f * mdy, t220, L[10]);
    L[7] = fmaf(mxx, t220, L[7]);
    L[4] = fmaf(myy, t220, L[4]);
    L[5] = fmaf(4.0f * mxy, t220, L[5]);
    L[2] = fmaf(2.0f * m210, t220, L[2]);
    L[1] = fmaf(2.0f * m120, t220, L[1]);
    const float t211 = -fmaf(c1_4, rx * t111 + ry * t201 + rz * t210, c2_4 * (t011));
    L[20] = fmaf(q, t211, L[20]);
    L[13] = fmaf(2.0f * mdx, t211, L[13]);
    L[11] = fmaf(mdy, t211, L[11]);
    L[10] = fmaf(mdz, t211, L[10]);
    L[8] = fmaf(mxx, t211, L[8]);
    L[6] = fmaf(2.0f * mxy, t211, L[6]);
    L[4] = fmaf(myz, t211, L[4]);
    L[5] = fmaf(2.0f * mzx, t211, L[5]);
    L[3] = fmaf(m210, t211, L[3]);
    L[2] = fmaf(m201, t211, L[2]);
    L[1] = fmaf(2.0f * m111, t211, L[1]);
    const float t202 = -fmaf(c1_4, rx * t102 + rz * t201, c2_4 * (t002 + t200));
    L[11] = fmaf(2.0f * mdz, t202, L[11]);
    L[4] = fmaf(mzz, t202, L[4]);
    L[6] = fmaf(4.0f * mzx, t202, L[6]);
    L[3] = fmaf(2.0f * m201, t202, L[3]);
    L[1] = fmaf(2.0f * m102, t202, L[1]);
    const float t130 = -fmaf(c1_4, rx * t030 + ry * t120, c2_4 * (t110));
    L[21] = fmaf(q, t130, L[21]);
    L[14] = fmaf(mdx, t130, L[14]);
    L[12] = fmaf(3.0f * mdy, t130, L[12]);
    L[5] = fmaf(3.0f * myy, t130, L[5]);
    L[7] = fmaf(3.0f * mxy, t130, L[7]);
    L[2] = fmaf(3.0f * m120, t130, L[2]);
    L[1] = fmaf(m030, t130, L[1]);
    const float t121 = -fmaf(c1_4, rx * t021 + ry * t111 + rz * t120, c2_4 * (t101));
    L[22] = fmaf(q, t121, L[22]);
    L[15] = fmaf(mdx, t121, L[15]);
    L[13] = fmaf(2.0f * mdy, t121, L[13]);
    L[12] = fmaf(mdz, t121, L[12]);
    L[6] = fmaf(myy, t121, L[6]);
    L[8] = fmaf(2.0f * mxy, t121, L[8]);
    L[5] = fmaf(2.0f * myz, t121, L[5]);
    L[7] = fmaf(mzx, t121, L[7]);
    L[3] = fmaf(m120, t121, L[3]);
    L[2] = fmaf(2.0f * m111, t121, L[2]);
    L[1] = fmaf(m021, t121, L[1]);
    const float t112 = -fmaf(c1_4, rx * t012 + ry * t102 + rz * t111, c2_4 * (t110));
    L[13] = fmaf(2.0f * mdz, t112, L[13]);
    L[5] = fmaf(mzz, t112, L[5]);
    L[6] = fmaf(2.0f * myz, t112, L[6]);
    L[8] = fmaf(2.0f * mzx, t112, L[8]);
    L[3] = fmaf(2.0f * m111, t112, L[3]);
    L[2] = fmaf(m102, t112, L[2]);
    L[1] = fmaf(m012, t112, L[1]);
    const float t103 = -fmaf(c1_4, rx * t003 + rz * t102, c2_4 * (t101));
    L[6] = fmaf(3.0f * mzz, t103, L[6]);
    L[3] = fmaf(3.0f * m102, t103, L[3]);
    L[1] = fmaf(m003, t103, L[1]);
    const float t040 = -fmaf(c1_4, ry * t030, c2_4 * (t020));
    L[23] = fmaf(q, t040, L[23]);
    L[14] = fmaf(4.0f * mdy, t040, L[14]);
    L[7] = fmaf(6.0f * myy, t040, L[7]);
    L[2] = fmaf(4.0f * m030, t040, L[2]);
    const float t031 = -fmaf(c1_4, ry * t021 + rz * t030, c2_4 * (t011));
    L[24] = fmaf(q, t031, L[24]);
    L[15] = fmaf(3.0f * mdy, t031, L[15]);
    L[14] = fmaf(mdz, t031, L[14]);
    L[8] = fmaf(3.0f * myy, t031, L[8]);
    L[7] = fmaf(3.0f * myz, t031, L[7]);
    L[3] = fmaf(m030, t031, L[3]);
    L[2] = fmaf(3.0f * m021, t031, L[2]);
    const float t022 = -fmaf(c1_4, ry * t012 + rz * t021, c2_4 * (t002 + t020));
    L[15] = fmaf(2.0f * mdz, t022, L[15]);
    L[7] = fmaf(mzz, t022, L[7]);
    L[8] = fmaf(4.0f * myz, t022, L[8]);
    L[3] = fmaf(2.0f * m021, t022, L[3]);
    L[2] = fmaf(2.0f * m012, t022, L[2]);
    const float t013 = -fmaf(c1_4, ry * t003 + rz * t012, c2_4 * (t011));
    L[8] = fmaf(3.0f * mzz, t013, L[8]);
    L[3] = fmaf(3.0f * m012, t013, L[3]);
    L[2] = fmaf(m003, t013, L[2]);
    const float t004 = -fmaf(c1_4, rz * t003, c2_4 * (t002));
    L[3] = fmaf(4.0f * m003, t004, L[3]);
    const float c1_5 = 1.8f * ir2, c2_5 = 0.8f * ir2;
    const float t500 = -fmaf(c1_5, rx * t400, c2_5 * (t300));
    L[25] = fmaf(q, t500, L[25]);
    L[16] = fmaf(5.0f * mdx, t500, L[16]);
    L[9] = fmaf(10.0f * mxx, t500, L[9]);
    L[4] = fmaf(10.0f * m300, t500, L[4]);
    const float t410 = -fmaf(c1_5, rx * t310 + ry * t400, c2_5 * (t210));
    L[26] = fmaf(q, t410, L[26]);
    L[17] = fmaf(4.0f * mdx, t410, L[17]);
    L[16] = fmaf(mdy, t410, L[16]);
    L[10] = fmaf(6.0f * mxx, t410, L[10]);
    L[9] = fmaf(4.0f * mxy, t410, L[9]);
    L[5] = fmaf(4.0f * m300, t410, L[5]);
    L[4] = fmaf(6.0f * m210, t410, L[4]);
    const float t401 = -fmaf(c1_5, rx * t301 + rz * t400, c2_5 * (t201));
    L[27] = fmaf(q, t401, L[27]);
    L[18] = fmaf(4.0f * mdx, t401, L[18]);
    L[16] = fmaf(mdz, t401, L[16]);
    L[11] = fmaf(6.0f * mxx, t401, L[11]);
    L[9] = fmaf(4.0f * mzx, t401, L[9]);
    L[6] = fmaf(4.0f * m300, t401, L[6]);
    L[4] = fmaf(6.0f * m201, t401, L[4]);
    const float t320 = -fmaf(c1_5, rx * t220 + ry * t310, c2_5 * (t120 + t300));
    L[28] = fmaf(q, t320, L[28]);
    L[19] = fmaf(3.0f * mdx, t320, L[19]);
    L[17] = fmaf(2.0f * mdy, t320, L[17]);
    L[12] = fmaf(3.0f * mxx, t320, L[12]);
    L[9] = fmaf(myy, t320, L[9]);
    L[10] = fmaf(6.0f * mxy, t320, L[10]);
    L[7] = fmaf(m300, t320, L[7]);
    L[5] = fmaf(6.0f * m210, t320, L[5]);
    L[4] = fmaf(3.0f * m120, t320, L[4]);
    const float t311 = -fmaf(c1_5, rx * t211 + ry * t301 + rz * t310, c2_5 * (t111));
    L[29] = fmaf(q, t311, L[29]);
    L[20] = fmaf(3.0f * mdx, t311, L[20]);
    L[18] = fmaf(mdy, t311, L[18]);
    L[17] = fmaf(mdz, t311, L[17]);
    L[13] = fmaf(3.0f * mxx, t311, L[13]);
    L[11] = fmaf(3.0f * mxy, t311, L[11]);
    L[9] = fmaf(myz, t311, L[9]);
    L[10] = fmaf(3.0f * mzx, t311, L[10]);
    L[8] = fmaf(m300, t311, L[8]);
    L[6] = fmaf(3.0f * m210, t311, L[6]);
    L[5] = fmaf(3.0f * m201, t311, L[5]);
    L[4] = fmaf(3.0f * m111, t311, L[4]);
    const float t302 = -fmaf(c1_5, rx * t202 + rz * t301, c2_5 * (t102 + t300));
    L[18] = fmaf(2.0f * mdz, t302, L[18]);
    L[9] = fmaf(mzz, t302, L[9]);
    L[11] = fmaf(6.0f * mzx, t302, L[11]);
    L[6] = fmaf(6.0f * m201, t302, L[6]);
    L[4] = fmaf(3.0f * m102, t302, L[4]);
    const float t230 = -fmaf(c1_5, rx * t130 + ry * t220, c2_5 * (t030 + t210));
    L[30] = fmaf(q, t230, L[30]);
    L[21] = fmaf(2.0f * mdx, t230, L[21]);
    L[19] = fmaf(3.0f * mdy, t230, L[19]);
    L[14] = fmaf(mxx, t230, L[14]);
    L[10] = fmaf(3.0f * myy, t230, L[10]);
    L[12] = fmaf(6.0f * mxy, t230, L[12]);
    L[7] = fmaf(3.0f * m210, t230, L[7]);
    L[5] = fmaf(6.0f * m120, t230, L[5]);
    L[4] = fmaf(m030, t230, L[4]);
    const float t221 = -fmaf(c1_5, rx * t121 + ry * t211 + rz * t220, c2_5 * (t021 + t201));
    L[31] = fmaf(q, t221, L[31]);
    L[22] = fmaf(2.0f * mdx, t221, L[22]);
    L[20] = fmaf(2.0f * mdy, t221, L[20]);
    L[19] = fmaf(mdz, t221, L[19]);
    L[15] = fmaf(mxx, t221, L[15]);
    L[11] = fmaf(myy, t221, L[11]);
    L[13] = fmaf(4.0f * mxy, t221, L[13]);
    L[10] = fmaf(2.0f * myz, t221, L[10]);
    L[12] = fmaf(2.0f * mzx, t221, L[12]);
    L[8] = fmaf(2.0f * m210, t221, L[8]);
    L[7] = fmaf(m201, t221, L[7]);
    L[6] = fmaf(2.0f * m120, t221, L[6]);
    L[5] = fmaf(4.0f * m111, t221, L[5]);
    L[4] = fmaf(m021, t221, L[4]);
    const float t212 = -fmaf(c1_5, rx * t112 + ry * t202 + rz * t211, c2_5 * (t012 + t210));
    L[20] = fmaf(2.0f * mdz, t212, L[20]);
    L[10] = fmaf(mzz, t212, L[10]);
    L[11] = fmaf(2.0f * myz, t212, L[11]);
    L[13] = fmaf(4.0f * mzx, t212, L[13]);
    L[8] = fmaf(2.0f * m201, t212, L[8]);
    L[6] = fmaf(4.0f * m111, t212, L[6]);
    L[5] = fmaf(2.0f * m102, t212, L[5]);
    L[4] = fmaf(m012, t212, L[4]);
    const float t203 = -fmaf(c1_5, rx * t103 + rz * t202, c2_5 * (t003 + t201));
    L[11] = fmaf(3.0f * mzz, t203, L[11]);
    L[6] = fmaf(6.0f * m102, t203, L[6]);
    L[4] = fmaf(m003, t203, L[4]);
    const float t140 = -fmaf(c1_5, rx * t040 + ry * t130, c2_5 * (t120));
    L[32] = fmaf(q, t140, L[32]);
    L[23] = fmaf(mdx, t140, L[23]);
    L[21] = fmaf(4.0f * mdy, t140, L[21]);
    L[12] = fmaf(6.0f * myy, t140, L[12]);
    L[14] = fmaf(4.0f * mxy, t140, L[14]);
    L[7] = fmaf(6.0f * m120, t140, L[7]);
    L[5] = fmaf(4.0f * m030, t140, L[5]);
    const float t131 = -fmaf(c1_5, rx * t031 + ry * t121 + rz * t130, c2_5 * (t111));
    L[33] = fmaf(q, t131, L[33]);
    L[24] = fmaf(mdx, t131, L[24]);
    L[22] = fmaf(3.0f * mdy, t131, L[22]);
    L[21] = fmaf(mdz, t131, L[21]);
    L[13] = fmaf(3.0f * myy, t131, L[13]);
    L[15] = fmaf(3.0f * mxy, t131, L[15]);
    L[12] = fmaf(3.0f * myz, t131, L[12]);
    L[14] = fmaf(mzx, t131, L[14]);
    L[8] = fmaf(3.0f * m120, t131, L[8]);
    L[7] = fmaf(3.0f * m111, t131, L[7]);
    L[6] = fmaf(m030, t131, L[6]);
    L[5] = fmaf(3.0f * m021, t131, L[5]);
    const float t122 = -fmaf(c1_5, rx * t022 + ry * t112 + rz * t121, c2_5 * (t102 + t120));
    L[22] = fmaf(2.0f * mdz, t122, L[22]);
    L[12] = fmaf(mzz, t122, L[12]);
    L[13] = fmaf(4.0f * myz, t122, L[13]);
    L[15] = fmaf(2.0f * mzx, t122, L[15]);
    L[8] = fmaf(4.0f * m111, t122, L[8]);
    L[7] = fmaf(m102, t122, L[7]);
    L[6] = fmaf(2.0f * m021, t122, L[6]);
    L[5] = fmaf(2.0f * m012, t122, L[5]);
    const float t113 = -fmaf(c1_5, rx * t013 + ry * t103 + rz * t112, c2_5 * (t111));
    L[13] = fmaf(3.0f * mzz, t113, L[13]);
    L[8] = fmaf(3.0f * m102, t113, L[8]);
    L[6] = fmaf(3.0f * m012, t113, L[6]);
    L[5] = fmaf(m003, t113, L[5]);
    const float t104 = -fmaf(c1_5, rx * t004 + rz * t103, c2_5 * (t102));
    L[6] = fmaf(4.0f * m003, t104, L[6]);
    const float t050 = -fmaf(c1_5, ry * t040, c2_5 * (t030));
    L[34] = fmaf(q, t050, L[34]);
    L[23] = fmaf(5.0f * mdy, t050, L[23]);
    L[14] = fmaf(10.0f * myy, t050, L[14]);
    L[7] = fmaf(10.0f * m030, t050, L[7]);
    const float t041 = -fmaf(c1_5, ry * t031 + rz * t040, c2_5 * (t021));
    L[35] = fmaf(q, t041, L[35]);
    L[24] = fmaf(4.0f * mdy, t041, L[24]);
    L[23] = fmaf(mdz, t041, L[23]);
    L[15] = fmaf(6.0f * myy, t041, L[15]);
    L[14] = fmaf(4.0f * myz, t041, L[14]);
    L[8] = fmaf(4.0f * m030, t041, L[8]);
    L[7] = fmaf(6.0f * m021, t041, L[7]);
    const float t032 = -fmaf(c1_5, ry * t022 + rz * t031, c2_5 * (t012 + t030));
    L[24] = fmaf(2.0f * mdz, t032, L[24]);
    L[14] = fmaf(mzz, t032, L[14]);
    L[15] = fmaf(6.0f * myz, t032, L[15]);
    L[8] = fmaf(6.0f * m021, t032, L[8]);
    L[7] = fmaf(3.0f * m012, t032, L[7]);
    const float t023 = -fmaf(c1_5, ry * t013 + rz * t022, c2_5 * (t003 + t021));
    L[15] = fmaf(3.0f * mzz, t023, L[15]);
    L[8] = fmaf(6.0f * m012, t023, L[8]);
    L[7] = fmaf(m003, t023, L[7]);
    const float t014 = -fmaf(c1_5, ry * t004 + rz * t013, c2_5 * (t012));
    L[8] = fmaf(4.0f * m003, t014, L[8]);
    const float c1_6 = 1.8333333333333333f * ir2, c2_6 = 0.8333333333333334f * ir2;
    const float t600 = -fmaf(c1_6, rx * t500, c2_6 * (t400));
    L[25] = fmaf(6.0f * mdx, t600, L[25]);
    L[16] = fmaf(15.0f * mxx, t600, L[16]);
    L[9] = fmaf(20.0f * m300, t600, L[9]);
    const float t510 = -fmaf(c1_6, rx * t410 + ry * t500, c2_6 * (t310));
    L[26] = fmaf(5.0f * mdx, t510, L[26]);
    L[25] = fmaf(mdy, t510, L[25]);
    L[17] = fmaf(10.0f * mxx, t510, L[17]);
    L[16] = fmaf(5.0f * mxy, t510, L[16]);
    L[10] = fmaf(10.0f * m300, t510, L[10]);
    L[9] = fmaf(10.0f * m210, t510, L[9]);
    const float t501 = -fmaf(c1_6, rx * t401 + rz * t500, c2_6 * (t301));
    L[27] = fmaf(5.0f * mdx, t501, L[27]);
    L[25] = fmaf(mdz, t501, L[25]);
    L[18] = fmaf(10.0f * mxx, t501, L[18]);
    L[16] = fmaf(5.0f * mzx, t501, L[16]);
    L[11] = fmaf(10.0f * m300, t501, L[11]);
    L[9] = fmaf(10.0f * m201, t501, L[9]);
    const float t420 = -fmaf(c1_6, rx * t320 + ry * t410, c2_6 * (t220 + t400));
    L[28] = fmaf(4.0f * mdx, t420, L[28]);
    L[26] = fmaf(2.0f * mdy, t420, L[26]);
    L[19] = fmaf(6.0f * mxx, t420, L[19]);
    L[16] = fmaf(myy, t420, L[16]);
    L[17] = fmaf(8.0f * mxy, t420, L[17]);
    L[12] = fmaf(4.0f * m300, t420, L[12]);
    L[10] = fmaf(12.0f * m210, t420, L[10]);
    L[9] = fmaf(4.0f * m120, t420, L[9]);
    const float t411 = -fmaf(c1_6, rx * t311 + ry * t401 + rz * t410, c2_6 * (t211));
    L[29] = fmaf(4.0f * mdx, t411, L[29]);
    L[27] = fmaf(mdy, t411, L[27]);
    L[26] = fmaf(mdz, t411, L[26]);
    L[20] = fmaf(6.0f * mxx, t411, L[20]);
    L[18] = fmaf(4.0f * mxy, t411, L[18]);
    L[16] = fmaf(myz, t411, L[16]);
    L[17] = fmaf(4.0f * mzx, t411, L[17]);
    L[13] = fmaf(4.0f * m300, t411, L[13]);
    L[11] = fmaf(6.0f * m210, t411, L[11]);
    L[10] = fmaf(6.0f * m201, t411, L[10]);
    L[9] = fmaf(4.0f * m111, t411, L[9]);
    const float t402 = -fmaf(c1_6, rx * t302 + rz * t401, c2_6 * (t202 + t400));
    L[27] = fmaf(2.0f * mdz, t402, L[27]);
    L[16] = fmaf(mzz, t402, L[16]);
    L[18] = fmaf(8.0f * mzx, t402, L[18]);
    L[11] = fmaf(12.0f * m201, t402, L[11]);
    L[9] = fmaf(4.0f * m102, t402, L[9]);
    const float t330 = -fmaf(c1_6, rx * t230 + ry * t320, c2_6 * (t130 + t310));
    L[30] = fmaf(3.0f * mdx, t330, L[30]);
    L[28] = fmaf(3.0f * mdy, t330, L[28]);
    L[21] = fmaf(3.0f * mxx, t330, L[21]);
    L[17] = fmaf(3.0f * myy, t330, L[17]);
    L[19] = fmaf(9.0f * mxy, t330, L[19]);
    L[14] = fmaf(m300, t330, L[14]);
    L[12] = fmaf(9.0f * m210, t330, L[12]);
    L[10] = fmaf(9.0f * m120, t330, L[10]);
    L[9] = fmaf(m030, t330, L[9]);
    const float t321 = -fmaf(c1_6, rx * t221 + ry * t311 + rz * t320, c2_6 * (t121 + t301));
    L[31] = fmaf(3.0f * mdx, t321, L[31]);
    L[29] = fmaf(2.0f * mdy, t321, L[29]);
    L[28] = fmaf(mdz, t321, L[28]);
    L[22] = fmaf(3.0f * mxx, t321, L[22]);
    L[18] = fmaf(myy, t321, L[18]);
    L[20] = fmaf(6.0f * mxy, t321, L[20]);
    L[17] = fmaf(2.0f * myz, t321, L[17]);
    L[19] = fmaf(3.0f * mzx, t321, L[19]);
    L[15] = fmaf(m300, t321, L[15]);
    L[13] = fmaf(6.0f * m210, t321, L[13]);
    L[12] = fmaf(3.0f * m201, t321, L[12]);
    L[11] = fmaf(3.0f * m120, t321, L[11]);
    L[10] = fmaf(6.0f * m111, t321, L[10]);
    L[9] = fmaf(m021, t321, L[9]);
    const float t312 = -fmaf(c1_6, rx * t212 + ry * t302 + rz * t311, c2_6 * (t112 + t310));
    L[29] = fmaf(2.0f * mdz, t312, L[29]);
    L[17] = fmaf(mzz, t312, L[17]);
    L[18] = fmaf(2.0f * myz, t312, L[18]);
    L[20] = fmaf(6.0f * mzx, t312, L[20]);
    L[13] = fmaf(6.0f * m201, t312, L[13]);
    L[11] = fmaf(6.0f * m111, t312, L[11]);
    L[10] = fmaf(3.0f * m102, t312, L[10]);
    L[9] = fmaf(m012, t312, L[9]);
    const float t303 = -fmaf(c1_6, rx * t203 + rz * t302, c2_6 * (t103 + t301));
    L[18] = fmaf(3.0f * mzz, t303, L[18]);
    L[11] = fmaf(9.0f * m102, t303, L[11]);
    L[9] = fmaf(m003, t303, L[9]);
    const float t240 = -fmaf(c1_6, rx * t140 + ry * t230, c2_6 * (t040 + t220));
    L[32] = fmaf(2.0f * mdx, t240, L[32]);
    L[30] = fmaf(4.0f * mdy, t240, L[30]);
    L[23] = fmaf(mxx, t240, L[23]);
    L[19] = fmaf(6.0f * myy, t240, L[19]);
    L[21] = fmaf(8.0f * mxy, t240, L[21]);
    L[14] = fmaf(4.0f * m210, t240, L[14]);
    L[12] = fmaf(12.0f * m120, t240, L[12]);
    L[10] = fmaf(4.0f * m030, t240, L[10]);
    const float t231 = -fmaf(c1_6, rx * t131 + ry * t221 + rz * t230, c2_6 * (t031 + t211));
    L[33] = fmaf(2.0f * mdx, t231, L[33]);
    L[31] = fmaf(3.0f * mdy, t231, L[31]);
    L[30] = fmaf(mdz, t231, L[30]);
    L[24] = fmaf(mxx, t231, L[24]);
    L[20] = fmaf(3.0f * myy, t231, L[20]);
    L[22] = fmaf(6.0f * mxy, t231, L[22]);
    L[19] = fmaf(3.0f * myz, t231, L[19]);
    L[21] = fmaf(2.0f * mzx, t231, L[21]);
    L[15] = fmaf(3.0f * m210, t231, L[15]);
    L[14] = fmaf(m201, t231, L[14]);
    L[13] = fmaf(6.0f * m120, t231, L[13]);
    L[12] = fmaf(6.0f * m111, t231, L[12]);
    L[11] = fmaf(m030, t231, L[11]);
    L[10] = fmaf(3.0f * m021, t231, L[10]);
    const float t222 = -fmaf(c1_6, rx * t122 + ry * t212 + rz * t221, c2_6 * (t022 + t202 + t220));
    L[31] = fmaf(2.0f * mdz, t222, L[31]);
    L[19] = fmaf(mzz, t222, L[19]);
    L[20] = fmaf(4.0f * myz, t222, L[20]);
    L[22] = fmaf(4.0f * mzx, t222, L[22]);
    L[15] = fmaf(2.0f * m201, t222, L[15]);
    L[13] = fmaf(8.0f * m111, t222, L[13]);
    L[12] = fmaf(2.0f * m102, t222, L[12]);
    L[11] = fmaf(2.0f * m021, t222, L[11]);
    L[10] = fmaf(2.0f * m012, t222, L[10]);
    const float t213 = -fmaf(c1_6, rx * t113 + ry * t203 + rz * t212, c2_6 * (t013 + t211));
    L[20] = fmaf(3.0f * mzz, t213, L[20]);
    L[13] = fmaf(6.0f * m102, t213, L[13]);
    L[11] = fmaf(3.0f * m012, t213, L[11]);
    L[10] = fmaf(m003, t213, L[10]);
    const float t204 = -fmaf(c1_6, rx * t104 + rz * t203, c2_6 * (t004 + t202));
    L[11] = fmaf(4.0f * m003, t204, L[11]);
    const float t150 = -fmaf(c1_6, rx * t050 + ry * t140, c2_6 * (t130));
    L[34] = fmaf(mdx, t150, L[34]);
    L[32] = fmaf(5.0f * mdy, t150, L[32]);
    L[21] = fmaf(10.0f * myy, t150, L[21]);
    L[23] = fmaf(5.0f * mxy, t150, L[23]);
    L[14] = fmaf(10.0f * m120, t150, L[14]);
    L[12] = fmaf(10.0f * m030, t150, L[12]);
    const float t141 = -fmaf(c1_6, rx * t041 + ry * t131 + rz * t140, c2_6 * (t121));
    L[35] = fmaf(mdx, t141, L[35]);
    L[33] = fmaf(4.0f * mdy, t141, L[33]);
    L[32] = fmaf(mdz, t141, L[32]);
    L[22] = fmaf(6.0f * myy, t141, L[22]);
    L[24] = fmaf(4.0f * mxy, t141, L[24]);
    L[21] = fmaf(4.0f * myz, t141, L[21]);
    L[23] = fmaf(mzx, t141, L[23]);
    L[15] = fmaf(6.0f * m120, t141, L[15]);
    L[14] = fmaf(4.0f * m111, t141, L[14]);
    L[13] = fmaf(4.0f * m030, t141, L[13]);
    L[12] = fmaf(6.0f * m021, t141, L[12]);
    const float t132 = -fmaf(c1_6, rx * t032 + ry * t122 + rz * t131, c2_6 * (t112 + t130));
    L[33] = fmaf(2.0f * mdz, t132, L[33]);
    L[21] = fmaf(mzz, t132, L[21]);
    L[22] = fmaf(6.0f * myz, t132, L[22]);
    L[24] = fmaf(2.0f * mzx, t132, L[24]);
    L[15] = fmaf(6.0f * m111, t132, L[15]);
    L[14] = fmaf(m102, t132, L[14]);
    L[13] = fmaf(6.0f * m021, t132, L[13]);
    L[12] = fmaf(3.0f * m012, t132, L[12]);
    const float t123 = -fmaf(c1_6, rx * t023 + ry * t113 + rz * t122, c2_6 * (t103 + t121));
    L[22] = fmaf(3.0f * mzz, t123, L[22]);
    L[15] = fmaf(3.0f * m102, t123, L[15]);
    L[13] = fmaf(6.0f * m012, t123, L[13]);
    L[12] = fmaf(m003, t123, L[12]);
    const float t114 = -fmaf(c1_6, rx * t014 + ry * t104 + rz * t113, c2_6 * (t112));
    L[13] = fmaf(4.0f * m003, t114, L[13]);
    const float t060 = -fmaf(c1_6, ry * t050, c2_6 * (t040));
    L[34] = fmaf(6.0f * mdy, t060, L[34]);
    L[23] = fmaf(15.0f * myy, t060, L[23]);
    L[14] = fmaf(20.0f * m030, t060, L[14]);
    const float t051 = -fmaf(c1_6, ry * t041 + rz * t050, c2_6 * (t031));
    L[35] = fmaf(5.0f * mdy, t051, L[35]);
    L[34] = fmaf(mdz, t051, L[34]);
    L[24] = fmaf(10.0f * myy, t051, L[24]);
    L[23] = fmaf(5.0f * myz, t051, L[23]);
    L[15] = fmaf(10.0f * m030, t051, L[15]);
    L[14] = fmaf(10.0f * m021, t051, L[14]);
    const float t042 = -fmaf(c1_6, ry * t032 + rz * t041, c2_6 * (t022 + t040));
    L[35] = fmaf(2.0f * mdz, t042, L[35]);
    L[23] = fmaf(mzz, t042, L[23]);
    L[24] = fmaf(8.0f * myz, t042, L[24]);
    L[15] = fmaf(12.0f * m021, t042, L[15]);
    L[14] = fmaf(4.0f * m012, t042, L[14]);
    const float t033 = -fmaf(c1_6, ry * t023 + rz * t032, c2_6 * (t013 + t031));
    L[24] = fmaf(3.0f * mzz, t033, L[24]);
    L[15] = fmaf(9.0f * m012, t033, L[15]);
    L[14] = fmaf(m003, t033, L[14]);
    const float t024 = -fmaf(c1_6, ry * t014 + rz * t023, c2_6 * (t004 + t022));
    L[15] = fmaf(4.0f * m003, t024, L[15]);
    const float c1_7 = 1.8571428571428572f * ir2, c2_7 = 0.8571428571428571f * ir2;
    const float t700 = -fmaf(c1_7, rx * t600, c2_7 * (t500));
    L[25] = fmaf(21.0f * mxx, t700, L[25]);
    L[16] = fmaf(35.0f * m300, t700, L[16]);
    const float t610 = -fmaf(c1_7, rx * t510 + ry * t600, c2_7 * (t410));
    L[26] = fmaf(15.0f * mxx, t610, L[26]);
    L[25] = fmaf(6.0f * mxy, t610, L[25]);
    L[17] = fmaf(20.0f * m300, t610, L[17]);
    L[16] = fmaf(15.0f * m210, t610, L[16]);
    const float t601 = -fmaf(c1_7, rx * t501 + rz * t600, c2_7 * (t401));
    L[27] = fmaf(15.0f * mxx, t601, L[27]);
    L[25] = fmaf(6.0f * mzx, t601, L[25]);
    L[18] = fmaf(20.0f * m300, t601, L[18]);
    L[16] = fmaf(15.0f * m201, t601, L[16]);
    const float t520 = -fmaf(c1_7, rx * t420 + ry * t510, c2_7 * (t320 + t500));
    L[28] = fmaf(10.0f * mxx, t520, L[28]);
    L[25] = fmaf(myy, t520, L[25]);
    L[26] = fmaf(10.0f * mxy, t520, L[26]);
    L[19] = fmaf(10.0f * m300, t520, L[19]);
    L[17] = fmaf(20.0f * m210, t520, L[17]);
    L[16] = fmaf(5.0f * m120, t520, L[16]);
    const float t511 = -fmaf(c1_7, rx * t411 + ry * t501 + rz * t510, c2_7 * (t311));
    L[29] = fmaf(10.0f * mxx, t511, L[29]);
    L[27] = fmaf(5.0f * mxy, t511, L[27]);
    L[25] = fmaf(myz, t511, L[25]);
    L[26] = fmaf(5.0f * mzx, t511, L[26]);
    L[20] = fmaf(10.0f * m300, t511, L[20]);
    L[18] = fmaf(10.0f * m210, t511, L[18]);
    L[17] = fmaf(10.0f * m201, t511, L[17]);
    L[16] = fmaf(5.0f * m111, t511, L[16]);
    const float t502 = -fmaf(c1_7, rx * t402 + rz * t501, c2_7 * (t302 + t500));
    L[25] = fmaf(mzz, t502, L[25]);
    L[27] = fmaf(10.0f * mzx, t502, L[27]);
    L[18] = fmaf(20.0f * m201, t502, L[18]);
    L[16] = fmaf(5.0f * m102, t502, L[16]);
    const float t430 = -fmaf(c1_7, rx * t330 + ry * t420, c2_7 * (t230 + t410));
    L[30] = fmaf(6.0f * mxx, t430, L[30]);
    L[26] = fmaf(3.0f * myy, t430, L[26]);
    L[28] = fmaf(12.0f * mxy, t430, L[28]);
    L[21] = fmaf(4.0f * m300, t430, L[21]);
    L[19] = fmaf(18.0f * m210, t430, L[19]);
    L[17] = fmaf(12.0f * m120, t430, L[17]);
    L[16] = fmaf(m030, t430, L[16]);
    const float t421 = -fmaf(c1_7, rx * t321 + ry * t411 + rz * t420, c2_7 * (t221 + t401));
    L[31] = fmaf(6.0f * mxx, t421, L[31]);
    L[27] = fmaf(myy, t421, L[27]);
    L[29] = fmaf(8.0f * mxy, t421, L[29]);
    L[26] = fmaf(2.0f * myz, t421, L[26]);
    L[28] = fmaf(4.0f * mzx, t421, L[28]);
    L[22] = fmaf(4.0f * m300, t421, L[22]);
    L[20] = fmaf(12.0f * m210, t421, L[20]);
    L[19] = fmaf(6.0f * m201, t421, L[19]);
    L[18] = fmaf(4.0f * m120, t421, L[18]);
    L[17] = fmaf(8.0f * m111, t421, L[17]);
    L[16] = fmaf(m021, t421, L[16]);
    const float t412 = -fmaf(c1_7, rx * t312 + ry * t402 + rz * t411, c2_7 * (t212 + t410));
    L[26] = fmaf(mzz, t412, L[26]);
    L[27] = fmaf(2.0f * myz, t412, L[27]);
    L[29] = fmaf(8.0f * mzx, t412, L[29]);
    L[20] = fmaf(12.0f * m201, t412, L[20]);
    L[18] = fmaf(8.0f * m111, t412, L[18]);
    L[17] = fmaf(4.0f * m102, t412, L[17]);
    L[16] = fmaf(m012, t412, L[16]);
    const float t403 = -fmaf(c1_7, rx * t303 + rz * t402, c2_7 * (t203 + t401));
    L[27] = fmaf(3.0f * mzz, t403, L[27]);
    L[18] = fmaf(12.0f * m102, t403, L[18]);
    L[16] = fmaf(m003, t403, L[16]);
    const float t340 = -fmaf(c1_7, rx * t240 + ry * t330, c2_7 * (t140 + t320));
    L[32] = fmaf(3.0f * mxx, t340, L[32]);
    L[28] = fmaf(6.0f * myy, t340, L[28]);
    L[30] = fmaf(12.0f * mxy, t340, L[30]);
    L[23] = fmaf(m300, t340, L[23]);
    L[21] = fmaf(12.0f * m210, t340, L[21]);
    L[19] = fmaf(18.0f * m120, t340, L[19]);
    L[17] = fmaf(4.0f * m030, t340, L[17]);
    const float t331 = -fmaf(c1_7, rx * t231 + ry * t321 + rz * t330, c2_7 * (t131 + t311));
    L[33] = fmaf(3.0f * mxx, t331, L[33]);
    L[29] = fmaf(3.0f * myy, t331, L[29]);
    L[31] = fmaf(9.0f * mxy, t331, L[31]);
    L[28] = fmaf(3.0f * myz, t331, L[28]);
    L[30] = fmaf(3.0f * mzx, t331, L[30]);
    L[24] = fmaf(m300, t331, L[24]);
    L[22] = fmaf(9.0f * m210, t331, L[22]);
    L[21] = fmaf(3.0f * m201, t331, L[21]);
    L[20] = fmaf(9.0f * m120, t331, L[20]);
    L[19] = fmaf(9.0f * m111, t331, L[19]);
    L[18] = fmaf(m030, t331, L[18]);
    L[17] = fmaf(3.0f * m021, t331, L[17]);
    const float t322 = -fmaf(c1_7, rx * t222 + ry * t312 + rz * t321, c2_7 * (t122 + t302 + t320));
    L[28] = fmaf(mzz, t322, L[28]);
    L[29] = fmaf(4.0f * myz, t322, L[29]);
    L[31] = fmaf(6.0f * mzx, t322, L[31]);
    L[22] = fmaf(6.0f * m201, t322, L[22]);
    L[20] = fmaf(12.0f * m111, t322, L[20]);
    L[19] = fmaf(3.0f * m102, t322, L[19]);
    L[18] = fmaf(2.0f * m021, t322, L[18]);
    L[17] = fmaf(2.0f * m012, t322, L[17]);
    const float t313 = -fmaf(c1_7, rx * t213 + ry * t303 + rz * t312, c2_7 * (t113 + t311));
    L[29] = fmaf(3.0f * mzz, t313, L[29]);
    L[20] = fmaf(9.0f * m102, t313, L[20]);
    L[18] = fmaf(3.0f * m012, t313, L[18]);
    L[17] = fmaf(m003, t313, L[17]);
    const float t304 = -fmaf(c1_7, rx * t204 + rz * t303, c2_7 * (t104 + t302));
    L[18] = fmaf(4.0f * m003, t304, L[18]);
    const float t250 = -fmaf(c1_7, rx * t150 + ry * t240, c2_7 * (t050 + t230));
    L[34] = fmaf(mxx, t250, L[34]);
    L[30] = fmaf(10.0f * myy, t250, L[30]);
    L[32] = fmaf(10.0f * mxy, t250, L[32]);
    L[23] = fmaf(5.0f * m210, t250, L[23]);
    L[21] = fmaf(20.0f * m120, t250, L[21]);
    L[19] = fmaf(10.0f * m030, t250, L[19]);
    const float t241 = -fmaf(c1_7, rx * t141 + ry * t231 + rz * t240, c2_7 * (t041 + t221));
    L[35] = fmaf(mxx, t241, L[35]);
    L[31] = fmaf(6.0f * myy, t241, L[31]);
    L[33] = fmaf(8.0f * mxy, t241, L[33]);
    L[30] = fmaf(4.0f * myz, t241, L[30]);
    L[32] = fmaf(2.0f * mzx, t241, L[32]);
    L[24] = fmaf(4.0f * m210, t241, L[24]);
    L[23] = fmaf(m201, t241, L[23]);
    L[22] = fmaf(12.0f * m120, t241, L[22]);
    L[21] = fmaf(8.0f * m111, t241, L[21]);
    L[20] = fmaf(4.0f * m030, t241, L[20]);
    L[19] = fmaf(6.0f * m021, t241, L[19]);
    const float t232 = -fmaf(c1_7, rx * t132 + ry * t222 + rz * t231, c2_7 * (t032 + t212 + t230));
    L[30] = fmaf(mzz, t232, L[30]);
    L[31] = fmaf(6.0f * myz, t232, L[31]);
    L[33] = fmaf(4.0f * mzx, t232, L[33]);
    L[24] = fmaf(2.0f * m201, t232, L[24]);
    L[22] = fmaf(12.0f * m111, t232, L[22]);
    L[21] = fmaf(2.0f * m102, t232, L[21]);
    L[20] = fmaf(6.0f * m021, t232, L[20]);
    L[19] = fmaf(3.0f * m012, t232, L[19]);
    const float t223 = -fmaf(c1_7, rx * t123 + ry * t213 + rz * t222, c2_7 * (t023 + t203 + t221));
    L[31] = fmaf(3.0f * mzz, t223, L[31]);
    L[22] = fmaf(6.0f * m102, t223, L[22]);
    L[20] = fmaf(6.0f * m012, t223, L[20]);
    L[19] = fmaf(m003, t223, L[19]);
    const float t214 = -fmaf(c1_7, rx * t114 + ry * t204 + rz * t213, c2_7 * (t014 + t212));
    L[20] = fmaf(4.0f * m003, t214, L[20]);
    const float t160 = -fmaf(c1_7, rx * t060 + ry * t150, c2_7 * (t140));
    L[32] = fmaf(15.0f * myy, t160, L[32]);
    L[34] = fmaf(6.0f * mxy, t160, L[34]);
    L[23] = fmaf(15.0f * m120, t160, L[23]);
    L[21] = fmaf(20.0f * m030, t160, L[21]);
    const float t151 = -fmaf(c1_7, rx * t051 + ry * t141 + rz * t150, c2_7 * (t131));
    L[33] = fmaf(10.0f * myy, t151, L[33]);
    L[35] = fmaf(5.0f * mxy, t151, L[35]);
    L[32] = fmaf(5.0f * myz, t151, L[32]);
    L[34] = fmaf(mzx, t151, L[34]);
    L[24] = fmaf(10.0f * m120, t151, L[24]);
    L[23] = fmaf(5.0f * m111, t151, L[23]);
    L[22] = fmaf(10.0f * m030, t151, L[22]);
    L[21] = fmaf(10.0f * m021, t151, L[21]);
    const float t142 = -fmaf(c1_7, rx * t042 + ry * t132 + rz * t141, c2_7 * (t122 + t140));
    L[32] = fmaf(mzz, t142, L[32]);
    L[33] = fmaf(8.0f * myz, t142, L[33]);
    L[35] = fmaf(2.0f * mzx, t142, L[35]);
    L[24] = fmaf(8.0f * m111, t142, L[24]);
    L[23] = fmaf(m102, t142, L[23]);
    L[22] = fmaf(12.0f * m021, t142, L[22]);
    L[21] = fmaf(4.0f * m012, t142, L[21]);
    const float t133 = -fmaf(c1_7, rx * t033 + ry * t123 + rz * t132, c2_7 * (t113 + t131));
    L[33] = fmaf(3.0f * mzz, t133, L[33]);
    L[24] = fmaf(3.0f * m102, t133, L[24]);
    L[22] = fmaf(9.0f * m012, t133, L[22]);
    L[21] = fmaf(m003, t133, L[21]);
    const float t124 = -fmaf(c1_7, rx * t024 + ry * t114 + rz * t123, c2_7 * (t104 + t122));
    L[22] = fmaf(4.0f * m003, t124, L[22]);
    const float t070 = -fmaf(c1_7, ry * t060, c2_7 * (t050));
    L[34] = fmaf(21.0f * myy, t070, L[34]);
    L[23] = fmaf(35.0f * m030, t070, L[23]);
    const float t061 = -fmaf(c1_7, ry * t051 + rz * t060, c2_7 * (t041));
    L[35] = fmaf(15.0f * myy, t061, L[35]);
    L[34] = fmaf(6.0f * myz, t061, L[34]);
    L[24] = fmaf(20.0f * m030, t061, L[24]);
    L[23] = fmaf(15.0f * m021, t061, L[23]);
    const float t052 = -fmaf(c1_7, ry * t042 + rz * t051, c2_7 * (t032 + t050));
    L[34] = fmaf(mzz, t052, L[34]);
    L[35] = fmaf(10.0f * myz, t052, L[35]);
    L[24] = fmaf(20.0f * m021, t052, L[24]);
    L[23] = fmaf(5.0f * m012, t052, L[23]);
    const float t043 = -fmaf(c1_7, ry * t033 + rz * t042, c2_7 * (t023 + t041));
    L[35] = fmaf(3.0f * mzz, t043, L[35]);
    L[24] = fmaf(12.0f * m012, t043, L[24]);
    L[23] = fmaf(m003, t043, L[23]);
    const float t034 = -fmaf(c1_7, ry * t024 + rz * t033, c2_7 * (t014 + t032));
    L[24] = fmaf(4.0f * m003, t034, L[24]);
    const float c1_8 = 1.875f * ir2, c2_8 = 0.875f * ir2;
    const float t800 = -fmaf(c1_8, rx * t700, c2_8 * (t600));
    L[25] = fmaf(56.0f * m300, t800, L[25]);
    const float t710 = -fmaf(c1_8, rx * t610 + ry * t700, c2_8 * (t510));
    L[26] = fmaf(35.0f * m300, t710, L[26]);
    L[25] = fmaf(21.0f * m210, t710, L[25]);
    const float t701 = -fmaf(c1_8, rx * t601 + rz * t700, c2_8 * (t501));
    L[27] = fmaf(35.0f * m300, t701, L[27]);
    L[25] = fmaf(21.0f * m201, t701, L[25]);
    const float t620 = -fmaf(c1_8, rx * t520 + ry * t610, c2_8 * (t420 + t600));
    L[28] = fmaf(20.0f * m300, t620, L[28]);
    L[26] = fmaf(30.0f * m210, t620, L[26]);
    L[25] = fmaf(6.0f * m120, t620, L[25]);
    const float t611 = -fmaf(c1_8, rx * t511 + ry * t601 + rz * t610, c2_8 * (t411));
    L[29] = fmaf(20.0f * m300, t611, L[29]);
    L[27] = fmaf(15.0f * m210, t611, L[27]);
    L[26] = fmaf(15.0f * m201, t611, L[26]);
    L[25] = fmaf(6.0f * m111, t611, L[25]);
    const float t602 = -fmaf(c1_8, rx * t502 + rz * t601, c2_8 * (t402 + t600));
    L[27] = fmaf(30.0f * m201, t602, L[27]);
    L[25] = fmaf(6.0f * m102, t602, L[25]);
    const float t530 = -fmaf(c1_8, rx * t430 + ry * t520, c2_8 * (t330 + t510));
    L[30] = fmaf(10.0f * m300, t530, L[30]);
    L[28] = fmaf(30.0f * m210, t530, L[28]);
    L[26] = fmaf(15.0f * m120, t530, L[26]);
    L[25] = fmaf(m030, t530, L[25]);
    const float t521 = -fmaf(c1_8, rx * t421 + ry * t511 + rz * t520, c2_8 * (t321 + t501));
    L[31] = fmaf(10.0f * m300, t521, L[31]);
    L[29] = fmaf(20.0f * m210, t521, L[29]);
    L[28] = fmaf(10.0f * m201, t521, L[28]);
    L[27] = fmaf(5.0f * m120, t521, L[27]);
    L[26] = fmaf(10.0f * m111, t521, L[26]);
    L[25] = fmaf(m021, t521, L[25]);
    const float t512 = -fmaf(c1_8, rx * t412 + ry * t502 + rz * t511, c2_8 * (t312 + t510));
    L[29] = fmaf(20.0f * m201, t512, L[29]);
    L[27] = fmaf(10.0f * m111, t512, L[27]);
    L[26] = fmaf(5.0f * m102, t512, L[26]);
    L[25] = fmaf(m012, t512, L[25]);
    const float t503 = -fmaf(c1_8, rx * t403 + rz * t502, c2_8 * (t303 + t501));
    L[27] = fmaf(15.0f * m102, t503, L[27]);
    L[25] = fmaf(m003, t503, L[25]);
    const float t440 = -fmaf(c1_8, rx * t340 + ry * t430, c2_8 * (t240 + t420));
    L[32] = fmaf(4.0f * m300, t440, L[32]);
    L[30] = fmaf(24.0f * m210, t440, L[30]);
    L[28] = fmaf(24.0f * m120, t440, L[28]);
    L[26] = fmaf(4.0f * m030, t440, L[26]);
    const float t431 = -fmaf(c1_8, rx * t331 + ry * t421 + rz * t430, c2_8 * (t231 + t411));
    L[33] = fmaf(4.0f * m300, t431, L[33]);
    L[31] = fmaf(18.0f * m210, t431, L[31]);
    L[30] = fmaf(6.0f * m201, t431, L[30]);
    L[29] = fmaf(12.0f * m120, t431, L[29]);
    L[28] = fmaf(12.0f * m111, t431, L[28]);
    L[27] = fmaf(m030, t431, L[27]);
    L[26] = fmaf(3.0f * m021, t431, L[26]);
    const float t422 = -fmaf(c1_8, rx * t322 + ry * t412 + rz * t421, c2_8 * (t222 + t402 + t420));
    L[31] = fmaf(12.0f * m201, t422, L[31]);
    L[29] = fmaf(16.0f * m111, t422, L[29]);
    L[28] = fmaf(4.0f * m102, t422, L[28]);
    L[27] = fmaf(2.0f * m021, t422, L[27]);
    L[26] = fmaf(2.0f * m012, t422, L[26]);
    const float t413 = -fmaf(c1_8, rx * t313 + ry * t403 + rz * t412, c2_8 * (t213 + t411));
    L[29] = fmaf(12.0f * m102, t413, L[29]);
    L[27] = fmaf(3.0f * m012, t413, L[27]);
    L[26] = fmaf(m003, t413, L[26]);
    const float t404 = -fmaf(c1_8, rx * t304 + rz * t403, c2_8 * (t204 + t402));
    L[27] = fmaf(4.0f * m003, t404, L[27]);
    const float t350 = -fmaf(c1_8, rx * t250 + ry * t340, c2_8 * (t150 + t330));
    L[34] = fmaf(m300, t350, L[34]);
    L[32] = fmaf(15.0f * m210, t350, L[32]);
    L[30] = fmaf(30.0f * m120, t350, L[30]);
    L[28] = fmaf(10.0f * m030, t350, L[28]);
    const float t341 = -fmaf(c1_8, rx * t241 + ry * t331 + rz * t340, c2_8 * (t141 + t321));
    L[35] = fmaf(m300, t341, L[35]);
    L[33] = fmaf(12.0f * m210, t341, L[33]);
    L[32] = fmaf(3.0f * m201, t341, L[32]);
    L[31] = fmaf(18.0f * m120, t341, L[31]);
    L[30] = fmaf(12.0f * m111, t341, L[30]);
    L[29] = fmaf(4.0f * m030, t341, L[29]);
    L[28] = fmaf(6.0f * m021, t341, L[28]);
    const float t332 = -fmaf(c1_8, rx * t232 + ry * t322 + rz * t331, c2_8 * (t132 + t312 + t330));
    L[33] = fmaf(6.0f * m201, t332, L[33]);
    L[31] = fmaf(18.0f * m111, t332, L[31]);
    L[30] = fmaf(3.0f * m102, t332, L[30]);
    L[29] = fmaf(6.0f * m021, t332, L[29]);
    L[28] = fmaf(3.0f * m012, t332, L[28]);
    const float t323 = -fmaf(c1_8, rx * t223 + ry * t313 + rz * t322, c2_8 * (t123 + t303 + t321));
    L[31] = fmaf(9.0f * m102, t323, L[31]);
    L[29] = fmaf(6.0f * m012, t323, L[29]);
    L[28] = fmaf(m003, t323, L[28]);
    const float t314 = -fmaf(c1_8, rx * t214 + ry * t304 + rz * t313, c2_8 * (t114 + t312));
    L[29] = fmaf(4.0f * m003, t314, L[29]);
    const float t260 = -fmaf(c1_8, rx * t160 + ry * t250, c2_8 * (t060 + t240));
    L[34] = fmaf(6.0f * m210, t260, L[34]);
    L[32] = fmaf(30.0f * m120, t260, L[32]);
    L[30] = fmaf(20.0f * m030, t260, L[30]);
    const float t251 = -fmaf(c1_8, rx * t151 + ry * t241 + rz * t250, c2_8 * (t051 + t231));
    L[35] = fmaf(5.0f * m210, t251, L[35]);
    L[34] = fmaf(m201, t251, L[34]);
    L[33] = fmaf(20.0f * m120, t251, L[33]);
    L[32] = fmaf(10.0f * m111, t251, L[32]);
    L[31] = fmaf(10.0f * m030, t251, L[31]);
    L[30] = fmaf(10.0f * m021, t251, L[30]);
    const float t242 = -fmaf(c1_8, rx * t142 + ry * t232 + rz * t241, c2_8 * (t042 + t222 + t240));
    L[35] = fmaf(2.0f * m201, t242, L[35]);
    L[33] = fmaf(16.0f * m111, t242, L[33]);
    L[32] = fmaf(2.0f * m102, t242, L[32]);
    L[31] = fmaf(12.0f * m021, t242, L[31]);
    L[30] = fmaf(4.0f * m012, t242, L[30]);
    const float t233 = -fmaf(c1_8, rx * t133 + ry * t223 + rz * t232, c2_8 * (t033 + t213 + t231));
    L[33] = fmaf(6.0f * m102, t233, L[33]);
    L[31] = fmaf(9.0f * m012, t233, L[31]);
    L[30] = fmaf(m003, t233, L[30]);
    const float t224 = -fmaf(c1_8, rx * t124 + ry * t214 + rz * t223, c2_8 * (t024 + t204 + t222));
    L[31] = fmaf(4.0f * m003, t224, L[31]);
    const float t170 = -fmaf(c1_8, rx * t070 + ry * t160, c2_8 * (t150));
    L[34] = fmaf(21.0f * m120, t170, L[34]);
    L[32] = fmaf(35.0f * m030, t170, L[32]);
    const float t161 = -fmaf(c1_8, rx * t061 + ry * t151 + rz * t160, c2_8 * (t141));
    L[35] = fmaf(15.0f * m120, t161, L[35]);
    L[34] = fmaf(6.0f * m111, t161, L[34]);
    L[33] = fmaf(20.0f * m030, t161, L[33]);
    L[32] = fmaf(15.0f * m021, t161, L[32]);
    const float t152 = -fmaf(c1_8, rx * t052 + ry * t142 + rz * t151, c2_8 * (t132 + t150));
    L[35] = fmaf(10.0f * m111, t152, L[35]);
    L[34] = fmaf(m102, t152, L[34]);
    L[33] = fmaf(20.0f * m021, t152, L[33]);
    L[32] = fmaf(5.0f * m012, t152, L[32]);
    const float t143 = -fmaf(c1_8, rx * t043 + ry * t133 + rz * t142, c2_8 * (t123 + t141));
    L[35] = fmaf(3.0f * m102, t143, L[35]);
    L[33] = fmaf(12.0f * m012, t143, L[33]);
    L[32] = fmaf(m003, t143, L[32]);
    const float t134 = -fmaf(c1_8, rx * t034 + ry * t124 + rz * t133, c2_8 * (t114 + t132));
    L[33] = fmaf(4.0f * m003, t134, L[33]);
    const float t080 = -fmaf(c1_8, ry * t070, c2_8 * (t060));
    L[34] = fmaf(56.0f * m030, t080, L[34]);
    const float t071 = -fmaf(c1_8, ry * t061 + rz * t070, c2_8 * (t051));
    L[35] = fmaf(35.0f * m030, t071, L[35]);
    L[34] = fmaf(21.0f * m021, t071, L[34]);
    const float t062 = -fmaf(c1_8, ry * t052 + rz * t061, c2_8 * (t042 + t060));
    L[35] = fmaf(30.0f * m021, t062, L[35]);
    L[34] = fmaf(6.0f * m012, t062, L[34]);
    const float t053 = -fmaf(c1_8, ry * t043 + rz * t052, c2_8 * (t033 + t051));
    L[35] = fmaf(15.0f * m012, t053, L[35]);
    L[34] = fmaf(m003, t053, L[34]);
    const float t044 = -fmaf(c1_8, ry * t034 + rz * t043, c2_8 * (t024 + t042));
    L[35] = fmaf(4.0f * m003, t044, L[35]);
}

// potential and field at offset y from the expansion centre (order P = 5, from the
// independent coefficients b_z <= 1; each homogeneous part of a harmonic
// expansion is harmonic, which fixes the others)
__device__ inline void l2p_p5h(const float (&Lh)[36], float x, float y, float z, float &phi, float &ex, float &ey, float &ez)
{
    float L[56];
    L[0] = Lh[0];
    L[1] = Lh[1];
    L[2] = Lh[2];
    L[3] = Lh[3];
    L[4] = Lh[4];
    L[5] = Lh[5];
    L[6] = Lh[6];
    L[7] = Lh[7];
    L[8] = Lh[8];
    L[10] = Lh[9];
    L[11] = Lh[10];
    L[12] = Lh[11];
    L[13] = Lh[12];
    L[14] = Lh[13];
    L[16] = Lh[14];
    L[17] = Lh[15];
    L[20] = Lh[16];
    L[21] = Lh[17];
    L[22] = Lh[18];
    L[23] = Lh[19];
    L[24] = Lh[20];
    L[26] = Lh[21];
    L[27] = Lh[22];
    L[30] = Lh[23];
    L[31] = Lh[24];
    L[35] = Lh[25];
    L[36] = Lh[26];
    L[37] = Lh[27];
    L[38] = Lh[28];
    L[39] = Lh[29];
    L[41] = Lh[30];
    L[42] = Lh[31];
    L[45] = Lh[32];
    L[46] = Lh[33];
    L[50] = Lh[34];
    L[51] = Lh[35];
    L[9] = -(1.0f * L[4] + 1.0f * L[7]);
    L[15] = -(3.0f * L[10] + 1.0f * L[13]);
    L[18] = -(1.0f * L[11] + 3.0f * L[16]);
    L[25] = -(6.0f * L[20] + 1.0f * L[23]);
    L[28] = -(3.0f * L[21] + 3.0f * L[26]);
    L[32] = -(1.0f * L[23] + 6.0f * L[30]);
    L[40] = -(10.0f * L[35] + 1.0f * L[38]);
    L[43] = -(6.0f * L[36] + 3.0f * L[41]);
    L[47] = -(3.0f * L[38] + 6.0f * L[45]);
    L[52] = -(1.0f * L[41] + 10.0f * L[50]);
    L[19] = -(0.3333333333333333f * L[12] + 0.3333333333333333f * L[17]);
    L[29] = -(1.0f * L[22] + 0.3333333333333333f * L[27]);
    L[33] = -(0.3333333333333333f * L[24] + 1.0f * L[31]);
    L[44] = -(2.0f * L[37] + 0.3333333333333333f * L[42]);
    L[48] = -(1.0f * L[39] + 1.0f * L[46]);
    L[53] = -(0.3333333333333333f * L[42] + 2.0f * L[51]);
    L[34] = -(0.16666666666666666f * L[25] + 0.16666666666666666f * L[32]);
    L[49] = -(0.5f * L[40] + 0.16666666666666666f * L[47]);
    L[54] = -(0.16666666666666666f * L[43] + 0.5f * L[52]);
    L[55] = -(0.1f * L[44] + 0.1f * L[53]);
    const float x0 = 1.0f, x1 = x;
    const float x2 = x1 * x;
    const float x3 = x2 * x;
    const float x4 = x3 * x;
    const float x5 = x4 * x;
    const float y0 = 1.0f, y1 = y;
    const float y2 = y1 * y;
    const float y3 = y2 * y;
    const float y4 = y3 * y;
    const float y5 = y4 * y;
    const float z0 = 1.0f, z1 = z;
    const float z2 = z1 * z;
    const float z3 = z2 * z;
    const float z4 = z3 * z;
    const float z5 = z4 * z;
    phi = L[0] * (x0 * y0 * z0)
        + L[1] * (x1 * y0 * z0)
        + L[2] * (x0 * y1 * z0)
        + L[3] * (x0 * y0 * z1)
        + L[4] * (x2 * y0 * z0)
        + L[5] * (x1 * y1 * z0)
        + L[6] * (x1 * y0 * z1)
        + L[7] * (x0 * y2 * z0)
        + L[8] * (x0 * y1 * z1)
        + L[9] * (x0 * y0 * z2)
        + L[10] * (x3 * y0 * z0)
        + L[11] * (x2 * y1 * z0)
        + L[12] * (x2 * y0 * z1)
        + L[13] * (x1 * y2 * z0)
        + L[14] * (x1 * y1 * z1)
        + L[15] * (x1 * y0 * z2)
        + L[16] * (x0 * y3 * z0)
        + L[17] * (x0 * y2 * z1)
        + L[18] * (x0 * y1 * z2)
        + L[19] * (x0 * y0 * z3)
        + L[20] * (x4 * y0 * z0)
        + L[21] * (x3 * y1 * z0)
        + L[22] * (x3 * y0 * z1)
        + L[23] * (x2 * y2 * z0)
        + L[24] * (x2 * y1 * z1)
        + L[25] * (x2 * y0 * z2)
        + L[26] * (x1 * y3 * z0)
        + L[27] * (x1 * y2 * z1)
        + L[28] * (x1 * y1 * z2)
        + L[29] * (x1 * y0 * z3)
        + L[30] * (x0 * y4 * z0)
        + L[31] * (x0 * y3 * z1)
        + L[32] * (x0 * y2 * z2)
        + L[33] * (x0 * y1 * z3)
        + L[34] * (x0 * y0 * z4)
        + L[35] * (x5 * y0 * z0)
        + L[36] * (x4 * y1 * z0)
        + L[37] * (x4 * y0 * z1)
        + L[38] * (x3 * y2 * z0)
        + L[39] * (x3 * y1 * z1)
        + L[40] * (x3 * y0 * z2)
        + L[41] * (x2 * y3 * z0)
        + L[42] * (x2 * y2 * z1)
        + L[43] * (x2 * y1 * z2)
        + L[44] * (x2 * y0 * z3)
        + L[45] * (x1 * y4 * z0)
        + L[46] * (x1 * y3 * z1)
        + L[47] * (x1 * y2 * z2)
        + L[48] * (x1 * y1 * z3)
        + L[49] * (x1 * y0 * z4)
        + L[50] * (x0 * y5 * z0)
        + L[51] * (x0 * y4 * z1)
        + L[52] * (x0 * y3 * z2)
        + L[53] * (x0 * y2 * z3)
        + L[54] * (x0 * y1 * z4)
        + L[55] * (x0 * y0 * z5);
    ex = -(1.0f * L[1] * (x0 * y0 * z0)
        + 2.0f * L[4] * (x1 * y0 * z0)
        + 1.0f * L[5] * (x0 * y1 * z0)
        + 1.0f * L[6] * (x0 * y0 * z1)
        + 3.0f * L[10] * (x2 * y0 * z0)
        + 2.0f * L[11] * (x1 * y1 * z0)
        + 2.0f * L[12] * (x1 * y0 * z1)
        + 1.0f * L[13] * (x0 * y2 * z0)
        + 1.0f * L[14] * (x0 * y1 * z1)
        + 1.0f * L[15] * (x0 * y0 * z2)
        + 4.0f * L[20] * (x3 * y0 * z0)
        + 3.0f * L[21] * (x2 * y1 * z0)
        + 3.0f * L[22] * (x2 * y0 * z1)
        + 2.0f * L[23] * (x1 * y2 * z0)
        + 2.0f * L[24] * (x1 * y1 * z1)
        + 2.0f * L[25] * (x1 * y0 * z2)
        + 1.0f * L[26] * (x0 * y3 * z0)
        + 1.0f * L[27] * (x0 * y2 * z1)
        + 1.0f * L[28] * (x0 * y1 * z2)
        + 1.0f * L[29] * (x0 * y0 * z3)
        + 5.0f * L[35] * (x4 * y0 * z0)
        + 4.0f * L[36] * (x3 * y1 * z0)
        + 4.0f * L[37] * (x3 * y0 * z1)
        + 3.0f * L[38] * (x2 * y2 * z0)
        + 3.0f * L[39] * (x2 * y1 * z1)
        + 3.0f * L[40] * (x2 * y0 * z2)
        + 2.0f * L[41] * (x1 * y3 * z0)
        + 2.0f * L[42] * (x1 * y2 * z1)
        + 2.0f * L[43] * (x1 * y1 * z2)
        + 2.0f * L[44] * (x1 * y0 * z3)
        + 1.0f * L[45] * (x0 * y4 * z0)
        + 1.0f * L[46] * (x0 * y3 * z1)
        + 1.0f * L[47] * (x0 * y2 * z2)
        + 1.0f * L[48] * (x0 * y1 * z3)
        + 1.0f * L[49] * (x0 * y0 * z4));
    ey = -(1.0f * L[2] * (x0 * y0 * z0)
        + 1.0f * L[5] * (x1 * y0 * z0)
        + 2.0f * L[7] * (x0 * y1 * z0)
        + 1.0f * L[8] * (x0 * y0 * z1)
        + 1.0f * L[11] * (x2 * y0 * z0)
        + 2.0f * L[13] * (x1 * y1 * z0)
        + 1.0f * L[14] * (x1 * y0 * z1)
        + 3.0f * L[16] * (x0 * y2 * z0)
        + 2.0f * L[17] * (x0 * y1 * z1)
        + 1.0f * L[18] * (x0 * y0 * z2)
        + 1.0f * L[21] * (x3 * y0 * z0)
        + 2.0f * L[23] * (x2 * y1 * z0)
        + 1.0f * L[24] * (x2 * y0 * z1)
        + 3.0f * L[26] * (x1 * y2 * z0)
        + 2.0f * L[27] * (x1 * y1 * z1)
        + 1.0f * L[28] * (x1 * y0 * z2)
        + 4.0f * L[30] * (x0 * y3 * z0)
        + 3.0f * L[31] * (x0 * y2 * z1)
        + 2.0f * L[32] * (x0 * y1 * z2)
        + 1.0f * L[33] * (x0 * y0 * z3)
        + 1.0f * L[36] * (x4 * y0 * z0)
        + 2.0f * L[38] * (x3 * y1 * z0)
        + 1.0f * L[39] * (x3 * y0 * z1)
        + 3.0f * L[41] * (x2 * y2 * z0)
        + 2.0f * L[42] * (x2 * y1 * z1)
        + 1.0f * L[43] * (x2 * y0 * z2)
        + 4.0f * L[45] * (x1 * y3 * z0)
        + 3.0f * L[46] * (x1 * y2 * z1)
        + 2.0f * L[47] * (x1 * y1 * z2)
        + 1.0f * L[48] * (x1 * y0 * z3)
        + 5.0f * L[50] * (x0 * y4 * z0)
        + 4.0f * L[51] * (x0 * y3 * z1)
        + 3.0f * L[52] * (x0 * y2 * z2)
        + 2.0f * L[53] * (x0 * y1 * z3)
        + 1.0f * L[54] * (x0 * y0 * z4));
    ez = -(1.0f * L[3] * (x0 * y0 * z0)
        + 1.0f * L[6] * (x1 * y0 * z0)
        + 1.0f * L[8] * (x0 * y1 * z0)
        + 2.0f * L[9] * (x0 * y0 * z1)
        + 1.0f * L[12] * (x2 * y0 * z0)
        + 1.0f * L[14] * (x1 * y1 * z0)
        + 2.0f * L[15] * (x1 * y0 * z1)
        + 1.0f * L[17] * (x0 * y2 * z0)
        + 2.0f * L[18] * (x0 * y1 * z1)
        + 3.0f * L[19] * (x0 * y0 * z2)
        + 1.0f * L[22] * (x3 * y0 * z0)
        + 1.0f * L[24] * (x2 * y1 * z0)
        + 2.0f * L[25] * (x2 * y0 * z1)
        + 1.0f * L[27] * (x1 * y2 * z0)
        + 2.0f * L[28] * (x1 * y1 * z1)
        + 3.0f * L[29] * (x1 * y0 * z2)
        + 1.0f * L[31] * (x0 * y3 * z0)
        + 2.0f * L[32] * (x0 * y2 * z1)
        + 3.0f * L[33] * (x0 * y1 * z2)
        + 4.0f * L[34] * (x0 * y0 * z3)
        + 1.0f * L[37] * (x4 * y0 * z0)
        + 1.0f * L[39] * (x3 * y1 * z0)
        + 2.0f * L[40] * (x3 * y0 * z1)
        + 1.0f * L[42] * (x2 * y2 * z0)
        + 2.0f * L[43] * (x2 * y1 * z1)
        + 3.0f * L[44] * (x2 * y0 * z2)
        + 1.0f * L[46] * (x1 * y3 * z0)
        + 2.0f * L[47] * (x1 * y2 * z1)
        + 3.0f * L[48] * (x1 * y1 * z2)
        + 4.0f * L[49] * (x1 * y0 * z3)
        + 1.0f * L[51] * (x0 * y4 * z0)
        + 2.0f * L[52] * (x0 * y3 * z1)
        + 3.0f * L[53] * (x0 * y2 * z2)
        + 4.0f * L[54] * (x0 * y1 * z3)
        + 5.0f * L[55] * (x0 * y0 * z4));
}

// adds one cell's local coefficients about the expansion centre to L[49]
// (order P = 6; R = centre - cell centre; Q' traceless, Q'zz = -Q'xx - Q'yy)
__device__ inline void m2l_acc_p6hxt7(float rx, float ry, float rz, float q, float dx, float dy, float dz,
                                  float qxx, float qyy, float qxy, float qyz, float qzx,
                                  float oxxx, float oxxy, float oxxz, float oxyy, float oxyz, float oyyy, float oyyz,
                                  float hxxxx, float hxxxy, float hxxxz, float hxxyy, float hxxyz, float hxyyy, float hxyyz,
                                  float hyyyy, float hyyyz, float (&L)[49])
{
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float mdx = -dx, mdy = -dy, mdz = -dz;
    const float mxx = qxx * (1.0f / 3.0f), myy = qyy * (1.0f / 3.0f), mzz = -(qxx + qyy) * (1.0f / 3.0f);
    const float mxy = qxy * (1.0f / 3.0f), myz = qyz * (1.0f / 3.0f), mzx = qzx * (1.0f / 3.0f);
    const float m300 = -oxxx, m210 = -oxxy, m201 = -oxxz, m120 = -oxyy, m111 = -oxyz;
    const float m030 = -oyyy, m021 = -oyyz, m102 = oxxx + oxyy, m012 = oxxy + oyyy, m003 = oxxz + oyyz;
    const float m400 = hxxxx, m310 = hxxxy, m301 = hxxxz, m220 = hxxyy, m211 = hxxyz;
    const float m130 = hxyyy, m121 = hxyyz, m040 = hyyyy, m031 = hyyyz;
    const float m202 = -(hxxxx + hxxyy), m112 = -(hxxxy + hxyyy), m022 = -(hxxyy + hyyyy);
    const float m103 = -(hxxxz + hxyyz), m013 = -(hxxyz + hyyyz), m004 = hxxxx + 2.0f * hxxyy + hyyyy;
    const float t000 = __builtin_amdgcn_rsqf(r2);
    const float ir2 = t000 * t000;
    L[0] = fmaf(q, t000, L[0]);
    const float c1_1 = 1.0f * ir2, c2_1 = 0.0f * ir2;
    const float t100 = -c1_1 * (rx * t000);
    L[1] = fmaf(q, t100, L[1]);
    L[0] = fmaf(mdx, t100, L[0]);
    const float t010 = -c1_1 * (ry * t000);
    L[2] = fmaf(q, t010, L[2]);
    L[0] = fmaf(mdy, t010, L[0]);
    const float t001 = -c1_1 * (rz * t000);
    L[3] = fmaf(q, t001, L[3]);
    L[0] = fmaf(mdz, t001, L[0]);
    const float c1_2 = 1.5f * ir2, c2_2 = 0.5f * ir2;
    const float t200 = -fmaf(c1_2, rx * t100, c2_2 * (t000));
    L[4] = fmaf(q, t200, L[4]);
    L[1] = fmaf(2.0f * mdx, t200, L[1]);
    L[0] = fmaf(mxx, t200, L[0]);
    const float t110 = -c1_2 * (rx * t010 + ry * t100);
    L[5] = fmaf(q, t110, L[5]);
    L[2] = fmaf(mdx, t110, L[2]);
    L[1] = fmaf(mdy, t110, L[1]);
    L[0] = fmaf(mxy, t110, L[0]);
    const float t101 = -c1_2 * (rx * t001 + rz * t100);
    L[6] = fmaf(q, t101, L[6]);
    L[3] = fmaf(mdx, t101, L[3]);
    L[1] = fmaf(mdz, t101, L[1]);
    L[0] = fmaf(mzx, t101, L[0]);
    const float t020 = -fmaf(c1_2, ry * t010, c2_2 * (t000));
    L[7] = fmaf(q, t020, L[7]);
    L[2] = fmaf(2.0f * mdy, t020, L[2]);
    L[0] = fmaf(myy, t020, L[0]);
    const float t011 = -c1_2 * (ry * t001 + rz * t010);
    L[8] = fmaf(q, t011, L[8]);
    L[3] = fmaf(mdy, t011, L[3]);
    L[2] = fmaf(mdz, t011, L[2]);
    L[0] = fmaf(myz, t011, L[0]);
    const float t002 = -fmaf(c1_2, rz * t001, c2_2 * (t000));
    L[3] = fmaf(2.0f * mdz, t002, L[3]);
    L[0] = fmaf(mzz, t002, L[0]);
    const float c1_3 = 1.6666666666666667f * ir2, c2_3 = 0.6666666666666666f * ir2;
    const float t300 = -fmaf(c1_3, rx * t200, c2_3 * (t100));
    L[9] = fmaf(q, t300, L[9]);
    L[4] = fmaf(3.0f * mdx, t300, L[4]);
    L[1] = fmaf(3.0f * mxx, t300, L[1]);
    L[0] = fmaf(m300, t300, L[0]);
    const float t210 = -fmaf(c1_3, rx * t110 + ry * t200, c2_3 * (t010));
    L[10] = fmaf(q, t210, L[10]);
    L[5] = fmaf(2.0f * mdx, t210, L[5]);
    L[4] = fmaf(mdy, t210, L[4]);
    L[2] = fmaf(mxx, t210, L[2]);
    L[1] = fmaf(2.0f * mxy, t210, L[1]);
    L[0] = fmaf(m210, t210, L[0]);
    const float t201 = -fmaf(c1_3, rx * t101 + rz * t200, c2_3 * (t001));
    L[11] = fmaf(q, t201, L[11]);
    L[6] = fmaf(2.0f * mdx, t201, L[6]);
    L[4] = fmaf(mdz, t201, L[4]);
    L[3] = fmaf(mxx, t201, L[3]);
    L[1] = fmaf(2.0f * mzx, t201, L[1]);
    L[0] = fmaf(m201, t201, L[0]);
    const float t120 = -fmaf(c1_3, rx * t020 + ry * t110, c2_3 * (t100));
    L[12] = fmaf(q, t120, L[12]);
    L[7] = fmaf(mdx, t120, L[7]);
    L[5] = fmaf(2.0f * mdy, t120, L[5]);
    L[1] = fmaf(myy, t120, L[1]);
    L[2] = fmaf(2.0f * mxy, t120, L[2]);
    L[0] = fmaf(m120, t120, L[0]);
    const float t111 = -c1_3 * (rx * t011 + ry * t101 + rz * t110);
    L[13] = fmaf(q, t111, L[13]);
    L[8] = fmaf(mdx, t111, L[8]);
    L[6] = fmaf(mdy, t111, L[6]);
    L[5] = fmaf(mdz, t111, L[5]);
    L[3] = fmaf(mxy, t111, L[3]);
    L[1] = fmaf(myz, t111, L[1]);
    L[2] = fmaf(mzx, t111, L[2]);
    L[0] = fmaf(m111, t111, L[0]);
    const float t102 = -fmaf(c1_3, rx * t002 + rz * t101, c2_3 * (t100));
    L[6] = fmaf(2.0f * mdz, t102, L[6]);
    L[1] = fmaf(mzz, t102, L[1]);
    L[3] = fmaf(2.0f * mzx, t102, L[3]);
    L[0] = fmaf(m102, t102, L[0]);
    const float t030 = -fmaf(c1_3, ry * t020, c2_3 * (t010));
    L[14] = fmaf(q, t030, L[14]);
    L[7] = fmaf(3.0f * mdy, t030, L[7]);
    L[2] = fmaf(3.0f * myy, t030, L[2]);
    L[0] = fmaf(m030, t030, L[0]);
    const float t021 = -fmaf(c1_3, ry * t011 + rz * t020, c2_3 * (t001));
    L[15] = fmaf(q, t021, L[15]);
    L[8] = fmaf(2.0f * mdy, t021, L[8]);
    L[7] = fmaf(mdz, t021, L[7]);
    L[3] = fmaf(myy, t021, L[3]);
    L[2] = fmaf(2.0f * myz, t021, L[2]);
    L[0] = fmaf(m021, t021, L[0]);
    const float t012 = -fmaf(c1_3, ry * t002 + rz * t011, c2_3 * (t010));
    L[8] = fmaf(2.0f * mdz, t012, L[8]);
    L[2] = fmaf(mzz, t012, L[2]);
    L[3] = fmaf(2.0f * myz, t012, L[3]);
    L[0] = fmaf(m012, t012, L[0]);
    const float t003 = -fmaf(c1_3, rz * t002, c2_3 * (t001));
    L[3] = fmaf(3.0f * mzz, t003, L[3]);
    L[0] = fmaf(m003, t003, L[0]);
    const float c1_4 = 1.75f * ir2, c2_4 = 0.75f * ir2;
    const float t400 = -fmaf(c1_4, rx * t300, c2_4 * (t200));
    L[16] = fmaf(q, t400, L[16]);
    L[9] = fmaf(4.0f * mdx, t400, L[9]);
    L[4] = fmaf(6.0f * mxx, t400, L[4]);
    L[1] = fmaf(4.0f * m300, t400, L[1]);
    L[0] = fmaf(m400, t400, L[0]);
    const float t310 = -fmaf(c1_4, rx * t210 + ry * t300, c2_4 * (t110));
    L[17] = fmaf(q, t310, L[17]);
    L[10] = fmaf(3.0f * mdx, t310, L[10]);
    L[9] = fmaf(mdy, t310, L[9]);
    L[5] = fmaf(3.0f * mxx, t310, L[5]);
    L[4] = fmaf(3.0f * mxy, t310, L[4]);
    L[2] = fmaf(m300, t310, L[2]);
    L[1] = fmaf(3.0f * m210, t310, L[1]);
    L[0] = fmaf(m310, t310, L[0]);
    const float t301 = -fmaf(c1_4, rx * t201 + rz * t300, c2_4 * (t101));
    L[18] = fmaf(q, t301, L[18]);
    L[11] = fmaf(3.0f * mdx, t301, L[11]);
    L[9] = fmaf(mdz, t301, L[9]);
    L[6] = fmaf(3.0f * mxx, t301, L[6]);
    L[4] = fmaf(3.0f * mzx, t301, L[4]);
    L[3] = fmaf(m300, t301, L[3]);
    L[1] = fmaf(3.0f * m201, t301, L[1]);
    L[0] = fmaf(m301, t301, L[0]);
    const float t220 = -fmaf(c1_4, rx * t120 + ry * t210, c2_4 * (t020 + t200));
    L[19] = fmaf(q, t220, L[19]);
    L[12] = fmaf(2.0f * mdx, t220, L[12]);
    L[10] = fmaf(2.0f * mdy, t220, L[10]);
    L[7] = fmaf(mxx, t220, L[7]);
    L[4] = fmaf(myy, t220, L[4]);
    L[5] = fmaf(4.0f * mxy, t220, L[5]);
    L[2] = fmaf(2.0f * m210, t220, L[2]);
    L[1] = fmaf(2.0f * m120, t220, L[1]);
    L[0] = fmaf(m220, t220, L[0]);
    const float t211 = -fmaf(c1_4, rx * t111 + ry * t201 + rz * t210, c2_4 * (t011));
    L[20] = fmaf(q, t211, L[20]);
    L[13] = fmaf(2.0f * mdx, t211, L[13]);
    L[11] = fmaf(mdy, t211, L[11]);
    L[10] = fmaf(mdz, t211, L[10]);
    L[8] = fmaf(mxx, t211, L[8]);
    L[6] = fmaf(2.0f * mxy, t211, L[6]);
    L[4] = fmaf(myz, t211, L[4]);
    L[5] = fmaf(2.0f * mzx, t211, L[5]);
    L[3] = fmaf(m210, t211, L[3]);
    L[2] = fmaf(m201, t211, L[2]);
    L[1] = fmaf(2.0f * m111, t211, L[1]);
    L[0] = fmaf(m211, t211, L[0]);
    const float t202 = -fmaf(c1_4, rx * t102 + rz * t201, c2_4 * (t002 + t200));
    L[11] = fmaf(2.0f * mdz, t202, L[11]);
    L[4] = fmaf(mzz, t202, L[4]);
    L[6] = fmaf(4.0f * mzx, t202, L[6]);
    L[3] = fmaf(2.0f * m201, t202, L[3]);
    L[1] = fmaf(2.0f * m102, t202, L[1]);
    L[0] = fmaf(m202, t202, L[0]);
    const float t130 = -fmaf(c1_4, rx * t030 + ry * t120, c2_4 * (t110));
    L[21] = fmaf(q, t130, L[21]);
    L[14] = fmaf(mdx, t130, L[14]);
    L[12] = fmaf(3.0f * mdy, t130, L[12]);
    L[5] = fmaf(3.0f * myy, t130, L[5]);
    L[7] = fmaf(3.0f * mxy, t130, L[7]);
    L[2] = fmaf(3.0f * m120, t130, L[2]);
    L[1] = fmaf(m030, t130, L[1]);
    L[0] = fmaf(m130, t130, L[0]);
    const float t121 = -fmaf(c1_4, rx * t021 + ry * t111 + rz * t120, c2_4 * (t101));
    L[22] = fmaf(q, t121, L[22]);
    L[15] = fmaf(mdx, t121, L[15]);
    L[13] = fmaf(2.0f * mdy, t121, L[13]);
    L[12] = fmaf(mdz, t121, L[12]);
    L[6] = fmaf(myy, t121, L[6]);
    L[8] = fmaf(2.0f * mxy, t121, L[8]);
    L[5] = fmaf(2.0f * myz, t121, L[5]);
    L[7] = fmaf(mzx, t121, L[7]);
    L[3] = fmaf(m120, t121, L[3]);
    L[2] = fmaf(2.0f * m111, t121, L[2]);
    L[1] = fmaf(m021, t121, L[1]);
    L[0] = fmaf(m121, t121, L[0]);
    const float t112 = -fmaf(c1_4, rx * t012 + ry * t102 + rz * t111, c2_4 * (t110));
    L[13] = fmaf(2.0f * mdz, t112, L[13]);
    L[5] = fmaf(mzz, t112, L[5]);
    L[6] = fmaf(2.0f * myz, t112, L[6]);
    L[8] = fmaf(2.0f * mzx, t112, L[8]);
    L[3] = fmaf(2.0f * m111, t112, L[3]);
    L[2] = fmaf(m102, t112, L[2]);
    L[1] = fmaf(m012, t112, L[1]);
    L[0] = fmaf(m112, t112, L[0]);
    const float t103 = -fmaf(c1_4, rx * t003 + rz * t102, c2_4 * (t101));
    L[6] = fmaf(3.0f * mzz, t103, L[6]);
    L[3] = fmaf(3.0f * m102, t103, L[3]);
    L[1] = fmaf(m003, t103, L[1]);
    L[0] = fmaf(m103, t103, L[0]);
    const float t040 = -fmaf(c1_4, ry * t030, c2_4 * (t020));
    L[23] = fmaf(q, t040, L[23]);
    L[14] = fmaf(4.0f * mdy, t040, L[14]);
    L[7] = fmaf(6.0f * myy, t040, L[7]);
    L[2] = fmaf(4.0f * m030, t040, L[2]);
    L[0] = fmaf(m040, t040, L[0]);
    const float t031 = -fmaf(c1_4, ry * t021 + rz * t030, c2_4 * (t011));
    L[24] = fmaf(q, t031, L[24]);
    L[15] = fmaf(3.0f * mdy, t031, L[15]);
    L[14] = fmaf(mdz, t031, L[14]);
    L[8] = fmaf(3.0f * myy, t031, L[8]);
    L[7] = fmaf(3.0f * myz, t031, L[7]);
    L[3] = fmaf(m030, t031, L[3]);
    L[2] = fmaf(3.0f * m021, t031, L[2]);
    L[0] = fmaf(m031, t031, L[0]);
    const float t022 = -fmaf(c1_4, ry * t012 + rz * t021, c2_4 * (t002 + t020));
    L[15] = fmaf(2.0f * mdz, t022, L[15]);
    L[7] = fmaf(mzz, t022, L[7]);
    L[8] = fmaf(4.0f * myz, t022, L[8]);
    L[3] = fmaf(2.0f * m021, t022, L[3]);
    L[2] = fmaf(2.0f * m012, t022, L[2]);
    L[0] = fmaf(m022, t022, L[0]);
    const float t013 = -fmaf(c1_4, ry * t003 + rz * t012, c2_4 * (t011));
    L[8] = fmaf(3.0f * mzz, t013, L[8]);
    L[3] = fmaf(3.0f * m012, t013, L[3]);
    L[2] = fmaf(m003, t013, L[2]);
    L[0] = fmaf(m013, t013, L[0]);
    const float t004 = -fmaf(c1_4, rz * t003, c2_4 * (t002));
    L[3] = fmaf(4.0f * m003, t004, L[3]);
    L[0] = fmaf(m004, t004, L[0]);
    const float c1_5 = 1.8f * ir2, c2_5 = 0.8f * ir2;
    const float t500 = -fmaf(c1_5, rx * t400, c2_5 * (t300));
    L[25] = fmaf(q, t500, L[25]);
    L[16] = fmaf(5.0f * mdx, t500, L[16]);
    L[9] = fmaf(10.0f * mxx, t500, L[9]);
    L[4] = fmaf(10.0f * m300, t500, L[4]);
    L[1] = fmaf(5.0f * m400, t500, L[1]);
    const float t410 = -fmaf(c1_5, rx * t310 + ry * t400, c2_5 * (t210));
    L[26] = fmaf(q, t410, L[26]);
    L[17] = fmaf(4.0f * mdx, t410, L[17]);
    L[16] = fmaf(mdy, t410, L[16]);
    L[10] = fmaf(6.0f * mxx, t410, L[10]);
    L[9] = fmaf(4.0f * mxy, t410, L[9]);
    L[5] = fmaf(4.0f * m300, t410, L[5]);
    L[4] = fmaf(6.0f * m210, t410, L[4]);
    L[2] = fmaf(m400, t410, L[2]);
    L[1] = fmaf(4.0f * m310, t410, L[1]);
    const float t401 = -fmaf(c1_5, rx * t301 + rz * t400, c2_5 * (t201));
    L[27] = fmaf(q, t401, L[27]);
    L[18] = fmaf(4.0f * mdx, t401, L[18]);
    L[16] = fmaf(mdz, t401, L[16]);
    L[11] = fmaf(6.0f * mxx, t401, L[11]);
    L[9] = fmaf(4.0f * mzx, t401, L[9]);
    L[6] = fmaf(4.0f * m300, t401, L[6]);
    L[4] = fmaf(6.0f * m201, t401, L[4]);
    L[3] = fmaf(m400, t401, L[3]);
    L[1] = fmaf(4.0f * m301, t401, L[1]);
    const float t320 = -fmaf(c1_5, rx * t220 + ry * t310, c2_5 * (t120 + t300));
    L[28] = fmaf(q, t320, L[28]);
    L[19] = fmaf(3.0f * mdx, t320, L[19]);
    L[17] = fmaf(2.0f * mdy, t320, L[17]);
    L[12] = fmaf(3.0f * mxx, t320, L[12]);
    L[9] = fmaf(myy, t320, L[9]);
    L[10] = fmaf(6.0f * mxy, t320, L[10]);
    L[7] = fmaf(m300, t320, L[7]);
    L[5] = fmaf(6.0f * m210, t320, L[5]);
    L[4] = fmaf(3.0f * m120, t320, L[4]);
    L[2] = fmaf(2.0f * m310, t320, L[2]);
    L[1] = fmaf(3.0f * m220, t320, L[1]);
    const float t311 = -fmaf(c1_5, rx * t211 + ry * t301 + rz * t310, c2_5 * (t111));
    L[29] = fmaf(q, t311, L[29]);
    L[20] = fmaf(3.0f * mdx, t311, L[20]);
    L[18] = fmaf(mdy, t311, L[18]);
    L[17] = fmaf(mdz, t311, L[17]);
    L[13] = fmaf(3.0f * mxx, t311, L[13]);
    L[11] = fmaf(3.0f * mxy, t311, L[11]);
    L[9] = fmaf(myz, t311, L[9]);
    L[10] = fmaf(3.0f * mzx, t311, L[10]);
    L[8] = fmaf(m300, t311, L[8]);
    L[6] = fmaf(3.0f * m210, t311, L[6]);
    L[5] = fmaf(3.0f * m201, t311, L[5]);
    L[4] = fmaf(3.0f * m111, t311, L[4]);
    L[3] = fmaf(m310, t311, L[3]);
    L[2] = fmaf(m301, t311, L[2]);
    L[1] = fmaf(3.0f * m211, t311, L[1]);
    const float t302 = -fmaf(c1_5, rx * t202 + rz * t301, c2_5 * (t102 + t300));
    L[18] = fmaf(2.0f * mdz, t302, L[18]);
    L[9] = fmaf(mzz, t302, L[9]);
    L[11] = fmaf(6.0f * mzx, t302, L[11]);
    L[6] = fmaf(6.0f * m201, t302, L[6]);
    L[4] = fmaf(3.0f * m102, t302, L[4]);
    L[3] = fmaf(2.0f * m301, t302, L[3]);
    L[1] = fmaf(3.0f * m202, t302, L[1]);
    const float t230 = -fmaf(c1_5, rx * t130 + ry * t220, c2_5 * (t030 + t210));
    L[30] = fmaf(q, t230, L[30]);
    L[21] = fmaf(2.0f * mdx, t230, L[21]);
    L[19] = fmaf(3.0f * mdy, t230, L[19]);
    L[14] = fmaf(mxx, t230, L[14]);
    L[10] = fmaf(3.0f * myy, t230, L[10]);
    L[12] = fmaf(6.0f * mxy, t230, L[12]);
    L[7] = fmaf(3.0f * m210, t230, L[7]);
    L[5] = fmaf(6.0f * m120, t230, L[5]);
    L[4] = fmaf(m030, t230, L[4]);
    L[2] = fmaf(3.0f * m220, t230, L[2]);
    L[1] = fmaf(2.0f * m130, t230, L[1]);
    const float t221 = -fmaf(c1_5, rx * t121 + ry * t211 + rz * t220, c2_5 * (t021 + t201));
    L[31] = fmaf(q, t221, L[31]);
    L[22] = fmaf(2.0f * mdx, t221, L[22]);
    L[20] = fmaf(2.0f * mdy, t221, L[20]);
    L[19] = fmaf(mdz, t221, L[19]);
    L[15] = fmaf(mxx, t221, L[15]);
    L[11] = fmaf(myy, t221, L[11]);
    L[13] = fmaf(4.0f * mxy, t221, L[13]);
    L[10] = fmaf(2.0f * myz, t221, L[10]);
    L[12] = fmaf(2.0f * mzx, t221, L[12]);
    L[8] = fmaf(2.0f * m210, t221, L[8]);
    L[7] = fmaf(m201, t221, L[7]);
    L[6] = fmaf(2.0f * m120, t221, L[6]);
    L[5] = fmaf(4.0f * m111, t221, L[5]);
    L[4] = fmaf(m021, t221, L[4]);
    L[3] = fmaf(m220, t221, L[3]);
    L[2] = fmaf(2.0f * m211, t221, L[2]);
    L[1] = fmaf(2.0f * m121, t221, L[1]);
    const float t212 = -fmaf(c1_5, rx * t112 + ry * t202 + rz * t211, c2_5 * (t012 + t210));
    L[20] = fmaf(2.0f * mdz, t212, L[20]);
    L[10] = fmaf(mzz, t212, L[10]);
    L[11] = fmaf(2.0f * myz, t212, L[11]);
    L[13] = fmaf(4.0f * mzx, t212, L[13]);
    L[8] = fmaf(2.0f * m201, t212, L[8]);
    L[6] = fmaf(4.0f * m111, t212, L[6]);
    L[5] = fmaf(2.0f * m102, t212, L[5]);
    L[4] = fmaf(m012, t212, L[4]);
    L[3] = fmaf(2.0f * m211, t212, L[3]);
    L[2] = fmaf(m202, t212, L[2]);
    L[1] = fmaf(2.0f * m112, t212, L[1]);
    const float t203 = -fmaf(c1_5, rx * t103 + rz * t202, c2_5 * (t003 + t201));
    L[11] = fmaf(3.0f * mzz, t203, L[11]);
    L[6] = fmaf(6.0f * m102, t203, L[6]);
    L[4] = fmaf(m003, t203, L[4]);
    L[3] = fmaf(3.0f * m202, t203, L[3]);
    L[1] = fmaf(2.0f * m103, t203, L[1]);
    const float t140 = -fmaf(c1_5, rx * t040 + ry * t130, c2_5 * (t120));
    L[32] = fmaf(q, t140, L[32]);
    L[23] = fmaf(mdx, t140, L[23]);
    L[21] = fmaf(4.0f * mdy, t140, L[21]);
    L[12] = fmaf(6.0f * myy, t140, L[12]);
    L[14] = fmaf(4.0f * mxy, t140, L[14]);
    L[7] = fmaf(6.0f * m120, t140, L[7]);
    L[5] = fmaf(4.0f * m030, t140, L[5]);
    L[2] = fmaf(4.0f * m130, t140, L[2]);
    L[1] = fmaf(m040, t140, L[1]);
    const float t131 = -fmaf(c1_5, rx * t031 + ry * t121 + rz * t130, c2_5 * (t111));
    L[33] = fmaf(q, t131, L[33]);
    L[24] = fmaf(mdx, t131, L[24]);
    L[22] = fmaf(3.0f * mdy, t131, L[22]);
    L[21] = fmaf(mdz, t131, L[21]);
    L[13] = fmaf(3.0f * myy, t131, L[13]);
    L[15] = fmaf(3.0f * mxy, t131, L[15]);
    L[12] = fmaf(3.0f * myz, t131, L[12]);
    L[14] = fmaf(mzx, t131, L[14]);
    L[8] = fmaf(3.0f * m120, t131, L[8]);
    L[7] = fmaf(3.0f * m111, t131, L[7]);
    L[6] = fmaf(m030, t131, L[6]);
    L[5] = fmaf(3.0f * m021, t131, L[5]);
    L[3] = fmaf(m130, t131, L[3]);
    L[2] = fmaf(3.0f * m121, t131, L[2]);
    L[1] = fmaf(m031, t131, L[1]);
    const float t122 = -fmaf(c1_5, rx * t022 + ry * t112 + rz * t121, c2_5 * (t102 + t120));
    L[22] = fmaf(2.0f * mdz, t122, L[22]);
    L[12] = fmaf(mzz, t122, L[12]);
    L[13] = fmaf(4.0f * myz, t122, L[13]);
    L[15] = fmaf(2.0f * mzx, t122, L[15]);
    L[8] = fmaf(4.0f * m111, t122, L[8]);
    L[7] = fmaf(m102, t122, L[7]);
    L[6] = fmaf(2.0f * m021, t122, L[6]);
    L[5] = fmaf(2.0f * m012, t122, L[5]);
    L[3] = fmaf(2.0f * m121, t122, L[3]);
    L[2] = fmaf(2.0f * m112, t122, L[2]);
    L[1] = fmaf(m022, t122, L[1]);
    const float t113 = -fmaf(c1_5, rx * t013 + ry * t103 + rz * t112, c2_5 * (t111));
    L[13] = fmaf(3.0f * mzz, t113, L[13]);
    L[8] = fmaf(3.0f * m102, t113, L[8]);
    L[6] = fmaf(3.0f * m012, t113, L[6]);
    L[5] = fmaf(m003, t113, L[5]);
    L[3] = fmaf(3.0f * m112, t113, L[3]);
    L[2] = fmaf(m103, t113, L[2]);
    L[1] = fmaf(m013, t113, L[1]);
    const float t104 = -fmaf(c1_5, rx * t004 + rz * t103, c2_5 * (t102));
    L[6] = fmaf(4.0f * m003, t104, L[6]);
    L[3] = fmaf(4.0f * m103, t104, L[3]);
    L[1] = fmaf(m004, t104, L[1]);
    const float t050 = -fmaf(c1_5, ry * t040, c2_5 * (t030));
    L[34] = fmaf(q, t050, L[34]);
    L[23] = fmaf(5.0f * mdy, t050, L[23]);
    L[14] = fmaf(10.0f * myy, t050, L[14]);
    L[7] = fmaf(10.0f * m030, t050, L[7]);
    L[2] = fmaf(5.0f * m040, t050, L[2]);
    const float t041 = -fmaf(c1_5, ry * t031 + rz * t040, c2_5 * (t021));
    L[35] = fmaf(q, t041, L[35]);
    L[24] = fmaf(4.0f * mdy, t041, L[24]);
    L[23] = fmaf(mdz, t041, L[23]);
    L[15] = fmaf(6.0f * myy, t041, L[15]);
    L[14] = fmaf(4.0f * myz, t041, L[14]);
    L[8] = fmaf(4.0f * m030, t041, L[8]);
    L[7] = fmaf(6.0f * m021, t041, L[7]);
    L[3] = fmaf(m040, t041, L[3]);
    L[2] = fmaf(4.0f * m031, t041, L[2]);
    const float t032 = -fmaf(c1_5, ry * t022 + rz * t031, c2_5 * (t012 + t030));
    L[24] = fmaf(2.0f * mdz, t032, L[24]);
    L[14] = fmaf(mzz, t032, L[14]);
    L[15] = fmaf(6.0f * myz, t032, L[15]);
    L[8] = fmaf(6.0f * m021, t032, L[8]);
    L[7] = fmaf(3.0f * m012, t032, L[7]);
    L[3] = fmaf(2.0f * m031, t032, L[3]);
    L[2] = fmaf(3.0f * m022, t032, L[2]);
    const float t023 = -fmaf(c1_5, ry * t013 + rz * t022, c2_5 * (t003 + t021));
    L[15] = fmaf(3.0f * mzz, t023, L[15]);
    L[8] = fmaf(6.0f * m012, t023, L[8]);
    L[7] = fmaf(m003, t023, L[7]);
    L[3] = fmaf(3.0f * m022, t023, L[3]);
    L[2] = fmaf(2.0f * m013, t023, L[2]);
    const float t014 = -fmaf(c1_5, ry * t004 + rz * t013, c2_5 * (t012));
    L[8] = fmaf(4.0f * m003, t014, L[8]);
    L[3] = fmaf(4.0f * m013, t014, L[3]);
    L[2] = fmaf(m004, t014, L[2]);
    const float t005 = -fmaf(c1_5, rz * t004, c2_5 * (t003));
    L[3] = fmaf(5.0f * m004, t005, L[3]);
    const float c1_6 = 1.8333333333333333f * ir2, c2_6 = 0.8333333333333334f * ir2;
    const float t600 = -fmaf(c1_6, rx * t500, c2_6 * (t400));
    L[36] = fmaf(q, t600, L[36]);
    L[25] = fmaf(6.0f * mdx, t600, L[25]);
    L[16] = fmaf(15.0f * mxx, t600, L[16]);
    L[9] = fmaf(20.0f * m300, t600, L[9]);
    L[4] = fmaf(15.0f * m400, t600, L[4]);
    const float t510 = -fmaf(c1_6, rx * t410 + ry * t500, c2_6 * (t310));
    L[37] = fmaf(q, t510, L[37]);
    L[26] = fmaf(5.0f * mdx, t510, L[26]);
    L[25] = fmaf(mdy, t510, L[25]);
    L[17] = fmaf(10.0f * mxx, t510, L[17]);
    L[16] = fmaf(5.0f * mxy, t510, L[16]);
    L[10] = fmaf(10.0f * m300, t510, L[10]);
    L[9] = fmaf(10.0f * m210, t510, L[9]);
    L[5] = fmaf(5.0f * m400, t510, L[5]);
    L[4] = fmaf(10.0f * m310, t510, L[4]);
    const float t501 = -fmaf(c1_6, rx * t401 + rz * t500, c2_6 * (t301));
    L[38] = fmaf(q, t501, L[38]);
    L[27] = fmaf(5.0f * mdx, t501, L[27]);
    L[25] = fmaf(mdz, t501, L[25]);
    L[18] = fmaf(10.0f * mxx, t501, L[18]);
    L[16] = fmaf(5.0f * mzx, t501, L[16]);
    L[11] = fmaf(10.0f * m300, t501, L[11]);
    L[9] = fmaf(10.0f * m201, t501, L[9]);
    L[6] = fmaf(5.0f * m400, t501, L[6]);
    L[4] = fmaf(10.0f * m301, t501, L[4]);
    const float t420 = -fmaf(c1_6, rx * t320 + ry * t410, c2_6 * (t220 + t400));
    L[39] = fmaf(q, t420, L[39]);
    L[28] = fmaf(4.0f * mdx, t420, L[28]);
    L[26] = fmaf(2.0f * mdy, t420, L[26]);
    L[19] = fmaf(6.0f * mxx, t420, L[19]);
    L[16] = fmaf(myy, t420, L[16]);
    L[17] = fmaf(8.0f * mxy, t420, L[17]);
    L[12] = fmaf(4.0f * m300, t420, L[12]);
    L[10] = fmaf(12.0f * m210, t420, L[10]);
    L[9] = fmaf(4.0f * m120, t420, L[9]);
    L[7] = fmaf(m400, t420, L[7]);
    L[5] = fmaf(8.0f * m310, t420, L[5]);
    L[4] = fmaf(6.0f * m220, t420, L[4]);
    const float t411 = -fmaf(c1_6, rx * t311 + ry * t401 + rz * t410, c2_6 * (t211));
    L[40] = fmaf(q, t411, L[40]);
    L[29] = fmaf(4.0f * mdx, t411, L[29]);
    L[27] = fmaf(mdy, t411, L[27]);
    L[26] = fmaf(mdz, t411, L[26]);
    L[20] = fmaf(6.0f * mxx, t411, L[20]);
    L[18] = fmaf(4.0f * mxy, t411, L[18]);
    L[16] = fmaf(myz, t411, L[16]);
    L[17] = fmaf(4.0f * mzx, t411, L[17]);
    L[13] = fmaf(4.0f * m300, t411, L[13]);
    L[11] = fmaf(6.0f * m210, t411, L[11]);
    L[10] = fmaf(6.0f * m201, t411, L[10]);
    L[9] = fmaf(4.0f * m111, t411, L[9]);
    L[8] = fmaf(m400, t411, L[8]);
    L[6] = fmaf(4.0f * m310, t411, L[6]);
    L[5] = fmaf(4.0f * m301, t411, L[5]);
    L[4] = fmaf(6.0f * m211, t411, L[4]);
    const float t402 = -fmaf(c1_6, rx * t302 + rz * t401, c2_6 * (t202 + t400));
    L[27] = fmaf(2.0f * mdz, t402, L[27]);
    L[16] = fmaf(mzz, t402, L[16]);
    L[18] = fmaf(8.0f * mzx, t402, L[18]);
    L[11] = fmaf(12.0f * m201, t402, L[11]);
    L[9] = fmaf(4.0f * m102, t402, L[9]);
    L[6] = fmaf(8.0f * m301, t402, L[6]);
    L[4] = fmaf(6.0f * m202, t402, L[4]);
    const float t330 = -fmaf(c1_6, rx * t230 + ry * t320, c2_6 * (t130 + t310));
    L[41] = fmaf(q, t330, L[41]);
    L[30] = fmaf(3.0f * mdx, t330, L[30]);
    L[28] = fmaf(3.0f * mdy, t330, L[28]);
    L[21] = fmaf(3.0f * mxx, t330, L[21]);
    L[17] = fmaf(3.0f * myy, t330, L[17]);
    L[19] = fmaf(9.0f * mxy, t330, L[19]);
    L[14] = fmaf(m300, t330, L[14]);
    L[12] = fmaf(9.0f * m210, t330, L[12]);
    L[10] = fmaf(9.0f * m120, t330, L[10]);
    L[9] = fmaf(m030, t330, L[9]);
    L[7] = fmaf(3.0f * m310, t330, L[7]);
    L[5] = fmaf(9.0f * m220, t330, L[5]);
    L[4] = fmaf(3.0f * m130, t330, L[4]);
    const float t321 = -fmaf(c1_6, rx * t221 + ry * t311 + rz * t320, c2_6 * (t121 + t301));
    L[42] = fmaf(q, t321, L[42]);
    L[31] = fmaf(3.0f * mdx, t321, L[31]);
    L[29] = fmaf(2.0f * mdy, t321, L[29]);
    L[28] = fmaf(mdz, t321, L[28]);
    L[22] = fmaf(3.0f * mxx, t321, L[22]);
    L[18] = fmaf(myy, t321, L[18]);
    L[20] = fmaf(6.0f * mxy, t321, L[20]);
    L[17] = fmaf(2.0f * myz, t321, L[17]);
    L[19] = fmaf(3.0f * mzx, t321, L[19]);
    L[15] = fmaf(m300, t321, L[15]);
    L[13] = fmaf(6.0f * m210, t321, L[13]);
    L[12] = fmaf(3.0f * m201, t321, L[12]);
    L[11] = fmaf(3.0f * m120, t321, L[11]);
    L[10] = fmaf(6.0f * m111, t321, L[10]);
    L[9] = fmaf(m021, t321, L[9]);
    L[8] = fmaf(2.0f * m310, t321, L[8]);
    L[7] = fmaf(m301, t321, L[7]);
    L[6] = fmaf(3.0f * m220, t321, L[6]);
    L[5] = fmaf(6.0f * m211, t321, L[5]);
    L[4] = fmaf(3.0f * m121, t321, L[4]);
    const float t312 = -fmaf(c1_6, rx * t212 + ry * t302 + rz * t311, c2_6 * (t112 + t310));
    L[29] = fmaf(2.0f * mdz, t312, L[29]);
    L[17] = fmaf(mzz, t312, L[17]);
    L[18] = fmaf(2.0f * myz, t312, L[18]);
    L[20] = fmaf(6.0f * mzx, t312, L[20]);
    L[13] = fmaf(6.0f * m201, t312, L[13]);
    L[11] = fmaf(6.0f * m111, t312, L[11]);
    L[10] = fmaf(3.0f * m102, t312, L[10]);
    L[9] = fmaf(m012, t312, L[9]);
    L[8] = fmaf(2.0f * m301, t312, L[8]);
    L[6] = fmaf(6.0f * m211, t312, L[6]);
    L[5] = fmaf(3.0f * m202, t312, L[5]);
    L[4] = fmaf(3.0f * m112, t312, L[4]);
    const float t303 = -fmaf(c1_6, rx * t203 + rz * t302, c2_6 * (t103 + t301));
    L[18] = fmaf(3.0f * mzz, t303, L[18]);
    L[11] = fmaf(9.0f * m102, t303, L[11]);
    L[9] = fmaf(m003, t303, L[9]);
    L[6] = fmaf(9.0f * m202, t303, L[6]);
    L[4] = fmaf(3.0f * m103, t303, L[4]);
    const float t240 = -fmaf(c1_6, rx * t140 + ry * t230, c2_6 * (t040 + t220));
    L[43] = fmaf(q, t240, L[43]);
    L[32] = fmaf(2.0f * mdx, t240, L[32]);
    L[30] = fmaf(4.0f * mdy, t240, L[30]);
    L[23] = fmaf(mxx, t240, L[23]);
    L[19] = fmaf(6.0f * myy, t240, L[19]);
    L[21] = fmaf(8.0f * mxy, t240, L[21]);
    L[14] = fmaf(4.0f * m210, t240, L[14]);
    L[12] = fmaf(12.0f * m120, t240, L[12]);
    L[10] = fmaf(4.0f * m030, t240, L[10]);
    L[7] = fmaf(6.0f * m220, t240, L[7]);
    L[5] = fmaf(8.0f * m130, t240, L[5]);
    L[4] = fmaf(m040, t240, L[4]);
    const float t231 = -fmaf(c1_6, rx * t131 + ry * t221 + rz * t230, c2_6 * (t031 + t211));
    L[44] = fmaf(q, t231, L[44]);
    L[33] = fmaf(2.0f * mdx, t231, L[33]);
    L[31] = fmaf(3.0f * mdy, t231, L[31]);
    L[30] = fmaf(mdz, t231, L[30]);
    L[24] = fmaf(mxx, t231, L[24]);
    L[20] = fmaf(3.0f * myy, t231, L[20]);
    L[22] = fmaf(6.0f * mxy, t231, L[22]);
    L[19] = fmaf(3.0f * myz, t231, L[19]);
    L[21] = fmaf(2.0f * mzx, t231, L[21]);
    L[15] = fmaf(3.0f * m210, t231, L[15]);
    L[14] = fmaf(m201, t231, L[14]);
    L[13] = fmaf(6.0f * m120, t231, L[13]);
    L[12] = fmaf(6.0f * m111, t231, L[12]);
    L[11] = fmaf(m030, t231, L[11]);
    L[10] = fmaf(3.0f * m021, t231, L[10]);
    L[8] = fmaf(3.0f * m220, t231, L[8]);
    L[7] = fmaf(3.0f * m211, t231, L[7]);
    L[6] = fmaf(2.0f * m130, t231, L[6]);
    L[5] = fmaf(6.0f * m121, t231, L[5]);
    L[4] = fmaf(m031, t231, L[4]);
    const float t222 = -fmaf(c1_6, rx * t122 + ry * t212 + rz * t221, c2_6 * (t022 + t202 + t220));
    L[31] = fmaf(2.0f * mdz, t222, L[31]);
    L[19] = fmaf(mzz, t222, L[19]);
    L[20] = fmaf(4.0f * myz, t222, L[20]);
    L[22] = fmaf(4.0f * mzx, t222, L[22]);
    L[15] = fmaf(2.0f * m201, t222, L[15]);
    L[13] = fmaf(8.0f * m111, t222, L[13]);
    L[12] = fmaf(2.0f * m102, t222, L[12]);
    L[11] = fmaf(2.0f * m021, t222, L[11]);
    L[10] = fmaf(2.0f * m012, t222, L[10]);
    L[8] = fmaf(4.0f * m211, t222, L[8]);
    L[7] = fmaf(m202, t222, L[7]);
    L[6] = fmaf(4.0f * m121, t222, L[6]);
    L[5] = fmaf(4.0f * m112, t222, L[5]);
    L[4] = fmaf(m022, t222, L[4]);
    const float t213 = -fmaf(c1_6, rx * t113 + ry * t203 + rz * t212, c2_6 * (t013 + t211));
    L[20] = fmaf(3.0f * mzz, t213, L[20]);
    L[13] = fmaf(6.0f * m102, t213, L[13]);
    L[11] = fmaf(3.0f * m012, t213, L[11]);
    L[10] = fmaf(m003, t213, L[10]);
    L[8] = fmaf(3.0f * m202, t213, L[8]);
    L[6] = fmaf(6.0f * m112, t213, L[6]);
    L[5] = fmaf(2.0f * m103, t213, L[5]);
    L[4] = fmaf(m013, t213, L[4]);
    const float t204 = -fmaf(c1_6, rx * t104 + rz * t203, c2_6 * (t004 + t202));
    L[11] = fmaf(4.0f * m003, t204, L[11]);
    L[6] = fmaf(8.0f * m103, t204, L[6]);
    L[4] = fmaf(m004, t204, L[4]);
    const float t150 = -fmaf(c1_6, rx * t050 + ry * t140, c2_6 * (t130));
    L[45] = fmaf(q, t150, L[45]);
    L[34] = fmaf(mdx, t150, L[34]);
    L[32] = fmaf(5.0f * mdy, t150, L[32]);
    L[21] = fmaf(10.0f * myy, t150, L[21]);
    L[23] = fmaf(5.0f * mxy, t150, L[23]);
    L[14] = fmaf(10.0f * m120, t150, L[14]);
    L[12] = fmaf(10.0f * m030, t150, L[12]);
    L[7] = fmaf(10.0f * m130, t150, L[7]);
    L[5] = fmaf(5.0f * m040, t150, L[5]);
    const float t141 = -fmaf(c1_6, rx * t041 + ry * t131 + rz * t140, c2_6 * (t121));
    L[46] = fmaf(q, t141, L[46]);
    L[35] = fmaf(mdx, t141, L[35]);
    L[33] = fmaf(4.0f * mdy, t141, L[33]);
    L[32] = fmaf(mdz, t141, L[32]);
    L[22] = fmaf(6.0f * myy, t141, L[22]);
    L[24] = fmaf(4.0f * mxy, t141, L[24]);
    L[21] = fmaf(4.0f * myz, t141, L[21]);
    L[23] = fmaf(mzx, t141, L[23]);
    L[15] = fmaf(6.0f * m120, t141, L[15]);
    L[14] = fmaf(4.0f * m111, t141, L[14]);
    L[13] = fmaf(4.0f * m030, t141, L[13]);
    L[12] = fmaf(6.0f * m021, t141, L[12]);
    L[8] = fmaf(4.0f * m130, t141, L[8]);
    L[7] = fmaf(6.0f * m121, t141, L[7]);
    L[6] = fmaf(m040, t141, L[6]);
    L[5] = fmaf(4.0f * m031, t141, L[5]);
    const float t132 = -fmaf(c1_6, rx * t032 + ry * t122 + rz * t131, c2_6 * (t112 + t130));
    L[33] = fmaf(2.0f * mdz, t132, L[33]);
    L[21] = fmaf(mzz, t132, L[21]);
    L[22] = fmaf(6.0f * myz, t132, L[22]);
    L[24] = fmaf(2.0f * mzx, t132, L[24]);
    L[15] = fmaf(6.0f * m111, t132, L[15]);
    L[14] = fmaf(m102, t132, L[14]);
    L[13] = fmaf(6.0f * m021, t132, L[13]);
    L[12] = fmaf(3.0f * m012, t132, L[12]);
    L[8] = fmaf(6.0f * m121, t132, L[8]);
    L[7] = fmaf(3.0f * m112, t132, L[7]);
    L[6] = fmaf(2.0f * m031, t132, L[6]);
    L[5] = fmaf(3.0f * m022, t132, L[5]);
    const float t123 = -fmaf(c1_6, rx * t023 + ry * t113 + rz * t122, c2_6 * (t103 + t121));
    L[22] = fmaf(3.0f * mzz, t123, L[22]);
    L[15] = fmaf(3.0f * m102, t123, L[15]);
    L[13] = fmaf(6.0f * m012, t123, L[13]);
    L[12] = fmaf(m003, t123, L[12]);
    L[8] = fmaf(6.0f * m112, t123, L[8]);
    L[7] = fmaf(m103, t123, L[7]);
    L[6] = fmaf(3.0f * m022, t123, L[6]);
    L[5] = fmaf(2.0f * m013, t123, L[5]);
    const float t114 = -fmaf(c1_6, rx * t014 + ry * t104 + rz * t113, c2_6 * (t112));
    L[13] = fmaf(4.0f * m003, t114, L[13]);
    L[8] = fmaf(4.0f * m103, t114, L[8]);
    L[6] = fmaf(4.0f * m013, t114, L[6]);
    L[5] = fmaf(m004, t114, L[5]);
    const float t105 = -fmaf(c1_6, rx * t005 + rz * t104, c2_6 * (t103));
    L[6] = fmaf(5.0f * m004, t105, L[6]);
    const float t060 = -fmaf(c1_6, ry * t050, c2_6 * (t040));
    L[47] = fmaf(q, t060, L[47]);
    L[34] = fmaf(6.0f * mdy, t060, L[34]);
    L[23] = fmaf(15.0f * myy, t060, L[23]);
    L[14] = fmaf(20.0f * m030, t060, L[14]);
    L[7] = fmaf(15.0f * m040, t060, L[7]);
    const float t051 = -fmaf(c1_6, ry * t041 + rz * t050, c2_6 * (t031));
    L[48] = fmaf(q, t051, L[48]);
    L[35] = fmaf(5.0f * mdy, t051, L[35]);
    L[34] = fmaf(mdz, t051, L[34]);
    L[24] = fmaf(10.0f * myy, t051, L[24]);
    L[23] = fmaf(5.0f * myz, t051, L[23]);
    L[15] = fmaf(10.0f * m030, t051, L[15]);
    L[14] = fmaf(10.0f * m021, t051, L[14]);
    L[8] = fmaf(5.0f * m040, t051, L[8]);
    L[7] = fmaf(10.0f * m031, t051, L[7]);
    const float t042 = -fmaf(c1_6, ry * t032 + rz * t041, c2_6 * (t022 + t040));
    L[35] = fmaf(2.0f * mdz, t042, L[35]);
    L[23] = fmaf(mzz, t042, L[23]);
    L[24] = fmaf(8.0f * myz, t042, L[24]);
    L[15] = fmaf(12.0f * m021, t042, L[15]);
    L[14] = fmaf(4.0f * m012, t042, L[14]);
    L[8] = fmaf(8.0f * m031, t042, L[8]);
    L[7] = fmaf(6.0f * m022, t042, L[7]);
    const float t033 = -fmaf(c1_6, ry * t023 + rz * t032, c2_6 * (t013 + t031));
    L[24] = fmaf(3.0f * mzz, t033, L[24]);
    L[15] = fmaf(9.0f * m012, t033, L[15]);
    L[14] = fmaf(m003, t033, L[14]);
    L[8] = fmaf(9.0f * m022, t033, L[8]);
    L[7] = fmaf(3.0f * m013, t033, L[7]);
    const float t024 = -fmaf(c1_6, ry * t014 + rz * t023, c2_6 * (t004 + t022));
    L[15] = fmaf(4.0f * m003, t024, L[15]);
    L[8] = fmaf(8.0f * m013, t024, L[8]);
    L[7] = fmaf(m004, t024, L[7]);
    const float t015 = -fmaf(c1_6, ry * t005 + rz * t014, c2_6 * (t013));
    L[8] = fmaf(5.0f * m004, t015, L[8]);
    const float c1_7 = 1.8571428571428572f * ir2, c2_7 = 0.8571428571428571f * ir2;
    const float t700 = -fmaf(c1_7, rx * t600, c2_7 * (t500));
    L[36] = fmaf(7.0f * mdx, t700, L[36]);
    L[25] = fmaf(21.0f * mxx, t700, L[25]);
    L[16] = fmaf(35.0f * m300, t700, L[16]);
    L[9] = fmaf(35.0f * m400, t700, L[9]);
    const float t610 = -fmaf(c1_7, rx * t510 + ry * t600, c2_7 * (t410));
    L[37] = fmaf(6.0f * mdx, t610, L[37]);
    L[36] = fmaf(mdy, t610, L[36]);
    L[26] = fmaf(15.0f * mxx, t610, L[26]);
    L[25] = fmaf(6.0f * mxy, t610, L[25]);
    L[17] = fmaf(20.0f * m300, t610, L[17]);
    L[16] = fmaf(15.0f * m210, t610, L[16]);
    L[10] = fmaf(15.0f * m400, t610, L[10]);
    L[9] = fmaf(20.0f * m310, t610, L[9]);
    const float t601 = -fmaf(c1_7, rx * t501 + rz * t600, c2_7 * (t401));
    L[38] = fmaf(6.0f * mdx, t601, L[38]);
    L[36] = fmaf(mdz, t601, L[36]);
    L[27] = fmaf(15.0f * mxx, t601, L[27]);
    L[25] = fmaf(6.0f * mzx, t601, L[25]);
    L[18] = fmaf(20.0f * m300, t601, L[18]);
    L[16] = fmaf(15.0f * m201, t601, L[16]);
    L[11] = fmaf(15.0f * m400, t601, L[11]);
    L[9] = fmaf(20.0f * m301, t601, L[9]);
    const float t520 = -fmaf(c1_7, rx * t420 + ry * t510, c2_7 * (t320 + t500));
    L[39] = fmaf(5.0f * mdx, t520, L[39]);
    L[37] = fmaf(2.0f * mdy, t520, L[37]);
    L[28] = fmaf(10.0f * mxx, t520, L[28]);
    L[25] = fmaf(myy, t520, L[25]);
    L[26] = fmaf(10.0f * mxy, t520, L[26]);
    L[19] = fmaf(10.0f * m300, t520, L[19]);
    L[17] = fmaf(20.0f * m210, t520, L[17]);
    L[16] = fmaf(5.0f * m120, t520, L[16]);
    L[12] = fmaf(5.0f * m400, t520, L[12]);
    L[10] = fmaf(20.0f * m310, t520, L[10]);
    L[9] = fmaf(10.0f * m220, t520, L[9]);
    const float t511 = -fmaf(c1_7, rx * t411 + ry * t501 + rz * t510, c2_7 * (t311));
    L[40] = fmaf(5.0f * mdx, t511, L[40]);
    L[38] = fmaf(mdy, t511, L[38]);
    L[37] = fmaf(mdz, t511, L[37]);
    L[29] = fmaf(10.0f * mxx, t511, L[29]);
    L[27] = fmaf(5.0f * mxy, t511, L[27]);
    L[25] = fmaf(myz, t511, L[25]);
    L[26] = fmaf(5.0f * mzx, t511, L[26]);
    L[20] = fmaf(10.0f * m300, t511, L[20]);
    L[18] = fmaf(10.0f * m210, t511, L[18]);
    L[17] = fmaf(10.0f * m201, t511, L[17]);
    L[16] = fmaf(5.0f * m111, t511, L[16]);
    L[13] = fmaf(5.0f * m400, t511, L[13]);
    L[11] = fmaf(10.0f * m310, t511, L[11]);
    L[10] = fmaf(10.0f * m301, t511, L[10]);
    L[9] = fmaf(10.0f * m211, t511, L[9]);
    const float t502 = -fmaf(c1_7, rx * t402 + rz * t501, c2_7 * (t302 + t500));
    L[38] = fmaf(2.0f * mdz, t502, L[38]);
    L[25] = fmaf(mzz, t502, L[25]);
    L[27] = fmaf(10.0f * mzx, t502, L[27]);
    L[18] = fmaf(20.0f * m201, t502, L[18]);
    L[16] = fmaf(5.0f * m102, t502, L[16]);
    L[11] = fmaf(20.0f * m301, t502, L[11]);
    L[9] = fmaf(10.0f * m202, t502, L[9]);
    const float t430 = -fmaf(c1_7, rx * t330 + ry * t420, c2_7 * (t230 + t410));
    L[41] = fmaf(4.0f * mdx, t430, L[41]);
    L[39] = fmaf(3.0f * mdy, t430, L[39]);
    L[30] = fmaf(6.0f * mxx, t430, L[30]);
    L[26] = fmaf(3.0f * myy, t430, L[26]);
    L[28] = fmaf(12.0f * mxy, t430, L[28]);
    L[21] = fmaf(4.0f * m300, t430, L[21]);
    L[19] = fmaf(18.0f * m210, t430, L[19]);
    L[17] = fmaf(12.0f * m120, t430, L[17]);
    L[16] = fmaf(m030, t430, L[16]);
    L[14] = fmaf(m400, t430, L[14]);
    L[12] = fmaf(12.0f * m310, t430, L[12]);
    L[10] = fmaf(18.0f * m220, t430, L[10]);
    L[9] = fmaf(4.0f * m130, t430, L[9]);
    const float t421 = -fmaf(c1_7, rx * t321 + ry * t411 + rz * t420, c2_7 * (t221 + t401));
    L[42] = fmaf(4.0f * mdx, t421, L[42]);
    L[40] = fmaf(2.0f * mdy, t421, L[40]);
    L[39] = fmaf(mdz, t421, L[39]);
    L[31] = fmaf(6.0f * mxx, t421, L[31]);
    L[27] = fmaf(myy, t421, L[27]);
    L[29] = fmaf(8.0f * mxy, t421, L[29]);
    L[26] = fmaf(2.0f * myz, t421, L[26]);
    L[28] = fmaf(4.0f * mzx, t421, L[28]);
    L[22] = fmaf(4.0f * m300, t421, L[22]);
    L[20] = fmaf(12.0f * m210, t421, L[20]);
    L[19] = fmaf(6.0f * m201, t421, L[19]);
    L[18] = fmaf(4.0f * m120, t421, L[18]);
    L[17] = fmaf(8.0f * m111, t421, L[17]);
    L[16] = fmaf(m021, t421, L[16]);
    L[15] = fmaf(m400, t421, L[15]);
    L[13] = fmaf(8.0f * m310, t421, L[13]);
    L[12] = fmaf(4.0f * m301, t421, L[12]);
    L[11] = fmaf(6.0f * m220, t421, L[11]);
    L[10] = fmaf(12.0f * m211, t421, L[10]);
    L[9] = fmaf(4.0f * m121, t421, L[9]);
    const float t412 = -fmaf(c1_7, rx * t312 + ry * t402 + rz * t411, c2_7 * (t212 + t410));
    L[40] = fmaf(2.0f * mdz, t412, L[40]);
    L[26] = fmaf(mzz, t412, L[26]);
    L[27] = fmaf(2.0f * myz, t412, L[27]);
    L[29] = fmaf(8.0f * mzx, t412, L[29]);
    L[20] = fmaf(12.0f * m201, t412, L[20]);
    L[18] = fmaf(8.0f * m111, t412, L[18]);
    L[17] = fmaf(4.0f * m102, t412, L[17]);
    L[16] = fmaf(m012, t412, L[16]);
    L[13] = fmaf(8.0f * m301, t412, L[13]);
    L[11] = fmaf(12.0f * m211, t412, L[11]);
    L[10] = fmaf(6.0f * m202, t412, L[10]);
    L[9] = fmaf(4.0f * m112, t412, L[9]);
    const float t403 = -fmaf(c1_7, rx * t303 + rz * t402, c2_7 * (t203 + t401));
    L[27] = fmaf(3.0f * mzz, t403, L[27]);
    L[18] = fmaf(12.0f * m102, t403, L[18]);
    L[16] = fmaf(m003, t403, L[16]);
    L[11] = fmaf(18.0f * m202, t403, L[11]);
    L[9] = fmaf(4.0f * m103, t403, L[9]);
    const float t340 = -fmaf(c1_7, rx * t240 + ry * t330, c2_7 * (t140 + t320));
    L[43] = fmaf(3.0f * mdx, t340, L[43]);
    L[41] = fmaf(4.0f * mdy, t340, L[41]);
    L[32] = fmaf(3.0f * mxx, t340, L[32]);
    L[28] = fmaf(6.0f * myy, t340, L[28]);
    L[30] = fmaf(12.0f * mxy, t340, L[30]);
    L[23] = fmaf(m300, t340, L[23]);
    L[21] = fmaf(12.0f * m210, t340, L[21]);
    L[19] = fmaf(18.0f * m120, t340, L[19]);
    L[17] = fmaf(4.0f * m030, t340, L[17]);
    L[14] = fmaf(4.0f * m310, t340, L[14]);
    L[12] = fmaf(18.0f * m220, t340, L[12]);
    L[10] = fmaf(12.0f * m130, t340, L[10]);
    L[9] = fmaf(m040, t340, L[9]);
    const float t331 = -fmaf(c1_7, rx * t231 + ry * t321 + rz * t330, c2_7 * (t131 + t311));
    L[44] = fmaf(3.0f * mdx, t331, L[44]);
    L[42] = fmaf(3.0f * mdy, t331, L[42]);
    L[41] = fmaf(mdz, t331, L[41]);
    L[33] = fmaf(3.0f * mxx, t331, L[33]);
    L[29] = fmaf(3.0f * myy, t331, L[29]);
    L[31] = fmaf(9.0f * mxy, t331, L[31]);
    L[28] = fmaf(3.0f * myz, t331, L[28]);
    L[30] = fmaf(3.0f * mzx, t331, L[30]);
    L[24] = fmaf(m300, t331, L[24]);
    L[22] = fmaf(9.0f * m210, t331, L[22]);
    L[21] = fmaf(3.0f * m201, t331, L[21]);
    L[20] = fmaf(9.0f * m120, t331, L[20]);
    L[19] = fmaf(9.0f * m111, t331, L[19]);
    L[18] = fmaf(m030, t331, L[18]);
    L[17] = fmaf(3.0f * m021, t331, L[17]);
    L[15] = fmaf(3.0f * m310, t331, L[15]);
    L[14] = fmaf(m301, t331, L[14]);
    L[13] = fmaf(9.0f * m220, t331, L[13]);
    L[12] = fmaf(9.0f * m211, t331, L[12]);
    L[11] = fmaf(3.0f * m130, t331, L[11]);
    L[10] = fmaf(9.0f * m121, t331, L[10]);
    L[9] = fmaf(m031, t331, L[9]);
    const float t322 = -fmaf(c1_7, rx * t222 + ry * t312 + rz * t321, c2_7 * (t122 + t302 + t320));
    L[42] = fmaf(2.0f * mdz, t322, L[42]);
    L[28] = fmaf(mzz, t322, L[28]);
    L[29] = fmaf(4.0f * myz, t322, L[29]);
    L[31] = fmaf(6.0f * mzx, t322, L[31]);
    L[22] = fmaf(6.0f * m201, t322, L[22]);
    L[20] = fmaf(12.0f * m111, t322, L[20]);
    L[19] = fmaf(3.0f * m102, t322, L[19]);
    L[18] = fmaf(2.0f * m021, t322, L[18]);
    L[17] = fmaf(2.0f * m012, t322, L[17]);
    L[15] = fmaf(2.0f * m301, t322, L[15]);
    L[13] = fmaf(12.0f * m211, t322, L[13]);
    L[12] = fmaf(3.0f * m202, t322, L[12]);
    L[11] = fmaf(6.0f * m121, t322, L[11]);
    L[10] = fmaf(6.0f * m112, t322, L[10]);
    L[9] = fmaf(m022, t322, L[9]);
    const float t313 = -fmaf(c1_7, rx * t213 + ry * t303 + rz * t312, c2_7 * (t113 + t311));
    L[29] = fmaf(3.0f * mzz, t313, L[29]);
    L[20] = fmaf(9.0f * m102, t313, L[20]);
    L[18] = fmaf(3.0f * m012, t313, L[18]);
    L[17] = fmaf(m003, t313, L[17]);
    L[13] = fmaf(9.0f * m202, t313, L[13]);
    L[11] = fmaf(9.0f * m112, t313, L[11]);
    L[10] = fmaf(3.0f * m103, t313, L[10]);
    L[9] = fmaf(m013, t313, L[9]);
    const float t304 = -fmaf(c1_7, rx * t204 + rz * t303, c2_7 * (t104 + t302));
    L[18] = fmaf(4.0f * m003, t304, L[18]);
    L[11] = fmaf(12.0f * m103, t304, L[11]);
    L[9] = fmaf(m004, t304, L[9]);
    const float t250 = -fmaf(c1_7, rx * t150 + ry * t240, c2_7 * (t050 + t230));
    L[45] = fmaf(2.0f * mdx, t250, L[45]);
    L[43] = fmaf(5.0f * mdy, t250, L[43]);
    L[34] = fmaf(mxx, t250, L[34]);
    L[30] = fmaf(10.0f * myy, t250, L[30]);
    L[32] = fmaf(10.0f * mxy, t250, L[32]);
    L[23] = fmaf(5.0f * m210, t250, L[23]);
    L[21] = fmaf(20.0f * m120, t250, L[21]);
    L[19] = fmaf(10.0f * m030, t250, L[19]);
    L[14] = fmaf(10.0f * m220, t250, L[14]);
    L[12] = fmaf(20.0f * m130, t250, L[12]);
    L[10] = fmaf(5.0f * m040, t250, L[10]);
    const float t241 = -fmaf(c1_7, rx * t141 + ry * t231 + rz * t240, c2_7 * (t041 + t221));
    L[46] = fmaf(2.0f * mdx, t241, L[46]);
    L[44] = fmaf(4.0f * mdy, t241, L[44]);
    L[43] = fmaf(mdz, t241, L[43]);
    L[35] = fmaf(mxx, t241, L[35]);
    L[31] = fmaf(6.0f * myy, t241, L[31]);
    L[33] = fmaf(8.0f * mxy, t241, L[33]);
    L[30] = fmaf(4.0f * myz, t241, L[30]);
    L[32] = fmaf(2.0f * mzx, t241, L[32]);
    L[24] = fmaf(4.0f * m210, t241, L[24]);
    L[23] = fmaf(m201, t241, L[23]);
    L[22] = fmaf(12.0f * m120, t241, L[22]);
    L[21] = fmaf(8.0f * m111, t241, L[21]);
    L[20] = fmaf(4.0f * m030, t241, L[20]);
    L[19] = fmaf(6.0f * m021, t241, L[19]);
    L[15] = fmaf(6.0f * m220, t241, L[15]);
    L[14] = fmaf(4.0f * m211, t241, L[14]);
    L[13] = fmaf(8.0f * m130, t241, L[13]);
    L[12] = fmaf(12.0f * m121, t241, L[12]);
    L[11] = fmaf(m040, t241, L[11]);
    L[10] = fmaf(4.0f * m031, t241, L[10]);
    const float t232 = -fmaf(c1_7, rx * t132 + ry * t222 + rz * t231, c2_7 * (t032 + t212 + t230));
    L[44] = fmaf(2.0f * mdz, t232, L[44]);
    L[30] = fmaf(mzz, t232, L[30]);
    L[31] = fmaf(6.0f * myz, t232, L[31]);
    L[33] = fmaf(4.0f * mzx, t232, L[33]);
    L[24] = fmaf(2.0f * m201, t232, L[24]);
    L[22] = fmaf(12.0f * m111, t232, L[22]);
    L[21] = fmaf(2.0f * m102, t232, L[21]);
    L[20] = fmaf(6.0f * m021, t232, L[20]);
    L[19] = fmaf(3.0f * m012, t232, L[19]);
    L[15] = fmaf(6.0f * m211, t232, L[15]);
    L[14] = fmaf(m202, t232, L[14]);
    L[13] = fmaf(12.0f * m121, t232, L[13]);
    L[12] = fmaf(6.0f * m112, t232, L[12]);
    L[11] = fmaf(2.0f * m031, t232, L[11]);
    L[10] = fmaf(3.0f * m022, t232, L[10]);
    const float t223 = -fmaf(c1_7, rx * t123 + ry * t213 + rz * t222, c2_7 * (t023 + t203 + t221));
    L[31] = fmaf(3.0f * mzz, t223, L[31]);
    L[22] = fmaf(6.0f * m102, t223, L[22]);
    L[20] = fmaf(6.0f * m012, t223, L[20]);
    L[19] = fmaf(m003, t223, L[19]);
    L[15] = fmaf(3.0f * m202, t223, L[15]);
    L[13] = fmaf(12.0f * m112, t223, L[13]);
    L[12] = fmaf(2.0f * m103, t223, L[12]);
    L[11] = fmaf(3.0f * m022, t223, L[11]);
    L[10] = fmaf(2.0f * m013, t223, L[10]);
    const float t214 = -fmaf(c1_7, rx * t114 + ry * t204 + rz * t213, c2_7 * (t014 + t212));
    L[20] = fmaf(4.0f * m003, t214, L[20]);
    L[13] = fmaf(8.0f * m103, t214, L[13]);
    L[11] = fmaf(4.0f * m013, t214, L[11]);
    L[10] = fmaf(m004, t214, L[10]);
    const float t205 = -fmaf(c1_7, rx * t105 + rz * t204, c2_7 * (t005 + t203));
    L[11] = fmaf(5.0f * m004, t205, L[11]);
    const float t160 = -fmaf(c1_7, rx * t060 + ry * t150, c2_7 * (t140));
    L[47] = fmaf(mdx, t160, L[47]);
    L[45] = fmaf(6.0f * mdy, t160, L[45]);
    L[32] = fmaf(15.0f * myy, t160, L[32]);
    L[34] = fmaf(6.0f * mxy, t160, L[34]);
    L[23] = fmaf(15.0f * m120, t160, L[23]);
    L[21] = fmaf(20.0f * m030, t160, L[21]);
    L[14] = fmaf(20.0f * m130, t160, L[14]);
    L[12] = fmaf(15.0f * m040, t160, L[12]);
    const float t151 = -fmaf(c1_7, rx * t051 + ry * t141 + rz * t150, c2_7 * (t131));
    L[48] = fmaf(mdx, t151, L[48]);
    L[46] = fmaf(5.0f * mdy, t151, L[46]);
    L[45] = fmaf(mdz, t151, L[45]);
    L[33] = fmaf(10.0f * myy, t151, L[33]);
    L[35] = fmaf(5.0f * mxy, t151, L[35]);
    L[32] = fmaf(5.0f * myz, t151, L[32]);
    L[34] = fmaf(mzx, t151, L[34]);
    L[24] = fmaf(10.0f * m120, t151, L[24]);
    L[23] = fmaf(5.0f * m111, t151, L[23]);
    L[22] = fmaf(10.0f * m030, t151, L[22]);
    L[21] = fmaf(10.0f * m021, t151, L[21]);
    L[15] = fmaf(10.0f * m130, t151, L[15]);
    L[14] = fmaf(10.0f * m121, t151, L[14]);
    L[13] = fmaf(5.0f * m040, t151, L[13]);
    L[12] = fmaf(10.0f * m031, t151, L[12]);
    const float t142 = -fmaf(c1_7, rx * t042 + ry * t132 + rz * t141, c2_7 * (t122 + t140));
    L[46] = fmaf(2.0f * mdz, t142, L[46]);
    L[32] = fmaf(mzz, t142, L[32]);
    L[33] = fmaf(8.0f * myz, t142, L[33]);
    L[35] = fmaf(2.0f * mzx, t142, L[35]);
    L[24] = fmaf(8.0f * m111, t142, L[24]);
    L[23] = fmaf(m102, t142, L[23]);
    L[22] = fmaf(12.0f * m021, t142, L[22]);
    L[21] = fmaf(4.0f * m012, t142, L[21]);
    L[15] = fmaf(12.0f * m121, t142, L[15]);
    L[14] = fmaf(4.0f * m112, t142, L[14]);
    L[13] = fmaf(8.0f * m031, t142, L[13]);
    L[12] = fmaf(6.0f * m022, t142, L[12]);
    const float t133 = -fmaf(c1_7, rx * t033 + ry * t123 + rz * t132, c2_7 * (t113 + t131));
    L[33] = fmaf(3.0f * mzz, t133, L[33]);
    L[24] = fmaf(3.0f * m102, t133, L[24]);
    L[22] = fmaf(9.0f * m012, t133, L[22]);
    L[21] = fmaf(m003, t133, L[21]);
    L[15] = fmaf(9.0f * m112, t133, L[15]);
    L[14] = fmaf(m103, t133, L[14]);
    L[13] = fmaf(9.0f * m022, t133, L[13]);
    L[12] = fmaf(3.0f * m013, t133, L[12]);
    const float t124 = -fmaf(c1_7, rx * t024 + ry * t114 + rz * t123, c2_7 * (t104 + t122));
    L[22] = fmaf(4.0f * m003, t124, L[22]);
    L[15] = fmaf(4.0f * m103, t124, L[15]);
    L[13] = fmaf(8.0f * m013, t124, L[13]);
    L[12] = fmaf(m004, t124, L[12]);
    const float t115 = -fmaf(c1_7, rx * t015 + ry * t105 + rz * t114, c2_7 * (t113));
    L[13] = fmaf(5.0f * m004, t115, L[13]);
    const float t070 = -fmaf(c1_7, ry * t060, c2_7 * (t050));
    L[47] = fmaf(7.0f * mdy, t070, L[47]);
    L[34] = fmaf(21.0f * myy, t070, L[34]);
    L[23] = fmaf(35.0f * m030, t070, L[23]);
    L[14] = fmaf(35.0f * m040, t070, L[14]);
    const float t061 = -fmaf(c1_7, ry * t051 + rz * t060, c2_7 * (t041));
    L[48] = fmaf(6.0f * mdy, t061, L[48]);
    L[47] = fmaf(mdz, t061, L[47]);
    L[35] = fmaf(15.0f * myy, t061, L[35]);
    L[34] = fmaf(6.0f * myz, t061, L[34]);
    L[24] = fmaf(20.0f * m030, t061, L[24]);
    L[23] = fmaf(15.0f * m021, t061, L[23]);
    L[15] = fmaf(15.0f * m040, t061, L[15]);
    L[14] = fmaf(20.0f * m031, t061, L[14]);
    const float t052 = -fmaf(c1_7, ry * t042 + rz * t051, c2_7 * (t032 + t050));
    L[48] = fmaf(2.0f * mdz, t052, L[48]);
    L[34] = fmaf(mzz, t052, L[34]);
    L[35] = fmaf(10.0f * myz, t052, L[35]);
    L[24] = fmaf(20.0f * m021, t052, L[24]);
    L[23] = fmaf(5.0f * m012, t052, L[23]);
    L[15] = fmaf(20.0f * m031, t052, L[15]);
    L[14] = fmaf(10.0f * m022, t052, L[14]);
    const float t043 = -fmaf(c1_7, ry * t033 + rz * t042, c2_7 * (t023 + t041));
    L[35] = fmaf(3.0f * mzz, t043, L[35]);
    L[24] = fmaf(12.0f * m012, t043, L[24]);
    L[23] = fmaf(m003, t043, L[23]);
    L[15] = fmaf(18.0f * m022, t043, L[15]);
    L[14] = fmaf(4.0f * m013, t043, L[14]);
    const float t034 = -fmaf(c1_7, ry * t024 + rz * t033, c2_7 * (t014 + t032));
    L[24] = fmaf(4.0f * m003, t034, L[24]);
    L[15] = fmaf(12.0f * m013, t034, L[15]);
    L[14] = fmaf(m004, t034, L[14]);
    const float t025 = -fmaf(c1_7, ry * t015 + rz * t024, c2_7 * (t005 + t023));
    L[15] = fmaf(5.0f * m004, t025, L[15]);
}

// potential and field at offset y from the expansion centre (order P = 6, from the
// independent coefficients b_z <= 1; each homogeneous part of a harmonic
// expansion is harmonic, which fixes the others)
__device__ inline void l2p_p6h(const float (&Lh)[49], float x, float y, float z, float &phi, float &ex, float &ey, float &ez)
{
    float L[84];
    L[0] = Lh[0];
    L[1] = Lh[1];
    L[2] = Lh[2];
    L[3] = Lh[3];
    L[4] = Lh[4];
    L[5] = Lh[5];
    L[6] = Lh[6];
    L[7] = Lh[7];
    L[8] = Lh[8];
    L[10] = Lh[9];
    L[11] = Lh[10];
    L[12] = Lh[11];
    L[13] = Lh[12];
    L[14] = Lh[13];
    L[16] = Lh[14];
    L[17] = Lh[15];
    L[20] = Lh[16];
    L[21] = Lh[17];
    L[22] = Lh[18];
    L[23] = Lh[19];
    L[24] = Lh[20];
    L[26] = Lh[21];
    L[27] = Lh[22];
    L[30] = Lh[23];
    L[31] = Lh[24];
    L[35] = Lh[25];
    L[36] = Lh[26];
    L[37] = Lh[27];
    L[38] = Lh[28];
    L[39] = Lh[29];
    L[41] = Lh[30];
    L[42] = Lh[31];
    L[45] = Lh[32];
    L[46] = Lh[33];
    L[50] = Lh[34];
    L[51] = Lh[35];
    L[56] = Lh[36];
    L[57] = Lh[37];
    L[58] = Lh[38];
    L[59] = Lh[39];
    L[60] = Lh[40];
    L[62] = Lh[41];
    L[63] = Lh[42];
    L[66] = Lh[43];
    L[67] = Lh[44];
    L[71] = Lh[45];
    L[72] = Lh[46];
    L[77] = Lh[47];
    L[78] = Lh[48];
    L[9] = -(1.0f * L[4] + 1.0f * L[7]);
    L[15] = -(3.0f * L[10] + 1.0f * L[13]);
    L[18] = -(1.0f * L[11] + 3.0f * L[16]);
    L[25] = -(6.0f * L[20] + 1.0f * L[23]);
    L[28] = -(3.0f * L[21] + 3.0f * L[26]);
    L[32] = -(1.0f * L[23] + 6.0f * L[30]);
    L[40] = -(10.0f * L[35] + 1.0f * L[38]);
    L[43] = -(6.0f * L[36] + 3.0f * L[41]);
    L[47] = -(3.0f * L[38] + 6.0f * L[45]);
    L[52] = -(1.0f * L[41] + 10.0f * L[50]);
    L[61] = -(15.0f * L[56] + 1.0f * L[59]);
    L[64] = -(10.0f * L[57] + 3.0f * L[62]);
    L[68] = -(6.0f * L[59] + 6.0f * L[66]);
    L[73] = -(3.0f * L[62] + 10.0f * L[71]);
    L[79] = -(1.0f * L[66] + 15.0f * L[77]);
    L[19] = -(0.3333333333333333f * L[12] + 0.3333333333333333f * L[17]);
    L[29] = -(1.0f * L[22] + 0.3333333333333333f * L[27]);
    L[33] = -(0.3333333333333333f * L[24] + 1.0f * L[31]);
    L[44] = -(2.0f * L[37] + 0.3333333333333333f * L[42]);
    L[48] = -(1.0f * L[39] + 1.0f * L[46]);
    L[53] = -(0.3333333333333333f * L[42] + 2.0f * L[51]);
    L[65] = -(3.3333333333333335f * L[58] + 0.3333333333333333f * L[63]);
    L[69] = -(2.0f * L[60] + 1.0f * L[67]);
    L[74] = -(1.0f * L[63] + 2.0f * L[72]);
    L[80] = -(0.3333333333333333f * L[67] + 3.3333333333333335f * L[78]);
    L[34] = -(0.16666666666666666f * L[25] + 0.16666666666666666f * L[32]);
    L[49] = -(0.5f * L[40] + 0.16666666666666666f * L[47]);
    L[54] = -(0.16666666666666666f * L[43] + 0.5f * L[52]);
    L[70] = -(1.0f * L[61] + 0.16666666666666666f * L[68]);
    L[75] = -(0.5f * L[64] + 0.5f * L[73]);
    L[81] = -(0.16666666666666666f * L[68] + 1.0f * L[79]);
    L[55] = -(0.1f * L[44] + 0.1f * L[53]);
    L[76] = -(0.3f * L[65] + 0.1f * L[74]);
    L[82] = -(0.1f * L[69] + 0.3f * L[80]);
    L[83] = -(0.06666666666666667f * L[70] + 0.06666666666666667f * L[81]);
    const float x0 = 1.0f, x1 = x;
    const float x2 = x1 * x;
    const float x3 = x2 * x;
    const float x4 = x3 * x;
    const float x5 = x4 * x;
    const float x6 = x5 * x;
    const float y0 = 1.0f, y1 = y;
    const float y2 = y1 * y;
    const float y3 = y2 * y;
    const float y4 = y3 * y;
    const float y5 = y4 * y;
    const float y6 = y5 * y;
    const float z0 = 1.0f, z1 = z;
    const float z2 = z1 * z;
    const float z3 = z2 * z;
    const float z4 = z3 * z;
    const float z5 = z4 * z;
    const float z6 = z5 * z;
    phi = L[0] * (x0 * y0 * z0)
        + L[1] * (x1 * y0 * z0)
        + L[2] * (x0 * y1 * z0)
        + L[3] * (x0 * y0 * z1)
        + L[4] * (x2 * y0 * z0)
        + L[5] * (x1 * y1 * z0)
        + L[6] * (x1 * y0 * z1)
        + L[7] * (x0 * y2 * z0)
        + L[8] * (x0 * y1 * z1)
        + L[9] * (x0 * y0 * z2)
        + L[10] * (x3 * y0 * z0)
        + L[11] * (x2 * y1 * z0)
        + L[12] * (x2 * y0 * z1)
        + L[13] * (x1 * y2 * z0)
        + L[14] * (x1 * y1 * z1)
        + L[15] * (x1 * y0 * z2)
        + L[16] * (x0 * y3 * z0)
        + L[17] * (x0 * y2 * z1)
        + L[18] * (x0 * y1 * z2)
        + L[19] * (x0 * y0 * z3)
        + L[20] * (x4 * y0 * z0)
        + L[21] * (x3 * y1 * z0)
        + L[22] * (x3 * y0 * z1)
        + L[23] * (x2 * y2 * z0)
        + L[24] * (x2 * y1 * z1)
        + L[25] * (x2 * y0 * z2)
        + L[26] * (x1 * y3 * z0)
        + L[27] * (x1 * y2 * z1)
        + L[28] * (x1 * y1 * z2)
        + L[29] * (x1 * y0 * z3)
        + L[30] * (x0 * y4 * z0)
        + L[31] * (x0 * y3 * z1)
        + L[32] * (x0 * y2 * z2)
        + L[33] * (x0 * y1 * z3)
        + L[34] * (x0 * y0 * z4)
        + L[35] * (x5 * y0 * z0)
        + L[36] * (x4 * y1 * z0)
        + L[37] * (x4 * y0 * z1)
        + L[38] * (x3 * y2 * z0)
        + L[39] * (x3 * y1 * z1)
        + L[40] * (x3 * y0 * z2)
        + L[41] * (x2 * y3 * z0)
        + L[42] * (x2 * y2 * z1)
        + L[43] * (x2 * y1 * z2)
        + L[44] * (x2 * y0 * z3)
        + L[45] * (x1 * y4 * z0)
        + L[46] * (x1 * y3 * z1)
        + L[47] * (x1 * y2 * z2)
        + L[48] * (x1 * y1 * z3)
        + L[49] * (x1 * y0 * z4)
        + L[50] * (x0 * y5 * z0)
        + L[51] * (x0 * y4 * z1)
        + L[52] * (x0 * y3 * z2)
        + L[53] * (x0 * y2 * z3)
        + L[54] * (x0 * y1 * z4)
        + L[55] * (x0 * y0 * z5)
        + L[56] * (x6 * y0 * z0)
        + L[57] * (x5 * y1 * z0)
        + L[58] * (x5 * y0 * z1)
        + L[59] * (x4 * y2 * z0)
        + L[60] * (x4 * y1 * z1)
        + L[61] * (x4 * y0 * z2)
        + L[62] * (x3 * y3 * z0)
        + L[63] * (x3 * y2 * z1)
        + L[64] * (x3 * y1 * z2)
        + L[65] * (x3 * y0 * z3)
        + L[66] * (x2 * y4 * z0)
        + L[67] * (x2 * y3 * z1)
        + L[68] * (x2 * y2 * z2)
        + L[69] * (x2 * y1 * z3)
        + L[70] * (x2 * y0 * z4)
        + L[71] * (x1 * y5 * z0)
        + L[72] * (x1 * y4 * z1)
        + L[73] * (x1 * y3 * z2)
        + L[74] * (x1 * y2 * z3)
        + L[75] * (x1 * y1 * z4)
        + L[76] * (x1 * y0 * z5)
        + L[77] * (x0 * y6 * z0)
        + L[78] * (x0 * y5 * z1)
        + L[79] * (x0 * y4 * z2)
        + L[80] * (x0 * y3 * z3)
        + L[81] * (x0 * y2 * z4)
        + L[82] * (x0 * y1 * z5)
        + L[83] * (x0 * y0 * z6);
    ex = -(1.0f * L[1] * (x0 * y0 * z0)
        + 2.0f * L[4] * (x1 * y0 * z0)
        + 1.0f * L[5] * (x0 * y1 * z0)
        + 1.0f * L[6] * (x0 * y0 * z1)
        + 3.0f * L[10] * (x2 * y0 * z0)
        + 2.0f * L[11] * (x1 * y1 * z0)
        + 2.0f * L[12] * (x1 * y0 * z1)
        + 1.0f * L[13] * (x0 * y2 * z0)
        + 1.0f * L[14] * (x0 * y1 * z1)
        + 1.0f * L[15] * (x0 * y0 * z2)
        + 4.0f * L[20] * (x3 * y0 * z0)
        + 3.0f * L[21] * (x2 * y1 * z0)
        + 3.0f * L[22] * (x2 * y0 * z1)
        + 2.0f * L[23] * (x1 * y2 * z0)
        + 2.0f * L[24] * (x1 * y1 * z1)
        + 2.0f * L[25] * (x1 * y0 * z2)
        + 1.0f * L[26] * (x0 * y3 * z0)
        + 1.0f * L[27] * (x0 * y2 * z1)
        + 1.0f * L[28] * (x0 * y1 * z2)
        + 1.0f * L[29] * (x0 * y0 * z3)
        + 5.0f * L[35] * (x4 * y0 * z0)
        + 4.0f * L[36] * (x3 * y1 * z0)
        + 4.0f * L[37] * (x3 * y0 * z1)
        + 3.0f * L[38] * (x2 * y2 * z0)
        + 3.0f * L[39] * (x2 * y1 * z1)
        + 3.0f * L[40] * (x2 * y0 * z2)
        + 2.0f * L[41] * (x1 * y3 * z0)
        + 2.0f * L[42] * (x1 * y2 * z1)
        + 2.0f * L[43] * (x1 * y1 * z2)
        + 2.0f * L[44] * (x1 * y0 * z3)
        + 1.0f * L[45] * (x0 * y4 * z0)
        + 1.0f * L[46] * (x0 * y3 * z1)
        + 1.0f * L[47] * (x0 * y2 * z2)
        + 1.0f * L[48] * (x0 * y1 * z3)
        + 1.0f * L[49] * (x0 * y0 * z4)
        + 6.0f * L[56] * (x5 * y0 * z0)
        + 5.0f * L[57] * (x4 * y1 * z0)
        + 5.0f * L[58] * (x4 * y0 * z1)
        + 4.0f * L[59] * (x3 * y2 * z0)
        + 4.0f * L[60] * (x3 * y1 * z1)
        + 4.0f * L[61] * (x3 * y0 * z2)
        + 3.0f * L[62] * (x2 * y3 * z0)
        + 3.0f * L[63] * (x2 * y2 * z1)
        + 3.0f * L[64] * (x2 * y1 * z2)
        + 3.0f * L[65] * (x2 * y0 * z3)
        + 2.0f * L[66] * (x1 * y4 * z0)
        + 2.0f * L[67] * (x1 * y3 * z1)
        + 2.0f * L[68] * (x1 * y2 * z2)
        + 2.0f * L[69] * (x1 * y1 * z3)
        + 2.0f * L[70] * (x1 * y0 * z4)
        + 1.0f * L[71] * (x0 * y5 * z0)
        + 1.0f * L[72] * (x0 * y4 * z1)
        + 1.0f * L[73] * (x0 * y3 * z2)
        + 1.0f * L[74] * (x0 * y2 * z3)
        + 1.0f * L[75] * (x0 * y1 * z4)
        + 1.0f * L[76] * (x0 * y0 * z5));
    ey = -(1.0f * L[2] * (x0 * y0 * z0)
        + 1.0f * L[5] * (x1 * y0 * z0)
        + 2.0f * L[7] * (x0 * y1 * z0)
        + 1.0f * L[8] * (x0 * y0 * z1)
        + 1.0f * L[11] * (x2 * y0 * z0)
        + 2.0f * L[13] * (x1 * y1 * z0)
        + 1.0f * L[14] * (x1 * y0 * z1)
        + 3.0f * L[16] * (x0 * y2 * z0)
        + 2.0f * L[17] * (x0 * y1 * z1)
        + 1.0f * L[18] * (x0 * y0 * z2)
        + 1.0f * L[21] * (x3 * y0 * z0)
        + 2.0f * L[23] * (x2 * y1 * z0)
        + 1.0f * L[24] * (x2 * y0 * z1)
        + 3.0f * L[26] * (x1 * y2 * z0)
        + 2.0f * L[27] * (x1 * y1 * z1)
        + 1.0f * L[28] * (x1 * y0 * z2)
        + 4.0f * L[30] * (x0 * y3 * z0)
        + 3.0f * L[31] * (x0 * y2 * z1)
        + 2.0f * L[32] * (x0 * y1 * z2)
        + 1.0f * L[33] * (x0 * y0 * z3)
        + 1.0f * L[36] * (x4 * y0 * z0)
        + 2.0f * L[38] * (x3 * y1 * z0)
        + 1.0f * L[39] * (x3 * y0 * z1)
        + 3.0f * L[41] * (x2 * y2 * z0)
        + 2.0f * L[42] * (x2 * y1 * z1)
        + 1.0f * L[43] * (x2 * y0 * z2)
        + 4.0f * L[45] * (x1 * y3 * z0)
        + 3.0f * L[46] * (x1 * y2 * z1)
        + 2.0f * L[47] * (x1 * y1 * z2)
        + 1.0f * L[48] * (x1 * y0 * z3)
        + 5.0f * L[50] * (x0 * y4 * z0)
        + 4.0f * L[51] * (x0 * y3 * z1)
        + 3.0f * L[52] * (x0 * y2 * z2)
        + 2.0f * L[53] * (x0 * y1 * z3)
        + 1.0f * L[54] * (x0 * y0 * z4)
        + 1.0f * L[57] * (x5 * y0 * z0)
        + 2.0f * L[59] * (x4 * y1 * z0)
        + 1.0f * L[60] * (x4 * y0 * z1)
        + 3.0f * L[62] * (x3 * y2 * z0)
        + 2.0f * L[63] * (x3 * y1 * z1)
        + 1.0f * L[64] * (x3 * y0 * z2)
        + 4.0f * L[66] * (x2 * y3 * z0)
        + 3.0f * L[67] * (x2 * y2 * z1)
        + 2.0f * L[68] * (x2 * y1 * z2)
        + 1.0f * L[69] * (x2 * y0 * z3)
        + 5.0f * L[71] * (x1 * y4 * z0)
        + 4.0f * L[72] * (x1 * y3 * z1)
        + 3.0f * L[73] * (x1 * y2 * z2)
        + 2.0f * L[74] * (x1 * y1 * z3)
        + 1.0f * L[75] * (x1 * y0 * z4)
        + 6.0f * L[77] * (x0 * y5 * z0)
        + 5.0f * L[78] * (x0 * y4 * z1)
        + 4.0f * L[79] * (x0 * y3 * z2)
        + 3.0f * L[80] * (x0 * y2 * z3)
        + 2.0f * L[81] * (x0 * y1 * z4)
        + 1.0f * L[82] * (x0 * y0 * z5));
    ez = -(1.0f * L[3] * (x0 * y0 * z0)
        + 1.0f * L[6] * (x1 * y0 * z0)
        + 1.0f * L[8] * (x0 * y1 * z0)
        + 2.0f * L[9] * (x0 * y0 * z1)
        + 1.0f * L[12] * (x2 * y0 * z0)
        + 1.0f * L[14] * (x1 * y1 * z0)
        + 2.0f * L[15] * (x1 * y0 * z1)
        + 1.0f * L[17] * (x0 * y2 * z0)
        + 2.0f * L[18] * (x0 * y1 * z1)
        + 3.0f * L[19] * (x0 * y0 * z2)
        + 1.0f * L[22] * (x3 * y0 * z0)
        + 1.0f * L[24] * (x2 * y1 * z0)
        + 2.0f * L[25] * (x2 * y0 * z1)
        + 1.0f * L[27] * (x1 * y2 * z0)
        + 2.0f * L[28] * (x1 * y1 * z1)
        + 3.0f * L[29] * (x1 * y0 * z2)
        + 1.0f * L[31] * (x0 * y3 * z0)
        + 2.0f * L[32] * (x0 * y2 * z1)
        + 3.0f * L[33] * (x0 * y1 * z2)
        + 4.0f * L[34] * (x0 * y0 * z3)
        + 1.0f * L[37] * (x4 * y0 * z0)
        + 1.0f * L[39] * (x3 * y1 * z0)
        + 2.0f * L[40] * (x3 * y0 * z1)
        + 1.0f * L[42] * (x2 * y2 * z0)
        + 2.0f * L[43] * (x2 * y1 * z1)
        + 3.0f * L[44] * (x2 * y0 * z2)
        + 1.0f * L[46] * (x1 * y3 * z0)
        + 2.0f * L[47] * (x1 * y2 * z1)
        + 3.0f * L[48] * (x1 * y1 * z2)
        + 4.0f * L[49] * (x1 * y0 * z3)
        + 1.0f * L[51] * (x0 * y4 * z0)
        + 2.0f * L[52] * (x0 * y3 * z1)
        + 3.0f * L[53] * (x0 * y2 * z2)
        + 4.0f * L[54] * (x0 * y1 * z3)
        + 5.0f * L[55] * (x0 * y0 * z4)
        + 1.0f * L[58] * (x5 * y0 * z0)
        + 1.0f * L[60] * (x4 * y1 * z0)
        + 2.0f * L[61] * (x4 * y0 * z1)
        + 1.0f * L[63] * (x3 * y2 * z0)
        + 2.0f * L[64] * (x3 * y1 * z1)
        + 3.0f * L[65] * (x3 * y0 * z2)
        + 1.0f * L[67] * (x2 * y3 * z0)
        + 2.0f * L[68] * (x2 * y2 * z1)
        + 3.0f * L[69] * (x2 * y1 * z2)
        + 4.0f * L[70] * (x2 * y0 * z3)
        + 1.0f * L[72] * (x1 * y4 * z0)
        + 2.0f * L[73] * (x1 * y3 * z1)
        + 3.0f * L[74] * (x1 * y2 * z2)
        + 4.0f * L[75] * (x1 * y1 * z3)
        + 5.0f * L[76] * (x1 * y0 * z4)
        + 1.0f * L[78] * (x0 * y5 * z0)
        + 2.0f * L[79] * (x0 * y4 * z1)
        + 3.0f * L[80] * (x0 * y3 * z2)
        + 4.0f * L[81] * (x0 * y2 * z3)
        + 5.0f * L[82] * (x0 * y1 * z4)
        + 6.0f * L[83] * (x0 * y0 * z5));
}
