// walk32.hip -- the mixed-precision force walk (same result as LocalCoulombSum,
// tree.f90:662-717, to a stated force error; not its FP64 operation order).
//
// The walk runs over one joint tree of both species (build.hip): its cells
// carry the net charge, dipole and traceless quadrupole about their geometric
// centre, so in a quasi-neutral plasma the far field is expanded from the
// small net charge distribution rather than as the difference of two large
// per-species fields.
//
// One wavefront per target group: 64*TPL consecutive targets of one species
// in key order.  Lanes test up to 64 candidate cells at once against the
// group's bounding box (a cell is accepted for every target when its distance
// to the box exceeds max(lx)/theta), ballots compact accepted cells and
// leaves into interaction lists in LDS, and opened cells push their
// contiguous child range onto an LDS stack.  Full lists are evaluated by all
// lanes for their own targets in FP32 with group-relative coordinates (formed
// in FP64 once per cell and group); each flush is added into
// Neumaier-compensated FP32 sums.
//
// FP64 is kept only for: cell moments (build.hip), positions and the
// group-relative differences, and the energy sums.
//
// P2P follows the reference's pair rules (tree.f90:806-815, 905-912): Coulomb
// beyond 0.1 um for like charges, simplified Kelbg for unlike ones; the rule
// is uniform across a wave because a group holds one target species.  Work is
// handed out by an atomic queue in one persistent launch over every target
// group of every domain.
#include <hip/hip_runtime.h>

#include "local_exp.h"
#include "tb_mixed.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace {

constexpr int BT = 256;
constexpr int WPB = BT / 64;
constexpr int STACK = 512;
// Interaction lists are flushed once FLUSH_M (cells) or FLUSH_P (leaves)
// entries are pending.  Each pass of the walk loop (a traversal step, or a
// learned band decision) adds at most 64 entries to a list before the flush
// checks, so a list of capacity FLUSH - 1 + 64 holds them.  The queues (band cells, far
// cells) are flushed 64 at a time: QCAP = 63 + 64.  The LDS per wave
// (13.0 KB) keeps three 4-wave blocks on a CU; cell entries (80 bytes, up to
// the third moment) are the largest part.
// FLUSH_M / FLUSH_P are the largest list lengths (they size the lists); the
// walk flushes at the lengths of one of TB_LIST_VARIANTS, chosen by the
// autotuner with the group size (shorter lists: more flushes of fewer
// entries, each flush's compensated sums and loop set-up paid more often).
constexpr int FLUSH_M = 11;
constexpr int LCAP_M = FLUSH_M - 1 + 64;
constexpr int FLUSH_P = 24;
constexpr int LCAP_P = FLUSH_P - 1 + 64;
constexpr int LIST_M[TB_LIST_VARIANTS] = {11, 9, 7};
constexpr int LIST_P[TB_LIST_VARIANTS] = {24, 16, 12};
constexpr int QCAP = 128;
// Unlike leaves within 22 rs of some target of the group: up to NEXACT of
// them per group are evaluated after the walk with separations from the FP64
// positions; more (a cluster of bound pairs near one group) fall back to
// FP32 group-relative coordinates, accurate to ~6e-8 of the group's size.
constexpr int NEXACT = 8;
constexpr int NXCD = 8;       // XCDs of an MI355X (work chunks of the walk queue)
constexpr int CTR_CHUNK = 8;  // first chunk counter in the counter block (16 ints)
constexpr int CTR_FAR = 4;    // far-chunk allocation counter
constexpr int FAR_SLAB = 256; // far chunks a wave reserves at a time
constexpr int BAND_MAXCNT = 4096;  // largest band cell that is labelled / decided by the network
// A cell of at most SMALL_OPEN particles is always opened: its particles'
// P2P (15 VALU instructions per pair) is exact and cheaper than one M2P to
// the fourth moment (~120 per target).  Measured at N = 4M, 1e-4 rms, far
// threshold 4, theta re-tuned (0.395-0.397 for each): walk 40.0 ms at 2
// particles, 37.6 at 4, 37.5 at 6, 38.1 at 8.  Far cells are opened too, so
// that the far path changes how accepted cells are evaluated, never which
// are accepted.
constexpr int SMALL_OPEN = 5;
// An accepted cell farther than OCT_RADIUS of its opening radii (s/theta)
// from the group's box takes the M2P to the third moment: there its fourth
// moment's term, falling as (s/d)^5, is below the truncation error of the
// cells at the opening radius.  Measured at N = 4M, groups of 128, theta
// bisected to 1e-4 for each: with the fourth moment dropped beyond 1.26 /
// 1.5 / 1.75 / 2.0 radii, theta 0.369 / 0.393 / 0.395 / 0.395 (0.395
// without), 84 / 46 / 26 / 9% of the M2P pairs to the third moment.
constexpr float OCT_RADIUS = 1.5f;
// In the learned walk the children of a band cell the network opened take the
// full M2P whatever their distance, so that opening a band cell never costs
// accuracy (with the third-moment M2P for its children it could: one
// fourth-moment M2P of the cell against eight third-moment ones).  The flag
// rides in bit 31 of the cell's child range on the stack; child indices are
// below 2^28 (bit 31 of a twig's record is clear: it marks leaves).
constexpr uint32_t FULL4 = 0x80000000u;

const char *g_err = "";

// far-field local expansion: order and coefficient count
constexpr int FAR_ORDER = 6;
constexpr int NLOC = FAR_ORDER == 3   ? TB_LOCAL_N3
                     : FAR_ORDER == 4 ? TB_LOCAL_N4H
                     : FAR_ORDER == 5 ? TB_LOCAL_N5H
                                      : TB_LOCAL_N6H;
// default threshold: cell distance / group radius.  The field's relative
// truncation error of the order-P expansion falls as (1/K)^P at the group's
// edge, and a higher order costs more per far cell.  Measured at N = 4M,
// groups of 128, theta re-tuned to 1e-4 rms for each (order, K), walk + far
// evaluation, cells to the third moment: order 4 at K = 6, 64.5 ms; order 5
// at K = 4, 58.4 ms; order 6 at K = 3, 55.5 ms (K = 3.5: 57.1; K = 2.5 needs
// theta below 0.22); order 7 at K = 2.5 or 3, 65 ms (its far evaluation
// costs more than it saves).  Order 6 keeping only the terms of total order
// |b| + |c| <= 7 (2 of the 3 orders of moment-times-coefficient products
// beyond it dropped): 53.0 ms at K = 3 against 56.2 for the full order 6 on
// the same GPU (<= 8: 53.9; <= 6: theta falls to 0.256 and the gain is lost).
// With the cells' fourth moment (in the M2P and in the far cells' M2L) theta
// rises to ~0.39 and the balance moves: K = 3, 38.8 ms; K = 4, 37.5 ms; K = 5,
// 38.6 ms (cells of up to 6 particles opened); K = 2.5 reaches 1e-4 at no theta
// below 0.5.  Without the fourth moment in the M2L, theta stayed at 0.28.  With
// k_far_eval at two waves per SIMD (was one) the far path costs less and the
// balance moves again (groups of 128): K = 3, 34.6 ms; 3.25, 33.5; 3.5, 32.7-33.2;
// 3.75, 34.5; 4, 34.1; 5, 35.5 (groups of 64: 39.6 / 34.8 / 35.0 / 36.6 at K = 3 /
// 3.5 / 4 / 5).
constexpr double FAR_RATIO_DEFAULT = 3.5;

struct WaveLDS {
    uint32_t stack[STACK];
    uint32_t batch[64];
    uint32_t bandq[QCAP];                 // learned: band cells waiting for a decision
    uint32_t farq[QCAP];                  // far accepted cells waiting for the local expansion
    // accepted cells: x y z q | Dx Dy Dz Q'xx | Q'yy Q'xy Q'yz Q'zx |
    // O'xxx O'xxy O'xxz O'xyy | O'xyz O'yyy O'yyz H'xxxx |
    // H'xxxy H'xxxz H'xxyy H'xxyz | H'xyyy H'xyyz H'yyyy H'yyyz
    float4 m0[LCAP_M], m1[LCAP_M], m2[LCAP_M], m3[LCAP_M], m4[LCAP_M], m5[LCAP_M], m6[LCAP_M];
    float4 p[LCAP_P];                     // leaves x y z, record index: like charges from 0 up, unlike from LCAP_P - 1 down
    uint32_t exact[NEXACT];               // leaves within 22 rs of some target: FP64 separations after the walk
};
// measured: 4 x 13056 bytes per block kept 3 blocks on a CU, 4 x 13568 did not
static_assert(sizeof(WaveLDS) * 4 <= 4 * 13056, "walk LDS would drop a CU to two blocks");

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline int lane_id() { return threadIdx.x & 63; }

// A value that every lane holds alike, moved to a scalar register: the
// per-group constants of the walk then take no vector registers, and values
// derived from a scalar group index (the domain descriptor) load as scalars.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ inline double uni(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline uint64_t lanemask_lt()
{
    int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ inline int wave_incl_scan(int v)
{
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (l >= o) v += t;
    }
    return v;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Neumaier compensated add: (s, c) += v
__device__ inline void nadd(float &s, float &c, float v)
{
    float t = s + v;
    c += (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
    s = t;
}

// A group's box: half extents from its min / max, rounded outwards, and the
// squared distance of a point (offset from the box centre) to the box.  The
// walk and the segment walk form them alike (no contraction left to the
// compiler), so their acceptance tests agree bit for bit.
__device__ inline float half_extent(double mn, double mx) { return fmaf((float)(0.5 * (mx - mn)), 1.000001f, 1e-30f); }

__device__ inline float box_d2(float cx, float cy, float cz, float hx, float hy, float hz)
{
    const float ddx = fmaxf(fabsf(cx) - hx, 0.0f);
    const float ddy = fmaxf(fabsf(cy) - hy, 0.0f);
    const float ddz = fmaxf(fabsf(cz) - hz, 0.0f);
    return fmaf(ddx, ddx, fmaf(ddy, ddy, ddz * ddz));
}

// 1 - e^{-u}(1+u) and 1 - e^{-u}, with series below u = 1/8 where the
// subtraction would cancel; both round to 1 in FP32 from u = 22 on
// (e^{-22} (1 + 22) < 2^-25, half an ulp below 1), which skips the
// exponential for all but close unlike pairs
__device__ inline void kelbg_gh(float u, float &g, float &h)
{
    if (u >= 22.0f) {
        g = h = 1.0f;
    } else if (u < 0.125f) {
        g = u * u * (0.5f + u * (-1.0f / 3.0f + u * (0.125f + u * (-1.0f / 30.0f + u * (1.0f / 144.0f +
                                                                                    u * (-1.0f / 840.0f))))));
        h = u * (1.0f + u * (-0.5f + u * (1.0f / 6.0f + u * (-1.0f / 24.0f + u * (1.0f / 120.0f +
                                                                                 u * (-1.0f / 720.0f))))));
    } else {
        float e = __expf(-u);
        g = 1.0f - e * (1.0f + u);
        h = 1.0f - e;
    }
}

// The domain descriptors are read-only while a walk runs: read through the
// constant address space, so their (wave-uniform) fields come by scalar loads
// and are not re-read from global memory after each of the walk's stores
typedef const __attribute__((address_space(4))) Domain CDomain;
__device__ inline CDomain *cdom(const Domain *p) { return (CDomain *)p; }

struct WalkArgs {
    const Domain *dom;    // ndom descriptors, ordered by gbase
    int ndom, ngt;        // domains, groups in all domains
    float lo_scale;       // learned / collect: theta / theta_lo; cells beyond lo_scale * s/theta are always accepted
    float log_err;        // learned: accept a band cell when the predicted log10 error is below this
    float cost_w;         // learned: ... plus cost_w * log10 of the work its acceptance saves (in M2Ps)
    float far_k2;         // square of the local-expansion threshold (distance / group radius); 0: off
    uint32_t *far_cells;  // far cells, 64 per chunk (0xffffffff: none)
    int32_t *far_link;    // per chunk: the group's previous chunk, or -1
    int32_t *far_head;    // per global group: its last chunk, or -1
    float *gemin;         // learned / collect: per global group, emin of the first attempt
    int emin_saved;       // 1 on a re-run: emin comes from gemin (the forces were overwritten)
    int far_cap;          // chunks in far_cells
    const TbMlpPack *mlpk;  // learned: the bf16 parameters packed as matrix-core operands
    BandRec *rec;         // collect: band pairs
    int rec_cap;
    int32_t *ctr;         // [1] error flag, [2] prep error, [3] band pair count, [4] far chunks, [8..15] work chunks
    int flush_m, flush_p; // list lengths at which the M2P / P2P lists are evaluated (<= FLUSH_M / FLUSH_P)
    // two-level far field: segments (super-groups of <= 8 groups inside one
    // cell) walk the top of the tree once (k_swalk); their groups start from
    // the segment's frontier, and cells far from the whole segment go through
    // the segment's own local expansion
    int two_level;
    int nseg0, nsegt;     // electron segments, all segments
    const int2 *sgrp[2];  // per species: segment -> first group, group count (combined indices)
    const int2 *cgrp[2];  // per species: the combined target groups
    int32_t *sfront;      // per segment: last frontier chunk (-1: none)
    int32_t *sfar;        // per segment: last chunk of its far cells (-1: none)
    double4 *scen;        // per segment: centre, half-diagonal
    double *gen;          // 2 per global group: energy of electron targets / of ion targets
    unsigned long long *cnt;
};

constexpr int MODE_FIXED = 0, MODE_LEARNED = 1, MODE_COLLECT = 2;

// features of a (cell, target group) pair, log2 scaled by 1/8 (k):
//   f0 s/d; f1 |q|/count; f2 |D|/(count s); f3 |Q'|/(count s^2); f4 count;
//   f5 group half-diagonal h / d; f6 emin d^2 / count; f7 +-targets in the
//   group (+ when the cell's net charge is like the targets').
// f1..f4 depend on the cell only and come precomputed in FP64 (bf, prep);
// k log2 s sits in bq2.w.  emin is the smallest net field |E| = |F| / eps
// over the group's targets from the previous force: the label is an error
// relative to the target's net force, which in a quasi-neutral plasma is far
// below the cell's gross field count / d^2.  Per group: klemin = k log2 emin,
// knact = k log2 nact.
__device__ inline void band_features(float4 bf, float kls, float d, float h, float klemin, bool like_net,
                                     float knact, float *f)
{
    const float k = 0.125f;
    const float kld = k * __log2f(d);
    f[0] = kls - kld;
    f[1] = bf.x;
    f[2] = bf.y;
    f[3] = bf.z;
    f[4] = bf.w;
    f[5] = k * __log2f(h / d + 9.5367431640625e-07f);
    f[6] = klemin + 2.0f * kld - bf.w;
    f[7] = like_net ? knact : -knact;
}

// P2P per pair against one M2P to the fourth moment, per target, in VALU
// instructions (15 against ~120, from the compiled flush loops)
constexpr float P2P_PER_M2P = 0.125f;

// log10 of the work that accepting band cell nd saves, in units of one M2P
// per target: opening it costs its children's evaluation (a leaf or a cell
// of at most SMALL_OPEN particles: P2P per particle; a larger cell: one
// M2P, as it is then half the size at the same distance) against the one
// M2P of accepting it.  No saving (opening is no dearer) gives -30, so the
// cell is opened.  A child range that is not strictly after the cell is
// left to the caller's check (no children read).
template <class Dom>
__device__ inline float band_saving_log10(const Dom &D, uint32_t nd)
{
    const uint32_t inf = D.binfo[nd];
    const uint32_t st = inf >> 3, nch = (inf & 7u) + 1u;
    if (st <= nd || st + nch > (uint32_t)D.nn) return 0.0f;
    float open = 0.0f;
    for (uint32_t k = 0; k < nch; k++) {
        const uint32_t ci = D.binfo[st + k];
        const float cnt = (ci & TB_LEAF) ? 1.0f : D.bq2[st + k].z;
        open += cnt <= (float)SMALL_OPEN ? cnt * P2P_PER_M2P : 1.0f;
    }
    const float save = open - 1.0f;
    return save > 0.0f ? __log10f(save) : -30.0f;
}

template <int TPL>
struct Acc {
    float s[TPL][3], c[TPL][3];  // compensated field sums
    float ps[TPL], pc[TPL];      // compensated potential
};

// monopole + dipole + traceless quadrupole + traceless third and fourth
// moments about the cell centre: phi = q/r + D.r/r^3 + r.Q'.r/(2 r^5) +
// 5/2 O'rrr/r^7 + 35/8 H'rrrr/r^9, E = -grad phi (the O' part: -15/2 v/r^7 +
// 35/2 (r.v) r/r^9, v_i = O'_ijk r_j r_k; the H' part: -35/2 w/r^9 + 315/8
// (r.w) r/r^11, w_i = H'_ijkl r_j r_k r_l)
// One M2P loop over list entries [j0, j1): to the fourth moment (HEX) or to
// the third (cells far enough from the group that their H' term is below the
// walk's error, see OCT_RADIUS; entries without m5 / m6)
template <int TPL, bool HEX>
__device__ inline void m2p_loop(const WaveLDS &L, int j0, int j1, const float (&rx)[TPL], const float (&ry)[TPL],
                                const float (&rz)[TPL], float (&ax)[TPL], float (&ay)[TPL], float (&az)[TPL],
                                float (&ap)[TPL])
{
    for (int j = j0; j < j1; j++) {
        const float4 a = L.m0[j], b = L.m1[j], c = L.m2[j], o1 = L.m3[j], o2 = L.m4[j];
        float4 h1 = make_float4(0.f, 0.f, 0.f, 0.f), h2 = h1;
        if constexpr (HEX) {
            h1 = L.m5[j];
            h2 = L.m6[j];
        }
        const float qxx = b.w, qyy = c.x, qzz = -b.w - c.x, qxy = c.y, qyz = c.z, qzx = c.w;
        // O' from its 7 independent components (tr = 0: O'_xzz = -(O'_xxx +
        // O'_xyy), O'_yzz = -(O'_xxy + O'_yyy), O'_zzz = -(O'_xxz + O'_yyz)),
        // so v = O'rr needs xx - zz, yy - zz and the doubled mixed products
        const float oxxx = o1.x, oxxy = o1.y, oxxz = o1.z, oxyy = o1.w, oxyz = o2.x, oyyy = o2.y, oyyz = o2.z;
        const float sxy = oxxx + oxyy, syz = oxxy + oyyy;
        // H' from its 9 independent components; w = H'rrr in the basis of
        // the cubic harmonics A1..A7 below (tr = 0 folded in)
        const float hxxxx = o2.w, hxxxy = h1.x, hxxxz = h1.y, hxxyy = h1.z, hxxyz = h1.w, hxyyy = h2.x,
                    hxyyz = h2.y, hyyyy = h2.z, hyyyz = h2.w;
        const float s1 = hxxxx + hxxyy, s2 = hxxxy + hxyyy, s3 = hxxyy + hyyyy;
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            const float dx = rx[t] - a.x, dy = ry[t] - a.y, dz = rz[t] - a.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            // raw v_rsq_f32: an accepted cell is at least s/theta away, so r2
            // is a normal number and the denormal-safe rsqrtf is not needed
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float u = ri * ri, ri3 = ri * u;
            const float dp = b.x * dx + b.y * dy + b.z * dz;
            const float qx = qxx * dx + qxy * dy + qzx * dz;
            const float qy = qxy * dx + qyy * dy + qyz * dz;
            const float qz = qzx * dx + qyz * dy + qzz * dz;
            const float xqx = dx * qx + dy * qy + dz * qz;
            const float zz = dz * dz, xmz = dx * dx - zz, ymz = dy * dy - zz;
            const float dx2 = dx + dx, xy2 = dx2 * dy, xz2 = dx2 * dz, yz2 = (dy + dy) * dz;
            const float vx = oxxx * xmz + oxyy * ymz + oxxy * xy2 + oxxz * xz2 + oxyz * yz2;
            const float vy = oxxy * xmz + oyyy * ymz + oxyy * xy2 + oxyz * xz2 + oyyz * yz2;
            const float vz = oxxz * xmz + oyyz * ymz + oxyz * xy2 - sxy * xz2 - syz * yz2;
            const float t3 = u * (dx * vx + dy * vy + dz * vz);
            const float u75 = 7.5f * u;
            if constexpr (HEX) {
                const float zz2 = zz + zz;
                const float A1 = dx * fmaf(-2.0f, zz, xmz), A2 = 3.0f * (dy * xmz), A3 = dz * fmaf(3.0f, xmz, zz2);
                const float A4 = 3.0f * (dx * ymz), A5 = 3.0f * (xy2 * dz), A6 = dy * fmaf(-2.0f, zz, ymz);
                const float A7 = dz * fmaf(3.0f, ymz, zz2);
                const float wx = hxxxx * A1 + hxxxy * A2 + hxxxz * A3 + hxxyy * A4 + hxxyz * A5 + hxyyy * A6 + hxyyz * A7;
                const float wy = hxxxy * A1 + hxxyy * A2 + hxxyz * A3 + hxyyy * A4 + hxyyz * A5 + hyyyy * A6 + hyyyz * A7;
                const float wz = hxxxz * A1 + hxxyz * A2 + hxyyz * A4 + hyyyz * A6 - s1 * A3 - s2 * A5 - s3 * A7;
                const float t4 = u * u * (dx * wx + dy * wy + dz * wz);
                // E = dx ri3 (q + u (3 dp + u (2.5 xqx + 17.5 t3 + 39.375 t4))) - ri3 (D + u (Q.dx + u (7.5 v + 17.5 u w)))
                // phi = ri (q + u (dp + u (0.5 xqx + 2.5 t3 + 4.375 t4)))
                const float a5 = u * fmaf(39.375f, t4, fmaf(17.5f, t3, 2.5f * xqx));
                const float w = ri3 * fmaf(u, fmaf(3.0f, dp, a5), a.w);
                const float u175 = 17.5f * u * u;
                ax[t] = fmaf(-ri3, fmaf(u, fmaf(u175, wx, fmaf(u75, vx, qx)), b.x), fmaf(dx, w, ax[t]));
                ay[t] = fmaf(-ri3, fmaf(u, fmaf(u175, wy, fmaf(u75, vy, qy)), b.y), fmaf(dy, w, ay[t]));
                az[t] = fmaf(-ri3, fmaf(u, fmaf(u175, wz, fmaf(u75, vz, qz)), b.z), fmaf(dz, w, az[t]));
                ap[t] = fmaf(ri, fmaf(u, dp + u * fmaf(4.375f, t4, fmaf(2.5f, t3, 0.5f * xqx)), a.w), ap[t]);
            } else {
                // E = dx ri3 (q + u (3 dp + u (2.5 xqx + 17.5 t3))) - ri3 (D + u (Q.dx + 7.5 u v))
                // phi = ri (q + u (dp + u (0.5 xqx + 2.5 t3)))
                const float a5 = u * fmaf(17.5f, t3, 2.5f * xqx);
                const float w = ri3 * fmaf(u, fmaf(3.0f, dp, a5), a.w);
                ax[t] = fmaf(-ri3, fmaf(u, fmaf(u75, vx, qx), b.x), fmaf(dx, w, ax[t]));
                ay[t] = fmaf(-ri3, fmaf(u, fmaf(u75, vy, qy), b.y), fmaf(dy, w, ay[t]));
                az[t] = fmaf(-ri3, fmaf(u, fmaf(u75, vz, qz), b.z), fmaf(dz, w, az[t]));
                ap[t] = fmaf(ri, fmaf(u, dp + u * fmaf(2.5f, t3, 0.5f * xqx), a.w), ap[t]);
            }
        }
    }
}

// The M2P list: nh cells to the fourth moment from the bottom of the list,
// nl cells to the third from the top ([LCAP_M - nl, LCAP_M)); one
// compensated add per list.
template <int TPL>
__device__ inline void flush_m2p(const WaveLDS &L, int nh, int nl, const float (&rx)[TPL], const float (&ry)[TPL],
                                 const float (&rz)[TPL], Acc<TPL> &A)
{
    float ax[TPL], ay[TPL], az[TPL], ap[TPL];
#pragma unroll
    for (int t = 0; t < TPL; t++) ax[t] = ay[t] = az[t] = ap[t] = 0.0f;
    m2p_loop<TPL, true>(L, 0, nh, rx, ry, rz, ax, ay, az, ap);
    m2p_loop<TPL, false>(L, LCAP_M - nl, LCAP_M, rx, ry, rz, ax, ay, az, ap);
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        nadd(A.s[t][0], A.c[t][0], ax[t]);
        nadd(A.s[t][1], A.c[t][1], ay[t]);
        nadd(A.s[t][2], A.c[t][2], az[t]);
        nadd(A.ps[t], A.pc[t], ap[t]);
    }
}

// Leaves (one particle each, charge +-1).  Like charges (the targets'
// species) sit at the bottom of L.p, [0, nl), unlike ones at the top,
// [LCAP_P - nu, LCAP_P): the Coulomb loop (with the 0.1 um cut) and the Kelbg
// loop then run without a per-entry branch on the charge.  Every like source
// has charge qt and every unlike one -qt, so both loops sum q-free terms and
// the sign goes on once at the end.  An entry's w holds its leaf's record
// index (uint32 bits): the charge follows from the list it is in.
//
// The close test of an unlike pair (r < 22 rs) is one expression, repeated
// alike wherever a pair is classified, so every pair is counted once.
__device__ inline bool close_pair(float dx, float dy, float dz, float close2, float &r2, float &rq)
{
    r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    rq = __builtin_amdgcn_rsqf(r2);
    return r2 < close2;  // r < 22 rs: close2 = (22 rs)^2
}

template <int TPL>
__device__ inline void flush_p2p(WaveLDS &L, int nl, int nu, const float (&rx)[TPL], const float (&ry)[TPL],
                                 const float (&rz)[TPL], float qt, float inv_rs, float close2, Acc<TPL> &A,
                                 int &nx)
{
    float ax[TPL], ay[TPL], az[TPL], ap[TPL];
#pragma unroll
    for (int t = 0; t < TPL; t++) ax[t] = ay[t] = az[t] = ap[t] = 0.0f;
#pragma unroll 4
    for (int j = 0; j < nl; j++) {
        const float4 q = L.p[j];
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            const float dx = rx[t] - q.x, dy = ry[t] - q.y, dz = rz[t] - q.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            // also drops the target itself; beyond the cut r2 is normal (raw v_rsq_f32)
            const float ri = r2 > 0.01f ? __builtin_amdgcn_rsqf(r2) : 0.0f;
            const float w = ri * ri * ri;
            ax[t] = fmaf(dx, w, ax[t]);
            ay[t] = fmaf(dy, w, ay[t]);
            az[t] = fmaf(dz, w, az[t]);
            ap[t] += ri;
        }
    }
    // unlike pairs: beyond 22 rs the Kelbg factors are 1 in FP32 and the term
    // is Coulomb's; a closer pair (rare: at 1e14-1e15 m^-3 about one per
    // thousand flushes) is left out here and added by the pass below, which
    // runs only when some lane of the wave has one
    bool close = false;
#pragma unroll 4
    for (int j = LCAP_P - nu; j < LCAP_P; j++) {
        const float4 q = L.p[j];
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            const float dx = rx[t] - q.x, dy = ry[t] - q.y, dz = rz[t] - q.z;
            float r2, rq;
            const bool cl = close_pair(dx, dy, dz, close2, r2, rq);
            close = close || cl;
            const float ri = cl ? 0.0f : rq;
            const float w = ri * ri * ri;
            ax[t] = fmaf(-dx, w, ax[t]);
            ay[t] = fmaf(-dy, w, ay[t]);
            az[t] = fmaf(-dz, w, az[t]);
            ap[t] -= ri;
        }
    }
    if (__ballot(close)) {
        // leaves close to some target: queued for the exact pass after the
        // walk while it has room, else evaluated here in FP32
        for (int j = LCAP_P - nu; j < LCAP_P; j++) {
            const float4 q = L.p[j];
            bool any = false;
#pragma unroll
            for (int t = 0; t < TPL; t++) {
                float r2, rq;
                any = any || close_pair(rx[t] - q.x, ry[t] - q.y, rz[t] - q.z, close2, r2, rq);
            }
            if (!__ballot(any)) continue;  // wave-uniform
            if (nx < NEXACT) {
                L.exact[nx] = __float_as_uint(q.w);  // every lane stores the same value
                nx++;
                continue;
            }
#pragma unroll
            for (int t = 0; t < TPL; t++) {
                const float dx = rx[t] - q.x, dy = ry[t] - q.y, dz = rz[t] - q.z;
                float r2, rq;
                if (close_pair(dx, dy, dz, close2, r2, rq)) {
                    float g, h;
                    kelbg_gh(r2 * rq * inv_rs, g, h);
                    const float w = g * rq * rq * rq;
                    ax[t] = fmaf(-dx, w, ax[t]);
                    ay[t] = fmaf(-dy, w, ay[t]);
                    az[t] = fmaf(-dz, w, az[t]);
                    ap[t] = fmaf(-h, rq, ap[t]);
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        nadd(A.s[t][0], A.c[t][0], qt * ax[t]);
        nadd(A.s[t][1], A.c[t][1], qt * ay[t]);
        nadd(A.s[t][2], A.c[t][2], qt * az[t]);
        nadd(A.ps[t], A.pc[t], qt * ap[t]);
    }
}

// The exact pass over the group's queued close leaves, after the walk: the
// same close test as in flush_p2p (same FP32 operands: the leaf's offset from
// the group centre rounded as when it was listed), then the pair's Kelbg term
// from the FP64 positions of the leaf and the target.  An electron bound a
// few nm from an ion is far below the resolution of FP32 coordinates relative
// to a group centre some um away.
template <int TPL>
__device__ inline void exact_pairs(const WaveLDS &L, int nx, const float (&rx)[TPL], const float (&ry)[TPL],
                                   const float (&rz)[TPL], float qt, float inv_rs, float close2, double gx,
                                   double gy, double gz,
                                   const double4 *__restrict__ bmac, const double *__restrict__ px,
                                   const double *__restrict__ py, const double *__restrict__ pz, int tbase, int nact,
                                   Acc<TPL> &A)
{
    float ax[TPL], ay[TPL], az[TPL], ap[TPL];
#pragma unroll
    for (int t = 0; t < TPL; t++) ax[t] = ay[t] = az[t] = ap[t] = 0.0f;
    for (int e = 0; e < nx; e++) {
        const double4 m = bmac[L.exact[e]];
        const float cx = (float)(m.x - gx), cy = (float)(m.y - gy), cz = (float)(m.z - gz);
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            float r2, rq;
            if (close_pair(rx[t] - cx, ry[t] - cy, rz[t] - cz, close2, r2, rq)) {
                const int k = lane_id() + 64 * t;
                const int ti = tbase + (k < nact ? k : 0);
                const float ex = (float)(px[ti] - m.x), ey = (float)(py[ti] - m.y), ez = (float)(pz[ti] - m.z);
                const float e2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
                const float eq = e2 > 0.0f ? __builtin_amdgcn_rsqf(e2) : 0.0f;
                float g, h;
                kelbg_gh(e2 * eq * inv_rs, g, h);
                const float w = g * eq * eq * eq;
                ax[t] = fmaf(-ex, w, ax[t]);
                ay[t] = fmaf(-ey, w, ay[t]);
                az[t] = fmaf(-ez, w, az[t]);
                // coincident (r = 0): phi = 1 / rs and no force
                ap[t] = e2 > 0.0f ? fmaf(-h, eq, ap[t]) : ap[t] - inv_rs;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < TPL; t++) {
        nadd(A.s[t][0], A.c[t][0], qt * ax[t]);
        nadd(A.s[t][1], A.c[t][1], qt * ay[t]);
        nadd(A.s[t][2], A.c[t][2], qt * az[t]);
        nadd(A.ps[t], A.pc[t], qt * ap[t]);
    }
}

// q0..q2: q, D, traceless quadrupole; q3, q4: traceless third moment (orders 4 to 6);
// q4.w, q5, q6: traceless fourth moment (order 6)
template <int N>
__device__ inline void m2l_acc(float rx, float ry, float rz, float4 q0, float4 q1, float4 q2, float4 q3, float4 q4,
                               float4 q5, float4 q6, float (&L)[N])
{
    if constexpr (N == TB_LOCAL_N3)
        m2l_acc_p3(rx, ry, rz, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, L);
    else if constexpr (N == TB_LOCAL_N4H)
        m2l_acc_p4ho(rx, ry, rz, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q3.x, q3.y, q3.z, q3.w, q4.x,
                     q4.y, q4.z, L);
    else if constexpr (N == TB_LOCAL_N5H)
        m2l_acc_p5ho(rx, ry, rz, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q3.x, q3.y, q3.z, q3.w, q4.x,
                     q4.y, q4.z, L);
    else
        m2l_acc_p6hxt7(rx, ry, rz, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q3.x, q3.y, q3.z, q3.w, q4.x,
                       q4.y, q4.z, q4.w, q5.x, q5.y, q5.z, q5.w, q6.x, q6.y, q6.z, q6.w, L);
}

template <int N>
__device__ inline void l2p(const float (&L)[N], float x, float y, float z, float &phi, float &ex, float &ey, float &ez)
{
    if constexpr (N == TB_LOCAL_N3)
        l2p_p3(L, x, y, z, phi, ex, ey, ez);
    else if constexpr (N == TB_LOCAL_N4H)
        l2p_p4h(L, x, y, z, phi, ex, ey, ez);
    else if constexpr (N == TB_LOCAL_N5H)
        l2p_p5h(L, x, y, z, phi, ex, ey, ez);
    else
        l2p_p6h(L, x, y, z, phi, ex, ey, ez);
}

// Far cells are evaluated after the walk (k_far_eval), so their local
// expansion takes no registers here: the walk appends each group's far cells
// to a linked list of 64-entry chunks in global memory.  A wave reserves
// chunks FAR_SLAB at a time; a chunk past the buffer's capacity is not
// written, and the host re-runs the force with a larger buffer.
struct FarSlab {
    int next = 0, end = 0;  // the wave's reserved chunk range
    int head = -1;          // the group's last chunk (lists run backwards)
};

__device__ inline void push_chunk(const uint32_t *q, int cnt, uint32_t *__restrict__ cells,
                                  int32_t *__restrict__ link, int cap, int32_t *ctr, FarSlab &F)
{
    const int lane = lane_id();
    if (F.next >= F.end) {
        // every lane issues the atomic (lane 0 adds, the rest add 0): no lane-index branch
        const int v = uni(atomicAdd(ctr, lane == 0 ? FAR_SLAB : 0));
        F.next = v;
        F.end = v + FAR_SLAB;
    }
    const int c = F.next++;
    if (c < cap) {
        cells[(size_t)c * 64 + lane] = lane < cnt ? q[lane] : 0xffffffffu;
        link[c] = F.head;  // every lane stores the same value
        F.head = c;
    }
}

__device__ inline void push_far(const WaveLDS &L, int cnt, uint32_t *__restrict__ cells, int32_t *__restrict__ link,
                                int cap, int32_t *ctr, FarSlab &F)
{
    const int lane = lane_id();
    if (F.next >= F.end) {
        // every lane issues the atomic (lane 0 adds, the rest add 0): no lane-index branch
        const int v = uni(atomicAdd(ctr, lane == 0 ? FAR_SLAB : 0));
        F.next = v;
        F.end = v + FAR_SLAB;
    }
    const int c = F.next++;
    if (c < cap) {
        cells[(size_t)c * 64 + lane] = lane < cnt ? L.farq[lane] : 0xffffffffu;
        link[c] = F.head;  // every lane stores the same value
        F.head = c;
    }
}

// At most 3 blocks of 4 waves fit a CU's LDS, i.e. 3 waves per SIMD: the
// register budget is capped to match (512 / 3 = 170 VGPRs).  The learned
// walk fits it without spilling since its band decisions and traversal steps
// share one set of flushes; when it spilled (an earlier loop shape) it was
// still faster under the cap (257 ms at N = 4M) than with 256 registers at 2
// waves per SIMD (290 ms).
template <int TPL, int MODE, bool COUNT>
__global__ __launch_bounds__(BT, 3) void k_walk32(WalkArgs W)
{
    __shared__ WaveLDS lds[WPB];
    const int lane = lane_id();
    WaveLDS &L = lds[threadIdx.x >> 6];
    const int G = 64 * TPL;
    // per target species: cells tested, M2P and like / unlike P2P interactions
    unsigned long long c_visit[2] = {0, 0}, c_m2p[2] = {0, 0}, c_like[2] = {0, 0}, c_unlike[2] = {0, 0};
    unsigned long long c_m2l[2] = {0, 0}, c_far[2] = {0, 0}, c_m2p3[2] = {0, 0};

    // Work queue, XCD-aware: workgroups go to the 8 XCDs round-robin, so
    // block b's waves first take groups from chunk b % 8 of the group range
    // (contiguous in key order, hence in space: that XCD's L2 keeps the tree
    // region its groups share), then from the other chunks in turn.
    int chunk = blockIdx.x & (NXCD - 1), tried = 0;
    FarSlab F;
    for (;;) {
        // No branch on the lane index inside this loop: the compiler may
        // thread such a (loop-invariant) branch across the back edge and split
        // the wave, leaving lanes 1..63 to walk without lane 0.  So every lane
        // issues the fetch (lane 0 adds 1, the rest 0) and the stores below
        // are made by every lane with the same value.
        int g = -1;
        while (tried < NXCD) {
            const int c0 = (int)((long)W.ngt * chunk / NXCD), c1 = (int)((long)W.ngt * (chunk + 1) / NXCD);
            int v = atomicAdd(&W.ctr[CTR_CHUNK + chunk], lane == 0 ? 1 : 0);
            v = uni(v);  // lane 0's count (every lane is active here)
            if (c0 + v < c1) {
                g = c0 + v;
                break;
            }
            chunk = (chunk + 1) & (NXCD - 1);
            tried++;
        }
        if (g < 0) break;
        // the domain holding global group g (descriptors ordered by gbase)
        int lo = 0, hi = W.ndom - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cdom(W.dom)[mid].gbase <= g) lo = mid;
            else hi = mid - 1;
        }
        CDomain &D = cdom(W.dom)[lo];
        const int gl = g - D.gbase;
        const int tsp = gl < D.ng[0] ? 0 : 1;
        const float qt = tsp ? 1.0f : -1.0f;
        const int2 gr = D.grp[tsp][tsp ? gl - D.ng[0] : gl];
        const int base = uni(gr.x);
        const int nact = uni(gr.y);  // targets in this group (at most G)
        const double *px = D.px[tsp], *py = D.py[tsp], *pz = D.pz[tsp];

        // targets and the group's bounding box (FP64)
        double tx[TPL], ty[TPL], tz[TPL];
        bool act[TPL];
        double mnx = 1e300, mny = 1e300, mnz = 1e300, mxx = -1e300, mxy = -1e300, mxz = -1e300;
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            act[t] = lane + 64 * t < nact;
            const int ii = act[t] ? base + lane + 64 * t : base;
            tx[t] = px[ii];
            ty[t] = py[ii];
            tz[t] = pz[ii];
            mnx = fmin(mnx, tx[t]); mxx = fmax(mxx, tx[t]);
            mny = fmin(mny, ty[t]); mxy = fmax(mxy, ty[t]);
            mnz = fmin(mnz, tz[t]); mxz = fmax(mxz, tz[t]);
        }
        mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
        mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
        // learned / collect: the group's weakest net field from the previous force
        float emin = 0.0f;
        if (MODE != MODE_FIXED) {
            if (W.emin_saved) {
                emin = W.gemin[g];
            } else {
                double e2 = 1e300;
#pragma unroll
                for (int t = 0; t < TPL; t++) {
                    if (!act[t]) continue;
                    const int i = base + lane + 64 * t;
                    const double fx = D.fx[tsp][i], fy = D.fy[tsp][i], fz = D.fz[tsp][i];
                    e2 = fmin(e2, fx * fx + fy * fy + fz * fz);
                }
                emin = (float)(sqrt(wave_min(e2)) / TB_EPS);
                // kept for a re-run of this walk, which finds the forces overwritten
                W.gemin[g] = uni(emin);  // every lane stores the same value
            }
        }
        emin = uni(emin);
        const double gx = uni(0.5 * (mnx + mxx)), gy = uni(0.5 * (mny + mxy)), gz = uni(0.5 * (mnz + mxz));
        // half extents, rounded outwards
        const float hx = uni(half_extent(mnx, mxx));
        const float hy = uni(half_extent(mny, mxy));
        const float hz = uni(half_extent(mnz, mxz));
        // per-group terms of the learned criterion's features
        const float hdiag = uni(sqrtf(hx * hx + hy * hy + hz * hz));
        const float klemin = uni(0.125f * __log2f(fmaxf(emin, 1e-30f)));
        const float knact = uni(0.125f * __log2f((float)nact));
        float rx[TPL], ry[TPL], rz[TPL];
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            rx[t] = (float)(tx[t] - gx);
            ry[t] = (float)(ty[t] - gy);
            rz[t] = (float)(tz[t] - gz);
        }
        Acc<TPL> A;
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            for (int k = 0; k < 3; k++) A.s[t][k] = A.c[t][k] = 0.0f;
            A.ps[t] = A.pc[t] = 0.0f;
        }
        // accepted cells farther than far_k2^(1/2) group radii from the
        // centre go to the group's far list (k_far_eval)
        const float far_r2 = uni(W.far_k2 > 0.0f ? W.far_k2 * hdiag * hdiag : 3.0e38f);
        F.head = -1;

        // stack, M2P, P2P (like, unlike), band-queue, far-queue lengths
        int sp = 1, nm = 0, nm3 = 0, npl = 0, npu = 0, nb = 0, nf = 0, nx = 0;
        // two-level: start from the segment's frontier (chunks of cells its
        // segment walk left to the groups) instead of the root
        int fc = -1;
        if (W.two_level) {
            const int gc = D.gfirst[tsp] + (tsp ? gl - D.ng[0] : gl);  // combined group index
            fc = uni(W.sfront[D.gsid[tsp][gc] + (tsp ? W.nseg0 : 0)]);
            sp = 0;
        }
        L.stack[0] = D.root << 3;  // the domain's root: a range of one cell
        wave_sync();
        // the lists are evaluated once they reach their flush lengths (and
        // the far queue handed on once 64 wait), after every pass
        auto flushes = [&]() {
            if (nm + nm3 >= W.flush_m) {
                if (COUNT) {
                    c_m2p[tsp] += (unsigned long long)(nm + nm3) * nact;
                    c_m2p3[tsp] += (unsigned long long)nm3 * nact;
                }
                flush_m2p<TPL>(L, nm, nm3, rx, ry, rz, A);
                nm = nm3 = 0;
                wave_sync();
            }
            if (npl + npu >= W.flush_p) {
                flush_p2p<TPL>(L, npl, npu, rx, ry, rz, qt, D.inv_rs, D.close2, A, nx);
                npl = npu = 0;
                wave_sync();
            }
            if (nf >= 64) {
                push_far(L, 64, W.far_cells, W.far_link, W.far_cap, W.ctr + CTR_FAR, F);
                // the rest of the queue (at most 63 entries) moves down
                const uint32_t rest = lane + 64 < nf ? L.farq[lane + 64] : 0u;
                wave_sync();
                if (lane + 64 < nf) L.farq[lane] = rest;
                nf -= 64;
                wave_sync();
            }
        };
        // learned: band cells wait in a queue.  Once 64 wait, or nothing
        // else is left, one pass decides up to 64 of them between the
        // traversal steps; the traversal loop itself is the fixed walk's.
        // Each pass (a traversal step or a decision) adds at most 64 entries
        // to each list before the flushes.
        for (;;) {
            while (sp > 0 || fc >= 0) {
                bool have, full4 = false;
                uint32_t node;
                if (sp > 0) {
                    // a batch of <= 64 cells from the top ranges of the stack
                    uint32_t R = 0;
                    int len = 0;
                    if (lane < sp) {
                        R = L.stack[sp - 1 - lane];
                        len = (int)(R & 7u) + 1;
                    }
                    const int incl = wave_incl_scan(len);
                    const bool take = lane < sp && incl <= 64;
                    const int K = __popcll(__ballot(take));
                    const int total = K > 0 ? __shfl(incl, K - 1, 64) : 0;
                    if (take) {
                        // learned: FULL4 on a range marks the children of a band
                        // cell the network opened (carried to its batch entries)
                        const uint32_t st = (R & ~FULL4) >> 3, fb = MODE == MODE_LEARNED ? R & FULL4 : 0u;
                        const int ex = incl - len;
                        for (int o = 0; o < len; o++) L.batch[ex + o] = (st + (uint32_t)o) | fb;
                    }
                    sp -= K;
                    wave_sync();
                    have = lane < total;
                    node = have ? L.batch[lane] : 0u;
                    wave_sync();
                    if (MODE == MODE_LEARNED) {
                        full4 = (node & FULL4) != 0u;
                        node &= ~FULL4;
                    }
                } else {
                    // the next frontier chunk: up to 64 cells
                    const uint32_t v = W.far_cells[(size_t)fc * 64 + lane];
                    have = v != 0xffffffffu;
                    node = have ? v : 0u;
                    fc = uni(W.far_link[fc]);
                }
                bool leaf = false, acc = false, open = false, band = false, oct = false;
                float cx = 0.f, cy = 0.f, cz = 0.f, d2 = 0.f;
                uint32_t info = 0;
                if (have) {
                    const double4 m = D.bmac[node];
                    info = D.binfo[node];
                    cx = (float)(m.x - gx);
                    cy = (float)(m.y - gy);
                    cz = (float)(m.z - gz);
                    if (info & TB_LEAF) {
                        leaf = true;
                    } else {
                        d2 = box_d2(cx, cy, cz, hx, hy, hz);
                        const float rc = (float)m.w;  // s / theta_mixed
                        const bool far_hi = d2 > rc * rc * 1.000004f;
                        // accepted beyond OCT_RADIUS opening radii: M2P to the third
                        // moment (the learned walk: radii of its band's outer edge
                        // s/theta_lo, so that it is never less accurate than the
                        // fixed walk at any theta in its band)
                        const float roct = (MODE == MODE_FIXED ? rc : rc * W.lo_scale) * OCT_RADIUS;
                        oct = !full4 && d2 > roct * roct;
                        const float cnt = D.bq2[node].z;
                        // a cell of at most SMALL_OPEN particles is opened: its
                        // particles' P2P is exact and cheaper than its M2P
                        const bool small = cnt <= (float)SMALL_OPEN;
                        if (MODE == MODE_FIXED) {
                            acc = far_hi && !small;
                            open = !acc;
                        } else {
                            // as selects, not branches: accepted beyond s/theta_lo,
                            // opened within s/theta_mixed, the band between.
                            // Larger cells are never labelled (the label sums the
                            // cell's particles), and far cells are cheap to accept
                            // (the group's local expansion): in the learned walk
                            // the band's geometric-mean radius decides both
                            const float rlo = rc * W.lo_scale;  // s / theta_lo
                            const bool far_lo = d2 > rlo * rlo * 1.000004f;
                            const bool far_mid = d2 > rlo * rc * 1.000004f;
                            const bool net = MODE == MODE_COLLECT ||
                                             (cnt <= (float)BAND_MAXCNT && cx * cx + cy * cy + cz * cz <= far_r2);
                            band = !small && !far_lo && far_hi && net;
                            acc = !small && (far_lo || (far_hi && !net && far_mid));
                            open = !acc && !band;
                        }
                    }
                }
                // learned: band cells are queued and decided 64 at a time below
                if (MODE == MODE_COLLECT && band) {
                    const float4 q0 = D.bq0[node], q1 = D.bq1[node], q2 = D.bq2[node];
                    float f[TB_MLP_NIN];
                    band_features(D.bf[node], q2.w, sqrtf(d2), hdiag, klemin, q0.x * qt >= 0.0f, knact, f);
                    {
                        open = true;  // the conservative walk opens it; the pair is recorded
                        if (q2.z <= (float)BAND_MAXCNT) {
                            const int k = atomicAdd(&W.ctr[3], 1);
                            if (k < W.rec_cap) {
                                BandRec r;
                                r.cell = node;
                                r.group = (uint32_t)g;
                                r.tree = 0;
                                r.pad = 0;
                                for (int j = 0; j < TB_MLP_NIN; j++) r.f[j] = f[j];
                                W.rec[k] = r;
                            }
                        }
                    }
                }
                if (COUNT) c_visit[tsp] += (unsigned long long)__popcll(__ballot(have));
                // accepted cells -> M2P list, or the far queue
                const bool far = acc && cx * cx + cy * cy + cz * cz > far_r2;
                const uint64_t fk = __ballot(far);
                if (far) L.farq[nf + __popcll(fk & lanemask_lt())] = node;
                nf += __popcll(fk);
                if (COUNT) {
                    c_m2l[tsp] += (unsigned long long)__popcll(fk) * nact;
                    c_far[tsp] += (unsigned long long)__popcll(fk);
                }
                acc = acc && !far;
                // to the fourth moment from the bottom of the M2P list, to the
                // third (no m5, m6) from the top
                const bool acc4 = acc && !oct, acc3 = acc && oct;
                uint64_t mk = __ballot(acc4);
                const uint64_t mk3 = __ballot(acc3);
                if (acc) {
                    const int k = acc4 ? nm + __popcll(mk & lanemask_lt()) : LCAP_M - 1 - nm3 - __popcll(mk3 & lanemask_lt());
                    const float4 q0 = D.bq0[node], q1 = D.bq1[node], q2 = D.bq2[node];
                    L.m0[k] = make_float4(cx, cy, cz, q0.x);
                    L.m1[k] = make_float4(q0.y, q0.z, q0.w, q1.x);
                    L.m2[k] = make_float4(q1.y, q1.z, q1.w, q2.x);
                    L.m3[k] = D.bq3[node];
                    L.m4[k] = D.bq4[node];
                    if (acc4) {
                        L.m5[k] = D.bq5[node];
                        L.m6[k] = D.bq6[node];
                    }
                }
                nm += __popcll(mk);
                nm3 += __popcll(mk3);
                // leaves -> P2P list: like charges from the bottom, unlike from the top
                const float qleaf = leaf ? D.bq0[node].x : 0.0f;
                const bool lk = leaf && qleaf == qt;
                const uint64_t ml = __ballot(lk), mu = __ballot(leaf && !lk);
                const float nodebits = __uint_as_float(node);
                if (lk) L.p[npl + __popcll(ml & lanemask_lt())] = make_float4(cx, cy, cz, nodebits);
                else if (leaf) L.p[LCAP_P - 1 - npu - __popcll(mu & lanemask_lt())] = make_float4(cx, cy, cz, nodebits);
                npl += __popcll(ml);
                npu += __popcll(mu);
                if (COUNT) {
                    c_like[tsp] += (unsigned long long)__popcll(ml) * nact;
                    c_unlike[tsp] += (unsigned long long)__popcll(mu) * nact;
                }
                // opened cells -> their child ranges.  Children follow their
                // parent in breadth-first order, so a range not strictly after the
                // cell (or past the table) is a malformed tree: flagged, never
                // pushed, which bounds every path through the traversal.
                if (open) {
                    const uint32_t st = info >> 3, len = (info & 7u) + 1u;
                    if (st <= node || st + len > (uint32_t)D.nn) {
                        atomicOr(&W.ctr[1], 2);
                        open = false;
                    }
                }
                mk = __ballot(open);
                const int npush = __popcll(mk);
                if (sp + npush > STACK) {
                    atomicOr(&W.ctr[1], 1);
                    sp = 0;
                    fc = -1;
                    nm = nm3 = npl = npu = nb = 0;
                    break;
                }
                if (open) L.stack[sp + __popcll(mk & lanemask_lt())] = info;
                sp += npush;
                wave_sync();
                if (MODE == MODE_LEARNED) {
                    const uint64_t bm = __ballot(band);
                    if (bm) {  // wave-uniform: most steps meet no band cell
                        if (band) L.bandq[nb + __popcll(bm & lanemask_lt())] = node;
                        nb += __popcll(bm);
                        wave_sync();
                    }
                }
                flushes();
                if (MODE == MODE_LEARNED && nb >= 64) break;
            }
            if (MODE != MODE_LEARNED || nb == 0) break;
            {
                // Decide up to 64 queued cells with one batched MLP pass on
                // the matrix cores (tb_mlp_eval32_mfma, two tiles of 32).
                // Entry e's features are recomputed from its records.
                const int ne = min(nb, 64);
                // lane l forms entry l's features; each tile's MFMA
                // operand takes features 0..3 of its entry in lanes 0..31
                // and 4..7 in lanes 32..63, exchanged across the halves
                float f[TB_MLP_NIN];
#pragma unroll
                for (int j = 0; j < TB_MLP_NIN; j++) f[j] = 0.0f;
                if (lane < ne) {
                    const uint32_t nd = L.bandq[lane];
                    const double4 m = D.bmac[nd];
                    const float ex = fmaxf(fabsf((float)(m.x - gx)) - hx, 0.0f);
                    const float ey = fmaxf(fabsf((float)(m.y - gy)) - hy, 0.0f);
                    const float ez = fmaxf(fabsf((float)(m.z - gz)) - hz, 0.0f);
                    band_features(D.bf[nd], D.bq2[nd].w, sqrtf(ex * ex + ey * ey + ez * ez), hdiag, klemin,
                                  D.bq0[nd].x * qt >= 0.0f, knact, f);
                }
                float xa[4], xb[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float lo4 = __shfl_xor(f[q], 32, 64), hi4 = __shfl_xor(f[4 + q], 32, 64);
                    xa[q] = lane < 32 ? f[q] : hi4;   // tile 0: entries 0..31
                    xb[q] = lane < 32 ? lo4 : f[4 + q];  // tile 1: entries 32..63
                }
                float pr[2];
                pr[0] = tb_mlp_eval32_mfma(W.mlpk, xa);
                pr[1] = tb_mlp_eval32_mfma(W.mlpk, xb);
                // lane l decides entry l (tile l / 32 holds it); without a
                // previous force (emin = 0) band cells are opened
                const bool mine = lane < ne;
                const uint32_t nd = mine ? L.bandq[lane] : 0u;
                // cost-aware threshold: the error budget goes where accepting
                // saves the most work (log_err + cost_w log10 S, S the
                // saving in M2Ps, see band_saving)
                float thr = W.log_err;
                if (W.cost_w > 0.0f && mine) thr += W.cost_w * band_saving_log10(D, nd);
                bool bacc = mine && emin > 0.0f && (lane < 32 ? pr[0] : pr[1]) < thr;
                bool bopen = mine && !bacc;
                // accepted band cells: far ones to the far queue, the rest to M2P
                float bx = 0.0f, by = 0.0f, bz = 0.0f;
                if (bacc) {
                    const double4 m = D.bmac[nd];
                    bx = (float)(m.x - gx);
                    by = (float)(m.y - gy);
                    bz = (float)(m.z - gz);
                }
                const bool bfar = bacc && bx * bx + by * by + bz * bz > far_r2;
                const uint64_t bf = __ballot(bfar);
                if (bfar) L.farq[nf + __popcll(bf & lanemask_lt())] = nd;
                nf += __popcll(bf);
                if (COUNT) {
                    c_m2l[tsp] += (unsigned long long)__popcll(bf) * nact;
                    c_far[tsp] += (unsigned long long)__popcll(bf);
                }
                bacc = bacc && !bfar;
                uint64_t bk = __ballot(bacc);
                if (bacc) {
                    const int k = nm + __popcll(bk & lanemask_lt());
                    const float4 q0 = D.bq0[nd], q1 = D.bq1[nd], q2 = D.bq2[nd];
                    L.m0[k] = make_float4(bx, by, bz, q0.x);
                    L.m1[k] = make_float4(q0.y, q0.z, q0.w, q1.x);
                    L.m2[k] = make_float4(q1.y, q1.z, q1.w, q2.x);
                    L.m3[k] = D.bq3[nd];
                    L.m4[k] = D.bq4[nd];
                    L.m5[k] = D.bq5[nd];
                    L.m6[k] = D.bq6[nd];
                }
                nm += __popcll(bk);
                const uint32_t binf = bopen ? D.binfo[nd] : 0u;
                if (bopen) {
                    const uint32_t st = binf >> 3, len = (binf & 7u) + 1u;
                    if (st <= nd || st + len > (uint32_t)D.nn) {
                        atomicOr(&W.ctr[1], 2);
                        bopen = false;
                    }
                }
                bk = __ballot(bopen);
                const int bpush = __popcll(bk);
                if (sp + bpush > STACK) {
                    atomicOr(&W.ctr[1], 1);
                    sp = 0;
                    nm = nm3 = npl = npu = nb = 0;
                    break;
                }
                if (bopen) L.stack[sp + __popcll(bk & lanemask_lt())] = binf | FULL4;
                sp += bpush;
                // the rest of the queue (at most 63 entries) moves down
                const uint32_t rest = lane + 64 < nb ? L.bandq[lane + 64] : 0u;
                wave_sync();
                if (lane + 64 < nb) L.bandq[lane] = rest;
                nb -= ne;
                wave_sync();
            }
            flushes();
        }
        if (nm + nm3 > 0) {
            if (COUNT) {
                c_m2p[tsp] += (unsigned long long)(nm + nm3) * nact;
                c_m2p3[tsp] += (unsigned long long)nm3 * nact;
            }
            flush_m2p<TPL>(L, nm, nm3, rx, ry, rz, A);
        }
        if (npl + npu > 0) {
            flush_p2p<TPL>(L, npl, npu, rx, ry, rz, qt, D.inv_rs, D.close2, A, nx);
        }
        if (nx > 0) {
            wave_sync();
            exact_pairs<TPL>(L, nx, rx, ry, rz, qt, D.inv_rs, D.close2, gx, gy, gz, D.bmac, px, py, pz, base, nact, A);
        }
        if (nf > 0) push_far(L, nf, W.far_cells, W.far_link, W.far_cap, W.ctr + CTR_FAR, F);
        W.far_head[g] = F.head;  // every lane stores the same value
        wave_sync();
        // forces F = q_t eps E; energy 0.5 q_t eps phi (FP64 from here on)
        const double ff = (double)qt * TB_EPS;
        double eg = 0.0;
#pragma unroll
        for (int t = 0; t < TPL; t++) {
            if (!act[t]) continue;
            const int i = base + lane + 64 * t;
            D.fx[tsp][i] = ff * ((double)A.s[t][0] + (double)A.c[t][0]);
            D.fy[tsp][i] = ff * ((double)A.s[t][1] + (double)A.c[t][1]);
            D.fz[tsp][i] = ff * ((double)A.s[t][2] + (double)A.c[t][2]);
            eg += 0.5 * ff * ((double)A.ps[t] + (double)A.pc[t]);
        }
        // lane 0's sum (butterfly sums may differ in the last bit between lanes)
        eg = __shfl(wave_sum(eg), 0, 64);
        W.gen[2 * g + tsp] = eg;
        W.gen[2 * g + 1 - tsp] = 0.0;
    }
    // counter slots (visits, m2p, p2p) x (ee, ei, ie, ii): like P2P and all
    // cell counts go to ee / ii by target species, unlike P2P to ei / ie
    if (COUNT && lane == 0) {
        atomicAdd(&W.cnt[0], c_visit[0]);
        atomicAdd(&W.cnt[1], c_m2p[0]);
        atomicAdd(&W.cnt[2], c_like[0]);
        atomicAdd(&W.cnt[5], c_unlike[0]);
        atomicAdd(&W.cnt[8], c_unlike[1]);
        atomicAdd(&W.cnt[9], c_visit[1]);
        atomicAdd(&W.cnt[10], c_m2p[1]);
        atomicAdd(&W.cnt[11], c_like[1]);
        atomicAdd(&W.cnt[12], c_m2l[0]);
        atomicAdd(&W.cnt[13], c_m2l[1]);
        atomicAdd(&W.cnt[14], c_far[0]);
        atomicAdd(&W.cnt[15], c_far[1]);
        atomicAdd(&W.cnt[18], c_m2p3[0]);
        atomicAdd(&W.cnt[19], c_m2p3[1]);
    }
}

// The far cells of every group, after the walk: one wavefront per group,
// one far cell per lane per chunk into the lane's partial local expansion
// about the group's centre (the centre the walk used: the middle of the
// targets' bounding box), summed across the wavefront and evaluated at each
// target; field and energy are added to the walk's.
// ------------------------------------------------------ segment walk
//
// One wavefront per segment S (a cell of at most 8 G targets of one species,
// cut into at most 8 groups).  It walks the top of the tree once for all its
// groups, with every group's own acceptance test:
//   - a cell that every group of S would open (its test fails for each group
//     box) is opened here;
//   - a cell far from the whole of S (accepted for S's box, which implies
//     every group accepts it, and beyond K times S's half-diagonal from S's
//     centre) goes to S's far list: its field comes from S's own local
//     expansion, evaluated at each of S's targets (k_far_eval);
//   - every other cell (some group may accept it, or a leaf, or a small cell
//     the groups open themselves) goes to S's frontier, where the groups'
//     walks start.
// So each group accepts and opens exactly the cells it would from the root,
// except that cells far from all of S share one expansion; the local
// expansion's truncation bound is the group level's, (1/K)^7 of the cell's
// field at the targets, with S's half-diagonal for the group's.
struct SegLDS {
    uint32_t stack[STACK];
    uint32_t batch[64];
    uint32_t qf[QCAP];      // frontier cells waiting for a chunk
    uint32_t qs[QCAP];      // far cells waiting for a chunk
    double gc[8][3];        // group centres
    float gh[8][3];         // group half extents
};

template <int MODE, bool COUNT>
__global__ __launch_bounds__(BT) void k_swalk(WalkArgs W)
{
    __shared__ SegLDS lds[WPB];
    const int lane = lane_id();
    SegLDS &L = lds[threadIdx.x >> 6];
    unsigned long long c_pairs[2] = {0, 0}, c_tgt[2] = {0, 0};
    for (int sg = uni((int)(blockIdx.x * WPB + (threadIdx.x >> 6))); sg < W.nsegt; sg += gridDim.x * WPB) {
        const int k = sg < W.nseg0 ? 0 : 1;
        const int2 sr = W.sgrp[k][sg - (k ? W.nseg0 : 0)];
        const int g0 = uni(sr.x), ng = uni(sr.y);
        // the domain holding the segment (its groups' combined indices)
        int lo = 0, hi = W.ndom - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (W.dom[mid].gfirst[k] <= g0) lo = mid;
            else hi = mid - 1;
        }
        const Domain &D = W.dom[lo];
        const double *px = D.px[k], *py = D.py[k], *pz = D.pz[k];
        // each group's box, formed as the walk forms it; the segment's box
        double smn[3] = {1e300, 1e300, 1e300}, smx[3] = {-1e300, -1e300, -1e300};
        int ntgt = 0;
        for (int j = 0; j < ng; j++) {
            const int2 gr = W.cgrp[k][g0 + j];
            const int base = uni(gr.x), nact = uni(gr.y);
            ntgt += nact;
            double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
            for (int t = 0; t < 4; t++) {
                const int ii = lane + 64 * t < nact ? base + lane + 64 * t : base;
                const double v[3] = {px[ii], py[ii], pz[ii]};
                for (int a = 0; a < 3; a++) {
                    mn[a] = fmin(mn[a], v[a]);
                    mx[a] = fmax(mx[a], v[a]);
                }
            }
            for (int a = 0; a < 3; a++) {
                mn[a] = wave_min(mn[a]);
                mx[a] = wave_max(mx[a]);
                smn[a] = fmin(smn[a], mn[a]);
                smx[a] = fmax(smx[a], mx[a]);
                const double cen = uni(0.5 * (mn[a] + mx[a]));
                const float he = uni(half_extent(mn[a], mx[a]));
                L.gc[j][a] = cen;  // every lane stores the same value
                L.gh[j][a] = he;
            }
        }
        const double sx = uni(0.5 * (smn[0] + smx[0])), sy = uni(0.5 * (smn[1] + smx[1])),
                     sz = uni(0.5 * (smn[2] + smx[2]));
        const float shx = uni(half_extent(smn[0], smx[0])), shy = uni(half_extent(smn[1], smx[1])),
                    shz = uni(half_extent(smn[2], smx[2]));
        const float sdiag = uni(sqrtf(shx * shx + shy * shy + shz * shz));
        const float far_r2 = W.far_k2 * sdiag * sdiag;
        W.scen[sg] = make_double4(sx, sy, sz, (double)sdiag);  // every lane stores the same value
        FarSlab Ff, Fs;
        Ff.head = Fs.head = -1;
        int sp = 1, nq = 0, ns = 0;
        L.stack[0] = D.root << 3;
        wave_sync();
        while (sp > 0) {
            uint32_t R = 0;
            int len = 0;
            if (lane < sp) {
                R = L.stack[sp - 1 - lane];
                len = (int)(R & 7u) + 1;
            }
            const int incl = wave_incl_scan(len);
            const bool take = lane < sp && incl <= 64;
            const int K = __popcll(__ballot(take));
            const int total = K > 0 ? __shfl(incl, K - 1, 64) : 0;
            if (take)
                for (int o = 0; o < len; o++) L.batch[incl - len + o] = (R >> 3) + (uint32_t)o;
            sp -= K;
            wave_sync();
            const bool have = lane < total;
            const uint32_t node = have ? L.batch[lane] : 0u;
            wave_sync();
            bool front = false, sfar = false, open = false;
            uint32_t info = 0;
            if (have) {
                const double4 m = D.bmac[node];
                info = D.binfo[node];
                if ((info & TB_LEAF) || D.bq2[node].z <= (float)SMALL_OPEN) {
                    front = true;
                } else {
                    const float rc = (float)m.w;  // s / theta_mixed
                    // the radius beyond which every group accepts outright
                    const float ra = MODE == MODE_FIXED ? rc : rc * W.lo_scale;
                    const float cx = (float)(m.x - sx), cy = (float)(m.y - sy), cz = (float)(m.z - sz);
                    if (box_d2(cx, cy, cz, shx, shy, shz) > ra * ra * 1.000004f &&
                        cx * cx + cy * cy + cz * cz > far_r2) {
                        sfar = true;
                    } else {
                        bool any = false;
                        for (int j = 0; j < ng; j++) {
                            const float gx = (float)(m.x - L.gc[j][0]), gy = (float)(m.y - L.gc[j][1]),
                                        gz = (float)(m.z - L.gc[j][2]);
                            any = any || box_d2(gx, gy, gz, L.gh[j][0], L.gh[j][1], L.gh[j][2]) >
                                             rc * rc * 1.000004f;
                        }
                        if (any) front = true;
                        else open = true;
                    }
                }
            }
            if (open) {
                const uint32_t st = info >> 3, n = (info & 7u) + 1u;
                if (st <= node || st + n > (uint32_t)D.nn) {
                    atomicOr(&W.ctr[1], 2);
                    open = false;
                }
            }
            const uint64_t om = __ballot(open);
            const int npush = __popcll(om);
            if (sp + npush > STACK) {
                atomicOr(&W.ctr[1], 1);
                break;
            }
            if (open) L.stack[sp + __popcll(om & lanemask_lt())] = info;
            sp += npush;
            const uint64_t fm = __ballot(front), sm = __ballot(sfar);
            if (front) L.qf[nq + __popcll(fm & lanemask_lt())] = node;
            if (sfar) L.qs[ns + __popcll(sm & lanemask_lt())] = node;
            nq += __popcll(fm);
            ns += __popcll(sm);
            if (COUNT) {
                c_pairs[k] += (unsigned long long)__popcll(sm);
                c_tgt[k] += (unsigned long long)__popcll(sm) * ntgt;
            }
            wave_sync();
            for (int pass = 0; pass < 2; pass++) {
                uint32_t *q = pass ? L.qs : L.qf;
                int &n = pass ? ns : nq;
                FarSlab &F = pass ? Fs : Ff;
                if (n >= 64) {
                    push_chunk(q, 64, W.far_cells, W.far_link, W.far_cap, W.ctr + CTR_FAR, F);
                    const uint32_t rest = lane + 64 < n ? q[lane + 64] : 0u;
                    wave_sync();
                    if (lane + 64 < n) q[lane] = rest;
                    n -= 64;
                    wave_sync();
                }
            }
        }
        if (nq > 0) push_chunk(L.qf, nq, W.far_cells, W.far_link, W.far_cap, W.ctr + CTR_FAR, Ff);
        if (ns > 0) push_chunk(L.qs, ns, W.far_cells, W.far_link, W.far_cap, W.ctr + CTR_FAR, Fs);
        W.sfront[sg] = Ff.head;  // every lane stores the same value
        W.sfar[sg] = Fs.head;
        wave_sync();
    }
    // segment far cells: (cell, segment) pairs [16, 17] and (cell, target) pairs into [12, 13]
    if (COUNT && lane == 0) {
        atomicAdd(&W.cnt[16], c_pairs[0]);
        atomicAdd(&W.cnt[17], c_pairs[1]);
        atomicAdd(&W.cnt[12], c_tgt[0]);
        atomicAdd(&W.cnt[13], c_tgt[1]);
    }
}

// each segment's far cells into its local expansion about its centre
struct SegFarArgs {
    const Domain *dom;
    int ndom, nseg0, nsegt;
    const int2 *sgrp[2];
    const uint32_t *far_cells;
    const int32_t *far_link, *sfar;
    const double4 *scen;
    float *sloc;
};

__global__ __launch_bounds__(BT) void k_sfar_eval(SegFarArgs A)
{
    const int lane = lane_id();
    const int sg = uni((int)(blockIdx.x * WPB + (threadIdx.x >> 6)));
    if (sg >= A.nsegt) return;
    const int k = sg < A.nseg0 ? 0 : 1;
    const int g0 = uni(A.sgrp[k][sg - (k ? A.nseg0 : 0)].x);
    int lo = 0, hi = A.ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.dom[mid].gfirst[k] <= g0) lo = mid;
        else hi = mid - 1;
    }
    const Domain &D = A.dom[lo];
    const double4 sc = A.scen[sg];
    float Lp[NLOC];
#pragma unroll
    for (int q = 0; q < NLOC; q++) Lp[q] = 0.0f;
    for (int c = uni(A.sfar[sg]); c >= 0; c = uni(A.far_link[c])) {
        const uint32_t nd = A.far_cells[(size_t)c * 64 + lane];
        float rx = 1.0f, ry = 0.0f, rz = 0.0f;
        float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0, q3 = q0, q4 = q0, q5 = q0, q6 = q0;
        if (nd != 0xffffffffu) {
            const double4 m = D.bmac[nd];
            rx = (float)(sc.x - m.x);
            ry = (float)(sc.y - m.y);
            rz = (float)(sc.z - m.z);
            q0 = D.bq0[nd];
            q1 = D.bq1[nd];
            q2 = D.bq2[nd];
            q3 = D.bq3[nd];
            q4 = D.bq4[nd];
            q5 = D.bq5[nd];
            q6 = D.bq6[nd];
        }
        m2l_acc(rx, ry, rz, q0, q1, q2, q3, q4, q5, q6, Lp);
    }
#pragma unroll
    for (int q = 0; q < NLOC; q++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) Lp[q] += __shfl_xor(Lp[q], o, 64);
        A.sloc[(size_t)sg * NLOC + q] = Lp[q];  // every lane stores the same value
    }
}

struct FarArgs {
    const Domain *dom;
    int ndom, ngt;
    const uint32_t *far_cells;
    const int32_t *far_link, *far_head;
    double *gen;
    // two-level: each target also evaluates its segment's expansion
    int two_level, nseg0;
    const double4 *scen;
    const float *sloc;
};

__global__ __launch_bounds__(BT, 2) void k_far_eval(FarArgs A)
{
    const int lane = lane_id();
    const int g = uni((int)(blockIdx.x * WPB + (threadIdx.x >> 6)));
    if (g >= A.ngt) return;
    int c = uni(A.far_head[g]);
    if (c < 0 && !A.two_level) return;
    int lo = 0, hi = A.ndom - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.dom[mid].gbase <= g) lo = mid;
        else hi = mid - 1;
    }
    const Domain &D = A.dom[lo];
    const int gl = g - D.gbase;
    const int tsp = gl < D.ng[0] ? 0 : 1;
    const int2 gr = D.grp[tsp][tsp ? gl - D.ng[0] : gl];
    const double *px = D.px[tsp], *py = D.py[tsp], *pz = D.pz[tsp];
    double mnx = 1e300, mny = 1e300, mnz = 1e300, mxx = -1e300, mxy = -1e300, mxz = -1e300;
    for (int t = lane; t < gr.y; t += 64) {
        const double x = px[gr.x + t], y = py[gr.x + t], z = pz[gr.x + t];
        mnx = fmin(mnx, x); mxx = fmax(mxx, x);
        mny = fmin(mny, y); mxy = fmax(mxy, y);
        mnz = fmin(mnz, z); mxz = fmax(mxz, z);
    }
    mnx = wave_min(mnx); mny = wave_min(mny); mnz = wave_min(mnz);
    mxx = wave_max(mxx); mxy = wave_max(mxy); mxz = wave_max(mxz);
    const double gx = uni(0.5 * (mnx + mxx)), gy = uni(0.5 * (mny + mxy)), gz = uni(0.5 * (mnz + mxz));
    float Lp[NLOC];
#pragma unroll
    for (int k = 0; k < NLOC; k++) Lp[k] = 0.0f;
    // the chunk list is walked one chunk ahead: the next chunk's link and
    // cell ids load while this chunk's cells are expanded, so only the
    // records' loads wait on a list entry
    uint32_t ndn = c >= 0 ? A.far_cells[(size_t)c * 64 + lane] : 0xffffffffu;
    while (c >= 0) {
        const uint32_t nd = ndn;
        c = uni(A.far_link[c]);
        if (c >= 0) ndn = A.far_cells[(size_t)c * 64 + lane];
        float rx = 1.0f, ry = 0.0f, rz = 0.0f;
        float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0, q3 = q0, q4 = q0, q5 = q0, q6 = q0;
        if (nd != 0xffffffffu) {
            const double4 m = D.bmac[nd];
            rx = (float)(gx - m.x);
            ry = (float)(gy - m.y);
            rz = (float)(gz - m.z);
            q0 = D.bq0[nd];
            q1 = D.bq1[nd];
            q2 = D.bq2[nd];
            q3 = D.bq3[nd];
            q4 = D.bq4[nd];
            q5 = D.bq5[nd];
            q6 = D.bq6[nd];
        }
        m2l_acc(rx, ry, rz, q0, q1, q2, q3, q4, q5, q6, Lp);
    }
#pragma unroll
    for (int k = 0; k < NLOC; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) Lp[k] += __shfl_xor(Lp[k], o, 64);
    const double ff = (tsp ? 1.0 : -1.0) * TB_EPS;
    double eg = 0.0;
    for (int t = lane; t < gr.y; t += 64) {
        const int i = gr.x + t;
        float phi, ex, ey, ez;
        l2p(Lp, (float)(px[i] - gx), (float)(py[i] - gy), (float)(pz[i] - gz), phi, ex, ey, ez);
        D.fx[tsp][i] += ff * (double)ex;
        D.fy[tsp][i] += ff * (double)ey;
        D.fz[tsp][i] += ff * (double)ez;
        eg += 0.5 * ff * (double)phi;
    }
    if (A.two_level) {
        // the segment's expansion (cells far from every group of the segment)
        const int gc = D.gfirst[tsp] + (tsp ? gl - D.ng[0] : gl);
        const int sg = uni(D.gsid[tsp][gc] + (tsp ? A.nseg0 : 0));
        const double4 sc = A.scen[sg];
#pragma unroll
        for (int q = 0; q < NLOC; q++) Lp[q] = A.sloc[(size_t)sg * NLOC + q];
        for (int t = lane; t < gr.y; t += 64) {
            const int i = gr.x + t;
            float phi, ex, ey, ez;
            l2p(Lp, (float)(px[i] - sc.x), (float)(py[i] - sc.y), (float)(pz[i] - sc.z), phi, ex, ey, ez);
            D.fx[tsp][i] += ff * (double)ex;
            D.fy[tsp][i] += ff * (double)ey;
            D.fz[tsp][i] += ff * (double)ez;
            eg += 0.5 * ff * (double)phi;
        }
    }
    eg = __shfl(wave_sum(eg), 0, 64);
    A.gen[2 * g + tsp] += eg;  // every lane: the same value to the same address
}

template <int MODE, bool COUNT>
void dispatch(hipStream_t s, int tpl, int blocks, const WalkArgs &W)
{
    switch (tpl) {
    case 1: hipLaunchKernelGGL((k_walk32<1, MODE, COUNT>), dim3(blocks), dim3(BT), 0, s, W); break;
    case 4: hipLaunchKernelGGL((k_walk32<4, MODE, COUNT>), dim3(blocks), dim3(BT), 0, s, W); break;
    default: hipLaunchKernelGGL((k_walk32<2, MODE, COUNT>), dim3(blocks), dim3(BT), 0, s, W); break;
    }
}

// ---------------------------------------------------------------- prep

__global__ void k_zero_ctr(int32_t *ctr, int lo, int hi)
{
    const int i = lo + (int)threadIdx.x;
    if (i < hi) ctr[i] = 0;
}

// per domain, deterministic: esum[4d + 2*tsp + tree] = sum over the domain's groups
__global__ __launch_bounds__(256) void k_domain_energy(const Domain *__restrict__ dom,
                                                       const double *__restrict__ gen, double *__restrict__ esum)
{
    __shared__ double red[2][256];
    const Domain &D = dom[blockIdx.x];
    double a[2] = {0.0, 0.0};
    const int ng = D.ng[0] + D.ng[1];
    for (int k = threadIdx.x; k < ng; k += 256) {
        a[0] += gen[2 * (D.gbase + k)];
        a[1] += gen[2 * (D.gbase + k) + 1];
    }
    for (int c = 0; c < 2; c++) red[c][threadIdx.x] = a[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 2; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    // split by target species: electron targets -> the ee slot, ion targets -> the ii slot
    if (threadIdx.x == 0) {
        esum[4 * blockIdx.x] = red[0][0];
        esum[4 * blockIdx.x + 1] = 0.0;
        esum[4 * blockIdx.x + 2] = 0.0;
        esum[4 * blockIdx.x + 3] = red[1][0];
    }
}

// a single domain's pairing sums and the walk's error flag into the handle's sums
__global__ void k_finish(const double *__restrict__ esum, const int32_t *__restrict__ ctr, double *__restrict__ out)
{
    if (threadIdx.x < 4) out[threadIdx.x] = esum[threadIdx.x];
    if (threadIdx.x == 0) out[15] = (double)(ctr[1] | ctr[2]);
}

// ------------------------------------------------------------ labels

struct LabelArgs {
    const BandRec *rec;
    int nrec;
    const int2 *grp[2];    // target groups (start, length) per species
    int ng0;               // electron groups
    const double *px[2], *py[2], *pz[2];  // key-ordered target positions
    const double *fx[2], *fy[2], *fz[2];  // forces of the collecting walk
    int n[2];
    const uint32_t *bfs;                  // breadth-first -> preorder (joint tree)
    const double4 *bmac;                  // breadth-first: cell centres
    const double *bnq;                    // breadth-first: cells' FP64 moments
    const int32_t *nstart, *ncount;
    const double4 *pos;                   // joint particles: x, y, z, charge
    double rs;
    float *lab;
};

// label of a band pair: log10 of the rms, over the group's targets, of
// |F_M2P - F_exact| / |F_target| with F_M2P the cell's monopole + dipole +
// quadrupole + third- and fourth-moment field and F_exact the FP64 sum over the cell's particles with the
// reference's pair rules (like: Coulomb beyond 0.1 um; unlike: Kelbg).  The
// rms (not the largest) is the cell's share of the rms force error that the
// walk is tuned to.
__global__ __launch_bounds__(64) void k_labels(LabelArgs A)
{
    const int r = blockIdx.x;
    if (r >= A.nrec) return;
    const BandRec R = A.rec[r];
    const int tsp = (int)R.group < A.ng0 ? 0 : 1;
    const double qt = tsp ? 1.0 : -1.0;
    const int2 gr = A.grp[tsp][(int)R.group - (tsp ? A.ng0 : 0)];
    const uint32_t p = A.bfs[R.cell];
    const double4 c = A.bmac[R.cell];
    const double *Q = A.bnq + (size_t)R.cell * TB_JQ;
    const double q = Q[0], Dx = Q[1], Dy = Q[2], Dz = Q[3];
    const double tr = Q[4] + Q[5] + Q[6];
    const double Qxx = 3.0 * Q[4] - tr, Qyy = 3.0 * Q[5] - tr, Qzz = 3.0 * Q[6] - tr;
    const double Qxy = 3.0 * Q[7], Qyz = 3.0 * Q[8], Qzx = 3.0 * Q[9];
    // traceless third moment (raw O: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
    const double *O = Q + 10;
    const double tx = O[0] + O[3] + O[5], ty = O[1] + O[6] + O[8], tz = O[2] + O[7] + O[9];
    const double oxxx = O[0] - 0.6 * tx, oxxy = O[1] - 0.2 * ty, oxxz = O[2] - 0.2 * tz, oxyy = O[3] - 0.2 * tx;
    const double oxyz = O[4], oxzz = O[5] - 0.2 * tx, oyyy = O[6] - 0.6 * ty, oyyz = O[7] - 0.2 * tz;
    const double oyzz = O[8] - 0.2 * ty, ozzz = O[9] - 0.6 * tz;
    // traceless fourth moment (raw H: xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz
    // xyzz xzzz yyyy yyyz yyzz yzzz zzzz), its 9 independent components
    const double *H = Q + 20;
    const double txx = H[0] + H[3] + H[5], tyy = H[3] + H[10] + H[12], tzz = H[5] + H[12] + H[14];
    const double txy = H[1] + H[6] + H[8], tyz = H[4] + H[11] + H[13], tzx = H[2] + H[7] + H[9];
    const double t0 = txx + tyy + tzz;
    const double hxxxx = H[0] - 6.0 * txx / 7.0 + 3.0 * t0 / 35.0, hxxxy = H[1] - 3.0 * txy / 7.0;
    const double hxxxz = H[2] - 3.0 * tzx / 7.0, hxxyy = H[3] - (txx + tyy) / 7.0 + t0 / 35.0;
    const double hxxyz = H[4] - tyz / 7.0, hxyyy = H[6] - 3.0 * txy / 7.0, hxyyz = H[7] - tzx / 7.0;
    const double hyyyy = H[10] - 6.0 * tyy / 7.0 + 3.0 * t0 / 35.0, hyyyz = H[11] - 3.0 * tyz / 7.0;
    const int s0 = A.nstart[p], cnt = A.ncount[p];
    double sq = 0.0;  // sum over the group's targets of the squared relative error
    for (int t = threadIdx.x; t < gr.y; t += 64) {
        const int i = gr.x + t;
        const double x = A.px[tsp][i], y = A.py[tsp][i], z = A.pz[tsp][i];
        const double dx = x - c.x, dy = y - c.y, dz = z - c.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ri = 1.0 / sqrt(r2), ri2 = ri * ri, ri3 = ri * ri2, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
        const double dp = Dx * dx + Dy * dy + Dz * dz;
        const double qx = Qxx * dx + Qxy * dy + Qzx * dz;
        const double qy = Qxy * dx + Qyy * dy + Qyz * dz;
        const double qz = Qzx * dx + Qyz * dy + Qzz * dz;
        const double xqx = dx * qx + dy * qy + dz * qz;
        const double vx = oxxx * dx * dx + oxyy * dy * dy + oxzz * dz * dz + 2.0 * (oxxy * dx * dy + oxxz * dx * dz + oxyz * dy * dz);
        const double vy = oxxy * dx * dx + oyyy * dy * dy + oyzz * dz * dz + 2.0 * (oxyy * dx * dy + oxyz * dx * dz + oyyz * dy * dz);
        const double vz = oxxz * dx * dx + oyyz * dy * dy + ozzz * dz * dz + 2.0 * (oxyz * dx * dy + oxzz * dx * dz + oyzz * dy * dz);
        const double s3 = dx * vx + dy * vy + dz * vz, ri9 = ri7 * ri2, ri11 = ri9 * ri2;
        // w = H'rrr in the cubic harmonics A1..A7 (the walk's form, in FP64)
        const double zz = dz * dz, xmz = dx * dx - zz, ymz = dy * dy - zz;
        const double A1 = dx * (xmz - 2.0 * zz), A2 = 3.0 * dy * xmz, A3 = dz * (3.0 * xmz + 2.0 * zz);
        const double A4 = 3.0 * dx * ymz, A5 = 6.0 * dx * dy * dz, A6 = dy * (ymz - 2.0 * zz);
        const double A7 = dz * (3.0 * ymz + 2.0 * zz);
        double wv[3];
        wv[0] = hxxxx * A1 + hxxxy * A2 + hxxxz * A3 + hxxyy * A4 + hxxyz * A5 + hxyyy * A6 + hxyyz * A7;
        wv[1] = hxxxy * A1 + hxxyy * A2 + hxxyz * A3 + hxyyy * A4 + hxyyz * A5 + hyyyy * A6 + hyyyz * A7;
        wv[2] = hxxxz * A1 + hxxyz * A2 + hxyyz * A4 + hyyyz * A6 - (hxxxx + hxxyy) * A3 - (hxxxy + hxyyy) * A5 -
                (hxxyy + hyyyy) * A7;
        const double s4 = dx * wv[0] + dy * wv[1] + dz * wv[2];
        const double k = q * ri3 + 3.0 * dp * ri5 + 2.5 * xqx * ri7 + 17.5 * s3 * ri9 + 39.375 * s4 * ri11;
        double ex = dx * k - Dx * ri3 - qx * ri5 - 7.5 * vx * ri7 - 17.5 * wv[0] * ri9;
        double ey = dy * k - Dy * ri3 - qy * ri5 - 7.5 * vy * ri7 - 17.5 * wv[1] * ri9;
        double ez = dz * k - Dz * ri3 - qz * ri5 - 7.5 * vz * ri7 - 17.5 * wv[2] * ri9;
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int j = s0; j < s0 + cnt; j++) {
            const double4 sj = A.pos[j];
            const double ax = x - sj.x, ay = y - sj.y, az = z - sj.z;
            const double a2 = ax * ax + ay * ay + az * az;
            const double ar = sqrt(a2);
            double w;
            if (sj.w == qt) {
                if (!(a2 > 0.01)) continue;
                w = sj.w / (a2 * ar);
            } else {
                const double e = exp(-ar / A.rs);
                w = sj.w * (1.0 / ar - e * (1.0 / ar + 1.0 / A.rs)) / a2;
            }
            gx += ax * w; gy += ay * w; gz += az * w;
        }
        ex -= gx; ey -= gy; ez -= gz;
        const double fx = A.fx[tsp][i], fy = A.fy[tsp][i], fz = A.fz[tsp][i];
        const double fn = sqrt(fx * fx + fy * fy + fz * fz);
        const double rel = TB_EPS * sqrt(ex * ex + ey * ey + ez * ez) / fmax(fn, 1e-300);
        sq += rel * rel;
    }
    sq = wave_sum(sq);
    if (threadIdx.x == 0) A.lab[r] = (float)log10(sqrt(sq / gr.y) + 1e-12);
}

inline int nblk(long n, int per) { return (int)((n + per - 1) / per); }

template <typename T>
int grow(T **p, int *cap, int want, size_t *bytes)
{
    if (want <= *cap) return TB_OK;
    if (*p) {
        hipFree(*p);
        *bytes -= (size_t)*cap * sizeof(T);
    }
    *p = nullptr;
    *cap = 0;
    int c = want + want / 8 + 64;
    if (hipMalloc((void **)p, (size_t)c * sizeof(T)) != hipSuccess) {
        g_err = "mixed walk: device allocation failed";
        return TB_ERR_NOMEM;
    }
    *bytes += (size_t)c * sizeof(T);
    *cap = c;
    return TB_OK;
}

// targets per lane the species' groups were cut for (0 with g_err set if
// they were not cut, or not for one size)
int group_tpl(const SpeciesDev &e, const SpeciesDev &i)
{
    const int G = e.gG;
    if (G != i.gG || (G != 64 && G != 128 && G != 256) || e.ng < 1 || i.ng < 1) {
        g_err = "mixed walk: target groups were not built for this walk";
        return 0;
    }
    return G / 64;
}

int resident_waves()
{
    static int w = 0;
    if (w == 0) {
        int dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (cus <= 0) cus = 256;
        w = cus * 12;  // ~3 blocks of 4 waves per CU under the LDS budget
    }
    return w;
}

Domain make_domain(SpeciesDev *const sp[2], const MixedTree *t, double rs, int gbase)
{
    Domain d;
    for (int k = 0; k < 2; k++) {
        const int c = sp[k]->cur;
        d.px[k] = sp[k]->x[c][0]; d.py[k] = sp[k]->x[c][1]; d.pz[k] = sp[k]->x[c][2];
        d.fx[k] = sp[k]->f[c][0]; d.fy[k] = sp[k]->f[c][1]; d.fz[k] = sp[k]->f[c][2];
        d.n[k] = sp[k]->n;
        d.ng[k] = sp[k]->ng;
        d.grp[k] = sp[k]->grp;
    }
    d.bmac = t->bmac; d.binfo = t->binfo; d.bq0 = t->bq0; d.bq1 = t->bq1; d.bq2 = t->bq2; d.bf = t->bf;
    d.bq3 = t->bq3; d.bq4 = t->bq4; d.bq5 = t->bq5; d.bq6 = t->bq6;
    d.inv_rs = (float)(1.0 / rs);
    d.close2 = (float)(484.0 * rs * rs);  // (22 rs)^2: closer unlike pairs take the Kelbg term
    d.rs = (float)rs;
    d.gbase = gbase;
    d.nn = t->nn;
    d.root = 0;
    for (int k = 0; k < 2; k++) {
        d.gfirst[k] = 0;
        d.gsid[k] = sp[k]->gsid;
    }
    return d;
}

// far-cell chunk buffer of at least `chunks` chunks (64 cells each)
int grow_far(MixedDev &M, long chunks)
{
    if (chunks > (long)INT32_MAX - (1L << 20)) {
        g_err = "mixed walk: far-cell buffer would exceed 2^31 chunks";
        return TB_ERR_NOMEM;
    }
    if (M.far_cells) hipFree(M.far_cells);
    if (M.far_link) hipFree(M.far_link);
    M.far_cells = nullptr;
    M.far_link = nullptr;
    M.far_cap = 0;
    if (hipMalloc((void **)&M.far_cells, (size_t)chunks * 64 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void **)&M.far_link, (size_t)chunks * sizeof(int32_t)) != hipSuccess) {
        g_err = "mixed walk: far-cell buffer allocation failed";
        return TB_ERR_NOMEM;
    }
    M.far_cap = (int)chunks;
    return TB_OK;
}

// Uploads the domain descriptors (synchronously: they change only after a
// build, which leaves the stream idle) and launches one persistent walk.
// per-segment buffers of the two-level far field for nseg segments
int grow_seg(MixedDev &M, int nseg)
{
    if (nseg <= M.seg_cap) return TB_OK;
    hipFree(M.sfront); hipFree(M.sfar); hipFree(M.scen); hipFree(M.sloc);
    M.sfront = M.sfar = nullptr;
    M.scen = nullptr;
    M.sloc = nullptr;
    M.seg_cap = 0;
    const int cap = nseg + nseg / 8 + 64;
    if (hipMalloc((void **)&M.sfront, (size_t)cap * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&M.sfar, (size_t)cap * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&M.scen, (size_t)cap * sizeof(double4)) != hipSuccess ||
        hipMalloc((void **)&M.sloc, (size_t)cap * NLOC * sizeof(float)) != hipSuccess) {
        g_err = "mixed walk: segment buffers allocation failed";
        return TB_ERR_NOMEM;
    }
    M.seg_cap = cap;
    return TB_OK;
}

template <int MODE, bool COUNT>
void dispatch_seg(hipStream_t s, int nsegt, const WalkArgs &W)
{
    const int blocks = std::max(1, std::min(nblk(nsegt, WPB), nblk(resident_waves(), WPB)));
    hipLaunchKernelGGL((k_swalk<MODE, COUNT>), dim3(blocks), dim3(BT), 0, s, W);
}

// sp: the species arrays the domains' groups and segments index (a batched
// ensemble's combined arrays)
int launch_domains(hipStream_t s, MixedDev &M, const std::vector<Domain> &doms, int tpl, const tb_config &cfg,
                   int64_t *d_cnt, int mode, const SpeciesDev *sp)
{
    const int nd = (int)doms.size();
    if (nd > M.ddom_cap) {
        if (M.ddom) hipFree(M.ddom);
        M.ddom = nullptr;
        M.ddom_cap = 0;
        if (hipMalloc((void **)&M.ddom, (size_t)nd * sizeof(Domain)) != hipSuccess) {
            g_err = "mixed walk: descriptor allocation failed";
            return TB_ERR_NOMEM;
        }
        M.ddom_cap = nd;
    }
    const int ngt = doms.back().gbase + doms.back().ng[0] + doms.back().ng[1];
    if (4 * nd + 2 * ngt > M.esum_cap) {
        if (M.esum) hipFree(M.esum);
        M.esum = nullptr;
        M.esum_cap = 0;
        const int cap = 4 * nd + 2 * ngt + 1024;
        if (hipMalloc((void **)&M.esum, (size_t)cap * sizeof(double)) != hipSuccess) {
            g_err = "mixed walk: energy sums allocation failed";
            return TB_ERR_NOMEM;
        }
        M.esum_cap = cap;
    }
    if (hipMemcpy(M.ddom, doms.data(), (size_t)nd * sizeof(Domain), hipMemcpyHostToDevice) != hipSuccess) {
        g_err = "mixed walk: descriptor upload failed";
        return TB_ERR_HIP;
    }
    WalkArgs W;
    W.dom = M.ddom;
    W.ndom = nd;
    W.ngt = ngt;
    const double th_hi = cfg.theta_mixed > 0.0 ? cfg.theta_mixed : cfg.theta;
    const double th_lo = cfg.theta_lo > 0.0 ? cfg.theta_lo : th_hi;
    W.lo_scale = (float)(th_hi / th_lo);
    W.log_err = (float)log10(cfg.err_target > 0.0 ? cfg.err_target : 1e-5);
    W.cost_w = cfg.cost_weight;
    const double kfar = cfg.far_ratio == 0.0 ? FAR_RATIO_DEFAULT : cfg.far_ratio;
    W.far_k2 = kfar > 0.0 ? (float)(kfar * kfar) : 0.0f;
    W.mlpk = M.mlp.Ppk;
    const int lv = M.lists >= 0 && M.lists < TB_LIST_VARIANTS ? M.lists : 0;
    W.flush_m = LIST_M[lv];
    W.flush_p = LIST_P[lv];
    W.rec = M.rec;
    W.rec_cap = M.rec_cap;
    W.ctr = M.ctr;
    W.gen = M.esum + 4 * nd;
    W.cnt = (unsigned long long *)d_cnt;
    const int waves = std::min(ngt, resident_waves());
    const int blocks = nblk(waves, WPB);
    M.tpl = tpl;
    if (!M.ev[0] && (hipEventCreate(&M.ev[0]) != hipSuccess || hipEventCreate(&M.ev[1]) != hipSuccess)) {
        g_err = "mixed walk: event creation failed";
        return TB_ERR_HIP;
    }
    if (ngt > M.head_cap) {
        if (M.far_head) hipFree(M.far_head);
        if (M.gemin) hipFree(M.gemin);
        M.far_head = nullptr;
        M.gemin = nullptr;
        M.head_cap = 0;
        if (hipMalloc((void **)&M.far_head, (size_t)ngt * sizeof(int32_t)) != hipSuccess ||
            hipMalloc((void **)&M.gemin, (size_t)ngt * sizeof(float)) != hipSuccess) {
            g_err = "mixed walk: per-group buffers allocation failed";
            return TB_ERR_NOMEM;
        }
        M.head_cap = ngt;
    }
    // Far-cell chunks: a first guess, then what the last walk asked for,
    // each time plus every wave's partly used slab.  A walk that asked for
    // more than the buffer holds is run again with a larger one (its far
    // lists are incomplete; the near forces it wrote are overwritten).
    const bool far_on = W.far_k2 > 0.0f;
    // two-level far field: segments walk the top of the tree for their groups
    const bool two = far_on && mode != MODE_COLLECT && cfg.far_levels == 2 && sp && sp[0].nseg > 0 && sp[1].nseg > 0;
    W.two_level = two ? 1 : 0;
    W.nseg0 = two ? sp[0].nseg : 0;
    W.nsegt = two ? sp[0].nseg + sp[1].nseg : 0;
    for (int k = 0; k < 2; k++) {
        W.sgrp[k] = sp ? sp[k].sgrp : nullptr;
        W.cgrp[k] = sp ? sp[k].grp : nullptr;
    }
    if (two) {
        const int rc = grow_seg(M, W.nsegt);
        if (rc) return rc;
    }
    W.sfront = M.sfront;
    W.sfar = M.sfar;
    W.scen = M.scen;
    // every wave of the segment walk and of the group walk may hold a part-used slab
    const long slack = (long)(blocks + (two ? nblk(std::min(W.nsegt, resident_waves()), WPB) : 0)) * WPB * FAR_SLAB;
    // TREEMI_FAR_CHUNKS caps the buffer the first attempt of every walk may
    // use (tests use it to take the re-run path on any force)
    const char *env = getenv("TREEMI_FAR_CHUNKS");
    const long env_cap = env ? atol(env) : 0;
    if (far_on && M.far_cap == 0) {
        const long first = env_cap > 0 ? env_cap : (long)ngt * 512 + slack;
        const int rc = grow_far(M, first);
        if (rc) return rc;
    }
    if (far_on && env_cap > 0 && M.far_cap > env_cap) M.far_cap = (int)env_cap;  // the allocation stays larger
    for (int attempt = 0;; attempt++) {
        W.far_cells = M.far_cells;
        W.far_link = M.far_link;
        W.far_head = M.far_head;
        W.far_cap = M.far_cap;
        W.gemin = M.gemin;
        W.emin_saved = attempt > 0 ? 1 : 0;
        hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, M.ctr, 0, 2);
        hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, M.ctr, CTR_FAR, CTR_FAR + 1);
        hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, M.ctr, CTR_CHUNK, CTR_CHUNK + NXCD);
        hipEventRecord(M.ev[0], s);
        if (two) {
            if (mode == MODE_LEARNED) {
                if (d_cnt) dispatch_seg<MODE_LEARNED, true>(s, W.nsegt, W);
                else dispatch_seg<MODE_LEARNED, false>(s, W.nsegt, W);
            } else {
                if (d_cnt) dispatch_seg<MODE_FIXED, true>(s, W.nsegt, W);
                else dispatch_seg<MODE_FIXED, false>(s, W.nsegt, W);
            }
        }
        if (mode == MODE_COLLECT) {
            hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, M.ctr, 3, 4);
            dispatch<MODE_COLLECT, false>(s, tpl, blocks, W);
        } else if (mode == MODE_LEARNED) {
            if (d_cnt) dispatch<MODE_LEARNED, true>(s, tpl, blocks, W);
            else dispatch<MODE_LEARNED, false>(s, tpl, blocks, W);
        } else {
            if (d_cnt) dispatch<MODE_FIXED, true>(s, tpl, blocks, W);
            else dispatch<MODE_FIXED, false>(s, tpl, blocks, W);
        }
        if (!far_on) break;
        int32_t used = 0;
        if (hipMemcpyAsync(&used, M.ctr + CTR_FAR, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            g_err = "mixed walk: far-chunk count read failed";
            return TB_ERR_HIP;
        }
        if (used <= M.far_cap) break;
        if (attempt > 0) {
            g_err = "mixed walk: far-cell buffer still too small after growing";
            return TB_ERR_NOMEM;
        }
        const int rc = grow_far(M, (long)used + used / 4 + slack);
        if (rc) return rc;
    }
    if (two) {
        SegFarArgs SA;
        SA.dom = M.ddom;
        SA.ndom = nd;
        SA.nseg0 = W.nseg0;
        SA.nsegt = W.nsegt;
        for (int k = 0; k < 2; k++) SA.sgrp[k] = W.sgrp[k];
        SA.far_cells = M.far_cells;
        SA.far_link = M.far_link;
        SA.sfar = M.sfar;
        SA.scen = M.scen;
        SA.sloc = M.sloc;
        hipLaunchKernelGGL(k_sfar_eval, dim3(nblk(W.nsegt, WPB)), dim3(BT), 0, s, SA);
    }
    if (far_on) {
        FarArgs FA;
        FA.dom = M.ddom;
        FA.ndom = nd;
        FA.ngt = ngt;
        FA.far_cells = M.far_cells;
        FA.far_link = M.far_link;
        FA.far_head = M.far_head;
        FA.gen = W.gen;
        FA.two_level = W.two_level;
        FA.nseg0 = W.nseg0;
        FA.scen = M.scen;
        FA.sloc = M.sloc;
        hipLaunchKernelGGL(k_far_eval, dim3(nblk(ngt, WPB)), dim3(BT), 0, s, FA);
    }
    hipEventRecord(M.ev[1], s);
    hipLaunchKernelGGL(k_domain_energy, dim3(nd), dim3(256), 0, s, M.ddom, W.gen, M.esum);
    if (hipGetLastError() != hipSuccess) {
        g_err = "mixed walk: launch failed";
        return TB_ERR_HIP;
    }
    return TB_OK;
}

int launch(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *d_sums, int64_t *d_cnt,
           int mode)
{
    const int tpl = group_tpl(sp[0], sp[1]);
    if (!tpl) return TB_ERR_STATE;
    SpeciesDev *const spp[2] = {&sp[0], &sp[1]};
    std::vector<Domain> doms(1, make_domain(spp, &M.t[0], cfg.rs, 0));
    int rc = launch_domains(s, M, doms, tpl, cfg, d_cnt, mode, sp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, s, M.esum, M.ctr, d_sums);
    return hipGetLastError() == hipSuccess ? TB_OK : TB_ERR_HIP;
}

int walk_mode(const MixedDev &M, const tb_config &cfg)
{
    // the fixed-theta walk is the fallback while no trained network is loaded
    return (cfg.mac == TB_MAC_LEARNED && M.mlp.ready) ? MODE_LEARNED : MODE_FIXED;
}

}  // namespace

int tb_mixed_prep(hipStream_t s, MixedDev &M, const JointDev *jt, size_t *bytes)
{
    if (!M.ctr) {
        if (hipMalloc((void **)&M.ctr, 64) != hipSuccess) {
            g_err = "mixed walk: counter allocation failed";
            return TB_ERR_NOMEM;
        }
        *bytes += 64;
    }
    hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, s, M.ctr, 2, 3);
    const JointDev &J = *jt;
    if (J.nn >= (1 << 28)) {
        // a cell's record holds its first child's index in bits 3..30
        g_err = "mixed walk: joint tree of 2^28 nodes or more";
        return TB_ERR_NOMEM;
    }
    MixedTree &T = M.t[0];
    int cap = T.cap, rc;
    int c1 = cap, c2 = cap, c3 = cap, c4 = cap, c5 = cap, c6 = cap, c7 = cap, c8 = cap, c9 = cap;
    int c10 = cap, c11 = cap, c12 = cap, c13 = cap * TB_JQ;  // bnq: TB_JQ doubles per node
    int c14 = cap, c15 = cap;
    if ((rc = grow(&T.bmac, &c1, J.nn, bytes))) return rc;
    if ((rc = grow(&T.binfo, &c2, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq0, &c3, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq1, &c4, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq2, &c5, J.nn, bytes))) return rc;
    if ((rc = grow(&T.inv, &c6, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bf, &c7, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq3, &c8, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq4, &c9, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq5, &c14, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bq6, &c15, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bflag, &c10, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bkey, &c11, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bcnt, &c12, J.nn, bytes))) return rc;
    if ((rc = grow(&T.bnq, &c13, J.nn * TB_JQ, bytes))) return rc;
    T.cap = std::min(std::min(std::min(std::min(c1, c2), std::min(c3, c4)), std::min(std::min(c5, c6), c7)),
                     std::min(std::min(std::min(c8, c9), std::min(c14, c15)),
                              std::min(std::min(c10, c11), std::min(c12, c13 / TB_JQ))));
    T.nn = J.nn;
    if (hipGetLastError() != hipSuccess) {
        g_err = "mixed walk: prep launch failed";
        return TB_ERR_HIP;
    }
    return TB_OK;
}

CellRecs tb_mixed_recs(MixedDev &M, const tb_config &cfg)
{
    MixedTree &T = M.t[0];
    CellRecs R;
    R.inv = T.inv;
    R.bmac = T.bmac;
    R.binfo = T.binfo;
    R.bq0 = T.bq0;
    R.bq1 = T.bq1;
    R.bq2 = T.bq2;
    R.bq3 = T.bq3;
    R.bq4 = T.bq4;
    R.bq5 = T.bq5;
    R.bq6 = T.bq6;
    R.bf = T.bf;
    R.bflag = T.bflag;
    R.bkey = T.bkey;
    R.bcnt = T.bcnt;
    R.bnq = T.bnq;
    R.theta = cfg.theta_mixed > 0.0 ? cfg.theta_mixed : cfg.theta;
    R.flag = M.ctr + 2;  // read with the walk's flags after the next force
    return R;
}

int tb_mixed_force(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *, double *d_sums)
{
    return launch(s, M, sp, cfg, d_sums, nullptr, walk_mode(M, cfg));
}

int tb_mixed_count(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *, double *d_sums,
                   int64_t *d_cnt)
{
    hipMemsetAsync(d_cnt, 0, TB_NCNT * sizeof(int64_t), s);
    return launch(s, M, sp, cfg, d_sums, d_cnt, walk_mode(M, cfg));
}

int tb_mixed_collect(hipStream_t s, MixedDev &M, SpeciesDev *sp, const JointDev *jt, const tb_config &cfg,
                     double *d_sums, int cap, float *X, float *y, int *nout)
{
    if (cap > M.rec_cap) {
        if (M.rec) hipFree(M.rec);
        if (M.lab) hipFree(M.lab);
        M.rec = nullptr;
        M.lab = nullptr;
        M.rec_cap = 0;
        if (hipMalloc((void **)&M.rec, (size_t)cap * sizeof(BandRec)) != hipSuccess ||
            hipMalloc((void **)&M.lab, (size_t)cap * sizeof(float)) != hipSuccess) {
            g_err = "collect: allocation failed";
            return TB_ERR_NOMEM;
        }
        M.rec_cap = cap;
    }
    int rc = launch(s, M, sp, cfg, d_sums, nullptr, MODE_COLLECT);
    if (rc) return rc;
    int32_t nrec = 0;
    if (hipMemcpyAsync(&nrec, M.ctr + 3, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        g_err = "collect: count read failed";
        return TB_ERR_HIP;
    }
    nrec = std::min(nrec, cap);
    LabelArgs A;
    A.rec = M.rec;
    A.nrec = nrec;
    A.ng0 = sp[0].ng;
    for (int k = 0; k < 2; k++) {
        const int c = sp[k].cur;
        A.px[k] = sp[k].x[c][0]; A.py[k] = sp[k].x[c][1]; A.pz[k] = sp[k].x[c][2];
        A.fx[k] = sp[k].f[c][0]; A.fy[k] = sp[k].f[c][1]; A.fz[k] = sp[k].f[c][2];
        A.n[k] = sp[k].n;
        A.grp[k] = sp[k].grp;
    }
    A.bfs = jt->bfs;
    A.bmac = M.t[0].bmac;
    A.bnq = M.t[0].bnq;
    A.nstart = jt->nstart;
    A.ncount = jt->ncount;
    A.pos = jt->pos;
    A.rs = cfg.rs;
    A.lab = M.lab;
    if (nrec > 0) hipLaunchKernelGGL(k_labels, dim3(nrec), dim3(64), 0, s, A);
    std::vector<BandRec> h((size_t)nrec);
    if (nrec > 0 && (hipMemcpyAsync(h.data(), M.rec, (size_t)nrec * sizeof(BandRec), hipMemcpyDeviceToHost, s) !=
                         hipSuccess ||
                     hipMemcpyAsync(y, M.lab, (size_t)nrec * 4, hipMemcpyDeviceToHost, s) != hipSuccess)) {
        g_err = "collect: read-back failed";
        return TB_ERR_HIP;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        g_err = "collect: labels failed";
        return TB_ERR_HIP;
    }
    for (int r = 0; r < nrec; r++)
        for (int j = 0; j < TB_MLP_NIN; j++) X[(size_t)r * TB_MLP_NIN + j] = h[r].f[j];
    *nout = nrec;
    return TB_OK;
}

int tb_mixed_force_ens(hipStream_t s, MixedDev &M, SpeciesDev *sp, int nreal, const uint32_t *roots,
                       const double *rs, const int32_t *const gfirst[2], const int32_t *const gcount[2],
                       const tb_config &cfg, double *esum_out_dev)
{
    const int tpl = group_tpl(sp[0], sp[1]);
    if (!tpl) return TB_ERR_STATE;
    SpeciesDev *const spp[2] = {&sp[0], &sp[1]};
    std::vector<Domain> doms;
    doms.reserve(nreal);
    int gb = 0;
    for (int r = 0; r < nreal; r++) {
        Domain d = make_domain(spp, &M.t[0], rs[r], gb);
        for (int k = 0; k < 2; k++) {
            d.grp[k] = sp[k].grp + gfirst[k][r];
            d.ng[k] = gcount[k][r];
            d.gfirst[k] = gfirst[k][r];
        }
        d.root = roots[r];
        doms.push_back(d);
        gb += d.ng[0] + d.ng[1];
    }
    int rc = launch_domains(s, M, doms, tpl, cfg, nullptr, MODE_FIXED, sp);
    if (rc) return rc;
    if (esum_out_dev &&
        hipMemcpyAsync(esum_out_dev, M.esum, (size_t)4 * nreal * sizeof(double), hipMemcpyDeviceToDevice, s) !=
            hipSuccess) {
        g_err = "mixed walk: energy copy failed";
        return TB_ERR_HIP;
    }
    return TB_OK;
}

void tb_mixed_free(MixedDev &M, size_t *bytes)
{
    MixedTree &T = M.t[0];
    hipFree(T.bmac); hipFree(T.binfo); hipFree(T.bq0); hipFree(T.bq1); hipFree(T.bq2); hipFree(T.inv);
    hipFree(T.bf); hipFree(T.bq3); hipFree(T.bq4); hipFree(T.bq5); hipFree(T.bq6);
    hipFree(T.bflag); hipFree(T.bkey); hipFree(T.bcnt); hipFree(T.bnq);
    T = MixedTree();
    if (M.far_cells) hipFree(M.far_cells);
    if (M.far_link) hipFree(M.far_link);
    if (M.far_head) hipFree(M.far_head);
    if (M.gemin) hipFree(M.gemin);
    M.gemin = nullptr;
    hipFree(M.sfront); hipFree(M.sfar); hipFree(M.scen); hipFree(M.sloc);
    M.sfront = M.sfar = nullptr;
    M.scen = nullptr;
    M.sloc = nullptr;
    M.seg_cap = 0;
    M.far_cells = nullptr;
    M.far_link = M.far_head = nullptr;
    M.far_cap = M.head_cap = 0;
    hipFree(M.ctr);
    if (M.ddom) hipFree(M.ddom);
    if (M.esum) hipFree(M.esum);
    if (M.rec) hipFree(M.rec);
    if (M.lab) hipFree(M.lab);
    tb_mlp_free(M.mlp);
    for (int k = 0; k < 2; k++)
        if (M.ev[k]) hipEventDestroy(M.ev[k]);
    M.ev[0] = M.ev[1] = nullptr;
    M.ctr = nullptr;
    M.ddom = nullptr;
    M.esum = nullptr;
    M.ddom_cap = M.esum_cap = 0;
    M.rec = nullptr;
    M.lab = nullptr;
    M.rec_cap = 0;
    (void)bytes;
}

const char *tb_mixed_error() { return g_err; }

float tb_mixed_walk_ms(MixedDev &M)
{
    float ms = -1.0f;
    if (!M.ev[0] || hipEventElapsedTime(&ms, M.ev[0], M.ev[1]) != hipSuccess) return -1.0f;
    return ms;
}

// autotuner table: log2 bucket of the total particle count -> targets per lane
static int g_tpl_table[64] = {0};

static int size_bucket(int n)
{
    int b = 0;
    while ((1 << (b + 1)) <= n && b < 62) b++;
    return b;
}

int tb_mixed_tpl(const tb_config &cfg, int n_total)
{
    if (cfg.group == 64 || cfg.group == 128 || cfg.group == 256) return cfg.group / 64;
    return tb_autotune_tpl(n_total);
}

int tb_autotune_tpl(int n_total)
{
    const int v = g_tpl_table[size_bucket(n_total)];
    return v ? v : 2;
}

void tb_autotune_record(int n_total, int tpl) { g_tpl_table[size_bucket(n_total)] = tpl; }

// list-length variant per size bucket (0, the longest lists, until autotuned)
static int g_list_table[64] = {0};

int tb_autotune_lists(int n_total) { return g_list_table[size_bucket(n_total)]; }

void tb_autotune_record_lists(int n_total, int variant) { g_list_table[size_bucket(n_total)] = variant; }

void tb_list_lengths(int variant, int *m, int *p)
{
    const int v = variant >= 0 && variant < TB_LIST_VARIANTS ? variant : 0;
    *m = LIST_M[v];
    *p = LIST_P[v];
}
