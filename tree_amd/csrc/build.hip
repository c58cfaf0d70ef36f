// build.hip -- tree construction on the GPU (replaces BuildLocalTree,
// tree.f90:39-100, Insert_Tree_e/_i 103-447 and PSeudoE/I 450-659).
//
// Pipeline per species (all on the handle's stream):
//   bbox      -> root centre (max+min)/2 and per-axis length max-min   (tree.f90:52-63)
//   keys      -> 21 octal digits from the reference's FP64 centre recurrence
//                c' = c + (+-0.25)*L, L' = L*0.5 with '<' going low   (tree.f90:114-173)
//   sort      -> LSD radix sort, 8-bit digits, LDS-staged wave-ranked scatter
//   lcp       -> leaf depth = 1 + max common-prefix with a neighbour; one
//                particle per leaf, single-child chains kept (insertion semantics)
//   emit      -> preorder node table (start particle, depth) = reference DFS
//                order; skip pointers replace the 8 child pointers
//   upward    -> level by level, child-point moments exactly as PSeudoE/I
//                (centroid, dx, D, Q, QQ over the children's centroids),
//                plus the exact second moment for the mixed walk.
// Contraction is off in this file so that every FP64 operation rounds as the
// reference's does; node tables are bit-identical to the reference's.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "tb_internal.h"

namespace {

constexpr int BT = 256;  // block threads

__device__ inline int lcp_digits(uint64_t a, uint64_t b)
{
    uint64_t x = a ^ b;
    if (x == 0) return TB_LEVELS;
    return (__clzll((long long)x) - 1) / 3;
}

// ------------------------------------------------------------------ bbox

__global__ __launch_bounds__(BT) void k_bbox_part(int n, const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ z, double *__restrict__ part)
{
    double mn0 = DBL_MAX, mn1 = DBL_MAX, mn2 = DBL_MAX, mx0 = -DBL_MAX, mx1 = -DBL_MAX, mx2 = -DBL_MAX;
    for (int i = blockIdx.x * BT + threadIdx.x; i < n; i += gridDim.x * BT) {
        double a = x[i], b = y[i], c = z[i];
        mn0 = fmin(mn0, a); mx0 = fmax(mx0, a);
        mn1 = fmin(mn1, b); mx1 = fmax(mx1, b);
        mn2 = fmin(mn2, c); mx2 = fmax(mx2, c);
    }
    __shared__ double s[6][BT];
    s[0][threadIdx.x] = mn0; s[1][threadIdx.x] = mn1; s[2][threadIdx.x] = mn2;
    s[3][threadIdx.x] = mx0; s[4][threadIdx.x] = mx1; s[5][threadIdx.x] = mx2;
    __syncthreads();
    for (int w = BT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int k = 0; k < 3; k++) s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
            for (int k = 3; k < 6; k++) s[k][threadIdx.x] = fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// min and max are exact and independent of order, so the block's strided
// reduction gives the same box as a serial pass over the partials
__global__ __launch_bounds__(BT) void k_bbox_final(int nb, const double *__restrict__ part, double *__restrict__ root)
{
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = threadIdx.x; b < nb; b += BT)
        for (int k = 0; k < 3; k++) {
            v[k] = fmin(v[k], part[b * 6 + k]);
            v[3 + k] = fmax(v[3 + k], part[b * 6 + 3 + k]);
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; k++) {
            v[k] = fmin(v[k], __shfl_xor(v[k], o));
            v[3 + k] = fmax(v[3 + k], __shfl_xor(v[3 + k], o));
        }
    __shared__ double red[BT / 64][6];
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) red[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < BT / 64; w++)
        for (int k = 0; k < 3; k++) {
            v[k] = fmin(v[k], red[w][k]);
            v[3 + k] = fmax(v[3 + k], red[w][3 + k]);
        }
    for (int k = 0; k < 3; k++) {
        root[k] = (v[3 + k] + v[k]) / 2.0;   // Roote%xx, tree.f90:58-60
        root[3 + k] = v[3 + k] - v[k];       // Roote%lx, tree.f90:61-63
    }
}

// ------------------------------------------------------------------ keys

__global__ __launch_bounds__(BT) void k_keys(int n, const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ z, const double *__restrict__ root,
                                             uint64_t *__restrict__ key, uint32_t *__restrict__ val,
                                             int32_t *__restrict__ err)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i], pz = z[i];
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(err, 2);
    double cx = root[0], cy = root[1], cz = root[2];
    double lx = root[3], ly = root[4], lz = root[5];
    uint64_t k = 0;
#pragma unroll
    for (int l = 0; l < TB_LEVELS; l++) {
        // strict '<' goes to the low half, ties and NaN to the high half
        uint32_t bx = (px < cx) ? 0u : 1u;
        uint32_t by = (py < cy) ? 0u : 1u;
        uint32_t bz = (pz < cz) ? 0u : 1u;
        k = (k << 3) | (uint64_t)(bx | (by << 1) | (bz << 2));
        cx = cx + (bx ? 0.25 : -0.25) * lx;
        cy = cy + (by ? 0.25 : -0.25) * ly;
        cz = cz + (bz ? 0.25 : -0.25) * lz;
        lx = lx * 0.5;
        ly = ly * 0.5;
        lz = lz * 0.5;
    }
    key[i] = k;
    val[i] = (uint32_t)i;
}

// ------------------------------------------------------------ radix sort

constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = BT * RS_ITEMS;  // 4096 keys per block
constexpr int RS_WAVES = BT / 64;
constexpr int RS_WSEG = 64 * RS_ITEMS;  // keys per wave

// lanes holding the same 8-bit digit
__device__ inline uint64_t match8(uint32_t d)
{
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

template<typename K>
__global__ __launch_bounds__(BT) void k_rs_up(int n, const K *__restrict__ key, int shift,
                                              uint32_t *__restrict__ hist, int nb)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long base = (long)blockIdx.x * RS_TILE + wave * RS_WSEG;
    // the tile's keys loaded together (all in flight at once), then a plain
    // LDS add per key: cheaper than grouping equal digits with ballots first
    // (k_rs_down needs those for its ranks, the counts do not)
    K k[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const long i = base + r * 64 + lane;
        k[r] = i < n ? key[i] : (K)~(K)0;  // pad: digit 255, sorts after all real keys
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) atomicAdd(&h[(uint32_t)(k[r] >> shift) & 255u], 1u);
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of 256 values held one per thread; returns the block total
__device__ inline uint32_t scan256_excl(uint32_t v, uint32_t *lds, uint32_t *total)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t t = threadIdx.x >= off ? lds[threadIdx.x - off] : 0u;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t incl = lds[threadIdx.x];
    if (total) *total = lds[255];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(BT) void k_rs_scan(int nb, uint32_t *__restrict__ hist, uint32_t *__restrict__ totals)
{
    __shared__ uint32_t lds[256];
    uint32_t *row = hist + (long)blockIdx.x * nb;
    int per = (nb + BT - 1) / BT;
    int beg = threadIdx.x * per, end = min(beg + per, nb);
    uint32_t sum = 0;
    for (int j = beg; j < end; j++) sum += row[j];
    uint32_t tot;
    uint32_t run = scan256_excl(sum, lds, &tot);
    for (int j = beg; j < end; j++) {
        uint32_t v = row[j];
        row[j] = run;
        run += v;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

// kout may be null: the permutation alone (vout) is wanted; vinv, if given,
// receives its inverse (vinv[vout[g]] = g)
template<typename K>
__global__ __launch_bounds__(BT) void k_rs_down(int n, const K *__restrict__ kin,
                                                const uint32_t *__restrict__ vin, K *__restrict__ kout,
                                                uint32_t *__restrict__ vout, int shift,
                                                const uint32_t *__restrict__ hist, const uint32_t *__restrict__ totals,
                                                int nb, int iota, uint32_t *__restrict__ vinv)
{
    __shared__ K skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t bstart[256];
    __shared__ uint32_t gstart[256];
    __shared__ uint32_t lds[256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int w = 0; w < RS_WAVES; w++) wcnt[w][threadIdx.x] = 0;
    __syncthreads();

    const long tile = (long)blockIdx.x * RS_TILE;
    const long base = tile + wave * RS_WSEG;
    K k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t pos[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        long i = base + r * 64 + lane;
        k[r] = i < n ? kin[i] : (K)~(K)0;
        v[r] = i < n ? (iota ? (uint32_t)i : vin[i]) : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint64_t m = match8(d);
        uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        uint32_t rank = (uint32_t)__popcll(m & lt);
        uint32_t prior = wcnt[wave][d];
        __builtin_amdgcn_wave_barrier();
        if ((int)__builtin_ctzll(m) == lane) wcnt[wave][d] = prior + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        pos[r] = prior + rank;
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
        for (int w = 0; w < RS_WAVES; w++) {
            uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
        uint32_t bs = scan256_excl(run, lds, nullptr);
        bstart[d] = bs;
        uint32_t gb = scan256_excl(totals[d], lds, nullptr);
        gstart[d] = gb + hist[(long)d * nb + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        uint32_t d = (uint32_t)(k[r] >> shift) & 255u;
        uint32_t lp = bstart[d] + wcnt[wave][d] + pos[r];
        skey[lp] = k[r];
        sval[lp] = v[r];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < RS_TILE; j += BT) {
        K kk = skey[j];
        uint32_t d = (uint32_t)(kk >> shift) & 255u;
        long g = (long)gstart[d] + (j - (long)bstart[d]);
        if (g < n) {
            if (kout) kout[g] = kk;
            vout[g] = sval[j];
            if (vinv) vinv[sval[j]] = (uint32_t)g;
        }
    }
}

// -------------------------------------------------------------- permute

struct PermCols {
    const double *src[9];
    double *dst[9];
};

__global__ __launch_bounds__(BT) void k_permute(int n, const uint32_t *__restrict__ perm, PermCols cols,
                                                const int32_t *__restrict__ osrc, int32_t *__restrict__ odst,
                                                int32_t *__restrict__ err)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    uint32_t j = perm[i];
    if (j >= (uint32_t)n) {  // never for a valid sort; reported instead of read out of bounds
        atomicOr(err, 4);
        return;
    }
#pragma unroll
    for (int a = 0; a < 9; a++) cols.dst[a][i] = cols.src[a][j];
    odst[i] = osrc[j];
}

// ------------------------------------------------------------------ scan

constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = BT * SC_ITEMS;

__global__ __launch_bounds__(BT) void k_scan_part(int n, const int32_t *__restrict__ in, int32_t *__restrict__ bsum)
{
    __shared__ uint32_t lds[256];
    long base = (long)blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    int32_t s = 0;
    for (int r = 0; r < SC_ITEMS; r++) {
        long i = base + r;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    scan256_excl((uint32_t)s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = (int32_t)tot;
}

__global__ __launch_bounds__(BT) void k_scan_bsum(int nb, int32_t *__restrict__ bsum, int32_t *__restrict__ total)
{
    __shared__ uint32_t lds[256];
    int per = (nb + BT - 1) / BT;
    int beg = threadIdx.x * per, end = min(beg + per, nb);
    uint32_t s = 0;
    for (int j = beg; j < end; j++) s += (uint32_t)bsum[j];
    uint32_t tot;
    uint32_t run = scan256_excl(s, lds, &tot);
    for (int j = beg; j < end; j++) {
        uint32_t v = (uint32_t)bsum[j];
        bsum[j] = (int32_t)run;
        run += v;
    }
    if (threadIdx.x == 0) *total = (int32_t)tot;
}

__global__ __launch_bounds__(BT) void k_scan_final(int n, const int32_t *__restrict__ in,
                                                   const int32_t *__restrict__ bsum, int32_t *__restrict__ out)
{
    __shared__ uint32_t lds[256];
    long base = (long)blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
    int32_t v[SC_ITEMS];
    int32_t s = 0;
    for (int r = 0; r < SC_ITEMS; r++) {
        long i = base + r;
        v[r] = i < n ? in[i] : 0;
        s += v[r];
    }
    uint32_t run = scan256_excl((uint32_t)s, lds, nullptr) + (uint32_t)bsum[blockIdx.x];
    for (int r = 0; r < SC_ITEMS; r++) {
        long i = base + r;
        if (i < n) out[i] = (int32_t)run;
        run += (uint32_t)v[r];
    }
}

// ------------------------------------------------------------------- lcp

// A block walks chunks of BT keys (block-strided, so every wave runs the same
// number of passes) and keeps its depth histogram in LDS across them: the
// grid is capped, and the global histogram takes one set of atomics per
// block instead of one per BT keys.
//
// Particles whose keys agree in all TB_LEVELS digits: a species tree (the
// reference's one particle per leaf) raises the coincidence flag.  The joint
// tree of the mixed walk instead keeps the finest cell as a twig whose
// children are its particles' leaves at depth TB_LEVELS + 1, for up to 8
// particles (an electron-ion pair a few nm apart is physical: the Kelbg
// potential is finite at r = 0); 9 or more raise the flag.
__global__ __launch_bounds__(BT) void k_lcp(int n, const uint64_t *__restrict__ key, int32_t *__restrict__ cnt,
                                            int32_t *__restrict__ hist, int32_t *__restrict__ err, int joint)
{
    __shared__ int32_t lh[2][TB_LEVELS + 2];
    if (threadIdx.x < 2 * (TB_LEVELS + 2)) (&lh[0][0])[threadIdx.x] = 0;
    __syncthreads();
    for (long base = (long)blockIdx.x * BT; base < n; base += (long)gridDim.x * BT) {
        const long i = base + threadIdx.x;
        int lm = TB_LEVELS + 2, leafd = 0;  // no depth counted for i >= n
        if (i < n) {
            uint64_t k = key[i];
            lm = i > 0 ? lcp_digits(key[i - 1], k) : 0;
            int lp = i < n - 1 ? lcp_digits(k, key[i + 1]) : 0;
            int mx = max(lm, lp);
            if (mx >= TB_LEVELS) {
                if (joint) {
                    if (i >= 8 && lcp_digits(key[i - 8], k) >= TB_LEVELS) atomicOr(err, 1);
                    mx = TB_LEVELS;
                } else {
                    atomicOr(err, 1);
                    mx = TB_LEVELS - 1;
                    lm = min(lm, mx);
                }
            }
            leafd = 1 + mx;
            cnt[i] = leafd - lm;
        }
        // particle i opens the twigs at depths lm+1 .. leafd-1 and a leaf at
        // depth leafd: counted per wavefront by ballots, one LDS add per depth
        for (int d = 1; d < TB_LEVELS + 2; d++) {
            const uint64_t tw = __ballot(lm < d && d < leafd);
            const uint64_t lf = __ballot(leafd == d);
            if ((threadIdx.x & 63) == 0) {
                if (tw) atomicAdd(&lh[0][d], (int32_t)__popcll(tw));
                if (lf) atomicAdd(&lh[1][d], (int32_t)__popcll(lf));
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * (TB_LEVELS + 2)) {
        int32_t v = (&lh[0][0])[threadIdx.x];
        if (v) atomicAdd(&hist[threadIdx.x], v);
    }
}

// ------------------------------------------------------------------ emit

__global__ __launch_bounds__(BT) void k_emit(int n, const uint64_t *__restrict__ key, const int32_t *__restrict__ offs,
                                             int total, const double *__restrict__ px, const double *__restrict__ py,
                                             const double *__restrict__ pz, double sign, double4 *__restrict__ nmac,
                                             uint32_t *__restrict__ nnext, int32_t *__restrict__ nstart,
                                             int32_t *__restrict__ ncount,
                                             int32_t *__restrict__ ndepth)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    if (i == 0) {
        ndepth[0] = 0; nstart[0] = 0; ncount[0] = n; nnext[0] = (uint32_t)total;
    }
    uint64_t k = key[i];
    int lm = i > 0 ? lcp_digits(key[i - 1], k) : 0;
    int lp = i < n - 1 ? lcp_digits(k, key[i + 1]) : 0;
    int mx = min(max(lm, lp), TB_LEVELS - 1);
    lm = min(lm, mx);
    int leafd = 1 + mx;
    int base = offs[i] + 1;
    // deepest first: a cell holds the one a level deeper that starts at the
    // same particle, so its end is at or past that one's and the search for
    // it starts there (one gallop over all the levels, not one per level)
    int end = i + 1;
    for (int d = leafd; d >= lm + 1; d--) {
        int id = base + (d - lm - 1);
        ndepth[id] = d;
        nstart[id] = i;
        if (d < leafd) {
            // first particle past the cell: gallop from the deeper cell's last
            // particle, then bisect on the key prefix
            int sh = 3 * (TB_LEVELS - d);
            uint64_t p = k >> sh;
            const int b = end - 1;
            int lo = b, hi = end, step = 1;
            while (hi < n && (key[hi] >> sh) == p) {
                lo = hi;
                step <<= 1;
                hi = b + step;
            }
            if (hi > n) hi = n;
            while (hi - lo > 1) {
                int mid = lo + ((hi - lo) >> 1);
                if ((key[mid] >> sh) == p) lo = mid;
                else hi = mid;
            }
            end = hi;
        }
        ncount[id] = end - i;
        uint32_t nxt = end >= n ? (uint32_t)total : (uint32_t)(offs[end] + 1);
        if (d == leafd) {
            nnext[id] = nxt | TB_LEAF;
            // leaf xx = particle, tree.f90:461; its moments (qeff = sign,
            // no D, Q, QQ, dx) are implied: the walk takes leaves by P2P and
            // the upward pass and the export know a leaf's values
            nmac[id] = make_double4(px[i], py[i], pz[i], -1.0);
        } else {
            nnext[id] = nxt;
        }
    }
}

// ---------------------------------------------------------------- upward

// The inverse of the breadth-first order (preorder node -> breadth-first
// index) for the species trees.  The joint tree's comes from its depth-order
// scatter instead (faster there; here, beside the other species' stream, the
// scatter inside the sort was slower: 1.61 against 1.59 ms per build).
__global__ __launch_bounds__(BT) void k_bfs_inv(int nn, const uint32_t *__restrict__ bfs, uint32_t *__restrict__ binv)
{
    const int j = blockIdx.x * BT + threadIdx.x;
    if (j < nn) binv[bfs[j]] = (uint32_t)j;
}

// Breadth-first companions for the upward pass, one thread per node in
// preorder (reads in order, writes to the node's breadth-first slot):
// whether the node is its parent's first child (in preorder a first child
// directly follows its parent, one level up; any other node follows a deeper
// one) and whether it is a leaf; a leaf's position and qeff (its particle,
// the species' sign); a cell's first particle's key.
__global__ __launch_bounds__(BT) void k_bfs_prep(int nn, const uint32_t *__restrict__ binv,
                                                 const uint32_t *__restrict__ nnext,
                                                 const int32_t *__restrict__ ndepth,
                                                 const int32_t *__restrict__ nstart,
                                                 const uint64_t *__restrict__ key,
                                                 const double4 *__restrict__ nmac, double sign,
                                                 uint8_t *__restrict__ bflag, uint64_t *__restrict__ bkey,
                                                 double4 *__restrict__ bxq)
{
    const int q = blockIdx.x * BT + threadIdx.x;
    if (q >= nn) return;
    const uint32_t j = binv[q];
    const bool leaf = (nnext[q] & TB_LEAF) != 0;
    const bool first = q > 0 && ndepth[q - 1] == ndepth[q] - 1;
    bflag[j] = (uint8_t)((first ? 1 : 0) | (leaf ? 2 : 0));
    if (leaf) {
        const double4 m = nmac[q];
        bxq[j] = make_double4(m.x, m.y, m.z, sign);
    } else {
        bkey[j] = key[nstart[q]];
    }
}

// Each twig's children: the run of breadth-first nodes from its first child
// (p + 1 in preorder) to the next first child.  The first node of every
// level is a first child, so a run never crosses into the next level.  At
// most 8 children, the first one flagged so, else a malformed table (flag 8).
__global__ __launch_bounds__(BT) void k_bfs_kids(int nn, const uint32_t *__restrict__ bfs,
                                                 const uint32_t *__restrict__ binv,
                                                 const uint8_t *__restrict__ bflag, uint32_t *__restrict__ bkids,
                                                 int32_t *__restrict__ err)
{
    const int b = blockIdx.x * BT + threadIdx.x;
    if (b >= nn || (bflag[b] & 2)) return;
    const uint32_t p = bfs[b];
    if (p + 1 >= (uint32_t)nn) {
        atomicOr(err, 8);
        return;
    }
    const int j0 = (int)binv[p + 1];
    if (j0 <= b || !(bflag[j0] & 1)) {
        atomicOr(err, 8);
        return;
    }
    int nch = 1;
    while (j0 + nch < nn && !(bflag[j0 + nch] & 1)) {
        if (++nch > 8) {
            atomicOr(err, 8);
            return;
        }
    }
    bkids[b] = ((uint32_t)j0 << 3) | (uint32_t)(nch - 1);
}

// One level of twigs, deepest level first.  A twig's children are the run
// of the next level [binv[p + 1], next first child): preorder order, which
// is the reference's child order 1..8.  Each twig
// - replays its geometric centre from its first particle's key (Insert_Tree's
//   placement);
// - computes centroid, dx, D, Q, QQ exactly as PSeudoE/I (tree.f90:450-553),
//   from its children's centroids and qeff = sign * count (its own qeff is
//   their sum: integers, exact);
// - writes the preorder table and its own breadth-first companion record.
// Children outside the next level are a malformed table: flag 8.
__global__ __launch_bounds__(BT) void k_upward(int beg, int cnt, int cbeg, int cend, int d,
                                               const uint32_t *__restrict__ bfs, const uint32_t *__restrict__ bkids,
                                               const uint8_t *__restrict__ bflag, const uint64_t *__restrict__ bkey,
                                               const double *__restrict__ root, double theta,
                                               double4 *__restrict__ nmac, double *__restrict__ nmom,
                                               double *__restrict__ ndx, double4 *__restrict__ bxq,
                                               int32_t *__restrict__ err)
{
    const int t = blockIdx.x * BT + threadIdx.x;
    if (t >= cnt) return;
    const int b = beg + t;
    if (bflag[b] & 2) return;  // a leaf
    const uint32_t p = bfs[b];
    const uint32_t kids = bkids[b];
    const int j0 = (int)(kids >> 3), nch = (int)(kids & 7u) + 1;
    if (j0 < cbeg || j0 + nch > cend) {
        atomicOr(err, 8);
        return;
    }
    // geometric centre and size of the cell, replaying Insert_Tree's placement
    double ox = root[0], oy = root[1], oz = root[2];
    double lx = root[3], ly = root[4], lz = root[5];
    const uint64_t k = bkey[b];
    for (int l = 0; l < d; l++) {
        uint32_t dig = (uint32_t)(k >> (3 * (TB_LEVELS - 1 - l))) & 7u;
        ox = ox + ((dig & 1u) ? 0.25 : -0.25) * lx;
        oy = oy + ((dig & 2u) ? 0.25 : -0.25) * ly;
        oz = oz + ((dig & 4u) ? 0.25 : -0.25) * lz;
        lx = lx * 0.5;
        ly = ly * 0.5;
        lz = lz * 0.5;
    }
    // PSeudoE: centroid over the children, children in order 1..8
    double qsum = 0.0, X = 0.0, Y = 0.0, Z = 0.0;
    for (int c = 0; c < nch; c++) {
        const double4 m = bxq[j0 + c];
        const double q = m.w;
        qsum = qsum + q;
        X = X + q * m.x;
        Y = Y + q * m.y;
        Z = Z + q * m.z;
    }
    X = X / qsum;
    Y = Y / qsum;
    Z = Z / qsum;
    double ax = X - ox, ay = Y - oy, az = Z - oz;
    double dx = sqrt(ax * ax + ay * ay + az * az);
    double D0 = 0.0, D1 = 0.0, D2 = 0.0, Q0 = 0.0, Q1 = 0.0, Q2 = 0.0, R0 = 0.0, R1 = 0.0, R2 = 0.0;
    for (int c = 0; c < nch; c++) {
        const double4 m = bxq[j0 + c];
        const double q = m.w;
        double d0 = m.x - X, d1 = m.y - Y, d2 = m.z - Z;
        D0 = D0 + q * d0;
        D1 = D1 + q * d1;
        D2 = D2 + q * d2;
        Q0 = Q0 + q * (d0 * d0) / 2.0;
        Q1 = Q1 + q * (d1 * d1) / 2.0;
        Q2 = Q2 + q * (d2 * d2) / 2.0;
        R0 = R0 + q * d0 * d1;
        R1 = R1 + q * d1 * d2;
        R2 = R2 + q * d2 * d0;
    }
    double s = lx;
    if (ly > s) s = ly;
    if (lz > s) s = lz;
    double rcrit = s / theta + dx;   // MAC left side, tree.f90:742
    nmac[p] = make_double4(X, Y, Z, rcrit);
    const double qeff = qsum;  // sign * count
    double *mo = nmom + (size_t)p * 10;
    mo[0] = qeff;
    mo[1] = D0; mo[2] = D1; mo[3] = D2;
    mo[4] = Q0; mo[5] = Q1; mo[6] = Q2;
    mo[7] = R0; mo[8] = R1; mo[9] = R2;
    ndx[p] = dx;
    bxq[b] = make_double4(X, Y, Z, qeff);
}

// ------------------------------------------------------------ export

// a leaf's moments are implied (qeff = sign, D = Q = QQ = 0, dx = 0): its
// records are not written by the build
__global__ __launch_bounds__(BT) void k_export(int nn, double sign, const double4 *__restrict__ nmac,
                                               const uint32_t *__restrict__ nnext, const double *__restrict__ nmom,
                                               const double *__restrict__ ndx, const int32_t *__restrict__ nstart,
                                               const int32_t *__restrict__ ncount, const int32_t *__restrict__ ndepth,
                                               const int32_t *__restrict__ orig, int32_t *depth, int32_t *leaf,
                                               int32_t *pt, int32_t *count, int32_t *next, double *xx, double *qeff,
                                               double *dx, double *D, double *Q, double *QQ)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= nn) return;
    uint32_t nx = nnext[i];
    bool lf = (nx & TB_LEAF) != 0;
    depth[i] = ndepth[i];
    leaf[i] = lf ? 1 : 0;
    pt[i] = lf ? orig[nstart[i]] : -1;
    count[i] = ncount[i];
    next[i] = (int32_t)(nx & ~TB_LEAF);
    double4 m = nmac[i];
    xx[3 * i] = m.x; xx[3 * i + 1] = m.y; xx[3 * i + 2] = m.z;
    if (lf) {
        qeff[i] = sign;
        for (int k = 0; k < 3; k++) D[3 * i + k] = Q[3 * i + k] = QQ[3 * i + k] = 0.0;
        dx[i] = 0.0;
        return;
    }
    const double *mo = nmom + (long)i * 10;
    qeff[i] = mo[0];
    for (int k = 0; k < 3; k++) {
        D[3 * i + k] = mo[1 + k];
        Q[3 * i + k] = mo[4 + k];
        QQ[3 * i + k] = mo[7 + k];
    }
    dx[i] = ndx[i];
}

// ------------------------------------------------ joint (both species) tree
//
// The mixed walk uses one tree over electrons and ions together: its cells
// carry the net charge, dipole and quadrupole of both species about the cell's
// geometric centre.  In a quasi-neutral plasma the two species' far fields
// nearly cancel, so expanding the net charge distribution keeps the far-field
// error proportional to the (small) net field instead of to each species'
// (large) field.  Both species are keyed over the joint box with the same FP64
// centre recurrence (k_keys), so the joint order is a merge of their sorted
// keys.

// the two species' position columns (k_jgather's sources)
struct JSrc {
    const double *x[2], *y[2], *z[2];
};

// The joint key order from the two species' sorted keys (both formed over the
// joint box): a merge.  Electron e at rank i goes to i + (ions with a smaller
// key), ion j to j + (electrons with a key not larger), so equal keys keep
// the electrons first and each species' own order: the order a stable sort
// of the electrons' keys followed by the ions' gives.  The payload is the
// species-tagged index in the species' key order.  One block merges 256
// consecutive entries of one species: two searches bracket the other
// species' range its keys fall in, then each lane searches that bracket.
__global__ __launch_bounds__(BT) void k_jmerge(int ne, int ni, const uint64_t *__restrict__ ke,
                                               const uint64_t *__restrict__ ki, uint64_t *__restrict__ kout,
                                               uint32_t *__restrict__ vout)
{
    __shared__ int brk[2];
    const int nbe = (ne + BT - 1) / BT;
    const bool ion = (int)blockIdx.x >= nbe;
    const int base = (ion ? (int)blockIdx.x - nbe : (int)blockIdx.x) * BT;
    const int nself = ion ? ni : ne, nother = ion ? ne : ni;
    const uint64_t *self = ion ? ki : ke, *other = ion ? ke : ki;
    const int last = min(base + BT, nself) - 1;
    // count of other keys < key (electrons) or <= key (ions) in [lo, hi)
    auto rank = [&](uint64_t key, int lo, int hi) {
        while (lo < hi) {
            const int mid = lo + ((hi - lo) >> 1);
            const uint64_t o = other[mid];
            if (ion ? (o <= key) : (o < key)) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    if (threadIdx.x == 0) brk[0] = rank(self[base], 0, nother);
    if (threadIdx.x == 1) brk[1] = rank(self[last], 0, nother);
    __syncthreads();
    const int i = base + (int)threadIdx.x;
    if (i > last) return;
    const uint64_t key = self[i];
    const int pos = i + rank(key, brk[0], brk[1]);
    kout[pos] = key;
    vout[pos] = (ion ? 0x80000000u : 0u) | (uint32_t)i;
}

// sorted joint particles: position and charge (electrons -1, ions +1)
__global__ __launch_bounds__(BT) void k_jgather(int n, const uint32_t *__restrict__ val, JSrc src,
                                                double4 *__restrict__ pos)
{
    int j = blockIdx.x * BT + threadIdx.x;
    if (j >= n) return;
    const uint32_t v = val[j];
    const int sp = (int)(v >> 31);
    const uint32_t i = v & 0x7fffffffu;
    pos[j] = make_double4(src.x[sp][i], src.y[sp][i], src.z[sp][i], sp ? 1.0 : -1.0);
}

__global__ __launch_bounds__(BT) void k_jemit(int n, const uint64_t *__restrict__ key, const int32_t *__restrict__ offs,
                                              int total, const double4 *__restrict__ pos,
                                              double4 *__restrict__ nc, uint32_t *__restrict__ nnext,
                                              int32_t *__restrict__ nstart,
                                              int32_t *__restrict__ ncount, int32_t *__restrict__ ndepth)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    if (i == 0) {
        ndepth[0] = 0; nstart[0] = 0; ncount[0] = n; nnext[0] = (uint32_t)total;
    }
    uint64_t k = key[i];
    int lm = i > 0 ? lcp_digits(key[i - 1], k) : 0;
    int lp = i < n - 1 ? lcp_digits(k, key[i + 1]) : 0;
    // identical keys: leaves one level below the finest cell (see k_lcp)
    int mx = min(max(lm, lp), TB_LEVELS);
    lm = min(lm, mx);
    int leafd = 1 + mx;
    int base = offs[i] + 1;
    // deepest first, each search from the deeper cell's end (as k_emit)
    int end = i + 1;
    for (int d = leafd; d >= lm + 1; d--) {
        int id = base + (d - lm - 1);
        ndepth[id] = d;
        nstart[id] = i;
        if (d < leafd) {
            int sh = 3 * (TB_LEVELS - d);
            uint64_t p = k >> sh;
            const int b = end - 1;
            int lo = b, hi = end, step = 1;
            while (hi < n && (key[hi] >> sh) == p) {
                lo = hi;
                step <<= 1;
                hi = b + step;
            }
            if (hi > n) hi = n;
            while (hi - lo > 1) {
                int mid = lo + ((hi - lo) >> 1);
                if ((key[mid] >> sh) == p) lo = mid;
                else hi = mid;
            }
            end = hi;
        }
        ncount[id] = end - i;
        uint32_t nxt = end >= n ? (uint32_t)total : (uint32_t)(offs[end] + 1);
        if (d == leafd) {
            // a leaf: its particle's position and charge
            nnext[id] = nxt | TB_LEAF;
            nc[id] = pos[i];
        } else {
            nnext[id] = nxt;
        }
    }
}

// index tuples of the raw fourth moment's components, and the raw third
// moment's component of a (sorted or not) index triple
__constant__ const int kHi[15][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2},
                                     {0, 0, 2, 2}, {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 2, 2},
                                     {1, 1, 1, 1}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}, {2, 2, 2, 2}};
__constant__ const int kO3[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}},
                                       {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                                       {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};

// The mixed walk's breadth-first records.  A leaf: its particle and charge.
// A cell: centre and s / theta, first child (breadth-first) << 3 | children
// - 1, q, D, the traceless quadrupole 3 M - tr M, the traceless third moment
// O' = O - (d_ij t_k + d_ik t_j + d_jk t_i) / 5 with t_k = O_iik, the
// traceless fourth moment H' = H - (d_ij T_kl + 5 more pairings) / 7 +
// (d_ij d_kl + d_ik d_jl + d_il d_jk) T0 / 35 with T_kl = H_mmkl, T0 = T_mm,
// and the learned criterion's per-cell features (FP64, then rounded; see
// band_features in walk32.hip).
// (A leaf's other records, bq1..bq6 and bf, are never read.)
__device__ inline void leaf_record(const CellRecs &R, int b, double4 c, int32_t start)
{
    R.binfo[b] = TB_LEAF | (uint32_t)start;
    R.bmac[b] = make_double4(c.x, c.y, c.z, 0.0);
    R.bq0[b] = make_float4((float)c.w, 0.f, 0.f, 0.f);
}

__device__ inline void cell_record(const CellRecs &R, int b, double cx, double cy, double cz, double sz,
                                   const double *Q, int32_t count)
{
#pragma clang fp contract(fast)
    R.bmac[b] = make_double4(cx, cy, cz, sz / R.theta);
    const double tr = Q[4] + Q[5] + Q[6];
    R.bq0[b] = make_float4((float)Q[0], (float)Q[1], (float)Q[2], (float)Q[3]);
    R.bq1[b] = make_float4((float)(3.0 * Q[4] - tr), (float)(3.0 * Q[5] - tr), (float)(3.0 * Q[7]),
                           (float)(3.0 * Q[8]));
    // raw O: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    const double *O = Q + 10;
    const double tx = O[0] + O[3] + O[5], ty = O[1] + O[6] + O[8], tz = O[2] + O[7] + O[9];
    R.bq3[b] = make_float4((float)(O[0] - 0.6 * tx), (float)(O[1] - 0.2 * ty), (float)(O[2] - 0.2 * tz),
                           (float)(O[3] - 0.2 * tx));
    // bq4.w (H'xxxx) is the fourth-moment record's (cell_record4)
    float *q4 = (float *)&R.bq4[b];
    q4[0] = (float)O[4];
    q4[1] = (float)(O[6] - 0.6 * ty);
    q4[2] = (float)(O[7] - 0.2 * tz);
    const double kf = 0.125, cn = (double)count, tiny = 9.5367431640625e-07;
    const double qxx = 3.0 * Q[4] - tr, qyy = 3.0 * Q[5] - tr, qzz = 3.0 * Q[6] - tr;
    const double qn = sqrt(qxx * qxx + qyy * qyy + qzz * qzz +
                           2.0 * (9.0 * Q[7] * Q[7] + 9.0 * Q[8] * Q[8] + 9.0 * Q[9] * Q[9]));
    const double dn = sqrt(Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]);
    R.bq2[b] = make_float4((float)(3.0 * Q[9]), (float)sz, (float)cn, (float)(kf * log2(sz)));
    R.bf[b] = make_float4((float)(kf * log2(fabs(Q[0]) / cn + tiny)), (float)(kf * log2(dn / (cn * sz) + tiny)),
                          (float)(kf * log2(qn / (cn * sz * sz) + tiny)), (float)(kf * log2(cn)));
}

// the traceless fourth moment's 9 independent components from the raw
// H (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz yyyy yyyz yyzz yzzz zzzz)
__device__ inline void cell_record4(const CellRecs &R, int b, const double (&H)[15])
{
#pragma clang fp contract(fast)
    const double txx = H[0] + H[3] + H[5], tyy = H[3] + H[10] + H[12], tzz = H[5] + H[12] + H[14];
    const double txy = H[1] + H[6] + H[8], tyz = H[4] + H[11] + H[13], tzx = H[2] + H[7] + H[9];
    const double t0 = txx + tyy + tzz, s7 = 1.0 / 7.0, s35 = 1.0 / 35.0;
    ((float *)&R.bq4[b])[3] = (float)(H[0] - 6.0 * s7 * txx + 3.0 * s35 * t0);
    R.bq5[b] = make_float4((float)(H[1] - 3.0 * s7 * txy), (float)(H[2] - 3.0 * s7 * tzx),
                           (float)(H[3] - s7 * (txx + tyy) + s35 * t0), (float)(H[4] - s7 * tyz));
    R.bq6[b] = make_float4((float)(H[6] - 3.0 * s7 * txy), (float)(H[7] - s7 * tzx),
                           (float)(H[10] - 6.0 * s7 * tyy + 3.0 * s35 * t0), (float)(H[11] - 3.0 * s7 * tyz));
}

// Breadth-first companions of the joint tree, one thread per node in
// preorder (reads in order, writes to the node's breadth-first slot b):
// first-child and leaf flags (a first child directly follows its parent in
// preorder); a leaf's record; a cell's first particle's key and particle
// count.
__global__ __launch_bounds__(BT) void k_jprep(int nn, const uint32_t *__restrict__ nnext,
                                              const int32_t *__restrict__ ndepth,
                                              const int32_t *__restrict__ nstart,
                                              const int32_t *__restrict__ ncount, const uint64_t *__restrict__ key,
                                              const double4 *__restrict__ nc, CellRecs R)
{
    const int q = blockIdx.x * BT + threadIdx.x;
    if (q >= nn) return;
    const int b = (int)R.inv[q];
    const bool leaf = (nnext[q] & TB_LEAF) != 0;
    const bool first = q > 0 && ndepth[q - 1] == ndepth[q] - 1;
    R.bflag[b] = (uint8_t)((first ? 1 : 0) | (leaf ? 2 : 0));
    const int32_t st = nstart[q];
    if (leaf) {
        leaf_record(R, b, nc[q], st);
    } else {
        R.bkey[b] = key[st];
        R.bcnt[b] = ncount[q];
    }
}

// Each cell's children: the run of breadth-first nodes from its first child
// (p + 1 in preorder) to the next first child (the first node of a level is
// a first child, so a run stays in its level); 1 to 8 children, else a
// malformed table.
__global__ __launch_bounds__(BT) void k_jkids(int nn, const uint32_t *__restrict__ bfs, CellRecs R)
{
    const int b = blockIdx.x * BT + threadIdx.x;
    if (b >= nn || (R.bflag[b] & 2)) return;
    const uint32_t p = bfs[b];
    if (p + 1 >= (uint32_t)nn) {
        atomicOr(R.flag, 16);
        return;
    }
    const int j0 = (int)R.inv[p + 1];
    if (j0 <= b || !(R.bflag[j0] & 1)) {
        atomicOr(R.flag, 16);
        return;
    }
    int nch = 1;
    while (j0 + nch < nn && !(R.bflag[j0 + nch] & 1)) {
        if (++nch > 8) {
            atomicOr(R.flag, 16);
            return;
        }
    }
    R.binfo[b] = ((uint32_t)j0 << 3) | (uint32_t)(nch - 1);
}

// net charge q, dipole D, second moment M and third and fourth moments O, H
// of the cells of one level (deepest first) about their geometric centres,
// shifted up from the children's own expansion centres (a child's offsets s
// become s + a, a = child centre - cell centre); the children are a
// contiguous run of the next level, read from their breadth-first records.
// A block takes JUP_CELLS cells: its first half of wavefronts forms q, D, M,
// O and the records built from them, the second half (the same cells) the
// fourth moment and its records.  Split by wavefront, the two parts run side
// by side without divergence, and each thread's serial chain over the
// children is half as long (the top levels are latency-bound).
//
// A batched ensemble's keys start with T octal digits naming the realisation
// (its tag): the cells of realisation r start at depth T from its own box
// root[6 r ..]; the T levels above hold no cell of any realisation and are
// never walked (their moments and records are zero).  One domain: T = 0.
constexpr int JUP_CELLS = BT / 2;

__global__ __launch_bounds__(BT) void k_jupward(int beg, int cnt, int cbeg, int cend, int d, int T,
                                                const double *__restrict__ root, CellRecs R)
{
    const int part = threadIdx.x / JUP_CELLS;  // wave-uniform: 0 q..O, 1 H
    const int t = blockIdx.x * JUP_CELLS + (int)(threadIdx.x % JUP_CELLS);
    if (t >= cnt) return;
    const int b = beg + t;
    if (R.bflag[b] & 2) return;  // a leaf: its record is written
    const uint32_t kids = R.binfo[b];
    const int j0 = (int)(kids >> 3), nch = (int)(kids & 7u) + 1;
    if (j0 < cbeg || j0 + nch > cend) {
        atomicOr(R.flag, 16);
        return;
    }
    double *o = R.bnq + (size_t)b * TB_JQ;
    if (d < T) {
        if (part == 0) {
            for (int c = 0; c < 20; c++) o[c] = 0.0;
            R.bmac[b] = make_double4(0.0, 0.0, 0.0, 0.0);
            R.bq0[b] = R.bq1[b] = R.bq2[b] = R.bq3[b] = R.bf[b] = make_float4(0.f, 0.f, 0.f, 0.f);
            float *q4 = (float *)&R.bq4[b];
            q4[0] = q4[1] = q4[2] = 0.f;
        } else {
            for (int c = 20; c < TB_JQ; c++) o[c] = 0.0;
            ((float *)&R.bq4[b])[3] = 0.f;
            R.bq5[b] = R.bq6[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    const uint64_t k = R.bkey[b];
    const double *rt = root + (T > 0 ? 6 * (size_t)(k >> (3 * (TB_LEVELS - T))) : 0);
    double ox = rt[0], oy = rt[1], oz = rt[2];
    double lx = rt[3], ly = rt[4], lz = rt[5];
    for (int l = T; l < d; l++) {
        uint32_t dig = (uint32_t)(k >> (3 * (TB_LEVELS - 1 - l))) & 7u;
        ox = ox + ((dig & 1u) ? 0.25 : -0.25) * lx;
        oy = oy + ((dig & 2u) ? 0.25 : -0.25) * ly;
        oz = oz + ((dig & 4u) ? 0.25 : -0.25) * lz;
        lx = lx * 0.5;
        ly = ly * 0.5;
        lz = lz * 0.5;
    }
    if (part == 1) {
        // H_ijkl += H_c + (O_jkl a_i + 3 more) + (M_ij a_k a_l + 5 more) +
        //           (D_i a_j a_k a_l + 3 more) + q a_i a_j a_k a_l, with the
        //           child's moments before this child's own shift (a particle:
        //           q a_i a_j a_k a_l alone, its other terms exact zeros)
        double H[15];
#pragma unroll
        for (int t2 = 0; t2 < 15; t2++) H[t2] = 0.0;
        for (int j = j0; j < j0 + nch; j++) {
            const double4 cc = R.bmac[j];
            const double av[3] = {cc.x - ox, cc.y - oy, cc.z - oz};
            if (R.bflag[j] & 2) {
                const double qc = (double)R.bq0[j].x;
#pragma unroll
                for (int t2 = 0; t2 < 15; t2++)
                    H[t2] += qc * av[kHi[t2][0]] * av[kHi[t2][1]] * av[kHi[t2][2]] * av[kHi[t2][3]];
                continue;
            }
            const double *cq = R.bnq + (size_t)j * TB_JQ;
            const double qc = cq[0], dv[3] = {cq[1], cq[2], cq[3]};
            const double mm[3][3] = {{cq[4], cq[7], cq[9]}, {cq[7], cq[5], cq[8]}, {cq[9], cq[8], cq[6]}};
            const double *oc = cq + 10;
#pragma unroll
            for (int t2 = 0; t2 < 15; t2++) {
                const int i = kHi[t2][0], jj = kHi[t2][1], kk = kHi[t2][2], ll = kHi[t2][3];
                H[t2] += cq[20 + t2] + oc[kO3[jj][kk][ll]] * av[i] + oc[kO3[i][kk][ll]] * av[jj] +
                         oc[kO3[i][jj][ll]] * av[kk] + oc[kO3[i][jj][kk]] * av[ll] +
                         mm[i][jj] * av[kk] * av[ll] + mm[i][kk] * av[jj] * av[ll] + mm[i][ll] * av[jj] * av[kk] +
                         mm[jj][kk] * av[i] * av[ll] + mm[jj][ll] * av[i] * av[kk] + mm[kk][ll] * av[i] * av[jj] +
                         dv[i] * av[jj] * av[kk] * av[ll] + dv[jj] * av[i] * av[kk] * av[ll] +
                         dv[kk] * av[i] * av[jj] * av[ll] + dv[ll] * av[i] * av[jj] * av[kk] +
                         qc * av[i] * av[jj] * av[kk] * av[ll];
            }
        }
        for (int c = 0; c < 15; c++) o[20 + c] = H[c];
        cell_record4(R, b, H);
        return;
    }
    double q = 0.0, D[3] = {0.0, 0.0, 0.0}, M[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double O[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int oi[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                           {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    for (int j = j0; j < j0 + nch; j++) {
        const double4 cc = R.bmac[j];
        const double a = cc.x - ox, b_ = cc.y - oy, e = cc.z - oz;
        if (R.bflag[j] & 2) {
            // a particle (charge +-1, exact in the FP32 record): the shift
            // below with D = M = O = 0, term for term (the dropped terms are
            // exact zeros)
            const double qc = (double)R.bq0[j].x, av[3] = {a, b_, e};
#pragma unroll
            for (int t2 = 0; t2 < 10; t2++) O[t2] += qc * av[oi[t2][0]] * av[oi[t2][1]] * av[oi[t2][2]];
            q += qc;
            D[0] += qc * a;
            D[1] += qc * b_;
            D[2] += qc * e;
            M[0] += qc * a * a;
            M[1] += qc * b_ * b_;
            M[2] += qc * e * e;
            M[3] += qc * a * b_;
            M[4] += qc * b_ * e;
            M[5] += qc * e * a;
            continue;
        }
        const double *cq = R.bnq + (size_t)j * TB_JQ;
        const double qc = cq[0], dx = cq[1], dy = cq[2], dz = cq[3];
        // O_ijk += O_c,ijk + (M_ij a_k + M_ik a_j + M_jk a_i) + (D_i a_j a_k + D_j a_i a_k + D_k a_i a_j)
        //          + q a_i a_j a_k, with the child's M and D (before this child's own shift)
        {
            const double av[3] = {a, b_, e}, dv[3] = {dx, dy, dz};
            const double mm[3][3] = {{cq[4], cq[7], cq[9]}, {cq[7], cq[5], cq[8]}, {cq[9], cq[8], cq[6]}};
#pragma unroll
            for (int t2 = 0; t2 < 10; t2++) {
                const int i = oi[t2][0], jj = oi[t2][1], kk = oi[t2][2];
                O[t2] += cq[10 + t2] + mm[i][jj] * av[kk] + mm[i][kk] * av[jj] + mm[jj][kk] * av[i] +
                         dv[i] * av[jj] * av[kk] + dv[jj] * av[i] * av[kk] + dv[kk] * av[i] * av[jj] +
                         qc * av[i] * av[jj] * av[kk];
            }
        }
        q += qc;
        D[0] += dx + qc * a;
        D[1] += dy + qc * b_;
        D[2] += dz + qc * e;
        M[0] += cq[4] + 2.0 * dx * a + qc * a * a;
        M[1] += cq[5] + 2.0 * dy * b_ + qc * b_ * b_;
        M[2] += cq[6] + 2.0 * dz * e + qc * e * e;
        M[3] += cq[7] + dx * b_ + dy * a + qc * a * b_;
        M[4] += cq[8] + dy * e + dz * b_ + qc * b_ * e;
        M[5] += cq[9] + dz * a + dx * e + qc * e * a;
    }
    double s = lx;
    if (ly > s) s = ly;
    if (lz > s) s = lz;
    double Q[20];
    Q[0] = q;
    for (int c = 0; c < 3; c++) Q[1 + c] = D[c];
    for (int c = 0; c < 6; c++) Q[4 + c] = M[c];
    for (int c = 0; c < 10; c++) Q[10 + c] = O[c];
    for (int c = 0; c < 20; c++) o[c] = Q[c];
    cell_record(R, b, ox, oy, oz, s, Q, R.bcnt[b]);
}

// ------------------------------------------------ mixed-walk target groups
//
// The mixed walk treats 64*TPL targets of one species as one group and
// accepts a cell only when it is far from the group's bounding box.  Groups
// of consecutive key-ordered particles that straddle a coarse cell boundary
// get boxes spanning a large part of the domain and walk almost every leaf.
// So the key-ordered sequence is first cut into segments, each inside one
// cell of at most W = 8G particles: a cut goes between j-1 and j when the
// deepest cell holding both has more than W particles.  Each segment (at most
// W long) is then split into ceil(m/G) near-equal groups.
//
// The cell at level d holding j has more than W particles exactly when some
// window key[s..s+W] with s in [j-W, j] shares d leading digits, so with
// w[s] = lcp(key[s], key[s+W]) the cut test is lcp(key[j-1], key[j]) <=
// max(w[j-W..j]).  That window maximum is max(suffix max of w over j-W's
// block of W, prefix max of w over j's block), both from k_gwin.

__global__ __launch_bounds__(BT) void k_gwin(int n, int W, const uint64_t *__restrict__ key,
                                             int8_t *__restrict__ pre, int8_t *__restrict__ suf)
{
    __shared__ int tp[BT], ts[BT];
    const int E = W / BT;  // 2, 4 or 8 entries per thread
    const long base = (long)blockIdx.x * W + (long)threadIdx.x * E;
    int w[8];
    int run = -1;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        w[e] = -1;
        const long q = base + e;
        if (e < E && q + W < n) w[e] = lcp_digits(key[q], key[q + W]);
        run = max(run, w[e]);
    }
    tp[threadIdx.x] = run;
    ts[threadIdx.x] = run;
    __syncthreads();
    // inclusive max-scans of the per-thread maxima, forwards and backwards
    for (int o = 1; o < BT; o <<= 1) {
        const int a = threadIdx.x >= o ? tp[threadIdx.x - o] : -1;
        const int c = threadIdx.x + o < BT ? ts[threadIdx.x + o] : -1;
        __syncthreads();
        tp[threadIdx.x] = max(tp[threadIdx.x], a);
        ts[threadIdx.x] = max(ts[threadIdx.x], c);
        __syncthreads();
    }
    int m = threadIdx.x > 0 ? tp[threadIdx.x - 1] : -1;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        m = max(m, w[e]);
        if (e < E && base + e < n) pre[base + e] = (int8_t)m;
    }
    m = threadIdx.x + 1 < BT ? ts[threadIdx.x + 1] : -1;
#pragma unroll
    for (int e = 7; e >= 0; e--) {
        m = max(m, w[e]);
        if (e < E && base + e < n) suf[base + e] = (int8_t)m;
    }
}

// flag[j] = 1 where a segment starts
// tagmask: key bits naming the realisation of a batched ensemble (0: one
// domain); a segment never spans two realisations
__global__ __launch_bounds__(BT) void k_gflag(int n, int W, const uint64_t *__restrict__ key,
                                              const int8_t *__restrict__ pre, const int8_t *__restrict__ suf,
                                              uint64_t tagmask, int32_t *__restrict__ flag)
{
    const int j = blockIdx.x * BT + threadIdx.x;
    if (j >= n) return;
    int f = 1;
    if (j > 0) {
        int mx = pre[j];
        if (j >= W) mx = max(mx, (int)suf[j - W]);
        f = (lcp_digits(key[j - 1], key[j]) <= mx || ((key[j - 1] ^ key[j]) & tagmask) != 0) ? 1 : 0;
    }
    flag[j] = f;
}

// batched ensemble: first target group of each realisation (its tag, the key
// bits from tagshift up, of the group's first particle)
__global__ __launch_bounds__(BT) void k_gfirst(int n, const int32_t *__restrict__ ng, const int2 *__restrict__ grp,
                                               const uint64_t *__restrict__ key, int tagshift,
                                               int32_t *__restrict__ first)
{
    const int g = blockIdx.x * BT + threadIdx.x;
    if (g >= n || g >= *ng) return;
    const uint64_t r = key[grp[g].x] >> tagshift;
    if (g == 0 || (key[grp[g - 1].x] >> tagshift) != r) first[r] = g;
}

__global__ __launch_bounds__(BT) void k_gseg(int n, const int32_t *__restrict__ flag,
                                             const int32_t *__restrict__ idx, int32_t *__restrict__ seg)
{
    const int j = blockIdx.x * BT + threadIdx.x;
    if (j < n && flag[j]) seg[idx[j]] = j;
}

// groups per segment (zero past the last segment, so the scan can run over n)
__global__ __launch_bounds__(BT) void k_gcount(int n, int G, const int32_t *__restrict__ nseg,
                                               const int32_t *__restrict__ seg, int32_t *__restrict__ cnt)
{
    const int k = blockIdx.x * BT + threadIdx.x;
    if (k >= n) return;
    const int S = *nseg;
    if (k >= S) {
        cnt[k] = 0;
        return;
    }
    const int m = (k + 1 < S ? seg[k + 1] : n) - seg[k];
    cnt[k] = (m + G - 1) / G;
}

__global__ __launch_bounds__(BT) void k_gemit(int n, const int32_t *__restrict__ nseg,
                                              const int32_t *__restrict__ seg, const int32_t *__restrict__ cnt,
                                              const int32_t *__restrict__ off, int2 *__restrict__ grp,
                                              int2 *__restrict__ sgrp, int32_t *__restrict__ gsid)
{
    const int k = blockIdx.x * BT + threadIdx.x;
    if (k >= n || k >= *nseg) return;
    const int st = seg[k], m = (k + 1 < *nseg ? seg[k + 1] : n) - st, p = cnt[k], o = off[k];
    for (int q = 0; q < p; q++) {
        const int a = st + (int)((long)q * m / p), b = st + (int)((long)(q + 1) * m / p);
        grp[o + q] = make_int2(a, b - a);
        gsid[o + q] = k;
    }
    sgrp[k] = make_int2(o, p);
}

// ----------------------------------------------- batched ensemble (segments)
//
// Realisation r holds particles [off[r], off[r+1]) of a combined array.

// partial min/max of each realisation: block (c, r) strides over r's slice
__global__ __launch_bounds__(BT) void k_sbbox_part(const int32_t *__restrict__ off, const double *__restrict__ x,
                                                   const double *__restrict__ y, const double *__restrict__ z,
                                                   double *__restrict__ part)
{
    const int r = blockIdx.y, b = off[r], e = off[r + 1];
    double mn0 = DBL_MAX, mn1 = DBL_MAX, mn2 = DBL_MAX, mx0 = -DBL_MAX, mx1 = -DBL_MAX, mx2 = -DBL_MAX;
    for (int i = b + blockIdx.x * BT + threadIdx.x; i < e; i += gridDim.x * BT) {
        double a = x[i], bb = y[i], c = z[i];
        mn0 = fmin(mn0, a); mx0 = fmax(mx0, a);
        mn1 = fmin(mn1, bb); mx1 = fmax(mx1, bb);
        mn2 = fmin(mn2, c); mx2 = fmax(mx2, c);
    }
    __shared__ double sm[6][BT];
    sm[0][threadIdx.x] = mn0; sm[1][threadIdx.x] = mn1; sm[2][threadIdx.x] = mn2;
    sm[3][threadIdx.x] = mx0; sm[4][threadIdx.x] = mx1; sm[5][threadIdx.x] = mx2;
    __syncthreads();
    for (int w = BT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int k = 0; k < 3; k++) sm[k][threadIdx.x] = fmin(sm[k][threadIdx.x], sm[k][threadIdx.x + w]);
            for (int k = 3; k < 6; k++) sm[k][threadIdx.x] = fmax(sm[k][threadIdx.x], sm[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[((size_t)r * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// box of each realisation (block r) from its partials of one or two arrays;
// min and max are exact, so the box is the one a single domain computes
__global__ __launch_bounds__(64) void k_sbbox_final(int chunks, const double *__restrict__ pa,
                                                    const double *__restrict__ pb, double *__restrict__ root)
{
    const int r = blockIdx.x;
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int src = 0; src < 2; src++) {
        const double *p = src == 0 ? pa : pb;
        if (!p) continue;
        for (int c = threadIdx.x; c < chunks; c += 64)
            for (int k = 0; k < 3; k++) {
                v[k] = fmin(v[k], p[((size_t)r * chunks + c) * 6 + k]);
                v[3 + k] = fmax(v[3 + k], p[((size_t)r * chunks + c) * 6 + 3 + k]);
            }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; k++) {
            v[k] = fmin(v[k], __shfl_xor(v[k], o));
            v[3 + k] = fmax(v[3 + k], __shfl_xor(v[3 + k], o));
        }
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; k++) {
        root[6 * r + k] = (v[3 + k] + v[k]) / 2.0;   // as k_bbox_final
        root[6 * r + 3 + k] = v[3 + k] - v[k];
    }
}

// keys of a batched ensemble: the realisation r as T octal digits, then 21 - T
// levels of the centre recurrence (k_keys) from r's own box
__global__ __launch_bounds__(BT) void k_ekeys(int n, const double *__restrict__ x, const double *__restrict__ y,
                                              const double *__restrict__ z, const int32_t *__restrict__ off,
                                              int nreal, const double *__restrict__ roots, int T,
                                              uint64_t *__restrict__ key, uint32_t *__restrict__ val, uint32_t tag,
                                              int32_t *__restrict__ err)
{
    int i = blockIdx.x * BT + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nreal - 1;  // the realisation holding i: off[lo] <= i < off[lo + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const double *rt = roots + 6 * (size_t)lo;
    double px = x[i], py = y[i], pz = z[i];
    if (!isfinite(px) || !isfinite(py) || !isfinite(pz)) atomicOr(err, 2);
    double cx = rt[0], cy = rt[1], cz = rt[2];
    double lx = rt[3], ly = rt[4], lz = rt[5];
    uint64_t k = (uint64_t)lo;
    for (int l = T; l < TB_LEVELS; l++) {
        uint32_t bx = (px < cx) ? 0u : 1u;
        uint32_t by = (py < cy) ? 0u : 1u;
        uint32_t bz = (pz < cz) ? 0u : 1u;
        k = (k << 3) | (uint64_t)(bx | (by << 1) | (bz << 2));
        cx = cx + (bx ? 0.25 : -0.25) * lx;
        cy = cy + (by ? 0.25 : -0.25) * ly;
        cz = cz + (bz ? 0.25 : -0.25) * lz;
        lx = lx * 0.5;
        ly = ly * 0.5;
        lz = lz * 0.5;
    }
    key[i] = k;
    val[i] = tag | (uint32_t)i;
}

inline int nblk(long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

// ====================================================================== host

size_t tb_sort_temp_bytes(int n)
{
    int nb = nblk(n, RS_TILE);
    return (size_t)256 * nb * sizeof(uint32_t) + 256 * sizeof(uint32_t) + 256;
}

size_t tb_scan_temp_bytes(int n) { return (size_t)(nblk(n, SC_TILE) + 64) * sizeof(int32_t); }

void tb_bbox(hipStream_t s, int n, const double *x, const double *y, const double *z, double *root, double *tmp)
{
    int nb = std::min(nblk(n, BT), 1024);
    hipLaunchKernelGGL(k_bbox_part, dim3(nb), dim3(BT), 0, s, n, x, y, z, tmp);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BT), 0, s, nb, tmp, root);
}

void tb_keys(hipStream_t s, int n, const double *x, const double *y, const double *z, const double *root,
             uint64_t *key, uint32_t *val, int32_t *err)
{
    hipLaunchKernelGGL(k_keys, dim3(nblk(n, BT)), dim3(BT), 0, s, n, x, y, z, root, key, val, err);
}

void tb_radix_sort(hipStream_t s, int n, uint64_t *key[2], uint32_t *val[2], int lo_bit, int hi_bit, void *tmp,
                   int *which, bool init_iota)
{
    int nb = nblk(n, RS_TILE);
    uint32_t *hist = (uint32_t *)tmp;
    uint32_t *totals = hist + (size_t)256 * nb;
    int src = 0;
    bool iota = init_iota;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        hipLaunchKernelGGL(k_rs_up<uint64_t>, dim3(nb), dim3(BT), 0, s, n, key[src], sh, hist, nb);
        hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(BT), 0, s, nb, hist, totals);
        hipLaunchKernelGGL(k_rs_down<uint64_t>, dim3(nb), dim3(BT), 0, s, n, key[src], val[src], key[1 - src], val[1 - src],
                           sh, hist, totals, nb, iota ? 1 : 0, (uint32_t *)nullptr);
        iota = false;
        src = 1 - src;
    }
    *which = src;
}

void tb_radix_sort2(const hipStream_t s[2], const int n[2], uint64_t **key[2], uint32_t **val[2], int lo_bit,
                    int hi_bit, void *const tmp[2], int which[2], bool init_iota)
{
    int nb[2], src[2] = {0, 0};
    for (int k = 0; k < 2; k++) nb[k] = nblk(n[k], RS_TILE);
    bool iota = init_iota;
    for (int sh = lo_bit; sh < hi_bit; sh += 8) {
        for (int k = 0; k < 2; k++) {
            uint32_t *hist = (uint32_t *)tmp[k];
            uint32_t *totals = hist + (size_t)256 * nb[k];
            uint64_t **kk = key[k];
            uint32_t **vv = val[k];
            hipLaunchKernelGGL(k_rs_up<uint64_t>, dim3(nb[k]), dim3(BT), 0, s[k], n[k], kk[src[k]], sh, hist, nb[k]);
            hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(BT), 0, s[k], nb[k], hist, totals);
            hipLaunchKernelGGL(k_rs_down<uint64_t>, dim3(nb[k]), dim3(BT), 0, s[k], n[k], kk[src[k]], vv[src[k]],
                               kk[1 - src[k]], vv[1 - src[k]], sh, hist, totals, nb[k], iota ? 1 : 0,
                               (uint32_t *)nullptr);
            src[k] = 1 - src[k];
        }
        iota = false;
    }
    for (int k = 0; k < 2; k++) which[k] = src[k];
}

void tb_depth_order(hipStream_t s, int n, const int32_t *depth, uint32_t *order, uint32_t *inv, void *tmp)
{
    // one stable 8-bit pass over the nodes' depths (< 2^8), node ids as
    // values; the scatter also writes the inverse order
    const int nb = nblk(n, RS_TILE);
    uint32_t *hist = (uint32_t *)tmp;
    uint32_t *totals = hist + (size_t)256 * nb;
    const uint32_t *key = (const uint32_t *)depth;
    hipLaunchKernelGGL(k_rs_up<uint32_t>, dim3(nb), dim3(BT), 0, s, n, key, 0, hist, nb);
    hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(BT), 0, s, nb, hist, totals);
    hipLaunchKernelGGL(k_rs_down<uint32_t>, dim3(nb), dim3(BT), 0, s, n, key, (const uint32_t *)nullptr,
                       (uint32_t *)nullptr, order, 0, hist, totals, nb, 1, inv);
}

void tb_permute(hipStream_t s, SpeciesDev &S, const uint32_t *perm, int32_t *err)
{
    const int c = S.cur, o = 1 - S.cur;
    PermCols cols;
    for (int k = 0; k < 3; k++) {
        cols.src[k] = S.x[c][k]; cols.src[3 + k] = S.v[c][k]; cols.src[6 + k] = S.f[c][k];
        cols.dst[k] = S.x[o][k]; cols.dst[3 + k] = S.v[o][k]; cols.dst[6 + k] = S.f[o][k];
    }
    hipLaunchKernelGGL(k_permute, dim3(nblk(S.n, BT)), dim3(BT), 0, s, S.n, perm, cols, (const int32_t *)S.orig[c],
                       S.orig[o], err);
    S.cur = o;
}

void tb_lcp(hipStream_t s, int n, const uint64_t *key, int32_t *cnt, int32_t *hist, int32_t *err, bool joint)
{
    hipLaunchKernelGGL(k_lcp, dim3(std::min(nblk(n, BT), 1024)), dim3(BT), 0, s, n, key, cnt, hist, err,
                       joint ? 1 : 0);
}

void tb_scan_excl(hipStream_t s, int n, const int32_t *in, int32_t *out, int32_t *total, void *tmp)
{
    int nb = nblk(n, SC_TILE);
    int32_t *bsum = (int32_t *)tmp;
    hipLaunchKernelGGL(k_scan_part, dim3(nb), dim3(BT), 0, s, n, in, bsum);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(BT), 0, s, nb, bsum, total);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(BT), 0, s, n, in, bsum, out);
}

void tb_emit(hipStream_t s, SpeciesDev &S, const uint64_t *key, int sign)
{
    const int c = S.cur;
    hipLaunchKernelGGL(k_emit, dim3(nblk(S.n, BT)), dim3(BT), 0, s, S.n, key, S.offs, S.nn, S.x[c][0], S.x[c][1],
                       S.x[c][2], (double)sign, S.nmac, S.nnext, S.nstart, S.ncount, S.ndepth);
}

void tb_upward_prep(hipStream_t s, SpeciesDev &S, const uint64_t *key, int sign, int32_t *err)
{
    hipLaunchKernelGGL(k_bfs_inv, dim3(nblk(S.nn, BT)), dim3(BT), 0, s, S.nn, S.bfs, S.binv);
    hipLaunchKernelGGL(k_bfs_prep, dim3(nblk(S.nn, BT)), dim3(BT), 0, s, S.nn, S.binv, S.nnext, S.ndepth, S.nstart,
                       key, S.nmac, (double)sign, S.bflag, S.bkey, S.bxq);
    hipLaunchKernelGGL(k_bfs_kids, dim3(nblk(S.nn, BT)), dim3(BT), 0, s, S.nn, S.bfs, S.binv, S.bflag, S.bkids, err);
}

void tb_upward_level(hipStream_t s, SpeciesDev &S, int d, double theta, int32_t *err)
{
    if (d < 0 || d > S.maxdepth) return;
    int beg = S.lvl_off[d], cnt = S.lvl_off[d + 1] - beg;
    if (cnt <= 0 || S.twig_hist[d] == 0) return;
    hipLaunchKernelGGL(k_upward, dim3(nblk(cnt, BT)), dim3(BT), 0, s, beg, cnt, S.lvl_off[d + 1], S.lvl_off[d + 2], d,
                       S.bfs, S.bkids, S.bflag, S.bkey, S.root, theta, S.nmac, S.nmom, S.ndx, S.bxq, err);
}


void tb_export_gather(hipStream_t s, const SpeciesDev &S, int sign, int32_t *depth, int32_t *leaf, int32_t *pt,
                      int32_t *count, int32_t *next, double *xx, double *qeff, double *dx, double *D, double *Q,
                      double *QQ)
{
    hipLaunchKernelGGL(k_export, dim3(nblk(S.nn, BT)), dim3(BT), 0, s, S.nn, (double)sign, S.nmac, S.nnext, S.nmom,
                       S.ndx,
                       S.nstart, S.ncount, S.ndepth, S.orig[S.cur], depth, leaf, pt, count, next, xx, qeff, dx, D, Q,
                       QQ);
}

void tb_jbbox(hipStream_t s, int ne, const double *ex, const double *ey, const double *ez, int ni, const double *ix,
              const double *iy, const double *iz, double *root, double *tmp)
{
    int nbe = std::min(nblk(ne, BT), 1024), nbi = std::min(nblk(ni, BT), 1024);
    hipLaunchKernelGGL(k_bbox_part, dim3(nbe), dim3(BT), 0, s, ne, ex, ey, ez, tmp);
    hipLaunchKernelGGL(k_bbox_part, dim3(nbi), dim3(BT), 0, s, ni, ix, iy, iz, tmp + 6 * nbe);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BT), 0, s, nbe + nbi, tmp, root);
}

void tb_jmerge(hipStream_t s, int ne, int ni, const uint64_t *ke, const uint64_t *ki, uint64_t *kout, uint32_t *vout)
{
    const int blocks = nblk(ne, BT) + nblk(ni, BT);
    hipLaunchKernelGGL(k_jmerge, dim3(blocks), dim3(BT), 0, s, ne, ni, ke, ki, kout, vout);
}

void tb_jgather(hipStream_t s, int n, const uint32_t *val, const double *const x[2], const double *const y[2],
                const double *const z[2], double4 *pos)
{
    JSrc src;
    for (int k = 0; k < 2; k++) {
        src.x[k] = x[k];
        src.y[k] = y[k];
        src.z[k] = z[k];
    }
    hipLaunchKernelGGL(k_jgather, dim3(nblk(n, BT)), dim3(BT), 0, s, n, val, src, pos);
}

void tb_groups(hipStream_t s, SpeciesDev &S, int G, int32_t *totals, void *scan_tmp, int T, int32_t *first)
{
    const int n = S.n, W = 8 * G;
    const int tagshift = 3 * (TB_LEVELS - T);
    const uint64_t tagmask = T > 0 ? ~0ull << tagshift : 0ull;
    hipLaunchKernelGGL(k_gwin, dim3(nblk(n, W)), dim3(BT), 0, s, n, W, S.skey, S.gpre, S.gsuf);
    hipLaunchKernelGGL(k_gflag, dim3(nblk(n, BT)), dim3(BT), 0, s, n, W, S.skey, S.gpre, S.gsuf, tagmask, S.cnt);
    tb_scan_excl(s, n, S.cnt, S.offs, totals, scan_tmp);
    hipLaunchKernelGGL(k_gseg, dim3(nblk(n, BT)), dim3(BT), 0, s, n, S.cnt, S.offs, S.gseg);
    hipLaunchKernelGGL(k_gcount, dim3(nblk(n, BT)), dim3(BT), 0, s, n, G, totals, S.gseg, S.cnt);
    tb_scan_excl(s, n, S.cnt, S.offs, totals + 1, scan_tmp);
    hipLaunchKernelGGL(k_gemit, dim3(nblk(n, BT)), dim3(BT), 0, s, n, totals, S.gseg, S.cnt, S.offs, S.grp, S.sgrp,
                       S.gsid);
    if (first)
        hipLaunchKernelGGL(k_gfirst, dim3(nblk(n, BT)), dim3(BT), 0, s, n, totals + 1, S.grp, S.skey,
                           T > 0 ? tagshift : 63, first);
    S.gG = G;
}

int tb_sbbox_chunks(int n_max) { return std::max(1, std::min(64, nblk(n_max, 4 * BT))); }

void tb_sbbox(hipStream_t s, int nreal, int chunks, const int32_t *off, const double *x, const double *y,
              const double *z, double *part)
{
    hipLaunchKernelGGL(k_sbbox_part, dim3(chunks, nreal), dim3(BT), 0, s, off, x, y, z, part);
}

void tb_sbbox_final(hipStream_t s, int nreal, int chunks, const double *pa, const double *pb, double *roots)
{
    hipLaunchKernelGGL(k_sbbox_final, dim3(nreal), dim3(64), 0, s, chunks, pa, pb, roots);
}

void tb_ekeys(hipStream_t s, int n, const double *x, const double *y, const double *z, const int32_t *off,
              int nreal, const double *roots, int T, uint64_t *key, uint32_t *val, uint32_t tag, int32_t *err)
{
    hipLaunchKernelGGL(k_ekeys, dim3(nblk(n, BT)), dim3(BT), 0, s, n, x, y, z, off, nreal, roots, T, key, val, tag,
                       err);
}

void tb_jemit(hipStream_t s, JointDev &J, const uint64_t *key)
{
    hipLaunchKernelGGL(k_jemit, dim3(nblk(J.n, BT)), dim3(BT), 0, s, J.n, key, J.offs, J.nn, J.pos, J.nc, J.nnext,
                       J.nstart, J.ncount, J.ndepth);
}

void tb_jupward(hipStream_t s, JointDev &J, const uint64_t *key, const CellRecs &R, int T, const double *roots)
{
    hipLaunchKernelGGL(k_jprep, dim3(nblk(J.nn, BT)), dim3(BT), 0, s, J.nn, J.nnext, J.ndepth, J.nstart, J.ncount,
                       key, J.nc, R);
    hipLaunchKernelGGL(k_jkids, dim3(nblk(J.nn, BT)), dim3(BT), 0, s, J.nn, J.bfs, R);
    for (int d = J.maxdepth; d >= 0; d--) {
        int beg = J.lvl_off[d], cnt = J.lvl_off[d + 1] - beg;
        if (cnt <= 0 || J.twig_hist[d] == 0) continue;
        const int cend = d + 2 <= TB_LEVELS + 2 ? J.lvl_off[d + 2] : J.nn;
        hipLaunchKernelGGL(k_jupward, dim3(nblk(cnt, JUP_CELLS)), dim3(BT), 0, s, beg, cnt, J.lvl_off[d + 1], cend, d, T,
                           T > 0 ? roots : J.root, R);
    }
}
