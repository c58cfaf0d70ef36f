// rdf32.hip -- the radial distribution histograms for the mixed walk.
//
// The reference accumulates RDF_ei / RDF_ii over the leaf pairs of every
// LocalCoulombSum (tree.f90:916-919, 1021-1024, 1125-1129): an unlike pair
// counts once from each side (the ie and ei walks), a like ion pair beyond the
// 0.1 um cut once from each ion, in bin INT(r / dx) + 1.  Which pairs are
// leaf pairs depends on the reference's per-target opening angle; the mixed
// walk's group MAC has no such set.  So for the mixed walk the histograms
// hold every pair within the histogram's range: the reference's quantity in
// its theta -> 0 limit, where every pair is a leaf pair (bit for bit: r is
// formed in FP64 as the FP64 walk forms it).
//
// One wavefront per ion target group (the walk's groups) searches the joint
// tree (breadth-first cell records) for cells within r_max = nbin dx of the
// group's bounding box; each particle leaf found is paired with every target
// of the group.  An electron leaf adds 2 to RDF_ei (the pair seen from both
// sides), an ion leaf beyond the cut 1 to RDF_ii (every ion is a target, so
// each ordered pair is seen).  Bins are 32-bit counters in LDS for the whole
// histogram (nbin <= TB_RDF32_MAXBIN), added to the global 64-bit histograms
// once per block.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>

#include "tb_mixed.h"

namespace {

constexpr int RB = 512;           // threads per block
constexpr int RW = RB / 64;       // waves per block
constexpr int RSTACK = 512;       // child ranges per wave
constexpr int RTPL = 4;           // targets per lane: groups of up to 256

struct RdfArgs {
    const double4 *bmac;          // cell records: centre (a leaf: its particle)
    const uint32_t *binfo;        // leaf | particle, or first child << 3 | nchild - 1
    const float4 *bq0;            // a leaf's charge in x
    const float4 *bq2;            // a cell's size max(lx) in y
    int nn;
    const int2 *grp;              // ion target groups (start, length)
    int ng;
    const double *px, *py, *pz;   // ion positions (key order)
    double dx_ei, dx_ii;
    int nbin;
    double rmax2;                 // square of the search radius
    unsigned long long *hist;     // [2][nbin]
    int32_t *flag;                // 1: stack overflow, 2: malformed table
};

struct RdfWave {
    uint32_t stack[RSTACK];
    uint32_t batch[64];
    double4 leaf[64];
};

__device__ inline int wave_incl_scan(int v)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (l >= o) v += t;
    }
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

__device__ inline uint64_t lanemask_lt()
{
    const int l = threadIdx.x & 63;
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ inline void sync_wave()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(RB) void k_rdf_pairs(RdfArgs A)
{
    __shared__ uint32_t hl[2 * TB_RDF32_MAXBIN];
    __shared__ RdfWave waves[RW];
    for (int k = threadIdx.x; k < 2 * A.nbin; k += RB) hl[k] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    RdfWave &L = waves[threadIdx.x >> 6];
    const int nwaves = gridDim.x * RW;
    for (int g = blockIdx.x * RW + (threadIdx.x >> 6); g < A.ng; g += nwaves) {
        const int2 gr = A.grp[g];
        const int base = __builtin_amdgcn_readfirstlane(gr.x), nact = __builtin_amdgcn_readfirstlane(gr.y);
        double tx[RTPL], ty[RTPL], tz[RTPL];
        bool act[RTPL];
        double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
#pragma unroll
        for (int t = 0; t < RTPL; t++) {
            act[t] = lane + 64 * t < nact;
            const int i = act[t] ? base + lane + 64 * t : base;
            tx[t] = A.px[i];
            ty[t] = A.py[i];
            tz[t] = A.pz[i];
            mn[0] = fmin(mn[0], tx[t]); mx[0] = fmax(mx[0], tx[t]);
            mn[1] = fmin(mn[1], ty[t]); mx[1] = fmax(mx[1], ty[t]);
            mn[2] = fmin(mn[2], tz[t]); mx[2] = fmax(mx[2], tz[t]);
        }
        for (int k = 0; k < 3; k++) {
            mn[k] = wave_min(mn[k]);
            mx[k] = wave_max(mx[k]);
        }
        int sp = 1;
        L.stack[0] = 0u;  // the root: a range of one cell
        sync_wave();
        while (sp > 0) {
            // up to 64 cells from the top ranges of the stack
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
            sync_wave();
            const bool have = lane < total;
            const uint32_t node = have ? L.batch[lane] : 0u;
            sync_wave();
            bool leaf = false, open = false;
            double4 m = make_double4(0.0, 0.0, 0.0, 0.0);
            uint32_t info = 0;
            if (have) {
                m = A.bmac[node];
                info = A.binfo[node];
                leaf = (info & TB_LEAF) != 0;
                // distance from the cell (a leaf: its particle) to the group's box
                const double h = leaf ? 0.0 : 0.5 * (double)A.bq2[node].y;
                const double c[3] = {m.x, m.y, m.z};
                double d2 = 0.0;
                for (int k = 0; k < 3; k++) {
                    const double dd = fmax(fmax(mn[k] - (c[k] + h), (c[k] - h) - mx[k]), 0.0);
                    d2 += dd * dd;
                }
                const bool near = d2 <= A.rmax2;
                if (!near) leaf = false;
                else if (!leaf) open = true;
            }
            if (open) {
                const uint32_t st = info >> 3, n = (info & 7u) + 1u;
                if (st <= node || st + n > (uint32_t)A.nn) {
                    atomicOr(A.flag, 2);
                    open = false;
                }
            }
            const uint64_t om = __ballot(open);
            const int npush = __popcll(om);
            if (sp + npush > RSTACK) {
                atomicOr(A.flag, 1);
                break;
            }
            if (open) L.stack[sp + __popcll(om & lanemask_lt())] = info;
            sp += npush;
            // the near leaves, paired with every target of the group
            const uint64_t lm = __ballot(leaf);
            if (leaf) L.leaf[__popcll(lm & lanemask_lt())] = make_double4(m.x, m.y, m.z, (double)A.bq0[node].x);
            sync_wave();
            const int nl = __popcll(lm);
            for (int j = 0; j < nl; j++) {
                const double4 p = L.leaf[j];
                const bool electron = p.w < 0.0;
#pragma unroll
                for (int t = 0; t < RTPL; t++) {
                    if (!act[t]) continue;
                    // as the FP64 walk forms it: target - leaf, r2 = x x + y y + z z
                    const double x = tx[t] - p.x, y = ty[t] - p.y, z = tz[t] - p.z;
                    const double r2 = x * x + y * y + z * z;
                    if (electron) {
                        const int id = (int)(sqrt(r2) / A.dx_ei) + 1;
                        if (id > 0 && id <= A.nbin) atomicAdd(&hl[id - 1], 2u);
                    } else if (r2 > 0.01) {
                        const int id = (int)(sqrt(r2) / A.dx_ii) + 1;
                        if (id > 0 && id <= A.nbin) atomicAdd(&hl[A.nbin + id - 1], 1u);
                    }
                }
            }
            sync_wave();
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * A.nbin; k += RB)
        if (hl[k]) atomicAdd(&A.hist[k], (unsigned long long)hl[k]);
}

}  // namespace

int tb_rdf_pairs(hipStream_t s, const MixedTree &T, const SpeciesDev &I, const TbRdf &R, int32_t *flag)
{
    if (R.nbin < 1 || R.nbin > TB_RDF32_MAXBIN || I.gG > 64 * RTPL) return TB_ERR_ARG;
    RdfArgs A;
    A.bmac = T.bmac;
    A.binfo = T.binfo;
    A.bq0 = T.bq0;
    A.bq2 = T.bq2;
    A.nn = T.nn;
    A.grp = I.grp;
    A.ng = I.ng;
    const int c = I.cur;
    A.px = I.x[c][0];
    A.py = I.x[c][1];
    A.pz = I.x[c][2];
    A.dx_ei = R.dx[0];
    A.dx_ii = R.dx[1];
    A.nbin = R.nbin;
    // every pair with r / dx < nbin lies within nbin max(dx) (with a margin for rounding)
    const double rmax = (double)R.nbin * (R.dx[0] > R.dx[1] ? R.dx[0] : R.dx[1]) * (1.0 + 1e-9);
    A.rmax2 = rmax * rmax;
    A.hist = (unsigned long long *)R.hist;
    A.flag = flag;
    int dev = 0, cus = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int blocks = std::max(1, std::min((I.ng + RW - 1) / RW, (cus > 0 ? cus : 256)));
    hipLaunchKernelGGL(k_rdf_pairs, dim3(blocks), dim3(RB), 0, s, A);
    return hipGetLastError() == hipSuccess ? TB_OK : TB_ERR_HIP;
}
