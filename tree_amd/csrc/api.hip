// api.hip -- the C-ABI of libtreemi.so (declared in include/treemi.h):
// handle lifetime, device memory, and the order of launches that stands in
// for the reference's Force (main.f90:506-541), BuildLocalTree
// (tree.f90:39-100), LocalCoulombSum (tree.f90:662-717) and
// VVerletNotemp1/2 (vverlet.f90:9-112).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tb_internal.h"
#include "tb_mixed.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(TB_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// ---------------------------------------------- allocation-free direct sum
//
// tb_direct_sample_noalloc stages its target list and forces in `scratch`, a
// fixed device array (one copy per GPU, held until the process ends), instead
// of device memory allocated per call.  Host code never names `scratch` as
// an address: its device address is looked up once per device, with that
// device current (scratch_ptr), and only that address reaches kernels and
// copies.  Layout: tspecies, reference indices and key-order indices (int32,
// TB_NOALLOC_MAX_TARGETS each), then forces (3 doubles per target).
constexpr size_t NOALLOC_INT_BYTES = (size_t)TB_NOALLOC_MAX_TARGETS * 4;
constexpr size_t NOALLOC_BYTES = 3 * NOALLOC_INT_BYTES + (size_t)TB_NOALLOC_MAX_TARGETS * 24;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

extern "C" void tb_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);  // gen.cpp

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// device address of this device's `scratch` (the device must be current);
// nullptr if it cannot be looked up
void *scratch_ptr(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// reference index -> key-order index of each listed target, through the
// species' inverse permutations
__global__ __launch_bounds__(256) void k_map_targets(int nt, const int32_t *__restrict__ tsp,
                                                     const int32_t *__restrict__ tref,
                                                     const int32_t *__restrict__ inv_e,
                                                     const int32_t *__restrict__ inv_i, int32_t *__restrict__ tout)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < nt) tout[t] = tsp[t] == 0 ? inv_e[tref[t]] : inv_i[tref[t]];
}

}  // namespace

struct tb_handle {
    tb_config cfg{};
    hipStream_t s = nullptr;
    // the ions' build work runs on s1 beside the electrons' on s (joined
    // back with fork / join events); each stream has its own temporaries
    hipStream_t s1 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    SpeciesDev sp[2];
    JointDev jt;                   // mixed walk: one tree over both species
    MixedDev mx;
    void *sort_tmp = nullptr, *sort_tmp1 = nullptr;
    size_t sort_tmp_bytes = 0;
    void *scan_tmp = nullptr, *scan_tmp1 = nullptr;
    size_t scan_tmp_bytes = 0;
    double *bbox_tmp = nullptr, *bbox_tmp1 = nullptr;
    int32_t *d_misc = nullptr;     // [0..1] err per species, [2..3] totals, [4..] hist 2 x 2 x (L+2)
    int32_t *h_misc = nullptr;     // pinned mirror
    double *epart = nullptr;       // per-block partial sums
    size_t epart_n = 0;
    double *d_sums = nullptr;      // 16 doubles
    double *h_sums = nullptr;      // pinned
    int64_t *d_cnt = nullptr;      // 16 counters
    size_t bytes = 0;              // device bytes held
    bool built = false, have_force = false;
    int tree_walk = -1;            // which walk the current trees were built for
    double epot[2] = {0, 0};
    tb_stats st{};
    hipEvent_t ev[6] = {};
    TbRdf rdf;                     // leaf-pair RDF histograms (FP64 walk), when enabled
    bool rdf_on = false;
    std::vector<void *> allocs;
};

namespace {

constexpr int MISC_HIST = 4;
constexpr int HW = TB_LEVELS + 2;  // histogram width

int dalloc(tb_handle *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (h->cfg.mem_cap > 0 && (int64_t)(h->bytes + bytes) > h->cfg.mem_cap)
        return fail(TB_ERR_NOMEM, "device allocation of %zu bytes exceeds the handle cap %lld", bytes,
                    (long long)h->cfg.mem_cap);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(TB_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    h->bytes += bytes;
    return TB_OK;
}

void dfree(tb_handle *h, void *p, size_t bytes)
{
    if (!p) return;
    hipFree(p);
    h->bytes -= std::min(h->bytes, bytes == 0 ? (size_t)16 : bytes);
}

template <typename T>
int dalloc_t(tb_handle *h, T **p, size_t count)
{
    return dalloc(h, (void **)p, count * sizeof(T));
}

void free_species(tb_handle *h, SpeciesDev &S)
{
    size_t n = (size_t)S.n;
    for (int b = 0; b < 2; b++) {
        for (int k = 0; k < 3; k++) {
            dfree(h, S.x[b][k], n * 8); dfree(h, S.v[b][k], n * 8); dfree(h, S.f[b][k], n * 8);
        }
        dfree(h, S.orig[b], n * 4);
        dfree(h, S.key[b], n * 8);
        dfree(h, S.val[b], n * 4);
    }
    dfree(h, S.cnt, n * 4);
    dfree(h, S.offs, n * 4);
    dfree(h, S.root, 64 * 8);
    dfree(h, S.gpre, n);
    dfree(h, S.gsuf, n);
    dfree(h, S.gseg, n * 4);
    dfree(h, S.sgrp, n * 8);
    dfree(h, S.gsid, n * 4);
    dfree(h, S.grp, n * 8);
    size_t c = (size_t)S.cap;
    dfree(h, S.nmac, c * 32); dfree(h, S.nnext, c * 4); dfree(h, S.nmom, c * 80);
    dfree(h, S.ndx, c * 8); dfree(h, S.nstart, c * 4); dfree(h, S.ncount, c * 4); dfree(h, S.ndepth, c * 4);
    dfree(h, S.binv, c * 4); dfree(h, S.bflag, c); dfree(h, S.bkids, c * 4); dfree(h, S.bxq, c * 32);
    dfree(h, S.bkey, c * 8);
    for (int b = 0; b < 2; b++) dfree(h, S.dval[b], c * 4);
    S = SpeciesDev();
}

int alloc_species(tb_handle *h, SpeciesDev &S, int n)
{
    S.n = n;
    int rc;
    for (int b = 0; b < 2; b++) {
        for (int k = 0; k < 3; k++) {
            if ((rc = dalloc_t(h, &S.x[b][k], n))) return rc;
            if ((rc = dalloc_t(h, &S.v[b][k], n))) return rc;
            if ((rc = dalloc_t(h, &S.f[b][k], n))) return rc;
        }
        if ((rc = dalloc_t(h, &S.orig[b], n))) return rc;
        if ((rc = dalloc_t(h, &S.key[b], n))) return rc;
        if ((rc = dalloc_t(h, &S.val[b], n))) return rc;
    }
    if ((rc = dalloc_t(h, &S.cnt, n))) return rc;
    if ((rc = dalloc_t(h, &S.offs, n))) return rc;
    if ((rc = dalloc_t(h, &S.root, 64))) return rc;
    if ((rc = dalloc_t(h, &S.gpre, n))) return rc;
    if ((rc = dalloc_t(h, &S.gsuf, n))) return rc;
    if ((rc = dalloc_t(h, &S.gseg, n))) return rc;
    if ((rc = dalloc_t(h, &S.sgrp, n))) return rc;
    if ((rc = dalloc_t(h, &S.gsid, n))) return rc;
    if ((rc = dalloc_t(h, &S.grp, n))) return rc;
    return TB_OK;
}

int ensure_nodes(tb_handle *h, SpeciesDev &S, int nn)
{
    if (nn <= S.cap) return TB_OK;
    size_t c = (size_t)S.cap;
    dfree(h, S.nmac, c * 32); dfree(h, S.nnext, c * 4); dfree(h, S.nmom, c * 80);
    dfree(h, S.ndx, c * 8); dfree(h, S.nstart, c * 4); dfree(h, S.ncount, c * 4); dfree(h, S.ndepth, c * 4);
    dfree(h, S.binv, c * 4); dfree(h, S.bflag, c); dfree(h, S.bkids, c * 4); dfree(h, S.bxq, c * 32);
    dfree(h, S.bkey, c * 8);
    for (int b = 0; b < 2; b++) dfree(h, S.dval[b], c * 4);
    int cap = (int)std::min<long>((long)nn + nn / 8 + 64, (long)INT32_MAX / 2);
    S.cap = 0;
    int rc;
    if ((rc = dalloc_t(h, &S.nmac, cap))) return rc;
    if ((rc = dalloc_t(h, &S.nnext, cap))) return rc;
    if ((rc = dalloc_t(h, &S.nmom, (size_t)cap * 10))) return rc;
    if ((rc = dalloc_t(h, &S.ndx, cap))) return rc;
    if ((rc = dalloc_t(h, &S.nstart, cap))) return rc;
    if ((rc = dalloc_t(h, &S.ncount, cap))) return rc;
    if ((rc = dalloc_t(h, &S.ndepth, cap))) return rc;
    if ((rc = dalloc_t(h, &S.binv, cap))) return rc;
    if ((rc = dalloc_t(h, &S.bflag, cap))) return rc;
    if ((rc = dalloc_t(h, &S.bkids, cap))) return rc;
    if ((rc = dalloc_t(h, &S.bkey, cap))) return rc;
    if ((rc = dalloc_t(h, &S.bxq, cap))) return rc;
    for (int b = 0; b < 2; b++) {
        if ((rc = dalloc_t(h, &S.dval[b], cap))) return rc;
    }
    S.cap = cap;
    return TB_OK;
}

void free_joint(tb_handle *h, JointDev &J)
{
    size_t n = (size_t)J.n, c = (size_t)J.cap;
    for (int b = 0; b < 2; b++) {
        dfree(h, J.key[b], n * 8);
        dfree(h, J.val[b], n * 4);
        dfree(h, J.dval[b], c * 4);
    }
    dfree(h, J.pos, n * 32);
    dfree(h, J.cnt, n * 4);
    dfree(h, J.offs, n * 4);
    dfree(h, J.root, 64 * 8);
    dfree(h, J.nc, c * 32); dfree(h, J.nnext, c * 4);
    dfree(h, J.nstart, c * 4); dfree(h, J.ncount, c * 4); dfree(h, J.ndepth, c * 4);
    J = JointDev();
}

int ensure_joint(tb_handle *h, int n)
{
    JointDev &J = h->jt;
    if (J.n == n && J.key[0]) return TB_OK;
    free_joint(h, J);
    J.n = n;
    int rc;
    for (int b = 0; b < 2; b++) {
        if ((rc = dalloc_t(h, &J.key[b], n))) return rc;
        if ((rc = dalloc_t(h, &J.val[b], n))) return rc;
    }
    if ((rc = dalloc_t(h, &J.pos, n))) return rc;
    if ((rc = dalloc_t(h, &J.cnt, n))) return rc;
    if ((rc = dalloc_t(h, &J.offs, n))) return rc;
    if ((rc = dalloc_t(h, &J.root, 64))) return rc;
    return TB_OK;
}

int ensure_joint_nodes(tb_handle *h, int nn)
{
    JointDev &J = h->jt;
    if (nn <= J.cap) return TB_OK;
    size_t c = (size_t)J.cap;
    for (int b = 0; b < 2; b++) dfree(h, J.dval[b], c * 4);
    dfree(h, J.nc, c * 32); dfree(h, J.nnext, c * 4);
    dfree(h, J.nstart, c * 4); dfree(h, J.ncount, c * 4); dfree(h, J.ndepth, c * 4);
    int cap = (int)std::min<long>((long)nn + nn / 8 + 64, (long)INT32_MAX / 2);
    J.cap = 0;
    int rc;
    if ((rc = dalloc_t(h, &J.nc, cap))) return rc;
    if ((rc = dalloc_t(h, &J.nnext, cap))) return rc;
    if ((rc = dalloc_t(h, &J.nstart, cap))) return rc;
    if ((rc = dalloc_t(h, &J.ncount, cap))) return rc;
    if ((rc = dalloc_t(h, &J.ndepth, cap))) return rc;
    for (int b = 0; b < 2; b++) {
        if ((rc = dalloc_t(h, &J.dval[b], cap))) return rc;
    }
    J.cap = cap;
    return TB_OK;
}

int ensure_tmp(tb_handle *h, int n)
{
    int rc;
    size_t sb = tb_sort_temp_bytes(n);
    if (sb > h->sort_tmp_bytes) {
        dfree(h, h->sort_tmp, h->sort_tmp_bytes);
        dfree(h, h->sort_tmp1, h->sort_tmp_bytes);
        h->sort_tmp = h->sort_tmp1 = nullptr; h->sort_tmp_bytes = 0;
        if ((rc = dalloc(h, &h->sort_tmp, sb)) || (rc = dalloc(h, &h->sort_tmp1, sb))) return rc;
        h->sort_tmp_bytes = sb;
    }
    size_t cb = tb_scan_temp_bytes(n);
    if (cb > h->scan_tmp_bytes) {
        dfree(h, h->scan_tmp, h->scan_tmp_bytes);
        dfree(h, h->scan_tmp1, h->scan_tmp_bytes);
        h->scan_tmp = h->scan_tmp1 = nullptr; h->scan_tmp_bytes = 0;
        if ((rc = dalloc(h, &h->scan_tmp, cb)) || (rc = dalloc(h, &h->scan_tmp1, cb))) return rc;
        h->scan_tmp_bytes = cb;
    }
    size_t pe = (size_t)std::max(tb_walk64_blocks(n), tb_vv_blocks(n)) * 4 + 4096;
    if (pe > h->epart_n) {
        dfree(h, h->epart, h->epart_n * 8);
        h->epart = nullptr; h->epart_n = 0;
        if ((rc = dalloc_t(h, &h->epart, pe))) return rc;
        h->epart_n = pe;
    }
    return TB_OK;
}

int check_cfg(const tb_config *c)
{
    if (!c) return fail(TB_ERR_ARG, "null config");
    if (!(c->theta > 0.0)) return fail(TB_ERR_ARG, "theta must be > 0 (got %g)", c->theta);
    if (c->theta_mixed < 0.0) return fail(TB_ERR_ARG, "theta_mixed must be >= 0");
    if (!(c->rs > 0.0)) return fail(TB_ERR_ARG, "rs must be > 0 (got %g)", c->rs);
    if (!(c->xm[0] > 0.0) || !(c->xm[1] > 0.0)) return fail(TB_ERR_ARG, "masses must be > 0");
    if (c->walk != TB_WALK_REF64 && c->walk != TB_WALK_MIXED) return fail(TB_ERR_ARG, "unknown walk %d", c->walk);
    if (c->mac != TB_MAC_FIXED && c->mac != TB_MAC_LEARNED) return fail(TB_ERR_ARG, "unknown mac %d", c->mac);
    if (c->group != 0 && c->group != 64 && c->group != 128 && c->group != 256)
        return fail(TB_ERR_ARG, "group must be 0, 64, 128 or 256 (got %d)", c->group);
    if (!std::isfinite(c->far_ratio) || (c->far_ratio > 0.0 && c->far_ratio < 2.0))
        return fail(TB_ERR_ARG, "far_ratio must be 0 (default), negative (off) or >= 2 (got %g)", c->far_ratio);
    if (c->far_levels < 0 || c->far_levels > 2)
        return fail(TB_ERR_ARG, "far_levels must be 0 or 1 (default, group level) or 2 (got %d)", c->far_levels);
    if (!(c->cost_weight >= 0.0f && c->cost_weight <= 4.0f))
        return fail(TB_ERR_ARG, "cost_weight must be in [0, 4] (got %g)", (double)c->cost_weight);
    return TB_OK;
}

// BuildLocalTree for both species, in two phases around one small host
// read-back (node totals, depth histograms, error flags).  An ensemble runs
// phase A for every realisation, synchronises once, then runs phase B.
constexpr int MISC_J = MISC_HIST + 2 * 2 * HW;  // joint tree: err, total, 2 x HW histograms
constexpr int MISC_G = MISC_J + 2 + 2 * HW;      // per species: segments, target groups
constexpr int MISC_N = MISC_G + 4;               // ints read back after phase A
constexpr int MISC_RDF = 240;                    // the mixed-walk RDF's error flag (not read back)
static_assert(MISC_N <= MISC_RDF, "d_misc layout");

// A batched ensemble builds all its realisations as one combined state:
// realisation r holds slice [off[k][r], off[k][r+1]) of species k's arrays,
// keys carry r as their first T octal digits (tb_ekeys), each realisation's
// boxes are its own (d_roots: electrons, ions, joint, [nreal][6] each), and
// its cells form the subtree at depth T of one joint table.
struct EnsBuild {
    int nreal = 0, T = 0, chunks = 1;
    int32_t *d_off[2] = {nullptr, nullptr};
    double *d_roots = nullptr;   // [3][nreal][6]
    double *d_part = nullptr;    // [2][nreal][chunks][6] box partials
    int32_t *d_gfirst = nullptr; // [2][nreal] first target group of each realisation
};

// Phase A.  tpl = walk-group size / 64 for the mixed walk's target groups
// (0: chosen from the configuration and the autotuner's table).
int build_phase_a(tb_handle *h, int *which, int tpl = 0, const EnsBuild *eb = nullptr)
{
    hipStream_t s = h->s;
    int rc;
    const bool mixed = h->cfg.walk == TB_WALK_MIXED;
    const int nj = h->sp[0].n + h->sp[1].n;
    if ((rc = ensure_tmp(h, mixed ? nj : std::max(h->sp[0].n, h->sp[1].n)))) return rc;
    if (mixed && (rc = ensure_joint(h, nj))) return rc;
    HIPCHK(hipMemsetAsync(h->d_misc, 0, sizeof(int32_t) * MISC_N, s));
    const size_t pstride = eb ? (size_t)eb->nreal * eb->chunks * 6 : 0;
    // The mixed walk keys both species over the joint box (each
    // realisation's, in an ensemble), so that the joint order is a merge of
    // the two species' sorted keys.
    double *jroots = eb ? eb->d_roots + (size_t)2 * eb->nreal * 6 : nullptr;
    if (mixed) {
        SpeciesDev &E = h->sp[0], &I = h->sp[1];
        if (eb) {
            for (int sp = 0; sp < 2; sp++) {
                SpeciesDev &S = h->sp[sp];
                tb_sbbox(s, eb->nreal, eb->chunks, eb->d_off[sp], S.x[S.cur][0], S.x[S.cur][1], S.x[S.cur][2],
                         eb->d_part + sp * pstride);
            }
            tb_sbbox_final(s, eb->nreal, eb->chunks, eb->d_part, eb->d_part + pstride, jroots);
        } else {
            tb_jbbox(s, E.n, E.x[E.cur][0], E.x[E.cur][1], E.x[E.cur][2], I.n, I.x[I.cur][0], I.x[I.cur][1],
                     I.x[I.cur][2], h->jt.root, h->bbox_tmp);
        }
    }
    // the two species' keys, sorts and groups are independent: ions on s1,
    // each stage enqueued for both species before the next, so that both
    // streams have work from the start
    HIPCHK(hipEventRecord(h->fork, s));
    HIPCHK(hipStreamWaitEvent(h->s1, h->fork, 0));
    const hipStream_t st2[2] = {s, h->s1};
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        const int c = S.cur;
        hipStream_t ss = st2[sp];
        if (eb) {
            tb_ekeys(ss, S.n, S.x[c][0], S.x[c][1], S.x[c][2], eb->d_off[sp], eb->nreal, jroots, eb->T, S.key[0],
                     S.val[0], 0u, h->d_misc + sp);
        } else if (mixed) {
            tb_keys(ss, S.n, S.x[c][0], S.x[c][1], S.x[c][2], h->jt.root, S.key[0], S.val[0], h->d_misc + sp);
        } else {
            tb_bbox(ss, S.n, S.x[c][0], S.x[c][1], S.x[c][2], S.root, sp ? h->bbox_tmp1 : h->bbox_tmp);
            tb_keys(ss, S.n, S.x[c][0], S.x[c][1], S.x[c][2], S.root, S.key[0], S.val[0], h->d_misc + sp);
        }
    }
    {
        const int n2[2] = {h->sp[0].n, h->sp[1].n};
        uint64_t **k2[2] = {h->sp[0].key, h->sp[1].key};
        uint32_t **v2[2] = {h->sp[0].val, h->sp[1].val};
        void *const t2[2] = {h->sort_tmp, h->sort_tmp1};
        tb_radix_sort2(st2, n2, k2, v2, 0, 64, t2, which, false);
    }
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        hipStream_t ss = st2[sp];
        void *scan_t = sp ? h->scan_tmp1 : h->scan_tmp;
        tb_permute(ss, S, S.val[which[sp]], h->d_misc + sp);
        S.skey = S.key[which[sp]];
        if (mixed)
            tb_groups(ss, S, 64 * (tpl ? tpl : tb_mixed_tpl(h->cfg, eb ? nj / eb->nreal : nj)),
                      h->d_misc + MISC_G + 2 * sp, scan_t, eb ? eb->T : 0,
                      eb ? eb->d_gfirst + (size_t)sp * eb->nreal : nullptr);
        if (!mixed) {
            tb_lcp(ss, S.n, S.key[which[sp]], S.cnt, h->d_misc + MISC_HIST + sp * 2 * HW, h->d_misc + sp);
            tb_scan_excl(ss, S.n, S.cnt, S.offs, h->d_misc + 2 + sp, scan_t);
        }
    }
    HIPCHK(hipEventRecord(h->join, h->s1));
    HIPCHK(hipStreamWaitEvent(s, h->join, 0));
    if (mixed) {
        // joint tree over both species, on the freshly permuted positions:
        // its key order merged from the species' sorted keys
        JointDev &J = h->jt;
        SpeciesDev &E = h->sp[0], &I = h->sp[1];
        const int ce = E.cur, ci = I.cur;
        tb_jmerge(s, E.n, I.n, E.key[which[0]], I.key[which[1]], J.key[0], J.val[0]);
        which[2] = 0;
        const double *px[2] = {E.x[ce][0], I.x[ci][0]}, *py[2] = {E.x[ce][1], I.x[ci][1]},
                     *pz[2] = {E.x[ce][2], I.x[ci][2]};
        tb_jgather(s, nj, J.val[which[2]], px, py, pz, J.pos);
        tb_lcp(s, nj, J.key[which[2]], J.cnt, h->d_misc + MISC_J + 2, h->d_misc + MISC_J, true);
        tb_scan_excl(s, nj, J.cnt, J.offs, h->d_misc + MISC_J + 1, h->scan_tmp);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->h_misc, h->d_misc, sizeof(int32_t) * MISC_N, hipMemcpyDeviceToHost, s));
    return TB_OK;
}

// after the stream has been synchronised
int build_phase_b(tb_handle *h, const int *which, const EnsBuild *eb = nullptr)
{
    hipStream_t s = h->s;
    int rc;
    const bool mixed = h->cfg.walk == TB_WALK_MIXED;
    for (int sp = 0; sp < 3; sp++) {
        int e = sp < 2 ? h->h_misc[sp] : (mixed ? h->h_misc[MISC_J] : 0);
        if (e & 4) return fail(TB_ERR_HIP, "species %d: internal sort check failed", sp);
        if (e & 2) return fail(TB_ERR_NONFINITE, "non-finite position in species %d", sp % 2);
        if (e & 1) {
            if (sp == 2)
                return fail(TB_ERR_DEPTH, "joint tree: more than 8 particles share all %d key levels", TB_LEVELS);
            return fail(TB_ERR_DEPTH, "%s: two particles share all %d key levels (coincident)",
                        sp == 0 ? "electrons" : "ions", TB_LEVELS);
        }
    }
    if (mixed) {
        for (int sp = 0; sp < 2; sp++) {
            h->sp[sp].nseg = h->h_misc[MISC_G + 2 * sp];
            h->sp[sp].ng = h->h_misc[MISC_G + 2 * sp + 1];
        }
        JointDev &J = h->jt;
        const int32_t *hist = h->h_misc + MISC_J + 2;
        J.nn = 1 + h->h_misc[MISC_J + 1];
        if ((rc = ensure_joint_nodes(h, J.nn))) return rc;
        if ((rc = ensure_tmp(h, std::max(J.nn, J.n)))) return rc;
        int all[HW + 1];
        for (int d = 0; d < HW; d++) {
            J.twig_hist[d] = hist[d] + (d == 0 ? 1 : 0);
            all[d] = hist[d] + hist[HW + d] + (d == 0 ? 1 : 0);
        }
        J.lvl_off[0] = 0;
        J.maxdepth = 0;
        for (int d = 0; d < HW; d++) {
            J.lvl_off[d + 1] = J.lvl_off[d] + all[d];
            if (all[d]) J.maxdepth = d;
        }
        const uint64_t *key = J.key[which[2]];
        tb_jemit(s, J, key);
        if (eb && J.lvl_off[eb->T + 1] - J.lvl_off[eb->T] != eb->nreal)
            // realisation r's root cell is the r-th cell at depth T (the tag digits)
            return fail(TB_ERR_HIP, "ensemble tree: %d cells at depth %d for %d realisations",
                        J.lvl_off[eb->T + 1] - J.lvl_off[eb->T], eb->T, eb->nreal);
        // the walk's breadth-first records, filled by the upward pass
        if ((rc = tb_mixed_prep(h->s, h->mx, &h->jt, &h->bytes))) return fail(rc, "%s", tb_mixed_error());
        const CellRecs R = tb_mixed_recs(h->mx, h->cfg);
        tb_depth_order(s, J.nn, J.ndepth, J.dval[0], R.inv, h->sort_tmp);
        J.bfs = J.dval[0];
        if (eb) tb_jupward(s, J, key, R, eb->T, eb->d_roots + (size_t)2 * eb->nreal * 6);
        else tb_jupward(s, J, key, R);
        h->st.nodes[0] = J.nn;
        h->st.nodes[1] = 0;
        h->st.max_depth[0] = J.maxdepth;
        h->st.max_depth[1] = 0;
        HIPCHK(hipGetLastError());
        h->mx.lists = tb_autotune_lists(eb ? (h->sp[0].n + h->sp[1].n) / eb->nreal : h->sp[0].n + h->sp[1].n);
        h->built = true;
        h->tree_walk = TB_WALK_MIXED;
        return TB_OK;
    }
    // all (re)allocation happens here, while the stream is idle
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        S.nn = 1 + h->h_misc[2 + sp];
        if ((rc = ensure_nodes(h, S, S.nn))) return rc;
    }
    if ((rc = ensure_tmp(h, std::max(h->sp[0].nn, h->sp[1].nn)))) return rc;
    // each species' linking, depth sort and upward pass: ions on s1, stage
    // by stage for both species
    HIPCHK(hipEventRecord(h->fork, s));
    HIPCHK(hipStreamWaitEvent(h->s1, h->fork, 0));
    const hipStream_t st2[2] = {s, h->s1};
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        hipStream_t ss = st2[sp];
        const int32_t *hist = h->h_misc + MISC_HIST + sp * 2 * HW;
        int all[HW + 1];
        for (int d = 0; d < HW; d++) {
            S.twig_hist[d] = hist[d] + (d == 0 ? 1 : 0);
            all[d] = hist[d] + hist[HW + d] + (d == 0 ? 1 : 0);
        }
        S.lvl_off[0] = 0;
        S.maxdepth = 0;
        for (int d = 0; d < HW; d++) {
            S.lvl_off[d + 1] = S.lvl_off[d] + all[d];
            if (all[d]) S.maxdepth = d;
        }
        tb_emit(ss, S, S.key[which[sp]], sp == 0 ? -1 : 1);
        h->st.nodes[sp] = S.nn;
        h->st.max_depth[sp] = S.maxdepth;
    }
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        tb_depth_order(st2[sp], S.nn, S.ndepth, S.dval[0], nullptr, sp ? h->sort_tmp1 : h->sort_tmp);
        S.bfs = S.dval[0];
    }
    for (int sp = 0; sp < 2; sp++)
        tb_upward_prep(st2[sp], h->sp[sp], h->sp[sp].key[which[sp]], sp == 0 ? -1 : 1, h->d_misc + sp);
    for (int d = std::max(h->sp[0].maxdepth, h->sp[1].maxdepth); d >= 0; d--)
        for (int sp = 0; sp < 2; sp++) tb_upward_level(st2[sp], h->sp[sp], d, h->cfg.theta, h->d_misc + sp);
    HIPCHK(hipEventRecord(h->join, h->s1));
    HIPCHK(hipStreamWaitEvent(s, h->join, 0));
    HIPCHK(hipGetLastError());
    h->built = true;
    h->tree_walk = TB_WALK_REF64;
    return TB_OK;
}

// Cuts both species' target groups again, for walk-group size 64 * tpl, from
// the sorted keys the last build kept; synchronises to read the counts.
int regroup(tb_handle *h, int tpl)
{
    for (int sp = 0; sp < 2; sp++)
        tb_groups(h->s, h->sp[sp], 64 * tpl, h->d_misc + MISC_G + 2 * sp, h->scan_tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->h_misc + MISC_G, h->d_misc + MISC_G, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    for (int sp = 0; sp < 2; sp++) {
        h->sp[sp].nseg = h->h_misc[MISC_G + 2 * sp];
        h->sp[sp].ng = h->h_misc[MISC_G + 2 * sp + 1];
    }
    return TB_OK;
}

// walk-kernel time of the last force (after the stream has synchronised)
void note_walk_ms(tb_handle *h)
{
    const float w = h->cfg.walk == TB_WALK_MIXED ? tb_mixed_walk_ms(h->mx) : -1.0f;
    h->st.ms_walk = w >= 0.0f ? (double)w : h->st.ms_force;
}

// integrity flags raised during phase B (bad permutation / malformed table)
int build_check(tb_handle *h)
{
    HIPCHK(hipMemcpyAsync(h->h_misc, h->d_misc, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    for (int sp = 0; sp < 2; sp++)
        if (h->h_misc[sp] & 12) {
            h->built = false;
            return fail(TB_ERR_HIP, "species %d: internal tree integrity check failed (%d)", sp, h->h_misc[sp]);
        }
    return TB_OK;
}

int build_all(tb_handle *h)
{
    int which[3], rc;
    if ((rc = build_phase_a(h, which))) return rc;
    HIPCHK(hipStreamSynchronize(h->s));
    if ((rc = build_phase_b(h, which))) return rc;
    return build_check(h);
}

int force_all(tb_handle *h, bool counting)
{
    hipStream_t s = h->s;
    if (h->cfg.walk == TB_WALK_MIXED && !counting) {
        int rc = tb_mixed_force(s, h->mx, h->sp, h->cfg, h->epart, h->d_sums);
        if (rc) return fail(rc, "%s", tb_mixed_error());
        // the mixed walk's RDF: every pair within the histograms' range
        if (h->rdf_on && (rc = tb_rdf_pairs(s, h->mx.t[0], h->sp[1], h->rdf, h->d_misc + MISC_RDF)))
            return fail(rc, "mixed-walk RDF launch failed (nbin %d, ion groups of %d)", h->rdf.nbin, h->sp[1].gG);
    } else {
        if (counting) HIPCHK(hipMemsetAsync(h->d_cnt, 0, TB_NCNT * sizeof(int64_t), s));
        const int nb0 = tb_walk64_blocks(h->sp[0].n);
        for (int tsp = 0; tsp < 2; tsp++) {
            double *part = h->epart + (tsp == 0 ? 0 : 2 * nb0);
            // the counting pass re-runs a force already counted in the RDF: it adds nothing
            tb_walk64(s, tsp, h->sp[tsp], h->sp[0], h->sp[1], h->cfg.rs, part,
                      counting ? h->d_cnt + 6 * tsp : nullptr, h->rdf_on && !counting ? &h->rdf : nullptr);
            tb_reduce_partials(s, tb_walk64_blocks(h->sp[tsp].n), 2, part, h->d_sums + 2 * tsp);
        }
    }
    HIPCHK(hipGetLastError());
    h->have_force = true;
    return TB_OK;
}

// walk integrity flag (d_sums[15], written by the mixed walk's reduction)
int check_walk_flags(tb_handle *h)
{
    if (h->cfg.walk == TB_WALK_MIXED && h->h_sums[15] != 0.0)
        return fail(TB_ERR_HIP, "mixed walk: traversal stack overflow or malformed node table (flag %g)",
                    h->h_sums[15]);
    return TB_OK;
}

// d_sums[0..3] = ee, ei, ie, ii (electron-target launch, then ion-target launch)
void epot_from_sums(const double *v, double *ee)
{
    // sys%Epot_e = Epot(ee) + Epot(ie) + Epot(ei); sys%Epot_i = Epot(ii) (tree.f90:688-712)
    ee[0] = v[0] + v[2] + v[1];
    ee[1] = v[3];
}

int vv2_all(tb_handle *h, double dt, int sample, bool fuse_next)
{
    hipStream_t s = h->s;
    const int nb0 = tb_vv_blocks(h->sp[0].n);
    for (int sp = 0; sp < 2; sp++) {
        double *part = h->epart + (sp == 0 ? 0 : 2 * nb0);
        if (fuse_next) tb_kdk(s, h->sp[sp], h->cfg.xm[sp], dt, sample, part);
        else tb_vv2(s, h->sp[sp], h->cfg.xm[sp], dt, sample, part);
        if (sample) tb_reduce_partials(s, tb_vv_blocks(h->sp[sp].n), 2, part, h->d_sums + 4 + 2 * sp);
    }
    HIPCHK(hipGetLastError());
    return TB_OK;
}

}  // namespace

// ============================================================== C-ABI

extern "C" {

const char *tb_last_error(void) { return g_err.c_str(); }

int tb_abi_sizes(int32_t *pt, int32_t *config, int32_t *stats)
{
    if (!pt || !config || !stats) return fail(TB_ERR_ARG, "null argument");
    *pt = (int32_t)sizeof(tb_pt);
    *config = (int32_t)sizeof(tb_config);
    *stats = (int32_t)sizeof(tb_stats);
    return TB_OK;
}

int tb_create(tb_handle **out, const tb_config *cfg)
{
    if (!out) return fail(TB_ERR_ARG, "null handle pointer");
    *out = nullptr;
    int rc = check_cfg(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(TB_ERR_NODEV, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(TB_ERR_ARG, "device %d out of range", cfg->device);
    HIPCHK(hipSetDevice(cfg->device));
    // this device's scratch array, looked up once per device with it current
    if (!scratch_ptr(cfg->device)) return fail(TB_ERR_HIP, "device %d: scratch array address lookup failed", cfg->device);
    tb_handle *h = new tb_handle();
    h->cfg = *cfg;
    if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(TB_ERR_HIP, "stream creation failed");
    }
    bool ok = hipStreamCreateWithFlags(&h->s1, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&h->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&h->join, hipEventDisableTiming) == hipSuccess;
    ok = ok && dalloc_t(h, &h->bbox_tmp, 2048 * 6) == TB_OK && dalloc_t(h, &h->bbox_tmp1, 2048 * 6) == TB_OK &&
         dalloc_t(h, &h->d_misc, 256) == TB_OK &&
              dalloc_t(h, &h->d_sums, 16) == TB_OK && dalloc_t(h, &h->d_cnt, TB_NCNT) == TB_OK;
    ok = ok && hipHostMalloc((void **)&h->h_misc, 256 * sizeof(int32_t), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&h->h_sums, 16 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    for (int k = 0; k < 6 && ok; k++) ok = hipEventCreate(&h->ev[k]) == hipSuccess;
    if (!ok) {
        std::string msg = g_err;
        tb_destroy(h);
        return fail(TB_ERR_NOMEM, "handle allocation failed %s", msg.c_str());
    }
    *out = h;
    return TB_OK;
}

int tb_destroy(tb_handle *h)
{
    if (!h) return TB_OK;
    hipSetDevice(h->cfg.device);
    if (h->s) hipStreamSynchronize(h->s);
    if (h->s1) hipStreamSynchronize(h->s1);
    for (int sp = 0; sp < 2; sp++) free_species(h, h->sp[sp]);
    free_joint(h, h->jt);
    tb_mixed_free(h->mx, &h->bytes);
    dfree(h, h->sort_tmp, h->sort_tmp_bytes);
    dfree(h, h->sort_tmp1, h->sort_tmp_bytes);
    dfree(h, h->scan_tmp, h->scan_tmp_bytes);
    dfree(h, h->scan_tmp1, h->scan_tmp_bytes);
    dfree(h, h->bbox_tmp, 0);
    dfree(h, h->bbox_tmp1, 0);
    dfree(h, h->d_misc, 0);
    dfree(h, h->epart, 0);
    dfree(h, h->d_sums, 0);
    dfree(h, h->d_cnt, 0);
    dfree(h, h->rdf.hist, 0);
    if (h->h_misc) hipHostFree(h->h_misc);
    if (h->h_sums) hipHostFree(h->h_sums);
    for (int k = 0; k < 6; k++)
        if (h->ev[k]) hipEventDestroy(h->ev[k]);
    if (h->fork) hipEventDestroy(h->fork);
    if (h->join) hipEventDestroy(h->join);
    if (h->s1) hipStreamDestroy(h->s1);
    if (h->s) hipStreamDestroy(h->s);
    delete h;
    return TB_OK;
}

int tb_set_config(tb_handle *h, const tb_config *cfg)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    int rc = check_cfg(cfg);
    if (rc) return rc;
    if (cfg->device != h->cfg.device) return fail(TB_ERR_ARG, "device cannot change on a live handle");
    if (h->rdf_on && cfg->walk == TB_WALK_MIXED && h->rdf.nbin > TB_RDF32_MAXBIN)
        return fail(TB_ERR_STATE, "the mixed walk's RDF holds at most %d bins: disable it before switching walks",
                    TB_RDF32_MAXBIN);
    h->cfg = *cfg;
    h->built = false;  // rcrit and walk records depend on theta / walk
    return TB_OK;
}

int tb_set_particles(tb_handle *h, int32_t ne, const tb_pt *e, int32_t ni, const tb_pt *i)
{
    if (!h || !e || !i) return fail(TB_ERR_ARG, "null argument");
    if (ne < 1 || ni < 1) return fail(TB_ERR_ARG, "need at least one particle per species (ne=%d ni=%d)", ne, ni);
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipStreamSynchronize(h->s));
    const int nn[2] = {ne, ni};
    const tb_pt *src[2] = {e, i};
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        if (S.n != nn[sp]) {
            free_species(h, S);
            int rc = alloc_species(h, S, nn[sp]);
            if (rc) return rc;
        }
        S.cur = 0;
        const int n = nn[sp];
        std::vector<double> col((size_t)n);
        // forces are not taken from the caller: they start at zero (the
        // learned criterion reads the previous force, and none exists yet)
        for (int k = 0; k < 9; k++) {
            for (int j = 0; j < n; j++) {
                const tb_pt &p = src[sp][j];
                col[j] = k < 3 ? p.xx[k] : (k < 6 ? p.xv[k - 3] : 0.0);
            }
            double *dst = k < 3 ? S.x[0][k] : (k < 6 ? S.v[0][k - 3] : S.f[0][k - 6]);
            HIPCHK(hipMemcpy(dst, col.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        }
        std::vector<int32_t> id((size_t)n);
        for (int j = 0; j < n; j++) id[j] = j;
        HIPCHK(hipMemcpy(S.orig[0], id.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    h->built = false;
    h->have_force = false;
    return TB_OK;
}

int tb_get_particles(tb_handle *h, tb_pt *e, tb_pt *i)
{
    if (!h || !e || !i) return fail(TB_ERR_ARG, "null argument");
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "no particles set");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipStreamSynchronize(h->s));
    tb_pt *dst[2] = {e, i};
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        const int n = S.n, c = S.cur;
        std::vector<int32_t> orig((size_t)n);
        HIPCHK(hipMemcpy(orig.data(), S.orig[c], (size_t)n * 4, hipMemcpyDeviceToHost));
        std::vector<double> col((size_t)n);
        for (int k = 0; k < 9; k++) {
            const double *src = k < 3 ? S.x[c][k] : (k < 6 ? S.v[c][k - 3] : S.f[c][k - 6]);
            HIPCHK(hipMemcpy(col.data(), src, (size_t)n * 8, hipMemcpyDeviceToHost));
            for (int j = 0; j < n; j++) {
                tb_pt &p = dst[sp][orig[j]];
                if (k < 3) p.xx[k] = col[j];
                else if (k < 6) p.xv[k - 3] = col[j];
                else p.ff[k - 6] = col[j];
            }
        }
    }
    return TB_OK;
}

int tb_build(tb_handle *h)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "tb_build before tb_set_particles");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipEventRecord(h->ev[0], h->s));
    int rc = build_all(h);
    if (rc) {
        h->built = false;
        return rc;
    }
    HIPCHK(hipEventRecord(h->ev[1], h->s));
    // the build ends synchronised (its integrity read-back), so this is its latency
    HIPCHK(hipEventSynchronize(h->ev[1]));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h->ev[0], h->ev[1]) == hipSuccess) h->st.ms_build = ms;
    return TB_OK;
}

int tb_force(tb_handle *h, double *epot_e, double *epot_i)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (!h->built) return fail(TB_ERR_STATE, "tb_force before tb_build");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipEventRecord(h->ev[2], h->s));
    HIPCHK(hipMemsetAsync(h->d_sums + 15, 0, sizeof(double), h->s));
    int rc = force_all(h, false);
    if (rc) return rc;
    HIPCHK(hipEventRecord(h->ev[3], h->s));
    HIPCHK(hipMemcpyAsync(h->h_sums, h->d_sums, 16 * sizeof(double), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    if ((rc = check_walk_flags(h))) return rc;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h->ev[2], h->ev[3]) == hipSuccess) h->st.ms_force = ms;
    note_walk_ms(h);
    epot_from_sums(h->h_sums, h->epot);
    if (epot_e) *epot_e = h->epot[0];
    if (epot_i) *epot_i = h->epot[1];
    return TB_OK;
}

int tb_vv1(tb_handle *h, double dt)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "no particles set");
    HIPCHK(hipSetDevice(h->cfg.device));
    for (int sp = 0; sp < 2; sp++) tb_vv1(h->s, h->sp[sp], h->cfg.xm[sp], dt);
    HIPCHK(hipGetLastError());
    h->built = false;  // positions moved
    return TB_OK;
}

int tb_vv2(tb_handle *h, double dt, int32_t sample, double *sums)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "no particles set");
    HIPCHK(hipSetDevice(h->cfg.device));
    int rc = vv2_all(h, dt, sample ? 1 : 0, false);
    if (rc) return rc;
    if (sample) {
        HIPCHK(hipMemcpyAsync(h->h_sums + 4, h->d_sums + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, h->s));
        HIPCHK(hipStreamSynchronize(h->s));
        if (sums) {
            sums[0] = h->h_sums[4];  // mv2el
            sums[1] = h->h_sums[6];  // mv2ion
            sums[2] = h->h_sums[5];  // Te_rad
            sums[3] = h->h_sums[7];  // Ti_rad
        }
    }
    return TB_OK;
}

int tb_step(tb_handle *h, double dt, int32_t nsteps, double *out)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (nsteps < 1) return fail(TB_ERR_ARG, "nsteps must be >= 1");
    if (!h->have_force) return fail(TB_ERR_STATE, "tb_step needs forces: call tb_build + tb_force first");
    HIPCHK(hipSetDevice(h->cfg.device));
    for (int sp = 0; sp < 2; sp++) tb_vv1(h->s, h->sp[sp], h->cfg.xm[sp], dt);
    HIPCHK(hipMemsetAsync(h->d_sums + 15, 0, sizeof(double), h->s));
    int rc;
    for (int k = 0; k < nsteps; k++) {
        HIPCHK(hipEventRecord(h->ev[0], h->s));
        if ((rc = build_all(h))) return rc;
        // the build synchronised: the walk flag of the previous step's force
        // (copied after it) is on the host; stop before integrating past it
        if (k > 0 && (rc = check_walk_flags(h))) return rc;
        HIPCHK(hipEventRecord(h->ev[1], h->s));
        if ((rc = force_all(h, false))) return rc;
        HIPCHK(hipMemcpyAsync(h->h_sums + 15, h->d_sums + 15, sizeof(double), hipMemcpyDeviceToHost, h->s));
        HIPCHK(hipEventRecord(h->ev[3], h->s));
        const bool last = k == nsteps - 1;
        if ((rc = vv2_all(h, dt, last && out ? 1 : 0, !last))) return rc;
        HIPCHK(hipEventRecord(h->ev[5], h->s));
    }
    HIPCHK(hipMemcpyAsync(h->h_sums, h->d_sums, 16 * sizeof(double), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    if ((rc = check_walk_flags(h))) return rc;
    epot_from_sums(h->h_sums, h->epot);
    h->built = false;
    if (out) {
        out[0] = h->epot[0];
        out[1] = h->epot[1];
        out[2] = h->h_sums[4];
        out[3] = h->h_sums[6];
        out[4] = h->h_sums[5];
        out[5] = h->h_sums[7];
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, h->ev[0], h->ev[1]) == hipSuccess) h->st.ms_build = ms;
    if (hipEventElapsedTime(&ms, h->ev[1], h->ev[3]) == hipSuccess) h->st.ms_force = ms;
    note_walk_ms(h);
    if (hipEventElapsedTime(&ms, h->ev[3], h->ev[5]) == hipSuccess) h->st.ms_vv = ms;
    return TB_OK;
}

int tb_export_tree(tb_handle *h, int32_t species, int32_t cap, int32_t *nnodes, int32_t *depth, int32_t *leaf,
                   int32_t *pt, int32_t *count, int32_t *next, double *xx, double *lx, double *qeff, double *dx,
                   double *D, double *Q, double *QQ)
{
    if (!h || !nnodes) return fail(TB_ERR_ARG, "null argument");
    if (species != 0 && species != 1) return fail(TB_ERR_ARG, "species must be 0 or 1");
    if (!h->built) return fail(TB_ERR_STATE, "tb_export_tree before tb_build");
    if (h->tree_walk != TB_WALK_REF64) return fail(TB_ERR_STATE, "tb_export_tree needs trees built for the FP64 walk");
    SpeciesDev &S = h->sp[species];
    *nnodes = S.nn;
    if (cap < S.nn) return fail(TB_ERR_ARG, "cap %d < %d nodes", cap, S.nn);
    if (!depth || !leaf || !pt || !count || !next || !xx || !lx || !qeff || !dx || !D || !Q || !QQ)
        return fail(TB_ERR_ARG, "null output array");
    HIPCHK(hipSetDevice(h->cfg.device));
    const size_t n = (size_t)S.nn;
    // 5 int32 columns, then 14 doubles per node (xx 3, qeff, dx, D 3, Q 3, QQ 3)
    const size_t tmp_bytes = n * 5 * 4 + 16 + n * 14 * 8;
    void *tmp = nullptr;
    int rc = dalloc(h, &tmp, tmp_bytes);
    if (rc) return rc;
    char *b = (char *)tmp;
    int32_t *d_depth = (int32_t *)b, *d_leaf = d_depth + n, *d_pt = d_leaf + n, *d_cnt = d_pt + n,
            *d_next = d_cnt + n;
    double *dd = (double *)(((uintptr_t)(d_next + n) + 15) & ~(uintptr_t)15);
    double *d_xx = dd, *d_q = d_xx + 3 * n, *d_dx = d_q + n, *d_D = d_dx + n, *d_Q = d_D + 3 * n, *d_QQ = d_Q + 3 * n;
    tb_export_gather(h->s, S, species ? 1 : -1, d_depth, d_leaf, d_pt, d_cnt, d_next, d_xx, d_q, d_dx, d_D, d_Q, d_QQ);
    double root[6];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    if (e == hipSuccess) e = hipMemcpy(root, S.root, sizeof(root), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(depth, d_depth, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(leaf, d_leaf, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(pt, d_pt, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(count, d_cnt, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(next, d_next, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(xx, d_xx, n * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(qeff, d_q, n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(dx, d_dx, n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(D, d_D, n * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(Q, d_Q, n * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(QQ, d_QQ, n * 24, hipMemcpyDeviceToHost);
    dfree(h, tmp, tmp_bytes);
    if (e != hipSuccess) return fail(TB_ERR_HIP, "tb_export_tree: %s", hipGetErrorString(e));
    // cell size: the root length halved once per level (exact), tree.f90:126
    for (size_t j = 0; j < n; j++) {
        double l[3] = {root[3], root[4], root[5]};
        for (int d = 0; d < depth[j]; d++)
            for (int k = 0; k < 3; k++) l[k] = l[k] * 0.5;
        for (int k = 0; k < 3; k++) lx[3 * j + k] = l[k];
    }
    return TB_OK;
}

int tb_direct_sample(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f)
{
    if (!h || !tspecies || !tidx || !f) return fail(TB_ERR_ARG, "null argument");
    if (nt < 1) return fail(TB_ERR_ARG, "nt must be >= 1");
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "no particles set");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipStreamSynchronize(h->s));
    // reference index -> current (key-ordered) index
    std::vector<int32_t> sidx((size_t)nt);
    std::vector<int32_t> inv[2];
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        std::vector<int32_t> orig((size_t)S.n);
        HIPCHK(hipMemcpy(orig.data(), S.orig[S.cur], (size_t)S.n * 4, hipMemcpyDeviceToHost));
        inv[sp].assign((size_t)S.n, 0);
        for (int j = 0; j < S.n; j++) inv[sp][orig[j]] = j;
    }
    for (int t = 0; t < nt; t++) {
        int sp = tspecies[t];
        if (sp != 0 && sp != 1) return fail(TB_ERR_ARG, "tspecies[%d] = %d", t, sp);
        if (tidx[t] < 0 || tidx[t] >= h->sp[sp].n) return fail(TB_ERR_ARG, "tidx[%d] = %d out of range", t, tidx[t]);
        sidx[t] = inv[sp][tidx[t]];
    }
    void *tmp = nullptr;
    int rc = dalloc(h, &tmp, (size_t)nt * (8 + 24) + 64);
    if (rc) return rc;
    int32_t *d_sp = (int32_t *)tmp, *d_ix = d_sp + nt;
    double *d_f = (double *)(((uintptr_t)(d_ix + nt) + 15) & ~(uintptr_t)15);
    hipError_t e = hipMemcpy(d_sp, tspecies, (size_t)nt * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ix, sidx.data(), (size_t)nt * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        tb_direct(h->s, nt, d_sp, d_ix, h->sp[0], h->sp[1], h->cfg.rs, d_f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    if (e == hipSuccess) e = hipMemcpy(f, d_f, (size_t)nt * 24, hipMemcpyDeviceToHost);
    dfree(h, tmp, (size_t)nt * (8 + 24) + 64);
    if (e != hipSuccess) return fail(TB_ERR_HIP, "tb_direct_sample: %s", hipGetErrorString(e));
    return TB_OK;
}

int tb_direct_sample_noalloc(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f)
{
    if (!h || !tspecies || !tidx || !f) return fail(TB_ERR_ARG, "null argument");
    if (nt < 1 || nt > TB_NOALLOC_MAX_TARGETS)
        return fail(TB_ERR_ARG, "nt must be in 1 .. %d (got %d)", TB_NOALLOC_MAX_TARGETS, nt);
    if (h->sp[0].n == 0) return fail(TB_ERR_STATE, "no particles set");
    for (int t = 0; t < nt; t++) {
        const int sp = tspecies[t];
        if (sp != 0 && sp != 1) return fail(TB_ERR_ARG, "tspecies[%d] = %d", t, sp);
        if (tidx[t] < 0 || tidx[t] >= h->sp[sp].n) return fail(TB_ERR_ARG, "tidx[%d] = %d out of range", t, tidx[t]);
    }
    HIPCHK(hipSetDevice(h->cfg.device));
    unsigned char *base = (unsigned char *)scratch_ptr(h->cfg.device);
    if (!base) return fail(TB_ERR_HIP, "scratch array address lookup failed");
    int32_t *d_sp = (int32_t *)base, *d_ref = (int32_t *)(base + NOALLOC_INT_BYTES),
            *d_ix = (int32_t *)(base + 2 * NOALLOC_INT_BYTES);
    double *d_f = (double *)(base + 3 * NOALLOC_INT_BYTES);
    hipStream_t s = h->s;
    HIPCHK(hipMemcpyAsync(d_sp, tspecies, (size_t)nt * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d_ref, tidx, (size_t)nt * 4, hipMemcpyHostToDevice, s));
    // inverse permutations of the current key order, into each species'
    // build-time scratch column (rewritten by every build)
    for (int sp = 0; sp < 2; sp++) tb_inverse_perm(s, h->sp[sp].n, h->sp[sp].orig[h->sp[sp].cur], h->sp[sp].cnt);
    hipLaunchKernelGGL(k_map_targets, dim3((nt + 255) / 256), dim3(256), 0, s, nt, d_sp, d_ref, h->sp[0].cnt,
                       h->sp[1].cnt, d_ix);
    tb_direct(s, nt, d_sp, d_ix, h->sp[0], h->sp[1], h->cfg.rs, d_f);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(f, d_f, (size_t)nt * 24, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return TB_OK;
}

int tb_local_eval(tb_handle *h, int32_t order, int32_t ncell, const float *cells, int32_t ny, const float *y,
                  float *out)
{
    if (!h || !cells || !y || !out) return fail(TB_ERR_ARG, "null argument");
    if (order < 3 || order > 6) return fail(TB_ERR_ARG, "order must be 3 to 6 (got %d)", order);
    if (ncell < 1 || ny < 1 || ncell > (1 << 20) || ny > (1 << 20))
        return fail(TB_ERR_ARG, "ncell and ny must be in 1 .. 2^20");
    HIPCHK(hipSetDevice(h->cfg.device));
    const size_t nb = (size_t)ncell * TB_LOCAL_CELL * 4 + (size_t)ny * 12 + (size_t)ny * 16 + 64;
    void *tmp = nullptr;
    int rc = dalloc(h, &tmp, nb);
    if (rc) return rc;
    float *d_c = (float *)tmp, *d_y = d_c + (size_t)ncell * TB_LOCAL_CELL, *d_o = d_y + (size_t)ny * 3;
    hipError_t e = hipMemcpy(d_c, cells, (size_t)ncell * TB_LOCAL_CELL * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_y, y, (size_t)ny * 12, hipMemcpyHostToDevice);
    if (e == hipSuccess && tb_local_run(h->s, order, ncell, d_c, ny, d_y, d_o) != TB_OK) e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    if (e == hipSuccess) e = hipMemcpy(out, d_o, (size_t)ny * 16, hipMemcpyDeviceToHost);
    dfree(h, tmp, nb);
    if (e != hipSuccess) return fail(TB_ERR_HIP, "tb_local_eval: %s", hipGetErrorString(e));
    return TB_OK;
}

// reference index -> key-order index for sampled targets
static int map_targets(tb_handle *h, int nt, const int32_t *tspecies, const int32_t *tidx, std::vector<int32_t> &sidx)
{
    sidx.assign((size_t)nt, 0);
    std::vector<int32_t> inv[2];
    for (int sp = 0; sp < 2; sp++) {
        SpeciesDev &S = h->sp[sp];
        std::vector<int32_t> orig((size_t)S.n);
        HIPCHK(hipMemcpy(orig.data(), S.orig[S.cur], (size_t)S.n * 4, hipMemcpyDeviceToHost));
        inv[sp].assign((size_t)S.n, 0);
        for (int j = 0; j < S.n; j++) inv[sp][orig[j]] = j;
    }
    for (int t = 0; t < nt; t++) {
        int sp = tspecies[t];
        if (sp != 0 && sp != 1) return fail(TB_ERR_ARG, "tspecies[%d] = %d", t, sp);
        if (tidx[t] < 0 || tidx[t] >= h->sp[sp].n) return fail(TB_ERR_ARG, "tidx[%d] = %d out of range", t, tidx[t]);
        sidx[t] = inv[sp][tidx[t]];
    }
    return TB_OK;
}

int tb_force_targets(tb_handle *h, int32_t nt, const int32_t *tspecies, const int32_t *tidx, double *f)
{
    if (!h || !tspecies || !tidx || !f || nt < 1) return fail(TB_ERR_ARG, "bad argument");
    if (!h->built || h->tree_walk != TB_WALK_REF64)
        return fail(TB_ERR_STATE, "tb_force_targets needs trees built for the FP64 walk");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipStreamSynchronize(h->s));
    std::vector<int32_t> sidx;
    int rc = map_targets(h, nt, tspecies, tidx, sidx);
    if (rc) return rc;
    void *tmp = nullptr;
    const size_t bytes = (size_t)nt * (8 + 24) + 64;
    if ((rc = dalloc(h, &tmp, bytes))) return rc;
    int32_t *d_sp = (int32_t *)tmp, *d_ix = d_sp + nt;
    double *d_f = (double *)(((uintptr_t)(d_ix + nt) + 15) & ~(uintptr_t)15);
    hipError_t e = hipMemcpy(d_sp, tspecies, (size_t)nt * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_ix, sidx.data(), (size_t)nt * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        tb_walk64_list(h->s, nt, d_sp, d_ix, 1, h->sp[0], h->sp[1], h->cfg.rs, d_f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    if (e == hipSuccess) e = hipMemcpy(f, d_f, (size_t)nt * 24, hipMemcpyDeviceToHost);
    dfree(h, tmp, bytes);
    return e == hipSuccess ? TB_OK : fail(TB_ERR_HIP, "tb_force_targets: %s", hipGetErrorString(e));
}

int tb_time_walk(tb_handle *h, int32_t stride, float *ms, int64_t *nwalked)
{
    if (!h || !ms || !nwalked || stride < 1) return fail(TB_ERR_ARG, "bad argument");
    if (!h->built || h->tree_walk != TB_WALK_REF64)
        return fail(TB_ERR_STATE, "tb_time_walk needs trees built for the FP64 walk");
    HIPCHK(hipSetDevice(h->cfg.device));
    // slots: every stride-th block of 64 targets per species; the last block may be short
    int nt = 0;
    int64_t walked = 0;
    for (int sp = 0; sp < 2; sp++) {
        const int n = h->sp[sp].n, blocks = (n + 63) / 64, sampled = (blocks + stride - 1) / stride;
        nt += sampled * 64;
        for (int b = 0; b < blocks; b += stride) walked += std::min(64, n - 64 * b);
    }
    int32_t *dummy = nullptr;
    int rc = dalloc_t(h, &dummy, 16);
    if (rc) return rc;
    HIPCHK(hipEventRecord(h->ev[2], h->s));
    tb_walk64_list(h->s, nt, nullptr, dummy, stride, h->sp[0], h->sp[1], h->cfg.rs, nullptr);
    HIPCHK(hipEventRecord(h->ev[3], h->s));
    HIPCHK(hipEventSynchronize(h->ev[3]));
    HIPCHK(hipEventElapsedTime(ms, h->ev[2], h->ev[3]));
    dfree(h, dummy, 64);
    *nwalked = walked;
    return TB_OK;
}

int tb_count_interactions(tb_handle *h, tb_stats *st)
{
    if (!h || !st) return fail(TB_ERR_ARG, "null argument");
    if (!h->built) return fail(TB_ERR_STATE, "tb_count_interactions before tb_build");
    HIPCHK(hipSetDevice(h->cfg.device));
    int64_t c[TB_NCNT] = {};
    if (h->cfg.walk == TB_WALK_MIXED) {
        int rc = tb_mixed_count(h->s, h->mx, h->sp, h->cfg, h->epart, h->d_sums, h->d_cnt);
        if (rc) return fail(rc, "%s", tb_mixed_error());
    } else {
        int rc = force_all(h, true);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(c, h->d_cnt, sizeof(c), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    // launch order: electron targets (e-tree = ee, i-tree = ei), ion targets (ie, ii)
    const int map[4] = {0, 6, 3, 9};  // ee, ie, ei, ii -> offset in c
    for (int p = 0; p < 4; p++) {
        h->st.visits[p] = c[map[p]];
        h->st.m2p[p] = c[map[p] + 1];
        h->st.p2p[p] = c[map[p] + 2];
    }
    // far cells of the mixed walk's local expansions, by target species (ee, ii)
    h->st.m2l[0] = c[12];
    h->st.m2l[1] = h->st.m2l[2] = 0;
    h->st.m2l[3] = c[13];
    h->st.far_pairs[0] = c[14];
    h->st.far_pairs[1] = c[15];
    // two-level far field: (far cell, segment) pairs by target species
    h->st.seg_pairs[0] = c[16];
    h->st.seg_pairs[1] = c[17];
    // M2P pairs to the third moment only, by target species
    h->st.m2p3[0] = c[18];
    h->st.m2p3[1] = c[19];
    *st = h->st;
    return TB_OK;
}

int tb_mlp_set(tb_handle *h, const float *params, int32_t n)
{
    if (!h || !params) return fail(TB_ERR_ARG, "null argument");
    if (n != TB_MLP_NPARAMS) return fail(TB_ERR_ARG, "expected %d parameters, got %d", TB_MLP_NPARAMS, n);
    HIPCHK(hipSetDevice(h->cfg.device));
    if (tb_mlp_upload(h->s, h->mx.mlp, params)) return fail(TB_ERR_NOMEM, "MLP upload failed");
    HIPCHK(hipStreamSynchronize(h->s));
    return TB_OK;
}

int tb_mlp_init(tb_handle *h, uint64_t seed)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    // He-uniform weights (limit sqrt(6 / fan_in)), zero biases, from
    // Philox4x32-10 keyed by the seed: counter (k, 0x4d4c5000, 0, 0) gives
    // the uniforms of weights 4k .. 4k+3 in layout order
    const int nin = TB_MLP_NIN, nh = 32;
    std::vector<float> p(TB_MLP_NPARAMS, 0.0f);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int k = 0;
    uint32_t buf[4] = {0, 0, 0, 0};
    auto uniform = [&](void) {
        if ((k & 3) == 0) {
            const uint32_t ctr[4] = {(uint32_t)(k >> 2), 0x4d4c5000u, 0u, 0u};
            tb_philox4x32_10(ctr, key, buf);
        }
        const uint32_t b = buf[k & 3];
        k++;
        return (double)(b >> 8) * (1.0 / 16777216.0);  // [0, 1)
    };
    const int fan_in[3] = {nin, nh, nh}, fan_out[3] = {nh, nh, 1};
    int o = 0;
    for (int l = 0; l < 3; l++) {
        const double lim = std::sqrt(6.0 / fan_in[l]);
        for (int j = 0; j < fan_in[l] * fan_out[l]; j++) p[o + j] = (float)(lim * (2.0 * uniform() - 1.0));
        o += fan_in[l] * fan_out[l] + fan_out[l];  // biases stay zero
    }
    return tb_mlp_set(h, p.data(), TB_MLP_NPARAMS);
}

int tb_mlp_get(tb_handle *h, float *params, int32_t n)
{
    if (!h || !params) return fail(TB_ERR_ARG, "null argument");
    if (n != TB_MLP_NPARAMS) return fail(TB_ERR_ARG, "expected %d parameters, got %d", TB_MLP_NPARAMS, n);
    if (!h->mx.mlp.P) return fail(TB_ERR_STATE, "no network set or trained");
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipMemcpyAsync(params, h->mx.mlp.P, (size_t)n * 4, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return TB_OK;
}

int tb_mlp_forward(tb_handle *h, int32_t n, const float *X, int32_t bf16, float *out)
{
    if (!h || !X || !out || n < 1) return fail(TB_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(h->cfg.device));
    int rc = tb_mlp_forward(h->s, h->mx.mlp, n, X, bf16, out);
    return rc ? fail(rc, "MLP forward failed (network set?)") : TB_OK;
}

int tb_mlp_loss_grad(tb_handle *h, int32_t n, const float *X, const float *y, float *loss, float *grad)
{
    if (!h || !X || !y || !loss || !grad || n < 1) return fail(TB_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(h->cfg.device));
    int rc = tb_mlp_loss_grad(h->s, h->mx.mlp, n, X, y, loss, grad);
    return rc ? fail(rc, "MLP gradient failed (network set?)") : TB_OK;
}

int tb_mlp_train(tb_handle *h, int32_t n, const float *X, const float *y, int32_t steps, int32_t batch, float lr,
                 uint64_t seed, float *loss_hist)
{
    if (!h || !X || !y || n < 1 || steps < 1) return fail(TB_ERR_ARG, "bad argument");
    HIPCHK(hipSetDevice(h->cfg.device));
    int rc = tb_mlp_train(h->s, h->mx.mlp, n, X, y, steps, batch, lr, seed, loss_hist);
    return rc ? fail(rc, "MLP training failed (network set?)") : TB_OK;
}

int tb_mlp_collect(tb_handle *h, int32_t cap, float *X, float *y, int32_t *n)
{
    if (!h || !X || !y || !n || cap < 1) return fail(TB_ERR_ARG, "bad argument");
    if (h->cfg.walk != TB_WALK_MIXED) return fail(TB_ERR_STATE, "tb_mlp_collect needs the mixed walk");
    if (!h->built) return fail(TB_ERR_STATE, "tb_mlp_collect before tb_build");
    if (!(h->cfg.theta_lo > 0.0) || !(h->cfg.theta_lo < (h->cfg.theta_mixed > 0 ? h->cfg.theta_mixed : h->cfg.theta)))
        return fail(TB_ERR_ARG, "need 0 < theta_lo < theta_mixed");
    HIPCHK(hipSetDevice(h->cfg.device));
    // a fixed-theta force first: its forces are the features' previous forces
    tb_config c = h->cfg;
    c.mac = TB_MAC_FIXED;
    std::swap(c, h->cfg);
    int rc = force_all(h, false);
    std::swap(c, h->cfg);
    if (rc) return rc;
    rc = tb_mixed_collect(h->s, h->mx, h->sp, &h->jt, h->cfg, h->d_sums, cap, X, y, n);
    if (rc) return fail(rc, "%s", tb_mixed_error());
    h->have_force = true;
    return TB_OK;
}

int tb_autotune(tb_handle *h, int32_t *group_out, float *ms_out)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (h->cfg.walk != TB_WALK_MIXED) return fail(TB_ERR_STATE, "tb_autotune needs the mixed walk");
    if (!h->built) return fail(TB_ERR_STATE, "tb_autotune before tb_build");
    HIPCHK(hipSetDevice(h->cfg.device));
    const int saved = h->cfg.group;
    float best = 1e30f;
    int best_tpl = 2, best_lv = 0;
    int rc;
    float tab[3][TB_LIST_VARIANTS];
    for (int tpl = 1; tpl <= 4; tpl *= 2) {
        h->cfg.group = 64 * tpl;
        if ((rc = regroup(h, tpl))) {
            h->cfg.group = saved;
            return rc;
        }
        for (int lv = 0; lv < TB_LIST_VARIANTS; lv++) {
            h->mx.lists = lv;
            float ms = 0.0f;
            for (int rep = 0; rep < 3; rep++) {
                HIPCHK(hipEventRecord(h->ev[2], h->s));
                rc = force_all(h, false);
                if (rc) {
                    h->cfg.group = saved;
                    return rc;
                }
                HIPCHK(hipEventRecord(h->ev[3], h->s));
                HIPCHK(hipEventSynchronize(h->ev[3]));
                float t = 0.0f;
                HIPCHK(hipEventElapsedTime(&t, h->ev[2], h->ev[3]));
                if (rep > 0) ms += t;  // first run warms caches
            }
            ms *= 0.5f;
            tab[tpl == 1 ? 0 : (tpl == 2 ? 1 : 2)][lv] = ms;
            if (ms < best) {
                best = ms;
                best_tpl = tpl;
                best_lv = lv;
            }
        }
    }
    // per group size: its time at the chosen list lengths
    if (ms_out)
        for (int g = 0; g < 3; g++) ms_out[g] = tab[g][best_lv];
    h->cfg.group = saved;
    tb_autotune_record(h->sp[0].n + h->sp[1].n, best_tpl);
    tb_autotune_record_lists(h->sp[0].n + h->sp[1].n, best_lv);
    h->mx.lists = best_lv;
    if ((rc = regroup(h, tb_mixed_tpl(h->cfg, h->sp[0].n + h->sp[1].n)))) return rc;
    if (group_out) *group_out = 64 * best_tpl;
    return TB_OK;
}

// =================================================== batched ensemble

}  // extern "C"

// The batched ensemble: every realisation in one combined state (one
// tb_handle C), built by one pipeline per step (EnsBuild: keys tagged by
// realisation, boxes per realisation) and walked by one persistent launch
// whose domains are the realisations' subtrees.
struct tb_ensemble {
    int nreal = 0, T = 0;
    tb_handle *C = nullptr;
    std::vector<double> rs;                 // Kelbg length per realisation
    std::vector<std::vector<tb_pt>> pe, pi; // the realisations' particles as last set (host)
    std::vector<int32_t> off[2];            // slices of the combined arrays [nreal + 1]
    bool dirty = true;                      // set particles not yet on the device
    bool have_force = false;
    EnsBuild eb;
    int32_t *h_gfirst = nullptr;            // pinned [2][nreal]
    std::vector<int32_t> gfirst[2], gcount[2];
    std::vector<uint32_t> roots;            // root cell of each realisation
    double *d_esum = nullptr;               // [nreal][4]
    double *d_ke = nullptr;                 // [2][nreal][2] mv2, radial sum
    double *d_kpart = nullptr;              // [nreal][chunks][2]
    double *h_buf = nullptr;                // pinned: 4 energies + 4 kinetic sums per realisation
};

namespace {

constexpr int ENS_MAX_CHUNKS = 64;

// group ranges (first, count) of every realisation after its groups were cut
void ens_groups_from_host(tb_ensemble *E)
{
    for (int k = 0; k < 2; k++) {
        E->gfirst[k].assign(E->h_gfirst + (size_t)k * E->nreal, E->h_gfirst + (size_t)(k + 1) * E->nreal);
        E->gcount[k].resize(E->nreal);
        for (int r = 0; r < E->nreal; r++) {
            const int end = r + 1 < E->nreal ? E->gfirst[k][r + 1] : E->C->sp[k].ng;
            E->gcount[k][r] = end - E->gfirst[k][r];
        }
    }
}

int ens_upload(tb_ensemble *E)
{
    std::vector<tb_pt> all[2];
    int nmax = 1;
    for (int k = 0; k < 2; k++) {
        E->off[k].assign(E->nreal + 1, 0);
        for (int r = 0; r < E->nreal; r++) {
            const std::vector<tb_pt> &v = k == 0 ? E->pe[r] : E->pi[r];
            if (v.empty()) return fail(TB_ERR_STATE, "realisation %d has no particles", r);
            E->off[k][r + 1] = E->off[k][r] + (int32_t)v.size();
            nmax = std::max(nmax, (int)v.size());
            all[k].insert(all[k].end(), v.begin(), v.end());
        }
    }
    int rc = tb_set_particles(E->C, E->off[0][E->nreal], all[0].data(), E->off[1][E->nreal], all[1].data());
    if (rc) return rc;
    for (int k = 0; k < 2; k++)
        HIPCHK(hipMemcpy(E->eb.d_off[k], E->off[k].data(), (E->nreal + 1) * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    E->eb.chunks = std::min(ENS_MAX_CHUNKS, tb_sbbox_chunks(nmax));
    E->dirty = false;
    return TB_OK;
}

// one build of every realisation: keys, sorts, target groups and the joint
// table, with one host read-back (node totals, group starts)
int ens_build(tb_ensemble *E)
{
    tb_handle *C = E->C;
    int which[3], rc;
    if ((rc = build_phase_a(C, which, 0, &E->eb))) return rc;
    HIPCHK(hipMemcpyAsync(E->h_gfirst, E->eb.d_gfirst, 2 * E->nreal * sizeof(int32_t), hipMemcpyDeviceToHost,
                          C->s));
    HIPCHK(hipStreamSynchronize(C->s));
    if ((rc = build_phase_b(C, which, &E->eb))) return rc;
    if ((rc = build_check(C))) return rc;
    ens_groups_from_host(E);
    E->roots.resize(E->nreal);
    for (int r = 0; r < E->nreal; r++) E->roots[r] = (uint32_t)(C->jt.lvl_off[E->T] + r);
    return TB_OK;
}

int ens_force(tb_ensemble *E)
{
    const int32_t *gf[2] = {E->gfirst[0].data(), E->gfirst[1].data()};
    const int32_t *gc[2] = {E->gcount[0].data(), E->gcount[1].data()};
    int rc = tb_mixed_force_ens(E->C->s, E->C->mx, E->C->sp, E->nreal, E->roots.data(), E->rs.data(), gf, gc,
                                E->C->cfg, E->d_esum);
    if (rc) return fail(rc, "%s", tb_mixed_error());
    E->C->have_force = true;
    return TB_OK;
}

// cuts the target groups again for walk-group size 64 * tpl (synchronises)
int ens_regroup(tb_ensemble *E, int tpl)
{
    tb_handle *C = E->C;
    for (int k = 0; k < 2; k++)
        tb_groups(C->s, C->sp[k], 64 * tpl, C->d_misc + MISC_G + 2 * k, C->scan_tmp, E->T,
                  E->eb.d_gfirst + (size_t)k * E->nreal);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(C->h_misc + MISC_G, C->d_misc + MISC_G, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, C->s));
    HIPCHK(hipMemcpyAsync(E->h_gfirst, E->eb.d_gfirst, 2 * E->nreal * sizeof(int32_t), hipMemcpyDeviceToHost,
                          C->s));
    HIPCHK(hipStreamSynchronize(C->s));
    for (int k = 0; k < 2; k++) {
        C->sp[k].nseg = C->h_misc[MISC_G + 2 * k];
        C->sp[k].ng = C->h_misc[MISC_G + 2 * k + 1];
    }
    ens_groups_from_host(E);
    return TB_OK;
}

// energies (and, sampled, the kinetic sums of the current state) of every
// realisation back to the host; the walk's and the build's integrity flags
int ens_finish(tb_ensemble *E, double *out, bool sampled)
{
    tb_handle *C = E->C;
    const int n = E->nreal;
    if (sampled)
        for (int k = 0; k < 2; k++)
            tb_ke_seg(C->s, C->sp[k], C->cfg.xm[k], n, E->eb.chunks, E->eb.d_off[k], E->d_kpart,
                      E->d_ke + (size_t)k * n * 2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(E->h_buf, E->d_esum, (size_t)4 * n * sizeof(double), hipMemcpyDeviceToHost, C->s));
    if (sampled)
        HIPCHK(hipMemcpyAsync(E->h_buf + 4 * n, E->d_ke, (size_t)4 * n * sizeof(double), hipMemcpyDeviceToHost,
                              C->s));
    HIPCHK(hipStreamSynchronize(C->s));
    int32_t fl[3];
    HIPCHK(hipMemcpy(fl, C->mx.ctr, 3 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (fl[1] || fl[2])
        return fail(TB_ERR_HIP, "ensemble walk: traversal stack overflow or malformed node table (flags %d %d)",
                    fl[1], fl[2]);
    HIPCHK(hipMemcpy(fl, C->d_misc, 2 * sizeof(int32_t), hipMemcpyDeviceToHost));
    if ((fl[0] | fl[1]) & 12) return fail(TB_ERR_HIP, "ensemble: tree integrity check failed");
    for (int r = 0; r < n; r++) {
        double ep[2];
        epot_from_sums(E->h_buf + 4 * r, ep);
        if (out) {
            double *o = out + 6 * r;
            o[0] = ep[0];
            o[1] = ep[1];
            if (sampled) {
                const double *ke = E->h_buf + 4 * n;
                o[2] = ke[2 * r];            // mv2el
                o[3] = ke[2 * n + 2 * r];    // mv2ion
                o[4] = ke[2 * r + 1];        // Te_rad
                o[5] = ke[2 * n + 2 * r + 1];// Ti_rad
            } else {
                o[2] = o[3] = o[4] = o[5] = 0.0;
            }
        }
    }
    return TB_OK;
}

}  // namespace

extern "C" {

int tb_ensemble_create(tb_ensemble **out, int32_t n, const tb_config *cfgs)
{
    if (!out || !cfgs || n < 1) return fail(TB_ERR_ARG, "bad argument");
    *out = nullptr;
    int T = 0;
    while ((1L << (3 * T)) < n) T++;
    if (T > 4) return fail(TB_ERR_ARG, "at most 4096 realisations (got %d)", n);
    const tb_config &c0 = cfgs[0];
    for (int r = 0; r < n; r++) {
        const tb_config &c = cfgs[r];
        int rc = check_cfg(&c);
        if (rc) return rc;
        if (c.walk != TB_WALK_MIXED) return fail(TB_ERR_ARG, "ensemble realisations use the mixed walk");
        if (c.device != c0.device) return fail(TB_ERR_ARG, "ensemble realisations share one device");
        // one walk launch over one combined table: these are shared, only rs may differ
        if (c.xm[0] != c0.xm[0] || c.xm[1] != c0.xm[1] || c.theta != c0.theta || c.theta_mixed != c0.theta_mixed ||
            c.group != c0.group || c.far_ratio != c0.far_ratio)
            return fail(TB_ERR_ARG, "realisation %d: masses, theta, group and far_ratio must equal realisation 0's", r);
    }
    tb_ensemble *E = new tb_ensemble();
    E->nreal = n;
    E->T = T;
    tb_config cfg = c0;
    cfg.mac = TB_MAC_FIXED;
    int rc = tb_create(&E->C, &cfg);
    if (rc) {
        delete E;
        return rc;
    }
    E->rs.resize(n);
    for (int r = 0; r < n; r++) E->rs[r] = cfgs[r].rs;
    E->pe.resize(n);
    E->pi.resize(n);
    E->eb.nreal = n;
    E->eb.T = T;
    tb_handle *C = E->C;
    bool ok = dalloc_t(C, &E->eb.d_off[0], n + 1) == TB_OK && dalloc_t(C, &E->eb.d_off[1], n + 1) == TB_OK &&
              dalloc_t(C, &E->eb.d_roots, (size_t)3 * n * 6) == TB_OK &&
              dalloc_t(C, &E->eb.d_part, (size_t)2 * n * ENS_MAX_CHUNKS * 6) == TB_OK &&
              dalloc_t(C, &E->eb.d_gfirst, (size_t)2 * n) == TB_OK && dalloc_t(C, &E->d_esum, (size_t)4 * n) == TB_OK &&
              dalloc_t(C, &E->d_ke, (size_t)4 * n) == TB_OK &&
              dalloc_t(C, &E->d_kpart, (size_t)2 * n * ENS_MAX_CHUNKS) == TB_OK;
    ok = ok && hipHostMalloc((void **)&E->h_gfirst, (size_t)2 * n * sizeof(int32_t), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&E->h_buf, (size_t)8 * n * sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        tb_ensemble_destroy(E);
        return fail(TB_ERR_NOMEM, "ensemble allocation failed");
    }
    *out = E;
    return TB_OK;
}

int tb_ensemble_destroy(tb_ensemble *E)
{
    if (!E) return TB_OK;
    if (E->h_gfirst) hipHostFree(E->h_gfirst);
    if (E->h_buf) hipHostFree(E->h_buf);
    if (E->C) {
        hipSetDevice(E->C->cfg.device);
        hipStreamSynchronize(E->C->s);
        void *dev[8] = {E->eb.d_off[0], E->eb.d_off[1], E->eb.d_roots, E->eb.d_part, E->eb.d_gfirst, E->d_esum,
                        E->d_ke, E->d_kpart};
        for (void *p : dev) dfree(E->C, p, 0);
        tb_destroy(E->C);
    }
    delete E;
    return TB_OK;
}

int tb_ensemble_set_particles(tb_ensemble *E, int32_t r, int32_t ne, const tb_pt *e, int32_t ni, const tb_pt *i)
{
    if (!E || r < 0 || r >= E->nreal) return fail(TB_ERR_ARG, "bad realisation index");
    if (!e || !i) return fail(TB_ERR_ARG, "null argument");
    if (ne < 1 || ni < 1) return fail(TB_ERR_ARG, "need at least one particle per species (ne=%d ni=%d)", ne, ni);
    E->pe[r].assign(e, e + ne);
    E->pi[r].assign(i, i + ni);
    for (auto &p : E->pe[r]) p.ff[0] = p.ff[1] = p.ff[2] = 0.0;
    for (auto &p : E->pi[r]) p.ff[0] = p.ff[1] = p.ff[2] = 0.0;
    E->dirty = true;
    E->have_force = false;
    return TB_OK;
}

int tb_ensemble_get_particles(tb_ensemble *E, int32_t r, tb_pt *e, tb_pt *i)
{
    if (!E || r < 0 || r >= E->nreal) return fail(TB_ERR_ARG, "bad realisation index");
    if (!e || !i) return fail(TB_ERR_ARG, "null argument");
    if (E->dirty) {
        std::copy(E->pe[r].begin(), E->pe[r].end(), e);
        std::copy(E->pi[r].begin(), E->pi[r].end(), i);
        return TB_OK;
    }
    tb_handle *C = E->C;
    HIPCHK(hipSetDevice(C->cfg.device));
    HIPCHK(hipStreamSynchronize(C->s));
    tb_pt *dst[2] = {e, i};
    for (int k = 0; k < 2; k++) {
        // realisation r's slice: its caller-order slice, and (keys tagged by
        // r first) its slice of the key order
        SpeciesDev &S = C->sp[k];
        const int b = E->off[k][r], n = E->off[k][r + 1] - b, c = S.cur;
        std::vector<int32_t> orig((size_t)n);
        HIPCHK(hipMemcpy(orig.data(), S.orig[c] + b, (size_t)n * 4, hipMemcpyDeviceToHost));
        std::vector<double> col((size_t)n);
        for (int q = 0; q < 9; q++) {
            const double *src = q < 3 ? S.x[c][q] : (q < 6 ? S.v[c][q - 3] : S.f[c][q - 6]);
            HIPCHK(hipMemcpy(col.data(), src + b, (size_t)n * 8, hipMemcpyDeviceToHost));
            for (int j = 0; j < n; j++) {
                const int at = orig[j] - b;
                if (at < 0 || at >= n) return fail(TB_ERR_HIP, "ensemble: realisation %d slice is not its own", r);
                tb_pt &p = dst[k][at];
                if (q < 3) p.xx[q] = col[j];
                else if (q < 6) p.xv[q - 3] = col[j];
                else p.ff[q - 6] = col[j];
            }
        }
        const std::vector<tb_pt> &src = k == 0 ? E->pe[r] : E->pi[r];
        for (int j = 0; j < n; j++) {
            dst[k][j].id = src[j].id;
            dst[k][j].pad_ = 0;
        }
    }
    return TB_OK;
}

int tb_ensemble_force(tb_ensemble *E, double *out)
{
    if (!E) return fail(TB_ERR_ARG, "null ensemble");
    HIPCHK(hipSetDevice(E->C->cfg.device));
    int rc;
    if (E->dirty && (rc = ens_upload(E))) return rc;
    if ((rc = ens_build(E))) return rc;
    if ((rc = ens_force(E))) return rc;
    E->have_force = true;
    std::vector<double> o(6 * (size_t)E->nreal);
    if ((rc = ens_finish(E, o.data(), false))) return rc;
    if (out)
        for (int r = 0; r < E->nreal; r++) {
            out[2 * r] = o[6 * r];
            out[2 * r + 1] = o[6 * r + 1];
        }
    return TB_OK;
}

int tb_ensemble_step(tb_ensemble *E, double dt, int32_t nsteps, double *out)
{
    if (!E || nsteps < 1) return fail(TB_ERR_ARG, "bad argument");
    if (!E->have_force || E->dirty)
        return fail(TB_ERR_STATE, "tb_ensemble_step needs forces: call tb_ensemble_force first");
    tb_handle *C = E->C;
    HIPCHK(hipSetDevice(C->cfg.device));
    int rc;
    for (int k = 0; k < 2; k++) tb_vv1(C->s, C->sp[k], C->cfg.xm[k], dt);
    for (int k = 0; k < nsteps; k++) {
        if ((rc = ens_build(E))) return rc;
        if ((rc = ens_force(E))) return rc;
        const bool last = k == nsteps - 1;
        if ((rc = vv2_all(C, dt, 0, !last))) return rc;
    }
    C->built = false;
    return ens_finish(E, out, out != nullptr);
}

int tb_ensemble_autotune(tb_ensemble *E, int32_t *group_out, float *ms_out)
{
    if (!E || !E->have_force) return fail(TB_ERR_STATE, "tb_ensemble_autotune needs tb_ensemble_force first");
    tb_handle *C = E->C;
    HIPCHK(hipSetDevice(C->cfg.device));
    float best = 1e30f;
    int best_tpl = 2, best_lv = 0;
    const int ntot = C->sp[0].n + C->sp[1].n;
    int rc;
    float tab[3][TB_LIST_VARIANTS];
    for (int tpl = 1; tpl <= 4; tpl *= 2) {
        if ((rc = ens_regroup(E, tpl))) return rc;
        for (int lv = 0; lv < TB_LIST_VARIANTS; lv++) {
            C->mx.lists = lv;
            float ms = 0.0f;
            for (int rep = 0; rep < 3; rep++) {
                HIPCHK(hipEventRecord(C->ev[4], C->s));
                if ((rc = ens_force(E))) return rc;
                HIPCHK(hipEventRecord(C->ev[5], C->s));
                HIPCHK(hipEventSynchronize(C->ev[5]));
                float t = 0.0f;
                HIPCHK(hipEventElapsedTime(&t, C->ev[4], C->ev[5]));
                if (rep > 0) ms += t;  // the first run warms caches
            }
            ms *= 0.5f;
            tab[tpl == 1 ? 0 : (tpl == 2 ? 1 : 2)][lv] = ms;
            if (ms < best) {
                best = ms;
                best_tpl = tpl;
                best_lv = lv;
            }
        }
    }
    if (ms_out)
        for (int g = 0; g < 3; g++) ms_out[g] = tab[g][best_lv];
    tb_autotune_record(ntot / E->nreal, best_tpl);
    tb_autotune_record_lists(ntot / E->nreal, best_lv);
    C->mx.lists = best_lv;
    if ((rc = ens_regroup(E, tb_mixed_tpl(C->cfg, ntot / E->nreal)))) return rc;
    // the forces on the host side belong to the last timed launch, same trees
    if ((rc = ens_force(E))) return rc;
    if (group_out) *group_out = 64 * best_tpl;
    return TB_OK;
}

int tb_walk_lists(tb_handle *h, int32_t *flush_m, int32_t *flush_p)
{
    if (!h || !flush_m || !flush_p) return fail(TB_ERR_ARG, "null argument");
    int m, p;
    tb_list_lengths(h->mx.lists, &m, &p);
    *flush_m = m;
    *flush_p = p;
    return TB_OK;
}

int tb_ensemble_walk_lists(tb_ensemble *E, int32_t *flush_m, int32_t *flush_p)
{
    if (!E) return fail(TB_ERR_ARG, "null ensemble");
    return tb_walk_lists(E->C, flush_m, flush_p);
}

// ------------------------------------------------------------------ RDF

int tb_rdf_enable(tb_handle *h, double dx_ei, double dx_ii, int32_t nbin)
{
    if (!h) return fail(TB_ERR_ARG, "null handle");
    if (nbin == 0) {
        HIPCHK(hipSetDevice(h->cfg.device));
        HIPCHK(hipStreamSynchronize(h->s));
        dfree(h, h->rdf.hist, 0);
        h->rdf = TbRdf();
        h->rdf_on = false;
        return TB_OK;
    }
    if (!(dx_ei > 0.0) || !(dx_ii > 0.0) || nbin < 1)
        return fail(TB_ERR_ARG, "RDF needs dx_ei > 0, dx_ii > 0 and nbin >= 1");
    if (h->cfg.walk == TB_WALK_MIXED && nbin > TB_RDF32_MAXBIN)
        return fail(TB_ERR_ARG, "the mixed walk's RDF holds at most %d bins (got %d)", TB_RDF32_MAXBIN, nbin);
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipStreamSynchronize(h->s));
    dfree(h, h->rdf.hist, 0);
    h->rdf = TbRdf();
    h->rdf_on = false;
    int rc = dalloc_t(h, &h->rdf.hist, (size_t)2 * nbin);
    if (rc) return fail(rc, "RDF allocation failed");
    HIPCHK(hipMemsetAsync(h->rdf.hist, 0, (size_t)2 * nbin * sizeof(int64_t), h->s));
    HIPCHK(hipMemsetAsync(h->d_misc + MISC_RDF, 0, sizeof(int32_t), h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->rdf.dx[0] = dx_ei;
    h->rdf.dx[1] = dx_ii;
    h->rdf.nbin = nbin;
    h->rdf_on = true;
    return TB_OK;
}

int tb_rdf_get(tb_handle *h, int32_t nbin, int64_t *rdf_ei, int64_t *rdf_ii, int32_t reset)
{
    if (!h || !rdf_ei || !rdf_ii) return fail(TB_ERR_ARG, "null argument");
    if (!h->rdf_on) return fail(TB_ERR_STATE, "tb_rdf_get before tb_rdf_enable");
    if (nbin != h->rdf.nbin) return fail(TB_ERR_ARG, "nbin %d differs from the enabled %d", nbin, h->rdf.nbin);
    HIPCHK(hipSetDevice(h->cfg.device));
    HIPCHK(hipMemcpyAsync(rdf_ei, h->rdf.hist, (size_t)nbin * 8, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipMemcpyAsync(rdf_ii, h->rdf.hist + nbin, (size_t)nbin * 8, hipMemcpyDeviceToHost, h->s));
    int32_t flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, h->d_misc + MISC_RDF, sizeof(int32_t), hipMemcpyDeviceToHost, h->s));
    if (reset) HIPCHK(hipMemsetAsync(h->rdf.hist, 0, (size_t)2 * nbin * sizeof(int64_t), h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    if (flag) return fail(TB_ERR_HIP, "mixed-walk RDF: pair search stack overflow or malformed table (%d)", flag);
    return TB_OK;
}

int tb_get_stats(tb_handle *h, tb_stats *st)
{
    if (!h || !st) return fail(TB_ERR_ARG, "null argument");
    *st = h->st;
    return TB_OK;
}

}  // extern "C"
