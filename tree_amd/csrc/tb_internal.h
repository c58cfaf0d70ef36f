// tb_internal.h -- device data layout and launcher declarations shared by the
// HIP translation units of libtreemi.so.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/treemi.h"

// Key depth: 21 octal digits in bits 62..0 of a 64-bit key.
#define TB_LEVELS 21
// Leaf marker in the preorder skip-pointer array.
#define TB_LEAF 0x80000000u
// Interaction counters of a counting pass (int64 each).
#define TB_NCNT 32
// Coulomb constant eV*um (data.f90:40).
#define TB_EPS 0.0014399835661145824

// Per-species device buffers.  Particle state is SoA FP64 kept in the key
// order of the last build (double-buffered for the permutation); orig[] maps
// back to the order the caller handed in.
struct SpeciesDev {
    int n = 0;
    int cur = 0;
    double *x[2][3] = {}, *v[2][3] = {}, *f[2][3] = {};
    int32_t *orig[2] = {};
    uint64_t *key[2] = {};
    uint32_t *val[2] = {};
    int32_t *cnt = nullptr, *offs = nullptr;   // nodes started per particle, exclusive scan
    double *root = nullptr;                    // c0[3], L0[3]
    // node table in preorder (reference DFS order)
    int nn = 0, cap = 0, maxdepth = 0;
    double4 *nmac = nullptr;    // centroid xyz, rcrit = max(lx)/theta + dx (leaf: -1)
    uint32_t *nnext = nullptr;  // preorder index after the subtree | TB_LEAF
    double *nmom = nullptr;     // [nn][10] qeff, D(3), Q(3), QQ(3)  (child-point, tree.f90:537-551)
    double *ndx = nullptr;      // centroid offset from the geometric centre
    int32_t *nstart = nullptr, *ncount = nullptr, *ndepth = nullptr;
    uint32_t *dval[2] = {};     // node ids in breadth-first order (dval[0])
    uint32_t *bfs = nullptr;    // preorder ids grouped by depth, preorder within a depth
    int lvl_off[TB_LEVELS + 3] = {};
    int twig_hist[TB_LEVELS + 2] = {};
    // breadth-first companions of the upward pass, so that a cell reads its
    // children (a contiguous run of the next level) without chasing pointers:
    // preorder -> breadth-first index; per breadth-first node bit 0 = first
    // child of its parent, bit 1 = leaf; per cell, first child << 3 |
    // children - 1; centroid and qeff
    uint32_t *binv = nullptr;
    uint8_t *bflag = nullptr;
    uint32_t *bkids = nullptr;
    uint64_t *bkey = nullptr;   // a cell's first particle's key (its centre replays from it)
    double4 *bxq = nullptr;
    // mixed walk: target groups of at most gG consecutive key-ordered
    // particles, each inside one cell of at most 8 gG particles (tb_groups)
    const uint64_t *skey = nullptr;  // this build's sorted keys
    int8_t *gpre = nullptr, *gsuf = nullptr;  // windowed lcp maxima
    int32_t *gseg = nullptr;    // segment starts
    int2 *grp = nullptr;        // group start, length
    int ng = 0, gG = 0;         // groups, and the size bound they were cut for
    // segments (cells of at most 8 gG particles, each split into <= 8 groups):
    // first group and group count per segment, segment of each group
    int2 *sgrp = nullptr;
    int32_t *gsid = nullptr;
    int nseg = 0;
};

// The joint tree over both species (mixed walk): nodes in preorder, expansion
// about each cell's geometric centre (a leaf: its particle).
// joint-tree moments per cell: net charge, dipole, second moment (xx yy zz
// xy yz zx) and third moment (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz)
constexpr int TB_JQ = 35;

struct JointDev {
    int n = 0;                  // ne + ni
    uint64_t *key[2] = {};
    uint32_t *val[2] = {};      // species << 31 | index in the species' key order
    double4 *pos = nullptr;     // sorted: x, y, z, charge
    int32_t *cnt = nullptr, *offs = nullptr;
    double *root = nullptr;     // joint box c0[3], L0[3]
    int nn = 0, cap = 0, maxdepth = 0;
    double4 *nc = nullptr;      // leaves: the particle's position and charge (cells: unused)
    uint32_t *nnext = nullptr;
    int32_t *nstart = nullptr, *ncount = nullptr, *ndepth = nullptr;
    uint32_t *dval[2] = {};
    uint32_t *bfs = nullptr;
    int lvl_off[TB_LEVELS + 3] = {};
    int twig_hist[TB_LEVELS + 2] = {};
};

// scratch shared by a handle's launches
struct Scratch {
    void *p = nullptr;
    size_t bytes = 0;
};

// ----- build.hip
size_t tb_sort_temp_bytes(int n);
void tb_bbox(hipStream_t s, int n, const double *x, const double *y, const double *z, double *root, double *tmp);
void tb_keys(hipStream_t s, int n, const double *x, const double *y, const double *z, const double *root,
             uint64_t *key, uint32_t *val, int32_t *err);
// sorts key[0]/val[0] by bits [lo_bit, hi_bit); result ends in key[*which]/val[*which]
void tb_radix_sort(hipStream_t s, int n, uint64_t *key[2], uint32_t *val[2], int lo_bit, int hi_bit, void *tmp,
                   int *which, bool init_iota);
// two independent sorts (key sets k = 0, 1 with their own temporaries), their
// passes enqueued alternately on two streams so both run from the start
void tb_radix_sort2(const hipStream_t s[2], const int n[2], uint64_t **key[2], uint32_t **val[2], int lo_bit,
                    int hi_bit, void *const tmp[2], int which[2], bool init_iota);
// breadth-first order: node ids stably sorted by depth (one pass over ndepth),
// and, if inv is given, its inverse
void tb_depth_order(hipStream_t s, int n, const int32_t *depth, uint32_t *order, uint32_t *inv, void *tmp);
void tb_permute(hipStream_t s, SpeciesDev &S, const uint32_t *perm, int32_t *err);
// joint: the mixed walk's joint tree (up to 8 particles may share a full key)
void tb_lcp(hipStream_t s, int n, const uint64_t *key, int32_t *cnt, int32_t *hist /*2*(TB_LEVELS+2)*/,
            int32_t *err, bool joint = false);
void tb_scan_excl(hipStream_t s, int n, const int32_t *in, int32_t *out, int32_t *total, void *tmp);
size_t tb_scan_temp_bytes(int n);
void tb_emit(hipStream_t s, SpeciesDev &S, const uint64_t *key, int sign);
// the upward pass: the breadth-first companions, then each level (deepest first)
void tb_upward_prep(hipStream_t s, SpeciesDev &S, const uint64_t *key, int sign, int32_t *err);
void tb_upward_level(hipStream_t s, SpeciesDev &S, int d, double theta, int32_t *err);
// sign: the species' charge (-1 electrons, +1 ions), a leaf's qeff
void tb_export_gather(hipStream_t s, const SpeciesDev &S, int sign, int32_t *depth, int32_t *leaf, int32_t *pt,
                      int32_t *count, int32_t *next, double *xx, double *qeff, double *dx, double *D, double *Q,
                      double *QQ);

void tb_jbbox(hipStream_t s, int ne, const double *ex, const double *ey, const double *ez, int ni, const double *ix,
              const double *iy, const double *iz, double *root, double *tmp);
// the joint key order as a merge of the two species' sorted keys (formed over
// the joint box); payload: species << 31 | index in the species' key order
void tb_jmerge(hipStream_t s, int ne, int ni, const uint64_t *ke, const uint64_t *ki, uint64_t *kout, uint32_t *vout);
void tb_jgather(hipStream_t s, int n, const uint32_t *val, const double *const x[2], const double *const y[2],
                const double *const z[2], double4 *pos);
void tb_jemit(hipStream_t s, JointDev &J, const uint64_t *key);
// The mixed walk's breadth-first cell records, which the joint tree's upward
// pass fills, and the breadth-first companions it works from.
struct CellRecs {
    uint32_t *inv = nullptr;    // preorder node -> breadth-first index
    double4 *bmac = nullptr;    // centre (a leaf: its particle), s / theta (a leaf: 0)
    uint32_t *binfo = nullptr;  // first child << 3 | children - 1, or leaf | joint particle
    float4 *bq0 = nullptr, *bq1 = nullptr, *bq2 = nullptr, *bq3 = nullptr, *bq4 = nullptr, *bf = nullptr;
    float4 *bq5 = nullptr, *bq6 = nullptr;
    uint8_t *bflag = nullptr;   // bit 0: first child of its parent; bit 1: leaf
    uint64_t *bkey = nullptr;   // a cell's first particle's key (its centre replays from it)
    int32_t *bcnt = nullptr;    // a cell's particle count
    double *bnq = nullptr;      // [nn][TB_JQ] a cell's FP64 q, D(3), M(6) (xx yy zz xy yz zx),
                                // O(10), H(15) (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz xyzz xzzz
                                // yyyy yyyz yyzz yzzz zzzz) about its centre
    double theta = 0.0;
    int32_t *flag = nullptr;    // 16: malformed node table
};
// T > 0: a batched ensemble's joint tree (keys tagged by tb_ekeys), roots [nreal][6]
void tb_jupward(hipStream_t s, JointDev &J, const uint64_t *key, const CellRecs &R, int T = 0,
                const double *roots = nullptr);
// target groups of at most G particles over S.skey; totals[0] = segments,
// totals[1] = groups (device); uses S.cnt and S.offs as scratch.  A batched
// ensemble (keys tagged with T digits) never groups across realisations, and
// first[r] (device) receives realisation r's first group.
void tb_groups(hipStream_t s, SpeciesDev &S, int G, int32_t *totals, void *scan_tmp, int T = 0,
               int32_t *first = nullptr);
// batched ensemble: realisation r holds [off[r], off[r+1]) of a combined array
int tb_sbbox_chunks(int n_max);
void tb_sbbox(hipStream_t s, int nreal, int chunks, const int32_t *off, const double *x, const double *y,
              const double *z, double *part /*[nreal][chunks][6]*/);
void tb_sbbox_final(hipStream_t s, int nreal, int chunks, const double *pa, const double *pb, double *roots);
void tb_ekeys(hipStream_t s, int n, const double *x, const double *y, const double *z, const int32_t *off,
              int nreal, const double *roots, int T, uint64_t *key, uint32_t *val, uint32_t tag, int32_t *err);

// ----- walk64.hip
// RDF accumulation of the FP64 walk: hist = [2][nbin] int64 (e-i, i-i) on the device
struct TbRdf {
    int64_t *hist = nullptr;
    double dx[2] = {0.0, 0.0};
    int nbin = 0;
};
void tb_walk64(hipStream_t s, int tsp, const SpeciesDev &T, const SpeciesDev &E, const SpeciesDev &I, double rs,
               double *epart, int64_t *counters, const TbRdf *rdf = nullptr);
void tb_reduce_partials(hipStream_t s, int nparts, int width, const double *part, double *out);
// FP64 walk for a target list (tsp/idx, key order) or, with tsp == nullptr, every stride-th target
void tb_walk64_list(hipStream_t s, int nt, const int32_t *tsp, const int32_t *idx, int stride, const SpeciesDev &E,
                    const SpeciesDev &I, double rs, double *fout);
int tb_walk64_blocks(int n);
// local.hip: the far-field local expansion of cells about one centre, evaluated at offsets
// (TB_LOCAL_CELL floats per cell)
constexpr int TB_LOCAL_CELL = 28;
int tb_local_run(hipStream_t s, int order, int ncell, const float *cells, int ny, const float *y, float *out);

void tb_direct(hipStream_t s, int nt, const int32_t *tsp, const int32_t *tidx_sorted, const SpeciesDev &E,
               const SpeciesDev &I, double rs, double *fout);
void tb_vv1(hipStream_t s, SpeciesDev &S, double xm, double dt);
void tb_vv2(hipStream_t s, SpeciesDev &S, double xm, double dt, int sample, double *part);
void tb_kdk(hipStream_t s, SpeciesDev &S, double xm, double dt, int sample, double *part);
int tb_vv_blocks(int n);
void tb_inverse_perm(hipStream_t s, int n, const int32_t *orig, int32_t *inv);
// batched ensemble: mv2 and radial sums per realisation (slices off[r] .. off[r+1]), out [nreal][2];
// part holds nreal * chunks * 2 doubles
void tb_ke_seg(hipStream_t s, SpeciesDev &S, double xm, int nreal, int chunks, const int32_t *off, double *part,
               double *out);
