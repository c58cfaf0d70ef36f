// tb_mixed.h -- the mixed-precision wavefront-group walk (walk32.hip) over
// the joint tree of both species (JointDev, build.hip).
//
// Tree records in breadth-first order (children of a cell contiguous):
//   bmac  double4  expansion centre (FP64; a leaf: its particle), rcrit = max(lx)/theta_mixed
//   binfo uint32   leaf: TB_LEAF | joint particle index; twig: first child << 3 | (nchild - 1)
//   bq0   float4   q, Dx, Dy, Dz          net charge and dipole about the centre
//   bq1   float4   Q'xx, Q'yy, Q'xy, Q'yz traceless quadrupole 3M - tr(M) I
//   bq2   float4   Q'zx, s, count, log2(s)/8      (Q'zz = -Q'xx - Q'yy)
//   bf    float4   the learned criterion's per-cell features (log2 / 8 of
//                  |q|/count, |D|/(count s), |Q'|/(count s^2), count)
//   bq3   float4   O'xxx, O'xxy, O'xxz, O'xyy   traceless third moment
//   bq4   float4   O'xyz, O'yyy, O'yyz, H'xxxx  (the other O' follow from tr = 0)
//   bq5   float4   H'xxxy, H'xxxz, H'xxyy, H'xxyz   traceless fourth moment
//   bq6   float4   H'xyyy, H'xyyz, H'yyyy, H'yyyz   (the others follow from tr = 0)
#pragma once
#include "tb_internal.h"
#include "tb_mlp.h"

struct MixedTree {
    int nn = 0, cap = 0;
    double4 *bmac = nullptr;
    uint32_t *binfo = nullptr;
    float4 *bq0 = nullptr, *bq1 = nullptr, *bq2 = nullptr, *bf = nullptr;
    float4 *bq3 = nullptr, *bq4 = nullptr;   // traceless third moment (and H'xxxx)
    float4 *bq5 = nullptr, *bq6 = nullptr;   // traceless fourth moment
    uint32_t *inv = nullptr;
    // the upward pass's breadth-first companions (CellRecs)
    uint8_t *bflag = nullptr;
    uint64_t *bkey = nullptr;
    int32_t *bcnt = nullptr;
    double *bnq = nullptr;
};

// a (cell, target group) pair of the learned criterion's band, with its features
struct BandRec {
    uint32_t cell, group, tree, pad;
    float f[TB_MLP_NIN];
};

// one realisation (joint tree and particle state) of a walk launch
struct Domain {
    const double *px[2], *py[2], *pz[2];
    double *fx[2], *fy[2], *fz[2];
    const int2 *grp[2];   // target groups (start, length) per species
    int n[2], ng[2];
    const double4 *bmac;
    const uint32_t *binfo;
    const float4 *bq0, *bq1, *bq2, *bf, *bq3, *bq4, *bq5, *bq6;
    float inv_rs, rs;
    float close2;     // (22 rs)^2: unlike pairs closer than this take the Kelbg term
    int gbase;        // first global group index of this domain
    int nn;           // cells in the tree
    uint32_t root;    // the domain's root cell (a batched ensemble's trees share one table)
    int gfirst[2];    // combined index of the domain's first group per species (grp[k] = combined + gfirst[k])
    const int32_t *gsid[2];  // combined: segment of each group
};

// list-length variants of the walk (walk32.hip LIST_M / LIST_P)
#define TB_LIST_VARIANTS 3

struct MixedDev {
    MixedTree t[1];
    int lists = 0;                // list-length variant of the next walks
    int32_t *ctr = nullptr;       // 16 ints: error flags, band pair count, work-queue chunk counters
    int tpl = 0;                  // targets per lane used by the last launch
    hipEvent_t ev[2] = {};        // around the last walk dispatch
    Domain *ddom = nullptr;       // device copy of the launch's domain descriptors
    int ddom_cap = 0;
    double *esum = nullptr;       // 4 energy sums (ee, ei, ie, ii) per domain
    int esum_cap = 0;
    // far cells of the last walk: chunk lists per target group (k_far_eval)
    uint32_t *far_cells = nullptr;
    int32_t *far_link = nullptr, *far_head = nullptr;
    // two-level far field, per segment: frontier and far-chunk heads, centre
    // and half-diagonal, local expansion coefficients
    int32_t *sfront = nullptr, *sfar = nullptr;
    double4 *scen = nullptr;
    float *sloc = nullptr;
    int seg_cap = 0;
    float *gemin = nullptr;       // per target group: the learned walk's emin (kept for a re-run)
    int far_cap = 0, head_cap = 0;
    // the learned opening criterion
    MlpDev mlp;
    BandRec *rec = nullptr;       // collected band pairs
    float *lab = nullptr;         // their labels
    int rec_cap = 0;
};

// the cell records for the joint tree *jt: allocated and their error flag
// cleared; tb_jupward with tb_mixed_recs then fills them
int tb_mixed_prep(hipStream_t s, MixedDev &M, const JointDev *jt, size_t *bytes);
CellRecs tb_mixed_recs(MixedDev &M, const tb_config &cfg);
int tb_mixed_force(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums);
int tb_mixed_count(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums,
                   int64_t *d_cnt);
void tb_mixed_free(MixedDev &M, size_t *bytes);
// run the conservative walk (theta_lo) recording band pairs (theta_lo .. theta_mixed),
// then label them with the FP64 error of accepting the cell; returns the pair count
int tb_mixed_collect(hipStream_t s, MixedDev &M, SpeciesDev *sp, const JointDev *jt, const tb_config &cfg,
                     double *d_sums, int cap,
                     float *X, float *y, int *nout);
const char *tb_mixed_error();
// duration of the last walk dispatch (after the stream has synchronised); -1 if none
float tb_mixed_walk_ms(MixedDev &M);

// one persistent launch over the realisations of a batched ensemble, all in
// one combined state: species arrays sp[2] (realisation r's targets in groups
// [gfirst[k][r], gfirst[k][r] + gcount[k][r]) of species k) and one table of
// cells (r's root at roots[r]); rs[r] its Kelbg length; fills esum[4 r]
int tb_mixed_force_ens(hipStream_t s, MixedDev &M, SpeciesDev *sp, int nreal, const uint32_t *roots,
                       const double *rs, const int32_t *const gfirst[2], const int32_t *const gcount[2],
                       const tb_config &cfg, double *esum_out_dev);

// rdf32.hip: the mixed walk's RDF, every pair within nbin dx of an ion target
// (nbin <= TB_RDF32_MAXBIN) added to R.hist; flag gets 1 (stack) / 2 (table)
#define TB_RDF32_MAXBIN 10240
int tb_rdf_pairs(hipStream_t s, const MixedTree &T, const SpeciesDev &I, const TbRdf &R, int32_t *flag);

// walk-group size chosen by the autotuner for a problem size (targets per lane: 1, 2 or 4)
int tb_autotune_tpl(int n_total);
// targets per lane for a configuration: its group size, else the autotuner's choice
int tb_mixed_tpl(const tb_config &cfg, int n_total);
void tb_autotune_record(int n_total, int tpl);
// list-length variant chosen by the autotuner for a problem size, its record, and its lengths
int tb_autotune_lists(int n_total);
void tb_autotune_record_lists(int n_total, int variant);
void tb_list_lengths(int variant, int *m, int *p);
