// tb_mixed.h -- the mixed-precision wavefront-group walk (walk32.hip).
//
// Tree records in breadth-first order (children of a cell contiguous):
//   bmac  double4  centroid (FP64), rcrit = max(lx)/theta_mixed + dx
//   binfo uint32   leaf: TB_LEAF | particle index; twig: first child << 3 | (nchild - 1)
//   bq0   float4   q, Q'xx, Q'yy, Q'zz   (traceless quadrupole 3M - tr(M) I of the
//   bq1   float4   Q'xy, Q'yz, Q'zx, 0    exact second moment M about the centroid)
#pragma once
#include "tb_internal.h"

struct MixedTree {
    int nn = 0, cap = 0;
    double4 *bmac = nullptr;
    uint32_t *binfo = nullptr;
    float4 *bq0 = nullptr, *bq1 = nullptr;
    uint32_t *inv = nullptr;
};

struct MixedDev {
    MixedTree t[2];
    int32_t *ctr = nullptr;       // work-queue counter (+ error flag)
    double *part = nullptr;       // per-wave energy partials, 4 per wave
    int part_n = 0;
    int tpl = 0;                  // targets per lane used by the last launch
    // MLP gate for the learned opening criterion (bf16 weights + fp32 biases)
    void *mlp = nullptr;
    int mlp_ready = 0;
};

int tb_mixed_prep(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, size_t *bytes);
int tb_mixed_force(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums);
int tb_mixed_count(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums,
                   int64_t *d_cnt);
void tb_mixed_free(MixedDev &M, size_t *bytes);
const char *tb_mixed_error();
