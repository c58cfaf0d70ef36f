// tb_mixed.h -- the mixed-precision wavefront-group walk (walk32.hip).
#pragma once
#include "tb_internal.h"

struct MixedDev {
    int ready = 0;
};

int tb_mixed_prep(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, size_t *bytes);
int tb_mixed_force(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums);
int tb_mixed_count(hipStream_t s, MixedDev &M, SpeciesDev *sp, const tb_config &cfg, double *epart, double *d_sums,
                   int64_t *d_cnt);
void tb_mixed_free(MixedDev &M, size_t *bytes);
const char *tb_mixed_error();
