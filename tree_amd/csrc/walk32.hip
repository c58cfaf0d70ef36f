// walk32.hip -- mixed-precision wavefront-group walk.
#include "tb_mixed.h"

static const char *g_mixed_err = "";

int tb_mixed_prep(hipStream_t, MixedDev &, SpeciesDev *, const tb_config &, size_t *)
{
    g_mixed_err = "mixed walk not built yet";
    return TB_ERR_ARG;
}
int tb_mixed_force(hipStream_t, MixedDev &, SpeciesDev *, const tb_config &, double *, double *)
{
    g_mixed_err = "mixed walk not built yet";
    return TB_ERR_ARG;
}
int tb_mixed_count(hipStream_t, MixedDev &, SpeciesDev *, const tb_config &, double *, double *, int64_t *)
{
    g_mixed_err = "mixed walk not built yet";
    return TB_ERR_ARG;
}
void tb_mixed_free(MixedDev &, size_t *) {}
const char *tb_mixed_error() { return g_mixed_err; }
