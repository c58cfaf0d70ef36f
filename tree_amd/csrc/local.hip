// local.hip -- a direct check of the far-field local expansion
// (local_exp.h) outside the walk: the local coefficients of a set of cells
// about one centre, summed across a wavefront the way the walk sums them,
// then evaluated at given offsets.  Order 3 keeps every coefficient and the
// moments up to the quadrupole; orders 4 to 6 (6 is the walk's) keep the
// independent coefficients of the harmonic expansion and add the third
// moment, and order 6 the fourth.
#include <hip/hip_runtime.h>

#include "local_exp.h"
#include "tb_internal.h"

namespace {

template <int N>
__device__ inline void wave_sum_all(float (&L)[N])
{
#pragma unroll
    for (int k = 0; k < N; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) L[k] += __shfl_xor(L[k], o, 64);
}

// cells: TB_LOCAL_CELL (28) floats each, R = centre - cell centre (3), q,
// D (3), Q'xx, Q'yy, Q'xy, Q'yz, Q'zx, then the traceless third moment O'xxx,
// O'xxy, O'xxz, O'xyy, O'xyz, O'yyy, O'yyz (orders 4 to 6) and the traceless
// fourth moment H'xxxx, H'xxxy, H'xxxz, H'xxyy, H'xxyz, H'xyyy, H'xyyz,
// H'yyyy, H'yyyz (order 6).  out: phi, Ex, Ey, Ez per offset.
template <int P>
__global__ __launch_bounds__(64) void k_local_eval(int ncell, const float *__restrict__ cells, int ny,
                                                   const float *__restrict__ y, float *__restrict__ out)
{
    constexpr int N = P == 3 ? TB_LOCAL_N3 : P == 4 ? TB_LOCAL_N4H : P == 5 ? TB_LOCAL_N5H : TB_LOCAL_N6H;
    float L[N];
#pragma unroll
    for (int k = 0; k < N; k++) L[k] = 0.0f;
    for (int c = threadIdx.x; c < ncell; c += 64) {
        const float *e = cells + TB_LOCAL_CELL * c;
        if constexpr (P == 3)
            m2l_acc_p3(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10], e[11], L);
        else if constexpr (P == 4)
            m2l_acc_p4ho(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10], e[11], e[12], e[13],
                         e[14], e[15], e[16], e[17], e[18], L);
        else if constexpr (P == 5)
            m2l_acc_p5ho(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10], e[11], e[12], e[13],
                         e[14], e[15], e[16], e[17], e[18], L);
        else
            m2l_acc_p6hxt7(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], e[10], e[11], e[12], e[13],
                           e[14], e[15], e[16], e[17], e[18], e[19], e[20], e[21], e[22], e[23], e[24], e[25], e[26],
                           e[27], L);
    }
    wave_sum_all(L);
    for (int j = threadIdx.x; j < ny; j += 64) {
        float phi, ex, ey, ez;
        if constexpr (P == 3)
            l2p_p3(L, y[3 * j], y[3 * j + 1], y[3 * j + 2], phi, ex, ey, ez);
        else if constexpr (P == 4)
            l2p_p4h(L, y[3 * j], y[3 * j + 1], y[3 * j + 2], phi, ex, ey, ez);
        else if constexpr (P == 5)
            l2p_p5h(L, y[3 * j], y[3 * j + 1], y[3 * j + 2], phi, ex, ey, ez);
        else
            l2p_p6h(L, y[3 * j], y[3 * j + 1], y[3 * j + 2], phi, ex, ey, ez);
        out[4 * j] = phi;
        out[4 * j + 1] = ex;
        out[4 * j + 2] = ey;
        out[4 * j + 3] = ez;
    }
}

}  // namespace

int tb_local_run(hipStream_t s, int order, int ncell, const float *cells, int ny, const float *y, float *out)
{
    if (order == 3)
        hipLaunchKernelGGL(k_local_eval<3>, dim3(1), dim3(64), 0, s, ncell, cells, ny, y, out);
    else if (order == 4)
        hipLaunchKernelGGL(k_local_eval<4>, dim3(1), dim3(64), 0, s, ncell, cells, ny, y, out);
    else if (order == 5)
        hipLaunchKernelGGL(k_local_eval<5>, dim3(1), dim3(64), 0, s, ncell, cells, ny, y, out);
    else if (order == 6)
        hipLaunchKernelGGL(k_local_eval<6>, dim3(1), dim3(64), 0, s, ncell, cells, ny, y, out);
    else
        return TB_ERR_ARG;
    return hipGetLastError() == hipSuccess ? TB_OK : TB_ERR_HIP;
}
