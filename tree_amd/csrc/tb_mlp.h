// tb_mlp.h -- the opening-criterion MLP (mlp.hip): device evaluation shared
// with the walk, training entry points used by the C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/treemi.h"

#define TB_MLP_NIN 8
#define TB_MLP_NH 32
#define TB_MLP_NPAR (TB_MLP_NIN * TB_MLP_NH + TB_MLP_NH + TB_MLP_NH * TB_MLP_NH + TB_MLP_NH + TB_MLP_NH + 1)

struct MlpDev {
    void *base = nullptr;
    float *P = nullptr;     // fp32 master parameters
    float *G = nullptr;     // gradient (+ loss at [NPAR])
    float *m = nullptr, *v = nullptr;  // Adam moments
    float *Pbf = nullptr;   // parameters rounded to bf16 (stored as fp32) for inference
    float *part = nullptr;  // per-block gradient partials
    void *data = nullptr;   // uploaded samples
    size_t data_bytes = 0;
    int ready = 0;          // trained or set
};

// round to the nearest bfloat16 (ties to even), kept in a float
__device__ inline float tb_bf16_round(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// forward pass with the parameter vector P (fp32 or bf16-rounded values)
__device__ inline float tb_mlp_eval(const float *P, const float *x)
{
    constexpr int NIN = TB_MLP_NIN, NH = TB_MLP_NH;
    constexpr int O_B1 = NIN * NH, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH, O_B3 = O_W3 + NH;
    float h1[NH];
#pragma unroll
    for (int j = 0; j < NH; j++) {
        float a = P[O_B1 + j];
#pragma unroll
        for (int k = 0; k < NIN; k++) a += x[k] * P[k * NH + j];
        h1[j] = fmaxf(a, 0.0f);
    }
    float y = P[O_B3];
    for (int j = 0; j < NH; j++) {
        float a = P[O_B2 + j];
#pragma unroll
        for (int k = 0; k < NH; k++) a += h1[k] * P[O_W2 + k * NH + j];
        y += fmaxf(a, 0.0f) * P[O_W3 + j];
    }
    return y;
}

int tb_mlp_alloc(MlpDev &M);
void tb_mlp_free(MlpDev &M);
int tb_mlp_upload(hipStream_t s, MlpDev &M, const float *params);
void tb_mlp_refresh_bf16(hipStream_t s, MlpDev &M);
int tb_mlp_loss_grad(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, float *loss, float *grad);
int tb_mlp_forward(hipStream_t s, MlpDev &M, int n, const float *X, int bf16, float *out);
int tb_mlp_train(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, int steps, int batch, float lr,
                 uint64_t seed, float *loss_hist);
