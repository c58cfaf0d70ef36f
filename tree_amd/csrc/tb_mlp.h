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
    struct TbMlpPack *Ppk = nullptr;  // the same, packed as matrix-core operands
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

// ---- batched inference on the matrix cores (the walk's band decisions)
//
// 32 samples per call, one wavefront, BF16 operands (weights already
// bf16-rounded; features and the first hidden layer rounded to bf16) with FP32
// accumulation, as transposed products so that no data moves between layers:
//   H1^T (32 hidden x 32 samples) = W1^T (32 x 8) . X^T (8 x 32)      1 MFMA
//   H2^T (32 x 32)                = W2^T (32 x 32) . H1^T (32 x 32)   4 MFMAs
// with v_mfma_f32_32x32x8_bf16 layouts: A[i][k] in lane i + 32 (k / 4),
// element k % 4; B[k][j] in lane j + 32 (k / 4), element k % 4; D[i][j] in
// lane j + 32 ((i / 4) % 2), register (i % 4) + 4 (i / 8).  Layer 1's
// registers 4c .. 4c+3 are then exactly layer 2's B operand for K-chunk c.
typedef short tb_v4s __attribute__((ext_vector_type(4)));
typedef float tb_v16f __attribute__((ext_vector_type(16)));

__device__ inline short tb_bf16_bits(float f) { return (short)(__float_as_uint(tb_bf16_round(f)) >> 16); }

// The bf16 parameters laid out as the matrix-core operands need them, so each
// lane loads its operands with a few vector loads and no conversion (built by
// tb_mlp_refresh_bf16 whenever the parameters change).
struct TbMlpPack {
    tb_v4s a1[64];                  // lane l: W1[4 (l / 32) + q][l % 32]
    tb_v4s a2[64][4];               // lane l, K-chunk c: W2[8 c + 4 (l / 32) + q][l % 32]
    float b1[2][16], b2[2][16], w3[2][16];  // half h, register v: unit 8 (v / 4) + 4 h + v % 4
    float b3;
};

// lane l supplies features 4 (l / 32) .. 4 (l / 32) + 3 of sample l % 32 in
// xq and receives the output for sample l % 32 (lanes l and l ^ 32 both).
// All 64 lanes must call it together.
__device__ inline float tb_mlp_eval32_mfma(const TbMlpPack *__restrict__ K, const float (&xq)[4])
{
    const int l = threadIdx.x & 63, h = l >> 5;
    tb_v4s b;
    tb_v16f c, d;
#pragma unroll
    for (int q = 0; q < 4; q++) b[q] = tb_bf16_bits(xq[q]);  // X[sample][k]
#pragma unroll
    for (int v = 0; v < 16; v++) {
        c[v] = K->b1[h][v];
        d[v] = K->b2[h][v];
    }
    c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(K->a1[l], b, c, 0, 0, 0);
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
#pragma unroll
        for (int q = 0; q < 4; q++) b[q] = tb_bf16_bits(fmaxf(c[4 * cc + q], 0.0f));  // H1^T[in][sample]
        d = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(K->a2[l][cc], b, d, 0, 0, 0);
    }
    float y = 0.0f;
#pragma unroll
    for (int v = 0; v < 16; v++) y += fmaxf(d[v], 0.0f) * K->w3[h][v];
    y += __shfl_xor(y, 32, 64);
    return y + K->b3;
}

int tb_mlp_alloc(MlpDev &M);
void tb_mlp_free(MlpDev &M);
int tb_mlp_upload(hipStream_t s, MlpDev &M, const float *params);
void tb_mlp_refresh_bf16(hipStream_t s, MlpDev &M);
int tb_mlp_loss_grad(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, float *loss, float *grad);
int tb_mlp_forward(hipStream_t s, MlpDev &M, int n, const float *X, int bf16, float *out);
int tb_mlp_train(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, int steps, int batch, float lr,
                 uint64_t seed, float *loss_hist);
