// mlp.hip -- the opening-criterion network: 8 inputs -> 32 -> 32 -> 1 (ReLU),
// trained on the GPU by these kernels (mean-squared error, Adam), evaluated
// with BF16 weights and FP32 accumulation.  The reference has no learned part;
// the network predicts log10 of the relative force error that accepting a
// cell for a target group would make.
//
// Parameter vector (fp32 master copy, 1377 values):
//   W1[8][32] (input-major), b1[32], W2[32][32], b2[32], w3[32], b3
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "tb_mlp.h"

namespace {

constexpr int NIN = TB_MLP_NIN, NH = TB_MLP_NH, NPAR = TB_MLP_NPAR;
constexpr int O_W1 = 0, O_B1 = NIN * NH, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH,
              O_B3 = O_W3 + NH;
constexpr int GB = 64;  // samples per gradient block (one wave)

// per-block gradient partials: forward, keep activations in LDS, then each
// lane sums its share of the 1377 parameter gradients over the block's samples
__global__ __launch_bounds__(GB) void k_mlp_grad(int B, const float *__restrict__ X, const float *__restrict__ T,
                                                 const float *__restrict__ P, float *__restrict__ part,
                                                 float inv_total)
{
    __shared__ float sP[NPAR];
    __shared__ float sX[GB][NIN];
    __shared__ float sH1[GB][NH];
    __shared__ float sH2[GB][NH];
    __shared__ float sD1[GB][NH];
    __shared__ float sD2[GB][NH];
    __shared__ float sDy[GB];
    __shared__ float sL[GB];
    const int t = threadIdx.x;
    for (int j = t; j < NPAR; j += GB) sP[j] = P[j];
    __syncthreads();
    const int s = blockIdx.x * GB + t;
    const bool act = s < B;
    float x[NIN];
    for (int k = 0; k < NIN; k++) x[k] = act ? X[(size_t)s * NIN + k] : 0.0f;
    float z1[NH], h1[NH];
    for (int j = 0; j < NH; j++) {
        float a = sP[O_B1 + j];
        for (int k = 0; k < NIN; k++) a += x[k] * sP[O_W1 + k * NH + j];
        z1[j] = a;
        h1[j] = fmaxf(a, 0.0f);
    }
    float z2[NH];
    float y = sP[O_B3];
    for (int j = 0; j < NH; j++) {
        float a = sP[O_B2 + j];
        for (int k = 0; k < NH; k++) a += h1[k] * sP[O_W2 + k * NH + j];
        z2[j] = a;
        y += fmaxf(a, 0.0f) * sP[O_W3 + j];
    }
    const float r = act ? y - T[s] : 0.0f;
    const float dy = 2.0f * r * inv_total;
    float d2[NH];
    for (int j = 0; j < NH; j++) d2[j] = z2[j] > 0.0f ? dy * sP[O_W3 + j] : 0.0f;
    for (int k = 0; k < NH; k++) {
        float a = 0.0f;
        for (int j = 0; j < NH; j++) a += d2[j] * sP[O_W2 + k * NH + j];
        sD1[t][k] = z1[k] > 0.0f ? a : 0.0f;
        sH1[t][k] = h1[k];
        sH2[t][k] = fmaxf(z2[k], 0.0f);
        sD2[t][k] = d2[k];
    }
    for (int k = 0; k < NIN; k++) sX[t][k] = x[k];
    sDy[t] = dy;
    sL[t] = r * r;
    __syncthreads();
    float *out = part + (size_t)blockIdx.x * (NPAR + 1);
    for (int j = t; j < NPAR; j += GB) {
        float g = 0.0f;
        if (j < O_B1) {
            const int k = j / NH, c = j % NH;
            for (int q = 0; q < GB; q++) g += sX[q][k] * sD1[q][c];
        } else if (j < O_W2) {
            const int c = j - O_B1;
            for (int q = 0; q < GB; q++) g += sD1[q][c];
        } else if (j < O_B2) {
            const int k = (j - O_W2) / NH, c = (j - O_W2) % NH;
            for (int q = 0; q < GB; q++) g += sH1[q][k] * sD2[q][c];
        } else if (j < O_W3) {
            const int c = j - O_B2;
            for (int q = 0; q < GB; q++) g += sD2[q][c];
        } else if (j < O_B3) {
            const int c = j - O_W3;
            for (int q = 0; q < GB; q++) g += sH2[q][c] * sDy[q];
        } else {
            for (int q = 0; q < GB; q++) g += sDy[q];
        }
        out[j] = g;
    }
    if (t == 0) {
        float l = 0.0f;
        for (int q = 0; q < GB; q++) l += sL[q];
        out[NPAR] = l;
    }
}

// deterministic sum of the block partials (fixed order)
__global__ __launch_bounds__(256) void k_mlp_reduce(int nb, const float *__restrict__ part, float *__restrict__ g,
                                                    float inv_total)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j > NPAR) return;
    float s = 0.0f;
    for (int b = 0; b < nb; b++) s += part[(size_t)b * (NPAR + 1) + j];
    g[j] = j == NPAR ? s * inv_total : s;  // g[NPAR] = mean squared error
}

// Adam (Kingma & Ba, 2015)
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= NPAR) return;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.0f - b1) * gj;
    const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    P[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_fwd(int B, const float *__restrict__ X, const float *__restrict__ P,
                                                 const TbMlpPack *__restrict__ K, int bf16, float *__restrict__ out)
{
    __shared__ float sP[NPAR];
    for (int j = threadIdx.x; j < NPAR; j += 256) sP[j] = bf16 ? tb_bf16_round(P[j]) : P[j];
    __syncthreads();
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (bf16) {
        // the walk's inference path: 64 samples per wavefront on the matrix cores
        const int l = threadIdx.x & 63, h = l >> 5;
        const int w0 = s - l;
        float pr[2];
#pragma unroll
        for (int tile = 0; tile < 2; tile++) {
            const int si = w0 + 32 * tile + (l & 31);
            float xq[4];
#pragma unroll
            for (int q = 0; q < 4; q++) xq[q] = si < B ? X[(size_t)si * NIN + 4 * h + q] : 0.0f;
            pr[tile] = tb_mlp_eval32_mfma(K, xq);
        }
        if (s < B) out[s] = l < 32 ? pr[0] : pr[1];
        return;
    }
    if (s >= B) return;
    float x[NIN];
    for (int k = 0; k < NIN; k++) x[k] = X[(size_t)s * NIN + k];
    out[s] = tb_mlp_eval(sP, x);
}

inline int nblk(long n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

// ====================================================================== host

int tb_mlp_alloc(MlpDev &M)
{
    if (M.P) return TB_OK;
    size_t bytes = (size_t)5 * (NPAR + 1) * sizeof(float) + (size_t)NPAR * sizeof(uint16_t) + 64 + 256 +
                   sizeof(TbMlpPack);
    if (hipMalloc(&M.base, bytes) != hipSuccess) return TB_ERR_NOMEM;
    float *f = (float *)M.base;
    M.P = f;
    M.G = f + (NPAR + 1);
    M.m = f + 2 * (NPAR + 1);
    M.v = f + 3 * (NPAR + 1);
    M.Pbf = f + 4 * (NPAR + 1);
    const size_t po = ((size_t)5 * (NPAR + 1) * sizeof(float) + 255) & ~(size_t)255;
    M.Ppk = (TbMlpPack *)((char *)M.base + po);
    hipMemset(M.base, 0, bytes);
    return TB_OK;
}

void tb_mlp_free(MlpDev &M)
{
    if (M.base) hipFree(M.base);
    if (M.part) hipFree(M.part);
    if (M.data) hipFree(M.data);
    M = MlpDev();
}

int tb_mlp_upload(hipStream_t s, MlpDev &M, const float *params)
{
    int rc = tb_mlp_alloc(M);
    if (rc) return rc;
    if (hipMemcpyAsync(M.P, params, NPAR * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    tb_mlp_refresh_bf16(s, M);
    M.ready = 1;
    return TB_OK;
}

namespace {
__global__ void k_bf16_copy(const float *__restrict__ P, float *__restrict__ Q)
{
    for (int j = threadIdx.x; j < NPAR; j += blockDim.x) Q[j] = tb_bf16_round(P[j]);
}

// TbMlpPack from the bf16-rounded parameters (one block of 64 threads)
__global__ void k_mlp_pack(const float *__restrict__ Pb, TbMlpPack *__restrict__ K)
{
    constexpr int O_B1 = NIN * NH, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH, O_B3 = O_W3 + NH;
    const int l = threadIdx.x, i = l & 31, h = l >> 5;
    for (int q = 0; q < 4; q++) {
        K->a1[l][q] = tb_bf16_bits(Pb[(4 * h + q) * NH + i]);
        for (int c = 0; c < 4; c++) K->a2[l][c][q] = tb_bf16_bits(Pb[O_W2 + (8 * c + 4 * h + q) * NH + i]);
    }
    if (l < 32) {
        const int hh = l >> 4, v = l & 15, r = 8 * (v >> 2) + 4 * hh + (v & 3);
        K->b1[hh][v] = Pb[O_B1 + r];
        K->b2[hh][v] = Pb[O_B2 + r];
        K->w3[hh][v] = Pb[O_W3 + r];
    }
    if (l == 0) K->b3 = Pb[O_B3];
}
}  // namespace

void tb_mlp_refresh_bf16(hipStream_t s, MlpDev &M)
{
    hipLaunchKernelGGL(k_bf16_copy, dim3(1), dim3(256), 0, s, M.P, M.Pbf);
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, s, M.Pbf, M.Ppk);
}

static int ensure_data(MlpDev &M, int n)
{
    size_t need = (size_t)n * (NIN + 2) * sizeof(float);
    if (need <= M.data_bytes) return TB_OK;
    if (M.data) hipFree(M.data);
    M.data = nullptr;
    M.data_bytes = 0;
    if (hipMalloc(&M.data, need) != hipSuccess) return TB_ERR_NOMEM;
    M.data_bytes = need;
    int nb = nblk(n, GB);
    if (M.part) hipFree(M.part);
    M.part = nullptr;
    if (hipMalloc((void **)&M.part, (size_t)nb * (NPAR + 1) * sizeof(float)) != hipSuccess) return TB_ERR_NOMEM;
    return TB_OK;
}

int tb_mlp_loss_grad_dev(hipStream_t s, MlpDev &M, int n, const float *dX, const float *dT)
{
    const int nb = nblk(n, GB);
    const float inv = 1.0f / (float)n;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(GB), 0, s, n, dX, dT, M.P, M.part, inv);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(nblk(NPAR + 1, 256)), dim3(256), 0, s, nb, M.part, M.G, inv);
    return hipGetLastError() == hipSuccess ? TB_OK : TB_ERR_HIP;
}

int tb_mlp_loss_grad(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, float *loss, float *grad)
{
    if (!M.P) return TB_ERR_STATE;
    int rc = ensure_data(M, n);
    if (rc) return rc;
    float *dX = (float *)M.data, *dT = dX + (size_t)n * NIN;
    if (hipMemcpyAsync(dX, X, (size_t)n * NIN * 4, hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    if (hipMemcpyAsync(dT, T, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    if ((rc = tb_mlp_loss_grad_dev(s, M, n, dX, dT))) return rc;
    std::vector<float> g(NPAR + 1);
    if (hipMemcpyAsync(g.data(), M.G, (NPAR + 1) * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return TB_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return TB_ERR_HIP;
    for (int j = 0; j < NPAR; j++) grad[j] = g[j];
    *loss = g[NPAR];
    return TB_OK;
}

int tb_mlp_forward(hipStream_t s, MlpDev &M, int n, const float *X, int bf16, float *out)
{
    if (!M.P) return TB_ERR_STATE;
    int rc = ensure_data(M, n);
    if (rc) return rc;
    float *dX = (float *)M.data, *dY = dX + (size_t)n * NIN;
    if (hipMemcpyAsync(dX, X, (size_t)n * NIN * 4, hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(nblk(n, 256)), dim3(256), 0, s, n, dX, M.P, M.Ppk, bf16, dY);
    if (hipMemcpyAsync(out, dY, (size_t)n * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return TB_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? TB_OK : TB_ERR_HIP;
}

int tb_mlp_train(hipStream_t s, MlpDev &M, int n, const float *X, const float *T, int steps, int batch, float lr,
                 uint64_t seed, float *loss_hist)
{
    if (!M.P) return TB_ERR_STATE;
    if (batch < 1 || batch > n) batch = n;
    int rc = ensure_data(M, n);
    if (rc) return rc;
    // shuffled copy of the data on the host, uploaded once; batches are contiguous slices
    std::vector<int> perm((size_t)n);
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    for (int i = n - 1; i > 0; i--) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        int j = (int)(st % (uint64_t)(i + 1));
        std::swap(perm[i], perm[j]);
    }
    std::vector<float> hx((size_t)n * NIN), ht((size_t)n);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < NIN; k++) hx[(size_t)i * NIN + k] = X[(size_t)perm[i] * NIN + k];
        ht[i] = T[perm[i]];
    }
    float *dX = (float *)M.data, *dT = dX + (size_t)n * NIN;
    if (hipMemcpyAsync(dX, hx.data(), hx.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    if (hipMemcpyAsync(dT, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess) return TB_ERR_HIP;
    hipMemsetAsync(M.m, 0, NPAR * 4, s);
    hipMemsetAsync(M.v, 0, NPAR * 4, s);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nbatches = n / batch;
    double c1 = 1.0, c2 = 1.0;
    for (int k = 0; k < steps; k++) {
        const int off = (k % nbatches) * batch;
        if ((rc = tb_mlp_loss_grad_dev(s, M, batch, dX + (size_t)off * NIN, dT + off))) return rc;
        c1 *= b1;
        c2 *= b2;
        hipLaunchKernelGGL(k_adam, dim3(nblk(NPAR, 256)), dim3(256), 0, s, M.P, M.G, M.m, M.v, lr, b1, b2, eps,
                           (float)(1.0 - c1), (float)(1.0 - c2));
        if (loss_hist) hipMemcpyAsync(loss_hist + k, M.G + NPAR, 4, hipMemcpyDeviceToHost, s);
    }
    tb_mlp_refresh_bf16(s, M);
    M.ready = 1;
    return hipStreamSynchronize(s) == hipSuccess ? TB_OK : TB_ERR_HIP;
}
