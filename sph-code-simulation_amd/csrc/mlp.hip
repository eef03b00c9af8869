// mlp.hip -- learned Riemann starting guess: parameters, pair-state log, API.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include <algorithm>

#include "ctx.hpp"
#include "mlp_infer.hpp"

namespace sphgpu {

struct MLP {
    float *w = nullptr;      // FP32 master parameters (device)
    MlpWeightsBF16 *img = nullptr;  // bf16 image used by the inference
    bool img_dirty = true;   // the image lags w
    float *m = nullptr, *v = nullptr;   // Adam moments
    float *grad = nullptr;   // reduced gradient + loss
    float *gpart = nullptr, *lpart = nullptr;  // per-tile partials
    int part_blocks = 0;
    int step = 0;
    bool trained = false;
};

int mlp_create(sphgpu_ctx *c)
{
    c->mlp = new MLP;
    if (hipMalloc((void **)&c->mlp->w, sizeof(float) * kMlpParams) != hipSuccess) return SPHGPU_E_NOMEM;
    if (hipMemset(c->mlp->w, 0, sizeof(float) * kMlpParams) != hipSuccess) return SPHGPU_E_HIP;
    if (hipMalloc((void **)&c->mlp->img, sizeof(MlpWeightsBF16)) != hipSuccess) return SPHGPU_E_NOMEM;
    return 0;
}

__global__ __launch_bounds__(256) void k_mlp_image(const float *__restrict__ w, MlpWeightsBF16 *img)
{
    mlp_load_weights(*img, w);
}

void mlp_destroy(sphgpu_ctx *c)
{
    if (!c->mlp) return;
    MLP *M = c->mlp;
    void *ps[] = {M->w, M->img, M->m, M->v, M->grad, M->gpart, M->lpart};
    for (void *q : ps)
        if (q) hipFree(q);
    delete c->mlp;
    c->mlp = nullptr;
}

MLPDev mlp_dev(sphgpu_ctx *c)
{
    MLPDev m;
    m.w = (c->mlp && c->mlp->trained) ? c->mlp->w : nullptr;
    m.img = nullptr;
    if (m.w) {
        if (c->mlp->img_dirty) {
            hipLaunchKernelGGL(k_mlp_image, dim3(1), dim3(256), 0, c->stream, (const float *)c->mlp->w, c->mlp->img);
            c->mlp->img_dirty = false;
        }
        m.img = c->mlp->img;
    }
    m.gamma = (float)c->prm.gamma;
    return m;
}

int riemann_refine64(sphgpu_ctx *c)
{
    (void)c;
    const char *e = std::getenv("SPHGPU_RIEMANN_REFINE64");
    return (e && e[0] == '0') ? 0 : 1;
}

namespace {

__global__ void k_mlp_predict(int n, const float *__restrict__ rows, MLPDev m, float *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *r = rows + (size_t)k * 7;
    out[k] = mlp_guess(m, r[0], r[1], r[2], r[3], r[4], r[5]);
}


// ---- bf16 MFMA training --------------------------------------------------------
// One workgroup (4 waves) takes a tile of 64 rows through forward and backward
// with v_mfma_f32_16x16x32_bf16: operands bf16 in LDS, accumulation FP32.
// Lane layout of the 16x16x32 tile (CDNA4): A[16][32] -- lane l holds row
// l%16, k = 8*(l/16)..+7; B[32][16] -- lane l holds column l%16, the same k
// range; D[16][16] -- lane l holds rows 4*(l/16)..+3 of column l%16.
typedef __bf16 bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // rows per workgroup
constexpr int kK1 = 32;     // layer-1 inputs padded to one MFMA k-step
constexpr int kLd1 = kK1 + 8;
constexpr int kLd = kMlpH + 8;

struct TrainLDS {
    bf16_t X[kTile][kLd1];      // features [row][in]
    bf16_t XT[kK1][kLd];        // [in][row]
    bf16_t W1[kMlpH][kLd1];     // [out][in]
    bf16_t H1[kTile][kLd];      // relu(z1) [row][h]
    bf16_t H1T[kMlpH][kLd];     // [h][row]
    bf16_t H2[kTile][kLd];      // relu(z2) [row][h]
    bf16_t H2T[kMlpH][kLd];     // [h][row]
    bf16_t W2[kMlpH][kLd];      // [out][in]
    bf16_t W2T[kMlpH][kLd];     // [in][out]
    bf16_t W3[kMlpH][kLd];      // [out][in]
    bf16_t W3T[kMlpH][kLd];     // [in][out]
    bf16_t DA[kTile][kLd];      // dz3 [row][out]
    bf16_t DAT[kMlpH][kLd];     // dz3, later dz1 [out][row]
    bf16_t DB[kTile][kLd];      // dz2 [row][out]
    bf16_t DBT[kMlpH][kLd];     // dz2 [out][row]
    float F[kTile][kMlpH + 1];  // relu(z3), later dz2, then dz1 (FP32)
    float y[kTile], dout[kTile], wt[kTile], loss[kTile];
    float b1[kMlpH], b2[kMlpH], b3[kMlpH], w4[kMlpH];
};

// C (M x N) = A (M x K) * B (K x N) with A row-major [M][lda] and B given by
// its transpose Bt row-major [N][ldb]; tiles of 16x16 spread over the waves.
template <typename Store>
__device__ void wave_gemm(const bf16_t *A, int lda, const bf16_t *Bt, int ldb, int M, int N, int K, Store store)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int tn_count = N / 16;
    for (int t = wave; t < (M / 16) * tn_count; t += nw) {
        const int tm = t / tn_count, tn = t % tn_count;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < K; k0 += 32) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(A + (tm * 16 + (lane & 15)) * lda + k0 + 8 * (lane >> 4));
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(Bt + (tn * 16 + (lane & 15)) * ldb + k0 + 8 * (lane >> 4));
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        for (int i = 0; i < 4; ++i) store(tm * 16 + 4 * (lane >> 4) + i, tn * 16 + (lane & 15), acc[i]);
    }
}

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float *__restrict__ rows, long long n_rows, long long row0,
                                                      long long batch, const float *__restrict__ w, float gamma,
                                                      float *__restrict__ gpart, float *__restrict__ lpart)
{
    __shared__ __align__(16) TrainLDS L;
    const int tid = threadIdx.x;
    const long long base = row0 + (long long)blockIdx.x * kTile;
    // parameters -> LDS (bf16 copies for the MFMA operands)
    for (int e = tid; e < kMlpH * kK1; e += blockDim.x) {
        const int o = e / kK1, k = e % kK1;
        L.W1[o][k] = (bf16_t)(k < kMlpIn ? w[kOffW1 + o * kMlpIn + k] : 0.f);
    }
    for (int e = tid; e < kMlpH * kMlpH; e += blockDim.x) {
        const int o = e / kMlpH, k = e % kMlpH;
        const bf16_t v2 = (bf16_t)w[kOffW2 + o * kMlpH + k];
        L.W2[o][k] = v2;
        L.W2T[k][o] = v2;
        const bf16_t v3 = (bf16_t)w[kOffW3 + o * kMlpH + k];
        L.W3[o][k] = v3;
        L.W3T[k][o] = v3;
    }
    if (tid < kMlpH) {
        L.b1[tid] = w[kOffB1 + tid];
        L.b2[tid] = w[kOffB2 + tid];
        L.b3[tid] = w[kOffB3 + tid];
        L.w4[tid] = w[kOffW4 + tid];
    }
    if (tid < kTile) {
        const long long r = base + tid;
        float x[6] = {0, 0, 0, 0, 0, 0};
        float y = 0.f, wt = 0.f;
        if (r < row0 + batch && r < n_rows) {
            const float *q = rows + r * 7;
            float p_m;
            mlp_features(gamma, q[0], q[1], q[2], q[3], q[4], q[5], x, p_m);
            y = __logf(q[6] / p_m);
            wt = 1.f;
        }
        for (int k = 0; k < kK1; ++k) {
            const bf16_t v = (bf16_t)(k < kMlpIn ? x[k] : 0.f);
            L.X[tid][k] = v;
            L.XT[k][tid] = v;
        }
        L.y[tid] = y;
        L.wt[tid] = wt;
    }
    __syncthreads();
    // forward: z1 = X W1^T + b1, z2 = H1 W2^T + b2, z3 = H2 W3^T + b3
    wave_gemm(&L.X[0][0], kLd1, &L.W1[0][0], kLd1, kTile, kMlpH, kK1, [&](int r, int c, float v) {
        const float h = fmaxf(v + L.b1[c], 0.f);
        L.H1[r][c] = (bf16_t)h;
        L.H1T[c][r] = (bf16_t)h;
    });
    __syncthreads();
    wave_gemm(&L.H1[0][0], kLd, &L.W2[0][0], kLd, kTile, kMlpH, kMlpH, [&](int r, int c, float v) {
        const float h = fmaxf(v + L.b2[c], 0.f);
        L.H2[r][c] = (bf16_t)h;
        L.H2T[c][r] = (bf16_t)h;
    });
    __syncthreads();
    wave_gemm(&L.H2[0][0], kLd, &L.W3[0][0], kLd, kTile, kMlpH, kMlpH, [&](int r, int c, float v) {
        L.F[r][c] = fmaxf(v + L.b3[c], 0.f);
    });
    __syncthreads();
    const float inv_b = 1.0f / (float)batch;
    if (tid < kTile) {
        float out = w[kOffB4];
        for (int o = 0; o < kMlpH; ++o) out = __builtin_fmaf(L.w4[o], L.F[tid][o], out);
        const float d = out - L.y[tid];
        L.loss[tid] = L.wt[tid] * d * d;
        L.dout[tid] = L.wt[tid] * 2.f * d * inv_b;
    }
    __syncthreads();
    float *gp = gpart + (size_t)blockIdx.x * kMlpParams;
    // output layer: dw4 = sum dout relu(z3), db4 = sum dout; db3 = sum dz3
    if (tid < kMlpH) {
        float s = 0.f, sb = 0.f;
        for (int r = 0; r < kTile; ++r) {
            s = __builtin_fmaf(L.dout[r], L.F[r][tid], s);
            sb += (L.F[r][tid] > 0.f) ? L.dout[r] * L.w4[tid] : 0.f;
        }
        gp[kOffW4 + tid] = s;
        gp[kOffB3 + tid] = sb;
    }
    if (tid == 0) {
        float s = 0.f, ls = 0.f;
        for (int r = 0; r < kTile; ++r) { s += L.dout[r]; ls += L.loss[r]; }
        gp[kOffB4] = s;
        lpart[blockIdx.x] = ls * inv_b;
    }
    for (int e = tid; e < kTile * kMlpH; e += blockDim.x) {
        const int r = e / kMlpH, o = e % kMlpH;
        const float dz = (L.F[r][o] > 0.f) ? L.dout[r] * L.w4[o] : 0.f;
        const bf16_t v = (bf16_t)dz;
        L.DA[r][o] = v;
        L.DAT[o][r] = v;
    }
    __syncthreads();
    // dW3 = dZ3^T H2 ; dH2 = dZ3 W3, masked by relu'(z2) -> dz2
    wave_gemm(&L.DAT[0][0], kLd, &L.H2T[0][0], kLd, kMlpH, kMlpH, kTile, [&](int r, int c, float v) {
        gp[kOffW3 + r * kMlpH + c] = v;
    });
    wave_gemm(&L.DA[0][0], kLd, &L.W3T[0][0], kLd, kTile, kMlpH, kMlpH, [&](int r, int c, float v) {
        const float dz = ((float)L.H2[r][c] > 0.f) ? v : 0.f;
        L.F[r][c] = dz;
        L.DB[r][c] = (bf16_t)dz;
        L.DBT[c][r] = (bf16_t)dz;
    });
    __syncthreads();
    // db2 = sum dz2 ; dW2 = dZ2^T H1
    if (tid < kMlpH) {
        float s = 0.f;
        for (int r = 0; r < kTile; ++r) s += L.F[r][tid];
        gp[kOffB2 + tid] = s;
    }
    wave_gemm(&L.DBT[0][0], kLd, &L.H1T[0][0], kLd, kMlpH, kMlpH, kTile, [&](int r, int c, float v) {
        gp[kOffW2 + r * kMlpH + c] = v;
    });
    __syncthreads();
    // dH1 = dZ2 W2, masked by relu'(z1) -> dz1
    wave_gemm(&L.DB[0][0], kLd, &L.W2T[0][0], kLd, kTile, kMlpH, kMlpH, [&](int r, int c, float v) {
        const float dz = ((float)L.H1[r][c] > 0.f) ? v : 0.f;
        L.F[r][c] = dz;
        L.DAT[c][r] = (bf16_t)dz;
    });
    __syncthreads();
    // db1 = sum dz1 ; dW1 = dZ1^T X (first 16 input columns computed, 6 used)
    wave_gemm(&L.DAT[0][0], kLd, &L.XT[0][0], kLd, kMlpH, 16, kTile, [&](int r, int c, float v) {
        if (c < kMlpIn) gp[kOffW1 + r * kMlpIn + c] = v;
    });
    if (tid < kMlpH) {
        float s = 0.f;
        for (int r = 0; r < kTile; ++r) s += L.F[r][tid];
        gp[kOffB1 + tid] = s;
    }
}

// sum of the per-tile partial gradients in a fixed order (deterministic), then
// optionally one Adam step on the FP32 master parameters
__global__ void k_mlp_reduce_adam(const float *__restrict__ gpart, const float *__restrict__ lpart, int nblk,
                                  float *__restrict__ grad, float *__restrict__ loss, float *__restrict__ w,
                                  float *__restrict__ m, float *__restrict__ v, float lr, float b1c, float b2c,
                                  int do_update)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float s = 0.f;
        for (int b = 0; b < nblk; ++b) s += lpart[b];
        *loss = s;
    }
    if (p >= kMlpParams) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += gpart[(size_t)b * kMlpParams + p];
    grad[p] = g;
    if (!do_update) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float mm = beta1 * m[p] + (1.f - beta1) * g;
    const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / b1c, vh = vv / b2c;
    w[p] -= lr * mh / (__builtin_sqrtf(vh) + eps);
}

} // namespace

} // namespace sphgpu

using namespace sphgpu;

extern "C" {

int64_t sphgpu_mlp_param_count(void) { return kMlpParams; }

int sphgpu_mlp_init(sphgpu_ctx *c, uint64_t seed)
{
    if (!c || !c->mlp) return SPHGPU_E_ARG;
    std::mt19937_64 rng(seed);
    std::vector<float> w(kMlpParams, 0.f);
    std::normal_distribution<float> n1(0.f, std::sqrt(2.0f / kMlpIn)), n2(0.f, std::sqrt(2.0f / kMlpH)),
        n3(0.f, std::sqrt(1.0f / kMlpH));
    for (int i = 0; i < kMlpH * kMlpIn; ++i) w[kOffW1 + i] = n1(rng);
    for (int i = 0; i < kMlpH * kMlpH; ++i) w[kOffW2 + i] = n2(rng);
    for (int i = 0; i < kMlpH * kMlpH; ++i) w[kOffW3 + i] = n2(rng);
    for (int i = 0; i < kMlpH; ++i) w[kOffW4 + i] = n3(rng);
    HIPCHK(c, hipMemcpy(c->mlp->w, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    c->mlp->trained = true;
    c->mlp->img_dirty = true;
    c->mlp->step = 0;
    if (c->mlp->m) {
        HIPCHK(c, hipMemset(c->mlp->m, 0, sizeof(float) * kMlpParams));
        HIPCHK(c, hipMemset(c->mlp->v, 0, sizeof(float) * kMlpParams));
    }
    return 0;
}

int sphgpu_mlp_set_params(sphgpu_ctx *c, const float *p, int64_t n)
{
    if (!c || !p || n != kMlpParams) return SPHGPU_E_ARG;
    HIPCHK(c, hipMemcpy(c->mlp->w, p, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    c->mlp->trained = true;
    c->mlp->img_dirty = true;
    return 0;
}

int sphgpu_mlp_get_params(sphgpu_ctx *c, float *p, int64_t n)
{
    if (!c || !p || n != kMlpParams) return SPHGPU_E_ARG;
    HIPCHK(c, hipMemcpy(p, c->mlp->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return 0;
}

int sphgpu_mlp_predict(sphgpu_ctx *c, const float *rows, int64_t n_rows, float *out)
{
    if (!c || !rows || !out || n_rows < 0) return SPHGPU_E_ARG;
    if (n_rows == 0) return 0;
    float *d = nullptr;
    HIPCHK(c, hipMalloc((void **)&d, sizeof(float) * 8 * (size_t)n_rows));
    float *dout = d + 7 * (size_t)n_rows;
    HIPCHK(c, hipMemcpy(d, rows, sizeof(float) * 7 * (size_t)n_rows, hipMemcpyHostToDevice));
    MLPDev m;
    m.w = c->mlp->w;
    m.gamma = (float)c->prm.gamma;
    hipLaunchKernelGGL(k_mlp_predict, dim3(grid_for(n_rows, 256)), dim3(256), 0, c->stream, (int)n_rows,
                       (const float *)d, m, dout);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, dout, sizeof(float) * (size_t)n_rows, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipFree(d);
    return 0;
}

// one pass of forward + backward over rows[row0, row0+batch) (device rows)
static int mlp_pass(sphgpu_ctx *c, const float *drows, long long n_rows, long long row0, long long batch,
                    float *dgrad, float *dloss, bool update, float lr, int step)
{
    MLP *M = c->mlp;
    const int nblk = (int)((batch + kTile - 1) / kTile);
    if (nblk > M->part_blocks) {
        if (M->gpart) hipFree(M->gpart);
        if (M->lpart) hipFree(M->lpart);
        M->part_blocks = nblk;
        HIPCHK(c, hipMalloc((void **)&M->gpart, sizeof(float) * kMlpParams * (size_t)nblk));
        HIPCHK(c, hipMalloc((void **)&M->lpart, sizeof(float) * (size_t)nblk));
    }
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nblk), dim3(256), 0, c->stream, drows, n_rows, row0, batch,
                       (const float *)M->w, (float)c->prm.gamma, M->gpart, M->lpart);
    const float b1c = 1.f - std::pow(0.9f, (float)step), b2c = 1.f - std::pow(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(grid_for(kMlpParams, 256)), dim3(256), 0, c->stream,
                       (const float *)M->gpart, (const float *)M->lpart, nblk, dgrad, dloss, M->w, M->m, M->v, lr,
                       b1c, b2c, update ? 1 : 0);
    HIPCHK(c, hipGetLastError());
    return 0;
}

static int mlp_ensure_state(sphgpu_ctx *c)
{
    MLP *M = c->mlp;
    if (!M->m) {
        HIPCHK(c, hipMalloc((void **)&M->m, sizeof(float) * kMlpParams));
        HIPCHK(c, hipMalloc((void **)&M->v, sizeof(float) * kMlpParams));
        HIPCHK(c, hipMalloc((void **)&M->grad, sizeof(float) * (kMlpParams + 1)));
        HIPCHK(c, hipMemset(M->m, 0, sizeof(float) * kMlpParams));
        HIPCHK(c, hipMemset(M->v, 0, sizeof(float) * kMlpParams));
    }
    return 0;
}

int sphgpu_mlp_loss_grad(sphgpu_ctx *c, const float *rows, int64_t n_rows, float *loss, float *grad)
{
    if (!c || !rows || !loss || !grad || n_rows <= 0) return SPHGPU_E_ARG;
    int rc = mlp_ensure_state(c);
    if (rc) return rc;
    float *d = nullptr;
    HIPCHK(c, hipMalloc((void **)&d, sizeof(float) * 7 * (size_t)n_rows));
    HIPCHK(c, hipMemcpy(d, rows, sizeof(float) * 7 * (size_t)n_rows, hipMemcpyHostToDevice));
    MLP *M = c->mlp;
    rc = mlp_pass(c, d, n_rows, 0, n_rows, M->grad, M->grad + kMlpParams, false, 0.f, 1);
    if (!rc) {
        HIPCHK(c, hipMemcpyAsync(grad, M->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(loss, M->grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    hipFree(d);
    return rc;
}

int sphgpu_mlp_train(sphgpu_ctx *c, const float *rows, int64_t n_rows, int epochs, int batch, float lr,
                     float *final_loss)
{
    if (!c || !rows || n_rows <= 0 || epochs <= 0 || batch <= 0) return SPHGPU_E_ARG;
    int rc = mlp_ensure_state(c);
    if (rc) return rc;
    MLP *M = c->mlp;
    float *d = nullptr;
    HIPCHK(c, hipMalloc((void **)&d, sizeof(float) * 7 * (size_t)n_rows));
    HIPCHK(c, hipMemcpy(d, rows, sizeof(float) * 7 * (size_t)n_rows, hipMemcpyHostToDevice));
    const long long nb = (n_rows + batch - 1) / batch;
    for (int e = 0; e < epochs && !rc; ++e) {
        for (long long b = 0; b < nb && !rc; ++b) {
            const long long r0 = b * batch;
            const long long bs = std::min<long long>(batch, n_rows - r0);
            rc = mlp_pass(c, d, n_rows, r0, bs, M->grad, M->grad + kMlpParams, true, lr, ++M->step);
        }
    }
    M->trained = true;
    M->img_dirty = true;
    if (!rc && final_loss) {
        // loss over the whole set with the final parameters
        rc = mlp_pass(c, d, n_rows, 0, n_rows, M->grad, M->grad + kMlpParams, false, 0.f, 1);
        HIPCHK(c, hipMemcpyAsync(final_loss, M->grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hipFree(d);
    return rc;
}

int sphgpu_riemann_log_enable(sphgpu_ctx *c, int64_t capacity, int stride)
{
    if (!c || capacity < 0 || stride < 1) return SPHGPU_E_ARG;
    if (c->rlog) { hipFree(c->rlog); c->rlog = nullptr; }
    if (!c->rlog_count) HIPCHK(c, hipMalloc((void **)&c->rlog_count, sizeof(unsigned long long)));
    HIPCHK(c, hipMemset(c->rlog_count, 0, sizeof(unsigned long long)));
    c->rlog_cap = capacity;
    c->rlog_stride = stride;
    c->rlog_on = capacity > 0;
    if (capacity > 0) HIPCHK(c, hipMalloc((void **)&c->rlog, sizeof(float) * 7 * (size_t)capacity));
    return 0;
}

int sphgpu_riemann_log_read(sphgpu_ctx *c, float *rows, int64_t max_rows, int64_t *n_rows)
{
    if (!c || !n_rows) return SPHGPU_E_ARG;
    *n_rows = 0;
    if (!c->rlog) return 0;
    unsigned long long cnt = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&cnt, c->rlog_count, sizeof(cnt), hipMemcpyDeviceToHost));
    long long logged = (long long)((cnt + c->rlog_stride - 1) / c->rlog_stride);
    if (logged > c->rlog_cap) logged = c->rlog_cap;
    if (logged > max_rows) logged = max_rows;
    if (rows && logged > 0)
        HIPCHK(c, hipMemcpy(rows, c->rlog, sizeof(float) * 7 * (size_t)logged, hipMemcpyDeviceToHost));
    *n_rows = logged;
    return 0;
}

} // extern "C"
