// mlp.hip -- learned Riemann starting guess: parameters, pair-state log, API.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "ctx.hpp"
#include "mlp_infer.hpp"

namespace sphgpu {

struct MLP {
    float *w = nullptr;      // FP32 master parameters (device)
    bool trained = false;
};

int mlp_create(sphgpu_ctx *c)
{
    c->mlp = new MLP;
    if (hipMalloc((void **)&c->mlp->w, sizeof(float) * kMlpParams) != hipSuccess) return SPHGPU_E_NOMEM;
    if (hipMemset(c->mlp->w, 0, sizeof(float) * kMlpParams) != hipSuccess) return SPHGPU_E_HIP;
    return 0;
}

void mlp_destroy(sphgpu_ctx *c)
{
    if (!c->mlp) return;
    if (c->mlp->w) hipFree(c->mlp->w);
    delete c->mlp;
    c->mlp = nullptr;
}

MLPDev mlp_dev(sphgpu_ctx *c)
{
    MLPDev m;
    m.w = (c->mlp && c->mlp->trained) ? c->mlp->w : nullptr;
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
    for (int i = 0; i < kMlpH; ++i) w[kOffW3 + i] = n3(rng);
    HIPCHK(c, hipMemcpy(c->mlp->w, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    c->mlp->trained = true;
    return 0;
}

int sphgpu_mlp_set_params(sphgpu_ctx *c, const float *p, int64_t n)
{
    if (!c || !p || n != kMlpParams) return SPHGPU_E_ARG;
    HIPCHK(c, hipMemcpy(c->mlp->w, p, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    c->mlp->trained = true;
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

int sphgpu_mlp_loss_grad(sphgpu_ctx *c, const float *rows, int64_t n_rows, float *loss, float *grad)
{
    (void)rows; (void)n_rows; (void)loss; (void)grad;
    if (!c) return SPHGPU_E_ARG;
    return fail(c, SPHGPU_E_STATE, "MLP training kernels are not built in this library version");
}

int sphgpu_mlp_train(sphgpu_ctx *c, const float *rows, int64_t n_rows, int epochs, int batch, float lr,
                     float *final_loss)
{
    (void)rows; (void)n_rows; (void)epochs; (void)batch; (void)lr; (void)final_loss;
    if (!c) return SPHGPU_E_ARG;
    return fail(c, SPHGPU_E_STATE, "MLP training kernels are not built in this library version");
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
