// context.hip -- lifecycle, host exchange and dispatch of the C-ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "ctx.hpp"

using namespace sphgpu;

namespace sphgpu {

int fail(sphgpu_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

// the kernel groups that timing mode 2 keeps
static bool pair_kernel(const char *name)
{
    static const char *const keep[] = {"search_pre", "search_force", "pre_interaction", "fluid_force", "gravity"};
    for (const char *k : keep)
        if (std::strcmp(name, k) == 0) return true;
    return false;
}

void time_begin(sphgpu_ctx *c, const char *name)
{
    c->timing_open = false;
    if (!c->timing || (c->timing == 2 && !pair_kernel(name))) return;
    c->timing_open = true;
    Timer t;
    if (c->event_pool.size() >= 2) {
        t.a = c->event_pool.back(); c->event_pool.pop_back();
        t.b = c->event_pool.back(); c->event_pool.pop_back();
    } else {
        hipEventCreate(&t.a);
        hipEventCreate(&t.b);
    }
    t.name = name;
    hipEventRecord(t.a, c->stream);
    c->pending.push_back(t);
}

void time_end(sphgpu_ctx *c)
{
    if (!c->timing_open || c->pending.empty()) return;
    c->timing_open = false;
    hipEventRecord(c->pending.back().b, c->stream);
}

void collect_times(sphgpu_ctx *c)
{
    if (c->pending.empty()) return;
    hipEventSynchronize(c->pending.back().b);
    for (auto &t : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            auto &e = c->ktimes[t.name];
            e.first += ms;
            e.second += 1;
        }
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
}

template <> RecView<float> rec_view<float>(sphgpu_ctx *c)
{
    RecView<float> v;
    for (int r = 0; r < R_NREC; ++r) v.r[r] = c->rec32[r];
    return v;
}

template <> RecView<double> rec_view<double>(sphgpu_ctx *c)
{
    RecView<double> v;
    for (int r = 0; r < R_NREC; ++r) v.r[r] = c->rec64[r];
    return v;
}

} // namespace sphgpu

// --------------------------------------------------------------------------
// plain copy of known size (counter calibration and attainable bandwidth)
__global__ __launch_bounds__(256) void k_copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                   long long n16)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// --------------------------------------------------------------------------
// host-exchange kernels: row f of the staging block is in id order
struct FieldPtrs {
    double *p[SPHGPU_NFIELDS];
    int *ip[SPHGPU_NFIELDS];
};

__global__ void k_gather_fields(FieldPtrs dst, const double *__restrict__ stage, const int *__restrict__ id,
                                int n, unsigned long long mask)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int src = id[k];
    for (int f = 0; f < SPHGPU_NFIELDS; ++f) {
        if (!(mask & (1ull << f))) continue;
        const double v = stage[(size_t)f * n + src];
        if (dst.p[f]) dst.p[f][k] = v;
        else if (dst.ip[f]) dst.ip[f][k] = (int)v;
    }
}

__global__ void k_scatter_fields(FieldPtrs src, double *__restrict__ stage, const int *__restrict__ id, int n,
                                 unsigned long long mask)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int dst = id[k];
    for (int f = 0; f < SPHGPU_NFIELDS; ++f) {
        if (!(mask & (1ull << f))) continue;
        double v = 0.0;
        if (src.p[f]) v = src.p[f][k];
        else if (src.ip[f]) v = (double)src.ip[f][k];
        stage[(size_t)f * n + dst] = v;
    }
}

__global__ void k_iota(int *a, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = k;
}

static FieldPtrs field_ptrs(sphgpu_ctx *c)
{
    FieldPtrs f;
    for (int i = 0; i < SPHGPU_NFIELDS; ++i) { f.p[i] = nullptr; f.ip[i] = nullptr; }
    for (int k = 0; k < c->dim; ++k) {
        f.p[SPHGPU_F_POS + k] = c->pos[k];
        f.p[SPHGPU_F_VEL + k] = c->vel[k];
        f.p[SPHGPU_F_VEL_P + k] = c->vel_p[k];
        f.p[SPHGPU_F_ACC + k] = c->acc[k];
        f.p[SPHGPU_F_GRAD_D + k] = c->grad[k];
        f.p[SPHGPU_F_GRAD_P + k] = c->grad[3 + k];
        for (int a = 0; a < c->dim; ++a) f.p[SPHGPU_F_GRAD_V + 3 * a + k] = c->grad[6 + 3 * a + k];
    }
    const int smap[S_NSCALAR] = {SPHGPU_F_MASS, SPHGPU_F_DENS, SPHGPU_F_PRES, SPHGPU_F_ENE, SPHGPU_F_ENE_P,
                                 SPHGPU_F_DENE, SPHGPU_F_SML, SPHGPU_F_SOUND, SPHGPU_F_BALSARA, SPHGPU_F_ALPHA,
                                 SPHGPU_F_GRADH, SPHGPU_F_PHI};
    for (int s = 0; s < S_NSCALAR; ++s) f.p[smap[s]] = c->sc[s];
    f.ip[SPHGPU_F_NEIGHBOR] = c->neighbor;
    return f;
}

// --------------------------------------------------------------------------
extern "C" {

void sphgpu_params_default(sphgpu_params *p, int dim)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dim = dim;
    p->sph_type = SPHGPU_SSPH;
    p->kernel = SPHGPU_CUBIC_SPLINE;
    p->riemann_solver = SPHGPU_RIEMANN_HLL;
    p->neighbor_number = 50;
    p->leaf_particle_num = 1;
    p->max_tree_level = 20;
    p->iterative_sml = 1;
    p->use_balsara_switch = 1;
    p->use_time_dependent_av = 0;
    p->use_ac = 0;
    p->periodic = 0;
    p->use_gravity = 0;
    p->gsph_2nd_order = 1;
    p->pair_precision = SPHGPU_PAIR_FP32;
    p->riemann_guess = SPHGPU_GUESS_PLAIN;
    p->gamma = 1.6666666666666667;
    p->cfl_sound = 0.3;
    p->cfl_force = 0.125;
    p->av_alpha = 1.0;
    p->alpha_max = 2.0;
    p->alpha_min = 0.1;
    p->epsilon_av = 0.2;
    p->alpha_ac = 1.0;
    p->gravity_constant = 1.0;
    p->theta = 0.5;
}

int sphgpu_device_info(char *name, int name_len, int *cu_count, double *hbm_gb)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SPHGPU_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return SPHGPU_E_HIP;
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_gb) *hbm_gb = (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0);
    return 0;
}

static int check_params(const sphgpu_params *p, std::string &why)
{
    if (p->dim < 1 || p->dim > 3) { why = "dim must be 1, 2 or 3"; return 0; }
    if (p->sph_type < SPHGPU_SSPH || p->sph_type > SPHGPU_GDISPH) { why = "unsupported SPHType"; return 0; }
    if (p->kernel < SPHGPU_CUBIC_SPLINE || p->kernel > SPHGPU_GAUSSIAN) { why = "unknown kernel"; return 0; }
    if (p->kernel == SPHGPU_WENDLAND && p->dim == 1) { why = "Wendland C4 is undefined in 1D"; return 0; }
    const int bits = p->dim;
    if (p->max_tree_level < 1 || p->max_tree_level * bits > 63) { why = "maxTreeLevel too deep for 64-bit keys"; return 0; }
    if (p->leaf_particle_num < 1 || p->leaf_particle_num > 4096) { why = "leafParticleNumber out of range"; return 0; }
    if (p->neighbor_number < 1) { why = "neighborNumber must be positive"; return 0; }
    return 1;
}

int sphgpu_create(const sphgpu_params *p, int n, sphgpu_ctx **out)
{
    if (!p || !out || n <= 0) return SPHGPU_E_ARG;
    std::string why;
    if (!check_params(p, why)) { *out = nullptr; return SPHGPU_E_ARG; }
    sphgpu_ctx *c = new sphgpu_ctx;
    c->prm = *p;
    c->n = n;
    c->dim = p->dim;
    c->box.periodic = p->periodic;
    for (int k = 0; k < 3; ++k) {
        c->box.lo[k] = p->range_min[k];
        c->box.hi[k] = p->range_max[k];
        c->box.range[k] = p->range_max[k] - p->range_min[k];
        c->box.rangef[k] = (float)c->box.range[k];
        c->box.rinvf[k] = c->box.range[k] > 0.0 ? (float)(1.0 / c->box.range[k]) : 0.f;
    }
    auto alloc = [&](void **ptr, size_t bytes) -> bool {
        if (hipMalloc(ptr, bytes) != hipSuccess) return false;
        return hipMemset(*ptr, 0, bytes) == hipSuccess;
    };
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    const size_t dn = sizeof(double) * (size_t)n;
    for (int k = 0; k < 3 && ok; ++k) {
        ok = ok && alloc((void **)&c->pos[k], dn) && alloc((void **)&c->vel[k], dn) &&
             alloc((void **)&c->vel_p[k], dn) && alloc((void **)&c->acc[k], dn);
    }
    for (int s = 0; s < S_NSCALAR && ok; ++s) ok = alloc((void **)&c->sc[s], dn);
    for (int g = 0; g < kNGrad && ok; ++g) ok = alloc((void **)&c->grad[g], dn);
    for (int a = 0; a < 16 && ok; ++a) ok = alloc((void **)&c->alt[a], dn);
    ok = ok && alloc((void **)&c->id, sizeof(int) * n) && alloc((void **)&c->id_alt, sizeof(int) * n) &&
         alloc((void **)&c->neighbor, sizeof(int) * n);
    for (int r = 0; r < R_NREC && ok; ++r) ok = alloc((void **)&c->rec32[r], sizeof(float4) * (size_t)n);
    if (p->pair_precision == SPHGPU_PAIR_FP64)
        for (int r = 0; r < R_NREC && ok; ++r) ok = alloc((void **)&c->rec64[r], sizeof(double4) * (size_t)n);
    ok = ok && alloc((void **)&c->hs, dn) && alloc((void **)&c->stage, dn * SPHGPU_NFIELDS);
    if (p->use_time_dependent_av) ok = ok && alloc((void **)&c->alpha_snap, dn);
    c->red_blocks = 1024;
    ok = ok && alloc((void **)&c->red, sizeof(double) * 8 * c->red_blocks);
    ok = ok && alloc((void **)&c->dscal, sizeof(double) * 16) && alloc((void **)&c->dmin, sizeof(unsigned long long) * kStatCopies) &&
         alloc((void **)&c->dstats, sizeof(long long) * kStatSlots * kStatCopies) && alloc((void **)&c->dflag, sizeof(int) * 8);
    ok = ok && hipHostMalloc((void **)&c->hscal, sizeof(double) * 16, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&c->hflag, sizeof(int) * 8, hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess &&
         hipHostGetDevicePointer((void **)&c->hflag_dev, c->hflag, 0) == hipSuccess &&
         hipEventCreateWithFlags(&c->flag_ready, hipEventDisableTiming) == hipSuccess;
    ok = ok && tree_alloc(c) == 0;
    // neighbour lists: about 5x the expected candidate count nn * 1.2^D
    // (a shock compresses the gas ahead of the smoothing lengths, and a list
    // that overflows costs a second search in that step; the slot-major
    // layout touches only the slots a particle fills, so capacity costs
    // memory, 4 B per slot and particle, not traffic)
    const double kap = p->iterative_sml ? std::pow(1.2, p->dim) : 1.0;
    int cap = (int)std::ceil(p->neighbor_number * kap * 5.0 / 64.0) * 64;
    cap = std::max(cap, 64);
    ok = ok && lists_alloc(c, c->lpre, cap) == 0 && lists_alloc(c, c->lforce, cap) == 0;
    ok = ok && mlp_create(c) == 0;
    if (!ok) {
        sphgpu_destroy(c);
        *out = nullptr;
        return SPHGPU_E_NOMEM;
    }
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->id, n);
    hipStreamSynchronize(c->stream);
    *out = c;
    return 0;
}

int sphgpu_destroy(sphgpu_ctx *c)
{
    if (!c) return SPHGPU_E_ARG;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (int k = 0; k < 3; ++k) {
        hipFree(c->pos[k]); hipFree(c->vel[k]); hipFree(c->vel_p[k]); hipFree(c->acc[k]);
    }
    for (int s = 0; s < S_NSCALAR; ++s) hipFree(c->sc[s]);
    for (int g = 0; g < kNGrad; ++g) hipFree(c->grad[g]);
    for (int a = 0; a < 16; ++a) hipFree(c->alt[a]);
    hipFree(c->id); hipFree(c->id_alt); hipFree(c->neighbor);
    for (int r = 0; r < R_NREC; ++r) {
        hipFree(c->rec32[r]);
        hipFree(c->rec64[r]);
    }
    hipFree(c->hs); hipFree(c->alpha_snap); hipFree(c->stage); hipFree(c->red);
    hipFree(c->dscal); hipFree(c->dmin); hipFree(c->dstats); hipFree(c->dflag);
    if (c->hscal) hipHostFree(c->hscal);
    if (c->hflag) hipHostFree(c->hflag);
    if (c->flag_ready) hipEventDestroy(c->flag_ready);
    tree_free(c);
    lists_free(c->lpre);
    lists_free(c->lforce);
    mlp_destroy(c);
    if (c->rlog) hipFree(c->rlog);
    if (c->rlog_count) hipFree(c->rlog_count);
    for (auto &t : c->pending) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    for (auto e : c->event_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char *sphgpu_last_error(const sphgpu_ctx *c) { return c ? c->err.c_str() : "null context"; }

int sphgpu_get_params(const sphgpu_ctx *c, sphgpu_params *p, int32_t *n)
{
    if (!c) return SPHGPU_E_ARG;
    if (p) *p = c->prm;
    if (n) *n = c->n;
    return 0;
}

int sphgpu_upload(sphgpu_ctx *c, const double *fields, uint64_t mask)
{
    if (!c || !fields) return SPHGPU_E_ARG;
    mask &= SPHGPU_MASK_ALL & ~SPHGPU_MASK(SPHGPU_F_ID);
    const size_t dn = sizeof(double) * (size_t)c->n;
    for (int f = 0; f < SPHGPU_NFIELDS; ++f)
        if (mask & SPHGPU_MASK(f))
            HIPCHK(c, hipMemcpyAsync(c->stage + (size_t)f * c->n, fields + (size_t)f * c->n, dn,
                                     hipMemcpyHostToDevice, c->stream));
    if (mask & SPHGPU_MASK(SPHGPU_F_MASS)) {
        // masses are inputs only: whether they are all equal is decided here
        const double *m = fields + (size_t)SPHGPU_F_MASS * c->n;
        bool same = c->n > 0;
        for (int p = 1; same && p < c->n; ++p) same = (m[p] == m[0]);
        c->mass_uniform = same;
        c->mass_u = c->n > 0 ? m[0] : 0.0;
    }
    FieldPtrs fp = field_ptrs(c);
    hipLaunchKernelGGL(k_gather_fields, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, fp,
                       (const double *)c->stage, (const int *)c->id, c->n, (unsigned long long)mask);
    HIPCHK(c, hipGetLastError());
    // new positions make the tree stale (BHTree::make runs again before the
    // next search); any host change of inputs invalidates the FP32 mirrors
    // and with them a derived force list
    const uint64_t pos_rows = SPHGPU_MASK(SPHGPU_F_POS) | SPHGPU_MASK(SPHGPU_F_POS + 1) | SPHGPU_MASK(SPHGPU_F_POS + 2);
    if (mask & pos_rows) c->tree.valid = false;
    if (refresh_mirrors(c)) return SPHGPU_E_HIP;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sphgpu_download(sphgpu_ctx *c, double *fields, uint64_t mask)
{
    if (!c || !fields) return SPHGPU_E_ARG;
    mask &= SPHGPU_MASK_ALL;
    FieldPtrs fp = field_ptrs(c);
    fp.ip[SPHGPU_F_ID] = c->id;
    hipLaunchKernelGGL(k_scatter_fields, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, fp, c->stage,
                       (const int *)c->id, c->n, (unsigned long long)mask);
    HIPCHK(c, hipGetLastError());
    const size_t dn = sizeof(double) * (size_t)c->n;
    for (int f = 0; f < SPHGPU_NFIELDS; ++f) {
        if (!(mask & SPHGPU_MASK(f))) continue;
        HIPCHK(c, hipMemcpyAsync(fields + (size_t)f * c->n, c->stage + (size_t)f * c->n, dn, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int get_scalar(sphgpu_ctx *c, int idx, double *v)
{
    if (!c || !v) return SPHGPU_E_ARG;
    HIPCHK(c, hipMemcpyAsync(c->hscal + idx, c->dscal + idx, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *v = c->hscal[idx];
    return 0;
}

static int set_scalar(sphgpu_ctx *c, int idx, double v)
{
    if (!c) return SPHGPU_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->hscal[idx] = v;
    HIPCHK(c, hipMemcpyAsync(c->dscal + idx, c->hscal + idx, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int sphgpu_set_dt(sphgpu_ctx *c, double dt) { return set_scalar(c, 0, dt); }
int sphgpu_get_dt(sphgpu_ctx *c, double *dt) { return get_scalar(c, 0, dt); }
int sphgpu_set_time(sphgpu_ctx *c, double t) { return set_scalar(c, 1, t); }
int sphgpu_get_time(sphgpu_ctx *c, double *t) { return get_scalar(c, 1, t); }
int sphgpu_get_h_per_v_sig(sphgpu_ctx *c, double *v) { return get_scalar(c, 2, v); }
int sphgpu_set_h_per_v_sig(sphgpu_ctx *c, double v) { return set_scalar(c, 2, v); }

#define CTX_CALL(expr)                                   \
    do {                                                 \
        if (!c) return SPHGPU_E_ARG;                     \
        int rc_ = (expr);                                \
        if (rc_) return rc_;                             \
        HIPCHK(c, hipStreamSynchronize(c->stream));      \
        collect_times(c);                                \
        return 0;                                        \
    } while (0)

int sphgpu_build_tree(sphgpu_ctx *c) { CTX_CALL(tree_build(c)); }
int sphgpu_set_kernel(sphgpu_ctx *c) { CTX_CALL(tree_set_kernel(c)); }
int sphgpu_pre_interaction(sphgpu_ctx *c) { CTX_CALL(pre_interaction(c)); }
int sphgpu_fluid_force(sphgpu_ctx *c) { CTX_CALL(fluid_force(c)); }
int sphgpu_gravity(sphgpu_ctx *c) { CTX_CALL(gravity(c)); }
int sphgpu_timestep(sphgpu_ctx *c) { CTX_CALL(timestep(c)); }
int sphgpu_predict(sphgpu_ctx *c) { CTX_CALL(predict(c)); }
int sphgpu_correct(sphgpu_ctx *c) { CTX_CALL(correct(c)); }

static int initialize_impl(sphgpu_ctx *c)
{
    int rc;
    if ((rc = init_fields(c))) return rc;
    if ((rc = tree_build(c))) return rc;
    if ((rc = pre_interaction(c))) return rc;
    if ((rc = fluid_force(c))) return rc;
    return gravity(c);
}

int sphgpu_initialize(sphgpu_ctx *c) { CTX_CALL(initialize_impl(c)); }

static int integrate_impl(sphgpu_ctx *c, int nsteps)
{
    int rc;
    for (int s = 0; s < nsteps; ++s) {
        if ((rc = timestep(c))) return rc;
        if ((rc = predict(c))) return rc;
        c->in_step = true;
        rc = tree_build(c);
        c->in_step = false;
        if (rc) return rc;
        if ((rc = pre_interaction_deferred(c))) return rc;
        if ((rc = fluid_force(c))) return rc;  // settles the deferred search first
        if ((rc = gravity(c))) return rc;
        if ((rc = correct(c))) return rc;
        if ((rc = advance_time(c))) return rc;
    }
    return 0;
}

int sphgpu_integrate(sphgpu_ctx *c, int nsteps) { CTX_CALL(integrate_impl(c, nsteps)); }

int sphgpu_energy_momentum(sphgpu_ctx *c, double out[7])
{
    if (!c || !out) return SPHGPU_E_ARG;
    return energy_momentum(c, out);
}

int sphgpu_get_stats(sphgpu_ctx *c, sphgpu_stats *out)
{
    if (!c || !out) return SPHGPU_E_ARG;
    std::vector<long long> all((size_t)kStatSlots * kStatCopies);
    HIPCHK(c, hipMemcpyAsync(all.data(), c->dstats, sizeof(long long) * all.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // sum the copies (slot 12, the tree depth, is a maximum)
    long long h[kStatSlots] = {};
    for (int k = 0; k < kStatCopies; ++k)
        for (int s = 0; s < kStatSlots; ++s) {
            const long long v = all[(size_t)k * kStatSlots + s];
            h[s] = (s == 12) ? std::max(h[s], v) : h[s] + v;
        }
    std::memset(out, 0, sizeof(*out));
    out->pre_candidates = h[0];
    out->pre_newton_iters = h[1];
    out->pre_newton_fail = h[2];
    out->force_pairs = h[3];
    out->force_candidates = h[4];
    out->riemann_iters = h[5];
    out->riemann_wave_max_iters = h[6];
    out->riemann_calls = h[7];
    out->riemann_fallbacks = h[8];
    out->riemann_learned = h[9];
    out->gravity_interactions = h[10];
    out->nodes = c->tree.n_nodes;
    out->leaves = h[11];
    out->max_depth = (int32_t)h[12];
    out->search_node_visits = h[13];
    out->search_member_tests = h[14];
    out->search_windows = h[15];
    out->search_leaves = h[16];
    out->search_exact = h[17];
    out->flist_h_beyond = h[18];
    out->flist_other = h[19];
    out->flist_searches = c->flist_searches;
    out->search_reach_tests = h[20];
    return 0;
}

int sphgpu_reset_stats(sphgpu_ctx *c)
{
    if (!c) return SPHGPU_E_ARG;
    HIPCHK(c, hipMemsetAsync(c->dstats, 0, sizeof(long long) * kStatSlots * kStatCopies, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->flist_searches = 0;
    return 0;
}

int sphgpu_neighbor_sets(sphgpu_ctx *c, const double *h, int is_ij, int cap, int32_t *counts, int32_t *lists)
{
    if (!c || !counts || !lists || cap <= 0 || is_ij < 0 || is_ij > 2 || (is_ij != 2 && !h)) return SPHGPU_E_ARG;
    if (!c->tree.valid) return fail(c, SPHGPU_E_STATE, "neighbor_sets needs a built tree");
    int rc;
    if (is_ij == 2) {
        // the force pair lists the last pre-interaction derived
        if ((rc = flist_check(c)) < 0) return rc;
        if (rc) return fail(c, SPHGPU_E_STATE, "neighbor_sets: no force pair list from a pre-interaction");
    } else {
        // smoothing lengths in id order -> sorted order
        std::vector<double> stage((size_t)c->n * SPHGPU_NFIELDS, 0.0);
        std::memcpy(stage.data() + (size_t)SPHGPU_F_SML * c->n, h, sizeof(double) * c->n);
        if ((rc = sphgpu_upload(c, stage.data(), SPHGPU_MASK(SPHGPU_F_SML)))) return rc;
        if ((rc = tree_set_kernel(c))) return rc;
        // pre-interaction radius: h itself (r2 < h^2, src/bhtree.cpp:280-299)
        HIPCHK(c, hipMemcpyAsync(c->hs, c->sc[S_SML], sizeof(double) * c->n, hipMemcpyDeviceToDevice, c->stream));
        if ((rc = search(c, is_ij ? c->lforce : c->lpre, is_ij, c->hs))) return rc;
    }
    Lists &l = is_ij ? c->lforce : c->lpre;
    const size_t list_len = (size_t)l.cap * (((size_t)c->n + 63) / 64) * 64;
    std::vector<int> ids(c->n), cnt(c->n), idx(list_len);
    HIPCHK(c, hipMemcpyAsync(ids.data(), c->id, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), l.cnt, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(idx.data(), l.idx, sizeof(int) * list_len, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->n; ++k) {
        const int i = ids[k];
        const int m = cnt[k];
        if (m > cap) return fail(c, SPHGPU_E_NEIGHBOR_OVERFLOW, "neighbor_sets: cap too small");
        counts[i] = m;
        for (int s = 0; s < m; ++s) lists[(size_t)i * cap + s] = ids[idx[list_slot(k, s, l.cap)]];
    }
    return 0;
}

int sphgpu_tree_dump(sphgpu_ctx *c, int64_t *n_nodes, int32_t *level, int32_t *count, int32_t *first,
                     int32_t *is_leaf, double *center, double *edge, double *mass, double *com,
                     double *kernel_size, int32_t *members)
{
    if (!c || !n_nodes) return SPHGPU_E_ARG;
    return tree_dump(c, n_nodes, level, count, first, is_leaf, center, edge, mass, com, kernel_size, members);
}

int sphgpu_riemann_batch(sphgpu_ctx *c, int solver, int precision, int n, const double *left, const double *right,
                         double *pstar, double *vstar, int32_t *iters, int32_t *learned_used)
{
    if (!c || n < 0 || (n > 0 && (!left || !right || !pstar || !vstar))) return SPHGPU_E_ARG;
    if (n == 0) return 0;
    return riemann_batch(c, solver, precision, n, left, right, pstar, vstar, iters, learned_used);
}

int sphgpu_riemann_batch_static(sphgpu_ctx *c, int solver, int precision, int n, const double *left,
                                const double *right, double *pstar, double *vstar, int32_t *iters,
                                int32_t *learned_used)
{
    if (!c || n < 0 || (n > 0 && (!left || !right || !pstar || !vstar))) return SPHGPU_E_ARG;
    if (n == 0) return 0;
    return riemann_batch_static(c, solver, precision, n, left, right, pstar, vstar, iters, learned_used);
}

int sphgpu_enable_timing(sphgpu_ctx *c, int on)
{
    if (!c) return SPHGPU_E_ARG;
    collect_times(c);
    c->timing = on < 0 ? 0 : on > 2 ? 1 : on;
    c->ktimes.clear();
    return 0;
}

int sphgpu_kernel_times(sphgpu_ctx *c, char *name_buf, int buf_len, double *avg_ms, int64_t *count, int max_kernels,
                        int *n_kernels)
{
    if (!c || !name_buf || buf_len <= 0 || !n_kernels) return SPHGPU_E_ARG;
    collect_times(c);
    std::string names;
    int k = 0;
    for (auto &kv : c->ktimes) {
        if (k >= max_kernels) break;
        if (avg_ms) avg_ms[k] = kv.second.first / (double)std::max(1LL, kv.second.second);
        if (count) count[k] = kv.second.second;
        names += kv.first;
        names += ";";
        ++k;
    }
    std::snprintf(name_buf, buf_len, "%s", names.c_str());
    *n_kernels = k;
    return 0;
}

int sphgpu_copy_probe(sphgpu_ctx *c, int64_t nbytes, double *ms)
{
    if (!c || !ms || nbytes < 16) return SPHGPU_E_ARG;
    const long long n16 = nbytes / 16;
    float4 *src = nullptr, *dst = nullptr;
    HIPCHK(c, hipMalloc((void **)&src, 16ull * n16));
    HIPCHK(c, hipMalloc((void **)&dst, 16ull * n16));
    HIPCHK(c, hipMemsetAsync(src, 0, 16ull * n16, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    const int blocks = 4096;
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, (const float4 *)src, dst, n16);
    HIPCHK(c, hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, (const float4 *)src, dst, n16);
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(src);
    hipFree(dst);
    return 0;
}

} // extern "C"
