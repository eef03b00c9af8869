// integrate.hip -- FP64 integrator, timestep and global reductions.
//
// References: Solver::predict / Solver::correct, standard branch
// (src/solver.cpp:1047-1060, 1106-1111); TimeStep::calculation
// (src/timestep.cpp:18-38); Solver::compute_total_energies
// (src/solver.cpp:1149-1175); Solver::initialize tail (src/solver.cpp:941-953).
// All arithmetic here is FP64 with the reference's expressions.  Reductions
// are two-pass with a fixed combination order, so they are deterministic.
#include <hip/hip_runtime.h>

#include <cstring>

#include "ctx.hpp"

namespace sphgpu {

namespace {

struct IntArgs {
    int n, dim;
    double *pos[3], *vel[3], *vel_p[3];
    const double *acc[3];
    double *ene, *ene_p, *sound;
    const double *dene;
    const double *dscal;
    double c_sound;
    Box box;
};

__global__ void k_predict(IntArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double dt = a.dscal[0];
    for (int k = 0; k < a.dim; ++k) a.vel_p[k][i] = a.vel[k][i] + a.acc[k][i] * (0.5 * dt);
    a.ene_p[i] = a.ene[i] + a.dene[i] * (0.5 * dt);
    for (int k = 0; k < a.dim; ++k) a.pos[k][i] += a.vel_p[k][i] * dt;
    for (int k = 0; k < a.dim; ++k) a.vel[k][i] += a.acc[k][i] * dt;
    a.ene[i] += a.dene[i] * dt;
    a.sound[i] = __builtin_sqrt(a.c_sound * a.ene[i]);
    if (a.box.periodic) {
        // Periodic::apply, include/periodic.hpp:61-72
        for (int k = 0; k < a.dim; ++k) {
            double r = a.pos[k][i];
            if (r < a.box.lo[k]) r += a.box.range[k];
            else if (r > a.box.hi[k]) r -= a.box.range[k];
            a.pos[k][i] = r;
        }
    }
}

__global__ void k_correct(IntArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double dt = a.dscal[0];
    for (int k = 0; k < a.dim; ++k) a.vel[k][i] = a.vel_p[k][i] + a.acc[k][i] * (0.5 * dt);
    a.ene[i] = a.ene_p[i] + a.dene[i] * (0.5 * dt);
    a.sound[i] = __builtin_sqrt(a.c_sound * a.ene[i]);
}

__global__ void k_init_fields(int n, double alpha0, double c_sound, const double *__restrict__ ene,
                              double *__restrict__ alpha, double *__restrict__ balsara, double *__restrict__ sound)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    alpha[i] = alpha0;
    balsara[i] = 1.0;
    sound[i] = __builtin_sqrt(c_sound * ene[i]);
}

// dt_force_i = c_force * sqrt(h / |a|) over |a| > 0, block minima
__global__ void k_dt_partial(int n, int dim, const double *ax, const double *ay, const double *az,
                             const double *__restrict__ sml, double c_force, double *__restrict__ part)
{
    __shared__ double s[256];
    double m = 1.7976931348623157e308;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a2 = ax[i] * ax[i];
        if (dim > 1) a2 = a2 + ay[i] * ay[i];
        if (dim > 2) a2 = a2 + az[i] * az[i];
        const double acc_abs = __builtin_sqrt(a2);
        if (acc_abs > 0.0) {
            const double d = c_force * __builtin_sqrt(sml[i] / acc_abs);
            if (d < m) m = d;
        }
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && s[threadIdx.x + o] < s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// final minimum over the block minima (one workgroup of 256; min is exact in
// any order)
__global__ __launch_bounds__(256) void k_dt_final(const double *__restrict__ part, int nb, double c_sound,
                                                  double *dscal)
{
    __shared__ double s[256];
    double m = 1.7976931348623157e308;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (part[b] < m) m = part[b];
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && s[threadIdx.x + o] < s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dt_sound = c_sound * dscal[2];
        dscal[0] = (s[0] < dt_sound) ? s[0] : dt_sound;  // std::min(dt_sound, dt_force)
    }
}

__global__ void k_advance_time(double *dscal) { dscal[1] = dscal[1] + dscal[0]; }

// kinetic, thermal, potential, P_x, P_y, P_z partial sums per block
__global__ void k_energy_partial(int n, int dim, int grav, const double *vx, const double *vy, const double *vz,
                                 const double *__restrict__ m, const double *__restrict__ u,
                                 const double *__restrict__ phi, double *__restrict__ part)
{
    __shared__ double s[6][256];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v0 = vx[i], v1 = dim > 1 ? vy[i] : 0.0, v2 = dim > 2 ? vz[i] : 0.0;
        double vsq = v0 * v0;
        if (dim > 1) vsq = vsq + v1 * v1;
        if (dim > 2) vsq = vsq + v2 * v2;
        a[0] += 0.5 * m[i] * vsq;
        a[1] += m[i] * u[i];
        if (grav) a[2] += 0.5 * m[i] * phi[i];
        a[3] += m[i] * v0;
        a[4] += m[i] * v1;
        a[5] += m[i] * v2;
    }
    for (int q = 0; q < 6; ++q) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 6; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) part[blockIdx.x * 6 + q] = s[q][0];
}

// final sums over the block partials: fixed thread-to-partial assignment and a
// fixed tree order, so the result is deterministic
__global__ __launch_bounds__(256) void k_energy_final(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    __shared__ double s[6][256];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int q = 0; q < 6; ++q) a[q] += part[b * 6 + q];
    for (int q = 0; q < 6; ++q) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 6; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s[0][0];
        out[1] = s[1][0];
        out[2] = s[2][0];
        out[3] = s[0][0] + s[1][0] + s[2][0];
        out[4] = s[3][0];
        out[5] = s[4][0];
        out[6] = s[5][0];
    }
}

struct RecArgs {
    int n, nrec;                  // records written: R_AUX + 1, or all with gradients
    const double *src[4 * R_NREC];
    float4 *r32[R_NREC];          // null entries are not written
    double4 *r64[R_NREC];
};

__global__ void k_records(RecArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    for (int r = 0; r < a.nrec; ++r) {
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = a.src[4 * r + q] ? a.src[4 * r + q][i] : 0.0;
        if (a.r64[r]) a.r64[r][i] = make_double4(v[0], v[1], v[2], v[3]);
        if (a.r32[r]) a.r32[r][i] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    }
}

IntArgs int_args(sphgpu_ctx *c)
{
    IntArgs a;
    a.n = c->n;
    a.dim = c->dim;
    for (int k = 0; k < 3; ++k) {
        a.pos[k] = c->pos[k];
        a.vel[k] = c->vel[k];
        a.vel_p[k] = c->vel_p[k];
        a.acc[k] = c->acc[k];
    }
    a.ene = c->sc[S_ENE];
    a.ene_p = c->sc[S_ENE_P];
    a.sound = c->sc[S_SOUND];
    a.dene = c->sc[S_DENE];
    a.dscal = c->dscal;
    a.c_sound = c->prm.gamma * (c->prm.gamma - 1.0);
    a.box = c->box;
    return a;
}

} // namespace

int predict(sphgpu_ctx *c)
{
    time_begin(c, "predict");
    hipLaunchKernelGGL(k_predict, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, int_args(c));
    time_end(c);
    HIPCHK(c, hipGetLastError());
    c->tree.valid = false;  // positions moved: the tree is stale
    c->mirror_dirty = true;
    return 0;
}

int correct(sphgpu_ctx *c)
{
    time_begin(c, "correct");
    hipLaunchKernelGGL(k_correct, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, int_args(c));
    time_end(c);
    HIPCHK(c, hipGetLastError());
    c->mirror_dirty = true;
    return 0;
}

int init_fields(sphgpu_ctx *c)
{
    hipLaunchKernelGGL(k_init_fields, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->n, c->prm.av_alpha,
                       c->prm.gamma * (c->prm.gamma - 1.0), (const double *)c->sc[S_ENE], c->sc[S_ALPHA],
                       c->sc[S_BALSARA], c->sc[S_SOUND]);
    HIPCHK(c, hipGetLastError());
    c->mirror_dirty = true;
    return 0;
}

int timestep(sphgpu_ctx *c)
{
    const int nb = std::min(c->red_blocks, grid_for(c->n, 256));
    const double *ax = c->acc[0], *ay = c->acc[c->dim > 1 ? 1 : 0], *az = c->acc[c->dim > 2 ? 2 : 0];
    time_begin(c, "timestep");
    hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(256), 0, c->stream, c->n, c->dim, ax, ay, az,
                       (const double *)c->sc[S_SML], c->prm.cfl_force, c->red);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, c->stream, (const double *)c->red, nb, c->prm.cfl_sound,
                       c->dscal);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int advance_time(sphgpu_ctx *c)
{
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, c->stream, c->dscal);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int energy_momentum(sphgpu_ctx *c, double out[7])
{
    const int nb = std::min(c->red_blocks, grid_for(c->n, 256));
    const double *vx = c->vel[0], *vy = c->vel[c->dim > 1 ? 1 : 0], *vz = c->vel[c->dim > 2 ? 2 : 0];
    hipLaunchKernelGGL(k_energy_partial, dim3(nb), dim3(256), 0, c->stream, c->n, c->dim, c->prm.use_gravity, vx, vy,
                       vz, (const double *)c->sc[S_MASS], (const double *)c->sc[S_ENE], (const double *)c->sc[S_PHI],
                       c->red);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, c->stream, (const double *)c->red, nb, c->dscal + 8);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->hscal + 8, c->dscal + 8, sizeof(double) * 7, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memcpy(out, c->hscal + 8, sizeof(double) * 7);
    return 0;
}

int refresh_mirrors(sphgpu_ctx *c)
{
    c->flist = 0;  // records change: a force list derived earlier is stale
    RecArgs a;
    a.n = c->n;
    const bool gsph2 = (c->prm.sph_type == SPHGPU_GSPH || c->prm.sph_type == SPHGPU_GDISPH) && c->prm.gsph_2nd_order;
    // the gradient records are rebuilt from the gradients unless the time
    // step's pre-interaction is about to recompute both
    a.nrec = (gsph2 && !c->in_step) ? R_NREC : R_AUX + 1;
    for (int f = 0; f < 4 * R_NREC; ++f) a.src[f] = nullptr;
    for (int k = 0; k < c->dim; ++k) {
        a.src[4 * R_POSH + k] = c->pos[k];
        a.src[4 * R_VELM + k] = c->vel[k];
    }
    a.src[4 * R_POSH + 3] = c->sc[S_SML];
    a.src[4 * R_VELM + 3] = c->sc[S_MASS];
    a.src[4 * R_THERMO + 0] = c->sc[S_DENS];
    a.src[4 * R_THERMO + 1] = c->sc[S_PRES];
    a.src[4 * R_THERMO + 2] = c->sc[S_SOUND];
    a.src[4 * R_THERMO + 3] = c->sc[S_ENE];
    a.src[4 * R_AUX + 0] = c->sc[S_GRADH];
    a.src[4 * R_AUX + 1] = c->sc[S_BALSARA];
    a.src[4 * R_AUX + 2] = c->sc[S_ALPHA];
    // gradients: flat index q of (gd[3], gp[3], gv[9]) goes to slot 4*R_G0 + q
    for (int k = 0; k < c->dim; ++k) {
        a.src[4 * R_G0 + k] = c->grad[k];
        a.src[4 * R_G0 + 3 + k] = c->grad[3 + k];
        for (int b = 0; b < c->dim; ++b) a.src[4 * R_G0 + 6 + 3 * b + k] = c->grad[6 + 3 * b + k];
    }
    const bool fp64 = c->prm.pair_precision == SPHGPU_PAIR_FP64;
    for (int r = 0; r < R_NREC; ++r) {
        a.r64[r] = fp64 ? c->rec64[r] : nullptr;
        // the FP32 position/h record always exists: it feeds the search prefilter
        a.r32[r] = (!fp64 || r == R_POSH) ? c->rec32[r] : nullptr;
    }
    time_begin(c, "records");
    hipLaunchKernelGGL(k_records, dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, a);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    c->mirror_dirty = false;
    return 0;
}

} // namespace sphgpu
