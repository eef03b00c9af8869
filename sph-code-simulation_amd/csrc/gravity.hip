// gravity.hip -- Barnes-Hut tree walk for self-gravity.
//
// Reference: GravityForce::calculation -> BHTree::tree_force ->
// BHNode::calc_force (src/gravity_force.cpp:52-89, src/bhtree.cpp:146-150,
// 319-393).  For particle i a node is opened when edge^2 > theta^2 d^2 or
// d^2 < edge^2, with d the minimum-image vector from r_i to the node's centre
// of mass; an opened leaf contributes Hernquist & Katz (1989) softened pair
// terms for every member j != i, an unopened node its monopole.  phi is
// assigned, acc is added to (the walk runs after FluidForce).
//
// One thread per particle walks the pre-ordered node array stacklessly
// (descend = index + 1, prune = skip).  Threads of a wave hold neighbouring
// particles in key order, so their walks are largely coherent.  The opening
// decision is taken in FP64 exactly as the reference takes it; the pair and
// monopole terms run in the pair precision T with compensated sums.
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace sphgpu {

namespace {

template <typename T> SPH_DEV T hk_f(T r, T h)
{
    const T e = h * T(0.5);
    const T u = r / e;
    if (u < T(1.0)) return (T(-0.5) * u * u * (T(1.0 / 3.0) - T(3.0 / 20) * u * u + u * u * u / T(20)) + T(1.4)) / e;
    else if (u < T(2.0))
        return T(-1.0) / (T(15) * r) + (-u * u * (T(4.0 / 3.0) - u + T(0.3) * u * u - u * u * u / T(30)) + T(1.6)) / e;
    else return T(1) / r;
}

template <typename T> SPH_DEV T hk_g(T r, T h)
{
    const T e = h * T(0.5);
    const T u = r / e;
    if (u < T(1.0)) return (T(4.0 / 3.0) - T(1.2) * u * u + T(0.5) * u * u * u) / (e * e * e);
    else if (u < T(2.0))
        return (T(-1.0 / 15) + T(8.0 / 3) * u * u * u - T(3) * u * u * u * u + T(1.2) * u * u * u * u * u -
                u * u * u * u * u * u / T(6.0)) / (r * r * r);
    else return T(1) / (r * r * r);
}

template <typename T> struct GravArgs {
    int n, n_nodes, node_cap;
    const double *x64[3];
    RecView<T> rv;
    const int *nstart, *nend, *nskip, *nlevel;
    const double *com, *edge, *mass;
    double *acc[3];
    double *phi;
    double theta2, G;
    Box box;
    long long *stats;
};

template <int D, typename T>
__global__ __launch_bounds__(256) void k_gravity(GravArgs<T> a)
{
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = i0 < a.n;
    const int i = valid ? i0 : a.n - 1;
    typedef typename V4T<T>::type V4;
    const V4 *__restrict__ RP = a.rv.r[R_POSH];
    const V4 *__restrict__ RV = a.rv.r[R_VELM];
    double ri[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) ri[k] = a.x64[k][i];
    const V4 pi_ = RP[i];
    const T xi[3] = {pi_.x, pi_.y, pi_.z};
    const T hi = pi_.w;
    const T G = T(a.G);
    Acc<T> phi, ac[3];
    long long inter = 0;
    int node = valid ? 0 : a.n_nodes;
    while (node < a.n_nodes) {
        const double e = a.edge[node];
        const double l2 = e * e;
        double d[3];
        double d2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            d[k] = rij_axis<double>(ri[k], a.com[(size_t)k * a.node_cap + node], a.box, k);
            d2 = (k == 0) ? d[k] * d[k] : d2 + d[k] * d[k];
        }
        const bool open = (l2 > a.theta2 * d2) || (d2 < l2);
        if (open) {
            if (a.nlevel[node] < 0) {
                const int s = a.nstart[node], en = a.nend[node];
                for (int j = s; j < en; ++j) {
                    if (j == i) continue;
                    const V4 pj = RP[j];
                    const T xj[3] = {pj.x, pj.y, pj.z};
                    T rij[3];
                    T r2 = T(0);
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        rij[k] = rij_axis<T>(xi[k], xj[k], a.box, k);
                        r2 = (k == 0) ? rij[k] * rij[k] : r2 + rij[k] * rij[k];
                    }
                    const T r = tsqrt(r2);
                    const T hj = pj.w;
                    const T gm = G * RV[j].w;
                    phi.add(-(gm * (hk_f<T>(r, hi) + hk_f<T>(r, hj)) * T(0.5)));
                    const T s2 = gm * (hk_g<T>(r, hi) + hk_g<T>(r, hj)) * T(0.5);
#pragma unroll
                    for (int k = 0; k < D; ++k) ac[k].add(-(rij[k] * s2));
                    ++inter;
                }
            }
            ++node;
        } else {
            const T r_inv = T(1.0) / tsqrt(T(d2));
            const T gm = G * T(a.mass[node]);
            phi.add(-(gm * r_inv));
            const T s = gm * (r_inv * r_inv * r_inv);
#pragma unroll
            for (int k = 0; k < D; ++k) ac[k].add(-(T(d[k]) * s));
            ++inter;
            node = a.nskip[node];
        }
    }
    if (valid) {
        a.phi[i] = (double)phi.get();
#pragma unroll
        for (int k = 0; k < D; ++k) a.acc[k][i] += (double)ac[k].get();
    }
    for (int off = kWave / 2; off > 0; off >>= 1) inter += __shfl_xor(inter, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(stat_ptr(a.stats, 10), (unsigned long long)inter);
}

template <typename T> void launch_gravity(sphgpu_ctx *c)
{
    GravArgs<T> a;
    a.n = c->n;
    a.n_nodes = c->tree.n_nodes;
    a.node_cap = c->tree.node_cap;
    for (int k = 0; k < 3; ++k) a.x64[k] = c->pos[k < c->dim ? k : 0];
    a.rv = rec_view<T>(c);
    a.nstart = c->tree.start; a.nend = c->tree.end; a.nskip = c->tree.skip; a.nlevel = c->tree.level;
    a.com = c->tree.com; a.edge = c->tree.edge; a.mass = c->tree.mass;
    for (int k = 0; k < 3; ++k) a.acc[k] = c->acc[k];
    a.phi = c->sc[S_PHI];
    a.theta2 = c->prm.theta * c->prm.theta;
    a.G = c->prm.gravity_constant;
    a.box = c->box;
    a.stats = c->dstats;
    const dim3 g(grid_for(c->n, 256)), b(256);
    if (c->dim == 1) hipLaunchKernelGGL((k_gravity<1, T>), g, b, 0, c->stream, a);
    else if (c->dim == 2) hipLaunchKernelGGL((k_gravity<2, T>), g, b, 0, c->stream, a);
    else hipLaunchKernelGGL((k_gravity<3, T>), g, b, 0, c->stream, a);
}

} // namespace

int gravity(sphgpu_ctx *c)
{
    if (!c->prm.use_gravity) return 0;
    if (!c->tree.valid) return fail(c, SPHGPU_E_STATE, "gravity needs a built tree");
    if (c->mirror_dirty) {
        int rc = refresh_mirrors(c);
        if (rc) return rc;
    }
    time_begin(c, "gravity");
    if (c->prm.pair_precision == SPHGPU_PAIR_FP64) launch_gravity<double>(c);
    else launch_gravity<float>(c);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    return 0;
}

} // namespace sphgpu
