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
// A wavefront owns 64 consecutive particles in key order (a compact region)
// and walks the pre-ordered node array once for all of them: at each node
// every lane takes the reference's opening decision for its own particle in
// FP64; lanes that accept the node add its monopole and sit out its subtree,
// and the wave descends while any lane opens.  An opened leaf's members are
// loaded once per wave into LDS and every opening lane sums its pair terms
// from there.  Each particle therefore interacts with exactly the nodes and
// leaf members of its own reference walk, in the same pre-order, while node
// records are wave-uniform loads and member records are loaded once per wave
// instead of once per lane.  The pair and monopole terms run in the pair
// precision T with compensated sums taken in blocks of at most 16 terms.
#include <hip/hip_runtime.h>

#include <climits>

#include "ctx.hpp"

namespace sphgpu {

namespace {

// order LDS accesses between the lanes of one wavefront
SPH_DEV void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Hernquist & Katz (1989) softened potential and force factors
// (src/bhtree.cpp:319-345), epsilon = h/2: f(r, h) for the potential, g(r, h)
// for the force (acceleration = -r_ij G m g).  FP64: the reference's
// expressions.
SPH_DEV double hk_f(double r, double h)
{
    const double e = h * 0.5;
    const double u = r / e;
    if (u < 1.0) return (-0.5 * u * u * (1.0 / 3.0 - 3.0 / 20 * u * u + u * u * u / 20) + 1.4) / e;
    else if (u < 2.0) return -1.0 / (15 * r) + (-u * u * (4.0 / 3.0 - u + 0.3 * u * u - u * u * u / 30) + 1.6) / e;
    else return 1 / r;
}

SPH_DEV double hk_g(double r, double h)
{
    const double e = h * 0.5;
    const double u = r / e;
    if (u < 1.0) return (4.0 / 3.0 - 1.2 * u * u + 0.5 * u * u * u) / (e * e * e);
    else if (u < 2.0)
        return (-1.0 / 15 + 8.0 / 3 * u * u * u - 3 * u * u * u * u + 1.2 * u * u * u * u * u - u * u * u * u * u * u / 6.0) /
               (r * r * r);
    else return 1 / (r * r * r);
}

// FP32: the same polynomials with the reciprocals 1/e and 1/r formed once per
// pair side and every branch evaluated (the lanes of a wave take different
// branches for the same neighbour), selected by u
SPH_DEV void hk_fg(float r, float inv_r, float inv_e, float &f, float &g)
{
    const float u = r * inv_e;
    const float u2 = u * u, u3 = u2 * u;
    const float inv_r3 = inv_r * inv_r * inv_r;
    const float inv_e3 = inv_e * inv_e * inv_e;
    const float fa = (-0.5f * u2 * (1.0f / 3.0f - (3.0f / 20.0f) * u2 + u3 * (1.0f / 20.0f)) + 1.4f) * inv_e;
    const float fb = -(1.0f / 15.0f) * inv_r + (-u2 * (4.0f / 3.0f - u + 0.3f * u2 - u3 * (1.0f / 30.0f)) + 1.6f) * inv_e;
    const float ga = (4.0f / 3.0f - 1.2f * u2 + 0.5f * u3) * inv_e3;
    const float gb = (-1.0f / 15.0f + u3 * (8.0f / 3.0f - 3.0f * u + 1.2f * u2 - u3 * (1.0f / 6.0f))) * inv_r3;
    f = u < 1.0f ? fa : (u < 2.0f ? fb : inv_r);
    g = u < 1.0f ? ga : (u < 2.0f ? gb : inv_r3);
}

template <typename T> struct GravArgs {
    int n, n_nodes;
    const double *x64[3];
    RecView<T> rv;
    const GravNode *__restrict__ node;
    double *acc[3];
    double *phi;
    double theta2, G;
    Box box;
    long long *stats;
};

// one record per node, pre-order: centre of mass, edge^2 (the reference's
// e * e), mass, member range, subtree end and signed level
__global__ void k_grav_nodes(int n_nodes, int cap, const double *__restrict__ com, const double *__restrict__ edge,
                             const double *__restrict__ mass, const int *__restrict__ start,
                             const int *__restrict__ end, const int *__restrict__ skip, const int *__restrict__ level,
                             GravNode *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_nodes) return;
    GravNode g;
    g.cx = com[k];
    g.cy = com[(size_t)cap + k];
    g.cz = com[2 * (size_t)cap + k];
    g.l2 = edge[k] * edge[k];
    g.mass = mass[k];
    g.pad = 0.0;
    g.start = start[k];
    g.end = end[k];
    g.skip = skip[k];
    g.level = level[k];
    out[k] = g;
}

template <typename T> struct GravStage;
template <> struct GravStage<float> {
    float4 p[kWave];   // x, y, z, 1/e_j
    float gm[kWave];   // G m_j
};
template <> struct GravStage<double> {
    double4 p[kWave];  // x, y, z, h_j
    double gm[kWave];
};

// minimum-image component for a periodic box, the plain difference otherwise
// (P is the box's periodic flag, fixed per launch)
template <bool P, typename T> SPH_DEV T grav_rij(T xi, T xj, const Box &b, int k)
{
    return P ? rij_axis<T>(xi, xj, b, k) : xi - xj;
}

template <int D, typename T, bool P>
__global__ __launch_bounds__(256) void k_gravity(GravArgs<T> a)
{
    __shared__ GravStage<T> stage[4];
    GravStage<T> &st = stage[(threadIdx.x >> 6) & 3];
    const int lane = threadIdx.x & (kWave - 1);
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if ((i0 & ~(kWave - 1)) >= a.n) return;  // whole wave beyond the particles
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
    // half-box ties of the FP32 minimum image: the FP32 positions and their
    // difference carry at most 4 u X of rounding (X the largest coordinate
    // magnitude); 2^-18 (X + L) covers it
    float tie_band[3] = {0.f, 0.f, 0.f};
    if constexpr (P) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double X = fmax(fabs(a.box.lo[k]), fabs(a.box.hi[k]));
            tie_band[k] = (float)((X + a.box.range[k]) * 3.814697265625e-06);
        }
    }
    const T inv_ei = T(2) / hi;
    const T G = T(a.G);
    // compensated sums in blocks of at most 16 terms (fold below)
    BlockAcc<T> phi, ac[3];
    auto fold = [&]() {
        phi.fold();
#pragma unroll
        for (int k = 0; k < D; ++k) ac[k].fold();
    };
    int since_fold = 0;  // node steps since the last fold (wave-uniform)
    long long inter = 0;
    // the lane takes part in the walk from node `resume` on; a lane that
    // accepted a node's monopole resumes at the node's subtree end
    int resume = valid ? 0 : INT_MAX;
    int node = 0;
    while (node < a.n_nodes) {
        // a lane adds at most one monopole per node step
        if (++since_fold == 16) {
            fold();
            since_fold = 0;
        }
        const GravNode g = a.node[node];
        const bool act = node >= resume;
        double d[3] = {0, 0, 0};
        const double cm[3] = {g.cx, g.cy, g.cz};
        double d2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            d[k] = grav_rij<P, double>(ri[k], cm[k], a.box, k);
            d2 = (k == 0) ? d[k] * d[k] : d2 + d[k] * d[k];
        }
        // BHNode::calc_force opening test (src/bhtree.cpp:350-358)
        const bool open = act && ((g.l2 > a.theta2 * d2) || (d2 < g.l2));
        if (act && !open) {
            // monopole (src/bhtree.cpp:387-392)
            const T r_inv = (sizeof(T) == 4) ? trsqrt(T(d2)) : tdiv(T(1.0), tsqrt(T(d2)));
            const T gm = G * T(g.mass);
            phi.add(-(gm * r_inv));
            const T s = gm * (r_inv * r_inv * r_inv);
#pragma unroll
            for (int k = 0; k < D; ++k) ac[k].add(-(T(d[k]) * s));
            ++inter;
            resume = g.skip;
        }
        if (__ballot(open) == 0ull) {
            node = g.skip;
            continue;
        }
        if (g.level < 0) {
            // opened leaf: members staged 64 at a time, pair terms with every
            // member j != i (src/bhtree.cpp:362-378)
            for (int s0 = g.start; s0 < g.end; s0 += kWave) {
                const int cnt = min(kWave, g.end - s0);
                if (lane < cnt) {
                    const V4 pj = RP[s0 + lane];
                    const T mj = RV[s0 + lane].w;
                    if constexpr (sizeof(T) == 4) st.p[lane] = make_float4(pj.x, pj.y, pj.z, 2.0f / pj.w);
                    else st.p[lane] = pj;
                    st.gm[lane] = G * mj;
                }
                wave_lds_fence();
                if (open) {
                    fold();
                    for (int t = 0; t < cnt; ++t) {
                        if ((t & 15) == 15) fold();
                        if (s0 + t == i) continue;
                        const typename V4T<T>::type q = st.p[t];
                        const T xj[3] = {q.x, q.y, q.z};
                        T rij[3] = {0, 0, 0};
                        T r2 = T(0);
#pragma unroll
                        for (int k = 0; k < D; ++k) rij[k] = grav_rij<P, T>(xi[k], xj[k], a.box, k);
                        if constexpr (P && sizeof(T) == 4) {
                            // a component within FP32 rounding of half the
                            // box takes its image from the FP64 positions, as
                            // the reference chooses it (the image, and with
                            // it the sign, flips there)
                            bool tie = false;
#pragma unroll
                            for (int k = 0; k < D; ++k)
                                tie = tie || fabsf(fabsf(xi[k] - xj[k]) - 0.5f * a.box.rangef[k]) <= tie_band[k];
                            if (tie)
#pragma unroll
                                for (int k = 0; k < D; ++k)
                                    rij[k] = (T)rij_axis<double>(ri[k], a.x64[k][s0 + t], a.box, k);
                        }
#pragma unroll
                        for (int k = 0; k < D; ++k) r2 = (k == 0) ? rij[k] * rij[k] : r2 + rij[k] * rij[k];
                        const T gm = st.gm[t];
                        T fsum, gsum;
                        if constexpr (sizeof(T) == 4) {
                            // 1/r from one reciprocal square root, r = r^2 / r
                            // (r = 0 for coincident particles: the polynomial
                            // branch u < 1 takes no 1/r)
                            const float inv_r = trsqrt(r2);
                            const float r = r2 > 0.f ? r2 * inv_r : 0.f;
                            // beyond both softening lengths (u >= 2 on both
                            // sides, the usual case in an opened leaf) the
                            // terms are 1/r and 1/r^3; the polynomials run
                            // only when some lane of the wave has a near pair
                            const bool near = r * inv_ei < 2.0f || r * q.w < 2.0f;
                            if (__ballot(near)) {
                                float fi, gi, fj, gj;
                                hk_fg(r, inv_r, inv_ei, fi, gi);
                                hk_fg(r, inv_r, q.w, fj, gj);
                                fsum = fi + fj;
                                gsum = gi + gj;
                            } else {
                                fsum = 2.0f * inv_r;
                                gsum = 2.0f * (inv_r * inv_r * inv_r);
                            }
                        } else {
                            const T r = tsqrt(r2);
                            fsum = hk_f(r, hi) + hk_f(r, q.w);
                            gsum = hk_g(r, hi) + hk_g(r, q.w);
                        }
                        phi.add(-(gm * fsum * T(0.5)));
                        const T s2 = gm * gsum * T(0.5);
#pragma unroll
                        for (int k = 0; k < D; ++k) ac[k].add(-(rij[k] * s2));
                        ++inter;
                    }
                    fold();
                }
                wave_lds_fence();
            }
        }
        ++node;
    }
    if (valid) {
        a.phi[i] = (double)phi.get();
#pragma unroll
        for (int k = 0; k < D; ++k) a.acc[k][i] += (double)ac[k].get();
    }
    for (int off = kWave / 2; off > 0; off >>= 1) inter += __shfl_xor(inter, off, kWave);
    if (lane == 0) atomicAdd(stat_ptr(a.stats, 10), (unsigned long long)inter);
}

template <typename T> void launch_gravity(sphgpu_ctx *c)
{
    GravArgs<T> a;
    a.n = c->n;
    a.n_nodes = c->tree.n_nodes;
    for (int k = 0; k < 3; ++k) a.x64[k] = c->pos[k < c->dim ? k : 0];
    a.rv = rec_view<T>(c);
    a.node = c->tree.gnode;
    for (int k = 0; k < 3; ++k) a.acc[k] = c->acc[k];
    a.phi = c->sc[S_PHI];
    a.theta2 = c->prm.theta * c->prm.theta;
    a.G = c->prm.gravity_constant;
    a.box = c->box;
    a.stats = c->dstats;
    const dim3 g(grid_for(c->n, 256)), b(256);
#define LAUNCH(D)                                                                          \
    if (c->box.periodic) hipLaunchKernelGGL((k_gravity<D, T, true>), g, b, 0, c->stream, a); \
    else hipLaunchKernelGGL((k_gravity<D, T, false>), g, b, 0, c->stream, a);
    if (c->dim == 1) { LAUNCH(1) } else if (c->dim == 2) { LAUNCH(2) } else { LAUNCH(3) }
#undef LAUNCH
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
    Tree &t = c->tree;
    if (t.gnode_cap < t.n_nodes) {
        if (t.gnode) hipFree(t.gnode);
        t.gnode = nullptr;
        t.gnode_cap = 0;
        const int cap = t.n_nodes + t.n_nodes / 4 + 64;
        if (hipMalloc((void **)&t.gnode, sizeof(GravNode) * (size_t)cap) != hipSuccess)
            return fail(c, SPHGPU_E_NOMEM, "gravity node records");
        t.gnode_cap = cap;
    }
    time_begin(c, "gravity_nodes");
    hipLaunchKernelGGL(k_grav_nodes, dim3(grid_for(t.n_nodes, 256)), dim3(256), 0, c->stream, t.n_nodes, t.node_cap,
                       t.com, t.edge, t.mass, t.start, t.end, t.skip, t.level, t.gnode);
    time_end(c);
    time_begin(c, "gravity");
    if (c->prm.pair_precision == SPHGPU_PAIR_FP64) launch_gravity<double>(c);
    else launch_gravity<float>(c);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    return 0;
}

} // namespace sphgpu
