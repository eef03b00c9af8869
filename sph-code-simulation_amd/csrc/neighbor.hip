// neighbor.hip -- neighbour lists from a wavefront-per-group tree walk.
//
// Reference: BHTree::neighbor_search / BHNode::neighbor_search
// (src/bhtree.cpp:122-144, 277-316).  The reference walks the tree once per
// particle, visiting a node when the Chebyshev distance from the particle to
// the node centre is <= edge/2 + h, and keeps j when r_ij^2 < h^2 (FP64).
// FluidForce searches with h = max(h_i, kernel_size) and then keeps the pair
// when |r_ij| < max(h_i, h_j) and r_ij != 0 (src/fluid_force.cpp:62,
// src/gsph/g_fluid_force.cpp:93).
//
// Here a 64-lane wavefront owns 64 consecutive particles in key order (a
// compact region of space).  The wave walks the pre-ordered node array
// stacklessly: a node is entered when the group's bounding box, widened by
// the group's largest radius (and the node's kernel_size for is_ij), can reach
// the node cell; otherwise the walk jumps to the node's skip index.  In an
// entered leaf every lane evaluates the reference's FP64 predicate for its own
// particle (behind an FP32 prefilter that only settles candidates whose FP32
// distance lies outside a proven rounding band), so the resulting sets are
// exactly the reference's:
//   is_ij = 0 :  r_ij^2 < hs_i^2
//   is_ij = 1 :  sqrt(r_ij^2) < max(h_i, h_j)  and  r_ij^2 != 0
// The group test only prunes; it never decides membership.
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace sphgpu {

namespace {

struct SearchArgs {
    int n;
    int n_nodes;
    const double *x[3];
    const double *radius;     // hs_i (is_ij = 0) or h_i (is_ij = 1)
    const double *sml;        // h_j for is_ij
    const float4 *posf;       // FP32 (x, y, z, h) records for the prefilter
    const double *root;       // tree root geometry (bounds |x|)
    const int *nstart, *nend, *nskip, *nlevel;
    const double *ncenter;    // [3][cap]
    const double *nedge, *nks;
    int node_cap;
    Box box;
    int *idx;                 // [cap][n]
    int *cnt;
    int cap;
    int *overflow;
    long long *stats;
};

template <int D> SPH_DEV double r2_exact(const double xi[3], const double xj[3], const Box &b)
{
    // inner_product(r_ij, r_ij) in the reference order (vector_type.hpp:192-201)
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double d = rij_axis<double>(xi[k], xj[k], b, k);
        s = (k == 0) ? d * d : s + d * d;
    }
    return s;
}

template <int D> SPH_DEV double wave_min(double v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}
template <int D> SPH_DEV double wave_max(double v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// particles per "chunk" node: a wavefront walks once per chunk it touches
constexpr int kChunk = 128;

template <int D, int IS_IJ>
__global__ __launch_bounds__(256) void k_search(SearchArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int group = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int i = group * kWave + lane;
    if (group * kWave >= a.n) return;
    const bool valid = i < a.n;
    const int ii = valid ? i : (a.n - 1);
    double xi[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = a.x[k][ii];
    const double R = a.radius[ii];
    const float4 pfi = a.posf[ii];
    const float xf[3] = {pfi.x, pfi.y, pfi.z};
    const float Rf = (float)R;
    const float mk = fp32_margin_k(a.root, D);
    int count = 0;
    bool over = false;
    long long visits = 0, tests = 0;
    const int p_end = min(group * kWave + kWave, a.n);
    int sub = group * kWave;
    // Consecutive particles in key order can straddle the boundary of a large
    // cell, which would make the group's bounding box span that cell.  The
    // wave therefore splits its particles at the boundaries of "chunk" nodes
    // (the deepest nodes holding their start particle with <= kChunk members)
    // and walks once per piece.
    while (sub < p_end) {
        int cn = 0;
        while (a.nlevel[cn] > 0 && a.nend[cn] - a.nstart[cn] > kChunk) {
            int ch = cn + 1;
            while (!(a.nstart[ch] <= sub && sub < a.nend[ch])) ch = a.nskip[ch];
            cn = ch;
        }
        const int sub_end = min(p_end, a.nend[cn]);
        const bool mine = valid && i >= sub && i < sub_end;
        // bounding box and radius of this piece
        double g[3] = {0, 0, 0}, hb[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double lo = wave_min<D>(mine ? xi[k] : 1.7976931348623157e308);
            const double hi = wave_max<D>(mine ? xi[k] : -1.7976931348623157e308);
            g[k] = 0.5 * (lo + hi);
            hb[k] = 0.5 * (hi - lo);
        }
        const double Rg = wave_max<D>(mine ? R : 0.0);
        int node = 0;
        while (node < a.n_nodes) {
            ++visits;
            const double e = a.nedge[node];
            double rn = Rg;
            if (IS_IJ) {
                const double ks = a.nks[node];
                rn = ks > rn ? ks : rn;
            }
            // conservative reach test, padded against rounding
            const double lim = (0.5 * e + rn) * (1.0 + 1e-12) + 1e-300;
            bool reach = true;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double c = a.ncenter[(size_t)k * a.node_cap + node];
                const double d = tabs(rij_axis<double>(c, g[k], a.box, k)) - hb[k];
                reach = reach && (d <= lim);
            }
            const int lv = a.nlevel[node];
            if (!reach) {
                node = a.nskip[node];
                continue;
            }
            if (lv > 0) {  // internal node: descend
                ++node;
                continue;
            }
            const int s = a.nstart[node], en = a.nend[node];
            tests += en - s;
            for (int j = s; j < en; ++j) {
                if (!mine) continue;
                // FP32 prefilter; the FP64 predicate decides inside its band
                const float4 pj = a.posf[j];
                const float pjx[3] = {pj.x, pj.y, pj.z};
                float r2f = 0.f;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const float d = rij_axis<float>(xf[k], pjx[k], a.box, k);
                    r2f = (k == 0) ? d * d : r2f + d * d;
                }
                int band = fp32_band(r2f, IS_IJ ? fmaxf(Rf, pj.w) : Rf, mk);
                if (IS_IJ && r2f == 0.f) band = 0;  // r_ij != 0 needs the FP64 test
                bool keep = band > 0;
                if (band == 0) {
                    double xj[3] = {0, 0, 0};
#pragma unroll
                    for (int k = 0; k < D; ++k) xj[k] = a.x[k][j];
                    const double r2 = r2_exact<D>(xi, xj, a.box);
                    if (IS_IJ) {
                        const double hj = a.sml[j];
                        const double H = (R < hj) ? hj : R;  // std::max(h_i, h_j)
                        keep = r_less(r2, H) && r2 != 0.0;
                    } else {
                        keep = r2 < R * R;
                    }
                }
                if (keep) {
                    if (count < a.cap) a.idx[list_slot(i, count, a.cap)] = j;
                    else over = true;
                    ++count;
                }
            }
            ++node;
        }
        sub = sub_end;
    }
    if (valid) a.cnt[i] = count < a.cap ? count : a.cap;
    if (over) atomicOr(a.overflow, 1);
    // candidate statistics
    long long c = valid ? count : 0;
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    if (lane == 0) {
        atomicAdd((unsigned long long *)&a.stats[IS_IJ ? 4 : 0], (unsigned long long)c);
        atomicAdd((unsigned long long *)&a.stats[13], (unsigned long long)visits);
        atomicAdd((unsigned long long *)&a.stats[14], (unsigned long long)tests);
    }
}

} // namespace

int lists_alloc(sphgpu_ctx *c, Lists &l, int cap)
{
    lists_free(l);
    l.cap = cap;
    const size_t blocks = ((size_t)c->n + 63) / 64;
    if (hipMalloc((void **)&l.idx, sizeof(int) * (size_t)cap * blocks * 64) != hipSuccess) return SPHGPU_E_NOMEM;
    if (hipMalloc((void **)&l.cnt, sizeof(int) * (size_t)c->n) != hipSuccess) return SPHGPU_E_NOMEM;
    if (hipMemset(l.cnt, 0, sizeof(int) * (size_t)c->n) != hipSuccess) return SPHGPU_E_HIP;
    return 0;
}

void lists_free(Lists &l)
{
    if (l.idx) hipFree(l.idx);
    if (l.cnt) hipFree(l.cnt);
    l = Lists();
}

int search(sphgpu_ctx *c, Lists &l, int is_ij, const double *radius)
{
    Tree &t = c->tree;
    if (!t.valid) return fail(c, SPHGPU_E_STATE, "neighbour search without a tree");
    for (int attempt = 0; attempt < 6; ++attempt) {
        SearchArgs a;
        a.n = c->n;
        a.n_nodes = t.n_nodes;
        for (int k = 0; k < 3; ++k) a.x[k] = c->pos[k < c->dim ? k : 0];
        a.radius = radius;
        a.sml = c->sc[S_SML];
        a.posf = c->rec32[R_POSH];
        a.root = t.root;
        a.nstart = t.start; a.nend = t.end; a.nskip = t.skip; a.nlevel = t.level;
        a.ncenter = t.center; a.nedge = t.edge; a.nks = t.ks; a.node_cap = t.node_cap;
        a.box = c->box;
        a.idx = l.idx; a.cnt = l.cnt; a.cap = l.cap;
        a.overflow = c->dflag;
        a.stats = c->dstats;
        HIPCHK(c, hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
        const int groups = (c->n + kWave - 1) / kWave;
        const dim3 grid(grid_for((long long)groups * kWave, 256)), block(256);
        time_begin(c, is_ij ? "search_force" : "search_pre");
#define LAUNCH(D)                                                                                   \
    if (is_ij) hipLaunchKernelGGL((k_search<D, 1>), grid, block, 0, c->stream, a);                  \
    else hipLaunchKernelGGL((k_search<D, 0>), grid, block, 0, c->stream, a);
        if (c->dim == 1) { LAUNCH(1) } else if (c->dim == 2) { LAUNCH(2) } else { LAUNCH(3) }
#undef LAUNCH
        time_end(c);
        HIPCHK(c, hipGetLastError());
        int over = 0;
        HIPCHK(c, hipMemcpyAsync(&over, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!over) return 0;
        // grow the list and search again (the reference's capacity is nn * 1000)
        const long long want = 2ll * l.cap;
        if (want > 1000ll * c->prm.neighbor_number)
            return fail(c, SPHGPU_E_NEIGHBOR_OVERFLOW, "Neighbor list overflow: increase neighbor_list_size in defines.hpp");
        int rc = lists_alloc(c, l, (int)want);
        if (rc) return fail(c, rc, "neighbour list reallocation failed");
    }
    return fail(c, SPHGPU_E_NEIGHBOR_OVERFLOW, "Neighbor list overflow");
}

} // namespace sphgpu
