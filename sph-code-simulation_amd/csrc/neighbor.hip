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
// stacklessly, 64 nodes per step: lane q tests node w0 + q against the
// group's FP32 bounding box widened by the group's largest radius (and the
// node's kernel_size for is_ij), conservatively padded for FP32 rounding; the
// walk then advances through the window on the wave-uniform reach mask
// (entered -> next node, pruned -> skip) and loads the next window where it
// leaves this one.  Members of entered leaves are queued, 64 at a time, in a
// per-lane candidate buffer; a flush loads the 64 candidate records in one
// gather and broadcasts them lane by lane, and every lane evaluates the
// reference's FP64 predicate for its own particle (behind an FP32 prefilter
// that only settles candidates whose FP32 distance lies outside a proven
// rounding band), so the resulting sets are exactly the reference's:
//   is_ij = 0 :  r_ij^2 < hs_i^2
//   is_ij = 1 :  sqrt(r_ij^2) < max(h_i, h_j)  and  r_ij^2 != 0
// The group test only prunes; it never decides membership.
#include <hip/hip_runtime.h>
#include <algorithm>

#include "ctx.hpp"

namespace sphgpu {

namespace {

struct SearchArgs {
    int n;
    int n_nodes;
    const double *x[3];
    const double *radius;     // hs_i (is_ij = 0) or h_i (is_ij = 1)
    double rscale;            // the search radius is radius[i] * rscale
    const double *sml;        // h_j for is_ij
    const float4 *posf;       // FP32 (x, y, z, h) records for the prefilter
    const double *root;       // tree root geometry (bounds |x|)
    const float4 *nbox;       // FP32 node boxes: centre, half extent x (rounded up)
    const float2 *nhalf;      // half extents y, z (rounded up)
    const int4 *nlink;        // start, end, skip, leaf flag | kernel_size
    const int2 *nchild;       // (offset into child, number of children)
    const int *child;         // children of each node, contiguous
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

// particles per "chunk" node: a wavefront walks once per chunk it touches
constexpr int kChunk = 128;
// nodes a breadth-first frontier can hold (per wavefront, in LDS)
constexpr int kFrontier = 512;
// relative widening of FP32 reach radii (covers rounding of radii and extents)
constexpr float kReachRel = 1.000003814697265625f;

// order LDS accesses between the lanes of one wavefront
SPH_DEV void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

SPH_DEV float rdl(float v, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane)); }
SPH_DEV int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// |minimum-image distance| along one axis in FP32: r^2 only needs |d|, and
// |d| < L inside the box, so min(|d|, L - |d|) is the minimum image (the
// reference's tie order only decides the sign)
SPH_DEV float abs_min_image_f(float d, int periodic, float L)
{
    const float a = fabsf(d);
    return periodic ? fminf(a, fabsf(L - a)) : a;
}

// FP32 test of the broadcast candidates [t0, t1) (at most 32) against this
// lane's particle: bit t - t0 of km is set when the candidate is a member for
// certain, of ku when the FP32 distance falls inside the rounding band (or,
// for is_ij, is near zero) and the FP64 predicate has to decide.  SHIFTED:
// candidates were moved to their image nearest the group centre.
template <int D, int IS_IJ, bool SHIFTED>
SPH_DEV void test_range(const float4 *srec, int t0, int t1, const float xf[3], const float Lf[3], float Rf,
                        float R2lo, float R2hi, float delta, float z2, unsigned &km, unsigned &ku)
{
    unsigned kk = 0, uu = 0;
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
        const float4 c = srec[t];
        float r2f;
        if (SHIFTED) {
            const float dx = xf[0] - c.x;
            const float dy = D > 1 ? xf[1] - c.y : 0.f;
            const float dz = D > 2 ? xf[2] - c.z : 0.f;
            r2f = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        } else {
            const float cj[3] = {c.x, c.y, c.z};
            r2f = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float dk = abs_min_image_f(xf[k] - cj[k], 1, Lf[k]);
                r2f = (k == 0) ? dk * dk : r2f + dk * dk;
            }
        }
        float lo2 = R2lo, hi2 = R2hi;
        if (IS_IJ) {
            const float Hf = fmaxf(Rf, c.w);
            lo2 = Hf * Hf * (1.0f - delta);
            hi2 = Hf * Hf * (1.0f + delta);
        }
        const unsigned bit = 1u << (t - t0);
        bool keep = r2f < lo2;
        bool und = !keep && !(r2f > hi2);
        if (IS_IJ) {
            // r_ij != 0: the pre-shift need not reproduce a coincident
            // position exactly, so near-zero FP32 distances go to FP64
            const bool n0 = !(r2f > z2);
            und = und || n0;
            keep = keep && !n0;
        }
        kk |= keep ? bit : 0u;
        uu |= und ? bit : 0u;
    }
    km = kk;
    ku = uu;
}

// The same test for the radius search (is_ij = 0) on candidates stored
// relative to the group centre g: record (y, |y|^2) with y = x_j - g (its
// image nearest g).  r^2 = |y_i|^2 + |y_j|^2 - 2 y_i . y_j, and the lane's
// |y_i|^2 moves to the bounds: s = |y_j|^2 - 2 y_i . y_j is D fused
// multiply-adds (m2y = -2 y_i), tested against Tlo = R2lo - |y_i|^2 and
// Thi = R2hi - |y_i|^2 (formed once per piece).  The candidates run from
// last to first and each decision is shifted in at bit 0 with a carry-in add
// (v_addc_co_u32 d, d + d + vcc), one instruction per mask; km gets
// s < Tlo, kle gets !(s > Thi), the band is kle & ~km.
template <int D>
SPH_DEV void test_range_expand(const float4 *srec, int t0, int t1, const float m2y[3], float Tlo, float Thi,
                               unsigned &km, unsigned &ku)
{
    unsigned kk = 0, kl = 0;
#pragma unroll 4
    for (int t = t1 - 1; t >= t0; --t) {
        const float4 c = srec[t];
        float sf = c.w;
        if (D > 2) sf = __builtin_fmaf(m2y[2], c.z, sf);
        if (D > 1) sf = __builtin_fmaf(m2y[1], c.y, sf);
        sf = __builtin_fmaf(m2y[0], c.x, sf);
        asm("v_cmp_gt_f32 vcc, %[lo], %[s]\n\t"
            "v_addc_co_u32 %[kk], vcc, %[kk], %[kk], vcc\n\t"
            "v_cmp_nlt_f32 vcc, %[hi], %[s]\n\t"
            "v_addc_co_u32 %[kl], vcc, %[kl], %[kl], vcc"
            : [kk] "+v"(kk), [kl] "+v"(kl)
            : [s] "v"(sf), [lo] "v"(Tlo), [hi] "v"(Thi)
            : "vcc");
    }
    km = kk;
    ku = kl & ~kk;
}

template <int D, int IS_IJ>
__global__ __launch_bounds__(256, 4) void k_search(SearchArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int group = (xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) / kWave;
    const int i = group * kWave + lane;
    if (group * kWave >= a.n) return;
    const bool valid = i < a.n;
    const int ii = valid ? i : (a.n - 1);
    const float4 pfi = a.posf[ii];
    const float xf[3] = {pfi.x, pfi.y, pfi.z};
    const float Rf = (float)(a.radius[ii] * a.rscale);
    // three times the pair kernels' band: pre-shifting a candidate to the
    // image nearest the group centre adds up to 5 u X of rounding to each
    // component (10 u X + u |r| in all, against the 5 u X + u |r| the band
    // is sized for with a factor 2)
    const float mk = 3.0f * fp32_margin_k(a.root, D);
    // candidate band of this lane (for is_ij the band of H = max(h_i, h_j) >= h_i
    // is at most this wide relative to H^2)
    const float delta = mk / Rf + kMarginC;
    const bool fast = (Rf * Rf > 1e-30f) && (delta < 0.5f);
    float Lf[3] = {1.f, 1.f, 1.f};
#pragma unroll
    for (int k = 0; k < D; ++k) Lf[k] = a.box.rangef[k];
    // reach-test padding: FP32 rounding of centres, positions and the box
    // shift is below 10 u X; 2^-18 X and a relative 2^-18 cover it
    double Xb = 0.0;
    for (int k = 0; k < D; ++k) Xb = fmax(Xb, fabs(a.root[k]));
    const float padX = (float)((Xb + a.root[3]) * 3.814697265625e-06);
    // FP32 r^2 below this may come from a coincident pair (rounding of the
    // pre-shift is below 9 u X per component): (2^-18 X)^2 leaves room
    const float z2 = padX * padX;
    int count = 0;
    bool over = false;
    // this particle's column of the list: entry q at col[64 q]
    int *const col = a.idx + list_slot(ii, 0, a.cap);
    int visits = 0, tests = 0, windows = 0, leaves = 0;
    int reach = 0;
    const int p_end = min(group * kWave + kWave, a.n);
    int sub = group * kWave;
    // candidate buffer: lane q holds the q-th pending candidate index
    int cand = 0, nb = 0;

    // per-wave LDS broadcast buffer of the pending candidates
    __shared__ float4 s_rec[4][kWave];
    __shared__ int s_idx[4][kWave];
    float4 *srec = s_rec[(threadIdx.x / kWave) & 3];
    int *sidx = s_idx[(threadIdx.x / kWave) & 3];
    // per-wave breadth-first frontiers (current and next level)
    __shared__ int s_fr[4][2][kFrontier];
    int *sfr[2] = {s_fr[(threadIdx.x / kWave) & 3][0], s_fr[(threadIdx.x / kWave) & 3][1]};
    // per-piece state used by flush: group centre, and whether candidates can
    // be pre-shifted to the minimum image seen from the group centre
    float g[3] = {0, 0, 0};
    // and the piece's half extents and largest radius (the staging cull)
    float hb[3] = {0, 0, 0};
    float Rg = 0.f;
    int shift = 1;
    // radius search with candidates relative to g (EXPAND): this lane's
    // -2 (x_i - g), |x_i - g|^2 and its band widened for that form
    const bool expand_ok = !IS_IJ;
    bool expand = false;
    float m2y[3] = {0, 0, 0}, yyi = 0.f;
    // this piece's band bounds on r^2 (direct or relative form)
    float R2lo = 0.f, R2hi = 0.f;
    // and the relative form's bounds on r^2 - |y_i|^2
    float Tlo = 0.f, Thi = 0.f;

    // the FP64 predicate (the reference's) for candidate j; the particle's
    // own FP64 position and radius are reloaded here (a rare path) rather
    // than held in registers for the whole walk
    int nexact = 0;
    auto exact = [&](int j) -> bool {
        ++nexact;
        double xi[3] = {0, 0, 0}, xj64[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            xi[k] = a.x[k][ii];
            xj64[k] = a.x[k][j];
        }
        const double R = a.radius[ii] * a.rscale;
        const double r2 = r2_exact<D>(xi, xj64, a.box);
        if (IS_IJ) {
            const double hj = a.sml[j];
            const double H = (R < hj) ? hj : R;  // std::max(h_i, h_j)
            return r_less(r2, H) && r2 != 0.0;
        }
        return r2 < R * R;
    };

    // test the nb pending candidates against this lane's particle
    auto flush = [&](bool mine) {
        // each loading lane prepares its candidate's record (x, y, z, h);
        // with `shift` the position is moved to its image nearest the group
        // centre, which for every pair closer than the search radius is the
        // minimum image seen from the group's particles; the relative form
        // stores (y, |y|^2) with y that image minus the centre.  A candidate
        // farther from the piece's box than the piece's largest radius (with
        // the node reach test's padding) is within no lane's radius and is
        // dropped; the kept ones are compacted in candidate order (the
        // leaves' members beyond the group's reach are most of what the leaf
        // granularity adds)
        float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
        bool keepc = false;
        if (lane < nb) {
            const float4 pj = a.posf[cand];
            const float q[3] = {pj.x, pj.y, pj.z};
            float dd[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float d = q[k] - g[k];
                dd[k] = !a.box.periodic ? d : (fabsf(d) <= 0.5f * Lf[k] ? d : (d > 0.f ? d - Lf[k] : d + Lf[k]));
            }
            if (expand) {
                rec = make_float4(dd[0], dd[1], dd[2], __builtin_fmaf(dd[0], dd[0], __builtin_fmaf(dd[1], dd[1], dd[2] * dd[2])));
            } else {
                rec = pj;
                if (a.box.periodic && shift) {
                    rec.x = g[0] + dd[0];
                    if (D > 1) rec.y = g[1] + dd[1];
                    if (D > 2) rec.z = g[2] + dd[2];
                }
            }
            const float rc = (IS_IJ ? fmaxf(Rg, pj.w) : Rg) * kReachRel + padX;
            float e2 = 0.f;  // squared distance to the piece's box
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float e = fmaxf(fabsf(dd[k]) - hb[k], 0.f);
                e2 = __builtin_fmaf(e, e, e2);
            }
            keepc = e2 <= rc * rc;
        }
        const unsigned long long kb = __ballot(keepc);
        if (keepc) {
            const int pos = __popcll(kb & ((1ull << lane) - 1ull));
            srec[pos] = rec;
            sidx[pos] = cand;
        }
        nb = __popcll(kb);
        tests += nb;
        wave_lds_fence();
        unsigned km0 = 0, km1 = 0, ku0 = 0, ku1 = 0;
        const int n0 = min(nb, 32);
        if (expand) {
            test_range_expand<D>(srec, 0, n0, m2y, Tlo, Thi, km0, ku0);
            if (nb > 32) test_range_expand<D>(srec, 32, nb, m2y, Tlo, Thi, km1, ku1);
        } else if (!a.box.periodic || shift) {
            test_range<D, IS_IJ, true>(srec, 0, n0, xf, Lf, Rf, R2lo, R2hi, delta, z2, km0, ku0);
            if (nb > 32) test_range<D, IS_IJ, true>(srec, 32, nb, xf, Lf, Rf, R2lo, R2hi, delta, z2, km1, ku1);
        } else {
            test_range<D, IS_IJ, false>(srec, 0, n0, xf, Lf, Rf, R2lo, R2hi, delta, z2, km0, ku0);
            if (nb > 32) test_range<D, IS_IJ, false>(srec, 32, nb, xf, Lf, Rf, R2lo, R2hi, delta, z2, km1, ku1);
        }
        unsigned long long km = ((unsigned long long)km1 << 32) | km0;
        unsigned long long ku = ((unsigned long long)ku1 << 32) | ku0;
        if (!fast) {  // no FP32 band for this lane: FP64 decides every candidate
            km = 0;
            ku = (nb == 64) ? ~0ull : ((1ull << nb) - 1ull);
        }
        if (!mine) km = ku = 0;
        if (__ballot(ku != 0ull) != 0ull) {
            // inside the rounding band: the reference's FP64 predicate decides
            while (ku) {
                const int t = __ffsll((long long)ku) - 1;
                ku &= ku - 1;
                if (exact(sidx[t])) km |= 1ull << t;
            }
        }
        // append the kept candidates in candidate order, four ranks per
        // round: the four LDS reads are issued together, so a round waits
        // once for LDS instead of once per kept candidate
        const int nkl = __popcll(km);
        int nk = nkl;
        for (int off = kWave / 2; off > 0; off >>= 1) nk = max(nk, __shfl_xor(nk, off, kWave));
        if (count + nkl > a.cap) over = true;
        for (int r = 0; r < nk; r += 4) {
            int t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                t[u] = km ? __ffsll((long long)km) - 1 : 0;
                km &= km - 1;  // no-op once empty
            }
            int j[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) j[u] = sidx[t[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = count + u;
                if (r + u < nkl && q < a.cap) col[q * kWave] = j[u];
            }
            count += max(0, min(4, nkl - r));
        }
        wave_lds_fence();
        nb = 0;
    };

    // Consecutive particles in key order can straddle the boundary of a large
    // cell, which would make the group's bounding box span that cell.  The
    // wave therefore splits its particles at the boundaries of "chunk" nodes
    // (the deepest nodes holding their start particle with <= kChunk members)
    // and walks once per piece.
    while (sub < p_end) {
        int cn = 0;
        int4 lc = a.nlink[0];
        while (lc.w >= 0 && lc.y - lc.x > kChunk) {
            int ch = cn + 1;
            int4 l2 = a.nlink[ch];
            while (!(l2.x <= sub && sub < l2.y)) {
                ch = l2.z;
                l2 = a.nlink[ch];
            }
            cn = ch;
            lc = l2;
        }
        const int sub_end = min(p_end, lc.y);
        const bool mine = valid && i >= sub && i < sub_end;
        // FP32 bounding box and radius of this piece
#pragma unroll
        for (int k = 0; k < D; ++k) {
            float lo = mine ? xf[k] : 3.0e38f, hi = mine ? xf[k] : -3.0e38f;
            for (int off = kWave / 2; off > 0; off >>= 1) {
                lo = fminf(lo, __shfl_xor(lo, off, kWave));
                hi = fmaxf(hi, __shfl_xor(hi, off, kWave));
            }
            g[k] = 0.5f * (lo + hi);
            hb[k] = 0.5f * (hi - lo);
        }
        Rg = mine ? Rf : 0.f;
        for (int off = kWave / 2; off > 0; off >>= 1) Rg = fmaxf(Rg, __shfl_xor(Rg, off, kWave));
        // pre-shifting candidates is exact for pairs closer than R when
        // R + (group half extent) < L/2 on every axis (with room for rounding);
        // for is_ij the radius is max(h_i, h_j), bounded by the root's
        // kernel_size
        {
            float rmax = Rg;
            if (IS_IJ) rmax = fmaxf(rmax, fabsf(__int_as_float(a.nlink[0].w)));
            bool ok = true;
#pragma unroll
            for (int k = 0; k < D; ++k) ok = ok && (rmax + hb[k] + 2.f * padX < 0.49f * Lf[k]);
            shift = __builtin_amdgcn_readfirstlane(ok ? 1 : 0);  // wave-uniform by construction
        }
        // relative form: y_i = x_i - g.  Its FP32 sum |y_i|^2 + |y_j|^2 -
        // 2 y_i . y_j rounds by at most 6 u (|y_i| + |y_j|)^2 (first order),
        // and a candidate whose decision can change has r < 2R, so |y_j| <
        // |y_i| + 2R: the band grows by 8 u (2 |y_i| + 2R)^2 / R^2.  The
        // components y carry at most 7 u X of rounding (the positions, the
        // differences and the box range), inside the 15 u X the search's
        // band allows.
        expand = expand_ok && (!a.box.periodic || shift);
        if (expand) {
            float yy = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float y = xf[k] - g[k];
                m2y[k] = -2.0f * y;
                yy = __builtin_fmaf(y, y, yy);
            }
            yyi = yy;
            // the r > 2R part of that bound needs |y_i| well below 2^-24 / (6u)
            // times R; pieces far wider than their radii keep the direct form
            expand = __ballot(mine && !(yy <= 4096.0f * Rf * Rf)) == 0ull;
        }
        if (expand) {
            const float yy = yyi;
            const float ext = 2.0f * sqrtf(yy) + 2.0f * Rf;
            const float de = delta + 8.0f * 5.9604644775390625e-08f * (ext * ext) / (Rf * Rf);
            R2lo = Rf * Rf * (1.0f - de);
            R2hi = Rf * Rf * (1.0f + de);
            // the bounds on s = r^2 - |y_i|^2: their rounding (u (R^2 +
            // |y_i|^2)) and the FMA chain's (3 u (|y_j|^2 + 2 |y_i| |y_j|))
            // with those of |y_i|^2 and |y_j|^2 stay below 6 u (|y_i| +
            // |y_j|)^2 + u R^2, inside the 8 u (2 |y_i| + 2R)^2 above
            Tlo = R2lo - yy;
            Thi = R2hi - yy;
            if (!(fast && de < 0.5f)) {
                // no FP32 decision: every candidate is in the band
                Tlo = -__builtin_inff();
                Thi = __builtin_inff();
            }
        } else {
            R2lo = Rf * Rf * (1.0f - delta);
            R2hi = Rf * Rf * (1.0f + delta);
        }

        // reach test of node k against this piece's box (lane-parallel)
        auto reach_node = [&](const float4 &bx, const float2 &bh, const int4 &lk) -> bool {
            float rn = Rg;
            if (IS_IJ) rn = fmaxf(rn, fabsf(__int_as_float(lk.w)));
            const float bc[3] = {bx.x, bx.y, bx.z};
            const float he[3] = {bx.w, bh.x, bh.y};
            bool r = true;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                const float d = abs_min_image_f(bc[q] - g[q], a.box.periodic, Lf[q]) - hb[q];
                r = r && (d <= (he[q] + rn) * kReachRel + padX);
            }
            return r;
        };
        // an entered leaf held by lane o: queue its members if some lane's
        // own sphere reaches the leaf's box
        auto visit_leaf = [&](int o, const float4 &bx, const float2 &bh, const int4 &lk) {
            const float lc3[3] = {rdl(bx.x, o), rdl(bx.y, o), rdl(bx.z, o)};
            const float lh[3] = {rdl(bx.w, o), rdl(bh.x, o), rdl(bh.y, o)};
            float rl = Rf;
            if (IS_IJ) rl = fmaxf(rl, fabsf(__int_as_float(rdl(lk.w, o))));
            rl = rl * kReachRel + padX;
            float dist2 = 0.f;
#pragma unroll
            for (int q = 0; q < D; ++q) {
                const float e = fmaxf(abs_min_image_f(lc3[q] - xf[q], a.box.periodic, Lf[q]) - lh[q], 0.f);
                dist2 = (q == 0) ? e * e : dist2 + e * e;
            }
            const bool own = mine && dist2 <= rl * rl;
            if (__ballot(own)) {
                ++leaves;
                // members this lane's own sphere would have to test in a
                // per-particle walk (the search's algorithmic work)
                if (own) reach += rdl(lk.y, o) - rdl(lk.x, o);
                int s = rdl(lk.x, o);
                const int e = rdl(lk.y, o);
                while (s < e) {
                    const int take = min(e - s, kWave - nb);
                    if (lane >= nb && lane < nb + take) cand = s + (lane - nb);
                    nb += take;
                    s += take;
                    if (nb == kWave) flush(mine);
                }
            }
        };
        // Depth-first walk of the pre-order node range [wb, we) (a whole
        // subtree), evaluated 64 nodes at a time: lane q tests node w0 + q; the
        // walk advances through the window on the wave-uniform reach mask
        // (descend = +1, prune = skip) and loads the next window where it
        // leaves this one.
        auto dfs = [&](int wb, int we) {
            int w0 = wb;
            while (w0 < we) {
                ++windows;
                const int k = w0 + lane;
                int4 lk = make_int4(0, 0, we, 0);
                float4 bx = make_float4(0.f, 0.f, 0.f, 0.f);
                float2 bh = make_float2(0.f, 0.f);
                bool reach = false;
                if (k < we) {
                    bx = a.nbox[k];
                    bh = a.nhalf[k];
                    lk = a.nlink[k];
                    reach = reach_node(bx, bh, lk);
                }
                const unsigned long long mask = __ballot(reach);
                // events: pruned nodes and entered leaves; entered internal
                // nodes between events are simply descended through
                const unsigned long long evt = __ballot(k < we && (!reach || lk.w < 0));
                int node = w0;
                const int wend = min(w0 + kWave, we);
                while (node < wend) {
                    const int off = node - w0;
                    const unsigned long long m = evt >> off;
                    if (m == 0ull) {
                        visits += wend - node;
                        node = wend;
                        break;
                    }
                    const int step = __ffsll((long long)m) - 1;
                    visits += step + 1;
                    const int o = off + step;
                    node = w0 + o;
                    if ((mask >> o) & 1ull) {
                        visit_leaf(o, bx, bh, lk);
                        ++node;
                    } else {
                        node = rdl(lk.z, o);
                    }
                }
                w0 = node;
            }
        };

        // Breadth-first walk: the frontier of nodes to test (children of the
        // nodes entered on the previous level) is tested 64 at a time, one node
        // per lane; entered leaves are visited, entered internal nodes append
        // their children (stored contiguously) to the next frontier.  If the
        // next frontier would overflow its LDS buffer, the subtrees of that
        // batch's entered internal nodes are walked depth-first instead.
        int *cur = sfr[0], *nxt = sfr[1];
        if (lane == 0) cur[0] = 0;  // the root
        int nf = 1;
        wave_lds_fence();
        while (nf > 0) {
            int nn = 0;
            for (int base = 0; base < nf; base += kWave) {
                ++windows;
                const int q = base + lane;
                const bool act = q < nf;
                const int node = act ? cur[q] : 0;
                int4 lk = make_int4(0, 0, 0, 0);
                float4 bx = make_float4(0.f, 0.f, 0.f, 0.f);
                float2 bh = make_float2(0.f, 0.f);
                int2 ch = make_int2(0, 0);
                bool reach = false;
                if (act) {
                    // the child list is loaded with the node (no dependent load)
                    bx = a.nbox[node];
                    bh = a.nhalf[node];
                    lk = a.nlink[node];
                    ch = a.nchild[node];
                    reach = reach_node(bx, bh, lk);
                }
                visits += min(kWave, nf - base);
                const bool is_leaf = lk.w < 0;
                unsigned long long lm = __ballot(act && reach && is_leaf);
                while (lm) {
                    const int o = __ffsll((long long)lm) - 1;
                    lm &= lm - 1;
                    visit_leaf(o, bx, bh, lk);
                }
                const bool expand = act && reach && !is_leaf;
                const int cnt = expand ? ch.y : 0;
                int incl = cnt;
                for (int off = 1; off < kWave; off <<= 1) {
                    const int t = __shfl_up(incl, off, kWave);
                    if (lane >= off) incl += t;
                }
                const int total = __shfl(incl, kWave - 1, kWave);
                if (nn + total <= kFrontier) {
                    const int pos = nn + incl - cnt;
                    if (expand) {
                        // at most 2^D children; the child array is padded by
                        // 8 entries, so the 2^D loads are independent and in bounds
                        int cv[1 << D];
#pragma unroll
                        for (int c = 0; c < (1 << D); ++c) cv[c] = a.child[ch.x + c];
#pragma unroll
                        for (int c = 0; c < (1 << D); ++c)
                            if (c < cnt) nxt[pos + c] = cv[c];
                    }
                    nn += total;
                } else {
                    unsigned long long em = __ballot(expand);
                    while (em) {
                        const int o = __ffsll((long long)em) - 1;
                        em &= em - 1;
                        dfs(rdl(node, o) + 1, rdl(lk.z, o));
                    }
                }
            }
            wave_lds_fence();
            int *t = cur;
            cur = nxt;
            nxt = t;
            nf = nn;
        }
        if (nb) flush(mine);
        sub = sub_end;
    }
    if (valid) a.cnt[i] = count < a.cap ? count : a.cap;
    if (over) atomicOr(a.overflow, 1);
    // candidate statistics
    int c = valid ? count : 0;
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    for (int off = kWave / 2; off > 0; off >>= 1) nexact += __shfl_xor(nexact, off, kWave);
    long long reach_w = reach;
    for (int off = kWave / 2; off > 0; off >>= 1) reach_w += __shfl_xor(reach_w, off, kWave);
    if (lane == 0) {
        atomicAdd(stat_ptr(a.stats, 17), (unsigned long long)nexact);
        atomicAdd(stat_ptr(a.stats, IS_IJ ? 4 : 0), (unsigned long long)c);
        atomicAdd(stat_ptr(a.stats, 13), (unsigned long long)visits);
        atomicAdd(stat_ptr(a.stats, 14), (unsigned long long)tests);
        atomicAdd(stat_ptr(a.stats, 15), (unsigned long long)windows);
        atomicAdd(stat_ptr(a.stats, 16), (unsigned long long)leaves);
        atomicAdd(stat_ptr(a.stats, 20), (unsigned long long)reach_w);
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
    if (hipMalloc((void **)&l.cnt_a, sizeof(int) * (size_t)c->n) != hipSuccess) return SPHGPU_E_NOMEM;
    if (hipMemset(l.cnt, 0, sizeof(int) * (size_t)c->n) != hipSuccess) return SPHGPU_E_HIP;
    return 0;
}

int lists_grow(sphgpu_ctx *c, Lists &l)
{
    // the reference's list holds nn * neighbor_list_size = nn * 1000 entries
    // and fails only when a particle has more (src/bhtree.cpp:300-302,
    // include/defines.hpp:35): double up to that size, then stop there
    const long long limit = 1000ll * c->prm.neighbor_number;
    if (l.cap >= limit)
        return fail(c, SPHGPU_E_NEIGHBOR_OVERFLOW, "Neighbor list overflow: increase neighbor_list_size in defines.hpp");
    const long long want = std::min(2ll * l.cap, limit);
    const int rc = lists_alloc(c, l, (int)want);
    return rc ? fail(c, rc, "neighbour list reallocation failed") : 0;
}

void lists_free(Lists &l)
{
    if (l.idx) hipFree(l.idx);
    if (l.cnt) hipFree(l.cnt);
    if (l.cnt_a) hipFree(l.cnt_a);
    l = Lists();
}

int flist_check(sphgpu_ctx *c)
{
    if (c->pre_search_pending) {
        const int rc = pre_resolve(c);
        if (rc) return rc;
    }
    if (c->flist == 1) {
        int bad = 1;
        HIPCHK(c, hipMemcpyAsync(&bad, c->dflag + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->flist = bad ? 0 : 2;
    }
    return c->flist == 2 ? 0 : 1;
}

int search(sphgpu_ctx *c, Lists &l, int is_ij, const double *radius, bool defer, bool flag_cleared, double rscale)
{
    Tree &t = c->tree;
    if (!t.valid) return fail(c, SPHGPU_E_STATE, "neighbour search without a tree");
    if (&l == &c->lforce) c->flist = 0;  // the derived force list is overwritten
    // the list grows by doubling until the candidates fit; the reference's
    // own capacity, nn * 1000 entries, bounds it (the loop ends there)
    for (;;) {
        SearchArgs a;
        a.n = c->n;
        a.n_nodes = t.n_nodes;
        for (int k = 0; k < 3; ++k) a.x[k] = c->pos[k < c->dim ? k : 0];
        a.radius = radius;
        a.rscale = rscale;
        a.sml = c->sc[S_SML];
        a.posf = c->rec32[R_POSH];
        a.root = t.root;
        a.nbox = t.nbox; a.nhalf = t.nhalf; a.nlink = t.nlink; a.nchild = t.nchild; a.child = t.child;
        a.box = c->box;
        a.idx = l.idx; a.cnt = l.cnt; a.cap = l.cap;
        a.overflow = c->dflag;
        a.stats = c->dstats;
        if (!flag_cleared) HIPCHK(c, hipMemsetAsync(c->dflag, 0, sizeof(int), c->stream));
        flag_cleared = false;  // a repeated search after growing the lists clears it here
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
        if (defer) {
            c->pre_search_pending = true;
            return 0;
        }
        int over = 0;
        HIPCHK(c, hipMemcpyAsync(&over, c->dflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!over) return 0;
        // grow the list and search again
        int rc = lists_grow(c, l);
        if (rc) return rc;
    }
}

} // namespace sphgpu
