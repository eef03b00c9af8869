// hydro.hip -- PreInteraction and FluidForce pair loops for SSPH, DISPH, GSPH.
//
// One thread per particle i walks its neighbour list (built by neighbor.hip
// with the reference's exact FP64 predicates) and reads neighbour data from
// key-ordered SoA arrays.  Pair arithmetic runs in T: float with
// Neumaier-compensated sums for densities and forces, or double with the
// reference's expressions.  Set membership (r < h) is always decided in FP64.
//
// References:
//   PreInteraction (SSPH)   src/pre_interaction.cpp:39-283
//   PreInteraction (DISPH)  src/disph/d_pre_interaction.cpp:21-230
//   PreInteraction (GSPH)   src/gsph/g_pre_interaction.cpp:27-144
//   FluidForce (SSPH)       src/fluid_force.cpp:26-116
//   FluidForce (DISPH)      src/disph/d_fluid_force.cpp:26-86
//   FluidForce (GSPH)       src/gsph/g_fluid_force.cpp:44-304
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "mlp_infer.hpp"

namespace sphgpu {

namespace {

constexpr double kPi = 3.1415926535897932384626433832795028841971693993751;
template <int D> constexpr double area_const() { return D == 1 ? 2.0 : D == 2 ? kPi : 4.0 * kPi / 3.0; }

template <typename T> struct PreArgs {
    typedef typename V4T<T>::type V4;
    int n;
    const int *idx;
    const int *cnt;
    int cap;
    RecView<T> rv;                // neighbour records (read)
    T *rec_w[R_NREC];             // the same records as scalars (own-record writes)
    float *posf_w;                // FP32 position/h records, written in FP64 mode
    const double *x64[3];
    const double *hs;
    const double *mass64;
    double mass_u;     // the common mass when mass_uniform
    int mass_uniform;
    double *sml, *dens, *pres, *gradh, *balsara, *alpha;
    int *neighbor;
    double *grad[kNGrad];
    unsigned long long *hvs_min;
    const double *dscal;
    long long *stats;
    Box box;
    int nn, iterative, balsara_on, td_av, gsph2;
    double gamma, kappa, alpha_min, alpha_max, eps_av;
    const double *root;           // tree root geometry (bounds |x| for the FP32 prefilter)
    // force pair list, first part (members with r != 0), written when
    // fl_idx is set
    int *fl_idx;
    int *fl_cnt, *fl_cnt_a;
    int fl_cap;
    int *fl_flag;                 // set when the derived list cannot be used
    // the candidates were searched within hs_i * hscale; with tight set
    // (hscale below 1), a particle whose Newton iterates or final h reach
    // that radius sets bit 1 of pre_flag (the pre-interaction must run again
    // at the reference's radius hs_i)
    double hscale;
    int tight;
    int *pre_flag;
    // time-dependent AV: alpha before the step's first pre-interaction (so a
    // repeated run updates it from the same value), or null: the record's
    const double *alpha_in;
};

// neighbour loops handle candidates in groups of this many
constexpr int kGroup = 4;
static_assert(16 % kGroup == 0, "blocked sums fold every 16 candidates, at group starts");

// entries q0 .. q0+kGroup-1 of particle i's list (entries at or past n
// repeat entry n - 1); the kGroup loads are issued together
SPH_DEV void list_group(const int *__restrict__ idx, int i, int q0, int n, int cap, int jg[kGroup])
{
#pragma unroll
    for (int u = 0; u < kGroup; ++u) jg[u] = idx[list_slot(i, min(q0 + u, n - 1), cap)];
}

// single components of a particle record: the mass (velocity record, w) and
// the specific internal energy (thermo record, w)
template <typename V4> SPH_DEV auto rec_mass(const V4 *rv, int j) -> decltype(rv->w)
{
    return reinterpret_cast<const decltype(rv->w) *>(rv)[4 * (size_t)j + 3];
}
template <typename V4> SPH_DEV auto rec_u(const V4 *rt, int j) -> decltype(rt->w)
{
    return reinterpret_cast<const decltype(rt->w) *>(rt)[4 * (size_t)j + 3];
}

template <int D, typename T>
SPH_DEV T rec_r2(const typename V4T<T>::type &pj, const T xi[3], const Box &b, T d[3])
{
    const T xj[3] = {pj.x, pj.y, pj.z};
    T s = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        d[k] = rij_axis<T>(xi[k], xj[k], b, k);
        s = (k == 0) ? d[k] * d[k] : s + d[k] * d[k];
    }
    return s;
}

SPH_DEV double r2_64(const double *const x[3], const double xi[3], int j, const Box &b, int D)
{
    double s = 0.0;
    for (int k = 0; k < D; ++k) {
        const double d = rij_axis<double>(xi[k], x[k][j], b, k);
        s = (k == 0) ? d * d : s + d * d;
    }
    return s;
}

SPH_DEV void wave_min_atomic(unsigned long long *dst, double v)
{
    dst += blockIdx.x & (kStatCopies - 1);  // one of kStatCopies copies
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMin(dst, dbits(v));
}

SPH_DEV void wave_add_atomic(long long *stats, int slot, long long v)
{
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(stat_ptr(stats, slot), (unsigned long long)v);
}

template <int D, int K, int TYPE, typename T>
__global__ __launch_bounds__(256) void k_pre(PreArgs<T> a)
{
    typedef typename V4T<T>::type V4;
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = i0 < a.n;
    const int i = valid ? i0 : a.n - 1;
    const V4 *__restrict__ RP = a.rv.r[R_POSH];
    const V4 *__restrict__ RV = a.rv.r[R_VELM];
    const V4 *__restrict__ RT = a.rv.r[R_THERMO];
    const int nc = valid ? a.cnt[i] : 0;
    const V4 pi_ = RP[i], vi_ = RV[i], ti_ = RT[i];
    const T xi[3] = {pi_.x, pi_.y, pi_.z};
    const T vi[3] = {vi_.x, vi_.y, vi_.z};
    double xi64[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) xi64[k] = a.x64[k][i];
    const T ui = ti_.w, ci = ti_.z;
    const double hs64 = a.hs[i];
    const double A = area_const<D>();
    long long newton_iters = 0, newton_fail = 0;

    // ---- smoothing length: newton_raphson (pre_interaction.cpp:227-283) ----
    T h;
    // the radius the candidates were searched within
    const double hst = hs64 * a.hscale;
    if (a.iterative) {
        h = T(hs64 / a.kappa);
        const T b = (TYPE == SPHGPU_DISPH || TYPE == SPHGPU_GDISPH) ? T(a.nn / A) : T(a.mass64[i] * a.nn / A);
        bool conv = false;
        int it = 0;
        // the largest h any sum of this particle uses (the iterates and the
        // final h)
        T hmax = h;
        for (; it < 10;) {
            const T hb = h;
            BlockAcc<T> dn, ddn;  // compensated in blocks of 16 candidates
            // candidates in groups of kGroup: the group's position and mass
            // loads are issued together, and the next group's indices are
            // loaded one group ahead, so a group waits for memory once; the
            // sums run in candidate order
            // (in the FP32 form the next group's positions and masses are
            // loaded one group ahead as well, its indices two ahead)
            constexpr bool kAhead = sizeof(T) == 4;
            constexpr bool kMassW = !(TYPE == SPHGPU_DISPH || TYPE == SPHGPU_GDISPH);
            // equal masses: m_j is the common value (the record holds the same
            // rounded number), one record load per candidate instead of two
            const bool mload = kMassW && !a.mass_uniform;
            const T mu = kMassW ? T(a.mass_u) : T(1);
            int jn[kGroup];
            V4 pn[kGroup];
            T wn[kGroup];
            if (nc > 0) list_group(a.idx, i, 0, nc, a.cap, jn);
            if (kAhead && nc > 0) {
#pragma unroll
                for (int u = 0; u < kGroup; ++u) pn[u] = RP[jn[u]];
                if (mload) {
#pragma unroll
                    for (int u = 0; u < kGroup; ++u) wn[u] = rec_mass(RV, jn[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < kGroup; ++u) wn[u] = mu;
                }
                if (kGroup < nc) list_group(a.idx, i, kGroup, nc, a.cap, jn);
            }
            for (int q0 = 0; q0 < nc; q0 += kGroup) {
                if ((q0 & 15) == 0) {
                    dn.fold();
                    ddn.fold();
                }
                V4 pg[kGroup];
                T wg[kGroup];
                if (kAhead) {
#pragma unroll
                    for (int u = 0; u < kGroup; ++u) {
                        pg[u] = pn[u];
                        wg[u] = wn[u];
                    }
                    if (q0 + kGroup < nc) {
#pragma unroll
                        for (int u = 0; u < kGroup; ++u) pn[u] = RP[jn[u]];
                        if (mload) {
#pragma unroll
                            for (int u = 0; u < kGroup; ++u) wn[u] = rec_mass(RV, jn[u]);
                        } else {
#pragma unroll
                            for (int u = 0; u < kGroup; ++u) wn[u] = mu;
                        }
                        if (q0 + 2 * kGroup < nc) list_group(a.idx, i, q0 + 2 * kGroup, nc, a.cap, jn);
                    }
                } else {
                    int jg[kGroup];
#pragma unroll
                    for (int u = 0; u < kGroup; ++u) jg[u] = jn[u];
#pragma unroll
                    for (int u = 0; u < kGroup; ++u) pg[u] = RP[jg[u]];
                    if (q0 + kGroup < nc) list_group(a.idx, i, q0 + kGroup, nc, a.cap, jn);
                    if (mload) {
#pragma unroll
                        for (int u = 0; u < kGroup; ++u) wg[u] = rec_mass(RV, jg[u]);
                    } else {
#pragma unroll
                        for (int u = 0; u < kGroup; ++u) wg[u] = mu;
                    }
                }
#pragma unroll
                for (int u = 0; u < kGroup; ++u) {
                    T r2;
                    if constexpr (sizeof(T) == 4) {
                        // only |r| is needed here
                        r2 = T(0);
#pragma unroll
                        for (int k = 0; k < D; ++k) {
                            const T dk = abs_rij_axis(xi[k], k == 0 ? pg[u].x : k == 1 ? pg[u].y : pg[u].z, a.box, k);
                            r2 = (k == 0) ? dk * dk : r2 + dk * dk;
                        }
                    } else {
                        T d[3];
                        r2 = rec_r2<D, T>(pg[u], xi, a.box, d);
                    }
                    const T r = tsqrt(r2);
                    if (q0 + u >= nc || r >= h) continue;
                    dn.add(wg[u] * Kern<D, K, T>::w(r, h));
                    ddn.add(wg[u] * Kern<D, K, T>::dhw(r, h));
                }
            }
            const T dens = dn.get(), ddens = ddn.get();
            const T f = dens * powh<D>(h) - b;
            const T df = ddens * powh<D>(h) + T(D) * dens * powh_m1<D>(h);
            h -= tdiv(f, df);
            ++it;
            hmax = h > hmax ? h : hmax;
            if (tabs(h - hb) < (h + hb) * T(1e-4)) { conv = true; break; }
        }
        newton_iters = it;
        // with the tight radius every candidate r < h of each sum must have
        // been searched: h stays below hst with room for the pair
        // precision's rounding of r (FP32 distances: 2^-16 relative)
        if (a.tight && valid) {
            const double m = (sizeof(T) == 8) ? 9.094947017729282e-13 : 1.52587890625e-05;
            if (!((double)hmax < hst * (1.0 - m))) atomicOr(a.pre_flag, 2);
        }
        if (!conv) {
            h = T(hs64 / a.kappa);
            newton_fail = 1;
        }
    } else {
        h = T(hs64);
    }
    const double h64 = (double)h;
    // FP32 form: membership by the FP32 prefilter, FP64 inside its band
    const float mk = (sizeof(T) == 4) ? fp32_margin_k(a.root, D) : 0.f;

    // ---- density, pressure, grad-h, v_sig (members r < h, decided exactly) ----
    // membership r < h, decided exactly (FP32 prefilter, FP64 inside its band)
    auto member = [&](int j, T r2) -> bool {
        if (sizeof(T) == 8) return tsqrt(r2) < h;
        const int band = fp32_band((float)r2, (float)h, mk);
        return band > 0 || (band == 0 && r_less(r2_64(a.x64, xi64, j, a.box, D), h64));
    };
    // membership bits of the first 256 candidates, for the force-list rows
    // written after this pass
    unsigned long long mb[4] = {0, 0, 0, 0};
    // compensated in blocks of 16 candidates
    BlockAcc<T> dens_a, dhd_a, n_a, pres_a, dhp_a, dhn_a;
    T vsig = ci * T(2.0);
    int w = 0;
    // members other than coincident ones start the force pair list: with
    // h_i below the search radius they are exactly {j : r_ij < h_i, r_ij != 0}.
    // They are written to the force list after this pass, one row per
    // member rank, from the membership bits (past 256 candidates, during the
    // pass); a coincident pair j != i (r_ij = 0) makes the derived list
    // unusable for this step (the force searches instead).
    const bool fl = a.fl_idx != nullptr;
    const bool fl_rows = fl && nc > 256;

    int qself = -1, nrow = 0;
    bool coincident = false;
    // FP32 r^2 at or below z2 may come from coincident positions (their
    // rounding moves each component by a few u X): (2^-18 X)^2 leaves room
    float z2 = 0.f;
    if (fl && sizeof(T) == 4) {
        double Xb = 0.0;
        for (int k = 0; k < D; ++k) Xb = fmax(Xb, fabs(a.root[k]));
        const float padX = (float)((Xb + a.root[3]) * 3.814697265625e-06);
        z2 = padX * padX;
    }
    // second-order GSPH / GDISPH: the MUSCL gradient sums run in this pass
    // too (members other than the particle itself, in candidate order; the
    // particle itself would add exactly zero)
    constexpr bool kGrad = TYPE == SPHGPU_GSPH || TYPE == SPHGPU_GDISPH;
    const bool grad_on = kGrad && a.gsph2;
    T gdd[3] = {0, 0, 0}, gdu[3] = {0, 0, 0}, gdv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    // and the Balsara / time-dependent AV sums (pre_interaction.cpp:115-161;
    // GDISPH, gd_pre_interaction.cpp:150-197: weighted by mass, or by mass x
    // u as in DISPH when the AV runs without Balsara)
    constexpr bool kBal = TYPE != SPHGPU_GSPH;
    const bool do_bal = a.balsara_on && D != 1;
    const bool bal_on = kBal && (do_bal || a.td_av);
    const bool disph_w = (TYPE == SPHGPU_DISPH) || (TYPE == SPHGPU_GDISPH && !do_bal);
    T bdiv = T(0), brot[3] = {0, 0, 0};
    // (FP32 form: the next group's positions are loaded one group ahead and
    // its indices two ahead)
    constexpr bool kAheadP = sizeof(T) == 4;
    int jn[kGroup], jc[kGroup];
    V4 pn[kGroup];
    if (nc > 0) list_group(a.idx, i, 0, nc, a.cap, jn);
    if (kAheadP && nc > 0) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            jc[u] = jn[u];
            pn[u] = RP[jn[u]];
        }
        if (kGroup < nc) list_group(a.idx, i, kGroup, nc, a.cap, jn);
    }
    for (int q0 = 0; q0 < nc; q0 += kGroup) {
    if ((q0 & 15) == 0) {
        dens_a.fold(); dhd_a.fold(); n_a.fold(); pres_a.fold(); dhp_a.fold(); dhn_a.fold();
    }
    // a group of candidates: positions first (the indices were loaded one
    // group ahead), then the records of the members, each set of loads
    // issued together
    int jg[kGroup];
    V4 pg[kGroup], vg[kGroup];
    T cg[kGroup], ug[kGroup];
    bool memb[kGroup];
    if (kAheadP) {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) {
            jg[u] = jc[u];
            pg[u] = pn[u];
        }
        if (q0 + kGroup < nc) {
#pragma unroll
            for (int u = 0; u < kGroup; ++u) {
                jc[u] = jn[u];
                pn[u] = RP[jn[u]];
            }
            if (q0 + 2 * kGroup < nc) list_group(a.idx, i, q0 + 2 * kGroup, nc, a.cap, jn);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kGroup; ++u) jg[u] = jn[u];
#pragma unroll
        for (int u = 0; u < kGroup; ++u) pg[u] = RP[jg[u]];
        if (q0 + kGroup < nc) list_group(a.idx, i, q0 + kGroup, nc, a.cap, jn);
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
        T d[3];
        memb[u] = q0 + u < nc && member(jg[u], rec_r2<D, T>(pg[u], xi, a.box, d));
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
        if (!memb[u]) continue;
        vg[u] = RV[jg[u]];
        const V4 t = RT[jg[u]];
        cg[u] = t.z;
        ug[u] = t.w;
    }

#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
        const int q = q0 + u;
        if (!memb[u]) continue;
        const int j = jg[u];
        T d[3];
        const T r2 = rec_r2<D, T>(pg[u], xi, a.box, d);
        const unsigned long long bit = 1ull << (q & 63);
#pragma unroll
        for (int b = 0; b < 4; ++b) mb[b] |= ((q >> 6) == b) ? bit : 0ull;
        ++w;
        if (fl) {
            if (j == i) {
                qself = q;
            } else {
                // r_ij != 0 as the reference decides it (FP64); an FP32 r^2
                // above z2 cannot come from coincident positions
                bool nz = r2 != T(0);
                if (sizeof(T) == 4 && !((float)r2 > z2)) nz = r2_64(a.x64, xi64, j, a.box, D) != 0.0;
                coincident = coincident || !nz;
                if (fl_rows && nz) {
                    if (nrow < a.fl_cap) a.fl_idx[list_slot(i, nrow, a.fl_cap)] = j;
                    ++nrow;
                }
            }
        }
        const T r = tsqrt(r2);
        const T wk = Kern<D, K, T>::w(r, h);
        const V4 vj = vg[u];
        const T mj = vj.w;
        if (TYPE == SPHGPU_DISPH || TYPE == SPHGPU_GDISPH) {
            const T dh = Kern<D, K, T>::dhw(r, h);
            const T uj = ug[u];
            dens_a.add(mj * wk);
            n_a.add(wk);
            pres_a.add(mj * uj * wk);
            dhp_a.add(mj * uj * dh);
            dhn_a.add(dh);
        } else if (TYPE == SPHGPU_SSPH) {
            dens_a.add(mj * wk);
            dhd_a.add(mj * Kern<D, K, T>::dhw(r, h));
        } else {
            dens_a.add(mj * wk);
        }
        if (j != i) {
            const T vjv[3] = {vj.x, vj.y, vj.z};
            T dot = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) dot = (k == 0) ? d[k] * (vi[k] - vjv[k]) : dot + d[k] * (vi[k] - vjv[k]);
            const T vs = ci + cg[u] - tdiv(T(3.0) * dot, r);
            if (vs > vsig) vsig = vs;
            if (kGrad && grad_on) {
                // g_pre_interaction.cpp:112-136 (GDISPH: gd_pre_interaction.cpp:125-148)
                const T cw = Kern<D, K, T>::dwc(r, h);
                const T mdu = mj * (ug[u] - ui);
                T mdv[3];
#pragma unroll
                for (int b2 = 0; b2 < D; ++b2) mdv[b2] = mj * (vjv[b2] - vi[b2]);
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const T dwk = d[k] * cw;
                    gdd[k] += dwk * mj;
                    gdu[k] += dwk * mdu;
#pragma unroll
                    for (int b2 = 0; b2 < D; ++b2) gdv[b2][k] += dwk * mdv[b2];
                }
            }
            if (kBal && bal_on) {
                const T cw = Kern<D, K, T>::dwc(r, h);
                T dw[3], vij[3];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    dw[k] = d[k] * cw;
                    vij[k] = vi[k] - vjv[k];
                }
                const T wt = disph_w ? mj * ug[u] : mj;
                T bdot = T(0);
#pragma unroll
                for (int k = 0; k < D; ++k) bdot = (k == 0) ? vij[k] * dw[k] : bdot + vij[k] * dw[k];
                bdiv -= wt * bdot;
                if (D == 3) {
                    brot[0] += (vij[1] * dw[2 % D] - vij[2 % D] * dw[1 % D]) * wt;
                    brot[1] += (vij[2 % D] * dw[0] - vij[0] * dw[2 % D]) * wt;
                    brot[2] += (vij[0] * dw[1 % D] - vij[1 % D] * dw[0]) * wt;
                } else if (D == 2) {
                    brot[0] += (vij[0] * dw[1 % D] - vij[1 % D] * dw[0]) * wt;
                }
            }
        }
    }
    }
    const T dens = dens_a.get();
    const T gm1 = T(a.gamma - 1.0);
    T pres, gradh = T(0);
    if (TYPE == SPHGPU_DISPH || TYPE == SPHGPU_GDISPH) {
        const T ni = n_a.get();
        pres = gm1 * pres_a.get();
        gradh = tdiv(tdiv(h, T(D) * ni) * dhp_a.get(), T(1.0) + tdiv(h, T(D) * ni) * dhn_a.get());
    } else if (TYPE == SPHGPU_SSPH) {
        pres = gm1 * dens * ui;
        gradh = tdiv(T(1.0), T(1.0) + tdiv(h, T(D) * dens) * dhd_a.get());
    } else {
        pres = gm1 * dens * ui;
    }
    // members other than the particle itself, as bits over the first 256
    // candidates
    unsigned long long am[4];
    int na_bits = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        am[b] = mb[b] & ~(((qself >> 6) == b) ? 1ull << (qself & 63) : 0ull);
        na_bits += __popcll(am[b]);
    }
    if (valid) {
        a.sml[i] = h64;
        a.dens[i] = (double)dens;
        a.pres[i] = (double)pres;
        if (TYPE != SPHGPU_GSPH) a.gradh[i] = (double)gradh;
        a.neighbor[i] = w;
        // own record: h, dens, pres, gradh (components other threads do not read here)
        a.rec_w[R_POSH][4 * (size_t)i + 3] = h;
        a.rec_w[R_THERMO][4 * (size_t)i + 0] = dens;
        a.rec_w[R_THERMO][4 * (size_t)i + 1] = pres;
        if (TYPE != SPHGPU_GSPH) a.rec_w[R_AUX][4 * (size_t)i + 0] = gradh;
        if (a.posf_w) a.posf_w[4 * (size_t)i + 3] = (float)h;
        if (fl) {
            const int na = fl_rows ? nrow : na_bits;
            a.fl_cnt_a[i] = min(na, a.fl_cap);
            a.fl_cnt[i] = min(na, a.fl_cap);
            // complete only when every j with r < h_i was a candidate
            // (r^2 < hst^2): h_i must stay below hst with room for rounding
            const bool beyond = !(h64 < hst * (1.0 - 9.094947017729282e-13));
            const bool other = (!fl_rows && qself < 0) || coincident || na > a.fl_cap;
            if (beyond || other) {
                atomicOr(a.fl_flag, 1);
                atomicAdd(stat_ptr(a.stats, beyond ? 18 : 19), 1ull);
            }
        }
    }
    wave_min_atomic(a.hvs_min, valid ? (double)(h / vsig) : 1.7976931348623157e308);
    wave_add_atomic(a.stats, 1, valid ? newton_iters : 0);
    wave_add_atomic(a.stats, 2, valid ? newton_fail : 0);
    if (!valid) return;

    // The member sums (density, v_sig, MUSCL gradients, Balsara / AV) were
    // taken in the density pass; what is left is writing the force-list rows
    // (row r of every lane = its r-th member, from the member bits, so a
    // store covers one 256-byte row) and finishing the sums.
    bool rows_pending = fl && !fl_rows;

    if (kGrad && grad_on) {
        // MUSCL gradients (summed in the density pass)
        T dd[3], du[3], dv[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dd[k] = gdd[k];
            du[k] = gdu[k];
#pragma unroll
            for (int b2 = 0; b2 < 3; ++b2) dv[b2][k] = gdv[b2][k];
        }
        const T rho_inv = tdiv(T(1.0), dens);
        T g[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) g[q] = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T gp = (dd[k] * ui + du[k]) * gm1;
            g[k] = dd[k];
            g[3 + k] = gp;
            a.grad[k][i] = (double)dd[k];
            a.grad[3 + k][i] = (double)gp;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) {
                const T gv = dv[b2][k] * rho_inv;
                g[6 + 3 * b2 + k] = gv;
                a.grad[6 + 3 * b2 + k][i] = (double)gv;
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) a.rec_w[R_G0 + (q >> 2)][4 * (size_t)i + (q & 3)] = g[q];
    }
    if (rows_pending) {
        // the force-list rows: row r of every lane = its r-th member
        int wd = 0;
        unsigned long long m = am[0];
        const int nw = min(na_bits, a.fl_cap);
        for (int r = 0; r < nw; ++r) {
            while (m == 0ull && wd < 3) {
                ++wd;
                m = (wd == 1) ? am[1] : (wd == 2) ? am[2] : am[3];
            }
            const int qc = (wd << 6) + __ffsll((long long)m) - 1;
            m &= m - 1ull;
            a.fl_idx[list_slot(i, r, a.fl_cap)] = a.idx[list_slot(i, qc, a.cap)];
        }
        rows_pending = false;
    }
    if (TYPE == SPHGPU_GSPH) return;

    // ---- Balsara switch and time-dependent AV (pre_interaction.cpp:115-161;
    // GDISPH, gd_pre_interaction.cpp:150-197: Balsara sums weighted by mass
    // and divided by density, the AV without Balsara as in DISPH) ----
    const T dt = T(a.dscal[0]);
    T alpha = a.alpha_in ? T(a.alpha_in[i]) : a.rv.r[R_AUX][i].z;
    if (do_bal || a.td_av) {
        // sums from the density pass
        T div = bdiv, rot[3] = {brot[0], brot[1], brot[2]};
        T rot_abs;
        if (disph_w) {
            const T p_inv = gm1 / pres;
            div *= p_inv;
            for (int k = 0; k < 3; ++k) rot[k] *= p_inv;
        } else {
            div /= dens;
            for (int k = 0; k < 3; ++k) rot[k] /= dens;
        }
        if (D == 3) rot_abs = tsqrt(rot[0] * rot[0] + rot[1] * rot[1] + rot[2] * rot[2]);
        else rot_abs = tabs(rot[0]);
        const T tau_inv = tdiv(T(a.eps_av) * ci, h);
        if (do_bal) {
            const T bal = tdiv(tabs(div), tabs(div) + rot_abs + tdiv(T(1e-4) * ci, h));
            a.balsara[i] = (double)bal;
            a.rec_w[R_AUX][4 * (size_t)i + 1] = bal;
            if (a.td_av) {
                const T mdiv = -div > T(0) ? -div : T(0);
                const T dalpha = (-(alpha - T(a.alpha_min)) * tau_inv + mdiv * (T(a.alpha_max) - alpha)) * dt;
                alpha += dalpha;
            }
        } else {
            const T s_i = -div > T(0) ? -div : T(0);
            alpha = (alpha + dt * tau_inv * T(a.alpha_min) + s_i * dt * T(a.alpha_max)) /
                    (T(1.0) + dt * tau_inv + s_i * dt);
        }
        if (a.td_av) {
            a.alpha[i] = (double)alpha;
            a.rec_w[R_AUX][4 * (size_t)i + 2] = alpha;
        }
    }
}

// ---- force pair lists from the pre-interaction lists ----------------------------
// The FluidForce pair set of i is {j : r_ij < max(h_i, h_j), r_ij != 0}
// (src/fluid_force.cpp:43-62, src/gsph/g_fluid_force.cpp:93): the members
// with r < h_i, which k_pre wrote, plus every j with h_j > r_ij >= h_i.  The
// latter pairs are exactly the transposes of k_pre's entries (i in the list
// of j) that fail r < h_i; since the FP64 r_ij^2 is symmetric in i and j,
// the same predicate decides them.  This pass appends them to the other
// particle's list; the next one puts each appended run in index order so the
// lists do not depend on the order of the atomics.
struct FlistArgs {
    int n;
    int *idx, *cnt;
    const int *cnt_a;
    int cap;

    const float4 *posf;       // FP32 (x, y, z, h)
    const double *x64[3];
    const double *sml;
    const double *root;
    Box box;
    int *flag;
    long long *stats;
    // k_flist_order also hands the list flags to the host (mapped pinned
    // memory) and stores the h / v_sig minimum
    const int *dflag;
    int *hflag;
    const unsigned long long *minbits;
    double *dscal;
};

template <int D, bool FP64>
__global__ __launch_bounds__(256) void k_flist_scatter(FlistArgs a)
{
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int na = a.cnt_a[i];
    const float4 pi = a.posf[i];
    const double hi = a.sml[i];
    double xi64[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) xi64[k] = a.x64[k][i];
    const float xf[3] = {pi.x, pi.y, pi.z};
    const float mk = fp32_margin_k(a.root, D);
    // members in groups of kGroup, each group's loads issued together
    for (int q0 = 0; q0 < na; q0 += kGroup) {
    int jg[kGroup];
    float4 pg[kGroup];
    double hg[kGroup];
    list_group(a.idx, i, q0, na, a.cap, jg);
#pragma unroll
    for (int u = 0; u < kGroup; ++u) pg[u] = a.posf[jg[u]];
    // FP32 rounding is monotone: rounded values that differ order h_i and
    // h_j as the FP64 values do, so the FP32 form reads the FP64 h_j only
    // for equal rounded values or an undecided FP32 distance test
#pragma unroll
    for (int u = 0; u < kGroup; ++u)
        hg[u] = (FP64 ? pg[u].w <= pi.w : pg[u].w == pi.w) ? a.sml[jg[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
        if (q0 + u >= na) break;
        const int j = jg[u];
        const float4 pj = pg[u];
        // h_j > h_i: r < h_i <= h_j puts i in the list of j already
        if (pj.w > pi.w) continue;
        const bool tie = FP64 || pj.w == pi.w;
        if (tie && !(hg[u] < hi)) continue;
        bool in_j;  // r_ij < h_j
        if (FP64) {
            in_j = r_less(r2_64(a.x64, xi64, j, a.box, D), hg[u]);
        } else {
            float d[3];
            const float r2f = rec_r2<D, float>(pj, xf, a.box, d);
            const int band = fp32_band(r2f, pj.w, mk);
            in_j = band > 0 || (band == 0 && r_less(r2_64(a.x64, xi64, j, a.box, D), tie ? hg[u] : a.sml[j]));
        }
        if (in_j) continue;
        const int s = atomicAdd(&a.cnt[j], 1);
        if (s < a.cap) a.idx[list_slot(j, s, a.cap)] = i;
        else {
            atomicOr(a.flag, 1);
            atomicAdd(stat_ptr(a.stats, 19), 1ull);
        }
    }
    }
}

constexpr int kRunSort = 16;

__global__ __launch_bounds__(256) void k_flist_order(FlistArgs a)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // every writer of the flags (search, pre-interaction, scatter) and of
        // the minima ran in earlier launches: their values are final here
        unsigned long long v = a.minbits[0];
        for (int k = 1; k < kStatCopies; ++k) v = a.minbits[k] < v ? a.minbits[k] : v;
        a.dscal[2] = bitsd(v);
        for (int k = 0; k < 3; ++k) __hip_atomic_store(&a.hflag[k], a.dflag[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __threadfence_system();
    }
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int nt = 0;
    if (i < a.n) {
        const int na = a.cnt_a[i];
        nt = min(a.cnt[i], a.cap);
        const int m = nt - na;  // length of the appended run
        if (m > 1 && m <= kRunSort) {
            // short runs (the usual case): all loads issued together, an
            // odd-even transposition network in registers (kRunSort rounds
            // sort kRunSort entries; the padding sorts to the end)
            int v[kRunSort];
#pragma unroll
            for (int u = 0; u < kRunSort; ++u) v[u] = u < m ? a.idx[list_slot(i, na + u, a.cap)] : INT_MAX;
#pragma unroll
            for (int r = 0; r < kRunSort; ++r) {
#pragma unroll
                for (int u = r & 1; u + 1 < kRunSort; u += 2) {
                    const int lo = min(v[u], v[u + 1]), hi = max(v[u], v[u + 1]);
                    v[u] = lo;
                    v[u + 1] = hi;
                }
            }
#pragma unroll
            for (int u = 0; u < kRunSort; ++u)
                if (u < m) a.idx[list_slot(i, na + u, a.cap)] = v[u];
        }
        // insertion sort of a longer appended run
        for (int q = na + 1; m > kRunSort && q < nt; ++q) {
            const int v = a.idx[list_slot(i, q, a.cap)];
            int p = q - 1;
            int u = a.idx[list_slot(i, p, a.cap)];
            while (p >= na && u > v) {
                a.idx[list_slot(i, p + 1, a.cap)] = u;
                --p;
                if (p >= na) u = a.idx[list_slot(i, p, a.cap)];
            }
            a.idx[list_slot(i, p + 1, a.cap)] = v;
        }
        a.cnt[i] = nt;
    }
    long long c = nt;
    for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(stat_ptr(a.stats, 4), (unsigned long long)c);
}

// ---- initial_smoothing (pre_interaction.cpp:171-215) ----------------------------
template <int D, int K>
__global__ __launch_bounds__(256) void k_initial_density(int n, int cap, const int *__restrict__ idx, const int *__restrict__ cnt,
                                                         const double *const *x, const double *__restrict__ mass,
                                                         const double *__restrict__ sml, double *__restrict__ dens,
                                                         Box box)
{
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xi[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) xi[k] = x[k][i];
    const double h = sml[i];
    double s = 0.0;
    const int nc = cnt[i];
    for (int q = 0; q < nc; ++q) {
        const int j = idx[list_slot(i, q, cap)];
        const double r = __builtin_sqrt(r2_64(x, xi, j, box, D));
        if (r >= h) continue;
        s += mass[j] * Kern<D, K, double>::w(r, h);
    }
    dens[i] = s;
}

// smoothing-length guess (pre_interaction.cpp:61-64, 186-189)
template <int D>
__global__ void k_hguess(int n, const double *__restrict__ mass, const double *__restrict__ dens, int nn,
                         double kappa, double *__restrict__ hs, unsigned long long *minbits, int *flags)
{
    // with flags: the pre-interaction's resets (search overflow, force-list
    // flag and the h / v_sig minima) ride on this launch
    if (flags && blockIdx.x == 0) {
        if (threadIdx.x < kStatCopies) minbits[threadIdx.x] = 0x7fefffffffffffffull;  // DBL_MAX
        if (threadIdx.x == 0) {
            flags[0] = 0;
            flags[2] = 0;
        }
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double A = area_const<D>();
    hs[i] = pow(nn * mass[i] / (dens[i] * A), 1.0 / D) * kappa;
}

// resets the h / v_sig minimum copies and, when given, the force-list flag
__global__ void k_minbits_init(unsigned long long *m, int *flag)
{
    if (threadIdx.x < kStatCopies) m[threadIdx.x] = 0x7fefffffffffffffull;  // DBL_MAX
    if (flag && threadIdx.x == 0) *flag = 0;
}

__global__ void k_store_hvs(const unsigned long long *m, double *dscal)
{
    unsigned long long v = m[0];
    for (int k = 1; k < kStatCopies; ++k) v = m[k] < v ? m[k] : v;
    dscal[2] = bitsd(v);
}

// ---- FluidForce --------------------------------------------------------------------
template <typename T> struct ForceArgs {
    // when set: the launch does nothing if skip_if[0] or skip_if[2] is
    // nonzero (the overflow flags of the lists it would read)
    const int *skip_if = nullptr;
    int n;
    const int *idx, *cnt;
    int cap;
    RecView<T> rv;
    double *acc[3];
    double *dene;
    const double *dscal;
    long long *stats;
    Box box;
    double gamma, alpha_ac;
    int use_ac, use_gravity, gsph2, refine64, guess;
    MLPDev mlp;
    float *rlog;
    long long rlog_cap;
    int rlog_stride;
    unsigned long long *rlog_count;
};

template <int D, typename T>
SPH_DEV T av_pi(const T vij[3], const T d[3], T r, T ai, T aj, T bi, T bj, T ci, T cj, T rhoi, T rhoj)
{
    // FluidForce::artificial_viscosity, src/fluid_force.cpp:89-106
    T vr = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) vr = (k == 0) ? vij[k] * d[k] : vr + vij[k] * d[k];
    if (vr < T(0)) {
        const T alpha = T(0.5) * (ai + aj);
        const T balsara = T(0.5) * (bi + bj);
        const T w_ij = tdiv(vr, r);
        const T v_sig = ci + cj - T(3.0) * w_ij;
        const T rho_ij_inv = tdiv(T(2.0), rhoi + rhoj);
        return T(-0.5) * balsara * alpha * v_sig * w_ij * rho_ij_inv;
    }
    return T(0);
}

// GSPH Riemann states of a pair, right = i and left = j: MUSCL-reconstructed
// with the van Leer limiter (src/gsph/g_fluid_force.cpp:103-140, the same in
// src/gdisph/gd_fluid_force.cpp:101-138) or the particles' own.  d = r_ij,
// g = (grad rho, grad p, grad v) flattened as in the gradient records.
template <int D, typename T>
SPH_DEV void pair_states(bool second, T gamma, T dt, const T d[3], T r, const T vi[3],
                         const typename V4T<T>::type &vj, const typename V4T<T>::type &ti,
                         const typename V4T<T>::type &tj, const T gi[16], const T gj[16], T left[4], T right[4])
{
    const T r_inv = tdiv(T(1.0), r);
    T e[3] = {0, 0, 0};
    T ve_i = T(0), ve_j = T(0);
    const T vjv[3] = {vj.x, vj.y, vj.z};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        e[k] = d[k] * r_inv;
        ve_i = (k == 0) ? vi[k] * e[k] : ve_i + vi[k] * e[k];
        ve_j = (k == 0) ? vjv[k] * e[k] : ve_j + vjv[k] * e[k];
    }
    const T rhoi = ti.x, presi = ti.y, ci = ti.z;
    const T rhoj = tj.x, presj = tj.y, cj = tj.z;
    if (!second) {
        right[0] = ve_i; right[1] = rhoi; right[2] = presi; right[3] = ci;
        left[0] = ve_j; left[1] = rhoj; left[2] = presj; left[3] = cj;
        return;
    }
    const T delta_i = T(0.5) * (T(1.0) - ci * dt * r_inv);
    const T delta_j = T(0.5) * (T(1.0) - cj * dt * r_inv);
    T dve_i = T(0), dve_j = T(0), dd_i = T(0), dd_j = T(0), dp_i = T(0), dp_j = T(0);
#pragma unroll
    for (int b = 0; b < D; ++b) {
        T si = T(0), sj = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            si = (k == 0) ? gi[6 + 3 * b + k] * e[k] : si + gi[6 + 3 * b + k] * e[k];
            sj = (k == 0) ? gj[6 + 3 * b + k] * e[k] : sj + gj[6 + 3 * b + k] * e[k];
        }
        dve_i = (b == 0) ? si * e[b] : dve_i + si * e[b];
        dve_j = (b == 0) ? sj * e[b] : dve_j + sj * e[b];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dd_i = (k == 0) ? gi[k] * e[k] : dd_i + gi[k] * e[k];
        dd_j = (k == 0) ? gj[k] * e[k] : dd_j + gj[k] * e[k];
        dp_i = (k == 0) ? gi[3 + k] * e[k] : dp_i + gi[3 + k] * e[k];
        dp_j = (k == 0) ? gj[3 + k] * e[k] : dp_j + gj[3 + k] * e[k];
    }
    dve_i = dve_i * r; dve_j = dve_j * r;
    dd_i = dd_i * r; dd_j = dd_j * r;
    dp_i = dp_i * r; dp_j = dp_j * r;
    const T dv_ij = ve_i - ve_j;
    right[0] = ve_i - limiter(dv_ij, dve_i) * delta_i;
    left[0] = ve_j + limiter(dv_ij, dve_j) * delta_j;
    const T dd_ij = rhoi - rhoj;
    right[1] = rhoi - limiter(dd_ij, dd_i) * delta_i;
    left[1] = rhoj + limiter(dd_ij, dd_j) * delta_j;
    const T dp_ij = presi - presj;
    right[2] = presi - limiter(dp_ij, dp_i) * delta_i;
    left[2] = presj + limiter(dp_ij, dp_j) * delta_j;
    right[3] = tsqrt(gamma * right[2] / right[1]);
    left[3] = tsqrt(gamma * left[2] / left[1]);
}

template <int D, int K, int TYPE, typename T>
__global__ __launch_bounds__(256) void k_force_sph(ForceArgs<T> a)
{
    if (a.skip_if && (a.skip_if[0] | a.skip_if[2])) return;  // whole grid
    typedef typename V4T<T>::type V4;
    const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const V4 *__restrict__ RP = a.rv.r[R_POSH];
    const V4 *__restrict__ RV = a.rv.r[R_VELM];
    const V4 *__restrict__ RT = a.rv.r[R_THERMO];
    const V4 *__restrict__ RA = a.rv.r[R_AUX];
    const V4 pi_ = RP[i], vi_ = RV[i], ti_ = RT[i], ai_ = RA[i];
    const T xi[3] = {pi_.x, pi_.y, pi_.z};
    const T vi[3] = {vi_.x, vi_.y, vi_.z};
    const T hi = pi_.w, rhoi = ti_.x, presi = ti_.y, gradhi = ai_.x;
    const T ci = ti_.z, ai = ai_.z, bi = ai_.y, mi = vi_.w, ui = ti_.w;
    const T gm1 = T(a.gamma - 1.0);
    const T p_per_rho2_i = tdiv(presi, sqr(rhoi));
    const T g2u_i = sqr(gm1) * ui;
    const T g2upp_i = tdiv(g2u_i, presi);
    const T m_u_inv = tdiv(T(1.0), mi * ui);
    const T ihi = tdiv(T(1.0), hi);
    // GDISPH: the particle's own MUSCL gradients (gd, gp, gv records)
    T gi[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) gi[q] = T(0);
    if (TYPE == SPHGPU_GDISPH && a.gsph2) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const V4 g = a.rv.r[R_G0 + b][i];
            gi[4 * b] = g.x; gi[4 * b + 1] = g.y; gi[4 * b + 2] = g.z; gi[4 * b + 3] = g.w;
        }
    }
    const T dt = (TYPE == SPHGPU_GDISPH) ? T(a.dscal[0]) : T(0);
    const T gamma = T(a.gamma);
    // compensated sums in blocks of 16 pairs
    BlockAcc<T> ac[3], de;
    const int nc = a.cnt[i];
    // the next pair index is loaded one pair ahead; in the FP32 form the
    // next pair's four records are loaded one pair ahead too and the index
    // two ahead
    constexpr bool kAhead = sizeof(T) == 4;
    int j_next = nc > 0 ? a.idx[list_slot(i, 0, a.cap)] : 0;
    V4 n_p{}, n_v{}, n_t{}, n_a{};
    int j_after = 0;
    if (kAhead && nc > 0) {
        n_p = RP[j_next];
        n_v = RV[j_next];
        n_t = RT[j_next];
        n_a = RA[j_next];
        if (nc > 1) j_after = a.idx[list_slot(i, 1, a.cap)];
    }
    for (int q = 0; q < nc; ++q) {
        if ((q & 15) == 15) {
#pragma unroll
            for (int k = 0; k < D; ++k) ac[k].fold();
            de.fold();
        }
        const int j = j_next;
        V4 pj, vj, tj, aj;
        if (kAhead) {
            pj = n_p;
            vj = n_v;
            tj = n_t;
            aj = n_a;
            if (q + 1 < nc) {
                j_next = j_after;
                n_p = RP[j_next];
                n_v = RV[j_next];
                n_t = RT[j_next];
                n_a = RA[j_next];
                if (q + 2 < nc) j_after = a.idx[list_slot(i, q + 2, a.cap)];
            }
        } else {
            if (q + 1 < nc) j_next = a.idx[list_slot(i, q + 1, a.cap)];
            pj = RP[j];
            vj = RV[j];
            tj = RT[j];
            aj = RA[j];
        }
        T d[3];
        const T r = tsqrt(rec_r2<D, T>(pj, xi, a.box, d));
        // GDISPH: HLL p* of the pair (right = i, left = j)
        T pstar = T(0);
        if (TYPE == SPHGPU_GDISPH) {
            T gj[16];
#pragma unroll
            for (int q2 = 0; q2 < 16; ++q2) gj[q2] = T(0);
            if (a.gsph2) {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const V4 g = a.rv.r[R_G0 + b][j];
                    gj[4 * b] = g.x; gj[4 * b + 1] = g.y; gj[4 * b + 2] = g.z; gj[4 * b + 3] = g.w;
                }
            }
            T left[4], right[4], vstar;
            pair_states<D, T>(a.gsph2 != 0, gamma, dt, d, r, vi, vj, ti_, tj, gi, gj, left, right);
            hll_solver<T>(left, right, pstar, vstar);
        }
        const T hj = pj.w, rhoj = tj.x, presj = tj.y, mj = vj.w, uj = tj.w;
        T cwi, cwj;
        if constexpr (sizeof(T) == 4) {
            // the Wendland gradient needs no 1/r
            const T r_inv = (K == SPHGPU_WENDLAND) ? T(0) : tdiv(T(1.0), r);
            cwi = dwc_rcp<D, K>(r, r_inv, ihi);
            cwj = dwc_rcp<D, K>(r, r_inv, tdiv(T(1.0), hj));
        } else {
            cwi = Kern<D, K, T>::dwc(r, hi);
            cwj = Kern<D, K, T>::dwc(r, hj);
        }
        const T vjv[3] = {vj.x, vj.y, vj.z};
        T dwi[3], dwj[3], dwij[3], vij[3];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            dwi[k] = d[k] * cwi;
            dwj[k] = d[k] * cwj;
            dwij[k] = (dwi[k] + dwj[k]) * T(0.5);
            vij[k] = vi[k] - vjv[k];
        }
        const T pi_ij = av_pi<D, T>(vij, d, r, ai, aj.z, bi, aj.y, ci, tj.z, rhoi, rhoj);
        T dene_ac = T(0);
        if (a.use_ac) {
            // FluidForce::artificial_conductivity, src/fluid_force.cpp:108-116
            T vsig;
            if (a.use_gravity) {
                T vr = T(0);
#pragma unroll
                for (int k = 0; k < D; ++k) vr = (k == 0) ? vij[k] * d[k] : vr + vij[k] * d[k];
                vsig = tabs(tdiv(vr, r));
            } else {
                vsig = tsqrt(tdiv(T(2.0) * tabs(presi - presj), rhoi + rhoj));
            }
            T dr = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) dr = (k == 0) ? dwij[k] * d[k] : dr + dwij[k] * d[k];
            dene_ac = tdiv(T(a.alpha_ac) * mj * vsig * (ui - uj) * dr, r);
        }
        T s1, s2, e1;
        T vdwi = T(0), vdwij = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            vdwi = (k == 0) ? vij[k] * dwi[k] : vdwi + vij[k] * dwi[k];
            vdwij = (k == 0) ? vij[k] * dwij[k] : vdwij + vij[k] * dwij[k];
        }
        if (TYPE == SPHGPU_SSPH) {
            s1 = mj * (p_per_rho2_i * gradhi + T(0.5) * pi_ij);
            s2 = mj * (tdiv(presj, sqr(rhoj)) * aj.x + T(0.5) * pi_ij);
            e1 = mj * p_per_rho2_i * gradhi * vdwi + T(0.5) * mj * pi_ij * vdwij + dene_ac;
        } else if (TYPE == SPHGPU_DISPH) {
            const T f_ij = T(1.0) - tdiv(gradhi, mj * uj);
            const T f_ji = T(1.0) - aj.x * m_u_inv;
            const T u_per_pres_j = tdiv(uj, presj);
            s1 = mj * (g2upp_i * uj * f_ij + T(0.5) * pi_ij);
            s2 = mj * (g2u_i * u_per_pres_j * f_ji + T(0.5) * pi_ij);
            e1 = mj * g2upp_i * uj * f_ij * vdwi + T(0.5) * mj * pi_ij * vdwij + dene_ac;
        } else {
            // GDISPH, gd_fluid_force.cpp:158-185: the DISPH terms scaled by
            // p* / P_i and p* / P_j
            const T f_ij = T(1.0) - tdiv(gradhi, mj * uj);
            const T f_ji = T(1.0) - aj.x * m_u_inv;
            const T u_per_pres_j = tdiv(uj, presj);
            const T pst_i = tdiv(pstar, presi), pst_j = tdiv(pstar, presj);
            s1 = mj * (g2upp_i * uj * f_ij * pst_i + T(0.5) * pi_ij);
            s2 = mj * (g2u_i * u_per_pres_j * f_ji * pst_j + T(0.5) * pi_ij);
            e1 = mj * g2upp_i * uj * f_ij * pst_i * vdwi + T(0.5) * mj * pi_ij * vdwij + dene_ac;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) ac[k].add(-(dwi[k] * s1 + dwj[k] * s2));
        de.add(e1);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) a.acc[k][i] = (double)ac[k].get();
    a.dene[i] = (double)de.get();
}

// Riemann problem of one pair: iterative two-shock solver or HLL.  In the
// FP32 form the Newton loop runs in FP32 and, with refine64, one final Newton
// step and v* are evaluated in FP64.
template <typename T, int SOLVER>
SPH_DEV int solve_pair(const T l[4], const T r[4], T gamma, int refine64, const T *guess, T &pstar, T &vstar,
                       int &fallback)
{
    if (SOLVER == SPHGPU_RIEMANN_HLL) {
        hll_solver<T>(l, r, pstar, vstar);
        return 0;
    }
    TwoShock<T> ts;
    ts.setup(gamma, l, r);
    if (ts.bad) {
        pstar = T(0);
        vstar = T(0);
        return 0;
    }
    T p;
    int it;
    fallback = 0;
    if (sizeof(T) == 4 && refine64) {
        // FP32 iterations that keep their last Newton factor F; the final
        // step is taken in FP64 with that F (the factor at the last iterate's
        // start, within the iteration tolerance of the one at p: it only
        // scales the small correction), and v* comes from the same FP64
        // evaluation
        T F = T(0);
        ts.prep_f(gamma, l[1], r[1]);
        if (guess) {
            const T g = *guess;
            bool conv = false;
            it = 0;
            if (g > T(1e-25) && g < T(1e30)) {
                p = g;
                for (int k = 0; k < 4; ++k) {
                    const T p_old = p;
                    p = ts.step_f(p, F);
                    ++it;
                    if (tabs(p - p_old) < T(1e-6) * p) { conv = true; break; }
                }
            }
            if (!conv) {
                fallback = 1;
                p = ts.plain_guess();
                it += ts.iterate_f(p, F);
            }
        } else {
            p = ts.plain_guess();
            it = ts.iterate_f(p, F);
        }
        const float lf[4] = {(float)l[0], (float)l[1], (float)l[2], (float)l[3]};
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        const double g64 = (double)gamma;
        double v64 = 0.0;
        const double pd = refine_step64_v(lf, rf, g64, 0.5 * (1.0 + g64) / g64, (float)p, (float)F, v64);
        pstar = (T)pd;
        vstar = (T)v64;
        return it;
    }
    if (guess) {
        // learned starting point: Newton from it for at most 4 steps; if that
        // does not meet the solver's tolerance, restart from the plain guess
        const T g = *guess;
        bool conv = false;
        it = 0;
        if (g > T(1e-25) && g < T(1e30)) {
            p = g;
            for (int k = 0; k < 4; ++k) {
                const T p_old = p;
                p = ts.step(p);
                ++it;
                if (tdiv(tabs(p - p_old), p) < T(1e-6)) { conv = true; break; }
            }
        }
        if (!conv) {
            fallback = 1;
            p = ts.plain_guess();
            it += ts.iterate(p);
        }
    } else {
        p = ts.plain_guess();
        it = ts.iterate(p);
    }
    if (sizeof(T) == 4 && refine64) {
        // final Newton step with the residual in FP64; v* from the refined p*
        const float lf[4] = {(float)l[0], (float)l[1], (float)l[2], (float)l[3]};
        const float rf[4] = {(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
        const double g64 = (double)gamma;
        const double pd = refine_step64(lf, rf, g64, 0.5 * (1.0 + g64) / g64, (float)p, (float)ts.newton_factor(p));
        pstar = (T)pd;
        vstar = ts.vstar(pstar);
    } else {
        pstar = p;
        vstar = ts.vstar(p);
    }
    return it;
}

// The pair-slot loop of the GSPH force.  Slot q holds each lane's q-th pair;
// the reconstruction and the force run per lane, the learned guess (when
// enabled) runs for the whole wavefront on MFMA between them.
template <int D, int K, typename T, int SOLVER, bool LEARNED>
SPH_DEV void force_gsph_slots(const ForceArgs<T> &a, int i, int nc, int nmax, const T xi[3], const T vi[3],
                              const T gdi[3], const T gpi[3], const T gvi[3][3], T hi, T rhoi, T presi, T ci,
                              T gamma, T rho2_inv_i, T dt, BlockAcc<T> ac[3], BlockAcc<T> &de, const MlpWeightsBF16 *wl,
                              MlpWaveLDS *wsc, bool valid)
{
    typedef typename V4T<T>::type V4;
    const V4 *__restrict__ RP = a.rv.r[R_POSH];
    const V4 *__restrict__ RV = a.rv.r[R_VELM];
    const V4 *__restrict__ RT = a.rv.r[R_THERMO];
    long long iters = 0, wave_max_sum = 0, calls = 0, fallbacks = 0, learned = 0;
    // The next slot's pair index is loaded one slot ahead, so its record
    // gathers do not wait for it; in the FP32 form the next slot's records
    // are loaded one slot ahead too (the slot's arithmetic covers their
    // latency), and the index two slots ahead.
    constexpr bool kAhead = sizeof(T) == 4;
    const T ihi = tdiv(T(1.0), hi);
    int j_next = nc > 0 ? a.idx[list_slot(i, 0, a.cap)] : 0;
    V4 n_p{}, n_v{}, n_t{}, n_g[4] = {};
    int j_after = 0;
    if (kAhead && nc > 0) {
        n_p = RP[j_next];
        n_v = RV[j_next];
        n_t = RT[j_next];
        if (a.gsph2)
#pragma unroll
            for (int b = 0; b < 4; ++b) n_g[b] = a.rv.r[R_G0 + b][j_next];
        if (nc > 1) j_after = a.idx[list_slot(i, 1, a.cap)];
    }
    for (int q = 0; q < nmax; ++q) {
        // the sums are compensated in blocks of 16 slots
        if ((q & 15) == 15) {
#pragma unroll
            for (int k = 0; k < D; ++k) ac[k].fold();
            de.fold();
        }
        const bool act = q < nc;
        int it = 0;
        int j = 0;
        const int j_cur = j_next;
        V4 c_p{}, c_v{}, c_t{}, c_g[4] = {};
        if (kAhead) {
            c_p = n_p;
            c_v = n_v;
            c_t = n_t;
#pragma unroll
            for (int b = 0; b < 4; ++b) c_g[b] = n_g[b];
            if (q + 1 < nc) {
                j_next = j_after;
                n_p = RP[j_next];
                n_v = RV[j_next];
                n_t = RT[j_next];
                if (a.gsph2)
#pragma unroll
                    for (int b = 0; b < 4; ++b) n_g[b] = a.rv.r[R_G0 + b][j_next];
                if (q + 2 < nc) j_after = a.idx[list_slot(i, q + 2, a.cap)];
            }
        } else if (q + 1 < nc) {
            j_next = a.idx[list_slot(i, q + 1, a.cap)];
        }
        T d[3] = {0, 0, 0}, e[3] = {0, 0, 0}, r = T(1), r_inv = T(1);
        T left[4] = {0, 1, 1, 1}, right[4] = {0, 1, 1, 1};
        T rhoj = T(1), hj = T(1), mj = T(0);
        if (act) {
            j = j_cur;
            const V4 pj = kAhead ? c_p : RP[j], vjr = kAhead ? c_v : RV[j], tj = kAhead ? c_t : RT[j];
            const T r2 = rec_r2<D, T>(pj, xi, a.box, d);
            if constexpr (sizeof(T) == 4) {
                // 1/r from one reciprocal square root, r = r^2 / r (list
                // pairs have r > 0)
                r_inv = trsqrt(r2);
                r = r2 > T(0) ? r2 * r_inv : T(0);
            } else {
                r = tsqrt(r2);
                r_inv = tdiv(T(1.0), r);
            }
            const T vj[3] = {vjr.x, vjr.y, vjr.z};
            T ve_i = T(0), ve_j = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                e[k] = d[k] * r_inv;
                ve_i = (k == 0) ? vi[k] * e[k] : ve_i + vi[k] * e[k];
                ve_j = (k == 0) ? vj[k] * e[k] : ve_j + vj[k] * e[k];
            }
            hj = pj.w;
            mj = vjr.w;
            rhoj = tj.x;
            const T presj = tj.y, cj = tj.z;
            if (a.gsph2) {
                // neighbour gradients: (gd, gp, gv) packed in four records
                const V4 g0 = kAhead ? c_g[0] : a.rv.r[R_G0][j], g1 = kAhead ? c_g[1] : a.rv.r[R_G1][j],
                         g2 = kAhead ? c_g[2] : a.rv.r[R_G2][j], g3 = kAhead ? c_g[3] : a.rv.r[R_G3][j];
                const T gj[16] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w,
                                  g2.x, g2.y, g2.z, g2.w, g3.x, g3.y, g3.z, g3.w};
                // MUSCL reconstruction, g_fluid_force.cpp:103-140
                const T delta_i = T(0.5) * (T(1.0) - ci * dt * r_inv);
                const T delta_j = T(0.5) * (T(1.0) - cj * dt * r_inv);
                const T dv_ij = ve_i - ve_j;
                T dvi[3], dvj[3];
#pragma unroll
                for (int b2 = 0; b2 < D; ++b2) {
                    T si = T(0), sj = T(0);
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        si = (k == 0) ? gvi[b2][k] * e[k] : si + gvi[b2][k] * e[k];
                        const T gvj = gj[6 + 3 * b2 + k];
                        sj = (k == 0) ? gvj * e[k] : sj + gvj * e[k];
                    }
                    dvi[b2] = si;
                    dvj[b2] = sj;
                }
                T dve_i = T(0), dve_j = T(0), dd_i = T(0), dd_j = T(0), dp_i = T(0), dp_j = T(0);
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    dve_i = (k == 0) ? dvi[k] * e[k] : dve_i + dvi[k] * e[k];
                    dve_j = (k == 0) ? dvj[k] * e[k] : dve_j + dvj[k] * e[k];
                    dd_i = (k == 0) ? gdi[k] * e[k] : dd_i + gdi[k] * e[k];
                    const T gdj = gj[k];
                    dd_j = (k == 0) ? gdj * e[k] : dd_j + gdj * e[k];
                    dp_i = (k == 0) ? gpi[k] * e[k] : dp_i + gpi[k] * e[k];
                    const T gpj = gj[3 + k];
                    dp_j = (k == 0) ? gpj * e[k] : dp_j + gpj * e[k];
                }
                dve_i = dve_i * r;
                dve_j = dve_j * r;
                dd_i = dd_i * r;
                dd_j = dd_j * r;
                dp_i = dp_i * r;
                dp_j = dp_j * r;
                right[0] = ve_i - limiter(dv_ij, dve_i) * delta_i;
                left[0] = ve_j + limiter(dv_ij, dve_j) * delta_j;
                const T dd_ij = rhoi - rhoj;
                right[1] = rhoi - limiter(dd_ij, dd_i) * delta_i;
                left[1] = rhoj + limiter(dd_ij, dd_j) * delta_j;
                const T dp_ij = presi - presj;
                right[2] = presi - limiter(dp_ij, dp_i) * delta_i;
                left[2] = presj + limiter(dp_ij, dp_j) * delta_j;
                right[3] = tsqrt(gamma * right[2] / right[1]);
                left[3] = tsqrt(gamma * left[2] / left[1]);
            } else {
                right[0] = ve_i; right[1] = rhoi; right[2] = presi; right[3] = ci;
                left[0] = ve_j; left[1] = rhoj; left[2] = presj; left[3] = cj;
            }
        }
        // learned starting guess for the "hard" pairs of this slot: strong
        // pressure ratio or velocity jump above the mean sound speed
        T guess = T(0);
        bool use_guess = false;
        if constexpr (LEARNED) {
            const float pl = (float)left[2], pr = (float)right[2];
            const float pmin = fminf(pl, pr), pmax = fmaxf(pl, pr);
            const float du = (float)(left[0] - right[0]);
            const float r_m = 0.5f * (float)(left[1] + right[1]), p_m = 0.5f * (pl + pr);
            const bool hard = a.mlp.w && act && pmin > 0.f && r_m > 0.f &&
                              (pmax > 4.f * pmin || du * du * r_m > (float)gamma * p_m);
            if (__ballot(hard)) {
                float x[6] = {0, 0, 0, 0, 0, 0}, pm = 1.f;
                if (hard)
                    mlp_features((float)gamma, (float)left[0], (float)left[1], pl, (float)right[0],
                                 (float)right[1], pr, x, pm);
                const float y = mlp_wave_infer(*wl, *wsc, x, hard);
                if (hard) {
                    guess = (T)(pm * __expf(y));
                    use_guess = true;
                }
            }
        }
        if (act) {
            T pstar, vstar;
            int fb = 0;
            it = solve_pair<T, SOLVER>(left, right, gamma, a.refine64, use_guess ? &guess : nullptr, pstar, vstar, fb);
            if (SOLVER == SPHGPU_RIEMANN_ITERATIVE) {
                ++calls;
                fallbacks += fb;
                learned += (use_guess && !fb) ? 1 : 0;
                if (a.rlog) {
                    const unsigned long long slot = atomicAdd(a.rlog_count, 1ull);
                    if ((long long)(slot % (unsigned long long)a.rlog_stride) == 0) {
                        const long long row = (long long)(slot / (unsigned long long)a.rlog_stride) % a.rlog_cap;
                        float *o = a.rlog + row * 7;
                        o[0] = (float)left[0]; o[1] = (float)left[1]; o[2] = (float)left[2];
                        o[3] = (float)right[0]; o[4] = (float)right[1]; o[5] = (float)right[2];
                        o[6] = (float)pstar;
                    }
                }
            }
            // Godunov force, g_fluid_force.cpp:158-165
            T cwi, cwj;
            if constexpr (sizeof(T) == 4) {
                cwi = dwc_rcp<D, K>(r, r_inv, ihi);
                cwj = dwc_rcp<D, K>(r, r_inv, tdiv(T(1.0), hj));
            } else {
                cwi = Kern<D, K, T>::dwc(r, hi);
                cwj = Kern<D, K, T>::dwc(r, hj);
            }
            const T rho2_inv_j = tdiv(T(1.0), sqr(rhoj));
            const T s1 = mj * pstar * rho2_inv_i;
            const T s2 = mj * pstar * rho2_inv_j;
            T fv = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const T f = d[k] * cwi * s1 + d[k] * cwj * s2;
                ac[k].add(-f);
                const T dvk = e[k] * vstar - vi[k];
                fv = (k == 0) ? f * dvk : fv + f * dvk;
            }
            de.add(-fv);
        }
        if (SOLVER == SPHGPU_RIEMANN_ITERATIVE) {
            iters += it;
            // the wavefront's largest count (at most 24: 5 bits), bit by bit
            // from the top with ballots
            int m = 0;
#pragma unroll
            for (int b = 4; b >= 0; --b) {
                const int cand = m | (1 << b);
                if (__ballot(it >= cand)) m = cand;
            }
            wave_max_sum += m;
        }
    }
    wave_add_atomic(a.stats, 3, valid ? nc : 0);
    if (SOLVER == SPHGPU_RIEMANN_ITERATIVE) {
        wave_add_atomic(a.stats, 5, iters);
        if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(stat_ptr(a.stats, 6), (unsigned long long)wave_max_sum);
        wave_add_atomic(a.stats, 7, calls);
        wave_add_atomic(a.stats, 8, fallbacks);
        wave_add_atomic(a.stats, 9, learned);
    }
}

template <int D, int K, typename T, int SOLVER, int LEARNED>
__global__ __launch_bounds__(256) void k_force_gsph(ForceArgs<T> a)
{
    if (a.skip_if && (a.skip_if[0] | a.skip_if[2])) return;  // whole grid
    const int i0 = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool valid = i0 < a.n;
    const int i = valid ? i0 : a.n - 1;
    typedef typename V4T<T>::type V4;
    const V4 pi_ = a.rv.r[R_POSH][i], vi_ = a.rv.r[R_VELM][i], ti_ = a.rv.r[R_THERMO][i];
    const T xi[3] = {pi_.x, pi_.y, pi_.z};
    const T vi[3] = {vi_.x, vi_.y, vi_.z};
    T gdi[3] = {0, 0, 0}, gpi[3] = {0, 0, 0}, gvi[3][3] = {};
    if (a.gsph2) {
        const V4 g0 = a.rv.r[R_G0][i], g1 = a.rv.r[R_G1][i], g2 = a.rv.r[R_G2][i], g3 = a.rv.r[R_G3][i];
        const T g[16] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w,
                         g2.x, g2.y, g2.z, g2.w, g3.x, g3.y, g3.z, g3.w};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            gdi[k] = g[k];
            gpi[k] = g[3 + k];
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) gvi[b2][k] = g[6 + 3 * b2 + k];
        }
    }
    const T hi = pi_.w, rhoi = ti_.x, presi = ti_.y, ci = ti_.z;
    const T gamma = T(a.gamma);
    const T rho2_inv_i = tdiv(T(1.0), sqr(rhoi));
    const T dt = T(a.dscal[0]);
    BlockAcc<T> ac[3], de;
    const int nc = valid ? a.cnt[i] : 0;
    // the wave runs as many pair slots as its longest list
    int nmax = nc;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(nmax, off, kWave);
        nmax = o > nmax ? o : nmax;
    }
    if constexpr (LEARNED) {
        // the network's bf16 image is read from global memory (L2-resident);
        // LDS holds each wavefront's activation tiles
        __shared__ MlpWaveLDS wsc[4];
        force_gsph_slots<D, K, T, SOLVER, true>(a, i, nc, nmax, xi, vi, gdi, gpi, gvi, hi, rhoi, presi, ci, gamma,
                                               rho2_inv_i, dt, ac, de, a.mlp.img, &wsc[(threadIdx.x >> 6) & 3], valid);
    } else {
        force_gsph_slots<D, K, T, SOLVER, false>(a, i, nc, nmax, xi, vi, gdi, gpi, gvi, hi, rhoi, presi, ci, gamma,
                                                rho2_inv_i, dt, ac, de, nullptr, nullptr, valid);
    }
    if (valid) {
#pragma unroll
        for (int k = 0; k < D; ++k) a.acc[k][i] = (double)ac[k].get();
        a.dene[i] = (double)de.get();
    }
}

// batched Riemann problems (parity hook)
template <typename T, int SOLVER>
__global__ void k_riemann_batch(int n, const double *__restrict__ L, const double *__restrict__ R, double gamma,
                                int refine64, double *__restrict__ ps, double *__restrict__ vs, int *__restrict__ it)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    T l[4], r[4];
    for (int q = 0; q < 4; ++q) { l[q] = (T)L[4 * k + q]; r[q] = (T)R[4 * k + q]; }
    T p, v;
    int fb = 0;
    const int iters = solve_pair<T, SOLVER>(l, r, (T)gamma, refine64, nullptr, p, v, fb);
    ps[k] = (double)p;
    vs[k] = (double)v;
    it[k] = iters;
}

// batched Riemann problems with the learned starting guess (FP32 form): the
// same gate and wave-cooperative MFMA inference as the GSPH force kernel
__global__ __launch_bounds__(256) void k_riemann_batch_learned(int n, const double *__restrict__ L,
                                                               const double *__restrict__ R, float gamma,
                                                               int refine64, const MlpWeightsBF16 *wl,
                                                               double *__restrict__ ps, double *__restrict__ vs,
                                                               int *__restrict__ it, int *__restrict__ used)
{
    __shared__ MlpWaveLDS wsc[4];
    const int k0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = k0 < n;
    const int k = act ? k0 : n - 1;
    float l[4], r[4];
    for (int q = 0; q < 4; ++q) { l[q] = (float)L[4 * k + q]; r[q] = (float)R[4 * k + q]; }
    const float pmin = fminf(l[2], r[2]), pmax = fmaxf(l[2], r[2]);
    const float du = l[0] - r[0];
    const float r_m = 0.5f * (l[1] + r[1]), p_m = 0.5f * (l[2] + r[2]);
    const bool hard = act && pmin > 0.f && r_m > 0.f && (pmax > 4.f * pmin || du * du * r_m > gamma * p_m);
    float guess = 0.f;
    if (__ballot(hard)) {
        float x[6] = {0, 0, 0, 0, 0, 0}, pm = 1.f;
        if (hard) mlp_features(gamma, l[0], l[1], l[2], r[0], r[1], r[2], x, pm);
        const float y = mlp_wave_infer(*wl, wsc[(threadIdx.x >> 6) & 3], x, hard);
        guess = pm * __expf(y);
    }
    if (!act) return;
    float p, v;
    int fb = 0;
    const int iters = solve_pair<float, SPHGPU_RIEMANN_ITERATIVE>(l, r, gamma, refine64, hard ? &guess : nullptr, p, v, fb);
    ps[k] = (double)p;
    vs[k] = (double)v;
    it[k] = iters;
    used[k] = hard ? (fb ? 2 : 1) : 0;
}

// ---------------------------------------------------------------------------------
template <int D, int K, int TYPE, typename T> void launch_pre(sphgpu_ctx *c)
{
    PreArgs<T> a;
    a.n = c->n;
    a.idx = c->lpre.idx;
    a.cnt = c->lpre.cnt;
    a.cap = c->lpre.cap;
    a.rv = rec_view<T>(c);
    for (int r = 0; r < R_NREC; ++r) a.rec_w[r] = (T *)a.rv.r[r];
    a.posf_w = (sizeof(T) == 8) ? (float *)c->rec32[R_POSH] : nullptr;
    a.root = c->tree.root;
    for (int k = 0; k < 3; ++k) a.x64[k] = c->pos[k < c->dim ? k : 0];
    a.hs = c->hs;
    a.mass64 = c->sc[S_MASS];
    a.mass_u = c->mass_u;
    a.mass_uniform = c->mass_uniform ? 1 : 0;
    a.sml = c->sc[S_SML]; a.dens = c->sc[S_DENS]; a.pres = c->sc[S_PRES]; a.gradh = c->sc[S_GRADH];
    a.balsara = c->sc[S_BALSARA]; a.alpha = c->sc[S_ALPHA];
    a.neighbor = c->neighbor;
    for (int g = 0; g < kNGrad; ++g) a.grad[g] = c->grad[g];
    a.hvs_min = c->dmin;
    a.dscal = c->dscal;
    a.stats = c->dstats;
    a.box = c->box;
    a.nn = c->prm.neighbor_number;
    a.iterative = c->prm.iterative_sml;
    a.balsara_on = c->prm.use_balsara_switch;
    a.td_av = c->prm.use_time_dependent_av;
    a.gsph2 = c->prm.gsph_2nd_order;
    a.gamma = c->prm.gamma;
    a.kappa = c->prm.iterative_sml ? 1.2 : 1.0;
    a.alpha_min = c->prm.alpha_min;
    a.alpha_max = c->prm.alpha_max;
    a.eps_av = c->prm.epsilon_av;
    // with the iterative h the final h stays below the search radius, so the
    // force pair lists follow from these lists (non-iterative: h = hs, and
    // r < h need not imply r^2 < hs^2 at the last bit: search instead)
    const bool fl = c->prm.iterative_sml != 0;
    a.fl_idx = fl ? c->lforce.idx : nullptr;
    a.fl_cnt = c->lforce.cnt;
    a.fl_cnt_a = c->lforce.cnt_a;
    a.fl_cap = c->lforce.cap;
    a.fl_flag = c->dflag + 2;
    a.hscale = c->pre_rscale;
    a.tight = c->pre_rscale < 1.0 && c->prm.iterative_sml;
    a.pre_flag = c->dflag;
    a.alpha_in = c->alpha_snap_on ? c->alpha_snap : nullptr;
    hipLaunchKernelGGL((k_pre<D, K, TYPE, T>), dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, a);
}

template <int D> void launch_flist(sphgpu_ctx *c)
{
    FlistArgs a;
    a.n = c->n;
    a.idx = c->lforce.idx;
    a.cnt = c->lforce.cnt;
    a.cnt_a = c->lforce.cnt_a;
    a.cap = c->lforce.cap;

    a.posf = c->rec32[R_POSH];
    for (int k = 0; k < 3; ++k) a.x64[k] = c->pos[k < c->dim ? k : 0];
    a.sml = c->sc[S_SML];
    a.root = c->tree.root;
    a.box = c->box;
    a.flag = c->dflag + 2;
    a.stats = c->dstats;
    a.dflag = c->dflag;
    a.hflag = c->hflag_dev;
    a.minbits = c->dmin;
    a.dscal = c->dscal;
    const dim3 g(grid_for(c->n, 256)), b(256);
    time_begin(c, "force_lists");
    if (c->prm.pair_precision == SPHGPU_PAIR_FP64) hipLaunchKernelGGL((k_flist_scatter<D, true>), g, b, 0, c->stream, a);
    else hipLaunchKernelGGL((k_flist_scatter<D, false>), g, b, 0, c->stream, a);
    time_end(c);
    time_begin(c, "force_lists_order");
    hipLaunchKernelGGL(k_flist_order, g, b, 0, c->stream, a);
    time_end(c);
}

template <int D, int K, typename T> void dispatch_pre_type(sphgpu_ctx *c)
{
    switch (c->prm.sph_type) {
    case SPHGPU_SSPH: launch_pre<D, K, SPHGPU_SSPH, T>(c); break;
    case SPHGPU_DISPH: launch_pre<D, K, SPHGPU_DISPH, T>(c); break;
    case SPHGPU_GDISPH: launch_pre<D, K, SPHGPU_GDISPH, T>(c); break;
    default: launch_pre<D, K, SPHGPU_GSPH, T>(c); break;
    }
}

template <int D, typename T> void dispatch_pre_kernel(sphgpu_ctx *c)
{
    switch (c->prm.kernel) {
    case SPHGPU_CUBIC_SPLINE: dispatch_pre_type<D, SPHGPU_CUBIC_SPLINE, T>(c); break;
    case SPHGPU_WENDLAND: if (D > 1) dispatch_pre_type<D, SPHGPU_WENDLAND, T>(c); break;
    default: dispatch_pre_type<D, SPHGPU_GAUSSIAN, T>(c); break;
    }
}

template <typename T> void dispatch_pre(sphgpu_ctx *c)
{
    if (c->dim == 1) dispatch_pre_kernel<1, T>(c);
    else if (c->dim == 2) dispatch_pre_kernel<2, T>(c);
    else dispatch_pre_kernel<3, T>(c);
}

template <int D> void launch_initial_density(sphgpu_ctx *c, const double *const *xdev)
{
    const dim3 g(grid_for(c->n, 256)), b(256);
    switch (c->prm.kernel) {
    case SPHGPU_CUBIC_SPLINE:
        hipLaunchKernelGGL((k_initial_density<D, SPHGPU_CUBIC_SPLINE>), g, b, 0, c->stream, c->n, c->lpre.cap, c->lpre.idx,
                           c->lpre.cnt, xdev, c->sc[S_MASS], c->sc[S_SML], c->sc[S_DENS], c->box);
        break;
    case SPHGPU_WENDLAND:
        if (D > 1)
            hipLaunchKernelGGL((k_initial_density<D, SPHGPU_WENDLAND>), g, b, 0, c->stream, c->n, c->lpre.cap, c->lpre.idx,
                               c->lpre.cnt, xdev, c->sc[S_MASS], c->sc[S_SML], c->sc[S_DENS], c->box);
        break;
    default:
        hipLaunchKernelGGL((k_initial_density<D, SPHGPU_GAUSSIAN>), g, b, 0, c->stream, c->n, c->lpre.cap, c->lpre.idx,
                           c->lpre.cnt, xdev, c->sc[S_MASS], c->sc[S_SML], c->sc[S_DENS], c->box);
    }
}

template <int D> void launch_hguess(sphgpu_ctx *c, double kappa, double *out, bool resets)
{
    hipLaunchKernelGGL((k_hguess<D>), dim3(grid_for(c->n, 256)), dim3(256), 0, c->stream, c->n,
                       (const double *)c->sc[S_MASS], (const double *)c->sc[S_DENS], c->prm.neighbor_number, kappa, out,
                       resets ? c->dmin : nullptr, resets ? c->dflag : nullptr);
}

static void hguess(sphgpu_ctx *c, double kappa, double *out, bool resets = false)
{
    if (c->dim == 1) launch_hguess<1>(c, kappa, out, resets);
    else if (c->dim == 2) launch_hguess<2>(c, kappa, out, resets);
    else launch_hguess<3>(c, kappa, out, resets);
}

// ---- force dispatch ----
template <typename T> ForceArgs<T> force_args(sphgpu_ctx *c)
{
    ForceArgs<T> a;
    a.n = c->n;
    a.idx = c->lforce.idx;
    a.cnt = c->lforce.cnt;
    a.cap = c->lforce.cap;
    a.rv = rec_view<T>(c);
    for (int k = 0; k < 3; ++k) a.acc[k] = c->acc[k];
    a.dene = c->sc[S_DENE];
    a.dscal = c->dscal;
    a.stats = c->dstats;
    a.box = c->box;
    a.gamma = c->prm.gamma;
    a.alpha_ac = c->prm.alpha_ac;
    a.use_ac = c->prm.use_ac;
    a.use_gravity = c->prm.use_gravity;
    a.gsph2 = c->prm.gsph_2nd_order;
    a.refine64 = riemann_refine64(c);
    a.guess = c->prm.riemann_guess;
    a.mlp = mlp_dev(c);
    if (c->prm.riemann_guess != SPHGPU_GUESS_LEARNED) {  // plain guess
        a.mlp.w = nullptr;
        a.mlp.img = nullptr;
    }
    a.rlog = c->rlog_on ? c->rlog : nullptr;
    a.rlog_cap = c->rlog_cap;
    a.rlog_stride = c->rlog_stride;
    a.rlog_count = c->rlog_count;
    return a;
}

template <int D, int K, typename T> void launch_force(sphgpu_ctx *c, const int *skip_if)
{
    ForceArgs<T> a = force_args<T>(c);
    a.skip_if = skip_if;
    const dim3 g(grid_for(c->n, 256)), b(256);
    if (c->prm.sph_type == SPHGPU_SSPH) {
        hipLaunchKernelGGL((k_force_sph<D, K, SPHGPU_SSPH, T>), g, b, 0, c->stream, a);
    } else if (c->prm.sph_type == SPHGPU_DISPH) {
        hipLaunchKernelGGL((k_force_sph<D, K, SPHGPU_DISPH, T>), g, b, 0, c->stream, a);
    } else if (c->prm.sph_type == SPHGPU_GDISPH) {
        hipLaunchKernelGGL((k_force_sph<D, K, SPHGPU_GDISPH, T>), g, b, 0, c->stream, a);
    } else if (c->prm.riemann_solver == SPHGPU_RIEMANN_HLL) {
        hipLaunchKernelGGL((k_force_gsph<D, K, T, SPHGPU_RIEMANN_HLL, 0>), g, b, 0, c->stream, a);
    } else {
        // one kernel for both guesses (the network is off when a.mlp.w is
        // null); this layout measured faster than a guess-free instantiation
        hipLaunchKernelGGL((k_force_gsph<D, K, T, SPHGPU_RIEMANN_ITERATIVE, 1>), g, b, 0, c->stream, a);
    }
}

template <int D, typename T> void dispatch_force_kernel(sphgpu_ctx *c, const int *skip_if)
{
    switch (c->prm.kernel) {
    case SPHGPU_CUBIC_SPLINE: launch_force<D, SPHGPU_CUBIC_SPLINE, T>(c, skip_if); break;
    case SPHGPU_WENDLAND: if (D > 1) launch_force<D, SPHGPU_WENDLAND, T>(c, skip_if); break;
    default: launch_force<D, SPHGPU_GAUSSIAN, T>(c, skip_if); break;
    }
}

template <typename T> void dispatch_force(sphgpu_ctx *c, const int *skip_if)
{
    if (c->dim == 1) dispatch_force_kernel<1, T>(c, skip_if);
    else if (c->dim == 2) dispatch_force_kernel<2, T>(c, skip_if);
    else dispatch_force_kernel<3, T>(c, skip_if);
}

} // namespace

static int pre_core(sphgpu_ctx *c, bool defer, bool guess_h);

int pre_interaction(sphgpu_ctx *c) { return pre_core(c, false, true); }

int pre_interaction_deferred(sphgpu_ctx *c)
{
    return pre_core(c, !c->prm.use_time_dependent_av || c->alpha_snap != nullptr, true);
}

int pre_resolve(sphgpu_ctx *c)
{
    if (!c->pre_search_pending) return 0;
    c->pre_search_pending = false;
    int w[3] = {1, 0, 1};
    HIPCHK(c, hipMemcpyAsync(w, c->dflag, sizeof(w), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (w[0]) {
        // bit 0: the search overflowed: grow its lists (the reference's
        // capacity is nn * 1000).  Bit 1: a particle's h reached the tight
        // search radius: the next steps search wider.  Either way the
        // pre-interaction runs again, at the reference's radius
        if (w[0] & 2) {
            c->pre_eps = std::min(2.0 * c->pre_eps, 0.2);
            c->pre_quiet = 0;
        }
        if (w[0] & 1) {
            int rc = lists_grow(c, c->lpre);
            if (rc) return rc;
        }
        c->pre_rscale = 1.0;
        return pre_core(c, false, false);
    }
    if (c->flist == 1) c->flist = w[2] ? 0 : 2;
    return 0;
}

static int pre_core(sphgpu_ctx *c, bool defer, bool guess_h)
{
    if (!c->tree.valid) return fail(c, SPHGPU_E_STATE, "pre_interaction needs a built tree");
    int rc;
    if (c->mirror_dirty && (rc = refresh_mirrors(c))) return rc;
    const double kappa = c->prm.iterative_sml ? 1.2 : 1.0;
    if (c->pre_first) {
        // PreInteraction::initial_smoothing: h = (nn m / (rho A))^(1/D), density
        hguess(c, 1.0, c->sc[S_SML]);
        if ((rc = search(c, c->lpre, 0, c->sc[S_SML]))) return rc;
        const double *xd[3];
        for (int k = 0; k < 3; ++k) xd[k] = c->pos[k < c->dim ? k : 0];
        double **xdev = nullptr;
        // the pointer table lives in the reduction scratch (3 entries)
        xdev = (double **)c->red;
        HIPCHK(c, hipMemcpyAsync(xdev, xd, sizeof(xd), hipMemcpyHostToDevice, c->stream));
        time_begin(c, "initial_smoothing");
        if (c->dim == 1) launch_initial_density<1>(c, (const double *const *)xdev);
        else if (c->dim == 2) launch_initial_density<2>(c, (const double *const *)xdev);
        else launch_initial_density<3>(c, (const double *const *)xdev);
        time_end(c);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->pre_first = false;
    }
    const bool fl = c->prm.iterative_sml != 0;
    if (guess_h) {
        // the guess launch also clears the search's overflow flag, the
        // force-list flag and the h / v_sig minima
        time_begin(c, "pre_hguess");
        hguess(c, kappa, c->hs, true);
        time_end(c);
        // time-dependent AV: a repeated run of this step's pre-interaction
        // (list overflow, tight radius) starts alpha from the step's value
        c->alpha_snap_on = defer && c->prm.use_time_dependent_av && c->alpha_snap;
        if (c->alpha_snap_on)
            HIPCHK(c, hipMemcpyAsync(c->alpha_snap, c->sc[S_ALPHA], sizeof(double) * c->n, hipMemcpyDeviceToDevice,
                                     c->stream));
        // The time step's deferred pre-interaction searches within
        // (1 + pre_eps) h_guess instead of kappa h_guess: a step's h stays
        // within a few per cent of its guess (the Newton iteration starts
        // there and mostly stops after one step), and every sum of the
        // reference runs over the candidates with r < h only, so the results
        // are the same as long as no h reaches the tight radius (checked per
        // particle; such a step runs again at the reference's radius).
        c->pre_rscale = 1.0;
        if (defer && fl) {
            if (++c->pre_quiet >= 64) {
                c->pre_quiet = 0;
                c->pre_eps = std::max(0.05, 0.5 * c->pre_eps);
            }
            if (1.0 + c->pre_eps < kappa) c->pre_rscale = (1.0 + c->pre_eps) / kappa;
        }
    }
    if ((rc = search(c, c->lpre, 0, c->hs, defer, guess_h, c->pre_rscale))) return rc;
    if (!guess_h)
        hipLaunchKernelGGL(k_minbits_init, dim3(1), dim3(kStatCopies), 0, c->stream, c->dmin, fl ? c->dflag + 2 : nullptr);
    time_begin(c, "pre_interaction");
    if (c->prm.pair_precision == SPHGPU_PAIR_FP64) dispatch_pre<double>(c);
    else dispatch_pre<float>(c);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    if (fl) {
        // the list-ordering launch stores h / v_sig and hands the list flags
        // to the host; flag_ready marks them readable
        if (c->dim == 1) launch_flist<1>(c);
        else if (c->dim == 2) launch_flist<2>(c);
        else launch_flist<3>(c);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->flag_ready, c->stream));
        c->flist = 1;  // checked (one flag read) by the first fluid_force
    } else {
        hipLaunchKernelGGL(k_store_hvs, dim3(1), dim3(1), 0, c->stream, (const unsigned long long *)c->dmin, c->dscal);
        HIPCHK(c, hipGetLastError());
    }
    c->have_pre = true;
    return tree_set_kernel(c);
}

int fluid_force(sphgpu_ctx *c)
{
    if (!c->tree.valid) return fail(c, SPHGPU_E_STATE, "fluid_force needs a built tree");
    int rc;
    if (c->mirror_dirty && (rc = refresh_mirrors(c))) return rc;
    if (c->flist == 1) {
        // Derived lists not yet checked: the force runs at once over them and
        // skips itself on the device when a list flag is set (the search's
        // overflow, or a force list that needs the is_ij search); the host
        // reads the same flags meanwhile.  After a skip, nothing was written
        // and the checked path below runs.
        time_begin(c, "fluid_force");
        if (c->prm.pair_precision == SPHGPU_PAIR_FP64) dispatch_force<double>(c, c->dflag);
        else dispatch_force<float>(c, c->dflag);
        time_end(c);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventSynchronize(c->flag_ready));
        if (c->hflag[0] == 0 && c->hflag[2] == 0) {
            c->pre_search_pending = false;
            c->flist = 2;
            return 0;
        }
    }
    // the pair lists derived by the pre-interaction when they are complete,
    // the is_ij tree search otherwise
    const int fc = flist_check(c);
    if (fc < 0) return fc;
    if (fc) {
        ++c->flist_searches;
        if ((rc = search(c, c->lforce, 1, c->sc[S_SML]))) return rc;
    }
    time_begin(c, "fluid_force");
    if (c->prm.pair_precision == SPHGPU_PAIR_FP64) dispatch_force<double>(c, nullptr);
    else dispatch_force<float>(c, nullptr);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    return 0;
}

namespace {

// Stages the states in d (10 n doubles: L, R, p*, v*) and di (2 n ints:
// iterations, learned-guess use), solves and copies the results back.
int riemann_batch_staged(sphgpu_ctx *c, int solver, int precision, int n, const double *left, const double *right,
                         double *pstar, double *vstar, int32_t *iters, int32_t *learned_used, double *d, int *di)
{
    const size_t b = sizeof(double) * (size_t)n;
    double *L = d, *R = d + 4ull * n, *P = d + 8ull * n, *V = d + 9ull * n;
    HIPCHK(c, hipMemcpyAsync(L, left, b * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(R, right, b * 4, hipMemcpyHostToDevice, c->stream));
    const dim3 g(grid_for(n, 256)), bl(256);
    const int ref = riemann_refine64(c);
    const MLPDev m = mlp_dev(c);
    const bool use_mlp = precision != SPHGPU_PAIR_FP64 && solver == SPHGPU_RIEMANN_ITERATIVE &&
                         c->prm.riemann_guess == SPHGPU_GUESS_LEARNED && m.w;
    if (use_mlp) {
        hipLaunchKernelGGL(k_riemann_batch_learned, g, bl, 0, c->stream, n, (const double *)L, (const double *)R,
                           (float)c->prm.gamma, ref, m.img, P, V, di, di + n);
    } else if (precision == SPHGPU_PAIR_FP64) {
        if (solver == SPHGPU_RIEMANN_HLL)
            hipLaunchKernelGGL((k_riemann_batch<double, SPHGPU_RIEMANN_HLL>), g, bl, 0, c->stream, n, L, R, c->prm.gamma, ref, P, V, di);
        else
            hipLaunchKernelGGL((k_riemann_batch<double, SPHGPU_RIEMANN_ITERATIVE>), g, bl, 0, c->stream, n, L, R, c->prm.gamma, ref, P, V, di);
    } else {
        if (solver == SPHGPU_RIEMANN_HLL)
            hipLaunchKernelGGL((k_riemann_batch<float, SPHGPU_RIEMANN_HLL>), g, bl, 0, c->stream, n, L, R, c->prm.gamma, ref, P, V, di);
        else
            hipLaunchKernelGGL((k_riemann_batch<float, SPHGPU_RIEMANN_ITERATIVE>), g, bl, 0, c->stream, n, L, R, c->prm.gamma, ref, P, V, di);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(pstar, P, b, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(vstar, V, b, hipMemcpyDeviceToHost, c->stream));
    if (iters) HIPCHK(c, hipMemcpyAsync(iters, di, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (learned_used) {
        if (use_mlp)
            HIPCHK(c, hipMemcpyAsync(learned_used, di + n, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        else
            std::memset(learned_used, 0, sizeof(int32_t) * (size_t)n);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Staging area of the allocation-free batch: 10 doubles and 2 ints per
// problem for up to SPHGPU_RIEMANN_STATIC_MAX problems (2.75 MiB), one copy
// per device for the life of the process.  Its name is a host-side stand-in;
// device code reaches it only through the address hipGetSymbolAddress returns
// for the current device.
constexpr size_t kStaticBytes = (size_t)SPHGPU_RIEMANN_STATIC_MAX * (10 * sizeof(double) + 2 * sizeof(int));
__device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// this device's address of `scratch`, looked up once per device (with the
// device current) and kept for later calls
int static_scratch(sphgpu_ctx *c, void **out)
{
    int dev = -1;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(c, SPHGPU_E_ARG, "device index beyond the static staging table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(c, SPHGPU_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

} // namespace

int riemann_batch(sphgpu_ctx *c, int solver, int precision, int n, const double *left, const double *right,
                  double *pstar, double *vstar, int32_t *iters, int32_t *learned_used)
{
    double *d = nullptr;
    int *di = nullptr;
    HIPCHK(c, hipMalloc((void **)&d, sizeof(double) * 10 * (size_t)n));
    if (hipMalloc((void **)&di, 2 * sizeof(int) * (size_t)n) != hipSuccess) {
        hipFree(d);
        return fail(c, SPHGPU_E_NOMEM, "riemann batch staging allocation failed");
    }
    const int rc = riemann_batch_staged(c, solver, precision, n, left, right, pstar, vstar, iters, learned_used, d, di);
    hipFree(d);
    hipFree(di);
    return rc;
}

int riemann_batch_static(sphgpu_ctx *c, int solver, int precision, int n, const double *left, const double *right,
                         double *pstar, double *vstar, int32_t *iters, int32_t *learned_used)
{
    if (n > SPHGPU_RIEMANN_STATIC_MAX)
        return fail(c, SPHGPU_E_ARG, "sphgpu_riemann_batch_static: n = " + std::to_string(n) +
                                         " exceeds SPHGPU_RIEMANN_STATIC_MAX = " +
                                         std::to_string(SPHGPU_RIEMANN_STATIC_MAX));
    void *base = nullptr;
    const int rc = static_scratch(c, &base);
    if (rc) return rc;
    double *d = (double *)base;
    int *di = (int *)(d + 10ull * n);
    return riemann_batch_staged(c, solver, precision, n, left, right, pstar, vstar, iters, learned_used, d, di);
}

} // namespace sphgpu
