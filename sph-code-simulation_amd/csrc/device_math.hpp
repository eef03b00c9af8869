// device_math.hpp -- per-pair arithmetic of the SPH hot path, for gfx950.
//
// Everything here is templated on the pair precision T (float or double) and
// the dimension D.  With T = double the expressions keep the association of
// the reference code they restate, so the FP64 form differs from the
// reference only through summation order; with T = float the same formulas
// run at FP32 rate.  All HIP sources are compiled with -ffp-contract=off;
// functions that may use FMA say so with a local pragma.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sphgpu.h"

#define SPH_DEV __device__ __forceinline__

namespace sphgpu {

constexpr int kWave = 64;

// Counters: kStatCopies copies of kStatSlots slots.  A launch's per-wave
// atomics go to the copy of their workgroup, which spreads them over many
// addresses instead of serialising them on one; the host sums the copies.
constexpr int kStatSlots = 32, kStatCopies = 64;
SPH_DEV unsigned long long *stat_ptr(long long *stats, int slot)
{
    return (unsigned long long *)&stats[kStatSlots * (blockIdx.x & (kStatCopies - 1)) + slot];
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Particle kernels remap the workgroup index so that XCD x
// runs one contiguous range of logical workgroups: particles consecutive in
// key order (spatial neighbours, whose pair data overlap) then share an L2.
constexpr int kNumXcd = 8;
SPH_DEV int xcd_block(int b, int nb)
{
    const int q = nb / kNumXcd, r = nb % kNumXcd;
    const int x = b % kNumXcd, k = b / kNumXcd;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

template <typename T> SPH_DEV T tabs(T x) { return x < T(0) ? -x : x; }
// FP32 square root: the hardware v_sqrt_f32 (1 ulp) without the compiler's
// rescaling of denormal inputs, which pair quantities (r^2 of distinct
// particles, pressures and densities) never are; sqrt(0) = 0 stays exact
SPH_DEV float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
SPH_DEV double tsqrt(double x) { return __builtin_sqrt(x); }
// 1 / sqrt(x): the hardware estimate in FP32 (about 1 ulp), IEEE in FP64
SPH_DEV float trsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
SPH_DEV double trsqrt(double x) { return 1.0 / __builtin_sqrt(x); }
SPH_DEV float texp(float x) { return __expf(x); }
SPH_DEV double texp(double x) { return exp(x); }

// Division in the pair precision: IEEE division for double (the reference's
// expressions), a * rcp(b) for float (v_rcp_f32, 1 ulp; the FP32 form trades
// the correctly rounded quotient for a 5x shorter sequence).
SPH_DEV double tdiv(double a, double b) { return a / b; }
SPH_DEV float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Periodic::calc_r_ij per axis (include/periodic.hpp:34-59): choose among
// d, d + L, d - L the one of smallest magnitude, ties in the order (d, d+L, d-L).
template <typename T> SPH_DEV T min_image(T d, T L)
{
    const T r1 = d, r2 = d + L, r3 = d - L;
    const T a1 = tabs(r1), a2 = tabs(r2), a3 = tabs(r3);
    if (a1 <= a2 && a1 <= a3) return r1;
    if (a2 <= a3 && a2 <= a1) return r2;
    if (a3 <= a1 && a3 <= a2) return r3;
    return r1;
}

// Box geometry shared by all kernels (FP64 master copy plus an FP32 copy).
struct Box {
    int periodic;
    double range[3];
    double lo[3];
    double hi[3];
    float rangef[3];
    float rinvf[3];   // 1 / rangef
};

template <typename T> SPH_DEV T box_range(const Box &b, int k);
template <> SPH_DEV double box_range<double>(const Box &b, int k) { return b.range[k]; }
template <> SPH_DEV float box_range<float>(const Box &b, int k) { return b.rangef[k]; }

// r_i - r_j with the reference's minimum-image choice (FP64).
template <typename T> SPH_DEV T rij_axis(T xi, T xj, const Box &b, int k)
{
    const T d = xi - xj;
    return b.periodic ? min_image(d, box_range<T>(b, k)) : d;
}

// FP32: d - L rint(d / L), three instructions.  It differs from the
// reference's choice only at |d| = L/2 (the tie order decides the sign there),
// and r^2 not at all; membership tests near such ties go through the FP64
// predicate anyway.
template <> SPH_DEV float rij_axis<float>(float xi, float xj, const Box &b, int k)
{
    const float d = xi - xj;
    return b.periodic ? __builtin_fmaf(-b.rangef[k], __builtin_rintf(d * b.rinvf[k]), d) : d;
}

// include/defines.hpp:28-32
template <typename T> SPH_DEV T sqr(T x) { return x * x; }
template <typename T> SPH_DEV T pow3(T x) { return x * x * x; }
template <typename T> SPH_DEV T pow5(T x) { return x * x * x * x * x; }
template <typename T> SPH_DEV T pow6(T x) { return x * x * x * x * x * x; }

// (a)_+ as the kernels write it, 0.5 (a + |a|): for a > 0 the sum 2a and the
// halving are exact and for a < 0 the sum is +0, so it equals max(a, 0) bit
// for bit (one instruction; a = c - q is never -0)
SPH_DEV float pos_part(float a) { return __builtin_fmaxf(a, 0.f); }
SPH_DEV double pos_part(double a) { return __builtin_fmax(a, 0.0); }
// ((|2 - q| + 2) - q) / 2, the cubic spline's dhw form: the sum rounds in
// another order, so FP64 keeps the reference's expression; FP32 uses the max
template <typename T> SPH_DEV T pos_part_dhw(T c, T q) { return (tabs(c - q) + c - q) * T(0.5); }
template <> SPH_DEV float pos_part_dhw<float>(float c, float q) { return fmaxf(c - q, 0.f); }

// |minimum-image d| in FP32 for r^2 only: min(|d|, L - |d|) equals |d - L
// rint(d / L)| for |d| < L (the box keeps positions there), two instructions
SPH_DEV float abs_rij_axis(float xi, float xj, const Box &b, int k)
{
    const float d = fabsf(xi - xj);
    return b.periodic ? fminf(d, fabsf(b.rangef[k] - d)) : d;
}

// include/kernel/kernel_function.hpp:8-16 and pre_interaction.cpp:217-225
template <int D, typename T> SPH_DEV T powh(T h) { return D == 1 ? h : D == 2 ? h * h : h * h * h; }
template <int D, typename T> SPH_DEV T powh_m1(T h) { return D == 1 ? T(1) : D == 2 ? h : h * h; }

template <int D> struct Sigma;
template <> struct Sigma<1> {
    static constexpr double cubic = 2.0 / 3.0, c4 = 0.0, gauss = 0.564189583547756;
};
template <> struct Sigma<2> {
    static constexpr double cubic = 10.0 / (7.0 * 3.1415926535897932384626433832795028841971693993751);
    static constexpr double c4 = 9.0 / 3.1415926535897932384626433832795028841971693993751;
    static constexpr double gauss = 1.0 / 3.1415926535897932384626433832795028841971693993751;
};
template <> struct Sigma<3> {
    static constexpr double cubic = 1.0 / 3.1415926535897932384626433832795028841971693993751;
    static constexpr double c4 = 495. / (32 * 3.1415926535897932384626433832795028841971693993751);
    static constexpr double gauss = 0.179587122125167;
};

// SPH kernels, support r < h.  w(r,h); dwc(r,h) with grad W = r_ij * dwc;
// dhw(r,h) = dW/dh.  Cubic spline: include/kernel/cubic_spline.hpp:27-52.
// Wendland C4: include/kernel/wendland_kernel.hpp:30-48.  Gaussian:
// include/kernel/gaussian_kernel.hpp:27-51.
template <int D, int K, typename T> struct Kern;

template <int D, typename T> struct Kern<D, SPHGPU_CUBIC_SPLINE, T> {
    static SPH_DEV T w(T r, T h)
    {
        const T sig = T(Sigma<D>::cubic);
        const T h_ = h * T(0.5);
        const T q = tdiv(r, h_);
        return tdiv(sig, powh<D>(h_)) * (T(0.25) * pow3(pos_part(T(2.0) - q)) - pow3(pos_part(T(1.0) - q)));
    }
    static SPH_DEV T dwc(T r, T h)
    {
        if (r == T(0)) return T(0);
        const T sig = T(Sigma<D>::cubic);
        const T h_ = h * T(0.5);
        const T q = tdiv(r, h_);
        return tdiv(-sig, powh<D>(h_) * h_ * r) *
               (T(0.75) * sqr(pos_part(T(2.0) - q)) - T(3.0) * sqr(pos_part(T(1.0) - q)));
    }
    static SPH_DEV T dhw(T r, T h)
    {
        const T sig = T(Sigma<D>::cubic);
        const T h_ = h * T(0.5);
        const T q = tdiv(r, h_);
        return tdiv(T(0.5) * sig, powh<D>(h_) * h_) *
               (sqr(pos_part_dhw(T(2.0), q)) * (T((3. + D) * 0.25) * q - T(0.5 * D)) +
                sqr(pos_part_dhw(T(1.0), q)) * (T(-3.0 - D) * q + T(D)));
    }
};

template <int D, typename T> struct Kern<D, SPHGPU_WENDLAND, T> {
    static SPH_DEV T w(T r, T h)
    {
        const T sig = T(Sigma<D>::c4);
        const T q = tdiv(r, h);
        return tdiv(sig, powh<D>(h)) * pow6(pos_part(T(1.0) - q)) * (T(1.0) + T(6.0) * q + T(35.0 / 3.0) * q * q);
    }
    static SPH_DEV T dwc(T r, T h)
    {
        const T sig = T(Sigma<D>::c4);
        const T q = tdiv(r, h);
        return tdiv(T(-56.0 / 3.0) * sig, powh<D>(h) * sqr(h)) * pow5(pos_part(T(1.0) - q)) * (T(1.0) + T(5.0) * q);
    }
    static SPH_DEV T dhw(T r, T h)
    {
        const T sig = T(Sigma<D>::c4);
        const T q = tdiv(r, h);
        return tdiv(-sig, powh<D>(h) * h * T(3.0)) * pow5(pos_part(T(1.0) - q)) *
               (T(3.0 * D) + T(15.0 * D) * q + T(-56.0 + 17.0 * D) * q * q - T(35.0 * (8.0 + D)) * pow3(q));
    }
};

template <int D, typename T> struct Kern<D, SPHGPU_GAUSSIAN, T> {
    static SPH_DEV T w(T r, T h)
    {
        const T q = tdiv(r, h);
        return tdiv(T(Sigma<D>::gauss), powh<D>(h)) * texp(-(q * q));
    }
    static SPH_DEV T dwc(T r, T h)
    {
        if (r == T(0)) return T(0);
        const T q = tdiv(r, h);
        const T w_val = tdiv(T(Sigma<D>::gauss), powh<D>(h)) * texp(-(q * q));
        return tdiv(T(-2.0) * w_val * q, h * r);
    }
    static SPH_DEV T dhw(T r, T h)
    {
        const T q = tdiv(r, h);
        const T q2 = q * q;
        const T w_val = tdiv(T(Sigma<D>::gauss), powh<D>(h)) * texp(-q2);
        return tdiv(w_val * (T(2.0) * q2 - T(D)), h);
    }
};

// dwc for the FP32 pair loops with the reciprocals 1/r and 1/h supplied by
// the caller (1/h_i is formed once per particle, 1/r once per pair): the same
// expressions as Kern::dwc with their divisions written as products.
template <int D, int K> SPH_DEV float dwc_rcp(float r, float r_inv, float ih)
{
    if (K == SPHGPU_CUBIC_SPLINE) {
        if (r == 0.f) return 0.f;
        const float ih_ = 2.0f * ih;  // 1 / (h / 2)
        const float q = r * ih_;
        const float c = float(-Sigma<D>::cubic) * powh<D>(ih_) * ih_ * r_inv;
        return c * (0.75f * sqr(pos_part(2.0f - q)) - 3.0f * sqr(pos_part(1.0f - q)));
    } else if (K == SPHGPU_WENDLAND) {
        const float q = r * ih;
        const float c = float(-56.0 / 3.0 * Sigma<D>::c4) * powh<D>(ih) * sqr(ih);
        return c * pow5(pos_part(1.0f - q)) * (1.0f + 5.0f * q);
    } else {
        if (r == 0.f) return 0.f;
        const float q = r * ih;
        const float w_val = float(Sigma<D>::gauss) * powh<D>(ih) * texp(-(q * q));
        return -2.0f * w_val * q * ih * r_inv;
    }
}

// van Leer limiter, src/gsph/g_fluid_force.cpp:33-41
template <typename T> SPH_DEV T limiter(T dq1, T dq2)
{
    const T p = dq1 * dq2;
    if (p <= T(0)) return T(0);
    return tdiv(T(2.0) * p, dq1 + dq2);
}

// HLL solver, src/gsph/g_fluid_force.cpp:173-204.  State = (u, rho, p, c).
template <typename T> SPH_DEV void hll_solver(const T l[4], const T r[4], T &pstar, T &vstar)
{
    const T u_l = l[0], rho_l = l[1], p_l = l[2], c_l = l[3];
    const T u_r = r[0], rho_r = r[1], p_r = r[2], c_r = r[3];
    const T roe_l = tsqrt(rho_l);
    const T roe_r = tsqrt(rho_r);
    const T roe_inv = tdiv(T(1.0), roe_l + roe_r);
    const T u_t = (roe_l * u_l + roe_r * u_r) * roe_inv;
    const T c_t = (roe_l * c_l + roe_r * c_r) * roe_inv;
    const T a = u_l - c_l, b = u_t - c_t;
    const T s_l = (b < a) ? b : a;      // std::min
    const T c = u_r + c_r, d = u_t + c_t;
    const T s_r = (c < d) ? d : c;      // std::max
    const T c1 = rho_l * (s_l - u_l);
    const T c2 = rho_r * (s_r - u_r);
    const T c3 = tdiv(T(1.0), c1 - c2);
    const T c4 = p_l - u_l * c1;
    const T c5 = p_r - u_r * c2;
    vstar = (c5 - c4) * c3;
    pstar = (c1 * c5 - c2 * c4) * c3;
}

// Two-shock Newton iteration of src/gsph/g_fluid_force.cpp:209-304, split so
// the starting guess can be supplied (the learned guess replaces :244-246).
template <typename T> struct TwoShock {
    T u_l, p_l, u_r, p_r, V_l, V_r, cl, cr, gamma1, gamma2;
    bool bad;
    SPH_DEV void setup(T gamma, const T l[4], const T r[4])
    {
        u_l = l[0]; p_l = l[2]; u_r = r[0]; p_r = r[2];
        const T rho_l = l[1], rho_r = r[1];
        bad = (rho_l < T(0) || rho_r < T(0) || p_l < T(0) || p_r < T(0));
        gamma2 = T(1.0) + gamma;
        gamma1 = tdiv(T(0.5) * gamma2, gamma);
        V_l = tdiv(T(1.0), rho_l);
        V_r = tdiv(T(1.0), rho_r);
        cl = tsqrt(gamma * p_l * rho_l);
        cr = tsqrt(gamma * p_r * rho_r);
    }
    // the reference's linearised impedance guess, g_fluid_force.cpp:244-246
    SPH_DEV T plain_guess() const
    {
        T g = p_r - p_l - cr * (u_r - u_l);
        g = p_l + tdiv(g * cl, cl + cr);
        return (g < T(1e-25)) ? T(1e-25) : g;  // std::max(g, smallp)
    }
    // one Newton step, g_fluid_force.cpp:257-279
    SPH_DEV T step(T p) const
    {
        T w_l = T(1.0) + tdiv(gamma1 * (p - p_l), p_l);
        w_l = cl * tsqrt(w_l);
        T w_r = T(1.0) + tdiv(gamma1 * (p - p_r), p_r);
        w_r = cr * tsqrt(w_r);
        T z_l = T(4.0) * V_l * w_l * w_l;
        z_l = tdiv(-z_l * w_l, z_l - gamma2 * (p - p_l));
        T z_r = T(4.0) * V_r * w_r * w_r;
        z_r = tdiv(z_r * w_r, z_r - gamma2 * (p - p_r));
        const T ustar_l = u_l - tdiv(p - p_l, w_l);
        const T ustar_r = u_r + tdiv(p - p_r, w_r);
        T pn = p + tdiv((ustar_r - ustar_l) * (z_l * z_r), z_r - z_l);
        return (T(1e-25) < pn) ? pn : T(1e-25);  // std::max(smallp, p)
    }
    // The FP32 form's Newton step, which also returns its Newton factor
    // F = z_l z_r / (z_r - z_l) (the last iterate is refined in FP64 with F;
    // the update is written p + (u*_R - u*_L) F).  The same quantities as
    // step() with three reciprocal-type instructions instead of nine and no
    // per-pair reciprocals: w_K^2 = c_K^2 (1 + gamma1 (p - p_K) / p_K) =
    // Y_K = gamma rho_K (p_K + gamma1 (p - p_K)), so with s_K = 1/sqrt(Y_K),
    // 1/w_K = s_K and w_K = Y_K s_K.  With A_K = 4 V_K w_K^3 and
    // B_K = 4 V_K w_K^2 - gamma2 (p - p_K), z_l = -A_l / B_l, z_r = A_r / B_r
    // and F = -A_l A_r / (A_r B_l + A_l B_r), which is unchanged when A_K and
    // B_K are both scaled by rho_K: rho_K A_K = 4 Y_K w_K and rho_K B_K =
    // 4 Y_K - gamma2 rho_K (p - p_K).  Y_K > 0 for p >= 0 (p_K + gamma1
    // (p - p_K) = (1 - gamma1) p_K + gamma1 p with gamma1 < 1).  prep_f()
    // forms the per-pair products first.
    T gr_l, gr_r, g2r_l, g2r_r;
    SPH_DEV void prep_f(T gamma, T rho_l, T rho_r)
    {
        gr_l = gamma * rho_l;
        gr_r = gamma * rho_r;
        g2r_l = gamma2 * rho_l;
        g2r_r = gamma2 * rho_r;
    }
    SPH_DEV T step_f(T p, T &F) const
    {
        const T dl = p - p_l, dr = p - p_r;
        const T Y_l = gr_l * (p_l + gamma1 * dl);
        const T Y_r = gr_r * (p_r + gamma1 * dr);
        const T s_l = trsqrt(Y_l), s_r = trsqrt(Y_r);
        const T w_l = Y_l * s_l, w_r = Y_r * s_r;
        const T A_l = T(4.0) * Y_l * w_l, A_r = T(4.0) * Y_r * w_r;
        const T B_l = T(4.0) * Y_l - g2r_l * dl;
        const T B_r = T(4.0) * Y_r - g2r_r * dr;
        F = -(A_l * A_r) * tdiv(T(1.0), A_r * B_l + A_l * B_r);
        const T ustar_l = u_l - dl * s_l;
        const T ustar_r = u_r + dr * s_r;
        const T pn = p + (ustar_r - ustar_l) * F;
        return (T(1e-25) < pn) ? pn : T(1e-25);
    }
    // the convergence test |dp| / p < 1e-6 as |dp| < 1e-6 p (p > 0)
    SPH_DEV int iterate_f(T &p, T &F) const
    {
        int it = 0;
        for (int iter = 0; iter < 20; ++iter) {
            const T p_old = p;
            p = step_f(p, F);
            it = iter + 1;
            if (tabs(p - p_old) < T(1e-6) * p) break;
        }
        return it;
    }
    // the Newton factor z_l z_r / (z_r - z_l) of step() at p
    SPH_DEV T newton_factor(T p) const
    {
        T w_l = T(1.0) + tdiv(gamma1 * (p - p_l), p_l);
        w_l = cl * tsqrt(w_l);
        T w_r = T(1.0) + tdiv(gamma1 * (p - p_r), p_r);
        w_r = cr * tsqrt(w_r);
        T z_l = T(4.0) * V_l * w_l * w_l;
        z_l = tdiv(-z_l * w_l, z_l - gamma2 * (p - p_l));
        T z_r = T(4.0) * V_r * w_r * w_r;
        z_r = tdiv(z_r * w_r, z_r - gamma2 * (p - p_r));
        return tdiv(z_l * z_r, z_r - z_l);
    }
    // g_fluid_force.cpp:254-286; returns the iteration count
    SPH_DEV int iterate(T &p) const
    {
        int it = 0;
        for (int iter = 0; iter < 20; ++iter) {
            const T p_old = p;
            p = step(p);
            it = iter + 1;
            if (tdiv(tabs(p - p_old), p) < T(1e-6)) break;
        }
        return it;
    }
    // g_fluid_force.cpp:288-301
    SPH_DEV T vstar(T p) const
    {
        T w_l = T(1.0) + tdiv(gamma1 * (p - p_l), p_l);
        w_l = cl * tsqrt(w_l);
        T w_r = T(1.0) + tdiv(gamma1 * (p - p_r), p_r);
        w_r = cr * tsqrt(w_r);
        const T ustar_l = u_l - tdiv(p - p_l, w_l);
        const T ustar_r = u_r + tdiv(p - p_r, w_r);
        return T(0.5) * (ustar_l + ustar_r);
    }
};

// 1 / sqrt(x) in FP64: the hardware estimate (v_rsq_f64, about 2^-22
// relative) and two Newton steps y (3 - x y^2) / 2, each squaring the
// relative error; a non-positive x gives NaN or inf as sqrt and division do
SPH_DEV double rsqrt64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    y = y * (1.5 - 0.5 * x * y * y);
    return y;
}

// refine_step64 with v* = (u*_L + u*_R) / 2 of the same FP64 evaluation at
// p32 (g_fluid_force.cpp:288-301 evaluates v* at the final iterate); the
// quotients (p - p0) / w are products with 1 / w = rsqrt(gamma rho (p0 +
// gamma1 (p - p0))) (a few ulp, far below the FP32 p* this refines)
SPH_DEV double refine_step64_v(const float l[4], const float r[4], double gamma, double gamma1, float p32, float F,
                               double &vstar)
{
    const double p = p32;
    const double u_l = l[0], rho_l = l[1], p_l = l[2];
    const double u_r = r[0], rho_r = r[1], p_r = r[2];
    const double iw_l = rsqrt64(gamma * rho_l * (p_l + gamma1 * (p - p_l)));
    const double iw_r = rsqrt64(gamma * rho_r * (p_r + gamma1 * (p - p_r)));
    const double ustar_l = u_l - (p - p_l) * iw_l;
    const double ustar_r = u_r + (p - p_r) * iw_r;
    vstar = 0.5 * (ustar_l + ustar_r);
    const double pn = p + (ustar_r - ustar_l) * (double)F;
    return (1e-25 < pn) ? pn : 1e-25;  // std::max(smallp, p)
}

// One FP64 Newton step of the two-shock iteration from an FP32 iterate p32:
// the residual u*_R(p) - u*_L(p), which cancels near the root, is evaluated
// in FP64 (w = c sqrt(1 + gamma1 (p - p0)/p0) written as
// sqrt(gamma rho (p0 + gamma1 (p - p0)))), while the Newton factor F =
// z_l z_r / (z_r - z_l), which only scales the small correction, comes from
// FP32.  The step's relative error is about (FP32 rounding) x (correction).
SPH_DEV double refine_step64(const float l[4], const float r[4], double gamma, double gamma1, float p32, float F)
{
    const double p = p32;
    const double u_l = l[0], rho_l = l[1], p_l = l[2];
    const double u_r = r[0], rho_r = r[1], p_r = r[2];
    const double w_l = __builtin_sqrt(gamma * rho_l * (p_l + gamma1 * (p - p_l)));
    const double w_r = __builtin_sqrt(gamma * rho_r * (p_r + gamma1 * (p - p_r)));
    const double ustar_l = u_l - (p - p_l) / w_l;
    const double ustar_r = u_r + (p - p_r) / w_r;
    const double pn = p + (ustar_r - ustar_l) * (double)F;
    return (1e-25 < pn) ? pn : 1e-25;  // std::max(smallp, p)
}

// Exact inequality of the reference's FP64 neighbour predicates.
// sqrt(r2) < h, with the correctly rounded FP64 sqrt.
SPH_DEV bool r_less(double r2, double h) { return __builtin_sqrt(r2) < h; }

// FP32 prefilter for the FP64 membership tests r^2 < h^2.  With positions
// bounded by X (|x_k| <= X, and the box range L <= X) and u = 2^-24,
// rounding the two positions (2uX), their difference (2uX), the box range
// (uX) and the periodic correction (u|r_ij|) moves each r_ij component by at
// most 5uX + u|r_ij|, and the FP32 sum of squares adds (D+1)u r^2, so near
// r = h
//   |r2_fp32 - r2_fp64| <= 10 sqrt(3) u X h + 6 u h^2   (first order).
// The band half-width used, delta * h^2 with delta = k / h + c,
//   k = 20 sqrt(3) u X,  c = 32 u,
// is at least twice that bound (and also covers rounding h, h^2 and
// h^2 (1 +- delta) in FP32).  Outside the band the FP32 comparison decides exactly as the
// FP64 one would; inside it the caller evaluates the FP64 predicate.
// root = tree root (center[3], edge); X is taken as max |center_k| + edge.
SPH_DEV float fp32_margin_k(const double *root, int D)
{
    double X = 0.0;
    for (int k = 0; k < D; ++k) {
        const double a = root[k] < 0.0 ? -root[k] : root[k];
        X = a > X ? a : X;
    }
    X += root[3];
    // beyond 1e18 the FP32 squares could overflow: no band (always FP64)
    return X < 1e18 ? (float)(20.0 * 1.7320508075688772 * 5.9604644775390625e-08 * X) : __builtin_inff();
}
constexpr float kMarginC = 32.0f * 5.9604644775390625e-08f;

// band of the test r2 < h2: returns +1 (inside for certain), -1 (outside for
// certain) or 0 (undecided: run the FP64 test)
SPH_DEV int fp32_band(float r2, float h, float margin_k)
{
    const float h2 = h * h;
    const float delta = margin_k / h + kMarginC;
    // tiny h: the FP32 squares could be denormal
    if (!(h2 > 1e-30f) || !(delta < 0.5f)) return 0;
    if (r2 < h2 * (1.0f - delta)) return 1;
    if (r2 > h2 * (1.0f + delta)) return -1;
    return 0;
}

// Neumaier-compensated accumulator for FP32 sums (plain sum for FP64).
template <typename T> struct Acc {
    T s, c;
    SPH_DEV Acc() : s(T(0)), c(T(0)) {}
    SPH_DEV void add(T x)
    {
        // Knuth's TwoSum: e = (s - (t - z)) + (x - z) is the exact rounding
        // error of t = s + x whatever the magnitudes (the same value as
        // Neumaier's ordered form, without its comparison)
        const T t = s + x;
        const T z = t - s;
        c += (s - (t - z)) + (x - z);
        s = t;
    }
    SPH_DEV T get() const { return s + c; }
};
template <> struct Acc<double> {
    double s;
    SPH_DEV Acc() : s(0.0) {}
    SPH_DEV void add(double x) { s += x; }
    SPH_DEV double get() const { return s; }
};

// Compensated sum in blocks, for long sums whose terms come one per loop
// step: each term costs one add into a plain partial p, and fold() adds the
// partial into the TwoSum pair (s, c) as Acc does.  With at most B terms per
// block the partial's rounding stays below (B - 1) u sum |x_k| (first order),
// against 2 u for Acc and (n - 1) u for a plain sum of all n terms.
template <typename T> struct BlockAcc {
    T s, c, p;
    SPH_DEV BlockAcc() : s(T(0)), c(T(0)), p(T(0)) {}
    SPH_DEV void add(T x) { p += x; }
    SPH_DEV void fold()
    {
        const T t = s + p;
        const T z = t - s;
        c += (s - (t - z)) + (p - z);
        s = t;
        p = T(0);
    }
    SPH_DEV T get() const { return s + (c + p); }
};
template <> struct BlockAcc<double> {
    double s;
    SPH_DEV BlockAcc() : s(0.0) {}
    SPH_DEV void add(double x) { s += x; }
    SPH_DEV void fold() {}
    SPH_DEV double get() const { return s; }
};

// Order-preserving integer image of a non-negative double, for atomicMin/Max.
SPH_DEV unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
SPH_DEV double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

} // namespace sphgpu
