// osb_common.hpp -- shared definitions for the gfx950 kernels of the
// OpenSBLI explicit-solver hot path.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "../../include/osb.h"
#include "osb_eigzero.hpp"
#include "osb_order.hpp"

namespace osb {

constexpr double INV_SQRT2 = 0.707106781186547;   // sqrt(2.0) printed to 15 s.f. (euler_eigensystem.py:109,139)

enum Scheme { SCH_WENOZ = 0, SCH_TENO5 = 1, SCH_TENO6 = 2 };

// Geometry of one padded block: SoA, x fastest, `halo` ghost layers in active
// directions.  Interior point (i,j,k) lives at base + i + s1*j + s2*k.
struct Geom {
    int n[3];          // interior extent
    int P[3];          // padded extent
    int o[3];          // offset of interior index 0 (halo or 0)
    int halo;
    int ndim;
    long s[3];         // strides (1, P0, P0*P1)
    long Ptot;         // points per variable
    // multiplicative inverses (cf_divisor: mp, l of an extent e0, then of
    // e0 e1) that split flat indices without integer division: dv[0] the
    // interior points (n0, n0 n1), dv[1 + d] the faces along d (n_d + 1 along
    // d), dv[4 + w] the fused face-residual tiles at width RESID_WS[w]
    // (columns, columns x segments), dv[6 + d] the boundary layers of
    // direction d (bc_work).  Set by geom_divisors.
    unsigned dv[9][4];
};

__host__ __device__ inline long gidx(const Geom& g, int i, int j, int k) {
    return (long)(i + g.o[0]) + g.s[1] * (long)(j + g.o[1]) + g.s[2] * (long)(k + g.o[2]);
}

struct Consts {
    double gama, gamma_m1, minf;
    double sfc;            // shock_filter_control
    double weno_eps, teno_eps, teno_ct;
    double invRe, cheat;   // 1/Re, 1/((gama-1) Minf^2 Pr Re)
    double invd[3];        // 1/Delta
    double inv2d[3];       // 1/Delta^2
    double cfl, dt;
    int avg_roe;
    int viscous;           // 0 none, 1 constant, 2 Sutherland
    // GLF splitting: per direction d, glf[3d..3d+2] = max over the grid of
    // |u_d|, |u_d + a|, |u_d - a| for the current stage; nullptr: LLF
    const double* glf;
    // Carpenter boundary closure of the central derivatives per direction
    // (bit 0: side 0, bit 1: side 1), set for extrapolation and Dirichlet faces
    int carp[3];
    // WENO filter: Ducros thresholds of the hybrid selection
    double ducros_check, ducros_select;
    // fused 3-D stage kernels (fill_fused_consts): central weights pre-scaled
    // by 1/Delta per direction (w1: first, w2: second derivative), products of
    // first-derivative weights for the mixed derivatives of the pairs (0,1),
    // (0,2), (1,2), and the factors of theta = p/rho: E = cE theta + |u|^2/2,
    // H = cH theta + |u|^2/2, T = cT theta, heat coefficient on theta
    double w1[3][2], w2[3][3], wm[3][4];
    double cE, cH, cT, cheat_th;
};

// ---------------------------------------------------------------------------
// Smoothness indicators and reconstructions.  Values arrive as T (double in
// the FP64 path, float in the mixed path); smoothness indicators, non-linear
// weights and the TENO cut-off are always evaluated in double.
// f[s] holds the split flux at offset s-2 from point i (interface i+1/2).
// The expression sequence (including every fused multiply-add, written out
// as fma) is fixed so that the FP64 path rounds exactly like the CPU oracle;
// the library is compiled with -ffp-contract=off.
// ---------------------------------------------------------------------------
// 13/12 s1^2 + 1/4 s2^2 (JS_smoothness, weno.py:139-206; teno.py:147-182)
__device__ __forceinline__ double js_sum(double s1, double s2) { return fma((13.0 / 12.0) * s1, s1, (0.25 * s2) * s2); }
__device__ __forceinline__ double js_first(double a, double b, double c) { return js_sum(a - 2.0 * b + c, 3.0 * a - 4.0 * b + c); }
__device__ __forceinline__ double js_centre(double a, double b, double c) { return js_sum(a - 2.0 * b + c, a - c); }
__device__ __forceinline__ double js_last(double a, double b, double c) { return js_sum(a - 2.0 * b + c, a - 4.0 * b + 3.0 * c); }

// Mixed precision: the three indicators over five consecutive values from
// their FP64 first and second differences (the FP32 inputs make the first
// differences exact): the same quantities as js_last(f0, f1, f2),
// js_centre(f1, f2, f3) and js_first(f2, f3, f4), with h the half of each
// second combination (h^2 = s2^2 / 4), in 20 operations instead of 30.
__device__ __forceinline__ void js_diff3(double f0, double f1, double f2, double f3, double f4, double& b_last,
                                         double& b_centre, double& b_first) {
    const double d1 = f1 - f0, d2 = f2 - f1, d3 = f3 - f2, d4 = f4 - f3;
    const double sl = d2 - d1, sc = d3 - d2, sf = d4 - d3;
    const double hl = fma(0.5, sl, d2);           // (f0 - 4 f1 + 3 f2) / 2
    const double hc = 0.5 * (d2 + d3);            // (f3 - f1) / 2
    const double hf = fma(0.5, sf, -d3);          // (3 f2 - 4 f3 + f4) / 2
    b_last = fma((13.0 / 12.0) * sl, sl, hl * hl);
    b_centre = fma((13.0 / 12.0) * sc, sc, hc * hc);
    b_first = fma((13.0 / 12.0) * sf, sf, hf * hf);
}

template <typename T>
__device__ __forceinline__ T eno3(T c0, T c1, T c2, T x0, T x1, T x2) { return fma(c2, x2, fma(c1, x1, c0 * x0)); }

// Reciprocals.  FP64 path: IEEE division, as the oracle.  rcp_nr: the hardware
// reciprocal refined by two Newton steps (the IEEE quotient on every input
// measured, without the scaled division sequence), used by the fused central
// stage kernel.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}
// One Newton step: the hardware reciprocal is good to 2^-24 relative on
// gfx950, one step to about 1e-15 (two reproduce the IEEE quotient).  For the
// mixed-precision path, whose FP64 results are rounded to FP32 downstream.
__device__ __forceinline__ double rcp_nr1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}
template <typename T>
__device__ __forceinline__ double wrecip(double x) {
    if constexpr (std::is_same<T, double>::value) return 1.0 / x;
    else return rcp_nr1(x);
}
template <typename T>
__device__ __forceinline__ double wdiv(double a, double b) {
    if constexpr (std::is_same<T, double>::value) return a / b;
    else return a * rcp_nr1(b);
}

template <typename T>
__device__ __forceinline__ double zalpha(double d, double t2, double b, double eps) {
    const double eb = eps + b;
    return d + wdiv<T>(d * t2, eb * eb);
}

// Non-linear WENO-Z weights omega_r (times 1/sum) from the smoothness
// indicators.  FP64 path: alpha_r = d_r + d_r tau^2/((eps+b_r)(eps+b_r)) and
// omega_r = alpha_r / sum(alpha), as printed by the reference.  Mixed path:
// the same weights with the three alpha denominators e_r = (eps+b_r)^2
// cleared -- omega_r = d_r (e_r + tau^2) e_s e_t / sum -- so one reciprocal
// replaces four divisions (e_r >= eps^2 = 1e-80 keeps the products in range).
template <typename T>
__device__ __forceinline__ void wenoz_weights(double d0, double d1, double d2, double b0, double b1, double b2, double eps,
                                              double& w0, double& w1, double& w2) {
    const double tau = fabs(b0 - b2), t2 = tau * tau;
    if constexpr (std::is_same<T, double>::value) {
        const double a0 = zalpha<T>(d0, t2, b0, eps);
        const double a1 = zalpha<T>(d1, t2, b1, eps);
        const double a2 = zalpha<T>(d2, t2, b2, eps);
        const double inv = 1.0 / (a0 + a1 + a2);
        w0 = a0 * inv; w1 = a1 * inv; w2 = a2 * inv;
    } else {
        const double e0 = (eps + b0) * (eps + b0), e1 = (eps + b1) * (eps + b1), e2 = (eps + b2) * (eps + b2);
        const double n0 = d0 * (e0 + t2) * (e1 * e2);
        const double n1 = d1 * (e1 + t2) * (e0 * e2);
        const double n2 = d2 * (e2 + t2) * (e0 * e1);
        const double inv = rcp_nr1(n0 + n1 + n2);
        w0 = n0 * inv; w1 = n1 * inv; w2 = n2 * inv;
    }
}

// shock sensor of the WENO filter (weno.py:443-459, normalised):
// sum_r |omega_r/d_r - 1| / (|1/min(d) - 1| + (k - 1)) = (...) / 11 for WENO5
__device__ __forceinline__ double weno_sensor(double d0, double d1, double d2, double w0, double w1, double w2) {
    return (fabs(w0 * (1.0 / d0) - 1.0) + fabs(w1 * (1.0 / d1) - 1.0) + fabs(w2 * (1.0 / d2) - 1.0)) / 11.0;
}

// WENO-Z, order 5 (weno.py:290-345): tau5 = |b0 - b2|,
// alpha_r = d_r + d_r tau^2/((eps+b_r)(eps+b_r)), omega_r = alpha_r / sum(alpha).
// sensor (FP64 only): the filter's shock sensor of these weights.
template <typename T>
__device__ __forceinline__ T wenoz_right(const T* f, double eps, double* sensor = nullptr) {
    double b0, b1, b2;
    if constexpr (std::is_same<T, double>::value) {
        b0 = js_first(f[2], f[3], f[4]);     // stencil (i, i+1, i+2), d = 3/10
        b1 = js_centre(f[1], f[2], f[3]);    // stencil (i-1, i, i+1), d = 6/10
        b2 = js_last(f[0], f[1], f[2]);      // stencil (i-2, i-1, i), d = 1/10
    } else {
        js_diff3(f[0], f[1], f[2], f[3], f[4], b2, b1, b0);
    }
    double w0, w1, w2;
    wenoz_weights<T>(3.0 / 10.0, 6.0 / 10.0, 1.0 / 10.0, b0, b1, b2, eps, w0, w1, w2);
    if (sensor) *sensor = weno_sensor(3.0 / 10.0, 6.0 / 10.0, 1.0 / 10.0, w0, w1, w2);
    const T e0 = eno3<T>(T(1.0 / 3.0), T(5.0 / 6.0), T(-1.0 / 6.0), f[2], f[3], f[4]);
    const T e1 = eno3<T>(T(-1.0 / 6.0), T(5.0 / 6.0), T(1.0 / 3.0), f[1], f[2], f[3]);
    const T e2 = eno3<T>(T(1.0 / 3.0), T(-7.0 / 6.0), T(11.0 / 6.0), f[0], f[1], f[2]);
    return fma(T(w2), e2, fma(T(w1), e1, T(w0) * e0));
}
template <typename T>
__device__ __forceinline__ T wenoz_left(const T* f, double eps, double* sensor = nullptr) {
    double b0, b1, b2;
    if constexpr (std::is_same<T, double>::value) {
        b0 = js_first(f[3], f[4], f[5]);     // (i+1, i+2, i+3), d = 1/10
        b1 = js_centre(f[2], f[3], f[4]);    // (i, i+1, i+2),   d = 6/10
        b2 = js_last(f[1], f[2], f[3]);      // (i-1, i, i+1),   d = 3/10
    } else {
        js_diff3(f[1], f[2], f[3], f[4], f[5], b2, b1, b0);
    }
    double w0, w1, w2;
    wenoz_weights<T>(1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0, b0, b1, b2, eps, w0, w1, w2);
    if (sensor) *sensor = weno_sensor(1.0 / 10.0, 6.0 / 10.0, 3.0 / 10.0, w0, w1, w2);
    const T e0 = eno3<T>(T(11.0 / 6.0), T(-7.0 / 6.0), T(1.0 / 3.0), f[3], f[4], f[5]);
    const T e1 = eno3<T>(T(1.0 / 3.0), T(5.0 / 6.0), T(-1.0 / 6.0), f[2], f[3], f[4]);
    const T e2 = eno3<T>(T(-1.0 / 6.0), T(5.0 / 6.0), T(1.0 / 3.0), f[1], f[2], f[3]);
    return fma(T(w2), e2, fma(T(w1), e1, T(w0) * e0));
}

// WENO3-Z for the WENO filter of order 3 (weno.py:36-136, 290-345 with k = 2:
// tau = |b0 - b1|, p = 2; generated form in
// apps/performance_benchmark/Central4-WENO5/opensbliblock00_kernels.h:930-990),
// FP64, f[s] at offset s-2; beta_r is the squared difference over stencil r and
// the sensor sum_r |omega_r/d_r - 1| / 3 (weno.py:443-459).
__device__ __forceinline__ double wenoz3_right(const double* f, double eps, double* sensor) {
    const double db0 = f[3] - f[2], db1 = f[2] - f[1];          // (i, i+1) d = 2/3; (i-1, i) d = 1/3
    const double b0 = db0 * db0, b1 = db1 * db1;
    const double tau = fabs(b0 - b1), t2 = tau * tau;
    const double a0 = zalpha<double>(2.0 / 3.0, t2, b0, eps), a1 = zalpha<double>(1.0 / 3.0, t2, b1, eps);
    const double inv = 1.0 / (a0 + a1);
    const double w0 = a0 * inv, w1 = a1 * inv;
    *sensor = (fabs(3.0 * w1 - 1.0) + fabs(1.5 * w0 - 1.0)) * (1.0 / 3.0);
    const double e0 = 0.5 * f[2] + 0.5 * f[3];
    const double e1 = fma(-0.5, f[1], 1.5 * f[2]);
    return fma(w1, e1, w0 * e0);
}
__device__ __forceinline__ double wenoz3_left(const double* f, double eps, double* sensor) {
    const double db0 = f[4] - f[3], db1 = f[3] - f[2];          // (i+1, i+2) d = 1/3; (i, i+1) d = 2/3
    const double b0 = db0 * db0, b1 = db1 * db1;
    const double tau = fabs(b0 - b1), t2 = tau * tau;
    const double a0 = zalpha<double>(1.0 / 3.0, t2, b0, eps), a1 = zalpha<double>(2.0 / 3.0, t2, b1, eps);
    const double inv = 1.0 / (a0 + a1);
    const double w0 = a0 * inv, w1 = a1 * inv;
    *sensor = (fabs(3.0 * w0 - 1.0) + fabs(1.5 * w1 - 1.0)) * (1.0 / 3.0);
    const double e0 = fma(-0.5, f[4], 1.5 * f[3]);
    const double e1 = 0.5 * f[2] + 0.5 * f[3];
    return fma(w1, e1, w0 * e0);
}

// (C + tau/beta)**6 printed as six repeated multiplies (opsc.py:222-241)
__device__ __forceinline__ double pow6(double x) { return x * x * x * x * x * x; }

// TENO5 / TENO6 (teno.py:59-217, 419-477).  SIDE +1: f+ (right), -1: f- (left).
template <int ORDER, int SIDE, typename T>
__device__ __forceinline__ T teno(const T* f, double eps, double ct) {
    constexpr int NS = ORDER == 5 ? 3 : 4;
    const double m2 = f[0], m1 = f[1], z0 = f[2], p1 = f[3], p2 = f[4], p3 = f[5];
    double beta[4];
    T eno[4];
    if (SIDE > 0) {
        // stencils [-1,0,1], [0,1,2], [-2,-1,0], [0,1,2,3]
        beta[0] = js_sum(m1 - 2.0 * z0 + p1, m1 - p1);
        beta[1] = js_sum(z0 - 2.0 * p1 + p2, 3.0 * z0 - 4.0 * p1 + p2);
        beta[2] = js_sum(m2 - 2.0 * m1 + z0, m2 - 4.0 * m1 + 3.0 * z0);
        eno[0] = fma(T(2.0 / 6.0), f[3], fma(T(5.0 / 6.0), f[2], T(-1.0 / 6.0) * f[1]));
        eno[1] = fma(T(-1.0 / 6.0), f[4], fma(T(5.0 / 6.0), f[3], T(2.0 / 6.0) * f[2]));
        eno[2] = fma(T(11.0 / 6.0), f[2], fma(T(-7.0 / 6.0), f[1], T(2.0 / 6.0) * f[0]));
        if (NS == 4) {
            const double s1 = -11.0 * z0 + 18.0 * p1 - 9.0 * p2 + 2.0 * p3;
            const double s2 = 2.0 * z0 - 5.0 * p1 + 4.0 * p2 - p3;
            const double s3 = -z0 + 3.0 * p1 - 3.0 * p2 + p3;
            // the right-side beta_3 keeps the 781/720 term unsquared (teno.py:168-169)
            beta[3] = fma((1.0 / 36.0) * s1, s1, fma((13.0 / 12.0) * s2, s2, (781.0 / 720.0) * s3));
            eno[3] = fma(T(1.0 / 12.0), f[5], fma(T(-5.0 / 12.0), f[4], fma(T(13.0 / 12.0), f[3], T(3.0 / 12.0) * f[2])));
        }
    } else {
        // stencils [0,1,2], [-1,0,1], [1,2,3], [-2,-1,0,1]
        beta[0] = js_sum(z0 - 2.0 * p1 + p2, z0 - p2);
        beta[1] = js_sum(m1 - 2.0 * z0 + p1, m1 - 4.0 * z0 + 3.0 * p1);
        beta[2] = js_sum(p1 - 2.0 * p2 + p3, 3.0 * p1 - 4.0 * p2 + p3);
        eno[0] = fma(T(-1.0 / 6.0), f[4], fma(T(5.0 / 6.0), f[3], T(2.0 / 6.0) * f[2]));
        eno[1] = fma(T(2.0 / 6.0), f[3], fma(T(5.0 / 6.0), f[2], T(-1.0 / 6.0) * f[1]));
        eno[2] = fma(T(1.0 / 3.0), f[5], fma(T(-7.0 / 6.0), f[4], T(11.0 / 6.0) * f[3]));
        if (NS == 4) {
            const double s1 = -2.0 * m2 + 9.0 * m1 - 18.0 * z0 + 11.0 * p1;
            const double s2 = -m2 + 4.0 * m1 - 5.0 * z0 + 2.0 * p1;
            const double s3 = -m2 + 3.0 * m1 - 3.0 * z0 + p1;
            beta[3] = fma((1.0 / 36.0) * s1, s1, fma((13.0 / 12.0) * s2, s2, ((781.0 / 720.0) * s3) * s3));
            eno[3] = fma(T(3.0 / 12.0), f[3], fma(T(13.0 / 12.0), f[2], fma(T(-5.0 / 12.0), f[1], T(1.0 / 12.0) * f[0])));
        }
    }
    double dopt[4];
    if (ORDER == 5) { dopt[0] = 11.0 / 20.0; dopt[1] = 4.0 / 10.0; dopt[2] = 1.0 / 20.0; dopt[3] = 0.0; }
    else { dopt[0] = 231.0 / 500.0; dopt[1] = 3.0 / 10.0; dopt[2] = 27.0 / 500.0; dopt[3] = 23.0 / 125.0; }
    const double tau = NS == 3 ? fabs(beta[0] - beta[2])
                               : fabs(beta[3] - (1.0 / 6.0) * (beta[0] + beta[2] - 4.0 * beta[1]));
    double alpha[4], asum = 0.0;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const double inv_beta = wrecip<T>(eps + beta[r]);
        alpha[r] = pow6(1.0 + tau * inv_beta);
        asum += alpha[r];
    }
    const double inv_asum = wrecip<T>(asum);
    double delta[4], wsum = 0.0;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        delta[r] = (alpha[r] * inv_asum < ct) ? 0.0 : 1.0;
        wsum += dopt[r] * delta[r];
    }
    const double inv_wsum = wrecip<T>(wsum);
    T rec = T(0);
#pragma unroll
    for (int r = 0; r < NS; ++r) rec = fma(T(dopt[r] * delta[r] * inv_wsum), eno[r], rec);
    return rec;
}

template <int SCH, typename T>
__device__ __forceinline__ T recon_pair(const T* fp, const T* fm, const Consts& c) {
    if (SCH == SCH_WENOZ) return wenoz_right<T>(fp, c.weno_eps) + wenoz_left<T>(fm, c.weno_eps);
    else if (SCH == SCH_TENO5) return teno<5, 1, T>(fp, c.teno_eps, c.teno_ct) + teno<5, -1, T>(fm, c.teno_eps, c.teno_ct);
    else return teno<6, 1, T>(fp, c.teno_eps, c.teno_ct) + teno<6, -1, T>(fm, c.teno_eps, c.teno_ct);
}

// ---------------------------------------------------------------------------
// Euler eigensystems (euler_eigensystem.py:53-165) for direction DIR, written
// out with the unit normal e_DIR folded in at compile time.
// L[k][m]: left eigenvectors (rows), R[m][k]: right eigenvectors (columns).
// ---------------------------------------------------------------------------
// 1/x for the eigenvector entries: exact in FP64, the hardware reciprocal in FP32
template <typename T>
__device__ __forceinline__ T frecip(T x) {
    if constexpr (std::is_same<T, double>::value) return T(1) / x;
    else return __builtin_amdgcn_rcpf(x);
}

// HALF: L is returned times 1/2, folded into the factors only L uses (1/a^2,
// 1/rho, beta, 1/D and the literal 1).  Scaling by a power of two commutes with
// rounding, so every entry is exactly half the unscaled one (no subnormals
// arise here), and so are the characteristic values L q built from it.
template <int NDIM, int DIR, typename T, bool HALF = false>
__device__ __forceinline__ void eigensystem(T gm1, T rho, const T* u, T a, T (&L)[NDIM + 2][NDIM + 2],
                                            T (&R)[NDIM + 2][NDIM + 2]) {
    constexpr int NV = NDIM + 2;
    constexpr T HL = HALF ? T(0.5) : T(1);
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) { L[i][j] = T(0); R[i][j] = T(0); }
    if constexpr (NDIM == 1) {
        const T u0 = u[0];
        const T inva = frecip<T>(a);
        const T H = (a * a) * frecip<T>(gm1) + (u0 * u0) * T(0.5);
        const T D = T(2) * H - u0 * u0;
        const T invD = frecip<T>(D) * HL;
        R[0][0] = 1; R[0][1] = 1; R[0][2] = 1;
        R[1][0] = u0 - a; R[1][1] = u0; R[1][2] = u0 + a;
        R[2][0] = H - u0 * a; R[2][1] = (u0 * u0) * T(0.5); R[2][2] = H + u0 * a;
        L[0][0] = u0 * (T(2) * H + a * u0 - u0 * u0) * T(0.5) * inva * invD;
        L[0][1] = (-H - a * u0 + (u0 * u0) * T(0.5)) * inva * invD;
        L[0][2] = invD;
        L[1][0] = T(2) * (H - u0 * u0) * invD;
        L[1][1] = T(2) * u0 * invD;
        L[1][2] = T(-2) * invD;
        L[2][0] = u0 * (-T(2) * H + a * u0 + u0 * u0) * T(0.5) * inva * invD;
        L[2][1] = (H - a * u0 - (u0 * u0) * T(0.5)) * inva * invD;
        L[2][2] = invD;
    } else {
    constexpr T k0 = DIR == 0 ? 1 : 0, k1 = DIR == 1 ? 1 : 0, k2 = DIR == 2 ? 1 : 0;
    T q2 = T(0);
#pragma unroll
    for (int d = 0; d < NDIM; ++d) q2 += u[d] * u[d];
    const T phi_sq = gm1 * (q2 * T(0.5));
    const T theta = u[DIR];
    const T inva = frecip<T>(a), invrho0 = frecip<T>(rho);
    const T inva2 = inva * inva * HL;
    const T alpha = rho * T(INV_SQRT2) * inva;
    const T bta = T(INV_SQRT2) * invrho0 * inva * HL;
    const T invrho = invrho0 * HL;
    const T igm1 = frecip<T>(gm1);
    const T a2 = a * a;
    if constexpr (NDIM == 2) {
        const T u0 = u[0], u1 = u[1];
        R[0][0] = 1; R[0][2] = alpha; R[0][3] = alpha;
        R[1][0] = u0; R[1][1] = k1 * rho; R[1][2] = alpha * (u0 + k0 * a); R[1][3] = alpha * (u0 - k0 * a);
        R[2][0] = u1; R[2][1] = -k0 * rho; R[2][2] = alpha * (u1 + k1 * a); R[2][3] = alpha * (u1 - k1 * a);
        R[3][0] = phi_sq * igm1; R[3][1] = rho * (k1 * u0 - k0 * u1);
        R[3][2] = alpha * ((phi_sq + a2) * igm1 + a * theta);
        R[3][3] = alpha * ((phi_sq + a2) * igm1 - a * theta);
        L[0][0] = HL - phi_sq * inva2; L[0][1] = gm1 * u0 * inva2; L[0][2] = gm1 * u1 * inva2; L[0][3] = -gm1 * inva2;
        L[1][0] = -(k1 * u0 - k0 * u1) * invrho; L[1][1] = k1 * invrho; L[1][2] = -k0 * invrho;
        L[2][0] = bta * (phi_sq - a * theta); L[2][1] = bta * (k0 * a - gm1 * u0);
        L[2][2] = bta * (k1 * a - gm1 * u1); L[2][3] = bta * gm1;
        L[3][0] = bta * (phi_sq + a * theta); L[3][1] = -bta * (k0 * a + gm1 * u0);
        L[3][2] = -bta * (k1 * a + gm1 * u1); L[3][3] = bta * gm1;
    } else {
    const T u0 = u[0], u1 = u[1], u2 = u[2];
    R[0][0] = k0; R[0][1] = k1; R[0][2] = k2; R[0][3] = alpha; R[0][4] = alpha;
    R[1][0] = k0 * u0; R[1][1] = k1 * u0 - k2 * rho; R[1][2] = k2 * u0 + k1 * rho;
    R[1][3] = alpha * (u0 + k0 * a); R[1][4] = alpha * (u0 - k0 * a);
    R[2][0] = k0 * u1 + k2 * rho; R[2][1] = k1 * u1; R[2][2] = k2 * u1 - k0 * rho;
    R[2][3] = alpha * (u1 + k1 * a); R[2][4] = alpha * (u1 - k1 * a);
    R[3][0] = k0 * u2 - k1 * rho; R[3][1] = k1 * u2 + k0 * rho; R[3][2] = k2 * u2;
    R[3][3] = alpha * (u2 + k2 * a); R[3][4] = alpha * (u2 - k2 * a);
    R[4][0] = k0 * phi_sq * igm1 + rho * (k2 * u1 - k1 * u2);
    R[4][1] = k1 * phi_sq * igm1 + rho * (k0 * u2 - k2 * u0);
    R[4][2] = k2 * phi_sq * igm1 + rho * (k1 * u0 - k0 * u1);
    R[4][3] = alpha * ((phi_sq + a2) * igm1 + theta * a);
    R[4][4] = alpha * ((phi_sq + a2) * igm1 - theta * a);
    const T c1 = HL - phi_sq * inva2;
    L[0][0] = k0 * c1 - (k2 * u1 - k1 * u2) * invrho;
    L[0][1] = k0 * gm1 * u0 * inva2;
    L[0][2] = k0 * gm1 * u1 * inva2 + k2 * invrho;
    L[0][3] = k0 * gm1 * u2 * inva2 - k1 * invrho;
    L[0][4] = -k0 * gm1 * inva2;
    L[1][0] = k1 * c1 - (k0 * u2 - k2 * u0) * invrho;
    L[1][1] = k1 * gm1 * u0 * inva2 - k2 * invrho;
    L[1][2] = k1 * gm1 * u1 * inva2;
    L[1][3] = k1 * gm1 * u2 * inva2 + k0 * invrho;
    L[1][4] = -k1 * gm1 * inva2;
    L[2][0] = k2 * c1 - (k1 * u0 - k0 * u1) * invrho;
    L[2][1] = k2 * gm1 * u0 * inva2 + k1 * invrho;
    L[2][2] = k2 * gm1 * u1 * inva2 - k0 * invrho;
    L[2][3] = k2 * gm1 * u2 * inva2;
    L[2][4] = -k2 * gm1 * inva2;
    L[3][0] = bta * (phi_sq - theta * a);
    L[3][1] = -bta * (gm1 * u0 - k0 * a);
    L[3][2] = -bta * (gm1 * u1 - k1 * a);
    L[3][3] = -bta * (gm1 * u2 - k2 * a);
    L[3][4] = bta * gm1;
    L[4][0] = bta * (phi_sq + theta * a);
    L[4][1] = -bta * (gm1 * u0 + k0 * a);
    L[4][2] = -bta * (gm1 * u1 + k1 * a);
    L[4][3] = -bta * (gm1 * u2 + k2 * a);
    L[4][4] = bta * gm1;
    }
    }
}

// ---------------------------------------------------------------------------
// Primitive variables and the physical flux along DIR of one point
// (constituent relations + flux vector, shock_capturing.py:786-836).  FAST
// (mixed precision) uses a refined reciprocal of rho instead of divisions.
// ---------------------------------------------------------------------------
template <int NDIM, int DIR, bool FAST>
__device__ __forceinline__ void point_prims(const double (&v)[NDIM + 2], const Consts& c, double (&uu)[NDIM], double& p,
                                            double (&f)[NDIM + 2], double& a) {
    constexpr int NV = NDIM + 2;
    const double r = v[0];
    double ke = 0.0;
    const double inv_r = FAST ? rcp_nr1(r) : 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        uu[d] = FAST ? v[1 + d] * inv_r : v[1 + d] / r;
        ke -= 0.5 * (uu[d] * uu[d]) * r;
    }
    p = c.gamma_m1 * (ke + v[NV - 1]);
    const double u_dir = uu[DIR];
    f[0] = v[1 + DIR];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) f[1 + d] = v[1 + d] * u_dir + (d == DIR ? p : 0.0);
    f[NV - 1] = (p + v[NV - 1]) * u_dir;
    // mixed precision: the sound speed only enters the FP32 wave speeds
    a = FAST ? (double)__builtin_amdgcn_sqrtf((float)(c.gama * p * inv_r)) : sqrt(c.gama * p / r);
}

// ---------------------------------------------------------------------------
// Characteristic LLF flux at interface i+1/2 along DIR from the six stencil
// points i-2..i+3 (shock_capturing.py:786-891, 957-1028).
//   pt(s, v, f, uu, p, a): FP64 state, flux and primitives of stencil point s.
// T = double: the reference's arithmetic.  T = float (mixed precision): the
// characteristic work runs on deviations from the state and flux at point
// i (s = 2), formed in FP64 and then rounded to FP32, so FP32 holds only
// the local variation.  WENO-Z and TENO are translation invariant
// (rec(f + c) = rec(f) + c: smoothness indicators see differences, the
// candidate weights sum to one), and f+ + f- = L F, so the interface flux
// is F_i + R rec(deviations) with the FP64 F_i added back at the end.
// ---------------------------------------------------------------------------
template <int NDIM, int DIR, int SCH, typename T, class PT, int FILT = 0>
__device__ __forceinline__ void face_core(PT&& pt, const Consts& c, double (&out)[NDIM + 2]) {
    constexpr int NV = NDIM + 2;
    constexpr bool DEV = !std::is_same<T, double>::value;
    T qs[6][NV];
    T F[6][NV];
    double base_q[NV], base_F[NV];
    double rho23[2], u23[2][NDIM], p23[2], E23[2], a23[2];
    auto point = [&](const int s) {
        double v[NV], f[NV], uu[NDIM], p, a;
        pt(s, v, f, uu, p, a);
        if constexpr (DEV) {
            if (s == 2) {
#pragma unroll
                for (int m = 0; m < NV; ++m) { base_q[m] = v[m]; base_F[m] = f[m]; }
            }
#pragma unroll
            for (int m = 0; m < NV; ++m) { qs[s][m] = T(v[m] - base_q[m]); F[s][m] = T(f[m] - base_F[m]); }
        } else {
#pragma unroll
            for (int m = 0; m < NV; ++m) { qs[s][m] = T(v[m]); F[s][m] = T(f[m]); }
        }
        if (s == 2 || s == 3) {
            const int w = s - 2;
            rho23[w] = v[0]; p23[w] = p; E23[w] = v[NV - 1];
#pragma unroll
            for (int d = 0; d < NDIM; ++d) u23[w][d] = uu[d];
            a23[w] = a;
        }
    };
    point(2);              // the base point first (mixed precision subtracts it)
    if (FILT != 3) point(0);
    point(1);
    point(3);
    point(4);
    if (FILT != 3) point(5);   // the order-3 filter reads offsets -1..2 only
    // averaged state (averaging.py:31-124)
    T rho_av, u_av[NDIM], a_av;
    if (DEV && c.avg_roe) {
        // mixed precision: the averaged state only builds the FP32 eigenvectors,
        // so it is formed in FP32 (hardware square roots and reciprocals, 1 ulp
        // on the positive normal densities and enthalpies here; sqrt(rho_L rho_R)
        // as sqrt(rho_L) sqrt(rho_R))
        const float sl = __builtin_amdgcn_sqrtf((float)rho23[0]), sr = __builtin_amdgcn_sqrtf((float)rho23[1]);
        const float inv = __builtin_amdgcn_rcpf(sl + sr);
        float ke = 0.f;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
            const float ua = (sl * (float)u23[0][d] + sr * (float)u23[1][d]) * inv;
            ke += ua * ua;
            u_av[d] = T(ua);
        }
        const float Hr = ((float)(p23[0] + E23[0]) * __builtin_amdgcn_rcpf(sl) +
                          (float)(p23[1] + E23[1]) * __builtin_amdgcn_rcpf(sr)) * inv;
        rho_av = T(sl * sr);
        a_av = T(__builtin_amdgcn_sqrtf((float)c.gamma_m1 * (Hr - 0.5f * ke)));
    } else if (c.avg_roe) {
        const double sl = sqrt(rho23[0]), sr = sqrt(rho23[1]);
        const double inv = wrecip<T>(sl + sr);
        double ke = 0.0;
        double ua[NDIM];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) { ua[d] = (sl * u23[0][d] + sr * u23[1][d]) * inv; ke += ua[d] * ua[d]; }
        const double Hr = (wdiv<T>(p23[0] + E23[0], sl) + wdiv<T>(p23[1] + E23[1], sr)) * inv;
        rho_av = T(sqrt(rho23[0] * rho23[1]));
#pragma unroll
        for (int d = 0; d < NDIM; ++d) u_av[d] = T(ua[d]);
        a_av = T(sqrt(c.gamma_m1 * (Hr - 0.5 * ke)));
    } else {
        rho_av = T(0.5 * (rho23[0] + rho23[1]));
#pragma unroll
        for (int d = 0; d < NDIM; ++d) u_av[d] = T(0.5 * (u23[0][d] + u23[1][d]));
        a_av = T(0.5 * (a23[0] + a23[1]));
    }
    // L times 1/2 unless the filter needs the unscaled characteristic fluxes:
    // f+- = (L F +- lambda L q) / 2 then need no multiplication
    constexpr bool HALF = FILT == 0;
    T L[NV][NV], R[NV][NV];
    eigensystem<NDIM, DIR, T, HALF>(T(c.gamma_m1), rho_av, u_av, a_av, L, R);
    // wave speeds: global (GLF, the stage's grid maxima in the eigenvalue order
    // of the direction, shock_capturing.py:838-857) or local Lax-Friedrichs over
    // offsets {0,1} (shock_capturing.py:1000-1028)
    T lam[NV];
    if (c.glf) {
        const double* gl = c.glf + 3 * DIR;
        if (NDIM == 1) {
            lam[0] = T(gl[2]);
            lam[1] = T(gl[0]);
            lam[2] = T(gl[1]);
        } else {
#pragma unroll
            for (int d = 0; d < NDIM; ++d) lam[d] = T(gl[0]);
            lam[NDIM] = T(gl[1]);
            lam[NDIM + 1] = T(gl[2]);
        }
    } else {
        const double ua = u23[0][DIR], ub = u23[1][DIR];
        if (NDIM == 1) {
            lam[0] = T(c.sfc * fmax(fabs(ua - a23[0]), fabs(ub - a23[1])));
            lam[1] = T(c.sfc * fmax(fabs(ua), fabs(ub)));
            lam[2] = T(c.sfc * fmax(fabs(ua + a23[0]), fabs(ub + a23[1])));
        } else {
            const T l0 = T(c.sfc * fmax(fabs(ua), fabs(ub)));
#pragma unroll
            for (int d = 0; d < NDIM; ++d) lam[d] = l0;
            lam[NDIM] = T(c.sfc * fmax(fabs(ua + a23[0]), fabs(ub + a23[1])));
            lam[NDIM + 1] = T(c.sfc * fmax(fabs(ua - a23[0]), fabs(ub - a23[1])));
        }
    }
    T acc[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) acc[m] = T(0);
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        T fp[6], fm[6], cfv[6];
        constexpr int S0 = FILT == 3 ? 1 : 0, S1 = FILT == 3 ? 5 : 6;
#pragma unroll
        for (int s = S0; s < S1; ++s) {
            T cs = T(0), cf = T(0);
#pragma unroll
            for (int m = 0; m < NV; ++m) {
                // structural zeros skipped (mixed precision, and FP64 in 3-D): their
                // products are zeros (of either sign) for every finite state, so
                // only the sign of a zero sum can differ.  The 2-D FP64 faces keep
                // them: there the skip raised the register count past 3 waves/SIMD
                // and slowed the fused face-residual kernel 7 % (KH 1024^2).
                if ((DEV || NDIM == 3) && l_zero<NDIM, DIR>(k, m)) continue;
                cs = fma(L[k][m], qs[s][m], cs);
                cf = fma(L[k][m], F[s][m], cf);
            }
            if constexpr (HALF) {
                fp[s] = fma(lam[k], cs, cf);
                fm[s] = fma(-lam[k], cs, cf);
            } else {
                fp[s] = T(0.5) * fma(lam[k], cs, cf);
                fm[s] = T(0.5) * fma(-lam[k], cs, cf);
            }
            cfv[s] = cf;
        }
        T rec;
        if constexpr (FILT == 5) {
            // WENO filter (WENO_filter.py:406-470, shock_capturing.py:554-581): the dissipative part
            // of WENO-Z -- its interface value less the 6th-order central interpolation of the
            // characteristic fluxes -- times the sensor max(right, left); FP64
            double s_r = 0.0, s_l = 0.0;
            const double r0 = wenoz_right<double>(fp, c.weno_eps, &s_r) + wenoz_left<double>(fm, c.weno_eps, &s_l);
            const double central = (((((cfv[0] - 8.0 * cfv[1]) + 37.0 * cfv[2]) + 37.0 * cfv[3]) - 8.0 * cfv[4]) + cfv[5]) *
                                   (1.0 / 60.0);
            rec = fmax(s_r, s_l) * (r0 - central);
        } else if constexpr (FILT == 3) {
            // order 3: WENO3-Z less the 4th-order central interpolation (-1, 7, 7, -1)/12
            double s_r = 0.0, s_l = 0.0;
            const double r0 = wenoz3_right(fp, c.weno_eps, &s_r) + wenoz3_left(fm, c.weno_eps, &s_l);
            const double central = (7.0 * (cfv[2] + cfv[3]) - (cfv[1] + cfv[4])) * (1.0 / 12.0);
            rec = fmax(s_r, s_l) * (r0 - central);
        } else {
            rec = recon_pair<SCH, T>(fp, fm, c);
        }
#pragma unroll
        for (int m = 0; m < NV; ++m)
            if (NDIM != 3 || !r_zero<NDIM, DIR>(m, k)) acc[m] = fma(R[m][k], rec, acc[m]);
    }
    if (DEV) {
#pragma unroll
        for (int m = 0; m < NV; ++m) out[m] = base_F[m] + double(acc[m]);
    } else {
#pragma unroll
        for (int m = 0; m < NV; ++m) out[m] = double(acc[m]);
    }
}

// sin(x) for x in [0, pi/2] by its Taylor series through x^21 (truncation
// below 1.2e-18 there), Horner in x^2 with fused multiply-adds: the HLLC-LM
// Mach switch evaluated identically on every platform.
__device__ __forceinline__ double hllc_sin(double x) {
    const double x2 = x * x;
    double s = -1.0 / 51090942171709440000.0;                          // -1/21!
    s = fma(s, x2, 1.0 / 121645100408832000.0);                         //  1/19!
    s = fma(s, x2, -1.0 / 355687428096000.0);                           // -1/17!
    s = fma(s, x2, 1.0 / 1307674368000.0);                              //  1/15!
    s = fma(s, x2, -1.0 / 6227020800.0);                                // -1/13!
    s = fma(s, x2, 1.0 / 39916800.0);                                   //  1/11!
    s = fma(s, x2, -1.0 / 362880.0);                                    // -1/9!
    s = fma(s, x2, 1.0 / 5040.0);                                       //  1/7!
    s = fma(s, x2, -1.0 / 120.0);                                       // -1/5!
    s = fma(s, x2, 1.0 / 6.0);                                          //  1/3!
    return fma(-(x * x2), s, x);                                        // x - x^3 (1/3! - x^2/5! + ...)
}

// HLLC / HLLC-LM interface flux (shock_capturing.py:1031-1256) at x0 + 1/2
// along DIR: WENO-Z of the characteristic state CS = L q at the six points
// (the right reconstruction is the face's left state, the left one its right
// state), back-projection, then the HLLC Riemann flux of the two states with
// sL = min(ubar - abar, uL - aL), sR = max(ubar + abar, uR + aR), the contact
// speed s*, the star states (sL in both star energies, as the reference
// writes them) and, for LM, sL and sR scaled by
// phi = sin(min(1, max(|uL/aL|, |uR/aR|)/0.1) pi/2).  FP64, the oracle's
// operation order.
template <int NDIM, int DIR, bool LM>
__device__ __forceinline__ void face_hllc(const double* __restrict__ q, long Ptot, long st, long x0, const Consts& c,
                                          double (&out)[NDIM + 2]) {
    constexpr int NV = NDIM + 2;
    double qs[6][NV];
    double rho23[2], u23[2][NDIM], p23[2], E23[2], a23[2];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const long xs = x0 + (long)(s - 2) * st;
#pragma unroll
        for (int m = 0; m < NV; ++m) qs[s][m] = q[(long)m * Ptot + xs];
        if (s == 2 || s == 3) {
            double f[NV], uu[NDIM], p, a;
            point_prims<NDIM, DIR, false>(qs[s], c, uu, p, f, a);
            const int w = s - 2;
            rho23[w] = qs[s][0]; p23[w] = p; E23[w] = qs[s][NV - 1]; a23[w] = a;
#pragma unroll
            for (int d = 0; d < NDIM; ++d) u23[w][d] = uu[d];
        }
    }
    double rho_av, u_av[NDIM], a_av;
    if (c.avg_roe) {
        const double sl = sqrt(rho23[0]), sr = sqrt(rho23[1]);
        const double inv = 1.0 / (sl + sr);
        double ke = 0.0;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) { u_av[d] = (sl * u23[0][d] + sr * u23[1][d]) * inv; ke += u_av[d] * u_av[d]; }
        const double Hr = ((p23[0] + E23[0]) / sl + (p23[1] + E23[1]) / sr) * inv;
        rho_av = sqrt(rho23[0] * rho23[1]);
        a_av = sqrt(c.gamma_m1 * (Hr - 0.5 * ke));
    } else {
        rho_av = 0.5 * (rho23[0] + rho23[1]);
#pragma unroll
        for (int d = 0; d < NDIM; ++d) u_av[d] = 0.5 * (u23[0][d] + u23[1][d]);
        a_av = 0.5 * (a23[0] + a23[1]);
    }
    double L[NV][NV], R[NV][NV];
    eigensystem<NDIM, DIR, double>(c.gamma_m1, rho_av, u_av, a_av, L, R);
    double qL[NV], qR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { qL[v] = 0.0; qR[v] = 0.0; }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double cs[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            double a = 0.0;
#pragma unroll
            for (int m = 0; m < NV; ++m) a = fma(L[k][m], qs[s][m], a);
            cs[s] = a;
        }
        const double wr = wenoz_right<double>(cs, c.weno_eps), wl = wenoz_left<double>(cs, c.weno_eps);
#pragma unroll
        for (int v = 0; v < NV; ++v) { qL[v] = fma(R[v][k], wr, qL[v]); qR[v] = fma(R[v][k], wl, qR[v]); }
    }
    const double gama = c.gama;
    const double irL = 1.0 / qL[0], irR = 1.0 / qR[0];
    double uL[NDIM], uR[NDIM], keL = 0.0, keR = 0.0;
#pragma unroll
    for (int m = 0; m < NDIM; ++m) {
        uL[m] = qL[1 + m] * irL;
        uR[m] = qR[1 + m] * irR;
        keL = fma(uL[m], uL[m], keL);
        keR = fma(uR[m], uR[m], keR);
    }
    const double pL = (gama - 1.0) * (qL[NV - 1] - 0.5 * qL[0] * keL);
    const double pR = (gama - 1.0) * (qR[NV - 1] - 0.5 * qR[0] * keR);
    const double aL = sqrt(gama * pL * irL), aR = sqrt(gama * pR * irR);
    const double ul = uL[DIR], ur = uR[DIR];
    double sL = fmin(u_av[DIR] - a_av, ul - aL);
    double sR = fmax(u_av[DIR] + a_av, ur + aR);
    const double ss = ((pR - pL) + (qL[0] * ul * (sL - ul) - qR[0] * ur * (sR - ur))) / (qL[0] * (sL - ul) - qR[0] * (sR - ur));
    if (LM) {
        const double ml = fmax(fabs(ul / aL), fabs(ur / aR));
        const double phi = hllc_sin(fmin(1.0, ml / 0.1) * (0.5 * 3.14159265358979323846));
        sL = sL * phi;
        sR = sR * phi;
    }
    double FL[NV], FR[NV], SL[NV], SR[NV];
    FL[0] = qL[1 + DIR];
    FR[0] = qR[1 + DIR];
#pragma unroll
    for (int m = 0; m < NDIM; ++m) {
        FL[1 + m] = qL[1 + m] * ul + (m == DIR ? pL : 0.0);
        FR[1 + m] = qR[1 + m] * ur + (m == DIR ? pR : 0.0);
    }
    FL[NV - 1] = (pL + qL[NV - 1]) * ul;
    FR[NV - 1] = (pR + qR[NV - 1]) * ur;
    const double fl = qL[0] * ((sL - ul) / (sL - ss)), fr = qR[0] * ((sR - ur) / (sR - ss));
    SL[0] = fl;
    SR[0] = fr;
#pragma unroll
    for (int m = 0; m < NDIM; ++m) {
        SL[1 + m] = fl * (m == DIR ? ss : uL[m]);
        SR[1 + m] = fr * (m == DIR ? ss : uR[m]);
    }
    SL[NV - 1] = fl * fma(ss - ul, ss + pL / (qL[0] * (sL - ul)), qL[NV - 1] * irL);
    SR[NV - 1] = fr * fma(ss - ur, ss + pR / (qR[0] * (sL - ur)), qR[NV - 1] * irR);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double f;
        if (LM) {
            if (sL >= 0.0) f = FL[v];
            else if (sR <= 0.0) f = FR[v];
            else f = 0.5 * (FL[v] + FR[v]) + 0.5 * ((sL * (SL[v] - qL[v]) + fabs(ss) * (SL[v] - SR[v])) + sR * (SR[v] - qR[v]));
        } else {
            if (sL >= 0.0) f = FL[v];
            else if (sL <= 0.0 && 0.0 <= ss) f = fma(sL, SL[v] - qL[v], FL[v]);
            else if (ss <= 0.0 && 0.0 <= sR) f = fma(sR, SR[v] - qR[v], FR[v]);
            else if (sR <= 0.0) f = FR[v];
            else f = 0.0;
        }
        out[v] = f;
    }
}

// face flux at interface x0 + 1/2 along DIR reading the six points from the
// state (q: FP64 state, `st` the stride along DIR)
template <int NDIM, int DIR, int SCH, typename T, int FILT = 0>
__device__ __forceinline__ void face_flux(const double* __restrict__ q, long Ptot, long st, long x0,
                                          const Consts& c, double (&out)[NDIM + 2]) {
    constexpr int NV = NDIM + 2;
    constexpr bool FAST = !std::is_same<T, double>::value;
    static_assert(!FILT || !FAST, "the WENO filter is FP64");
    // FP64: the stencil points' states are all loaded, behind a scheduling
    // fence, before any is used (the 3-D TENO5 faces measured 1.5 % faster;
    // the mixed-precision faces, which the scheduler orders well, slower)
    double qs[6][NV];
    if (!FAST) {
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            if (FILT == 3 && (s == 0 || s == 5)) continue;      // the order-3 filter reads offsets -1..2
            const long xs = x0 + (long)(s - 2) * st;
#pragma unroll
            for (int m = 0; m < NV; ++m) qs[s][m] = q[(long)m * Ptot + xs];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    auto pt = 
        [&](int s, double (&v)[NV], double (&f)[NV], double (&uu)[NDIM], double& p, double& a) {
            const long xs = x0 + (long)(s - 2) * st;
#pragma unroll
            for (int m = 0; m < NV; ++m) v[m] = FAST ? q[(long)m * Ptot + xs] : qs[s][m];
            point_prims<NDIM, DIR, FAST>(v, c, uu, p, f, a);
        };
    face_core<NDIM, DIR, SCH, T, decltype(pt)&, FILT>(pt, c, out);
}

// Every face periodic (n >= 2h in every direction): store the values qv of
// interior point (i, j, k) at its images in the halo of q, i.e. the halo
// points whose index wraps to it (i + n where i < h, i - n where i >= n - h,
// per direction, and their combinations at edges and corners). Each halo
// point has exactly one such source, so the stores leave what the halo fill
// would copy there.
template <int NV>
__device__ __forceinline__ void store_halo_images(double* __restrict__ q, const Geom& g, int i, int j, int k,
                                                  const double (&qv)[NV]) {
    const int h = g.halo;
    const int oi = i < h ? g.n[0] : (i >= g.n[0] - h ? -g.n[0] : 0);
    const int oj = j < h ? g.n[1] : (j >= g.n[1] - h ? -g.n[1] : 0);
    const int ok = k < h ? g.n[2] : (k >= g.n[2] - h ? -g.n[2] : 0);
    if (!(oi | oj | ok)) return;
    for (int m = 1; m < 8; ++m) {
        if (((m & 1) && !oi) || ((m & 2) && !oj) || ((m & 4) && !ok)) continue;
        const long xh = gidx(g, i + ((m & 1) ? oi : 0), j + ((m & 2) ? oj : 0), k + ((m & 4) ? ok : 0));
#pragma unroll
        for (int v = 0; v < NV; ++v) q[(long)v * g.Ptot + xh] = qv[v];
    }
}

}  // namespace osb
