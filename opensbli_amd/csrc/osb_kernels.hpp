// osb_kernels.hpp -- gfx950 kernels of the OpenSBLI explicit-solver hot path.
#pragma once

#include "osb_common.hpp"

namespace osb {

constexpr int BLOCK = 256;

__device__ __forceinline__ long gtid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }
// XCD-aware thread index (the WENO filter's face kernels, where it measured
// 9 % faster; the scheme's face kernels did not gain): workgroups are dealt round-robin
// to the 8 XCDs (workgroup w on XCD w % 8), so XCD x is given the contiguous
// range of blocks [x*B/8, (x+1)*B/8) and the rows a y or z stencil reads share
// its L2 (a permutation of the blocks when B % 8 == 0, the identity otherwise)
__device__ __forceinline__ long xcd_gtid() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned bb = (nb & 7u) == 0 ? (b & 7u) * (nb >> 3) + (b >> 3) : b;
    return (long)bb * blockDim.x + threadIdx.x;
}

// decompose t over extents e[0..2] (x fastest) with lower bounds lo[]
__device__ __forceinline__ void unflatten(long t, const int (&e)[3], const int (&lo)[3], int& i, int& j, int& k) {
    const int e0 = e[0], e1 = e[1];
    i = (int)(t % e0) + lo[0];
    const long r = t / e0;
    j = (int)(r % e1) + lo[1];
    k = (int)(r / e1) + lo[2];
}
// (unflatten_dv, osb_order.hpp: the same split from the host's inverses)

// ---------------------------------------------------------------------------
// Interface fluxes along DIR for faces i+1/2, i in [-1, n) along DIR and the
// interior across (TENO6/opensbli.cpp:263-316 ranges).
// ---------------------------------------------------------------------------
template <int NDIM, int DIR, int SCH, typename T>
__device__ __forceinline__ void face_flux_item(long t, const double* __restrict__ q, double* __restrict__ F, const Geom& g,
                                               const Consts& c) {
    constexpr int NV = NDIM + 2;
    int e[3] = {g.n[0], g.n[1], g.n[2]};
    e[DIR] += 1;
    int lo[3] = {0, 0, 0};
    lo[DIR] = -1;
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[1 + DIR], lo, i, j, k);
    const long x0 = gidx(g, i, j, k);
    double out[NV];
    face_flux<NDIM, DIR, SCH, T>(q, g.Ptot, g.s[DIR], x0, c, out);
#pragma unroll
    for (int v = 0; v < NV; ++v) F[(long)v * g.Ptot + x0] = out[v];
}
template <int NDIM, int DIR>
__host__ __device__ inline long face_flux_work(const Geom& g) {
    long w = 1;
    for (int d = 0; d < 3; ++d) w *= (long)g.n[d] + (d == DIR ? 1 : 0);
    return w;
}

template <int NDIM, int DIR, int SCH, typename T>
__global__ void __launch_bounds__(BLOCK) k_face_flux(const double* __restrict__ q, double* __restrict__ F, Geom g, Consts c) {
    face_flux_item<NDIM, DIR, SCH, T>(gtid(), q, F, g, c);
}

// HLLC / HLLC-LM interface fluxes along DIR over faces [-1, n) (HLLCWeno,
// shock_capturing.py:1031-1256)
template <int NDIM, int DIR, bool LM>
__global__ void __launch_bounds__(BLOCK) k_face_hllc(const double* __restrict__ q, double* __restrict__ F, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    int e[3] = {g.n[0], g.n[1], g.n[2]};
    e[DIR] += 1;
    int lo[3] = {0, 0, 0};
    lo[DIR] = -1;
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[1 + DIR], lo, i, j, k);
    const long x0 = gidx(g, i, j, k);
    double out[NV];
    face_hllc<NDIM, DIR, LM>(q, g.Ptot, g.s[DIR], x0, c, out);
#pragma unroll
    for (int v = 0; v < NV; ++v) F[(long)v * g.Ptot + x0] = out[v];
}

// WENO filter fluxes along DIR (WENO_filter.py:406-470): faces [-1, n) as the
// reconstruction kernels of the scheme; c carries RoeAverage and LLF.  ORDER 5
// or 3; HYB: the Ducros hybrid selection (WENO_filter.py:561-584) -- a face is
// reconstructed only where max kappa over offsets -3..2 along DIR exceeds
// ducros_check, else its fluxes are zero.
template <int NDIM, int DIR, int ORDER, bool HYB>
__global__ void __launch_bounds__(BLOCK) k_filter_face(const double* __restrict__ q, const double* __restrict__ kappa,
                                                       double* __restrict__ F, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    int e[3] = {g.n[0], g.n[1], g.n[2]};
    e[DIR] += 1;
    int lo[3] = {0, 0, 0};
    lo[DIR] = -1;
    const long t = xcd_gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[1 + DIR], lo, i, j, k);
    const long x0 = gidx(g, i, j, k);
    double out[NV];
    bool active = true;
    if (HYB) {
        const long st = g.s[DIR];
        double km = kappa[x0];
#pragma unroll
        for (int o = -3; o <= 2; ++o) km = fmax(km, kappa[x0 + o * st]);
        active = km > c.ducros_check;
    }
    if (active) {
        face_flux<NDIM, DIR, SCH_WENOZ, double, ORDER>(q, g.Ptot, g.s[DIR], x0, c, out);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) out[v] = 0.0;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) F[(long)v * g.Ptot + x0] = out[v];
}

// u_m = rhou_m / rho over [-2, n+2) in the active directions, for the sensor
template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_velocity(const double* __restrict__ q, double* __restrict__ u, Geom g) {
    int e[3], lo[3];
    for (int d = 0; d < 3; ++d) {
        const bool act = d < NDIM;
        e[d] = act ? g.n[d] + 4 : 1;
        lo[d] = act ? -2 : 0;
    }
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten(t, e, lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const double r = q[x];
#pragma unroll
    for (int m = 0; m < NDIM; ++m) u[(long)m * g.Ptot + x] = q[(long)(1 + m) * g.Ptot + x] / r;
}

// Modified Ducros sensor (shock_sensors.py:16-59; generated
// Central4-WENO5/opensbliblock00_kernels.h:572-617) over [0, n):
// kappa = div^2 (1/2 - 1/2 tanh(2.5 + 500 div / sqrt(sum Delta^2))) / (div^2 + |curl u|^2 + 1e-40)
// with central derivatives of u (k_velocity); the halos of kappa stay zero.
template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_ducros(const double* __restrict__ u, double* __restrict__ kappa, Geom g, Consts c,
                                                  double sd) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = xcd_gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const int ii[3] = {i, j, k};
    const long x = gidx(g, i, j, k);
    const long P = g.Ptot;
    double G[NDIM][NDIM];
#pragma unroll
    for (int m = 0; m < NDIM; ++m)
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
            const long st = g.s[d];
            const double* um = u + (long)m * P;
            G[m][d] = d1c([&](int s) { return um[x + s * st]; }, c.invd[d], ii[d], g.n[d], c.carp[d]);
        }
    double div = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) div += G[d][d];
    double w2;
    if constexpr (NDIM == 2) {
        const double wz = G[1][0] - G[0][1];
        w2 = wz * wz;
    } else {
        const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
        w2 = (wx * wx + wy * wy) + wz * wz;
    }
    const double th = 0.5 - 0.5 * tanh(2.5 + (500.0 * div) / sd);
    kappa[x] = ((div * div) * th) / ((div * div + w2) + 1.0e-40);
}

// filter application (WENO_filter.py:479-536): q += dt * (-sum_d (F_d(i) - F_d(i-1)) / Delta_d), in
// place; HYB: only where max kappa over the point and its -1, +1, +2 neighbours reaches ducros_select
template <int NDIM, bool HYB>
__global__ void __launch_bounds__(BLOCK) k_filter_update(const double* __restrict__ F, const double* __restrict__ kappa,
                                                         double* __restrict__ q, const double* __restrict__ scal, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = xcd_gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    if (HYB) {
        double km = kappa[x];
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
            const long st = g.s[d];
            km = fmax(km, fmax(kappa[x - st], fmax(kappa[x + st], kappa[x + 2 * st])));
        }
        if (!(km >= c.ducros_select)) return;
    }
    const long nvP = (long)NV * g.Ptot;
    const double dt = scal[1];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double acc = 0.0;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) {
            const double* w = F + d * nvP + (long)v * g.Ptot;
            acc -= (w[x] - w[x - g.s[d]]) * c.invd[d];
        }
        const long y = (long)v * g.Ptot + x;
        q[y] = q[y] + dt * acc;
    }
}

// flux difference at interior point x: -sum_d (F_d(x) - F_d(x - s_d)) / Delta_d
template <int NDIM>
__device__ __forceinline__ double flux_difference(const double* __restrict__ F, long nvP, long P, int v, long x, const Geom& g,
                                                  const Consts& c) {
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const double* w = F + d * nvP + (long)v * P;
        acc -= (w[x] - w[x - g.s[d]]) * c.invd[d];
    }
    return acc;
}

// R_v = -sum_d (F_d(i) - F_d(i-1)) / Delta_d   (shock_capturing.py:23-36)
template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_flux_residual(const double* __restrict__ F, double* __restrict__ R, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const long nvP = (long)NV * g.Ptot;
#pragma unroll
    for (int v = 0; v < NV; ++v) R[(long)v * g.Ptot + x] = flux_difference<NDIM>(F, nvP, g.Ptot, v, x, g, c);
}

// Flux-difference residual fused with the RK stage update (no residual array):
// TIME 1, RK-LS (rk_LS.py:204-215): du = A du + dt R; q_out = q_in + B du
// TIME 2, classic RK3 (rk_sbli.py:129-159): q_out = qold + dt*A*R; qold += dt*B*R  (du holds qold)
template <int NDIM, int TIME>
__device__ __forceinline__ void resid_rk_item(long t, const double* __restrict__ F, const double* __restrict__ qin,
                                              double* __restrict__ qout, double* __restrict__ du, double dt, const Geom& g,
                                              const Consts& c, double A, double B) {
    constexpr int NV = NDIM + 2;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const long nvP = (long)NV * g.Ptot;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double r = flux_difference<NDIM>(F, nvP, g.Ptot, v, x, g, c);
        const long y = (long)v * g.Ptot + x;
        if (TIME == 1) {
            const double d = fma(A, du[y], dt * r);
            du[y] = d;
            qout[y] = fma(B, d, qin[y]);
        } else {
            const double o = du[y];
            qout[y] = fma(dt * A, r, o);
            du[y] = fma(dt * B, r, o);
        }
    }
}

template <int NDIM, int TIME>
__global__ void __launch_bounds__(BLOCK) k_resid_rk(const double* __restrict__ F, const double* __restrict__ qin,
                                                    double* __restrict__ qout, double* __restrict__ du,
                                                    const double* __restrict__ scal, Geom g, Consts c, double A, double B) {
    resid_rk_item<NDIM, TIME>(gtid(), F, qin, qout, du, scal[1], g, c, A, B);
}

// ---------------------------------------------------------------------------
// Last direction's face fluxes fused with the flux-difference residual and
// the RK update.  A workgroup owns FW lines along D = NDIM-1 and FL faces on
// each (faces j0-1 .. j0+FL-2, consecutive threads across the lines).  Each
// thread evaluates one face as k_face_flux does and leaves it in LDS; the
// threads of faces 1..FL-1 then update the points between them, with the
// fluxes of the other directions read from F.  One face in FL is evaluated
// twice (by two neighbouring tiles); the D-direction fluxes never reach HBM.
// The arithmetic is that of face_flux + resid_rk_item, term for term.
// ---------------------------------------------------------------------------
// lines per fused tile (2-D/3-D): 32 lines x 8 faces measured ahead of 16 x 16
// and 8 x 32 on KH 256^2 and 1024^2 (coalescing along x outweighs the
// recomputed face, one in eight)
constexpr int FUSED_FW = 32;
// widths offered to the launch-shape timing of contexts and batches (2-D/3-D)
constexpr int RESID_WS[] = {32, 16};
constexpr int RESID_NW = 2;
// rounds in which the autotuners time each candidate (alternating between the
// candidates, keeping each one's best round)
constexpr int TUNE_ROUNDS = 3;

template <int NDIM, int W = FUSED_FW>
struct FusedTile {
    static constexpr int D = NDIM - 1;
    static constexpr int FW = NDIM == 1 ? 1 : W;        // lines (consecutive x for D > 0)
    static constexpr int FL = BLOCK / FW;               // faces per line
    static constexpr int PL = FL - 1;                   // points per line
    static_assert(FL * FW == BLOCK, "one face per thread");
};

template <int NDIM, int W = FUSED_FW>
__host__ __device__ inline long fused_tiles(const Geom& g) {
    using FT = FusedTile<NDIM, W>;
    constexpr int D = NDIM - 1;
    const long nseg = (g.n[D] + FT::PL - 1) / FT::PL;
    if (NDIM == 1) return nseg;
    const long ncol = (g.n[0] + FT::FW - 1) / FT::FW;
    return ncol * nseg * (NDIM == 3 ? g.n[1] : 1);
}

// Geom::dv: the inverses the kernels split flat indices with
__host__ inline void geom_divisors(Geom& g) {
    static_assert(RESID_NW == 2, "index_divisors takes two widths");
    const int n[3] = {g.n[0], g.n[1], g.n[2]};
    const int ws[2] = {RESID_WS[0], RESID_WS[1]};
    index_divisors(g.ndim, n, g.halo, ws, BLOCK, g.dv);
}

template <int NDIM, int SCH, typename T, int TIME, int W = FUSED_FW>
__device__ __forceinline__ void face_resid_tile(long tile, const double* __restrict__ q, const double* __restrict__ F,
                                                double* __restrict__ qout, double* __restrict__ du, double dt, const Geom& g,
                                                const Consts& c, double A, double B) {
    constexpr int NV = NDIM + 2;
    using FT = FusedTile<NDIM, W>;
    constexpr int D = FT::D, FW = FT::FW, PL = FT::PL;
    __shared__ double s_F[NV][BLOCK];
    const long nseg = (g.n[D] + PL - 1) / PL;
    int col0 = 0, j0, other = 0;
    if (NDIM == 1) {
        j0 = (int)tile * PL;
    } else {
        // column, segment and (3-D) j of the tile, from the host's inverses of
        // the column count and of columns x segments at this width
        static_assert(W == RESID_WS[0] || W == RESID_WS[1], "a width of RESID_WS");
        const unsigned(&dv)[4] = g.dv[4 + (W == RESID_WS[0] ? 0 : 1)];
        const unsigned ncol = (unsigned)((g.n[0] + FW - 1) / FW), u = (unsigned)tile;
        const unsigned a = cf_div(u, dv[0], dv[1]), b = cf_div(u, dv[2], dv[3]);
        col0 = (int)(u - a * ncol) * FW;
        j0 = (int)(a - b * (unsigned)nseg) * PL;
        other = (int)b;                   // k for NDIM 2 (always 0), j for NDIM 3
    }
    const int tid = (int)threadIdx.x;
    const int w = tid % FW, lf = tid / FW;
    const int f = j0 - 1 + lf;            // face f + 1/2, stored at point f
    int i, j, k;
    if (NDIM == 1) { i = f; j = 0; k = 0; }
    else if (NDIM == 2) { i = col0 + w; j = f; k = 0; }
    else { i = col0 + w; j = other; k = f; }
    const bool ok = (NDIM == 1 || col0 + w < g.n[0]) && f < g.n[D];
    const long x = ok ? gidx(g, i, j, k) : 0;
    __syncthreads();                      // the previous tile of a persistent loop is done with LDS
    if (ok) {
        double out[NV];
        face_flux<NDIM, D, SCH, T>(q, g.Ptot, g.s[D], x, c, out);
#pragma unroll
        for (int v = 0; v < NV; ++v) s_F[v][tid] = out[v];
    }
    __syncthreads();
    // point f lies between faces lf - 1 and lf (no early return: the persistent
    // batched loop calls this again and every thread must reach its barriers)
    if (ok && lf > 0) {
        const long nvP = (long)NV * g.Ptot;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double* wd = F + d * nvP + (long)v * g.Ptot;
                acc -= (wd[x] - wd[x - g.s[d]]) * c.invd[d];
            }
            acc -= (s_F[v][tid] - s_F[v][tid - FW]) * c.invd[D];
            const long y = (long)v * g.Ptot + x;
            if (TIME == 1) {
                const double dd = fma(A, du[y], dt * acc);
                du[y] = dd;
                qout[y] = fma(B, dd, q[y]);
            } else {
                const double o = du[y];
                qout[y] = fma(dt * A, acc, o);
                du[y] = fma(dt * B, acc, o);
            }
        }
    }
}

template <int NDIM, int SCH, typename T, int TIME, int W>
__global__ void __launch_bounds__(BLOCK) k_face_resid(const double* __restrict__ q, const double* __restrict__ F,
                                                      double* __restrict__ qout, double* __restrict__ du,
                                                      const double* __restrict__ scal, Geom g, Consts c, double A, double B) {
    face_resid_tile<NDIM, SCH, T, TIME, W>(blockIdx.x, q, F, qout, du, scal[1], g, c, A, B);
}

// Constituent relations u_i, p, T, mu on the box [-ext, n+ext)
// (compressible_TGV.py:78-83; TENO6/opensbliblock00_kernels.h:48-91)
template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_constituent(const double* __restrict__ q, double* __restrict__ cr, Geom g, Consts c, int ext) {
    constexpr int NV = NDIM + 2;
    int e[3], lo[3];
    for (int d = 0; d < 3; ++d) {
        const bool act = d < NDIM;
        e[d] = act ? g.n[d] + 2 * ext : 1;
        lo[d] = act ? -ext : 0;
    }
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten(t, e, lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const double r = q[x];
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const double u = q[(long)(1 + d) * g.Ptot + x] / r;
        cr[(long)d * g.Ptot + x] = u;
        ke -= 0.5 * (u * u) * r;
    }
    const double p = c.gamma_m1 * (ke + q[(long)(NV - 1) * g.Ptot + x]);
    const double TT = (c.minf * c.minf) * c.gama * p / r;
    cr[(long)NDIM * g.Ptot + x] = p;
    cr[(long)(NDIM + 1) * g.Ptot + x] = TT;
    cr[(long)(NDIM + 2) * g.Ptot + x] = 1.4042 * TT * sqrt(TT) / (0.40417 + TT);
}

template <class A>
__device__ __forceinline__ double d1(A f, double invd) {
    return fma(1.0 / 12.0, f(-2) - f(2), (2.0 / 3.0) * (f(1) - f(-1))) * invd;
}
template <class A>
__device__ __forceinline__ double d2(A f, double inv2d) {
    return fma(-1.0 / 12.0, f(-2) + f(2), fma(4.0 / 3.0, f(-1) + f(1), (-5.0 / 2.0) * f(0))) * inv2d;
}

// Carpenter's 4th-order boundary closure (Carpenter_scheme.py:6-125) at the
// points next to extrapolation / Dirichlet faces (opensblifunctions.py:486-534):
// first derivative at r = 0..3 points from the face from the 6 points nearest
// it (negated on side 1), second derivative at r = 0, 1 from 5 points.  The
// coefficients are al4^-1 ar4 of Carpenter_scheme.py:97-121, correctly rounded.
__device__ __forceinline__ double carp_c1(int r, int j) {
    constexpr double C[4][6] = {
        {-11.0 / 6.0, 3.0, -1.5, 1.0 / 3.0, 0.0, 0.0},
        {-0.3762836775133548, -0.3224849328821596, 0.719443173328853, 0.03941685243994666, -0.0658051057710398, 0.005713690397754591},
        {0.11344647038423929, -0.7912455927658653, 0.19718433388773493, 0.521455851089594, -0.0367146847001282, -0.004126377895574793},
        {-0.009325979850499882, 0.12193715322406413, -0.7278221477245911, 0.045103322334387165, 0.6521410848612417, -0.08203343284460206}};
    return C[r][j];
}
__device__ __forceinline__ double carp_c2(int r, int j) {
    constexpr double C[2][5] = {{35.0 / 12.0, -104.0 / 12.0, 114.0 / 12.0, -56.0 / 12.0, 11.0 / 12.0},
                                {11.0 / 12.0, -20.0 / 12.0, 6.0 / 12.0, 4.0 / 12.0, -1.0 / 12.0}};
    return C[r][j];
}
// dirn = +1 on side 0 (points j - r from x), -1 on side 1 (r - j)
template <class A>
__device__ __forceinline__ double carp1(A f, int r, int dirn) {
    double s = carp_c1(r, 0) * f(-dirn * r);
    for (int j = 1; j < 6; ++j) s = fma(carp_c1(r, j), f(dirn * (j - r)), s);
    return s;
}
template <class A>
__device__ __forceinline__ double carp2(A f, int r, int dirn) {
    double s = carp_c2(r, 0) * f(-dirn * r);
    for (int j = 1; j < 5; ++j) s = fma(carp_c2(r, j), f(dirn * (j - r)), s);
    return s;
}
// central derivatives at index idx of n along a direction with closure mask
template <class A>
__device__ __forceinline__ double d1c(A f, double invd, int idx, int n, int mask) {
    if (mask) {
        if ((mask & 1) && idx < 4) return carp1(f, idx, 1) * invd;
        if ((mask & 2) && idx >= n - 4) return -carp1(f, n - 1 - idx, -1) * invd;
    }
    return d1(f, invd);
}
template <class A>
__device__ __forceinline__ double d2c(A f, double inv2d, int idx, int n, int mask) {
    if (mask) {
        if ((mask & 1) && idx < 2) return carp2(f, idx, 1) * inv2d;
        if ((mask & 2) && idx >= n - 2) return carp2(f, n - 1 - idx, -1) * inv2d;
    }
    return d2(f, inv2d);
}

// Stored velocity gradients g[m*NDIM+j] = D_j u_m on [0,n) along j and
// [-2, n+2) across (StoreSome.py:87-135; TENO6/opensbli.cpp:340-384).
template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_gradients(const double* __restrict__ cr, double* __restrict__ gr, Geom g, Consts c) {
    int e[3], lo[3];
    for (int d = 0; d < 3; ++d) {
        const bool act = d < NDIM;
        e[d] = act ? g.n[d] + 4 : 1;
        lo[d] = act ? -2 : 0;
    }
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int ii[3];
    unflatten(t, e, lo, ii[0], ii[1], ii[2]);
    const long x = gidx(g, ii[0], ii[1], ii[2]);
#pragma unroll
    for (int jd = 0; jd < NDIM; ++jd) {
        if (ii[jd] < 0 || ii[jd] >= g.n[jd]) continue;
        const long st = g.s[jd];
#pragma unroll
        for (int m = 0; m < NDIM; ++m) {
            const double* um = cr + (long)m * g.Ptot;
            gr[(long)(m * NDIM + jd) * g.Ptot + x] = d1c([&](int s) { return um[x + s * st]; }, c.invd[jd], ii[jd], g.n[jd], c.carp[jd]);
        }
    }
}

// Central split-form convective residual, written first.
// SKEW = false: Feiereisen form (split_forms.py:111-144) with StoreSome gradients.
// SKEW = true : Skew(ab) = 1/2(D(ab) + b D(a) + a D(b)) terms of the tg_sym app
//               (parsing.py:96-112; verify_tg_sym.py:32-36).
enum CcForm { CC_FEIEREISEN = 0, CC_SKEW = 1, CC_KGP = 2 };

template <int NDIM, int FORM>
__global__ void __launch_bounds__(BLOCK) k_central_convective(const double* __restrict__ q, const double* __restrict__ cr,
                                                              const double* __restrict__ gr, double* __restrict__ R, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const int ii[3] = {i, j, k};
    const long x = gidx(g, i, j, k);
    const long P = g.Ptot;
    const double* rho = q;
    const double* rhoE = q + (NV - 1) * P;
    const double* pp = cr + NDIM * P;
    double div_rhou = 0.0, Rrho = 0.0, eterm = 0.0;
    double conv[NDIM];
#pragma unroll
    for (int m = 0; m < NDIM; ++m) conv[m] = 0.0;
#pragma unroll
    for (int jd = 0; jd < NDIM; ++jd) {
        const long st = g.s[jd];
        const double inv = c.invd[jd];
        const int id = ii[jd], nn = g.n[jd], cm = c.carp[jd];
        const double* rhou_j = q + (long)(1 + jd) * P;
        const double* u_j = cr + (long)jd * P;
        auto at = [&](const double* a, int s) { return a[x + s * st]; };
        if (FORM == CC_KGP) {
            // KGP, enthalpy splitting, all coefficients 1/4 (split_forms.py:178-234): the
            // same term sequence as the CPU oracle; H = (rhoE + p)/rho at each stencil point
            const double ujc = at(u_j, 0), rhoujc = at(rhou_j, 0);
            const double gjj = gr[(long)(jd * NDIM + jd) * P + x];
            auto Hs = [&](int s) { return (at(rhoE, s) + at(pp, s)) / at(rho, s); };
            const double d_rhouj = d1c([&](int s) { return at(rhou_j, s); }, inv, id, nn, cm);
            const double d_rho = d1c([&](int s) { return at(rho, s); }, inv, id, nn, cm);
            Rrho += 0.5 * d_rhouj + 0.5 * (ujc * d_rho + rho[x] * gjj);
#pragma unroll
            for (int m = 0; m < NDIM; ++m) {
                const double* rhou_m = q + (long)(1 + m) * P;
                const double* u_m = cr + (long)m * P;
                const double umc = at(u_m, 0), rhoumc = at(rhou_m, 0);
                const double g_mj = gr[(long)(m * NDIM + jd) * P + x];
                const double d_rhouj_um = d1c([&](int s) { return at(u_m, s) * at(rhou_j, s); }, inv, id, nn, cm);
                const double d_rhoum = d1c([&](int s) { return at(rhou_m, s); }, inv, id, nn, cm);
                const double d_ujum = d1c([&](int s) { return at(u_j, s) * at(u_m, s); }, inv, id, nn, cm);
                conv[m] += 0.25 * ((((((d_rhouj_um + umc * d_rhouj) + rhoujc * g_mj) + ujc * d_rhoum) + rhoumc * gjj) +
                                    rho[x] * d_ujum) + (umc * ujc) * d_rho);
            }
            const double Hc = Hs(0);
            const double d_Hrhouj = d1c([&](int s) { return Hs(s) * at(rhou_j, s); }, inv, id, nn, cm);
            const double d_H = d1c(Hs, inv, id, nn, cm);
            const double d_Hrho = d1c([&](int s) { return Hs(s) * at(rho, s); }, inv, id, nn, cm);
            const double d_Huj = d1c([&](int s) { return Hs(s) * at(u_j, s); }, inv, id, nn, cm);
            eterm += 0.25 * ((((((d_Hrhouj + Hc * d_rhouj) + rhoujc * d_H) + ujc * d_Hrho) + (rho[x] * Hc) * gjj) +
                              rho[x] * d_Huj) + (ujc * Hc) * d_rho);
        } else if (FORM == CC_FEIEREISEN) {
            const double d_rhouj = d1c([&](int s) { return at(rhou_j, s); }, inv, id, nn, cm);
            div_rhou += d_rhouj;
#pragma unroll
            for (int m = 0; m < NDIM; ++m) {
                const double* rhou_m = q + (long)(1 + m) * P;
                const double d_prod = d1c([&](int s) { return at(rhou_m, s) * at(u_j, s); }, inv, id, nn, cm);
                conv[m] += d_prod + at(rhou_j, 0) * gr[(long)(m * NDIM + jd) * P + x];
            }
            const double d_Eu = d1c([&](int s) { return at(rhoE, s) * at(u_j, s); }, inv, id, nn, cm);
            const double d_E = d1c([&](int s) { return at(rhoE, s) / at(rho, s); }, inv, id, nn, cm);
            const double d_pu = d1c([&](int s) { return at(pp, s) * at(u_j, s); }, inv, id, nn, cm);
            eterm += -0.5 * (d_Eu + at(rhou_j, 0) * d_E) - d_pu;
        } else {
            const double ujc = at(u_j, 0);
            const double d_rhou = d1c([&](int s) { return at(rho, s) * at(u_j, s); }, inv, id, nn, cm);
            const double d_rho = d1c([&](int s) { return at(rho, s); }, inv, id, nn, cm);
            const double d_u = d1c([&](int s) { return at(u_j, s); }, inv, id, nn, cm);
            Rrho -= 0.5 * (d_rhou + ujc * d_rho + rho[x] * d_u);
#pragma unroll
            for (int m = 0; m < NDIM; ++m) {
                const double* rhou_m = q + (long)(1 + m) * P;
                const double d_prod = d1c([&](int s) { return at(rhou_m, s) * at(u_j, s); }, inv, id, nn, cm);
                const double d_rm = d1c([&](int s) { return at(rhou_m, s); }, inv, id, nn, cm);
                conv[m] += 0.5 * (d_prod + ujc * d_rm + at(rhou_m, 0) * d_u);
            }
            const double d_Eu = d1c([&](int s) { return at(rhoE, s) * at(u_j, s); }, inv, id, nn, cm);
            const double d_rE = d1c([&](int s) { return at(rhoE, s); }, inv, id, nn, cm);
            const double d_pu = d1c([&](int s) { return at(pp, s) * at(u_j, s); }, inv, id, nn, cm);
            eterm += -0.5 * (d_Eu + ujc * d_rE + rhoE[x] * d_u) - d_pu;
        }
    }
#pragma unroll
    for (int m = 0; m < NDIM; ++m) {
        const long st = g.s[m];
        const double dp = d1c([&](int s) { return pp[x + s * st]; }, c.invd[m], ii[m], g.n[m], c.carp[m]);
        double rm;
        if (FORM == CC_FEIEREISEN) rm = -dp - 0.5 * (conv[m] + cr[(long)m * P + x] * div_rhou);
        else rm = -dp - conv[m];
        R[(long)(1 + m) * P + x] = rm;
    }
    if (FORM == CC_KGP) {
        R[x] = -Rrho;
        R[(long)(NV - 1) * P + x] = -eterm;
    } else if (FORM == CC_FEIEREISEN) {
        R[x] = -div_rhou;
        R[(long)(NV - 1) * P + x] = eterm - 0.5 * (rhoE[x] / rho[x]) * div_rhou;
    } else {
        R[x] = Rrho;
        R[(long)(NV - 1) * P + x] = eterm;
    }
}

// Viscous stresses and heat conduction added to R (split_forms.py:383-392;
// TENO6/opensbliblock00_kernels.h:4568-4707).  Mixed second derivatives are
// D_b(D_a u) with a < b on the stored gradients.
template <int NDIM, bool SUTH>
__global__ void __launch_bounds__(BLOCK) k_viscous(const double* __restrict__ cr, const double* __restrict__ gr,
                                                   double* __restrict__ R, Geom g, Consts c) {
    constexpr int NV = NDIM + 2;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const int ii[3] = {i, j, k};
    const long x = gidx(g, i, j, k);
    const long P = g.Ptot;
    const double* Ta = cr + (long)(NDIM + 1) * P;
    const double* mua = cr + (long)(NDIM + 2) * P;
    double G[NDIM][NDIM], D2u[NDIM][NDIM], dmu[NDIM], dT[NDIM];
    double lapT = 0.0;
#pragma unroll
    for (int m = 0; m < NDIM; ++m)
#pragma unroll
        for (int jd = 0; jd < NDIM; ++jd) G[m][jd] = gr[(long)(m * NDIM + jd) * P + x];
#pragma unroll
    for (int jd = 0; jd < NDIM; ++jd) {
        const long st = g.s[jd];
        const int id = ii[jd], nn = g.n[jd], cm = c.carp[jd];
        dmu[jd] = SUTH ? d1c([&](int s) { return mua[x + s * st]; }, c.invd[jd], id, nn, cm) : 0.0;
        dT[jd] = d1c([&](int s) { return Ta[x + s * st]; }, c.invd[jd], id, nn, cm);
        lapT += d2c([&](int s) { return Ta[x + s * st]; }, c.inv2d[jd], id, nn, cm);
#pragma unroll
        for (int m = 0; m < NDIM; ++m) {
            const double* um = cr + (long)m * P;
            D2u[m][jd] = d2c([&](int s) { return um[x + s * st]; }, c.inv2d[jd], id, nn, cm);
        }
    }
    double div = 0.0;
#pragma unroll
    for (int m = 0; m < NDIM; ++m) div += G[m][m];
    const double muv = SUTH ? mua[x] : 1.0;
    double Eacc = 0.0;
#pragma unroll
    for (int m = 0; m < NDIM; ++m) {
        double Lm = (4.0 / 3.0) * D2u[m][m];
#pragma unroll
        for (int jd = 0; jd < NDIM; ++jd) {
            if (jd == m) continue;
            const int a = jd < m ? jd : m, b = jd < m ? m : jd;
            const double* ga = gr + (long)(jd * NDIM + a) * P;
            const long sb = g.s[b];
            const double mixed = d1c([&](int s) { return ga[x + s * sb]; }, c.invd[b], ii[b], g.n[b], c.carp[b]);
            Lm += D2u[m][jd] + (1.0 / 3.0) * mixed;
        }
        double S[NDIM];
#pragma unroll
        for (int jd = 0; jd < NDIM; ++jd) S[jd] = G[m][jd] + G[jd][m] - (m == jd ? (2.0 / 3.0) * div : 0.0);
        double mom = muv * Lm;
        if (SUTH) {
#pragma unroll
            for (int jd = 0; jd < NDIM; ++jd) mom += S[jd] * dmu[jd];
        }
        R[(long)(1 + m) * P + x] += c.invRe * mom;
        double e = cr[(long)m * P + x] * mom;                 // u_i d(tau_ij)/dx_j
#pragma unroll
        for (int jd = 0; jd < NDIM; ++jd) e += muv * S[jd] * G[m][jd];   // tau_ij du_i/dx_j
        Eacc += e;
    }
    double heat = muv * lapT;
    if (SUTH) {
#pragma unroll
        for (int jd = 0; jd < NDIM; ++jd) heat += dT[jd] * dmu[jd];
    }
    R[(long)(NV - 1) * P + x] += c.invRe * Eacc + c.cheat * heat;
}

// Low-storage RK3 stage (rk_LS.py:204-215): du = A du + dt R; q_out = q_in + B du
static __global__ void __launch_bounds__(BLOCK) k_rkls(const double* __restrict__ qin, double* __restrict__ qout, double* __restrict__ du,
                                                const double* __restrict__ R, const double* __restrict__ scal, Geom g, int nv,
                                                double A, double B) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const double dt = scal[1];
    for (int v = 0; v < nv; ++v) {
        const long y = (long)v * g.Ptot + x;
        const double d = fma(A, du[y], dt * R[y]);
        du[y] = d;
        qout[y] = fma(B, d, qin[y]);
    }
}

// Classic RK3 (rk_sbli.py:129-159): q = qold + dt*rknew*R; qold += dt*rkold*R
static __global__ void __launch_bounds__(BLOCK) k_rk3_stage(double* __restrict__ qout, double* __restrict__ qold,
                                                     const double* __restrict__ R, const double* __restrict__ scal, Geom g, int nv,
                                                     double rknew, double rkold) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    const double dt = scal[1];
    for (int v = 0; v < nv; ++v) {
        const long y = (long)v * g.Ptot + x;
        const double r = R[y], o = qold[y];
        qout[y] = fma(dt * rknew, r, o);
        qold[y] = fma(dt * rkold, r, o);
    }
}

static __global__ void __launch_bounds__(BLOCK) k_copy_interior(const double* __restrict__ src, double* __restrict__ dst, Geom g, int nv) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const long x = gidx(g, i, j, k);
    for (int v = 0; v < nv; ++v) dst[(long)v * g.Ptot + x] = src[(long)v * g.Ptot + x];
}

// The two faces of one direction: OSB_BC_* type and, for OSB_BC_DIRICHLET,
// the conservative values written there.
struct BcSides {
    int type[2];
    double val[2][5];
};

// Boundary conditions along DIR (bc_core.py:128-202): side 0 layers t=0..h
// (t = 0 is the boundary point), side 1 likewise, over the halo-extended
// range of the other directions.
template <int DIR>
__host__ __device__ inline long bc_work(const Geom& g) {
    long w = 2L * (g.halo + 1);
    for (int d = 0; d < g.ndim; ++d)
        if (d != DIR) w *= (g.n[d] + 2 * g.halo);
    return w;
}

template <int DIR>
__device__ __forceinline__ void bc_item(long t, double* __restrict__ q, const Geom& g, int nv, const BcSides& b) {
    const int h = g.halo;
    int e[3], lo[3];
    for (int d = 0; d < 3; ++d) {
        const bool act = d < g.ndim;
        e[d] = act ? g.n[d] + 2 * h : 1;
        lo[d] = act ? -h : 0;
    }
    e[DIR] = 2 * (h + 1);
    lo[DIR] = 0;
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int ii[3];
    unflatten_dv(t, e, g.dv[6 + DIR], lo, ii[0], ii[1], ii[2]);
    const int layer = ii[DIR];                      // in [0, 2h + 2)
    const int side = layer > h ? 1 : 0;
    const int tt = layer - side * (h + 1);
    const int type = b.type[side];
    const int n = g.n[DIR];
    const int bnd = side == 0 ? 0 : n - 1;          // boundary point
    const int inw = side == 0 ? 1 : -1;             // towards the interior (set_side_factor)
    int a[3] = {ii[0], ii[1], ii[2]};
    auto at = [&](int i) { a[DIR] = i; return gidx(g, a[0], a[1], a[2]); };
    if (type == OSB_BC_EXTRAPOLATION1) {
        // extrapolation.py:41-54: layer t reads layers t-1 and t-2, so one
        // thread walks the layers of its line in order
        if (tt != 0) return;
        for (int s = 1; s <= h; ++s) {
            const long x0 = at(bnd - inw * s), x1 = at(bnd - inw * (s - 1)), x2 = at(bnd - inw * (s - 2));
            for (int v = 0; v < nv; ++v) {
                double* qv = q + (long)v * g.Ptot;
                qv[x0] = 2.0 * qv[x1] - qv[x2];
            }
        }
        return;
    }
    if (type == OSB_BC_DIRICHLET) {                 // dirichlet.py:28-41
        const long xd = at(bnd - inw * tt);
        for (int v = 0; v < nv; ++v) q[(long)v * g.Ptot + xd] = b.val[side][v];
        return;
    }
    int dst, src;
    bool negate = false, zero = false;
    if (tt == 0) {
        // zero-gradient, inviscid wall and order-0 extrapolation modify the boundary point
        if (type != OSB_BC_ZERO_GRADIENT && type != OSB_BC_INVISCID_WALL && type != OSB_BC_EXTRAPOLATION0) return;
        dst = bnd;
        src = bnd + inw;
        zero = (type == OSB_BC_INVISCID_WALL);
    } else if (type == OSB_BC_EXTRAPOLATION0) {     // extrapolation.py:36-40
        dst = bnd - inw * tt;
        src = bnd + inw;
    } else {
        if (side == 0) { dst = -tt; src = (type == OSB_BC_PERIODIC) ? n - tt : tt; }
        else { dst = n - 1 + tt; src = (type == OSB_BC_PERIODIC) ? tt - 1 : n - 1 - tt; }
        negate = (type == OSB_BC_INVISCID_WALL || type == OSB_BC_SYMMETRY);
    }
    const long xd = at(dst);
    const long xs = at(src);
    for (int v = 0; v < nv; ++v) {
        double val = q[(long)v * g.Ptot + xs];
        if (v == 1 + DIR) {
            if (negate) val = val - 2.0 * val;     // v - 2 (v.n) n   (inviscid_wall.py:37)
            if (zero) val = val - 1.0 * val;       // v - (v.n) n     (inviscid_wall.py:40)
        }
        q[(long)v * g.Ptot + xd] = val;
    }
}

template <int DIR>
__global__ void __launch_bounds__(BLOCK) k_bc(double* __restrict__ q, Geom g, int nv, BcSides b) {
    bc_item<DIR>(gtid(), q, g, nv, b);
}

// Halo points of the halo-extended box when every face is periodic, in one
// pass: region d holds the points with directions > d interior, direction d
// in its 2h halo layers and directions < d over the extended range, so the
// regions partition the halo. Each point copies the interior point at its
// index wrapped in every direction. With n[d] >= h in every direction that
// is the value the direction-by-direction fill (k_bc<0>, k_bc<1>, k_bc<2>,
// each over the extended range of the other directions) leaves at every halo
// point, corners included; sources are interior points only, so no thread
// reads what another writes.
__host__ __device__ inline long bc_periodic_work(const Geom& g) {
    long ext = 1, in = 1;
    for (int d = 0; d < g.ndim; ++d) {
        ext *= g.n[d] + 2 * g.halo;
        in *= g.n[d];
    }
    return ext - in;
}

static __global__ void __launch_bounds__(BLOCK) k_bc_periodic(double* __restrict__ q, Geom g, int nv) {
    long t = gtid();
    const int h = g.halo;
    for (int d = 0; d < g.ndim; ++d) {
        int e[3], lo[3];
        long cnt = 1;
        for (int a = 0; a < 3; ++a) {
            if (a >= g.ndim) { e[a] = 1; lo[a] = 0; }
            else if (a < d) { e[a] = g.n[a] + 2 * h; lo[a] = -h; }
            else if (a == d) { e[a] = 2 * h; lo[a] = 0; }
            else { e[a] = g.n[a]; lo[a] = 0; }
            cnt *= e[a];
        }
        if (t < cnt) {
            int c[3];
            unflatten(t, e, lo, c[0], c[1], c[2]);
            c[d] = c[d] < h ? c[d] - h : g.n[d] + (c[d] - h);
            int s[3] = {c[0], c[1], c[2]};
            for (int a = 0; a < g.ndim; ++a) s[a] = s[a] < 0 ? s[a] + g.n[a] : (s[a] >= g.n[a] ? s[a] - g.n[a] : s[a]);
            const long xd = gidx(g, c[0], c[1], c[2]), xs = gidx(g, s[0], s[1], s[2]);
            for (int v = 0; v < nv; ++v) q[(long)v * g.Ptot + xd] = q[(long)v * g.Ptot + xs];
            return;
        }
        t -= cnt;
    }
}

// host interior layout (v, k, j, i) <-> padded device layout
static __global__ void __launch_bounds__(BLOCK) k_scatter(const double* __restrict__ src, double* __restrict__ q, Geom g, int nv) {
    const long npts = (long)g.n[0] * g.n[1] * g.n[2];
    const long t = gtid();
    if (t >= npts * nv) return;
    const int v = (int)(t / npts);
    const long r = t % npts;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    int i, j, k;
    unflatten(r, e, lo, i, j, k);
    q[(long)v * g.Ptot + gidx(g, i, j, k)] = src[t];
}
static __global__ void __launch_bounds__(BLOCK) k_gather(const double* __restrict__ q, double* __restrict__ dst, Geom g, int nv) {
    const long npts = (long)g.n[0] * g.n[1] * g.n[2];
    const long t = gtid();
    if (t >= npts * nv) return;
    const int v = (int)(t / npts);
    const long r = t % npts;
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    int i, j, k;
    unflatten(r, e, lo, i, j, k);
    dst[t] = q[(long)v * g.Ptot + gidx(g, i, j, k)];
}

// ---------------------------------------------------------------------------
// Reductions.  Wave-level butterfly through DPP/permute shuffles, then LDS
// across the block's waves, then one atomic per block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

static __global__ void k_zero_scalar(double* scal) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { reinterpret_cast<unsigned long long*>(scal)[0] = 0ull; }
}

// max over interior points of sum_d (|u_d| + a)/Delta_d; stored as the bit
// pattern of a non-negative double in scal[0] (order-preserving for >= 0)
template <int NDIM>
__device__ __forceinline__ double wavespeed_point(const double* __restrict__ q, long x, const Geom& g, const Consts& c) {
    constexpr int NV = NDIM + 2;
    const double r = q[x];
    double uu[NDIM], ke = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) { uu[d] = q[(long)(1 + d) * g.Ptot + x] / r; ke -= 0.5 * (uu[d] * uu[d]) * r; }
    const double p = c.gamma_m1 * (ke + q[(long)(NV - 1) * g.Ptot + x]);
    const double a = sqrt(c.gama * p / r);
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) s += (fabs(uu[d]) + a) * c.invd[d];
    return s;
}

// block max of m, then one atomicMax on the bit pattern (order-preserving for
// non-negative doubles, so the result does not depend on the order)
__device__ __forceinline__ void block_max_to(double m, double* scal) {
    __shared__ double red[BLOCK / 64];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fmax(b, red[w]);
        atomicMax(reinterpret_cast<unsigned long long*>(scal), (unsigned long long)__double_as_longlong(b));
    }
    __syncthreads();
}

template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_wavespeed(const double* __restrict__ q, double* scal, Geom g, Consts c) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long tot = (long)e[0] * e[1] * e[2];
    double m = 0.0;
    for (long t = gtid(); t < tot; t += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        unflatten_dv(t, e, g.dv[0], lo, i, j, k);
        m = fmax(m, wavespeed_point<NDIM>(q, gidx(g, i, j, k), g, c));
    }
    block_max_to(m, scal);
}

// GLF wave speeds ("Global wave-speed reductions", weno.py:515-519,
// shock_capturing.py:838-857): per direction the maxima over the grid of
// |u_d|, |u_d + a|, |u_d - a|, from the same point primitives as the face
// fluxes.  Stored as bit patterns of non-negative doubles (order-preserving
// for atomicMax); k_glf_zero clears them first.
static __global__ void k_glf_zero(double* glf) {
    if (blockIdx.x == 0 && threadIdx.x < 9) reinterpret_cast<unsigned long long*>(glf)[threadIdx.x] = 0ull;
}

template <int NDIM>
__device__ __forceinline__ void glf_point(const double* __restrict__ q, long x, long Ptot, const Consts& c,
                                          double (&m)[3 * NDIM]) {
    double v[NDIM + 2], f[NDIM + 2], uu[NDIM], p, a;
#pragma unroll
    for (int k = 0; k < NDIM + 2; ++k) v[k] = q[(long)k * Ptot + x];
    point_prims<NDIM, 0, false>(v, c, uu, p, f, a);
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        m[3 * d] = fmax(m[3 * d], fabs(uu[d]));
        m[3 * d + 1] = fmax(m[3 * d + 1], fabs(uu[d] + a));
        m[3 * d + 2] = fmax(m[3 * d + 2], fabs(uu[d] - a));
    }
}

template <int NDIM>
__device__ __forceinline__ void glf_block_to(double (&m)[3 * NDIM], double* glf) {
    __shared__ double red[3 * NDIM][BLOCK / 64];
#pragma unroll
    for (int r = 0; r < 3 * NDIM; ++r) {
        const double w = wave_max(m[r]);
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 3 * NDIM) {
        double b = red[threadIdx.x][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fmax(b, red[threadIdx.x][w]);
        atomicMax(reinterpret_cast<unsigned long long*>(glf) + threadIdx.x, (unsigned long long)__double_as_longlong(b));
    }
    __syncthreads();
}

template <int NDIM>
__global__ void __launch_bounds__(BLOCK) k_glf(const double* __restrict__ q, double* glf, Geom g, Consts c) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long tot = (long)e[0] * e[1] * e[2];
    double m[3 * NDIM];
#pragma unroll
    for (int r = 0; r < 3 * NDIM; ++r) m[r] = 0.0;
    for (long t = gtid(); t < tot; t += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        unflatten_dv(t, e, g.dv[0], lo, i, j, k);
        glf_point<NDIM>(q, gidx(g, i, j, k), g.Ptot, c, m);
    }
    glf_block_to<NDIM>(m, glf);
}

// dt = cfl / max wave speed; time += dt
static __global__ void k_dt_finalize(double* scal, double cfl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double m = __longlong_as_double((long long)reinterpret_cast<unsigned long long*>(scal)[0]);
        const double dt = cfl / m;
        scal[1] = dt;
        scal[2] += dt;
    }
}

// Taylor-Green reductions: per-block partial sums (deterministic two-pass)
static __global__ void __launch_bounds__(BLOCK) k_tgv_partials(const double* __restrict__ qrho, const double* __restrict__ cr,
                                                        double* __restrict__ part, Geom g, Consts c) {
    __shared__ double red[4][BLOCK / 64];
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long tot = (long)e[0] * e[1] * e[2];
    double acc[4] = {0, 0, 0, 0};
    const long P = g.Ptot;
    for (long t = gtid(); t < tot; t += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        unflatten_dv(t, e, g.dv[0], lo, i, j, k);
        const long x = gidx(g, i, j, k);
        double G[3][3];
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d) {
                const double* um = cr + (long)m * P;
                const long st = g.s[d];
                G[m][d] = d1([&](int s) { return um[x + s * st]; }, c.invd[d]);
            }
        const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
        const double divV = G[0][0] + G[1][1] + G[2][2];
        const double rho = qrho[x], mu = cr[5 * P + x];
        const double u0 = cr[x], u1 = cr[P + x], u2 = cr[2 * P + x];
        acc[0] += 0.5 * (u0 * u0 + u1 * u1 + u2 * u2) * rho;
        acc[1] += (4.0 / 3.0) * (divV * divV) * mu;
        acc[2] += (wx * wx + wy * wy + wz * wz) * mu;
        acc[3] += rho;
    }
    for (int r = 0; r < 4; ++r) {
        const double s = wave_sum(acc[r]);
        if ((threadIdx.x & 63) == 0) red[r][threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
        part[(long)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

static __global__ void k_sum_partials(const double* __restrict__ part, double* __restrict__ out, int nblocks) {
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += part[(long)b * 4 + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

static __global__ void __launch_bounds__(BLOCK) k_nan_count(const double* __restrict__ q, unsigned int* cnt, Geom g) {
    const int e[3] = {g.n[0], g.n[1], g.n[2]};
    const int lo[3] = {0, 0, 0};
    const long t = gtid();
    if (t >= (long)e[0] * e[1] * e[2]) return;
    int i, j, k;
    unflatten_dv(t, e, g.dv[0], lo, i, j, k);
    const double r = q[gidx(g, i, j, k)];
    if (!isfinite(r)) atomicAdd(cnt, 1u);
}

}  // namespace osb
