// osb_central_fused.hpp -- one-launch RK stage for the central schemes on gfx950.
//
// One workgroup owns a TX x TY x TZ tile of grid points (16 x 4 x 4,
// 16 x 8 x 4 or 8 x 8 x 8, timed at a context's first step).  Each thread
// first issues the loads of its own point's RK operands (du and q), which are
// consumed only at the end.  Phase 1 evaluates the constituent relations
// once per point of the tile plus a 2-point halo box and stages them in LDS:
//   u0, u1, u2, theta = p/rho, rho        (mu as well for Sutherland's law)
// five FP64 fields = 51.2 KB at the default tile.  Every other pointwise
// quantity the stencils difference is rebuilt at the stencil point from
// these (rho u_m, p = rho theta, E = rhoE/rho = theta/(gama-1) + |u|^2/2,
// rhoE = rho E, H = (rhoE+p)/rho = gama theta/(gama-1) + |u|^2/2,
// T = gama Minf^2 theta), so phase 2 reads no stencil from HBM or L2.  Each
// derivative is still the central difference of the pointwise field (the
// same discrete operator as StoreSome.py:138-228 and split_forms.py:111-144,
// 178-234, 383-392), evaluated with one reciprocal of rho per box point and
// with stencil weights pre-scaled by 1/Delta.  The result rounds differently
// from the CPU oracle's sequence of divisions; it agrees with it to the FP64
// parity bar (relative L2 / L-inf <= 1e-11).
//
// Phase 2 gives every tile point its full right-hand side -- split-form
// convective terms, viscous stresses and heat conduction, or the
// interface-flux difference of a shock-capturing scheme -- followed directly
// by the RK update, so the stage reads q, du (and the face fluxes) and writes
// du and q once.  It runs in sections separated by scheduling fences (one per
// direction, one per momentum component of the viscous stress), which keeps
// the kernel within the 168 VGPRs of 3 waves/SIMD (160 for Feiereisen with
// constant viscosity at 16 x 4 x 4).
#pragma once

#include "osb_kernels.hpp"
#include "osb_order.hpp"

namespace osb {

// tile shapes offered to the tile selection: {TX, TY, TZ}
struct CfTile { int tx, ty, tz; };
constexpr CfTile CF_TILES[] = {{16, 4, 4}, {16, 8, 4}, {8, 8, 8}};
constexpr int CF_NTILES = 3;
// index of a tile shape in CF_TILES, -1 if it is not one
__host__ __device__ constexpr int cf_tile_index(int tx, int ty, int tz) {
    for (int t = 0; t < CF_NTILES; ++t)
        if (CF_TILES[t].tx == tx && CF_TILES[t].ty == ty && CF_TILES[t].tz == tz) return t;
    return -1;
}

enum CfConv { CF_NONE = 0, CF_FEIEREISEN = 1, CF_SKEW = 2, CF_KGP = 3 };

// Central differences with the pre-scaled weights of Consts (w1, w2, wm).
// f(s) is the field at offset s along the direction.
template <class A>
__device__ __forceinline__ double fd1(A f, const double (&w)[2]) {
    return fma(w[0], f(1) - f(-1), w[1] * (f(-2) - f(2)));
}
template <class A>
__device__ __forceinline__ double fd2(A f, const double (&w)[3]) {
    return fma(w[2], f(-2) + f(2), fma(w[1], f(-1) + f(1), w[0] * f(0)));
}
// D_b(D_a f) with f(r, s) the field at offset r along a and s along b: the
// product of the two antisymmetric 4th-order first-derivative stencils,
// grouped by |r|, |s| (wm = w_a(1) w_b(1), w_a(2) w_b(1), w_a(1) w_b(2), w_a(2) w_b(2))
template <class A>
__device__ __forceinline__ double fd11(A f, const double (&wm)[4]) {
    const double x11 = (f(1, 1) - f(-1, 1)) - (f(1, -1) - f(-1, -1));
    const double x21 = (f(2, 1) - f(-2, 1)) - (f(2, -1) - f(-2, -1));
    const double x12 = (f(1, 2) - f(-1, 2)) - (f(1, -2) - f(-1, -2));
    const double x22 = (f(2, 2) - f(-2, 2)) - (f(2, -2) - f(-2, -2));
    return fma(wm[0], x11, fma(wm[1], x21, fma(wm[2], x12, wm[3] * x22)));
}

// Host side: the pre-scaled weights and constants the fused kernels read.
// D1 f = (2/3)(f(1) - f(-1))/Delta + (1/12)(f(-2) - f(2))/Delta
// D2 f = (-5/2 f(0) + 4/3 (f(-1) + f(1)) - 1/12 (f(-2) + f(2)))/Delta^2
inline void fill_fused_consts(Consts& k) {
    for (int d = 0; d < 3; ++d) {
        k.w1[d][0] = (2.0 / 3.0) * k.invd[d];
        k.w1[d][1] = (1.0 / 12.0) * k.invd[d];
        k.w2[d][0] = (-5.0 / 2.0) * k.inv2d[d];
        k.w2[d][1] = (4.0 / 3.0) * k.inv2d[d];
        k.w2[d][2] = (-1.0 / 12.0) * k.inv2d[d];
    }
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    for (int p = 0; p < 3; ++p) {
        // stencil weights at offsets +1, +2: w(1) = w1[0], w(2) = -w1[1]
        const double a1 = k.w1[pa[p]][0], a2 = -k.w1[pa[p]][1];
        const double b1 = k.w1[pb[p]][0], b2 = -k.w1[pb[p]][1];
        k.wm[p][0] = a1 * b1;
        k.wm[p][1] = a2 * b1;
        k.wm[p][2] = a1 * b2;
        k.wm[p][3] = a2 * b2;
    }
    k.cE = 1.0 / k.gamma_m1;
    k.cH = k.gama / k.gamma_m1;
    k.cT = (k.minf * k.minf) * k.gama;
    k.cheat_th = k.cheat * k.cT;
}

// Waves per SIMD a fused-kernel variant is compiled for: as many workgroups
// as its LDS box allows on a CU (160 KB), at most 4 waves per SIMD (a 128-VGPR
// budget).  A 256-thread tile with five FP64 fields (51.2 KB) gets 3, with
// six (Sutherland's mu or KGP) 2; a 512-thread 16x8x4 tile with five fields
// (76.8 KB) gets 4.
__host__ __device__ constexpr int cf_nfield(int conv, int visc) { return 4 + (conv != 0 ? 1 : 0) + (visc == 2 ? 1 : 0); }
__host__ __device__ constexpr int cf_waves(int tx, int ty, int tz, int nfield) {
    const int blocks_per_cu = 163840 / (nfield * (tx + 4) * (ty + 4) * (tz + 4) * 8);
    const int waves = blocks_per_cu * (tx * ty * tz / 64) / 4;
    return waves > 4 ? 4 : (waves < 1 ? 1 : waves);
}

// Scheduling fence between the sections of phase 2 (one per direction, one
// per momentum component's mixed derivatives).  Without it the scheduler
// hoists the LDS reads of later sections ahead to hide their latency, which
// takes the kernel to 248 VGPRs and 2 waves per SIMD; with it each section's
// values are loaded, used and dead before the next (160 VGPRs with phase 1's
// loads issued together and the RK operands held from the start), and
// occupancy hides the latency instead.
__device__ __forceinline__ void cf_section() { __builtin_amdgcn_sched_barrier(0); }

// index of the (a, b) pair in Consts::wm, a < b
__host__ __device__ constexpr int cf_pair(int a, int b) { return a == 0 ? b - 1 : 2; }

// CONV: CfConv; VISC: 0 none, 1 constant, 2 Sutherland; FLUXRES: start R from
// the face-flux difference (shock-capturing schemes); TIME: OSB_TIME_RKLS3 (1)
// or OSB_TIME_RK3 (2).
// One RK stage over the tile whose first point is interior point (i0, j0, k0):
// the body of k_central_fused (one tile per workgroup) and of the batched
// kb_central_fused (tiles of many blocks).
template <int CF_TX, int CF_TY, int CF_TZ, int CONV, int VISC, bool FLUXRES, int TIME>
__device__ __forceinline__ void cf_tile_stage(const int i0, const int j0, const int k0, const double* __restrict__ q,
                                              const double* __restrict__ F, double* __restrict__ reg,
                                              double* __restrict__ qout, const double* __restrict__ scal, const Geom& g,
                                              const Consts& c, const double A, const double B, const int wrap) {
    constexpr int NV = 5;
    constexpr int CF_BX = CF_TX + 4, CF_BY = CF_TY + 4, CF_BZ = CF_TZ + 4;
    constexpr int CF_BOX = CF_BX * CF_BY * CF_BZ;
    constexpr int CF_THREADS = CF_TX * CF_TY * CF_TZ;
    constexpr bool HAS_CONV = CONV != CF_NONE;
    constexpr bool SUTH = VISC == 2;
    // LDS fields: u0, u1, u2, theta, [rho], [mu], one flat array
    constexpr int F_TH = 3, F_RHO = 4, F_MU = HAS_CONV ? 5 : 4;
    constexpr int NFIELD = cf_nfield(CONV, VISC);
    __shared__ double s_f[NFIELD * CF_BOX];

    const long P = g.Ptot;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];

    // The RK operands of the thread's own point (du and q, or q_old) are
    // loaded first, so that they arrive while phases 1 and 2 run.
    const int tx = threadIdx.x % CF_TX, ty = (threadIdx.x / CF_TX) % CF_TY, tz = threadIdx.x / (CF_TX * CF_TY);
    const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
    const bool active = i < n0 && j < n1 && k < n2;
    const long x = active ? gidx(g, i, j, k) : 0;
    double r_pt[NV], q_pt[NV];
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            r_pt[v] = reg[(long)v * P + x];
            if (TIME == 1) q_pt[v] = q[(long)v * P + x];
        }
    }
    cf_section();

    // ---- phase 1: constituent relations on the halo box -------------------
    // Every box point is evaluated and stored (the corners are never read);
    // loads are clamped into the stored range, so past the high end of a
    // partial tile they repeat the last halo layer, which only feeds points
    // the tile does not update.
    // Point offsets are 32-bit: a block's points per variable stay below 2^32
    // (checked when the context is created), and each variable has its own
    // base pointer.
    constexpr int CF_ITERS = (CF_BOX + CF_THREADS - 1) / CF_THREADS;
    const double* __restrict__ q1 = q + P;
    const double* __restrict__ q2 = q + 2 * P;
    const double* __restrict__ q3 = q + 3 * P;
    const double* __restrict__ q4 = q + 4 * P;
    const unsigned p0 = (unsigned)g.P[0], p01 = (unsigned)g.P[0] * (unsigned)g.P[1];
    // box point (0,0,0) is interior point (i0-2, j0-2, k0-2): padded index o - 2
    const int ib = i0 + g.o[0] - 2, jb = j0 + g.o[1] - 2, kb = k0 + g.o[2] - 2;
    const int imax = n0 + 1 + g.o[0], jmax = n1 + 1 + g.o[1], kmax = n2 + 1 + g.o[2];
#pragma unroll
    for (int it = 0; it < CF_ITERS; ++it) {
        const int b = threadIdx.x + it * CF_THREADS;
        if (CF_BOX % CF_THREADS != 0 && b >= CF_BOX) break;
        const int bx = b % CF_BX, by = (b / CF_BX) % CF_BY, bz = b / (CF_BX * CF_BY);
        const unsigned x = (unsigned)min(ib + bx, imax) + p0 * (unsigned)min(jb + by, jmax) + p01 * (unsigned)min(kb + bz, kmax);
        const double r = q[x];
        const double rm0 = q1[x], rm1 = q2[x], rm2 = q3[x];
        const double rhoE = q4[x];
        const double ri = rcp_nr(r);
        const double u0 = rm0 * ri, u1 = rm1 * ri, u2 = rm2 * ri;
        const double p = c.gamma_m1 * fma(-0.5, fma(rm0, u0, fma(rm1, u1, rm2 * u2)), rhoE);
        const double th = p * ri;
        s_f[b] = u0;
        s_f[CF_BOX + b] = u1;
        s_f[2 * CF_BOX + b] = u2;
        s_f[F_TH * CF_BOX + b] = th;
        if (HAS_CONV) s_f[F_RHO * CF_BOX + b] = r;
        if (SUTH) {
            const double TT = c.cT * th;
            s_f[F_MU * CF_BOX + b] = 1.4042 * TT * sqrt(TT) * rcp_nr(0.40417 + TT);
        }
    }
    __syncthreads();

    // ---- phase 2: right-hand side and RK update at one point --------------
    // Every stencil value comes from LDS: the momenta rho * u_m, rhoE = rho E
    // and p = rho theta are rebuilt at the stencil point.  The reads share one base
    // address (the box point at offset -2, -2, -2 from the thread's point), so
    // each is that base plus a constant.
    if (!active) return;
    constexpr int ls[3] = {1, CF_BX, CF_BX * CF_BY};
    constexpr int ORG = 2 * (ls[0] + ls[1] + ls[2]);
    const double* __restrict__ sb = s_f + (tx + CF_BX * (ty + CF_BY * tz));     // box index of the point, less ORG
    auto S = [&](int f, int off) { return sb[f * CF_BOX + ORG + off]; };
    // E or H at a stencil point from theta and the velocities
    auto EH = [&](double coef, int off) {
        const double a = S(0, off), b = S(1, off), d = S(2, off);
        return fma(coef, S(F_TH, off), 0.5 * fma(a, a, fma(b, b, d * d)));
    };

    double Rv[NV];
    if (FLUXRES) {
        const long nvP = (long)NV * P;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double* wf = F + d * nvP + (long)v * P;
                acc -= (wf[x] - wf[x - g.s[d]]) * c.invd[d];
            }
            Rv[v] = acc;
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) Rv[v] = 0.0;
    }

    // Phase 2 runs in sections: one per direction jd, which reads only the
    // fields along the jd line through the point (the gradient column
    // G[.][jd], the jd part of the convective terms, of the pressure gradient,
    // of the Laplacians and of the heat flux), then one per momentum component
    // for the mixed second derivatives of the viscous stress.
    constexpr int FR = HAS_CONV ? F_RHO : 0;
    constexpr int FM = SUTH ? F_MU : 0;
    double G[3][3];          // G[m][jd] = D_jd u_m
    double Lm[3] = {0.0, 0.0, 0.0};      // sum_jd c_jd D_jd^2 u_m, c = 4/3 along m, else 1
    double heat = 0.0, dth[3], dmu[3];
    double div_rhou = 0.0, Rrho = 0.0, eterm = 0.0, dp[3];
    double conv[3] = {0.0, 0.0, 0.0};
    const double rho = HAS_CONV ? S(FR, 0) : 0.0;
#pragma unroll
    for (int jd = 0; jd < 3; ++jd) {
        const int lj = ls[jd];
        const double(&w)[2] = c.w1[jd];
#pragma unroll
        for (int m = 0; m < 3; ++m) G[m][jd] = fd1([&](int s) { return S(m, s * lj); }, w);
        if (VISC != 0) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double d2 = fd2([&](int s) { return S(m, s * lj); }, c.w2[jd]);
                Lm[m] += jd == m ? (4.0 / 3.0) * d2 : d2;
            }
            heat += fd2([&](int s) { return S(F_TH, s * lj); }, c.w2[jd]);
            if (SUTH) {
                dmu[jd] = fd1([&](int s) { return S(FM, s * lj); }, w);
                dth[jd] = fd1([&](int s) { return S(F_TH, s * lj); }, w);
            }
        }
        if (HAS_CONV) {
            auto rs = [&](int s) { return S(FR, s * lj); };
            auto uj = [&](int s) { return S(jd, s * lj); };
            auto ruj = [&](int s) { return rs(s) * uj(s); };
            auto um_ = [&](int m, int s) { return S(m, s * lj); };
            dp[jd] = fd1([&](int s) { return rs(s) * S(F_TH, s * lj); }, w);
            if (CONV == CF_FEIEREISEN) {
                auto Es = [&](int s) { return EH(c.cE, s * lj); };
                const double d_rhouj = fd1(ruj, w);
                div_rhou += d_rhouj;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double d_prod = fd1([&](int s) { return ruj(s) * um_(m, s); }, w);
                    conv[m] += d_prod + ruj(0) * G[m][jd];
                }
                const double d_Eu = fd1([&](int s) { return ruj(s) * Es(s); }, w);
                const double d_E = fd1(Es, w);
                const double d_pu = fd1([&](int s) { return ruj(s) * S(F_TH, s * lj); }, w);
                eterm += -0.5 * (d_Eu + ruj(0) * d_E) - d_pu;
            } else if (CONV == CF_KGP) {
                // KGP, enthalpy splitting, all coefficients 1/4 (split_forms.py:178-234)
                const double ujc = uj(0), rhoujc = ruj(0);
                const double gjj = G[jd][jd];
                auto Hs = [&](int s) { return EH(c.cH, s * lj); };
                const double d_rhouj = fd1(ruj, w);
                const double d_rho = fd1(rs, w);
                Rrho += 0.5 * d_rhouj + 0.5 * (ujc * d_rho + rho * gjj);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double umc = um_(m, 0), rhoumc = rho * umc;
                    const double d_rhouj_um = fd1([&](int s) { return um_(m, s) * ruj(s); }, w);
                    const double d_rhoum = fd1([&](int s) { return rs(s) * um_(m, s); }, w);
                    const double d_ujum = fd1([&](int s) { return uj(s) * um_(m, s); }, w);
                    conv[m] += 0.25 * ((((((d_rhouj_um + umc * d_rhouj) + rhoujc * G[m][jd]) + ujc * d_rhoum) + rhoumc * gjj) +
                                        rho * d_ujum) + (umc * ujc) * d_rho);
                }
                const double Hc = Hs(0);
                const double d_Hrhouj = fd1([&](int s) { return Hs(s) * ruj(s); }, w);
                const double d_H = fd1(Hs, w);
                const double d_Hrho = fd1([&](int s) { return Hs(s) * rs(s); }, w);
                const double d_Huj = fd1([&](int s) { return Hs(s) * uj(s); }, w);
                eterm += 0.25 * ((((((d_Hrhouj + Hc * d_rhouj) + rhoujc * d_H) + ujc * d_Hrho) + (rho * Hc) * gjj) + rho * d_Huj) +
                                 (ujc * Hc) * d_rho);
            } else {
                // skew-symmetric form (parsing.py:76-112)
                auto rEs = [&](int s) { return rs(s) * EH(c.cE, s * lj); };
                const double ujc = uj(0);
                const double d_rhou = fd1(ruj, w);
                const double d_rho = fd1(rs, w);
                const double d_u = G[jd][jd];
                Rrho -= 0.5 * (d_rhou + ujc * d_rho + rho * d_u);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double d_prod = fd1([&](int s) { return ruj(s) * um_(m, s); }, w);
                    const double d_rm = fd1([&](int s) { return rs(s) * um_(m, s); }, w);
                    conv[m] += 0.5 * (d_prod + ujc * d_rm + (rho * um_(m, 0)) * d_u);
                }
                const double d_Eu = fd1([&](int s) { return rEs(s) * uj(s); }, w);
                const double d_rE = fd1(rEs, w);
                const double d_pu = fd1([&](int s) { return ruj(s) * S(F_TH, s * lj); }, w);
                eterm += -0.5 * (d_Eu + ujc * d_rE + rEs(0) * d_u) - d_pu;
            }
        }
        cf_section();
    }

    if (HAS_CONV) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (CONV == CF_FEIEREISEN) Rv[1 + m] = Rv[1 + m] - dp[m] - 0.5 * (conv[m] + S(m, 0) * div_rhou);
            else Rv[1 + m] = Rv[1 + m] - dp[m] - conv[m];
        }
        if (CONV == CF_FEIEREISEN) {
            Rv[0] -= div_rhou;
            Rv[4] += eterm - 0.5 * EH(c.cE, 0) * div_rhou;
        } else if (CONV == CF_KGP) {
            Rv[0] -= Rrho;
            Rv[4] -= eterm;
        } else {
            Rv[0] += Rrho;
            Rv[4] += eterm;
        }
    }

    if (VISC != 0) {
        // momentum: mu (lap u_m + 1/3 D_m div u) (+ S_mj D_j mu), energy:
        // u_m mom_m + mu S_mj G_mj, heat: mu lap T (+ D_j T D_j mu); T = cT theta
        const double muv = SUTH ? S(FM, 0) : 1.0;
        const double div = G[0][0] + G[1][1] + G[2][2];
        heat *= muv;
        if (SUTH) {
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) heat += dth[jd] * dmu[jd];
        }
        double Eacc = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            // the mixed derivatives D_m D_jd u_jd, jd != m, in the (a, b) = sorted (jd, m) plane
            double mixed = 0.0;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                if (jd == m) continue;
                const int a = jd < m ? jd : m, b = jd < m ? m : jd;
                const int la = ls[a], lb = ls[b];
                mixed += fd11([&](int r, int s) { return S(jd, r * la + s * lb); }, c.wm[cf_pair(a, b)]);
            }
            double mom = muv * fma(1.0 / 3.0, mixed, Lm[m]);
            double e = 0.0;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                const double Smj = G[m][jd] + G[jd][m] - (m == jd ? (2.0 / 3.0) * div : 0.0);
                if (SUTH) mom += Smj * dmu[jd];
                e += Smj * G[m][jd];
            }
            Rv[1 + m] += c.invRe * mom;
            Eacc += fma(S(m, 0), mom, muv * e);
            cf_section();
        }
        Rv[4] += c.invRe * Eacc + c.cheat_th * heat;
    }

    const double dt = scal[1];
    double qv[NV];
    if (TIME == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double d = fma(A, r_pt[v], dt * Rv[v]);
            reg[y] = d;
            qv[v] = fma(B, d, q_pt[v]);
            qout[y] = qv[v];
        }
    } else {
        // classic RK3: A = rknew, B = rkold; reg holds q_old
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double o = r_pt[v];
            qv[v] = fma(dt * A, Rv[v], o);
            qout[y] = qv[v];
            reg[y] = fma(dt * B, Rv[v], o);
        }
    }
    // every face periodic: the point's wrapped images in the halo of qout
    if (wrap) store_halo_images<NV>(qout, g, i, j, k, qv);
}

template <int CF_TX, int CF_TY, int CF_TZ, int CONV, int VISC, bool FLUXRES, int TIME>
__global__ void __launch_bounds__(CF_TX * CF_TY * CF_TZ, cf_waves(CF_TX, CF_TY, CF_TZ, cf_nfield(CONV, VISC)))
k_central_fused(const double* __restrict__ q, const double* __restrict__ F, double* __restrict__ reg, double* __restrict__ qout,
                const double* __restrict__ scal, Geom g, Consts c, double A, double B, CfOrder o, int wrap) {
    // one tile per workgroup, in the XCD-aware order o (a 1-D grid)
    int tbx, tby, tbz;
    cf_tile_of(o, blockIdx.x, tbx, tby, tbz);
    cf_tile_stage<CF_TX, CF_TY, CF_TZ, CONV, VISC, FLUXRES, TIME>(tbx * CF_TX, tby * CF_TY, tbz * CF_TZ, q, F, reg, qout, scal, g,
                                                                  c, A, B, wrap);
}

}  // namespace osb
