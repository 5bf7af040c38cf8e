// osb_central_mixed.hpp -- mixed-precision RK stage of the 3-D central path.
//
// Same tiling and phases as k_central_fused, with the derivative work in
// FP32: every field staged in LDS is stored as the FP32 deviation from a
// reference value (the FP64 state and primitives of the tile's first point),
// so FP32 holds only the variation across the tile.  Derivatives of a field
// are derivatives of its deviation (the stencil weights sum to zero); a
// derivative of a product a*b expands exactly as
//   D(ab) = a_ref D(b') + b_ref D(a') + D(a'b').
// The constituent relations are evaluated in FP64, the right-hand side is
// accumulated in FP32, and the RK update (du, q) is FP64.
#pragma once

#include "osb_kernels.hpp"
#include "osb_central_fused.hpp"

namespace osb {

template <class A>
__device__ __forceinline__ float d1f(A f, float invd) {
    return fmaf(1.0f / 12.0f, f(-2) - f(2), (2.0f / 3.0f) * (f(1) - f(-1))) * invd;
}
template <class A>
__device__ __forceinline__ float d2f(A f, float inv2d) {
    return fmaf(-1.0f / 12.0f, f(-2) + f(2), fmaf(4.0f / 3.0f, f(-1) + f(1), (-5.0f / 2.0f) * f(0))) * inv2d;
}

// CONV: CF_FEIEREISEN, CF_KGP (H = (rhoE + p)/rho staged in place of E and
// rho in place of rhoE) or CF_NONE (with FLUXRES: the shock-capturing flux
// difference, kept in FP64).  VISC: 0, 1 constant, 2 Sutherland.
template <int CF_TX, int CF_TY, int CF_TZ, int CONV, int VISC, bool FLUXRES, int TIME>
__global__ void __launch_bounds__(CF_TX * CF_TY * CF_TZ) k_central_mixed(const double* __restrict__ q, const double* __restrict__ F,
                                                              double* __restrict__ reg, double* __restrict__ qout,
                                                              const double* __restrict__ scal, Geom g, Consts c,
                                                              double A, double B, CfOrder o, int wrap) {
    static_assert(CONV == CF_FEIEREISEN || CONV == CF_KGP || CONV == CF_NONE, "mixed central kernel: Feiereisen, KGP or viscous-only");
    constexpr bool KGP = CONV == CF_KGP;
    constexpr int NV = 5;
    constexpr int CF_BX = CF_TX + 4, CF_BY = CF_TY + 4, CF_BZ = CF_TZ + 4;
    constexpr int CF_BOX = CF_BX * CF_BY * CF_BZ;
    constexpr int CF_THREADS = CF_TX * CF_TY * CF_TZ;
    constexpr bool HAS_CONV = CONV != CF_NONE;
    constexpr bool SUTH = VISC == 2;
    __shared__ float s_u[3][CF_BOX];
    __shared__ float s_T[VISC ? CF_BOX : 1];
    __shared__ float s_mu[SUTH ? CF_BOX : 1];
    __shared__ float s_p[HAS_CONV ? CF_BOX : 1];
    __shared__ float s_E[HAS_CONV ? CF_BOX : 1];
    __shared__ float s_r[HAS_CONV ? 4 : 1][HAS_CONV ? CF_BOX : 1];    // rho u0..2, rhoE (KGP: rho)

    // one tile per workgroup, in the XCD-aware order o (as k_central_fused)
    int tbx, tby, tbz;
    cf_tile_of(o, blockIdx.x, tbx, tby, tbz);
    const int i0 = tbx * CF_TX, j0 = tby * CF_TY, k0 = tbz * CF_TZ;
    const long P = g.Ptot;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];

    // RK operands of the thread's point, loaded first so that they arrive
    // while phases 1 and 2 run
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

    // ---- phase 1: constituent relations (FP64) -> FP32 deviations in LDS ----
    // The loads come first: the reference point's, then every box point's
    // (all in flight together), then the reference state and the box points'
    // relations.  Every box point is evaluated and stored (the corners, and
    // the scalar fields on the edges, are never read), with the loads clamped
    // into the stored range as in k_central_fused.
    constexpr int CF_ITERS = (CF_BOX + CF_THREADS - 1) / CF_THREADS;
    const double* __restrict__ qvar[NV] = {q, q + P, q + 2 * P, q + 3 * P, q + 4 * P};
    const unsigned p0 = (unsigned)g.P[0], p01 = (unsigned)g.P[0] * (unsigned)g.P[1];
    const int ib = i0 + g.o[0] - 2, jb = j0 + g.o[1] - 2, kb = k0 + g.o[2] - 2;
    const int imax = n0 + 1 + g.o[0], jmax = n1 + 1 + g.o[1], kmax = n2 + 1 + g.o[2];
    // reference state: the tile's first point (in the interior, identical in every thread)
    double refq[NV];
    {
        const unsigned xr = (unsigned)(ib + 2) + p0 * (unsigned)(jb + 2) + p01 * (unsigned)(kb + 2);
#pragma unroll
        for (int v = 0; v < NV; ++v) refq[v] = qvar[v][xr];
    }
    double bq[CF_ITERS][NV];
#pragma unroll
    for (int it = 0; it < CF_ITERS; ++it) {
        const int b = threadIdx.x + it * CF_THREADS;
        if (CF_BOX % CF_THREADS == 0 || b < CF_BOX) {
            const int bx = b % CF_BX, by = (b / CF_BX) % CF_BY, bz = b / (CF_BX * CF_BY);
            const unsigned xb = (unsigned)min(ib + bx, imax) + p0 * (unsigned)min(jb + by, jmax) + p01 * (unsigned)min(kb + bz, kmax);
#pragma unroll
            for (int v = 0; v < NV; ++v) bq[it][v] = qvar[v][xb];
        }
    }
    // The deviations' algebra is exact for any reference values (a derivative
    // of a' = a - a_ref is that of a, and the product expansion holds for any
    // a_ref, b_ref), so the reference state only has to be close to the
    // tile's values: it is formed with the refined hardware reciprocal.
    double ref_u[3], ref_p, ref_T, ref_E, ref_mu, ref_r[4];
    {
        const double r = refq[0], inv_r = rcp_nr1(r);
        double ke = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ref_r[d] = refq[1 + d];
            ref_u[d] = ref_r[d] * inv_r;
            ke -= 0.5 * (ref_u[d] * ref_u[d]) * r;
        }
        const double rE = refq[4];
        ref_p = c.gamma_m1 * (ke + rE);
        ref_T = (c.minf * c.minf) * c.gama * ref_p * inv_r;
        ref_E = KGP ? (rE + ref_p) * inv_r : rE * inv_r;
        ref_r[3] = KGP ? r : rE;
        ref_mu = SUTH ? 1.4042 * ref_T * sqrt(ref_T) * rcp_nr1(0.40417 + ref_T) : 1.0;
    }
#pragma unroll
    for (int it = 0; it < CF_ITERS; ++it) {
        const int b = threadIdx.x + it * CF_THREADS;
        if (CF_BOX % CF_THREADS != 0 && b >= CF_BOX) break;
        const double r = bq[it][0];
        const double inv_r = rcp_nr(r);
        double ke = 0.0;
        const double rv[3] = {bq[it][1], bq[it][2], bq[it][3]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double u = rv[d] * inv_r;
            s_u[d][b] = float(u - ref_u[d]);
            ke -= 0.5 * (u * u) * r;
        }
        const double rhoE = bq[it][4];
        const double p = c.gamma_m1 * (ke + rhoE);
        if (VISC) {
            const double TT = (c.minf * c.minf) * c.gama * p * inv_r;
            s_T[VISC ? b : 0] = float(TT - ref_T);
            if (SUTH) s_mu[SUTH ? b : 0] = float(1.4042 * TT * sqrt(TT) / (0.40417 + TT) - ref_mu);
        }
        if (HAS_CONV) {
            s_p[HAS_CONV ? b : 0] = float(p - ref_p);
            s_E[HAS_CONV ? b : 0] = float((KGP ? (rhoE + p) * inv_r : rhoE * inv_r) - ref_E);
#pragma unroll
            for (int d = 0; d < 3; ++d) s_r[HAS_CONV ? d : 0][HAS_CONV ? b : 0] = float(rv[d] - ref_r[d]);
            s_r[HAS_CONV ? 3 : 0][HAS_CONV ? b : 0] = float((KGP ? r : rhoE) - ref_r[3]);
        }
    }
    __syncthreads();

    // ---- phase 2: FP32 right-hand side, FP64 RK update --------------------
    // Fences between the phase-2 sections only where they keep the occupancy:
    // with the convective terms the kernel's occupancy is set by its LDS (3
    // workgroups per CU, or 2 with Sutherland's law) and the scheduler may use
    // the VGPRs that leaves (105 -> 154; the Feiereisen TGV stage 108.7 ->
    // 104 us); the viscous-only kernel (4 workgroups per CU) keeps them.
    auto phase2_fence = [] {
        if (!HAS_CONV) cf_section();
    };
    if (!active) return;
    const int l = (tx + 2) + CF_BX * ((ty + 2) + CF_BY * (tz + 2));
    const int ls[3] = {1, CF_BX, CF_BX * CF_BY};
    const float invd[3] = {float(c.invd[0]), float(c.invd[1]), float(c.invd[2])};
    const float inv2d[3] = {float(c.inv2d[0]), float(c.inv2d[1]), float(c.inv2d[2])};

    double Rd[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};       // FP64 part (face-flux difference)
    if (FLUXRES) {
        const long nvP = (long)NV * P;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double* wv = F + d * nvP + (long)v * P;
                acc -= (wv[x] - wv[x - g.s[d]]) * c.invd[d];
            }
            Rd[v] = acc;
        }
    }
    float Rv[NV] = {0.f, 0.f, 0.f, 0.f, 0.f};

    // point values (full) and gradients of u
    float uc[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) uc[m] = float(ref_u[m]) + s_u[m][l];
    float G[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) G[m][jd] = d1f([&](int s) { return s_u[m][l + s * ls[jd]]; }, invd[jd]);
    phase2_fence();

    if (KGP) {
        // KGP, enthalpy splitting, coefficients 1/4 (split_forms.py:178-234) on deviations:
        // D(ab) = a_ref D(b') + b_ref D(a') + D(a'b'); s_E holds H', s_r[3] rho'
        float Rrho = 0.f, eterm = 0.f;
        float conv[3] = {0.f, 0.f, 0.f};
        float rc[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) rc[v] = float(ref_r[v]) + s_r[HAS_CONV ? v : 0][HAS_CONV ? l : 0];
        const float rhoc = rc[3], Hc = float(ref_E) + s_E[HAS_CONV ? l : 0];
        const float rref = float(ref_r[3]), Href = float(ref_E);
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            const int lj = ls[jd];
            const float inv = invd[jd];
            const float urj = float(ref_u[jd]), rurj = float(ref_r[jd]);
            auto ud = [&](int s) { return s_u[jd][l + s * lj]; };
            auto rd = [&](int v, int s) { return s_r[HAS_CONV ? v : 0][HAS_CONV ? l + s * lj : 0]; };
            auto hd = [&](int s) { return s_E[HAS_CONV ? l + s * lj : 0]; };
            const float ujc = uc[jd], rhoujc = rc[jd];
            const float gjj = G[jd][jd];
            const float d_rhouj = d1f([&](int s) { return rd(jd, s); }, inv);
            const float d_rho = d1f([&](int s) { return rd(3, s); }, inv);
            Rrho += 0.5f * d_rhouj + 0.5f * (ujc * d_rho + rhoc * gjj);
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                auto umd = [&](int s) { return s_u[m][l + s * lj]; };
                const float urm = float(ref_u[m]);
                const float umc = uc[m], rhoumc = rc[m];
                const float d_um = G[m][jd];
                const float d_rhouj_um = urm * d_rhouj + rurj * d_um + d1f([&](int s) { return umd(s) * rd(jd, s); }, inv);
                const float d_rhoum = d1f([&](int s) { return rd(m, s); }, inv);
                const float d_ujum = urj * d_um + urm * gjj + d1f([&](int s) { return ud(s) * umd(s); }, inv);
                conv[m] += 0.25f * ((((((d_rhouj_um + umc * d_rhouj) + rhoujc * G[m][jd]) + ujc * d_rhoum) + rhoumc * gjj) +
                                     rhoc * d_ujum) + (umc * ujc) * d_rho);
            }
            const float d_H = d1f(hd, inv);
            const float d_Hrhouj = Href * d_rhouj + rurj * d_H + d1f([&](int s) { return hd(s) * rd(jd, s); }, inv);
            const float d_Hrho = Href * d_rho + rref * d_H + d1f([&](int s) { return hd(s) * rd(3, s); }, inv);
            const float d_Huj = Href * gjj + urj * d_H + d1f([&](int s) { return hd(s) * ud(s); }, inv);
            eterm += 0.25f * ((((((d_Hrhouj + Hc * d_rhouj) + rhoujc * d_H) + ujc * d_Hrho) + (rhoc * Hc) * gjj) + rhoc * d_Huj) +
                              (ujc * Hc) * d_rho);
            phase2_fence();
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int lm = ls[m];
            const float dp = d1f([&](int s) { return s_p[HAS_CONV ? l + s * lm : 0]; }, invd[m]);
            Rv[1 + m] = -dp - conv[m];
        }
        Rv[0] = -Rrho;
        Rv[4] = -eterm;
    } else if (HAS_CONV) {
        const float* su = nullptr;
        (void)su;
        float div_rhou = 0.f, eterm = 0.f;
        float conv[3] = {0.f, 0.f, 0.f};
        float rc[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) rc[v] = float(ref_r[v]) + s_r[HAS_CONV ? v : 0][HAS_CONV ? l : 0];
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            const int lj = ls[jd];
            const float inv = invd[jd];
            const float urj = float(ref_u[jd]);
            auto ud = [&](int s) { return s_u[jd][l + s * lj]; };                    // u_j deviation
            auto rd = [&](int v, int s) { return s_r[HAS_CONV ? v : 0][HAS_CONV ? l + s * lj : 0]; };
            const float d_rhouj = d1f([&](int s) { return rd(jd, s); }, inv);
            div_rhou += d_rhouj;
            const float d_uj = G[jd][jd];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                // D(rho u_m u_j) = ref(rho u_m) D(u_j') + ref(u_j) D(rho u_m') + D(rho u_m' u_j')
                const float d_rm = d1f([&](int s) { return rd(m, s); }, inv);
                const float d_cross = d1f([&](int s) { return rd(m, s) * ud(s); }, inv);
                const float d_prod = float(ref_r[m]) * d_uj + urj * d_rm + d_cross;
                conv[m] += d_prod + rc[jd] * G[m][jd];
            }
            const float d_rE = d1f([&](int s) { return rd(3, s); }, inv);
            const float d_Eu = float(ref_r[3]) * d_uj + urj * d_rE + d1f([&](int s) { return rd(3, s) * ud(s); }, inv);
            const float d_E = d1f([&](int s) { return s_E[HAS_CONV ? l + s * lj : 0]; }, inv);
            const float d_pp = d1f([&](int s) { return s_p[HAS_CONV ? l + s * lj : 0]; }, inv);
            const float d_pu = float(ref_p) * d_uj + urj * d_pp + d1f([&](int s) { return s_p[HAS_CONV ? l + s * lj : 0] * ud(s); }, inv);
            eterm += -0.5f * (d_Eu + rc[jd] * d_E) - d_pu;
            phase2_fence();
        }
        const float Ec = float(ref_E) + s_E[HAS_CONV ? l : 0];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int lm = ls[m];
            const float dp = d1f([&](int s) { return s_p[HAS_CONV ? l + s * lm : 0]; }, invd[m]);
            Rv[1 + m] = -dp - 0.5f * (conv[m] + uc[m] * div_rhou);
        }
        Rv[0] = -div_rhou;
        Rv[4] = eterm - 0.5f * Ec * div_rhou;
    }

    if (VISC != 0) {
        float D2u[3][3], dmu[3], dT[3];
        float lapT = 0.f;
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            const int lj = ls[jd];
            dmu[jd] = SUTH ? d1f([&](int s) { return s_mu[SUTH ? l + s * lj : 0]; }, invd[jd]) : 0.f;
            dT[jd] = d1f([&](int s) { return s_T[VISC ? l + s * lj : 0]; }, invd[jd]);
            lapT += d2f([&](int s) { return s_T[VISC ? l + s * lj : 0]; }, inv2d[jd]);
#pragma unroll
            for (int m = 0; m < 3; ++m) D2u[m][jd] = d2f([&](int s) { return s_u[m][l + s * lj]; }, inv2d[jd]);
            phase2_fence();
        }
        float div = 0.f;
#pragma unroll
        for (int m = 0; m < 3; ++m) div += G[m][m];
        const float muv = SUTH ? float(ref_mu) + s_mu[SUTH ? l : 0] : 1.f;
        float Eacc = 0.f;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float Lm = (4.0f / 3.0f) * D2u[m][m];
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                if (jd == m) continue;
                const int a = jd < m ? jd : m, b = jd < m ? m : jd;
                const int la = ls[a], lb = ls[b];
                auto ga = [&](int s) {
                    const int lc = l + s * lb;
                    return d1f([&](int r) { return s_u[jd][lc + r * la]; }, invd[a]);
                };
                Lm += D2u[m][jd] + (1.0f / 3.0f) * d1f(ga, invd[b]);
            }
            float S[3];
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) S[jd] = G[m][jd] + G[jd][m] - (m == jd ? (2.0f / 3.0f) * div : 0.f);
            float mom = muv * Lm;
            if (SUTH) {
#pragma unroll
                for (int jd = 0; jd < 3; ++jd) mom += S[jd] * dmu[jd];
            }
            Rv[1 + m] += float(c.invRe) * mom;
            float e = uc[m] * mom;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) e += muv * S[jd] * G[m][jd];
            Eacc += e;
            phase2_fence();
        }
        float heat = muv * lapT;
        if (SUTH) {
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) heat += dT[jd] * dmu[jd];
        }
        Rv[4] += float(c.invRe) * Eacc + float(c.cheat) * heat;
    }

    const double dt = scal[1];
    double qv[NV];
    if (TIME == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double R = Rd[v] + double(Rv[v]);
            const double d = fma(A, r_pt[v], dt * R);
            reg[y] = d;
            qv[v] = fma(B, d, q_pt[v]);
            qout[y] = qv[v];
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double R = Rd[v] + double(Rv[v]);
            const double o = r_pt[v];
            qv[v] = fma(dt * A, R, o);
            qout[y] = qv[v];
            reg[y] = fma(dt * B, R, o);
        }
    }
    // every face periodic: the point's wrapped images in the halo of qout
    if (wrap) store_halo_images<NV>(qout, g, i, j, k, qv);
}

}  // namespace osb
