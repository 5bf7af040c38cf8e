// osb_central_fused.hpp -- one-launch RK stage for the central schemes on gfx950.
//
// One workgroup owns a 16 x 4 x 4 tile of grid points.  Phase 1 stages the
// constituent relations of the tile plus a 2-point halo box in LDS:
//   u0,u1,u2                at tile, faces and edges of the box (the edges feed
//                           the mixed second derivatives D_b(D_a u))
//   p, T, E=rhoE/rho, mu    at tile and faces (single-direction stencils);
//                           the KGP form stores H=(rhoE+p)/rho in place of E
// each evaluated once per box point.  Phase 2 gives every point its full
// right-hand side -- split-form convective terms, viscous stresses and heat
// conduction, or the interface-flux difference of a shock-capturing scheme --
// followed directly by the RK update, so the stage reads q, du (and the face
// fluxes) and writes du and q once.  The expression sequence matches the CPU
// oracle (StoreSome.py:138-228, split_forms.py:111-144, 383-392).
#pragma once

#include "osb_kernels.hpp"

namespace osb {

// tile shapes offered to the autotuner: {TX, TY, TZ}
struct CfTile { int tx, ty, tz; };
constexpr CfTile CF_TILES[] = {{16, 4, 4}, {8, 8, 4}, {32, 4, 2}, {16, 8, 4}, {32, 4, 4}};
constexpr int CF_NTILES = 5;

enum CfConv { CF_NONE = 0, CF_FEIEREISEN = 1, CF_SKEW = 2, CF_KGP = 3 };

// CONV: CfConv; VISC: 0 none, 1 constant, 2 Sutherland; FLUXRES: start R from
// the face-flux difference (shock-capturing schemes); TIME: OSB_TIME_RKLS3 (1)
// or OSB_TIME_RK3 (2).
template <int CF_TX, int CF_TY, int CF_TZ, int CONV, int VISC, bool FLUXRES, int TIME>
__global__ void __launch_bounds__(CF_TX * CF_TY * CF_TZ) k_central_fused(const double* __restrict__ q, const double* __restrict__ F,
                                                              double* __restrict__ reg, double* __restrict__ qout,
                                                              const double* __restrict__ scal, Geom g, Consts c,
                                                              double A, double B, int xcd_remap, int wrap) {
    constexpr int NV = 5;
    constexpr int CF_BX = CF_TX + 4, CF_BY = CF_TY + 4, CF_BZ = CF_TZ + 4;
    constexpr int CF_BOX = CF_BX * CF_BY * CF_BZ;
    constexpr int CF_THREADS = CF_TX * CF_TY * CF_TZ;
    __shared__ double s_u[3][CF_BOX];
    __shared__ double s_p[CF_BOX];
    __shared__ double s_T[CF_BOX];
    __shared__ double s_E[CF_BOX];
    __shared__ double s_mu[VISC == 2 ? CF_BOX : 1];

    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
    // workgroup w runs on XCD w % 8.  Giving XCD x the contiguous range of
    // tiles [x*T/8, (x+1)*T/8) (x fastest, then y, then z) keeps neighbouring
    // tiles -- which share halo points -- in the same L2.
    const int ntx = gridDim.x, nty = gridDim.y;
    const int ntiles = ntx * nty * gridDim.z;
    const int w = blockIdx.x + ntx * (blockIdx.y + nty * blockIdx.z);
    int tile = w;
    if (xcd_remap && (ntiles & 7) == 0) tile = (w & 7) * (ntiles >> 3) + (w >> 3);
    int tbx = tile % ntx, tby = (tile / ntx) % nty, tbz = tile / (ntx * nty);
    if (xcd_remap == 2 && (ntx & 7) == 0 && (nty & 3) == 0 && (gridDim.z & 1) == 0) {
        // supertiles of 8 x 4 x 2 tiles (one XCD's 64 resident workgroups) walked in
        // x, y, z order: a squarer resident region, so less of its halo leaves L2
        const int in = tile & 63, sup = tile >> 6;
        const int sx = ntx >> 3, sy = nty >> 2;
        tbx = ((sup % sx) << 3) + (in & 7);
        tby = (((sup / sx) % sy) << 2) + ((in >> 3) & 3);
        tbz = ((sup / (sx * sy)) << 1) + (in >> 5);
    }
    const int i0 = tbx * CF_TX, j0 = tby * CF_TY, k0 = tbz * CF_TZ;
    const long P = g.Ptot;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];

    // ---- phase 1: constituent relations on the halo box -------------------
    // Unrolled over the box points of a thread; the loads are unconditional
    // (clamped into the stored range) so the iterations can issue them
    // together, and only the stores keep the skips.
    constexpr int CF_ITERS = (CF_BOX + CF_THREADS - 1) / CF_THREADS;
#pragma unroll
    for (int it = 0; it < CF_ITERS; ++it) {
        const int b = threadIdx.x + it * CF_THREADS;
        if (CF_BOX % CF_THREADS != 0 && b >= CF_BOX) break;
        const int bx = b % CF_BX, by = (b / CF_BX) % CF_BY, bz = b / (CF_BX * CF_BY);
        const int i = i0 + bx - 2, j = j0 + by - 2, k = k0 + bz - 2;
        const int out = (bx < 2 || bx >= CF_TX + 2) + (by < 2 || by >= CF_TY + 2) + (bz < 2 || bz >= CF_TZ + 2);
        // box corners are never read; points beyond the stored halo are not stored
        const bool keep = out != 3 && i < n0 + 2 && j < n1 + 2 && k < n2 + 2;
        const long x = gidx(g, min(i, n0 + 1), min(j, n1 + 1), min(k, n2 + 1));
        const double r = q[x];
        const double rm[3] = {q[P + x], q[2 * P + x], q[3 * P + x]};
        const double rhoE = q[4 * P + x];
        double ke = 0.0;
        double uu[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            uu[d] = rm[d] / r;
            ke -= 0.5 * (uu[d] * uu[d]) * r;
        }
        if (keep) {
#pragma unroll
            for (int d = 0; d < 3; ++d) s_u[d][b] = uu[d];
            if (out <= 1) {
                const double p = c.gamma_m1 * (ke + rhoE);
                const double TT = (c.minf * c.minf) * c.gama * p / r;
                s_p[b] = p;
                s_T[b] = TT;
                s_E[b] = CONV == CF_KGP ? (rhoE + p) / r : rhoE / r;
                if (VISC == 2) s_mu[VISC == 2 ? b : 0] = 1.4042 * TT * sqrt(TT) / (0.40417 + TT);
            }
        }
    }
    __syncthreads();

    // ---- phase 2: right-hand side and RK update at one point --------------
    const int tx = threadIdx.x % CF_TX, ty = (threadIdx.x / CF_TX) % CF_TY, tz = threadIdx.x / (CF_TX * CF_TY);
    const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
    if (i >= n0 || j >= n1 || k >= n2) return;
    const long x = gidx(g, i, j, k);
    const int l = (tx + 2) + CF_BX * ((ty + 2) + CF_BY * (tz + 2));
    const int ls[3] = {1, CF_BX, CF_BX * CF_BY};

    double Rv[NV];
    if (FLUXRES) {
        const long nvP = (long)NV * P;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double* w = F + d * nvP + (long)v * P;
                acc -= (w[x] - w[x - g.s[d]]) * c.invd[d];
            }
            Rv[v] = acc;
        }
    }

    // stored-gradient values G[m][jd] = D_jd u_m at the point
    double G[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int jd = 0; jd < 3; ++jd)
            G[m][jd] = d1([&](int s) { return s_u[m][l + s * ls[jd]]; }, c.invd[jd]);

    if (CONV != CF_NONE) {
        const double rho = q[x], rhoE = q[4 * P + x];
        double div_rhou = 0.0, Rrho = 0.0, eterm = 0.0;
        double conv[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            const long st = g.s[jd];
            const int lj = ls[jd];
            const double inv = c.invd[jd];
            const double* rhou_j = q + (long)(1 + jd) * P;
            auto qat = [&](const double* a, int s) { return a[x + s * st]; };
            auto uj = [&](int s) { return s_u[jd][l + s * lj]; };
            if (CONV == CF_FEIEREISEN) {
                const double d_rhouj = d1([&](int s) { return qat(rhou_j, s); }, inv);
                div_rhou += d_rhouj;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double* rhou_m = q + (long)(1 + m) * P;
                    const double d_prod = d1([&](int s) { return qat(rhou_m, s) * uj(s); }, inv);
                    conv[m] += d_prod + qat(rhou_j, 0) * G[m][jd];
                }
                const double d_Eu = d1([&](int s) { return qat(q + 4 * P, s) * uj(s); }, inv);
                const double d_E = d1([&](int s) { return s_E[l + s * lj]; }, inv);
                const double d_pu = d1([&](int s) { return s_p[l + s * lj] * uj(s); }, inv);
                eterm += -0.5 * (d_Eu + qat(rhou_j, 0) * d_E) - d_pu;
            } else if (CONV == CF_KGP) {
                // KGP, enthalpy splitting, all coefficients 1/4 (split_forms.py:178-234): the term
                // sequence of the CPU oracle, with u and H = (rhoE + p)/rho from LDS
                const double ujc = uj(0), rhoujc = qat(rhou_j, 0);
                const double gjj = G[jd][jd];
                auto Hs = [&](int s) { return s_E[l + s * lj]; };
                const double d_rhouj = d1([&](int s) { return qat(rhou_j, s); }, inv);
                const double d_rho = d1([&](int s) { return qat(q, s); }, inv);
                Rrho += 0.5 * d_rhouj + 0.5 * (ujc * d_rho + rho * gjj);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double* rhou_m = q + (long)(1 + m) * P;
                    auto um = [&](int s) { return s_u[m][l + s * lj]; };
                    const double umc = um(0), rhoumc = qat(rhou_m, 0);
                    const double d_rhouj_um = d1([&](int s) { return um(s) * qat(rhou_j, s); }, inv);
                    const double d_rhoum = d1([&](int s) { return qat(rhou_m, s); }, inv);
                    const double d_ujum = d1([&](int s) { return uj(s) * um(s); }, inv);
                    conv[m] += 0.25 * ((((((d_rhouj_um + umc * d_rhouj) + rhoujc * G[m][jd]) + ujc * d_rhoum) + rhoumc * gjj) +
                                        rho * d_ujum) + (umc * ujc) * d_rho);
                }
                const double Hc = Hs(0);
                const double d_Hrhouj = d1([&](int s) { return Hs(s) * qat(rhou_j, s); }, inv);
                const double d_H = d1(Hs, inv);
                const double d_Hrho = d1([&](int s) { return Hs(s) * qat(q, s); }, inv);
                const double d_Huj = d1([&](int s) { return Hs(s) * uj(s); }, inv);
                eterm += 0.25 * ((((((d_Hrhouj + Hc * d_rhouj) + rhoujc * d_H) + ujc * d_Hrho) + (rho * Hc) * gjj) + rho * d_Huj) +
                                 (ujc * Hc) * d_rho);
            } else {
                const double ujc = uj(0);
                const double d_rhou = d1([&](int s) { return qat(q, s) * uj(s); }, inv);
                const double d_rho = d1([&](int s) { return qat(q, s); }, inv);
                const double d_u = d1([&](int s) { return uj(s); }, inv);
                Rrho -= 0.5 * (d_rhou + ujc * d_rho + rho * d_u);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double* rhou_m = q + (long)(1 + m) * P;
                    const double d_prod = d1([&](int s) { return qat(rhou_m, s) * uj(s); }, inv);
                    const double d_rm = d1([&](int s) { return qat(rhou_m, s); }, inv);
                    conv[m] += 0.5 * (d_prod + ujc * d_rm + qat(rhou_m, 0) * d_u);
                }
                const double d_Eu = d1([&](int s) { return qat(q + 4 * P, s) * uj(s); }, inv);
                const double d_rE = d1([&](int s) { return qat(q + 4 * P, s); }, inv);
                const double d_pu = d1([&](int s) { return s_p[l + s * lj] * uj(s); }, inv);
                eterm += -0.5 * (d_Eu + ujc * d_rE + rhoE * d_u) - d_pu;
            }
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int lm = ls[m];
            const double dp = d1([&](int s) { return s_p[l + s * lm]; }, c.invd[m]);
            if (CONV == CF_FEIEREISEN) Rv[1 + m] = -dp - 0.5 * (conv[m] + s_u[m][l] * div_rhou);
            else Rv[1 + m] = -dp - conv[m];
        }
        if (CONV == CF_FEIEREISEN) {
            Rv[0] = -div_rhou;
            Rv[4] = eterm - 0.5 * (rhoE / rho) * div_rhou;
        } else if (CONV == CF_KGP) {
            Rv[0] = -Rrho;
            Rv[4] = -eterm;
        } else {
            Rv[0] = Rrho;
            Rv[4] = eterm;
        }
    }

    if (VISC != 0) {
        constexpr bool SUTH = VISC == 2;
        double D2u[3][3], dmu[3], dT[3];
        double lapT = 0.0;
#pragma unroll
        for (int jd = 0; jd < 3; ++jd) {
            const int lj = ls[jd];
            dmu[jd] = SUTH ? d1([&](int s) { return s_mu[SUTH ? l + s * lj : 0]; }, c.invd[jd]) : 0.0;
            dT[jd] = d1([&](int s) { return s_T[l + s * lj]; }, c.invd[jd]);
            lapT += d2([&](int s) { return s_T[l + s * lj]; }, c.inv2d[jd]);
#pragma unroll
            for (int m = 0; m < 3; ++m) D2u[m][jd] = d2([&](int s) { return s_u[m][l + s * lj]; }, c.inv2d[jd]);
        }
        double div = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) div += G[m][m];
        const double muv = SUTH ? s_mu[SUTH ? l : 0] : 1.0;
        double Eacc = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double Lm = (4.0 / 3.0) * D2u[m][m];
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                if (jd == m) continue;
                const int a = jd < m ? jd : m, b = jd < m ? m : jd;
                // D_b(D_a u_jd): the inner derivative at the four points along b
                const int la = ls[a], lb = ls[b];
                auto ga = [&](int s) {
                    const int lc = l + s * lb;
                    return d1([&](int r) { return s_u[jd][lc + r * la]; }, c.invd[a]);
                };
                const double mixed = d1(ga, c.invd[b]);
                Lm += D2u[m][jd] + (1.0 / 3.0) * mixed;
            }
            double S[3];
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) S[jd] = G[m][jd] + G[jd][m] - (m == jd ? (2.0 / 3.0) * div : 0.0);
            double mom = muv * Lm;
            if (SUTH) {
#pragma unroll
                for (int jd = 0; jd < 3; ++jd) mom += S[jd] * dmu[jd];
            }
            Rv[1 + m] += c.invRe * mom;
            double e = s_u[m][l] * mom;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) e += muv * S[jd] * G[m][jd];
            Eacc += e;
        }
        double heat = muv * lapT;
        if (SUTH) {
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) heat += dT[jd] * dmu[jd];
        }
        Rv[4] += c.invRe * Eacc + c.cheat * heat;
    }

    const double dt = scal[1];
    double qv[NV];
    if (TIME == 1) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double d = fma(A, reg[y], dt * Rv[v]);
            reg[y] = d;
            qv[v] = fma(B, d, q[y]);
            qout[y] = qv[v];
        }
    } else {
        // classic RK3: A = rknew, B = rkold; reg holds q_old
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const long y = (long)v * P + x;
            const double o = reg[y];
            qv[v] = fma(dt * A, Rv[v], o);
            qout[y] = qv[v];
            reg[y] = fma(dt * B, Rv[v], o);
        }
    }
    // every face periodic: the point's wrapped images in the halo of qout
    if (wrap) store_halo_images<NV>(qout, g, i, j, k, qv);
}

}  // namespace osb
