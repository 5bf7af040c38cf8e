// osb_central_march.hpp -- RK stage of the 3-D central schemes as a z-march.
//
// A workgroup owns a TX x TY column of grid points and walks it up a chunk of
// KZ planes.  Each plane of conservative variables is read from HBM once
// (plus a 2-point x/y halo) and its constituent relations are evaluated once:
//   * u0, u1, u2 of the plane with its full x/y halo go to a 3-plane LDS ring
//     (planes z, z+1, z+2): x/y derivatives, the x-y mixed derivatives, and
//     the inner x/y derivatives of the z-mixed terms read it;
//   * each thread keeps its own column's values of the five planes z-2..z+2
//     in registers (rho u_m, rhoE, rho, u_m, p, T, E, mu), so every
//     z-derivative is register arithmetic;
//   * the scalars p, T, E, mu and rho u_m, rhoE of the output plane's x/y halo
//     are staged in a second LDS buffer when the plane is output.
// The right-hand side of every point is then evaluated with the expression
// sequence of the single-tile kernel (k_central_fused), so both round
// identically; the RK update follows in the same pass.
#pragma once

#include "osb_kernels.hpp"
#include "osb_central_fused.hpp"

namespace osb {

// plane index of enumeration entry e: the TX*TY interior points in thread
// order, then the halo ring (two bottom rows, two top rows, then the two left
// and two right points of each interior row)
template <int TX, int TY>
__device__ __forceinline__ int march_point(int e) {
    constexpr int BX = TX + 4, BY = TY + 4;
    if (e < TX * TY) return (e % TX + 2) + BX * (e / TX + 2);
    int h = e - TX * TY;
    if (h < 2 * BX) return h;
    h -= 2 * BX;
    if (h < 2 * BX) return (h % BX) + BX * (BY - 2 + h / BX);
    h -= 2 * BX;
    const int r = h >> 2, cc = h & 3;
    return (cc < 2 ? cc : TX + cc) + BX * (r + 2);
}

template <int TX, int TY, int CONV, int VISC, bool FLUXRES, int TIME>
__global__ void __launch_bounds__(TX * TY) k_central_march(const double* __restrict__ q, const double* __restrict__ F,
                                                           double* __restrict__ reg, double* __restrict__ qout,
                                                           const double* __restrict__ scal, Geom g, Consts c,
                                                           double A, double B, int KZ) {
    constexpr int NV = 5;
    constexpr int BX = TX + 4, BY = TY + 4, BP = BX * BY;     // plane with x/y halo
    constexpr int NT = TX * TY;
    constexpr bool SUTH = VISC == 2;
    constexpr bool SKEW = CONV == CF_SKEW;
    constexpr bool FEI = CONV == CF_FEIEREISEN;
    constexpr bool HAS_CONV = CONV != CF_NONE;
    // LDS: u ring (3 planes) and the output plane's scalars
    __shared__ double s_u[3][3][BP];
    __shared__ double s_p[HAS_CONV ? BP : 1];
    __shared__ double s_T[BP];
    __shared__ double s_E[FEI ? BP : 1];
    __shared__ double s_mu[SUTH ? BP : 1];
    // conservative variables of the output plane's x/y halo: rho u_m, rhoE
    // (and rho for the skew form), slot v - Q0 holds variable v
    constexpr int Q0 = SKEW ? 0 : 1;
    __shared__ double s_q[HAS_CONV ? NV - Q0 : 1][HAS_CONV ? BP : 1];

    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY, k0 = blockIdx.z * KZ;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const long P = g.Ptot;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = i0 + tx, j = j0 + ty;
    const bool col_in = i < n0 && j < n1;
    const int lc = (tx + 2) + BX * (ty + 2);                 // own column in the plane
    const int lsx = 1, lsy = BX;
    const int kend = min(k0 + KZ, n2);

    // register rings, slot s holds plane z + s - 2 of the output plane z
    double r_q[NV][5];          // rho, rho u0..2, rhoE
    double r_u[3][5];
    double r_p[5], r_T[5], r_E[5], r_mu[5];
    double r_mx[4][5];          // D_x u2, D_y u2, D_x u0, D_y u1 (inner parts of the z-mixed terms)
#pragma unroll
    for (int s = 0; s < 5; ++s) {
#pragma unroll
        for (int v = 0; v < NV; ++v) r_q[v][s] = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) r_u[m][s] = 0.0;
        r_p[s] = r_T[s] = r_E[s] = r_mu[s] = 0.0;
#pragma unroll
        for (int m = 0; m < 4; ++m) r_mx[m][s] = 0.0;
    }

    for (int zl = k0 - 2; zl < kend + 2; ++zl) {
        const int slot = (zl + 3) % 3;                      // zl >= -2
        // ---- shift the register rings ------------------------------------
#pragma unroll
        for (int s = 0; s < 4; ++s) {
#pragma unroll
            for (int v = 0; v < NV; ++v) r_q[v][s] = r_q[v][s + 1];
#pragma unroll
            for (int m = 0; m < 3; ++m) r_u[m][s] = r_u[m][s + 1];
            r_p[s] = r_p[s + 1]; r_T[s] = r_T[s + 1]; r_E[s] = r_E[s + 1]; r_mu[s] = r_mu[s + 1];
#pragma unroll
            for (int m = 0; m < 4; ++m) r_mx[m][s] = r_mx[m][s + 1];
        }
        __syncthreads();                                    // previous output done with LDS
        // ---- load plane zl: u to the LDS ring; own column to registers -----
        const bool plane_ok = zl < n2 + 2;
        for (int e = threadIdx.x; e < BP; e += NT) {
            // enumeration: e < NT is the interior in thread order (thread t's
            // first point is its own column, whose values stay in registers),
            // then the 2-wide halo ring
            const int bb = march_point<TX, TY>(e);
            const int bx = bb % BX, by = bb / BX;
            const int ii = i0 + bx - 2, jj = j0 + by - 2;
            if (!plane_ok || ii >= n0 + 2 || jj >= n1 + 2) continue;
            const long x = gidx(g, ii, jj, zl);
            const double r = q[x];
            double uu[3];
            double ke = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                uu[d] = q[(long)(1 + d) * P + x] / r;
                s_u[slot][d][bb] = uu[d];
                ke -= 0.5 * (uu[d] * uu[d]) * r;
            }
            if (bb == lc) {
                const double rhoE = q[4 * P + x];
                const double p = c.gamma_m1 * (ke + rhoE);
                const double TT = (c.minf * c.minf) * c.gama * p / r;
                r_q[0][4] = r;
#pragma unroll
                for (int d = 0; d < 3; ++d) { r_q[1 + d][4] = q[(long)(1 + d) * P + x]; r_u[d][4] = uu[d]; }
                r_q[4][4] = rhoE;
                r_p[4] = p;
                r_T[4] = TT;
                r_E[4] = rhoE / r;
                if (SUTH) r_mu[4] = 1.4042 * TT * sqrt(TT) / (0.40417 + TT);
            }
        }
        __syncthreads();                                    // u of plane zl complete
        // inner derivatives of the z-mixed terms at plane zl
        {
            const double* u0 = s_u[slot][0];
            const double* u1 = s_u[slot][1];
            const double* u2 = s_u[slot][2];
            r_mx[0][4] = d1([&](int r) { return u2[lc + r * lsx]; }, c.invd[0]);
            r_mx[1][4] = d1([&](int r) { return u2[lc + r * lsy]; }, c.invd[1]);
            r_mx[2][4] = d1([&](int r) { return u0[lc + r * lsx]; }, c.invd[0]);
            r_mx[3][4] = d1([&](int r) { return u1[lc + r * lsy]; }, c.invd[1]);
        }
        const int k = zl - 2;                               // output plane
        if (k < k0) continue;                               // uniform over the block
        const int oslot = (k + 3) % 3;
        // ---- scalars of the output plane's x/y halo (cross) ----------------
        for (int b = threadIdx.x; b < BP; b += NT) {
            const int bx = b % BX, by = b / BX;
            const bool hx = bx < 2 || bx >= TX + 2, hy = by < 2 || by >= TY + 2;
            if (hx && hy) continue;                         // corners: not read
            if (!hx && !hy) continue;                       // interior: written from registers below
            const int ii = i0 + bx - 2, jj = j0 + by - 2;
            if (ii >= n0 + 2 || jj >= n1 + 2) continue;
            const long x = gidx(g, ii, jj, k);
            const double r = q[x];
            double ke = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double u = s_u[oslot][d][b];
                ke -= 0.5 * (u * u) * r;
            }
            const double rhoE = q[4 * P + x];
            const double p = c.gamma_m1 * (ke + rhoE);
            const double TT = (c.minf * c.minf) * c.gama * p / r;
            if (HAS_CONV) {
                s_p[HAS_CONV ? b : 0] = p;
#pragma unroll
                for (int v = Q0; v < NV; ++v) s_q[HAS_CONV ? v - Q0 : 0][HAS_CONV ? b : 0] = q[(long)v * P + x];
            }
            s_T[b] = TT;
            if (FEI) s_E[FEI ? b : 0] = rhoE / r;
            if (SUTH) s_mu[SUTH ? b : 0] = 1.4042 * TT * sqrt(TT) / (0.40417 + TT);
        }
        {   // own column at the output plane (ring slot 2)
            if (HAS_CONV) {
                s_p[HAS_CONV ? lc : 0] = r_p[2];
#pragma unroll
                for (int v = Q0; v < NV; ++v) s_q[HAS_CONV ? v - Q0 : 0][HAS_CONV ? lc : 0] = r_q[v][2];
            }
            s_T[lc] = r_T[2];
            if (FEI) s_E[FEI ? lc : 0] = r_E[2];
            if (SUTH) s_mu[SUTH ? lc : 0] = r_mu[2];
        }
        __syncthreads();
        if (!col_in || k >= kend) continue;

        // ---- right-hand side at (i, j, k) ----------------------------------
        const long x = gidx(g, i, j, k);
        const double* su[3] = {s_u[oslot][0], s_u[oslot][1], s_u[oslot][2]};
        const int ls[2] = {lsx, lsy};
        // value of field f at offset s along direction jd
        auto U = [&](int m, int jd, int s) { return jd < 2 ? su[m][lc + s * ls[jd < 2 ? jd : 0]] : r_u[m][s + 2]; };
        auto Pf = [&](int jd, int s) { return jd < 2 ? s_p[HAS_CONV ? lc + s * ls[jd < 2 ? jd : 0] : 0] : r_p[s + 2]; };
        auto Tf = [&](int jd, int s) { return jd < 2 ? s_T[lc + s * ls[jd < 2 ? jd : 0]] : r_T[s + 2]; };
        auto Ef = [&](int jd, int s) { return jd < 2 ? s_E[FEI ? lc + s * ls[jd < 2 ? jd : 0] : 0] : r_E[s + 2]; };
        auto MUf = [&](int jd, int s) { return jd < 2 ? s_mu[SUTH ? lc + s * ls[jd < 2 ? jd : 0] : 0] : r_mu[s + 2]; };
        auto Qf = [&](int v, int jd, int s) {
            return jd < 2 ? s_q[HAS_CONV ? v - Q0 : 0][HAS_CONV ? lc + s * ls[jd < 2 ? jd : 0] : 0] : r_q[v][s + 2];
        };

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

        double G[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) G[m][jd] = d1([&](int s) { return U(m, jd, s); }, c.invd[jd]);

        if (HAS_CONV) {
            const double rho = r_q[0][2], rhoE = r_q[4][2];
            double div_rhou = 0.0, Rrho = 0.0, eterm = 0.0;
            double conv[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                const double inv = c.invd[jd];
                auto uj = [&](int s) { return U(jd, jd, s); };
                if (FEI) {
                    const double d_rhouj = d1([&](int s) { return Qf(1 + jd, jd, s); }, inv);
                    div_rhou += d_rhouj;
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        const double d_prod = d1([&](int s) { return Qf(1 + m, jd, s) * uj(s); }, inv);
                        conv[m] += d_prod + Qf(1 + jd, jd, 0) * G[m][jd];
                    }
                    const double d_Eu = d1([&](int s) { return Qf(4, jd, s) * uj(s); }, inv);
                    const double d_E = d1([&](int s) { return Ef(jd, s); }, inv);
                    const double d_pu = d1([&](int s) { return Pf(jd, s) * uj(s); }, inv);
                    eterm += -0.5 * (d_Eu + Qf(1 + jd, jd, 0) * d_E) - d_pu;
                } else {
                    const double ujc = uj(0);
                    const double d_rhou = d1([&](int s) { return Qf(0, jd, s) * uj(s); }, inv);
                    const double d_rho = d1([&](int s) { return Qf(0, jd, s); }, inv);
                    const double d_u = d1([&](int s) { return uj(s); }, inv);
                    Rrho -= 0.5 * (d_rhou + ujc * d_rho + rho * d_u);
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        const double d_prod = d1([&](int s) { return Qf(1 + m, jd, s) * uj(s); }, inv);
                        const double d_rm = d1([&](int s) { return Qf(1 + m, jd, s); }, inv);
                        conv[m] += 0.5 * (d_prod + ujc * d_rm + Qf(1 + m, jd, 0) * d_u);
                    }
                    const double d_Eu = d1([&](int s) { return Qf(4, jd, s) * uj(s); }, inv);
                    const double d_rE = d1([&](int s) { return Qf(4, jd, s); }, inv);
                    const double d_pu = d1([&](int s) { return Pf(jd, s) * uj(s); }, inv);
                    eterm += -0.5 * (d_Eu + ujc * d_rE + rhoE * d_u) - d_pu;
                }
            }
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double dp = d1([&](int s) { return Pf(m, s); }, c.invd[m]);
                if (FEI) Rv[1 + m] = -dp - 0.5 * (conv[m] + U(m, 0, 0) * div_rhou);
                else Rv[1 + m] = -dp - conv[m];
            }
            if (FEI) {
                Rv[0] = -div_rhou;
                Rv[4] = eterm - 0.5 * (rhoE / rho) * div_rhou;
            } else {
                Rv[0] = Rrho;
                Rv[4] = eterm;
            }
        }

        if (VISC != 0) {
            double D2u[3][3], dmu[3], dT[3];
            double lapT = 0.0;
#pragma unroll
            for (int jd = 0; jd < 3; ++jd) {
                dmu[jd] = SUTH ? d1([&](int s) { return MUf(jd, s); }, c.invd[jd]) : 0.0;
                dT[jd] = d1([&](int s) { return Tf(jd, s); }, c.invd[jd]);
                lapT += d2([&](int s) { return Tf(jd, s); }, c.inv2d[jd]);
#pragma unroll
                for (int m = 0; m < 3; ++m) D2u[m][jd] = d2([&](int s) { return U(m, jd, s); }, c.inv2d[jd]);
            }
            double div = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) div += G[m][m];
            const double muv = SUTH ? r_mu[2] : 1.0;
            double Eacc = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                double Lm = (4.0 / 3.0) * D2u[m][m];
#pragma unroll
                for (int jd = 0; jd < 3; ++jd) {
                    if (jd == m) continue;
                    const int a = jd < m ? jd : m, b = jd < m ? m : jd;
                    // D_b(D_a u_jd)
                    double mixed;
                    if (b < 2) {
                        // x-y: both from the LDS plane
                        const int la = ls[a < 2 ? a : 0], lb = ls[b < 2 ? b : 0];
                        auto ga = [&](int s) {
                            const int lq = lc + s * lb;
                            return d1([&](int r) { return su[jd][lq + r * la]; }, c.invd[a]);
                        };
                        mixed = d1(ga, c.invd[b]);
                    } else {
                        // b = z: inner D_a u_jd at planes k-2..k+2 from the ring
                        const int ring = (jd == 2) ? (a == 0 ? 0 : 1) : (jd == 0 ? 2 : 3);
                        mixed = d1([&](int s) { return r_mx[ring][s + 2]; }, c.invd[2]);
                    }
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
                double e = U(m, 0, 0) * mom;
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
        if (TIME == 1) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const long y = (long)v * P + x;
                const double d = fma(A, reg[y], dt * Rv[v]);
                reg[y] = d;
                qout[y] = fma(B, d, r_q[v][2]);
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const long y = (long)v * P + x;
                const double o = reg[y];
                qout[y] = fma(dt * A, Rv[v], o);
                reg[y] = fma(dt * B, Rv[v], o);
            }
        }
    }
}

}  // namespace osb
