// osb_oracle.cpp -- CPU restatement of OpenSBLI's explicit-solver hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke entry
// and the benchmark's CPU-baseline leg load it (libosb_oracle.so) to check and
// time the HIP library.  The product (libosb_hip.so) never links or calls it.
//
// It restates, array by array, the kernel sequence the reference generates
// for one block (opensbli/code_generation/algorithm/algorithm.py:387-502;
// exemplar apps/performance_benchmark/TENO6/opensbli.cpp:216-430):
//   constituent relations on halo-extended ranges  -> u_i, p, a, T, mu arrays
//   characteristic reconstruction per direction     -> face-flux work arrays wk
//   residual kernel  R = -sum_d (wk_d(i) - wk_d(i-1)) / Delta_d
//   or central split-form convective residual
//   stored velocity gradients + viscous kernel      R += ...
//   RK update, then boundary conditions
// Every formula cites the reference lines it follows.  Arrays carry 5 halo
// points per side like the reference's ops_dat declarations
// (opensbli/schemes/spatial/scheme.py:48-61).
//
// Each function here is pinned by a check independent of this file: the
// reference's known answers (the TENO6 128^3 README line, tg_sym, the Sod
// exact solution), published tables (WENO5/TENO5 weights, RK-LS
// coefficients, Carpenter's closure rows) and mathematical identities
// (polynomial exactness and convergence order, L R = I and A R = R Lambda,
// the Roe property, conservation, shift invariance, the CFL definition
// recomputed from the state, a constant-viscosity shear wave); see
// tests/test_oracle.py.  No function is left parity-unpinned.
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).

#include "../include/osb.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

constexpr int HALO = 5;          // ops_dat halo depth in the reference
constexpr double INV_SQRT2 = 0.707106781186547;  // sqrt(2.0) printed to 15 s.f. (euler_eigensystem.py:109,139)

std::string g_create_error;

struct Range { int lo[3]; int hi[3]; };

}  // namespace

struct osb_ctx {
    osb_problem_desc d;
    int ndim = 0, nv = 0;
    int n[3] = {1, 1, 1};
    int P[3] = {1, 1, 1};
    int off[3] = {0, 0, 0};
    size_t Ptot = 0;
    int depth = 0;                 // boundary halo depth (symmetric max of scheme halos)
    bool shock = false, central = false, viscous = false;
    int carp[3] = {0, 0, 0};       // Carpenter closure per direction: bit 0 side 0, bit 1 side 1

    std::vector<double> q, rkold, R;
    std::vector<double> u, p, a, T, mu;   // constituent relations
    std::vector<double> wk;               // face fluxes, nv per direction
    std::vector<double> kappa;            // WENO filter's Ducros sensor, [0,N); 0 in the halos
    std::vector<double> g;                // stored velocity gradients g[i*ndim+j] = d u_i / d x_j
    double time = 0.0;
    int64_t iter = 0;
    double last_dt = 0.0;
    std::string err;
    int rk_stages = 3;                    // RK-LS stages and coefficients
    double rkA[5] = {0}, rkB[5] = {0};
    double glf[3][3] = {};                // GLF: max |u_d|, |u_d + a|, |u_d - a| over the grid, per stage
    int next_stage = 0;                   // osb_stage: the stage the current step continues with
    double stage_dt = 0.0;                // dt of the current step

    size_t I(int i, int j, int k) const {
        return (size_t)(i + off[0]) + (size_t)P[0] * ((size_t)(j + off[1]) + (size_t)P[1] * (size_t)(k + off[2]));
    }
    double* Q(int v) { return q.data() + (size_t)v * Ptot; }
    size_t stride(int dir) const { return dir == 0 ? 1 : (dir == 1 ? (size_t)P[0] : (size_t)P[0] * P[1]); }
};

namespace {

int fail(osb_ctx* c, int code, const std::string& m) {
    if (c) c->err = m; else g_create_error = m;
    return code;
}

Range make_range(const osb_ctx* c, int ext) {
    Range r;
    for (int d = 0; d < 3; ++d) {
        if (d < c->ndim) { r.lo[d] = -ext; r.hi[d] = c->n[d] + ext; }
        else { r.lo[d] = 0; r.hi[d] = 1; }
    }
    return r;
}

template <class F>
void for_range(const Range& r, F f) {
    const int nk = r.hi[2] - r.lo[2], nj = r.hi[1] - r.lo[1];
    const long total = (long)nk * nj;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < total; ++t) {
        const int k = r.lo[2] + (int)(t / nj);
        const int j = r.lo[1] + (int)(t % nj);
        for (int i = r.lo[0]; i < r.hi[0]; ++i) f(i, j, k);
    }
}

// ---------------------------------------------------------------------------
// Constituent relations (apps/compressible_taylor_green_vortex/compressible_TGV.py:78-83;
// generated TENO6/opensbliblock00_kernels.h:48-91)
// ---------------------------------------------------------------------------
void constituent_relations(osb_ctx* c, const Range& r, bool need_T_mu) {
    const int nd = c->ndim;
    const double gama = c->d.gama, minf = c->d.minf;
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        const double rho = c->q[x];
        double uu[3] = {0, 0, 0};
        for (int dd = 0; dd < nd; ++dd) {
            uu[dd] = c->q[(size_t)(1 + dd) * c->Ptot + x] / rho;   // u_i = rhou_i/rho
            c->u[(size_t)dd * c->Ptot + x] = uu[dd];
        }
        const double rhoE = c->q[(size_t)(nd + 1) * c->Ptot + x];
        // p = (gama-1)*(rhoE - (1/2)*rho*u_i*u_i)
        double ke = 0.0;
        for (int dd = 0; dd < nd; ++dd) ke -= 0.5 * (uu[dd] * uu[dd]) * rho;
        const double pp = (gama - 1.0) * (ke + rhoE);
        c->p[x] = pp;
        c->a[x] = std::sqrt(gama * pp / rho);                       // a = (gama*p/rho)**0.5
        if (need_T_mu) {
            const double TT = (minf * minf) * gama * pp / rho;       // T = p*gama*Minf^2/rho
            c->T[x] = TT;
            c->mu[x] = 1.4042 * TT * std::sqrt(TT) / (0.40417 + TT); // T**1.5*1.4042/(T+0.40417)
        }
    });
}

// ---------------------------------------------------------------------------
// Euler eigensystems (opensbli/physical_models/euler_eigensystem.py:53-165),
// evaluated on averaged quantities.  Direction enters through the unit normal
// k = e_dir (k_t = k / |k| = k).  Full dense matrices; zero entries stay zero.
// Repeated divisions are replaced by reciprocals of a and rho as the
// reference does (create_LEV_inverses / create_REV_inverses,
// shock_capturing.py:475-530) and of gamma-1 (invgamma_m1, TENO6/opensbli.cpp:140).
// ---------------------------------------------------------------------------
// Roe average of the states at offsets 0 and 1 (averaging.py:65-124):
// rho = sqrt(rho_a rho_b), u and H = (p + rhoE)/rho weighted by sqrt(rho),
// a = sqrt((gama-1)(H - |u|^2/2)).
inline void roe_average(int nd, double gamma_m1, double rl, double rr, const double* ua, const double* ub, double pa, double pb,
                        double rEa, double rEb, double& rho_av, double* u_av, double& a_av) {
    const double srl = std::sqrt(rl), srr = std::sqrt(rr);
    rho_av = std::sqrt(rl * rr);
    const double inv = 1.0 / (srl + srr);
    double ke = 0.0;
    for (int m = 0; m < nd; ++m) {
        u_av[m] = (srl * ua[m] + srr * ub[m]) * inv;
        ke += u_av[m] * u_av[m];
    }
    const double Hroe = ((pa + rEa) / srl + (pb + rEb) / srr) * inv;
    a_av = std::sqrt(gamma_m1 * (Hroe - 0.5 * ke));
}

struct Eig {
    double L[5][5];
    double Rm[5][5];
};

void eigensystem(int nd, int dir, double gm1, double rho, const double* uu, double a, Eig& e) {
    std::memset(&e, 0, sizeof(e));
    const double igm1 = 1.0 / gm1;
    const double ia = 1.0 / a;
    if (nd == 1) {
        // H = a**2/(gama-1) + u0**2/2 substituted (euler_eigensystem.py:75-76)
        const double u0 = uu[0];
        const double H = (a * a) * igm1 + (u0 * u0) * 0.5;
        const double D = 2.0 * H - u0 * u0;
        const double iD = 1.0 / D;
        // REV = [[1,1,1],[u0-a,u0,u0+a],[H-u0*a,u0**2/2,H+u0*a]]   (:89)
        e.Rm[0][0] = 1.0; e.Rm[0][1] = 1.0; e.Rm[0][2] = 1.0;
        e.Rm[1][0] = u0 - a; e.Rm[1][1] = u0; e.Rm[1][2] = u0 + a;
        e.Rm[2][0] = H - u0 * a; e.Rm[2][1] = (u0 * u0) * 0.5; e.Rm[2][2] = H + u0 * a;
        // LEV (:90), 1/(2H - u0**2) and 1/a as reciprocals
        e.L[0][0] = u0 * (2.0 * H + a * u0 - u0 * u0) * 0.5 * ia * iD;
        e.L[0][1] = (-H - a * u0 + (u0 * u0) * 0.5) * ia * iD;
        e.L[0][2] = iD;
        e.L[1][0] = 2.0 * (H - u0 * u0) * iD;
        e.L[1][1] = 2.0 * u0 * iD;
        e.L[1][2] = -2.0 * iD;
        e.L[2][0] = u0 * (-2.0 * H + a * u0 + u0 * u0) * 0.5 * ia * iD;
        e.L[2][1] = (H - a * u0 - (u0 * u0) * 0.5) * ia * iD;
        e.L[2][2] = iD;
        return;
    }
    const double k0 = dir == 0 ? 1.0 : 0.0, k1 = dir == 1 ? 1.0 : 0.0, k2 = dir == 2 ? 1.0 : 0.0;
    double q2 = 0.0;
    for (int dd = 0; dd < nd; ++dd) q2 += uu[dd] * uu[dd];
    const double phi_sq = gm1 * (q2 * 0.5);              // (gama-1)*((u0**2+u1**2+u2**2)/2)
    const double theta = uu[dir];                          // k0_t*u0 + k1_t*u1 + k2_t*u2
    const double irho = 1.0 / rho;
    const double ia2 = ia * ia;
    const double alpha = rho * INV_SQRT2 * ia;             // rho/(a*sqrt(2.0))
    const double bta = INV_SQRT2 * irho * ia;              // 1/(rho*a*sqrt(2.0))
    const double a2 = a * a;
    if (nd == 2) {
        const double u0 = uu[0], u1 = uu[1];
        // REV (:121)
        e.Rm[0][0] = 1.0; e.Rm[0][1] = 0.0; e.Rm[0][2] = alpha; e.Rm[0][3] = alpha;
        e.Rm[1][0] = u0; e.Rm[1][1] = k1 * rho; e.Rm[1][2] = alpha * (u0 + k0 * a); e.Rm[1][3] = alpha * (u0 - k0 * a);
        e.Rm[2][0] = u1; e.Rm[2][1] = -k0 * rho; e.Rm[2][2] = alpha * (u1 + k1 * a); e.Rm[2][3] = alpha * (u1 - k1 * a);
        e.Rm[3][0] = phi_sq * igm1; e.Rm[3][1] = rho * (k1 * u0 - k0 * u1);
        e.Rm[3][2] = alpha * ((phi_sq + a2) * igm1 + a * theta);
        e.Rm[3][3] = alpha * ((phi_sq + a2) * igm1 - a * theta);
        // LEV (:122)
        e.L[0][0] = 1.0 - phi_sq * ia2; e.L[0][1] = gm1 * u0 * ia2; e.L[0][2] = gm1 * u1 * ia2; e.L[0][3] = -gm1 * ia2;
        e.L[1][0] = -(k1 * u0 - k0 * u1) * irho; e.L[1][1] = k1 * irho; e.L[1][2] = -k0 * irho; e.L[1][3] = 0.0;
        e.L[2][0] = bta * (phi_sq - a * theta); e.L[2][1] = bta * (k0 * a - gm1 * u0);
        e.L[2][2] = bta * (k1 * a - gm1 * u1); e.L[2][3] = bta * gm1;
        e.L[3][0] = bta * (phi_sq + a * theta); e.L[3][1] = -bta * (k0 * a + gm1 * u0);
        e.L[3][2] = -bta * (k1 * a + gm1 * u1); e.L[3][3] = bta * gm1;
        return;
    }
    // ndim == 3 (:150-152)
    const double u0 = uu[0], u1 = uu[1], u2 = uu[2];
    e.Rm[0][0] = k0; e.Rm[0][1] = k1; e.Rm[0][2] = k2; e.Rm[0][3] = alpha; e.Rm[0][4] = alpha;
    e.Rm[1][0] = k0 * u0; e.Rm[1][1] = k1 * u0 - k2 * rho; e.Rm[1][2] = k2 * u0 + k1 * rho;
    e.Rm[1][3] = alpha * (u0 + k0 * a); e.Rm[1][4] = alpha * (u0 - k0 * a);
    e.Rm[2][0] = k0 * u1 + k2 * rho; e.Rm[2][1] = k1 * u1; e.Rm[2][2] = k2 * u1 - k0 * rho;
    e.Rm[2][3] = alpha * (u1 + k1 * a); e.Rm[2][4] = alpha * (u1 - k1 * a);
    e.Rm[3][0] = k0 * u2 - k1 * rho; e.Rm[3][1] = k1 * u2 + k0 * rho; e.Rm[3][2] = k2 * u2;
    e.Rm[3][3] = alpha * (u2 + k2 * a); e.Rm[3][4] = alpha * (u2 - k2 * a);
    e.Rm[4][0] = k0 * phi_sq * igm1 + rho * (k2 * u1 - k1 * u2);
    e.Rm[4][1] = k1 * phi_sq * igm1 + rho * (k0 * u2 - k2 * u0);
    e.Rm[4][2] = k2 * phi_sq * igm1 + rho * (k1 * u0 - k0 * u1);
    e.Rm[4][3] = alpha * ((phi_sq + a2) * igm1 + theta * a);
    e.Rm[4][4] = alpha * ((phi_sq + a2) * igm1 - theta * a);
    const double c1 = 1.0 - phi_sq * ia2;
    e.L[0][0] = k0 * c1 - (k2 * u1 - k1 * u2) * irho;
    e.L[0][1] = k0 * gm1 * u0 * ia2;
    e.L[0][2] = k0 * gm1 * u1 * ia2 + k2 * irho;
    e.L[0][3] = k0 * gm1 * u2 * ia2 - k1 * irho;
    e.L[0][4] = -k0 * gm1 * ia2;
    e.L[1][0] = k1 * c1 - (k0 * u2 - k2 * u0) * irho;
    e.L[1][1] = k1 * gm1 * u0 * ia2 - k2 * irho;
    e.L[1][2] = k1 * gm1 * u1 * ia2;
    e.L[1][3] = k1 * gm1 * u2 * ia2 + k0 * irho;
    e.L[1][4] = -k1 * gm1 * ia2;
    e.L[2][0] = k2 * c1 - (k1 * u0 - k0 * u1) * irho;
    e.L[2][1] = k2 * gm1 * u0 * ia2 + k1 * irho;
    e.L[2][2] = k2 * gm1 * u1 * ia2 - k0 * irho;
    e.L[2][3] = k2 * gm1 * u2 * ia2;
    e.L[2][4] = -k2 * gm1 * ia2;
    e.L[3][0] = bta * (phi_sq - theta * a);
    e.L[3][1] = -bta * (gm1 * u0 - k0 * a);
    e.L[3][2] = -bta * (gm1 * u1 - k1 * a);
    e.L[3][3] = -bta * (gm1 * u2 - k2 * a);
    e.L[3][4] = bta * gm1;
    e.L[4][0] = bta * (phi_sq + theta * a);
    e.L[4][1] = -bta * (gm1 * u0 + k0 * a);
    e.L[4][2] = -bta * (gm1 * u1 + k1 * a);
    e.L[4][3] = -bta * (gm1 * u2 + k2 * a);
    e.L[4][4] = bta * gm1;
}

// ---------------------------------------------------------------------------
// Reconstructions.  f[0..5] hold the split flux at offsets -2..3 from point i;
// the interface is i+1/2.  Right (f+) uses offsets -2..2, left (f-) -1..3
// (opensbli/schemes/spatial/weno.py:123-136; teno.py:95-107).
// Fused multiply-adds appear only where written as std::fma.
// ---------------------------------------------------------------------------

// Jiang-Shu smoothness of a 3-point stencil (a, b, c in ascending x):
// 13/12 (a-2b+c)^2 + 1/4 s2^2 with s2 = 3a-4b+c ('first'), a-c ('centre'),
// a-4b+3c ('last') -- the k=3 coefficients of JS_smoothness (weno.py:139-206).
inline double js_sum(double s1, double s2) { return std::fma((13.0 / 12.0) * s1, s1, (0.25 * s2) * s2); }
inline double js_first(double a, double b, double c) { return js_sum(a - 2.0 * b + c, 3.0 * a - 4.0 * b + c); }
inline double js_centre(double a, double b, double c) { return js_sum(a - 2.0 * b + c, a - c); }
inline double js_last(double a, double b, double c) { return js_sum(a - 2.0 * b + c, a - 4.0 * b + 3.0 * c); }

// Shu (1997) eq. (2.21) ENO coefficients c_rj (weno.py:71-99), evaluated in
// exact integer arithmetic for k = 3.  'dshift' is 1 for the right (side +1)
// and 0 for the left (side -1) reconstruction.
struct Frac { long num, den; };
Frac frac_add(Frac x, Frac y) {
    long n = x.num * y.den + y.num * x.den, dd = x.den * y.den;
    long g = std::abs(n), h = std::abs(dd);
    while (h) { long t = g % h; g = h; h = t; }
    if (g == 0) g = 1;
    n /= g; dd /= g;
    if (dd < 0) { n = -n; dd = -dd; }
    return {n, dd};
}
Frac eno_coefficient(int k, int r, int j, int dshift) {
    Frac sum{0, 1};
    for (int m = j + 1; m <= k; ++m) {
        long top_sum = 0, bottom = 1;
        for (int l = 0; l <= k; ++l) {
            if (l == m) continue;
            long prod = 1;
            for (int qq = 0; qq <= k; ++qq) {
                if (qq == m || qq == l) continue;
                prod *= (r - qq + dshift);
            }
            bottom *= (m - l);
            top_sum += prod;
        }
        sum = frac_add(sum, Frac{top_sum, bottom});
    }
    return sum;
}

struct WenoCoeffs {
    double c[2][3][3];   // [side 0=right,1=left][r][j]
    double dopt[2][3];
};
WenoCoeffs make_weno5() {
    WenoCoeffs w;
    for (int s = 0; s < 2; ++s)
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) {
                Frac f = eno_coefficient(3, r, j, s == 0 ? 1 : 0);
                w.c[s][r][j] = (double)f.num / (double)f.den;
            }
    // Optimal weights (weno.py:101-121).  Right: r=0 [0,1,2] 3/10, r=1 [-1,0,1] 6/10,
    // r=2 [-2,-1,0] 1/10; the left side mirrors them.
    w.dopt[0][0] = 3.0 / 10.0; w.dopt[0][1] = 6.0 / 10.0; w.dopt[0][2] = 1.0 / 10.0;
    w.dopt[1][0] = 1.0 / 10.0; w.dopt[1][1] = 6.0 / 10.0; w.dopt[1][2] = 3.0 / 10.0;
    return w;
}
const WenoCoeffs W5 = make_weno5();

// three-point ENO candidate, points in ascending x
inline double eno3(const double* c, double x0, double x1, double x2) {
    return std::fma(c[2], x2, std::fma(c[1], x1, c[0] * x0));
}
// alpha_r = d_r*(1 + (tau/(eps+b_r))**2), printed expanded as d_r + d_r*tau^2/((eps+b_r)(eps+b_r))
inline double zalpha(double d, double t2, double b, double eps) {
    const double eb = eps + b;
    return d + (d * t2) / (eb * eb);
}

// WENO-Z (weno.py:290-345): tau5 = |b0 - b2|.
// shock sensor of the WENO filter (weno.py:443-459, normalised): sum_r |omega_r/d_r - 1| /
// (|1/min(d) - 1| + (k - 1)) = (...) / 11 for WENO5
inline double weno_sensor(const double* d, double w0, double w1, double w2) {
    return (std::fabs(w0 * (1.0 / d[0]) - 1.0) + std::fabs(w1 * (1.0 / d[1]) - 1.0) + std::fabs(w2 * (1.0 / d[2]) - 1.0)) / 11.0;
}

double wenoz_right(const double* f, double eps, double* sensor = nullptr) {
    // points f[0]=i-2 ... f[4]=i+2 ; r=0 -> (i,i+1,i+2), r=1 -> (i-1,i,i+1), r=2 -> (i-2,i-1,i)
    const double b0 = js_first(f[2], f[3], f[4]);
    const double b1 = js_centre(f[1], f[2], f[3]);
    const double b2 = js_last(f[0], f[1], f[2]);
    const double tau = std::fabs(b0 - b2), t2 = tau * tau;
    const double* d = W5.dopt[0];
    const double a0 = zalpha(d[0], t2, b0, eps), a1 = zalpha(d[1], t2, b1, eps), a2 = zalpha(d[2], t2, b2, eps);
    const double inv = 1.0 / (a0 + a1 + a2);
    const double e0 = eno3(W5.c[0][0], f[2], f[3], f[4]);
    const double e1 = eno3(W5.c[0][1], f[1], f[2], f[3]);
    const double e2 = eno3(W5.c[0][2], f[0], f[1], f[2]);
    if (sensor) *sensor = weno_sensor(d, a0 * inv, a1 * inv, a2 * inv);
    return std::fma(a2 * inv, e2, std::fma(a1 * inv, e1, (a0 * inv) * e0));
}
double wenoz_left(const double* f, double eps, double* sensor = nullptr) {
    // points f[1]=i-1 ... f[5]=i+3 ; r=0 -> (i+1,i+2,i+3), r=1 -> (i,i+1,i+2), r=2 -> (i-1,i,i+1)
    const double b0 = js_first(f[3], f[4], f[5]);
    const double b1 = js_centre(f[2], f[3], f[4]);
    const double b2 = js_last(f[1], f[2], f[3]);
    const double tau = std::fabs(b0 - b2), t2 = tau * tau;
    const double* d = W5.dopt[1];
    const double a0 = zalpha(d[0], t2, b0, eps), a1 = zalpha(d[1], t2, b1, eps), a2 = zalpha(d[2], t2, b2, eps);
    const double inv = 1.0 / (a0 + a1 + a2);
    const double e0 = eno3(W5.c[1][0], f[3], f[4], f[5]);
    const double e1 = eno3(W5.c[1][1], f[2], f[3], f[4]);
    const double e2 = eno3(W5.c[1][2], f[1], f[2], f[3]);
    if (sensor) *sensor = weno_sensor(d, a0 * inv, a1 * inv, a2 * inv);
    return std::fma(a2 * inv, e2, std::fma(a1 * inv, e1, (a0 * inv) * e0));
}

// WENO3-Z, used by the WENO filter of order 3 (weno.py:36-136, 290-345 with
// k = 2: tau = |beta_0 - beta_1|, p = 2; generated form in
// apps/performance_benchmark/Central4-WENO5/opensbliblock00_kernels.h:930-990).
// f[s] is the split value at offset s-2 from point i; beta_r is the squared
// difference over stencil r (the generated 1/4 scaling is the split's 1/2
// squared).  The sensor is sum_r |omega_r/d_r - 1| / 3 (weno.py:443-459).
double wenoz3_right(const double* f, double eps, double* sensor) {
    // r=0 -> (i, i+1), d = 2/3; r=1 -> (i-1, i), d = 1/3
    const double db0 = f[3] - f[2], db1 = f[2] - f[1];
    const double b0 = db0 * db0, b1 = db1 * db1;
    const double tau = std::fabs(b0 - b1), t2 = tau * tau;
    const double a0 = zalpha(2.0 / 3.0, t2, b0, eps), a1 = zalpha(1.0 / 3.0, t2, b1, eps);
    const double inv = 1.0 / (a0 + a1);
    const double w0 = a0 * inv, w1 = a1 * inv;
    *sensor = (std::fabs(3.0 * w1 - 1.0) + std::fabs(1.5 * w0 - 1.0)) * (1.0 / 3.0);
    const double e0 = 0.5 * f[2] + 0.5 * f[3];
    const double e1 = std::fma(-0.5, f[1], 1.5 * f[2]);
    return std::fma(w1, e1, w0 * e0);
}
double wenoz3_left(const double* f, double eps, double* sensor) {
    // r=0 -> (i+1, i+2), d = 1/3; r=1 -> (i, i+1), d = 2/3
    const double db0 = f[4] - f[3], db1 = f[3] - f[2];
    const double b0 = db0 * db0, b1 = db1 * db1;
    const double tau = std::fabs(b0 - b1), t2 = tau * tau;
    const double a0 = zalpha(1.0 / 3.0, t2, b0, eps), a1 = zalpha(2.0 / 3.0, t2, b1, eps);
    const double inv = 1.0 / (a0 + a1);
    const double w0 = a0 * inv, w1 = a1 * inv;
    *sensor = (std::fabs(3.0 * w0 - 1.0) + std::fabs(1.5 * w1 - 1.0)) * (1.0 / 3.0);
    const double e0 = std::fma(-0.5, f[4], 1.5 * f[3]);
    const double e1 = 0.5 * f[2] + 0.5 * f[3];
    return std::fma(w1, e1, w0 * e0);
}

// TENO5 / TENO6 (teno.py:59-217, 419-477).
struct TenoParams { int order; double eps, ct; };
// (C + tau/beta)**6 printed as six repeated multiplies (opsc.py:222-241)
inline double pow6(double x) { return x * x * x * x * x * x; }

double teno_side(const double* f, int side, const TenoParams& tp) {
    auto F = [&](int o) { return f[o + 2]; };
    double beta[4], eno[4], dopt[4];
    const int ns = tp.order == 5 ? 3 : 4;
    if (tp.order == 5) { dopt[0] = 11.0 / 20.0; dopt[1] = 4.0 / 10.0; dopt[2] = 1.0 / 20.0; dopt[3] = 0.0; }
    else { dopt[0] = 231.0 / 500.0; dopt[1] = 3.0 / 10.0; dopt[2] = 27.0 / 500.0; dopt[3] = 23.0 / 125.0; }
    if (side == +1) {
        // stencils [-1,0,1], [0,1,2], [-2,-1,0], [0,1,2,3]  (teno.py:95-99, 110-116, 149-169)
        beta[0] = js_sum(F(-1) - 2.0 * F(0) + F(1), F(-1) - F(1));
        beta[1] = js_sum(F(0) - 2.0 * F(1) + F(2), 3.0 * F(0) - 4.0 * F(1) + F(2));
        beta[2] = js_sum(F(-2) - 2.0 * F(-1) + F(0), F(-2) - 4.0 * F(-1) + 3.0 * F(0));
        eno[0] = std::fma(2.0 / 6.0, F(1), std::fma(5.0 / 6.0, F(0), (-1.0 / 6.0) * F(-1)));
        eno[1] = std::fma(-1.0 / 6.0, F(2), std::fma(5.0 / 6.0, F(1), (2.0 / 6.0) * F(0)));
        eno[2] = std::fma(11.0 / 6.0, F(0), std::fma(-7.0 / 6.0, F(-1), (2.0 / 6.0) * F(-2)));
        if (ns == 4) {
            const double s1 = -11.0 * F(0) + 18.0 * F(1) - 9.0 * F(2) + 2.0 * F(3);
            const double s2 = 2.0 * F(0) - 5.0 * F(1) + 4.0 * F(2) - F(3);
            const double s3 = -F(0) + 3.0 * F(1) - 3.0 * F(2) + F(3);
            // the right-side beta_3 carries the 781/720 term unsquared (teno.py:168-169)
            beta[3] = std::fma((1.0 / 36.0) * s1, s1, std::fma((13.0 / 12.0) * s2, s2, (781.0 / 720.0) * s3));
            eno[3] = std::fma(1.0 / 12.0, F(3), std::fma(-5.0 / 12.0, F(2), std::fma(13.0 / 12.0, F(1), (3.0 / 12.0) * F(0))));
        }
    } else {
        // stencils [0,1,2], [-1,0,1], [1,2,3], [-2,-1,0,1]  (teno.py:100-107, 117-123, 170-180)
        beta[0] = js_sum(F(0) - 2.0 * F(1) + F(2), F(0) - F(2));
        beta[1] = js_sum(F(-1) - 2.0 * F(0) + F(1), F(-1) - 4.0 * F(0) + 3.0 * F(1));
        beta[2] = js_sum(F(1) - 2.0 * F(2) + F(3), 3.0 * F(1) - 4.0 * F(2) + F(3));
        eno[0] = std::fma(-1.0 / 6.0, F(2), std::fma(5.0 / 6.0, F(1), (2.0 / 6.0) * F(0)));
        eno[1] = std::fma(2.0 / 6.0, F(1), std::fma(5.0 / 6.0, F(0), (-1.0 / 6.0) * F(-1)));
        eno[2] = std::fma(1.0 / 3.0, F(3), std::fma(-7.0 / 6.0, F(2), (11.0 / 6.0) * F(1)));
        if (ns == 4) {
            const double s1 = -2.0 * F(-2) + 9.0 * F(-1) - 18.0 * F(0) + 11.0 * F(1);
            const double s2 = -F(-2) + 4.0 * F(-1) - 5.0 * F(0) + 2.0 * F(1);
            const double s3 = -F(-2) + 3.0 * F(-1) - 3.0 * F(0) + F(1);
            beta[3] = std::fma((1.0 / 36.0) * s1, s1, std::fma((13.0 / 12.0) * s2, s2, ((781.0 / 720.0) * s3) * s3));
            eno[3] = std::fma(3.0 / 12.0, F(1), std::fma(13.0 / 12.0, F(0), std::fma(-5.0 / 12.0, F(-1), (1.0 / 12.0) * F(-2))));
        }
    }
    double tau;
    if (ns == 3) tau = std::fabs(beta[0] - beta[2]);                                          // Teno5 (:210)
    else tau = std::fabs(beta[3] - (1.0 / 6.0) * (beta[0] + beta[2] - 4.0 * beta[1]));         // Teno6 (:228)
    double alpha[4], asum = 0.0;
    for (int r = 0; r < ns; ++r) {
        const double inv_beta = 1.0 / (tp.eps + beta[r]);
        alpha[r] = pow6(1.0 + tau * inv_beta);                                                  // (C + tau/beta)**6
        asum += alpha[r];
    }
    const double inv_asum = 1.0 / asum;
    double delta[4], wsum = 0.0;
    for (int r = 0; r < ns; ++r) {
        delta[r] = (alpha[r] * inv_asum < tp.ct) ? 0.0 : 1.0;                                  // cut-off (:457-477)
        wsum += dopt[r] * delta[r];
    }
    const double inv_wsum = 1.0 / wsum;
    double rec = 0.0;
    for (int r = 0; r < ns; ++r) rec = std::fma(dopt[r] * delta[r] * inv_wsum, eno[r], rec);
    return rec;
}

// ---------------------------------------------------------------------------
// sin(x) for x in [0, pi/2] by its Taylor series through x^21 (truncation
// below 1.2e-18 there), Horner in x^2 with fused multiply-adds: the HLLC-LM
// Mach switch evaluated identically on every platform.
// ---------------------------------------------------------------------------
inline double hllc_sin(double x) {
    const double x2 = x * x;
    double s = -1.0 / 51090942171709440000.0;                          // -1/21!
    s = std::fma(s, x2, 1.0 / 121645100408832000.0);                    //  1/19!
    s = std::fma(s, x2, -1.0 / 355687428096000.0);                      // -1/17!
    s = std::fma(s, x2, 1.0 / 1307674368000.0);                         //  1/15!
    s = std::fma(s, x2, -1.0 / 6227020800.0);                           // -1/13!
    s = std::fma(s, x2, 1.0 / 39916800.0);                              //  1/11!
    s = std::fma(s, x2, -1.0 / 362880.0);                               // -1/9!
    s = std::fma(s, x2, 1.0 / 5040.0);                                  //  1/7!
    s = std::fma(s, x2, -1.0 / 120.0);                                  // -1/5!
    s = std::fma(s, x2, 1.0 / 6.0);                                     //  1/3! (sign below)
    // sin x = x - x^3 (1/3! - x^2/5! + ...) = x - x^3 * s
    return std::fma(-(x * x2), s, x);
}

// HLLC / HLLC-LM interface flux of reconstructed left and right conservative
// states along dir (shock_capturing.py:1108-1256): wave speeds
// sL = min(ubar - abar, uL - aL), sR = max(ubar + abar, uR + aR) with the face's
// averaged state, the contact speed s*, the star states
// rho_K (s_K - u_K)/(s_K - s*) [1, u (s* along dir), E_K/rho_K + (s* - u_K)(s* + p_K/(rho_K (sL - u_K)))]
// (the reference writes sL in both star energies), then the four-region HLLC
// flux or, for HLLC-LM, the low-Mach form with sL, sR scaled by
// phi = sin(min(1, max(|uL/aL|, |uR/aR|)/0.1) pi/2).
void hllc_flux(int nd, int dir, double gama, bool lm, const double* qL, const double* qR, double ubar, double abar,
               double* out) {
    const int nv = nd + 2;
    const double irL = 1.0 / qL[0], irR = 1.0 / qR[0];
    double uL[3] = {0, 0, 0}, uR[3] = {0, 0, 0}, keL = 0.0, keR = 0.0;
    for (int m = 0; m < nd; ++m) {
        uL[m] = qL[1 + m] * irL;
        uR[m] = qR[1 + m] * irR;
        keL = std::fma(uL[m], uL[m], keL);
        keR = std::fma(uR[m], uR[m], keR);
    }
    const double pL = (gama - 1.0) * (qL[nv - 1] - 0.5 * qL[0] * keL);
    const double pR = (gama - 1.0) * (qR[nv - 1] - 0.5 * qR[0] * keR);
    const double aL = std::sqrt(gama * pL * irL), aR = std::sqrt(gama * pR * irR);
    const double ul = uL[dir], ur = uR[dir];
    double sL = std::min(ubar - abar, ul - aL);
    double sR = std::max(ubar + abar, ur + aR);
    const double ss = ((pR - pL) + (qL[0] * ul * (sL - ul) - qR[0] * ur * (sR - ur))) / (qL[0] * (sL - ul) - qR[0] * (sR - ur));
    if (lm) {
        const double ml = std::max(std::fabs(ul / aL), std::fabs(ur / aR));
        const double phi = hllc_sin(std::min(1.0, ml / 0.1) * (0.5 * 3.14159265358979323846));
        sL = sL * phi;
        sR = sR * phi;
    }
    double FL[5], FR[5], SL[5], SR[5];
    FL[0] = qL[1 + dir];
    FR[0] = qR[1 + dir];
    for (int m = 0; m < nd; ++m) {
        FL[1 + m] = qL[1 + m] * ul + (m == dir ? pL : 0.0);
        FR[1 + m] = qR[1 + m] * ur + (m == dir ? pR : 0.0);
    }
    FL[nv - 1] = (pL + qL[nv - 1]) * ul;
    FR[nv - 1] = (pR + qR[nv - 1]) * ur;
    const double fl = qL[0] * ((sL - ul) / (sL - ss)), fr = qR[0] * ((sR - ur) / (sR - ss));
    SL[0] = fl;
    SR[0] = fr;
    for (int m = 0; m < nd; ++m) {
        SL[1 + m] = fl * (m == dir ? ss : uL[m]);
        SR[1 + m] = fr * (m == dir ? ss : uR[m]);
    }
    SL[nv - 1] = fl * std::fma(ss - ul, ss + pL / (qL[0] * (sL - ul)), qL[nv - 1] * irL);
    SR[nv - 1] = fr * std::fma(ss - ur, ss + pR / (qR[0] * (sL - ur)), qR[nv - 1] * irR);
    for (int v = 0; v < nv; ++v) {
        double f;
        if (lm) {
            if (sL >= 0.0) f = FL[v];
            else if (sR <= 0.0) f = FR[v];
            else f = 0.5 * (FL[v] + FR[v]) + 0.5 * ((sL * (SL[v] - qL[v]) + std::fabs(ss) * (SL[v] - SR[v])) + sR * (SR[v] - qR[v]));
        } else {
            if (sL >= 0.0) f = FL[v];
            else if (sL <= 0.0 && 0.0 <= ss) f = std::fma(sL, SL[v] - qL[v], FL[v]);
            else if (ss <= 0.0 && 0.0 <= sR) f = std::fma(sR, SR[v] - qR[v], FR[v]);
            else if (sR <= 0.0) f = FR[v];
            else f = 0.0;
        }
        out[v] = f;
    }
}

// ---------------------------------------------------------------------------
// Characteristic flux reconstruction for one direction
// (shock_capturing.py:786-891, 957-1028; weno.py:505-551; teno.py:510-559).
// Writes the interface flux at i+1/2 into wk_dir[v] at index i over
// [-1, N+1) along dir and [0, N) across (TENO6/opensbli.cpp:263-316).
// ---------------------------------------------------------------------------
void reconstruct_direction(osb_ctx* c, int dir, bool filter = false) {
    const int nd = c->ndim, nv = c->nv;
    const double gama = c->d.gama, gamma_m1 = gama - 1.0;
    const double sfc = c->d.shock_filter_control;
    Range r = make_range(c, 0);
    r.lo[dir] = -1; r.hi[dir] = c->n[dir] + 1;
    const size_t st = c->stride(dir);
    double* wk = c->wk.data() + (size_t)dir * nv * c->Ptot;
    const bool teno = c->d.conv_scheme == OSB_CONV_TENO;
    TenoParams tp{c->d.order, c->d.teno_eps, c->d.teno_ct};
    const double weps = c->d.weno_eps;
    const bool hybrid = filter && c->d.filter_sensor == OSB_SENSOR_DUCROS;
    const bool hllc = !filter && (c->d.flux_type == OSB_FLUX_HLLC || c->d.flux_type == OSB_FLUX_HLLC_LM);
    for_range(r, [&](int i, int j, int k) {
        const size_t x0 = c->I(i, j, k);
        if (hybrid) {
            // WENO_filter.py:561-584: reconstruct only where max kappa over offsets
            // -3..2 along dir exceeds Ducros_check, else zero fluxes
            double km = c->kappa[x0];
            for (int o = -3; o <= 2; ++o) km = std::max(km, c->kappa[x0 + (size_t)((long)o * (long)st)]);
            if (!(km > c->d.ducros_check)) {
                for (int v = 0; v < nv; ++v) wk[(size_t)v * c->Ptot + x0] = 0.0;
                return;
            }
        }
        // point data at offsets -2..3
        double qv[6][5], F[6][5], ud[6], aa[6];
        for (int s = 0; s < 6; ++s) {
            const size_t x = x0 + (size_t)((long)(s - 2) * (long)st);
            for (int v = 0; v < nv; ++v) qv[s][v] = c->q[(size_t)v * c->Ptot + x];
            const double pp = c->p[x];
            const double u_dir = c->u[(size_t)dir * c->Ptot + x];
            ud[s] = u_dir;
            aa[s] = c->a[x];
            // flux vector of Conservative(rhou_j), Conservative(rhou_i*u_j + KD*p), Conservative((p+rhoE)*u_j)
            F[s][0] = qv[s][1 + dir];
            for (int m = 0; m < nd; ++m) F[s][1 + m] = qv[s][1 + m] * u_dir + (m == dir ? pp : 0.0);
            F[s][nd + 1] = (pp + qv[s][nd + 1]) * u_dir;
        }
        // averaged state between offsets 0 and 1 (averaging.py:31-124)
        const size_t xa = x0, xb = x0 + st;
        double rho_av, u_av[3] = {0, 0, 0}, a_av;
        if (filter || c->d.averaging == OSB_AVG_ROE) {   // the filter's scheme averages by Roe (WENO_filter.py:438)
            double ua[3], ub[3];
            for (int m = 0; m < nd; ++m) { ua[m] = c->u[(size_t)m * c->Ptot + xa]; ub[m] = c->u[(size_t)m * c->Ptot + xb]; }
            roe_average(nd, gamma_m1, c->q[xa], c->q[xb], ua, ub, c->p[xa], c->p[xb], c->q[(size_t)(nd + 1) * c->Ptot + xa],
                        c->q[(size_t)(nd + 1) * c->Ptot + xb], rho_av, u_av, a_av);
        } else {
            rho_av = 0.5 * (c->q[xa] + c->q[xb]);
            for (int m = 0; m < nd; ++m) u_av[m] = 0.5 * (c->u[(size_t)m * c->Ptot + xa] + c->u[(size_t)m * c->Ptot + xb]);
            a_av = 0.5 * (c->a[xa] + c->a[xb]);
        }
        Eig e;
        eigensystem(nd, dir, gamma_m1, rho_av, u_av, a_av, e);
        if (hllc) {
            // HLLCWeno: WENO-Z of the characteristic state CS = L q (no flux splitting); the
            // right reconstruction (offsets -2..2) is the face's left state, the left one
            // (-1..3) its right state (shock_capturing.py:1117-1135)
            double wr[5], wl[5];
            for (int kk = 0; kk < nv; ++kk) {
                double cs[6];
                for (int s = 0; s < 6; ++s) {
                    double a = 0.0;
                    for (int m = 0; m < nv; ++m) a = std::fma(e.L[kk][m], qv[s][m], a);
                    cs[s] = a;
                }
                wr[kk] = wenoz_right(cs, weps);
                wl[kk] = wenoz_left(cs, weps);
            }
            double qL[5], qR[5], fo[5];
            for (int v = 0; v < nv; ++v) {
                double sl = 0.0, sr = 0.0;
                for (int kk = 0; kk < nv; ++kk) { sl = std::fma(e.Rm[v][kk], wr[kk], sl); sr = std::fma(e.Rm[v][kk], wl[kk], sr); }
                qL[v] = sl;
                qR[v] = sr;
            }
            hllc_flux(nd, dir, gama, c->d.flux_type == OSB_FLUX_HLLC_LM, qL, qR, u_av[dir], a_av, fo);
            for (int v = 0; v < nv; ++v) wk[(size_t)v * c->Ptot + x0] = fo[v];
            return;
        }
        // local LF wave speeds: max over offsets {0, 1} of |eigenvalue| (shock_capturing.py:1000-1028)
        double lam[5];
        if (c->d.flux_type == OSB_FLUX_GLF) {
            // global wave speeds of this stage (shock_capturing.py:838-857): the
            // reductions follow the eigenvalue order of the direction
            const double* gl = c->glf[dir];
            if (nd == 1) { lam[0] = gl[2]; lam[1] = gl[0]; lam[2] = gl[1]; }
            else {
                for (int m = 0; m < nd; ++m) lam[m] = gl[0];
                lam[nd] = gl[1];
                lam[nd + 1] = gl[2];
            }
        } else if (nd == 1) {
            lam[0] = sfc * std::max(std::fabs(ud[2] - aa[2]), std::fabs(ud[3] - aa[3]));
            lam[1] = sfc * std::max(std::fabs(ud[2]), std::fabs(ud[3]));
            lam[2] = sfc * std::max(std::fabs(ud[2] + aa[2]), std::fabs(ud[3] + aa[3]));
        } else {
            const double l0 = sfc * std::max(std::fabs(ud[2]), std::fabs(ud[3]));
            for (int m = 0; m < nd; ++m) lam[m] = l0;
            lam[nd] = sfc * std::max(std::fabs(ud[2] + aa[2]), std::fabs(ud[3] + aa[3]));
            lam[nd + 1] = sfc * std::max(std::fabs(ud[2] - aa[2]), std::fabs(ud[3] - aa[3]));
        }
        double recon[5];
        for (int kk = 0; kk < nv; ++kk) {
            double fp[6], fm[6], cfv[6];
            for (int s = 0; s < 6; ++s) {
                double cs = 0.0, cf = 0.0;
                for (int m = 0; m < nv; ++m) { cs = std::fma(e.L[kk][m], qv[s][m], cs); cf = std::fma(e.L[kk][m], F[s][m], cf); }
                fp[s] = 0.5 * std::fma(lam[kk], cs, cf);   // f+ = (CF + lambda CS)/2
                fm[s] = 0.5 * std::fma(-lam[kk], cs, cf);  // f- = (CF - lambda CS)/2
                cfv[s] = cf;
            }
            if (filter) {
                // WENO filter (WENO_filter.py:406-470, shock_capturing.py:554-581): the dissipative
                // part of WENO-Z -- its interface value less the 6th-order central interpolation
                // of the characteristic fluxes -- times the sensor max(right, left)
                double s_r = 0.0, s_l = 0.0, rec, central;
                if (c->d.filter_order == 3) {
                    // WENO3-Z less the 4th-order central interpolation (-1, 7, 7, -1)/12
                    rec = wenoz3_right(fp, weps, &s_r) + wenoz3_left(fm, weps, &s_l);
                    central = (7.0 * (cfv[2] + cfv[3]) - (cfv[1] + cfv[4])) * (1.0 / 12.0);
                } else {
                    rec = wenoz_right(fp, weps, &s_r) + wenoz_left(fm, weps, &s_l);
                    central = (((((cfv[0] - 8.0 * cfv[1]) + 37.0 * cfv[2]) + 37.0 * cfv[3]) - 8.0 * cfv[4]) + cfv[5]) * (1.0 / 60.0);
                }
                recon[kk] = std::max(s_r, s_l) * (rec - central);
            } else if (teno) recon[kk] = teno_side(fp, +1, tp) + teno_side(fm, -1, tp);
            else recon[kk] = wenoz_right(fp, weps) + wenoz_left(fm, weps);
        }
        for (int v = 0; v < nv; ++v) {
            double s = 0.0;
            for (int kk = 0; kk < nv; ++kk) s = std::fma(e.Rm[v][kk], recon[kk], s);
            wk[(size_t)v * c->Ptot + x0] = s;
        }
    });
}

// residual R_v = -sum_d (wk_d(i) - wk_d(i-1)) * invDelta_d  (shock_capturing.py:23-36, opensblifunctions.py:577-591)
void shock_residual(osb_ctx* c) {
    const int nd = c->ndim, nv = c->nv;
    Range r = make_range(c, 0);
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        for (int v = 0; v < nv; ++v) {
            double acc = 0.0;
            for (int dir = 0; dir < nd; ++dir) {
                const double* w = c->wk.data() + ((size_t)dir * nv + v) * c->Ptot;
                acc -= (w[x] - w[x - c->stride(dir)]) * (1.0 / c->d.delta[dir]);
            }
            c->R[(size_t)v * c->Ptot + x] = acc;
        }
    });
}

// ---------------------------------------------------------------------------
// 4th-order central derivatives (scheme.py:83-87; opensblifunctions.py:486-512):
// d1 weights (1/12, -2/3, 0, 2/3, -1/12)/Delta, d2 weights (-1/12, 4/3, -5/2, 4/3, -1/12)/Delta^2
// ---------------------------------------------------------------------------
template <class G>
inline double d1(G f, double invd) {
    return std::fma(1.0 / 12.0, f(-2) - f(2), (2.0 / 3.0) * (f(1) - f(-1))) * invd;
}
template <class G>
inline double d2(G f, double inv2d) {
    return std::fma(-1.0 / 12.0, f(-2) + f(2), std::fma(4.0 / 3.0, f(-1) + f(1), (-5.0 / 2.0) * f(0))) * inv2d;
}

// Carpenter's 4th-order boundary closure (Carpenter_scheme.py:6-125; Carpenter,
// Nordstrom & Gottlieb, J. Comput. Phys. 148 (1999)), replacing the central
// derivatives at the grid points next to faces whose boundary condition is a
// ModifyCentralDerivative (ExtrapolationBC, DirichletBC; bc_core.py:15-18,
// opensblifunctions.py:486-534).  First derivative at boundary point r = 0..3
// (counted from the face): sum_j C1[r][j] f(point j) / Delta over points
// j = 0..5 from the face, negated on side 1; second derivative at r = 0, 1:
// sum_j C2[r][j] f(point j) / Delta^2 over j = 0..4.  C1 = al4^-1 ar4 of
// Carpenter_scheme.py:97-121 (row 0 simplified to the exact rationals as
// there), here correctly rounded from a 40-digit evaluation.
const double CARP_D1[4][6] = {
    {-11.0 / 6.0, 3.0, -1.5, 1.0 / 3.0, 0.0, 0.0},
    {-0.3762836775133548, -0.3224849328821596, 0.719443173328853, 0.03941685243994666, -0.0658051057710398, 0.005713690397754591},
    {0.11344647038423929, -0.7912455927658653, 0.19718433388773493, 0.521455851089594, -0.0367146847001282, -0.004126377895574793},
    {-0.009325979850499882, 0.12193715322406413, -0.7278221477245911, 0.045103322334387165, 0.6521410848612417, -0.08203343284460206}};
const double CARP_D2[2][5] = {{35.0 / 12.0, -104.0 / 12.0, 114.0 / 12.0, -56.0 / 12.0, 11.0 / 12.0},
                              {11.0 / 12.0, -20.0 / 12.0, 6.0 / 12.0, 4.0 / 12.0, -1.0 / 12.0}};

// the closure sums; dirn = +1 on side 0 (points j - r from x), -1 on side 1 (r - j)
template <class G>
inline double carp1(G f, int r, int dirn) {
    double s = CARP_D1[r][0] * f(-dirn * r);
    for (int j = 1; j < 6; ++j) s = std::fma(CARP_D1[r][j], f(dirn * (j - r)), s);
    return s;
}
template <class G>
inline double carp2(G f, int r, int dirn) {
    double s = CARP_D2[r][0] * f(-dirn * r);
    for (int j = 1; j < 5; ++j) s = std::fma(CARP_D2[r][j], f(dirn * (j - r)), s);
    return s;
}
// first / second derivative at index idx of n along a direction whose faces
// carry the closure where mask bit 0 (side 0) / bit 1 (side 1) is set
template <class G>
[[gnu::noinline]] double d1c_edge(G f, double invd, int idx, int n, int mask) {
    if ((mask & 1) && idx < 4) return carp1(f, idx, 1) * invd;
    if ((mask & 2) && idx >= n - 4) return -carp1(f, n - 1 - idx, -1) * invd;
    return d1(f, invd);
}
template <class G>
[[gnu::noinline]] double d2c_edge(G f, double inv2d, int idx, int n, int mask) {
    if ((mask & 1) && idx < 2) return carp2(f, idx, 1) * inv2d;
    if ((mask & 2) && idx >= n - 2) return carp2(f, n - 1 - idx, -1) * inv2d;
    return d2(f, inv2d);
}
template <class G>
inline double d1c(G f, double invd, int idx, int n, int mask) {
    if (__builtin_expect(mask != 0, 0)) return d1c_edge(f, invd, idx, n, mask);
    return d1(f, invd);
}
template <class G>
inline double d2c(G f, double inv2d, int idx, int n, int mask) {
    if (__builtin_expect(mask != 0, 0)) return d2c_edge(f, inv2d, idx, n, mask);
    return d2(f, inv2d);
}

// Central convective residual, written first (StoreSome.py:138-228; scheme.py:321-350).
void central_convective(osb_ctx* c) {
    const int nd = c->ndim;
    const size_t Pt = c->Ptot;
    const bool skew = c->d.conv_scheme == OSB_CONV_CENTRAL_SKEW;
    const bool kgp = c->d.conv_scheme == OSB_CONV_CENTRAL_KGP;
    Range r = make_range(c, 0);
    const double* q = c->q.data();
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        const double rho = q[x], rhoE = q[(size_t)(nd + 1) * Pt + x];
        double Rm[3] = {0, 0, 0}, Rrho = 0, RE = 0;
        double div_rhou = 0.0;            // sum_j D_j(rhou_j)
        double sum_conv[3] = {0, 0, 0};   // per momentum component
        double eterm = 0.0;
        for (int jd = 0; jd < nd; ++jd) {
            const size_t st = c->stride(jd);
            const double invd = 1.0 / c->d.delta[jd];
            auto at = [&](const double* arr, int s) { return arr[x + (long)s * (long)st]; };
            const int id = jd == 0 ? i : (jd == 1 ? j : k), nn = c->n[jd], cm = c->carp[jd];
            const double* rhou_j = q + (size_t)(1 + jd) * Pt;
            const double* u_j = c->u.data() + (size_t)jd * Pt;
            const double* rhoEa = q + (size_t)(nd + 1) * Pt;
            const double* pa = c->p.data();
            if (kgp) {
                // KGP with enthalpy splitting, alpha = beta = gamma = delta = 1/4
                // (split_forms.py:178-234; generated form in
                // apps/performance_benchmark/Central4-WENO5/opensbliblock00_kernels.h:157-382):
                // H = (rhoE + p)/rho, D_j u_j from the stored gradients
                const double ujc = at(u_j, 0), rhoujc = at(rhou_j, 0);
                const double gjj = c->g[(size_t)(jd * nd + jd) * Pt + x];
                auto Hs = [&](int s) { return (at(rhoEa, s) + at(pa, s)) / at(q, s); };
                const double d_rhouj = d1c([&](int s) { return at(rhou_j, s); }, invd, id, nn, cm);
                const double d_rho = d1c([&](int s) { return at(q, s); }, invd, id, nn, cm);
                Rrho += 0.5 * d_rhouj + 0.5 * (ujc * d_rho + rho * gjj);
                for (int m = 0; m < nd; ++m) {
                    const double* rhou_m = q + (size_t)(1 + m) * Pt;
                    const double* u_m = c->u.data() + (size_t)m * Pt;
                    const double umc = at(u_m, 0), rhoumc = at(rhou_m, 0);
                    const double g_mj = c->g[(size_t)(m * nd + jd) * Pt + x];
                    const double d_rhouj_um = d1c([&](int s) { return at(u_m, s) * at(rhou_j, s); }, invd, id, nn, cm);
                    const double d_rhoum = d1c([&](int s) { return at(rhou_m, s); }, invd, id, nn, cm);
                    const double d_ujum = d1c([&](int s) { return at(u_j, s) * at(u_m, s); }, invd, id, nn, cm);
                    sum_conv[m] += 0.25 * ((((((d_rhouj_um + umc * d_rhouj) + rhoujc * g_mj) + ujc * d_rhoum) + rhoumc * gjj) +
                                            rho * d_ujum) + (umc * ujc) * d_rho);
                }
                const double Hc = Hs(0);
                const double d_Hrhouj = d1c([&](int s) { return Hs(s) * at(rhou_j, s); }, invd, id, nn, cm);
                const double d_H = d1c(Hs, invd, id, nn, cm);
                const double d_Hrho = d1c([&](int s) { return Hs(s) * at(q, s); }, invd, id, nn, cm);
                const double d_Huj = d1c([&](int s) { return Hs(s) * at(u_j, s); }, invd, id, nn, cm);
                eterm += 0.25 * ((((((d_Hrhouj + Hc * d_rhouj) + rhoujc * d_H) + ujc * d_Hrho) + (rho * Hc) * gjj) + rho * d_Huj) +
                                 (ujc * Hc) * d_rho);
            } else if (!skew) {
                // Feiereisen (split_forms.py:111-144)
                const double d_rhouj = d1c([&](int s) { return at(rhou_j, s); }, invd, id, nn, cm);
                div_rhou += d_rhouj;
                for (int m = 0; m < nd; ++m) {
                    const double* rhou_m = q + (size_t)(1 + m) * Pt;
                    const double d_prod = d1c([&](int s) { return at(rhou_m, s) * at(u_j, s); }, invd, id, nn, cm);
                    const double g_mj = c->g[(size_t)(m * nd + jd) * Pt + x];   // stored D_j u_m
                    sum_conv[m] += d_prod + at(rhou_j, 0) * g_mj;
                }
                const double d_rhoEu = d1c([&](int s) { return at(rhoEa, s) * at(u_j, s); }, invd, id, nn, cm);
                const double d_E = d1c([&](int s) { return at(rhoEa, s) / at(q, s); }, invd, id, nn, cm);
                const double d_pu = d1c([&](int s) { return at(pa, s) * at(u_j, s); }, invd, id, nn, cm);
                eterm += -0.5 * (d_rhoEu + at(rhou_j, 0) * d_E) - d_pu;
            } else {
                // Skew(ab, x) = 1/2 (D(ab) + b D(a) + a D(b))  (parsing.py:96-112), tg_sym equations
                const double* rhoa = q;
                const double ujc = at(u_j, 0);
                const double d_rhou = d1c([&](int s) { return at(rhoa, s) * at(u_j, s); }, invd, id, nn, cm);
                const double d_rho = d1c([&](int s) { return at(rhoa, s); }, invd, id, nn, cm);
                const double d_u = d1c([&](int s) { return at(u_j, s); }, invd, id, nn, cm);
                Rrho -= 0.5 * (d_rhou + ujc * d_rho + rho * d_u);
                for (int m = 0; m < nd; ++m) {
                    const double* rhou_m = q + (size_t)(1 + m) * Pt;
                    const double d_prod = d1c([&](int s) { return at(rhou_m, s) * at(u_j, s); }, invd, id, nn, cm);
                    const double d_rm = d1c([&](int s) { return at(rhou_m, s); }, invd, id, nn, cm);
                    sum_conv[m] += 0.5 * (d_prod + ujc * d_rm + at(rhou_m, 0) * d_u);
                }
                const double d_Eu = d1c([&](int s) { return at(rhoEa, s) * at(u_j, s); }, invd, id, nn, cm);
                const double d_rE = d1c([&](int s) { return at(rhoEa, s); }, invd, id, nn, cm);
                const double d_pu = d1c([&](int s) { return at(pa, s) * at(u_j, s); }, invd, id, nn, cm);
                eterm += -0.5 * (d_Eu + ujc * d_rE + rhoE * d_u) - d_pu;
            }
        }
        for (int m = 0; m < nd; ++m) {
            const size_t st = c->stride(m);
            const double invd = 1.0 / c->d.delta[m];
            const double* pa = c->p.data();
            const int id = m == 0 ? i : (m == 1 ? j : k);
            const double dp = d1c([&](int s) { return pa[x + (long)s * (long)st]; }, invd, id, c->n[m], c->carp[m]);
            if (!skew && !kgp) {
                const double um = c->u[(size_t)m * Pt + x];
                Rm[m] = -dp - 0.5 * (sum_conv[m] + um * div_rhou);
            } else {
                Rm[m] = -dp - sum_conv[m];
            }
        }
        if (kgp) {
            Rrho = -Rrho;
            RE = -eterm;
        } else if (!skew) {
            Rrho = -div_rhou;
            RE = eterm - 0.5 * (rhoE / rho) * div_rhou;
        } else {
            RE = eterm;
        }
        c->R[x] = Rrho;
        for (int m = 0; m < nd; ++m) c->R[(size_t)(1 + m) * Pt + x] = Rm[m];
        c->R[(size_t)(nd + 1) * Pt + x] = RE;
    });
}

// Stored first derivatives of u_i (StoreSome.py:33-45, 87-135), evaluated on
// ranges extended by 2 across so the mixed derivatives can differentiate them
// (TENO6/opensbli.cpp:340-384).
void stored_gradients(osb_ctx* c) {
    const int nd = c->ndim;
    const size_t Pt = c->Ptot;
    Range r = make_range(c, 2);
    for (int jd = 0; jd < nd; ++jd) {
        Range rr = r;
        rr.lo[jd] = 0; rr.hi[jd] = c->n[jd];
        const size_t st = c->stride(jd);
        const double invd = 1.0 / c->d.delta[jd];
        for (int m = 0; m < nd; ++m) {
            const double* um = c->u.data() + (size_t)m * Pt;
            double* gm = c->g.data() + (size_t)(m * nd + jd) * Pt;
            for_range(rr, [&](int i, int j, int k) {
                const size_t x = c->I(i, j, k);
                const int id = jd == 0 ? i : (jd == 1 ? j : k);
                gm[x] = d1c([&](int s) { return um[x + (long)s * (long)st]; }, invd, id, c->n[jd], c->carp[jd]);
            });
        }
    }
}

// Viscous and heat-conduction terms, added to the residual
// (split_forms.py:383-392; generated TENO6/opensbliblock00_kernels.h:4568-4707).
// Mixed second derivatives are D_b(D_a u) with a < b applied to the stored
// first derivative; repeated-direction second derivatives use the d2 stencil.
void viscous_terms(osb_ctx* c) {
    const int nd = c->ndim;
    const size_t Pt = c->Ptot;
    const bool suth = c->d.viscous == OSB_VISC_SUTHERLAND;
    const double invRe = 1.0 / c->d.re;
    const double gama = c->d.gama, minf = c->d.minf, pr = c->d.pr;
    const double cheat = 1.0 / ((gama - 1.0) * minf * minf * pr * c->d.re);
    Range r = make_range(c, 0);
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        double G[3][3] = {{0}};                 // G[m][jd] = D_jd u_m at x
        for (int m = 0; m < nd; ++m)
            for (int jd = 0; jd < nd; ++jd) G[m][jd] = c->g[(size_t)(m * nd + jd) * Pt + x];
        double dmu[3] = {0, 0, 0}, dT[3] = {0, 0, 0}, lapT = 0.0;
        double D2u[3][3] = {{0}};                // D2u[m][jd] = d2 u_m / d x_jd^2
        for (int jd = 0; jd < nd; ++jd) {
            const size_t st = c->stride(jd);
            const double invd = 1.0 / c->d.delta[jd];
            const double inv2d = 1.0 / (c->d.delta[jd] * c->d.delta[jd]);
            auto at = [&](const double* arr, int s) { return arr[x + (long)s * (long)st]; };
            const int id = jd == 0 ? i : (jd == 1 ? j : k), nn = c->n[jd], cm = c->carp[jd];
            if (suth) dmu[jd] = d1c([&](int s) { return at(c->mu.data(), s); }, invd, id, nn, cm);
            dT[jd] = d1c([&](int s) { return at(c->T.data(), s); }, invd, id, nn, cm);
            lapT += d2c([&](int s) { return at(c->T.data(), s); }, inv2d, id, nn, cm);
            for (int m = 0; m < nd; ++m) {
                const double* um = c->u.data() + (size_t)m * Pt;
                D2u[m][jd] = d2c([&](int s) { return at(um, s); }, inv2d, id, nn, cm);
            }
        }
        // mixed(m, a, b) = D_b(D_a u_m), a < b
        auto mixed = [&](int m, int a, int b) {
            if (a > b) std::swap(a, b);
            const double* ga = c->g.data() + (size_t)(m * nd + a) * Pt;
            const size_t st = c->stride(b);
            const int id = b == 0 ? i : (b == 1 ? j : k);
            return d1c([&](int s) { return ga[x + (long)s * (long)st]; }, 1.0 / c->d.delta[b], id, c->n[b], c->carp[b]);
        };
        double div = 0.0;
        for (int m = 0; m < nd; ++m) div += G[m][m];
        double S[3][3];
        for (int m = 0; m < nd; ++m)
            for (int jd = 0; jd < nd; ++jd)
                S[m][jd] = G[m][jd] + G[jd][m] - (m == jd ? (2.0 / 3.0) * div : 0.0);
        // L_m = d(S_mj)/dx_j = (4/3) d2u_m/dx_m^2 + sum_{j!=m} d2u_m/dx_j^2 + (1/3) sum_{j!=m} d2u_j/dx_m dx_j
        double Lm[3] = {0, 0, 0};
        for (int m = 0; m < nd; ++m) {
            double s = (4.0 / 3.0) * D2u[m][m];
            for (int jd = 0; jd < nd; ++jd) {
                if (jd == m) continue;
                s += D2u[m][jd] + (1.0 / 3.0) * mixed(jd, jd, m);
            }
            Lm[m] = s;
        }
        const double muv = suth ? c->mu[x] : 1.0;
        double Eacc = 0.0;
        for (int m = 0; m < nd; ++m) {
            double mom = muv * Lm[m];
            if (suth) for (int jd = 0; jd < nd; ++jd) mom += S[m][jd] * dmu[jd];
            c->R[(size_t)(1 + m) * Pt + x] += invRe * mom;
            const double um = c->u[(size_t)m * Pt + x];
            double e = um * mom;                           // u_i d(tau_ij)/dx_j
            for (int jd = 0; jd < nd; ++jd) e += muv * S[m][jd] * G[m][jd];   // tau_ij du_i/dx_j
            Eacc += e;
        }
        double heat = muv * lapT;
        if (suth) for (int jd = 0; jd < nd; ++jd) heat += dT[jd] * dmu[jd];
        c->R[(size_t)(nd + 1) * Pt + x] += invRe * Eacc + cheat * heat;
    });
}

// ---------------------------------------------------------------------------
// Boundary conditions, applied direction by direction, side 0 then side 1,
// over the halo-extended range of the other directions
// (bc_core.py:128-202; algorithm.py:460-462).
// ---------------------------------------------------------------------------
void apply_bcs(osb_ctx* c) {
    const int nd = c->ndim, nv = c->nv;
    const int h = c->depth;
    const size_t Pt = c->Ptot;
    for (int dir = 0; dir < nd; ++dir) {
        const int N = c->n[dir];
        const size_t st = c->stride(dir);
        for (int side = 0; side < 2; ++side) {
            const int type = c->d.bc[dir][side];
            Range r = make_range(c, h);
            r.lo[dir] = side == 0 ? 0 : N - 1;
            r.hi[dir] = r.lo[dir] + 1;
            const int to = side == 0 ? 1 : -1;       // set_side_factor
            for_range(r, [&](int i, int j, int k) {
                const size_t xb = c->I(i, j, k);    // boundary point
                auto cp = [&](long dst, long src, bool reflect_half, bool reflect_full) {
                    for (int v = 0; v < nv; ++v) {
                        double val = c->q[(size_t)v * Pt + (size_t)((long)xb + src * (long)st)];
                        if (v == 1 + dir) {
                            if (reflect_full) val = val - 2. * val;     // v - 2(v.n)n
                            else if (reflect_half) val = val - 1. * val; // v - (v.n)n
                        }
                        c->q[(size_t)v * Pt + (size_t)((long)xb + dst * (long)st)] = val;
                    }
                };
                if (type == OSB_BC_PERIODIC) {
                    // ExchangeSelf: side 0 copies [0,h) -> [N,N+h); side 1 copies [N-h,N) -> [-h,0)
                    for (int t = 0; t < h; ++t) {
                        if (side == 0) cp((long)N + t, t, false, false);
                        else cp(-(long)(N - 1) - h + t, (long)(1 - h + t), false, false);
                    }
                } else if (type == OSB_BC_ZERO_GRADIENT) {
                    cp(0, to, false, false);                               // q[0] = q[1]
                    for (int t = 1; t <= h; ++t) cp(-(long)to * t, (long)to * t, false, false);
                } else if (type == OSB_BC_INVISCID_WALL) {
                    for (int t = 1; t <= h; ++t) cp(-(long)to * t, (long)to * t, false, true);
                    cp(0, to, true, false);                                // boundary point from q[1]
                } else if (type == OSB_BC_SYMMETRY) {
                    for (int t = 1; t <= h; ++t) cp(-(long)to * t, (long)to * t, false, true);
                } else if (type == OSB_BC_EXTRAPOLATION0) {
                    // extrapolation.py:36-40: points 0, -1, .., -h (side 0) take q[1]
                    for (int t = 0; t <= h; ++t) cp(-(long)to * t, to, false, false);
                } else if (type == OSB_BC_EXTRAPOLATION1) {
                    // extrapolation.py:41-54: q[-t] = 2.0*q[-t+1] - q[-t+2], t = 1..h in
                    // order, so each layer reads the one written before it
                    for (int t = 1; t <= h; ++t)
                        for (int v = 0; v < nv; ++v) {
                            double* qv = c->q.data() + (size_t)v * Pt;
                            const double one = qv[(long)xb - (long)to * (t - 1) * (long)st];
                            const double two = qv[(long)xb - (long)to * (t - 2) * (long)st];
                            qv[(long)xb - (long)to * t * (long)st] = 2.0 * one - two;
                        }
                } else if (type == OSB_BC_DIRICHLET) {
                    // dirichlet.py:28-41: the boundary point and the halo range of this side
                    for (int t = 0; t <= h; ++t)
                        for (int v = 0; v < nv; ++v)
                            c->q[(size_t)v * Pt + (size_t)((long)xb - (long)to * t * (long)st)] = c->d.bc_value[dir][side][v];
                }
            });
        }
    }
}

// ---------------------------------------------------------------------------
// WENO filter step after a full RK step (WENO_filter.py:406-536; placed in the
// time loop after the stage loop, algorithm.py:428-450): constituent relations
// of the current state, the filtered interface fluxes per direction, and
// q += dt * R_filter with R_filter = -sum_d (F_d(i) - F_d(i-1)) / Delta_d.
// ---------------------------------------------------------------------------
// Modified Ducros sensor (shock_sensors.py:16-59; generated
// Central4-WENO5/opensbliblock00_kernels.h:572-617) over [0, N):
// kappa = div^2 (1/2 - 1/2 tanh(2.5 + 500 div / sqrt(sum_d Delta_d^2))) /
//         (div^2 + |curl u|^2 + 1e-40), central derivatives of u.
void ducros_sensor(osb_ctx* c) {
    const int nd = c->ndim;
    const size_t Pt = c->Ptot;
    double sd = 0.0;
    for (int d = 0; d < nd; ++d) sd += c->d.delta[d] * c->d.delta[d];
    sd = std::sqrt(sd);
    Range r = make_range(c, 0);
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        const int ii[3] = {i, j, k};
        double G[3][3] = {{0}};              // G[m][d] = D_d u_m
        for (int m = 0; m < nd; ++m)
            for (int d = 0; d < nd; ++d) {
                const double* um = c->u.data() + (size_t)m * Pt;
                const size_t st = c->stride(d);
                G[m][d] = d1c([&](int s) { return um[x + (long)s * (long)st]; }, 1.0 / c->d.delta[d], ii[d], c->n[d], c->carp[d]);
            }
        double div = 0.0;
        for (int d = 0; d < nd; ++d) div += G[d][d];
        double w2;
        if (nd == 2) {
            const double wz = G[1][0] - G[0][1];
            w2 = wz * wz;
        } else {
            const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
            w2 = (wx * wx + wy * wy) + wz * wz;
        }
        const double th = 0.5 - 0.5 * std::tanh(2.5 + (500.0 * div) / sd);
        c->kappa[x] = ((div * div) * th) / ((div * div + w2) + 1.0e-40);
    });
}

void apply_weno_filter(osb_ctx* c, double dt) {
    Range rcr = make_range(c, 3);
    for (int d = 0; d < c->ndim; ++d) rcr.hi[d] = c->n[d] + 4;
    constituent_relations(c, rcr, false);
    const bool hybrid = c->d.filter_sensor == OSB_SENSOR_DUCROS;
    if (hybrid) ducros_sensor(c);
    for (int dir = 0; dir < c->ndim; ++dir) reconstruct_direction(c, dir, true);
    const int nd = c->ndim, nv = c->nv;
    Range r = make_range(c, 0);
    for_range(r, [&](int i, int j, int k) {
        const size_t x = c->I(i, j, k);
        if (hybrid) {
            // WENO_filter.py:486-500: filter where max kappa over the point and its
            // -1, +1, +2 neighbours in every direction reaches Ducros_select
            double km = c->kappa[x];
            for (int d = 0; d < nd; ++d) {
                const long st = (long)c->stride(d);
                km = std::max(km, std::max(c->kappa[x - st], std::max(c->kappa[x + st], c->kappa[x + 2 * st])));
            }
            if (!(km >= c->d.ducros_select)) return;
        }
        for (int v = 0; v < nv; ++v) {
            double acc = 0.0;
            for (int dir = 0; dir < nd; ++dir) {
                const double* w = c->wk.data() + ((size_t)dir * nv + v) * c->Ptot;
                acc -= (w[x] - w[x - c->stride(dir)]) * (1.0 / c->d.delta[dir]);
            }
            c->q[(size_t)v * c->Ptot + x] = c->q[(size_t)v * c->Ptot + x] + dt * acc;
        }
    });
}

// ---------------------------------------------------------------------------
// "Global wave-speed reductions" kernel of GLF splitting (weno.py:515-519,
// teno.py:520-524; shock_capturing.py:416-438, :838-857): per direction the
// maxima over the grid [0, N) of |u_d|, |u_d + a|, |u_d - a| from the
// constituent relations of this stage.
// ---------------------------------------------------------------------------
void glf_reductions(osb_ctx* c) {
    const Range r = make_range(c, 0);
    for (int d = 0; d < c->ndim; ++d) {
        double m0 = 0.0, m1 = 0.0, m2 = 0.0;
        const double* ud = c->u.data() + (size_t)d * c->Ptot;
        for (int k = r.lo[2]; k < r.hi[2]; ++k)          // serial: a max is order-free anyway
            for (int j = r.lo[1]; j < r.hi[1]; ++j)
                for (int i = r.lo[0]; i < r.hi[0]; ++i) {
                    const size_t x = c->I(i, j, k);
                    m0 = std::max(m0, std::fabs(ud[x]));
                    m1 = std::max(m1, std::fabs(ud[x] + c->a[x]));
                    m2 = std::max(m2, std::fabs(ud[x] - c->a[x]));
                }
        c->glf[d][0] = m0; c->glf[d][1] = m1; c->glf[d][2] = m2;
    }
}

// ---------------------------------------------------------------------------
// Right-hand side at the current state (halos already filled).
// ---------------------------------------------------------------------------
void compute_rhs(osb_ctx* c) {
    const int ext = c->shock ? 3 : 2;
    Range rcr = make_range(c, ext);
    if (c->shock) {
        // u,p,a over [-3, N+4) (TENO6/opensbli.cpp:219-240)
        for (int d = 0; d < c->ndim; ++d) rcr.hi[d] = c->n[d] + 4;
    }
    constituent_relations(c, rcr, c->viscous || c->central);
    if (c->shock) {
        if (c->d.flux_type == OSB_FLUX_GLF) glf_reductions(c);
        for (int dir = 0; dir < c->ndim; ++dir) reconstruct_direction(c, dir);
        shock_residual(c);
    }
    if (c->central || c->viscous) stored_gradients(c);
    if (c->central) central_convective(c);
    if (c->viscous) viscous_terms(c);
}

double max_wavespeed(osb_ctx* c) {
    // CFL sum over directions of (|u_d| + a)/Delta_d at interior points, from the current state
    const int nd = c->ndim;
    const double gama = c->d.gama;
    Range r = make_range(c, 0);
    const int nk = r.hi[2] - r.lo[2], nj = r.hi[1] - r.lo[1];
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long t = 0; t < (long)nk * nj; ++t) {
        const int k = (int)(t / nj), j = (int)(t % nj);
        for (int i = 0; i < c->n[0]; ++i) {
            const size_t x = c->I(i, j, k);
            const double rho = c->q[x];
            double uu[3] = {0, 0, 0}, ke = 0.0;
            for (int d = 0; d < nd; ++d) { uu[d] = c->q[(size_t)(1 + d) * c->Ptot + x] / rho; ke -= 0.5 * (uu[d] * uu[d]) * rho; }
            const double pp = (gama - 1.0) * (ke + c->q[(size_t)(nd + 1) * c->Ptot + x]);
            const double a = std::sqrt(gama * pp / rho);
            double s = 0.0;
            for (int d = 0; d < nd; ++d) s += (std::fabs(uu[d]) + a) * (1.0 / c->d.delta[d]);
            m = std::max(m, s);
        }
    }
    return m;
}

// RK point updates.  Low-storage RK3 (rk_LS.py:121-124, 204-215):
// du = A*du + dt*R; q = q + B*du.  Classic RK3 (rk_sbli.py:69-76, 129-159):
// q = qold + dt*rknew*R; qold = qold + dt*rkold*R.
// RungeKuttaLS(3) Williamson coefficients (rk_LS.py:121-124)
const double RKLS_A[3] = {0, (-5.0 / 9.0), (-153.0 / 128.0)};
const double RKLS_B[3] = {(1.0 / 3.0), (15.0 / 16.0), (8.0 / 15.0)};
const double RK3_OLD[3] = {1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0};
const double RK3_NEW[3] = {2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0};
// rk_LS.py:204-215: du = rkA*du + dt*R; q = q + rkB*du
inline void rkls_point(double A, double B, double dt, double R, double& du, double& q) {
    du = std::fma(A, du, dt * R);
    q = std::fma(B, du, q);
}
inline void rk3_point(int s, double dt, double R, double& qold, double& q) {
    q = std::fma(dt * RK3_NEW[s], R, qold);
    qold = std::fma(dt * RK3_OLD[s], R, qold);
}

int validate(const osb_problem_desc* d, std::string& why) {
    if (!d) { why = "null descriptor"; return OSB_E_INVALID; }
    if (d->struct_size != sizeof(osb_problem_desc) || d->abi_version != OSB_ABI_VERSION) { why = "descriptor size/abi mismatch"; return OSB_E_INVALID; }
    if (d->ndim < 1 || d->ndim > 3) { why = "ndim must be 1..3"; return OSB_E_INVALID; }
    for (int i = 0; i < 3; ++i) {
        if (i < d->ndim && (d->n[i] < 8 || !(d->delta[i] > 0))) { why = "n >= 8 and delta > 0 required in active directions"; return OSB_E_INVALID; }
        if (i >= d->ndim && d->n[i] != 1) { why = "inactive directions must have n == 1"; return OSB_E_INVALID; }
    }
    const int cs = d->conv_scheme;
    const bool central_cs = cs == OSB_CONV_CENTRAL_FEIEREISEN || cs == OSB_CONV_CENTRAL_SKEW || cs == OSB_CONV_CENTRAL_KGP;
    if (cs != OSB_CONV_WENO_Z && cs != OSB_CONV_TENO && !central_cs) { why = "unknown conv_scheme"; return OSB_E_UNSUPPORTED; }
    if (cs == OSB_CONV_WENO_Z && d->order != 5) { why = "WENO-Z supports order 5"; return OSB_E_UNSUPPORTED; }
    if (cs == OSB_CONV_TENO && d->order != 5 && d->order != 6) { why = "TENO supports orders 5 and 6"; return OSB_E_UNSUPPORTED; }
    if (central_cs && d->order != 4) { why = "central scheme supports order 4"; return OSB_E_UNSUPPORTED; }
    if ((cs == OSB_CONV_WENO_Z || cs == OSB_CONV_TENO) && d->averaging != OSB_AVG_SIMPLE && d->averaging != OSB_AVG_ROE) { why = "unknown averaging"; return OSB_E_UNSUPPORTED; }
    if (d->viscous != OSB_VISC_NONE && d->viscous != OSB_VISC_CONSTANT && d->viscous != OSB_VISC_SUTHERLAND) { why = "unknown viscous model"; return OSB_E_UNSUPPORTED; }
    if (d->viscous != OSB_VISC_NONE && d->ndim < 2) { why = "viscous terms need ndim >= 2"; return OSB_E_UNSUPPORTED; }
    if (d->time_scheme != OSB_TIME_RKLS3 && d->time_scheme != OSB_TIME_RK3) { why = "unknown time scheme"; return OSB_E_UNSUPPORTED; }
    if (d->precision != OSB_PREC_FP64 && d->precision != OSB_PREC_MIXED) { why = "unknown precision"; return OSB_E_UNSUPPORTED; }
    if (d->flux_type < OSB_FLUX_LLF || d->flux_type > OSB_FLUX_HLLC_LM) { why = "unknown flux type"; return OSB_E_UNSUPPORTED; }
    if (d->flux_type >= OSB_FLUX_HLLC && cs != OSB_CONV_WENO_Z) { why = "HLLC is built for the WENO-Z scheme (HLLCWeno)"; return OSB_E_UNSUPPORTED; }
    if (d->flux_type >= OSB_FLUX_HLLC && d->precision != OSB_PREC_FP64) { why = "HLLC runs in FP64 only"; return OSB_E_UNSUPPORTED; }
    if (d->shock_filter != OSB_FILTER_NONE && d->shock_filter != OSB_FILTER_WENOZ) { why = "unknown shock filter"; return OSB_E_UNSUPPORTED; }
    if (d->shock_filter == OSB_FILTER_WENOZ && d->precision != OSB_PREC_FP64) { why = "the WENO filter runs in FP64 only"; return OSB_E_UNSUPPORTED; }
    if (d->shock_filter == OSB_FILTER_WENOZ && d->flux_type != OSB_FLUX_LLF) { why = "the WENO filter runs with LLF splitting of the base scheme only"; return OSB_E_UNSUPPORTED; }
    if (d->shock_filter == OSB_FILTER_WENOZ) {
        if (d->filter_order != 0 && d->filter_order != 3 && d->filter_order != 5) { why = "WENO filter orders 3 and 5"; return OSB_E_UNSUPPORTED; }
        if (d->filter_sensor != OSB_SENSOR_NONE && d->filter_sensor != OSB_SENSOR_DUCROS) { why = "unknown filter sensor"; return OSB_E_UNSUPPORTED; }
        if (d->filter_sensor == OSB_SENSOR_DUCROS && d->ndim < 2) { why = "the Ducros sensor needs ndim >= 2 (shock_sensors.py:42-43)"; return OSB_E_UNSUPPORTED; }
        if (d->filter_sensor == OSB_SENSOR_DUCROS && (!std::isfinite(d->ducros_check) || !std::isfinite(d->ducros_select))) { why = "non-finite Ducros thresholds"; return OSB_E_INVALID; }
    }
    if (d->flux_type == OSB_FLUX_GLF && cs != OSB_CONV_WENO_Z && cs != OSB_CONV_TENO) { why = "GLF applies to shock capturing only"; return OSB_E_INVALID; }
    if (d->rk_stages < 0 || d->rk_stages > 5) { why = "rk_stages must be 0..5"; return OSB_E_INVALID; }
    if (d->rk_stages > 0 && d->time_scheme != OSB_TIME_RKLS3) { why = "rk_stages applies to OSB_TIME_RKLS3 only"; return OSB_E_INVALID; }
    for (int s = 0; s < d->rk_stages; ++s)
        if (!std::isfinite(d->rk_a[s]) || !std::isfinite(d->rk_b[s])) { why = "non-finite RK coefficient"; return OSB_E_INVALID; }
    for (int i = 0; i < d->ndim; ++i) {
        for (int s = 0; s < 2; ++s) {
            const int b = d->bc[i][s];
            if (b < OSB_BC_PERIODIC || b > OSB_BC_DIRICHLET) { why = "unknown boundary condition"; return OSB_E_UNSUPPORTED; }
            if (b >= OSB_BC_EXTRAPOLATION0 && (central_cs || d->viscous != OSB_VISC_NONE) && d->precision != OSB_PREC_FP64) {
                why = "the Carpenter boundary closure of the central derivatives runs in FP64 only";
                return OSB_E_UNSUPPORTED;
            }
            if (b == OSB_BC_DIRICHLET)
                for (int v = 0; v < d->ndim + 2; ++v)
                    if (!std::isfinite(d->bc_value[i][s][v])) { why = "non-finite Dirichlet value"; return OSB_E_INVALID; }
        }
        if ((d->bc[i][0] == OSB_BC_PERIODIC) != (d->bc[i][1] == OSB_BC_PERIODIC)) { why = "periodic boundaries must be paired"; return OSB_E_INVALID; }
    }
    if (!(d->cfl > 0) && !(d->dt > 0)) { why = "dt > 0 or cfl > 0 required"; return OSB_E_INVALID; }
    if (!(d->gama > 1.0)) { why = "gama > 1 required"; return OSB_E_INVALID; }
    if (d->viscous != OSB_VISC_NONE && (!(d->re > 0) || !(d->pr > 0) || !(d->minf > 0))) { why = "Re, Pr, Minf > 0 required"; return OSB_E_INVALID; }
    return OSB_OK;
}

}  // namespace

extern "C" {

int osb_abi_version(void) { return OSB_ABI_VERSION; }

// host bytes of the arrays a context allocates (the oracle's own layout)
int osb_estimate_bytes(const osb_problem_desc* desc, int64_t* out) {
    if (!out) return fail(nullptr, OSB_E_INVALID, "null output pointer");
    std::string why;
    int rc = validate(desc, why);
    if (rc) return fail(nullptr, rc, why);
    const int nd = desc->ndim, nv = nd + 2;
    size_t P = 1;
    for (int i = 0; i < nd; ++i) P *= (size_t)(desc->n[i] + 2 * HALO);
    const bool shock = desc->conv_scheme == OSB_CONV_WENO_Z || desc->conv_scheme == OSB_CONV_TENO;
    const bool cv = !shock || desc->viscous != OSB_VISC_NONE;
    size_t doubles = 3 * (size_t)nv * P + (size_t)(nd + 2) * P;
    if (cv) doubles += 2 * P + (size_t)nd * nd * P;
    if (shock) doubles += (size_t)nd * nv * P;
    *out = (int64_t)(doubles * sizeof(double));
    return OSB_OK;
}

const char* osb_last_error(const osb_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int osb_ctx_create(const osb_problem_desc* desc, osb_ctx** out) {
    if (!out) return fail(nullptr, OSB_E_INVALID, "null output pointer");
    *out = nullptr;
    std::string why;
    int rc = validate(desc, why);
    if (rc) return fail(nullptr, rc, why);
    osb_ctx* c = new osb_ctx();
    c->d = *desc;
    c->ndim = desc->ndim;
    c->nv = desc->ndim + 2;
    for (int i = 0; i < 3; ++i) {
        c->n[i] = desc->n[i];
        if (i < c->ndim) { c->P[i] = desc->n[i] + 2 * HALO; c->off[i] = HALO; }
    }
    c->Ptot = (size_t)c->P[0] * c->P[1] * c->P[2];
    c->shock = desc->conv_scheme == OSB_CONV_WENO_Z || desc->conv_scheme == OSB_CONV_TENO;
    c->central = !c->shock;
    c->viscous = desc->viscous != OSB_VISC_NONE;
    for (int d = 0; d < c->ndim; ++d)
        for (int sd = 0; sd < 2; ++sd)
            if (desc->bc[d][sd] >= OSB_BC_EXTRAPOLATION0) c->carp[d] |= 1 << sd;
    // WenoHalos/TenoHalos [-3,4] -> 4; CentralHalos(4) -> 2; the WENO filter adds
    // its WENO halos to the block's (WENO_filter.py:440-441)
    c->depth = (c->shock || desc->shock_filter == OSB_FILTER_WENOZ) ? 4 : 2;
    if (desc->rk_stages == 0) {
        c->rk_stages = 3;
        for (int s = 0; s < 3; ++s) { c->rkA[s] = RKLS_A[s]; c->rkB[s] = RKLS_B[s]; }
    } else {
        c->rk_stages = desc->rk_stages;
        for (int s = 0; s < desc->rk_stages; ++s) { c->rkA[s] = desc->rk_a[s]; c->rkB[s] = desc->rk_b[s]; }
    }
    const size_t nvP = (size_t)c->nv * c->Ptot;
    c->q.assign(nvP, 0.0);
    c->rkold.assign(nvP, 0.0);
    c->R.assign(nvP, 0.0);
    c->u.assign((size_t)c->ndim * c->Ptot, 0.0);
    c->p.assign(c->Ptot, 0.0);
    c->a.assign(c->Ptot, 0.0);
    if (c->viscous || c->central) { c->T.assign(c->Ptot, 0.0); c->mu.assign(c->Ptot, 0.0); }
    if (c->shock || desc->shock_filter == OSB_FILTER_WENOZ) c->wk.assign((size_t)c->ndim * nvP, 0.0);
    if (desc->shock_filter == OSB_FILTER_WENOZ && desc->filter_sensor == OSB_SENSOR_DUCROS) c->kappa.assign(c->Ptot, 0.0);
    if (c->viscous || c->central) c->g.assign((size_t)c->ndim * c->ndim * c->Ptot, 0.0);
    *out = c;
    return OSB_OK;
}

int osb_ctx_destroy(osb_ctx* ctx) { delete ctx; return OSB_OK; }

int osb_set_state(osb_ctx* c, const double* host, int layout) {
    if (!c || !host) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR && layout != OSB_LAYOUT_PADDED) return fail(c, OSB_E_INVALID, "unknown layout");
    c->next_stage = 0;
    std::fill(c->rkold.begin(), c->rkold.end(), 0.0);
    if (layout == OSB_LAYOUT_PADDED) {            // halos are refilled by the boundary conditions
        std::copy(host, host + c->q.size(), c->q.begin());
        apply_bcs(c);
        return OSB_OK;
    }
    std::fill(c->q.begin(), c->q.end(), 0.0);
    const size_t npts = (size_t)c->n[0] * c->n[1] * c->n[2];
    for (int v = 0; v < c->nv; ++v)
        for (int k = 0; k < c->n[2]; ++k)
            for (int j = 0; j < c->n[1]; ++j)
                for (int i = 0; i < c->n[0]; ++i)
                    c->q[(size_t)v * c->Ptot + c->I(i, j, k)] = host[(size_t)v * npts + (size_t)i + (size_t)c->n[0] * ((size_t)j + (size_t)c->n[1] * k)];
    apply_bcs(c);
    return OSB_OK;
}

int osb_get_state(osb_ctx* c, double* host, int layout) {
    if (!c || !host) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR && layout != OSB_LAYOUT_PADDED) return fail(c, OSB_E_INVALID, "unknown layout");
    if (layout == OSB_LAYOUT_PADDED) {
        std::copy(c->q.begin(), c->q.end(), host);
        return OSB_OK;
    }
    const size_t npts = (size_t)c->n[0] * c->n[1] * c->n[2];
    for (int v = 0; v < c->nv; ++v)
        for (int k = 0; k < c->n[2]; ++k)
            for (int j = 0; j < c->n[1]; ++j)
                for (int i = 0; i < c->n[0]; ++i)
                    host[(size_t)v * npts + (size_t)i + (size_t)c->n[0] * ((size_t)j + (size_t)c->n[1] * k)] = c->q[(size_t)v * c->Ptot + c->I(i, j, k)];
    return OSB_OK;
}

int osb_set_time(osb_ctx* c, double time, int64_t iter) {
    if (!c) return OSB_E_INVALID;
    c->time = time; c->iter = iter;
    return OSB_OK;
}

int osb_apply_bcs(osb_ctx* c) { if (!c) return OSB_E_INVALID; apply_bcs(c); return OSB_OK; }

int osb_residual(osb_ctx* c, double* host_out) {
    if (!c || !host_out) return OSB_E_INVALID;
    apply_bcs(c);
    compute_rhs(c);
    const size_t npts = (size_t)c->n[0] * c->n[1] * c->n[2];
    for (int v = 0; v < c->nv; ++v)
        for (int k = 0; k < c->n[2]; ++k)
            for (int j = 0; j < c->n[1]; ++j)
                for (int i = 0; i < c->n[0]; ++i)
                    host_out[(size_t)v * npts + (size_t)i + (size_t)c->n[0] * ((size_t)j + (size_t)c->n[1] * k)] = c->R[(size_t)v * c->Ptot + c->I(i, j, k)];
    return OSB_OK;
}

int osb_max_wavespeed(osb_ctx* c, double* out) {
    if (!c || !out) return OSB_E_INVALID;
    *out = max_wavespeed(c);
    return OSB_OK;
}

// RK stage s of the current step; the step's start (BCs, dt, the RK3 saved
// state) happens at s = 0 and its end (iteration count, time) after the last.
void run_stage(osb_ctx* c, int s) {
    const int nv = c->nv;
    const size_t Pt = c->Ptot;
    Range in = make_range(c, 0);
    const int nst = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
    if (s == 0) {
        apply_bcs(c);                                        // temporal_start BC kernels (algorithm.py:462)
        c->stage_dt = c->d.cfl > 0 ? c->d.cfl / max_wavespeed(c) : c->d.dt;
        c->last_dt = c->stage_dt;
        if (c->d.time_scheme == OSB_TIME_RK3)
            // rk_sbli.py:129-159: save q ("Save equations"), then the stage updates
            for_range(in, [&](int i, int j, int k) {
                const size_t x = c->I(i, j, k);
                for (int v = 0; v < nv; ++v) c->rkold[(size_t)v * Pt + x] = c->q[(size_t)v * Pt + x];
            });
    }
    const double dt = c->stage_dt;
    compute_rhs(c);
    if (c->d.time_scheme == OSB_TIME_RKLS3) {
        const double A = c->rkA[s], B = c->rkB[s];
        for_range(in, [&](int i, int j, int k) {
            const size_t x = c->I(i, j, k);
            for (int v = 0; v < nv; ++v) {
                const size_t y = (size_t)v * Pt + x;
                rkls_point(A, B, dt, c->R[y], c->rkold[y], c->q[y]);
            }
        });
    } else {
        for_range(in, [&](int i, int j, int k) {
            const size_t x = c->I(i, j, k);
            for (int v = 0; v < nv; ++v) {
                const size_t y = (size_t)v * Pt + x;
                rk3_point(s, dt, c->R[y], c->rkold[y], c->q[y]);
            }
        });
    }
    apply_bcs(c);
    c->next_stage = (s + 1) % nst;
    if (c->next_stage == 0) {
        if (c->d.shock_filter == OSB_FILTER_WENOZ) apply_weno_filter(c, dt);
        c->iter += 1;
        c->time += dt;
    }
}

int osb_step(osb_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return OSB_E_INVALID;
    if (c->next_stage != 0) return fail(c, OSB_E_INVALID, "a step is partly advanced by osb_stage; finish its stages first");
    const int nst = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
    for (int it = 0; it < nsteps; ++it)
        for (int s = 0; s < nst; ++s) run_stage(c, s);
    return OSB_OK;
}

int osb_stage(osb_ctx* c, int s) {
    if (!c) return OSB_E_INVALID;
    const int nst = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
    if (s < 0 || s >= nst) return fail(c, OSB_E_INVALID, "stage index out of range");
    if (s != c->next_stage) return fail(c, OSB_E_INVALID, "stages run in order: expected stage " + std::to_string(c->next_stage));
    run_stage(c, s);
    return OSB_OK;
}

int osb_sync(osb_ctx* c) { return c ? OSB_OK : OSB_E_INVALID; }

int osb_get_stats(osb_ctx* c, osb_stats* out) {
    if (!c || !out) return OSB_E_INVALID;
    out->struct_size = sizeof(osb_stats);
    out->iter = c->iter;
    out->time = c->time;
    out->last_dt = c->last_dt;
    out->halo = HALO;
    out->nv = c->nv;
    out->device_bytes = 0;
    return OSB_OK;
}

int osb_tgv_diagnostics(osb_ctx* c, double out[4]) {
    // TENO6/opensbliblock00_kernels.h:4737-4797; u and mu are the constituent-relation
    // arrays left by the last stage, rho is the current state.
    if (!c || !out) return OSB_E_INVALID;
    if (c->ndim != 3) return fail(c, OSB_E_UNSUPPORTED, "TGV diagnostics need ndim == 3");
    if (c->mu.empty()) return fail(c, OSB_E_UNSUPPORTED, "TGV diagnostics need a viscous configuration");
    const size_t Pt = c->Ptot;
    double ke = 0, dil = 0, ens = 0, rhom = 0;
    const int N0 = c->n[0], N1 = c->n[1], N2 = c->n[2];
#pragma omp parallel for reduction(+ : ke, dil, ens, rhom) schedule(static)
    for (long t = 0; t < (long)N2 * N1; ++t) {
        const int k = (int)(t / N1), j = (int)(t % N1);
        for (int i = 0; i < N0; ++i) {
            const size_t x = c->I(i, j, k);
            double G[3][3];
            for (int m = 0; m < 3; ++m)
                for (int dd = 0; dd < 3; ++dd) {
                    const double* um = c->u.data() + (size_t)m * Pt;
                    const size_t st = c->stride(dd);
                    G[m][dd] = d1([&](int s) { return um[x + (long)s * (long)st]; }, 1.0 / c->d.delta[dd]);
                }
            const double wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
            const double divV = G[0][0] + G[1][1] + G[2][2];
            const double rho = c->q[x], muv = c->mu[x];
            const double u0 = c->u[x], u1 = c->u[Pt + x], u2 = c->u[2 * Pt + x];
            rhom += rho;
            ke += 0.5 * (u0 * u0 + u1 * u1 + u2 * u2) * rho;
            dil += (4.0 / 3.0) * (divV * divV) * muv;
            ens += (wx * wx + wy * wy + wz * wz) * muv;
        }
    }
    out[0] = ke; out[1] = dil; out[2] = ens; out[3] = rhom;
    return OSB_OK;
}

int osb_probe(osb_ctx* c, int n, const int32_t* ijk, const int32_t* var, double* out) {
    if (!c || n < 0 || (n > 0 && (!ijk || !var || !out))) return OSB_E_INVALID;
    for (int p = 0; p < n; ++p) {
        for (int d = 0; d < 3; ++d)
            if (ijk[3 * p + d] < 0 || ijk[3 * p + d] >= c->n[d]) return fail(c, OSB_E_INVALID, "probe outside the grid");
        if (var[p] < 0 || var[p] >= c->nv) return fail(c, OSB_E_INVALID, "probe variable out of range");
        out[p] = c->q[(size_t)var[p] * c->Ptot + c->I(ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2])];
    }
    return OSB_OK;
}

int osb_nan_check(osb_ctx* c) {
    if (!c) return OSB_E_INVALID;
    for (int k = 0; k < c->n[2]; ++k)
        for (int j = 0; j < c->n[1]; ++j)
            for (int i = 0; i < c->n[0]; ++i)
                if (!std::isfinite(c->q[c->I(i, j, k)])) return fail(c, OSB_E_NONFINITE, "non-finite rho");
    return OSB_OK;
}

int osb_state_device_ptr(osb_ctx* c, int, void**) { return fail(c, OSB_E_UNSUPPORTED, "oracle has no device buffers"); }
int osb_set_stream(osb_ctx* c, void*) { return fail(c, OSB_E_UNSUPPORTED, "oracle has no streams"); }
int osb_set_state_device(osb_ctx* c, const double*, int) { return fail(c, OSB_E_UNSUPPORTED, "oracle has no device buffers"); }
int osb_step_noalloc(const osb_problem_desc*, double*, double*, int, void*) { return fail(nullptr, OSB_E_UNSUPPORTED, "oracle has no device buffers"); }
int osb_describe(osb_ctx* c, char* buf, int len) {
    if (!c || !buf || len < 1) return OSB_E_INVALID;
    std::snprintf(buf, (size_t)len, "{\"path\": \"oracle\"}");
    return OSB_OK;
}
int osb_get_state_device(osb_ctx* c, double*, int) { return fail(c, OSB_E_UNSUPPORTED, "oracle has no device buffers"); }

}  // extern "C"

// ---------------------------------------------------------------------------
// Test hooks (oracle only): direct access to the pieces that the tests pin
// against closed forms and textbook identities.
// ---------------------------------------------------------------------------
extern "C" {

// ENO coefficients c[side][r][j] and optimal weights d[side][r] of WENO5
int osb_oracle_weno5_tables(double* c18, double* d6) {
    for (int s = 0; s < 2; ++s)
        for (int r = 0; r < 3; ++r) {
            d6[s * 3 + r] = W5.dopt[s][r];
            for (int j = 0; j < 3; ++j) c18[(s * 3 + r) * 3 + j] = W5.c[s][r][j];
        }
    return OSB_OK;
}

// one reconstruction: scheme 0 WENO-Z, 5 TENO5, 6 TENO6; side +1 (f+, offsets
// -2..2) or -1 (f-, offsets -1..3); f[0..5] at offsets -2..3
double osb_oracle_reconstruct(int scheme, int side, const double* f, double eps, double ct) {
    if (scheme == 0) return side > 0 ? wenoz_right(f, eps) : wenoz_left(f, eps);
    TenoParams tp{scheme, eps, ct};
    return teno_side(f, side, tp);
}

// eigensystem for direction dir: L (left eigenvectors as rows) and R (right
// eigenvectors as columns), row-major nv x nv
int osb_oracle_eigensystem(int ndim, int dir, double gama, double rho, const double* u, double a, double* L, double* R) {
    Eig e;
    eigensystem(ndim, dir, gama - 1.0, rho, u, a, e);
    const int nv = ndim + 2;
    for (int i = 0; i < nv; ++i)
        for (int j = 0; j < nv; ++j) { L[i * nv + j] = e.L[i][j]; R[i * nv + j] = e.Rm[i][j]; }
    return OSB_OK;
}

// scalar RK step of dq/dt = lam*q with the solver's stage updates
// stages = 0: the default coefficients of `scheme`; else RK-LS with A/B given
// Roe average of two conservative states (the face routine's averaging):
// out = rho, u0, u1, u2, a
int osb_oracle_roe_average(int ndim, double gama, const double* qa, const double* qb, double* out) {
    double ua[3] = {0, 0, 0}, ub[3] = {0, 0, 0}, kea = 0.0, keb = 0.0;
    for (int m = 0; m < ndim; ++m) {
        ua[m] = qa[1 + m] / qa[0];
        ub[m] = qb[1 + m] / qb[0];
        kea -= 0.5 * (ua[m] * ua[m]) * qa[0];
        keb -= 0.5 * (ub[m] * ub[m]) * qb[0];
    }
    const double pa = (gama - 1.0) * (kea + qa[ndim + 1]), pb = (gama - 1.0) * (keb + qb[ndim + 1]);
    double u_av[3] = {0, 0, 0};
    roe_average(ndim, gama - 1.0, qa[0], qb[0], ua, ub, pa, pb, qa[ndim + 1], qb[ndim + 1], out[0], u_av, out[4]);
    for (int m = 0; m < 3; ++m) out[1 + m] = u_av[m];
    return OSB_OK;
}

double osb_oracle_rk_scalar(int scheme, int stages, const double* A, const double* B, double lam, double dt, double q0) {
    double q = q0, reg = 0.0;
    if (scheme == OSB_TIME_RKLS3 && stages > 0) {
        for (int s = 0; s < stages; ++s) rkls_point(A[s], B[s], dt, lam * q, reg, q);
    } else if (scheme == OSB_TIME_RKLS3) {
        for (int s = 0; s < 3; ++s) rkls_point(RKLS_A[s], RKLS_B[s], dt, lam * q, reg, q);
    } else {
        reg = q;
        for (int s = 0; s < 3; ++s) rk3_point(s, dt, lam * q, reg, q);
    }
    return q;
}

// the padded conservative state (nv x P2 x P1 x P0 with 5 halo layers in
// active directions), after the boundary conditions
int osb_oracle_padded_state(osb_ctx* c, double* out, int* dims) {
    if (!c) return OSB_E_INVALID;
    apply_bcs(c);
    dims[0] = c->P[0]; dims[1] = c->P[1]; dims[2] = c->P[2]; dims[3] = HALO;
    if (out) std::copy(c->q.begin(), c->q.end(), out);
    return OSB_OK;
}

// GLF wave speeds of the last right-hand side: out[3d..3d+2] = max |u_d|, |u_d + a|, |u_d - a|
int osb_oracle_glf(const osb_ctx* c, double* out) {
    if (!c || !out) return OSB_E_INVALID;
    for (int d = 0; d < 3; ++d)
        for (int r = 0; r < 3; ++r) out[3 * d + r] = c->glf[d][r];
    return OSB_OK;
}

// test hook: the Ducros sensor of the current state over the interior (x fastest)
int osb_oracle_kappa(osb_ctx* c, double* out) {
    if (!c || !out || c->ndim < 2) return OSB_E_INVALID;
    if (c->kappa.empty()) c->kappa.assign(c->Ptot, 0.0);
    Range rcr = make_range(c, 3);
    constituent_relations(c, rcr, false);
    ducros_sensor(c);
    size_t o = 0;
    for (int k = 0; k < c->n[2]; ++k)
        for (int j = 0; j < c->n[1]; ++j)
            for (int i = 0; i < c->n[0]; ++i) out[o++] = c->kappa[c->I(i, j, k)];
    return OSB_OK;
}

// test hook: the HLLC-LM switch's sine
double osb_oracle_hllc_sin(double x) { return hllc_sin(x); }

// test hook: the Carpenter closure tables (4 x 6 first-derivative rows, then 2 x 5)
int osb_oracle_carpenter(double* out) {
    if (!out) return OSB_E_INVALID;
    for (int r = 0; r < 4; ++r)
        for (int j = 0; j < 6; ++j) out[6 * r + j] = CARP_D1[r][j];
    for (int r = 0; r < 2; ++r)
        for (int j = 0; j < 5; ++j) out[24 + 5 * r + j] = CARP_D2[r][j];
    return OSB_OK;
}

}  // extern "C"
