// Drives the CPU oracle (oracle/osb_oracle.cpp, compiled in) through the C ABI
// over small configurations covering its paths -- 1-D/2-D/3-D, WENO-Z, TENO5,
// TENO6, HLLC, the central split forms with constant and Sutherland
// viscosity, every boundary type (with the Carpenter closure next to
// extrapolation and Dirichlet faces), GLF, CFL steps, the WENO filter with
// and without the Ducros selection, classic RK3 and 5-stage RK-LS -- plus the
// residual, wave-speed, diagnostics, probe and NaN-check entry points and the
// descriptor checks.  Built with AddressSanitizer and UndefinedBehaviorSanitizer
// by tests/test_sanitize.py; exits non-zero if a run fails or goes non-finite.
#include "osb.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

osb_problem_desc base(int ndim, int n0, int n1, int n2) {
    osb_problem_desc d;
    std::memset(&d, 0, sizeof d);
    d.struct_size = sizeof d;
    d.abi_version = OSB_ABI_VERSION;
    d.ndim = ndim;
    const int n[3] = {n0, n1, n2};
    for (int i = 0; i < 3; ++i) {
        d.n[i] = i < ndim ? n[i] : 1;
        d.delta[i] = i < ndim ? 2.0 * M_PI / n[i] : 1.0;
        for (int s = 0; s < 2; ++s) d.bc[i][s] = i < ndim ? OSB_BC_PERIODIC : 0;
    }
    d.conv_scheme = OSB_CONV_WENO_Z;
    d.order = 5;
    d.averaging = OSB_AVG_ROE;
    d.viscous = OSB_VISC_NONE;
    d.time_scheme = OSB_TIME_RKLS3;
    d.precision = OSB_PREC_FP64;
    d.gama = 1.4;
    d.minf = 0.5;
    d.re = 200.0;
    d.pr = 0.71;
    d.weno_eps = 1e-40;
    d.teno_eps = 1e-40;
    d.teno_ct = 1e-5;
    d.shock_filter_control = 1.0;
    d.dt = 1e-3;
    d.filter_order = 5;
    return d;
}

// interior-layout state: smooth waves in every active direction, plus a
// density and pressure jump at the middle of x0 when `jump`
std::vector<double> state(const osb_problem_desc& d, bool jump) {
    const int nv = d.ndim + 2, n0 = d.n[0], n1 = d.n[1], n2 = d.n[2];
    std::vector<double> q((size_t)nv * n0 * n1 * n2);
    const size_t P = (size_t)n0 * n1 * n2;
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const double x = (i + 0.5) / n0, y = (j + 0.5) / n1, z = (k + 0.5) / n2;
                double rho = 1.0 + 0.2 * std::sin(2 * M_PI * (x + 0.3 * y)) * std::cos(2 * M_PI * z);
                double p = 1.0 + 0.1 * std::cos(2 * M_PI * (x - y + z));
                if (jump && x > 0.5) { rho *= 0.25; p *= 0.2; }
                const double u[3] = {0.3 * std::sin(2 * M_PI * y), -0.2 * std::cos(2 * M_PI * x), 0.1 * std::sin(2 * M_PI * (x + z))};
                double ke = 0.0;
                const size_t at = (size_t)i + (size_t)n0 * ((size_t)j + (size_t)n1 * k);
                q[at] = rho;
                for (int m = 0; m < d.ndim; ++m) { q[(1 + m) * P + at] = rho * u[m]; ke += 0.5 * rho * u[m] * u[m]; }
                q[(nv - 1) * P + at] = p / (d.gama - 1.0) + ke;
            }
    return q;
}

int failures = 0;

void check(bool ok, const char* name, const char* what) {
    if (!ok) { std::printf("FAIL %s: %s\n", name, what); ++failures; }
}

void run(const char* name, osb_problem_desc d, int steps, bool jump) {
    osb_ctx* c = nullptr;
    int rc = osb_ctx_create(&d, &c);
    if (rc) { std::printf("FAIL %s: create %d %s\n", name, rc, osb_last_error(nullptr)); ++failures; return; }
    std::vector<double> q = state(d, jump);
    check(osb_set_state(c, q.data(), OSB_LAYOUT_INTERIOR) == OSB_OK, name, "set_state");
    check(osb_step(c, steps) == OSB_OK, name, "step");
    const int nst = d.time_scheme == OSB_TIME_RKLS3 && d.rk_stages ? d.rk_stages : 3;
    for (int s = 0; s < nst; ++s) check(osb_stage(c, s) == OSB_OK, name, "stage");
    std::vector<double> out(q.size()), r(q.size());
    check(osb_get_state(c, out.data(), OSB_LAYOUT_INTERIOR) == OSB_OK, name, "get_state");
    osb_stats st;
    st.struct_size = sizeof st;
    check(osb_get_stats(c, &st) == OSB_OK, name, "stats");
    size_t padded = (size_t)st.nv;
    for (int i = 0; i < 3; ++i) padded *= (size_t)d.n[i] + (i < d.ndim ? 2 * st.halo : 0);
    std::vector<double> qp(padded);
    check(osb_get_state(c, qp.data(), OSB_LAYOUT_PADDED) == OSB_OK, name, "get_state padded");
    check(osb_set_state(c, qp.data(), OSB_LAYOUT_PADDED) == OSB_OK, name, "set_state padded");
    check(osb_apply_bcs(c) == OSB_OK, name, "apply_bcs");
    check(osb_residual(c, r.data()) == OSB_OK, name, "residual");
    double w = 0.0;
    check(osb_max_wavespeed(c, &w) == OSB_OK && std::isfinite(w) && w > 0, name, "max_wavespeed");
    if (d.ndim == 3) {
        double diag[4];
        check(osb_tgv_diagnostics(c, diag) == OSB_OK, name, "tgv_diagnostics");
    }
    const int32_t ijk[6] = {0, 0, 0, d.n[0] - 1, d.n[1] - 1, d.n[2] - 1};
    const int32_t var[2] = {0, d.ndim + 1};
    double probe[2];
    check(osb_probe(c, 2, ijk, var, probe) == OSB_OK, name, "probe");
    check(osb_nan_check(c) == OSB_OK, name, "nan_check");
    bool finite = true;
    for (double v : out) finite = finite && std::isfinite(v);
    for (double v : r) finite = finite && std::isfinite(v);
    check(finite, name, "non-finite state or residual");
    // a NaN in rho is reported
    q[q.size() / 7 % (q.size() / (d.ndim + 2))] = NAN;
    check(osb_set_state(c, q.data(), OSB_LAYOUT_INTERIOR) == OSB_OK, name, "set_state nan");
    check(osb_nan_check(c) == OSB_E_NONFINITE, name, "nan_check on a NaN");
    // stages run in order; a partly advanced step refuses osb_step
    check(osb_stage(c, 1) != OSB_OK, name, "out-of-order stage refused");
    osb_ctx_destroy(c);
    std::printf("ok %s\n", name);
}

}  // namespace

int main() {
    {
        osb_problem_desc d = base(1, 64, 1, 1);
        d.averaging = OSB_AVG_SIMPLE;
        d.bc[0][0] = d.bc[0][1] = OSB_BC_ZERO_GRADIENT;
        run("sod-like WENO-Z, zero gradient", d, 10, true);
        d.conv_scheme = OSB_CONV_TENO;
        d.flux_type = OSB_FLUX_GLF;
        d.bc[0][0] = d.bc[0][1] = OSB_BC_DIRICHLET;
        for (int s = 0; s < 2; ++s) {
            d.bc_value[0][s][0] = 1.0;
            d.bc_value[0][s][2] = 2.5;
        }
        run("TENO5 GLF, Dirichlet", d, 10, true);
        d.conv_scheme = OSB_CONV_WENO_Z;
        d.flux_type = OSB_FLUX_HLLC_LM;
        d.bc[0][0] = OSB_BC_EXTRAPOLATION0;
        d.bc[0][1] = OSB_BC_EXTRAPOLATION1;
        run("HLLC-LM, extrapolation", d, 10, true);
        d = base(1, 48, 1, 1);
        d.conv_scheme = OSB_CONV_CENTRAL_FEIEREISEN;
        d.order = 4;
        d.shock_filter = OSB_FILTER_WENOZ;
        run("1-D central + WENO filter", d, 6, false);
    }
    {
        osb_problem_desc d = base(2, 32, 24, 1);
        d.bc[1][0] = d.bc[1][1] = OSB_BC_INVISCID_WALL;
        run("2-D WENO-Z, walls", d, 6, true);
        d.precision = OSB_PREC_MIXED;
        run("2-D WENO-Z mixed", d, 6, true);
        d = base(2, 32, 24, 1);
        d.bc[0][0] = OSB_BC_DIRICHLET;
        d.bc[0][1] = OSB_BC_EXTRAPOLATION1;
        d.bc[1][0] = OSB_BC_EXTRAPOLATION0;
        d.bc[1][1] = OSB_BC_SYMMETRY;
        for (int v = 0; v < 4; ++v) d.bc_value[0][0][v] = v == 0 ? 1.0 : (v == 3 ? 2.5 : 0.0);
        d.cfl = 0.4;
        d.dt = 0.0;
        run("2-D mixed faces, CFL", d, 6, true);
        d = base(2, 32, 24, 1);
        d.conv_scheme = OSB_CONV_CENTRAL_KGP;
        d.order = 4;
        d.viscous = OSB_VISC_CONSTANT;
        d.shock_filter = OSB_FILTER_WENOZ;
        d.filter_order = 3;
        d.filter_sensor = OSB_SENSOR_DUCROS;
        d.ducros_check = d.ducros_select = 0.05;
        run("2-D KGP viscous + WENO3 filter, Ducros", d, 6, false);
        d.conv_scheme = OSB_CONV_CENTRAL_SKEW;
        d.shock_filter = OSB_FILTER_NONE;
        d.time_scheme = OSB_TIME_RK3;
        d.bc[0][0] = d.bc[0][1] = OSB_BC_SYMMETRY;
        run("2-D skew, classic RK3, symmetry", d, 4, false);
    }
    {
        osb_problem_desc d = base(3, 12, 10, 8);
        d.conv_scheme = OSB_CONV_CENTRAL_FEIEREISEN;
        d.order = 4;
        d.viscous = OSB_VISC_CONSTANT;
        run("3-D Feiereisen, constant viscosity", d, 3, false);
        d.viscous = OSB_VISC_SUTHERLAND;
        d.bc[0][0] = OSB_BC_EXTRAPOLATION1;
        d.bc[0][1] = OSB_BC_EXTRAPOLATION0;
        run("3-D Sutherland, closure faces", d, 3, false);
        d = base(3, 12, 10, 8);
        d.conv_scheme = OSB_CONV_TENO;
        d.viscous = OSB_VISC_SUTHERLAND;
        d.cfl = 0.5;
        d.dt = 0.0;
        run("3-D TENO5 + viscous, CFL", d, 3, true);
        d.order = 6;
        d.cfl = 0.0;
        d.dt = 1e-3;
        d.rk_stages = 5;
        const double A[5] = {0.0, -4.344339134485095, 0.0, -3.770024161386381, -0.046347284573284};
        const double B[5] = {0.713497331193829, 0.133505249805329, 0.713497331193829, 0.149579395628565, 0.384471116121269};
        for (int s = 0; s < 5; ++s) { d.rk_a[s] = A[s]; d.rk_b[s] = B[s]; }
        run("3-D TENO6, 5-stage RK-LS", d, 2, true);
    }
    {
        // descriptor checks: every refusal returns a code and leaves a message
        osb_problem_desc d = base(3, 12, 10, 8);
        osb_ctx* c = nullptr;
        d.abi_version = 0;
        check(osb_ctx_create(&d, &c) == OSB_E_INVALID && !c, "descriptor", "abi version");
        d = base(3, 12, 10, 8);
        d.n[1] = 4;
        check(osb_ctx_create(&d, &c) == OSB_E_INVALID && std::strlen(osb_last_error(nullptr)) > 0, "descriptor", "n < 8");
        d = base(2, 16, 16, 1);
        d.bc[0][1] = OSB_BC_ZERO_GRADIENT;
        check(osb_ctx_create(&d, &c) == OSB_E_INVALID, "descriptor", "unpaired periodic face");
        check(osb_ctx_create(nullptr, &c) != OSB_OK, "descriptor", "null descriptor");
    }
    std::printf("%s (%d failure%s)\n", failures ? "FAILED" : "all ok", failures, failures == 1 ? "" : "s");
    return failures ? 1 : 0;
}
