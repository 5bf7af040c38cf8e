// Drives the HDF5 output/restart library (opensbli_amd/csrc/osb_io.cpp,
// compiled in) through include/osb_io.h: a 3-D and a 1-D block written with
// float64 and int32 constants, read back (shapes, arrays, constants) and
// compared, and the refusals (missing file, wrong element count).  Built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize.py.
#include "osb_io.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

static int failures = 0;
static void check(bool ok, const char* what) {
    if (!ok) { std::printf("FAIL %s: %s\n", what, osb_io_last_error()); ++failures; }
}

static void block_roundtrip(const std::string& path, int ndim, const int32_t (&size)[3], int halo) {
    osb_io_block b{"opensbliblock00", 0, ndim, {size[0], size[1], size[2]}, halo};
    size_t count = 1;
    for (int d = 0; d < 3; ++d) count *= (size_t)size[d] + (d < ndim ? 2 * halo : 0);
    std::vector<double> a(count), c(count);
    for (size_t i = 0; i < count; ++i) { a[i] = std::sin(0.01 * i); c[i] = 1.0 / (1.0 + i); }
    const char* names[2] = {"rho_B0", "rhoE_B0"};
    const double* data[2] = {a.data(), c.data()};
    const char* dn[2] = {"gama", "dt"};
    const double dv[2] = {1.4, 2e-4};
    const char* in[1] = {"niter"};
    const int32_t iv[1] = {1000};
    check(osb_io_write(path.c_str(), &b, 2, names, data, 2, dn, dv, 1, in, iv) == 0, "write");
    int32_t nd = 0, sz[3] = {0, 0, 0}, dm[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    check(osb_io_dat_info(path.c_str(), "opensbliblock00", "rhoE_B0", &nd, sz, dm, dp) == 0, "dat_info");
    check(nd == ndim, "ndim");
    for (int d = 0; d < ndim; ++d) check(sz[d] == size[d] && dm[d] == -halo && dp[d] == halo, "shape");
    std::vector<double> back(count);
    check(osb_io_read_dat(path.c_str(), "opensbliblock00", "rho_B0", back.data(), (int64_t)count) == 0, "read rho");
    check(back == a, "rho round trip");
    check(osb_io_read_dat(path.c_str(), "opensbliblock00", "rhoE_B0", back.data(), (int64_t)count) == 0, "read rhoE");
    check(back == c, "rhoE round trip");
    double v = 0.0;
    check(osb_io_read_const(path.c_str(), "gama", &v) == 0 && v == 1.4, "float constant");
    check(osb_io_read_const(path.c_str(), "niter", &v) == 0 && v == 1000.0, "integer constant");
    check(osb_io_read_dat(path.c_str(), "opensbliblock00", "rho_B0", back.data(), (int64_t)count - 1) != 0, "wrong count refused");
    check(osb_io_read_const(path.c_str(), "no_such_constant", &v) != 0, "missing constant refused");
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    block_roundtrip(dir + "/block3d.h5", 3, {12, 10, 8}, 2);
    block_roundtrip(dir + "/block1d.h5", 1, {64, 1, 1}, 4);
    double v = 0.0;
    check(osb_io_read_const((dir + "/missing.h5").c_str(), "gama", &v) != 0, "missing file refused");
    std::printf("%s (%d failure%s)\n", failures ? "FAILED" : "all ok", failures, failures == 1 ? "" : "s");
    return failures ? 1 : 0;
}
