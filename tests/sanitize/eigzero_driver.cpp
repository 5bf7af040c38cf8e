// Checks the structural zeros that the face kernels skip in their
// characteristic projections (opensbli_amd/csrc/osb_eigzero.hpp) against the
// oracle's Euler eigensystem (osb_oracle_eigensystem), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize.py.
// For 1-D, 2-D and 3-D and every direction, on random states: each entry of
// L and R that l_zero / r_zero flag is exactly zero, and each entry they do
// not flag is nonzero (so no entry the kernels need is skipped, and the
// predicates flag every zero the oracle's matrices hold).
// Exits non-zero on the first failed check.
#define __host__
#define __device__
#include "osb_eigzero.hpp"

#include <cstdio>
#include <random>

extern "C" int osb_oracle_eigensystem(int ndim, int dir, double gama, double rho, const double* u, double a, double* L,
                                      double* R);

static int fails = 0;

template <int NDIM, int DIR>
static void check(std::mt19937_64& rng) {
    constexpr int NV = NDIM + 2;
    std::uniform_real_distribution<double> pos(0.2, 3.0), vel(0.05, 2.0);
    long zeros_l = 0, zeros_r = 0;
    for (int t = 0; t < 2000; ++t) {
        double u[3];
        for (int d = 0; d < NDIM; ++d) u[d] = (rng() & 1 ? 1.0 : -1.0) * vel(rng);
        const double rho = pos(rng), a = pos(rng), gama = 1.0 + 0.8 * pos(rng) / 3.0;
        double L[NV * NV], R[NV * NV];
        osb_oracle_eigensystem(NDIM, DIR, gama, rho, u, a, L, R);
        for (int i = 0; i < NV; ++i)
            for (int j = 0; j < NV; ++j) {
                const bool lz = osb::l_zero<NDIM, DIR>(i, j), rz = osb::r_zero<NDIM, DIR>(i, j);
                zeros_l += lz;
                zeros_r += rz;
                if (lz != (L[i * NV + j] == 0.0) || rz != (R[i * NV + j] == 0.0)) {
                    if (fails++ < 10)
                        std::printf("FAIL ndim %d dir %d entry (%d,%d): L %g flagged %d, R %g flagged %d\n", NDIM, DIR, i, j,
                                    L[i * NV + j], (int)lz, R[i * NV + j], (int)rz);
                }
            }
    }
    std::printf("ndim %d dir %d: %ld L and %ld R zero entries per state\n", NDIM, DIR, zeros_l / 2000, zeros_r / 2000);
}

int main() {
    std::mt19937_64 rng(20261204);
    check<1, 0>(rng);
    check<2, 0>(rng);
    check<2, 1>(rng);
    check<3, 0>(rng);
    check<3, 1>(rng);
    check<3, 2>(rng);
    if (fails) {
        std::printf("%d failed checks\n", fails);
        return 1;
    }
    std::printf("ok structural zeros\nall ok\n");
    return 0;
}
