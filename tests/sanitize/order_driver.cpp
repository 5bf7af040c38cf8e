// Checks the tile order of the 3-D stage kernels (opensbli_amd/csrc/osb_order.hpp)
// on the host, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize.py:
//  - the multiplicative-inverse quotient cf_div against integer division, for
//    every divisor below 70000 on sampled and extreme numerators, and for
//    larger divisors up to 2^32 - 1 on random and extreme ones;
//  - for every tile grid up to 24 x 24 x 24 and the shapes of the 128^3-1024^3
//    grids, in each mode: every workgroup gets a tile in range, each tile
//    exactly once, the same tile as the plain division form of the order, and
//    the mode's locality property (mode 1: XCD x gets the linear tile range
//    [x T/8, (x+1) T/8); mode 2: XCD x gets the same range of the
//    supertile-major order, so with a supertile count divisible by 8 each
//    supertile lies on one XCD and each XCD's supertiles are contiguous in
//    z-first order), and the order before the XCD split (the batched
//    kernel's walk of one instance) visits every tile once, supertile by
//    supertile in mode 2.
//  - the kernels' flat index splits (unflatten_dv over index_divisors'
//    inverses, as Geom::dv holds them) against integer division for random
//    1-D/2-D/3-D blocks: interior points, faces along each direction, the
//    fused face-residual tiles at both widths and the boundary layers.
// Exits non-zero on the first failed check.
#define __host__
#define __device__
#define __forceinline__ inline
#include "osb_order.hpp"

#include <cstdio>
#include <random>
#include <vector>

using osb::CfOrder;

namespace {

long g_bad = 0;

void fail(const char* what, unsigned a, unsigned b, unsigned c, int m) {
    if (g_bad++ < 8) std::printf("FAIL %s (%u, %u, %u) mode arg %d\n", what, a, b, c, m);
}

void check_div() {
    std::mt19937_64 rng(20261204);
    for (unsigned d = 1; d < 70000; ++d) {
        unsigned mp, l;
        osb::cf_divisor(d, mp, l);
        const unsigned edge[] = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
        for (unsigned n : edge)
            if (osb::cf_div(n, mp, l) != n / d) fail("cf_div edge", d, n, 0, 0);
        for (int k = 0; k < 64; ++k) {
            const unsigned n = k < 32 ? (unsigned)(k * 104729u) % (1u << 24) : (unsigned)rng();
            if (osb::cf_div(n, mp, l) != n / d) fail("cf_div", d, n, 0, 0);
        }
    }
    const unsigned big[] = {65537u, 999983u, 1u << 20, (1u << 20) + 1, 0x7fffffffu, 0x80000000u, 0x80000001u, 0xc0000000u, 0xfffffffeu, 0xffffffffu};
    for (unsigned d : big) {
        unsigned mp, l;
        osb::cf_divisor(d, mp, l);
        for (int k = 0; k < 200000; ++k) {
            const unsigned n = (unsigned)rng();
            if (osb::cf_div(n, mp, l) != n / d) fail("cf_div big", d, n, 0, 0);
        }
        const unsigned edge[] = {0u, d - 1, d, 0xfffffffeu, 0xffffffffu};
        for (unsigned n : edge)
            if (osb::cf_div(n, mp, l) != n / d) fail("cf_div big edge", d, n, 0, 0);
    }
    std::printf("ok cf_div\n");
}

// the order written with plain integer division
void plain_tile(unsigned ntx, unsigned nty, unsigned ntz, int remap, unsigned w, int& x, int& y, int& z) {
    const unsigned T = ntx * nty * ntz;
    unsigned t = w;
    if (remap && T % 8 == 0) t = (w % 8) * (T / 8) + w / 8;
    x = t % ntx;
    y = (t / ntx) % nty;
    z = t / (ntx * nty);
    if (remap == 2 && ntx % 8 == 0 && nty % 4 == 0 && ntz % 2 == 0) {
        const unsigned in = t % 64, sup = t / 64, sz = ntz / 2, sy = nty / 4;
        z = 2 * (sup % sz) + in / 32;
        y = 4 * ((sup / sz) % sy) + (in / 8) % 4;
        x = 8 * (sup / (sz * sy)) + in % 8;
    }
}

void check_walk(unsigned ntx, unsigned nty, unsigned ntz, int remap) {
    const CfOrder o = osb::cf_order(ntx, nty, ntz, remap);
    const unsigned T = ntx * nty * ntz;
    const bool m2 = remap == 2 && ntx % 8 == 0 && nty % 4 == 0 && ntz % 2 == 0;
    const int want = m2 ? 2 : (remap && T % 8 == 0 ? 1 : 0);
    if (o.mode != want) return fail("mode", ntx, nty, ntz, remap);
    std::vector<unsigned char> seen(T, 0);
    std::vector<int> sup_xcd(T / 64 + 1, -1);
    std::vector<unsigned> sup_lo(8, ~0u), sup_hi(8, 0), sup_n(8, 0);
    for (unsigned w = 0; w < T; ++w) {
        int x, y, z, px, py, pz;
        osb::cf_tile_of(o, w, x, y, z);
        if (x < 0 || y < 0 || z < 0 || (unsigned)x >= ntx || (unsigned)y >= nty || (unsigned)z >= ntz)
            return fail("tile out of range", ntx, nty, ntz, remap);
        plain_tile(ntx, nty, ntz, remap, w, px, py, pz);
        if (x != px || y != py || z != pz) return fail("differs from the division form", ntx, nty, ntz, remap);
        const unsigned id = x + ntx * (y + nty * (unsigned)z);
        if (seen[id]++) return fail("tile visited twice", ntx, nty, ntz, remap);
        const unsigned xcd = w % 8;
        if (o.mode == 1 && (id < xcd * (T / 8) || id >= (xcd + 1) * (T / 8))) return fail("mode 1 range", ntx, nty, ntz, remap);
        if (o.mode == 2) {
            // position in the supertile-major order: z-first supertile index, then
            // the tile within its supertile (x fastest, then y, then z)
            const unsigned s = z / 2 + (ntz / 2) * (y / 4 + (nty / 4) * (x / 8));
            const unsigned pos = 64 * s + x % 8 + 8 * (y % 4) + 32 * (z % 2);
            if (pos < xcd * (T / 8) || pos >= (xcd + 1) * (T / 8)) return fail("mode 2 range", ntx, nty, ntz, remap);
            if ((T / 64) % 8 == 0) {
                if (sup_xcd[s] < 0) {
                    sup_xcd[s] = (int)xcd;
                    sup_lo[xcd] = s < sup_lo[xcd] ? s : sup_lo[xcd];
                    sup_hi[xcd] = s > sup_hi[xcd] ? s : sup_hi[xcd];
                    ++sup_n[xcd];
                } else if (sup_xcd[s] != (int)xcd) {
                    return fail("supertile split over XCDs", ntx, nty, ntz, remap);
                }
            }
        }
    }
    if (o.mode == 2 && (T / 64) % 8 == 0)
        for (int x = 0; x < 8; ++x)
            if (sup_n[x] != sup_hi[x] - sup_lo[x] + 1) return fail("XCD supertiles not contiguous", ntx, nty, ntz, remap);
    // the order before the XCD split (the batched kernel's walk of one instance):
    // each tile once; in mode 2 the 64 tiles of a supertile consecutive
    std::vector<unsigned char> seen2(T, 0);
    for (unsigned p = 0; p < T; ++p) {
        int x, y, z;
        osb::cf_tile_at(o, p, x, y, z);
        if (x < 0 || y < 0 || z < 0 || (unsigned)x >= ntx || (unsigned)y >= nty || (unsigned)z >= ntz)
            return fail("cf_tile_at out of range", ntx, nty, ntz, remap);
        if (seen2[x + ntx * (y + nty * (unsigned)z)]++) return fail("cf_tile_at visits a tile twice", ntx, nty, ntz, remap);
        if (o.mode == 2) {
            const unsigned s = z / 2 + (ntz / 2) * (y / 4 + (nty / 4) * (x / 8));
            if (s != p / 64) return fail("cf_tile_at supertile", ntx, nty, ntz, remap);
        } else if (x + ntx * (y + nty * (unsigned)z) != p) {
            return fail("cf_tile_at not linear", ntx, nty, ntz, remap);
        }
    }
}

// the kernels' flat index splits (unflatten_dv over index_divisors' inverses)
// against integer division, over every extent family the kernels split: the
// interior, the faces along each direction, the fused face-residual tiles at
// both widths and the boundary layers of each direction
void split_ref(long t, const int (&e)[3], const int (&lo)[3], int& i, int& j, int& k) {
    i = (int)(t % e[0]) + lo[0];
    j = (int)((t / e[0]) % e[1]) + lo[1];
    k = (int)(t / ((long)e[0] * e[1])) + lo[2];
}

void check_family(const int (&e)[3], const int (&lo)[3], const unsigned (&dv)[4], std::mt19937_64& rng, int ndim, int n0) {
    const long tot = (long)e[0] * e[1] * e[2];
    auto one = [&](long t) {
        int i, j, k, ri, rj, rk;
        osb::unflatten_dv(t, e, dv, lo, i, j, k);
        split_ref(t, e, lo, ri, rj, rk);
        if (i != ri || j != rj || k != rk) fail("unflatten_dv", (unsigned)ndim, (unsigned)n0, (unsigned)t, e[0]);
    };
    if (tot <= 200000) {
        for (long t = 0; t < tot; ++t) one(t);
    } else {
        for (int s = 0; s < 200000; ++s) one((long)(rng() % (unsigned long long)tot));
        one(tot - 1);
    }
}

void check_index_splits() {
    std::mt19937_64 rng(7);
    const int ws[2] = {32, 16};
    const int block = 256;
    long cases = 0;
    for (int ndim = 1; ndim <= 3; ++ndim)
        for (int c = 0; c < 60; ++c) {
            int n[3] = {1, 1, 1};
            for (int d = 0; d < ndim; ++d) n[d] = 8 + (int)(rng() % (ndim == 1 ? 5000 : (ndim == 2 ? 700 : 120)));
            if (c == 0)
                for (int d = 0; d < ndim; ++d) n[d] = ndim == 3 ? 256 : 1024;
            const int halo = (c & 1) ? 4 : 2;
            unsigned dv[9][4];
            osb::index_divisors(ndim, n, halo, ws, block, dv);
            const int zero[3] = {0, 0, 0};
            check_family(n, zero, dv[0], rng, ndim, n[0]);                       // interior
            for (int d = 0; d < ndim; ++d) {                                    // faces along d
                int e[3] = {n[0], n[1], n[2]}, lo[3] = {0, 0, 0};
                e[d] += 1;
                lo[d] = -1;
                check_family(e, lo, dv[1 + d], rng, ndim, n[0]);
            }
            if (ndim >= 2)
                for (int w = 0; w < 2; ++w) {                                   // face-residual tiles
                    const int fw = ws[w], pl = block / fw - 1, D = ndim - 1;
                    const int e[3] = {(n[0] + fw - 1) / fw, (n[D] + pl - 1) / pl, ndim == 3 ? n[1] : 1};
                    check_family(e, zero, dv[4 + w], rng, ndim, n[0]);
                }
            for (int d = 0; d < ndim; ++d) {                                    // boundary layers of d
                int e[3], lo[3];
                for (int a = 0; a < 3; ++a) {
                    e[a] = a < ndim ? n[a] + 2 * halo : 1;
                    lo[a] = a < ndim ? -halo : 0;
                }
                e[d] = 2 * (halo + 1);
                lo[d] = 0;
                check_family(e, lo, dv[6 + d], rng, ndim, n[0]);
            }
            ++cases;
        }
    std::printf("ok index splits %ld blocks\n", cases);
}

}  // namespace

int main() {
    check_div();
    check_index_splits();
    long walks = 0;
    for (unsigned x = 1; x <= 24; ++x)
        for (unsigned y = 1; y <= 24; ++y)
            for (unsigned z = 1; z <= 24; ++z)
                for (int r = 0; r < 3; ++r, ++walks) check_walk(x, y, z, r);
    // 128^3 .. 1024^3 at the 16x4x4, 16x8x4 and 8x8x8 tiles, and ragged grids
    const unsigned shapes[][3] = {{8, 32, 32}, {8, 16, 32}, {16, 16, 16}, {16, 64, 64}, {16, 32, 64}, {32, 32, 32},
                                  {64, 256, 256}, {64, 128, 256}, {128, 128, 128}, {7, 33, 65}, {9, 12, 30}};
    for (const auto& s : shapes)
        for (int r = 0; r < 3; ++r, ++walks) check_walk(s[0], s[1], s[2], r);
    if (g_bad) {
        std::printf("%ld failures\n", g_bad);
        return 1;
    }
    std::printf("ok tile walks %ld\nall ok\n", walks);
    return 0;
}
