// osb_order.hpp -- division by invariant integers, and the tile order of the
// 3-D stage kernels (the tile that workgroup w of a one-dimensional launch
// evaluates).  Integer code only, shared by the kernels and by the host code
// that launches them; the kernels also split their flat thread indices with
// these inverses (Geom::dv).
//
// Workgroups are dealt round-robin to the 8 XCDs, so workgroup w runs on XCD
// w % 8.
//   mode 0: tile w, x fastest.
//   mode 1 (tile count a multiple of 8): XCD x takes the contiguous range of
//     tiles [x*T/8, (x+1)*T/8), so neighbouring tiles -- which share halo
//     points -- stay in the same L2.
//   mode 2 (tiles along x, y, z multiples of 8, 4, 2): the same ranges over
//     supertiles of 8 x 4 x 2 tiles (about one XCD's resident workgroups)
//     walked z first.  With a supertile count divisible by 8, each XCD's range
//     is a slab of whole supertile columns, and consecutive supertiles share
//     their z-halo planes while they are in L2.
// The index n (a tile in modes 0 and 1, a supertile in mode 2) splits into
// digits of radix c0, then c1.  Its two quotients n / c0 and n / (c0 c1) come
// from the host's multiplicative inverses, so a workgroup's prologue is a few
// scalar multiplies instead of six integer-division sequences.
#pragma once

namespace osb {

struct CfOrder {
    int mode;
    unsigned tq;                  // tiles per XCD range (modes 1 and 2)
    unsigned c0, c1;              // modes 0/1: tiles along x, y; mode 2: supertiles along z, y
    unsigned mp0, l0, mp1, l1;    // inverses of c0 and of c0 * c1 (cf_divisor)
};

// n / d = (n + umulhi(n, mp)) >> l for every 32-bit n, with l = ceil(log2 d)
// and mp = ceil(2^(32+l) / d) - 2^32 (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, sec. 4).  d >= 1; for
// d > 2^31 (l = 32, never a power of two) ceil(2^64 / d) = (2^64 - 1) / d + 1.
__host__ inline void cf_divisor(unsigned d, unsigned& mp, unsigned& l) {
    l = 0;
    while ((1ull << l) < d) ++l;
    const unsigned long long m = l < 32 ? ((1ull << (32 + l)) + d - 1) / d : ~0ull / d + 1;
    mp = (unsigned)(m - (1ull << 32));
}
__host__ __device__ __forceinline__ unsigned cf_div(unsigned n, unsigned mp, unsigned l) {
    const unsigned t = (unsigned)(((unsigned long long)n * mp) >> 32);
    return (unsigned)(((unsigned long long)n + t) >> l);
}

// ntx, nty, ntz: tiles along x, y, z; xcd_remap: 0 mode 0, 1 mode 1 where it
// applies, 2 mode 2 where it applies (else mode 1, else mode 0)
__host__ inline CfOrder cf_order(unsigned ntx, unsigned nty, unsigned ntz, int xcd_remap) {
    CfOrder o{};
    const unsigned ntiles = ntx * nty * ntz;
    o.mode = xcd_remap == 2 && ntx % 8 == 0 && nty % 4 == 0 && ntz % 2 == 0 ? 2 : (xcd_remap && ntiles % 8 == 0 ? 1 : 0);
    o.tq = ntiles / 8;
    o.c0 = o.mode == 2 ? ntz / 2 : ntx;
    o.c1 = o.mode == 2 ? nty / 4 : nty;
    cf_divisor(o.c0, o.mp0, o.l0);
    cf_divisor(o.c0 * o.c1, o.mp1, o.l1);
    return o;
}

// t split over extents e (x fastest) with lower bounds lo, from the inverses
// dv of e[0] and e[0] e[1] (cf_divisor); t below the extents' product, < 2^32
__host__ __device__ __forceinline__ void unflatten_dv(long t, const int (&e)[3], const unsigned (&dv)[4], const int (&lo)[3],
                                                      int& i, int& j, int& k) {
    const unsigned u = (unsigned)t;
    const unsigned a = cf_div(u, dv[0], dv[1]), b = cf_div(u, dv[2], dv[3]);
    i = (int)(u - a * (unsigned)e[0]) + lo[0];
    j = (int)(a - b * (unsigned)e[1]) + lo[1];
    k = (int)b + lo[2];
}

// The inverses the kernels split flat indices with (Geom::dv), for a block of
// ndim dimensions, interior extents n and halo depth halo: dv[0] the interior
// points (n0, n0 n1); dv[1 + d] the faces along d (n_d + 1 along d); dv[4 + w]
// the fused face-residual tiles of width resid_ws[w] (block threads per tile:
// columns of the width, segments of block / width - 1 points along the last
// direction); dv[6 + d] the boundary layers of direction d (2 (halo + 1)
// along d, n + 2 halo across in the active directions).
__host__ inline void index_divisors(int ndim, const int (&n)[3], int halo, const int (&resid_ws)[2], int block,
                                    unsigned (&dv)[9][4]) {
    auto set = [](unsigned(&d)[4], long e0, long e1) {
        cf_divisor((unsigned)e0, d[0], d[1]);
        cf_divisor((unsigned)(e0 * e1), d[2], d[3]);
    };
    set(dv[0], n[0], n[1]);
    for (int d = 0; d < 3; ++d) set(dv[1 + d], n[0] + (d == 0 ? 1 : 0), n[1] + (d == 1 ? 1 : 0));
    const int D = ndim - 1;
    for (int w = 0; w < 2; ++w) {
        const int fw = ndim == 1 ? 1 : resid_ws[w], pl = block / fw - 1;
        set(dv[4 + w], (n[0] + fw - 1) / fw, (n[D] + pl - 1) / pl);
    }
    for (int d = 0; d < 3; ++d) {
        long e[3];
        for (int a = 0; a < 3; ++a) e[a] = a < ndim ? n[a] + 2 * halo : 1;
        e[d] = 2 * (halo + 1);
        set(dv[6 + d], e[0], e[1]);
    }
}

// tile coordinates (in tiles) of position p of the order before the XCD split:
// linear in modes 0 and 1, supertile-major in mode 2 (the batched stage kernel
// applies its own XCD split over all instances and walks each instance's tiles
// in this order)
__host__ __device__ __forceinline__ void cf_tile_at(const CfOrder& o, unsigned p, int& tbx, int& tby, int& tbz) {
    const unsigned n = o.mode == 2 ? p >> 6 : p;
    const unsigned a = cf_div(n, o.mp0, o.l0), b = cf_div(n, o.mp1, o.l1);
    const int r0 = (int)(n - a * o.c0), r1 = (int)(a - b * o.c1), r2 = (int)b;
    if (o.mode == 2) {
        const int in = (int)(p & 63);
        tbz = (r0 << 1) + (in >> 5);
        tby = (r1 << 2) + ((in >> 3) & 3);
        tbx = (r2 << 3) + (in & 7);
    } else {
        tbx = r0;
        tby = r1;
        tbz = r2;
    }
}

// tile coordinates of workgroup w of a single-instance launch
__host__ __device__ __forceinline__ void cf_tile_of(const CfOrder& o, unsigned w, int& tbx, int& tby, int& tbz) {
    cf_tile_at(o, o.mode ? (w & 7) * o.tq + (w >> 3) : w, tbx, tby, tbz);
}

}  // namespace osb
