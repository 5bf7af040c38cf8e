// osb_batch.hpp -- batched kernels: one launch advances every instance of a
// batch group (same dimension, scheme, precision and time scheme; any grid
// sizes).  Each instance's work is a contiguous run of virtual blocks; the
// launch is persistent (grid-stride over virtual blocks), and its grid size is
// a tuning parameter.  The per-point work is the same device code as the
// single-block kernels, so results are bitwise identical to stepping each
// instance alone.
#pragma once

#include "osb_kernels.hpp"
#include "osb_central_fused.hpp"

namespace osb {

struct BItem {
    double* q[2];          // state ping-pong; the input of a stage is q[parity]
    double* du;            // RK-LS register (q_old for the classic RK3)
    double* F;             // interface fluxes, ndim x nv x Ptot
    double* scal;          // [0] wave-speed bits, [1] dt, [2] device time
    Geom g;
    Consts c;
    BcSides bc[3];         // boundary types and Dirichlet values per direction
    double cfl;            // > 0: dt from the wave-speed reduction
    double rkA[5], rkB[5]; // stage coefficients (classic RK3: rknew, rkold)
    CfOrder ord[CF_NTILES];    // 3-D central groups: the walk of the instance's tiles at each tile shape
};

// Items are read through the constant address space: a pointer loaded from
// constant memory is taken to be a global one, so the kernels access the state
// with global loads and stores.  Read as plain memory, the pointers' address
// space is unknown and every access is a FLAT instruction, which also counts
// against the LDS wait counter (a wait for an LDS read then waits for every HBM
// load in flight).  The table is never written while a launch runs.  The host
// pass only type-checks kernel bodies, so the qualifier is device-side only.
#ifdef __HIP_DEVICE_COMPILE__
typedef const __attribute__((address_space(4))) BItem CBItem;
#else
typedef const BItem CBItem;
#endif
__device__ __forceinline__ CBItem& item_const(const BItem* items, int i) { return ((CBItem*)items)[i]; }

// index of the item owning virtual block vb: largest i with offs[i] <= vb
__device__ __forceinline__ int find_item(const int* __restrict__ offs, int n, int vb) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= vb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define OSB_BATCH_LOOP                                                   \
    const int total_ = offs[n];                                          \
    for (int vb = blockIdx.x; vb < total_; vb += gridDim.x)

template <int NDIM, int DIR, int SCH, typename T>
__global__ void __launch_bounds__(BLOCK) kb_face_flux(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                      int parity) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        // fluxes of direction DIR occupy the DIR-th nv x Ptot block of F
        face_flux_item<NDIM, DIR, SCH, T>(t, it.q[parity], it.F + (long)DIR * (NDIM + 2) * it.g.Ptot, it.g, it.c);
    }
}

// last direction's faces fused with the update: one virtual block is one tile
// of W lines (a width of RESID_WS, timed by the batch's autotuner)
template <int NDIM, int SCH, typename T, int TIME, int W>
__global__ void __launch_bounds__(BLOCK) kb_face_resid(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                       int parity, int stage) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        face_resid_tile<NDIM, SCH, T, TIME, W>(vb - offs[i], it.q[parity], it.F, it.q[parity ^ 1], it.du, it.scal[1], it.g, it.c,
                                               it.rkA[stage], it.rkB[stage]);
    }
}

template <int NDIM, int TIME>
__global__ void __launch_bounds__(BLOCK) kb_resid_rk(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                     int parity, int stage) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        resid_rk_item<NDIM, TIME>(t, it.F, it.q[parity], it.q[parity ^ 1], it.du, it.scal[1], it.g, it.c,
                                  it.rkA[stage], it.rkB[stage]);
    }
}

template <int DIR>
__global__ void __launch_bounds__(BLOCK) kb_bc(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                               int parity) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        bc_item<DIR>(t, it.q[parity], it.g, it.g.ndim + 2, it.bc[DIR]);
    }
}

// classic RK3: q_old = q at the start of the step (interior)
static __global__ void __launch_bounds__(BLOCK) kb_save_state(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                       int parity) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const Geom& g = it.g;
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        const int e[3] = {g.n[0], g.n[1], g.n[2]};
        const int lo[3] = {0, 0, 0};
        if (t >= (long)e[0] * e[1] * e[2]) continue;
        int a, b, c;
        unflatten_dv(t, e, g.dv[0], lo, a, b, c);
        const long x = gidx(g, a, b, c);
        for (int v = 0; v < g.ndim + 2; ++v) it.du[(long)v * g.Ptot + x] = it.q[parity][(long)v * g.Ptot + x];
    }
}

// CFL: reset, per-instance max wave speed, dt = cfl / max and time += dt
static __global__ void kb_zero_wavespeed(const BItem* __restrict__ items, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && item_const(items, i).cfl > 0) reinterpret_cast<unsigned long long*>(item_const(items, i).scal)[0] = 0ull;
}

template <int NDIM>
__global__ void __launch_bounds__(BLOCK) kb_wavespeed(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                      int parity) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const Geom& g = it.g;
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        const int e[3] = {g.n[0], g.n[1], g.n[2]};
        const int lo[3] = {0, 0, 0};
        double m = 0.0;
        if (it.cfl > 0 && t < (long)e[0] * e[1] * e[2]) {
            int a, b, c;
            unflatten_dv(t, e, g.dv[0], lo, a, b, c);
            m = wavespeed_point<NDIM>(it.q[parity], gidx(g, a, b, c), g, it.c);
        }
        if (it.cfl > 0) block_max_to(m, it.scal);     // uniform over the block
    }
}

// GLF wave speeds of every GLF item (it.c.glf set), as k_glf_zero / k_glf
static __global__ void kb_glf_zero(const BItem* __restrict__ items, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !item_const(items, i).c.glf) return;
    unsigned long long* z = reinterpret_cast<unsigned long long*>(const_cast<double*>(item_const(items, i).c.glf));
    for (int r = 0; r < 9; ++r) z[r] = 0ull;
}

template <int NDIM>
__global__ void __launch_bounds__(BLOCK) kb_glf(const BItem* __restrict__ items, const int* __restrict__ offs, int n,
                                                int parity) {
    OSB_BATCH_LOOP {
        const int i = find_item(offs, n, vb);
        CBItem& it = item_const(items, i);
        const Geom& g = it.g;
        const long t = (long)(vb - offs[i]) * BLOCK + threadIdx.x;
        const int e[3] = {g.n[0], g.n[1], g.n[2]};
        const int lo[3] = {0, 0, 0};
        double m[3 * NDIM];
#pragma unroll
        for (int r = 0; r < 3 * NDIM; ++r) m[r] = 0.0;
        const bool on = it.c.glf != nullptr;          // uniform over the block: one item per virtual block
        if (on && t < (long)e[0] * e[1] * e[2]) {
            int a, b, c;
            unflatten_dv(t, e, g.dv[0], lo, a, b, c);
            glf_point<NDIM>(it.q[parity], gidx(g, a, b, c), g.Ptot, it.c, m);
        }
        if (on) glf_block_to<NDIM>(m, const_cast<double*>(it.c.glf));
    }
}

// 3-D central RK-LS stage of every item, all faces periodic: one virtual
// block is one TX x TY x TZ tile of an item (a tile shape of CF_TILES, timed
// by the batch's autotuner), evaluated by the single-block kernel's tile code,
// which also writes the wrapped halo images.  An item's tiles are walked in
// its order it.ord (8 x 4 x 2 supertiles z first where the tile counts allow,
// else x, then y, then z), with host-computed quotients.
// Workgroup v of a launch runs on XCD v % 8.  This permutation of [0, total)
// gives XCD x the contiguous range of virtual blocks [x*(q+1), (x+1)*(q+1)) for
// x < r and the ranges of q after them (total = 8q + r), so neighbouring tiles
// of an instance, which share halo points, stay in one L2.
__device__ __forceinline__ int xcd_contiguous(int v, int total) {
    const int q = total >> 3, r = total & 7, x = v & 7, k = v >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
template <int TX, int TY, int TZ, int CONV, int VISC>
__global__ void __launch_bounds__(TX * TY * TZ, cf_waves(TX, TY, TZ, cf_nfield(CONV, VISC)))
kb_central_fused(const BItem* __restrict__ items, const int* __restrict__ offs, int n, int parity, int stage) {
    // one workgroup per tile of every member (the launch covers offs[n] tiles)
    const int vb = xcd_contiguous(blockIdx.x, offs[n]);
    const int i = find_item(offs, n, vb);
    CBItem& it = item_const(items, i);
    constexpr int TI = cf_tile_index(TX, TY, TZ);
    static_assert(TI >= 0, "a tile shape of CF_TILES");
    int tbx, tby, tbz;
    cf_tile_at(it.ord[TI], (unsigned)(vb - offs[i]), tbx, tby, tbz);
    cf_tile_stage<TX, TY, TZ, CONV, VISC, false, OSB_TIME_RKLS3>(tbx * TX, tby * TY, tbz * TZ, it.q[parity], nullptr, it.du,
                                                                 it.q[parity ^ 1], it.scal, it.g, it.c, it.rkA[stage],
                                                                 it.rkB[stage], 1);
}
__host__ inline long kb_central_tiles(const Geom& g, const CfTile& t) {
    return (long)((g.n[0] + t.tx - 1) / t.tx) * ((g.n[1] + t.ty - 1) / t.ty) * ((g.n[2] + t.tz - 1) / t.tz);
}
// an item's tile walk at each tile shape (BItem::ord)
__host__ inline void kb_central_orders(const Geom& g, CfOrder (&ord)[CF_NTILES]) {
    for (int t = 0; t < CF_NTILES; ++t) {
        const CfTile& s = CF_TILES[t];
        ord[t] = cf_order((g.n[0] + s.tx - 1) / s.tx, (g.n[1] + s.ty - 1) / s.ty, (g.n[2] + s.tz - 1) / s.tz, 2);
    }
}

static __global__ void kb_dt_finalize(const BItem* __restrict__ items, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !(item_const(items, i).cfl > 0)) return;
    double* scal = item_const(items, i).scal;
    const double m = __longlong_as_double((long long)reinterpret_cast<unsigned long long*>(scal)[0]);
    const double dt = item_const(items, i).cfl / m;
    scal[1] = dt;
    scal[2] += dt;
}

#undef OSB_BATCH_LOOP

}  // namespace osb
