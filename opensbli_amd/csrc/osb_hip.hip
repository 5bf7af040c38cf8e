// osb_hip.hip -- host runtime and C ABI of the gfx950 implementation of the
// OpenSBLI explicit-solver hot path (include/osb.h).
//
// One osb_ctx owns one block's resident FP64 state on one device:
//   q_a, q_b   conservative variables, ping-pong between RK stages
//   du         RK-LS register (or q_old for the classic RK3)
//   R          residual
//   F          interface fluxes, nv per direction (shock capturing)
//   cr, gr     constituent relations and stored velocity gradients
// all SoA with `halo` ghost layers, x fastest.  Every launch goes to the
// context's stream; nothing synchronises except the calls that return data.

#include "osb_runtime.hpp"
#include "osb_central_mixed.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace osb;

std::string g_osb_create_error;

namespace {

int set_err(osb_ctx* c, int code, const std::string& m) {
    if (c) c->err = m; else g_osb_create_error = m;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            return set_err((ctx), OSB_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
        }                                                                                     \
    } while (0)

inline unsigned nblocks(long work) { return (unsigned)((work + BLOCK - 1) / BLOCK); }

// kernel families reported by osb_profile_read
enum ProfFamily { PF_FACE_FLUX, PF_FLUX_RESIDUAL, PF_CONSTITUENT, PF_GRADIENTS, PF_CENTRAL, PF_VISCOUS,
                  PF_RK, PF_BC, PF_REDUCE, PF_CENTRAL_FUSED, PF_FACE_RESID, PF_FILTER, PF_COUNT };
const char* const kFamilyName[PF_COUNT] = {"face_flux", "flux_residual", "constituent", "gradients", "central_convective",
                                           "viscous", "rk_update", "boundary", "reduction", "central_fused",
                                           "face_resid_rk", "weno_filter"};

hipEvent_t pool_event(osb_ctx* c) {
    if (!c->prof_pool.empty()) { hipEvent_t e = c->prof_pool.back(); c->prof_pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// brackets the launches issued in its scope with events when timing is on
struct Timed {
    osb_ctx* c; int fam; hipEvent_t a = nullptr;
    Timed(osb_ctx* ctx, int f) : c(ctx), fam(f) {
        if (c->prof_on) { a = pool_event(c); (void)hipEventRecord(a, c->stream); }
    }
    ~Timed() {
        if (c->prof_on && a) {
            hipEvent_t b = pool_event(c);
            (void)hipEventRecord(b, c->stream);
            c->prof_pending.push_back({fam, a, b});
        }
    }
};

void prof_resolve(osb_ctx* c) {
    for (auto& r : c->prof_pending) {
        float ms = 0.f;
        (void)hipEventSynchronize(r.b);
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) { c->prof_ms[r.fam] += ms; c->prof_n[r.fam] += 1; }
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_pending.clear();
}

long interior_points(const Geom& g) { return (long)g.n[0] * g.n[1] * g.n[2]; }

long box_points(const Geom& g, int ext) {
    long t = 1;
    for (int d = 0; d < 3; ++d) t *= (d < g.ndim) ? (long)(g.n[d] + 2 * ext) : 1L;
    return t;
}

// ---------------------------------------------------------------------------
// kernel dispatch
// ---------------------------------------------------------------------------
template <int NDIM, int DIR, typename T>
void launch_flux_t(osb_ctx* c, const double* q) {
    double* F = c->F + (long)DIR * c->nv * c->g.Ptot;
    Timed tm(c, PF_FACE_FLUX);
    long work = interior_points(c->g) / c->g.n[DIR] * (long)(c->g.n[DIR] + 1);
    switch (c->sch) {
        case SCH_WENOZ: k_face_flux<NDIM, DIR, SCH_WENOZ, T><<<nblocks(work), BLOCK, 0, c->stream>>>(q, F, c->g, c->c); break;
        case SCH_TENO5: k_face_flux<NDIM, DIR, SCH_TENO5, T><<<nblocks(work), BLOCK, 0, c->stream>>>(q, F, c->g, c->c); break;
        default: k_face_flux<NDIM, DIR, SCH_TENO6, T><<<nblocks(work), BLOCK, 0, c->stream>>>(q, F, c->g, c->c); break;
    }
}

template <int NDIM, int DIR>
void launch_flux(osb_ctx* c, const double* q) {
    if (c->d.flux_type == OSB_FLUX_HLLC || c->d.flux_type == OSB_FLUX_HLLC_LM) {     // FP64 only
        double* F = c->F + (long)DIR * c->nv * c->g.Ptot;
        Timed tm(c, PF_FACE_FLUX);
        const long work = interior_points(c->g) / c->g.n[DIR] * (long)(c->g.n[DIR] + 1);
        if (c->d.flux_type == OSB_FLUX_HLLC_LM) k_face_hllc<NDIM, DIR, true><<<nblocks(work), BLOCK, 0, c->stream>>>(q, F, c->g, c->c);
        else k_face_hllc<NDIM, DIR, false><<<nblocks(work), BLOCK, 0, c->stream>>>(q, F, c->g, c->c);
        return;
    }
    if (c->mixed) launch_flux_t<NDIM, DIR, float>(c, q);
    else launch_flux_t<NDIM, DIR, double>(c, q);
}

// GLF: the stage's global wave speeds, before the face fluxes that read them
void launch_glf(osb_ctx* c, const double* q) {
    if (!c->c.glf) return;
    Timed tm(c, PF_REDUCE);
    double* glf = c->scal + OSB_SCAL_GLF;
    k_glf_zero<<<1, 64, 0, c->stream>>>(glf);
    const unsigned nb = std::min<unsigned>(nblocks(interior_points(c->g)), 2048u);
    switch (c->g.ndim) {
        case 1: k_glf<1><<<nb, BLOCK, 0, c->stream>>>(q, glf, c->g, c->c); break;
        case 2: k_glf<2><<<nb, BLOCK, 0, c->stream>>>(q, glf, c->g, c->c); break;
        default: k_glf<3><<<nb, BLOCK, 0, c->stream>>>(q, glf, c->g, c->c); break;
    }
}

template <int NDIM>
void rhs_nd(osb_ctx* c, const double* q) {
    const long nin = interior_points(c->g);
    if (c->shock) {
        launch_glf(c, q);
        launch_flux<NDIM, 0>(c, q);
        if (NDIM > 1) launch_flux<NDIM, (NDIM > 1 ? 1 : 0)>(c, q);
        if (NDIM > 2) launch_flux<NDIM, (NDIM > 2 ? 2 : 0)>(c, q);
        Timed tm(c, PF_FLUX_RESIDUAL);
        k_flux_residual<NDIM><<<nblocks(nin), BLOCK, 0, c->stream>>>(c->F, c->R, c->g, c->c);
    }
    if (c->central || c->viscous) {
        {
            Timed tm(c, PF_CONSTITUENT);
            k_constituent<NDIM><<<nblocks(box_points(c->g, 2)), BLOCK, 0, c->stream>>>(q, c->cr, c->g, c->c, 2);
        }
        Timed tm(c, PF_GRADIENTS);
        k_gradients<NDIM><<<nblocks(box_points(c->g, 2)), BLOCK, 0, c->stream>>>(c->cr, c->gr, c->g, c->c);
    }
    if (c->central) {
        Timed tm(c, PF_CENTRAL);
        if (c->d.conv_scheme == OSB_CONV_CENTRAL_SKEW)
            k_central_convective<NDIM, CC_SKEW><<<nblocks(nin), BLOCK, 0, c->stream>>>(q, c->cr, c->gr, c->R, c->g, c->c);
        else if (c->d.conv_scheme == OSB_CONV_CENTRAL_KGP)
            k_central_convective<NDIM, CC_KGP><<<nblocks(nin), BLOCK, 0, c->stream>>>(q, c->cr, c->gr, c->R, c->g, c->c);
        else
            k_central_convective<NDIM, CC_FEIEREISEN><<<nblocks(nin), BLOCK, 0, c->stream>>>(q, c->cr, c->gr, c->R, c->g, c->c);
    }
    if (c->viscous) {
        Timed tm(c, PF_VISCOUS);
        if (c->d.viscous == OSB_VISC_SUTHERLAND)
            k_viscous<NDIM, true><<<nblocks(nin), BLOCK, 0, c->stream>>>(c->cr, c->gr, c->R, c->g, c->c);
        else
            k_viscous<NDIM, false><<<nblocks(nin), BLOCK, 0, c->stream>>>(c->cr, c->gr, c->R, c->g, c->c);
    }
}

// Shock capturing without viscous terms: face fluxes, then one kernel that
// differences them and applies the RK update (no residual array).
template <int NDIM, int TIME, typename T, int W>
void launch_face_resid_w(osb_ctx* c, const double* q, double* qn, double A, double B) {
    const unsigned nt = (unsigned)fused_tiles<NDIM, W>(c->g);
    switch (c->sch) {
        case SCH_WENOZ:
            k_face_resid<NDIM, SCH_WENOZ, T, TIME, W><<<nt, BLOCK, 0, c->stream>>>(q, c->F, qn, c->du, c->scal, c->g, c->c, A, B);
            break;
        case SCH_TENO5:
            k_face_resid<NDIM, SCH_TENO5, T, TIME, W><<<nt, BLOCK, 0, c->stream>>>(q, c->F, qn, c->du, c->scal, c->g, c->c, A, B);
            break;
        default:
            k_face_resid<NDIM, SCH_TENO6, T, TIME, W><<<nt, BLOCK, 0, c->stream>>>(q, c->F, qn, c->du, c->scal, c->g, c->c, A, B);
            break;
    }
}

template <int NDIM, int TIME, typename T>
void launch_face_resid(osb_ctx* c, const double* q, double* qn, double A, double B) {
    Timed tm(c, PF_FACE_RESID);
    static_assert(RESID_NW == 2 && RESID_WS[0] == 32 && RESID_WS[1] == 16, "one case per width");
    if (NDIM > 1 && c->resid_w == 16) launch_face_resid_w<NDIM, TIME, T, 16>(c, q, qn, A, B);
    else launch_face_resid_w<NDIM, TIME, T, FUSED_FW>(c, q, qn, A, B);
}
// the fused face-residual launch alone (launch-shape timing)
void launch_face_resid_any(osb_ctx* c, const double* q, double* qn, double A, double B) {
    const bool rk3 = c->d.time_scheme == OSB_TIME_RK3;
    if (c->g.ndim == 2) {
        if (c->mixed) { if (rk3) launch_face_resid<2, 2, float>(c, q, qn, A, B); else launch_face_resid<2, 1, float>(c, q, qn, A, B); }
        else { if (rk3) launch_face_resid<2, 2, double>(c, q, qn, A, B); else launch_face_resid<2, 1, double>(c, q, qn, A, B); }
    } else {
        if (c->mixed) { if (rk3) launch_face_resid<3, 2, float>(c, q, qn, A, B); else launch_face_resid<3, 1, float>(c, q, qn, A, B); }
        else { if (rk3) launch_face_resid<3, 2, double>(c, q, qn, A, B); else launch_face_resid<3, 1, double>(c, q, qn, A, B); }
    }
}

template <int NDIM, int TIME>
void shock_stage_nd(osb_ctx* c, const double* q, double* qn, double A, double B) {
    launch_glf(c, q);
    if (c->shock_fused) {
        // the other directions' fluxes to F, then the last direction fused with the update
        if (NDIM > 1) launch_flux<NDIM, 0>(c, q);
        if (NDIM > 2) launch_flux<NDIM, (NDIM > 2 ? 1 : 0)>(c, q);
        if (c->mixed) launch_face_resid<NDIM, TIME, float>(c, q, qn, A, B);
        else launch_face_resid<NDIM, TIME, double>(c, q, qn, A, B);
        return;
    }
    launch_flux<NDIM, 0>(c, q);
    if (NDIM > 1) launch_flux<NDIM, (NDIM > 1 ? 1 : 0)>(c, q);
    if (NDIM > 2) launch_flux<NDIM, (NDIM > 2 ? 2 : 0)>(c, q);
    Timed tm(c, PF_RK);
    k_resid_rk<NDIM, TIME><<<nblocks(interior_points(c->g)), BLOCK, 0, c->stream>>>(c->F, q, qn, c->du, c->scal, c->g, c->c, A, B);
}

template <int TIME>
void shock_stage(osb_ctx* c, const double* q, double* qn, double A, double B) {
    switch (c->g.ndim) {
        case 1: shock_stage_nd<1, TIME>(c, q, qn, A, B); break;
        case 2: shock_stage_nd<2, TIME>(c, q, qn, A, B); break;
        default: shock_stage_nd<3, TIME>(c, q, qn, A, B); break;
    }
}

void rhs(osb_ctx* c, const double* q) {
    switch (c->g.ndim) {
        case 1: rhs_nd<1>(c, q); break;
        case 2: rhs_nd<2>(c, q); break;
        default: rhs_nd<3>(c, q); break;
    }
}

void apply_bcs(osb_ctx* c, double* q) {
    Timed tm(c, PF_BC);
    const Geom& g = c->g;
    if (c->bc_periodic) {
        k_bc_periodic<<<nblocks(bc_periodic_work(g)), BLOCK, 0, c->stream>>>(q, g, c->nv);
        return;
    }
    for (int dir = 0; dir < g.ndim; ++dir) {
        long work = 2L * (g.halo + 1);
        for (int d = 0; d < g.ndim; ++d)
            if (d != dir) work *= (g.n[d] + 2 * g.halo);
        const BcSides b = bc_sides(c->d, dir);
        if (dir == 0) k_bc<0><<<nblocks(work), BLOCK, 0, c->stream>>>(q, g, c->nv, b);
        else if (dir == 1) k_bc<1><<<nblocks(work), BLOCK, 0, c->stream>>>(q, g, c->nv, b);
        else k_bc<2><<<nblocks(work), BLOCK, 0, c->stream>>>(q, g, c->nv, b);
    }
}

void launch_wavespeed(osb_ctx* c, const double* q) {
    Timed tm(c, PF_REDUCE);
    k_zero_scalar<<<1, 64, 0, c->stream>>>(c->scal);
    const long nin = interior_points(c->g);
    const unsigned nb = std::min<unsigned>(nblocks(nin), 2048u);
    switch (c->g.ndim) {
        case 1: k_wavespeed<1><<<nb, BLOCK, 0, c->stream>>>(q, c->scal, c->g, c->c); break;
        case 2: k_wavespeed<2><<<nb, BLOCK, 0, c->stream>>>(q, c->scal, c->g, c->c); break;
        default: k_wavespeed<3><<<nb, BLOCK, 0, c->stream>>>(q, c->scal, c->g, c->c); break;
    }
}

template <int TX, int TY, int TZ, int CONV, int VISC, bool FLUXRES, int TIME>
void launch_fused_tile(osb_ctx* c, const double* q, double* qn, double A, double B) {
    const Geom& g = c->g;
    const unsigned ntx = (g.n[0] + TX - 1) / TX, nty = (g.n[1] + TY - 1) / TY, ntz = (g.n[2] + TZ - 1) / TZ;
    k_central_fused<TX, TY, TZ, CONV, VISC, FLUXRES, TIME><<<ntx * nty * ntz, TX * TY * TZ, 0, c->stream>>>(
        q, c->F, c->du, qn, c->scal, g, c->c, A, B, cf_order(ntx, nty, ntz, c->xcd_remap),
        (int)(c->bc_periodic && c->bc_in_stage));
}

template <int CONV, int VISC, bool FLUXRES, int TIME>
void launch_fused_t(osb_ctx* c, const double* q, double* qn, double A, double B) {
    switch (c->fused_tile) {
        case 1: launch_fused_tile<16, 8, 4, CONV, VISC, FLUXRES, TIME>(c, q, qn, A, B); break;
        case 2: launch_fused_tile<8, 8, 8, CONV, VISC, FLUXRES, TIME>(c, q, qn, A, B); break;
        default: launch_fused_tile<16, 4, 4, CONV, VISC, FLUXRES, TIME>(c, q, qn, A, B); break;
    }
}

template <int CONV, bool FLUXRES, int TIME>
void launch_fused_v(osb_ctx* c, const double* q, double* qn, double A, double B) {
    switch (c->d.viscous) {
        case OSB_VISC_NONE: if (!FLUXRES) launch_fused_t<CONV, 0, FLUXRES, TIME>(c, q, qn, A, B); break;
        case OSB_VISC_CONSTANT: launch_fused_t<CONV, 1, FLUXRES, TIME>(c, q, qn, A, B); break;
        default: launch_fused_t<CONV, 2, FLUXRES, TIME>(c, q, qn, A, B); break;
    }
}

template <int CONV, int VISC, bool FLUXRES, int TIME>
void launch_mixed_t(osb_ctx* c, const double* q, double* qn, double A, double B) {
    const Geom& g = c->g;
    const unsigned ntx = (g.n[0] + 15) / 16, nty = (g.n[1] + 3) / 4, ntz = (g.n[2] + 3) / 4;
    k_central_mixed<16, 4, 4, CONV, VISC, FLUXRES, TIME><<<ntx * nty * ntz, 256, 0, c->stream>>>(
        q, c->F, c->du, qn, c->scal, g, c->c, A, B, cf_order(ntx, nty, ntz, c->xcd_remap),
        (int)(c->bc_periodic && c->bc_in_stage));
}

template <int CONV, bool FLUXRES, int TIME>
void launch_mixed_v(osb_ctx* c, const double* q, double* qn, double A, double B) {
    switch (c->d.viscous) {
        case OSB_VISC_NONE: if (!FLUXRES) launch_mixed_t<CONV, 0, FLUXRES, TIME>(c, q, qn, A, B); break;
        case OSB_VISC_CONSTANT: launch_mixed_t<CONV, 1, FLUXRES, TIME>(c, q, qn, A, B); break;
        default: launch_mixed_t<CONV, 2, FLUXRES, TIME>(c, q, qn, A, B); break;
    }
}

// One RK stage through the fused kernel: 3-D central schemes, and 3-D
// shock-capturing schemes with viscous terms (face fluxes first).
// Returns whether the stage kernel also wrote qn's halos (the tile kernels,
// FP64 and mixed, with every face periodic), in which case the stage needs no
// halo fill.
template <int TIME>
bool fused_stage(osb_ctx* c, const double* q, double* qn, double A, double B) {
    if (c->shock) {
        launch_glf(c, q);
        launch_flux<3, 0>(c, q);
        launch_flux<3, 1>(c, q);
        launch_flux<3, 2>(c, q);
    }
    Timed tm(c, PF_CENTRAL_FUSED);
    const bool wrapped = c->bc_periodic && c->bc_in_stage;
    if (c->mixed && c->d.conv_scheme != OSB_CONV_CENTRAL_SKEW) {
        // FP32 derivatives on deviations (the skew form stays FP64)
        if (c->shock) launch_mixed_v<CF_NONE, true, TIME>(c, q, qn, A, B);
        else if (c->d.conv_scheme == OSB_CONV_CENTRAL_KGP) launch_mixed_v<CF_KGP, false, TIME>(c, q, qn, A, B);
        else launch_mixed_v<CF_FEIEREISEN, false, TIME>(c, q, qn, A, B);
        return wrapped;
    }
    if (c->d.conv_scheme == OSB_CONV_CENTRAL_KGP) {           // tile kernel only
        launch_fused_v<CF_KGP, false, TIME>(c, q, qn, A, B);
        return wrapped;
    }
    if (c->shock) launch_fused_v<CF_NONE, true, TIME>(c, q, qn, A, B);
    else if (c->d.conv_scheme == OSB_CONV_CENTRAL_SKEW) launch_fused_v<CF_SKEW, false, TIME>(c, q, qn, A, B);
    else launch_fused_v<CF_FEIEREISEN, false, TIME>(c, q, qn, A, B);
    return wrapped;
}

// the fused 3-D stage kernel implements the Feiereisen, skew and KGP forms
// without the Carpenter boundary closure; closure faces run the separate kernels
bool use_fused(const osb_ctx* c) {
    return c->g.ndim == 3 && (c->central || c->viscous) && !(c->c.carp[0] | c->c.carp[1] | c->c.carp[2]);
}
bool use_shock_stage(const osb_ctx* c) { return c->shock && !c->central && !c->viscous; }

int check_launch(osb_ctx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, OSB_E_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return OSB_OK;
}

// ---------------------------------------------------------------------------
// Tile autotuning of the FP64 fused 3-D stage kernel.  At a context's first
// step every tile shape of CF_TILES runs the stage kernel once untimed and
// three times timed (HIP events on the context's stream) on the context's own
// state.  The trial launches write only the first stage's output buffer and
// the RK register: the first real stage overwrites the former, and the
// register is cleared afterwards.  The fastest shape is kept and remembered
// for the process per kernel variant and grid, so contexts of one shape (a
// serving pool's) tune once.  OSB_FUSED_TILE=<index> forces a shape.
// ---------------------------------------------------------------------------
struct TileKey {
    int conv, visc, shock, time, n0, n1, n2;
    bool operator<(const TileKey& o) const {
        const int a[7] = {conv, visc, shock, time, n0, n1, n2}, b[7] = {o.conv, o.visc, o.shock, o.time, o.n0, o.n1, o.n2};
        return std::lexicographical_compare(a, a + 7, b, b + 7);
    }
};
std::map<TileKey, std::pair<int, std::vector<float>>> g_tile_cache;
std::mutex g_tile_mu;

template <int TIME>
void fused_kernel_only(osb_ctx* c, const double* q, double* qn, double A, double B) {
    if (c->d.conv_scheme == OSB_CONV_CENTRAL_KGP) launch_fused_v<CF_KGP, false, TIME>(c, q, qn, A, B);
    else if (c->shock) launch_fused_v<CF_NONE, true, TIME>(c, q, qn, A, B);
    else if (c->d.conv_scheme == OSB_CONV_CENTRAL_SKEW) launch_fused_v<CF_SKEW, false, TIME>(c, q, qn, A, B);
    else launch_fused_v<CF_FEIEREISEN, false, TIME>(c, q, qn, A, B);
}

// Width of the fused face-residual tiles (2-D/3-D shock capturing without
// viscous terms), timed like the tile shapes: each RESID_WS width runs the
// launch once untimed and three times timed on the context's state (writing
// the next buffer and the RK register, which is cleared afterwards), the
// fastest kept and remembered per kernel variant and grid.  OSB_RESID_W=<16|32>
// forces a width.
std::map<TileKey, std::pair<int, std::vector<float>>> g_resid_cache;

int tune_resid_w(osb_ctx* c) {
    const TileKey key{c->sch, (int)c->mixed, c->g.ndim, c->d.time_scheme, c->g.n[0], c->g.n[1], c->g.n[2]};
    {
        std::lock_guard<std::mutex> lock(g_tile_mu);
        auto it = g_resid_cache.find(key);
        if (it != g_resid_cache.end()) {
            c->resid_w = it->second.first;
            c->resid_w_us = it->second.second;
            return OSB_OK;
        }
    }
    double* qn = (c->q_cur == c->qa) ? c->qb : c->qa;
    const bool rk3 = c->d.time_scheme == OSB_TIME_RK3;
    const double A = rk3 ? 2.0 / 3.0 : c->rkA[0], B = rk3 ? 1.0 / 4.0 : c->rkB[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    hipError_t ec = hipEventCreate(&e1);
    if (ec != hipSuccess) { (void)hipEventDestroy(e0); HIP_TRY(c, ec); }
    // each width timed over whole stages (the other directions' face fluxes,
    // then the face-residual launch), in TUNE_ROUNDS rounds alternating
    // between the widths, keeping each width's best round
    auto stage = [&]() {
        if (rk3) shock_stage<2>(c, c->q_cur, qn, A, B);
        else shock_stage<1>(c, c->q_cur, qn, A, B);
    };
    c->resid_w_us.assign(RESID_NW, 1e30f);
    hipError_t e = hipSuccess;
    const bool prof = c->prof_on;            // the trial stages stay out of the kernel timings
    c->prof_on = false;
    for (int round = 0; round < TUNE_ROUNDS && e == hipSuccess; ++round)
        for (int t = 0; t < RESID_NW && e == hipSuccess; ++t) {
            c->resid_w = RESID_WS[t];
            stage();
            e = hipEventRecord(e0, c->stream);
            for (int r = 0; r < 3 && e == hipSuccess; ++r) stage();
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            c->resid_w_us[t] = std::min(c->resid_w_us[t], ms * 1000.f / 3.f);
        }
    c->prof_on = prof;
    int best_w = FUSED_FW;
    float best = 1e30f;
    for (int t = 0; t < RESID_NW; ++t)
        if (c->resid_w_us[t] < best) { best = c->resid_w_us[t]; best_w = RESID_WS[t]; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(c, e);
    c->resid_w = best_w;
    HIP_TRY(c, hipMemsetAsync(c->du, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
    int r = check_launch(c);
    if (r) return r;
    std::lock_guard<std::mutex> lock(g_tile_mu);
    g_resid_cache[key] = {best_w, c->resid_w_us};
    return OSB_OK;
}

int tune_fused_tile(osb_ctx* c) {
    c->tile_tuned = true;
    if (use_shock_stage(c) && c->shock_fused && c->g.ndim >= 2) return c->resid_forced ? OSB_OK : tune_resid_w(c);
    if (!use_fused(c) || c->mixed || c->tile_forced) return OSB_OK;
    const TileKey key{c->d.conv_scheme, c->d.viscous, (int)c->shock, c->d.time_scheme, c->g.n[0], c->g.n[1], c->g.n[2]};
    {
        std::lock_guard<std::mutex> lock(g_tile_mu);
        auto it = g_tile_cache.find(key);
        if (it != g_tile_cache.end()) {
            c->fused_tile = it->second.first;
            c->tile_us = it->second.second;
            return OSB_OK;
        }
    }
    double* qn = (c->q_cur == c->qa) ? c->qb : c->qa;
    const bool rk3 = c->d.time_scheme == OSB_TIME_RK3;
    const double A = rk3 ? 2.0 / 3.0 : c->rkA[0], B = rk3 ? 1.0 / 4.0 : c->rkB[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    hipError_t ec = hipEventCreate(&e1);
    if (ec != hipSuccess) { (void)hipEventDestroy(e0); HIP_TRY(c, ec); }
    auto launch = [&]() {
        if (rk3) fused_kernel_only<2>(c, c->q_cur, qn, A, B);
        else fused_kernel_only<1>(c, c->q_cur, qn, A, B);
    };
    // pass 0: one warm-up and three timed launches per shape (an estimate);
    // pass 1: each shape timed over enough launches for ~0.5 ms of work, so
    // that launch gaps and timer resolution do not decide between small grids
    float best = 1e30f;
    int best_t = 0;
    c->tile_us.assign(CF_NTILES, 0.f);
    hipError_t e = hipSuccess;
    auto time_shape = [&](int t, int reps) -> float {
        c->fused_tile = t;
        launch();
        e = hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) launch();
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        return ms * 1000.f / (float)reps;
    };
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
        for (int t = 0; t < CF_NTILES && e == hipSuccess; ++t) {
            const int reps = pass == 0 ? 3 : std::min(200, std::max(3, (int)(500.f / std::max(c->tile_us[t], 1.f))));
            c->tile_us[t] = time_shape(t, reps);
        }
    }
    for (int t = 0; t < CF_NTILES; ++t)
        if (c->tile_us[t] < best) { best = c->tile_us[t]; best_t = t; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(c, e);
    c->fused_tile = best_t;
    HIP_TRY(c, hipMemsetAsync(c->du, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
    int r = check_launch(c);
    if (r) return r;
    std::lock_guard<std::mutex> lock(g_tile_mu);
    g_tile_cache[key] = {best_t, c->tile_us};
    return OSB_OK;
}

// WENO filter after the full step, in place on q (WENO_filter.py:406-536):
// the Ducros sensor (HYB), the filter fluxes per direction, the update
template <int NDIM, int ORDER, bool HYB>
void filter_nd_t(osb_ctx* c, double* q) {
    Consts fc = c->c;
    fc.avg_roe = 1;                       // the filter's LFWeno uses RoeAverage (WENO_filter.py:438)
    fc.glf = nullptr;
    const Geom& g = c->g;
    Timed tm(c, PF_FILTER);
    if constexpr (HYB && NDIM >= 2) {
        double sd = 0.0;
        for (int d = 0; d < NDIM; ++d) sd += c->d.delta[d] * c->d.delta[d];
        // the face-flux array is free until the filter faces: it holds u for the sensor
        k_velocity<NDIM><<<nblocks(box_points(g, 2)), BLOCK, 0, c->stream>>>(q, c->F, g);
        k_ducros<NDIM><<<nblocks(interior_points(g)), BLOCK, 0, c->stream>>>(c->F, c->kappa, g, fc, std::sqrt(sd));
    }
    constexpr int D1 = NDIM > 1 ? 1 : 0, D2 = NDIM > 2 ? 2 : 0;
    const long P = (long)c->nv * g.Ptot;
    k_filter_face<NDIM, 0, ORDER, HYB><<<nblocks(face_flux_work<NDIM, 0>(g)), BLOCK, 0, c->stream>>>(q, c->kappa, c->F, g, fc);
    if (NDIM > 1)
        k_filter_face<NDIM, D1, ORDER, HYB><<<nblocks(face_flux_work<NDIM, D1>(g)), BLOCK, 0, c->stream>>>(q, c->kappa, c->F + P, g, fc);
    if (NDIM > 2)
        k_filter_face<NDIM, D2, ORDER, HYB><<<nblocks(face_flux_work<NDIM, D2>(g)), BLOCK, 0, c->stream>>>(q, c->kappa, c->F + 2 * P, g, fc);
    k_filter_update<NDIM, HYB><<<nblocks(interior_points(g)), BLOCK, 0, c->stream>>>(c->F, c->kappa, q, c->scal, g, fc);
}

template <int NDIM>
void filter_nd(osb_ctx* c, double* q) {
    if (c->d.filter_order == 3) {
        if (c->d.filter_sensor == OSB_SENSOR_DUCROS) filter_nd_t<NDIM, 3, true>(c, q);
        else filter_nd_t<NDIM, 3, false>(c, q);
    } else {
        if (c->d.filter_sensor == OSB_SENSOR_DUCROS) filter_nd_t<NDIM, 5, true>(c, q);
        else filter_nd_t<NDIM, 5, false>(c, q);
    }
}

void apply_filter(osb_ctx* c, double* q) {
    switch (c->g.ndim) {
        case 1: filter_nd<1>(c, q); break;
        case 2: filter_nd<2>(c, q); break;
        default: filter_nd<3>(c, q); break;
    }
}

// RK stage s of the current step (stage 0 also does the step's start: the
// CFL dt, the classic RK3 saved state), ending with the boundary conditions.
// With the WENO filter the step starts with the boundary conditions (the
// filter leaves the halos stale, algorithm.py:462) and ends with the filter.
void do_stage(osb_ctx* c, int s) {
    const Geom& g = c->g;
    const long nin = interior_points(g);
    const bool filter = c->d.shock_filter == OSB_FILTER_WENOZ;
    if (s == 0 && filter) apply_bcs(c, c->q_cur);
    if (s == 0 && c->d.cfl > 0) {
        launch_wavespeed(c, c->q_cur);
        k_dt_finalize<<<1, 64, 0, c->stream>>>(c->scal, c->d.cfl);
    }
    double* qn = (c->q_cur == c->qa) ? c->qb : c->qa;
    bool halos_done = false;
    if (c->d.time_scheme == OSB_TIME_RKLS3) {
        if (use_fused(c)) {
            halos_done = fused_stage<1>(c, c->q_cur, qn, c->rkA[s], c->rkB[s]);
        } else if (use_shock_stage(c)) {
            shock_stage<1>(c, c->q_cur, qn, c->rkA[s], c->rkB[s]);
        } else {
            rhs(c, c->q_cur);
            Timed tm(c, PF_RK);
            k_rkls<<<nblocks(nin), BLOCK, 0, c->stream>>>(c->q_cur, qn, c->du, c->R, c->scal, g, c->nv, c->rkA[s], c->rkB[s]);
        }
    } else {
        const double rkold[3] = {1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0};
        const double rknew[3] = {2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0};
        if (s == 0) k_copy_interior<<<nblocks(nin), BLOCK, 0, c->stream>>>(c->q_cur, c->du, g, c->nv);
        if (use_fused(c)) {
            halos_done = fused_stage<2>(c, c->q_cur, qn, rknew[s], rkold[s]);
        } else if (use_shock_stage(c)) {
            shock_stage<2>(c, c->q_cur, qn, rknew[s], rkold[s]);
        } else {
            rhs(c, c->q_cur);
            Timed tm(c, PF_RK);
            k_rk3_stage<<<nblocks(nin), BLOCK, 0, c->stream>>>(qn, c->du, c->R, c->scal, g, c->nv, rknew[s], rkold[s]);
        }
    }
    if (!halos_done) apply_bcs(c, qn);
    c->q_prev = c->q_cur;
    c->q_cur = qn;
    const int nst = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
    if (filter && s == nst - 1) apply_filter(c, c->q_cur);
}

int stage_count(const osb_ctx* c) { return c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3; }

int do_step(osb_ctx* c) {
    for (int s = 0; s < stage_count(c); ++s) do_stage(c, s);
    c->have_prev = true;
    return check_launch(c);
}

int validate(const osb_problem_desc* d, std::string& why) {
    if (!d) { why = "null descriptor"; return OSB_E_INVALID; }
    if (d->struct_size != sizeof(osb_problem_desc) || d->abi_version != OSB_ABI_VERSION) { why = "descriptor size/abi mismatch"; return OSB_E_INVALID; }
    if (d->ndim < 1 || d->ndim > 3) { why = "ndim must be 1..3"; return OSB_E_INVALID; }
    for (int i = 0; i < 3; ++i) {
        if (i < d->ndim && (d->n[i] < 8 || !(d->delta[i] > 0))) { why = "n >= 8 and delta > 0 required in active directions"; return OSB_E_INVALID; }
        if (i >= d->ndim && d->n[i] != 1) { why = "inactive directions must have n == 1"; return OSB_E_INVALID; }
    }
    {
        // points per variable, halo included (at most 4 deep), must fit the kernels' 32-bit point offsets
        double pts = 1.0;
        for (int i = 0; i < d->ndim; ++i) pts *= (double)d->n[i] + 8.0;
        if (pts >= 4294967296.0) { why = "more than 2^32 points per variable"; return OSB_E_UNSUPPORTED; }
    }
    const int cs = d->conv_scheme;
    const bool central_cs = cs == OSB_CONV_CENTRAL_FEIEREISEN || cs == OSB_CONV_CENTRAL_SKEW || cs == OSB_CONV_CENTRAL_KGP;
    if (cs != OSB_CONV_WENO_Z && cs != OSB_CONV_TENO && !central_cs) { why = "unknown conv_scheme"; return OSB_E_UNSUPPORTED; }
    if (cs == OSB_CONV_CENTRAL_KGP && d->precision == OSB_PREC_MIXED && d->ndim != 3) { why = "mixed precision for the KGP split form is built for 3-D blocks"; return OSB_E_UNSUPPORTED; }
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

int dalloc(osb_ctx* c, double** p, size_t count) {
    const size_t bytes = count * sizeof(double);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();      // clear the runtime's sticky error so later launch checks see only their own
        return set_err(nullptr, OSB_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemset(*p, 0, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_err(nullptr, OSB_E_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    c->bytes += bytes;
    return OSB_OK;
}

Geom make_geom(const osb_problem_desc* desc) {
    const int nd = desc->ndim;
    const bool shock = desc->conv_scheme == OSB_CONV_WENO_Z || desc->conv_scheme == OSB_CONV_TENO;
    Geom g{};
    g.ndim = nd;
    // symmetric scheme halo: WENO5/TENO [-3,4] -> 4, Central(4) -> 2; the WENO
    // filter's reconstruction needs the WENO halo
    g.halo = (shock || desc->shock_filter == OSB_FILTER_WENOZ) ? 4 : 2;
    for (int i = 0; i < 3; ++i) {
        g.n[i] = desc->n[i];
        g.P[i] = i < nd ? desc->n[i] + 2 * g.halo : 1;
        g.o[i] = i < nd ? g.halo : 0;
    }
    g.s[0] = 1;
    g.s[1] = g.P[0];
    g.s[2] = (long)g.P[0] * g.P[1];
    g.Ptot = (long)g.P[0] * g.P[1] * g.P[2];
    geom_divisors(g);
    return g;
}

constexpr int PART_BLOCKS = 1024;

// device arrays (in doubles) a context allocates at creation.  The residual
// and the constituent-relation / gradient arrays are needed for stepping only
// by the unfused central path (ndim < 3); otherwise osb_residual and
// osb_tgv_diagnostics allocate them on first use.
struct AllocPlan {
    size_t q = 0, R = 0, F = 0, kappa = 0, cr = 0, gr = 0, staging = 0;
    size_t bytes() const {
        return sizeof(double) * (3 * q + R + F + kappa + cr + gr + OSB_SCAL_SIZE + staging + (size_t)PART_BLOCKS * 4) +
               sizeof(unsigned int);
    }
};

AllocPlan plan_alloc(const osb_problem_desc* d) {
    const Geom g = make_geom(d);
    const int nd = d->ndim, nv = nd + 2;
    const bool shock = d->conv_scheme == OSB_CONV_WENO_Z || d->conv_scheme == OSB_CONV_TENO;
    const bool central = !shock, viscous = d->viscous != OSB_VISC_NONE;
    bool carp = false;
    for (int i = 0; i < nd; ++i) carp = carp || d->bc[i][0] >= OSB_BC_EXTRAPOLATION0 || d->bc[i][1] >= OSB_BC_EXTRAPOLATION0;
    const bool fused = nd == 3 && (central || viscous) && !carp;
    const bool unfused_central = (central || viscous) && !fused;
    const size_t P = (size_t)g.Ptot;
    AllocPlan a;
    a.q = (size_t)nv * P;
    a.R = unfused_central ? a.q : 0;
    a.F = (shock || d->shock_filter == OSB_FILTER_WENOZ) ? (size_t)nd * a.q : 0;
    a.kappa = (d->shock_filter == OSB_FILTER_WENOZ && d->filter_sensor == OSB_SENSOR_DUCROS) ? P : 0;
    a.cr = unfused_central ? (size_t)(nd + 3) * P : 0;
    a.gr = unfused_central ? (size_t)nd * nd * P : 0;
    a.staging = (size_t)nv * (size_t)g.n[0] * g.n[1] * g.n[2];
    return a;
}

int ensure_buf(osb_ctx* c, double** p, size_t count) {
    if (*p) return OSB_OK;
    int r = dalloc(c, p, count);
    if (r) c->err = g_osb_create_error;
    return r;
}

// arrays of the unfused right-hand side (osb_residual on any path)
int ensure_rhs_arrays(osb_ctx* c) {
    const size_t P = (size_t)c->g.Ptot;
    const int nd = c->g.ndim;
    int r = ensure_buf(c, &c->R, (size_t)c->nv * P);
    if (!r && (c->central || c->viscous)) r = ensure_buf(c, &c->cr, (size_t)(nd + 3) * P);
    if (!r && (c->central || c->viscous)) r = ensure_buf(c, &c->gr, (size_t)nd * nd * P);
    return r;
}

void drop_graphs(osb_ctx* c) {
    for (auto& e : c->gexec)
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
}

void free_ctx(osb_ctx* c) {
    if (!c) return;
    drop_graphs(c);
    double* bufs[] = {c->qa, c->qb, c->du, c->R, c->F, c->kappa, c->cr, c->gr, c->scal, c->staging, c->part};
    for (double* b : bufs) if (b) (void)hipFree(b);
    if (c->nan_cnt) (void)hipFree(c->nan_cnt);
    for (auto& r : c->prof_pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Every field of a context that needs no device memory: the descriptor,
// geometry, constants, RK coefficients and the kernel choices (with their
// environment overrides).  osb_ctx_create and osb_step_noalloc share it.
void init_ctx(osb_ctx* c, const osb_problem_desc* desc) {
    c->d = *desc;
    (void)hipGetDevice(&c->device);
    const int nd = desc->ndim;
    c->nv = nd + 2;
    c->shock = desc->conv_scheme == OSB_CONV_WENO_Z || desc->conv_scheme == OSB_CONV_TENO;
    c->central = !c->shock;
    c->viscous = desc->viscous != OSB_VISC_NONE;
    c->mixed = desc->precision == OSB_PREC_MIXED;
    if (desc->rk_stages == 0) {
        // RungeKuttaLS(3) Williamson coefficients (rk_LS.py:121-124)
        const double A[3] = {0, (-5.0 / 9.0), (-153.0 / 128.0)};
        const double B[3] = {(1.0 / 3.0), (15.0 / 16.0), (8.0 / 15.0)};
        c->rk_stages = 3;
        for (int s = 0; s < 3; ++s) { c->rkA[s] = A[s]; c->rkB[s] = B[s]; }
    } else {
        c->rk_stages = desc->rk_stages;
        for (int s = 0; s < desc->rk_stages; ++s) { c->rkA[s] = desc->rk_a[s]; c->rkB[s] = desc->rk_b[s]; }
    }
    c->sch = desc->conv_scheme == OSB_CONV_WENO_Z ? SCH_WENOZ : (desc->order == 6 ? SCH_TENO6 : SCH_TENO5);

    c->g = make_geom(desc);
    const Geom& g = c->g;

    Consts& k = c->c;
    k.gama = desc->gama;
    k.gamma_m1 = desc->gama - 1.0;
    k.minf = desc->minf;
    k.sfc = desc->shock_filter_control;
    k.weno_eps = desc->weno_eps;
    k.teno_eps = desc->teno_eps;
    k.teno_ct = desc->teno_ct;
    k.invRe = c->viscous ? 1.0 / desc->re : 0.0;
    k.cheat = c->viscous ? 1.0 / ((desc->gama - 1.0) * desc->minf * desc->minf * desc->pr * desc->re) : 0.0;
    for (int i = 0; i < 3; ++i) {
        k.invd[i] = 1.0 / desc->delta[i];
        k.inv2d[i] = 1.0 / (desc->delta[i] * desc->delta[i]);
    }
    k.cfl = desc->cfl;
    k.dt = desc->dt;
    k.avg_roe = desc->averaging == OSB_AVG_ROE;
    k.viscous = desc->viscous;
    k.glf = nullptr;
    k.ducros_check = desc->ducros_check;
    k.ducros_select = desc->ducros_select;
    fill_fused_consts(k);
    for (int i = 0; i < 3; ++i) {
        k.carp[i] = 0;
        for (int sd = 0; sd < 2; ++sd)
            if (i < desc->ndim && desc->bc[i][sd] >= OSB_BC_EXTRAPOLATION0) k.carp[i] |= 1 << sd;
    }

    if (const char* e = std::getenv("OSB_FUSED_TILE")) {
        c->fused_tile = std::max(0, std::min(CF_NTILES - 1, std::atoi(e)));
        c->tile_forced = true;
    }
    if (const char* e = std::getenv("OSB_RESID_W")) {
        const int w = std::atoi(e);
        for (int t = 0; t < RESID_NW; ++t)
            if (RESID_WS[t] == w) { c->resid_w = w; c->resid_forced = true; }
    }
    if (const char* e = std::getenv("OSB_GRAPH")) c->use_graph = std::atoi(e) != 0;
    if (desc->flux_type >= OSB_FLUX_HLLC) c->shock_fused = false;   // HLLC faces run their own kernel and the separate update
    c->bc_periodic = true;
    for (int d = 0; d < c->g.ndim; ++d)
        if (desc->bc[d][0] != OSB_BC_PERIODIC || desc->bc[d][1] != OSB_BC_PERIODIC || c->g.n[d] < c->g.halo) c->bc_periodic = false;
    for (int d = 0; d < c->g.ndim; ++d)
        if (c->g.n[d] < 2 * c->g.halo) c->bc_in_stage = false;
    if (const char* e = std::getenv("OSB_BC_KERNEL")) {
        if (std::strcmp(e, "dir") == 0) c->bc_periodic = false;
        if (std::strcmp(e, "pass") == 0) c->bc_in_stage = false;
    }
}


// Allocation-free stepping (osb_step_noalloc): every device buffer of a
// context beyond the caller's state and RK register comes from this array,
// one copy per GPU, held until the process ends.  Host code uses its
// per-device address from hipGetSymbolAddress, never the array's name.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = size_t(1) << 20;       // 8 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// the current device's copy of `scratch` (looked up once per device)
int noalloc_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { (void)hipGetLastError(); return set_err(nullptr, OSB_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e)); }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_err(nullptr, OSB_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return set_err(nullptr, OSB_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return OSB_OK;
}

// clear n doubles of the scratch, then set the scalars' dt (scal[1]): fill
// kernels rather than memsets, so the call can be captured into a hipGraph
__global__ void k_noalloc_clear(double* p, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0.0;
}
__global__ void k_set_dt(double* scal, double dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[1] = dt;
}
// the padded state back into the caller's buffer
__global__ void k_copy_all(double* __restrict__ dst, const double* __restrict__ src, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) dst[t] = src[t];
}

}  // namespace

extern "C" {

int osb_abi_version(void) { return OSB_ABI_VERSION; }

int osb_estimate_bytes(const osb_problem_desc* desc, int64_t* out) {
    if (!out) return set_err(nullptr, OSB_E_INVALID, "null output pointer");
    std::string why;
    int rc = validate(desc, why);
    if (rc) return set_err(nullptr, rc, why);
    *out = (int64_t)plan_alloc(desc).bytes();
    return OSB_OK;
}

const char* osb_last_error(const osb_ctx* ctx) { return ctx ? ctx->err.c_str() : g_osb_create_error.c_str(); }

int osb_ctx_create(const osb_problem_desc* desc, osb_ctx** out) {
    if (!out) return set_err(nullptr, OSB_E_INVALID, "null output pointer");
    *out = nullptr;
    std::string why;
    int rc = validate(desc, why);
    if (rc) return set_err(nullptr, rc, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(nullptr, OSB_E_DEVICE, "no HIP device available");

    osb_ctx* c = new osb_ctx();
    init_ctx(c, desc);
    Consts& k = c->c;
    const AllocPlan ap = plan_alloc(desc);
    int r = OSB_OK;
    if (!r) r = dalloc(c, &c->qa, ap.q);
    if (!r) r = dalloc(c, &c->qb, ap.q);
    if (!r) r = dalloc(c, &c->du, ap.q);
    if (!r && ap.R) r = dalloc(c, &c->R, ap.R);
    if (!r && ap.F) r = dalloc(c, &c->F, ap.F);
    if (!r && ap.kappa) r = dalloc(c, &c->kappa, ap.kappa);      // zeroed: the halos stay 0
    if (!r && ap.cr) r = dalloc(c, &c->cr, ap.cr);
    if (!r && ap.gr) r = dalloc(c, &c->gr, ap.gr);
    if (!r) r = dalloc(c, &c->scal, OSB_SCAL_SIZE);
    if (!r && c->shock && desc->flux_type == OSB_FLUX_GLF) k.glf = c->scal + OSB_SCAL_GLF;
    if (!r) r = dalloc(c, &c->staging, ap.staging);
    c->part_blocks = PART_BLOCKS;
    if (!r) r = dalloc(c, &c->part, (size_t)c->part_blocks * 4);
    if (!r) {
        hipError_t e = hipMalloc((void**)&c->nan_cnt, sizeof(unsigned int));
        if (e != hipSuccess) { c->nan_cnt = nullptr; (void)hipGetLastError(); r = set_err(nullptr, OSB_E_NOMEM, "hipMalloc nan counter"); }
    }
    if (!r) {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) r = set_err(nullptr, OSB_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        else c->own_stream = true;
    }
    if (!r) {
        double s[16] = {0};
        s[1] = desc->dt;
        hipError_t e = hipMemcpy(c->scal, s, sizeof(s), hipMemcpyHostToDevice);
        if (e != hipSuccess) r = set_err(nullptr, OSB_E_DEVICE, "scalar upload");
    }
    if (r) { free_ctx(c); return r; }
    c->q_cur = c->qa;
    c->q_prev = c->qb;
    *out = c;
    return OSB_OK;
}

int osb_ctx_destroy(osb_ctx* c) {
    if (!c) return OSB_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return OSB_OK;
}

int osb_set_stream(osb_ctx* c, void* s) {
    if (!c) return OSB_E_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->own_stream) { (void)hipStreamDestroy(c->stream); c->own_stream = false; }
    c->stream = (hipStream_t)s;
    drop_graphs(c);                        // recapture on the new stream
    return OSB_OK;
}

int osb_set_state(osb_ctx* c, const double* host, int layout) {
    if (!c || !host) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR && layout != OSB_LAYOUT_PADDED) return set_err(c, OSB_E_INVALID, "unknown layout");
    const long npts = interior_points(c->g);
    HIP_TRY(c, hipMemsetAsync(c->du, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
    if (layout == OSB_LAYOUT_PADDED) {
        // the whole padded array; the boundary conditions below refill the halos
        HIP_TRY(c, hipMemcpyAsync(c->qa, host, sizeof(double) * c->nv * c->g.Ptot, hipMemcpyHostToDevice, c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(c->staging, host, sizeof(double) * npts * c->nv, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemsetAsync(c->qa, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
        k_scatter<<<nblocks(npts * c->nv), BLOCK, 0, c->stream>>>(c->staging, c->qa, c->g, c->nv);
    }
    c->q_cur = c->qa;
    c->q_prev = c->qb;
    c->have_prev = false;
    c->next_stage = 0;
    apply_bcs(c, c->q_cur);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_get_state(osb_ctx* c, double* host, int layout) {
    if (!c || !host) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR && layout != OSB_LAYOUT_PADDED) return set_err(c, OSB_E_INVALID, "unknown layout");
    if (layout == OSB_LAYOUT_PADDED) {
        HIP_TRY(c, hipMemcpyAsync(host, c->q_cur, sizeof(double) * c->nv * c->g.Ptot, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return check_launch(c);
    }
    const long npts = interior_points(c->g);
    k_gather<<<nblocks(npts * c->nv), BLOCK, 0, c->stream>>>(c->q_cur, c->staging, c->g, c->nv);
    HIP_TRY(c, hipMemcpyAsync(host, c->staging, sizeof(double) * npts * c->nv, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_set_time(osb_ctx* c, double time, int64_t iter) {
    if (!c) return OSB_E_INVALID;
    c->time = time;
    c->iter = iter;
    double t = time;
    HIP_TRY(c, hipMemcpyAsync(c->scal + 2, &t, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return OSB_OK;
}

int osb_apply_bcs(osb_ctx* c) {
    if (!c) return OSB_E_INVALID;
    apply_bcs(c, c->q_cur);
    return check_launch(c);
}

int osb_residual(osb_ctx* c, double* host_out) {
    if (!c || !host_out) return OSB_E_INVALID;
    int r = ensure_rhs_arrays(c);
    if (r) return r;
    apply_bcs(c, c->q_cur);
    rhs(c, c->q_cur);
    const long npts = interior_points(c->g);
    k_gather<<<nblocks(npts * c->nv), BLOCK, 0, c->stream>>>(c->R, c->staging, c->g, c->nv);
    HIP_TRY(c, hipMemcpyAsync(host_out, c->staging, sizeof(double) * npts * c->nv, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_max_wavespeed(osb_ctx* c, double* out) {
    if (!c || !out) return OSB_E_INVALID;
    launch_wavespeed(c, c->q_cur);
    unsigned long long bits = 0;
    HIP_TRY(c, hipMemcpyAsync(&bits, c->scal, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *out = v;
    return check_launch(c);
}

// One time step replayed from a captured hipGraph (one graph per starting
// buffer).  The graph holds exactly the launches of do_step, so results are
// those of the eager path; per-launch timing needs the eager path.
int graph_step(osb_ctx* c) {
    const int p = c->q_cur == c->qa ? 0 : 1;
    if (!c->gexec[p]) {
        hipGraph_t graph = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int r = do_step(c);
        hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        if (e != hipSuccess) return set_err(c, OSB_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->gexec[p], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return set_err(c, OSB_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->g_cur[p] = c->q_cur;            // do_step advanced the host view
        c->g_prev[p] = c->q_prev;
    } else {
        c->q_cur = c->g_cur[p];
        c->q_prev = c->g_prev[p];
        c->have_prev = true;
    }
    HIP_TRY(c, hipGraphLaunch(c->gexec[p], c->stream));
    return OSB_OK;
}

int osb_step(osb_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return OSB_E_INVALID;
    if (c->next_stage != 0) return set_err(c, OSB_E_INVALID, "a step is partly advanced by osb_stage; finish its stages first");
    if (nsteps > 0 && !c->tile_tuned) {
        int r = tune_fused_tile(c);
        if (r) return r;
    }
    const bool graphs = c->use_graph && !c->prof_on;
    for (int i = 0; i < nsteps; ++i) {
        int r = graphs ? graph_step(c) : do_step(c);
        if (r) return r;
        c->iter += 1;
        if (!(c->d.cfl > 0)) { c->time += c->d.dt; c->last_dt = c->d.dt; }
    }
    return OSB_OK;
}

int osb_stage(osb_ctx* c, int s) {
    if (!c) return OSB_E_INVALID;
    if (s < 0 || s >= stage_count(c)) return set_err(c, OSB_E_INVALID, "stage index out of range");
    if (s != c->next_stage)
        return set_err(c, OSB_E_INVALID, "stages run in order: expected stage " + std::to_string(c->next_stage));
    if (s == 0 && !c->tile_tuned) {
        int r = tune_fused_tile(c);
        if (r) return r;
    }
    do_stage(c, s);                       // eager: single stages are not captured
    c->next_stage = (s + 1) % stage_count(c);
    if (c->next_stage == 0) {
        c->have_prev = true;
        c->iter += 1;
        if (!(c->d.cfl > 0)) { c->time += c->d.dt; c->last_dt = c->d.dt; }
    }
    return check_launch(c);
}

int osb_sync(osb_ctx* c) {
    if (!c) return OSB_E_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_get_stats(osb_ctx* c, osb_stats* out) {
    if (!c || !out) return OSB_E_INVALID;
    out->struct_size = sizeof(osb_stats);
    out->iter = c->iter;
    if (c->d.cfl > 0) {
        double s[3];
        HIP_TRY(c, hipMemcpyAsync(s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->time = s[2];
        c->last_dt = s[1];
    }
    out->time = c->time;
    out->last_dt = c->last_dt;
    out->halo = c->g.halo;
    out->nv = c->nv;
    out->device_bytes = (int64_t)c->bytes;
    return OSB_OK;
}

int osb_tgv_diagnostics(osb_ctx* c, double out[4]) {
    if (!c || !out) return OSB_E_INVALID;
    if (c->g.ndim != 3) return set_err(c, OSB_E_UNSUPPORTED, "TGV diagnostics need ndim == 3");
    if (!(c->central || c->viscous)) return set_err(c, OSB_E_UNSUPPORTED, "TGV diagnostics need a viscous configuration");
    int r = ensure_buf(c, &c->cr, (size_t)(c->g.ndim + 3) * c->g.Ptot);
    if (r) return r;
    // velocity and mu from the state at the start of the last stage (the
    // reference's constituent-relation arrays), rho from the current state
    const double* qcr = c->have_prev ? c->q_prev : c->q_cur;
    k_constituent<3><<<nblocks(box_points(c->g, 2)), BLOCK, 0, c->stream>>>(qcr, c->cr, c->g, c->c, 2);
    const long nin = interior_points(c->g);
    const unsigned nb = std::min<unsigned>(nblocks(nin), (unsigned)c->part_blocks);
    k_tgv_partials<<<nb, BLOCK, 0, c->stream>>>(c->q_cur, c->cr, c->part, c->g, c->c);
    k_sum_partials<<<1, 64, 0, c->stream>>>(c->part, c->scal + 4, (int)nb);
    HIP_TRY(c, hipMemcpyAsync(out, c->scal + 4, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_probe(osb_ctx* c, int n, const int32_t* ijk, const int32_t* var, double* out) {
    if (!c || n < 0 || (n > 0 && (!ijk || !var || !out))) return OSB_E_INVALID;
    for (int p = 0; p < n; ++p) {
        for (int d = 0; d < 3; ++d)
            if (ijk[3 * p + d] < 0 || ijk[3 * p + d] >= c->g.n[d]) return set_err(c, OSB_E_INVALID, "probe outside the grid");
        if (var[p] < 0 || var[p] >= c->nv) return set_err(c, OSB_E_INVALID, "probe variable out of range");
    }
    // one 8-byte copy per probe from the current state (monitor frequency is low)
    for (int p = 0; p < n; ++p) {
        const long x = (long)var[p] * c->g.Ptot + gidx(c->g, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2]);
        HIP_TRY(c, hipMemcpyAsync(out + p, c->q_cur + x, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_nan_check(osb_ctx* c) {
    if (!c) return OSB_E_INVALID;
    HIP_TRY(c, hipMemsetAsync(c->nan_cnt, 0, sizeof(unsigned int), c->stream));
    const long nin = interior_points(c->g);
    k_nan_count<<<nblocks(nin), BLOCK, 0, c->stream>>>(c->q_cur, c->nan_cnt, c->g);
    unsigned int cnt = 0;
    HIP_TRY(c, hipMemcpyAsync(&cnt, c->nan_cnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (cnt) return set_err(c, OSB_E_NONFINITE, "non-finite rho at " + std::to_string(cnt) + " points");
    return check_launch(c);
}

int osb_profile_enable(osb_ctx* c, int enable) {
    if (!c) return OSB_E_INVALID;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    prof_resolve(c);
    for (int i = 0; i < PF_COUNT; ++i) { c->prof_ms[i] = 0.0; c->prof_n[i] = 0; }
    c->prof_on = enable != 0;
    return OSB_OK;
}

int osb_profile_read(osb_ctx* c, int idx, char* name, int name_len, double* total_ms, int64_t* launches) {
    if (!c || idx < 0 || idx >= PF_COUNT) return OSB_E_INVALID;
    prof_resolve(c);
    if (name && name_len > 0) {
        std::strncpy(name, kFamilyName[idx], (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = c->prof_ms[idx];
    if (launches) *launches = c->prof_n[idx];
    return OSB_OK;
}

// Device-to-device state transfer (interior layout, nv x n2 x n1 x n0 FP64 in
// device memory, e.g. a torch tensor's data_ptr()).  The caller makes sure the
// source is ready; the call returns when the copy is done.
int osb_set_state_device(osb_ctx* c, const double* dev, int layout) {
    if (!c || !dev) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR) return set_err(c, OSB_E_UNSUPPORTED, "only the interior layout is accepted");
    const long npts = interior_points(c->g);
    HIP_TRY(c, hipMemsetAsync(c->qa, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->du, 0, sizeof(double) * c->nv * c->g.Ptot, c->stream));
    k_scatter<<<nblocks(npts * c->nv), BLOCK, 0, c->stream>>>(dev, c->qa, c->g, c->nv);
    c->q_cur = c->qa;
    c->q_prev = c->qb;
    c->have_prev = false;
    c->next_stage = 0;
    apply_bcs(c, c->q_cur);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_get_state_device(osb_ctx* c, double* dev, int layout) {
    if (!c || !dev) return OSB_E_INVALID;
    if (layout != OSB_LAYOUT_INTERIOR) return set_err(c, OSB_E_UNSUPPORTED, "only the interior layout is returned");
    const long npts = interior_points(c->g);
    k_gather<<<nblocks(npts * c->nv), BLOCK, 0, c->stream>>>(c->q_cur, dev, c->g, c->nv);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return check_launch(c);
}

int osb_state_device_ptr(osb_ctx* c, int var, void** out) {
    if (!c || !out || var < 0 || var >= c->nv) return OSB_E_INVALID;
    *out = (void*)(c->q_cur + (long)var * c->g.Ptot);
    return OSB_OK;
}


// The context's kernel path and launch shapes as JSON (no counterpart in the
// reference; the analogue of the generated code's choice of kernels).
int osb_describe(osb_ctx* c, char* buf, int len) {
    if (!c || !buf || len < 1) return OSB_E_INVALID;
    const char* path = use_fused(c) ? (c->mixed ? "central_mixed" : "central_fused")
                                    : (use_shock_stage(c) ? (c->shock_fused ? "face_flux+face_resid" : "face_flux+resid_rk")
                                                          : "unfused");
    std::string s = std::string("{\"path\": \"") + path + "\"";
    if (use_shock_stage(c) && c->shock_fused && c->g.ndim >= 2) {
        char tmp[160];
        std::snprintf(tmp, sizeof(tmp), ", \"resid_w\": %d, \"resid_w_choice\": \"%s\", \"resid_w_us\": [", c->resid_w,
                      c->resid_forced ? "forced" : (c->tile_tuned ? "autotuned" : "default"));
        s += tmp;
        for (size_t i = 0; i < c->resid_w_us.size(); ++i) {
            std::snprintf(tmp, sizeof(tmp), "%s%.2f", i ? ", " : "", (double)c->resid_w_us[i]);
            s += tmp;
        }
        s += "]";
    }
    if (use_fused(c) && !c->mixed) {
        const CfTile t = CF_TILES[c->fused_tile];
        char tmp[160];
        std::snprintf(tmp, sizeof(tmp), ", \"tile\": [%d, %d, %d], \"tile_choice\": \"%s\", \"tile_us\": [", t.tx, t.ty, t.tz,
                      c->tile_forced ? "forced" : (c->tile_tuned ? "autotuned" : "default"));
        s += tmp;
        for (size_t i = 0; i < c->tile_us.size(); ++i) {
            std::snprintf(tmp, sizeof(tmp), "%s%.2f", i ? ", " : "", (double)c->tile_us[i]);
            s += tmp;
        }
        s += "]";
    }
    s += ", \"graphs\": " + std::string(c->use_graph ? "true" : "false") + "}";
    std::strncpy(buf, s.c_str(), (size_t)len - 1);
    buf[len - 1] = '\0';
    return (int)s.size() < len ? OSB_OK : OSB_E_INVALID;
}

// Allocation-free stepping of a caller-owned block (no counterpart in the
// reference).  q: padded state (nv x padded points, halos filled), advanced in
// place; du: the RK-LS register (classic RK3: the saved state), same layout,
// zero before a run's first step.  The context's other buffers are carved from
// the library's scratch array; the steps are launched eagerly on `stream`.
int osb_step_noalloc(const osb_problem_desc* desc, double* q, double* du, int nsteps, void* stream) {
    std::string why;
    int rc = validate(desc, why);
    if (rc) return set_err(nullptr, rc, why);
    if (!q || !du || nsteps < 0) return set_err(nullptr, OSB_E_INVALID, "osb_step_noalloc: null buffer or negative step count");
    double* base = nullptr;
    rc = noalloc_scratch(&base);
    if (rc) return rc;
    const AllocPlan ap = plan_alloc(desc);
    const size_t need = ap.q + ap.R + ap.F + ap.kappa + ap.cr + ap.gr + OSB_SCAL_SIZE + (size_t)PART_BLOCKS * 4;
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return set_err(nullptr, OSB_E_UNSUPPORTED,
                       "osb_step_noalloc: this block needs " + std::to_string(need) + " doubles of scratch; the library holds " +
                           std::to_string(NOALLOC_SCRATCH_DOUBLES) + " (a smaller block, or a context from osb_ctx_create)");
    osb_ctx c;
    init_ctx(&c, desc);
    c.use_graph = false;
    c.stream = (hipStream_t)stream;
    c.qa = q;
    c.du = du;
    double* p = base;
    auto carve = [&](double** dst, size_t n) { *dst = n ? p : nullptr; p += n; };
    carve(&c.qb, ap.q);
    carve(&c.R, ap.R);
    carve(&c.F, ap.F);
    carve(&c.kappa, ap.kappa);
    carve(&c.cr, ap.cr);
    carve(&c.gr, ap.gr);
    carve(&c.scal, OSB_SCAL_SIZE);
    c.part_blocks = PART_BLOCKS;
    carve(&c.part, (size_t)PART_BLOCKS * 4);
    if (c.shock && desc->flux_type == OSB_FLUX_GLF) c.c.glf = c.scal + OSB_SCAL_GLF;
    // every buffer of a new context starts zeroed; the scratch keeps the last call's contents
    k_noalloc_clear<<<std::min<long>(nblocks((long)need), 2048), BLOCK, 0, c.stream>>>(base, (long)need);
    k_set_dt<<<1, 64, 0, c.stream>>>(c.scal, desc->dt);
    c.q_cur = c.qa;
    c.q_prev = c.qb;
    for (int i = 0; i < nsteps; ++i) do_step(&c);
    const long n = (long)c.nv * c.g.Ptot;
    if (c.q_cur != q) k_copy_all<<<std::min<long>(nblocks(n), 4096), BLOCK, 0, c.stream>>>(q, c.q_cur, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(nullptr, OSB_E_DEVICE, std::string("osb_step_noalloc launch: ") + hipGetErrorString(e));
    return OSB_OK;
}

}  // extern "C"
