// osb_batch.hip -- batched ensembles (osb_batch_*) of the C ABI.
//
// A batch owns one context per instance (any sizes) and one stream.
// Instances that use the shock-capturing path without viscous terms are
// grouped by (ndim, scheme, precision, time scheme, stages).  One group step
// is a fixed sequence of batched launches, each covering every instance of
// the group:
//   [CFL: zero, wave speed, dt]  [RK3: save q]
//   per stage: face fluxes per direction, flux difference + RK update, BCs per direction
// That sequence is captured once per parity into a hipGraph and replayed.
// 3-D central instances (FP64, RK-LS, every face periodic) form groups by
// (split form, viscosity, stages): per stage one launch of the fused stage
// kernel over the tiles of every member, which also writes the wrapped halo
// images.  Other instances (3-D shock capturing with viscous terms, mixed
// precision or boundary faces on the central path, the WENO filter) are
// stepped one after another on the same stream.
//
// Launch shapes: the batched kernels are persistent over virtual blocks; the
// grid of each face-flux family is picked by timing a few candidates on the
// batch's own data before the first step (osb_batch_autotune).

#include "osb_runtime.hpp"
#include "osb_batch.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace osb;

namespace {

enum BFam { BF_FACE0, BF_FACE1, BF_FACE2, BF_RESID, BF_BC0, BF_BC1, BF_BC2, BF_WAVE, BF_SAVE, BF_COUNT };
const char* const kBFamName[BF_COUNT] = {"face_flux_x", "face_flux_y", "face_flux_z", "resid_rk", "bc_x", "bc_y", "bc_z",
                                         "wavespeed", "save_state"};

enum GroupKind { GK_SHOCK = 0, GK_CENTRAL = 1 };

struct Group {
    int kind = GK_SHOCK;                   // shock-capturing stages, or the fused 3-D central stage
    int conv = 0, visc = 0;                // GK_CENTRAL: split form (CfConv) and viscosity model
    int ndim = 0, sch = 0, mixed = 0, time = 0, stages = 0;
    bool any_cfl = false;
    bool any_glf = false;
    bool fused = false;                    // last direction's faces fused with the update (kb_face_resid)
    std::vector<int> members;
    std::vector<int> offs;                 // BF_COUNT x (n + 1) virtual-block offsets
    BItem* d_items = nullptr;
    int* d_offs = nullptr;
    int cap[BF_COUNT];                     // persistent grid size; 0 = one block per virtual block
    float tuned_us[BF_COUNT];
    int parity = 0;
    int cf_tile = 0;                       // GK_CENTRAL: tile shape, an index into CF_TILES (timed by autotune)
    float tile_us[CF_NTILES];
    int resid_w = FUSED_FW;                // GK_SHOCK, fused: lines per face-residual tile (RESID_WS, timed by autotune)
    float resid_w_us[RESID_NW];
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    Group() {
        for (int f = 0; f < BF_COUNT; ++f) { cap[f] = 0; tuned_us[f] = 0.f; }
        for (int t = 0; t < CF_NTILES; ++t) tile_us[t] = 0.f;
        for (int t = 0; t < RESID_NW; ++t) resid_w_us[t] = 0.f;
    }
    int n() const { return (int)members.size(); }
    const int* dof(int f) const { return d_offs + (size_t)f * (members.size() + 1); }
    int total(int f) const { return offs[(size_t)f * (members.size() + 1) + members.size()]; }
};

// GK_SHOCK, GK_CENTRAL, or -1: stepped alone
int group_kind(const osb_ctx* c) {
    if (c->shock && !c->central && !c->viscous && c->d.shock_filter == OSB_FILTER_NONE && c->d.flux_type <= OSB_FLUX_GLF)
        return GK_SHOCK;
    bool closure = false;
    for (int d = 0; d < 3; ++d) closure = closure || c->c.carp[d] != 0;
    if (c->g.ndim == 3 && c->central && !c->shock && !c->mixed && !closure && c->bc_periodic && c->bc_in_stage &&
        c->d.shock_filter == OSB_FILTER_NONE && c->d.time_scheme == OSB_TIME_RKLS3)
        return GK_CENTRAL;
    return -1;
}

int cf_conv(const osb_ctx* c) {
    return c->d.conv_scheme == OSB_CONV_CENTRAL_SKEW ? CF_SKEW : (c->d.conv_scheme == OSB_CONV_CENTRAL_KGP ? CF_KGP : CF_FEIEREISEN);
}

// same group: kind and the signature of that kind
bool same_signature(const Group& gr, const osb_ctx* c, int kind, int stages) {
    if (gr.kind != kind || gr.ndim != c->g.ndim || gr.time != c->d.time_scheme || gr.stages != stages || gr.mixed != (int)c->mixed)
        return false;
    if (kind == GK_CENTRAL) return gr.conv == cf_conv(c) && gr.visc == c->d.viscous;
    return gr.sch == c->sch && gr.fused == c->shock_fused;
}

void set_signature(Group& gr, const osb_ctx* c, int kind, int stages) {
    gr.kind = kind;
    gr.ndim = c->g.ndim; gr.sch = c->sch; gr.mixed = c->mixed; gr.time = c->d.time_scheme; gr.stages = stages;
    gr.fused = c->shock_fused;
    gr.conv = cf_conv(c);
    gr.visc = c->d.viscous;
}

inline unsigned vblocks(long work) { return (unsigned)((work + BLOCK - 1) / BLOCK); }

}  // namespace

struct osb_batch {
    std::vector<osb_ctx*> ctx;
    std::vector<Group> groups;
    std::vector<int> solo;
    hipStream_t stream = nullptr;
    bool use_graph = true;
    bool items_dirty = true;
    bool tuned = false;
    bool autotune = true;
    std::string err;
};

namespace {

int berr(osb_batch* b, int code, const std::string& m) {
    if (b) b->err = m; else g_osb_create_error = m;
    return code;
}

#define B_TRY(b, expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError(); /* clear the runtime's sticky error */                   \
            return berr((b), OSB_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));  \
        }                                                                                     \
    } while (0)

// work of family f for one instance, in threads (BLOCK threads = one virtual
// block); cf_tile: the central group's tile shape; resid_w: the fused
// face-residual tiles' width
long family_work(const osb_ctx* c, int f, int cf_tile = 0, int resid_w = FUSED_FW) {
    const Geom& g = c->g;
    if (group_kind(c) == GK_CENTRAL) {
        if (f == BF_RESID) return kb_central_tiles(g, CF_TILES[cf_tile]) * BLOCK;   // one virtual block per tile
        if (f == BF_WAVE) return (long)g.n[0] * g.n[1] * g.n[2];
        return 0;
    }
    switch (f) {
        case BF_FACE0: return face_flux_work<3, 0>(g);
        case BF_FACE1: return g.ndim > 1 ? face_flux_work<3, 1>(g) : 0;
        case BF_FACE2: return g.ndim > 2 ? face_flux_work<3, 2>(g) : 0;
        case BF_BC0: return bc_work<0>(g);
        case BF_BC1: return g.ndim > 1 ? bc_work<1>(g) : 0;
        case BF_BC2: return g.ndim > 2 ? bc_work<2>(g) : 0;
        case BF_RESID:
            if (c->shock_fused) {                            // fused last-direction tiles
                const long t = g.ndim == 1 ? fused_tiles<1>(g)
                                           : (resid_w == 16 ? (g.ndim == 2 ? fused_tiles<2, 16>(g) : fused_tiles<3, 16>(g))
                                                            : (g.ndim == 2 ? fused_tiles<2>(g) : fused_tiles<3>(g)));
                return t * BLOCK;
            }
            return (long)g.n[0] * g.n[1] * g.n[2];
        default: return (long)g.n[0] * g.n[1] * g.n[2];     // wavespeed, save_state
    }
}

void free_group(Group& gr) {
    for (auto& e : gr.exec)
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    if (gr.d_items) (void)hipFree(gr.d_items);
    if (gr.d_offs) (void)hipFree(gr.d_offs);
    gr.d_items = nullptr;
    gr.d_offs = nullptr;
}

int build_groups(osb_batch* b) {
    for (int i = 0; i < (int)b->ctx.size(); ++i) {
        const osb_ctx* c = b->ctx[i];
        const int kind = group_kind(c);
        if (kind < 0) { b->solo.push_back(i); continue; }
        const int stages = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
        Group* gp = nullptr;
        for (auto& gr : b->groups)
            if (same_signature(gr, c, kind, stages)) { gp = &gr; break; }
        if (!gp) {
            b->groups.emplace_back();
            gp = &b->groups.back();
            set_signature(*gp, c, kind, stages);
        }
        gp->members.push_back(i);
        gp->any_cfl = gp->any_cfl || c->d.cfl > 0;
        gp->any_glf = gp->any_glf || c->c.glf != nullptr;
    }
    for (auto& gr : b->groups) {
        const int n = gr.n();
        gr.offs.assign((size_t)BF_COUNT * (n + 1), 0);
        for (int f = 0; f < BF_COUNT; ++f) {
            int acc = 0;
            for (int m = 0; m < n; ++m) {
                gr.offs[(size_t)f * (n + 1) + m] = acc;
                const long vb = vblocks(family_work(b->ctx[gr.members[m]], f, gr.cf_tile, gr.resid_w));
                if (acc + vb > 0x7fffffffL) return berr(b, OSB_E_INVALID, "batch too large for 32-bit block offsets");
                acc += (int)vb;
            }
            gr.offs[(size_t)f * (n + 1) + n] = acc;
        }
        B_TRY(b, hipMalloc((void**)&gr.d_items, sizeof(BItem) * n));
        B_TRY(b, hipMalloc((void**)&gr.d_offs, sizeof(int) * gr.offs.size()));
        B_TRY(b, hipMemcpy(gr.d_offs, gr.offs.data(), sizeof(int) * gr.offs.size(), hipMemcpyHostToDevice));
    }
    return OSB_OK;
}

// item table from the contexts' current buffers; q[parity] = current state
int upload_items(osb_batch* b) {
    for (auto& gr : b->groups) {
        std::vector<BItem> it(gr.n());
        for (int m = 0; m < gr.n(); ++m) {
            const osb_ctx* c = b->ctx[gr.members[m]];
            BItem& x = it[m];
            std::memset(&x, 0, sizeof(x));
            x.q[gr.parity] = c->q_cur;
            x.q[gr.parity ^ 1] = c->q_cur == c->qa ? c->qb : c->qa;
            x.du = c->du;
            x.F = c->F;
            x.scal = c->scal;
            x.g = c->g;
            x.c = c->c;
            kb_central_orders(c->g, x.ord);
            for (int d = 0; d < 3; ++d) x.bc[d] = bc_sides(c->d, d);
            x.cfl = c->d.cfl;
            if (c->d.time_scheme == OSB_TIME_RKLS3) {
                for (int s = 0; s < c->rk_stages; ++s) { x.rkA[s] = c->rkA[s]; x.rkB[s] = c->rkB[s]; }
            } else {
                // rk_sbli.py: rknew (q = qold + dt*rknew*R), rkold (qold += dt*rkold*R)
                const double rknew[3] = {2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0};
                const double rkold[3] = {1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0};
                for (int s = 0; s < 3; ++s) { x.rkA[s] = rknew[s]; x.rkB[s] = rkold[s]; }
            }
        }
        B_TRY(b, hipMemcpyAsync(gr.d_items, it.data(), sizeof(BItem) * it.size(), hipMemcpyHostToDevice, b->stream));
    }
    B_TRY(b, hipStreamSynchronize(b->stream));
    b->items_dirty = false;
    return OSB_OK;
}

unsigned grid_of(const Group& gr, int f) {
    const int tot = gr.total(f);
    const int cap = gr.cap[f];
    return (unsigned)std::max(1, cap > 0 ? std::min(tot, cap) : tot);
}

template <int NDIM, int DIR, typename T>
void launch_face(hipStream_t st, const Group& gr, int parity, unsigned grid) {
    const int f = BF_FACE0 + DIR;
    switch (gr.sch) {
        case SCH_WENOZ: kb_face_flux<NDIM, DIR, SCH_WENOZ, T><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(f), gr.n(), parity); break;
        case SCH_TENO5: kb_face_flux<NDIM, DIR, SCH_TENO5, T><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(f), gr.n(), parity); break;
        default: kb_face_flux<NDIM, DIR, SCH_TENO6, T><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(f), gr.n(), parity); break;
    }
}

template <int NDIM, int DIR>
void launch_face_p(hipStream_t st, const Group& gr, int parity, unsigned grid) {
    if (gr.mixed) launch_face<NDIM, DIR, float>(st, gr, parity, grid);
    else launch_face<NDIM, DIR, double>(st, gr, parity, grid);
}

template <int NDIM>
void launch_face_dir(hipStream_t st, const Group& gr, int dir, int parity, unsigned grid) {
    if (dir == 0) launch_face_p<NDIM, 0>(st, gr, parity, grid);
    else if (dir == 1) launch_face_p<NDIM, (NDIM > 1 ? 1 : 0)>(st, gr, parity, grid);
    else launch_face_p<NDIM, (NDIM > 2 ? 2 : 0)>(st, gr, parity, grid);
}

template <int NDIM, int TIME, typename T, int W>
void launch_face_resid_w(hipStream_t st, const Group& gr, int parity, int stage) {
    const unsigned grid = grid_of(gr, BF_RESID);
    switch (gr.sch) {
        case SCH_WENOZ:
            kb_face_resid<NDIM, SCH_WENOZ, T, TIME, W><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage);
            break;
        case SCH_TENO5:
            kb_face_resid<NDIM, SCH_TENO5, T, TIME, W><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage);
            break;
        default:
            kb_face_resid<NDIM, SCH_TENO6, T, TIME, W><<<grid, BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage);
            break;
    }
}
template <int NDIM, int TIME, typename T>
void launch_face_resid(hipStream_t st, const Group& gr, int parity, int stage) {
    static_assert(RESID_NW == 2 && RESID_WS[0] == 32 && RESID_WS[1] == 16, "one case per width");
    if (NDIM > 1 && gr.resid_w == 16) launch_face_resid_w<NDIM, TIME, T, 16>(st, gr, parity, stage);
    else launch_face_resid_w<NDIM, TIME, T, FUSED_FW>(st, gr, parity, stage);
}

template <int NDIM, int TIME>
void launch_face_resid_p(hipStream_t st, const Group& gr, int parity, int stage) {
    if (gr.mixed) launch_face_resid<NDIM, TIME, float>(st, gr, parity, stage);
    else launch_face_resid<NDIM, TIME, double>(st, gr, parity, stage);
}

template <int NDIM>
void launch_bcs(hipStream_t st, const Group& gr, int parity) {
    kb_bc<0><<<grid_of(gr, BF_BC0), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_BC0), gr.n(), parity);
    if (NDIM > 1) kb_bc<1><<<grid_of(gr, BF_BC1), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_BC1), gr.n(), parity);
    if (NDIM > 2) kb_bc<2><<<grid_of(gr, BF_BC2), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_BC2), gr.n(), parity);
}

template <int TX, int TY, int TZ, int CONV>
void launch_central_t(hipStream_t st, const Group& gr, int parity, int stage) {
    const unsigned grid = (unsigned)std::max(1, gr.total(BF_RESID));     // one workgroup per tile
    const dim3 block(TX * TY * TZ);
    switch (gr.visc) {
        case OSB_VISC_NONE: kb_central_fused<TX, TY, TZ, CONV, 0><<<grid, block, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage); break;
        case OSB_VISC_CONSTANT: kb_central_fused<TX, TY, TZ, CONV, 1><<<grid, block, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage); break;
        default: kb_central_fused<TX, TY, TZ, CONV, 2><<<grid, block, 0, st>>>(gr.d_items, gr.dof(BF_RESID), gr.n(), parity, stage); break;
    }
}
// the tile shapes of CF_TILES, in its order
template <int CONV>
void launch_central_v(hipStream_t st, const Group& gr, int parity, int stage) {
    static_assert(CF_NTILES == 3, "one case per tile shape");
    switch (gr.cf_tile) {
        case 1: launch_central_t<16, 8, 4, CONV>(st, gr, parity, stage); break;
        case 2: launch_central_t<8, 8, 8, CONV>(st, gr, parity, stage); break;
        default: launch_central_t<16, 4, 4, CONV>(st, gr, parity, stage); break;
    }
}

// one time step of a 3-D central group: [CFL] then per stage the fused stage
// kernel over every member's tiles (halo images included: no BC launches)
void enqueue_central_step(hipStream_t st, const Group& gr, int parity) {
    const int n = gr.n();
    if (gr.any_cfl) {
        kb_zero_wavespeed<<<(n + 63) / 64, 64, 0, st>>>(gr.d_items, n);
        kb_wavespeed<3><<<grid_of(gr, BF_WAVE), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_WAVE), n, parity);
        kb_dt_finalize<<<(n + 63) / 64, 64, 0, st>>>(gr.d_items, n);
    }
    for (int s = 0; s < gr.stages; ++s) {
        if (gr.conv == CF_SKEW) launch_central_v<CF_SKEW>(st, gr, parity, s);
        else if (gr.conv == CF_KGP) launch_central_v<CF_KGP>(st, gr, parity, s);
        else launch_central_v<CF_FEIEREISEN>(st, gr, parity, s);
        parity ^= 1;
    }
}

// one time step of a group, starting at `parity`
template <int NDIM>
void enqueue_group_step_nd(hipStream_t st, const Group& gr, int parity) {
    const int n = gr.n();
    if (gr.any_cfl) {
        kb_zero_wavespeed<<<(n + 63) / 64, 64, 0, st>>>(gr.d_items, n);
        kb_wavespeed<NDIM><<<grid_of(gr, BF_WAVE), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_WAVE), n, parity);
        kb_dt_finalize<<<(n + 63) / 64, 64, 0, st>>>(gr.d_items, n);
    }
    if (gr.time == OSB_TIME_RK3)
        kb_save_state<<<grid_of(gr, BF_SAVE), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_SAVE), n, parity);
    for (int s = 0; s < gr.stages; ++s) {
        if (gr.any_glf) {
            kb_glf_zero<<<(n + 63) / 64, 64, 0, st>>>(gr.d_items, n);
            kb_glf<NDIM><<<grid_of(gr, BF_WAVE), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_WAVE), n, parity);
        }
        if (gr.fused) {
            for (int d = 0; d + 1 < NDIM; ++d) launch_face_dir<NDIM>(st, gr, d, parity, grid_of(gr, BF_FACE0 + d));
            if (gr.time == OSB_TIME_RK3) launch_face_resid_p<NDIM, 2>(st, gr, parity, s);
            else launch_face_resid_p<NDIM, 1>(st, gr, parity, s);
        } else {
            for (int d = 0; d < NDIM; ++d) launch_face_dir<NDIM>(st, gr, d, parity, grid_of(gr, BF_FACE0 + d));
            if (gr.time == OSB_TIME_RK3)
                kb_resid_rk<NDIM, 2><<<grid_of(gr, BF_RESID), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_RESID), n, parity, s);
            else
                kb_resid_rk<NDIM, 1><<<grid_of(gr, BF_RESID), BLOCK, 0, st>>>(gr.d_items, gr.dof(BF_RESID), n, parity, s);
        }
        launch_bcs<NDIM>(st, gr, parity ^ 1);
        parity ^= 1;
    }
}

void enqueue_group_step(hipStream_t st, const Group& gr, int parity) {
    if (gr.kind == GK_CENTRAL) {
        enqueue_central_step(st, gr, parity);
        return;
    }
    switch (gr.ndim) {
        case 1: enqueue_group_step_nd<1>(st, gr, parity); break;
        case 2: enqueue_group_step_nd<2>(st, gr, parity); break;
        default: enqueue_group_step_nd<3>(st, gr, parity); break;
    }
}

int capture_graphs(osb_batch* b, Group& gr) {
    for (int p = 0; p < 2; ++p) {
        if (gr.exec[p]) { (void)hipGraphExecDestroy(gr.exec[p]); gr.exec[p] = nullptr; }
        hipGraph_t graph = nullptr;
        B_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_group_step(b->stream, gr, p);
        hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (e != hipSuccess) return berr(b, OSB_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&gr.exec[p], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return berr(b, OSB_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return OSB_OK;
}

// The central group's tile shape t: its stage launch covers every member's
// tiles of that shape, so the stage family's virtual-block offsets follow it.
int set_central_tile(osb_batch* b, Group& gr, int t) {
    gr.cf_tile = t;
    const int n = gr.n();
    int acc = 0;
    for (int m = 0; m < n; ++m) {
        gr.offs[(size_t)BF_RESID * (n + 1) + m] = acc;
        acc += (int)vblocks(family_work(b->ctx[gr.members[m]], BF_RESID, t));
    }
    gr.offs[(size_t)BF_RESID * (n + 1) + n] = acc;
    B_TRY(b, hipMemcpyAsync(gr.d_offs + (size_t)BF_RESID * (n + 1), gr.offs.data() + (size_t)BF_RESID * (n + 1),
                            sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
    return OSB_OK;
}

// The fused shock group's face-residual width w: its launch covers every
// member's tiles of that width, so that family's offsets follow it.
int set_resid_w(osb_batch* b, Group& gr, int w) {
    gr.resid_w = w;
    const int n = gr.n();
    int acc = 0;
    for (int m = 0; m < n; ++m) {
        gr.offs[(size_t)BF_RESID * (n + 1) + m] = acc;
        acc += (int)vblocks(family_work(b->ctx[gr.members[m]], BF_RESID, gr.cf_tile, w));
    }
    gr.offs[(size_t)BF_RESID * (n + 1) + n] = acc;
    B_TRY(b, hipMemcpyAsync(gr.d_offs + (size_t)BF_RESID * (n + 1), gr.offs.data() + (size_t)BF_RESID * (n + 1),
                            sizeof(int) * (n + 1), hipMemcpyHostToDevice, b->stream));
    return OSB_OK;
}

// Launch shapes, timed on the batch's own state before its first step:
// * shock groups: the persistent grid of each face-flux family (face fluxes
//   only read q and overwrite F);
// * 3-D central groups: the tile shape of the fused stage kernel (CF_TILES);
// * 2-D/3-D shock groups with the fused last direction: the width of the
//   face-residual tiles (RESID_WS), timed over a whole stage (the face-flux
//   launches at their chosen grids, then the face-residual launch), so that
//   the residual kernel reads the fluxes as a real stage leaves them.
// Each candidate is timed in TUNE_ROUNDS rounds that alternate between the
// candidates, and keeps its best round: a single pass of 3 launches per
// candidate drifted by more than the widths' difference from box to box.
// The trial launches that update write the members' next buffer, which the
// first real stage overwrites, and their RK register, which is cleared
// afterwards (every RK-LS scheme's first stage multiplies it by zero, so only
// its finiteness matters; classic RK3 saves the state into it first).
int autotune(osb_batch* b) {
    hipEvent_t e0, e1;
    B_TRY(b, hipEventCreate(&e0));
    B_TRY(b, hipEventCreate(&e1));
    // average microseconds of 3 launches of run() after one untimed launch
    auto time_us = [&](auto&& run, float& us) -> int {
        run();
        B_TRY(b, hipEventRecord(e0, b->stream));
        for (int k = 0; k < 3; ++k) run();
        B_TRY(b, hipEventRecord(e1, b->stream));
        B_TRY(b, hipEventSynchronize(e1));
        float ms = 0.f;
        B_TRY(b, hipEventElapsedTime(&ms, e0, e1));
        us = ms * 1000.f / 3.f;
        return OSB_OK;
    };
    auto clear_du = [&](const Group& gr) -> int {
        for (int m : gr.members) {
            const osb_ctx* c = b->ctx[m];
            B_TRY(b, hipMemsetAsync(c->du, 0, sizeof(double) * c->nv * c->g.Ptot, b->stream));
        }
        return OSB_OK;
    };
    for (auto& gr : b->groups) {
        if (gr.kind != GK_CENTRAL) continue;
        // OSB_FUSED_TILE forces the tile shape of every context, so of the group too
        const osb_ctx* c0 = b->ctx[gr.members[0]];
        if (c0->tile_forced) {
            int r = set_central_tile(b, gr, c0->fused_tile);
            if (r) return r;
            continue;
        }
        auto run = [&]() {
            if (gr.conv == CF_SKEW) launch_central_v<CF_SKEW>(b->stream, gr, gr.parity, 0);
            else if (gr.conv == CF_KGP) launch_central_v<CF_KGP>(b->stream, gr, gr.parity, 0);
            else launch_central_v<CF_FEIEREISEN>(b->stream, gr, gr.parity, 0);
        };
        for (int t = 0; t < CF_NTILES; ++t) gr.tile_us[t] = 1e30f;
        for (int round = 0; round < TUNE_ROUNDS; ++round)
            for (int t = 0; t < CF_NTILES; ++t) {
                int r = set_central_tile(b, gr, t);
                if (r) return r;
                float us = 0.f;
                r = time_us(run, us);
                if (r) return r;
                gr.tile_us[t] = std::min(gr.tile_us[t], us);
            }
        int best_t = 0;
        for (int t = 1; t < CF_NTILES; ++t)
            if (gr.tile_us[t] < gr.tile_us[best_t]) best_t = t;
        int r = set_central_tile(b, gr, best_t);
        if (r) return r;
        gr.tuned_us[BF_RESID] = gr.tile_us[best_t];
        r = clear_du(gr);
        if (r) return r;
    }
    for (auto& gr : b->groups) {
        if (gr.kind != GK_SHOCK) continue;                            // the fused stage kernels change du: not timed
        for (int d = 0; d < gr.ndim - (gr.fused ? 1 : 0); ++d) {     // the fused family changes du: not timed
            const int f = BF_FACE0 + d;
            const int tot = gr.total(f);
            const int cands[5] = {0, 8192, 4096, 2048, 1024};
            float best_us[5];
            for (int ci = 0; ci < 5; ++ci) best_us[ci] = 1e30f;
            for (int round = 0; round < TUNE_ROUNDS; ++round)
                for (int ci = 0; ci < 5; ++ci) {
                    const int cap = cands[ci];
                    if (cap > 0 && cap >= tot) continue;
                    const unsigned grid = (unsigned)(cap > 0 ? cap : std::max(tot, 1));
                    auto run = [&]() {
                        switch (gr.ndim) {
                            case 1: launch_face_dir<1>(b->stream, gr, d, gr.parity, grid); break;
                            case 2: launch_face_dir<2>(b->stream, gr, d, gr.parity, grid); break;
                            default: launch_face_dir<3>(b->stream, gr, d, gr.parity, grid); break;
                        }
                    };
                    float us = 0.f;
                    int r = time_us(run, us);
                    if (r) return r;
                    best_us[ci] = std::min(best_us[ci], us);
                }
            int best = 0;
            for (int ci = 1; ci < 5; ++ci)
                if (best_us[ci] < best_us[best]) best = ci;
            gr.cap[f] = cands[best];
            gr.tuned_us[f] = best_us[best];
        }
    }
    for (auto& gr : b->groups) {
        if (gr.kind != GK_SHOCK || !gr.fused || gr.ndim < 2) continue;
        const osb_ctx* c0 = b->ctx[gr.members[0]];
        if (c0->resid_forced) {                                       // OSB_RESID_W, as for the contexts
            int r = set_resid_w(b, gr, c0->resid_w);
            if (r) return r;
            continue;
        }
        // one stage's face fluxes (at their tuned grids) and its face-residual launch
        auto run = [&]() {
            for (int d = 0; d < gr.ndim - 1; ++d) {
                const unsigned grid = grid_of(gr, BF_FACE0 + d);
                if (gr.ndim == 2) launch_face_dir<2>(b->stream, gr, d, gr.parity, grid);
                else launch_face_dir<3>(b->stream, gr, d, gr.parity, grid);
            }
            if (gr.ndim == 2) {
                if (gr.time == OSB_TIME_RK3) launch_face_resid_p<2, 2>(b->stream, gr, gr.parity, 0);
                else launch_face_resid_p<2, 1>(b->stream, gr, gr.parity, 0);
            } else {
                if (gr.time == OSB_TIME_RK3) launch_face_resid_p<3, 2>(b->stream, gr, gr.parity, 0);
                else launch_face_resid_p<3, 1>(b->stream, gr, gr.parity, 0);
            }
        };
        for (int t = 0; t < RESID_NW; ++t) gr.resid_w_us[t] = 1e30f;
        for (int round = 0; round < TUNE_ROUNDS; ++round)
            for (int t = 0; t < RESID_NW; ++t) {
                int r = set_resid_w(b, gr, RESID_WS[t]);
                if (r) return r;
                float us = 0.f;
                r = time_us(run, us);
                if (r) return r;
                gr.resid_w_us[t] = std::min(gr.resid_w_us[t], us);
            }
        int best_t = 0;
        for (int t = 1; t < RESID_NW; ++t)
            if (gr.resid_w_us[t] < gr.resid_w_us[best_t]) best_t = t;
        int r = set_resid_w(b, gr, RESID_WS[best_t]);
        if (r) return r;
        gr.tuned_us[BF_RESID] = gr.resid_w_us[best_t];
        r = clear_du(gr);
        if (r) return r;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B_TRY(b, hipGetLastError());
    b->tuned = true;
    return OSB_OK;
}

int prepare(osb_batch* b) {
    if (b->items_dirty) {
        int r = upload_items(b);
        if (r) return r;
    }
    if (!b->tuned && b->autotune) {
        int r = autotune(b);
        if (r) return r;
    }
    if (b->use_graph) {
        for (auto& gr : b->groups)
            if (!gr.exec[0]) {
                int r = capture_graphs(b, gr);
                if (r) return r;
            }
    }
    return OSB_OK;
}

void finish_members(osb_batch* b, Group& gr, int nsteps) {
    // host view of each member after nsteps: buffers, counters and time
    for (int m = 0; m < gr.n(); ++m) {
        osb_ctx* c = b->ctx[gr.members[m]];
        double* cur_in = c->q_cur;
        double* other = c->q_cur == c->qa ? c->qb : c->qa;
        const int flips = (int)(((long)nsteps * gr.stages) & 1);
        if (flips) std::swap(cur_in, other);
        c->q_cur = cur_in;
        c->q_prev = other;            // input of the last stage is the other buffer
        c->have_prev = true;
        c->iter += nsteps;
        if (!(c->d.cfl > 0))
            for (int s = 0; s < nsteps; ++s) { c->time += c->d.dt; c->last_dt = c->d.dt; }
    }
}

}  // namespace

extern "C" {

const char* osb_batch_last_error(const osb_batch* b) { return b ? b->err.c_str() : g_osb_create_error.c_str(); }

int osb_batch_create(const osb_problem_desc* descs, int n, osb_batch** out) {
    if (!out || !descs || n < 1) return berr(nullptr, OSB_E_INVALID, "bad batch arguments");
    *out = nullptr;
    osb_batch* b = new osb_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return berr(nullptr, OSB_E_DEVICE, "hipStreamCreate"); }
    for (int i = 0; i < n; ++i) {
        osb_ctx* c = nullptr;
        int r = osb_ctx_create(&descs[i], &c);
        if (!r) r = osb_set_stream(c, b->stream);
        if (r) {
            if (c) osb_ctx_destroy(c);
            osb_batch_destroy(b);
            return r;
        }
        b->ctx.push_back(c);
    }
    if (const char* s = std::getenv("OSB_BATCH_GRAPH")) b->use_graph = std::atoi(s) != 0;
    if (const char* s = std::getenv("OSB_BATCH_TUNE")) b->autotune = std::atoi(s) != 0;
    int r = build_groups(b);
    if (r) {
        std::string m = b->err;
        osb_batch_destroy(b);
        return berr(nullptr, r, m);
    }
    *out = b;
    return OSB_OK;
}

int osb_batch_destroy(osb_batch* b) {
    if (!b) return OSB_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& gr : b->groups) free_group(gr);
    for (osb_ctx* c : b->ctx) osb_ctx_destroy(c);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return OSB_OK;
}

int osb_batch_set_state(osb_batch* b, int i, const double* host, int layout) {
    if (!b || i < 0 || i >= (int)b->ctx.size()) return OSB_E_INVALID;
    int r = osb_set_state(b->ctx[i], host, layout);
    if (r) return berr(b, r, b->ctx[i]->err);
    b->items_dirty = true;
    return OSB_OK;
}

int osb_batch_get_state(osb_batch* b, int i, double* host, int layout) {
    if (!b || i < 0 || i >= (int)b->ctx.size()) return OSB_E_INVALID;
    int r = osb_get_state(b->ctx[i], host, layout);
    if (r) return berr(b, r, b->ctx[i]->err);
    return OSB_OK;
}

int osb_batch_step(osb_batch* b, int nsteps) {
    if (!b || nsteps < 0) return OSB_E_INVALID;
    if (nsteps == 0) return OSB_OK;
    int r = prepare(b);
    if (r) return r;
    for (int s = 0; s < nsteps; ++s) {
        for (auto& gr : b->groups) {
            if (b->use_graph) B_TRY(b, hipGraphLaunch(gr.exec[gr.parity], b->stream));
            else enqueue_group_step(b->stream, gr, gr.parity);
            gr.parity ^= gr.stages & 1;
        }
        for (int i : b->solo) {
            r = osb_step(b->ctx[i], 1);
            if (r) return berr(b, r, b->ctx[i]->err);
        }
    }
    B_TRY(b, hipGetLastError());
    for (auto& gr : b->groups) finish_members(b, gr, nsteps);
    return OSB_OK;
}

int osb_batch_sync(osb_batch* b) {
    if (!b) return OSB_E_INVALID;
    B_TRY(b, hipStreamSynchronize(b->stream));
    B_TRY(b, hipGetLastError());
    return OSB_OK;
}

int osb_batch_get_stats(osb_batch* b, int i, osb_stats* out) {
    if (!b || i < 0 || i >= (int)b->ctx.size()) return OSB_E_INVALID;
    int r = osb_get_stats(b->ctx[i], out);
    if (r) return berr(b, r, b->ctx[i]->err);
    return OSB_OK;
}

int osb_batch_autotune(osb_batch* b) {
    if (!b) return OSB_E_INVALID;
    if (b->items_dirty) {
        int r = upload_items(b);
        if (r) return r;
    }
    int r = autotune(b);
    if (r) return r;
    for (auto& gr : b->groups)
        for (auto& e : gr.exec)
            if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    return OSB_OK;
}

int osb_batch_describe(osb_batch* b, char* buf, int len) {
    if (!b || !buf || len < 1) return OSB_E_INVALID;
    std::string s = "{\"groups\": [";
    for (size_t k = 0; k < b->groups.size(); ++k) {
        const Group& gr = b->groups[k];
        char tmp[256];
        std::snprintf(tmp, sizeof(tmp), "%s{\"kind\": \"%s\", \"ndim\": %d, \"scheme\": %d, \"mixed\": %d, \"time\": %d, \"stages\": %d, \"members\": %d, \"tuning\": {",
                      k ? ", " : "", gr.kind == GK_CENTRAL ? "central" : "shock", gr.ndim, gr.kind == GK_CENTRAL ? gr.conv : gr.sch,
                      gr.mixed, gr.time, gr.stages, gr.n());
        s += tmp;
        bool first = true;
        if (gr.kind == GK_SHOCK && gr.fused && gr.ndim >= 2) {
            std::snprintf(tmp, sizeof(tmp), "\"resid_w\": %d, \"resid_w_us\": [%.2f, %.2f]", gr.resid_w, (double)gr.resid_w_us[0],
                          (double)gr.resid_w_us[1]);
            s += tmp;
            first = false;
        }
        if (gr.kind == GK_CENTRAL) {
            const CfTile t = CF_TILES[gr.cf_tile];
            std::snprintf(tmp, sizeof(tmp), "\"tile\": [%d, %d, %d], \"tile_us\": [%.2f, %.2f, %.2f]", t.tx, t.ty, t.tz,
                          (double)gr.tile_us[0], (double)gr.tile_us[1], (double)gr.tile_us[2]);
            s += tmp;
            first = false;
        }
        for (int f = 0; f < BF_COUNT; ++f) {
            if (gr.total(f) == 0 || (f >= BF_FACE0 + gr.ndim && f <= BF_FACE2)) continue;
            if (gr.fused && f == BF_FACE0 + gr.ndim - 1) continue;          // evaluated inside face_resid_rk
            const char* name = gr.kind == GK_CENTRAL && f == BF_RESID ? "central_fused"
                               : ((gr.fused && f == BF_RESID) ? "face_resid_rk" : kBFamName[f]);
            std::snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"virtual_blocks\": %d, \"grid\": %u, \"us\": %.2f}", first ? "" : ", ",
                          name, gr.total(f), grid_of(gr, f), (double)gr.tuned_us[f]);
            s += tmp;
            first = false;
        }
        s += "}}";
    }
    char tmp[128];
    std::snprintf(tmp, sizeof(tmp), "], \"solo\": %d, \"graphs\": %s}", (int)b->solo.size(), b->use_graph ? "true" : "false");
    s += tmp;
    std::strncpy(buf, s.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return (int)s.size() < len ? OSB_OK : OSB_E_INVALID;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Serving pool (osb_pool_*): solves are admitted while others run, as slots
// and device memory free up.  Shock-capturing solves of one signature form a
// group that advances in chunks of steps through the batched launches; a
// solve admitted while its group's chunk is on the GPU joins at the next
// chunk.  Other solves (3-D central / viscous) run their chunks on side
// streams.  Every solve's arithmetic is that of stepping it alone.
// ---------------------------------------------------------------------------
namespace {

struct PoolSolve {
    osb_ctx* c = nullptr;
    int state = 0;                 // 0 free, 1 running, 2 done
    int remaining = 0;
    int group = -1;                // signature group, or -1 (solo)
    hipStream_t stream = nullptr;  // solo: its side stream
    hipEvent_t ev = nullptr;       // solo: end of the last chunk
    bool inflight = false;
    int64_t bytes = 0;
};

struct PoolGroup {
    Group g;                       // signature and device tables (members rebuilt per chunk)
    int cap_items = 0;             // capacity of g.d_items
    size_t cap_offs = 0;           // capacity of g.d_offs (ints)
    std::vector<BItem> staged;     // host copy of the items of the chunk in flight
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool inflight = false;
    std::vector<int> running;      // slots in the chunk in flight
    int k = 0;                     // steps of the chunk in flight
};

}  // namespace

struct osb_pool {
    int64_t budget = 0, used = 0;
    std::vector<PoolSolve> slots;
    std::vector<hipStream_t> streams;
    std::vector<PoolGroup> groups;
    int next_stream = 0;
    std::string err;
};

namespace {

int perr(osb_pool* p, int code, const std::string& m) {
    if (p) p->err = m; else g_osb_create_error = m;
    return code;
}

#define P_TRY(p, expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (void)hipGetLastError(); /* clear the runtime's sticky error */                   \
            return perr((p), OSB_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));  \
        }                                                                                     \
    } while (0)

int signature_group(osb_pool* p, const osb_ctx* c) {
    const int stages = c->d.time_scheme == OSB_TIME_RKLS3 ? c->rk_stages : 3;
    const int kind = group_kind(c);
    for (int i = 0; i < (int)p->groups.size(); ++i)
        if (same_signature(p->groups[i].g, c, kind, stages)) return i;
    p->groups.emplace_back();
    PoolGroup& pg = p->groups.back();
    set_signature(pg.g, c, kind, stages);
    pg.stream = p->streams[p->next_stream++ % p->streams.size()];
    if (hipEventCreateWithFlags(&pg.ev, hipEventDisableTiming) != hipSuccess) pg.ev = nullptr;
    return (int)p->groups.size() - 1;
}

// collect finished chunks: solo events and group events
int pool_refresh(osb_pool* p) {
    for (auto& s : p->slots) {
        if (s.state != 1 || s.group >= 0 || !s.inflight) continue;
        hipError_t e = hipEventQuery(s.ev);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) return perr(p, OSB_E_DEVICE, std::string("solve chunk: ") + hipGetErrorString(e));
        s.inflight = false;
        if (s.remaining == 0) s.state = 2;
    }
    for (auto& pg : p->groups) {
        if (!pg.inflight) continue;
        hipError_t e = hipEventQuery(pg.ev);
        if (e == hipErrorNotReady) continue;
        if (e != hipSuccess) return perr(p, OSB_E_DEVICE, std::string("group chunk: ") + hipGetErrorString(e));
        pg.inflight = false;
        for (int sl : pg.running) {
            PoolSolve& s = p->slots[sl];
            s.inflight = false;
            if (s.remaining == 0) s.state = 2;
        }
        pg.running.clear();
    }
    return OSB_OK;
}

int launch_group_chunk(osb_pool* p, PoolGroup& pg, const std::vector<int>& members, int chunk) {
    Group& g = pg.g;
    const int n = (int)members.size();
    int k = chunk;
    for (int sl : members) k = std::min(k, p->slots[sl].remaining);
    // virtual-block offsets and items of this chunk's members
    g.members.clear();
    for (int sl : members) g.members.push_back(sl);
    g.offs.assign((size_t)BF_COUNT * (n + 1), 0);
    g.any_cfl = false;
    g.any_glf = false;
    for (int f = 0; f < BF_COUNT; ++f) {
        long acc = 0;
        for (int m = 0; m < n; ++m) {
            g.offs[(size_t)f * (n + 1) + m] = (int)acc;
            acc += vblocks(family_work(p->slots[members[m]].c, f, g.cf_tile, g.resid_w));
            if (acc > 0x7fffffffL) return perr(p, OSB_E_INVALID, "pool group too large for 32-bit block offsets");
        }
        g.offs[(size_t)f * (n + 1) + n] = (int)acc;
    }
    if (n > pg.cap_items) {
        if (g.d_items) (void)hipFree(g.d_items);
        g.d_items = nullptr;
        const int cap = std::max(n, 2 * pg.cap_items);
        P_TRY(p, hipMalloc((void**)&g.d_items, sizeof(BItem) * cap));
        pg.cap_items = cap;
    }
    if (g.offs.size() > pg.cap_offs) {
        if (g.d_offs) (void)hipFree(g.d_offs);
        g.d_offs = nullptr;
        const size_t cap = std::max(g.offs.size(), 2 * pg.cap_offs);
        P_TRY(p, hipMalloc((void**)&g.d_offs, sizeof(int) * cap));
        pg.cap_offs = cap;
    }
    g.parity = 0;
    pg.staged.assign(n, BItem{});
    for (int m = 0; m < n; ++m) {
        const osb_ctx* c = p->slots[members[m]].c;
        BItem& x = pg.staged[m];
        std::memset(&x, 0, sizeof(x));
        x.q[0] = c->q_cur;
        x.q[1] = c->q_cur == c->qa ? c->qb : c->qa;
        x.du = c->du; x.F = c->F; x.scal = c->scal; x.g = c->g; x.c = c->c;
        kb_central_orders(c->g, x.ord);
        for (int d = 0; d < 3; ++d) x.bc[d] = bc_sides(c->d, d);
        x.cfl = c->d.cfl;
        g.any_cfl = g.any_cfl || c->d.cfl > 0;
        g.any_glf = g.any_glf || c->c.glf != nullptr;
        if (c->d.time_scheme == OSB_TIME_RKLS3) {
            for (int s = 0; s < c->rk_stages; ++s) { x.rkA[s] = c->rkA[s]; x.rkB[s] = c->rkB[s]; }
        } else {
            const double rknew[3] = {2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0};
            const double rkold[3] = {1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0};
            for (int s = 0; s < 3; ++s) { x.rkA[s] = rknew[s]; x.rkB[s] = rkold[s]; }
        }
    }
    // the group's previous chunk has completed, so its tables are free to overwrite
    P_TRY(p, hipMemcpyAsync(g.d_items, pg.staged.data(), sizeof(BItem) * n, hipMemcpyHostToDevice, pg.stream));
    P_TRY(p, hipMemcpyAsync(g.d_offs, g.offs.data(), sizeof(int) * g.offs.size(), hipMemcpyHostToDevice, pg.stream));
    int parity = 0;
    for (int s = 0; s < k; ++s) {
        enqueue_group_step(pg.stream, g, parity);
        parity ^= g.stages & 1;
    }
    P_TRY(p, hipGetLastError());
    P_TRY(p, hipEventRecord(pg.ev, pg.stream));
    pg.inflight = true;
    pg.k = k;
    pg.running = members;
    // host view of the members after the chunk
    Group view;
    view.stages = g.stages;
    for (int sl : members) {
        PoolSolve& s = p->slots[sl];
        s.inflight = true;
        s.remaining -= k;
        osb_ctx* c = s.c;
        const int flips = (int)(((long)k * g.stages) & 1);
        double* cur = c->q_cur;
        double* other = c->q_cur == c->qa ? c->qb : c->qa;
        if (flips) std::swap(cur, other);
        c->q_cur = cur;
        c->q_prev = other;
        c->have_prev = true;
        c->iter += k;
        if (!(c->d.cfl > 0))
            for (int t = 0; t < k; ++t) { c->time += c->d.dt; c->last_dt = c->d.dt; }
    }
    return OSB_OK;
}

}  // namespace

extern "C" {

int osb_pool_create(int64_t hbm_budget_bytes, int max_solves, int nstreams, osb_pool** out) {
    if (!out || hbm_budget_bytes <= 0 || max_solves < 1 || nstreams < 1 || nstreams > 16)
        return perr(nullptr, OSB_E_INVALID, "bad pool arguments (budget > 0, max_solves >= 1, 1 <= nstreams <= 16)");
    *out = nullptr;
    osb_pool* p = new osb_pool();
    p->budget = hbm_budget_bytes;
    p->slots.resize(max_solves);
    for (int i = 0; i < nstreams; ++i) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            osb_pool_destroy(p);
            return perr(nullptr, OSB_E_DEVICE, "hipStreamCreate");
        }
        p->streams.push_back(s);
    }
    *out = p;
    return OSB_OK;
}

int osb_pool_destroy(osb_pool* p) {
    if (!p) return OSB_OK;
    for (hipStream_t s : p->streams) (void)hipStreamSynchronize(s);
    for (auto& s : p->slots) {
        if (s.c) osb_ctx_destroy(s.c);
        if (s.ev) (void)hipEventDestroy(s.ev);
    }
    for (auto& pg : p->groups) {
        free_group(pg.g);
        if (pg.ev) (void)hipEventDestroy(pg.ev);
    }
    for (hipStream_t s : p->streams) (void)hipStreamDestroy(s);
    delete p;
    return OSB_OK;
}

const char* osb_pool_last_error(const osb_pool* p) { return p ? p->err.c_str() : g_osb_create_error.c_str(); }

int osb_pool_admit(osb_pool* p, const osb_problem_desc* desc, const double* host_q0, int nsteps, int* slot) {
    if (!p || !desc || !host_q0 || !slot || nsteps < 1) return OSB_E_INVALID;
    *slot = -1;
    int64_t need = 0;
    int r = osb_estimate_bytes(desc, &need);
    if (r) return perr(p, r, g_osb_create_error);
    if (p->used + need > p->budget) return perr(p, OSB_E_NOMEM, "device-memory budget: solve waits");
    int sl = -1;
    for (int i = 0; i < (int)p->slots.size(); ++i)
        if (p->slots[i].state == 0) { sl = i; break; }
    if (sl < 0) return perr(p, OSB_E_NOMEM, "all solve slots busy: solve waits");
    osb_ctx* c = nullptr;
    r = osb_ctx_create(desc, &c);
    if (r) return perr(p, r, g_osb_create_error);
    PoolSolve& s = p->slots[sl];
    s = PoolSolve{};
    s.c = c;
    s.group = group_kind(c) >= 0 ? signature_group(p, c) : -1;
    s.stream = s.group >= 0 ? p->groups[s.group].stream : p->streams[p->next_stream++ % p->streams.size()];
    r = osb_set_stream(c, s.stream);
    if (!r) r = osb_set_state(c, host_q0, OSB_LAYOUT_INTERIOR);
    if (!r && s.group < 0 && hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) r = OSB_E_DEVICE;
    if (r) {
        std::string m = c->err;
        osb_ctx_destroy(c);
        s = PoolSolve{};
        return perr(p, r, m);
    }
    s.state = 1;
    s.remaining = nsteps;
    s.bytes = (int64_t)c->bytes;
    p->used += s.bytes;
    *slot = sl;
    return OSB_OK;
}

// Enqueue the next chunk (at most `chunk` steps) of every solve whose previous
// chunk has finished.  Does not wait for the GPU.
int osb_pool_advance(osb_pool* p, int chunk) {
    if (!p || chunk < 1) return OSB_E_INVALID;
    int r = pool_refresh(p);
    if (r) return r;
    for (int i = 0; i < (int)p->slots.size(); ++i) {
        PoolSolve& s = p->slots[i];
        if (s.state != 1 || s.inflight || s.group >= 0 || s.remaining == 0) continue;
        const int k = std::min(chunk, s.remaining);
        r = osb_step(s.c, k);
        if (r) return perr(p, r, s.c->err);
        P_TRY(p, hipEventRecord(s.ev, s.stream));
        s.inflight = true;
        s.remaining -= k;
    }
    for (int gi = 0; gi < (int)p->groups.size(); ++gi) {
        PoolGroup& pg = p->groups[gi];
        if (pg.inflight) continue;
        std::vector<int> members;
        for (int i = 0; i < (int)p->slots.size(); ++i) {
            const PoolSolve& s = p->slots[i];
            if (s.state == 1 && s.group == gi && !s.inflight && s.remaining > 0) members.push_back(i);
        }
        if (members.empty()) continue;
        r = launch_group_chunk(p, pg, members, chunk);
        if (r) return r;
    }
    return OSB_OK;
}

// Slots whose solves have completed every step (not yet released).
int osb_pool_poll(osb_pool* p, int* done_slots, int cap, int* ndone) {
    if (!p || !ndone || (cap > 0 && !done_slots)) return OSB_E_INVALID;
    int r = pool_refresh(p);
    if (r) return r;
    int n = 0;
    for (int i = 0; i < (int)p->slots.size(); ++i)
        if (p->slots[i].state == 2) {
            if (n < cap) done_slots[n] = i;
            ++n;
        }
    *ndone = n;
    return OSB_OK;
}

int osb_pool_fetch(osb_pool* p, int slot, double* host, osb_stats* st) {
    if (!p || slot < 0 || slot >= (int)p->slots.size() || p->slots[slot].state == 0) return OSB_E_INVALID;
    PoolSolve& s = p->slots[slot];
    int r = OSB_OK;
    if (host) r = osb_get_state(s.c, host, OSB_LAYOUT_INTERIOR);
    if (!r && st) r = osb_get_stats(s.c, st);
    if (r) return perr(p, r, s.c->err);
    return OSB_OK;
}

int osb_pool_release(osb_pool* p, int slot) {
    if (!p || slot < 0 || slot >= (int)p->slots.size() || p->slots[slot].state == 0) return OSB_E_INVALID;
    PoolSolve& s = p->slots[slot];
    if (s.inflight) {
        P_TRY(p, hipStreamSynchronize(s.stream));
        int r = pool_refresh(p);
        if (r) return r;
    }
    osb_ctx_destroy(s.c);
    if (s.ev) (void)hipEventDestroy(s.ev);
    p->used -= s.bytes;
    s = PoolSolve{};
    return OSB_OK;
}

int osb_pool_usage(osb_pool* p, int64_t* used_bytes, int* running, int* done) {
    if (!p) return OSB_E_INVALID;
    int nr = 0, nd = 0;
    for (const auto& s : p->slots) { nr += s.state == 1; nd += s.state == 2; }
    if (used_bytes) *used_bytes = p->used;
    if (running) *running = nr;
    if (done) *done = nd;
    return OSB_OK;
}

}  // extern "C"
