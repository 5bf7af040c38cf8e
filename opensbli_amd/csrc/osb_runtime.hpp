// osb_runtime.hpp -- the per-block context shared by the single-block C ABI
// (osb_hip.hip) and the batch / pool runtime (osb_batch.hip).
#pragma once

#include "osb_kernels.hpp"
#include "osb_central_fused.hpp"
#include "../../include/osb.h"

#include <string>
#include <vector>

// per-context device scalars (see osb_ctx::scal)
constexpr int OSB_SCAL_SIZE = 32;
constexpr int OSB_SCAL_GLF = 16;

// message of the last failed create call (osb_last_error(NULL))
extern std::string g_osb_create_error;

struct osb_ctx {
    osb_problem_desc d{};
    osb::Geom g{};
    osb::Consts c{};
    int nv = 0;
    bool shock = false, central = false, viscous = false;
    int sch = osb::SCH_WENOZ;
    bool mixed = false;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    double* qa = nullptr;
    double* qb = nullptr;
    double* q_cur = nullptr;        // current state
    double* q_prev = nullptr;       // input of the last stage of the last step
    double* du = nullptr;
    double* R = nullptr;
    double* F = nullptr;
    double* kappa = nullptr;        // WENO filter's Ducros sensor over [0,n), zero halos
    double* cr = nullptr;
    double* gr = nullptr;
    double* scal = nullptr;         // OSB_SCAL_SIZE: [0] max wavespeed bits, [1] dt, [2] device time,
                                    // [4..7] TGV reduction out, [OSB_SCAL_GLF..+9) GLF wave speeds
    double* staging = nullptr;      // interior-layout transfer buffer
    double* part = nullptr;         // reduction partials
    unsigned int* nan_cnt = nullptr;
    int part_blocks = 0;
    size_t bytes = 0;

    double time = 0.0;
    int64_t iter = 0;
    double last_dt = 0.0;
    bool have_prev = false;
    std::string err;
    int next_stage = 0;             // osb_stage: the stage the current step continues with
    int rk_stages = 3;              // RK-LS stages and coefficients (rk_LS.py:97-132)
    double rkA[5] = {0}, rkB[5] = {0};
    int fused_tile = 0;             // index into CF_TILES: timed at the first step (tune_fused_tile)
    bool tile_tuned = false;        // the tile choice has been made
    bool tile_forced = false;       // OSB_FUSED_TILE chose it
    std::vector<float> tile_us;     // timed microseconds per launch of each CF_TILES shape
    int resid_w = osb::FUSED_FW;    // lines per tile of the fused face-residual kernel (RESID_WS), timed at the first step
    std::vector<float> resid_w_us;  // timed microseconds per launch of each RESID_WS width
    bool resid_forced = false;      // OSB_RESID_W chose it
    int xcd_remap = 2;              // XCD-aware workgroup-to-tile order: 8x4x2 supertiles per XCD range

    bool shock_fused = true;        // shock-only stages: last direction's faces fused with the update (HLLC: separate)
    bool bc_periodic = false;       // every face periodic and n >= halo: one-pass halo fill (OSB_BC_KERNEL=dir: per direction)
    bool bc_in_stage = true;        // with bc_periodic, the 3-D tile stage kernel writes the halos itself (n >= 2 halo;
                                    // OSB_BC_KERNEL=pass: the separate one-pass fill)

    // captured time steps: graph p starts from buffer qa (p = 0) or qb (p = 1)
    bool use_graph = true;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double* g_cur[2] = {nullptr, nullptr};     // q_cur / q_prev after the captured step
    double* g_prev[2] = {nullptr, nullptr};

    // kernel timing (osb_profile_*)
    bool prof_on = false;
    struct Rec { int fam; hipEvent_t a, b; };
    std::vector<Rec> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[16] = {0};
    int64_t prof_n[16] = {0};
};

// the two faces of direction dir as the BC kernels take them
inline osb::BcSides bc_sides(const osb_problem_desc& d, int dir) {
    osb::BcSides b;
    for (int s = 0; s < 2; ++s) {
        b.type[s] = d.bc[dir][s];
        for (int v = 0; v < 5; ++v) b.val[s][v] = d.bc_value[dir][s][v];
    }
    return b;
}


