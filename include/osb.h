/*
 * osb.h -- C ABI of the OpenSBLI explicit-solver hot path.
 *
 * One header, two implementations:
 *   libosb_hip.so     the product: hand-written HIP kernels for gfx950
 *   libosb_oracle.so  the CPU restatement used only by tests and the
 *                     CPU-baseline leg of the benchmark
 *
 * The reference has no C ABI: its drop-in seam is the Python back-end call
 * OPSC(alg, ...) (opensbli/code_generation/opsc.py:408) that prints an OPS
 * C program, plus the OPS runtime calls that program makes
 * (ops_par_loop / ops_halo_transfer / ops_reduction_result,
 *  opensbli/core/kernel.py:303-388, opensbli/code_generation/opsc.py:979-1064).
 * Each entry point below names the part of that seam it replaces.
 *
 * Conventions
 *   - every function returns int: 0 = OK, < 0 = error code (OSB_E_*);
 *     osb_last_error(ctx) holds a message.  Nothing aborts across the ABI.
 *   - state is structure-of-arrays, conservative variables in the order
 *     rho, rhou0 .. rhou{ndim-1}, rhoE; x fastest.
 *   - a context is externally synchronised; distinct contexts may be driven
 *     from distinct host threads.
 */
#ifndef OSB_H
#define OSB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSB_ABI_VERSION 5

/* error codes */
#define OSB_OK            0
#define OSB_E_INVALID    -1   /* bad descriptor / argument                  */
#define OSB_E_UNSUPPORTED -2  /* combination not implemented               */
#define OSB_E_DEVICE     -3   /* HIP runtime error                          */
#define OSB_E_NOMEM      -4   /* allocation failed / over budget            */
#define OSB_E_NONFINITE  -5   /* NaN/Inf detected (ops_NaNcheck analogue)   */

/* convective schemes (opensbli/schemes/spatial/{weno,teno}.py, optimisations/StoreSome.py) */
#define OSB_CONV_WENO_Z            1  /* LFWeno(order=5, formulation='Z')          weno.py:475   */
#define OSB_CONV_TENO              2  /* LFTeno(order=5|6)                          teno.py:480   */
#define OSB_CONV_CENTRAL_FEIEREISEN 3 /* NS_Split('Feiereisen') + StoreSome(4)      split_forms.py:111 */
#define OSB_CONV_CENTRAL_SKEW      4  /* Skew(...) convective terms + Central(4)    parsing.py:96 */
#define OSB_CONV_CENTRAL_KGP       5  /* NS_Split('KGP', energy_formulation='enthalpy') split_forms.py:178-234 */

/* characteristic averaging (opensbli/schemes/spatial/averaging.py) */
#define OSB_AVG_SIMPLE 1              /* SimpleAverage  averaging.py:31 */
#define OSB_AVG_ROE    2              /* RoeAverage     averaging.py:65 */

/* viscous terms (opensbli/physical_models/split_forms.py:383-392) */
#define OSB_VISC_NONE       0
#define OSB_VISC_CONSTANT   1         /* tau = (1.0/Re)(...)                         */
#define OSB_VISC_SUTHERLAND 2         /* tau = (mu/Re)(...), mu = 1.4042 T^1.5/(T+0.40417) */

/* flux splitting of the shock-capturing schemes (LFWeno/LFTeno flux_type) */
#define OSB_FLUX_LLF 0                /* local: lambda = max over {i, i+1}, shock_capturing.py:1000-1028 */
#define OSB_FLUX_GLF 1                /* global: lambda = domain max per stage, shock_capturing.py:416-438, :838-857 */
#define OSB_FLUX_HLLC 2               /* HLLCWeno(flux_type='HLLC'): WENO-Z of the characteristic state, HLLC
                                         Riemann flux of the reconstructed states, shock_capturing.py:1031-1256 */
#define OSB_FLUX_HLLC_LM 3            /* HLLCWeno(flux_type='HLLC-LM'): the low-Mach HLLC variant (same lines) */

/* non-linear filter step after each full RK step (WENOFilter, filters/WENO_filter.py:406) */
#define OSB_FILTER_NONE  0
#define OSB_FILTER_WENOZ 1            /* WENOFilter(order=5, formulation='Z', flux_type='LLF'):
                                         q += dt * (-div of the sensor-weighted dissipative
                                         part of characteristic WENO-Z, RoeAverage)   */

/* hybrid selection of the WENO filter (WENOFilter(optimize=...)) */
#define OSB_SENSOR_NONE   0           /* optimize=False: every face and point         */
#define OSB_SENSOR_DUCROS 1           /* optimize=True: modified Ducros sensor kappa,
                                         shock_sensors.py:16-59                       */

/* time integration */
#define OSB_TIME_RKLS3 1              /* RungeKuttaLS  rk_LS.py:17 (2N-storage; see rk_stages) */
#define OSB_TIME_RK3   2              /* RungeKutta(3)    rk_sbli.py:16 */

/* boundary conditions (opensbli/core/boundary_conditions/) */
#define OSB_BC_PERIODIC      1        /* PeriodicBC            periodic.py:7              */
#define OSB_BC_ZERO_GRADIENT 2        /* ZeroGradientOutletBC  zero_gradient_outlet.py:5  */
#define OSB_BC_INVISCID_WALL 3        /* InviscidWallBC        inviscid_wall.py:6         */
#define OSB_BC_SYMMETRY      4        /* SymmetryBC            symmetry.py:6              */
#define OSB_BC_EXTRAPOLATION0 5       /* ExtrapolationBC(order=0) extrapolation.py:36-40:
                                         boundary point and halos take q[1] (side 0)        */
#define OSB_BC_EXTRAPOLATION1 6       /* ExtrapolationBC(order=1) extrapolation.py:41-54:
                                         q[-t] = 2 q[-t+1] - q[-t+2], t = 1..h, in order    */
#define OSB_BC_DIRICHLET     7        /* DirichletBC dirichlet.py:28-41 with constant
                                         values: boundary point and halos = bc_value     */

/* arithmetic of the flux / derivative evaluation */
#define OSB_PREC_FP64  1
#define OSB_PREC_MIXED 2              /* FP32 flux/derivatives over an FP64 state, FP64 smoothness
                                         indicators, TENO cut-off and RK accumulation            */

/* state layouts for osb_set_state / osb_get_state */
#define OSB_LAYOUT_INTERIOR 0         /* nv x n[2] x n[1] x n[0]                          */
#define OSB_LAYOUT_PADDED   1         /* nv x (n+2h)... with the context's halo depth h  */

typedef struct osb_problem_desc {
    uint32_t struct_size;      /* sizeof(osb_problem_desc)                         */
    uint32_t abi_version;      /* OSB_ABI_VERSION                                  */
    int32_t  ndim;             /* 1, 2 or 3                                        */
    int32_t  n[3];             /* interior points per direction (1 if unused)      */
    double   delta[3];         /* grid spacing Delta{d}block0                      */
    int32_t  conv_scheme;      /* OSB_CONV_*                                       */
    int32_t  order;            /* 5 (WENO-Z, TENO5), 6 (TENO6), 4 (central)        */
    int32_t  averaging;        /* OSB_AVG_* (shock capturing only)                 */
    int32_t  viscous;          /* OSB_VISC_*                                       */
    int32_t  time_scheme;      /* OSB_TIME_*                                       */
    int32_t  precision;        /* OSB_PREC_*                                       */
    int32_t  bc[3][2];         /* OSB_BC_* per direction and side                  */
    double   gama, minf, re, pr;
    double   weno_eps;         /* WENO-Z epsilon, 1e-40 in weno.py:293             */
    double   teno_eps;         /* TENO 'eps' constant, teno.py:369                 */
    double   teno_ct;          /* TENO_CT, teno.py:373                             */
    double   shock_filter_control; /* LLF wave-speed scale, shock_capturing.py:1014 */
    double   dt;               /* constant time step (used when cfl <= 0)          */
    double   cfl;              /* > 0: dt = cfl / max(sum_d (|u_d|+a)/Delta_d)     */
    int32_t  rk_stages;        /* OSB_TIME_RKLS3 only.  0: RungeKuttaLS(3), the
                                  Williamson coefficients of rk_LS.py:121-124;
                                  1..5: stages with rk_a/rk_b below (RungeKuttaLS
                                  formulation='SSP' 3/4/5 stages, or order 4,
                                  rk_LS.py:97-132)                                 */
    int32_t  flux_type;        /* OSB_FLUX_* (shock capturing only)                */
    double   rk_a[5];          /* rkA: du = rk_a[s]*du + dt*R                       */
    double   rk_b[5];          /* rkB: q  = q + rk_b[s]*du                          */
    double   bc_value[3][2][5];/* OSB_BC_DIRICHLET: conservative values per face,
                                  (rho, rhou0.., rhoE); unused entries ignored     */
    int32_t  shock_filter;     /* OSB_FILTER_*                                     */
    int32_t  filter_order;     /* WENO filter order: 5, or 3 (WENO3-Z); 0 means 5   */
    int32_t  filter_sensor;    /* OSB_SENSOR_*: the filter's hybrid selection      */
    int32_t  reserved0;
    double   ducros_check;     /* OSB_SENSOR_DUCROS: a face is reconstructed where
                                  max kappa over its stencil > ducros_check
                                  (WENO_filter.py:561-584, 'Ducros_check')         */
    double   ducros_select;    /* ... and a point filtered where max kappa over
                                  it and its +-1, +2 neighbours >= ducros_select
                                  (WENO_filter.py:486-500, 'Ducros_select')        */
} osb_problem_desc;

typedef struct osb_stats {
    uint32_t struct_size;
    int64_t  iter;             /* completed iterations                             */
    double   time;             /* simulation_time                                  */
    double   last_dt;          /* dt used by the last step                         */
    int32_t  halo;             /* stored halo depth                                */
    int32_t  nv;               /* conservative variables                           */
    int64_t  device_bytes;     /* device memory held by the context                */
} osb_stats;

typedef struct osb_ctx osb_ctx;
typedef struct osb_batch osb_batch;
typedef struct osb_pool osb_pool;

/* Lifecycle.  Replaces OPSC(alg) + ops_init/ops_decl_block/ops_decl_dat
 * (opsc.py:408, opsc.py:838-960): one context owns one block's arrays.      */
int  osb_ctx_create(const osb_problem_desc* desc, osb_ctx** out);
int  osb_ctx_destroy(osb_ctx* ctx);
const char* osb_last_error(const osb_ctx* ctx);   /* ctx may be NULL: last create error */
int  osb_abi_version(void);
/* Device bytes a context for `desc` allocates at creation (admission control). */
int  osb_estimate_bytes(const osb_problem_desc* desc, int64_t* out);

/* State transfer.  Replaces GridBasedInitialisation / ops_decl_dat_hdf5
 * (initialisation/gridbasedinit.py:16, opsc.py:962-977).                   */
int  osb_set_state(osb_ctx* ctx, const double* host, int layout);
int  osb_get_state(osb_ctx* ctx, double* host, int layout);
int  osb_set_time(osb_ctx* ctx, double time, int64_t iter);

/* Stepping.  Replaces the generated time loop
 * (algorithm/algorithm.py:387-502; TENO6/opensbli.cpp:206-432).             */
int  osb_step(osb_ctx* ctx, int nsteps);
/* One RK stage of the current step, for debugging: the body of the stage loop
 * (rk_LS.py:229-233 -- the RHS, the stage update, the BCs).  Stages run in
 * order 0..stages-1; stage 0 also computes the CFL dt, and the iteration count
 * and time advance after the last.  osb_step refuses a partly advanced step. */
int  osb_stage(osb_ctx* ctx, int stage);
int  osb_apply_bcs(osb_ctx* ctx);                        /* ops_halo_transfer + BC kernels */
int  osb_residual(osb_ctx* ctx, double* host_out);       /* one RHS evaluation, interior layout */
int  osb_max_wavespeed(osb_ctx* ctx, double* out);       /* max_points sum_d (|u_d|+a)/Delta_d */
int  osb_sync(osb_ctx* ctx);
int  osb_get_stats(osb_ctx* ctx, osb_stats* out);

/* Taylor-Green post-processing reductions KE, dilatation_dissipation,
 * enstrophy_dissipation, rhom (TENO6/opensbliblock00_kernels.h:4737-4797,
 * compressible_TGV.py:176-206).  Velocity and mu come from the constituent
 * relations of the last stage, rho from the current state, as in the
 * reference's post-process kernel placement (TENO6/opensbli.cpp:433-455).   */
int  osb_tgv_diagnostics(osb_ctx* ctx, double out[4]);

/* Point probes of the state (SimulationMonitor, simulation_monitors.py:40-265:
 * the monitor_%d_<name> reduction kernels over one grid point).  Probe p reads
 * conservative variable var[p] (0 = rho ... nv-1 = rhoE) at interior grid
 * index ijk[3p..3p+2] (unused directions 0) into out[p].                     */
int  osb_probe(osb_ctx* ctx, int n, const int32_t* ijk, const int32_t* var, double* out);

/* NaN check of rho (ops_NaNcheck, helperfunctions.py:201-214):
 * returns OSB_E_NONFINITE if any interior rho is not finite.                */
int  osb_nan_check(osb_ctx* ctx);

/* Kernel timing (HIP library only).  While enabled, every kernel launch of
 * the context is bracketed by a pair of hipEvents on the context's stream and
 * its duration is accumulated per kernel family.  osb_profile_read returns
 * family `idx` (0-based); OSB_E_INVALID past the last family.  Replaces the
 * OPS per-kernel timers (ops_timing_output, opsc.py:826-836, 1018-1024).    */
int  osb_profile_enable(osb_ctx* ctx, int enable);
int  osb_profile_read(osb_ctx* ctx, int idx, char* name, int name_len, double* total_ms, int64_t* launches);

/* The context's kernel path and launch shapes as JSON in buf (HIP library:
 * the fused 3-D stage kernel's tile shape, timed at the first step unless
 * OSB_FUSED_TILE forces one, with each candidate's microseconds per launch).
 * No counterpart in the reference, whose generated code fixes its kernels. */
int  osb_describe(osb_ctx* ctx, char* buf, int len);

/* Device buffers (HIP library only): pointer to the FP64 conservative
 * variable `var` (padded layout) for zero-copy interop.                     */
int  osb_state_device_ptr(osb_ctx* ctx, int var, void** out);
/* device-to-device state transfer in the interior layout (e.g. a torch
 * tensor's data_ptr(); the caller makes sure the source is ready) */
int  osb_set_state_device(osb_ctx* ctx, const double* dev, int layout);
int  osb_get_state_device(osb_ctx* ctx, double* dev, int layout);
int  osb_set_stream(osb_ctx* ctx, void* hip_stream);

/* Allocation-free stepping: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): advances a caller-owned
 * block by nsteps on `hip_stream`, allocating nothing.  q is the padded state
 * (nv x padded points, variable-major, x fastest, halos filled -- the
 * OSB_LAYOUT_PADDED layout of osb_get_state), updated in place; du is the RK
 * register in the same layout (zero before a run's first step).  The other
 * buffers a context would hold come from a fixed 8 MiB array per GPU, held
 * until the process ends; a block needing more (osb_estimate_bytes less the
 * state, register and staging) is refused with OSB_E_UNSUPPORTED and
 * osb_last_error(NULL) says why.  Calls that overlap on one GPU share that
 * array and race: the caller serialises them.  The steps launch eagerly
 * (identical kernels to osb_step); with CFL > 0 each step's dt stays on the
 * device (time is the caller's to keep).                                     */
int  osb_step_noalloc(const osb_problem_desc* desc, double* q, double* du, int nsteps, void* hip_stream);

/* Batched ensembles: many instances of possibly different sizes advanced by
 * one launch per RK stage (persistent kernels, hipGraph replay).  No
 * counterpart in the reference, which runs one block per executable.        */
int  osb_batch_create(const osb_problem_desc* descs, int n, osb_batch** out);
int  osb_batch_destroy(osb_batch* b);
int  osb_batch_set_state(osb_batch* b, int i, const double* host, int layout);
int  osb_batch_get_state(osb_batch* b, int i, double* host, int layout);
int  osb_batch_step(osb_batch* b, int nsteps);
int  osb_batch_sync(osb_batch* b);
const char* osb_batch_last_error(const osb_batch* b);

/* Serving pool (HIP library only): continuous batching of independent solves.
 * osb_pool_admit creates a solve if `max_solves` slots and the device-memory
 * budget allow it (else OSB_E_NOMEM: the caller keeps it queued).  Each
 * osb_pool_advance enqueues, without waiting, the next chunk of at most
 * `chunk` steps of every solve whose previous chunk has finished: solves of
 * one shock-capturing signature advance together through the batched
 * launches (a solve admitted meanwhile joins the next chunk); the others run
 * on side streams.  osb_pool_poll lists finished solves; fetch their state,
 * then release the slot and its memory.  No counterpart in the reference.   */
int osb_pool_create(int64_t hbm_budget_bytes, int max_solves, int nstreams, osb_pool** out);
int osb_pool_destroy(osb_pool* p);
int osb_pool_admit(osb_pool* p, const osb_problem_desc* desc, const double* host_q0, int nsteps, int* slot);
int osb_pool_advance(osb_pool* p, int chunk);
int osb_pool_poll(osb_pool* p, int* done_slots, int cap, int* ndone);
int osb_pool_fetch(osb_pool* p, int slot, double* host, osb_stats* st);
int osb_pool_release(osb_pool* p, int slot);
int osb_pool_usage(osb_pool* p, int64_t* used_bytes, int* running, int* done);
const char* osb_pool_last_error(const osb_pool* p);

#ifdef __cplusplus
}
#endif
#endif /* OSB_H */
