/*
 * ts_fvm.h -- C ABI of the MI355X explicit finite-volume step library (libts_fvm.so).
 *
 * The library replaces one explicit step of TurboStructured's
 * ExplicitRungeKuttaFVM<Reconstruction> (src/Methods/ExplicitRungeKuttaFVM.h:1440-1491):
 * device ghost fill (src/kernel/kernel.h:693-915), reconstruction
 * (ExplicitRungeKuttaFVM.h:879-925), per-face Riemann flux and flux divergence
 * (ExplicitRungeKuttaFVM.h:1150-1424), the CFL time step (ExplicitRungeKuttaFVM.h:1427-1437)
 * and the update U += dt*R/V (src/kernel/kernel.h:319-331).  All of it runs as hand-written
 * HIP kernels for gfx950 on one GPU; the MPI halo exchange (kernel.h:334-690) is not needed
 * because the whole domain lives in one GPU's HBM.
 *
 * Conventions (no exceptions cross this boundary):
 *   - every entry point returns TS_OK (0) or a negative TS_ERR_* code;
 *   - ts_last_error() returns the message of the last failure on that context;
 *   - one context per host thread; a context owns one HIP stream and all device memory;
 *   - state arrays that cross the boundary use the reference layout: array-of-structures,
 *     5 doubles per cell (rho, rho*u, rho*v, rho*w, rho*E), ghost layers included,
 *     serial index (k*nYAll + j)*nXAll + i (src/kernel/grid.h:64-67, kernel.h:93-96).
 * Enumerations keep the numeric values of the reference's own enums so that a host can
 * static_cast them (src/Methods/MethodConfiguration.h:5-23,
 * src/BoundaryConditions/BoundaryConditionConfiguration.h:12-19).
 */
#ifndef TS_FVM_H
#define TS_FVM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TS_ABI_VERSION 1
#define TS_NVAR 5
#define TS_MAX_BC 8

/* ---- status codes ---------------------------------------------------------------------- */
enum {
    TS_OK = 0,
    TS_ERR_INVALID = -1,     /* bad argument or configuration */
    TS_ERR_HIP = -2,         /* a HIP runtime call failed */
    TS_ERR_UNPHYSICAL = -3,  /* NaN, non-positive density or pressure, or dt <= 0 */
    TS_ERR_UNSUPPORTED = -4, /* configuration not built into this library */
    TS_ERR_NOMEM = -5,       /* device memory exhausted or over the admission budget */
    TS_ERR_BUSY = -6         /* request not finished yet / queue full */
};

/* ---- method enums (values of the reference's enum class members) ----------------------- */
/* RPSolver (MethodConfiguration.h:12-16); HLLC is an extension with no reference counterpart */
enum { TS_RIEMANN_GODUNOV = 0, TS_RIEMANN_ROE = 1, TS_RIEMANN_NONE = 2, TS_RIEMANN_HLLC = 3 };
/* Reconstruction (MethodConfiguration.h:5-10) */
enum { TS_RECON_PWC = 0, TS_RECON_LINEAR = 1, TS_RECON_ENO2 = 2, TS_RECON_LINEAR_LIM = 3 };
/* LimiterType (MethodConfiguration.h:18-21) */
enum { TS_LIMITER_BJ = 0, TS_LIMITER_VENKAT = 1 };
/* BoundaryConditionType (BoundaryConditionConfiguration.h:12-19) */
enum {
    TS_BC_WALL = 0,
    TS_BC_SYMMETRY = 1,
    TS_BC_MOVING_WALL = 2,
    TS_BC_SUBSONIC_INLET = 3,
    TS_BC_SUBSONIC_OUTLET = 4,
    TS_BC_NATURAL = 5
};
/* arithmetic of the residual */
enum { TS_FP64 = 0, TS_FP32_RESIDUAL = 1 };
/* sides, in the order the ghost fill visits them */
enum { TS_XLEFT = 0, TS_XRIGHT = 1, TS_YLEFT = 2, TS_YRIGHT = 3, TS_ZLEFT = 4, TS_ZRIGHT = 5 };

/* One boundary condition: mirrors BoundaryConditionConfiguration (BoundaryConditionConfiguration.h:22-36). */
typedef struct ts_bc_config {
    int type;           /* TS_BC_* */
    double velocity[3]; /* MovingWall wall velocity / SubsonicInlet velocity */
    double pstatic;     /* SubsonicInlet static pressure */
    double density;     /* SubsonicInlet density */
} ts_bc_config;

/*
 * Solver configuration: the fields of KernelConfiguration (src/kernel/KernelConfiguration.h:21-93)
 * and MethodConfiguration (MethodConfiguration.h:26-37) that the inviscid step reads.
 */
typedef struct ts_config {
    int ndims;           /* 1, 2 or 3 */
    int nx, ny, nz;      /* interior cells per axis (ny = nz = 1 below 2-D / 3-D) */
    int ghost;           /* DummyLayerSize; 1 (every reference scenario uses 1) */
    double lx, ly, lz;   /* domain lengths; the domain starts at 0 on every axis */
    int periodic[3];     /* isPeriodicX/Y/Z */
    /* optional cell-centre coordinates and sizes per axis, ghost layers included
       (n + 2*ghost entries, grid.h:45-50); NULL = uniform grid built as grid.h:164-218 does */
    const double* coord[3];
    const double* h[3];
    double gamma;        /* Gamma */
    double cfl;          /* CFL */
    double eps;          /* Eps: Harten entropy-fix coefficient of the Roe solver */
    int riemann;         /* TS_RIEMANN_* */
    int recon;           /* TS_RECON_* */
    int limiter;         /* TS_LIMITER_* (TS_RECON_LINEAR_LIM only) */
    int rk_stages;       /* 1 = forward Euler (the reference's step), 3 = SSP-RK3 (extension) */
    int precision;       /* TS_FP64 or TS_FP32_RESIDUAL */
    /* boundary conditions: a table indexed by marker (MyConditions), a marker per side
       (DomainBCinfo::SetMarker), and optional per-boundary-face marker arrays for sides that
       carry several conditions (DomainBCinfo::getMarker evaluated at each face centre).
       Face arrays are ordered (k*ny + j) for x sides, (k*nx + i) for y sides, (j*nx + i) for z. */
    int n_bc;
    ts_bc_config bc[TS_MAX_BC];
    int side_marker[6];
    const uint8_t* face_marker[6];
    int device;          /* HIP device ordinal */
    /* viscous terms and body forces (KernelConfiguration.h:41-51, 86-88).  FP64 only; single
       solves only (a batch refuses them). */
    int viscous;         /* IsViscousFlow: viscous fluxes from cell velocity gradients (ERK:79-874) */
    double viscosity;    /* Viscosity; enters the viscous spectral-radius term (ERK:1239-1248)
                            whether or not `viscous` is set, as in the reference */
    int external_force;  /* IsExternalForceRequared: residual += V Sigma, V Sigma.u (kernel.h:941-965) */
    double sigma[3];     /* Sigma: potential force, e.g. a pressure gradient */
    int external_accel;  /* IsUnifromAccelerationRequared: residual += rho V g, rho V g.u (kernel.h:966-989) */
    double accel[3];     /* UniformAcceleration */
} ts_config;

/* StepInfo (kernel.h:56) plus the per-variable residual norms of UpdateSolution (kernel.h:326). */
typedef struct ts_step_info {
    double time;
    double dt;
    long long iteration;
    double residual_norm[5];
    int steps_done;      /* steps executed by this ts_step call */
} ts_step_info;

typedef struct ts_ctx ts_ctx;

/* Create a solver context: validates the configuration, builds the grid, allocates the
   device state.  Replaces Kernel::Init + InitializeMethod (kernel.h:178-281, ERK:39-75). */
int ts_create(const ts_config* cfg, ts_ctx** out);
int ts_destroy(ts_ctx* ctx);
/* Message of the context's last failure.  ts_last_error(NULL) returns the message of the calling
   thread's last failed call that has no context to hold it (ts_create, ts_batch_create,
   ts_step_noalloc, ts_eos_eval). */
const char* ts_last_error(const ts_ctx* ctx);

/* Number of doubles of the reference state array (nXAll*nYAll*nZAll*5, kernel.h:240). */
size_t ts_state_size(const ts_ctx* ctx);

/* Upload / download the conserved state in the reference AoS layout, ghosts included.
   Ghost values are ignored on upload; on download they hold the boundary-condition images of
   the current interior (kernel.h:693-915 applied to the returned state). */
int ts_set_state(ts_ctx* ctx, const double* aos_values, size_t n_doubles);
int ts_get_state(ts_ctx* ctx, double* aos_values, size_t n_doubles);

/* Values of selected interior cells, without downloading the whole state: the 5 conserved
   values of cell (ijk[3c], ijk[3c+1], ijk[3c+2]) go to out[5c..5c+4].  Indices are the
   reference's local ones (ghost layer included, as Kernel::getCellValues, kernel.h:93-96): i in
   1..nx, j in 1..ny (0 below 2-D), k in 1..nz (0 below 3-D).  For sensors (src/Sensors/Sensor.h). */
int ts_get_cells(ts_ctx* ctx, const int* ijk, size_t n, double* out);

/* Restore Time and Iteration after LoadSolution (kernel.h:1031-1079). */
int ts_set_time(ts_ctx* ctx, double time, long long iteration);

/*
 * Run up to max_steps explicit steps (ERK:1440-1491 each), stopping after the step at which
 * Time >= t_max (kernel.h:1841) or Time == next_snapshot_time.  dt is clamped so that
 * Time + dt lands on next_snapshot_time when that is > 0 (ERK:1433).  Steps are issued
 * without host synchronisation; the call returns after the last one has finished.
 */
int ts_step(ts_ctx* ctx, int max_steps, double t_max, double next_snapshot_time,
            ts_step_info* out);

/* Allocation-free step (no counterpart in the reference; for callers that cannot allocate device
   memory: code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
   The steps of ts_step on a context created from cfg, starting from (time, iteration), applied
   in place to d_state: device memory holding the interior state as 5 planes of nx*ny*nz doubles
   (rho, rho u, rho v, rho w, E; x fastest).  The grid tables, the control block and the stage
   arrays live in one fixed device region of TS_NOALLOC_SCRATCH_BYTES that the library holds on
   each GPU it runs on until the process ends; a problem that needs more returns TS_ERR_INVALID
   before any launch.  Calls no hipMalloc/hipFree/hipHostMalloc and builds no graph.  `stream`
   is a hipStream_t (NULL: the default stream); the call returns after its work has finished.
   Every call on one GPU shares that GPU's region: calls that overlap on one GPU race.  No dt
   log, residual norms or kernel timing; errors are read with ts_last_error(NULL). */
#define TS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ts_step_noalloc(const ts_config* cfg, double* d_state, double time, long long iteration,
                    int max_steps, double t_max, double next_snapshot_time, ts_step_info* out,
                    void* stream);

/* Enable the residual norms sum(V*|R|) per variable of UpdateSolution (kernel.h:326), reported
   in ts_step_info.residual_norm for the last step taken (R of U^n: the forward-Euler residual,
   or SSP-RK3 stage 1).  Off by default: they cost one extra residual pass per ts_step call
   (in its last step). */
int ts_enable_residual_norms(ts_ctx* ctx, int enable);

/* Device-side log of the dt of every step (for comparing dt sequences). */
int ts_enable_dt_log(ts_ctx* ctx, long long capacity);
int ts_get_dt_log(ts_ctx* ctx, double* out, long long capacity, long long* n_out);

/* Per-kernel timing with HIP events recorded on the context's stream around every launch.
   kind: 0 = dt pass, 1 = stage 1 / Euler stage, 2 = stage 2, 3 = stage 3, 4 = kinds 0-3,
   5 = residual-norm pass, 6 = viscous pre-pass (velocity gradients and viscous face fluxes).
   1-D forward Euler in FP64 on at most 64 strips runs chunks of steps in one workgroup: each
   chunk is one kind-1 launch (dt passes included) and kind 0 stays empty. */
int ts_enable_kernel_timing(ts_ctx* ctx, int enable);
int ts_get_kernel_timing(ts_ctx* ctx, int kind, double* total_ms, long long* launches);

/* Device bytes held by the context and the algorithmic HBM bytes per cell of each kernel
   kind of ts_get_kernel_timing except 5 (compulsory traffic: one state array = 5*sizeof(real)
   per cell; with viscous fluxes, 3*ndims velocity gradients and 4 values per face). */
int ts_memory_footprint(const ts_ctx* ctx, size_t* device_bytes);
int ts_algorithmic_bytes_per_cell(const ts_ctx* ctx, int kind, double* bytes);

/* The equation of state on the device: the GeneralEOS interface (generalEOS.h:6-23) as
   IdealGasEOS implements it (idealGasEOS.h:7-46), the functor the flux kernels evaluate
   pressures and sound speeds through.  out[i] = method(ro[i], x[i]) for n points of host arrays:
   method 0 GetPressure(ro, e), 1 GetSoundSpeed(ro, e), 2 GetInternalEnergy(ro, p),
   3 GetPressureEnergyDerivative(ro, e), 4 GetPressureDensityDerivative(ro, e), with e the
   specific internal energy.  Errors are read with ts_last_error(NULL). */
int ts_eos_eval(double gamma, int method, const double* ro, const double* x, size_t n, double* out,
                int device);

/* ---- batching and serving (extension: many independent problems on one GPU) ------------- */
typedef struct ts_batch ts_batch;

typedef struct ts_batch_config {
    int device;
    size_t hbm_budget_bytes; /* admission limit on the device memory of resident problems */
    int max_active;          /* maximum number of problems advanced together */
    int use_graphs;          /* capture the per-step launches of a fixed batch in a hipGraph */
    int autotune_tiles;      /* 1: time each candidate tiling per kernel family and load (floor of
                                log2 of its resident cells) over whole passes and keep the fastest
                                per cell-step; 0: a load rule.  Results do not depend on it. */
} ts_batch_config;

int ts_batch_create(const ts_batch_config* bcfg, ts_batch** out);
int ts_batch_destroy(ts_batch* b);
/* The text of the batch's last error, copied for the calling thread: valid until that thread
   calls this function again. */
const char* ts_batch_last_error(const ts_batch* b);

/* Queue one solve.  init_state is the reference AoS array of cfg's grid.  The solve runs until
   max_steps steps are done or Time reaches t_end (the last dt is clamped to land on t_end).
   Thread-safe with respect to ts_batch_run on another thread. */
int ts_batch_submit(ts_batch* b, const ts_config* cfg, const double* init_state,
                    int max_steps, double t_end, long long* ticket);

/* Admit queued solves that fit the budget, then advance every resident solve by up to n_steps
   steps (continuous batching: solves join and leave between steps). Returns the number of
   solves still resident or queued in *pending.  While the server thread runs
   (ts_batch_start_server) it is the only caller: this returns TS_ERR_INVALID then. */
int ts_batch_run(ts_batch* b, int n_steps, int* pending);

/* done = 1 when the ticket finished; then out_state (may be NULL) receives the final state in
   the reference AoS layout and info its step information; the ticket is released. */
int ts_batch_poll(ts_batch* b, long long ticket, int* done, double* out_state,
                  ts_step_info* info);

/* Serving: a background thread runs ts_batch_run continuously. */
int ts_batch_start_server(ts_batch* b);
int ts_batch_stop_server(ts_batch* b);

/* The batch autotuner's trials so far: one entry per (kernel family, load bucket, candidate
   tiling) that has been timed.  tile_param is the output rows per tile in 2-D and the 60-cell
   strips per work item in 1-D; chosen = 1 marks the winner of a bucket whose candidates are all
   timed.  Writes up to cap entries and their number to *n. */
typedef struct ts_tile_trial {
    int ndims, riemann, recon_variant, precision; /* the kernel family */
    int load_bucket;                              /* floor(log2(resident cells of the family)) */
    int tile_param;
    int passes;
    double ns_per_cell_step;
    int chosen;
} ts_tile_trial;
int ts_batch_tuning(ts_batch* b, ts_tile_trial* out, int cap, int* n);

/* Autotuner: time the candidate tile heights for cfg's kernel family at cfg's size (the
   occupancy-model choice, its neighbours and 4..64 rows) and store the winner in the
   process-wide tuning table used by later ts_create calls.  Results never depend on it. */
int ts_autotune(const ts_config* cfg, int* chosen_rows_per_tile, double* best_ms);

/* The same autotuner timed on the caller's own state (the reference AoS layout, n_doubles =
   ts_state_size of cfg's context; NULL = the quiescent state of ts_autotune) over `steps` steps
   per candidate, so that data-dependent cost (the exact solver's Newton iterations, limiter
   branches) counts.  Up to `cap` candidates are reported as (rows per tile, kernel ms per step);
   the winner is stored in the table and returned in *chosen. */
int ts_autotune_state(const ts_config* cfg, const double* aos_state, size_t n_doubles, int steps,
                      int* cand_rows_per_tile, double* cand_ms_per_step, int cap, int* n_cand,
                      int* chosen_rows_per_tile);

#ifdef __cplusplus
}
#endif

#endif /* TS_FVM_H */
