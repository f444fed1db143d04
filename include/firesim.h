/*
 * firesim.h — C ABI of the MI355X (gfx950) fire-simulation step.
 *
 * One handle = one simulation grid on one GPU, the equivalent of the reference's
 * `FireSim` object (src/sim.js:48-649).  Every entry point returns an int status
 * (FS_OK = 0) and never throws; the message of the last failure on the calling
 * thread is returned by fs_last_error().  Handles are not thread-safe: calls on one
 * handle must be serialised by the caller (the reference is single-threaded JS).
 *
 * Work is asynchronous on the handle's HIP stream, like WebGPU's encode/submit:
 * fs_step() enqueues, fs_sync() waits (the reference awaits
 * queue.onSubmittedWorkDone(), src/main.js:244).
 *
 * Field layout (fs_read_field / fs_write_field / fs_device_ptr): tightly packed,
 * x fastest, then y, then z; 4 interleaved channels (RGBA) for the vector fields,
 * 1 channel for pressure and divergence — the texel order of the reference's 3D
 * textures (src/sim.js:71-84).
 */
#ifndef FIRESIM_H
#define FIRESIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define FS_OK       0
#define FS_EINVAL  -1   /* bad argument, unknown parameter name (src/gpu.js:65 throws) */
#define FS_EDEVICE -2   /* HIP runtime error */
#define FS_ENOMEM  -3   /* device allocation failed */
#define FS_ESTATE  -4   /* call not valid in the handle's current state */

/* ---- storage precision of the advected fields ---------------------------- */
#define FS_PREC_FP32        0  /* plain FP32 step: float4 fields, FP32 V-cycles           */
#define FS_PREC_FP16        1  /* reference format: rgba16float fields, FP32 V-cycles     */
#define FS_PREC_BF16        2  /* bf16x4 fields, FP32 V-cycles                            */

/* ---- pressure solve ------------------------------------------------------- */
#define FS_SOLVE_VCYCLES    0  /* `vcycles` V(2,2) cycles on p, warm-started (src/sim.js:642) */
#define FS_SOLVE_REFINE64   1  /* FP32 V-cycles inside FP64 defect correction to div_tol  */
#define FS_SOLVE_SPECTRAL64 3  /* build extension: the exact separable inner solve (DST-II / DCT-IV
                                  transforms as FP32 matrix-core GEMMs, one division by the
                                  eigenvalues) inside the FP64 defect correction to div_tol, in place
                                  of the V-cycles; nx even; adaptive = 1; max_vcycles caps the
                                  corrections.  The z pass is one fused kernel for nz = 32, 64, 96
                                  or 128 (nx ny a multiple of 32), rocBLAS GEMMs otherwise */
#define FS_SOLVE_TOL32      2  /* FP32 V-cycles on p until the FP64-evaluated ||b - A p||/||b||
                                  <= div_tol or max_vcycles (no defect correction)        */

/* ---- field identifiers ----------------------------------------------------- */
#define FS_VEL    0   /* xyz velocity (m/s), w = expansion phi              (src/sim.js:3)   */
#define FS_S      1   /* theta, fuel, soot, q*1e-4                           (src/sim.js:4)   */
#define FS_S2     2   /* products, age*Z, smoke, CO fraction                 (src/sim.js:5-6) */
#define FS_VORT   3   /* curl * dx, |curl * dx|                              (src/sim.js:210) */
#define FS_GEO    4   /* band, bed, solid, sdf                               (src/sim.js:76)  */
#define FS_VELX   5   /* ping-pong intermediates (src/sim.js:72-74)                            */
#define FS_SX     6
#define FS_S2X    7
#define FS_VELY   8
#define FS_SY     9
#define FS_S2Y   10
#define FS_P     11   /* level-0 pressure, 1 channel (r32float, src/sim.js:84)                */
#define FS_B     12   /* level-0 divergence / right-hand side, 1 channel                      */
#define FS_NOISE 13   /* 128^3 rgba16float noise texture   (src/noise.js:6-64)                */
#define FS_CURL  14   /* 128^3 rgba16float curl texture    (src/noise.js:67-90)               */
#define FS_NUM_FIELDS 15

/* ---- host-side element type for field transfers -------------------------- */
#define FS_DT_F32     0   /* converted to/from float                      */
#define FS_DT_F64     1   /* converted to/from double                     */
#define FS_DT_NATIVE  2   /* raw storage bytes (half/bf16/float/double)   */

/* ---- stages (the reference's Profiler labels + individual dispatches) ---- */
#define FS_STAGE_GEO      0   /* src/sim.js:498-523 */
#define FS_STAGE_ADV1     1   /* src/sim.js:142-157 */
#define FS_STAGE_ADV2     2   /* src/sim.js:159-197 */
#define FS_STAGE_CURL     3   /* src/sim.js:200-211 */
#define FS_STAGE_FORCES   4   /* src/sim.js:214-396 */
#define FS_STAGE_DIV      5   /* src/sim.js:399-413 */
#define FS_STAGE_MG       6   /* src/sim.js:642 (all V-cycles of one step, or the FP64 refinement) */
#define FS_STAGE_PROJECT  7   /* src/sim.js:526-550 */
#define FS_NUM_STAGES     8

/* Number of named parameters: SIM_FIELDS (src/sim.js:14-23), in that order. */
#define FS_NUM_PARAMS 60

typedef struct fs_sim fs_sim;

typedef struct fs_config {
    int    res[3];        /* cells (nx, ny, nz); reference default 80x120x80 (src/sim.js:49)  */
    float  size[3];       /* metres; dx = size[0]/res[0] (src/sim.js:53)                      */
    int    vcycles;       /* V-cycles per step in FS_SOLVE_VCYCLES (src/sim.js:49, default 2) */
    int    max_levels;    /* 0 = the reference's automatic hierarchy (src/sim.js:80-87)        */
    int    precision;     /* FS_PREC_*                                                         */
    int    solve;         /* FS_SOLVE_*                                                        */
    double div_tol;       /* FS_SOLVE_REFINE64 / TOL32: target ||b - A p|| / ||b||             */
    int    max_vcycles;   /* FS_SOLVE_REFINE64 / TOL32: cap on V-cycles per step               */
    int    noise_n;       /* noise/curl texture edge (reference: 128, src/main.js:36)          */
    int    noise_seed;    /* reference: 1 (src/main.js:36)                                     */
    int    device;        /* HIP device ordinal                                                */
    void*  stream;        /* hipStream_t to run on, or NULL for a stream owned by the handle   */
    int    use_graph;     /* 1 = replay each step from a captured HIP graph                    */
    int    adaptive;      /* FS_SOLVE_REFINE64 / TOL32: 1 = the host enqueues as many V-cycles
                             as the previous step needed and checks the device's stopping test
                             once per step (one stream sync per step, no wasted launches);
                             0 = fully asynchronous: max_vcycles iterations are enqueued and the
                             ones after convergence exit at once.  FS_SOLVE_SPECTRAL64 needs 1
                             (its library transforms cannot exit on the device's flag)         */
    float  omega;         /* FS_SOLVE_REFINE64 / TOL32: over-relaxation factor of the red-black
                             smoother on every level but the coarsest (0 = default 1.25; 1 =
                             plain Gauss-Seidel).  FS_SOLVE_VCYCLES, the reference's solve,
                             always uses 1                                                    */
    int    pre_sweeps;    /* FS_SOLVE_REFINE64 / TOL32: red-black sweeps before / after the     */
    int    post_sweeps;   /* coarse correction on every level but the coarsest (pre -1 = none; 0 = default:
                             V(1,2), which with omega 1.25 converges faster per unit of work
                             than the reference's V(2,2)); FS_SOLVE_VCYCLES always uses V(2,2) */
    int    direct_level;  /* FS_SOLVE_REFINE64 / TOL32: the level whose coarse-grid problem is solved
                             exactly (separable DST-II / DCT-IV transforms of its operator) in place of
                             the V-cycle's coarser levels; 0 = automatic (the first level below the
                             finest with x a multiple of 16, y of 4, z of 8, each <= 128, and at most
                             2^19 cells; the plain V-cycle if none), -1 = none (the V-cycle to
                             the coarsest level), k >= 1 = level k.  FS_SOLVE_VCYCLES never uses it */
    int    first_cycles;  /* FS_SOLVE_REFINE64: V-cycles of a step's first FP64 correction (the FP32 V-cycles
                             iterate on the correction, its residual reduced in FP32 inside each restriction);
                             later corrections take one each.  0 = default 3; 1 = one V-cycle per FP64 residual */
    int    reserved;
} fs_config;

/* Fills a config with the reference defaults (res 80x120x80, size 1.2x1.8x1.2,
 * vcycles 2, fp16 storage, noise 128 seed 1). */
void        fs_config_default(fs_config* cfg);

const char* fs_last_error(void);
const char* fs_version(void);

/* Replaces `new FireSim(device, {res, size, logs, vcycles, noiseTex, curlTex})`
 * (src/sim.js:49-101).  `logs` is packLogs() output: n_logs records of 8 floats
 * (a.xyz, r, b.xyz, burn) (src/scene.js:92-98); the noise and curl textures are
 * baked on the device from cfg->noise_seed (src/noise.js:6-90, src/main.js:36-37).
 * The state starts zeroed, as after reset().  FS_EINVAL (before any device call) for a
 * resolution below 2 on an axis, or one whose 4-channel field reaches 4 GiB (fp32
 * storage: 2^28 cells; fp16 / bf16: 2^29 cells; the kernels address a field with
 * 32-bit byte offsets). */
int  fs_create(const fs_config* cfg, const float* logs, int n_logs, fs_sim** out);
int  fs_destroy(fs_sim* h);

/* Parameter block (SIM_FIELDS order, src/sim.js:14-23; defaults src/sim.js:25-44). */
int         fs_param_count(void);
const char* fs_param_name(int index);
int         fs_param_index(const char* name);        /* -1 if unknown */
int         fs_param_defaults(float out[FS_NUM_PARAMS]);

/* Replaces setParams(obj) (src/sim.js:601-604): unknown names give FS_EINVAL and
 * change nothing; hotR0/hotR1/hotY1/bedRadius invalidate the emitter geometry.
 * The grid/clock entries (nx..nLogs) are owned by the handle and are ignored. */
int  fs_set_params(fs_sim* h, const char* const* names, const float* values, int n);
int  fs_get_params(fs_sim* h, float out[FS_NUM_PARAMS]);

/* Rewrites the log buffer and invalidates the geometry (src/main.js:147-149). */
int  fs_set_logs(fs_sim* h, const float* logs, int n_logs);

/* ---- pyrolysis spots and floor fuel flux (build extension; the reference's
 *      sources are the log bands of src/sim.js:246-297 and have no mass flux) ----
 * A spot is a sphere of `radius` m at `pos` (m, sim coordinates) that adds fuel
 * mass at `flux` kg/s: each step every cell within the radius gains
 *   dY = flux * dt * w / (rho0 * sum_cells(w) * dx^3),  w = 1 - smoothstep(r/2, r, |p - pos|),
 * rho0 = 1.204 kg/m^3, so sum(dY) * rho0 * dx^3 = flux * dt exactly per step.  Inside the
 * spot the gas temperature is held at least at `temp` (theta, like srcTemp of a hot
 * patch, src/sim.js:268), the fuel is tagged as wood gas and its age restarts, and the
 * spot counts as a source for the turbulence/confinement gates (srcAll).  The floor
 * flux field adds face_flux[z][x] * dt / (rho0 * dx) (kg/(m^2 s)) to the fuel of the
 * y = 0 cells.  Both act inside the forces kernel, after the reference's sources. */
#define FS_MAX_SPOTS 64
#define FS_RHO0 1.204f
typedef struct fs_spot {
    float pos[3];
    float radius;
    float flux;     /* kg/s */
    float temp;     /* pilot theta = (T - 293 K) / 1000; 0 = no pilot */
    float norm;     /* filled by fs_set_spots: 1 / (rho0 * sum_cells(w) * dx^3) */
    float pad;
} fs_spot;
int  fs_set_spots(fs_sim* h, const fs_spot* spots, int n);                 /* n <= FS_MAX_SPOTS */
int  fs_get_spots(fs_sim* h, fs_spot* out, int max, int* n);              /* with `norm` filled */
/* floor fuel flux, nx*nz values (x fastest); NULL clears it */
int  fs_set_face_flux(fs_sim* h, const float* flux_xz);

/* Replaces reset() (src/sim.js:606-618): zero the 10 dynamic fields and every
 * level's p; b and geo keep their contents; time = 0, frame = 0. */
int  fs_reset(fs_sim* h);

/* Replaces `nsteps` calls of step(enc, dt) (src/sim.js:627-648), enqueued (with
 * cfg.adaptive on a tolerance-driven solve, each step waits once for its solve's
 * stopping test before enqueueing the projection). */
int  fs_step(fs_sim* h, double dt, int nsteps);
int  fs_sync(fs_sim* h);                                /* onSubmittedWorkDone() */

/* Runs one stage of the step on the current state (for tests and profiling);
 * FS_STAGE_MG runs the whole pressure solve of one step. Does not advance time. */
int  fs_run_stage(fs_sim* h, int stage, double dt);

int  fs_read_field(fs_sim* h, int field, int dtype, void* host_dst, size_t bytes);
int  fs_write_field(fs_sim* h, int field, int dtype, const void* host_src, size_t bytes);
int  fs_field_info(fs_sim* h, int field, int dims[3], int* channels, int* native_bytes);
/* Device address of a field.  FS_P and FS_B are stored parity-split on the device
 * (each x-row holds its even-x cells, then its odd-x cells; row stride 2*ceil(nx/2));
 * fs_read_field / fs_write_field convert them to and from x-fastest order. */
void* fs_device_ptr(fs_sim* h, int field);

/* time (s), frame, grid, multigrid hierarchy (src/sim.js:80-92). */
int  fs_get_clock(fs_sim* h, double* time, int64_t* frame);
/* Sets the clock (time in s, frame count), e.g. to restart from a stored state; no
 * counterpart in the reference, whose clock only advances (src/sim.js:646-647). */
int  fs_set_clock(fs_sim* h, double time, int64_t frame);
int  fs_grid_info(fs_sim* h, int res[3], double* dx, double origin[3]);
int  fs_mg_info(fs_sim* h, int* levels, int dims[][3], int* coarse_iters);

/* Relative Poisson residual ||b - A p||/||b|| of the last solve, FP64 on the device,
 * and the number of V-cycles it used. */
int  fs_residual(fs_sim* h, double* rel, int* vcycles_used);

/* Per-kernel device time (ms) of the last step while profiling is on (HIP events
 * on the handle's stream): adv1, adv2, curl, forces, div, mg, project — the
 * Profiler labels of src/sim.js:629-644 with `adv` split into its two passes.
 * Profiled steps are launched eagerly, not from the graph. */
#define FS_NUM_TIMINGS 7
int  fs_profile_enable(fs_sim* h, int on);
int  fs_timings(fs_sim* h, float ms[FS_NUM_TIMINGS]);

/* Replaces buildCampfire(seed).logs + packLogs() (src/scene.js:13-42, 92-98):
 * writes max_logs*8 floats (zero-padded) and returns the number of real logs. */
int  fs_build_campfire(int seed, float* out, int max_logs);

/* ---- diagnostics ------------------------------------------------------------ */
/* FIRE.stats(name) (src/main.js:268-294), reduced on the device, for a 4-channel
 * state field (FS_VEL .. FS_S2Y, FS_VORT, FS_GEO): for channel c, out[5c + 0..4] =
 * min, max, mean over all cells, count of non-finite values, and the linear index
 * (x fastest) of the first cell holding the maximum (-1 if the channel has no finite
 * value).  Min, max and mean are over the finite values; the mean divides by the
 * cell count, as the reference does.  Waits for the handle's stream. */
int  fs_stats(fs_sim* h, int field, double out[20]);
/* fs_stats without any allocation: no counterpart in the reference; for callers that cannot
 * allocate device or host memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  Same results as fs_stats, bit for bit: the per-block partials and
 * the combine run on the device in a fixed scratch array of the library (one per GPU, about
 * 330 KiB, held until the process ends), so the field may have at most
 * FS_STATS_NOALLOC_MAX_CELLS cells (FS_EINVAL beyond).  Every handle on one GPU shares that
 * scratch: calls that overlap on one GPU (from several threads or streams) race.  Waits for the
 * handle's stream. */
#define FS_STATS_NOALLOC_MAX_CELLS (1 << 24)
int  fs_stats_noalloc(fs_sim* h, int field, double out[20]);
/* FIRE.profile() (src/main.js:322-355), reduced on the device from S: *q_watts =
 * heat release sum(354800 q / T * dx^3) over cells with q = |S.w| 1e4 > 0, T = 293 +
 * 1000 |S.x|; then for y-planes j = 0, 4, 8, ... one row of 6 doubles: height
 * (j + 0.5) dx, mean T of the 5x5 centre columns, max T, max |soot|, max |fuel|,
 * cells above 1000 K.  *n_rows = ceil(ny / 4); at most max_rows rows are written.
 * Rounding for display is left to the caller.  Waits for the handle's stream. */
int  fs_profile(fs_sim* h, double* q_watts, double* rows, int max_rows, int* n_rows);
/* Per-step diagnostics record, no counterpart in the reference: with capacity > 0
 * every later step appends, after its last kernel, one record of (4 + ny) doubles
 * to a device ring of `capacity` records: time after the step (s), heat release
 * (W, as fs_profile), max T (K), flame tip (m: top of the highest y-plane whose max
 * T exceeds t_flame, 0 if none), then per y-plane the max T of the 5x5 centre
 * columns (the axis profile behind flame intermittency).  capacity 0 turns it off. */
int  fs_diag_enable(fs_sim* h, int capacity, double t_flame);
/* Copies the held records, oldest first (at most max_records, the newest kept). */
int  fs_diag_read(fs_sim* h, double* out, int max_records, int* n_records);

/* ---- FP64 mass and energy budget of the steps (no counterpart in the reference) ----
 * With the budget on, every later step runs its source kernel (src/sim.js:214-396) in a tracing
 * form that writes the same fields, bit for bit, and reduces on the device, in FP64, the fuel
 * fraction and the temperature at the boundaries of the cell update's sub-models.  Each term is
 * a difference of consecutive f32 values of one cell, so the terms telescope:
 *   FUEL = FUEL0 + TRANSPORT + SOURCES + INJECTED - BURNED - DISSIPATED - REMOVED + ROUNDING,
 *   HEAT = HEAT0 + TRANSPORT + SOURCES + RELEASED - RADIATED - ENTRAINED - REMOVED + ROUNDING
 * up to the FP64 summation.  Masses are rho0 dx^3 sum(Y) (kg, rho0 = FS_RHO0); heat is the
 * constant-density thermal energy rho0 cp dx^3 sum(T - 293 K) (J, cp = 1005 J/(kg K)).
 * TRANSPORT is the change by advection and projection (inflow and outflow through the open
 * faces and the semi-Lagrangian scheme's non-conservation); ROUNDING the f32 conversions and
 * the storage rounding of the fields.  FUEL_EXPECTED accumulates sum(spot flux) dt + sum(floor
 * flux) dx^2 dt, what INJECTED equals where no cell reaches a fuel fraction of 1.  EXPANSION and
 * DIVERGENCE are the last step's sum(phi) dx^3 and sum(div - phi) dx^3 (m^3/s): the volume the
 * flame's gas expansion asks for, and the net outflow the projection must realise.  Cells whose
 * trace is not finite are counted in BAD_CELLS and left out of the sums. */
enum {
    FS_BUDGET_STEPS = 0, FS_BUDGET_FUEL0, FS_BUDGET_FUEL, FS_BUDGET_FUEL_TRANSPORT, FS_BUDGET_FUEL_SOURCES,
    FS_BUDGET_FUEL_INJECTED, FS_BUDGET_FUEL_BURNED, FS_BUDGET_FUEL_DISSIPATED, FS_BUDGET_FUEL_REMOVED,
    FS_BUDGET_FUEL_ROUNDING, FS_BUDGET_HEAT0, FS_BUDGET_HEAT, FS_BUDGET_HEAT_TRANSPORT, FS_BUDGET_HEAT_SOURCES,
    FS_BUDGET_HEAT_RELEASED, FS_BUDGET_HEAT_RADIATED, FS_BUDGET_HEAT_ENTRAINED, FS_BUDGET_HEAT_REMOVED,
    FS_BUDGET_HEAT_ROUNDING, FS_BUDGET_FUEL_EXPECTED, FS_BUDGET_EXPANSION, FS_BUDGET_DIVERGENCE,
    FS_BUDGET_BAD_CELLS, FS_BUDGET_N = 24
};
/* on = 1: clears the record and starts it from the current state (FUEL0, HEAT0); on = 0 stops
 * recording (the record is kept for fs_budget_read). */
int  fs_budget_enable(fs_sim* h, int on);
/* Copies the FS_BUDGET_N doubles of the record (waits for the handle's stream). */
int  fs_budget_read(fs_sim* h, double out[FS_BUDGET_N]);

/* Measurement utility with no counterpart in the reference: times `iters`
 * device-to-device copies of `bytes` with a plain streaming kernel (the box's
 * achievable HBM bandwidth, and the calibration of bandwidth counters). */
int  fs_bench_copy(int device, size_t bytes, int iters, float* ms_per_copy);
/* The achievable-bandwidth probe, no counterpart in the reference: times `iters` passes over `bytes`
 * of device memory with kind 0 = the plain copy of fs_bench_copy, 1 = a tuned copy (8 16-byte words a
 * thread, all loads in flight before the stores, non-temporal loads and stores), 2 = reads only,
 * 3 = writes only.  A copy pass moves 2 x bytes, the others bytes. */
int  fs_bench_stream(int device, size_t bytes, int kind, int iters, float* ms_per_pass);

/* ---- isolated pressure Poisson solve (cell-centred 7-point operator,
 *      Dirichlet p=0 on the x/z faces and the top, Neumann floor) ------------- */
typedef struct fs_poisson fs_poisson;

int  fs_poisson_create(const int dims[3], double h, int max_levels, int device, void* stream,
                       fs_poisson** out);
int  fs_poisson_destroy(fs_poisson* ps);
int  fs_poisson_set_rhs(fs_poisson* ps, int dtype, const void* host_b);   /* FS_DT_F32 / F64 */
int  fs_poisson_set_x(fs_poisson* ps, int dtype, const void* host_x);
/* Runs FP32 V(2,2) cycles inside FP64 defect correction until ||b-Ax||/||b|| <= tol
 * or max_cycles.  hist (max_cycles+1 entries, may be NULL): hist[0] = relative
 * residual of the starting x, hist[k] = after cycle k (FP64, device-reduced). */
int  fs_poisson_solve(fs_poisson* ps, double tol, int max_cycles, double* hist, int* cycles);
/* Plain FP32 V-cycles directly on p (the reference's solve, src/sim.js:642);
 * hist[k] (may be NULL) = relative residual after cycle k+1. */
int  fs_poisson_vcycles(fs_poisson* ps, int cycles, double* hist);
int  fs_poisson_read_x(fs_poisson* ps, int dtype, void* host_x);
/* Over-relaxation factor of every smoothing pass except the coarsest level's sweeps,
 * for both fs_poisson_solve and fs_poisson_vcycles (default 1: the reference's
 * Gauss-Seidel).  No counterpart in the reference. */
int  fs_poisson_set_omega(fs_poisson* ps, float omega);
int  fs_poisson_info(fs_poisson* ps, int* levels, int dims[][3], int* coarse_iters);

/* ---- reverse-mode derivative of a run of steps -------------------------------
 * No counterpart in the reference (its gradients are planned, not built); these are
 * the entry points a torch.autograd.Function binds.  An fs_adjoint belongs to one
 * fs_sim; it runs the steps itself (fs_adjoint_forward, the same kernels as fs_step),
 * keeping the state (vel, S, S2 and the pressure) every `checkpoint_every` steps,
 * and fs_adjoint_backward sweeps the steps in reverse: each checkpoint interval is
 * recomputed from its checkpoint, keeping every step's intermediates, and the steps'
 * hand-written adjoint kernels run last to first (projection, the pressure solve,
 * divergence, sources with the hot-gas viscosity's velocity Laplacian, curl, MacCormack
 * correction with the limiter's active set frozen, semi-Lagrangian pass).  Device memory:
 * (ceil(max_steps/k) + k + 1) states of 56 B per cell, k step records of 136 B per cell
 * (k = checkpoint_every) and 12 adjoint fields of 16 B per cell.
 * The derivative is that of the step with a converged pressure solve (the solve is
 * x = A^-1 b for the symmetric 7-point operator A, so its adjoint is the same solve);
 * requirements: FS_PREC_FP32 storage, FS_SOLVE_REFINE64.  Float atomics
 * accumulate scattered adjoints, so results vary in the last bits run to run.
 * Diagnostics records are not written by these steps. */
typedef struct fs_adjoint fs_adjoint;

int  fs_adjoint_create(fs_sim* h, int max_steps, int checkpoint_every, fs_adjoint** out);
int  fs_adjoint_destroy(fs_adjoint* a);
/* nsteps (<= max_steps) steps of dt from the sim's current state; the sim ends in the
 * final state, its clock advanced, as after fs_step.  Clears any seed. */
int  fs_adjoint_forward(fs_adjoint* a, double dt, int nsteps);
/* The adjoint of the final state for field FS_VEL (xyz; w ignored), FS_S or FS_S2:
 * host array of nx*ny*nz*4 values, x fastest (FS_DT_F32 or FS_DT_F64).  Unseeded
 * fields are zero. */
int  fs_adjoint_seed(fs_adjoint* a, int field, int dtype, const void* host, size_t bytes);
/* Seeds the mismatch L = sum over cells of (theta - T*)^2 + (soot - soot*)^2 of the
 * final state (theta, soot = channels 0, 2 of FS_S; targets: host arrays of nx*ny*nz
 * floats, x fastest); returns L (FP64).  Replaces every seed. */
int  fs_adjoint_seed_mismatch(fs_adjoint* a, const float* target_theta, const float* target_soot, double* loss);
/* The reverse sweep; each adjoint pressure solve runs FP64 refinement to solve_tol
 * (<= 0: the sim's div_tol) within max_cycles (<= 0: the sim's max_vcycles).  The seed is
 * consumed: a later backward needs a new seed (unseeded = zero). */
int  fs_adjoint_backward(fs_adjoint* a, double solve_tol, int max_cycles);
/* Forward mode: runs nsteps steps of dt like fs_adjoint_forward (recording the same
 * checkpoints) and carries the tangent-linear derivative along the direction given by
 * spot_dir (one value per spot, kg/s), face_dir (nz*nx floats, kg/m^2/s) and the initial
 * state's direction vel_dir / S_dir / S2_dir (nx*ny*nz*4 floats each); NULL = zero.  The
 * tangent of the final state is then read with fs_adjoint_read.  Each tangent pressure
 * solve runs FP64 refinement to solve_tol / max_cycles (<= 0: the sim's settings).  The
 * kernels mirror the adjoint's branch for branch, so the dot products of a seed with the
 * tangent and of the gradients with the direction agree to rounding. */
int  fs_adjoint_tangent(fs_adjoint* a, double dt, int nsteps, const double* spot_dir, const float* face_dir,
                        const float* vel_dir, const float* S_dir, const float* S2_dir, double solve_tol,
                        int max_cycles);
/* After fs_adjoint_backward: the adjoint of the initial state (FS_VEL, FS_S or FS_S2);
 * after fs_adjoint_tangent: the tangent of the final state.  Layout as fs_adjoint_seed. */
int  fs_adjoint_read(fs_adjoint* a, int field, int dtype, void* host, size_t bytes);
/* dL/d(flux) of each pyrolysis spot (kg/s; up to max_spots entries) and, when a floor
 * flux is set (fs_set_face_flux), dL/d(face flux) per floor cell (nz*nx, x fastest;
 * may be NULL). */
int  fs_adjoint_grads(fs_adjoint* a, double* spot_flux, int max_spots, double* face_flux);

/* ---- spark (ember) particles carried by the velocity field ---------------------
 * Replaces `new Sparks(device, {count, sim, curlTex})` (src/sparks.js:17-36), its
 * reset/burst/step (src/sparks.js:38-46, 241-271) and the update kernel
 * (src/sparks.js:55-135).  Drawing (src/sparks.js:273-289) stays with the renderer.
 * A particle is FS_SPARK_FLOATS floats: position + age, velocity + burn time left,
 * previous position + temperature (K), diameter (mm) / seed / alive / flicker (Hz).
 * fs_sparks_step reads the sim's current velocity: call it right after fs_step with
 * the same dt, on the sim's stream.  Dead particles take the due spawn slots in index
 * order (the reference takes them in atomic-counter order), so runs are reproducible. */
#define FS_SPARK_FLOATS 16
typedef struct fs_sparks fs_sparks;
typedef struct fs_spark_burst {          /* Sparks.burst({...}) options (src/sparks.js:44-46) */
    int   n;                             /* sparks (60)                                      */
    float pos[3];                        /* centre (0, 0.1, 0) m                             */
    float speed;                         /* launch speed (3 m/s)                             */
    float radius;                        /* spawn box half-size (0.05 m)                     */
    double dur;                          /* emission time (0.12 s; double like the JS number) */
    float fan;                           /* outward fan (0.6)                                */
    float size;                          /* size factor of burst sparks (1.3)                */
    int   has_aim;                       /* 0 = no aim (the reference's aim: null)           */
    float aim[3];                        /* aim direction (x, z used)                        */
} fs_spark_burst;

void fs_spark_burst_default(fs_spark_burst* b);
int  fs_sparks_create(fs_sim* h, int count, fs_sparks** out);          /* count <= 2^24 */
/* May be called after the sim was destroyed (it touches only the particle memory). */
int  fs_sparks_destroy(fs_sparks* sp);
int  fs_sparks_reset(fs_sparks* sp);
/* "rate" (sparks/s, 3), "turb" (1), "lifeK" (1), "sizeK" (1), "upK" (1), "brightness" (13) */
int  fs_sparks_set(fs_sparks* sp, const char* name, float value);
int  fs_sparks_burst(fs_sparks* sp, const fs_spark_burst* b);
int  fs_sparks_step(fs_sparks* sp, double dt);
int  fs_sparks_read(fs_sparks* sp, float* host, size_t bytes);         /* count * 16 floats */
int  fs_sparks_write(fs_sparks* sp, const float* host, size_t bytes);
int  fs_sparks_clock(fs_sparks* sp, double* time, int64_t* frame);
/* Device address of the particle array (count * 64 B), for renderers on the same device. */
void* fs_sparks_device_ptr(fs_sparks* sp);

/* ---- render-volume consumers of the step's fields -------------------------------
 * Replaces the Renderer passes that read sim.S / sim.S2 (bind groups at
 * src/renderer.js:603-610, dispatch at 668-690): the fire-light extraction
 * (src/renderer.js:104-164), the flame-field prefilter (173-191) and the volume
 * ray-march (194-326), with the uniform block of uploadParams (634-648) and the
 * blackbody LUT of setWhiteBalance (57-63, src/blackbody.js:78-88).  The SDF surface
 * pass, composite, bloom and tone map stay with the caller's renderer: the surface
 * depth comes in through fs_render_set_depth (without one, every pixel sees the
 * surface pass's sky depth, 1e6 m), and the volume (L, Tr) and aux (haze, depth)
 * images go out through fs_render_read / fs_render_device_ptr.
 * A frame runs on the sim's stream and reads the sim's current S and S2.  Every call but
 * fs_render_destroy needs the sim alive; fs_render_destroy touches only the renderer's memory. */
#define FS_RP_COUNT 88                   /* RP_FIELDS entries (src/renderer.js:8-17) */
#define FS_RENDER_LIGHTS 8               /* extracted point lights (src/renderer.js:27) */
enum { FS_RENDER_VOL = 0,                /* vw*vh x (L.rgb, Tr), binary16 on the device, read as float */
       FS_RENDER_AUX = 1,                /* vw*vh x (haze, surface depth), float */
       FS_RENDER_LIGHTS_BUF = 2,         /* 8 x (pos.xyz, luminance, colour.rgb, spread), float */
       FS_RENDER_SHARP = 3,              /* the prefiltered S (binary16 on the device, read as float) */
       FS_RENDER_LUT = 4 };              /* 2 x bbN x 4 floats: blackbody row, soot row */
typedef struct fs_render fs_render;
int  fs_render_param_count(void);
int  fs_render_param_index(const char* name);          /* -1 if unknown */
int  fs_render_param_defaults(float* out);             /* RP_DEFAULTS, FS_RP_COUNT floats */
/* Renderer(...) + resize(width, height, vol_scale): vw = max(8, round(width * vol_scale)). */
int  fs_render_create(fs_sim* h, int width, int height, double vol_scale, fs_render** out);
int  fs_render_destroy(fs_render* r);
/* this.params[name] = value, for the user-settable RP_FIELDS entries (exposure, stepMul, ...). */
int  fs_render_set(fs_render* r, const char* name, float value);
/* The uniform block the last fs_render_frame uploaded (FS_RP_COUNT floats). */
int  fs_render_get_uniforms(fs_render* r, float* out);
/* setCamera(lookAt(...)): position, unit forward / right / up, vertical field of view (rad). */
int  fs_render_set_camera(fs_render* r, const double pos[3], const double fwd[3], const double right[3],
                          const double up[3], double fov_y);
int  fs_render_set_white_balance(fs_render* r, double twb);   /* blackbodyLUT(1024, 400, 3000, twb) */
/* Surface depth (m along the ray) of the surface pass, dw x dh floats; NULL = sky everywhere. */
int  fs_render_set_depth(fs_render* r, const float* depth, int dw, int dh);
/* uploadParams(time, frame, jitter) then the lights, prefilter (flameSharp > 0 and cubic,
 * once per sim state) and volume passes of renderHDR. */
int  fs_render_frame(fs_render* r, double time, double frame, float jit_x, float jit_y);
int  fs_render_size(fs_render* r, int* vw, int* vh);
/* Empty-space skipping of the march (default on; no counterpart in the reference): samples
 * whose texels cannot make them contribute are passed over.  Results are bit-identical
 * either way; off is for timing comparisons. */
int  fs_render_empty_skip(fs_render* r, int on);
/* HIP events around the passes of each frame on the sim's stream; fs_render_timings gives the
 * last frame's device time of the lights, prefilter (0 if it did not run) and volume passes (ms). */
int  fs_render_profile_enable(fs_render* r, int on);
int  fs_render_timings(fs_render* r, float ms[3]);
int  fs_render_read(fs_render* r, int what, float* host, size_t bytes);
void* fs_render_device_ptr(fs_render* r, int what);
/* Bases of the exact pressure solve: kind 0 = DST-II (Dirichlet faces at both ends), kind 1 =
 * DCT-IV (Neumann low end, Dirichlet high end).  T is n x n row-major (row k = mode k, orthonormal),
 * lam[k] the eigenvalue of the 1D second difference with that boundary rule (h = 1).  Host only. */
int  fs_spectral_basis(int n, int kind, double* T, double* lam);
/* blackbodyLUT({n, Tmin, Tmax, Twb}) (src/blackbody.js:78-88): 2 x n x 4 floats. */
int  fs_blackbody_lut(int n, double tmin, double tmax, double twb, float* out);

#ifdef __cplusplus
}
#endif
#endif /* FIRESIM_H */
