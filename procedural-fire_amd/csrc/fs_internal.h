// Internal declarations of the gfx950 fire-simulation library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/firesim.h"

// ---- parameter block: SIM_FIELDS order (src/sim.js:14-23) ------------------
enum {
    P_nx, P_ny, P_nz, P_dx, P_dt, P_time, P_frame, P_seed, P_ox, P_oy, P_oz, P_nLogs,
    P_fuelSrc, P_bedFuel, P_srcTemp, P_bedTemp, P_srcVel, P_srcFreq, P_srcSpeed, P_srcThresh,
    P_ignT, P_kMix, P_rs, P_heatOut, P_expansion, P_sootForm, P_sootOx, P_radCool,
    P_coolRate, P_sootDissip, P_smokeDissip, P_vitDecay, P_fuelDissip, P_buoyancy, P_vortEps, P_turbAmp,
    P_turbFreq, P_turbRise, P_turbEvolve, P_windX, P_windZ, P_gust, P_intensity, P_velDissip,
    P_bedRadius, P_logSrc, P_burstX, P_burstY, P_burstZ, P_burstAmt, P_smolder, P_sponge, P_hotR0, P_hotR1, P_hotY1,
    P_visc, P_lamAge, P_anchorFreq, P_srcFloor, P_anchorGain,
};

// Device-resident uniform block.  The f32 entries are what the reference uploads
// each step (src/sim.js:620-624); `time` and `dt` are kept in double as the host's
// JS numbers are, and rounded to f32 where a kernel reads them.  Living in device
// memory (not kernel arguments) keeps every launch of a step static, so the step
// can be replayed from a captured graph.
// Per-step values of the source kernel that depend on the parameters alone, computed once by the
// parameter upload (derive_params) with the same expressions the kernel would evaluate per cell:
// the five dissipation factors exp(-rate dt), turbAmp / 12, and the sponge factors of the five
// side-distance rows (side = 0.5 .. 4.5 cells) and of the ten top rows.
enum {
    D_eSootDissip, D_eFuelDissip, D_eVitDecay, D_eSmokeDissip, D_eVelDissip, D_turbAmp12,
    D_ks0, D_kt0 = D_ks0 + 5, D_COUNT = D_kt0 + 10,
};
struct DevParams {
    float f[64];
    float d[D_COUNT];            // derived by the upload kernel; the host leaves them zero
    double time;
    double dt;
    long long frame;
    int n_logs;
    int n_spots;                 // pyrolysis spots in FieldSet::spots
    int has_face_flux;           // FieldSet::face_flux holds nx*nz floor fluxes
    int pad[3];
};

// ---- storage of the 4-channel fields ------------------------------------------
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef unsigned short ushort4_t __attribute__((ext_vector_type(4)));

struct StF32  { typedef float4 raw; };
struct StF16  { typedef half4_t raw; };
struct StBF16 { typedef ushort4_t raw; };

// grid of a 3D field plus the block tiling of the per-cell kernels
struct Dims {
    int nx, ny, nz;
    int nbx, nby;      // tiles of 64 x 4 cells per block in x, y
    int nblocks;       // nbx * nby * nz
};
constexpr int TILE_X = 64;
constexpr int TILE_Y = 4;

// ---- multigrid (sim.js variant) --------------------------------------------
constexpr int MG_MAX_LEVELS = 12;
constexpr int MG_MAX_REFINE = 64;

struct RefineState {            // device-resident FP64 refinement bookkeeping
    double rel[MG_MAX_REFINE + 1];
    double bnorm2;
    int used;                   // V-cycles applied in the current solve
    int done;                   // 1 once rel <= tol or used == max
    int pending;                // partials not yet reduced: 0 none, 1 a residual after a V-cycle, 2 a starting residual
    int never;                  // always 0: the skip flag of launches that run unconditionally
    int inc;                    // V-cycles applied since the pending residual's predecessor
    // Steps enqueued ahead of their predecessor's stopping test: hold = 1 while the last finished step still
    // needs a correction (every kernel of a step enqueued behind it then exits at once); skipv = done | hold,
    // the stop flag of such a step's multigrid kernels
    int hold;
    int skipv;
};

// A pending residual reduction handed to the first smoothing pass of the next V-cycle, whose
// block (0, 0, 0) runs it before its own cells (the pass itself then cannot wait on the stop flag)
struct FinArgs {
    RefineState* st;
    const double* partials;
    int npart, max_cycles;
    double tol;
};

struct GpuMG {
    int nlev = 0;
    int dims[MG_MAX_LEVELS][3];
    float h[MG_MAX_LEVELS];
    int coarse_iters = 8;
    float* p[MG_MAX_LEVELS] = {};
    float* b[MG_MAX_LEVELS] = {};
    // FP64 defect correction
    double* x64 = nullptr;       // level-0 FP64 iterate
    float* bkeep = nullptr;      // level-0 right-hand side (the divergence)
    float* p0post = nullptr;     // level-0 V-cycle correction after the prolongation (refinement)
    double* partials = nullptr;  // per-block (r.r, b.b) pairs
    int npart = 0;
    RefineState* st = nullptr;
    int tail_from = -1;          // first level solved by the single-workgroup tail kernel
    bool use_tail = true;        // false: one launch per colour pass on every level
    bool fuse_prolong = true;    // prolongation folded into the first post-smoothing pass
    float omega = 1.f;           // over-relaxation of the smoother (all levels but the coarsest)
    int pre = 2, post = 2;       // sweeps before / after the coarse correction (all levels but the coarsest)
    int first = 1;               // V-cycles of a step's first FP64 correction (FS_SOLVE_REFINE64)
    int direct = 0;              // > 0: level solved exactly by separable transforms (the cycle stops there)
    // set (to &st->hold) while a step enqueued ahead of its predecessor's stopping test is recorded: its starting
    // residual and the first pass of each V-cycle skip on it, the other kernels on st->skipv
    const int* gate = nullptr;
    struct DirectSolve* ds = nullptr;
};

// Multigrid arrays are parity-split: each x-row stores its even-x cells, then its odd-x
// cells; rb_index is the offset of cell (x, y, z) in a level of width nx.
__host__ __device__ inline int rb_index(int x, int y, int z, int nx, int ny) {
    const int nh = (nx + 1) >> 1;
    return (z * ny + y) * 2 * nh + (x & 1) * nh + (x >> 1);
}
// An index the compiler cannot see through.  Stencil loads pick an in-bounds neighbour
// index with a select (the cell itself where the neighbour is outside); without this the
// compiler folds "load(select(i, own))" back into a conditional load of i and waits on
// each neighbour in turn instead of issuing all loads at once.
__device__ __forceinline__ int fs_opaque(int i) {
    asm("" : "+v"(i));
    return i;
}

int  mg_build(GpuMG& mg, int nx, int ny, int nz, float h, int max_levels, std::string& err);
// ---- exact coarse-level solve (fs_direct.hip) ------------------------------------
// Whether a level of X x Y x Z cells can be solved exactly by the separable-transform kernels.
bool direct_fits(int X, int Y, int Z);
int  direct_build(struct DirectSolve** out, int X, int Y, int Z, float h, std::string& err);
void direct_free(struct DirectSolve* d);
// p = A^-1 b on one level (parity-split arrays); the kernels exit at once when *skip is set
void direct_solve(const struct DirectSolve* d, hipStream_t s, const float* b, float* p, const int* skip);
// Chooses and sets up the exact coarse level of a tolerance solve: `want` = 0 automatic (the first
// level below the finest that direct_fits), -1 none, k >= 1 level k (FS_EINVAL if it does not fit).
int  mg_set_direct(GpuMG& mg, int want, std::string& err);
size_t mg_level_floats(const GpuMG& mg, int l);   // array length of level l (with row padding)
void mg_free(GpuMG& mg);
// FP32 V(2,2) cycle on level-0 p/b; `skip` (device int, may be null) = early exit flag.
// With `xupd` set, x += p (FP64) is fused into the cycle's last colour pass where
// possible; returns true if it was.
// With `red0_done`, level 0's first colour-0 pass has already been written (level 0 started
// from p = 0 and the kernel that zeroed it did that pass).
// With `fin`, the cycle's first kernel also runs that reduction.
// With p0 / p0alt, level 0's iterate is in p0 (default mg.p[0]) and the post-smoothing runs on p0alt, which the
// prolongation fills (a level-0 iterate that the cycle's caller treats as scratch); *p0out receives the array that
// holds level 0's iterate after the cycle.
bool mg_vcycle(GpuMG& mg, hipStream_t s, const int* skip, double* xupd = nullptr, bool red0_done = false,
               const FinArgs* fin = nullptr, float* p0 = nullptr, float* p0alt = nullptr, float** p0out = nullptr);
// FP64 refinement: r = bkeep - A x; sets level-0 b = r, p = 0; reduces norms;
// then up to max_cycles of {V-cycle, x += p, residual}; all device-side.
void mg_refine(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
// FP32 V-cycles on level-0 p until the FP64 residual <= tol or max_cycles (device-side test)
void mg_solve_tol32(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
// the same two solves split into their starting residual and one (skippable) iteration,
// for callers that decide on the host how many iterations to enqueue
// mg_refine_iter leaves its residual's reduction to the first kernel of the next iteration (or to
// mg_state_ready): call mg_state_ready before the host reads the state
void mg_refine_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
// one FP64 correction x <- x + e, e from `cycles` FP32 V-cycles on A e = r (the first from e = 0), then the residual;
// form_rhs: r is formed first (the previous residual kept only its norms); keep_rhs: the closing residual writes
// the next correction's right-hand side; close = false: no closing residual (the caller forms its norms, e.g. by
// launch_project_resid)
void mg_refine_iter(GpuMG& mg, hipStream_t s, double tol, int max_cycles, int cycles = 1, bool form_rhs = false,
                    bool keep_rhs = true, bool close = true);
// reduces a residual whose partials are still pending, with the solve's stopping test (a one-block
// kernel that exits at once when none is)
// With `host` (pinned host memory, device-accessible) the kernel also stores {used, done} there.  gate: the
// kernel exits at once while *gate; set_hold: it then sets st->hold = !done (a step's closing test).
void mg_state_ready(GpuMG& mg, hipStream_t s, double tol, int max_cycles, int* host = nullptr, const int* gate = nullptr,
                    bool set_hold = false);
void mg_tol32_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
void mg_tol32_iter(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
// relative FP64 residual of the FP32 level-0 p against level-0 b -> st->rel[0]
void mg_residual_of_p(GpuMG& mg, hipStream_t s);
// ---- exact separable inner solve (fs_spectral.hip) ----------------------------------
struct Spectral;
// transform bases, eigenvalues and a rocBLAS handle for an nx x ny x nz level 0 of cell size h
// (nx even); runs one eager solve on scratch_b / scratch_p (level-0 sized) to load the library
int  spectral_build(Spectral** out, int nx, int ny, int nz, double h, hipStream_t s, float* scratch_b, float* scratch_p,
                    std::string& err);
void spectral_free(Spectral* sp);
// p = A^-1 b on the parity-split level-0 layout (FP32 transforms; b is kept)
int  spectral_solve(Spectral* sp, hipStream_t s, float* b, float* p);
// FP64 defect correction with the spectral inner solve: begin = residual of x, iter = one correction
void mg_spectral_begin(GpuMG& mg, hipStream_t s, double tol, int max_cycles);
int  mg_spectral_iter(GpuMG& mg, Spectral* sp, hipStream_t s, double tol, int max_cycles);
// single-operator launches (tests)
void mg_smooth(GpuMG& mg, hipStream_t s, int level, int iters, const int* skip, double* xupd = nullptr,
               bool first_red_done = false, const FinArgs* fin = nullptr, float* pl = nullptr);

// ---- error handling ----------------------------------------------------------
void fs_set_error(const std::string& msg);
#define FS_HIP_CHECK(call)                                                         \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            fs_set_error(std::string(#call) + ": " + hipGetErrorString(e_));       \
            return FS_EDEVICE;                                                     \
        }                                                                          \
    } while (0)

// ---- field kernels (fs_fields.hip) -------------------------------------------
Dims make_dims(int nx, int ny, int nz);
struct FieldSet {
    void* vel; void* velX; void* velY;
    void* S; void* SX; void* SY;
    void* S2; void* S2X; void* S2Y;
    void* vort; void* geo;
    const int* hold = nullptr;   // set: the kernels exit at once while *hold (a step enqueued ahead of its predecessor's test)
    const half4_t* noise; const half4_t* curl; int noise_n;
    const float* logs;
    const fs_spot* spots;        // FS_MAX_SPOTS entries
    const float* face_flux;      // nx*nz
    // One word per cell from the first advection pass to the second: the limiter's corner cell and the
    // outside test of the backtrace (k_adv1 writes it, k_adv2 reads it; null: k_adv2 recomputes both)
    unsigned* btpack = nullptr;
};
// The backtrace's limiter corner floor(b - 1/2) per axis, clamped to [-1, n] (the texels it selects are
// the same as for the unclamped value), plus one, in 10 bits each; bit 30: the backtrace leaves the
// grid (outside_fade < 1).  Needs every dimension <= 1022.
constexpr int FS_BTPACK_MAX_DIM = 1022;
void launch_geo(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F);
void launch_adv1(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F);
void launch_adv2(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F);
void launch_curl(int prec, const Dims& g, hipStream_t s, const FieldSet& F);
void launch_forces(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F);
void launch_div(int prec, const Dims& g, hipStream_t s, DevParams* P, const FieldSet& F, float* b_out, int advance_clock);
void launch_project(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F,
                    const float* p32, const double* p64);
// the projection with the FP64 refinement's iterate, fused with the closing residual's norms (mg.partials,
// pending for mg_state_ready; `cycles`: the V-cycles of the correction it closes)
void launch_project_resid(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, GpuMG& mg, int cycles);
void launch_noise_bake(hipStream_t s, half4_t* noise, half4_t* curl, int n, int seed);
void launch_set_params(hipStream_t s, DevParams* dst, const DevParams& src);
void launch_fill_zero(hipStream_t s, void* ptr, size_t bytes);
void launch_convert(hipStream_t s, int kind, const void* src, void* dst, size_t n);
void launch_copy_probe(hipStream_t s, const void* src, void* dst, size_t bytes);
// kind 1 tuned copy, 2 read only, 3 write only (fs_bench_stream)
void launch_stream_probe(hipStream_t s, int kind, const void* src, void* dst, size_t bytes);
// ---- adjoint of one step (fs_adjoint.hip), float4 storage ------------------------
// Stage inputs are the forward's recorded fields; l* arguments are adjoint fields.
// Gathers store (lvIn of project, lvY/lsY/ls2Y and the xyz of lvort in forces); every
// other output is accumulated with atomics and must be cleared first.
void launch_adv1_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* lvX, const float4* lsX, const float4* ls2X, float4* lvel,
                     float4* ls, float4* ls2);
void launch_adv2_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* vH, const float4* sH, const float4* s2H, const float4* lvY,
                     const float4* lsY, const float4* ls2Y, float4* lvel, float4* ls, float4* ls2, float4* lvH, float4* lsH,
                     float4* ls2H);
void launch_curl_adj(const Dims& g, hipStream_t s, const float4* v, const float4* lvort, float4* lv);
void launch_forces_adj(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float4* velY,
                       const float4* sY, const float4* s2Y, const float4* vort, const float4* lvOut, const float4* lsOut,
                       const float4* ls2Out, float4* lvY, float4* lsY, float4* ls2Y, float4* lvort, float4* llap,
                       bool visc, double* gspot, double* gface);
void launch_div_adj(const Dims& g, hipStream_t s, const DevParams* P, const double* lb, float4* lvF);
void launch_project_adj(const Dims& g, hipStream_t s, const DevParams* P, const float4* vIn, const double* p,
                        const float4* geo, const float4* lvOut, float4* lvIn, float* lp);
void launch_loss_seed(const Dims& g, hipStream_t s, const float4* Sf, const float* tT, const float* tSoot, float4* lS,
                      double* partial);
// tangent-linear counterparts (direction in, direction out; every output is stored)
void launch_adv1_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* tvel, const float4* tS, const float4* tS2, float4* tvX, float4* tsX,
                     float4* ts2X);
void launch_adv2_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vel, const float4* sF,
                     const float4* s2F, const float4* vH, const float4* sH, const float4* s2H, const float4* tvel,
                     const float4* tS, const float4* tS2, const float4* tvH, const float4* tsH, const float4* ts2H,
                     float4* tvY, float4* tsY, float4* ts2Y);
void launch_curl_tan(const Dims& g, hipStream_t s, const float4* v, const float4* tv, float4* tvort);
void launch_forces_tan(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float4* velY,
                       const float4* sY, const float4* s2Y, const float4* vort, const float* spot_dir,
                       const float* face_dir, const float4* tvY, const float4* tsY, const float4* ts2Y,
                       const float4* tvort, float4* tvF, float4* tS, float4* tS2);
void launch_div_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* tv, float* tb);
void launch_project_tan(const Dims& g, hipStream_t s, const DevParams* P, const float4* vIn, const double* p,
                        const float4* geo, const float4* tvIn, const double* tp, float4* tvOut);
// ---- sparks (fs_sparks.hip) --------------------------------------------------------
// The update kernel's uniform block: SPARK_FIELDS of src/sparks.js:10-14, all f32 as the
// reference uploads them.
struct SparkUniform {
    float dt, time, frame, count, nBase, nBurst, burstX, burstY, burstZ, burstSpeed, burstRadius;
    float ox, oy, oz, sx, sy, sz, turb, lifeK, bedRadius, seed, sizeK, upK, burstFan, burstSize, aimX, aimZ;
};
int  spark_blocks(int count);
// dead particles per block (after a host write or reset)
void launch_spark_count(hipStream_t s, const float4* sp, int count, int* dead);
// scan `dead` into `off`, update (sp: count x 4 float4) and write the blocks' dead counts
// after the update back into `dead` (both arrays spark_blocks(count) ints)
void launch_spark_step(int prec, hipStream_t s, const SparkUniform& U, float4* sp, int* dead, int* off, const void* vel,
                       const Dims& g, const half4_t* curl, int nn);
// ---- render-volume consumers (fs_render.hip) --------------------------------------
// The renderer's uniform block: RP_FIELDS of src/renderer.js:8-17, all f32 as uploaded.
enum {
    RP_width, RP_height, RP_vw, RP_vh, RP_time, RP_frame, RP_nLogs, RP_nStones,
    RP_camX, RP_camY, RP_camZ, RP_tanFov, RP_fwdX, RP_fwdY, RP_fwdZ, RP_aspect, RP_rgtX, RP_rgtY, RP_rgtZ, RP_jitX,
    RP_upX, RP_upY, RP_upZ, RP_jitY,
    RP_ox, RP_oy, RP_oz, RP_dx, RP_sx, RP_sy, RP_sz, RP_tempScale, RP_nx, RP_ny, RP_nz, RP_nLights,
    RP_exposure, RP_stepMul, RP_sootAbs, RP_sootEmit, RP_emitK, RP_blueK, RP_smokeScat, RP_smokeG,
    RP_ambient, RP_detailAmt, RP_detailFreq, RP_detailSpeed, RP_hazeK, RP_glowK, RP_bloomK, RP_tonemap,
    RP_grain, RP_vignette, RP_saturation, RP_contrast, RP_bedGlow, RP_logGlow, RP_lightK, RP_accumN,
    RP_bbTmin, RP_bbTmax, RP_bbN, RP_emberFlicker, RP_sootEmitMin, RP_flameSharp, RP_moon, RP_fogK,
    RP_shadowSoft, RP_lift, RP_emitTempLo, RP_glowTemp, RP_inset, RP_knee, RP_surfDetail, RP_matA, RP_matB, RP_matC,
    RP_sootScat, RP_chroma, RP_lensX, RP_lensY, RP_focus, RP_lightDt, RP_flickerK, RP_cubic, RP_aperture, RP_sootEdge,
    RP_COUNT
};
static_assert(RP_COUNT == FS_RP_COUNT, "uniform block layout");
struct RenderU { float f[RP_COUNT]; };
constexpr int FS_NLIGHTS = 8;
// light extraction: one block per height band; S is the sim's field (its storage), lut 2 x bbN
// float4 texels, lstate 3 x 8 float4 of filter state, lights 8 x 2 float4 (pos, col)
void launch_lights(int prec, hipStream_t s, const RenderU& R, const Dims& g, const void* S_field, const float4* lut,
                   float4* lstate, float4* lights);
// flame-field prefilter of S into a binary16 copy (same cell order)
void launch_sharpen(int prec, hipStream_t s, float k, const Dims& g, const void* S_field, half4_t* out);
// volume march: S_field in the sim's storage, or the binary16 prefiltered copy when `sharpened`;
// depth is dw x dh floats; vol is vw x vh binary16 (L, Tr), aux vw x vh float2 (haze, depth)
// empty-space map: raw / flags are bx*by*bz bytes (bricks of 4^3 cells); a brick is flagged
// when a cell within rb bricks may make a march sample contribute
void launch_occupancy(int prec, bool sharpened, hipStream_t s, const Dims& g, const void* S_field, const void* S2_field,
                      bool smoke_on, unsigned char* raw, unsigned char* flags, int bx, int by, int bz, int rb);
// flags null: march every sample
void launch_volume(int prec, bool sharpened, hipStream_t s, const RenderU& R, const Dims& g, const void* S_field,
                   const void* S2_field, const float4* lut, const float4* lights, const float* depth, int dw, int dh,
                   const half4_t* curl, int nn, const unsigned char* flags, int bx, int by, int bz, float inv_dx,
                   half4_t* vol, float2* aux);
// ---- FP64 mass and energy budget (fs_budget.hip) ----------------------------------
size_t budget_partials(const Dims& g);   // doubles of the partials buffer
// the source kernel in its tracing form (same fields) writing per-block budget partials
void launch_forces_budget(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, double* part);
// after the divergence: reduce the partials (and sum(b)) into the record
void launch_budget_step(const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, const float* b, double* part,
                        double* rec);
// start the record from the current state
void launch_budget_init(int prec, const Dims& g, hipStream_t s, const DevParams* P, const FieldSet& F, double* part,
                        double* rec);
// ---- diagnostics (fs_diag.hip) -------------------------------------------------
int  stats_blocks(size_t n);
void launch_stats(int prec, hipStream_t s, const void* field, size_t n, double* part);
// fs_stats through the library's fixed device scratch (no allocation); FS_EINVAL above
// FS_STATS_NOALLOC_MAX_CELLS cells
int  stats_noalloc(int prec, hipStream_t s, int device, const void* field, size_t n, double out[20]);
void launch_planes(int prec, hipStream_t s, const void* S_field, int nx, int ny, int nz, double V, double* out);
void launch_record(hipStream_t s, const DevParams* P, const double* planes, int ny, double dx, double t_flame,
                   double* ring, int cap, int* count);
