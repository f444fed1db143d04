// C ABI of the gfx950 fire-simulation step (include/firesim.h).
//
// One fs_sim owns its device fields, the multigrid hierarchy, the noise textures
// and a device-resident parameter block.  A step is a fixed sequence of launches
// on the handle's stream (no host round trips, no allocation), optionally
// replayed from a captured HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fs_internal.h"

// ---- errors ------------------------------------------------------------------
static thread_local std::string g_last_error;
void fs_set_error(const std::string& msg) { g_last_error = msg; }
static int fail(int code, const std::string& msg) { g_last_error = msg; return code; }

// ---- SIM_FIELDS / SIM_DEFAULTS (src/sim.js:14-44) ------------------------------
static const char* const kParamNames[FS_NUM_PARAMS] = {
    "nx", "ny", "nz", "dx", "dt", "time", "frame", "seed",
    "ox", "oy", "oz", "nLogs",
    "fuelSrc", "bedFuel", "srcTemp", "bedTemp", "srcVel", "srcFreq", "srcSpeed", "srcThresh",
    "ignT", "kMix", "rs", "heatOut", "expansion", "sootForm", "sootOx", "radCool",
    "coolRate", "sootDissip", "smokeDissip", "vitDecay", "fuelDissip", "buoyancy", "vortEps", "turbAmp",
    "turbFreq", "turbRise", "turbEvolve", "windX", "windZ", "gust", "intensity", "velDissip",
    "bedRadius", "logSrc", "burstX", "burstY", "burstZ", "burstAmt", "smolder", "sponge", "hotR0", "hotR1", "hotY1",
    "visc", "lamAge", "anchorFreq", "srcFloor", "anchorGain",
};
static void default_params(float* p) {
    std::fill(p, p + FS_NUM_PARAMS, 0.f);
    p[P_fuelSrc] = 0.13f; p[P_bedFuel] = 0.16f; p[P_srcTemp] = 0.75f; p[P_bedTemp] = 0.82f; p[P_srcVel] = 0.25f;
    p[P_srcFreq] = 35.0f; p[P_srcSpeed] = 0.06f; p[P_srcThresh] = 0.0f;
    p[P_ignT] = 550.f; p[P_kMix] = 30.f; p[P_rs] = 4.7f; p[P_heatOut] = 7700.f; p[P_expansion] = 0.7f;
    p[P_sootForm] = 25.f; p[P_sootOx] = 30.f; p[P_radCool] = 0.25f;
    p[P_coolRate] = 3.0f; p[P_sootDissip] = 0.4f; p[P_smokeDissip] = 0.3f; p[P_vitDecay] = 1.5f; p[P_fuelDissip] = 0.3f;
    p[P_buoyancy] = 26.8f;
    p[P_vortEps] = 8.f; p[P_turbAmp] = 0.8f; p[P_turbFreq] = 1.5f; p[P_turbRise] = 1.5f; p[P_turbEvolve] = 0.6f;
    p[P_windX] = 0.f; p[P_windZ] = 0.f; p[P_gust] = 0.f; p[P_intensity] = 1.f; p[P_velDissip] = 0.05f;
    p[P_bedRadius] = 0.22f; p[P_hotR0] = 0.12f; p[P_hotR1] = 0.45f; p[P_hotY1] = 0.5f; p[P_logSrc] = 1.f;
    p[P_burstX] = 0.f; p[P_burstY] = 0.f; p[P_burstZ] = 0.f; p[P_burstAmt] = 0.f; p[P_smolder] = 0.02f; p[P_sponge] = 1.f;
    p[P_visc] = 0.f; p[P_lamAge] = 0.f; p[P_anchorFreq] = 1.6f; p[P_srcFloor] = 0.25f; p[P_anchorGain] = 0.95f;
}

// ---- the handle ------------------------------------------------------------------
namespace {
enum { F_VEL, F_VELX, F_VELY, F_S, F_SX, F_SY, F_S2, F_S2X, F_S2Y, F_VORT, F_GEO, F_COUNT };
int field_slot(int field) {
    switch (field) {
        case FS_VEL: return F_VEL; case FS_S: return F_S; case FS_S2: return F_S2; case FS_VORT: return F_VORT;
        case FS_GEO: return F_GEO; case FS_VELX: return F_VELX; case FS_SX: return F_SX; case FS_S2X: return F_S2X;
        case FS_VELY: return F_VELY; case FS_SY: return F_SY; case FS_S2Y: return F_S2Y;
    }
    return -1;
}
size_t texel_bytes(int prec) { return prec == FS_PREC_FP32 ? 16 : 8; }
}  // namespace

struct fs_sim {
    fs_config cfg;
    int nx, ny, nz;
    size_t ncell;
    double dx;
    double origin[3];
    int prec;
    Dims g;
    void* f[F_COUNT] = {};
    half4_t* noise = nullptr;
    half4_t* curl = nullptr;
    float* logs_d = nullptr;
    int logs_cap = 0;
    fs_spot* spots_d = nullptr;
    std::vector<fs_spot> spots_h;
    float* face_d = nullptr;
    bool has_face = false;
    unsigned* btpack_d = nullptr;    // the advection passes' per-cell limiter corner (null: grids above 1022)
    std::vector<float> logs_h;
    DevParams* P_d = nullptr;
    float prm[FS_NUM_PARAMS];
    bool params_dirty = true;
    double dev_dt = -1.0;
    GpuMG mg;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // captured graphs: the whole step (fixed solves); the whole step of an adaptive solve with the number of
    // iterations the last step needed (step_k) and the readback of the solve's state; its three parts
    // The refinement's adaptive step graph (rexec below) closes its last correction with the projection itself
    // (the fused projection + residual norms, then the stop state for the host); an extra correction forms its
    // own right-hand side (G_ITER_NC) and is closed the same way (pexec below).
    enum { G_FULL, G_STEP, G_PRE, G_FIRST, G_ITER, G_POST, G_ITER_NC, G_COUNT };
    int step_k = 0;
    hipGraphExec_t gexec[G_COUNT] = {};
    hipGraph_t graph[G_COUNT] = {};
    int predict = 1;                 // V-cycles the last adaptive solve needed
    int* host_state = nullptr;       // pinned: {used, done} of the last solve (the refinement: two slots, by step parity)
    int* host_state_dev = nullptr;   // its device address
    // The refinement's steps of one call run one ahead: a step's graph (per V-cycle count and host-state slot) is
    // enqueued behind the previous step before the host has read that step's stop state; its kernels exit at once
    // while that state asks for another correction (st->hold), and the host then enqueues the step again.
    hipGraphExec_t rexec[MG_MAX_REFINE + 1][2] = {};
    hipGraph_t rgraph[MG_MAX_REFINE + 1][2] = {};
    hipGraphExec_t pexec[2] = {};    // a further correction's projection and stop state, per slot
    hipGraph_t pgraph[2] = {};
    hipEvent_t ev_slot[2] = {};
    int slot = 0;                    // host-state slot of the next step
    int cur_slot = 0;                // the slot of the step being recorded
    bool inflight = false;           // the next step is enqueued already
    int inflight_k = 0;              // with this many V-cycles
    bool gated = false;              // recording a step that may run ahead: its kernels test st->hold
    bool geo_ready = false;
    double time = 0.0;
    int64_t frame = 0;
    bool profile = false;
    hipEvent_t ev[8] = {};
    hipEvent_t ev_state = nullptr;   // the stop-flag readback of an adaptive step has landed
    bool timings_valid = false;
    // diagnostics: per-plane partials, stats partials, the per-step record ring
    double* planes_d = nullptr;      // ny * 8
    double* stats_d = nullptr;       // stats blocks * 20
    double* ring_d = nullptr;        // diag_cap * (4 + ny)
    int* ring_count_d = nullptr;
    int diag_cap = 0;
    double diag_tflame = 1000.0;
    // FP64 budget: per-block partials and the device record
    bool budget_on = false;
    double* budget_part = nullptr;
    double* budget_rec = nullptr;
    uint64_t gen = 0;                // bumped by every call that may change the fields

    FieldSet fields() const {
        FieldSet F;
        F.vel = f[F_VEL]; F.velX = f[F_VELX]; F.velY = f[F_VELY];
        F.S = f[F_S]; F.SX = f[F_SX]; F.SY = f[F_SY];
        F.S2 = f[F_S2]; F.S2X = f[F_S2X]; F.S2Y = f[F_S2Y];
        F.vort = f[F_VORT]; F.geo = f[F_GEO];
        F.hold = gated ? &mg.st->hold : nullptr;
        F.noise = noise; F.curl = curl; F.noise_n = cfg.noise_n;
        F.logs = logs_d;
        F.spots = spots_d;
        F.face_flux = face_d;
        F.btpack = btpack_d;
        return F;
    }
    Spectral* spec = nullptr;        // FS_SOLVE_SPECTRAL64: the exact separable inner solve
    bool refine() const { return cfg.solve == FS_SOLVE_REFINE64 || cfg.solve == FS_SOLVE_SPECTRAL64; }
    bool tracked() const { return cfg.solve != FS_SOLVE_VCYCLES; }
    bool adaptive() const { return tracked() && cfg.adaptive; }
    void solve_begin(hipStream_t s) {
        if (cfg.solve == FS_SOLVE_SPECTRAL64) mg_spectral_begin(mg, s, cfg.div_tol, cfg.max_vcycles);
        else if (cfg.solve == FS_SOLVE_REFINE64) mg_refine_begin(mg, s, cfg.div_tol, cfg.max_vcycles);
        else mg_tol32_begin(mg, s, cfg.div_tol, cfg.max_vcycles);
    }
    // V-cycles of a step's first correction (FS_SOLVE_REFINE64; the other solves take one per iteration)
    int first_cycles() const {
        return cfg.solve == FS_SOLVE_REFINE64 ? std::max(1, std::min(mg.first, std::min(cfg.max_vcycles, MG_MAX_REFINE))) : 1;
    }
    // one iteration of the solve; a refinement's correction takes `cycles` V-cycles (form_rhs / keep_rhs: see
    // mg_refine_iter)
    int solve_iter(hipStream_t s, int cycles = 1, bool form_rhs = false, bool keep_rhs = true, bool close = true) {
        if (cfg.solve == FS_SOLVE_SPECTRAL64) return mg_spectral_iter(mg, spec, s, cfg.div_tol, cfg.max_vcycles);
        if (cfg.solve == FS_SOLVE_REFINE64)
            mg_refine_iter(mg, s, cfg.div_tol, cfg.max_vcycles, cycles, form_rhs, keep_rhs, close);
        else mg_tol32_iter(mg, s, cfg.div_tol, cfg.max_vcycles);
        return FS_OK;
    }
    // the pressure solve of one step (sim.js:642); begun: its starting residual is already formed
    int solve(hipStream_t s, bool begun = false) {
        if (cfg.solve == FS_SOLVE_SPECTRAL64) {   // corrections enqueued while the device's stop flag is clear
            if (!begun) mg_spectral_begin(mg, s, cfg.div_tol, cfg.max_vcycles);
            for (int k = 0; k < std::min(cfg.max_vcycles, MG_MAX_REFINE); k++) {
                FS_HIP_CHECK(hipMemcpyAsync(host_state, &mg.st->used, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
                FS_HIP_CHECK(hipStreamSynchronize(s));
                if (host_state[1]) break;
                if (int rc = mg_spectral_iter(mg, spec, s, cfg.div_tol, cfg.max_vcycles)) return rc;
            }
        } else if (cfg.solve == FS_SOLVE_REFINE64) {   // the first correction, then one V-cycle per correction
            if (!begun) mg_refine_begin(mg, s, cfg.div_tol, cfg.max_vcycles);
            const int f = first_cycles();
            mg_refine_iter(mg, s, cfg.div_tol, cfg.max_vcycles, f);
            for (int k = f; k < std::min(cfg.max_vcycles, MG_MAX_REFINE); k++) mg_refine_iter(mg, s, cfg.div_tol, cfg.max_vcycles);
            mg_state_ready(mg, s, cfg.div_tol, cfg.max_vcycles);
        }
        else if (cfg.solve == FS_SOLVE_TOL32) mg_solve_tol32(mg, s, cfg.div_tol, cfg.max_vcycles);
        else for (int i = 0; i < cfg.vcycles; i++) mg_vcycle(mg, s, nullptr);
        return FS_OK;
    }
};

namespace {

int set_device(const fs_sim* h) {
    FS_HIP_CHECK(hipSetDevice(h->cfg.device));
    return FS_OK;
}

// host parameter block -> device (stream ordered; the block travels as a kernel argument)
void push_params(fs_sim* h, double dt) {
    DevParams d;
    std::memset(&d, 0, sizeof(d));
    for (int i = 0; i < FS_NUM_PARAMS; i++) d.f[i] = h->prm[i];
    // grid and clock entries are owned by the handle (sim.js:620-624)
    d.f[P_nx] = (float)h->nx; d.f[P_ny] = (float)h->ny; d.f[P_nz] = (float)h->nz;
    d.f[P_dx] = (float)h->dx; d.f[P_dt] = (float)dt; d.f[P_time] = (float)h->time; d.f[P_frame] = (float)h->frame;
    d.f[P_ox] = (float)h->origin[0]; d.f[P_oy] = (float)h->origin[1]; d.f[P_oz] = (float)h->origin[2];
    d.f[P_nLogs] = (float)(h->logs_h.size() / 8);
    d.time = h->time;
    d.dt = dt;
    d.frame = h->frame;
    d.n_logs = (int)(h->logs_h.size() / 8);
    d.n_spots = (int)h->spots_h.size();
    d.has_face_flux = h->has_face ? 1 : 0;
    launch_set_params(h->stream, h->P_d, d);
    h->params_dirty = false;
    h->dev_dt = dt;
}
void sync_params(fs_sim* h, double dt) {
    if (h->params_dirty || dt != h->dev_dt) push_params(h, dt);
}

// the launches of one step after the geometry (sim.js:633-644), in three parts so a
// tolerance-driven solve can be continued from the host: pre = adv1 .. div (+ the
// solve's starting residual), iter = one V-cycle + residual, post = project
void enqueue_pre(fs_sim* h, bool prof, bool begin_solve) {
    const FieldSet F = h->fields();
    hipStream_t s = h->stream;
    if (prof) (void)hipEventRecord(h->ev[0], s);
    launch_adv1(h->prec, h->g, s, h->P_d, F);
    if (prof) (void)hipEventRecord(h->ev[1], s);
    launch_adv2(h->prec, h->g, s, h->P_d, F);
    if (prof) (void)hipEventRecord(h->ev[2], s);
    launch_curl(h->prec, h->g, s, F);
    if (prof) (void)hipEventRecord(h->ev[3], s);
    if (h->budget_on) launch_forces_budget(h->prec, h->g, s, h->P_d, F, h->budget_part);
    else launch_forces(h->prec, h->g, s, h->P_d, F);
    if (prof) (void)hipEventRecord(h->ev[4], s);
    float* const b_out = h->refine() ? h->mg.bkeep : h->mg.b[0];
    launch_div(h->prec, h->g, s, h->P_d, F, b_out, 1);
    if (h->budget_on) launch_budget_step(h->g, s, h->P_d, F, b_out, h->budget_part, h->budget_rec);
    if (prof) (void)hipEventRecord(h->ev[5], s);
    if (begin_solve) h->solve_begin(s);
}
void enqueue_post(fs_sim* h, bool prof) {
    const FieldSet F = h->fields();
    hipStream_t s = h->stream;
    if (prof) (void)hipEventRecord(h->ev[6], s);
    launch_project(h->prec, h->g, s, h->P_d, F, h->refine() ? nullptr : h->mg.p[0], h->refine() ? h->mg.x64 : nullptr);
    if (prof) (void)hipEventRecord(h->ev[7], s);
    if (h->diag_cap > 0) {           // the per-step record, after the step's last kernel
        launch_planes(h->prec, s, F.S, h->nx, h->ny, h->nz, std::pow((double)(float)h->dx, 3.0), h->planes_d);
        launch_record(s, h->P_d, h->planes_d, h->ny, (double)(float)h->dx, h->diag_tflame, h->ring_d, h->diag_cap,
                      h->ring_count_d);
    }
}
int enqueue_step(fs_sim* h, bool prof) {
    enqueue_pre(h, prof, false);
    if (int rc = h->solve(h->stream)) return rc;
    enqueue_post(h, prof);
    return FS_OK;
}

// the projection fused with the closing residual of the refinement's last correction (of `cycles` V-cycles),
// then the stop state for the host
void enqueue_post_resid(fs_sim* h, int cycles) {
    launch_project_resid(h->prec, h->g, h->stream, h->P_d, h->fields(), h->mg, cycles);
    mg_state_ready(h->mg, h->stream, h->cfg.div_tol, std::min(h->cfg.max_vcycles, MG_MAX_REFINE),
                   h->host_state_dev + 2 * h->cur_slot, h->gated ? &h->mg.st->hold : nullptr, true);
}
// an adaptive step with step_k V-cycles (the first correction's, then one per correction) and the stop-flag
// readback (the refinement's step ends with its projection; the other solves' projection follows as its own
// part; the host checks the flag afterwards)
int enqueue_adaptive_step(fs_sim* h) {
    const int f = h->first_cycles();
    if (h->cfg.solve == FS_SOLVE_REFINE64) {   // recorded held on st->hold (it may run ahead of its predecessor's test)
        h->gated = true;
        h->mg.gate = &h->mg.st->hold;
        enqueue_pre(h, false, true);
        int rc = h->solve_iter(h->stream, f, false, h->step_k > f, h->step_k > f);
        for (int k = f; k < h->step_k && !rc; k++) rc = h->solve_iter(h->stream, 1, false, k + 1 < h->step_k, k + 1 < h->step_k);
        if (!rc) enqueue_post_resid(h, h->step_k > f ? 1 : f);
        h->gated = false;
        h->mg.gate = nullptr;
        return rc;
    }
    enqueue_pre(h, false, true);
    if (int rc = h->solve_iter(h->stream, f)) return rc;
    for (int k = f; k < h->step_k; k++)
        if (int rc = h->solve_iter(h->stream)) return rc;
    // the stopping state goes to the host by a copy
    FS_HIP_CHECK(hipMemcpyAsync(h->host_state, &h->mg.st->used, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    return FS_OK;
}

// Records `enqueue` into (gr, ex) on first use and launches it; without graphs, enqueues it directly.
template <class Enq>
int run_graph(fs_sim* h, hipGraph_t& gr, hipGraphExec_t& ex, Enq enqueue) {
    if (!h->cfg.use_graph) return enqueue();
    if (!ex) {
        FS_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue();
        hipGraph_t g = nullptr;
        if (rc) {   // end the capture before reporting the failure
            (void)hipStreamEndCapture(h->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        FS_HIP_CHECK(hipStreamEndCapture(h->stream, &g));
        gr = g;
        FS_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    }
    FS_HIP_CHECK(hipGraphLaunch(ex, h->stream));
    return FS_OK;
}
// the refinement's step with k V-cycles, its stop state into host slot `slot`, held while st->hold
int run_refine_step(fs_sim* h, int k, int slot) {
    return run_graph(h, h->rgraph[k][slot], h->rexec[k][slot], [&] {
        h->step_k = k;
        h->cur_slot = slot;
        return enqueue_adaptive_step(h);
    });
}
// a further correction's projection with the closing residual and the stop state into slot `slot` (not held)
int run_refine_post(fs_sim* h, int slot) {
    return run_graph(h, h->pgraph[slot], h->pexec[slot], [&] {
        h->cur_slot = slot;
        enqueue_post_resid(h, 1);
        return FS_OK;
    });
}

int capture(fs_sim* h, int which) {
    if (h->gexec[which]) return FS_OK;
    FS_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = FS_OK;
    switch (which) {
        case fs_sim::G_FULL: rc = enqueue_step(h, false); break;
        case fs_sim::G_STEP: rc = enqueue_adaptive_step(h); break;
        case fs_sim::G_PRE: enqueue_pre(h, false, true); break;
        case fs_sim::G_FIRST: rc = h->solve_iter(h->stream, h->first_cycles()); break;
        case fs_sim::G_ITER: rc = h->solve_iter(h->stream); break;
        case fs_sim::G_POST: enqueue_post(h, false); break;
        case fs_sim::G_ITER_NC: rc = h->solve_iter(h->stream, 1, true, false, false); break;
    }
    hipGraph_t gr = nullptr;
    if (rc) {   // end the capture before reporting the failure
        (void)hipStreamEndCapture(h->stream, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    FS_HIP_CHECK(hipStreamEndCapture(h->stream, &gr));
    h->graph[which] = gr;
    FS_HIP_CHECK(hipGraphInstantiate(&h->gexec[which], gr, nullptr, nullptr, 0));
    return FS_OK;
}
int run_part(fs_sim* h, int which, bool prof) {
    if (h->cfg.use_graph && !prof) {
        int rc = capture(h, which);
        if (rc) return rc;
        FS_HIP_CHECK(hipGraphLaunch(h->gexec[which], h->stream));
        return FS_OK;
    }
    switch (which) {
        case fs_sim::G_FULL: return enqueue_step(h, prof);
        case fs_sim::G_STEP: return enqueue_adaptive_step(h);
        case fs_sim::G_PRE: enqueue_pre(h, prof, true); break;
        case fs_sim::G_FIRST: return h->solve_iter(h->stream, h->first_cycles());
        case fs_sim::G_ITER: return h->solve_iter(h->stream);
        case fs_sim::G_POST: enqueue_post(h, prof); break;
        case fs_sim::G_ITER_NC: return h->solve_iter(h->stream, 1, true, false, false);
    }
    return FS_OK;
}
void drop_graphs(fs_sim* h) {
    for (int i = 0; i < fs_sim::G_COUNT; i++) {
        if (h->gexec[i]) (void)hipGraphExecDestroy(h->gexec[i]);
        if (h->graph[i]) (void)hipGraphDestroy(h->graph[i]);
        h->gexec[i] = nullptr;
        h->graph[i] = nullptr;
    }
    for (int s = 0; s < 2; s++) {
        for (int k = 0; k <= MG_MAX_REFINE; k++) {
            if (h->rexec[k][s]) (void)hipGraphExecDestroy(h->rexec[k][s]);
            if (h->rgraph[k][s]) (void)hipGraphDestroy(h->rgraph[k][s]);
            h->rexec[k][s] = nullptr;
            h->rgraph[k][s] = nullptr;
        }
        if (h->pexec[s]) (void)hipGraphExecDestroy(h->pexec[s]);
        if (h->pgraph[s]) (void)hipGraphDestroy(h->pgraph[s]);
        h->pexec[s] = nullptr;
        h->pgraph[s] = nullptr;
    }
}

// One step with a tolerance-driven solve: enqueue as many iterations as the last step needed, the readback
// of the device's stopping state and the projection at once, before the host has read that state; with
// graphs all of it is one captured graph (per iteration count).  If the solve had not stopped yet, the
// host enqueues the further iterations and the projection again: the projection only reads the advected
// velocity (which nothing between the two runs writes) and the pressure, so its second run gives the state
// of the unspeculated order, bit for bit.  The FP64 refinement goes further: the projection itself forms the
// closing residual's norms (it reads the same iterate values) and the stop state follows it, and the next step
// of the same call is enqueued behind the step before the host reads that state: its kernels exit at once while
// the state asks for another correction (st->hold), in which case the host enqueues the correction and then the
// next step again.  Not speculated: profiled steps and steps that append a per-step diagnostics record (the
// projection part would record twice).  more: another step of the same call follows.
int adaptive_step(fs_sim* h, bool prof, bool more) {
    const int maxc = std::min(h->cfg.max_vcycles, MG_MAX_REFINE);
    const bool spec = !prof && h->diag_cap == 0;
    bool projected = false;
    int k = 0, rc = FS_OK;
    const int f = h->first_cycles();
    if (spec && h->cfg.solve == FS_SOLVE_REFINE64) {   // with graphs or one launch at a time
        const int kk = std::min(std::max(h->predict, f), maxc);
        const int s0 = h->slot;
        if (h->budget_on) more = false;                    // (the budget's kernels are not held)
        if (!h->inflight) {
            if ((rc = run_refine_step(h, kk, s0))) return rc;
            h->inflight_k = kk;
        }
        h->inflight = false;
        k = h->inflight_k;                                 // the V-cycles this step was enqueued with
        FS_HIP_CHECK(hipEventRecord(h->ev_slot[s0], h->stream));
        if (more) {                                        // the next step, held on the device until this one is done
            if ((rc = run_refine_step(h, kk, 1 - s0))) return rc;
            h->inflight = true;
            h->inflight_k = kk;
        }
        const int* hs = h->host_state + 2 * s0;
        FS_HIP_CHECK(hipEventSynchronize(h->ev_slot[s0]));
        bool corrected = false;
        while (!hs[1] && k < maxc) {
            // a correction more (forming its right-hand side), the projection and the stop state again, not held
            if ((rc = run_part(h, fs_sim::G_ITER_NC, prof))) return rc;
            if ((rc = run_refine_post(h, s0))) return rc;
            FS_HIP_CHECK(hipEventRecord(h->ev_slot[s0], h->stream));
            FS_HIP_CHECK(hipEventSynchronize(h->ev_slot[s0]));
            k++;
            corrected = true;
        }
        if (corrected) h->inflight = false;                // the next step ran held: it is enqueued again
        if (!hs[1]) {   // (not reached: a solve at its V-cycle cap is done) release the hold all the same
            launch_fill_zero(h->stream, &h->mg.st->hold, 2 * sizeof(int));
            h->inflight = false;
        }
        h->predict = std::max(1, hs[0]);
        h->slot = 1 - s0;
        return FS_OK;
    }
    if (spec && h->cfg.use_graph) {
        const int kk = std::min(std::max(h->predict, f), maxc);
        if (h->step_k != kk && h->gexec[fs_sim::G_STEP]) {
            (void)hipGraphExecDestroy(h->gexec[fs_sim::G_STEP]);
            (void)hipGraphDestroy(h->graph[fs_sim::G_STEP]);
            h->gexec[fs_sim::G_STEP] = nullptr;
            h->graph[fs_sim::G_STEP] = nullptr;
        }
        h->step_k = kk;
        if ((rc = run_part(h, fs_sim::G_STEP, prof))) return rc;
        k = kk;
        while (true) {   // the state after the step's corrections, the projection enqueued behind it
            FS_HIP_CHECK(hipEventRecord(h->ev_state, h->stream));
            if ((rc = run_part(h, fs_sim::G_POST, prof))) return rc;
            FS_HIP_CHECK(hipEventSynchronize(h->ev_state));
            if (h->host_state[1] || k >= maxc) break;
            // one more iteration, its state, and the projection again
            if ((rc = run_part(h, fs_sim::G_ITER, prof))) return rc;
            k++;
            FS_HIP_CHECK(hipMemcpyAsync(h->host_state, &h->mg.st->used, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        }
        h->predict = std::max(1, h->host_state[0]);
        return FS_OK;
    } else {
        if ((rc = run_part(h, fs_sim::G_PRE, prof))) return rc;
        if ((rc = run_part(h, fs_sim::G_FIRST, prof))) return rc;
        for (k = f; k < std::min(h->predict, maxc); k++) if ((rc = run_part(h, fs_sim::G_ITER, prof))) return rc;
    }
    while (true) {
        // the last residual's reduction (the refinement leaves it to the next iteration's first kernel)
        if (h->cfg.solve == FS_SOLVE_REFINE64) mg_state_ready(h->mg, h->stream, h->cfg.div_tol, maxc);
        FS_HIP_CHECK(hipMemcpyAsync(h->host_state, &h->mg.st->used, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        FS_HIP_CHECK(hipEventRecord(h->ev_state, h->stream));
        if (spec) {
            if ((rc = run_part(h, fs_sim::G_POST, prof))) return rc;
            projected = true;
        }
        FS_HIP_CHECK(hipEventSynchronize(h->ev_state));
        if (h->host_state[1] || k >= maxc) break;
        projected = false;
        if ((rc = run_part(h, fs_sim::G_ITER, prof))) return rc;
        k++;
    }
    h->predict = std::max(1, h->host_state[0]);
    return projected ? FS_OK : run_part(h, fs_sim::G_POST, prof);
}

void free_sim(fs_sim* h) {
    drop_graphs(h);
    if (h->host_state) (void)hipHostFree(h->host_state);
    for (auto& p : h->f) if (p) (void)hipFree(p);
    if (h->noise) (void)hipFree(h->noise);
    if (h->curl) (void)hipFree(h->curl);
    if (h->logs_d) (void)hipFree(h->logs_d);
    if (h->spots_d) (void)hipFree(h->spots_d);
    if (h->face_d) (void)hipFree(h->face_d);
    if (h->btpack_d) (void)hipFree(h->btpack_d);
    if (h->P_d) (void)hipFree(h->P_d);
    if (h->planes_d) (void)hipFree(h->planes_d);
    if (h->stats_d) (void)hipFree(h->stats_d);
    if (h->ring_d) (void)hipFree(h->ring_d);
    if (h->ring_count_d) (void)hipFree(h->ring_count_d);
    if (h->budget_part) (void)hipFree(h->budget_part);
    if (h->budget_rec) (void)hipFree(h->budget_rec);
    mg_free(h->mg);
    spectral_free(h->spec);
    for (auto& e : h->ev) if (e) (void)hipEventDestroy(e);
    if (h->ev_state) (void)hipEventDestroy(h->ev_state);
    for (auto& e : h->ev_slot) if (e) (void)hipEventDestroy(e);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int upload_logs(fs_sim* h, const float* logs, int n_logs) {
    h->logs_h.assign(logs, logs + (size_t)n_logs * 8);
    const int need = std::max(n_logs, 1) * 8;
    if (need > h->logs_cap) {
        if (h->logs_d) { FS_HIP_CHECK(hipStreamSynchronize(h->stream)); (void)hipFree(h->logs_d); h->logs_d = nullptr; }
        FS_HIP_CHECK(hipMalloc(&h->logs_d, need * sizeof(float)));
        h->logs_cap = need;
        drop_graphs(h);
    }
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (n_logs > 0) FS_HIP_CHECK(hipMemcpy(h->logs_d, logs, (size_t)n_logs * 8 * sizeof(float), hipMemcpyHostToDevice));
    h->geo_ready = false;
    h->params_dirty = true;
    return FS_OK;
}

// float <-> storage conversion kernels for field transfers
__global__ void k_pack_f16(const float* __restrict__ src, half4_t* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    half4_t h;
    h.x = (_Float16)src[4 * i]; h.y = (_Float16)src[4 * i + 1]; h.z = (_Float16)src[4 * i + 2]; h.w = (_Float16)src[4 * i + 3];
    dst[i] = h;
}
__global__ void k_unpack_f16(const half4_t* __restrict__ src, float* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const half4_t h = src[i];
    dst[4 * i] = (float)h.x; dst[4 * i + 1] = (float)h.y; dst[4 * i + 2] = (float)h.z; dst[4 * i + 3] = (float)h.w;
}
__device__ __forceinline__ unsigned short f2bf_io(float f) {
    unsigned x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return (unsigned short)(x >> 16);
}
__global__ void k_pack_bf16(const float* __restrict__ src, unsigned short* __restrict__ dst, size_t n4) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    dst[i] = f2bf_io(src[i]);
}
__global__ void k_unpack_bf16(const unsigned short* __restrict__ src, float* __restrict__ dst, size_t n4) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    dst[i] = __uint_as_float((unsigned)src[i] << 16);
}

struct DevField { void* ptr; size_t count; int channels; int kind; };   // kind: 0 f32, 1 f16, 2 bf16, 3 f64
bool resolve_field(fs_sim* h, int field, DevField& d) {
    const size_t n = h->ncell;
    const int sk = h->prec == FS_PREC_FP32 ? 0 : (h->prec == FS_PREC_FP16 ? 1 : 2);
    if (field == FS_NOISE || field == FS_CURL) {
        const size_t nn = (size_t)h->cfg.noise_n * h->cfg.noise_n * h->cfg.noise_n;
        d = {field == FS_NOISE ? (void*)h->noise : (void*)h->curl, nn, 4, 1};
        return true;
    }
    if (field == FS_P) { d = h->refine() ? DevField{h->mg.x64, n, 1, 3} : DevField{h->mg.p[0], n, 1, 0}; return true; }
    if (field == FS_B) { d = DevField{h->refine() ? (void*)h->mg.bkeep : (void*)h->mg.b[0], n, 1, 0}; return true; }
    const int slot = field_slot(field);
    if (slot < 0) return false;
    d = DevField{h->f[slot], n, 4, sk};
    return true;
}
size_t kind_bytes(int kind) { return kind == 0 ? 4 : kind == 3 ? 8 : 2; }

// the multigrid's parity-split order (rb_index) <-> x-fastest order
template <class T> void rb_to_natural(const T* src, T* dst, int nx, int ny, int nz) {
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) dst[((size_t)z * ny + y) * nx + x] = src[rb_index(x, y, z, nx, ny)];
}
template <class T> void natural_to_rb(const T* src, T* dst, int nx, int ny, int nz) {
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) dst[rb_index(x, y, z, nx, ny)] = src[((size_t)z * ny + y) * nx + x];
}
bool is_mg_field(int field) { return field == FS_P || field == FS_B; }

}  // namespace

extern "C" {

const char* fs_last_error(void) { return g_last_error.c_str(); }
const char* fs_version(void) { return "firesim-gfx950 1.0"; }

void fs_config_default(fs_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->res[0] = 80; c->res[1] = 120; c->res[2] = 80;          // src/sim.js:49
    c->size[0] = 1.2f; c->size[1] = 1.8f; c->size[2] = 1.2f;
    c->vcycles = 2;
    c->max_levels = 0;
    c->precision = FS_PREC_FP16;
    c->solve = FS_SOLVE_VCYCLES;
    c->div_tol = 1e-5;
    c->max_vcycles = 8;
    c->noise_n = 128;                                         // src/main.js:36
    c->noise_seed = 1;
    c->device = 0;
    c->stream = nullptr;
    c->use_graph = 0;
    c->adaptive = 1;
    c->omega = 0.f;                                           // tolerance solves: 1.25
    c->pre_sweeps = 0;                                        // tolerance solves: V(1,2)
    c->post_sweeps = 0;
}

int fs_param_count(void) { return FS_NUM_PARAMS; }
const char* fs_param_name(int i) { return (i >= 0 && i < FS_NUM_PARAMS) ? kParamNames[i] : nullptr; }
int fs_param_index(const char* name) {
    if (!name) return -1;
    for (int i = 0; i < FS_NUM_PARAMS; i++) if (std::strcmp(name, kParamNames[i]) == 0) return i;
    return -1;
}
int fs_param_defaults(float* out) {
    if (!out) return fail(FS_EINVAL, "null output");
    default_params(out);
    return FS_OK;
}

int fs_create(const fs_config* cfg, const float* logs, int n_logs, fs_sim** out) {
    if (!cfg || !out) return fail(FS_EINVAL, "null argument");
    *out = nullptr;
    const int nx = cfg->res[0], ny = cfg->res[1], nz = cfg->res[2];
    if (nx < 2 || ny < 2 || nz < 2) return fail(FS_EINVAL, "res must be >= 2 in every dimension");
    if ((long long)nx * ny * nz * 16 > (1LL << 31) * 8LL) return fail(FS_EINVAL, "grid too large");
    if ((long long)nx * ny * nz >= (1LL << 31) / 4) return fail(FS_EINVAL, "grid exceeds 32-bit texel indexing");
    if (cfg->precision < FS_PREC_FP32 || cfg->precision > FS_PREC_BF16) return fail(FS_EINVAL, "bad precision");
    if ((long long)nx * ny * nz * (long long)texel_bytes(cfg->precision) >= (1LL << 32))   // buffer-resource reads
        return fail(FS_EINVAL, "a field must stay below 4 GiB (fp32 storage: fewer than 2^28 cells)");
    if (cfg->solve < FS_SOLVE_VCYCLES || cfg->solve > FS_SOLVE_SPECTRAL64) return fail(FS_EINVAL, "bad solve mode");
    if (cfg->solve == FS_SOLVE_SPECTRAL64 && nx % 2 != 0) return fail(FS_EINVAL, "FS_SOLVE_SPECTRAL64 needs an even nx");
    // its transforms are library GEMMs that cannot skip on the device's stop flag: iterations are enqueued as the
    // host's check of that flag asks for them
    if (cfg->solve == FS_SOLVE_SPECTRAL64 && !cfg->adaptive)
        return fail(FS_EINVAL, "FS_SOLVE_SPECTRAL64 needs adaptive = 1 (the host enqueues each correction)");
    if (!(cfg->omega == 0.f || (cfg->omega > 0.f && cfg->omega < 2.f))) return fail(FS_EINVAL, "omega must be 0 or in (0, 2)");
    if (cfg->pre_sweeps < -1 || cfg->pre_sweeps > 8 || cfg->post_sweeps < 0 || cfg->post_sweeps > 8)
        return fail(FS_EINVAL, "pre sweeps must be in -1..8, post sweeps in 0..8");
    if (cfg->vcycles < 0 || cfg->max_vcycles < 0 || cfg->max_vcycles > MG_MAX_REFINE) return fail(FS_EINVAL, "bad V-cycle count");
    const int nn = cfg->noise_n;
    if (nn < 4 || (nn & (nn - 1)) != 0 || nn > 512) return fail(FS_EINVAL, "noise_n must be a power of two in [4, 512]");
    if (n_logs < 0 || (n_logs > 0 && !logs)) return fail(FS_EINVAL, "bad logs");
    if (!(cfg->size[0] > 0.f)) return fail(FS_EINVAL, "size must be positive");

    fs_sim* h = new fs_sim();
    h->cfg = *cfg;
    h->nx = nx; h->ny = ny; h->nz = nz;
    h->ncell = (size_t)nx * ny * nz;
    h->dx = (double)cfg->size[0] / nx;                                  // sim.js:53
    h->origin[0] = -(double)cfg->size[0] / 2; h->origin[1] = 0; h->origin[2] = -(double)cfg->size[2] / 2;   // sim.js:54
    h->prec = cfg->precision;
    h->g = make_dims(nx, ny, nz);
    default_params(h->prm);
    auto bail = [&](int code) { std::string e = g_last_error; free_sim(h); g_last_error = e; return code; };
    if (hipSetDevice(cfg->device) != hipSuccess) return bail(fail(FS_EDEVICE, "hipSetDevice failed"));
    if (cfg->stream) h->stream = (hipStream_t)cfg->stream;
    else {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(FS_EDEVICE, "stream create failed"));
        h->own_stream = true;
    }
    const size_t fb = h->ncell * texel_bytes(h->prec);
    for (int i = 0; i < F_COUNT; i++) {
        if (hipMalloc(&h->f[i], fb) != hipSuccess) return bail(fail(FS_ENOMEM, "hipMalloc failed for a field"));
        launch_fill_zero(h->stream, h->f[i], fb);
    }
    const size_t nb = (size_t)nn * nn * nn * sizeof(half4_t);
    if (hipMalloc(&h->noise, nb) != hipSuccess || hipMalloc(&h->curl, nb) != hipSuccess)
        return bail(fail(FS_ENOMEM, "hipMalloc failed for noise textures"));
    if (hipMalloc(&h->P_d, sizeof(DevParams)) != hipSuccess || hipMalloc(&h->spots_d, FS_MAX_SPOTS * sizeof(fs_spot)) != hipSuccess ||
        hipMalloc(&h->face_d, (size_t)nx * nz * sizeof(float)) != hipSuccess)
        return bail(fail(FS_ENOMEM, "hipMalloc failed"));
    launch_fill_zero(h->stream, h->spots_d, FS_MAX_SPOTS * sizeof(fs_spot));
    launch_fill_zero(h->stream, h->face_d, (size_t)nx * nz * sizeof(float));
    if (std::max(nx, std::max(ny, nz)) <= FS_BTPACK_MAX_DIM && hipMalloc(&h->btpack_d, h->ncell * sizeof(unsigned)) != hipSuccess)
        return bail(fail(FS_ENOMEM, "hipMalloc failed"));
    std::string err;
    int rc = mg_build(h->mg, nx, ny, nz, (float)h->dx, cfg->max_levels, err);
    if (rc) return bail(fail(rc, err));
    // over-relaxed smoothing in the tolerance solves only; the reference's solve is plain GS
    if (cfg->solve != FS_SOLVE_VCYCLES) {
        h->mg.omega = cfg->omega > 0.f ? cfg->omega : 1.25f;
        h->mg.pre = cfg->pre_sweeps > 0 ? cfg->pre_sweeps : cfg->pre_sweeps < 0 ? 0 : 1;
        h->mg.post = cfg->post_sweeps > 0 ? cfg->post_sweeps : 2;
        if (cfg->solve == FS_SOLVE_REFINE64) h->mg.first = cfg->first_cycles > 0 ? cfg->first_cycles : 3;
    }
    if (cfg->solve == FS_SOLVE_REFINE64 || cfg->solve == FS_SOLVE_TOL32) {
        rc = mg_set_direct(h->mg, cfg->direct_level, err);
        if (rc) return bail(fail(rc, err));
    }
    if (cfg->solve == FS_SOLVE_SPECTRAL64) {
        rc = spectral_build(&h->spec, nx, ny, nz, (double)h->mg.h[0], h->stream, h->mg.b[0], h->mg.p[0], err);
        if (rc) return bail(fail(rc, err));
    }
    for (auto& e : h->ev) if (hipEventCreate(&e) != hipSuccess) return bail(fail(FS_EDEVICE, "event create failed"));
    if (hipEventCreateWithFlags(&h->ev_state, hipEventDisableTiming) != hipSuccess)
        return bail(fail(FS_EDEVICE, "event create failed"));
    for (auto& e : h->ev_slot)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bail(fail(FS_EDEVICE, "event create failed"));
    if (hipHostMalloc((void**)&h->host_state, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return bail(fail(FS_ENOMEM, "pinned host allocation failed"));
    if (hipHostGetDevicePointer((void**)&h->host_state_dev, h->host_state, 0) != hipSuccess)
        return bail(fail(FS_EDEVICE, "pinned host memory is not device-accessible"));
    launch_noise_bake(h->stream, h->noise, h->curl, nn, cfg->noise_seed);   // src/main.js:36-37
    std::vector<float> none;
    rc = upload_logs(h, n_logs > 0 ? logs : none.data(), n_logs);
    if (rc) return bail(rc);
    push_params(h, 0.0);
    if (hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return bail(fail(FS_EDEVICE, "device error during create"));
    *out = h;
    return FS_OK;
}

int fs_destroy(fs_sim* h) {
    if (!h) return FS_OK;
    (void)hipSetDevice(h->cfg.device);
    (void)hipStreamSynchronize(h->stream);
    free_sim(h);
    return FS_OK;
}

int fs_set_params(fs_sim* h, const char* const* names, const float* values, int n) {
    if (!h || (n > 0 && (!names || !values))) return fail(FS_EINVAL, "null argument");
    int idx[FS_NUM_PARAMS * 4];
    if (n > (int)(sizeof(idx) / sizeof(idx[0]))) return fail(FS_EINVAL, "too many parameters");
    for (int i = 0; i < n; i++) {   // validate all first: an unknown key changes nothing (src/gpu.js:65)
        idx[i] = fs_param_index(names[i]);
        if (idx[i] < 0) return fail(FS_EINVAL, std::string("unknown uniform SimParams.") + (names[i] ? names[i] : "(null)"));
    }
    for (int i = 0; i < n; i++) {
        const int k = idx[i];
        if (k == P_hotR0 || k == P_hotR1 || k == P_hotY1 || k == P_bedRadius) h->geo_ready = false;   // sim.js:602
        if (k <= P_nLogs && k != P_seed) continue;   // grid/clock entries belong to the handle (sim.js:622-623)
        h->prm[k] = values[i];
    }
    h->params_dirty = true;
    return FS_OK;
}

int fs_get_params(fs_sim* h, float* out) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    std::memcpy(out, h->prm, sizeof(h->prm));
    out[P_nx] = (float)h->nx; out[P_ny] = (float)h->ny; out[P_nz] = (float)h->nz; out[P_dx] = (float)h->dx;
    out[P_time] = (float)h->time; out[P_frame] = (float)h->frame;
    out[P_ox] = (float)h->origin[0]; out[P_oy] = (float)h->origin[1]; out[P_oz] = (float)h->origin[2];
    out[P_nLogs] = (float)(h->logs_h.size() / 8);
    out[P_dt] = (float)(h->dev_dt < 0 ? 0.0 : h->dev_dt);
    return FS_OK;
}

int fs_set_logs(fs_sim* h, const float* logs, int n_logs) {
    if (!h || n_logs < 0 || (n_logs > 0 && !logs)) return fail(FS_EINVAL, "bad logs");
    if (set_device(h)) return FS_EDEVICE;
    return upload_logs(h, logs, n_logs);
}

// normalisation of a spot's mass injection over the grid (double, cell centres
// origin + (i + 1/2) dx, the kernel's weight w = 1 - smoothstep(r/2, r, d))
static double spot_weight_sum(const fs_sim* h, const fs_spot& sp) {
    const double dx = (double)(float)h->dx;
    const double ox = (double)(float)h->origin[0], oy = (double)(float)h->origin[1], oz = (double)(float)h->origin[2];
    const double r = sp.radius;
    auto cell = [&](double o, double pc) { return (int)std::floor((pc - o) / dx - 0.5); };
    const int x0 = std::max(0, cell(ox, sp.pos[0] - r) - 1), x1 = std::min(h->nx - 1, cell(ox, sp.pos[0] + r) + 2);
    const int y0 = std::max(0, cell(oy, sp.pos[1] - r) - 1), y1 = std::min(h->ny - 1, cell(oy, sp.pos[1] + r) + 2);
    const int z0 = std::max(0, cell(oz, sp.pos[2] - r) - 1), z1 = std::min(h->nz - 1, cell(oz, sp.pos[2] + r) + 2);
    double sum = 0.0;
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                const double px = ox + (x + 0.5) * dx, py = oy + (y + 0.5) * dx, pz = oz + (z + 0.5) * dx;
                const double d = std::sqrt((px - sp.pos[0]) * (px - sp.pos[0]) + (py - sp.pos[1]) * (py - sp.pos[1]) +
                                           (pz - sp.pos[2]) * (pz - sp.pos[2]));
                if (d >= r) continue;
                double t = (d - 0.5 * r) / (0.5 * r);
                t = std::min(std::max(t, 0.0), 1.0);
                sum += 1.0 - t * t * (3.0 - 2.0 * t);
            }
    return sum;
}

int fs_set_spots(fs_sim* h, const fs_spot* spots, int n) {
    if (!h || n < 0 || n > FS_MAX_SPOTS || (n > 0 && !spots)) return fail(FS_EINVAL, "bad spots (at most FS_MAX_SPOTS)");
    if (set_device(h)) return FS_EDEVICE;
    std::vector<fs_spot> v(spots, spots + n);
    const double dx3 = std::pow((double)(float)h->dx, 3.0);
    for (auto& sp : v) {
        if (!(sp.radius > 0.f) || !std::isfinite(sp.flux)) return fail(FS_EINVAL, "spot radius must be positive, flux finite");
        const double ws = spot_weight_sum(h, sp);
        sp.norm = ws > 0.0 ? (float)(1.0 / ((double)FS_RHO0 * ws * dx3)) : 0.f;
        sp.pad = 0.f;
    }
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (n > 0) FS_HIP_CHECK(hipMemcpy(h->spots_d, v.data(), n * sizeof(fs_spot), hipMemcpyHostToDevice));
    h->spots_h = v;
    h->params_dirty = true;
    return FS_OK;
}

int fs_get_spots(fs_sim* h, fs_spot* out, int max, int* n) {
    if (!h || (max > 0 && !out)) return fail(FS_EINVAL, "null argument");
    const int k = std::min(max, (int)h->spots_h.size());
    for (int i = 0; i < k; i++) out[i] = h->spots_h[i];
    if (n) *n = (int)h->spots_h.size();
    return FS_OK;
}

int fs_set_face_flux(fs_sim* h, const float* flux) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (flux) FS_HIP_CHECK(hipMemcpy(h->face_d, flux, (size_t)h->nx * h->nz * sizeof(float), hipMemcpyHostToDevice));
    h->has_face = flux != nullptr;
    h->params_dirty = true;
    return FS_OK;
}

int fs_reset(fs_sim* h) {   // sim.js:606-618
    if (!h) return fail(FS_EINVAL, "null handle");
    if (set_device(h)) return FS_EDEVICE;
    const size_t fb = h->ncell * texel_bytes(h->prec);
    for (int i = 0; i < F_COUNT; i++) if (i != F_GEO) launch_fill_zero(h->stream, h->f[i], fb);
    for (int l = 0; l < h->mg.nlev; l++)
        launch_fill_zero(h->stream, h->mg.p[l], (size_t)h->mg.dims[l][0] * h->mg.dims[l][1] * h->mg.dims[l][2] * sizeof(float));
    launch_fill_zero(h->stream, h->mg.x64, h->ncell * sizeof(double));
    h->time = 0.0;
    h->frame = 0;
    h->params_dirty = true;
    h->gen++;
    if (hipGetLastError() != hipSuccess) return fail(FS_EDEVICE, "launch failed in reset");
    return FS_OK;
}

int fs_step(fs_sim* h, double dt, int nsteps) {   // sim.js:627-648
    if (!h) return fail(FS_EINVAL, "null handle");
    if (nsteps < 0 || !(dt > 0.0) || !std::isfinite(dt)) return fail(FS_EINVAL, "dt must be positive and finite, nsteps >= 0");
    if (set_device(h)) return FS_EDEVICE;
    h->gen++;
    h->inflight = false;
    h->slot = 0;
    if (h->cfg.solve == FS_SOLVE_REFINE64 && nsteps > 0)   // no step is held from an earlier call
        launch_fill_zero(h->stream, &h->mg.st->hold, 2 * sizeof(int));
    for (int k = 0; k < nsteps; k++) {
        sync_params(h, dt);
        if (!h->geo_ready) {   // sim.js:632
            launch_geo(h->prec, h->g, h->stream, h->P_d, h->fields());
            h->geo_ready = true;
        }
        const int rc = h->adaptive() ? adaptive_step(h, h->profile, k + 1 < nsteps)
                                     : run_part(h, fs_sim::G_FULL, h->profile);
        if (rc) return rc;
        h->time += dt;   // sim.js:646-647 (the device clock advances inside the step)
        h->frame += 1;
        if (h->profile) h->timings_valid = true;
    }
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_sync(fs_sim* h) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    return FS_OK;
}

int fs_run_stage(fs_sim* h, int stage, double dt) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (stage < 0 || stage >= FS_NUM_STAGES) return fail(FS_EINVAL, "bad stage");
    if (set_device(h)) return FS_EDEVICE;
    h->gen++;
    sync_params(h, dt);
    FieldSet F = h->fields();
    F.btpack = nullptr;   // a stage on its own: the second advection pass forms its own backtrace
    hipStream_t s = h->stream;
    switch (stage) {
        case FS_STAGE_GEO: launch_geo(h->prec, h->g, s, h->P_d, F); h->geo_ready = true; break;
        case FS_STAGE_ADV1: launch_adv1(h->prec, h->g, s, h->P_d, F); break;
        case FS_STAGE_ADV2: launch_adv2(h->prec, h->g, s, h->P_d, F); break;
        case FS_STAGE_CURL: launch_curl(h->prec, h->g, s, F); break;
        case FS_STAGE_FORCES: launch_forces(h->prec, h->g, s, h->P_d, F); break;
        case FS_STAGE_DIV: launch_div(h->prec, h->g, s, h->P_d, F, h->refine() ? h->mg.bkeep : h->mg.b[0], 0); break;
        case FS_STAGE_MG: if (int rc = h->solve(s)) return rc; break;
        case FS_STAGE_PROJECT:
            launch_project(h->prec, h->g, s, h->P_d, F, h->refine() ? nullptr : h->mg.p[0], h->refine() ? h->mg.x64 : nullptr);
            break;
    }
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_field_info(fs_sim* h, int field, int dims[3], int* channels, int* native_bytes) {
    if (!h) return fail(FS_EINVAL, "null handle");
    DevField d;
    if (!resolve_field(h, field, d)) return fail(FS_EINVAL, "unknown field");
    const bool tex = field == FS_NOISE || field == FS_CURL;
    if (dims) {
        dims[0] = tex ? h->cfg.noise_n : h->nx; dims[1] = tex ? h->cfg.noise_n : h->ny; dims[2] = tex ? h->cfg.noise_n : h->nz;
    }
    if (channels) *channels = d.channels;
    if (native_bytes) *native_bytes = (int)kind_bytes(d.kind);
    return FS_OK;
}

void* fs_device_ptr(fs_sim* h, int field) {
    if (!h) return nullptr;
    DevField d;
    if (!resolve_field(h, field, d)) { fail(FS_EINVAL, "unknown field"); return nullptr; }
    return d.ptr;
}

int fs_read_field(fs_sim* h, int field, int dtype, void* dst, size_t bytes) {
    if (!h || !dst) return fail(FS_EINVAL, "null argument");
    if (set_device(h)) return FS_EDEVICE;
    DevField d;
    if (!resolve_field(h, field, d)) return fail(FS_EINVAL, "unknown field");
    const size_t nval = d.count * d.channels;
    const size_t want = dtype == FS_DT_F32 ? nval * 4 : dtype == FS_DT_F64 ? nval * 8 : nval * kind_bytes(d.kind);
    if (dtype != FS_DT_F32 && dtype != FS_DT_F64 && dtype != FS_DT_NATIVE) return fail(FS_EINVAL, "bad dtype");
    if (bytes != want) return fail(FS_EINVAL, "destination size mismatch: expected " + std::to_string(want) + " bytes");
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (is_mg_field(field)) {   // parity-split on the device: copy, reorder, convert
        const size_t nd = mg_level_floats(h->mg, 0);
        if (d.kind == 3) {
            std::vector<double> raw(nd), nat(nval);
            FS_HIP_CHECK(hipMemcpy(raw.data(), d.ptr, nd * 8, hipMemcpyDeviceToHost));
            rb_to_natural(raw.data(), nat.data(), h->nx, h->ny, h->nz);
            if (dtype == FS_DT_F32) for (size_t i = 0; i < nval; i++) ((float*)dst)[i] = (float)nat[i];
            else std::memcpy(dst, nat.data(), nval * 8);
        } else {
            std::vector<float> raw(nd), nat(nval);
            FS_HIP_CHECK(hipMemcpy(raw.data(), d.ptr, nd * 4, hipMemcpyDeviceToHost));
            rb_to_natural(raw.data(), nat.data(), h->nx, h->ny, h->nz);
            if (dtype == FS_DT_F64) for (size_t i = 0; i < nval; i++) ((double*)dst)[i] = (double)nat[i];
            else std::memcpy(dst, nat.data(), nval * 4);
        }
        return FS_OK;
    }
    if (dtype == FS_DT_NATIVE) { FS_HIP_CHECK(hipMemcpy(dst, d.ptr, want, hipMemcpyDeviceToHost)); return FS_OK; }
    // decode to f32 (or f64) on the device, then copy
    if (d.kind == 3) {
        std::vector<double> tmp(nval);
        FS_HIP_CHECK(hipMemcpy(tmp.data(), d.ptr, nval * 8, hipMemcpyDeviceToHost));
        if (dtype == FS_DT_F64) std::memcpy(dst, tmp.data(), nval * 8);
        else for (size_t i = 0; i < nval; i++) ((float*)dst)[i] = (float)tmp[i];
        return FS_OK;
    }
    float* dev = nullptr;
    if (d.kind == 0) dev = (float*)d.ptr;
    else {
        FS_HIP_CHECK(hipMalloc(&dev, nval * 4));
        if (d.kind == 1) k_unpack_f16<<<(unsigned)((d.count + 255) / 256), 256, 0, h->stream>>>((const half4_t*)d.ptr, dev, d.count);
        else k_unpack_bf16<<<(unsigned)((nval + 255) / 256), 256, 0, h->stream>>>((const unsigned short*)d.ptr, dev, nval);
        FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    std::vector<float> tmp(nval);
    hipError_t e = hipMemcpy(tmp.data(), dev, nval * 4, hipMemcpyDeviceToHost);
    if (d.kind != 0) (void)hipFree(dev);
    if (e != hipSuccess) return fail(FS_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if (dtype == FS_DT_F32) std::memcpy(dst, tmp.data(), nval * 4);
    else for (size_t i = 0; i < nval; i++) ((double*)dst)[i] = (double)tmp[i];
    return FS_OK;
}

int fs_write_field(fs_sim* h, int field, int dtype, const void* src, size_t bytes) {
    if (!h || !src) return fail(FS_EINVAL, "null argument");
    if (set_device(h)) return FS_EDEVICE;
    DevField d;
    if (!resolve_field(h, field, d)) return fail(FS_EINVAL, "unknown field");
    const size_t nval = d.count * d.channels;
    const size_t want = dtype == FS_DT_F32 ? nval * 4 : dtype == FS_DT_F64 ? nval * 8 : nval * kind_bytes(d.kind);
    if (dtype != FS_DT_F32 && dtype != FS_DT_F64 && dtype != FS_DT_NATIVE) return fail(FS_EINVAL, "bad dtype");
    if (bytes != want) return fail(FS_EINVAL, "source size mismatch: expected " + std::to_string(want) + " bytes");
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->gen++;
    if (is_mg_field(field)) {   // reorder into the device's parity-split layout (padding stays zero)
        const size_t nd = mg_level_floats(h->mg, 0);
        const bool src64 = dtype == FS_DT_F64 || (dtype == FS_DT_NATIVE && d.kind == 3);
        if (d.kind == 3) {
            std::vector<double> nat(nval), raw(nd, 0.0);
            for (size_t i = 0; i < nval; i++) nat[i] = src64 ? ((const double*)src)[i] : (double)((const float*)src)[i];
            natural_to_rb(nat.data(), raw.data(), h->nx, h->ny, h->nz);
            FS_HIP_CHECK(hipMemcpy(d.ptr, raw.data(), nd * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> nat(nval), raw(nd, 0.f);
            for (size_t i = 0; i < nval; i++) nat[i] = src64 ? (float)((const double*)src)[i] : ((const float*)src)[i];
            natural_to_rb(nat.data(), raw.data(), h->nx, h->ny, h->nz);
            FS_HIP_CHECK(hipMemcpy(d.ptr, raw.data(), nd * 4, hipMemcpyHostToDevice));
        }
    } else if (dtype == FS_DT_NATIVE) FS_HIP_CHECK(hipMemcpy(d.ptr, src, want, hipMemcpyHostToDevice));
    else if (d.kind == 3) {
        std::vector<double> tmp(nval);
        if (dtype == FS_DT_F64) std::memcpy(tmp.data(), src, nval * 8);
        else for (size_t i = 0; i < nval; i++) tmp[i] = (double)((const float*)src)[i];
        FS_HIP_CHECK(hipMemcpy(d.ptr, tmp.data(), nval * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(nval);
        if (dtype == FS_DT_F32) std::memcpy(tmp.data(), src, nval * 4);
        else for (size_t i = 0; i < nval; i++) tmp[i] = (float)((const double*)src)[i];
        if (d.kind == 0) FS_HIP_CHECK(hipMemcpy(d.ptr, tmp.data(), nval * 4, hipMemcpyHostToDevice));
        else {
            float* dev = nullptr;
            FS_HIP_CHECK(hipMalloc(&dev, nval * 4));
            hipError_t e = hipMemcpy(dev, tmp.data(), nval * 4, hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                if (d.kind == 1) k_pack_f16<<<(unsigned)((d.count + 255) / 256), 256, 0, h->stream>>>(dev, (half4_t*)d.ptr, d.count);
                else k_pack_bf16<<<(unsigned)((nval + 255) / 256), 256, 0, h->stream>>>(dev, (unsigned short*)d.ptr, nval);
                e = hipStreamSynchronize(h->stream);
            }
            (void)hipFree(dev);
            if (e != hipSuccess) return fail(FS_EDEVICE, std::string("field upload: ") + hipGetErrorString(e));
        }
    }
    // writing the pressure in FS_SOLVE_VCYCLES mode also seeds the FP64 iterate and vice versa
    if (field == FS_P) {
        const size_t nd = mg_level_floats(h->mg, 0);
        if (h->refine()) launch_convert(h->stream, 0, h->mg.x64, h->mg.p[0], nd);
        else launch_convert(h->stream, 1, h->mg.p[0], h->mg.x64, nd);
        FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    }
    if (field == FS_GEO) h->geo_ready = true;
    return FS_OK;
}

int fs_get_clock(fs_sim* h, double* time, int64_t* frame) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (time) *time = h->time;
    if (frame) *frame = h->frame;
    return FS_OK;
}

int fs_set_clock(fs_sim* h, double time, int64_t frame) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (!std::isfinite(time) || frame < 0) return fail(FS_EINVAL, "time must be finite and frame >= 0");
    h->time = time;
    h->frame = frame;
    h->params_dirty = true;
    return FS_OK;
}

int fs_grid_info(fs_sim* h, int res[3], double* dx, double origin[3]) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (res) { res[0] = h->nx; res[1] = h->ny; res[2] = h->nz; }
    if (dx) *dx = h->dx;
    if (origin) { origin[0] = h->origin[0]; origin[1] = h->origin[1]; origin[2] = h->origin[2]; }
    return FS_OK;
}

int fs_mg_info(fs_sim* h, int* levels, int dims[][3], int* coarse_iters) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (levels) *levels = h->mg.nlev;
    if (dims) for (int l = 0; l < h->mg.nlev; l++) for (int k = 0; k < 3; k++) dims[l][k] = h->mg.dims[l][k];
    if (coarse_iters) *coarse_iters = h->mg.coarse_iters;
    return FS_OK;
}

int fs_residual(fs_sim* h, double* rel, int* used) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (set_device(h)) return FS_EDEVICE;
    const bool tracked = h->cfg.solve != FS_SOLVE_VCYCLES;   // the solve itself recorded its residuals
    if (!tracked) mg_residual_of_p(h->mg, h->stream);
    else mg_state_ready(h->mg, h->stream, h->cfg.div_tol, h->cfg.max_vcycles);
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    RefineState st;
    FS_HIP_CHECK(hipMemcpy(&st, h->mg.st, sizeof(st), hipMemcpyDeviceToHost));
    if (rel) *rel = tracked ? st.rel[st.used] : st.rel[0];
    if (used) *used = tracked ? st.used : h->cfg.vcycles;
    return FS_OK;
}

int fs_profile_enable(fs_sim* h, int on) {
    if (!h) return fail(FS_EINVAL, "null handle");
    h->profile = on != 0;
    h->timings_valid = false;
    return FS_OK;
}

int fs_timings(fs_sim* h, float ms[FS_NUM_TIMINGS]) {
    if (!h || !ms) return fail(FS_EINVAL, "null argument");
    if (!h->timings_valid) return fail(FS_ESTATE, "no profiled step yet (fs_profile_enable, then fs_step)");
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    for (int i = 0; i < FS_NUM_TIMINGS; i++) FS_HIP_CHECK(hipEventElapsedTime(&ms[i], h->ev[i], h->ev[i + 1]));
    return FS_OK;
}

int fs_bench_copy(int device, size_t bytes, int iters, float* ms_per_copy) {
    if (!ms_per_copy || iters < 1 || bytes < 16) return fail(FS_EINVAL, "bad arguments");
    bytes &= ~(size_t)15;
    if (hipSetDevice(device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FS_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = fail(FS_ENOMEM, "hipMalloc failed");
    if (!rc && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = fail(FS_EDEVICE, "stream/event create failed");
    if (!rc) {
        launch_fill_zero(s, a, bytes);
        launch_copy_probe(s, a, b, bytes);   // warm-up
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < iters; i++) launch_copy_probe(s, a, b, bytes);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FS_EDEVICE, "copy probe failed");
        float ms = 0.f;
        if (!rc) { (void)hipEventElapsedTime(&ms, e0, e1); *ms_per_copy = ms / iters; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int fs_bench_stream(int device, size_t bytes, int kind, int iters, float* ms_per_pass) {
    if (!ms_per_pass || iters < 1 || bytes < 16 || kind < 0 || kind > 3) return fail(FS_EINVAL, "bad arguments");
    bytes &= ~(size_t)15;
    if (hipSetDevice(device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FS_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = fail(FS_ENOMEM, "hipMalloc failed");
    if (!rc && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = fail(FS_EDEVICE, "stream/event create failed");
    if (!rc) {
        launch_fill_zero(s, a, bytes);
        launch_fill_zero(s, b, bytes);
        auto pass = [&]() {
            if (kind == 0) launch_copy_probe(s, a, b, bytes);
            else launch_stream_probe(s, kind, a, b, bytes);
        };
        pass();   // warm-up
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < iters; i++) pass();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(FS_EDEVICE, "stream probe failed");
        float ms = 0.f;
        if (!rc) { (void)hipEventElapsedTime(&ms, e0, e1); *ms_per_pass = ms / iters; }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

// buildCampfire(seed).logs + packLogs (src/scene.js:13-42, 92-98)
int fs_build_campfire(int seed, float* out, int max_logs) {
    if (!out || max_logs < 0) return fail(FS_EINVAL, "bad arguments");
    uint32_t s = (uint32_t)seed;
    auto R = [&]() { s = s * 1664525u + 1013904223u; return (double)s / 4294967296.0; };   // scene.js:13-16
    std::vector<std::vector<double>> logs;
    auto add = [&](double ax, double ay, double az, double bx, double by, double bz, double r, double burn) {
        logs.push_back({ax, ay, az, r, bx, by, bz, burn});
    };
    add(-0.40, 0.055, -0.17, 0.38, 0.065, -0.13, 0.062, 0.8);   // two base logs
    add(-0.36, 0.060, 0.15, 0.41, 0.050, 0.19, 0.058, 0.8);
    for (int i = 0; i < 5; i++) {                                // teepee of split wood
        const double ang = ((double)i / 5) * M_PI * 2 + 0.35 + (R() - 0.5) * 0.4;
        const double rb = 0.30 + R() * 0.06;
        const double topR = 0.035 + R() * 0.03;
        const double ta = ang + M_PI + (R() - 0.5) * 0.8;
        const double ty = 0.36 + R() * 0.1;
        const double rr = 0.032 + R() * 0.012;
        add(std::cos(ang) * rb, 0.03, std::sin(ang) * rb, std::cos(ta) * topR, ty, std::sin(ta) * topR, rr, 1.0);
    }
    add(-0.12, 0.035, 0.02, 0.10, 0.03, -0.05, 0.035, 1.0);      // charred chunk in the coals
    std::fill(out, out + (size_t)max_logs * 8, 0.f);
    const int cnt = std::min((int)logs.size(), max_logs);
    for (int i = 0; i < cnt; i++) for (int k = 0; k < 8; k++) out[i * 8 + k] = (float)logs[i][k];
    return cnt;
}

// ---- isolated Poisson solve ---------------------------------------------------------
// ---- diagnostics (src/main.js:268-355) -------------------------------------------
namespace {
int ensure_planes(fs_sim* h) {
    if (!h->planes_d) FS_HIP_CHECK(hipMalloc(&h->planes_d, (size_t)h->ny * 8 * sizeof(double)));
    return FS_OK;
}
}  // namespace

int fs_stats(fs_sim* h, int field, double out[20]) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    const int slot = field_slot(field);
    if (slot < 0) return fail(FS_EINVAL, "stats needs a 4-channel state field");
    if (set_device(h)) return FS_EDEVICE;
    const int nb = stats_blocks(h->ncell);
    if (!h->stats_d) FS_HIP_CHECK(hipMalloc(&h->stats_d, (size_t)nb * 20 * sizeof(double)));
    launch_stats(h->prec, h->stream, h->f[slot], h->ncell, h->stats_d);
    FS_HIP_CHECK(hipGetLastError());
    std::vector<double> part((size_t)nb * 20);
    FS_HIP_CHECK(hipMemcpyAsync(part.data(), h->stats_d, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < 4; c++) {
        double mn = INFINITY, mx = -INFINITY, sum = 0.0, bad = 0.0, arg = -1.0;
        for (int b = 0; b < nb; b++) {   // blocks in index order: ties keep the earlier maximum
            const double* q = &part[((size_t)b * 4 + c) * 5];
            mn = std::min(mn, q[0]);
            if (q[4] >= 0 && (arg < 0 || q[1] > mx)) { mx = q[1]; arg = q[4]; }
            sum += q[2];
            bad += q[3];
        }
        out[c * 5 + 0] = mn;
        out[c * 5 + 1] = mx;
        out[c * 5 + 2] = sum / (double)h->ncell;
        out[c * 5 + 3] = bad;
        out[c * 5 + 4] = arg;
    }
    return FS_OK;
}

int fs_stats_noalloc(fs_sim* h, int field, double out[20]) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    const int slot = field_slot(field);
    if (slot < 0) return fail(FS_EINVAL, "stats needs a 4-channel state field");
    if (set_device(h)) return FS_EDEVICE;
    return stats_noalloc(h->prec, h->stream, h->cfg.device, h->f[slot], h->ncell, out);
}

int fs_profile(fs_sim* h, double* q_watts, double* rows, int max_rows, int* n_rows) {
    if (!h || (max_rows > 0 && !rows)) return fail(FS_EINVAL, "null argument");
    if (set_device(h)) return FS_EDEVICE;
    if (ensure_planes(h)) return FS_EDEVICE;
    const double dx = (double)(float)h->dx;
    launch_planes(h->prec, h->stream, h->f[F_S], h->nx, h->ny, h->nz, dx * dx * dx, h->planes_d);
    FS_HIP_CHECK(hipGetLastError());
    std::vector<double> pl((size_t)h->ny * 8);
    FS_HIP_CHECK(hipMemcpyAsync(pl.data(), h->planes_d, pl.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    double Q = 0.0;
    for (int j = 0; j < h->ny; j++) Q += pl[(size_t)j * 8 + 6];
    if (q_watts) *q_watts = Q;
    int r = 0;
    for (int j = 0; j < h->ny; j += 4, r++) {
        if (r >= max_rows) continue;
        const double* q = &pl[(size_t)j * 8];
        double* o = rows + (size_t)r * 6;
        o[0] = (j + 0.5) * h->dx;
        o[1] = q[2] > 0 ? q[1] / q[2] : 0.0;
        o[2] = q[0];
        o[3] = q[3];
        o[4] = q[4];
        o[5] = q[5];
    }
    if (n_rows) *n_rows = r;
    return FS_OK;
}

int fs_diag_enable(fs_sim* h, int capacity, double t_flame) {
    if (!h || capacity < 0 || !(t_flame > 0)) return fail(FS_EINVAL, "bad diagnostics arguments");
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    drop_graphs(h);                  // the record kernels join (or leave) the captured step
    if (h->ring_d) { (void)hipFree(h->ring_d); h->ring_d = nullptr; }
    h->diag_cap = 0;
    if (capacity == 0) return FS_OK;
    if (ensure_planes(h)) return FS_EDEVICE;
    if (!h->ring_count_d) FS_HIP_CHECK(hipMalloc(&h->ring_count_d, sizeof(int)));
    FS_HIP_CHECK(hipMalloc(&h->ring_d, (size_t)capacity * (4 + h->ny) * sizeof(double)));
    launch_fill_zero(h->stream, h->ring_count_d, sizeof(int));
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->diag_cap = capacity;
    h->diag_tflame = t_flame;
    return FS_OK;
}

int fs_diag_read(fs_sim* h, double* out, int max_records, int* n_records) {
    if (!h || (max_records > 0 && !out)) return fail(FS_EINVAL, "null argument");
    if (h->diag_cap == 0) { if (n_records) *n_records = 0; return FS_OK; }
    if (set_device(h)) return FS_EDEVICE;
    int count = 0;
    FS_HIP_CHECK(hipMemcpyAsync(&count, h->ring_count_d, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    const int have = std::min(count, h->diag_cap);
    const size_t rec = 4 + (size_t)h->ny;
    std::vector<double> ring((size_t)h->diag_cap * rec);
    FS_HIP_CHECK(hipMemcpy(ring.data(), h->ring_d, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int first = count - have;               // oldest record still held
    const int k0 = std::max(0, have - max_records);   // keep the newest max_records
    int w = 0;
    for (int k = k0; k < have; k++, w++) {
        const int idx = (first + k) % h->diag_cap;
        std::copy(&ring[(size_t)idx * rec], &ring[(size_t)idx * rec] + rec, out + (size_t)w * rec);
    }
    if (n_records) *n_records = w;
    return FS_OK;
}

int fs_budget_enable(fs_sim* h, int on) {
    if (!h) return fail(FS_EINVAL, "null handle");
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    drop_graphs(h);                  // the tracing source kernel joins (or leaves) the captured step
    h->budget_on = false;
    if (!on) return FS_OK;
    if (!h->budget_part) FS_HIP_CHECK(hipMalloc(&h->budget_part, budget_partials(h->g) * sizeof(double)));
    if (!h->budget_rec) FS_HIP_CHECK(hipMalloc(&h->budget_rec, FS_BUDGET_N * sizeof(double)));
    sync_params(h, h->dev_dt < 0 ? 0.0 : h->dev_dt);
    launch_budget_init(h->prec, h->g, h->stream, h->P_d, h->fields(), h->budget_part, h->budget_rec);
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->budget_on = true;
    return FS_OK;
}

int fs_budget_read(fs_sim* h, double* out) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    if (!h->budget_rec) return fail(FS_ESTATE, "the budget was never enabled");
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    FS_HIP_CHECK(hipMemcpy(out, h->budget_rec, FS_BUDGET_N * sizeof(double), hipMemcpyDeviceToHost));
    return FS_OK;
}

struct fs_poisson {
    GpuMG mg;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    int n[3];
};

int fs_poisson_create(const int dims[3], double h, int max_levels, int device, void* stream, fs_poisson** out) {
    if (!dims || !out || !(h > 0)) return fail(FS_EINVAL, "bad arguments");
    *out = nullptr;
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2) return fail(FS_EINVAL, "dims must be >= 2");
    if ((long long)dims[0] * dims[1] * dims[2] >= (1LL << 31) / 8) return fail(FS_EINVAL, "grid too large");
    fs_poisson* ps = new fs_poisson();
    ps->device = device;
    for (int k = 0; k < 3; k++) ps->n[k] = dims[k];
    if (hipSetDevice(device) != hipSuccess) { delete ps; return fail(FS_EDEVICE, "hipSetDevice failed"); }
    if (stream) ps->stream = (hipStream_t)stream;
    else {
        if (hipStreamCreateWithFlags(&ps->stream, hipStreamNonBlocking) != hipSuccess) { delete ps; return fail(FS_EDEVICE, "stream"); }
        ps->own_stream = true;
    }
    std::string err;
    int rc = mg_build(ps->mg, dims[0], dims[1], dims[2], (float)h, max_levels, err);
    if (rc) { mg_free(ps->mg); if (ps->own_stream) (void)hipStreamDestroy(ps->stream); delete ps; return fail(rc, err); }
    *out = ps;
    return FS_OK;
}

int fs_poisson_set_omega(fs_poisson* ps, float omega) {
    if (!ps || !(omega > 0.f && omega < 2.f)) return fail(FS_EINVAL, "omega must be in (0, 2)");
    ps->mg.omega = omega;
    return FS_OK;
}

int fs_poisson_destroy(fs_poisson* ps) {
    if (!ps) return FS_OK;
    (void)hipSetDevice(ps->device);
    (void)hipStreamSynchronize(ps->stream);
    mg_free(ps->mg);
    if (ps->own_stream) (void)hipStreamDestroy(ps->stream);
    delete ps;
    return FS_OK;
}

static int poisson_upload(fs_poisson* ps, int dtype, const void* src, float* f32dst, double* f64dst) {
    const int nx = ps->n[0], ny = ps->n[1], nz = ps->n[2];
    const size_t n = (size_t)nx * ny * nz, nd = mg_level_floats(ps->mg, 0);
    std::vector<double> d(n), drb(nd, 0.0);
    if (dtype == FS_DT_F64) std::memcpy(d.data(), src, n * 8);
    else if (dtype == FS_DT_F32) for (size_t i = 0; i < n; i++) d[i] = ((const float*)src)[i];
    else return fail(FS_EINVAL, "dtype must be FS_DT_F32 or FS_DT_F64");
    natural_to_rb(d.data(), drb.data(), nx, ny, nz);
    std::vector<float> f(nd);
    for (size_t i = 0; i < nd; i++) f[i] = (float)drb[i];
    FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
    if (f32dst) FS_HIP_CHECK(hipMemcpy(f32dst, f.data(), nd * 4, hipMemcpyHostToDevice));
    if (f64dst) FS_HIP_CHECK(hipMemcpy(f64dst, drb.data(), nd * 8, hipMemcpyHostToDevice));
    return FS_OK;
}

int fs_poisson_set_rhs(fs_poisson* ps, int dtype, const void* b) {
    if (!ps || !b) return fail(FS_EINVAL, "null argument");
    if (hipSetDevice(ps->device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    int rc = poisson_upload(ps, dtype, b, ps->mg.bkeep, nullptr);
    if (rc) return rc;
    FS_HIP_CHECK(hipMemcpy(ps->mg.b[0], ps->mg.bkeep, mg_level_floats(ps->mg, 0) * 4, hipMemcpyDeviceToDevice));
    return FS_OK;
}

int fs_poisson_set_x(fs_poisson* ps, int dtype, const void* x) {
    if (!ps || !x) return fail(FS_EINVAL, "null argument");
    if (hipSetDevice(ps->device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    return poisson_upload(ps, dtype, x, ps->mg.p[0], ps->mg.x64);
}

int fs_poisson_solve(fs_poisson* ps, double tol, int max_cycles, double* hist, int* cycles) {
    if (!ps || max_cycles < 0 || max_cycles > MG_MAX_REFINE) return fail(FS_EINVAL, "bad arguments");
    if (hipSetDevice(ps->device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    // one refinement iteration at a time, each followed by a read of the device's stopping
    // test: no V-cycles are launched past convergence (an iteration at this size costs far
    // more than the round trip)
    mg_refine_begin(ps->mg, ps->stream, tol, max_cycles);
    for (int k = 0; k <= max_cycles; k++) {
        int used_done[2];
        mg_state_ready(ps->mg, ps->stream, tol, max_cycles);
        FS_HIP_CHECK(hipMemcpyAsync(used_done, &ps->mg.st->used, sizeof(used_done), hipMemcpyDeviceToHost, ps->stream));
        FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
        if (used_done[1] || k == max_cycles) break;
        mg_refine_iter(ps->mg, ps->stream, tol, max_cycles);
    }
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
    RefineState st;
    FS_HIP_CHECK(hipMemcpy(&st, ps->mg.st, sizeof(st), hipMemcpyDeviceToHost));
    if (hist) for (int k = 0; k <= st.used; k++) hist[k] = st.rel[k];
    if (cycles) *cycles = st.used;
    return FS_OK;
}

int fs_poisson_vcycles(fs_poisson* ps, int cycles, double* hist) {
    if (!ps || cycles < 0) return fail(FS_EINVAL, "bad arguments");
    if (hipSetDevice(ps->device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    // the V-cycles act on (p, b) of level 0; b is the stored right-hand side
    FS_HIP_CHECK(hipMemcpyAsync(ps->mg.b[0], ps->mg.bkeep, mg_level_floats(ps->mg, 0) * 4, hipMemcpyDeviceToDevice,
                                ps->stream));
    for (int k = 0; k < cycles; k++) {
        mg_vcycle(ps->mg, ps->stream, nullptr);
        if (hist) {
            mg_residual_of_p(ps->mg, ps->stream);
            FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
            RefineState st;
            FS_HIP_CHECK(hipMemcpy(&st, ps->mg.st, sizeof(st), hipMemcpyDeviceToHost));
            hist[k] = st.rel[0];
        }
    }
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
    return FS_OK;
}

int fs_poisson_read_x(fs_poisson* ps, int dtype, void* x) {
    if (!ps || !x) return fail(FS_EINVAL, "null argument");
    if (hipSetDevice(ps->device) != hipSuccess) return fail(FS_EDEVICE, "hipSetDevice failed");
    const int nx = ps->n[0], ny = ps->n[1], nz = ps->n[2];
    const size_t nd = mg_level_floats(ps->mg, 0);
    FS_HIP_CHECK(hipStreamSynchronize(ps->stream));
    if (dtype == FS_DT_F64) {
        std::vector<double> raw(nd);
        FS_HIP_CHECK(hipMemcpy(raw.data(), ps->mg.x64, nd * 8, hipMemcpyDeviceToHost));
        rb_to_natural(raw.data(), (double*)x, nx, ny, nz);
    } else if (dtype == FS_DT_F32) {
        std::vector<float> raw(nd);
        FS_HIP_CHECK(hipMemcpy(raw.data(), ps->mg.p[0], nd * 4, hipMemcpyDeviceToHost));
        rb_to_natural(raw.data(), (float*)x, nx, ny, nz);
    } else return fail(FS_EINVAL, "dtype must be FS_DT_F32 (the FP32 iterate p) or FS_DT_F64 (the FP64 iterate x)");
    return FS_OK;
}

int fs_poisson_info(fs_poisson* ps, int* levels, int dims[][3], int* coarse_iters) {
    if (!ps) return fail(FS_EINVAL, "null handle");
    if (levels) *levels = ps->mg.nlev;
    if (dims) for (int l = 0; l < ps->mg.nlev; l++) for (int k = 0; k < 3; k++) dims[l][k] = ps->mg.dims[l][k];
    if (coarse_iters) *coarse_iters = ps->mg.coarse_iters;
    return FS_OK;
}

}  // extern "C"

// ---- reverse-mode derivative of a run of steps ------------------------------------
struct AdjSlot {                     // the state at the start of a step
    float4* vel = nullptr;
    float4* S = nullptr;
    float4* S2 = nullptr;
    double* p = nullptr;             // FP64 pressure of the previous step (the solve's starting guess)
};
struct StepRec {                     // one step's intermediates, kept for its reverse and tangent
    float4 *vX = nullptr, *sX = nullptr, *s2X = nullptr;   // adv1 out
    float4 *vY = nullptr, *sY = nullptr, *s2Y = nullptr;   // adv2 out
    float4* vort = nullptr;                                 // curl
    float4* vF = nullptr;                                   // forces' velocity, before the projection
    double* p = nullptr;                                    // the step's pressure (FP64, parity-split)
};

struct fs_adjoint {
    fs_sim* h = nullptr;
    int max_steps = 0, every = 1;
    std::vector<AdjSlot> ckpt;       // states 0, every, 2 every, ...
    std::vector<AdjSlot> seg;        // states inside one checkpoint interval (index = step % every; 0 unused)
    AdjSlot fin;                     // the final state (seg is reused by the reverse sweep)
    std::vector<StepRec> rec;        // intermediates of the steps of one interval (index = step % every)
    std::vector<double> times;       // clock at the start of each step
    std::vector<int64_t> frames;
    double dt = 0.0, t_end = 0.0;
    int64_t frame_end = 0;
    int nsteps = 0;
    bool have_forward = false, have_grads = false;
    bool have_tangent = false;       // l[L_VEL..L_S2] hold the final state's tangent
    bool seeds_stale = false;        // the seed buffers hold something else: clear on the first seed
    bool last_interval_held = false; // seg and rec hold the last interval (true after the forward)
    int predict_step = 1, predict_lin = 1;   // V-cycles the last step / adjoint-or-tangent solve needed
    // adjoint (or tangent) fields: the state (l_vel, l_S, l_S2 = seed, then the result) and
    // the intermediates of one step
    enum { L_VEL, L_S, L_S2, L_VF, L_VY, L_SY, L_S2Y, L_VORT, L_VH, L_SH, L_S2H, L_LAP, L_COUNT };
    float4* l[L_COUNT] = {};
    double* gspot = nullptr;         // FS_MAX_SPOTS
    double* gface = nullptr;         // nx * nz
    double* xlin = nullptr;          // the last adjoint / tangent pressure solution: the next one's start
    float* spot_dir = nullptr;       // FS_MAX_SPOTS (tangent runs)
    float* face_dir = nullptr;       // nx * nz (tangent runs)
    float* target = nullptr;         // 2 * ncell (mismatch targets)
    double* partial = nullptr;       // per-block loss partials
    std::vector<void*> owned;

    const AdjSlot& slot(int k) const { return k % every == 0 ? ckpt[k / every] : seg[k % every]; }
};

namespace {

int adj_alloc(fs_adjoint* a, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return fail(FS_ENOMEM, "hipMalloc failed for adjoint storage");
    a->owned.push_back(*p);
    return FS_OK;
}
void adj_free(fs_adjoint* a) {
    for (void* p : a->owned) (void)hipFree(p);
    a->owned.clear();
}
int adj_slot_alloc(fs_adjoint* a, AdjSlot& s) {
    const size_t fb = a->h->ncell * sizeof(float4), pb = mg_level_floats(a->h->mg, 0) * sizeof(double);
    if (adj_alloc(a, (void**)&s.vel, fb) || adj_alloc(a, (void**)&s.S, fb) || adj_alloc(a, (void**)&s.S2, fb) ||
        adj_alloc(a, (void**)&s.p, pb))
        return FS_ENOMEM;
    return FS_OK;
}
int adj_rec_alloc(fs_adjoint* a, StepRec& r) {
    const size_t fb = a->h->ncell * sizeof(float4), pb = mg_level_floats(a->h->mg, 0) * sizeof(double);
    float4** f[8] = {&r.vX, &r.sX, &r.s2X, &r.vY, &r.sY, &r.s2Y, &r.vort, &r.vF};
    for (float4** q : f) if (adj_alloc(a, (void**)q, fb)) return FS_ENOMEM;
    return adj_alloc(a, (void**)&r.p, pb);
}
int adj_check(const fs_sim* h) {
    if (h->prec != FS_PREC_FP32) return fail(FS_ESTATE, "the adjoint needs FS_PREC_FP32 storage");
    if (h->cfg.solve != FS_SOLVE_REFINE64) return fail(FS_ESTATE, "the adjoint needs the converged solve FS_SOLVE_REFINE64");
    return FS_OK;
}
// FP64-refined solve of A x = bkeep from the current x64, run from the host like the
// adaptive step: as many iterations as the last solve of this kind needed, then the
// device's stopping test is read and the solve continued if needed (the iterations past
// convergence would only launch and exit).  The result equals mg_refine's.
int refine_solve(fs_sim* h, double tol, int maxc, int* predict) {
    hipStream_t s = h->stream;
    maxc = std::min(maxc, MG_MAX_REFINE);
    mg_refine_begin(h->mg, s, tol, maxc);
    int k = 0;
    for (; k < std::min(*predict, maxc); k++) mg_refine_iter(h->mg, s, tol, maxc);
    while (true) {
        mg_state_ready(h->mg, s, tol, maxc);
        FS_HIP_CHECK(hipMemcpyAsync(h->host_state, &h->mg.st->used, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        FS_HIP_CHECK(hipStreamSynchronize(s));
        if (h->host_state[1] || k >= maxc) break;
        mg_refine_iter(h->mg, s, tol, maxc);
        k++;
    }
    *predict = std::max(1, h->host_state[0]);
    return FS_OK;
}

// copies between the sim's own state and a slot
void slot_copy(fs_sim* h, const AdjSlot& s, bool to_slot) {
    const size_t fb = h->ncell * sizeof(float4), pb = mg_level_floats(h->mg, 0) * sizeof(double);
    void* sim[4] = {h->f[F_VEL], h->f[F_S], h->f[F_S2], h->mg.x64};
    void* sl[4] = {s.vel, s.S, s.S2, s.p};
    const size_t nb[4] = {fb, fb, fb, pb};
    if (!to_slot) h->gen++;
    for (int i = 0; i < 4; i++)
        (void)hipMemcpyAsync(to_slot ? sl[i] : sim[i], to_slot ? sim[i] : sl[i], nb[i], hipMemcpyDeviceToDevice, h->stream);
}

// Step k from slot `in` to slot `out` (sim.js:633-644), its intermediates written to
// rec[k % every]: adv1 -> vX/sX/s2X, adv2 -> vY/sY/s2Y, curl -> vort, forces -> vF and
// out.S/S2, div -> bkeep, solve (from in.p) -> p, projection -> out.vel, out.p.
void replay_step(fs_adjoint* a, int k, const AdjSlot& in, const AdjSlot& out) {
    fs_sim* h = a->h;
    hipStream_t s = h->stream;
    const StepRec& r = a->rec[k % a->every];
    h->time = a->times[k];
    h->frame = a->frames[k];
    push_params(h, a->dt);
    FieldSet A = h->fields();
    A.vel = in.vel; A.S = in.S; A.S2 = in.S2;
    A.velX = r.vX; A.SX = r.sX; A.S2X = r.s2X;
    A.velY = r.vY; A.SY = r.sY; A.S2Y = r.s2Y;
    A.vort = r.vort;
    launch_adv1(h->prec, h->g, s, h->P_d, A);
    launch_adv2(h->prec, h->g, s, h->P_d, A);
    launch_curl(h->prec, h->g, s, A);
    FieldSet B = A;
    B.velX = r.vF; B.S = out.S; B.S2 = out.S2;
    launch_forces(h->prec, h->g, s, h->P_d, B);
    launch_div(h->prec, h->g, s, h->P_d, B, h->mg.bkeep, 0);
    const size_t pb = mg_level_floats(h->mg, 0) * sizeof(double);
    (void)hipMemcpyAsync(h->mg.x64, in.p, pb, hipMemcpyDeviceToDevice, s);
    refine_solve(h, h->cfg.div_tol, h->cfg.max_vcycles, &a->predict_step);
    FieldSet C = B;
    C.vel = out.vel;
    launch_project(h->prec, h->g, s, h->P_d, C, nullptr, h->mg.x64);
    (void)hipMemcpyAsync(r.p, h->mg.x64, pb, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(out.p, h->mg.x64, pb, hipMemcpyDeviceToDevice, s);
}

// The adjoint of step k: l_vel/l_S/l_S2 hold the adjoint of the state after the step on
// entry and of the state before it on return; parameter gradients accumulate.
void reverse_step(fs_adjoint* a, int k, double tol, int maxc) {
    fs_sim* h = a->h;
    hipStream_t s = h->stream;
    const Dims& g = h->g;
    const AdjSlot& in = a->slot(k);
    const StepRec& r = a->rec[k % a->every];
    h->time = a->times[k];
    h->frame = a->frames[k];
    push_params(h, a->dt);
    float4** l = a->l;
    const size_t fb = h->ncell * sizeof(float4), nd = mg_level_floats(h->mg, 0);
    const FieldSet F = h->fields();
    // projection: vF, p -> vel (sim.js:526-550); the pressure adjoint goes to bkeep
    launch_fill_zero(s, h->mg.bkeep, nd * sizeof(float));
    launch_project_adj(g, s, h->P_d, r.vF, r.p, (const float4*)F.geo, l[fs_adjoint::L_VEL], l[fs_adjoint::L_VF],
                       h->mg.bkeep);
    // the pressure solve: A lb = lp, started from the previous step's solution (the
    // adjoint pressure changes little from step to step)
    (void)hipMemcpyAsync(h->mg.x64, a->xlin, nd * sizeof(double), hipMemcpyDeviceToDevice, s);
    refine_solve(h, tol, maxc, &a->predict_lin);
    (void)hipMemcpyAsync(a->xlin, h->mg.x64, nd * sizeof(double), hipMemcpyDeviceToDevice, s);
    // divergence (sim.js:399-413)
    launch_div_adj(g, s, h->P_d, h->mg.x64, l[fs_adjoint::L_VF]);
    // sources (sim.js:214-396)
    launch_fill_zero(s, l[fs_adjoint::L_VORT], fb);
    launch_forces_adj(g, s, h->P_d, F, r.vY, r.sY, r.s2Y, r.vort, l[fs_adjoint::L_VF], l[fs_adjoint::L_S],
                      l[fs_adjoint::L_S2], l[fs_adjoint::L_VY], l[fs_adjoint::L_SY], l[fs_adjoint::L_S2Y],
                      l[fs_adjoint::L_VORT], l[fs_adjoint::L_LAP], h->prm[P_visc] > 0.f, a->gspot, a->gface);
    // curl (sim.js:200-211)
    launch_curl_adj(g, s, r.vY, l[fs_adjoint::L_VORT], l[fs_adjoint::L_VY]);
    // MacCormack correction and semi-Lagrangian pass (sim.js:142-197)
    for (int i : {fs_adjoint::L_VEL, fs_adjoint::L_S, fs_adjoint::L_S2, fs_adjoint::L_VH, fs_adjoint::L_SH, fs_adjoint::L_S2H})
        launch_fill_zero(s, l[i], fb);
    launch_adv2_adj(g, s, h->P_d, in.vel, in.S, in.S2, r.vX, r.sX, r.s2X, l[fs_adjoint::L_VY], l[fs_adjoint::L_SY], l[fs_adjoint::L_S2Y], l[fs_adjoint::L_VEL],
                    l[fs_adjoint::L_S], l[fs_adjoint::L_S2], l[fs_adjoint::L_VH], l[fs_adjoint::L_SH], l[fs_adjoint::L_S2H]);
    launch_adv1_adj(g, s, h->P_d, in.vel, in.S, in.S2, l[fs_adjoint::L_VH], l[fs_adjoint::L_SH], l[fs_adjoint::L_S2H],
                    l[fs_adjoint::L_VEL], l[fs_adjoint::L_S], l[fs_adjoint::L_S2]);
}

float4* adj_field(fs_adjoint* a, int field) {
    switch (field) {
        case FS_VEL: return a->l[fs_adjoint::L_VEL];
        case FS_S: return a->l[fs_adjoint::L_S];
        case FS_S2: return a->l[fs_adjoint::L_S2];
    }
    return nullptr;
}

}  // namespace

extern "C" {

int fs_adjoint_create(fs_sim* h, int max_steps, int checkpoint_every, fs_adjoint** out) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    *out = nullptr;
    if (max_steps < 1 || checkpoint_every < 1) return fail(FS_EINVAL, "max_steps and checkpoint_every must be >= 1");
    if (adj_check(h)) return FS_ESTATE;
    if (set_device(h)) return FS_EDEVICE;
    fs_adjoint* a = new fs_adjoint();
    a->h = h;
    a->max_steps = max_steps;
    a->every = std::min(checkpoint_every, max_steps);
    auto bail = [&](int code) { std::string e = g_last_error; adj_free(a); delete a; g_last_error = e; return code; };
    a->ckpt.resize((max_steps + a->every - 1) / a->every + 1);
    a->seg.resize(a->every);
    for (auto& s : a->ckpt) if (adj_slot_alloc(a, s)) return bail(FS_ENOMEM);
    for (int i = 1; i < a->every; i++) if (adj_slot_alloc(a, a->seg[i])) return bail(FS_ENOMEM);
    if (adj_slot_alloc(a, a->fin)) return bail(FS_ENOMEM);
    a->rec.resize(a->every);
    for (auto& r : a->rec) if (adj_rec_alloc(a, r)) return bail(FS_ENOMEM);
    const size_t fb = h->ncell * sizeof(float4);
    for (int i = 0; i < fs_adjoint::L_COUNT; i++) {
        if (adj_alloc(a, (void**)&a->l[i], fb)) return bail(FS_ENOMEM);
        launch_fill_zero(h->stream, a->l[i], fb);
    }
    if (adj_alloc(a, (void**)&a->gspot, FS_MAX_SPOTS * sizeof(double)) ||
        adj_alloc(a, (void**)&a->gface, (size_t)h->nx * h->nz * sizeof(double)) ||
        adj_alloc(a, (void**)&a->spot_dir, FS_MAX_SPOTS * sizeof(float)) ||
        adj_alloc(a, (void**)&a->xlin, mg_level_floats(h->mg, 0) * sizeof(double)) ||
        adj_alloc(a, (void**)&a->face_dir, (size_t)h->nx * h->nz * sizeof(float)) ||
        adj_alloc(a, (void**)&a->target, 2 * h->ncell * sizeof(float)) ||
        adj_alloc(a, (void**)&a->partial, (size_t)h->g.nblocks * sizeof(double)))
        return bail(FS_ENOMEM);
    launch_fill_zero(h->stream, a->gspot, FS_MAX_SPOTS * sizeof(double));
    launch_fill_zero(h->stream, a->gface, (size_t)h->nx * h->nz * sizeof(double));
    a->times.resize(max_steps + 1);
    a->frames.resize(max_steps + 1);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(FS_EDEVICE, "device error in adjoint create"));
    *out = a;
    return FS_OK;
}

int fs_adjoint_destroy(fs_adjoint* a) {
    if (!a) return FS_OK;
    if (a->h) { (void)set_device(a->h); (void)hipStreamSynchronize(a->h->stream); }
    adj_free(a);
    delete a;
    return FS_OK;
}

}  // extern "C"

namespace {

// The tangent of step k, after replay_step(k, in, &out): l[L_VEL/L_S/L_S2] hold the state
// tangent before the step on entry and after it on return.
void tangent_step(fs_adjoint* a, int k, const AdjSlot& in, double tol, int maxc) {
    fs_sim* h = a->h;
    hipStream_t s = h->stream;
    const Dims& g = h->g;
    float4** l = a->l;
    const FieldSet F = h->fields();
    const StepRec& r = a->rec[k % a->every];
    const size_t nd = mg_level_floats(h->mg, 0);
    launch_adv1_tan(g, s, h->P_d, in.vel, in.S, in.S2, l[fs_adjoint::L_VEL], l[fs_adjoint::L_S], l[fs_adjoint::L_S2],
                    l[fs_adjoint::L_VH], l[fs_adjoint::L_SH], l[fs_adjoint::L_S2H]);
    launch_adv2_tan(g, s, h->P_d, in.vel, in.S, in.S2, r.vX, r.sX, r.s2X, l[fs_adjoint::L_VEL], l[fs_adjoint::L_S], l[fs_adjoint::L_S2], l[fs_adjoint::L_VH],
                    l[fs_adjoint::L_SH], l[fs_adjoint::L_S2H], l[fs_adjoint::L_VY], l[fs_adjoint::L_SY],
                    l[fs_adjoint::L_S2Y]);
    launch_curl_tan(g, s, r.vY, l[fs_adjoint::L_VY], l[fs_adjoint::L_VORT]);
    launch_forces_tan(g, s, h->P_d, F, r.vY, r.sY, r.s2Y, r.vort, a->spot_dir, a->face_dir, l[fs_adjoint::L_VY], l[fs_adjoint::L_SY],
                      l[fs_adjoint::L_S2Y], l[fs_adjoint::L_VORT], l[fs_adjoint::L_VF], l[fs_adjoint::L_S],
                      l[fs_adjoint::L_S2]);
    launch_div_tan(g, s, h->P_d, l[fs_adjoint::L_VF], h->mg.bkeep);
    (void)hipMemcpyAsync(h->mg.x64, a->xlin, nd * sizeof(double), hipMemcpyDeviceToDevice, s);
    refine_solve(h, tol, maxc, &a->predict_lin);
    (void)hipMemcpyAsync(a->xlin, h->mg.x64, nd * sizeof(double), hipMemcpyDeviceToDevice, s);
    launch_project_tan(g, s, h->P_d, r.vF, r.p, (const float4*)F.geo, l[fs_adjoint::L_VF], h->mg.x64,
                       l[fs_adjoint::L_VEL]);
}

// nsteps steps from the sim's state, recording the checkpoints; with `tangent`, the state
// tangent in l[L_VEL/L_S/L_S2] and the flux directions are carried along
int run_forward(fs_adjoint* a, double dt, int nsteps, bool tangent, double tol, int maxc) {
    fs_sim* h = a->h;
    a->dt = dt;
    a->nsteps = nsteps;
    a->have_forward = a->have_grads = a->have_tangent = false;
    sync_params(h, dt);
    if (!h->geo_ready) {   // sim.js:632
        launch_geo(h->prec, h->g, h->stream, h->P_d, h->fields());
        h->geo_ready = true;
    }
    slot_copy(h, a->slot(0), true);
    for (int k = 0; k < nsteps; k++) {
        a->times[k] = h->time;
        a->frames[k] = h->frame;
        replay_step(a, k, a->slot(k), a->slot(k + 1));
        if (tangent) tangent_step(a, k, a->slot(k), tol, maxc);
        h->time = a->times[k] + dt;   // sim.js:646-647
        h->frame = a->frames[k] + 1;
    }
    a->t_end = h->time;
    a->frame_end = h->frame;
    slot_copy(h, a->slot(nsteps), false);
    slot_copy(h, a->fin, true);
    h->params_dirty = true;
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    a->have_forward = true;
    a->have_tangent = tangent;
    a->seeds_stale = true;
    a->last_interval_held = true;
    return FS_OK;
}

int forward_args(fs_adjoint* a, double dt, int nsteps) {
    if (!a) return fail(FS_EINVAL, "null handle");
    if (nsteps < 1 || nsteps > a->max_steps) return fail(FS_EINVAL, "nsteps must be in 1..max_steps");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(FS_EINVAL, "dt must be positive and finite");
    if (adj_check(a->h)) return FS_ESTATE;
    if (set_device(a->h)) return FS_EDEVICE;
    return FS_OK;
}

}  // namespace

extern "C" {

int fs_adjoint_forward(fs_adjoint* a, double dt, int nsteps) {
    if (int rc = forward_args(a, dt, nsteps)) return rc;
    return run_forward(a, dt, nsteps, false, 0.0, 0);
}

int fs_adjoint_tangent(fs_adjoint* a, double dt, int nsteps, const double* spot_dir, const float* face_dir,
                       const float* vel_dir, const float* S_dir, const float* S2_dir, double solve_tol, int max_cycles) {
    if (int rc = forward_args(a, dt, nsteps)) return rc;
    fs_sim* h = a->h;
    const size_t n = h->ncell, fb = n * sizeof(float4);
    const int ns = (int)h->spots_h.size();
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    std::vector<float> sd(FS_MAX_SPOTS, 0.f);
    if (spot_dir) for (int q = 0; q < ns; q++) sd[q] = (float)spot_dir[q];
    FS_HIP_CHECK(hipMemcpy(a->spot_dir, sd.data(), FS_MAX_SPOTS * sizeof(float), hipMemcpyHostToDevice));
    if (face_dir) FS_HIP_CHECK(hipMemcpy(a->face_dir, face_dir, (size_t)h->nx * h->nz * sizeof(float), hipMemcpyHostToDevice));
    else launch_fill_zero(h->stream, a->face_dir, (size_t)h->nx * h->nz * sizeof(float));
    const float* dirs[3] = {vel_dir, S_dir, S2_dir};
    const int slots[3] = {fs_adjoint::L_VEL, fs_adjoint::L_S, fs_adjoint::L_S2};
    for (int i = 0; i < 3; i++) {
        if (dirs[i]) FS_HIP_CHECK(hipMemcpy(a->l[slots[i]], dirs[i], fb, hipMemcpyHostToDevice));
        else launch_fill_zero(h->stream, a->l[slots[i]], fb);
    }
    launch_fill_zero(h->stream, a->xlin, mg_level_floats(h->mg, 0) * sizeof(double));
    const double tol = solve_tol > 0.0 ? solve_tol : h->cfg.div_tol;
    const int maxc = std::min(max_cycles > 0 ? max_cycles : h->cfg.max_vcycles, MG_MAX_REFINE);
    return run_forward(a, dt, nsteps, true, tol, maxc);
}

int fs_adjoint_seed(fs_adjoint* a, int field, int dtype, const void* host, size_t bytes) {
    if (!a || !host) return fail(FS_EINVAL, "null argument");
    float4* dst = adj_field(a, field);
    if (!dst) return fail(FS_EINVAL, "the seed field must be FS_VEL, FS_S or FS_S2");
    if (!a->have_forward) return fail(FS_ESTATE, "run fs_adjoint_forward first");
    fs_sim* h = a->h;
    const size_t nval = h->ncell * 4;
    if (dtype != FS_DT_F32 && dtype != FS_DT_F64) return fail(FS_EINVAL, "dtype must be FS_DT_F32 or FS_DT_F64");
    if (bytes != nval * (dtype == FS_DT_F32 ? 4 : 8)) return fail(FS_EINVAL, "seed size mismatch");
    if (set_device(h)) return FS_EDEVICE;
    std::vector<float> tmp(nval);
    if (dtype == FS_DT_F32) std::memcpy(tmp.data(), host, nval * 4);
    else for (size_t i = 0; i < nval; i++) tmp[i] = (float)((const double*)host)[i];
    if (a->seeds_stale) {
        for (int i : {fs_adjoint::L_VEL, fs_adjoint::L_S, fs_adjoint::L_S2}) launch_fill_zero(h->stream, a->l[i], nval * 4);
        a->seeds_stale = false;
    }
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    FS_HIP_CHECK(hipMemcpy(dst, tmp.data(), nval * 4, hipMemcpyHostToDevice));
    a->have_grads = a->have_tangent = false;
    return FS_OK;
}

int fs_adjoint_seed_mismatch(fs_adjoint* a, const float* target_theta, const float* target_soot, double* loss) {
    if (!a || !target_theta || !target_soot) return fail(FS_EINVAL, "null argument");
    if (!a->have_forward) return fail(FS_ESTATE, "run fs_adjoint_forward first");
    fs_sim* h = a->h;
    if (set_device(h)) return FS_EDEVICE;
    const size_t n = h->ncell, fb = n * sizeof(float4);
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    FS_HIP_CHECK(hipMemcpy(a->target, target_theta, n * sizeof(float), hipMemcpyHostToDevice));
    FS_HIP_CHECK(hipMemcpy(a->target + n, target_soot, n * sizeof(float), hipMemcpyHostToDevice));
    launch_fill_zero(h->stream, a->l[fs_adjoint::L_VEL], fb);
    launch_fill_zero(h->stream, a->l[fs_adjoint::L_S2], fb);
    launch_loss_seed(h->g, h->stream, a->fin.S, a->target, a->target + n, a->l[fs_adjoint::L_S], a->partial);
    std::vector<double> part(h->g.nblocks);
    FS_HIP_CHECK(hipMemcpyAsync(part.data(), a->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    double sum = 0.0;
    for (double v : part) sum += v;
    if (loss) *loss = sum;
    a->have_grads = a->have_tangent = false;
    a->seeds_stale = false;
    return FS_OK;
}

int fs_adjoint_backward(fs_adjoint* a, double solve_tol, int max_cycles) {
    if (!a) return fail(FS_EINVAL, "null handle");
    if (!a->have_forward) return fail(FS_ESTATE, "run fs_adjoint_forward first");
    fs_sim* h = a->h;
    if (adj_check(h)) return FS_ESTATE;
    if (set_device(h)) return FS_EDEVICE;
    const double tol = solve_tol > 0.0 ? solve_tol : h->cfg.div_tol;
    const int maxc = std::min(max_cycles > 0 ? max_cycles : h->cfg.max_vcycles, MG_MAX_REFINE);
    launch_fill_zero(h->stream, a->gspot, FS_MAX_SPOTS * sizeof(double));
    launch_fill_zero(h->stream, a->gface, (size_t)h->nx * h->nz * sizeof(double));
    launch_fill_zero(h->stream, a->xlin, mg_level_floats(h->mg, 0) * sizeof(double));
    if (a->seeds_stale) {   // nothing seeded since the forward: the seed is zero
        for (int i : {fs_adjoint::L_VEL, fs_adjoint::L_S, fs_adjoint::L_S2})
            launch_fill_zero(h->stream, a->l[i], h->ncell * sizeof(float4));
        a->seeds_stale = false;
    }
    a->have_tangent = false;
    const int N = a->nsteps, E = a->every;
    const int nseg = (N + E - 1) / E;
    for (int j = nseg - 1; j >= 0; j--) {
        const int k0 = j * E, k1 = std::min(k0 + E, N);
        // the states inside the interval; the forward left the last interval's in place
        if (j != nseg - 1 || !a->last_interval_held) {
            // deterministic kernels: the states and checkpoints rewritten here are bit-identical
            for (int k = k0; k < k1; k++) replay_step(a, k, a->slot(k), a->slot(k + 1));
            a->last_interval_held = j == nseg - 1;
        }
        for (int k = k1 - 1; k >= k0; k--) reverse_step(a, k, tol, maxc);
        FS_HIP_CHECK(hipGetLastError());
    }
    // the sim's state, pressure and clock as the forward left them
    slot_copy(h, a->fin, false);
    h->time = a->t_end;
    h->frame = a->frame_end;
    h->params_dirty = true;
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    a->have_grads = true;
    a->seeds_stale = true;   // the seed buffers now hold the initial state's adjoint
    return FS_OK;
}

int fs_adjoint_read(fs_adjoint* a, int field, int dtype, void* host, size_t bytes) {
    if (!a || !host) return fail(FS_EINVAL, "null argument");
    float4* src = adj_field(a, field);
    if (!src) return fail(FS_EINVAL, "the field must be FS_VEL, FS_S or FS_S2");
    if (!a->have_grads && !a->have_tangent) return fail(FS_ESTATE, "run fs_adjoint_backward (or fs_adjoint_tangent) first");
    fs_sim* h = a->h;
    const size_t nval = h->ncell * 4;
    if (dtype != FS_DT_F32 && dtype != FS_DT_F64) return fail(FS_EINVAL, "dtype must be FS_DT_F32 or FS_DT_F64");
    if (bytes != nval * (dtype == FS_DT_F32 ? 4 : 8)) return fail(FS_EINVAL, "destination size mismatch");
    if (set_device(h)) return FS_EDEVICE;
    std::vector<float> tmp(nval);
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    FS_HIP_CHECK(hipMemcpy(tmp.data(), src, nval * 4, hipMemcpyDeviceToHost));
    if (dtype == FS_DT_F32) std::memcpy(host, tmp.data(), nval * 4);
    else for (size_t i = 0; i < nval; i++) ((double*)host)[i] = (double)tmp[i];
    return FS_OK;
}

int fs_adjoint_grads(fs_adjoint* a, double* spot_flux, int max_spots, double* face_flux) {
    if (!a) return fail(FS_EINVAL, "null handle");
    if (!a->have_grads) return fail(FS_ESTATE, "run fs_adjoint_backward first");
    fs_sim* h = a->h;
    if (set_device(h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(h->stream));
    const int n = std::min(max_spots, (int)h->spots_h.size());
    if (spot_flux && n > 0) FS_HIP_CHECK(hipMemcpy(spot_flux, a->gspot, n * sizeof(double), hipMemcpyDeviceToHost));
    if (face_flux) {
        if (!h->has_face) return fail(FS_ESTATE, "no floor flux is set");
        FS_HIP_CHECK(hipMemcpy(face_flux, a->gface, (size_t)h->nx * h->nz * sizeof(double), hipMemcpyDeviceToHost));
    }
    return FS_OK;
}

}  // extern "C"

// ---- spark particles (src/sparks.js) -----------------------------------------------
struct fs_sparks {
    fs_sim* h = nullptr;
    int device = 0;                  // for destroy, which must not touch the sim (it may be gone)
    int count = 0;
    float4* sp = nullptr;            // count x 4 float4
    int* dead = nullptr;             // per-block dead counts (kept current by the update kernel)
    int* off = nullptr;              // their exclusive scan: the blocks' first spawn slots
    bool counts_valid = false;       // false after a host write / reset: count before the next step
    double rate = 3.0;               // baseline sparks per second (src/sparks.js:24)
    double carry = 0.0;
    float turb = 1.f, lifeK = 1.f, sizeK = 1.f, upK = 1.f, brightness = 13.f;   // src/sparks.js:29
    struct Burst { double t0; int n; float pos[3]; float speed, radius; double dur; float fan, size; bool has_aim; float aim[3]; int emitted; };
    std::vector<Burst> bursts;
    double time = 0.0;
    int64_t frame = 0;
};

extern "C" {

void fs_spark_burst_default(fs_spark_burst* b) {   // src/sparks.js:44 defaults
    if (!b) return;
    b->n = 60; b->pos[0] = 0.f; b->pos[1] = 0.1f; b->pos[2] = 0.f;
    b->speed = 3.f; b->radius = 0.05f; b->dur = 0.12; b->fan = 0.6f; b->size = 1.3f;
    b->has_aim = 0; b->aim[0] = b->aim[1] = b->aim[2] = 0.f;
}

int fs_sparks_create(fs_sim* h, int count, fs_sparks** out) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    *out = nullptr;
    if (count < 1 || count > (1 << 24)) return fail(FS_EINVAL, "count must be in 1..2^24");
    if (set_device(h)) return FS_EDEVICE;
    fs_sparks* s = new fs_sparks();
    s->h = h;
    s->device = h->cfg.device;
    s->count = count;
    const size_t nb = (size_t)spark_blocks(count) * sizeof(int);
    if (hipMalloc(&s->sp, (size_t)count * 64) != hipSuccess || hipMalloc(&s->dead, nb) != hipSuccess ||
        hipMalloc(&s->off, nb) != hipSuccess) {
        if (s->sp) (void)hipFree(s->sp);
        if (s->dead) (void)hipFree(s->dead);
        delete s;
        return fail(FS_ENOMEM, "hipMalloc failed for sparks");
    }
    launch_fill_zero(h->stream, s->sp, (size_t)count * 64);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        (void)hipFree(s->sp); (void)hipFree(s->dead); (void)hipFree(s->off); delete s;
        return fail(FS_EDEVICE, "device error in sparks create");
    }
    *out = s;
    return FS_OK;
}

int fs_sparks_destroy(fs_sparks* s) {
    if (!s) return FS_OK;
    (void)hipSetDevice(s->device);   // hipFree waits for the device's work
    (void)hipFree(s->sp);
    (void)hipFree(s->dead);
    (void)hipFree(s->off);
    delete s;
    return FS_OK;
}

int fs_sparks_reset(fs_sparks* s) {   // src/sparks.js:38-41
    if (!s) return fail(FS_EINVAL, "null handle");
    if (set_device(s->h)) return FS_EDEVICE;
    launch_fill_zero(s->h->stream, s->sp, (size_t)s->count * 64);
    s->counts_valid = false;
    s->carry = 0.0; s->bursts.clear(); s->time = 0.0; s->frame = 0;
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_sparks_set(fs_sparks* s, const char* name, float value) {
    if (!s || !name) return fail(FS_EINVAL, "null argument");
    const std::string n(name);
    if (n == "rate") s->rate = value;
    else if (n == "turb") s->turb = value;
    else if (n == "lifeK") s->lifeK = value;
    else if (n == "sizeK") s->sizeK = value;
    else if (n == "upK") s->upK = value;
    else if (n == "brightness") s->brightness = value;
    else return fail(FS_EINVAL, "unknown spark parameter: " + n);
    return FS_OK;
}

int fs_sparks_burst(fs_sparks* s, const fs_spark_burst* b) {   // src/sparks.js:44-46
    if (!s || !b) return fail(FS_EINVAL, "null argument");
    if (b->n < 0 || !(b->dur > 0.0)) return fail(FS_EINVAL, "burst needs n >= 0 and dur > 0");
    fs_sparks::Burst q;
    q.t0 = s->time; q.n = b->n;
    for (int k = 0; k < 3; k++) { q.pos[k] = b->pos[k]; q.aim[k] = b->aim[k]; }
    q.speed = b->speed; q.radius = b->radius; q.dur = b->dur; q.fan = b->fan; q.size = b->size;
    q.has_aim = b->has_aim != 0;
    q.emitted = 0;
    s->bursts.push_back(q);
    return FS_OK;
}

int fs_sparks_step(fs_sparks* s, double dt) {   // src/sparks.js:241-271
    if (!s) return fail(FS_EINVAL, "null handle");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(FS_EINVAL, "dt must be positive and finite");
    fs_sim* h = s->h;
    if (set_device(h)) return FS_EDEVICE;
    s->carry += s->rate * dt;
    const double nBase = std::floor(s->carry);
    s->carry -= nBase;
    int nBurst = 0;
    const fs_sparks::Burst* B = nullptr;
    for (auto& b : s->bursts) {
        const double frac = std::min((s->time + dt - b.t0) / b.dur, 1.0);
        const int due = (int)std::min((double)b.n, std::floor((double)b.n * frac + 0.5));   // Math.round
        const int k = due - b.emitted;
        if (k > 0 && !B) { B = &b; nBurst = k; b.emitted += k; }
    }
    fs_sparks::Burst Bc{};
    if (B) Bc = *B;
    s->bursts.erase(std::remove_if(s->bursts.begin(), s->bursts.end(),
                                   [](const fs_sparks::Burst& b) { return b.emitted >= b.n; }),
                    s->bursts.end());
    SparkUniform U;
    U.dt = (float)dt; U.time = (float)s->time; U.frame = (float)s->frame; U.count = (float)s->count;
    U.nBase = (float)nBase; U.nBurst = (float)nBurst;
    U.burstX = B ? Bc.pos[0] : 0.f; U.burstY = B ? Bc.pos[1] : 0.f; U.burstZ = B ? Bc.pos[2] : 0.f;
    U.burstSpeed = B ? Bc.speed : 0.f; U.burstRadius = B ? Bc.radius : 0.f;
    U.burstFan = B ? Bc.fan : 0.f; U.burstSize = B ? Bc.size : 1.f;
    U.aimX = B && Bc.has_aim ? Bc.aim[0] : 0.f; U.aimZ = B && Bc.has_aim ? Bc.aim[2] : 0.f;
    U.ox = (float)h->origin[0]; U.oy = (float)h->origin[1]; U.oz = (float)h->origin[2];
    U.sx = h->cfg.size[0]; U.sy = h->cfg.size[1]; U.sz = h->cfg.size[2];
    U.turb = s->turb; U.lifeK = s->lifeK; U.bedRadius = 0.2f; U.seed = 1234.f; U.sizeK = s->sizeK; U.upK = s->upK;
    if (!s->counts_valid) launch_spark_count(h->stream, s->sp, s->count, s->dead);
    launch_spark_step(h->prec, h->stream, U, s->sp, s->dead, s->off, h->f[F_VEL], h->g, h->curl, h->cfg.noise_n);
    s->counts_valid = true;
    s->time += dt;
    s->frame += 1;
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_sparks_read(fs_sparks* s, float* host, size_t bytes) {
    if (!s || !host) return fail(FS_EINVAL, "null argument");
    if (bytes != (size_t)s->count * 64) return fail(FS_EINVAL, "size mismatch: expected count * 64 bytes");
    if (set_device(s->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(s->h->stream));
    FS_HIP_CHECK(hipMemcpy(host, s->sp, bytes, hipMemcpyDeviceToHost));
    return FS_OK;
}

int fs_sparks_write(fs_sparks* s, const float* host, size_t bytes) {
    if (!s || !host) return fail(FS_EINVAL, "null argument");
    if (bytes != (size_t)s->count * 64) return fail(FS_EINVAL, "size mismatch: expected count * 64 bytes");
    if (set_device(s->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(s->h->stream));
    FS_HIP_CHECK(hipMemcpy(s->sp, host, bytes, hipMemcpyHostToDevice));
    s->counts_valid = false;
    return FS_OK;
}

int fs_sparks_clock(fs_sparks* s, double* time, int64_t* frame) {
    if (!s) return fail(FS_EINVAL, "null handle");
    if (time) *time = s->time;
    if (frame) *frame = s->frame;
    return FS_OK;
}

void* fs_sparks_device_ptr(fs_sparks* s) { return s ? (void*)s->sp : nullptr; }

}  // extern "C"

// ---- render-volume consumers (src/renderer.js) ---------------------------------------
namespace {
// RP_FIELDS (src/renderer.js:8-17)
const char* const kRenderNames[FS_RP_COUNT] = {
    "width", "height", "vw", "vh", "time", "frame", "nLogs", "nStones",
    "camX", "camY", "camZ", "tanFov", "fwdX", "fwdY", "fwdZ", "aspect", "rgtX", "rgtY", "rgtZ", "jitX", "upX", "upY", "upZ", "jitY",
    "ox", "oy", "oz", "dx", "sx", "sy", "sz", "tempScale", "nx", "ny", "nz", "nLights",
    "exposure", "stepMul", "sootAbs", "sootEmit", "emitK", "blueK", "smokeScat", "smokeG",
    "ambient", "detailAmt", "detailFreq", "detailSpeed", "hazeK", "glowK", "bloomK", "tonemap",
    "grain", "vignette", "saturation", "contrast", "bedGlow", "logGlow", "lightK", "accumN",
    "bbTmin", "bbTmax", "bbN", "emberFlicker", "sootEmitMin", "flameSharp", "moon", "fogK",
    "shadowSoft", "lift", "emitTempLo", "glowTemp", "inset", "knee", "surfDetail", "matA", "matB", "matC", "sootScat",
    "chroma", "lensX", "lensY", "focus", "lightDt", "flickerK", "cubic", "aperture", "sootEdge",
};
// RP_DEFAULTS (src/renderer.js:19-25); entries without a default start at 0
void render_defaults(float* p) {
    std::fill(p, p + FS_RP_COUNT, 0.f);
    p[RP_exposure] = 15.0f; p[RP_stepMul] = 0.7f; p[RP_sootAbs] = 4.0f; p[RP_sootEmit] = 10.0f; p[RP_emitK] = 4.0f;
    p[RP_blueK] = 0.05f; p[RP_smokeScat] = 0.02f; p[RP_smokeG] = 0.5f;
    p[RP_ambient] = 1.0f; p[RP_detailAmt] = 0.6f; p[RP_detailFreq] = 2.2f; p[RP_detailSpeed] = 0.9f; p[RP_hazeK] = 1.0f;
    p[RP_glowK] = 1.0f; p[RP_bloomK] = 0.04f; p[RP_tonemap] = 2.f;
    p[RP_grain] = 0.02f; p[RP_vignette] = 0.6f; p[RP_saturation] = 1.0f; p[RP_contrast] = 1.0f; p[RP_bedGlow] = 1.0f;
    p[RP_logGlow] = 1.0f; p[RP_lightK] = 1.0f;
    p[RP_emberFlicker] = 1.0f; p[RP_sootEmitMin] = 0.0f; p[RP_flameSharp] = 0.0f; p[RP_moon] = 1.0f; p[RP_fogK] = 1.0f;
    p[RP_shadowSoft] = 6.0f; p[RP_lift] = 0.0f;
    p[RP_emitTempLo] = 900.f; p[RP_glowTemp] = 1.0f; p[RP_tempScale] = 1000.f; p[RP_inset] = 0.06f; p[RP_knee] = 0.6f;
    p[RP_surfDetail] = 1.f; p[RP_sootScat] = 0.015f; p[RP_chroma] = 0.0025f; p[RP_lightDt] = -1.f; p[RP_flickerK] = 0.35f;
}
// entries uploadParams owns (src/renderer.js:640-647); setting them through this.params has no effect
bool render_owned(int i) {
    switch (i) {
        case RP_width: case RP_height: case RP_vw: case RP_vh: case RP_time: case RP_frame: case RP_nLogs: case RP_nStones:
        case RP_nLights: case RP_lightDt: case RP_bbTmin: case RP_bbTmax: case RP_bbN: case RP_tanFov: case RP_aspect:
        case RP_jitX: case RP_jitY: case RP_ox: case RP_oy: case RP_oz: case RP_dx: case RP_sx: case RP_sy: case RP_sz:
        case RP_nx: case RP_ny: case RP_nz:
            return true;
    }
    return (i >= RP_camX && i <= RP_camZ) || (i >= RP_fwdX && i <= RP_fwdZ) || (i >= RP_rgtX && i <= RP_rgtZ) ||
           (i >= RP_upX && i <= RP_upZ);
}

// ---- blackbody LUT (src/blackbody.js): Planck's law against the multi-lobe Gaussian fit of the
// CIE 1931 2-degree observer (Wyman, Sloan & Shirley 2013), XYZ -> linear sRGB, Bradford white
// balance from a blackbody at Twb to D65.  All in double, like the reference's JS numbers.
namespace bb {
const double kH = 6.62607015e-34, kC = 2.99792458e8, kB = 1.380649e-23;
double lobe(double x, double mu, double s1, double s2) {
    const double t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5 * t * t);
}
void cie(double l, double o[3]) {
    o[0] = 1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) - 0.065 * lobe(l, 501.1, 20.4, 26.2);
    o[1] = 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
    o[2] = 1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8);
}
double planck(double lnm, double T) {   // spectral radiance, W sr^-1 m^-3
    const double l = lnm * 1e-9;
    return (2 * kH * kC * kC) / (std::pow(l, 5) * (std::exp((kH * kC) / (l * kB * T)) - 1));
}
// 380..780 nm in 2 nm steps; soot: emissivity ~ (lambda / 550 nm)^-1.39 (Hottel-Broughton)
void xyz_of(double T, bool soot, double o[3]) {
    double X = 0, Y = 0, Z = 0;
    for (int l = 380; l <= 780; l += 2) {
        const double p = planck(l, T) * (soot ? std::pow(l / 550.0, -1.39) : 1.0);
        double c[3];
        cie(l, c);
        X += p * c[0]; Y += p * c[1]; Z += p * c[2];
    }
    o[0] = X; o[1] = Y; o[2] = Z;
}
const double kXYZ2RGB[3][3] = {{3.2404542, -1.5371385, -0.4985314}, {-0.9692660, 1.8760108, 0.0415560},
                               {0.0556434, -0.2040259, 1.0572252}};
const double kBradford[3][3] = {{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}};
const double kD65[3] = {0.95047, 1.0, 1.08883};
void mul(const double M[3][3], const double v[3], double o[3]) {
    for (int r = 0; r < 3; r++) o[r] = M[r][0] * v[0] + M[r][1] * v[1] + M[r][2] * v[2];
}
void inverse(const double m[3][3], double o[3][3]) {   // adjugate / determinant
    const double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2];
    const double g = m[2][0], h = m[2][1], i = m[2][2];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    o[0][0] = A / det; o[0][1] = -(b * i - c * h) / det; o[0][2] = (b * f - c * e) / det;
    o[1][0] = B / det; o[1][1] = (a * i - c * g) / det; o[1][2] = -(a * f - c * d) / det;
    o[2][0] = C / det; o[2][1] = -(a * h - b * g) / det; o[2][2] = (a * e - b * d) / det;
}
}  // namespace bb
}  // namespace

struct fs_render {
    fs_sim* h = nullptr;
    int device = 0;
    int w = 0, hgt = 0, vw = 0, vh = 0;
    float params[FS_RP_COUNT];         // this.params
    RenderU U{};                       // the block of the last frame
    bool have_U = false;
    bool have_cam = false;
    double cam_pos[3], cam_fwd[3], cam_rgt[3], cam_up[3], fov_y = 0.0;
    int bbN = 1024;
    double bbTmin = 400.0, bbTmax = 3000.0;
    float4* lut = nullptr;             // 2 x bbN
    float4* lstate = nullptr;          // 3 x 8 light filter state
    float4* lights = nullptr;          // 8 x (pos, col)
    float* depth = nullptr;            // dw x dh
    int dw = 1, dh = 1;
    size_t depth_cap = 1;
    half4_t* sharp = nullptr;          // prefiltered S
    half4_t* vol = nullptr;            // vw x vh
    float2* aux = nullptr;
    bool have_last_light = false;
    double last_light_time = 0.0;
    bool sharp_valid = false;          // prefilter stamp (src/renderer.js:679)
    uint64_t sharp_gen = 0;
    int64_t sharp_frame = 0;
    double sharp_time = 0.0;
    float sharp_k = 0.f;
    bool skip = true;                  // empty-space map of the march
    int bx = 0, by = 0, bz = 0;        // bricks of 4^3 cells
    unsigned char* raw = nullptr;      // bricks holding an active cell
    unsigned char* flags = nullptr;    // ... within the march's reach
    float curl_max = 0.f;              // largest |component| of the curl texture's xyz
    bool profile = false;              // HIP events around the passes of each frame
    hipEvent_t ev[4] = {};
    bool timings_valid = false;
    bool sharp_ran = false;            // the last frame ran the prefilter
};

namespace {
void render_free(fs_render* r) {
    void* p[] = {r->lut, r->lstate, r->lights, r->depth, r->sharp, r->vol, r->aux, r->raw, r->flags};
    for (void* q : p) if (q) (void)hipFree(q);
    for (hipEvent_t& e : r->ev) if (e) (void)hipEventDestroy(e);
}
size_t render_bytes(const fs_render* r, int what) {
    switch (what) {
        case FS_RENDER_VOL: return (size_t)r->vw * r->vh * 4 * sizeof(float);
        case FS_RENDER_AUX: return (size_t)r->vw * r->vh * 2 * sizeof(float);
        case FS_RENDER_LIGHTS_BUF: return (size_t)FS_NLIGHTS * 8 * sizeof(float);
        case FS_RENDER_SHARP: return r->h->ncell * 4 * sizeof(float);
        case FS_RENDER_LUT: return (size_t)2 * r->bbN * 4 * sizeof(float);
    }
    return 0;
}
float half_bits_to_float(uint16_t u) {   // binary16 -> binary32, exact
    const uint32_t s = (uint32_t)(u & 0x8000u) << 16, e = (u >> 10) & 0x1fu, m = u & 0x3ffu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = s;
        else {   // subnormal: m * 2^-24
            float f = (float)m * 5.9604644775390625e-8f;
            std::memcpy(&bits, &f, 4);
            bits |= s;
        }
    } else if (e == 31) bits = s | 0x7f800000u | (m << 13);
    else bits = s | ((e + 112u) << 23) | (m << 13);
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
}  // namespace

extern "C" {

int fs_blackbody_lut(int n, double tmin, double tmax, double twb, float* out) {   // src/blackbody.js:57-88
    if (!out) return fail(FS_EINVAL, "null output");
    if (n < 2 || n > (1 << 16)) return fail(FS_EINVAL, "n must be in 2..65536");
    if (!(tmin > 0.0) || !(tmax > tmin) || !(twb >= 0.0)) return fail(FS_EINVAL, "need 0 < tmin < tmax and twb >= 0");
    double lms2xyz[3][3];
    bb::inverse(bb::kBradford, lms2xyz);
    double src[3] = {1, 1, 1}, dst[3] = {1, 1, 1};
    if (twb > 0.0) {   // whiteBalance(Twb); Twb = 0 is the reference's `!Twb`: plain XYZ -> sRGB
        double w[3], wn[3];
        bb::xyz_of(twb, false, w);
        for (int k = 0; k < 3; k++) wn[k] = w[k] / w[1];
        bb::mul(bb::kBradford, wn, src);
        bb::mul(bb::kBradford, bb::kD65, dst);
    }
    double ref[3];
    bb::xyz_of(1500.0, false, ref);
    for (int row = 0; row < 2; row++) {
        for (int i = 0; i < n; i++) {
            const double T = tmin + (tmax - tmin) * i / (n - 1);
            double xyz[3], rgb[3];
            bb::xyz_of(T, row == 1, xyz);
            for (int k = 0; k < 3; k++) xyz[k] = xyz[k] / ref[1];
            if (twb > 0.0) {
                double lms[3], x2[3];
                bb::mul(bb::kBradford, xyz, lms);
                for (int k = 0; k < 3; k++) lms[k] = lms[k] * dst[k] / src[k];
                bb::mul(lms2xyz, lms, x2);
                bb::mul(bb::kXYZ2RGB, x2, rgb);
            } else {
                bb::mul(bb::kXYZ2RGB, xyz, rgb);
            }
            float* o = out + ((size_t)row * n + i) * 4;
            for (int k = 0; k < 3; k++) o[k] = (float)std::max(0.0, rgb[k]);
            o[3] = (float)xyz[1];
        }
    }
    return FS_OK;
}

int fs_render_param_count(void) { return FS_RP_COUNT; }
int fs_render_param_index(const char* name) {
    if (!name) return -1;
    for (int i = 0; i < FS_RP_COUNT; i++) if (std::strcmp(name, kRenderNames[i]) == 0) return i;
    return -1;
}
int fs_render_param_defaults(float* out) {
    if (!out) return fail(FS_EINVAL, "null output");
    render_defaults(out);
    return FS_OK;
}

int fs_render_set_white_balance(fs_render* r, double twb) {   // src/renderer.js:57-63
    if (!r) return fail(FS_EINVAL, "null handle");
    std::vector<float> lut((size_t)2 * r->bbN * 4);
    if (int rc = fs_blackbody_lut(r->bbN, r->bbTmin, r->bbTmax, twb, lut.data())) return rc;
    if (set_device(r->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(r->h->stream));
    FS_HIP_CHECK(hipMemcpy(r->lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice));
    return FS_OK;
}

int fs_render_create(fs_sim* h, int width, int height, double vol_scale, fs_render** out) {
    if (!h || !out) return fail(FS_EINVAL, "null argument");
    *out = nullptr;
    if (!(vol_scale > 0.0) || !std::isfinite(vol_scale)) return fail(FS_EINVAL, "vol_scale must be positive");
    if (h->ny / FS_NLIGHTS < 2 || h->nx < 2 || h->nz < 2)
        return fail(FS_EINVAL, "light extraction needs ny >= 16 (two rows per band) and nx, nz >= 2");
    // resize (src/renderer.js:575-581)
    const int w = std::max(16, width), hh = std::max(16, height);
    const int vw = std::max(8, (int)std::floor(w * vol_scale + 0.5)), vh = std::max(8, (int)std::floor(hh * vol_scale + 0.5));
    if ((long long)vw * vh > (1LL << 28)) return fail(FS_EINVAL, "volume image too large");
    if (set_device(h)) return FS_EDEVICE;
    fs_render* r = new fs_render();
    r->h = h;
    r->device = h->cfg.device;
    r->w = w; r->hgt = hh; r->vw = vw; r->vh = vh;
    r->bx = (h->nx + 3) / 4; r->by = (h->ny + 3) / 4; r->bz = (h->nz + 3) / 4;
    render_defaults(r->params);
    const float sky = 1e6f;   // depth the surface pass writes where no surface is hit
    bool ok = hipMalloc(&r->lut, (size_t)2 * r->bbN * sizeof(float4)) == hipSuccess &&
              hipMalloc(&r->lstate, 3 * FS_NLIGHTS * sizeof(float4)) == hipSuccess &&
              hipMalloc(&r->lights, 2 * FS_NLIGHTS * sizeof(float4)) == hipSuccess &&
              hipMalloc(&r->depth, sizeof(float)) == hipSuccess &&
              hipMalloc(&r->sharp, h->ncell * sizeof(half4_t)) == hipSuccess &&
              hipMalloc(&r->vol, (size_t)vw * vh * sizeof(half4_t)) == hipSuccess &&
              hipMalloc(&r->aux, (size_t)vw * vh * sizeof(float2)) == hipSuccess &&
              hipMalloc(&r->raw, (size_t)r->bx * r->by * r->bz) == hipSuccess &&
              hipMalloc(&r->flags, (size_t)r->bx * r->by * r->bz) == hipSuccess;
    if (!ok) {
        render_free(r);
        delete r;
        return fail(FS_ENOMEM, "hipMalloc failed for the renderer");
    }
    launch_fill_zero(h->stream, r->lstate, 3 * FS_NLIGHTS * sizeof(float4));
    launch_fill_zero(h->stream, r->lights, 2 * FS_NLIGHTS * sizeof(float4));
    const size_t nn3 = (size_t)h->cfg.noise_n * h->cfg.noise_n * h->cfg.noise_n;
    std::vector<uint16_t> curl(nn3 * 4);
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(r->depth, &sky, sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(curl.data(), h->curl, curl.size() * 2, hipMemcpyDeviceToHost) != hipSuccess) {
        render_free(r);
        delete r;
        return fail(FS_EDEVICE, "device error in renderer create");
    }
    for (size_t i = 0; i < nn3; i++)   // the curl warp moves a sample by at most detail * 0.12 * this
        for (int c = 0; c < 3; c++) r->curl_max = std::max(r->curl_max, std::fabs(half_bits_to_float(curl[i * 4 + c])));
    if (int rc = fs_render_set_white_balance(r, 5000.0)) {   // src/renderer.js:53
        render_free(r);
        delete r;
        return rc;
    }
    *out = r;
    return FS_OK;
}

int fs_render_destroy(fs_render* r) {
    if (!r) return FS_OK;
    (void)hipSetDevice(r->device);
    render_free(r);
    delete r;
    return FS_OK;
}

int fs_render_set(fs_render* r, const char* name, float value) {
    if (!r || !name) return fail(FS_EINVAL, "null argument");
    const int i = fs_render_param_index(name);
    if (i < 0) return fail(FS_EINVAL, std::string("unknown render parameter: ") + name);
    if (render_owned(i)) return fail(FS_EINVAL, std::string("render parameter is set per frame: ") + name);
    r->params[i] = value;
    return FS_OK;
}

int fs_render_get_uniforms(fs_render* r, float* out) {
    if (!r || !out) return fail(FS_EINVAL, "null argument");
    if (!r->have_U) return fail(FS_ESTATE, "no frame rendered yet");
    std::memcpy(out, r->U.f, sizeof(r->U.f));
    return FS_OK;
}

int fs_render_set_camera(fs_render* r, const double pos[3], const double fwd[3], const double right[3], const double up[3],
                         double fov_y) {
    if (!r || !pos || !fwd || !right || !up) return fail(FS_EINVAL, "null argument");
    if (!(fov_y > 0.0 && fov_y < 3.14159)) return fail(FS_EINVAL, "fov_y must be in (0, pi)");
    for (int k = 0; k < 3; k++) {
        r->cam_pos[k] = pos[k]; r->cam_fwd[k] = fwd[k]; r->cam_rgt[k] = right[k]; r->cam_up[k] = up[k];
    }
    r->fov_y = fov_y;
    r->have_cam = true;
    return FS_OK;
}

int fs_render_set_depth(fs_render* r, const float* depth, int dw, int dh) {
    if (!r) return fail(FS_EINVAL, "null handle");
    if (set_device(r->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(r->h->stream));
    if (!depth) {
        const float sky = 1e6f;
        FS_HIP_CHECK(hipMemcpy(r->depth, &sky, sizeof(float), hipMemcpyHostToDevice));
        r->dw = r->dh = 1;
        return FS_OK;
    }
    if (dw < 1 || dh < 1 || (long long)dw * dh > (1LL << 28)) return fail(FS_EINVAL, "bad depth size");
    const size_t n = (size_t)dw * dh;
    if (n > r->depth_cap) {
        float* nd = nullptr;
        if (hipMalloc(&nd, n * sizeof(float)) != hipSuccess) return fail(FS_ENOMEM, "hipMalloc failed for depth");
        (void)hipFree(r->depth);
        r->depth = nd;
        r->depth_cap = n;
    }
    FS_HIP_CHECK(hipMemcpy(r->depth, depth, n * sizeof(float), hipMemcpyHostToDevice));
    r->dw = dw;
    r->dh = dh;
    return FS_OK;
}

int fs_render_frame(fs_render* r, double time, double frame, float jit_x, float jit_y) {
    if (!r) return fail(FS_EINVAL, "null handle");
    if (!r->have_cam) return fail(FS_ESTATE, "set a camera first");
    fs_sim* h = r->h;
    if (set_device(h)) return FS_EDEVICE;
    // uploadParams (src/renderer.js:634-648): this.params plus the per-frame entries, each a
    // JS number rounded to f32
    const double lightDt = (!r->have_last_light || time < r->last_light_time || time - r->last_light_time > 0.5)
                               ? -1.0 : time - r->last_light_time;
    r->have_last_light = true;
    r->last_light_time = time;
    RenderU& U = r->U;
    std::memcpy(U.f, r->params, sizeof(U.f));
    U.f[RP_lightDt] = (float)lightDt;
    U.f[RP_width] = (float)r->w; U.f[RP_height] = (float)r->hgt; U.f[RP_vw] = (float)r->vw; U.f[RP_vh] = (float)r->vh;
    U.f[RP_time] = (float)time; U.f[RP_frame] = (float)frame;
    U.f[RP_nLogs] = (float)(h->logs_h.size() / 8); U.f[RP_nStones] = 0.f; U.f[RP_nLights] = (float)FS_NLIGHTS;
    U.f[RP_camX] = (float)r->cam_pos[0]; U.f[RP_camY] = (float)r->cam_pos[1]; U.f[RP_camZ] = (float)r->cam_pos[2];
    U.f[RP_fwdX] = (float)r->cam_fwd[0]; U.f[RP_fwdY] = (float)r->cam_fwd[1]; U.f[RP_fwdZ] = (float)r->cam_fwd[2];
    U.f[RP_rgtX] = (float)r->cam_rgt[0]; U.f[RP_rgtY] = (float)r->cam_rgt[1]; U.f[RP_rgtZ] = (float)r->cam_rgt[2];
    U.f[RP_upX] = (float)r->cam_up[0]; U.f[RP_upY] = (float)r->cam_up[1]; U.f[RP_upZ] = (float)r->cam_up[2];
    U.f[RP_tanFov] = (float)std::tan(r->fov_y / 2); U.f[RP_aspect] = (float)((double)r->w / r->hgt);
    U.f[RP_jitX] = jit_x; U.f[RP_jitY] = jit_y;
    U.f[RP_ox] = (float)h->origin[0]; U.f[RP_oy] = (float)h->origin[1]; U.f[RP_oz] = (float)h->origin[2];
    U.f[RP_dx] = (float)h->dx;
    U.f[RP_sx] = h->cfg.size[0]; U.f[RP_sy] = h->cfg.size[1]; U.f[RP_sz] = h->cfg.size[2];
    U.f[RP_nx] = (float)h->nx; U.f[RP_ny] = (float)h->ny; U.f[RP_nz] = (float)h->nz;
    U.f[RP_bbTmin] = (float)r->bbTmin; U.f[RP_bbTmax] = (float)r->bbTmax; U.f[RP_bbN] = (float)r->bbN;
    r->have_U = true;
    // renderHDR's passes on the step's fields (src/renderer.js:670-690)
    const FieldSet F = h->fields();
    if (r->profile) FS_HIP_CHECK(hipEventRecord(r->ev[0], h->stream));
    launch_lights(h->prec, h->stream, U, h->g, F.S, r->lut, r->lstate, r->lights);
    if (r->profile) FS_HIP_CHECK(hipEventRecord(r->ev[1], h->stream));
    const bool sharp = r->params[RP_flameSharp] > 0.f && r->params[RP_cubic] > 0.5f;
    r->sharp_ran = false;
    if (sharp) {   // once per simulation state
        const bool fresh = r->sharp_valid && r->sharp_gen == h->gen && r->sharp_frame == h->frame &&
                           r->sharp_time == h->time && r->sharp_k == r->params[RP_flameSharp];
        if (!fresh) {
            launch_sharpen(h->prec, h->stream, U.f[RP_flameSharp], h->g, F.S, r->sharp);
            r->sharp_valid = true;
            r->sharp_ran = true;
            r->sharp_gen = h->gen; r->sharp_frame = h->frame; r->sharp_time = h->time; r->sharp_k = r->params[RP_flameSharp];
        }
    }
    if (r->profile) FS_HIP_CHECK(hipEventRecord(r->ev[2], h->stream));
    const void* marched = sharp ? (const void*)r->sharp : F.S;
    const float detail = U.f[RP_detailAmt];
    // reach of a sample's texels from its unwarped cell: the cubic footprint (2), the warp and
    // one cell for the rounding of the map lookup
    const double warp = detail > 0.f ? (double)detail * 0.12 * r->curl_max : 0.0;
    const bool skip = r->skip && std::isfinite(warp) && warp < 64.0;
    if (skip) {
        const int reach = 3 + (int)std::ceil(warp);
        launch_occupancy(h->prec, sharp, h->stream, h->g, marched, F.S2, U.f[RP_smokeScat] > 0.f, r->raw, r->flags, r->bx,
                         r->by, r->bz, (reach + 3) / 4);
    }
    launch_volume(h->prec, sharp, h->stream, U, h->g, marched, F.S2, r->lut, r->lights, r->depth, r->dw, r->dh, h->curl,
                  h->cfg.noise_n, skip ? r->flags : nullptr, r->bx, r->by, r->bz, (float)(1.0 / h->dx), r->vol, r->aux);
    if (r->profile) FS_HIP_CHECK(hipEventRecord(r->ev[3], h->stream));
    r->timings_valid = r->profile;
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int fs_render_empty_skip(fs_render* r, int on) {
    if (!r) return fail(FS_EINVAL, "null handle");
    r->skip = on != 0;
    return FS_OK;
}

int fs_render_profile_enable(fs_render* r, int on) {
    if (!r) return fail(FS_EINVAL, "null handle");
    if (set_device(r->h)) return FS_EDEVICE;
    if (on && !r->ev[0])
        for (hipEvent_t& e : r->ev) FS_HIP_CHECK(hipEventCreate(&e));
    r->profile = on != 0;
    r->timings_valid = false;
    return FS_OK;
}

int fs_render_timings(fs_render* r, float ms[3]) {
    if (!r || !ms) return fail(FS_EINVAL, "null argument");
    if (!r->timings_valid) return fail(FS_ESTATE, "enable profiling and render a frame first");
    if (set_device(r->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipEventSynchronize(r->ev[3]));
    FS_HIP_CHECK(hipEventElapsedTime(&ms[0], r->ev[0], r->ev[1]));
    FS_HIP_CHECK(hipEventElapsedTime(&ms[1], r->ev[1], r->ev[2]));
    FS_HIP_CHECK(hipEventElapsedTime(&ms[2], r->ev[2], r->ev[3]));
    if (!r->sharp_ran) ms[1] = 0.f;
    return FS_OK;
}

int fs_render_size(fs_render* r, int* vw, int* vh) {
    if (!r) return fail(FS_EINVAL, "null handle");
    if (vw) *vw = r->vw;
    if (vh) *vh = r->vh;
    return FS_OK;
}

int fs_render_read(fs_render* r, int what, float* host, size_t bytes) {
    if (!r || !host) return fail(FS_EINVAL, "null argument");
    const size_t want = render_bytes(r, what);
    if (!want) return fail(FS_EINVAL, "unknown render buffer");
    if (bytes != want) return fail(FS_EINVAL, "size mismatch: expected " + std::to_string(want) + " bytes");
    if (what == FS_RENDER_SHARP && !r->sharp_valid) return fail(FS_ESTATE, "no prefiltered field yet");
    if (set_device(r->h)) return FS_EDEVICE;
    FS_HIP_CHECK(hipStreamSynchronize(r->h->stream));
    if (what == FS_RENDER_VOL || what == FS_RENDER_SHARP) {   // binary16 on the device
        const size_t nval = want / sizeof(float);
        std::vector<uint16_t> raw(nval);
        FS_HIP_CHECK(hipMemcpy(raw.data(), what == FS_RENDER_VOL ? (void*)r->vol : (void*)r->sharp, nval * 2, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nval; i++) host[i] = half_bits_to_float(raw[i]);
        return FS_OK;
    }
    const void* src = what == FS_RENDER_AUX ? (const void*)r->aux : what == FS_RENDER_LIGHTS_BUF ? (const void*)r->lights
                                                                                                  : (const void*)r->lut;
    FS_HIP_CHECK(hipMemcpy(host, src, want, hipMemcpyDeviceToHost));
    return FS_OK;
}

void* fs_render_device_ptr(fs_render* r, int what) {
    if (!r) return nullptr;
    switch (what) {
        case FS_RENDER_VOL: return r->vol;
        case FS_RENDER_AUX: return r->aux;
        case FS_RENDER_LIGHTS_BUF: return r->lights;
        case FS_RENDER_SHARP: return r->sharp;
        case FS_RENDER_LUT: return r->lut;
    }
    return nullptr;
}

}  // extern "C"
