#!/usr/bin/env python3
"""Throughput of the fire-simulation step on one MI355X.

Workload: 128 x 192 x 128 cells at dx = 5 mm (0.64 x 0.96 x 0.64 m), dt = 1 ms,
vorticity confinement eps 0.3, fed by 8 pyrolysis spots of U(1.5, 3) g/s drawn with seed 7
on a 0.3 m disc at 0.1 m height (scene "spots8_128"), noise 128^3 seed 1.  `--scene campfire`
feeds the same grid from the reference's seed-7 campfire logs instead.
Two forms of the same step run back to back from the same warmed state:
  * plain FP32:  float4 fields, 2 x V(2,2) FP32 V-cycles per step (the reference's solve);
  * mixed:       rgba16float fields (the reference's storage), FP32 math, FP32 V-cycles
                 inside FP64 defect correction to ||b - A p|| / ||b|| <= 1e-5.
`value` is the mixed step's voxel-steps/s; the FP32 step's rate and the ratio are
reported beside it.  Inputs are resident in HBM before the timed region.

Also reported: the roofline of the dominant kernel (algorithmic bytes / average
launch time from HIP events on the library's stream), the multigrid's achieved
bandwidth, and the CPU oracle (reference semantics: fp16 storage, f32 math,
2 V-cycles) timed on this host's cores on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "procedural-fire_amd"))

import numpy as np  # noqa: E402

import firesim as F  # noqa: E402
from firesim import scenes  # noqa: E402

RES = (128, 192, 128)
DX = 0.005
SIZE = tuple(n * DX for n in RES)
DT = 1e-3
PARAMS = {"vortEps": 0.3}
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per cell of each kernel: every input field read once, every
# output written once; E = bytes per 4-channel texel (8 fp16/bf16, 16 fp32); the
# advection passes also hand over one 4-byte word per cell (grids up to 1022 a side)
def kernel_bytes(E, p_bytes=4, pack=4):
    return {"adv1": 6 * E + pack, "adv2": 9 * E + pack, "curl": 2 * E, "forces": 8 * E, "div": E + 4,
            "project": 3 * E + p_bytes}


def vcycle_bytes(levels, pre, post, sor, zero_start0, direct=0, coarse_iters=0):
    """Algorithmic bytes of one V(pre, post) cycle per fine cell, kernel by kernel as they run (src/sim.js:588-599;
    float32 parity-split level arrays), per cell of each smoothed level:
      a colour pass reads the other colour's p and its own b, writes its own p: 6 B (8 B with over-relaxation,
        which also reads the own p); a level whose iterate starts at zero has its first colour-0 pass written by
        the kernel that zeroes it (restriction or FP64 residual), so one pre pass less;
      restriction: p and b at every fine cell read, the coarse b and p written: 9 B;
      prolongation fused with the first colour-0 post pass: p, colour-0 b and the coarse p read, every child
        written: 10.5 B;
    the coarsest level `coarse_iters` plain sweeps of 12 B, or with direct = l the exact solve of level l: three
    transform launches reading and writing one array each, 24 B per level-l cell."""
    n0 = float(np.prod(levels[0]))
    P = 8.0 if sor else 6.0
    last = direct if direct else len(levels) - 1
    tot = 0.0
    for l in range(last):
        w = np.prod(levels[l]) / n0
        pre_passes = 2 * pre - (1 if (l > 0 or zero_start0) else 0)
        tot += w * (pre_passes * P + 9.0 + 10.5 + (2 * post - 1) * P)
    if direct:
        tot += 24.0 * np.prod(levels[direct]) / n0
    else:
        tot += 12.0 * coarse_iters * np.prod(levels[last]) / n0
    return tot


# FP64 refinement per fine cell: the residual (x 8 and b 4 read, right-hand side and first colour-0 pass written)
# and the update in the last colour-1 pass (x read and written)
RESID64_BYTES = 20.0
UPDATE64_BYTES = 16.0


def refine_solve_bytes(levels, nv, first, direct=1, pre=1, post=2):
    """Algorithmic bytes per fine cell of the mixed step's pressure solve: FP64 refinement whose first correction
    takes `first` FP32 V(pre, post) over-relaxed cycles with the exact level-`direct` solve (the later ones one each),
    nv cycles in all, as the profiled (unspeculated) schedule runs it: the starting residual, the cycles (the 2nd
    and later of a correction start from a nonzero iterate: one colour-0 pass more), and per correction the update
    and the closing residual."""
    cyc = vcycle_bytes(levels, pre, post, True, True, direct=direct)
    ncorr = 1.0 + max(0.0, nv - first)
    return RESID64_BYTES + nv * cyc + 8.0 * (min(nv, first) - 1) + ncorr * (UPDATE64_BYTES + RESID64_BYTES)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by the
    counters' reading for a plain copy of known size), recorded in profiles/pmc_traffic.json."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None
    return {"bytes": k["fetch_bytes"] + k["write_bytes"], "source": "profiles/pmc_traffic.json"}


SCENE = "spots8_128"
SWEEPS = (0, 0)
# the step forms: storage precision and pressure solve
FORMS = {"fp32": (F.PREC_FP32, F.SOLVE_VCYCLES), "fp32_tol": (F.PREC_FP32, F.SOLVE_TOL32),
         "mixed_spectral": (F.PREC_FP16, F.SOLVE_SPECTRAL64), "mixed": (F.PREC_FP16, F.SOLVE_REFINE64)}        # red-black sweeps before / after the coarse correction in the tolerance solves (0: default)


# rocprofv3 exports its ROCPROF_* settings to the program it runs.  Under it every mode launches the
# step's kernels eagerly: a kernel-traced run with graph replay once ended in a host SIGSEGV inside the
# HIP runtime's graph launch path, and eager launches trace every dispatch with the same kernels.
UNDER_PROFILER = any(k.startswith("ROCPROF_") for k in os.environ)


def make_sim(precision, solve, use_graph=not UNDER_PROFILER):
    sw = SWEEPS if solve != F.SOLVE_VCYCLES else (0, 0)
    if SCENE == "campfire":
        logs, _ = F.build_campfire(7)
        s = F.FireSim(res=RES, size=SIZE, logs=logs, vcycles=2, precision=precision, solve=solve, div_tol=1e-5,
                      max_vcycles=8, noise_n=128, use_graph=use_graph, pre_sweeps=sw[0], post_sweeps=sw[1])
        s.set_params(PARAMS)
        return s
    return scenes.make_scene(SCENE, seed=7, precision=precision, solve=solve, div_tol=1e-5, max_vcycles=8,
                             noise_n=128, use_graph=use_graph, pre_sweeps=sw[0], post_sweeps=sw[1])


def time_steps(sim, steps):
    sim.sync()
    t0 = time.perf_counter()
    sim.step(DT, steps)
    sim.sync()
    return time.perf_counter() - t0


def profile_kernels(sim, steps=5):
    sim.profile_stages(True)
    acc = {k: 0.0 for k in F.STAGE_LABELS}
    for _ in range(steps):
        sim.step(DT, 1)
        t = sim.timings()
        for k in acc:
            acc[k] += t[k]
    sim.profile_stages(False)
    return {k: v / steps for k, v in acc.items()}


def cpu_baseline(gpu_sim, steps=5, steps_1core=3):
    """The oracle in reference semantics (fp16 storage, f32 math, 2 x V(2,2)) stepping the GPU's
    warmed state on the host: the median of `steps` steps on every core the process may use (after
    one warm-up step), then the median of `steps_1core` steps on one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O
    if SCENE == "campfire":
        logs, _ = F.build_campfire(7)
        orc = O.OracleSim(res=RES, size=SIZE, logs=logs, n_logs=16, precision=O.FP16, vcycles=2, noise_n=128)
        orc.set_params(**PARAMS)
    else:
        orc = O.OracleSim(res=RES, size=scenes.scene_size(SCENE), logs=np.zeros(8, np.float32), n_logs=0,
                          precision=O.FP16, vcycles=2, noise_n=128)
        orc.set_params(**scenes.scene_params(SCENE))
        orc.set_spots(scenes.scene_spots(SCENE, 7))
    for f in (F.VEL, F.S, F.S2, F.VORT, F.GEO, F.VELX, F.SX, F.S2X, F.VELY, F.SY, F.S2Y, F.NOISE, F.CURL):
        orc.write(f, gpu_sim.read(f, np.float64))
    orc.write(O.P, gpu_sim.read(F.P, np.float64))
    orc.set_clock(gpu_sim.time, gpu_sim.frame)

    def median_step(n):
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            orc.step(DT, 1)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    threads = O.num_threads()
    orc.step(DT, 1)   # warm-up
    med = median_step(steps)
    O.set_threads(1)
    med1 = median_step(steps_1core)
    O.set_threads(threads)
    ncell = float(np.prod(RES))
    return {"value": ncell / med, "unit": "voxel-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "value_1core": ncell / med1, "s_per_step": med, "s_per_step_1core": med1,
            "sample": f"median of {steps} steps (after 1 warm-up) on {threads} threads, then median of "
                      f"{steps_1core} steps on 1 thread, of the 128x192x128 step in reference semantics (fp16 "
                      f"storage, f32 math, 2 V(2,2)) from the warmed GPU state; {med:.2f} s/step on {threads} "
                      f"threads, {med1:.2f} s/step on 1 (nproc {os.cpu_count()})"}


def errors_vs_fp64(warmup):
    """Per-field relative L2 of the mixed and plain FP32 GPU steps against the FP64 oracle
    (float64 storage and math, pressure converged to 1e-10), after 1 and 10 steps from the
    GPU's warmed state; the FP64 oracle is the checker, run on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O

    def temp(a):
        return 293.0 + 1000.0 * np.asarray(a, np.float64)[..., 0]

    def rel(a, b):
        a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
        return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))

    base = make_sim(F.PREC_FP16, F.SOLVE_REFINE64, False)
    base.step(DT, warmup)
    base.sync()
    state = {f: base.read(f, np.float64) for f in (F.VEL, F.S, F.S2, F.VORT, F.GEO, F.VELX, F.SX, F.S2X, F.VELY,
                                                     F.SY, F.S2Y, F.P)}
    t0, fr0 = base.time, base.frame
    base.close()
    out = {}
    for name, prec, solve in (("mixed", F.PREC_FP16, F.SOLVE_REFINE64), ("fp32", F.PREC_FP32, F.SOLVE_VCYCLES)):
        g = make_sim(prec, solve, False)
        for f, a in state.items():
            g.write(f, a)
        g.step(DT, warmup)            # advances the clock to the warmed time; the fields are rewritten next
        g.sync()
        for f, a in state.items():
            g.write(f, a)
        orc = O.OracleSim(res=RES, size=scenes.scene_size(SCENE), logs=np.zeros(8, np.float32), n_logs=0,
                          precision=O.FP64, vcycles=2, solve=O.SOLVE_REFINE64, div_tol=1e-10, max_vcycles=40,
                          noise_n=128)
        orc.set_params(**scenes.scene_params(SCENE))
        orc.set_spots(scenes.scene_spots(SCENE, 7))
        for f, a in state.items():
            orc.write(f, a)
        orc.write(F.NOISE, g.read(F.NOISE, np.float64))
        orc.write(F.CURL, g.read(F.CURL, np.float64))
        orc.set_clock(t0, fr0)
        assert abs(g.time - t0) < 1e-12
        res = {}
        for done, n in ((1, 1), (10, 9)):
            for _ in range(n):
                g.step(DT, 1)
                orc.step(DT, 1)
            g.sync()
            e = {k: rel(g.read(f, np.float64)[..., :3], orc.read(f)[..., :3])
                 for k, f in (("vel", F.VEL), ("S", F.S), ("S2", F.S2))}
            e["T"] = rel(temp(g.read(F.S, np.float64)), temp(orc.read(F.S)))
            e["theta"] = rel(g.read(F.S, np.float64)[..., 0], orc.read(F.S)[..., 0])   # T - 293 K
            e["residual"] = g.residual()[0] if solve != F.SOLVE_VCYCLES else None
            res[str(done)] = e
            print(f"{name} after {done}: {e}", flush=True)
        out[name] = res
        g.close()
    print(json.dumps({"errors_vs_fp64": out, "workload": SCENE, "grid": list(RES), "warmup": warmup}))


def poisson_bench(args):
    """The isolated pressure solve: 256 x 384 x 256 cells, 64 seeded Gaussian blobs + white noise
    (zero mean), Dirichlet top/sides and Neumann floor, at most 6 levels, FP32 V(2,2) cycles in
    FP64 defect correction to ||b - A x|| / ||b|| <= 1e-8; error against the FP64 multigrid
    oracle converged to 1e-12, and the FP32 oracle V-cycle timed on the host as the CPU baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O
    dims = (256, 384, 256)
    h = 1.0 / 256
    n = float(np.prod(dims))
    b = scenes.blob_rhs(dims)
    zero = np.zeros_like(b)
    out = {"workload": "isolated Poisson 256x384x256, 64 blobs + N(0,1e-2) noise, 6 levels, tol 1e-8",
           "grid": list(dims)}
    for omega in (1.25, 1.0):
        ps = F.Poisson(dims, h, max_levels=6, omega=omega)
        ldims, coarse = ps.info()
        levels = len(ldims)
        ps.set_rhs(b)
        ts, hist = [], None
        for rep in range(1 + max(args.steps // 50, 3)):
            ps.set_x(zero)
            t0 = time.perf_counter()
            hist = ps.solve(1e-8, 40)
            t = time.perf_counter() - t0
            if rep:
                ts.append(t)
        cyc = len(hist) - 1
        ms = float(np.median(ts)) * 1e3
        alg = ((vcycle_bytes(ldims, 2, 2, omega != 1.0, True, coarse_iters=coarse) + UPDATE64_BYTES + RESID64_BYTES) * cyc
               + RESID64_BYTES) * n
        out[f"omega_{omega}"] = {"ms_per_solve": ms, "vcycles": cyc, "ms_per_vcycle": ms / cyc, "history": list(hist),
                                 "achieved_GBs": alg / (ms * 1e-3) / 1e9, "levels": levels, "coarse_iters": coarse}
        print(f"omega {omega}: {cyc} cycles, {ms:.3f} ms, final {hist[-1]:.2e}", flush=True)
        if omega == 1.25:
            x = ps.read_x()
        ps.close()
    out["roofline"] = {"bound": "hbm", "kernel": "V-cycle + refinement (omega 1.25)",
                       "achieved": out["omega_1.25"]["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": out["omega_1.25"]["achieved_GBs"] / HBM_PEAK_GBS}
    # FP64 multigrid oracle (the checker), converged
    t0 = time.perf_counter()
    ref = O.OracleMG(dims, h, max_levels=6, fp64=True, omega=1.25)
    ref.set(b=b, p=zero)
    rh = ref.vcycles(14)
    xr = ref.get()
    out["fp64_oracle"] = {"final_residual": float(rh[-1]), "cycles": 14, "s": time.perf_counter() - t0}
    out["error_vs_fp64"] = {"rel_l2": float(np.linalg.norm(x - xr) / np.linalg.norm(xr)),
                            "max_rel": float(np.abs(x - xr).max() / np.abs(xr).max())}
    del ref
    # CPU baseline: the FP32 oracle V(2,2) cycle (reference semantics, omega 1) on this host's cores
    cpu = O.OracleMG(dims, h, max_levels=6, fp64=False)
    cpu.set(b=b, p=zero)
    cpu.vcycles(1)
    t0 = time.perf_counter()
    cpu.vcycles(2)
    tc = (time.perf_counter() - t0) / 2
    out["cpu_baseline"] = {"value": n / tc, "unit": "cell-V-cycles/s", "cores": O.num_threads(), "kind": "port",
                           "sample": f"2 FP32 V(2,2) cycles of the same grid; {tc:.2f} s per cycle"}
    out["gpu_cell_vcycles_per_s"] = n / (out["omega_1.0"]["ms_per_vcycle"] * 1e-3)
    print(json.dumps(out))


def flame_statistics(args):
    """Long runs of the plain FP32 and the mixed step from rest with the per-step diagnostics
    record on: heat release, flame height (50 % axis intermittency) and puffing frequency
    (spectral peak of the heat release) of each, against f = 1.5 / sqrt(D)."""
    from firesim import estimators as E
    name = args.flame_scene
    if name == "campfire":
        # the reference app's "ultra" preset (src/main.js:19, 32-38): the seed-7 campfire in a
        # 1.2 x 1.8 x 1.2 m box at 128 x 192 x 128, 3 V-cycles, dt 1/120 s; its own calibration
        # targets (research/DESIGN.md:427-431): HRR 60 kW +-10 %, puffing 1.94 Hz +-20 %,
        # 50 % axis intermittency at 0.6 m +-0.1
        sc = {"res": (128, 192, 128), "dx": 1.2 / 128, "dt": 1.0 / 120, "vcycles": 3}
        D = 0.5
    else:
        sc = scenes.SCENES[name]
        D = scenes.DISC_DIAMETER
    steps = args.steps if args.steps != 200 else (2400 if name == "campfire" else 3500)
    skip = min(args.warmup if args.warmup != 100 else 500, steps // 2)
    out = {"workload": name, "grid": list(sc["res"]), "dt": sc["dt"], "steps": steps, "skipped": skip,
           "fire_diameter_m": D, "puffing_correlation_Hz": E.puffing_correlation(D)}
    # reference: the reference's own step (fp16 storage, fixed V-cycles); fp32: the same with
    # float4 storage; mixed: fp16 storage with the pressure refined to 1e-5 in FP64
    for form, prec, solve in (("reference", F.PREC_FP16, F.SOLVE_VCYCLES), ("fp32", F.PREC_FP32, F.SOLVE_VCYCLES),
                              ("mixed", F.PREC_FP16, F.SOLVE_REFINE64)):
        if name == "campfire":
            logs, _ = F.build_campfire(7)
            sim = F.FireSim(res=sc["res"], size=(1.2, 1.8, 1.2), logs=logs, vcycles=sc["vcycles"], precision=prec,
                            solve=solve, div_tol=1e-5, max_vcycles=8, noise_n=128, use_graph=not UNDER_PROFILER)
        else:
            sim = scenes.make_scene(name, seed=7, precision=prec, solve=solve, div_tol=1e-5, max_vcycles=8,
                                    noise_n=128, use_graph=not UNDER_PROFILER)
        sim.diag_enable(steps, t_flame=1000.0)
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(500, steps - done)
            sim.step(sc["dt"], n)
            sim.sync()
            done += n
            print(f"{form}: {done}/{steps} steps", flush=True)
        wall = time.perf_counter() - t0
        s = sim.diag_series()
        if args.save:
            os.makedirs(args.save, exist_ok=True)
            np.savez(os.path.join(args.save, f"flames_{name}_{form}.npz"), **s)
        out[form] = E.summary(s, float(np.float32(sc["dx"])), sc["dt"], 1000.0, skip)
        out[form]["wall_s"] = wall
        sim.close()
    print(json.dumps(out))


def adjoint_bench(args):
    """Gradients of a temperature/soot mismatch with respect to the pyrolysis-spot fluxes and the
    floor fuel flux: the adjoint scene (64 x 96 x 64 at 1 cm, dt 2 ms, 8 spots of U(1.5, 3) g/s
    plus a 64 x 64 floor-flux texture) run 200 steps from rest, checkpoints every 20 steps.  The
    target is the same run with every flux scaled by (1 + N(0, 0.2)).

    Checks reported: the tangent-linear (forward-mode) derivative along a spot + floor direction
    on the same trajectory (dot-product test of the adjoint); FP64 central differences of the
    FP64 oracle (fluxes perturbed in double) at 200 steps and, where differences are a stable
    reference, over the first 50 steps; FP32 central differences on the GPU (h = 1e-3); and
    the conditioning of the 200-step derivative (FP64 differences after a 1e-9 change of the
    initial temperature field)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O
    name = "adjoint8_64"
    sc = scenes.SCENES[name]
    steps = args.steps if args.steps != 200 else sc["steps"]
    every = 20
    dt = sc["dt"]
    tol = 1e-8
    sim = scenes.make_scene(name, precision=F.PREC_FP32, solve=F.SOLVE_REFINE64, div_tol=tol, max_vcycles=40,
                            noise_n=128)
    rows = sim.spots()[:, :6].astype(np.float32)
    face = scenes.scene_face_flux(name)
    base = sim.save_state()
    rng = np.random.default_rng(2024)
    pr = rows.copy()
    pr[:, 4] *= (1.0 + rng.normal(0.0, 0.2, len(rows))).astype(np.float32)
    pf = (face * (1.0 + rng.normal(0.0, 0.2, face.shape))).astype(np.float32)

    def make_target(n):
        sim.set_spots(pr)
        sim.set_face_flux(pf)
        sim.load_state(base)
        sim.step(dt, n)
        S = sim.read("S")
        sim.set_spots(rows)
        sim.set_face_flux(face)
        return np.ascontiguousarray(S[..., 0]), np.ascontiguousarray(S[..., 2])

    def loss_of(S, target):
        S = np.asarray(S, np.float64)
        return float(np.sum((S[..., 0] - target[0]) ** 2) + np.sum((S[..., 2] - target[1]) ** 2))

    target = make_target(steps)
    adj = F.Adjoint(sim, steps, every)
    tf, tb = [], []
    for rep in range(3):
        sim.load_state(base)
        sim.sync()
        t0 = time.perf_counter()
        adj.forward(dt, steps)
        t1 = time.perf_counter()
        loss = adj.seed_mismatch(*target)
        t2 = time.perf_counter()
        adj.backward(solve_tol=tol, max_cycles=40)
        t3 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t3 - t2)
        print(f"adjoint rep {rep}: forward {t1 - t0:.3f} s, backward {t3 - t2:.3f} s, loss {loss:.6e}", flush=True)
    g_spot, g_face = adj.grads(len(rows), face=True)
    out = {"workload": f"{name}: 64x96x64 dx=1cm dt=2ms, {steps} steps from rest, 8 spots + 64x64 floor flux, "
                       f"checkpoints every {every}, solve refined to {tol:g}",
           "loss": loss, "forward_ms_per_step": float(np.median(tf)) / steps * 1e3,
           "backward_ms_per_step": float(np.median(tb)) / steps * 1e3,
           "backward_over_forward": float(np.median(tb) / np.median(tf)),
           "grad_spot_flux": g_spot.tolist()}
    # tangent along a spot + floor direction on the recorded trajectory: <seed, J d> = <J^T seed, d>
    d_spot = rng.normal(0.0, 1.0, len(rows)) * 1e-3
    d_face = (np.where(face > 0, rng.normal(0.0, 1.0, face.shape), 0.0) * 1e-3).astype(np.float32)
    sim.load_state(base)
    sim.sync()
    t0 = time.perf_counter()
    tan = adj.tangent(dt, steps, spot_dir=d_spot, face_dir=d_face, solve_tol=tol, max_cycles=40)
    t_tan = time.perf_counter() - t0
    S_fin = sim.read("S").astype(np.float64)
    jd = float(np.sum(2.0 * (S_fin[..., 0] - target[0]) * tan["S"][..., 0]) +
               np.sum(2.0 * (S_fin[..., 2] - target[1]) * tan["S"][..., 2]))
    jtw = float(np.sum(g_spot * d_spot) + np.sum(g_face * d_face))
    out["tangent"] = {"ms_per_step": t_tan / steps * 1e3, "seed_dot_tangent": jd, "grad_dot_direction": jtw,
                      "rel_diff": abs(jd - jtw) / max(abs(jd), abs(jtw))}
    print(f"dot-product test: {jd:.9e} vs {jtw:.9e}", flush=True)
    if args.timing_only:
        print(json.dumps(out))
        return
    # FP32 central differences on the GPU, h = 1e-3 relative (float fluxes)
    fd32 = np.zeros(len(rows))
    for q in range(len(rows)):
        f0 = float(rows[q, 4])
        v = {}
        for sgn in (1, -1):
            r = rows.copy()
            r[q, 4] = np.float32(f0 * (1 + sgn * 1e-3))
            sim.set_spots(r)
            sim.load_state(base)
            sim.step(dt, steps)
            v[sgn] = (loss_of(sim.read("S"), target), float(r[q, 4]))
        fd32[q] = (v[1][0] - v[-1][0]) / (v[1][1] - v[-1][1])
    sim.set_spots(rows)
    out["fd32_spot_flux"] = fd32.tolist()
    out["rel_err_vs_fd32"] = float(np.linalg.norm(g_spot - fd32) / np.linalg.norm(fd32))
    print(f"spot gradient vs FD32: {out['rel_err_vs_fd32']:.3e}", flush=True)
    # FP64 oracle from the identical state, fluxes perturbed in double precision
    orc = O.OracleSim(res=sc["res"], size=scenes.scene_size(name), logs=np.zeros(8, np.float32), n_logs=0,
                      precision=O.FP64, vcycles=sc["vcycles"], max_levels=sc["max_levels"], solve=O.SOLVE_REFINE64,
                      div_tol=1e-11, max_vcycles=40, noise_n=128)
    orc.set_params(**scenes.scene_params(name))
    for f in (F.GEO, F.NOISE, F.CURL):
        orc.write(f, sim.read(f, np.float64))
    orc.set_spots(rows.astype(np.float64))
    orc.set_face_flux(face)
    vel0, S0, S20, p0, t00, fr0 = base

    def fd64(n, d, h, tgt, dS=None):
        v = {}
        for sgn in (1, -1):
            orc.set_flux64(spot=rows[:, 4].astype(np.float64) + sgn * h * d)
            orc.write(O.VEL, vel0.astype(np.float64))
            orc.write(O.S, S0.astype(np.float64) + (0.0 if dS is None else dS))
            orc.write(O.S2, S20.astype(np.float64))
            orc.write(O.P, p0)
            orc.set_clock(t00, fr0)
            t0 = time.perf_counter()
            for k in range(0, n, 50):
                orc.step(dt, min(50, n - k))
                print(f"  FP64 oracle: {min(k + 50, n)}/{n} steps ({time.perf_counter() - t0:.0f} s)", flush=True)
            v[sgn] = loss_of(orc.read(O.S), tgt)
        return (v[1] - v[-1]) / (2 * h)

    dg = g_spot / np.linalg.norm(g_spot) * np.linalg.norm(rows[:, 4])
    pred = float(np.sum(g_spot * dg))
    out["fp64_200"] = {"adjoint": pred}
    for h in (1e-6, 1e-7):
        out["fp64_200"][f"fd64_h{h:g}"] = fd64(steps, dg, h, target)
        print(f"FP64 {steps} steps along the gradient, h={h:g}: {out['fp64_200'][f'fd64_h{h:g}']:.6e} "
              f"(adjoint {pred:.6e})", flush=True)
    # conditioning: the same derivative from initial temperature fields changed by N(0, eps^2)
    # (theta units, 1 = 1000 K): eps 1e-7 is one float rounding of theta ~ 1
    for eps in (1e-9, 1e-7, 1e-5):
        nudge = np.zeros_like(S0, np.float64)
        nudge[..., 0] = eps * np.random.default_rng(7).normal(0.0, 1.0, S0.shape[:3])
        key = f"fd64_h1e-07_theta_nudged_{eps:g}"
        out["fp64_200"][key] = fd64(steps, dg, 1e-7, target, nudge)
        print(f"FP64 {steps} steps, initial theta nudged by {eps:g}: {out['fp64_200'][key]:.6e}", flush=True)
    # 50 steps: adjoint against converged FP64 differences
    n50 = 50
    t50 = make_target(n50)
    adj50 = F.Adjoint(sim, n50, every)
    sim.load_state(base)
    adj50.forward(dt, n50)
    adj50.seed_mismatch(*t50)
    adj50.backward(solve_tol=tol, max_cycles=40)
    g50 = adj50.grads(len(rows))[0]
    d50 = g50 / np.linalg.norm(g50) * np.linalg.norm(rows[:, 4])
    p50 = float(np.sum(g50 * d50))
    f50 = fd64(n50, d50, 1e-7, t50)
    out["fp64_50"] = {"adjoint": p50, "fd64": f50, "rel_err": abs(p50 - f50) / abs(f50)}
    print(f"FP64 50 steps: adjoint {p50:.6e} FD64 {f50:.6e} rel err {abs(p50 - f50) / abs(f50):.3e}", flush=True)
    print(json.dumps(out))


def sparks_bench(args):
    """Spark particle update (src/sparks.js:55-135) over the step's velocity: the 8-spot scene at
    128x192x128 (fp16 fields) after 100 steps from rest; N sparks spawned by one burst over the fire,
    then timed spark steps on the frozen velocity field.  Algorithmic bytes: one read and one write
    of the 64-B particle record per update (the velocity and curl samples hit the caches)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O
    sim = scenes.make_scene("spots8_128", precision=F.PREC_FP16, solve=F.SOLVE_REFINE64, div_tol=1e-5, max_vcycles=8,
                            noise_n=128, use_graph=not UNDER_PROFILER)
    sim.step(DT, 100)
    sim.sync()
    out = {"workload": "sparks over the 8-spot 128x192x128 fire (fp16 fields, 100 steps from rest), dt 1 ms"}
    steps = max(args.steps // 10, 20)
    for count in (8192, 1 << 22):
        sp = F.Sparks(sim, count)
        sp.set(rate=0.0)
        sp.burst(n=count, pos=(0.0, 0.25, 0.0), speed=1.0, radius=0.2, dur=1e-4, fan=0.3)
        sp.step(DT)
        for _ in range(3):
            sp.step(DT)
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            sp.step(DT)
        sim.sync()
        t = (time.perf_counter() - t0) / steps
        alive = int((sp.read()[:, 14] > 0.5).sum())
        gbs = count * 128 / t / 1e9
        out[f"n{count}"] = {"us_per_step": t * 1e6, "particle_updates_per_s": count / t, "alive": alive,
                            "achieved_GBs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
        print(f"{count} sparks: {t * 1e6:.1f} us/step, {count / t / 1e9:.2f} G updates/s, {gbs:.0f} GB/s, "
              f"{alive} alive", flush=True)
        sp.close()
    out["roofline"] = {"bound": "hbm", "kernel": "k_spark_update (4M sparks)", "achieved": out[f"n{1 << 22}"]["achieved_GBs"],
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": out[f"n{1 << 22}"]["frac_of_hbm_peak"]}
    # CPU baseline: the oracle's update (one thread) on 65536 sparks over the same velocity field
    res = RES
    orc = O.OracleSim(res=res, size=SIZE, logs=np.zeros(8, np.float32), n_logs=0, precision=O.FP16, noise_n=128)
    for f in (F.VEL, F.CURL):
        orc.write(f, sim.read(f, np.float64))
    so = O.OracleSparks(orc, 65536)
    so.set(rate=0.0)
    so.burst(n=65536, pos=(0.0, 0.25, 0.0), speed=1.0, radius=0.2, dur=1e-4, fan=0.3)
    so.step(DT)
    t0 = time.perf_counter()
    for _ in range(5):
        so.step(DT)
    tc = (time.perf_counter() - t0) / 5
    out["cpu_baseline"] = {"value": 65536 / tc, "unit": "particle-updates/s", "cores": 1, "kind": "port",
                           "sample": f"5 oracle spark steps of 65536 sparks; {tc * 1e3:.1f} ms per step"}
    print(json.dumps(out))


def capture_bench(args):
    """The timeline-driven capture loop (per-step parameter overrides, log shift, event bursts,
    sparks; firesim.driver) against plain stepping, on the reference app's default grid
    (80x120x80 in 1.2x1.8x1.2 m, fp16 fields, 2 V-cycles, campfire seed 7) at the capture's
    dt = 1/240 s: 1 s of pre-roll, then frames 0-59 (0-2 s of the cinematic, the hero pop at 0.3 s)."""
    from firesim import driver as D
    logs, _ = F.build_campfire(7)
    out = {"workload": "capture loop, 80x120x80 campfire (fp16, 2 V-cycles), dt 1/240 s, 8 steps per 30-fps frame"}
    for label, with_sparks in (("driver+sparks", True), ("driver", False)):
        sim = F.FireSim(res=(80, 120, 80), size=(1.2, 1.8, 1.2), logs=logs, vcycles=2, precision=F.PREC_FP16,
                        noise_n=128, use_graph=not UNDER_PROFILER)
        sparks = F.Sparks(sim, 8192) if with_sparks else None
        cap = D.Capture(sim, sparks, logs, sim_dt=1 / 240)
        cap.begin(preroll=1.0)
        sim.sync()
        t0 = time.perf_counter()
        for i in range(60):
            cap.frame(i)
        sim.sync()
        t = (time.perf_counter() - t0) / 60
        out[label] = {"ms_per_frame": t * 1e3, "ms_per_step": t * 1e3 / cap.n, "log_uploads": cap.log_uploads}
        if sparks is not None:
            out[label]["sparks_alive"] = int((sparks.read()[:, 14] > 0.5).sum())
            sparks.close()
        sim.close()
        print(f"{label}: {t * 1e3:.3f} ms per frame", flush=True)
    sim = F.FireSim(res=(80, 120, 80), size=(1.2, 1.8, 1.2), logs=logs, vcycles=2, precision=F.PREC_FP16,
                    noise_n=128, use_graph=not UNDER_PROFILER)
    sim.step(1 / 240, 240)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(1 / 240, 480)
    sim.sync()
    t = (time.perf_counter() - t0) / 60
    out["plain"] = {"ms_per_frame": t * 1e3, "ms_per_step": t * 1e3 / 8}
    out["driver_overhead_ms_per_step"] = out["driver"]["ms_per_step"] - out["plain"]["ms_per_step"]
    print(json.dumps(out))


def march_samples(U):
    """Upper bound of the march samples of one frame: (tf - tn) / ds per ray through the box with no
    surfaces, ignoring the early exit at Tr < 0.003 (src/renderer.js:245-259), in float64."""
    vw, vh = int(U[2]), int(U[3])
    px = np.arange(vw) + 0.5
    py = np.arange(vh) + 0.5
    ndx = (px / vw * 2 - 1)[None, :, None]
    ndy = (1 - py / vh * 2)[:, None, None]
    fwd, rgt, up = (np.array(U[i:i + 3], np.float64) for i in (12, 16, 20))
    d = fwd + rgt * ndx * U[11] * U[15] + up * ndy * U[11]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    ro = np.array(U[8:11], np.float64)
    bmin = np.array(U[24:27], np.float64)
    bmax = bmin + np.array(U[28:31], np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (bmin - ro) / d, (bmax - ro) / d
    tn = np.maximum(np.minimum(t0, t1).max(-1), 0)
    tf = np.maximum(t0, t1).min(-1)
    ds = U[27] * U[37]
    n = np.where(tf > tn, np.minimum((tf - tn) / ds + 0.5, 1200), 0)
    return float(n.sum()), int((tf > tn).sum())


def volume_bench(args):
    """Render-volume consumers of the step's fields (src/renderer.js:104-326): the fire-light extraction,
    flame-field prefilter and volume ray-march of one 1080 x 1920 frame at volume scale 1 (2.07 M rays,
    stepMul 0.5: the reference's quality preset for a 128-cell-wide grid, src/main.js:19; look of
    tools/render_final.mjs:23: cubic, flameSharp 1, detailAmt 0.3, sootEdge 0.12, lightK 0.6), on the
    mixed 128x192x128 8-spot fire after 1000 steps (1 s) from rest, when the plume fills the box.  One sim
    step between frames, so every frame runs all three passes (as the capture loop renders each sub-step).
    The march runs with its empty-space map (the default) and, for comparison, without it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import fsoracle as O
    sim = make_sim(F.PREC_FP16, F.SOLVE_REFINE64, not args.no_graph)
    sim.step(DT, 1000)
    sim.sync()
    W, H = 1080, 1920
    look = dict(stepMul=0.5, cubic=1, flameSharp=1, detailAmt=0.3, sootEdge=0.12, lightK=0.6)
    cam = F.orbit(yaw=0.35, pitch=0.12, dist=1.6, target=(0.0, 0.42, 0.0), fov=38.0)
    r = F.Render(sim, W, H, 1.0)
    r.set(**look)
    r.set_camera(cam)
    frames = max(10, args.steps // 10)
    for i in range(3):
        sim.step(DT, 1)
        r.frame(sim.time, i)
    sim.sync()
    t0 = time.perf_counter()
    for i in range(frames):
        sim.step(DT, 1)
    sim.sync()
    t_steps = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(frames):
        sim.step(DT, 1)
        r.frame(sim.time, 3 + i)
    sim.sync()
    t_frame = (time.perf_counter() - t0 - t_steps) / frames
    r.profile(True)
    acc = {"lights": 0.0, "sharpen": 0.0, "volume": 0.0}
    for i in range(frames):
        sim.step(DT, 1)
        r.frame(sim.time, 100 + i)
        for k, v in r.timings().items():
            acc[k] += v / frames
    r.empty_skip(False)
    ms_noskip = 0.0
    for i in range(frames):
        sim.step(DT, 1)
        r.frame(sim.time, 200 + i)
        ms_noskip += r.timings()["volume"] / frames
    r.empty_skip(True)
    r.profile(False)
    U = r.uniforms()
    vol = r.read_vol()
    samples, hit = march_samples(U)
    ncell = float(np.prod(RES))
    npix = W * H
    nn = 128
    # algorithmic bytes of one volume launch: the prefiltered S, S2 and the curl texture read once
    # (binary16 x 4), the image written once (binary16 x 4 + 2 x float)
    alg = ncell * 8 + ncell * 8 + nn ** 3 * 8 + npix * 16
    ms_v = acc["volume"]
    out = {"metric": "render-volume frame (lights + prefilter + ray-march) on the 128x192x128 fire",
           "workload": f"{W}x{H} frame, volume scale 1 ({npix} rays), stepMul 0.5, cubic + flameSharp 1, detailAmt 0.3, "
                       "sootEdge 0.12, lightK 0.6; mixed 8-spot fire after 1000 steps",
           "ms_per_frame": t_frame * 1e3, "frames_per_s": 1.0 / t_frame, "rays_per_s": npix / t_frame,
           "kernel_ms": acc, "volume_ms_without_empty_skip": ms_noskip, "rays_hitting_box": hit,
           "march_samples_upper": samples,
           "samples_per_s_upper": samples / (ms_v * 1e-3),
           "texel_bytes_requested_per_sample": 64 * 8 + 16 * 8 + 8 * 8,
           "requested_GBs_upper": samples * 704 / (ms_v * 1e-3) / 1e9,
           "mean_L": float(vol[..., :3].mean()), "min_Tr": float(vol[..., 3].min()),
           "roofline": {"bound": "hbm", "kernel": "k_volume", "achieved": alg / (ms_v * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": alg / (ms_v * 1e-3) / 1e9 / HBM_PEAK_GBS, "alg_bytes_per_launch": alg,
                        "avg_launch_ms": ms_v, "traffic": None},
           "sharpen_GBs": ncell * 16 / (acc["sharpen"] * 1e-3) / 1e9 if acc["sharpen"] > 0 else None}
    print(f"volume frame: {t_frame * 1e3:.3f} ms ({acc}); march without the empty-space map {ms_noskip:.3f} ms",
          flush=True)
    if args.no_cpu:
        r.close()
        sim.close()
        print(json.dumps(out))
        return
    # CPU baseline: the oracle's lights + prefilter + march of a 270 x 480 frame (1/16 of the rays, same
    # camera and look) from the GPU's fields, on this host's cores
    rs = F.Render(sim, W // 4, H // 4, 1.0)
    rs.set(**look)
    rs.set_camera(cam)
    rs.frame(sim.time, 0)
    Us = rs.uniforms()
    S, S2 = sim.read(F.S, np.float32), sim.read(F.S2, np.float32)
    curl = sim.read(F.CURL, np.float32)
    lut = rs.read_lut()
    t0 = time.perf_counter()
    lst = np.zeros((3, 8, 4), np.float32)
    lights = O.render_lights(Us, S, lut, lst)
    sharp = O.render_sharpen(1.0, S)
    O.render_volume(Us, sharp, S2, lut, lights, curl)
    tc = time.perf_counter() - t0
    nps = (W // 4) * (H // 4)
    out["cpu_baseline"] = {"value": nps / tc, "unit": "rays/s", "cores": O.num_threads(), "kind": "port",
                           "sample": f"oracle lights + prefilter + march of a {W // 4}x{H // 4} frame ({nps} rays) "
                                     f"from the GPU's fields; {tc:.2f} s"}
    out["vs_cpu"] = (npix / t_frame) / out["cpu_baseline"]["value"]
    rs.close()
    r.close()
    sim.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--scene", default="spots8_128", choices=["spots8_128", "campfire"])
    ap.add_argument("--sweeps", default="0,0",
                    help="pre,post red-black sweeps of the tolerance solves' V-cycles (0,0: the library default V(1,2))")
    ap.add_argument("--poisson", action="store_true",
                    help="instead of the step: the isolated 256x384x256 pressure solve (FP64 refinement to 1e-8)")
    ap.add_argument("--flames", action="store_true",
                    help="instead of timing: long FP32 and mixed runs with flame-height / puffing statistics")
    ap.add_argument("--save", default=None, help="with --flames: directory for the per-step series (.npz)")
    ap.add_argument("--flame-scene", default="spots8_128",
                    choices=["spots8_128", "segments12_192", "spots4_64", "campfire"])
    ap.add_argument("--errors", action="store_true",
                    help="instead of timing: per-field error of the mixed and FP32 steps vs the FP64 oracle "
                         "after 1 and 10 steps")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of the step: the adjoint workload (200 steps at 64x96x64, gradients of a "
                         "temperature/soot mismatch w.r.t. spot and floor fluxes, checked by FD32 and FD64)")
    ap.add_argument("--sparks", action="store_true",
                    help="instead of the step: spark-particle update throughput (8192 and 4M sparks) on the step's velocity")
    ap.add_argument("--capture", action="store_true",
                    help="instead of the step: the timeline-driven capture loop (firesim.driver) vs plain stepping")
    ap.add_argument("--volume", action="store_true",
                    help="instead of the step: the render-volume consumers (lights, prefilter, ray-march) of a 1080x1920 frame")
    ap.add_argument("--timing-only", action="store_true",
                    help="with --adjoint: forward, backward and tangent timings and the dot-product test only")
    ap.add_argument("--trace-form", default=None, choices=list(FORMS),
                    help="short run for kernel traces: --warmup then --steps steps of one step form, nothing else")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware-counter passes: a 256 MiB copy, then 3 steps of each step form")
    args = ap.parse_args()
    global SCENE, SWEEPS
    SCENE = args.scene
    SWEEPS = tuple(int(v) for v in args.sweeps.split(","))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the step runs on a single MI355X (no multi-GPU path is built); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.errors:
        errors_vs_fp64(args.warmup)
        return
    if args.poisson:
        poisson_bench(args)
        return
    if args.flames:
        flame_statistics(args)
        return
    if args.adjoint:
        adjoint_bench(args)
        return
    if args.sparks:
        sparks_bench(args)
        return
    if args.capture:
        capture_bench(args)
        return
    if args.volume:
        volume_bench(args)
        return
    ncell = float(np.prod(RES))
    graph = not args.no_graph and not UNDER_PROFILER
    if args.trace_form:
        prec, solve = FORMS[args.trace_form]
        sim = make_sim(prec, solve, graph)
        sim.step(DT, args.warmup)
        sim.sync()
        print(f"{args.trace_form}: {args.warmup} warm-up steps, residual {sim.residual()}", flush=True)
        sim.step(DT, args.steps)
        sim.sync()
        print(f"{args.trace_form}: {args.steps} steps, residual {sim.residual()}", flush=True)
        sim.close()
        return
    if args.probe:
        for k in F.STREAM_KINDS:
            print(f"{k} probe, 256 MiB: {F.bench_stream(k, 256 << 20, 3):.0f} GB/s", flush=True)
        for name, prec, solve in (("fp32", F.PREC_FP32, F.SOLVE_VCYCLES), ("mixed", F.PREC_FP16, F.SOLVE_REFINE64)):
            sim = make_sim(prec, solve, False)
            for _ in range(3):   # one step per call: no step runs ahead (held) of its predecessor's stopping test
                sim.step(DT, 1)
            sim.sync()
            print(f"{name}: 3 steps done", flush=True)
            sim.close()
        return
    probe = {k: F.bench_stream(k, 1 << 30, 10) for k in F.STREAM_KINDS}
    best_copy = max(probe["copy_plain"], probe["copy"])
    runs = {}
    # fp32: the reference's solve (2 x V(2,2)) on float4 fields; fp32_tol: plain FP32 V-cycles until the
    # residual reaches the same 1e-5 as the mixed step (or 8 cycles); mixed: the headline
    for name, (prec, solve) in FORMS.items():
        sim = make_sim(prec, solve, graph)
        sim.step(DT, args.warmup)
        t = time_steps(sim, args.steps)
        used, rels = [], []
        for _ in range(5):
            sim.step(DT, 1)
            r, u = sim.residual()
            used.append(u)
            rels.append(r)
        runs[name] = {"sim": sim, "t": t, "prof": profile_kernels(sim), "vcycles": float(np.mean(used)),
                      "rel": float(np.max(rels))}
        if name != "mixed":
            if name == "fp32":
                levels, coarse = sim.mg_info()
            sim.close()
    mix = runs["mixed"]["sim"]
    tmix, profm, nv = runs["mixed"]["t"], runs["mixed"]["prof"], runs["mixed"]["vcycles"]
    t32, prof32 = runs["fp32"]["t"], runs["fp32"]["prof"]
    v32 = ncell * args.steps / t32
    v32t = ncell * args.steps / runs["fp32_tol"]["t"]
    vspec = ncell * args.steps / runs["mixed_spectral"]["t"]
    vmix = ncell * args.steps / tmix
    # dominant single kernel of the mixed step (MG is many launches; reported separately)
    kb = kernel_bytes(8, 8)                                   # the mixed step projects with the FP64 pressure
    dom = max(kb, key=lambda k: profm[k])
    dom_ms = profm[dom]
    achieved = kb[dom] * ncell / (dom_ms * 1e-3) / 1e9
    # the mixed solve: V(1,2) cycles (or --sweeps) with level 1 solved exactly (the library's choice for this
    # grid), the first correction 3 cycles
    mg_cell = refine_solve_bytes(levels, nv, 3, pre=SWEEPS[0] if all(SWEEPS) else 1, post=SWEEPS[1] if all(SWEEPS) else 2)
    mg_alg = mg_cell * ncell
    mg_gbs = mg_alg / (profm["mg"] * 1e-3) / 1e9
    mg32_gbs = 2 * vcycle_bytes(levels, 2, 2, False, False, coarse_iters=coarse) * ncell / (prof32["mg"] * 1e-3) / 1e9
    out = {
        "metric": "sim voxel-steps/s @128x192x128 (mixed vs FP32) + MG HBM GB/s % of peak",
        "value": vmix,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmix / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 storage / f32 math / f64 refinement",
        "data": ("synthetic (8 seeded pyrolysis spots, 1.5-3 g/s, from rest; noise seed 1)" if SCENE != "campfire"
                 else "synthetic (seed-7 campfire emitter from rest, noise seed 1)"),
        "config": {"workload": "fire step 128x192x128 dx=5mm dt=1ms vortEps=0.3 "
                               + ("8 spots (seed 7)" if SCENE != "campfire" else "campfire logs (seed 7)"),
                   "grid": list(RES),
                   "solve_mixed": "FP64 defect correction to 1e-5: the first correction 3 FP32 V(1,2) cycles "
                                  "(red-black SOR, omega 1.25; level 1 solved exactly by DST-II/DCT-IV transforms on "
                                  "the matrix cores), later ones 1; the last correction's residual norms formed by the "
                                  "projection kernel; one graph per step",
                   "solve_fp32": "2 x V(2,2) FP32", "solve_fp32_tol": "FP32 V(1,2) (omega 1.25) until 1e-5 (cap 8)",
                   "graph": graph},
        "fp32_value": v32,
        "fp32_ms_per_step": t32 / args.steps * 1e3,
        "mixed_vs_fp32": vmix / v32,
        "fp32_tol_value": v32t,
        "mixed_vs_fp32_tol": vmix / v32t,
        "mixed_spectral_value": vspec,
        "mixed_spectral_ms_per_step": runs["mixed_spectral"]["t"] / args.steps * 1e3,
        "mixed_spectral_vs_fp32": vspec / v32,
        "mixed_spectral_solve": "exact separable inner solve (DST-II/DCT-IV transforms on the matrix cores: x and y "
                                "as FP32 rocBLAS GEMMs, z + eigenvalue division + inverse z as one fused MFMA "
                                "kernel) in FP64 defect correction, tol 1e-5 (FS_SOLVE_SPECTRAL64)",
        "vcycles_per_step": {k: v["vcycles"] for k, v in runs.items()},
        "rel_residual": {k: v["rel"] for k, v in runs.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_launch": kb[dom] * ncell, "avg_launch_ms": dom_ms},
        "mg": {"achieved_GBs": mg_gbs, "frac": mg_gbs / HBM_PEAK_GBS, "ms_per_step": profm["mg"],
               "alg_bytes_per_cell": mg_cell, "fp32_achieved_GBs": mg32_gbs},
        "kernel_ms_mixed": profm,
        "kernel_ms_fp32": prof32,
        "kernel_ms_fp32_tol": runs["fp32_tol"]["prof"],
        "kernel_ms_mixed_spectral": runs["mixed_spectral"]["prof"],
        "hbm_stream_GBs": best_copy,
        "hbm_probe_GBs": probe,
        "kernel_GBs_mixed": {k: {"GBs": kb[k] * ncell / (profm[k] * 1e-3) / 1e9,
                                 "frac_peak": kb[k] * ncell / (profm[k] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 "frac_probe": kb[k] * ncell / (profm[k] * 1e-3) / 1e9 / best_copy}
                             for k in kb if profm.get(k)},
    }
    out["roofline"]["frac_probe"] = achieved / best_copy
    out["mg"]["frac_probe"] = mg_gbs / best_copy
    traffic = measured_traffic(dom)
    if traffic:
        out["roofline"]["traffic"] = traffic["bytes"]
        out["roofline"]["traffic_source"] = traffic["source"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mix)
        out["vs_cpu"] = vmix / out["cpu_baseline"]["value"]
    mix.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
