#!/usr/bin/env python3
"""Benchmark of the FLIP/PIC step on one MI355X.

Workload (default): the 128^3 high-density tank - lower 40% filled at 8
particles per cell (6,684,672 particles, seed 1), velocity U[-0.5,0.5]^3,
dt 1/240, FLIP 0.95, CG to a 1e-5 relative residual (cap 200), warm start,
sort every substep.  Inputs are synthetic and resident in HBM before timing.

value     = particle-steps/s of the mixed-precision fused step (f32 vectors,
            f64 CG dots and true residual): N * K / seconds for K steps,
            device-synchronised on both sides.
plain     = the same for the reference-structured FP32 step (one kernel per
            WGSL pass, f32 reduction trees, host-driven CG loop).
roofline  = the persistent CG kernel: algorithmic bytes per launch over its
            event-timed average duration, against 8 TB/s.
cpu_baseline = the CPU restatement (oracle/, plain FP32, OpenMP on all host
            cores) timed on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("fluid-simulation_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0


def cg_bytes(cells, iters):
    """Compulsory CG traffic for one solve over `cells` cells and `iters`
    iterations: 45 B/cell/iteration (SpMV+d.q: read d 4 + flag 1, write q 4;
    update+r.r: read p,d,r,q 16, write p,r 8; direction: read r,d 8, write d
    4) plus 30 B/cell per solve (init 21 + f64 true residual 9)."""
    return 45.0 * cells * iters + 30.0 * cells


def load_traffic():
    path = os.path.join(ROOT, "profiles", "cg_persistent_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def copy_calibration(mib):
    """A plain elementwise copy of known size (torch.neg of `mib` MiB of f32
    into another buffer: that many bytes read and written; the source, larger
    than the L2 and Infinity Cache together, is not resident), run before the
    workload so that a counter run of this command reads the same counters
    for a kernel of known traffic and can scale the others by it."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "torch.neg (elementwise)", "dispatches": 3,
                                           "bytes_read": 4 * n, "bytes_written": 4 * n}}), file=sys.stderr, flush=True)
    del src, dst
    torch.cuda.empty_cache()


def run_gpu(fsm, sc, params, steps, warmup, timing_steps=3):
    pos, vel = sc["pos"], sc["vel"]
    t = fsm.FluidStep(params, pos, vel)
    t.set_timing(False)
    if warmup:
        t.step(warmup)
    t.synchronize()
    t0 = time.perf_counter()
    t.step(steps)
    t.synchronize()
    dt = time.perf_counter() - t0
    # Per-launch CG timing with device events (same stream as the kernels).
    t.set_timing(True)
    per = []
    for _ in range(timing_steps):
        t.step(1)
        per.append(t.stats())
    # One more substep with the reference's detailed sections (outside the
    # timed region: the section events add a wait per substep).
    t.set_timing(False)
    t.set_detailed_timing(True)
    t.step(1)
    per[-1]["detailed"] = t.detailed_timing()
    per[-1]["dispatches_per_substep"] = t.dispatches_per_substep()
    t.close()
    return dt, per


def section_model_bytes(N, C, F, c_act, iters):
    """Compulsory bytes of one substep per fine section, by the byte model of
    the step (140 N + 28 F + 35 C outside the CG, FP32 SoA particles, u8 cell
    flags): mark 16 N + 4 C (position read, key write; histogram); sort 12 C
    (scan and bin starts); classify 5 C; scatter (P2G and its face epilogue)
    24 N + 8 F; divergence (with CG init) 4 F + 9 C + 21 C; gradient 8 F + 5 C;
    g2p 48 N + 8 F plus the sort's 52 N particle permutation, which G2P
    applies as it writes; cg_solve 45 B per liquid-tile cell and iteration
    plus 9 B per cell for the f64 true residual."""
    return {"mark": 16 * N + 4 * C, "sort": 12 * C, "classify": 5 * C, "scatter": 24 * N + 8 * F,
            "divergence": 4 * F + 30 * C, "gradient_u": 8 * F + 5 * C, "g2p": 100 * N + 8 * F,
            "cg_solve": 45 * c_act * iters + 9 * c_act}


def section_gbs(fine, model):
    """Achieved GB/s of each fine section on its compulsory bytes."""
    return {k: round(model[k] / (fine[k] * 1e-3) / 1e9, 1) for k in model if fine.get(k, 0) > 0}


def fine_sections(d):
    """One substep's device time by the reference's FINE_SECTIONS names (the
    fused engine's absorbed sections omitted) and its CG as one solve."""
    out = {k: round(v, 4) for k, v in d["sections"].items() if k not in d["fused"]}
    out["cg_solve"] = round(d["cg_total_ms"], 4)
    return out


def cpu_baseline(sc, steps, threads=None):
    """The plain-FP32 CPU restatement over `steps` substeps from the
    workload's initial state: (seconds, cores used, ms per phase per substep).
    threads=None uses every OpenMP thread of the host."""
    import oracle

    import helpers  # tests/helpers.py

    P = helpers.oracle_params(sc, 0)
    full = oracle.num_threads()
    if threads:
        oracle.set_num_threads(threads)
    cores = oracle.num_threads()
    oracle.phase_reset()
    t0 = time.perf_counter()
    oracle.run(P, sc["pos"], sc["vel"], steps=steps, sort_period=1, want=False)
    dt = time.perf_counter() - t0
    phases = {k: round(v / steps, 3) for k, v in oracle.phase_ms().items()}
    if threads:
        oracle.set_num_threads(full)
    return dt, cores, phases


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


def batch_bench(fsm, scenes, gpu_params, args):
    """The 64-tank batch: aggregate particle-steps/s and per-tank step
    latency, against the same 64 tanks stepped one after another as single
    tanks (same fused engine), and against the plain FP32 single-tank step."""
    tanks = scenes.batched_tanks()
    counts = [len(sc["pos"]) for sc in tanks]
    N = int(sum(counts))
    params = [gpu_params(sc, fsm.ENGINE_FUSED) for sc in tanks]
    b = fsm.FluidBatch(params, [sc["pos"] for sc in tanks], [sc["vel"] for sc in tanks])
    b.step(max(args.warmup, 1))
    b.synchronize()
    tuning = None
    if not args.no_tune:
        tuning = b.autotune()
        print("batch tuning", json.dumps(tuning), file=sys.stderr, flush=True)
        b.step(1)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.synchronize()
    dt = time.perf_counter() - t0
    value = N * args.steps / dt
    # Per-step device time (every tank's step ends with the batch's step).
    b.set_timing(True)
    lat = []
    for _ in range(args.latency_steps):
        b.step(1)
        lat.append(b.stats(0)["ms_total"])
    iters = [b.stats(t)["cg_iterations"] for t in range(len(tanks))]
    b.close()

    # The same tanks one after another, each its own single-tank launch sequence.
    def sequential(engine, steps):
        hs = [fsm.FluidStep(gpu_params(sc, engine), sc["pos"], sc["vel"]) for sc in tanks]
        for h in hs:
            h.step(1)
            h.synchronize()
        per_tank = []
        t0 = time.perf_counter()
        for _ in range(steps):
            for h in hs:
                s0 = time.perf_counter()
                h.step(1)
                h.synchronize()
                per_tank.append((time.perf_counter() - s0) * 1e3)
        sdt = time.perf_counter() - t0
        for h in hs:
            h.close()
        return N * steps / sdt, per_tank

    seq_v, seq_lat = sequential(fsm.ENGINE_FUSED, args.seq_steps)
    line = {
        "metric": "batched tanks: aggregate FLIP particle-steps/s (64 tanks, one launch per phase, hipGraph)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "precision": "mixed: f32 vectors and particles, f64 CG dots/scalars/true residual",
        "data": "synthetic",
        "config": {"workload": "batched_tanks", "tanks": len(tanks), "particles": N,
                   "edges": [sc["n"] for sc in tanks], "dt": tanks[0]["physics"]["dt"], "cg_tolerance": 1e-5,
                   "cg_cap": 200, "sort_period": 1, "engine": "fused batch"},
        "tuning": tuning,
        "latency_ms": {"batch_step_p50": _pct(lat, 50), "batch_step_p99": _pct(lat, 99),
                       "note": "every tank of the batch completes its step when the batch step completes"},
        "cg_iterations": {"min": int(min(iters)), "max": int(max(iters)), "mean": float(np.mean(iters))},
        "sequential_fused": {"value": seq_v, "unit": "particle-steps/s",
                             "tank_step_ms_p50": _pct(seq_lat, 50), "tank_step_ms_p99": _pct(seq_lat, 99)},
        "speedup_vs_sequential_fused": value / seq_v,
    }
    if not args.no_plain:
        pv, plat = sequential(fsm.ENGINE_REFERENCE, 1)
        line["sequential_plain_fp32"] = {"value": pv, "unit": "particle-steps/s",
                                         "tank_step_ms_p50": _pct(plat, 50), "tank_step_ms_p99": _pct(plat, 99)}
        line["speedup_vs_sequential_plain_fp32"] = value / pv
    return line


def adjoint_report(fsm, scenes, gpu_params, horizons=(10, 20, 60), every=10, hs=(1e-3, 1e-5), fp64=True,
                   tolerance=None):
    """Differentiable control at 64^3: gradient of |centroid - (0.7,0.3,0.5)|^2
    after H substeps w.r.t. a uniform shift of the initial particle velocity
    (x, y, z) and gravity_y, for each horizon H.

    gpu:        the checkpointed forward + reverse pass on the GPU.
    fp64:       the FP64 reverse pass of the restatement along the GPU's own
                trajectory (its input state of every substep).
    frozen_fd:  central differences (each h) of the smooth forward with every
                decision of the GPU trajectory replayed (sort order, cell types,
                occupancies, face supports, CFL and wall-clamp branches; stencils
                from the current positions), real P2G sums and an exact pressure
                solve: the function the adjoint's conventions define.
    tlm:        forward-mode derivatives (dual numbers, oracle.tlm_from_states)
                at each input state of the GPU's own trajectory, carried
                through all H substeps, with no step size, so they exist at
                any horizon (central differences stop converging once the run
                is chaotic).  "replay": linearized where the reverse pass
                linearizes (P2G at 2^22, the substep's CG tolerance), the
                other product of the same Jacobian; "smooth": linearized at
                the smooth substep (real sums, exact solve), which shows how
                far the derivative moves with the linearization point.
    plain_fd:   central differences (h = 1e-3) of the GPU forward itself, whose
                discrete decisions move with the perturbation.
    Gravity is a float32 parameter: every difference in gravity divides by
    the step that the rounded values g0 +- h actually take."""
    import oracle

    import helpers

    sc = scenes.control_64()
    if tolerance is not None:  # a tighter pressure solve for the forward (cap raised to match)
        sc["physics"]["pressure_residual_tolerance"] = tolerance
        sc["physics"]["pressure_iterations"] = 2000
    target = np.array([0.7, 0.3, 0.5])
    params = gpu_params(sc, fsm.ENGINE_FUSED)
    P64 = helpers.oracle_params(sc, 1)
    v0 = sc["vel"].astype(np.float64)
    g0 = float(params.gravity[1])

    def loss(pos):
        c = pos.astype(np.float64).mean(0)
        return float(((c - target) ** 2).sum())

    def upstream(pos):
        c = pos.astype(np.float64).mean(0)
        return np.tile(2.0 * (c - target) / len(pos), (len(pos), 1))

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def grad4(gv, gg):
        return [float(gv[:, 0].sum()), float(gv[:, 1].sum()), float(gv[:, 2].sum()), float(gg[1])]

    t = fsm.FluidStep(params, sc["pos"], sc["vel"])
    t.forward_checkpointed(1, 1)  # warm-up of every kernel
    t.close()
    report = {"workload": "control_64", "particles": len(sc["pos"]), "checkpoint_every": every,
              "cg_tolerance": float(sc["physics"]["pressure_residual_tolerance"]),
              "cg_cap": int(sc["physics"]["pressure_iterations"]),
              "gradient": "d loss / d (uniform initial velocity shift x, y, z; gravity_y)", "horizons": []}
    for H in horizons:
        ent = {"steps": H}
        t = fsm.FluidStep(params, sc["pos"], sc["vel"])
        t0 = time.perf_counter()
        t.forward_checkpointed(H, every)
        t.synchronize()
        fwd_s = time.perf_counter() - t0
        pos, _ = t.particles()
        t0 = time.perf_counter()
        gx, gv, gg = t.adjoint(upstream(pos), None)
        bwd_s = time.perf_counter() - t0
        t.close()
        g_gpu = grad4(gv, gg)
        ent["gpu"] = {"loss": loss(pos), "grad": g_gpu, "forward_checkpointed_s": fwd_s, "adjoint_s": bwd_s}
        print(f"adjoint report: H={H} gpu grad {g_gpu}", file=sys.stderr, flush=True)

        def L_gpu(vel, gy):
            p = fsm.Params.from_buffer_copy(params)
            p.gravity[1] = gy
            tt = fsm.FluidStep(p, sc["pos"], vel.astype(np.float32))
            tt.step(H)
            pp, _ = tt.particles()
            tt.close()
            return loss(pp)

        def g_step(h):
            gp, gm = float(np.float32(g0 + h)), float(np.float32(g0 - h))
            return gp, gm, gp - gm

        hp = 1e-3
        fd = [(L_gpu(v0 + hp * np.eye(3)[a], g0) - L_gpu(v0 - hp * np.eye(3)[a], g0)) / (2 * hp) for a in range(3)]
        gp, gm, gs = g_step(hp)
        fd.append((L_gpu(v0, gp) - L_gpu(v0, gm)) / gs)
        ent["plain_fd"] = {"h": hp, "grad": fd, "rel_err_vs_gpu": rel(g_gpu, fd)}
        if fp64:
            # The GPU trajectory: every substep's input state (read-out order).
            tt = fsm.FluidStep(params, sc["pos"], sc["vel"])
            states = []
            for _ in range(H):
                p_, v_ = tt.particles()
                states.append((p_, v_, tt.pressure()))
                tt.step(1)
            pf, _ = tt.particles()
            tt.close()
            t0 = time.perf_counter()
            T = oracle.AdjointTape.from_states(P64, states)
            _, ogv, ogg = T.backward(upstream(pf), None)
            T.close()
            g64 = grad4(ogv, ogg)
            ent["fp64"] = {"grad": g64, "seconds": time.perf_counter() - t0, "rel_err_gpu_vs_fp64": rel(g_gpu, g64),
                           "rel_l2_per_particle_vel_grad": rel(gv, ogv)}
            t0 = time.perf_counter()
            n = len(pf)
            dpos0 = np.zeros((4, n, 3))
            dvel0 = np.zeros((4, n, 3))
            for a in range(3):
                dvel0[a, :, a] = 1.0
            dgrav = np.zeros((4, 3))
            dgrav[3, 1] = 1.0
            up = upstream(pf)
            ent["tlm"] = {}
            for name, replay in (("replay", True), ("smooth", False)):
                t0 = time.perf_counter()
                tp, _ = oracle.tlm_from_states(P64, states, dpos0, dvel0, dgrav, replay=replay)
                g_tlm = [float((up * tp[k]).sum()) for k in range(4)]
                del tp
                ent["tlm"][name] = {"grad": g_tlm, "seconds": time.perf_counter() - t0,
                                    "rel_err_gpu_vs_tlm": rel(g_gpu, g_tlm), "rel_err_fp64_vs_tlm": rel(g64, g_tlm)}
                print(f"adjoint report: H={H} forward mode ({name}) {g_tlm} rel err gpu "
                      f"{ent['tlm'][name]['rel_err_gpu_vs_tlm']:.3e}", file=sys.stderr, flush=True)
            del dpos0, dvel0
            rec = oracle.smooth_record_states(P64, states)
            del states
            x0 = sc["pos"].astype(np.float64)

            def L_frozen(vel, gy):
                Pg = oracle.make_params(P64.nx, P64.ny, P64.nz, P64.h, dt=P64.dt, flip_blend=P64.flip_blend,
                                        grav=(0.0, gy, 0.0), rest=P64.rest, stiff=P64.stiff,
                                        drift_clamp=P64.drift_clamp, cfl=P64.cfl,
                                        surface_dilation=P64.surface_dilation, reduce_mode=1)
                p_, _, _ = oracle.smooth_forward(Pg, x0, vel, rec, free_stencils=True)
                print(f"adjoint report: H={H} frozen forward done", file=sys.stderr, flush=True)
                return loss(p_) if np.all(np.isfinite(p_)) else float("nan")

            ent["frozen_fd"] = {}
            for h in hs:
                t0 = time.perf_counter()
                fd = [(L_frozen(v0 + h * np.eye(3)[a], g0) - L_frozen(v0 - h * np.eye(3)[a], g0)) / (2 * h)
                      for a in range(3)]
                gp, gm, gs = g_step(h)
                fd.append((L_frozen(v0, gp) - L_frozen(v0, gm)) / gs)
                ent["frozen_fd"][f"{h:g}"] = {"grad": fd, "rel_err_gpu_vs_fd": rel(g_gpu, fd),
                                              "rel_err_fp64_vs_fd": rel(g64, fd), "seconds": time.perf_counter() - t0}
                print(f"adjoint report: H={H} frozen fd h={h:g} {fd}", file=sys.stderr, flush=True)
            del rec
        report["horizons"].append(ent)
    return report


def _track(ids, perm):
    """Compose the sort permutation into the particle-identity array."""
    return ids[perm] if perm is not None else ids


def precision_report(steps, teacher_states):
    """Error of each GPU precision mode against the FP64 restatement on the
    128^3 workload: (a) free-running for `steps` steps (particles compared by
    identity through the sort permutations), (b) one step from the FP64
    oracle's own state, averaged over `teacher_states` states."""
    import fluidstep as fsm
    import oracle
    import scenes
    from helpers import gpu_params, oracle_params, rel_l2

    sc = scenes.high_density_128()
    N = len(sc["pos"])
    P64 = oracle_params(sc, 1)
    modes = {
        "mixed_f32_velocity": gpu_params(sc, fsm.ENGINE_FUSED),
        "mixed_bf16_velocity": gpu_params(sc, fsm.ENGINE_FUSED, velocity=fsm.VELOCITY_BF16),
        "plain_fp32": gpu_params(sc, fsm.ENGINE_REFERENCE),
    }
    # FP64 oracle trajectory, keeping a few states for the one-step errors.
    pos, vel = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    pr = np.zeros(sc["n"] ** 3)
    ids64 = np.arange(N)
    states = []
    keep = set(np.linspace(0, steps - 1, teacher_states).astype(int).tolist()) if teacher_states else set()
    t0 = time.perf_counter()
    for s in range(steps):
        if s in keep:
            states.append((pos.copy(), vel.copy(), pr.copy()))
        r = oracle.run(P64, pos, vel, pressure=pr, steps=1, sort_period=1, f64=True)
        ids64 = _track(ids64, r.perm)
        pos, vel, pr = r.pos, r.vel_particles, r.pressure
        print(f"fp64 oracle step {s + 1}/{steps} cg={r.stats.cg_iterations}", file=sys.stderr, flush=True)
    t_cpu = time.perf_counter() - t0
    order64 = np.argsort(ids64)
    ref_pos, ref_vel, ref_pr = pos[order64], vel[order64], pr
    report = {"workload": "high_density_128", "steps": steps, "particles": N,
              "fp64_oracle_seconds": t_cpu, "modes": {}}
    for name, prm in modes.items():
        t = fsm.FluidStep(prm, sc["pos"], sc["vel"])
        ids = np.arange(N)
        for s in range(steps):
            t.step(1)
            ids = _track(ids, t.permutation().astype(np.int64))
        gp, gv = t.particles()
        order = np.argsort(ids)
        ent = {
            "free_run_pos_rel_l2": rel_l2(gp[order], ref_pos),
            "free_run_vel_rel_l2": rel_l2(gv[order], ref_vel),
            "free_run_pressure_rel_l2": rel_l2(t.pressure(), ref_pr),
        }
        one = {"pos": [], "vel": [], "pressure": []}
        for (sp, sv, spr) in states:
            r = oracle.run(P64, sp, sv, pressure=spr, steps=1, sort_period=1, f64=True)
            t.load_state(sp.astype(np.float32), sv.astype(np.float32), spr.astype(np.float32))
            t.step(1)
            gp, gv = t.particles()
            # Both sides re-sorted from the same input order; undo each side's
            # permutation so particles are compared by identity.
            gperm, operm = t.permutation().astype(np.int64), r.perm.astype(np.int64)
            ug, uo = np.empty_like(gp), np.empty_like(r.pos)
            ug[gperm], uo[operm] = gp, r.pos
            uvg, uvo = np.empty_like(gv), np.empty_like(r.vel_particles)
            uvg[gperm], uvo[operm] = gv, r.vel_particles
            one["pos"].append(rel_l2(ug, uo))
            one["vel"].append(rel_l2(uvg, uvo))
            one["pressure"].append(rel_l2(t.pressure(), r.pressure))
        for k, v in one.items():
            if v:
                ent[f"one_step_{k}_rel_l2_max"] = float(np.max(v))
        t.close()
        report["modes"][name] = ent
        print(name, json.dumps(ent), file=sys.stderr, flush=True)
    return report


def scene_report(fsm, reps=5):
    """Initial particles of a scene (generate_particles, src/gpu/fluid.rs:
    1838-1913): fs_scene_seed on the GPU (device buffer, kernel, copy to the
    host) against the sequential CPU restatement, per case, with the bits
    compared.  The seeding kernel's own duration is in the rocprof trace of
    this command."""
    import oracle
    H = np.float32(2.0 / 64.0)
    cases = [
        ("default_falling_blob_80x40x80_d10", dict()),
        ("dam_break_80x40x80_fill50_d8", {"scene.preset": 1, "scene.fill_level": 50.0, "particles.density": 8.0}),
        ("falling_blob_128^3_fill50_d32", {"grid.res_x": 128, "grid.res_y": 128, "grid.res_z": 128,
                                           "scene.fill_level": 50.0, "particles.density": 32.0}),
    ]
    out = {"report": "scene_seed", "cases": []}
    for name, settings in cases:
        r = fsm.Registry()
        for k, v in settings.items():
            r.set(k, v)
        sc = fsm.scene_from_registry(r)
        origin = [-(np.float32(n) * H) / np.float32(2.0) for n in sc.res]
        pos = fsm.scene_seed(sc, H, origin)  # warm-up (module load, first allocation)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pos = fsm.scene_seed(sc, H, origin)
            times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        ref = oracle.scene_generate(sc.blocks_array(), sc.particle_count, H, origin, tuple(sc.res))
        cpu_s = time.perf_counter() - t0
        gpu_s = float(np.median(times))
        ent = {"case": name, "particles": int(len(pos)), "target": int(sc.particle_count),
               "gpu_ms_per_call": gpu_s * 1e3, "gpu_particles_per_s": len(pos) / gpu_s,
               "cpu_ms": cpu_s * 1e3, "cpu_particles_per_s": len(pos) / cpu_s, "cpu_cores": 1,
               "bit_exact": bool(pos.shape == ref.shape and np.array_equal(pos.view(np.uint32), ref.view(np.uint32)))}
        out["cases"].append(ent)
        print(name, json.dumps(ent), file=sys.stderr, flush=True)
    return out


def view_report(fsm, scenes, gpu_params, reps=20):
    """The live-buffer hand-off on the 128^3 tank: fs_get_device_view (no
    copy) and fs_pack_particles_vec4, the reference's 32-byte particle record
    packed from the live side (24 B read + 32 B written per particle).  The
    pack kernel's own duration is in the rocprof trace of this command."""
    import torch
    sc = scenes.high_density_128()
    t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
    t.step(1)
    t.synchronize()
    n = t.n
    dst = torch.empty(8 * n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        t.device_view()
    view_s = (time.perf_counter() - t0) / reps
    t.pack_particles_vec4(dst.data_ptr())
    t.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        t.pack_particles_vec4(dst.data_ptr())
    t.synchronize()
    pack_s = (time.perf_counter() - t0) / reps
    t.close()
    return {"report": "device_view", "workload": sc["name"], "particles": n,
            "device_view_us_per_call": view_s * 1e6,
            "pack_vec4_ms_per_call": pack_s * 1e3,
            "pack_vec4_algorithmic_GBps": 56.0 * n / pack_s / 1e9,
            "note": "pack timed as back-to-back calls on the tank's stream, host wall clock"}


def default_scene_report(fsm, steps=240, warmup=24, frames=120):
    """The reference's default configuration through the drop-in: registry
    defaults (falling blob in an 80x40x80 tank, particles.density 10, 30 CG
    iterations, sort every 4th substep), the tank built by
    fs_create_from_settings (GpuFluid::new) and stepped by fs_step, for the
    fused engine and for the reference-structured FP32 engine on the same
    seeded particles; then the frame driver (fs_frame_run) at a 60 Hz render
    delta, as FluidApp::frame runs it."""
    r = fsm.Registry()
    sc = fsm.scene_from_registry(r)
    out = {"report": "default_scene", "grid": list(sc.res), "preset": int(sc.preset),
           "particles": int(sc.generated_count), "pressure_iterations": int(r.get("solver.pressure_iterations")),
           "sort_period": int(r.get("dev.particle_sort_period")), "engines": {}}
    fused = fsm.FluidStep.from_settings(r, sc)
    pos0, _ = fused.particles()
    ref_params = fsm.Params.from_buffer_copy(fused.params)
    ref_params.engine, ref_params.precision = fsm.ENGINE_REFERENCE, fsm.PRECISION_F32
    tanks = {"fused": fused, "reference": fsm.FluidStep(ref_params, pos0)}
    for name, t in tanks.items():
        t.step(warmup)
        t.synchronize()
        t0 = time.perf_counter()
        t.step(steps)
        t.synchronize()
        dt = time.perf_counter() - t0
        st = t.stats()
        # Four substeps (one sort period) with the reference's detailed sections.
        t.set_detailed_timing(True)
        t.step(4)
        d = t.detailed_timing()
        t.set_detailed_timing(False)
        out["engines"][name] = {"particle_steps_per_s": t.n * steps / dt, "ms_per_substep": dt / steps * 1e3,
                                "cg_iterations_last": st["cg_iterations"], "liquid_cells": st["liquid_cells"],
                                "fine_sections_ms_4_substeps": fine_sections(d),
                                "dispatches_per_substep": t.dispatches_per_substep()}
        print(f"default scene: {name} done", file=sys.stderr, flush=True)
    tanks["reference"].close()
    f = fsm.Frame(fused, r)
    f.run(16.6667)
    fused.synchronize()
    subs = 0
    t0 = time.perf_counter()
    for _ in range(frames):
        subs += f.run(16.6667)["substeps"]
    fused.synchronize()
    dt = time.perf_counter() - t0
    out["frame_driver"] = {"frames": frames, "render_dt_ms": 16.6667, "substeps": subs,
                           "ms_per_frame": dt / frames * 1e3,
                           "stats_json": json.loads(fused.stats_json(f))}
    f.close()
    fused.close()
    return out


def sections_report(fsm, scenes, gpu_params, steps=2):
    """Per-section device time of both engines on the 128^3 tank under the
    reference's FINE_SECTIONS / CG_CATS names (fs_get_detailed_timing), each
    engine's coarse spans from a run without the detailed events beside it,
    its kernel dispatches per substep, and fs_stats_json's object."""
    sc = scenes.high_density_128()
    out = {"report": "sections", "workload": sc["name"], "steps": steps, "engines": {}}
    for name, engine in (("fused", fsm.ENGINE_FUSED), ("reference", fsm.ENGINE_REFERENCE)):
        t = fsm.FluidStep(gpu_params(sc, engine), sc["pos"], sc["vel"])
        t.set_timing(True)
        t.step(1)
        t.step(steps)
        coarse = t.stats()
        t.set_detailed_timing(True)
        t.step(steps)
        d = t.detailed_timing()
        out["engines"][name] = {
            "coarse_ms": {k: coarse[k] for k in ("ms_prep", "ms_pressure", "ms_finish", "ms_total", "ms_cg")},
            "detailed": d, "dispatches_per_substep": t.dispatches_per_substep(),
            "stats_json": json.loads(t.stats_json()),
        }
        print(f"sections: {name} done", file=sys.stderr, flush=True)
        t.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-steps-1t", type=int, default=3, help="substeps of the one-thread CPU row (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--bf16", action="store_true", help="bf16 particle-velocity storage")
    ap.add_argument("--precision-report", type=int, default=0, metavar="STEPS",
                    help="instead of timing, report each GPU precision mode's error against the FP64 "
                         "restatement after STEPS steps (and one-step errors from its states)")
    ap.add_argument("--teacher-states", type=int, default=3)
    ap.add_argument("--adjoint-report", type=int, default=0, metavar="STEPS",
                    help="instead of timing, run the differentiable-control check at 64^3 over STEPS substeps")
    ap.add_argument("--no-fp64", action="store_true", help="adjoint report: skip the FP64 reference gradient")
    ap.add_argument("--adjoint-tolerance", type=float, default=None,
                    help="adjoint report: pressure-solve tolerance of the forward (default: the workload's 1e-5)")
    ap.add_argument("--adjoint-horizons", default="10,20",
                    help="adjoint report: shorter horizons reported beside STEPS (comma-separated)")
    ap.add_argument("--view-report", action="store_true",
                    help="instead of timing the step, time the live-buffer hand-off and the vec4 particle pack")
    ap.add_argument("--default-scene-report", action="store_true",
                    help="instead of timing the step, step the reference's default scene built from the registry "
                         "defaults, with both engines and through the frame driver")
    ap.add_argument("--sections-report", action="store_true",
                    help="instead of timing the step, report both engines' per-section device time under the "
                         "reference's section names")
    ap.add_argument("--scene-report", action="store_true",
                    help="instead of timing the step, time scene seeding on the GPU against the CPU restatement")
    ap.add_argument("--workload", default="high_density_128",
                    choices=["high_density_128", "large_splash_256", "dam_break_32", "control_64", "batched_tanks"],
                    help="the metric is quoted on high_density_128; the others are for per-kernel profiles "
                         "and the batched-tank figures")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run an elementwise copy of MIB MiB before the workload (counter-run calibration)")
    ap.add_argument("--no-tune", action="store_true", help="batched_tanks: skip the autotuner")
    ap.add_argument("--latency-steps", type=int, default=20, help="batched_tanks: steps timed one by one")
    ap.add_argument("--seq-steps", type=int, default=3, help="batched_tanks: steps of the sequential baseline")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} (WORLD_SIZE={world}) is not supported; "
              "the step runs on a single MI355X and multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import fluidstep as fsm
    import scenes
    from helpers import gpu_params

    if args.precision_report:
        print(json.dumps(precision_report(args.precision_report, args.teacher_states)), flush=True)
        return

    if args.adjoint_report:
        hz = sorted({int(x) for x in args.adjoint_horizons.split(",") if 0 < int(x) <= args.adjoint_report}
                    | {args.adjoint_report})
        print(json.dumps(adjoint_report(fsm, scenes, gpu_params, horizons=hz, fp64=not args.no_fp64,
                                        tolerance=args.adjoint_tolerance)), flush=True)
        return

    if args.scene_report:
        print(json.dumps(scene_report(fsm)), flush=True)
        return

    if args.default_scene_report:
        print(json.dumps(default_scene_report(fsm)), flush=True)
        return

    if args.sections_report:
        print(json.dumps(sections_report(fsm, scenes, gpu_params)), flush=True)
        return

    if args.view_report:
        print(json.dumps(view_report(fsm, scenes, gpu_params)), flush=True)
        return

    if args.workload == "batched_tanks":
        print(json.dumps(batch_bench(fsm, scenes, gpu_params, args)), flush=True)
        return

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    sc = getattr(scenes, args.workload)()
    N = len(sc["pos"])
    vstore = fsm.VELOCITY_BF16 if args.bf16 else fsm.VELOCITY_F32
    fused = gpu_params(sc, fsm.ENGINE_FUSED, velocity=vstore)
    dt, per = run_gpu(fsm, sc, fused, args.steps, args.warmup)
    value = N * args.steps / dt
    ms_per_step = dt / args.steps * 1e3

    # Roofline of the persistent CG kernel (dominant kernel of the step).
    ach, ach_full, frac, cg_ms, iters, act = [], [], [], [], [], []
    C = sc["n"] ** 3
    for s in per:
        k = s["cg_iterations"]
        Ca = s["cg_active_cells"]
        sec = s["ms_cg"] / 1e3
        if sec > 0:
            ach.append(cg_bytes(Ca, k) / sec / 1e9)
            ach_full.append(cg_bytes(C, k) / sec / 1e9)
        cg_ms.append(s["ms_cg"])
        iters.append(k)
        act.append(Ca)
    achieved = float(np.mean(ach)) if ach else None
    traffic = load_traffic()
    if traffic and traffic.get("workload") != sc["name"]:
        traffic = None

    line = {
        "metric": ("FLIP particle-steps/s @128^3: mixed vs FP32 CG at 1e-5 residual; CG %HBM peak"
                   if sc["name"] == "high_density_128" else f"FLIP particle-steps/s ({sc['name']})"),
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "mixed: f32 vectors and particles%s, f64 CG dots/scalars/true residual"
                     % (" (bf16 particle velocity)" if args.bf16 else ""),
        "data": "synthetic",
        "config": {
            "workload": sc["name"], "grid": [sc["n"]] * 3, "particles": N, "dt": sc["physics"]["dt"],
            "flip": sc["physics"]["flip_blend"], "cg_tolerance": 1e-5, "cg_cap": 200, "sort_period": 1, "engine": "fused",
        },
        "cg": {"iterations_per_solve": iters, "active_cells": act, "ms_per_solve": cg_ms,
               "ms_per_iteration": [m / max(k, 1) for m, k in zip(cg_ms, iters)],
               "relative_residual_target": 1e-5,
               # |r_k| / |r_0| of the recurrence (the stop test) and |b - A p| / |r_0|
               # recomputed in f64 after the solve, per timed solve
               "relative_residual": [float(np.sqrt(s["rs_final"] / s["rs_initial"])) if s["rs_initial"] > 0
                                     else 0.0 for s in per],
               "relative_true_residual": [float(np.sqrt(s["true_residual_sq"] / s["rs_initial"]))
                                          if s["rs_initial"] > 0 else 0.0 for s in per]},
        "sections_ms": {k: per[-1][k] for k in ("ms_prep", "ms_pressure", "ms_finish", "ms_cg")},
        "fine_sections_ms": fine_sections(per[-1]["detailed"]),
        "fine_sections_gbs": section_gbs(fine_sections(per[-1]["detailed"]),
                                         section_model_bytes(N, C, 3 * (sc["n"] + 1) * sc["n"] ** 2,
                                                             per[-1]["cg_active_cells"], per[-1]["cg_iterations"])),
        "dispatches_per_substep": per[-1]["dispatches_per_substep"],
        "roofline": {
            "kernel": "k_cg_persistent",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": (achieved / PEAK_HBM_GBS) if achieved else None,
            "basis": "45 B/cell/iteration + 30 B/cell/solve over the cells of liquid tiles",
            "achieved_full_grid_basis": float(np.mean(ach_full)) if ach_full else None,
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "traffic_source": ("stored figure, not measured in this run: " + traffic.get("source", "")
                               + " (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, each scaled by the same "
                               "counter's reading for a copy of known size)") if traffic else None,
        },
    }

    if not args.no_plain:
        plain = gpu_params(sc, fsm.ENGINE_REFERENCE)
        pdt, pper = run_gpu(fsm, sc, plain, args.plain_steps, 1, timing_steps=1)
        pv = N * args.plain_steps / pdt
        line["plain_fp32"] = {"value": pv, "unit": "particle-steps/s", "ms_per_step": pdt / args.plain_steps * 1e3,
                              "cg_iterations": pper[-1]["cg_iterations"], "ms_cg": pper[-1]["ms_cg"]}
        line["speedup_vs_plain_fp32"] = value / pv

    if not args.no_cpu:
        cdt, cores, phases = cpu_baseline(sc, args.cpu_steps)
        line["cpu_baseline"] = {
            "value": N * args.cpu_steps / cdt, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{args.cpu_steps} substeps of the {sc['name']} workload from its initial state, plain FP32 "
                      f"restatement (oracle/), OpenMP {cores} threads, {cdt:.1f} s",
            "ms_per_substep_by_phase": phases,
        }
        if args.cpu_steps_1t > 0:
            c1, _, ph1 = cpu_baseline(sc, args.cpu_steps_1t, threads=1)
            line["cpu_baseline_1_thread"] = {
                "value": N * args.cpu_steps_1t / c1, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                "sample": f"{args.cpu_steps_1t} substeps of the {sc['name']} workload from its initial state, "
                          f"plain FP32 restatement, one thread, {c1:.1f} s",
                "ms_per_substep_by_phase": ph1,
            }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
