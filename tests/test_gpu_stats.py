"""The timing and stats contract on the device: detailed timing with the
reference's section and category names (GpuTimers' detailed mode,
src/gpu/timing.rs:32-116, src/gpu/mod.rs:1039-1212) tiles the substep as the
coarse spans do, marks the fused engine's absorbed sections, leaves the
results bit-identical, and fs_stats_json equals the restatement's text
(oracle/stats_oracle.py, src/profiler/mod.rs:416-612) built from the same
read-outs.  Kernel dispatch counts follow the reference-structured engine's
launch sequence."""
import json
import math

import numpy as np
import pytest

import fluidstep as fsm
import scenes
import stats_oracle as so
from helpers import bits_equal, gpu_params

pytestmark = pytest.mark.gpu

ITERS = 30


def tank(engine, sort=1, warm=1, seed=4):
    sc = scenes.small_tank(n=20, seed=seed)
    p = gpu_params(sc, engine, pressure_iterations=ITERS, particle_sort=sort, pressure_warm_start=warm)
    return fsm.FluidStep(p, sc["pos"], sc["vel"])


def close(a, b, rel=0.05, abs_ms=0.1):
    return abs(a - b) <= abs_ms + rel * max(abs(a), abs(b))


def readout_of(d):
    names, cats = fsm.fine_sections(), fsm.cg_categories()
    return dict(detailed=d["detailed"], substeps=d["substeps"], prep_ms=d["prep_ms"],
                pressure_ms=d["pressure_ms"], finish_ms=d["finish_ms"],
                sections=[d["sections"][n] for n in names], cg_cats=[d["cg"][c] for c in cats],
                cg_total_ms=d["cg_total_ms"], cg_iters=d["cg_iters"], fused=d["fused"])


def expected_json(t, frame, call_substeps, engine_name):
    st = t.stats()
    if t._detailed:
        gpu = so.gpu_object(readout_of(t.detailed_timing()), fsm.fine_sections(), fsm.cg_categories(),
                            st["liquid_cells"])
    elif t._coarse:
        gpu = so.gpu_object(dict(detailed=False, substeps=call_substeps, prep_ms=st["ms_prep"],
                                 pressure_ms=st["ms_pressure"], finish_ms=st["ms_finish"]),
                            fsm.fine_sections(), fsm.cg_categories(), st["liquid_cells"])
    else:
        gpu = None
    q = t.current_params()
    ts, dropped = frame.timestep_stats() if frame is not None else (None, 0.0)
    f = dict(substeps=call_substeps, nx=q.nx, ny=q.ny, nz=q.nz, h=float(np.float32(q.h)),
             liquid_cells=st["liquid_cells"], particles=t.n, pressure_iterations=q.pressure_iterations,
             buffer_bytes=t.buffer_bytes(), dispatches_per_substep=t.dispatches_per_substep(), engine=engine_name)
    return so.stats_json(f, ts, dropped, gpu)


def modes(t, coarse, detailed):
    t.set_timing(coarse)
    t.set_detailed_timing(detailed)
    t._coarse, t._detailed = coarse, detailed


def test_reference_engine_sections_tile_the_coarse_spans():
    t = tank(fsm.ENGINE_REFERENCE)
    modes(t, True, True)
    t.step(3)
    d, s = t.detailed_timing(), t.stats()
    names = fsm.fine_sections()
    secs = [d["sections"][n] for n in names]
    assert d["detailed"] and d["substeps"] == 3 and d["cg_iters"] == ITERS and d["cg_split"] and not d["fused"]
    assert min(secs) >= 0.0 and all(d["sections"][n] > 0.0 for n in names if n != "sort")
    cats = [d["cg"][c] for c in fsm.cg_categories()]
    assert min(cats) > 0.0 and math.isclose(sum(cats), d["cg_total_ms"], rel_tol=1e-9)
    prep, pres, fin = so.rollup(secs, names, d["cg_total_ms"])
    assert (prep, pres, fin) == (d["prep_ms"], d["pressure_ms"], d["finish_ms"])
    assert close(prep, s["ms_prep"]) and close(pres, s["ms_pressure"]) and close(fin, s["ms_finish"])
    assert close(d["cg_total_ms"], s["ms_cg"])
    t.close()


def test_fused_engine_marks_absorbed_sections():
    t = tank(fsm.ENGINE_FUSED)
    modes(t, True, True)
    t.step(2)
    d, s = t.detailed_timing(), t.stats()
    names = fsm.fine_sections()
    fused = [n for i, n in enumerate(names) if (6 <= i <= 16) or i == 18 or (20 <= i <= 24)]
    assert d["fused"] == fused and not d["cg_split"]
    assert all(d["sections"][n] == 0.0 for n in fused)
    assert all(d["sections"][n] > 0.0 for n in ("clear", "mark", "sort", "classify", "scatter", "normalize_u",
                                               "divergence", "gradient_u", "g2p"))
    assert all(v is None for v in d["cg"].values()) and d["cg_total_ms"] > 0.0
    assert close(d["cg_total_ms"], s["ms_cg"]) and close(d["prep_ms"], s["ms_prep"])
    assert close(d["finish_ms"], s["ms_finish"])
    t.close()


@pytest.mark.parametrize("engine", [fsm.ENGINE_REFERENCE, fsm.ENGINE_FUSED])
def test_detailed_timing_leaves_results_bit_identical(engine):
    a, b = tank(engine), tank(engine)
    modes(a, False, True)
    a.step(3)
    b.step(3)
    pa, va = a.particles()
    pb, vb = b.particles()
    assert bits_equal(pa, pb) and bits_equal(va, vb)
    a.close(), b.close()


@pytest.mark.parametrize("warm", [0, 1])
def test_reference_engine_dispatch_count(warm):
    # Per substep without a sort: 8 clears (+1 pressure clear when cold),
    # mark, classify, scatter, 12 face passes, divergence, 4 CG-init
    # dispatches, 10 per CG iteration, 6 finish passes and G2P.
    t = tank(fsm.ENGINE_REFERENCE, sort=0, warm=warm)
    t.step(1)
    assert t.dispatches_per_substep() == 35 + (1 - warm) + 10 * ITERS
    t.close()
    t = tank(fsm.ENGINE_REFERENCE, sort=1, warm=warm)
    t.step(1)
    assert t.dispatches_per_substep() > 35 + (1 - warm) + 10 * ITERS + 4  # key_hist, scans, radix, permute
    t.close()


def test_fused_engine_dispatches_fewer_kernels():
    f, r = tank(fsm.ENGINE_FUSED), tank(fsm.ENGINE_REFERENCE)
    f.step(1)
    r.step(1)
    assert 0 < f.dispatches_per_substep() < r.dispatches_per_substep() // 4
    f.close(), r.close()


@pytest.mark.parametrize("engine,coarse,detailed", [
    (fsm.ENGINE_REFERENCE, False, True), (fsm.ENGINE_FUSED, True, True), (fsm.ENGINE_FUSED, True, False),
    (fsm.ENGINE_REFERENCE, False, False),
])
def test_stats_json_equals_the_restatement(engine, coarse, detailed):
    t = tank(engine)
    modes(t, coarse, detailed)
    r = fsm.Registry()
    assert r.set("physics.max_substeps", 3)["ok"]
    fr = fsm.Frame(t, r)
    st = fr.run(33.0)
    assert st["substeps"] >= 1
    name = "fused" if engine == fsm.ENGINE_FUSED else "reference"
    got = t.stats_json(fr)
    assert got == expected_json(t, fr, st["substeps"], name)
    obj = json.loads(got)
    assert obj["substeps_this_frame"] == st["substeps"] and obj["particles"] == t.n
    assert (obj["gpu"] is None) == (not (coarse or detailed))
    if detailed:
        assert list(obj["gpu"]["sections"]) == fsm.fine_sections() and obj["gpu"]["cg"]["iters"] == ITERS
    # Without a frame: the profiler's default timestep fields.
    got = t.stats_json()
    assert got == expected_json(t, None, st["substeps"], name)
    fr.close(), t.close()
