"""CPU-side checks of the C-ABI library: it loads, exports every entry point
include/fluidstep.h declares, and its host-only settings registry follows the
reference's mutation semantics (src/settings/mod.rs tests).  No compute call
is made here."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import fluidstep as fsm
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "fluidstep.h")).read()
    return sorted(set(re.findall(r"FS_API\s+[\w\s\*]+?\b(fs_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(fsmod):
    syms = header_symbols()
    assert len(syms) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", fsm.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = fsmod.lib()
    for s in syms:
        assert hasattr(lib, s)
    assert set(fsm.EXPORTED) <= set(syms)
    # Nothing beyond the ABI leaks out of the library.
    assert {s for s in exported if not s.startswith("fs_")} <= {"_init", "_fini"} | set()


def test_version_and_last_error(fsmod):
    assert b"fluidstep" in fsmod.lib().fs_version()


def test_registry_defaults(fsmod):
    r = fsm.Registry()
    assert r.get("grid.res_x") == 80 and r.get("grid.res_y") == 40 and r.get("grid.res_z") == 80
    assert r.get("solver.pressure_warm_start") == 1.0
    assert r.get("solver.pressure_residual_tolerance") == 0.0
    assert r.get("solver.pressure_iterations") == 30
    assert r.get("dev.particle_sort_period") == 4
    assert abs(r.get("physics.flip_blend") - 0.9) < 1e-6
    assert abs(r.get("physics.gravity") + 9.81) < 1e-6


def test_finite_out_of_range_values_are_clamped(fsmod):  # src/settings/mod.rs:2174-2187
    r = fsm.Registry()
    assert r.set("physics.cfl", 1e100)["ok"]
    assert r.get("physics.cfl") == 6.0
    assert r.set("solver.pressure_iterations", 1e100)["ok"]
    assert r.get("solver.pressure_iterations") == 200
    assert r.set("solver.pressure_warm_start", 1e100)["ok"]
    assert r.get("solver.pressure_warm_start") == 1.0
    assert r.set("solver.pressure_residual_tolerance", 1e100)["ok"]
    assert r.get("solver.pressure_residual_tolerance") == pytest.approx(0.1, rel=1e-7)


def test_non_finite_values_are_rejected(fsmod):  # src/settings/mod.rs:2160-2172
    r = fsm.Registry()
    before = r.get("physics.flip_blend")
    for v in (math.nan, math.inf, -math.inf):
        res = r.set("physics.flip_blend", v)
        assert not res["ok"] and res["status"] == "non_finite_rejected" and res["stored_value"] is None
    assert r.get("physics.flip_blend") == before


def test_mutation_result_reports_clamp_and_live_apply_class(fsmod):  # :2270-2284
    res = fsm.Registry().set("physics.cfl", 1e100)
    assert res == {"ok": True, "status": "applied", "stored_value": 6.0, "clamped": True, "apply": "live",
                   "applied_live": False, "needs_reset": False}


def test_mutation_result_reports_stored_reset_class(fsmod):  # :2286-2299
    r = fsm.Registry()
    res = r.set("grid.res_x", 32.0)
    assert res["ok"] and res["status"] == "stored" and res["apply"] == "reset"
    assert res["stored_value"] == 32.0 and res["needs_reset"]
    assert r.get("grid.res_x") == 32


def test_unknown_id_is_rejected(fsmod):  # :2301-2309
    res = fsm.Registry().set("missing.setting", 1.0)
    assert not res["ok"] and res["status"] == "unknown_id" and res["apply"] is None


def test_u32_rounding_is_half_away_from_zero(fsmod):
    r = fsm.Registry()
    r.set("solver.pressure_iterations", 2.5)
    assert r.get("solver.pressure_iterations") == 3  # f64::round
    r.set("solver.pressure_iterations", -7.0)
    assert r.get("solver.pressure_iterations") == 1
    res = r.set("classify.liquid_threshold", 0.4)
    assert res["clamped"] and r.get("classify.liquid_threshold") == 1


def test_registry_to_params_matches_gpufluid_new(fsmod):
    """src/gpu/fluid.rs:258-318 with the all-64 tank of src/sim/mod.rs:587-602:
    extent 2, origin -1; Auto rest density = scene density."""
    r = fsm.Registry()
    for a in "xyz":
        r.set(f"grid.res_{a}", 64)
    H = np.float32(2.0) / np.float32(64.0)
    o = float(-np.float32(64) * H / np.float32(2))
    assert o == -1.0 and float(H * 64) == 2.0
    p = r.to_params(float(H), (o, o, o), 10.0)
    assert (p.nx, p.ny, p.nz) == (64, 64, 64)
    assert p.rest_density == 10.0 and p.surface_dilation == 0
    assert abs(p.dt - 1 / 120) < 1e-9 and p.rho == 1000.0
    p2 = r.to_params(float(H), (o, o, o), 4.0)
    assert p2.surface_dilation == 1  # effective_surface_dilation below 8/cell


def test_create_without_gpu_fails_loudly(fsmod):
    """No CPU fallback: on a host without a HIP device fs_create reports a
    device error instead of computing anything."""
    import ctypes
    n = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipGetDeviceCount(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    sc = scenes.small_tank(n=16)
    with pytest.raises(fsm.FluidStepError) as e:
        fsm.FluidStep(fsm.make_params(16, 16, 16, 1 / 16), sc["pos"], sc["vel"])
    assert e.value.code == -3


def test_invalid_arguments_are_rejected_before_the_device(fsmod):
    pos = np.zeros((4, 3), np.float32)
    with pytest.raises(fsm.FluidStepError) as e:
        fsm.FluidStep(fsm.make_params(2, 16, 16, 1 / 16), pos)
    assert e.value.code == -1
    bad = pos.copy()
    bad[1, 1] = np.nan
    with pytest.raises(fsm.FluidStepError) as e:
        fsm.FluidStep(fsm.make_params(16, 16, 16, 1 / 16), bad)
    assert e.value.code == -5
    with pytest.raises(fsm.FluidStepError) as e:
        fsm.FluidStep(fsm.make_params(16, 16, 16, 1 / 16, engine=0, precision=1), pos)
    assert e.value.code == -2


def test_batch_rejects_what_it_does_not_run_before_the_device(fsmod):
    """A batch runs the fused, mixed-precision, sort-every-substep step with
    one velocity storage; other requests fail with FS_ERR_UNSUPPORTED (and
    bad tanks with their single-tank codes) before any device call."""
    pos = np.full((4, 3), 0.5, np.float32)
    kw = dict(engine=1, precision=1, particle_sort=1, particle_sort_period=1)
    ok = fsm.make_params(16, 16, 16, 1 / 16, **kw)
    cases = [
        (fsm.make_params(16, 16, 16, 1 / 16, **dict(kw, engine=0, precision=0)), -2),
        (fsm.make_params(16, 16, 16, 1 / 16, **dict(kw, particle_sort=0)), -2),
        (fsm.make_params(16, 16, 16, 1 / 16, **dict(kw, particle_sort_period=4)), -2),
        (fsm.make_params(16, 16, 16, 1 / 16, **dict(kw, velocity_storage=1)), -2),
        (fsm.make_params(2, 16, 16, 1 / 16, **kw), -1),
    ]
    for second, code in cases:
        with pytest.raises(fsm.FluidStepError) as e:
            fsm.FluidBatch([ok, second], [pos, pos])
        assert e.value.code == code
    with pytest.raises(fsm.FluidStepError) as e:
        fsm.FluidBatch([], [])
    assert e.value.code == -1


def test_workload_particle_counts():
    assert len(scenes.dam_break_32()["pos"]) == 77824            # 16 x 19 x 32 x 8
    assert len(scenes.control_64()["pos"]) == 524288             # 32 x 32 x 64 x 8
    sc = scenes.dam_break_32()
    h = 1 / 32
    assert sc["pos"].min() >= np.float32(1.05 * h) and sc["pos"].max() <= np.float32(1 - 1.05 * h)
