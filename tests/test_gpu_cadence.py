"""Parity at the reference's sort cadence: substeps that re-sort the particles
and substeps that leave them in place (advance_sort_tick, src/gpu/fluid.rs:
1491-1498; default period 4, src/settings/mod.rs:1379-1391), for both
engines, against the CPU restatement run with the same sort_period and tick.

On a substep without a sort the reference keeps the particle order
(sh/scatter.wgsl / sh/g2p.wgsl run over the live buffer as it lies); the
fused engine then gathers P2G through a cell order of the live side and the
reference-structured engine scatters through the workgroup-local table of
sh/scatter_local.wgsl (per-tap atomics with the sort off).  Both must give the
restatement's integer sums bit for bit (`saved`), its particle order, and
faces, pressure and particles within 1e-5 relative L2 per substep.

Teacher-forced: every substep starts both sides from the restatement's
state; the GPU tank's sort tick advances on its own, exactly as the
restatement's tick0 + s.  Free run: both sides advance on their own and the
particles are matched by identity through each sort's permutation."""
import sys

import numpy as np
import pytest

import fluidstep as fsm
import oracle
import scenes
from helpers import bits_equal, gpu_params, oracle_params, rel_l2

pytestmark = pytest.mark.gpu

ENGINES = [fsm.ENGINE_REFERENCE, fsm.ENGINE_FUSED]
CADENCES = [(1, 4), (1, 3), (0, 1)]  # (dev.particle_sort, dev.particle_sort_period)
TOL = 1e-5
H = np.float32(2.0 / 64.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def brick_edge_misfit_24():
    """y sorting into cell 6 (y/h) but scattering from cell 7 (y*inv_h) at n=24."""
    h = np.float32(1.0) / np.float32(24)
    y = np.float32(0.29166666)
    assert np.floor(y / h) == 6 and np.floor(y * (np.float32(1.0) / h)) == 7
    return float(y)


def oracle_period(sort, period):
    return period if sort else 0


def check_substep(t, ref, sorted_now, engine, ctx):
    """One substep's outputs against the restatement's."""
    st = t.stats()
    assert st["sorted"] == int(sorted_now) == int(ref.sorted), ctx
    if sorted_now:
        assert np.array_equal(t.permutation(), ref.perm), ctx
        assert np.array_equal(t.cell_offset(), ref.cell_offset), ctx
    assert np.array_equal(t.occupancy(), ref.occupancy), ctx
    assert np.array_equal(t.cell_type(), ref.cell_type), ctx
    saved = t.faces(fsm.FACE_SAVED)
    vel = t.faces(fsm.FACE_VELOCITY)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a]), (ctx, "saved", a)
        assert rel_l2(vel[a], ref.vel_final[a]) <= TOL, (ctx, "faces", a)
    assert rel_l2(t.pressure(), ref.pressure) <= TOL, ctx
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL, ctx
    if engine == fsm.ENGINE_REFERENCE:
        # f32 trees in the WGSL order: the same bits as the restatement
        assert bits_equal(t.pressure(), ref.pressure), ctx
        assert bits_equal(gp, ref.pos) and bits_equal(gv, ref.vel_particles), ctx
        assert st["cg_iterations"] == ref.stats.cg_iterations, ctx
    else:
        assert abs(st["cg_iterations"] - ref.stats.cg_iterations) <= 1, ctx
        if not sorted_now:
            assert st["p2g_gather"] == 1, ctx


def teacher_forced(t, P, pos, vel, pr, sort, period, substeps, engine, name):
    for s in range(substeps):
        ref = oracle.run(P, pos, vel, pressure=pr, steps=1, sort_period=oracle_period(sort, period), tick0=s)
        t.load_state(pos, vel, pr)
        t.step(1)
        t.synchronize()
        sorted_now = bool(sort) and s % period == 0
        say("cadence", name, "engine", engine, "sort", sort, "period", period, "substep", s,
            "sorted", sorted_now)
        check_substep(t, ref, sorted_now, engine, (name, engine, sort, period, s))
        pos, vel, pr = ref.pos, ref.vel_particles, ref.pressure


@pytest.mark.parametrize("sort,period", CADENCES)
@pytest.mark.parametrize("engine", ENGINES)
def test_cadence_teacher_forced_small_tank(engine, sort, period):
    """24^3 tank with escaped particles and particles whose sort cell and
    P2G cell differ, 8 substeps: sorted and unsorted substeps alike."""
    sc = scenes.small_tank(n=24, seed=11, fill=0.45)
    pos = sc["pos"].copy()
    pos[0] = (-0.3, 0.5, 0.5)
    pos[1] = (1.7, 1.2, -0.1)
    pos[5, 1] = brick_edge_misfit_24()
    pos[9, 2] = brick_edge_misfit_24()
    P = oracle_params(sc, 0 if engine == fsm.ENGINE_REFERENCE else 1)
    t = fsm.FluidStep(gpu_params(sc, engine, particle_sort=sort, particle_sort_period=period), pos, sc["vel"])
    teacher_forced(t, P, pos, sc["vel"], np.zeros(24 ** 3, np.float32), sort, period, 8, engine, sc["name"])
    t.close()


def default_scene():
    """The registry's default scene (falling blob, 80x40x80, 10 per seeded
    cell, ~512k particles, 30 CG iterations with the tolerance off, sort
    period 4), seeded on the GPU and checked against the restatement's
    lattice."""
    r = fsm.Registry()
    sc = fsm.scene_from_registry(r)
    tank = fsm.FluidStep.from_settings(r, sc)
    p = tank.params
    origin = [-(np.float32(n) * H) / np.float32(2.0) for n in (p.nx, p.ny, p.nz)]
    pos = oracle.scene_generate(sc.blocks_array(), sc.particle_count, H, origin, (p.nx, p.ny, p.nz))
    gp, _ = tank.particles()
    assert bits_equal(gp, pos)
    return tank, p, pos


def oracle_params_of(p, reduce_mode):
    return oracle.make_params(
        p.nx, p.ny, p.nz, p.h, origin=tuple(p.origin), dt=p.dt, rho=p.rho, flip_blend=p.flip_blend,
        wall_friction=p.wall_friction, grav=tuple(p.gravity), rest=p.rest_density, stiff=p.volume_stiffness,
        drift_clamp=p.drift_clamp, liquid_threshold=p.liquid_threshold, surface_dilation=p.surface_dilation,
        cfl=p.cfl, pressure_iters=p.pressure_iterations, warm_start=p.pressure_warm_start,
        tol=p.pressure_residual_tolerance, reduce_mode=reduce_mode)


def tank_of(engine):
    """The default scene's tank for `engine` (the registry builds the fused
    engine's; the reference-structured engine gets the same parameters)."""
    tank, p, pos = default_scene()
    assert (p.particle_sort, p.particle_sort_period, p.pressure_iterations) == (1, 4, 30)
    assert p.pressure_residual_tolerance == 0.0
    if engine == fsm.ENGINE_REFERENCE:
        tank.close()
        p.engine, p.precision = fsm.ENGINE_REFERENCE, fsm.PRECISION_F32
        tank = fsm.FluidStep(p, pos)
    return tank, p, pos


@pytest.mark.parametrize("engine", ENGINES)
def test_cadence_teacher_forced_default_scene(engine):
    tank, p, pos = tank_of(engine)
    P = oracle_params_of(p, 0 if engine == fsm.ENGINE_REFERENCE else 1)
    say("default scene", (p.nx, p.ny, p.nz), "N", len(pos), "engine", engine)
    teacher_forced(tank, P, pos, np.zeros_like(pos), np.zeros(p.nx * p.ny * p.nz, np.float32), 1, 4, 8, engine,
                   "default_80x40x80")
    tank.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_cadence_free_run_default_scene(engine):
    """8 substeps advancing on their own.  Particles are matched by identity
    (each sort's permutation applied to an id column on both sides).  The
    reference-structured engine reproduces the restatement bit for bit; the
    fused engine (f64 dots) stays within 1e-5 on particles, faces and
    pressure."""
    tank, p, pos = tank_of(engine)
    P = oracle_params_of(p, 0 if engine == fsm.ENGINE_REFERENCE else 1)
    n = len(pos)
    ids_g = np.arange(n)
    ids_r = np.arange(n)
    rpos, rvel, rpr = pos, np.zeros_like(pos), np.zeros(p.nx * p.ny * p.nz, np.float32)
    for s in range(8):
        ref = oracle.run(P, rpos, rvel, pressure=rpr, steps=1, sort_period=4, tick0=s)
        tank.step(1)
        tank.synchronize()
        st = tank.stats()
        assert st["sorted"] == int(ref.sorted) == int(s % 4 == 0)
        if ref.sorted:
            ids_r = ids_r[ref.perm]
            ids_g = ids_g[tank.permutation()]
        gp, gv = tank.particles()
        og = np.argsort(ids_g, kind="stable")
        orr = np.argsort(ids_r, kind="stable")
        errs = dict(pos=rel_l2(gp[og], ref.pos[orr]), vel=rel_l2(gv[og], ref.vel_particles[orr]),
                    pressure=rel_l2(tank.pressure(), ref.pressure),
                    faces=max(rel_l2(f, ref.vel_final[a]) for a, f in enumerate(tank.faces())),
                    same_order=bool(np.array_equal(ids_g, ids_r)))
        say("free run default scene engine", engine, "substep", s, errs)
        if engine == fsm.ENGINE_REFERENCE:
            assert errs["same_order"]
            assert bits_equal(gp, ref.pos) and bits_equal(gv, ref.vel_particles)
            assert bits_equal(tank.pressure(), ref.pressure)
        for k in ("pos", "vel", "pressure", "faces"):
            assert errs[k] <= TOL, (s, k, errs)
        rpos, rvel, rpr = ref.pos, ref.vel_particles, ref.pressure
    tank.close()


def test_unsorted_substeps_keep_the_last_sorts_bin_starts():
    """cell_offset holds the bin starts of the last sort (the reference writes
    it only in record_sort, src/gpu/fluid.rs:1522-1571), also after substeps
    whose gather built its own cell order."""
    sc = scenes.small_tank(n=20, seed=4, fill=0.45)
    t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED, particle_sort=1, particle_sort_period=4), sc["pos"],
                      sc["vel"])
    t.step(1)
    first = t.cell_offset()
    occ0 = t.occupancy()
    assert np.array_equal(first, np.concatenate([[0], np.cumsum(occ0)[:-1]]).astype(np.uint32))
    t.step(3)
    assert t.stats()["sorted"] == 0 and t.stats()["p2g_gather"] == 1
    assert np.array_equal(t.cell_offset(), first)
    assert not np.array_equal(t.occupancy(), occ0)
    t.close()


def test_reference_engine_local_scatter_with_a_full_table(monkeypatch):
    """The reference-structured engine scatters through the workgroup-local
    table (sh/scatter_local.wgsl) whenever the sort is enabled.  Particles in
    random order on a substep without a sort make a 64-particle workgroup's
    taps overflow the 1024-slot table, so the rest go to global memory
    directly; the integer sums still equal the restatement's, and the per-tap
    atomic form's (FLUIDSTEP_REF_P2G=atomic), bit for bit."""
    sc = scenes.small_tank(n=24, seed=5, fill=0.5)
    rng = np.random.default_rng(17)
    order = rng.permutation(len(sc["pos"]))
    pos, vel = sc["pos"][order].copy(), sc["vel"][order].copy()
    pr = np.zeros(24 ** 3, np.float32)
    P = oracle_params(sc, 0)
    prm = gpu_params(sc, fsm.ENGINE_REFERENCE, particle_sort=1, particle_sort_period=4)
    outs = []
    for form in ("local", "atomic"):
        if form == "atomic":
            monkeypatch.setenv("FLUIDSTEP_REF_P2G", "atomic")
        t = fsm.FluidStep(prm, pos, vel)
        t.step(1)  # tick 0 sorts; tick 1 does not
        t.load_state(pos, vel, pr)
        t.step(1)
        say("full local table", form, "N", len(pos))
        assert t.stats()["sorted"] == 0
        outs.append((t.faces(fsm.FACE_SAVED), t.pressure(), *t.particles()))
        t.close()
    ref = oracle.run(P, pos, vel, pressure=pr, steps=1, sort_period=4, tick0=1)
    for saved, p, gp, gv in outs:
        for a in range(3):
            assert bits_equal(saved[a], ref.saved[a]), a
        assert bits_equal(p, ref.pressure)
        assert bits_equal(gp, ref.pos) and bits_equal(gv, ref.vel_particles)
