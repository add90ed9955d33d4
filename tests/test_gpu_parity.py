"""Parity of the HIP step (through the C-ABI) against the CPU restatement.

Bar: bit-exact for binning (occupancy, bin starts, permutation) and for every
integer/P2G result; within 1e-5 relative L2 for face velocities, pressure and
particles per step, starting both sides from the same state (teacher-forced).
The reference-structured engine mirrors the f32 WGSL reduction trees and is
additionally expected to match bit for bit; the fused engine computes its CG
dots in f64, like src/sim/pressure.rs:182-188.
"""
import sys

import numpy as np
import pytest

import fluidstep as fsm
import oracle
import scenes
from helpers import bits_equal, gpu_params, oracle_params, rel_l2

pytestmark = pytest.mark.gpu

ENGINES = [fsm.ENGINE_REFERENCE, fsm.ENGINE_FUSED]
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def reduce_mode(engine):
    return 0 if engine == fsm.ENGINE_REFERENCE else 1


def run_pair(scene, engine, steps=1, pos=None, vel=None, pressure=None, velocity=fsm.VELOCITY_F32, **over):
    """One teacher-forced comparison: both sides start from (pos, vel,
    pressure) and run `steps` substeps (sort every substep)."""
    pos = scene["pos"] if pos is None else pos
    vel = scene["vel"] if vel is None else vel
    P = oracle_params(scene, reduce_mode(engine), **over)
    ref = oracle.run(P, pos, vel, pressure=pressure, steps=steps, sort_period=1)
    t = fsm.FluidStep(gpu_params(scene, engine, velocity=velocity, **over), pos, vel)
    if pressure is not None:
        t.load_state(pos, vel, pressure)
    t.step(steps)
    t.synchronize()
    return ref, t


@pytest.mark.parametrize("engine", ENGINES)
def test_binning_bit_exact(engine):
    sc = scenes.dam_break_32()
    say("binning", sc["name"], "engine", engine, "N", len(sc["pos"]))
    ref, t = run_pair(sc, engine)
    assert np.array_equal(t.occupancy(), ref.occupancy)
    assert np.array_equal(t.cell_offset(), ref.cell_offset)
    assert np.array_equal(t.permutation(), ref.perm)
    assert t.stats()["sorted"] == 1
    t.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_step_parity_dam_break(engine):
    sc = scenes.dam_break_32()
    say("step parity", sc["name"], "engine", engine)
    ref, t = run_pair(sc, engine)
    st = t.stats()
    assert np.array_equal(t.cell_type(), ref.cell_type)
    assert st["liquid_cells"] == ref.stats.liquid_cells
    saved = t.faces(fsm.FACE_SAVED)
    vel = t.faces(fsm.FACE_VELOCITY)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a]), f"saved axis {a}"
        assert rel_l2(vel[a], ref.vel_final[a]) <= TOL
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    pos, pv = t.particles()
    assert rel_l2(pos, ref.pos) <= TOL
    assert rel_l2(pv, ref.vel_particles) <= TOL
    if engine == fsm.ENGINE_REFERENCE:
        assert st["cg_iterations"] == ref.stats.cg_iterations
        assert bits_equal(t.pressure(), ref.pressure)
        assert bits_equal(pos, ref.pos) and bits_equal(pv, ref.vel_particles)
    else:
        assert st["p2g_gather"] == 1
        assert abs(st["cg_iterations"] - ref.stats.cg_iterations) <= 1
    t.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_teacher_forced_steps(engine):
    sc = scenes.small_tank(n=16, seed=7)
    P = oracle_params(sc, reduce_mode(engine))
    pos, vel, pr = sc["pos"], sc["vel"], np.zeros(16 ** 3, np.float32)
    t = fsm.FluidStep(gpu_params(sc, engine), pos, vel)
    for s in range(4):
        say("teacher-forced", sc["name"], "engine", engine, "step", s)
        ref = oracle.run(P, pos, vel, pressure=pr, steps=1, sort_period=1)
        t.load_state(pos, vel, pr)
        t.step(1)
        gp, gv = t.particles()
        assert rel_l2(t.pressure(), ref.pressure) <= TOL
        assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
        for a, f in enumerate(t.faces()):
            assert rel_l2(f, ref.vel_final[a]) <= TOL
        pos, vel, pr = ref.pos, ref.vel_particles, ref.pressure
    t.close()


def _misfit_coordinate(n):
    """A float32 x in (0,1) where floor(x/h) != floor(x*f32(1/h)) for h=f32(1/n)."""
    h = np.float32(1.0) / np.float32(n)
    inv = np.float32(1.0) / h
    for c in range(2, n - 2):
        x0 = np.float32(c) * h
        x = x0
        for _ in range(64):
            if np.floor(x / h) != np.floor(x * inv):
                return float(x)
            x = np.nextafter(x, np.float32(0))
        x = x0
        for _ in range(64):
            if np.floor(x / h) != np.floor(x * inv):
                return float(x)
            x = np.nextafter(x, np.float32(1))
    return None


def _brick_edge_misfit_24():
    """At n = 24, y = f32(0.29166666) sorts into cell 6 (y/h) but scatters
    from cell 7 (y*inv_h): its v taps reach face j = 8, owned by the gather
    brick starting at y = 8, whose halo (cells 7..16) does not hold cell 6."""
    h = np.float32(1.0) / np.float32(24)
    y = np.float32(0.29166666)
    assert np.floor(y / h) == 6 and np.floor(y * (np.float32(1.0) / h)) == 7
    return float(y)


@pytest.mark.parametrize("engine", ENGINES)
def test_non_power_of_two_grid_with_misfit(engine):
    """24^3 (h = 1/24): particles whose sort cell (divide by h) and P2G cell
    (multiply by 1/h) differ, one of them reaching a gather brick from outside
    its halo.  The fused engine's gather adds them from the misfit list;
    results still match bit for bit on the P2G side."""
    sc = scenes.small_tank(n=24, seed=11, fill=0.45)
    x = _misfit_coordinate(24)
    assert x is not None
    pos = sc["pos"].copy()
    pos[5, 0] = x
    pos[6, 1] = _brick_edge_misfit_24()
    pos[7, 2] = _brick_edge_misfit_24()
    say("misfit", sc["name"], "engine", engine, "x", x)
    ref, t = run_pair(sc, engine, pos=pos)
    st = t.stats()
    if engine == fsm.ENGINE_FUSED:
        assert st["misfits"] >= 3 and st["p2g_gather"] == 1
    assert np.array_equal(t.permutation(), ref.perm)
    saved = t.faces(fsm.FACE_SAVED)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a])
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    t.close()


@pytest.mark.parametrize("form", ["owner", "particle"])
def test_many_misfits(monkeypatch, form):
    """More misfit particles than the face-owning gather's list holds (4096).
    The face-owning gather (the default) switches to the atomic P2G form for
    the step; the particle-owned brick scatter (FLUIDSTEP_P2G_FORM=particle)
    lists them per brick and adds their taps one by one.  Results match."""
    monkeypatch.setenv("FLUIDSTEP_P2G_FORM", form)
    sc = scenes.small_tank(n=24, seed=12, fill=0.45)
    pos = sc["pos"].copy()
    assert len(pos) > 6000
    pos[:5000, 1] = _brick_edge_misfit_24()
    say("misfit overflow", sc["name"], "N", len(pos), "form", form)
    ref, t = run_pair(sc, fsm.ENGINE_FUSED, pos=pos)
    st = t.stats()
    assert st["misfits"] >= 5000 and st["p2g_gather"] == (1 if form == "particle" else 0)
    assert np.array_equal(t.permutation(), ref.perm)
    saved = t.faces(fsm.FACE_SAVED)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a])
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    t.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_escaped_particles_clamp_into_edge_cells(engine):
    """Particles outside the tank bin into the clamped edge cells (src/lib.rs:
    1286-1300) and G2P pulls them back inside."""
    sc = scenes.small_tank(n=16, seed=3)
    pos = sc["pos"].copy()
    pos[0] = (-0.3, 0.5, 0.5)
    pos[1] = (1.7, 1.2, -0.1)
    say("escaped", engine)
    ref, t = run_pair(sc, engine, pos=pos)
    assert np.array_equal(t.occupancy(), ref.occupancy)
    assert np.array_equal(t.permutation(), ref.perm)
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    t.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_empty_tank(engine):
    sc = scenes.small_tank(n=16)
    t = fsm.FluidStep(gpu_params(sc, engine), np.zeros((0, 3), np.float32))
    say("empty", engine)
    t.step(2)
    st = t.stats()
    assert st["liquid_cells"] == 0
    assert np.all(t.pressure() == 0)
    t.close()


def test_fused_is_deterministic():
    sc = scenes.dam_break_32()
    say("determinism")
    outs = []
    for _ in range(2):
        t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
        t.step(5)
        pos, vel = t.particles()
        outs.append((pos, vel, t.pressure()))
        t.close()
    for a, b in zip(outs[0], outs[1]):
        assert bits_equal(a, b)


def test_bf16_velocity_storage():
    """BF16 particle velocities: P2G reads bf16, G2P stores bf16 (f32 math).
    Against the oracle fed the same bf16-rounded velocities: positions match
    to 1e-5 and the stored velocity is the bf16 rounding of the f32 result."""
    sc = scenes.dam_break_32()
    rng = np.random.default_rng(0)
    vel = (rng.random(sc["pos"].shape, dtype=np.float32) - 0.5).astype(np.float32)
    vb = (vel.view(np.uint32) + 0x7FFF + ((vel.view(np.uint32) >> 16) & 1)) & 0xFFFF0000
    vel_b = vb.astype(np.uint32).view(np.float32)
    say("bf16")
    ref, t = run_pair(sc, fsm.ENGINE_FUSED, vel=vel_b, velocity=fsm.VELOCITY_BF16)
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL
    rv = ref.vel_particles
    rb = ((rv.view(np.uint32) + 0x7FFF + ((rv.view(np.uint32) >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    assert rel_l2(gv, rb.view(np.float32)) <= 1e-4
    t.close()


def test_live_setters_change_the_step():
    sc = scenes.small_tank(n=16, seed=9)
    t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
    r = t.set_param("physics.flip_blend", 0.5)
    assert r["status"] == "applied" and r["applied_live"]
    r = t.set_param("physics.cfl", 1e100)
    assert r["stored_value"] == 6.0 and r["clamped"]
    say("setters")
    ref = oracle.run(oracle_params(sc, 1, flip_blend=0.5, cfl=6.0), sc["pos"], sc["vel"], steps=1)
    t.step(1)
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    t.close()


def test_impulse_and_reset():
    sc = scenes.small_tank(n=16, seed=2)
    t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
    t.apply_impulse(0.25, -0.5, 1.0)
    _, v = t.particles()
    exp = sc["vel"] + np.array([0.25, -0.5, 1.0], np.float32)
    assert bits_equal(v, exp.astype(np.float32))
    t.step(2)
    t.reset()
    p, v = t.particles()
    assert bits_equal(p, sc["pos"]) and bits_equal(v, sc["vel"])
    assert np.all(t.pressure() == 0)
    t.close()


def test_full_size_high_density_one_step():
    """128^3, ~6.7M particles: one teacher-forced step of the fused engine
    against the restatement (f64 dots), plus size-independent invariants."""
    sc = scenes.high_density_128()
    N = len(sc["pos"])
    say("full size", sc["name"], "N", N)
    ref, t = run_pair(sc, fsm.ENGINE_FUSED)
    occ = t.occupancy()
    assert int(occ.sum()) == N
    off = t.cell_offset()
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(occ, dtype=np.uint64)[:-1]]).astype(np.uint32))
    perm = t.permutation()
    assert np.array_equal(perm, ref.perm)
    st = t.stats()
    # The cold first solve of this tank stops at the 200-iteration cap on both
    # sides (the restatement reports the same count); a converged solve meets
    # the relative tolerance.
    assert abs(st["cg_iterations"] - ref.stats.cg_iterations) <= 1
    if st["cg_iterations"] < 200:
        assert st["rs_final"] <= (1e-5 ** 2) * st["rs_initial"] * 1.0001
    assert st["true_residual_sq"] <= 1.01 * st["rs_final"] + 1e-12 * st["rs_initial"]
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    gp, gv = t.particles()
    assert np.all(np.isfinite(gp)) and np.all(np.isfinite(gv))
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    t.close()


@pytest.mark.timeout(900)
def test_full_size_large_splash_256():
    """256^3, pool plus falling sphere (35,450,304 particles): one
    teacher-forced step of the fused engine against the restatement, then
    two more free-running substeps under size-independent invariants
    (occupancy sums to N, bin starts are the scan of the occupancy, the
    permutation is a bijection that orders the particles by cell, the state
    stays finite and inside the tank)."""
    sc = scenes.large_splash_256()
    N = len(sc["pos"])
    n = sc["n"]
    say("full size", sc["name"], "N", N)
    ref, t = run_pair(sc, fsm.ENGINE_FUSED)
    say("large splash: restatement step done, cg", ref.stats.cg_iterations)
    assert np.array_equal(t.occupancy(), ref.occupancy)
    assert np.array_equal(t.permutation(), ref.perm)
    assert np.array_equal(t.cell_type(), ref.cell_type)
    saved = t.faces(fsm.FACE_SAVED)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a]), a
    st = t.stats()
    assert abs(st["cg_iterations"] - ref.stats.cg_iterations) <= 1
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    del ref, saved
    h = np.float32(1.0) / np.float32(n)
    for s in range(2):
        t.step(1)
        say("large splash: free substep", s)
        occ = t.occupancy()
        assert int(occ.sum(dtype=np.uint64)) == N
        off = t.cell_offset()
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(occ, dtype=np.uint64)[:-1]]).astype(np.uint32))
        perm = t.permutation()
        seen = np.zeros(N, np.bool_)
        seen[perm] = True
        assert seen.all()
        gp, gv = t.particles()
        assert np.all(np.isfinite(gp)) and np.all(np.isfinite(gv))
        lo, hi = np.float32(1.05) * h, np.float32(1.0) - np.float32(1.05) * h
        assert gp.min() >= lo - 1e-6 and gp.max() <= hi + 1e-6
        assert np.all(np.isfinite(t.pressure()))
    t.close()


@pytest.mark.parametrize("cap", [30, 31])
def test_streaming_solve_parity(cap, monkeypatch):
    """The persistent solve's streaming form (a workgroup walks its tiles in
    batches instead of holding them in registers; p takes two iterations'
    updates at a time).  A launch capped at 8 workgroups gives every
    workgroup of this 64^3 tank more tiles than its waves hold.  With the
    tolerance off the solve runs exactly `cap` iterations; an odd count leaves
    one deferred p update for after the loop."""
    monkeypatch.setenv("FLUIDSTEP_CG_BLOCKS", "8")
    sc = scenes.control_64()
    say("streaming solve parity", sc["name"], "cap", cap)
    ref, t = run_pair(sc, fsm.ENGINE_FUSED, pressure_iterations=cap, pressure_residual_tolerance=0.0)
    st = t.stats()
    assert st["cg_iterations"] == cap and ref.stats.cg_iterations == cap
    assert rel_l2(t.pressure(), ref.pressure) <= TOL
    vel = t.faces(fsm.FACE_VELOCITY)
    for a in range(3):
        assert rel_l2(vel[a], ref.vel_final[a]) <= TOL
    pos, pv = t.particles()
    assert rel_l2(pos, ref.pos) <= TOL
    assert rel_l2(pv, ref.vel_particles) <= TOL
    t.close()


@pytest.mark.parametrize("engine", ENGINES)
def test_saturating_p2g_velocities(engine):
    """A few particles fast enough that w * v * 2^16 leaves the i32 range:
    the P2G conversion saturates as WGSL's i32() does, on the gather's
    saturating path (the plain conversion is taken only where it cannot
    saturate).  One particle at 3e38 overflows w * v * 2^16 to infinity.
    The integer sums, hence `saved`, stay bit-exact."""
    sc = scenes.small_tank(n=16, seed=11)
    vel = np.array(sc["vel"], np.float32, copy=True)
    fast = np.arange(0, len(vel), max(1, len(vel) // 7))[:7]
    vel[fast] = np.array([[2.0e5, -3.0e5, 1.0e9]] * len(fast), np.float32)
    vel[fast[1::2], 0] = -4.0e4
    vel[fast[2], 1] = -3.0e38
    say("saturating P2G", engine, "fast particles", len(fast))
    ref, t = run_pair(sc, engine, vel=vel)
    assert np.array_equal(t.cell_type(), ref.cell_type)
    saved = t.faces(fsm.FACE_SAVED)
    for a in range(3):
        assert bits_equal(saved[a], ref.saved[a]), f"saved axis {a}"
    t.close()


@pytest.mark.parametrize("scene_name,forms,blocks", [
    ("dam_break_32", ("tagged", "coherent", "counter"), None),
    ("control_64", ("tagged", "coherent"), None),
    ("control_64", ("tagged", "coherent", "counter"), 8),
    ("high_density_128", ("counter", "auto", "coherent", "tagged"), None),
])
def test_barrier_forms_give_identical_bits(scene_name, forms, blocks, monkeypatch):
    """The persistent solve's memory-ordering barrier after phase B has three
    forms: counters with release/acquire (grid_sum), tagged slots with fences
    (grid_sum_tagged), and the fence-free form of the register-resident solve
    (coherent_sum: write-through stores and coherent neighbour loads, no L2
    write-back or invalidate).  All run the same arithmetic in the same
    order, so a stale neighbour value read by any form would show as a bit
    difference.  The 128^3 tank solves on 208 workgroups, all resident, where
    the default ("auto") takes the fence-free form; the 64^3 tank capped at 8
    workgroups streams its vectors (the fence-free form then stores r and d
    through and reads neighbours coherently in the streaming phases)."""
    if blocks is not None:
        monkeypatch.setenv("FLUIDSTEP_CG_BLOCKS", str(blocks))
    sc = getattr(scenes, scene_name)()
    outs = []
    for form in forms:
        monkeypatch.setenv("FLUIDSTEP_CG_DATA_BARRIER", form)
        say("barrier form", scene_name, form)
        t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
        t.step(3)
        pos, vel = t.particles()
        faces = t.faces(fsm.FACE_VELOCITY)
        outs.append(([t.pressure(), *faces, pos, vel], t.stats()["cg_iterations"]))
        t.close()
    base, iters = outs[0]
    for (arrs, it), form in zip(outs[1:], forms[1:]):
        assert it == iters, form
        for i, (a, b) in enumerate(zip(base, arrs)):
            assert bits_equal(a, b), (form, i)


@pytest.mark.parametrize("n,seed,sort,period", [(20, 4, 1, 4), (24, 11, 1, 3), (16, 3, 0, 1)])
def test_unsorted_substeps_gather_equals_atomic(monkeypatch, n, seed, sort, period):
    """Substeps without a sort (the reference's default sorts every 4th): the
    fused engine's gather reads the particles through the live side's stable
    cell order, and must give the bits of the atomic P2G form
    (FLUIDSTEP_UNSORTED_P2G=atomic), with escaped and misfit particles in the
    tank, and keep the particle order (no move)."""
    sc = scenes.small_tank(n=n, seed=seed, fill=0.45)
    pos = sc["pos"].copy()
    pos[0] = (-0.3, 0.5, 0.5)
    pos[1] = (1.7, 1.2, -0.1)
    if n == 24:
        pos[5, 1] = _brick_edge_misfit_24()
    p = gpu_params(sc, fsm.ENGINE_FUSED, particle_sort=sort, particle_sort_period=period)
    say("unsorted gather", sc["name"], "sort", sort, "period", period)
    monkeypatch.setenv("FLUIDSTEP_UNSORTED_P2G", "atomic")
    a = fsm.FluidStep(p, pos, sc["vel"])
    monkeypatch.delenv("FLUIDSTEP_UNSORTED_P2G")
    g = fsm.FluidStep(p, pos, sc["vel"])
    for _ in range(5):
        a.step(1)
        g.step(1)
        sa, sg = a.stats(), g.stats()
        assert sa["sorted"] == sg["sorted"]
        if not sg["sorted"]:
            assert sa["p2g_gather"] == 0 and sg["p2g_gather"] == 1
        for k in (fsm.FACE_VELOCITY, fsm.FACE_SAVED):
            fa, fg = a.faces(k), g.faces(k)
            assert all(bits_equal(fa[i], fg[i]) for i in range(3))
        assert bits_equal(a.pressure(), g.pressure())
        pa, va = a.particles()
        pg, vg = g.particles()
        assert bits_equal(pa, pg) and bits_equal(va, vg)
    a.close(), g.close()


@pytest.mark.parametrize("period", [1, 4])
def test_gather_brick_order_gives_identical_bits(monkeypatch, period):
    """On grids of more bricks than are resident at once (128^3: 4096 bricks)
    the gather runs its bricks heavy-first, in an order formed from the
    previous gather's per-brick particle counts (and from zeroed counts on the
    first substep).  The order only decides which workgroup runs a brick
    when; each brick writes its own faces from integer sums, so faces,
    pressure and particles equal those of brick order bit for bit, on sorted
    substeps and (period 4) on the unsorted ones, whose gather walks a cell
    order of the particles."""
    sc = scenes.high_density_128()
    outs = []
    for order in ("1", "0"):
        monkeypatch.setenv("FLUIDSTEP_GATHER_ORDER", order)
        say("gather brick order", order, "period", period)
        t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED, particle_sort_period=period), sc["pos"], sc["vel"])
        arrs = []
        for _ in range(4):
            t.step(1)
            arrs += [*t.faces(fsm.FACE_SAVED), *t.faces(fsm.FACE_VELOCITY), t.pressure(), *t.particles()]
        assert t.stats()["p2g_gather"] == 1
        outs.append(arrs)
        t.close()
    for i, (a, b) in enumerate(zip(*outs)):
        assert bits_equal(a, b), i


@pytest.mark.parametrize("scene_name,period,velocity", [("high_density_128", 1, fsm.VELOCITY_F32),
                                                        ("high_density_128", 4, fsm.VELOCITY_F32),
                                                        ("dam_break_32", 1, fsm.VELOCITY_F32),
                                                        ("dam_break_32", 4, fsm.VELOCITY_BF16),
                                                        ("small_tank_24", 4, fsm.VELOCITY_F32)])
def test_g2p_row_chunks_give_identical_bits(monkeypatch, scene_name, period, velocity):
    """G2P by row chunks (the faces a chunk's particles read staged in LDS, a
    wave whose stencils leave the staged rows reading that axis from global
    memory) against one particle per thread reading every face from global
    memory: the same face values in the same arithmetic, so faces, pressure
    and particles agree bit for bit over four substeps, on sorted substeps
    (the permutation applied as G2P writes) and (period 4) on unsorted ones
    (written back in place through the cell order), with the row chunks
    forced on every substep, forced off, and as chosen by default (at 128^3:
    row chunks on the sort substeps only), with f32 and (32^3) bf16 particle
    velocities.  The 24^3 tank's h is not a power of two, so some particles'
    stencils leave their sort cell's rows."""
    if scene_name == "small_tank_24":
        sc = scenes.small_tank(n=24, seed=21, fill=0.5)
    else:
        sc = getattr(scenes, scene_name)()
    outs = []
    for rows in ("1", "0", None):
        if rows is None:
            monkeypatch.delenv("FLUIDSTEP_G2P_ROWS", raising=False)
        else:
            monkeypatch.setenv("FLUIDSTEP_G2P_ROWS", rows)
        say("g2p rows", scene_name, rows, "period", period, "velocity", velocity)
        t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED, velocity=velocity, particle_sort_period=period),
                          sc["pos"], sc["vel"])
        arrs = []
        for _ in range(4):
            t.step(1)
            arrs += [*t.faces(fsm.FACE_SAVED), *t.faces(fsm.FACE_VELOCITY), t.pressure(), *t.particles()]
        outs.append(arrs)
        t.close()
    for i, (a, b, c) in enumerate(zip(*outs)):
        assert bits_equal(a, b) and bits_equal(a, c), i


@pytest.mark.parametrize("scene_name,sort,period", [("dam_break_32", 1, 1), ("high_density_128", 1, 1),
                                                    ("small_tank_24", 1, 4)])
def test_p2g_forms_give_identical_bits(monkeypatch, scene_name, sort, period):
    """The fused engine's two P2G forms - the particle-owned brick scatter
    (each particle's taps added once, by the brick of its sort cell, then a
    face pass) and the face-owning gather (each brick sums the taps of the
    particles of its halo that reach its faces) - add the same integer terms,
    so faces, pressure and particles agree bit for bit over four substeps,
    on sorted and (period 4) unsorted substeps."""
    if scene_name == "small_tank_24":
        sc = scenes.small_tank(n=24, seed=21, fill=0.5)
    else:
        sc = getattr(scenes, scene_name)()
    outs = []
    for form in ("owner", "particle"):
        monkeypatch.setenv("FLUIDSTEP_P2G_FORM", form)
        say("p2g form", scene_name, form)
        t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED, particle_sort=sort, particle_sort_period=period),
                          sc["pos"], sc["vel"])
        arrs = []
        for _ in range(4):
            t.step(1)
            arrs += [*t.faces(fsm.FACE_SAVED), *t.faces(fsm.FACE_VELOCITY), t.pressure(), *t.particles()]
        outs.append(arrs)
        t.close()
    for i, (a, b) in enumerate(zip(*outs)):
        assert bits_equal(a, b), i
