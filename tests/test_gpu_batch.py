"""Batched tanks (fs_batch_*) against the CPU restatement, tank by tank.

Every tank of a batch runs the fused engine's step (mixed precision, sort
every substep), so each is held to the single-tank bar: binning, bin starts,
permutation and the P2G `saved` faces bit-exact; face velocities, pressure and
particles within 1e-5 relative L2 after one substep from the same state.
Tile-shape variants of the autotuner's menu must not change results beyond
the f64 reduction order (binning and P2G bit-exact; the rest within 1e-5).
"""
import os
import sys

import numpy as np
import pytest

import fluidstep as fsm
import oracle
import scenes
from helpers import bits_equal, gpu_params, oracle_params, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch():
    """Six tanks of uneven sizes: power-of-two and not (misfit particles),
    one nearly empty, random initial velocities."""
    specs = [(16, 0.5, 8, 1), (24, 0.45, 8, 2), (20, 0.3, 5, 3), (32, 0.6, 8, 4), (12, 0.2, 4, 5), (16, 0.15, 8, 6)]
    tanks = []
    for n, fill, ppc, seed in specs:
        sc = scenes.small_tank(n=n, seed=seed, fill=fill, ppc=ppc, vel_scale=0.5)
        tanks.append(sc)
    # The last tank holds a handful of particles only.
    tanks[-1]["pos"] = tanks[-1]["pos"][:7].copy()
    tanks[-1]["vel"] = tanks[-1]["vel"][:7].copy()
    return tanks


def make_batch(tanks, velocity=fsm.VELOCITY_F32):
    params = [gpu_params(sc, fsm.ENGINE_FUSED, velocity=velocity) for sc in tanks]
    return fsm.FluidBatch(params, [sc["pos"] for sc in tanks], [sc["vel"] for sc in tanks])


def oracle_step(sc, steps=1):
    P = oracle_params(sc, 1)
    return oracle.run(P, sc["pos"], sc["vel"], steps=steps, sort_period=1)


def test_batch_one_substep_matches_oracle_per_tank():
    tanks = small_batch()
    say("batch", [len(sc["pos"]) for sc in tanks])
    b = make_batch(tanks)
    b.step(1)
    b.synchronize()
    for t, sc in enumerate(tanks):
        say("tank", t, sc["name"], len(sc["pos"]))
        ref = oracle_step(sc)
        st = b.stats(t)
        assert st["p2g_gather"] == 1
        assert np.array_equal(b.occupancy(t), ref.occupancy)
        assert np.array_equal(b.cell_offset(t), ref.cell_offset)
        assert np.array_equal(b.permutation(t), ref.perm)
        assert st["liquid_cells"] == int((ref.cell_type == 1).sum())
        assert np.array_equal(b.cell_type(t), ref.cell_type)
        saved = b.faces(t, fsm.FACE_SAVED)
        for a in range(3):
            assert bits_equal(saved[a], ref.saved[a]), (t, a)
        vel = b.faces(t, fsm.FACE_VELOCITY)
        for a in range(3):
            assert rel_l2(vel[a], ref.vel_final[a]) <= TOL, (t, a)
        assert rel_l2(b.pressure(t), ref.pressure) <= TOL
        gp, gv = b.particles(t)
        if len(gp):
            assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    b.close()


def test_batch_tank_equals_single_tank_over_substeps():
    """Five substeps of a batch against the single-tank fused engine on each
    tank: the same kernels per tank, so results agree to the f64 reduction
    order of the CG."""
    tanks = small_batch()[:4]
    b = make_batch(tanks)
    b.step(5)
    b.synchronize()
    for t, sc in enumerate(tanks):
        s = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
        s.step(5)
        s.synchronize()
        say("tank", t, "iters", b.stats(t)["cg_iterations"], s.stats()["cg_iterations"])
        assert rel_l2(b.pressure(t), s.pressure()) <= 1e-4
        bp, bv = b.particles(t)
        sp, sv = s.particles()
        assert rel_l2(bp, sp) <= 1e-5 and rel_l2(bv, sv) <= 1e-4
        s.close()
    b.close()


@pytest.mark.parametrize("brick,sort_tile,unroll", [(1, 0, 0), (2, 0, 0), (0, 1, 0), (0, 2, 0), (0, 0, 1), (0, 0, 2)])
def test_batch_tuning_menu_entries_agree(brick, sort_tile, unroll):
    tanks = small_batch()[:4]
    say("menu", brick, sort_tile, unroll)
    base = make_batch(tanks)
    base.step(2)
    var = make_batch(tanks)
    var.set_tuning(gather_brick=brick, sort_tile=sort_tile, cg_unroll=unroll)
    var.step(2)
    for t in range(len(tanks)):
        assert np.array_equal(base.permutation(t), var.permutation(t))
        assert np.array_equal(base.cell_offset(t), var.cell_offset(t))
        assert rel_l2(var.pressure(t), base.pressure(t)) <= TOL
        bp, bv = base.particles(t)
        vp, vv = var.particles(t)
        assert rel_l2(vp, bp) <= TOL and rel_l2(vv, bv) <= TOL
    base.close()
    var.close()


def test_batch_autotune_keeps_the_state():
    """Tuning runs trial phases on the live state and puts it back: the next
    substep matches a batch that was never tuned."""
    tanks = small_batch()[:4]
    a = make_batch(tanks)
    b = make_batch(tanks)
    a.step(1)
    b.step(1)
    tuned = b.autotune()
    say("tuned", tuned)
    assert tuned["gather_brick"] in fsm.Tuning.BRICKS and tuned["ms_cg"] > 0
    a.step(1)
    b.step(1)
    for t in range(len(tanks)):
        assert np.array_equal(a.permutation(t), b.permutation(t))
        assert rel_l2(b.pressure(t), a.pressure(t)) <= TOL
        ap, av = a.particles(t)
        bp, bv = b.particles(t)
        assert rel_l2(bp, ap) <= TOL and rel_l2(bv, av) <= TOL
    a.close()
    b.close()


def test_batch_graph_replay_equals_direct_launches():
    """The captured substep graph and the direct launches run the same
    kernels in the same order: identical results, replay after replay."""
    tanks = small_batch()[:3]
    g = make_batch(tanks)
    os.environ["FLUIDSTEP_BATCH_GRAPH"] = "0"
    try:
        d = make_batch(tanks)
    finally:
        del os.environ["FLUIDSTEP_BATCH_GRAPH"]
    for _ in range(3):
        g.step(1)
        d.step(1)
        for t in range(len(tanks)):
            gp, gv = g.particles(t)
            dp, dv = d.particles(t)
            assert bits_equal(gp, dp) and bits_equal(gv, dv)
            assert bits_equal(g.pressure(t), d.pressure(t))
    g.close()
    d.close()


def test_batch_bf16_velocity_storage():
    tanks = small_batch()[:3]
    b = make_batch(tanks, velocity=fsm.VELOCITY_BF16)
    b.step(1)
    for t, sc in enumerate(tanks):
        s = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED, velocity=fsm.VELOCITY_BF16), sc["pos"], sc["vel"])
        s.step(1)
        assert np.array_equal(b.permutation(t), s.permutation())
        for a in range(3):
            assert bits_equal(b.faces(t, fsm.FACE_SAVED)[a], s.faces(fsm.FACE_SAVED)[a])
        assert rel_l2(b.pressure(t), s.pressure()) <= TOL
        s.close()
    b.close()


def test_batch_of_the_benchmark_tanks_one_substep():
    """The 64-tank benchmark batch: one substep, every tank's solve converged
    or at its cap, binning invariants hold (bin starts are the exclusive scan
    of occupancy; the permutation is a permutation)."""
    tanks = scenes.batched_tanks()
    say("64 tanks", sum(len(sc["pos"]) for sc in tanks))
    b = make_batch(tanks)
    b.step(1)
    b.synchronize()
    for t, sc in enumerate(tanks):
        st = b.stats(t)
        assert st["cg_iterations"] >= 1
        assert st["true_residual_sq"] <= max(1e-6, 1e-6 * st["rs_initial"]) or st["cg_iterations"] == 200
        occ = b.occupancy(t)
        off = b.cell_offset(t)
        assert occ.sum() == len(sc["pos"])
        assert np.array_equal(off[1:], np.cumsum(occ, dtype=np.uint64)[:-1].astype(np.uint32))
        perm = b.permutation(t)
        assert np.array_equal(np.sort(perm), np.arange(len(sc["pos"]), dtype=np.uint32))
    b.close()


@pytest.mark.parametrize("which,unroll", [("small", 0), ("small", 1), ("small", 2), ("benchmark", 0)])
def test_batch_barrier_forms_give_identical_bits(which, unroll, monkeypatch):
    """The batched solve's barrier after phase B: counters with
    release/acquire (lockstep_sum) or the fence-free form (lockstep_coherent:
    write-through r and d, coherent neighbour loads).  Both run the same
    arithmetic in the same order, so a stale neighbour value would show as a
    bit difference in some tank."""
    tanks = small_batch() if which == "small" else scenes.batched_tanks()
    outs = []
    for form in ("counter", "coherent"):
        monkeypatch.setenv("FLUIDSTEP_BATCH_DATA_BARRIER", form)
        say("batch barrier form", which, unroll, form)
        b = make_batch(tanks)
        if unroll:
            b.set_tuning(cg_unroll=unroll)
        b.step(2 if which == "small" else 1)
        b.synchronize()
        per = []
        for t in range(len(tanks)):
            gp, gv = b.particles(t)
            per.append(([b.pressure(t), *b.faces(t, fsm.FACE_VELOCITY), gp, gv], b.stats(t)["cg_iterations"]))
        outs.append(per)
        b.close()
    for t, ((a_arrs, a_it), (b_arrs, b_it)) in enumerate(zip(outs[0], outs[1])):
        assert a_it == b_it, t
        for i, (x, y) in enumerate(zip(a_arrs, b_arrs)):
            assert bits_equal(x, y), (t, i)
