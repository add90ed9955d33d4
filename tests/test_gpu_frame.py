"""The frame driver around a tank (fs_frame_*, FluidApp's frame loop,
src/lib.rs:352-398, 506-601, 907-937) against a twin tank driven directly
through fs_step / fs_apply_impulse / fs_set_gravity by the restatement's
accumulator and schedules (oracle/frame_oracle.py): same substep counts,
same impulses at the same frames, hence bit-identical particles; gravity in
the box frame follows the auto-roll pose; pause, queued steps, export frames
and reset behave as the reference's."""
import numpy as np
import pytest

import fluidstep as fsm
import frame_oracle as fo
import scenes
from helpers import bits_equal, gpu_params

pytestmark = pytest.mark.gpu


def pair():
    sc = scenes.small_tank(n=20, seed=4)
    p = gpu_params(sc, fsm.ENGINE_FUSED)
    return fsm.FluidStep(p, sc["pos"], sc["vel"]), fsm.FluidStep(p, sc["pos"], sc["vel"])


def settings(**kw):
    r = fsm.Registry()
    for k, v in kw.items():
        assert r.set(k.replace("__", "."), v)["ok"]
    return r


def same_particles(a, b):
    pa, va = a.particles()
    pb, vb = b.particles()
    return bits_equal(pa, pb) and bits_equal(va, vb)


def test_running_frames_step_like_the_accumulator():
    r = settings(physics__max_substeps=3)
    a, b = pair()
    f = fsm.Frame(a, r)
    ts = fo.Timestep(float(np.float32(r.get("physics.fixed_dt"))), 3)
    frames = [16.6667, 8.0, 33.4, 1000.0, 0.5, 7.0, 16.6667, float("nan"), -5.0, 25.0]
    total = 0
    for ms in frames:
        st = f.run(ms)
        dt_s = float(np.float32(np.float32(max(ms if np.isfinite(ms) else 0.0, 0.0)) / np.float32(1000.0)))
        n = ts.steps_for_frame(dt_s, 1.0)
        assert st["substeps"] == n
        if n:
            b.step(n)
        total += n
    assert f.tick() == total > 0
    assert same_particles(a, b)
    f.close(), a.close(), b.close()


def test_pause_queued_steps_and_export():
    r = settings()
    a, b = pair()
    f = fsm.Frame(a, r)
    f.set_paused(True)
    assert f.run(16.7)["substeps"] == 0 and f.tick() == 0
    f.queue_step()
    f.queue_step()
    st = f.run(16.7)
    assert st["substeps"] == 1 and f.tick() == 1
    assert f.run(16.7)["substeps"] == 1 and f.run(16.7)["substeps"] == 0 and f.tick() == 2
    b.step(2)
    st = f.export(5, 5.0 / 120.0)
    assert st["substeps"] == 5 and st["policy"] == 1 and f.tick() == 7  # "fixed-explicit-export"
    b.step(5)
    f.set_paused(False)
    assert same_particles(a, b)
    f.close(), a.close(), b.close()


def test_wave_impulses_match_the_schedule():
    r = settings(interaction__wave_enabled=1, interaction__wave_strength=0.75, interaction__wave_frequency=3.0)
    a, b = pair()
    f = fsm.Frame(a, r)
    ts = fo.Timestep(float(np.float32(r.get("physics.fixed_dt"))), int(r.get("physics.max_substeps")))
    wv = fo.Interactions(fo.INTERACTION_SEED)
    fired = 0
    for ms in [16.6667] * 40:
        f.run(ms)
        dt_s = np.float32(np.float32(ms) / np.float32(1000.0))
        idt = float(min(max(dt_s * np.float32(1.0), np.float32(0.0)), fo.INTERACTION_MAX_DT))
        dv = wv.wave(idt, True, r.get("interaction.wave_strength"), r.get("interaction.wave_frequency"))
        if dv is not None:
            b.apply_impulse(*[float(x) for x in dv])
            fired += 1
        n = ts.steps_for_frame(float(dt_s), 1.0)
        if n:
            b.step(n)
    assert fired >= 3 and same_particles(a, b)
    f.close(), a.close(), b.close()


def test_auto_roll_turns_gravity_into_the_box_frame():
    r = settings(interaction__auto_roll_enabled=1, interaction__auto_roll_strength=0.5,
                 interaction__auto_roll_cadence=0.6)
    a, _ = pair()
    f = fsm.Frame(a, r)
    o = fo.Interactions(fo.INTERACTION_SEED)
    qo = fo.IDENTITY.copy()
    g = r.get("physics.gravity")
    moved = 0
    for _ in range(60):
        f.run(16.6667)
        nxt = o.auto_roll(float(np.float32(np.float32(16.6667) / np.float32(1000.0))), qo, True, 0.5, 0.6)
        if abs(float(np.dot(qo, nxt))) < 0.999999:
            qo = nxt
        q = f.orientation()
        assert np.max(np.abs(q.astype(np.float64) - qo)) < 1e-5
        grav = np.array(list(a.current_params().gravity), np.float64)
        expect = fo.q_rotate(q.astype(np.float64) * np.array([-1, -1, -1, 1]), [0.0, g, 0.0])
        assert np.max(np.abs(grav - expect)) < 1e-5
        moved += int(abs(float(q[3])) < 0.999999)
    assert moved > 40
    f.close(), a.close()


def test_reset_restores_the_reference_state():
    r = settings(interaction__auto_roll_enabled=1)
    a, b = pair()
    f = fsm.Frame(a, r)
    for _ in range(10):
        f.run(16.6667)
    assert f.tick() > 0 and abs(float(f.orientation()[3])) < 1.0
    f.reset()
    assert f.tick() == 0 and np.array_equal(f.orientation(), np.array([0, 0, 0, 1], np.float32))
    assert list(a.current_params().gravity) == [0.0, np.float32(r.get("physics.gravity")), 0.0]
    assert same_particles(a, b)  # the initial particles again
    # The interactions are reseeded with interaction_seed_for_reset(1): the
    # poses after the reset are that generator's (the box takes a pose once it
    # differs from the current one, |dot| < 0.999999, src/lib.rs:922-927).
    o = fo.Interactions(fo.seed_for_reset(1))
    qo = fo.IDENTITY.copy()
    for _ in range(8):
        f.run(16.6667)
        nxt = o.auto_roll(float(np.float32(np.float32(16.6667) / np.float32(1000.0))), qo, True,
                          r.get("interaction.auto_roll_strength"), r.get("interaction.auto_roll_cadence"))
        if abs(float(np.dot(qo, nxt))) < 0.999999:
            qo = nxt
        assert np.max(np.abs(f.orientation().astype(np.float64) - qo)) < 1e-5
    assert abs(float(qo[3])) < 0.999999  # the box has moved
    f.close(), a.close(), b.close()
