"""The frame loop's host logic: the fixed-timestep accumulator
(fs_timestep_*, src/timestep.rs) and the interaction schedules
(fs_interactions_*, src/lib.rs:55-264), against the reference's own tests
(src/timestep.rs:268-438, src/lib.rs:1008-1058), the restatement in
oracle/frame_oracle.py (bit for bit for the accumulator, the generator and
the wave; to 1e-5 for the auto-roll rotation, whose quaternion arithmetic is
f32 in the product and f64 in the restatement), and quaternion identities.
No GPU."""
import math

import numpy as np
import pytest

import fluidstep as fsm
import frame_oracle as fo

F = np.float32
MAX_RENDER_DT = float(F(1.0) / F(30.0))


class OracleTS:
    name = "oracle"

    def __init__(self, dt, cap):
        self.t = fo.Timestep(dt, cap)

    def steps(self, dt, scale=1.0):
        return self.t.steps_for_frame(dt, scale)

    def stats(self):
        return {k: (float(v) if isinstance(v, np.floating) else v) for k, v in self.t.last.items()}

    def dropped(self):
        return float(self.t.dropped)

    def acc(self):
        return float(self.t.acc)


class ProductTS:
    name = "product"

    def __init__(self, dt, cap):
        self.t = fsm.Timestep(dt, cap)

    def steps(self, dt, scale=1.0):
        return self.t.steps_for_frame(dt, scale)

    def stats(self):
        return self.t.last_stats()

    def dropped(self):
        return self.t.total_dropped()

    def acc(self):
        return self.t.last_stats()["accumulated_after"]


TS = [OracleTS, ProductTS]


def close(a, b):
    assert abs(a - b) < 1e-6, (a, b)


# ---- the reference's timestep tests (src/timestep.rs:268-438) on both ----

@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_one_hundred_twentieth_frame_with_cap_one_runs_realtime(mk):
    t = mk(1.0 / 120.0, 1)
    assert t.steps(1.0 / 120.0) == 1
    s = t.stats()
    assert s["substeps"] == 1 and s["natural_substeps"] == 1 and s["max_substeps"] == 1 and not s["substep_cap_hit"]
    close(s["fixed_dt"], 1 / 120)
    close(s["wall_dt"], 1 / 120)
    close(s["sim_advanced"], 1 / 120)
    close(s["real_time_factor"], 1.0)
    assert s["policy"] == 0  # TIMESTEP_POLICY_LABEL "fixed-drop-cap"
    assert t.dropped() == 0.0


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_sixty_hz_frame_with_cap_one_reports_half_realtime(mk):
    t = mk(1.0 / 120.0, 1)
    assert t.steps(1.0 / 60.0) == 1
    s = t.stats()
    assert s["natural_substeps"] == 2 and s["substep_cap_hit"]
    close(s["wall_dt"], 1 / 60)
    close(s["sim_advanced"], 1 / 120)
    close(s["real_time_factor"], 0.5)
    close(s["dropped_this_frame"], 1 / 120)
    close(t.dropped(), 1 / 120)
    assert abs(s["accumulated_after"]) < 1e-6


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_large_hitch_uses_raw_wall_time_for_realtime_factor(mk):
    t = mk(1.0 / 120.0, 1)
    assert t.steps(1.0) == 1
    s = t.stats()
    assert s["natural_substeps"] == 4 and s["substep_cap_hit"]
    close(s["accumulated_before"], MAX_RENDER_DT)
    close(s["wall_dt"], 1.0)
    close(s["real_time_factor"], 1 / 120)
    close(s["dropped_this_frame"], 3 / 120)
    close(t.dropped(), 3 / 120)


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_scaled_time_advances_fractional_wall_time(mk):
    t = mk(1.0 / 120.0, 4)
    assert t.steps(1.0 / 60.0, 0.25) == 0
    s = t.stats()
    close(s["sim_time_scale"], 0.25)
    close(s["accumulated_before"], 1 / 240)
    close(s["real_time_factor"], 0.0)
    assert t.steps(1.0 / 60.0, 0.25) == 1
    s = t.stats()
    close(s["sim_advanced"], 1 / 120)
    close(s["real_time_factor"], 0.5)
    assert abs(s["accumulated_after"]) < 1e-6


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_two_steps_per_half_period_frame(mk):
    t = mk(1.0 / 120.0, 4)
    assert t.steps(1.0 / 60.0) == 2
    s = t.stats()
    assert abs(t.acc()) < 1e-6 and t.dropped() == 0.0 and s["dropped_this_frame"] == 0.0
    assert s["natural_substeps"] == 2 and not s["substep_cap_hit"]


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_large_frame_clamped_and_capped(mk):
    t = mk(1.0 / 120.0, 4)
    assert t.steps(1.0) == 4 and t.dropped() == 0.0 and t.stats()["dropped_this_frame"] == 0.0
    t2 = mk(1.0 / 120.0, 2)
    assert t2.steps(1.0) == 2
    assert t2.dropped() > 0.0 and abs(t2.acc()) < 1e-6 and t2.stats()["dropped_this_frame"] > 0.0


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_tiny_frames_accumulate(mk):
    t = mk(1.0 / 120.0, 4)
    assert sum(t.steps(0.002) for _ in range(5)) == 1 and t.dropped() == 0.0


@pytest.mark.parametrize("mk", TS, ids=lambda m: m.name)
def test_large_frame_drops_excess_and_records_stats(mk):
    t = mk(1.0 / 120.0, 1)
    assert t.steps(1.0) == 1
    s = t.stats()
    assert s["substeps"] == 1 and s["dropped_this_frame"] > 0 and abs(s["accumulated_after"]) < 1e-6
    assert t.dropped() > 0


# ---- product accumulator == restatement, bit for bit ----

def test_accumulator_matches_restatement_bit_for_bit():
    rng = np.random.default_rng(5)
    specials = [float("nan"), float("inf"), -float("inf"), -0.004, 0.0, 1.0, 1e-9, 1 / 30, 1 / 60, 1 / 144]
    for fixed_dt, cap in ((1 / 120, 2), (1 / 240, 16), (1 / 30, 1), (1 / 480, 3), (0.0, 4)):
        o, p = fo.Timestep(fixed_dt, cap), fsm.Timestep(fixed_dt, cap)
        for k in range(400):
            dt = specials[k % len(specials)] if k % 7 == 0 else float(rng.uniform(0, 0.05))
            scale = [1.0, 0.25, 0.05, float("nan"), 2.0][k % 5]
            op = k % 23
            if op == 5:
                o.record_manual(dt, 1, scale), p.record_manual(dt, 1, scale)
            elif op == 11:
                o.record_export(dt, 3), p.record_export(dt, 3)
            elif op == 17:
                o.record_idle(dt, scale), p.record_idle(dt, scale)
            elif op == 19:
                o.reset(), p.reset()
            else:
                assert o.steps_for_frame(dt, scale) == p.steps_for_frame(dt, scale)
            ps, os_ = p.last_stats(), o.last
            for key, v in os_.items():
                a, b = ps[key], v
                if isinstance(b, np.floating):
                    assert F(a).view(np.uint32) == F(b).view(np.uint32), (k, key, a, b)
                else:
                    assert a == int(b), (k, key, a, b)
            assert F(p.total_dropped()).view(np.uint32) == F(o.dropped).view(np.uint32)


# ---- interactions ----

def test_interaction_rng_is_repeatable_and_the_reference_generator():  # src/lib.rs:1008-1016
    for seed in (123, 0, fo.INTERACTION_SEED, fo.seed_for_reset(3)):
        a, b, o = fsm.Interactions(seed), fsm.Interactions(seed), fo.Rng(seed)
        for _ in range(16):
            x = a.next_u32()
            assert x == b.next_u32() == o.next_u32()


def test_seed_for_reset():  # src/lib.rs:992-994
    assert fo.seed_for_reset(0) == fo.INTERACTION_SEED
    assert fo.seed_for_reset(1) == fo.INTERACTION_SEED ^ 0x9E3779B97F4A7C15


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_wave_scheduler_fires_immediate_then_by_frequency(impl):  # src/lib.rs:1033-1050
    s = fsm.Interactions(7) if impl == "product" else fo.Interactions(7)

    def wave(*a):
        r = s.wave(*a)
        return None if r is None else [float(v) for v in r]

    first = wave(0.0, True, 0.5, 2.0)
    assert first == [0.5, 0.0, 0.0]
    assert wave(0.1, True, 0.5, 2.0) is None
    third = wave(0.4, True, 0.5, 2.0)
    assert third == [-0.5, 0.0, -0.0] and math.copysign(1.0, third[2]) == -1.0
    s.reset(7)
    assert wave(0.0, False, 0.5, 2.0) is None


def test_wave_matches_restatement_bit_for_bit():
    p, o = fsm.Interactions(11), fo.Interactions(11)
    rng = np.random.default_rng(2)
    for k in range(300):
        dt = float(rng.uniform(0, 0.2))
        en = k % 37 != 0
        st, fr = float(F(rng.uniform(0, 3))), float(F(rng.uniform(0.05, 4)))
        a, b = p.wave(dt, en, st, fr), o.wave(dt, en, st, fr)
        assert (a is None) == (b is None), k
        if a is not None:
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k


def angle_between_identity(q):
    return 2.0 * math.acos(min(abs(float(q[3])), 1.0))


def test_auto_roll_stays_within_strength_bound():  # src/lib.rs:1018-1031
    s = fsm.Interactions(42)
    q = [0.0, 0.0, 0.0, 1.0]
    for _ in range(64):
        q = s.auto_roll(0.1, q, True, 0.4, 1.0)
        assert q is not None and angle_between_identity(q) <= 0.4 + 1e-5
        assert abs(float(np.dot(q.astype(np.float64), q)) - 1.0) < 1e-6


def test_auto_roll_matches_restatement():
    """Same draws (the generator is exact), same schedule; the rotation agrees
    to 1e-5 with the f64 restatement over several cadences, strengths and dt."""
    for seed, strength, cadence, dt in ((42, 0.4, 1.0, 0.1), (9, 0.222, 0.8375, 1 / 60), (3, 1.2, 0.5, 0.033),
                                        (5, 0.05, 8.0, 0.25)):
        p, o = fsm.Interactions(seed), fo.Interactions(seed)
        qp = [0.0, 0.0, 0.0, 1.0]
        qo = fo.IDENTITY.copy()
        for k in range(200):
            en = k % 53 != 52
            a = p.auto_roll(dt, qp, en, strength, cadence)
            b = o.auto_roll(dt, qo, en, strength, cadence)
            assert (a is None) == (b is None)
            if a is None:
                continue
            assert np.max(np.abs(a.astype(np.float64) - b)) < 1e-5, (seed, k, a, b)
            assert angle_between_identity(a) <= strength + 1e-5
            qp, qo = a, b


def test_auto_roll_eases_from_start_to_target():
    """A schedule starts at its `from` pose (eased 0) and reaches the target
    rotation about a horizontal axis by strength * [0.35, 1.0] radians."""
    o = fo.Interactions(17)
    q0 = o.auto_roll(0.0, fo.IDENTITY, True, 0.6, 2.0)
    assert np.allclose(q0, fo.IDENTITY, atol=1e-12)
    axis, angle = fo.q_to_axis_angle(o.roll_target)
    assert abs(axis[1]) < 1e-12 and 0.35 * 0.6 - 1e-9 <= angle <= 0.6 + 1e-9
    assert 0.75 * 2.0 <= o.roll_duration <= 1.25 * 2.0
    end = o.auto_roll(o.roll_duration * 0.999999, q0, True, 0.6, 2.0)
    assert np.allclose(end, o.roll_target, atol=1e-5)


def test_gravity_in_the_box_frame_is_the_inverse_rotation():
    """push_gravity (src/lib.rs:907-912): g_local = q^-1 (0, g, 0); rotating
    it back by q gives the world gravity, and its length is |g|."""
    q = fo.q_normalize(fo.q_from_axis_angle(np.array([0.6, 0.0, 0.8]), 0.3))
    qi = q * np.array([-1, -1, -1, 1])
    g = fo.q_rotate(qi, [0.0, -9.81, 0.0])
    assert abs(np.linalg.norm(g) - 9.81) < 1e-9
    assert np.allclose(fo.q_rotate(q, g), [0.0, -9.81, 0.0], atol=1e-9)
