"""Restatement of fluid-lab's frame loop logic: the fixed-timestep
accumulator (src/timestep.rs:63-247) and the interaction schedules
(src/lib.rs:55-264, 914-937, 992-994).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
product's fs_timestep_* / fs_interactions_* / fs_frame_* entry points; the
product library never imports it.

The accumulator and the wave schedule are restated in numpy float32 scalars
in the reference's operation order (so they compare bit for bit); the
interaction generator in exact Python integers; the auto-roll rotation in
float64 quaternion arithmetic (an independent-precision cross-check of the
product's f32 one).  Pins (tests/test_oracle_frame.py): the reference's own
timestep and interaction tests (src/timestep.rs:268-438,
src/lib.rs:1008-1058) and quaternion identities.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32
MAX_RENDER_DT = F(1.0) / F(30.0)
INTERACTION_MAX_DT = F(1.0) / F(30.0)
INTERACTION_SEED = 0x464C5549445F5631
M64 = (1 << 64) - 1
F32_EPS = float(np.finfo(np.float32).eps)


def _fnn(v):
    """finite_nonnegative (src/timestep.rs:249-255)."""
    v = F(v)
    return max(v, F(0.0)) if np.isfinite(v) else F(0.0)


def _rtf(sim, wall):
    return sim / wall if wall > F(0.0) else F(0.0)


class Timestep:
    """TimestepController."""

    def __init__(self, fixed_dt, max_substeps):
        self.acc = F(0.0)
        self.fixed_dt = F(fixed_dt)
        self.max_substeps = int(max_substeps)
        self.dropped = F(0.0)
        self.last = self._idle(F(0.0), F(0.0))
        self.last["max_substeps"] = 0
        self.last["fixed_dt"] = F(0.0)
        self.last["policy"] = 0

    def _idle(self, wall, scale):
        return dict(substeps=0, natural_substeps=0, max_substeps=self.max_substeps, substep_cap_hit=0,
                    fixed_dt=self.fixed_dt, accumulated_before=F(0.0), accumulated_after=F(0.0),
                    dropped_this_frame=F(0.0), sim_advanced=F(0.0), wall_dt=wall, sim_time_scale=scale,
                    real_time_factor=F(0.0), policy=0)

    def steps_for_frame(self, render_dt, scale=1.0):
        wall, ts = _fnn(render_dt), _fnn(scale)
        self.acc = F(self.acc + min(F(wall * ts), MAX_RENDER_DT))
        before = self.acc
        natural = 0
        if self.fixed_dt > F(0.0):
            q = np.floor(F(self.acc / self.fixed_dt))
            natural = 0 if not q > 0 else int(min(q, 4294967295.0))
        n = min(natural, self.max_substeps)
        self.acc = F(self.acc - F(F(n) * self.fixed_dt))
        dropped = F(0.0)
        if natural > self.max_substeps:
            dropped = self.acc
            self.dropped = F(self.dropped + dropped)
            self.acc = F(0.0)
        adv = F(F(n) * self.fixed_dt)
        self.last = dict(substeps=n, natural_substeps=natural, max_substeps=self.max_substeps,
                         substep_cap_hit=int(natural > self.max_substeps), fixed_dt=self.fixed_dt,
                         accumulated_before=before, accumulated_after=self.acc, dropped_this_frame=dropped,
                         sim_advanced=adv, wall_dt=wall, sim_time_scale=ts, real_time_factor=_rtf(adv, wall),
                         policy=0)
        return n

    def record_manual(self, render_dt, substeps, scale):
        s = self._idle(_fnn(render_dt), _fnn(scale))
        s["substeps"] = substeps
        s["sim_advanced"] = F(F(substeps) * self.fixed_dt)
        s["real_time_factor"] = _rtf(s["sim_advanced"], s["wall_dt"])
        self.last = s

    def record_export(self, sim_dt, substeps):
        s = self._idle(_fnn(sim_dt), F(1.0))
        s.update(substeps=substeps, max_substeps=substeps, natural_substeps=substeps, policy=1)
        s["sim_advanced"] = F(F(substeps) * self.fixed_dt)
        s["real_time_factor"] = _rtf(s["sim_advanced"], s["wall_dt"])
        self.last = s

    def record_idle(self, render_dt, scale):
        self.last = self._idle(_fnn(render_dt), _fnn(scale))

    def reset(self):
        self.acc = F(0.0)
        self.last = self._idle(F(0.0), F(1.0))


# ---- interactions ----

class Rng:
    """InteractionRng (src/lib.rs:55-86) in exact integers."""

    def __init__(self, seed):
        self.state = seed if seed != 0 else 0x9E3779B97F4A7C15

    def next_u32(self):
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & M64
        return self.state >> 32

    def next_f32(self):
        return F(F(self.next_u32()) * F(2.0 ** -32))

    def signed_f32(self):
        return F(F(self.next_f32() * F(2.0)) - F(1.0))


def q_normalize(q):
    q = np.asarray(q, np.float64)
    return q / math.sqrt(float(q @ q))


def q_from_axis_angle(axis, angle):
    s, c = math.sin(angle / 2.0), math.cos(angle / 2.0)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, c])


def q_to_axis_angle(q):
    w = float(q[3])
    angle = 2.0 * math.acos(min(max(w, -1.0), 1.0))
    ss = max(1.0 - w * w, 0.0)
    axis = np.asarray(q[:3], np.float64) / math.sqrt(ss) if ss >= 1e-16 else np.array([1.0, 0.0, 0.0])
    return axis, angle


def q_slerp(a, b, s):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = float(a @ b)
    if d < 0:
        b, d = -b, -d
    if d > 1.0 - F32_EPS:
        return q_normalize(a + (b - a) * s)
    th = math.acos(min(d, 1.0))
    return (a * math.sin(th * (1 - s)) + b * math.sin(th * s)) / math.sin(th)


def q_rotate(q, v):
    """q v q^-1 for a unit quaternion (x, y, z, w)."""
    x, y, z, w = (float(c) for c in q)
    b = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return v * (w * w - b @ b) + b * (2.0 * (v @ b)) + np.cross(b, v) * (2.0 * w)


def clamp_rotation(q, max_angle):
    """clamp_rotation (src/lib.rs:246-255)."""
    max_angle = max(max_angle, 0.0)
    axis, angle = q_to_axis_angle(q)
    if angle > max_angle and float(axis @ axis) > 0.0:
        return q_normalize(q_from_axis_angle(axis, max_angle))
    return q_normalize(q)


IDENTITY = np.array([0.0, 0.0, 0.0, 1.0])
WAVE_DIRS = [(F(1.0), F(0.0)), (F(-1.0), F(-0.0)), (F(0.0), F(1.0)), (F(-0.0), F(-1.0))]  # (x, z)


class Interactions:
    """InteractionState (src/lib.rs:88-236)."""

    def __init__(self, seed):
        self.reset(seed)

    def reset(self, seed):
        self.rng = Rng(seed)
        self.roll_enabled_last, self.roll_elapsed, self.roll_duration = False, 0.0, 0.0
        self.roll_from, self.roll_target = IDENTITY.copy(), IDENTITY.copy()
        self.wave_enabled_last, self.wave_time, self.wave_phase = False, F(0.0), 0

    def _schedule(self, frm, strength, cadence):
        cadence = max(float(F(cadence)), 0.1)
        jitter = float(F(F(0.75) + F(self.rng.next_f32() * F(0.5))))
        hx = float(self.rng.signed_f32())
        hz = float(self.rng.signed_f32())
        h = np.array([hx, 0.0, hz])
        axis = h / math.sqrt(float(h @ h)) if float(h @ h) > 1e-6 else np.array([1.0, 0.0, 0.0])
        angle = strength * float(F(F(0.35) + F(self.rng.next_f32() * F(0.65))))
        self.roll_from = clamp_rotation(q_normalize(frm), strength)
        self.roll_target = q_normalize(q_from_axis_angle(axis, angle))
        self.roll_duration = cadence * jitter
        self.roll_elapsed = 0.0

    def auto_roll(self, dt, current, enabled, strength, cadence):
        strength = max(float(F(strength)), 0.0)
        if not enabled or strength <= F32_EPS:
            self.roll_enabled_last = False
            return None
        dt = float(_fnn(dt))
        if not self.roll_enabled_last or self.roll_duration <= F32_EPS:
            self._schedule(current, strength, cadence)
            self.roll_enabled_last = True
        self.roll_elapsed += dt
        while self.roll_duration > F32_EPS and self.roll_elapsed >= self.roll_duration:
            overflow = self.roll_elapsed - self.roll_duration
            self._schedule(self.roll_target, strength, cadence)
            self.roll_enabled_last = True
            self.roll_elapsed = min(overflow, self.roll_duration)
        t = min(max(self.roll_elapsed / self.roll_duration, 0.0), 1.0)
        eased = t * t * (3.0 - 2.0 * t)
        return clamp_rotation(q_normalize(q_slerp(self.roll_from, self.roll_target, eased)), strength)

    def wave(self, dt, enabled, strength, frequency):
        strength, frequency = max(F(strength), F(0.0)), max(F(frequency), F(0.0))
        if not enabled or strength <= F32_EPS or frequency <= F32_EPS:
            self.wave_enabled_last = False
            return None
        if not self.wave_enabled_last:
            self.wave_enabled_last, self.wave_time = True, F(0.0)
        self.wave_time = F(self.wave_time - _fnn(dt))
        if self.wave_time > F(0.0):
            return None
        dx, dz = WAVE_DIRS[self.wave_phase % 4]
        self.wave_phase = (self.wave_phase + 1) & 0xFFFFFFFF
        self.wave_time = F(self.wave_time + F(F(1.0) / frequency))
        if self.wave_time < F(0.0):
            self.wave_time = F(F(1.0) / frequency)
        return np.array([F(dx * strength), F(0.0), F(dz * strength)], np.float32)


def seed_for_reset(reset_count):
    """interaction_seed_for_reset (src/lib.rs:992-994)."""
    return INTERACTION_SEED ^ ((reset_count * 0x9E3779B97F4A7C15) & M64)
