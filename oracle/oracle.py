"""ctypes front end of the CPU restatement in oracle/ (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library in
fluid-simulation_amd/ never imports this module.

Each wrapper names the C function it calls; the C sources cite the reference
file:line they restate (app/crates/fluid-lab/src/...).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = C.POINTER(C.c_uint32)


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("pressure_iters", C.c_uint32), ("warm_start", C.c_uint32), ("reduce_mode", C.c_uint32),
        ("h", C.c_float), ("inv_h", C.c_float), ("dt", C.c_float), ("fixed_scale", C.c_float),
        ("rho", C.c_float), ("flip_blend", C.c_float), ("wall_friction", C.c_float),
        ("origin", C.c_float * 3), ("grav", C.c_float * 3),
        ("rest", C.c_float), ("stiff", C.c_float), ("drift_clamp", C.c_float),
        ("liquid_threshold", C.c_float), ("surface_dilation", C.c_float), ("cfl", C.c_float),
        ("tol_sq", C.c_float), ("tol", C.c_float),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("liquid_cells", C.c_uint32), ("cg_iterations", C.c_uint32),
        ("misc", C.c_uint32), ("pad", C.c_uint32),
        ("rs_initial", C.c_double), ("rs_final", C.c_double), ("true_residual_sq", C.c_double),
    ]


def _outputs_type(real, acc):
    rp = C.POINTER(real)
    ap = C.POINTER(acc)

    class Out(C.Structure):
        _fields_ = [
            ("occupancy", _u32p), ("cell_offset", _u32p), ("perm", _u32p), ("cell_type", _u32p),
            ("num", ap * 3), ("den", ap * 3),
            ("vel", rp * 3), ("saved", rp * 3), ("vel_final", rp * 3),
            ("divergence", rp), ("stats", C.POINTER(OrcStats)),
        ]
    return Out


OutF32 = _outputs_type(C.c_float, C.c_int32)
OutF64 = _outputs_type(C.c_double, C.c_int64)

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_cell_key_f32.restype = C.c_uint32
        _lib.orc_cell_key_f32.argtypes = [C.POINTER(OrcParams), C.c_float, C.c_float, C.c_float]
        _lib.orc_classify_u32.restype = C.c_uint32
        _lib.orc_pressure_solve_f32.restype = C.c_uint32
        _lib.orc_pressure_solve_f64.restype = C.c_uint32
        _lib.orc_run_f32.restype = C.c_int
        _lib.orc_run_f64.restype = C.c_int
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_sample_axis_f32.argtypes = [
            C.POINTER(OrcParams), C.c_int, C.c_float, C.c_float, C.c_float,
            C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _typed(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(nx, ny, nz, h, origin=(0.0, 0.0, 0.0), dt=1.0 / 120.0, rho=1000.0,
                flip_blend=0.95, wall_friction=0.0, grav=(0.0, -9.81, 0.0), rest=8.0,
                stiff=0.45, drift_clamp=0.5, liquid_threshold=1, surface_dilation=0, cfl=2.0,
                pressure_iters=200, warm_start=1, tol=1e-5, reduce_mode=0,
                fixed_scale=65536.0, clamp_tol=True) -> OrcParams:
    """Params uniform as src/gpu/fluid.rs:289-318 fills it (inv_h = f32(1/h))."""
    P = OrcParams()
    P.nx, P.ny, P.nz = nx, ny, nz
    P.pressure_iters = max(int(pressure_iters), 1)
    P.warm_start = int(bool(warm_start))
    P.reduce_mode = int(reduce_mode)
    h32 = np.float32(h)
    P.h = float(h32)
    P.inv_h = float(np.float32(1.0) / h32)
    P.dt = float(np.float32(dt))
    P.fixed_scale = fixed_scale
    P.rho = rho
    P.flip_blend = flip_blend
    P.wall_friction = wall_friction
    for a in range(3):
        P.origin[a] = float(np.float32(origin[a]))
        P.grav[a] = float(np.float32(grav[a]))
    P.rest, P.stiff, P.drift_clamp = rest, stiff, drift_clamp
    P.liquid_threshold = float(liquid_threshold)
    P.surface_dilation = float(surface_dilation)
    P.cfl = cfl
    t = np.float32(tol)
    if clamp_tol:  # src/gpu/fluid.rs:47-50
        t = np.float32(min(max(float(t), 0.0), 0.1))
    P.tol = float(t)
    P.tol_sq = float(t * t)
    return P


def face_count(P, a):
    d = [P.nx, P.ny, P.nz]
    d[a] += 1
    return d[0] * d[1] * d[2]


def cell_key(P, x, y, z) -> int:
    return int(lib().orc_cell_key_f32(C.byref(P), x, y, z))


def exclusive_scan(hist) -> np.ndarray:
    hist = np.ascontiguousarray(hist, dtype=np.uint32)
    out = np.zeros_like(hist)
    lib().orc_exclusive_scan_u32(C.c_uint32(hist.size), ptr(hist), ptr(out))
    return out


def stable_permutation(keys, cells) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.zeros(keys.size, dtype=np.uint32)
    lib().orc_stable_permutation(C.c_uint32(keys.size), C.c_uint32(cells), ptr(keys), ptr(perm))
    return perm


def histogram(P, pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    px, py, pz = (np.ascontiguousarray(pos[:, a]) for a in range(3))
    occ = np.zeros(P.nx * P.ny * P.nz, dtype=np.uint32)
    lib().orc_histogram_f32(C.byref(P), C.c_uint32(len(px)), ptr(px), ptr(py), ptr(pz), ptr(occ))
    return occ


def classify(P, occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint32)
    ct = np.zeros_like(occ)
    n = lib().orc_classify_u32(C.byref(P), ptr(occ), ptr(ct))
    return ct, int(n)


def p2g(P, pos, vel):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    cols = [np.ascontiguousarray(pos[:, a]) for a in range(3)] + \
           [np.ascontiguousarray(vel[:, a]) for a in range(3)]
    num = [np.zeros(face_count(P, a), np.int32) for a in range(3)]
    den = [np.zeros(face_count(P, a), np.int32) for a in range(3)]
    nump = (C.c_void_p * 3)(*[ptr(x) for x in num])
    denp = (C.c_void_p * 3)(*[ptr(x) for x in den])
    lib().orc_p2g_f32(C.byref(P), C.c_uint32(len(pos)), *[ptr(c) for c in cols], nump, denp)
    return num, den


def p2g_epilogue(P, ct, num, den, f64=False):
    """orc_p2g_epilogue_*: (vel, saved) per axis from P2G's num/den and the
    cell types (normalize, save, forces, pre-pressure enforce)."""
    real = np.float64 if f64 else np.float32
    acc = np.int64 if f64 else np.int32
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    num = [np.array(x, dtype=acc, copy=True) for x in num]
    den = [np.array(x, dtype=acc, copy=True) for x in den]
    vel = [np.zeros(face_count(P, a), real) for a in range(3)]
    saved = [np.zeros(face_count(P, a), real) for a in range(3)]
    arr = lambda xs: (C.c_void_p * 3)(*[ptr(x) for x in xs])
    fn = lib().orc_p2g_epilogue_f64 if f64 else lib().orc_p2g_epilogue_f32
    fn(C.byref(P), ptr(ct), arr(num), arr(den), arr(vel), arr(saved))
    return vel, saved


def divergence(P, vel, ct, occ=None):
    vel = [np.ascontiguousarray(v, dtype=np.float32) for v in vel]
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    if occ is None:
        occ = np.zeros(ct.size, np.uint32)
    occ = np.ascontiguousarray(occ, dtype=np.uint32)
    div = np.zeros(ct.size, np.float32)
    vp = (C.c_void_p * 3)(*[ptr(v) for v in vel])
    lib().orc_divergence_f32(C.byref(P), vp, ptr(ct), ptr(occ), ptr(div))
    return div


def apply_poisson(P, ct, x, f64=False):
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    dt = np.float64 if f64 else np.float32
    x = np.ascontiguousarray(x, dtype=dt)
    out = np.zeros_like(x)
    fn = lib().orc_apply_poisson_f64 if f64 else lib().orc_apply_poisson_f32
    fn(C.byref(P), ptr(ct), ptr(x), ptr(out))
    return out


def pressure_solve(P, ct, div, p0=None, f64=False):
    """Returns (pressure, OrcStats).  p0 is the warm-start state (pressure_a)."""
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    dt = np.float64 if f64 else np.float32
    div = np.ascontiguousarray(div, dtype=dt)
    p = np.zeros(ct.size, dt) if p0 is None else np.array(p0, dtype=dt, copy=True)
    st = OrcStats()
    fn = lib().orc_pressure_solve_f64 if f64 else lib().orc_pressure_solve_f32
    fn(C.byref(P), ptr(ct), ptr(div), ptr(p), C.byref(st))
    return p, st


def jacobi_solve(P, ct, div, iters):
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    div = np.ascontiguousarray(div, dtype=np.float32)
    p = np.zeros(ct.size, np.float32)
    lib().orc_jacobi_solve_f32(C.byref(P), ptr(ct), ptr(div), C.c_uint32(iters), ptr(p))
    return p


def gradient_enforce(P, p, ct, vel):
    vel = [np.array(v, dtype=np.float32, copy=True) for v in vel]
    p = np.ascontiguousarray(p, dtype=np.float32)
    ct = np.ascontiguousarray(ct, dtype=np.uint32)
    vp = (C.c_void_p * 3)(*[ptr(v) for v in vel])
    lib().orc_gradient_enforce_f32(C.byref(P), ptr(p), ptr(ct), vp)
    return vel


def sample_axis(P, axis, point, vel, saved, wall_aware=True):
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    saved = np.ascontiguousarray(saved, dtype=np.float32)
    f = C.c_float()
    s = C.c_float()
    lib().orc_sample_axis_f32(C.byref(P), axis, point[0], point[1], point[2], ptr(vel), ptr(saved),
                              int(wall_aware), C.byref(f), C.byref(s))
    return f.value, s.value


def g2p(P, pos, pvel, vel, saved):
    pos = np.array(pos, dtype=np.float32, copy=True)
    pvel = np.array(pvel, dtype=np.float32, copy=True)
    cols = [np.ascontiguousarray(pos[:, a]) for a in range(3)] + \
           [np.ascontiguousarray(pvel[:, a]) for a in range(3)]
    vel = [np.ascontiguousarray(v, dtype=np.float32) for v in vel]
    saved = [np.ascontiguousarray(v, dtype=np.float32) for v in saved]
    vp = (C.c_void_p * 3)(*[ptr(v) for v in vel])
    sp = (C.c_void_p * 3)(*[ptr(v) for v in saved])
    lib().orc_g2p_f32(C.byref(P), C.c_uint32(len(pos)), *[ptr(c) for c in cols], vp, sp)
    return np.stack(cols[:3], 1), np.stack(cols[3:], 1)


class StepResult:
    pass


def run(P, pos, vel, pressure=None, steps=1, sort_period=1, tick0=0, f64=False, want=True):
    """Run `steps` substeps of the restated reference step from (pos, vel,
    pressure); returns a StepResult with the final particles and the last
    substep's grid outputs."""
    real = np.float64 if f64 else np.float32
    acc = np.int64 if f64 else np.int32
    n = len(pos)
    Cc = P.nx * P.ny * P.nz
    cols = [np.array(pos[:, a], dtype=real, copy=True) for a in range(3)] + \
           [np.array(vel[:, a], dtype=real, copy=True) for a in range(3)]
    pr = np.zeros(Cc, real) if pressure is None else np.array(pressure, dtype=real, copy=True)
    res = StepResult()
    out = (OutF64 if f64 else OutF32)()
    st = OrcStats()
    if want:
        res.occupancy = np.zeros(Cc, np.uint32)
        res.cell_offset = np.zeros(Cc, np.uint32)
        res.perm = np.zeros(max(n, 1), np.uint32)
        res.cell_type = np.zeros(Cc, np.uint32)
        res.divergence = np.zeros(Cc, real)
        res.num = [np.zeros(face_count(P, a), acc) for a in range(3)]
        res.den = [np.zeros(face_count(P, a), acc) for a in range(3)]
        res.vel = [np.zeros(face_count(P, a), real) for a in range(3)]
        res.saved = [np.zeros(face_count(P, a), real) for a in range(3)]
        res.vel_final = [np.zeros(face_count(P, a), real) for a in range(3)]
        ct_real = C.c_double if f64 else C.c_float
        ct_acc = C.c_int64 if f64 else C.c_int32
        out.occupancy = _typed(res.occupancy, C.c_uint32)
        out.cell_offset = _typed(res.cell_offset, C.c_uint32)
        out.perm = _typed(res.perm, C.c_uint32)
        out.cell_type = _typed(res.cell_type, C.c_uint32)
        out.divergence = _typed(res.divergence, ct_real)
        for a in range(3):
            out.num[a] = _typed(res.num[a], ct_acc)
            out.den[a] = _typed(res.den[a], ct_acc)
            out.vel[a] = _typed(res.vel[a], ct_real)
            out.saved[a] = _typed(res.saved[a], ct_real)
            out.vel_final[a] = _typed(res.vel_final[a], ct_real)
    out.stats = C.pointer(st)
    fn = lib().orc_run_f64 if f64 else lib().orc_run_f32
    res.sorted = bool(fn(C.byref(P), C.c_uint32(n), *[ptr(c) for c in cols], ptr(pr),
                         C.c_uint32(steps), C.c_uint32(sort_period), C.c_uint32(tick0),
                         C.byref(out)))
    res.pos = np.stack(cols[:3], 1)
    res.vel_particles = np.stack(cols[3:], 1)
    res.pressure = pr
    res.stats = st
    if want:
        res.perm = res.perm[:n]
    return res


PHASES = ("mark", "sort", "classify", "scatter", "face_epilogue", "divergence", "cg", "gradient", "g2p")


def phase_ms(f64=False) -> dict:
    """Wall time per phase of run() since the last phase_reset (ms)."""
    out = np.zeros(len(PHASES), np.float64)
    getattr(lib(), "orc_phase_ms_f64" if f64 else "orc_phase_ms_f32")(ptr(out))
    return dict(zip(PHASES, out.tolist()))


def phase_reset(f64=False) -> None:
    getattr(lib(), "orc_phase_reset_f64" if f64 else "orc_phase_reset_f32")()


def num_threads() -> int:
    return int(lib().orc_num_threads())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(C.c_int(n))


class AdjointTape:
    """Checkpointed forward of the restated step (flip_oracle_adjoint.h) and
    its reverse pass.  After construction `pos`, `vel`, `pressure` hold the
    final state; `backward(gpos, gvel)` returns dL/d(initial pos), dL/d(initial
    vel) and dL/d(gravity) for dL/d(final pos, vel) given in the final order."""

    def __init__(self, P, pos, vel, pressure=None, steps=1, every=10, f64=True, record=False):
        self.real = np.float64 if f64 else np.float32
        self.sfx = "f64" if f64 else "f32"
        n = len(pos)
        C_ = P.nx * P.ny * P.nz
        self._cols_p = [np.array(pos[:, a], dtype=self.real, copy=True) for a in range(3)]
        self._cols_v = [np.array(vel[:, a], dtype=self.real, copy=True) for a in range(3)]
        self.pressure = np.zeros(C_, self.real) if pressure is None else np.array(pressure, self.real, copy=True)
        # Every substep's cell types and occupancies (the decisions the
        # reverse pass holds fixed), for run_frozen.
        self.ct_steps = np.zeros((steps, C_), np.uint32) if record else None
        self.occ_steps = np.zeros((steps, C_), np.uint32) if record else None
        L = lib()
        fwd = getattr(L, f"orc_adj_forward_{self.sfx}")
        fwd.restype = C.c_void_p
        vp = C.c_void_p * 3
        self._tape = fwd(C.byref(P), C.c_uint32(n), vp(*[ptr(c) for c in self._cols_p]),
                         vp(*[ptr(c) for c in self._cols_v]), ptr(self.pressure), C.c_uint32(steps),
                         C.c_uint32(every), ptr(self.ct_steps), ptr(self.occ_steps))
        self.n = n
        self.pos = np.stack(self._cols_p, 1)
        self.vel = np.stack(self._cols_v, 1)

    @classmethod
    def from_states(cls, P, states, f64=True):
        """A tape whose substeps start from the given states: a list of
        (pos [N,3], vel [N,3], pressure [C]) per substep, each in that
        substep's input order (a trajectory computed elsewhere, e.g. on the
        GPU).  `backward` then runs the reverse pass along it."""
        self = cls.__new__(cls)
        self.real = np.float64 if f64 else np.float32
        self.sfx = "f64" if f64 else "f32"
        self.n = len(states[0][0])
        flat = np.concatenate([np.concatenate([np.asarray(p, self.real).T.ravel(), np.asarray(v, self.real).T.ravel(),
                                               np.asarray(pr, self.real).ravel()]) for p, v, pr in states])
        fn = getattr(lib(), f"orc_adj_tape_from_states_{self.sfx}")
        fn.restype = C.c_void_p
        self._tape = fn(C.byref(P), C.c_uint32(self.n), C.c_uint32(len(states)), ptr(flat))
        self.pos = self.vel = self.pressure = None
        self.ct_steps = self.occ_steps = None
        return self

    def backward(self, gpos=None, gvel=None):
        n = self.n
        gp = [np.zeros(n, self.real) if gpos is None else np.array(gpos[:, a], self.real, copy=True)
              for a in range(3)]
        gv = [np.zeros(n, self.real) if gvel is None else np.array(gvel[:, a], self.real, copy=True)
              for a in range(3)]
        grav = np.zeros(3, np.float64)
        vp = C.c_void_p * 3
        getattr(lib(), f"orc_adj_backward_{self.sfx}")(C.c_void_p(self._tape), vp(*[ptr(c) for c in gp]),
                                                      vp(*[ptr(c) for c in gv]), ptr(grav))
        return np.stack(gp, 1), np.stack(gv, 1), grav

    def close(self):
        if getattr(self, "_tape", None):
            getattr(lib(), f"orc_adj_free_{self.sfx}")(C.c_void_p(self._tape))
            self._tape = None

    __del__ = close


def run_frozen(P, pos, vel, ct_steps, occ_steps, pressure=None, f64=True):
    """The forward over len(ct_steps) substeps with every substep's cell
    types and occupancies taken from ct_steps / occ_steps (as AdjointTape
    records them): the function the reverse pass differentiates.  Returns
    the final (pos, vel, pressure)."""
    real = np.float64 if f64 else np.float32
    sfx = "f64" if f64 else "f32"
    cols_p = [np.array(pos[:, a], dtype=real, copy=True) for a in range(3)]
    cols_v = [np.array(vel[:, a], dtype=real, copy=True) for a in range(3)]
    pr = np.zeros(P.nx * P.ny * P.nz, real) if pressure is None else np.array(pressure, real, copy=True)
    ct = np.ascontiguousarray(ct_steps, np.uint32)
    occ = np.ascontiguousarray(occ_steps, np.uint32)
    vp = C.c_void_p * 3
    getattr(lib(), f"orc_run_frozen_{sfx}")(C.byref(P), C.c_uint32(len(pos)), vp(*[ptr(c) for c in cols_p]),
                                            vp(*[ptr(c) for c in cols_v]), ptr(pr), C.c_uint32(len(ct)), ptr(ct),
                                            ptr(occ))
    return np.stack(cols_p, 1), np.stack(cols_v, 1), pr


class _SmoothRecC(C.Structure):
    _fields_ = [("perm", C.c_void_p), ("occ", C.c_void_p), ("ct", C.c_void_p), ("cell", C.c_void_p),
                ("bits", C.c_void_p), ("support", C.c_void_p)]


class SmoothRecord:
    """The decisions of `steps` substeps of the smooth forward
    (orc_smooth_rec): sort permutation, occupancies, cell types, stencil
    bases and the CFL / wall-clamp branches of every particle."""

    def __init__(self, P, n, steps):
        C_ = P.nx * P.ny * P.nz
        self.n, self.steps = n, steps
        self.perm = np.zeros((steps, max(n, 1)), np.uint32)
        self.occ = np.zeros((steps, C_), np.uint32)
        self.ct = np.zeros((steps, C_), np.uint32)
        self.cell = np.zeros((steps, 3 * max(n, 1)), np.int32)
        self.bits = np.zeros((steps, max(n, 1)), np.uint16)
        self.support = np.zeros((steps, sum(face_count(P, a) for a in range(3))), np.uint8)
        self._c = _SmoothRecC(*[a.ctypes.data for a in (self.perm, self.occ, self.ct, self.cell, self.bits,
                                                        self.support)])


def smooth_forward(P, pos, vel, rec, pressure=None, record=False, free_stencils=False, f64=True):
    """orc_smooth_forward_*: rec.steps substeps of the smooth forward from
    (pos, vel, pressure), taking every discrete decision from `rec` (or, with
    record=True, taking them itself and writing them to `rec`).  With
    free_stencils the replay takes each particle's stencil from its current
    position (the true tent weights) instead of the record.  Returns the
    final (pos, vel, pressure), particles in the last sort's order."""
    real = np.float64 if f64 else np.float32
    cols_p = [np.array(pos[:, a], dtype=real, copy=True) for a in range(3)]
    cols_v = [np.array(vel[:, a], dtype=real, copy=True) for a in range(3)]
    pr = np.zeros(P.nx * P.ny * P.nz, real) if pressure is None else np.array(pressure, real, copy=True)
    vp = C.c_void_p * 3
    getattr(lib(), "orc_smooth_forward_f64" if f64 else "orc_smooth_forward_f32")(
        C.byref(P), C.c_uint32(len(pos)), vp(*[ptr(c) for c in cols_p]), vp(*[ptr(c) for c in cols_v]), ptr(pr),
        C.c_uint32(rec.steps), C.byref(rec._c), C.c_int(int(record) | (2 if free_stencils else 0)))
    return np.stack(cols_p, 1), np.stack(cols_v, 1), pr


def smooth_record_states(P, states, f64=True):
    """The smooth forward's decisions along a trajectory computed elsewhere:
    substep s taken (in record mode) from states[s] = (pos, vel, pressure) in
    that substep's input order.  Returns a SmoothRecord."""
    real = np.float64 if f64 else np.float32
    n = len(states[0][0])
    rec = SmoothRecord(P, n, len(states))
    flat = np.concatenate([np.concatenate([np.asarray(p, real).T.ravel(), np.asarray(v, real).T.ravel(),
                                           np.asarray(pr, real).ravel()]) for p, v, pr in states])
    getattr(lib(), "orc_smooth_record_states_f64" if f64 else "orc_smooth_record_states_f32")(
        C.byref(P), C.c_uint32(n), C.c_uint32(len(states)), ptr(flat), C.byref(rec._c))
    return rec


def tlm_from_states(P, states, dpos0, dvel0, dgrav, replay=False, f64=True):
    """orc_tlm_from_states_*: forward-mode (tangent-linear) derivatives of the
    smooth forward along a trajectory computed elsewhere (states as for
    smooth_record_states).  dpos0, dvel0: [K, N, 3] tangent directions of the
    first input state (its order); dgrav: [K, 3] tangents of the gravity
    vector.  replay=False linearizes each substep at the smooth substep (real
    P2G sums, exact pressure solve); replay=True at the reverse pass's replay
    of it (P2G at 2^22 fixed point, the substep's own CG tolerance), so that
    forward mode and the reverse pass are products of one Jacobian.  Returns
    (dpos, dvel), [K, N, 3], the tangents of the final state in the last
    substep's sort order."""
    real = np.float64 if f64 else np.float32
    n = len(states[0][0])
    K = len(dgrav)
    cols_p = [np.array(dpos0[k][:, a], dtype=real, copy=True) for k in range(K) for a in range(3)]
    cols_v = [np.array(dvel0[k][:, a], dtype=real, copy=True) for k in range(K) for a in range(3)]
    g = np.ascontiguousarray(np.asarray(dgrav, np.float32).reshape(K * 3))
    flat = np.concatenate([np.concatenate([np.asarray(p, real).T.ravel(), np.asarray(v, real).T.ravel(),
                                           np.asarray(pr, real).ravel()]) for p, v, pr in states])
    vp = C.c_void_p * (3 * K)
    getattr(lib(), "orc_tlm_from_states_f64" if f64 else "orc_tlm_from_states_f32")(
        C.byref(P), C.c_uint32(n), C.c_uint32(len(states)), ptr(flat), C.c_uint32(K), vp(*[ptr(c) for c in cols_p]),
        vp(*[ptr(c) for c in cols_v]), ptr(g), C.c_int(1 if replay else 0))
    dpos = np.stack([np.stack(cols_p[3 * k:3 * k + 3], 1) for k in range(K)])
    dvel = np.stack([np.stack(cols_v[3 * k:3 * k + 3], 1) for k in range(K)])
    return dpos, dvel


# ---------------------------------------------------------------------------
# Scene presets, spawn counts and the initial lattice (flip_oracle_scene.c).

def _scene_lib():
    L = lib()
    if not getattr(L, "_scene_typed", False):
        L.orc_scene_preset_blocks.restype = C.c_uint32
        L.orc_scene_preset_blocks.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_void_p]
        L.orc_scene_seeded_fraction.restype = C.c_float
        L.orc_scene_seeded_fraction.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_scene_resolved_count.restype = C.c_uint32
        L.orc_scene_resolved_count.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_uint32]
        L.orc_scene_effective_density.restype = C.c_float
        L.orc_scene_effective_density.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_scene_generate.restype = C.c_uint64
        L.orc_scene_generate.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_uint64, C.c_void_p]
        L._scene_typed = True
    return L


def scene_preset_blocks(preset, drop_height, fill_level) -> np.ndarray:
    """preset_blocks (src/scene/mod.rs:208-238): (n, 6) float32, min xyz then max xyz."""
    out = np.zeros((8, 6), np.float32)
    n = _scene_lib().orc_scene_preset_blocks(preset, np.float32(drop_height), np.float32(fill_level), ptr(out))
    return out[:n].copy()


def scene_seeded_fraction(blocks) -> np.float32:
    """seeded_volume_fraction (src/scene/mod.rs:134-146)."""
    b = np.ascontiguousarray(blocks, np.float32)
    return np.float32(_scene_lib().orc_scene_seeded_fraction(ptr(b), len(b)))


def scene_resolved_count(res, blocks, density, override=0) -> int:
    """resolved_particle_count (src/scene/mod.rs:179-193)."""
    b = np.ascontiguousarray(blocks, np.float32)
    r = np.asarray(res, np.uint32)
    return int(_scene_lib().orc_scene_resolved_count(ptr(r), ptr(b), len(b), np.float32(density), override))


def scene_effective_density(res, blocks, count) -> np.float32:
    """effective_particle_density_for_count (src/scene/mod.rs:278-289)."""
    b = np.ascontiguousarray(blocks, np.float32)
    r = np.asarray(res, np.uint32)
    return np.float32(_scene_lib().orc_scene_effective_density(ptr(r), ptr(b), len(b), count))


def scene_generate(blocks, particle_count, h, origin, res, want=True, info=False):
    """generate_particles (src/gpu/fluid.rs:1838-1913): (k, 3) float32 positions
    (want=False: the count k only).  info=True also returns the per-block
    lattice as an (n, 7) array: wmin xyz, spacing, counts xyz."""
    b = np.ascontiguousarray(blocks, np.float32)
    r = np.asarray(res, np.uint32)
    o = np.asarray(origin, np.float32)
    L = _scene_lib()
    inf = np.zeros((max(len(b), 1), 7), np.float32)
    k = int(L.orc_scene_generate(ptr(b), len(b), particle_count, np.float32(h), ptr(o), ptr(r), None, 0, ptr(inf)))
    if not want:
        return (k, inf[:len(b)]) if info else k
    out = np.zeros((k, 3), np.float32)
    L.orc_scene_generate(ptr(b), len(b), particle_count, np.float32(h), ptr(o), ptr(r), ptr(out), k, None)
    return (out, inf[:len(b)]) if info else out
