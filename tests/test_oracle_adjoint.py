"""The FP64 reverse pass of the restated step (oracle/flip_oracle_adjoint.h)
against central finite differences of the FP64 forward.

The adjoint holds the forward's discrete decisions fixed and counts P2G's
fixed-point rounding as the identity, so the comparison runs where the
forward is smooth in the perturbed input: one or two substeps, P2G fixed
point at 2^28 (its rounding far below the perturbation: at h = 1e-4 the
central difference then agrees with the adjoint to ~1e-5, while at 2^16 the
rounding alone moves it by ~50%), a tight pressure solve, and perturbations
small enough that the forward's discrete decisions do not change.  Within that
regime a wrong sign, a dropped term or a transposed stencil shows up as an
O(1) relative error; the bar is 1e-3 relative (2e-2 for gravity, whose
derivative is a small residual of the hydrostatic balance).
"""
import numpy as np
import pytest

import oracle
import scenes
from helpers import oracle_params

TARGET = np.array([0.7, 0.3, 0.5])


def smooth_params(sc, stiff=None, g=-9.81):
    P = oracle_params(sc, 1)
    P.fixed_scale = 2.0 ** 28
    P.tol, P.tol_sq, P.pressure_iters = 1e-15, 1e-30, 20000
    P.grav[1] = g
    if stiff is not None:
        P.stiff = stiff
    return P


def loss_and_tape(P, pos, vel, steps):
    T = oracle.AdjointTape(P, pos, vel, steps=steps, every=1, f64=True)
    c = T.pos.mean(0)
    return float(((c - TARGET) ** 2).sum()), T


def centroid_grad(T):
    c = T.pos.mean(0)
    return np.tile(2.0 * (c - TARGET) / len(T.pos), (len(T.pos), 1))


@pytest.mark.parametrize("steps,stiff", [(1, 0.0), (1, 0.45), (2, 0.0)])
def test_velocity_directional_derivatives_match_central_differences(steps, stiff):
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    x0 = sc["pos"].astype(np.float64)
    v0 = sc["vel"].astype(np.float64)
    P = smooth_params(sc, stiff)
    _, T = loss_and_tape(P, x0, v0, steps)
    _, gv, _ = T.backward(centroid_grad(T), None)
    # Directions whose derivative stands well above the forward's rounding
    # floor (~1e-12 in the loss; from the second substep on every particle's
    # P2G rounding moves with the perturbation): a uniform shift along x,
    # and each of the three particles with the largest gradient moved along
    # its own gradient.  Random directions sum per-particle terms that mostly
    # cancel, so their central difference sits at that floor.
    hu, hp = (1e-4, 1e-3) if steps == 1 else (1e-3, 1e-2)
    dirs = [(np.tile([1.0, 0.0, 0.0], (len(v0), 1)), hu)]
    for p in np.argsort(-np.abs(gv).sum(1))[:3]:
        d = np.zeros_like(v0)
        d[p] = gv[p] / np.linalg.norm(gv[p])
        dirs.append((d, hp))
    for d, h in dirs:
        lp, _ = loss_and_tape(P, x0, v0 + h * d, steps)
        lm, _ = loss_and_tape(P, x0, v0 - h * d, steps)
        fd = (lp - lm) / (2 * h)
        ad = float((gv * d).sum())
        assert abs(fd - ad) <= 1e-3 * abs(fd), (fd, ad)


def test_gravity_derivative_matches_central_differences():
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    x0 = sc["pos"].astype(np.float64)
    v0 = sc["vel"].astype(np.float64)
    _, T = loss_and_tape(smooth_params(sc, 0.0), x0, v0, 2)
    _, _, gg = T.backward(centroid_grad(T), None)
    h = 1e-4
    lp, _ = loss_and_tape(smooth_params(sc, 0.0, -9.81 + h), x0, v0, 2)
    lm, _ = loss_and_tape(smooth_params(sc, 0.0, -9.81 - h), x0, v0, 2)
    fd = (lp - lm) / (2 * h)
    assert abs(fd - gg[1]) <= 2e-2 * abs(fd), (fd, gg[1])


def test_position_derivative_matches_central_differences():
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    x0 = sc["pos"].astype(np.float64)
    v0 = sc["vel"].astype(np.float64)
    P = smooth_params(sc, 0.0)
    _, T = loss_and_tape(P, x0, v0, 1)
    gx, _, _ = T.backward(centroid_grad(T), None)
    d = np.random.default_rng(2).standard_normal(x0.shape)
    h = 1e-7
    lp, _ = loss_and_tape(P, x0 + h * d, v0, 1)
    lm, _ = loss_and_tape(P, x0 - h * d, v0, 1)
    fd = (lp - lm) / (2 * h)
    ad = float((gx * d).sum())
    assert abs(fd - ad) <= 1e-2 * abs(fd), (fd, ad)


def test_checkpoint_interval_does_not_change_the_gradient():
    """Every interval recomputes the same forward: the reverse pass is the
    same for every = 1, 2 and 5 over five substeps."""
    sc = scenes.small_tank(n=12, seed=7, fill=0.5, ppc=8, vel_scale=0.5)
    x0 = sc["pos"].astype(np.float64)
    v0 = sc["vel"].astype(np.float64)
    P = oracle_params(sc, 1)
    out = []
    for every in (1, 2, 5):
        T = oracle.AdjointTape(P, x0, v0, steps=5, every=every, f64=True)
        out.append(T.backward(centroid_grad(T), None))
    for gx, gv, gg in out[1:]:
        assert np.allclose(gx, out[0][0], rtol=0, atol=1e-12 * np.abs(out[0][0]).max())
        assert np.allclose(gv, out[0][1], rtol=0, atol=1e-12 * np.abs(out[0][1]).max())
        assert np.allclose(gg, out[0][2], rtol=1e-10, atol=1e-18)


def test_frozen_decision_central_differences_over_ten_substeps():
    """Over many substeps the +h and -h runs of the plain forward classify a
    few surface cells differently, and each flip makes the loss jump; the
    adjoint differentiates the forward with those decisions fixed.  Replaying
    the base run's cell types and occupancies (oracle.run_frozen) gives that
    function, and its central difference matches the adjoint at the default
    physics (volume stiffness, 1e-5 CG, 2^16 fixed point): ~1e-3 at h = 1e-3,
    ~2e-4 at h = 1e-4 (bar 2e-3 / 1e-3)."""
    sc = scenes.small_tank(n=24, seed=9, fill=0.5, ppc=8, vel_scale=0.3)
    x0 = sc["pos"].astype(np.float64)
    v0 = sc["vel"].astype(np.float64)
    P = oracle_params(sc, 1)
    T = oracle.AdjointTape(P, x0, v0, steps=10, every=5, f64=True, record=True)
    _, gv, _ = T.backward(centroid_grad(T), None)

    def L(vel):
        pos, _, _ = oracle.run_frozen(P, x0, vel, T.ct_steps, T.occ_steps)
        return float(((pos.mean(0) - TARGET) ** 2).sum())

    d = np.zeros_like(v0)
    d[:, 0] = 1.0
    ad = float((gv * d).sum())
    for h, bar in ((1e-3, 2e-3), (1e-4, 1e-3)):
        fd = (L(v0 + h * d) - L(v0 - h * d)) / (2 * h)
        assert abs(fd - ad) <= bar * abs(fd), (h, fd, ad)


def test_zero_upstream_gradient_gives_zero():
    sc = scenes.small_tank(n=12, seed=7, fill=0.5, ppc=8, vel_scale=0.5)
    T = oracle.AdjointTape(oracle_params(sc, 1), sc["pos"].astype(np.float64), sc["vel"].astype(np.float64),
                           steps=2, every=1)
    gx, gv, gg = T.backward(None, None)
    assert not gx.any() and not gv.any() and not gg.any()


def _smooth_trajectory(P, x0, v0, steps):
    """The smooth forward's own trajectory, one recorded substep at a time:
    the input state of every substep and the record of all of them."""
    rec = oracle.SmoothRecord(P, len(x0), steps)
    states = []
    pos, vel, pr = x0, v0, np.zeros(P.nx * P.ny * P.nz)
    for s in range(steps):
        states.append((pos, vel, pr))
        r1 = oracle.SmoothRecord(P, len(x0), 1)
        pos, vel, pr = oracle.smooth_forward(P, pos, vel, r1, pressure=pr, record=True)
        for k in ("perm", "occ", "ct", "cell", "bits", "support"):
            getattr(rec, k)[s] = getattr(r1, k)[0]
    return rec, states, pos


def _centroid_jacobians(P, x0, v0, states, rec, h, free_stencils):
    """d(final centroid)/d(uniform velocity shift x, y, z; gravity_y): the
    FP64 reverse pass along `states` (one upstream gradient per centroid
    component) and central differences (h) of the smooth forward replaying
    `rec` from (x0, v0)."""
    T = oracle.AdjointTape.from_states(P, states)
    n = len(x0)
    J_ad = np.zeros((3, 4))
    for i in range(3):
        up = np.zeros((n, 3))
        up[:, i] = 1.0 / n
        _, gv, gg = T.backward(up, None)
        J_ad[i] = [gv[:, 0].sum(), gv[:, 1].sum(), gv[:, 2].sum(), gg[1]]
    T.close()

    def centroid(vel, gy):
        Pg = oracle.make_params(P.nx, P.ny, P.nz, P.h, dt=P.dt, flip_blend=P.flip_blend, grav=(0.0, gy, 0.0),
                                rest=P.rest, stiff=P.stiff, drift_clamp=P.drift_clamp, cfl=P.cfl,
                                surface_dilation=P.surface_dilation, pressure_iters=P.pressure_iters,
                                tol=P.tol, reduce_mode=1)
        p_, _, _ = oracle.smooth_forward(Pg, x0, vel, rec, free_stencils=free_stencils)
        assert np.all(np.isfinite(p_))
        return p_.mean(0)

    J_fd = np.zeros((3, 4))
    g0 = float(P.grav[1])
    for j in range(3):
        d = np.zeros_like(v0)
        d[:, j] = 1.0
        J_fd[:, j] = (centroid(v0 + h * d, g0) - centroid(v0 - h * d, g0)) / (2 * h)
    J_fd[:, 3] = (centroid(v0, g0 + h) - centroid(v0, g0 - h)) / (2 * h)
    return J_ad, J_fd


@pytest.mark.parametrize("free_stencils,h,bar", [(False, 1e-3, 2e-4), (True, 1e-4, 3e-4)])
def test_reverse_pass_is_the_derivative_of_the_frozen_decision_forward(free_stencils, h, bar):
    """Twenty substeps at 24^3 at the default physics (volume stiffness, 1e-5
    CG, 2^16 fixed point): the FP64 reverse pass, as a Jacobian of the final
    centroid w.r.t. uniform velocity shifts and gravity, against central
    differences of the smooth forward with every discrete decision of the
    trajectory replayed (sort order, cell types and occupancies, face
    supports, CFL and wall-clamp branches, and - unless free_stencils - each
    particle's stencil bases), no fixed-point rounding and an exact pressure
    solve: the function the adjoint's conventions define.  Measured 2.9e-5
    (recorded stencils, h = 1e-3) and 6.6e-5 (free stencils, h = 1e-4).  A
    dropped term, a wrong sign or a transposed stencil in the reverse pass
    moves the Jacobian by O(1)."""
    sc = scenes.small_tank(n=24, seed=9, fill=0.5, ppc=8, vel_scale=0.3)
    P = oracle_params(sc, 1)
    x0, v0 = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    rec, states, _ = _smooth_trajectory(P, x0, v0, 20)
    J_ad, J_fd = _centroid_jacobians(P, x0, v0, states, rec, h, free_stencils)
    assert np.linalg.norm(J_ad - J_fd) <= bar * np.linalg.norm(J_fd), (J_ad, J_fd)


def test_smooth_forward_record_then_replay_reproduces_the_trajectory():
    """Replaying a record from the same start reproduces the recorded run
    (to the parallel sums' rounding), and the smooth forward stays close to
    the restated step it smooths (the difference is the 1e-5 CG tolerance
    and the 2^16 rounding of the step)."""
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    P = oracle_params(sc, 1)
    x0, v0 = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    rec = oracle.SmoothRecord(P, len(x0), 4)
    pa, va, _ = oracle.smooth_forward(P, x0, v0, rec, record=True)
    pb, vb, _ = oracle.smooth_forward(P, x0, v0, rec)
    assert np.abs(pa - pb).max() <= 1e-12 and np.abs(va - vb).max() <= 1e-10
    r = oracle.run(P, x0, v0, steps=4, sort_period=1, f64=True)
    assert np.abs(pa - r.pos).max() <= 1e-3
    assert rec.perm.shape == (4, len(x0)) and rec.support.any()


def _tlm_directions(n, rng):
    """Tangent directions: uniform velocity shifts x, y, z, gravity_y, and a
    random per-particle position and velocity perturbation."""
    dpos = np.zeros((5, n, 3))
    dvel = np.zeros((5, n, 3))
    for a in range(3):
        dvel[a, :, a] = 1.0
    dpos[4] = rng.standard_normal((n, 3)) * 1e-2
    dvel[4] = rng.standard_normal((n, 3))
    dgrav = np.zeros((5, 3))
    dgrav[3, 1] = 1.0
    return dpos, dvel, dgrav


def test_tangent_linear_model_is_the_derivative_of_the_frozen_decision_forward():
    """Forward mode (orc_tlm_from_states: every operation of the smooth
    substep differentiated as a dual number, decisions taken from each input
    state) against central differences of the smooth forward replaying the
    same decisions, three substeps at 16^3, every particle's final position
    and velocity.  Measured ~1e-9 relative at h = 1e-5 (x and z velocity
    shifts, a random per-particle direction) and 2e-8 for gravity at 1e-3.
    Gravity is a float32 parameter, so its step is the one the rounded
    values g0 +- h actually take (at h = 1e-5 the f32 spacing near 9.81,
    9.5e-7, moves an unrounded quotient by ~5%).  A uniform y shift is
    removed by the projection (the liquid rests on the floor), so both sides
    are ~0 there.  A dropped product-rule term, the wrong sign of a
    tent-weight derivative or a missing gravity source each leave an O(1)
    error."""
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    P = oracle_params(sc, 1)
    x0, v0 = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    rec, states, _ = _smooth_trajectory(P, x0, v0, 3)
    dpos0, dvel0, dgrav = _tlm_directions(len(x0), np.random.default_rng(3))
    tp, tv = oracle.tlm_from_states(P, states, dpos0, dvel0, dgrav)
    g0 = float(P.grav[1])

    def forward(k, s, h):
        gy = float(np.float32(g0 + s * h * dgrav[k, 1]))
        Pg = oracle.make_params(P.nx, P.ny, P.nz, P.h, dt=P.dt, flip_blend=P.flip_blend, grav=(0.0, gy, 0.0),
                                rest=P.rest, stiff=P.stiff, drift_clamp=P.drift_clamp, cfl=P.cfl,
                                surface_dilation=P.surface_dilation, pressure_iters=P.pressure_iters, tol=P.tol,
                                reduce_mode=1)
        p_, v_, _ = oracle.smooth_forward(Pg, x0 + s * h * dpos0[k], v0 + s * h * dvel0[k], rec)
        return p_, v_, gy

    for k in range(5):
        h = 1e-3 if k == 3 else 1e-5
        (pa, va, ga), (pb, vb, gb) = forward(k, 1.0, h), forward(k, -1.0, h)
        step = (ga - gb) if k == 3 else 2 * h
        fp, fv = (pa - pb) / step, (va - vb) / step
        if k == 1:
            assert np.linalg.norm(tp[k]) <= 1e-8 and np.linalg.norm(fp) <= 1e-8
            continue
        assert np.linalg.norm(tp[k] - fp) <= 1e-6 * np.linalg.norm(fp), (k, np.linalg.norm(tp[k] - fp),
                                                                         np.linalg.norm(fp))
        assert np.linalg.norm(tv[k] - fv) <= 1e-6 * np.linalg.norm(fv), (k, np.linalg.norm(tv[k] - fv),
                                                                         np.linalg.norm(fv))


def test_tangent_linear_model_agrees_with_the_reverse_pass():
    """The FP64 reverse pass and forward mode along the same trajectory
    (24^3, twenty substeps at the default physics): the reverse pass's
    Jacobian of the final centroid w.r.t. uniform velocity shifts and gravity
    against the tangents' centroid.  Linearized at the same point (replay:
    P2G at the reverse pass's 2^22 fixed point, the substep's own CG
    tolerance) the two are products of one Jacobian and agree to rounding
    (measured 1e-14): the reverse pass is the exact transpose of the forward
    linearization.  Linearized at the smooth substep instead (real P2G sums,
    exact solve) they differ by how much the Jacobian moves with the
    linearization point, measured 1.5e-5 here."""
    sc = scenes.small_tank(n=24, seed=9, fill=0.5, ppc=8, vel_scale=0.3)
    P = oracle_params(sc, 1)
    x0, v0 = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    _, states, _ = _smooth_trajectory(P, x0, v0, 20)
    T = oracle.AdjointTape.from_states(P, states)
    n = len(x0)
    J_ad = np.zeros((3, 4))
    for i in range(3):
        up = np.zeros((n, 3))
        up[:, i] = 1.0 / n
        _, gv, gg = T.backward(up, None)
        J_ad[i] = [gv[:, 0].sum(), gv[:, 1].sum(), gv[:, 2].sum(), gg[1]]
    T.close()
    dpos0, dvel0, dgrav = _tlm_directions(n, np.random.default_rng(4))
    for replay, bar in ((True, 1e-11), (False, 5e-5)):
        tp, _ = oracle.tlm_from_states(P, states, dpos0[:4], dvel0[:4], dgrav[:4], replay=replay)
        J_tlm = np.stack([tp[k].mean(0) for k in range(4)], 1)
        assert np.linalg.norm(J_ad - J_tlm) <= bar * np.linalg.norm(J_tlm), (replay, J_ad, J_tlm)


def test_dot_product_of_forward_mode_and_reverse_pass_per_substep():
    """<w, J u> from forward mode equals <J^T w, u> from the reverse pass for
    random full-state directions u (positions and velocities of every
    particle) and cotangents w, one substep at each of several states of a
    sixteen-substep run at 16^3, both linearized at the reverse pass's replay.
    A transposed stencil, a missing term of one side or a wrong sign breaks
    the identity by O(1); measured 5e-14 to 1.5e-12 relative."""
    sc = scenes.small_tank(n=16, seed=5, fill=0.5, ppc=8, vel_scale=0.5)
    P = oracle_params(sc, 1)
    x0, v0 = sc["pos"].astype(np.float64), sc["vel"].astype(np.float64)
    _, states, _ = _smooth_trajectory(P, x0, v0, 16)
    n = len(x0)
    rng = np.random.default_rng(11)
    for s in (0, 7, 15):
        u_p, u_v = rng.standard_normal((n, 3)) * 1e-2, rng.standard_normal((n, 3))
        w_p, w_v = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)) * 1e-2
        tp, tv = oracle.tlm_from_states(P, [states[s]], u_p[None], u_v[None], np.zeros((1, 3)), replay=True)
        lhs = (w_p * tp[0]).sum() + (w_v * tv[0]).sum()
        T = oracle.AdjointTape.from_states(P, [states[s]])
        gx, gv, _ = T.backward(w_p, w_v)
        T.close()
        rhs = (gx * u_p).sum() + (gv * u_v).sum()
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (s, lhs, rhs)
