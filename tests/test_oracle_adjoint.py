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
