"""Closed-form pins for the parts of the CPU restatement that the reference's
own tests do not cover: P2G taps and weights, the face epilogue, divergence of
a linear field and the drift source, the Poisson operator (SPD, dense solve),
a hydrostatic column, and the G2P blend / CFL / wall clamp / friction.  Each
expected value follows from the mathematics, not from the code under test."""
import numpy as np
import pytest

N16 = 16
H16 = 1.0 / 16.0


def fidx(dims, i, j, k):
    return i + dims[0] * (j + dims[1] * k)


def face_dims(n, a):
    d = [n, n, n]
    d[a] += 1
    return d


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_p2g_particle_on_a_face_hits_only_that_face(orc, axis):
    """A particle at the centre of face (i,j,k) of `axis` has g = (i,j,k)
    exactly: weight 1 on that face, 0 elsewhere (zero weights are skipped)."""
    P = orc.make_params(N16, N16, N16, H16)
    i, j, k = 5, 7, 9
    g = [i + 0.5, j + 0.5, k + 0.5]
    g[axis] -= 0.5
    pos = np.array([[c * H16 for c in g]], np.float32)
    v = np.array([[0.3, -1.25, 2.5]], np.float32)
    num, den = orc.p2g(P, pos, v)
    d = face_dims(N16, axis)
    f = fidx(d, i, j, k)
    assert den[axis][f] == 65536 and den[axis].sum() == 65536
    assert num[axis][f] == int(np.rint(np.float32(v[0, axis]) * np.float32(65536)))
    assert np.count_nonzero(den[axis]) == 1


def test_p2g_tent_weights_split_linearly(orc):
    """Offset a quarter cell along x from a u face: weights 0.75 / 0.25."""
    P = orc.make_params(N16, N16, N16, H16)
    pos = np.array([[4.25 * H16, 6.5 * H16, 3.5 * H16]], np.float32)
    num, den = orc.p2g(P, pos, np.array([[1.0, 0.0, 0.0]], np.float32))
    d = face_dims(N16, 0)
    assert den[0][fidx(d, 4, 6, 3)] == 49152 and den[0][fidx(d, 5, 6, 3)] == 16384
    assert num[0][fidx(d, 4, 6, 3)] == 49152


def test_p2g_partition_of_unity_and_uniform_field(orc):
    rng = np.random.default_rng(0)
    n = 2000
    P = orc.make_params(N16, N16, N16, H16)
    pos = (2 + rng.random((n, 3)) * 12).astype(np.float32) * np.float32(H16)
    vel = np.tile(np.array([[0.75, -2.0, 1.5]], np.float32), (n, 1))
    num, den = orc.p2g(P, pos, vel)
    for a in range(3):
        tot = int(den[a].astype(np.int64).sum())
        assert abs(tot - n * 65536) <= 4 * n  # 8 rounded taps per particle
        m = den[a] > 0
        ratio = num[a][m].astype(np.float64) / den[a][m]
        # |num - v*den| <= 8 * 0.5 rounding per tap sum
        assert np.all(np.abs(ratio - vel[0, a]) <= 4.0 * (1 + abs(vel[0, a])) / den[a][m])


def test_p2g_is_linear_in_velocity_up_to_rounding(orc):
    rng = np.random.default_rng(1)
    P = orc.make_params(N16, N16, N16, H16)
    pos = (1.5 + rng.random((500, 3)) * 13).astype(np.float32) * np.float32(H16)
    v1 = rng.standard_normal((500, 3)).astype(np.float32)
    v2 = rng.standard_normal((500, 3)).astype(np.float32)
    n1, d1 = orc.p2g(P, pos, v1)
    n2, _ = orc.p2g(P, pos, v2)
    n12, d12 = orc.p2g(P, pos, (v1 + v2).astype(np.float32))
    for a in range(3):
        assert np.array_equal(d1[a], d12[a])
        assert np.max(np.abs(n12[a].astype(np.int64) - n1[a] - n2[a])) <= 2 * 8 * 8


def _interior_liquid(n):
    occ = np.zeros(n ** 3, np.uint32)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                occ[i + n * (j + n * k)] = 8
    return occ


def test_divergence_of_linear_field_and_drift_source(orc):
    n, h = 8, 0.25
    alpha = 1.5
    P = orc.make_params(n, n, n, h, stiff=0.0)
    occ = _interior_liquid(n)
    ct, _ = orc.classify(P, occ)
    d = face_dims(n, 0)
    u = np.zeros(d[0] * d[1] * d[2], np.float32)
    for k in range(n):
        for j in range(n):
            for i in range(n + 1):
                u[fidx(d, i, j, k)] = alpha * i * h
    vel = [u, np.zeros((n + 1) * n * n, np.float32), np.zeros((n + 1) * n * n, np.float32)]
    div = orc.divergence(P, vel, ct, occ)
    assert np.allclose(div[ct == 1], alpha, rtol=0, atol=1e-6)
    assert np.all(div[ct != 1] == 0)
    # Drift: occ = 2 * rest -> over = 1 -> drift = min(stiff, clamp).
    P2 = orc.make_params(n, n, n, h, rest=4.0, stiff=0.45, drift_clamp=0.3)
    div2 = orc.divergence(P2, vel, ct, occ)
    assert np.allclose(div2[ct == 1], alpha - 0.3, atol=1e-6)
    P3 = orc.make_params(n, n, n, h, rest=4.0, stiff=0.2, drift_clamp=0.5)
    assert np.allclose(orc.divergence(P3, vel, ct, occ)[ct == 1], alpha - 0.2, atol=1e-6)


def _random_liquid_tank(orc, n, seed):
    rng = np.random.default_rng(seed)
    P = orc.make_params(n, n, n, 1.0 / n, stiff=0.0, liquid_threshold=1, surface_dilation=0)
    occ = np.zeros(n ** 3, np.uint32)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                if j < n - 3 or rng.random() < 0.4:
                    occ[i + n * (j + n * k)] = 1
    ct, _ = orc.classify(P, occ)
    return P, ct


def test_poisson_operator_is_spd_with_the_stencil_structure(orc):
    n = 7
    P, ct = _random_liquid_tank(orc, n, 3)
    C = n ** 3
    liq = np.nonzero(ct == 1)[0]
    A = np.zeros((len(liq), len(liq)))
    for col, c in enumerate(liq):
        e = np.zeros(C, np.float64)
        e[c] = 1.0
        A[:, col] = orc.apply_poisson(P, ct, e, f64=True)[liq]
    assert np.array_equal(A, A.T)
    assert np.min(np.linalg.eigvalsh(A)) > 0
    # diagonal = non-solid neighbours, off-diagonal -1 per liquid neighbour
    pos = {c: r for r, c in enumerate(liq)}
    for r, c in enumerate(liq):
        nbs = [c - 1, c + 1, c - n, c + n, c - n * n, c + n * n]
        assert A[r, r] == sum(1 for b in nbs if ct[b] != 0)
        for b in nbs:
            if ct[b] == 1:
                assert A[r, pos[b]] == -1


@pytest.mark.parametrize("f64", [False, True])
def test_cg_matches_a_dense_solve(orc, f64):
    n = 7
    P, ct = _random_liquid_tank(orc, n, 4)
    P.reduce_mode = 1
    P.pressure_iters = 400
    P.tol = 1e-7
    P.tol_sq = float(np.float32(1e-7) ** 2)
    P.warm_start = 0
    rng = np.random.default_rng(5)
    div = np.where(ct == 1, rng.standard_normal(n ** 3), 0.0)
    p, st = orc.pressure_solve(P, ct, div, f64=f64)
    liq = np.nonzero(ct == 1)[0]
    A = np.zeros((len(liq), len(liq)))
    for col, c in enumerate(liq):
        e = np.zeros(n ** 3)
        e[c] = 1
        A[:, col] = orc.apply_poisson(P, ct, e, f64=True)[liq]
    scale = P.rho * P.h * P.h / P.dt
    x = np.linalg.solve(A, -scale * div[liq])
    err = np.linalg.norm(p[liq] - x) / np.linalg.norm(x)
    assert err < (1e-5 if not f64 else 1e-7)
    assert st.cg_iterations < 400
    assert abs(st.true_residual_sq - st.rs_final) <= 1e-3 * st.rs_initial * 1e-10 + 1e-6 * st.rs_initial


def test_hydrostatic_column_pressure_is_linear_in_depth(orc):
    """Liquid up to layer top, air above, gravity on v faces: after the
    projection p = rho*|g|*h*depth (depth 1 at the top liquid layer)."""
    n, top = 8, 5
    h, dt, rho, g = 0.125, 1.0 / 120.0, 1000.0, -9.81
    P = orc.make_params(n, n, n, h, dt=dt, rho=rho, stiff=0.0, reduce_mode=1, warm_start=0,
                        pressure_iters=400, tol=1e-6, clamp_tol=False)
    occ = np.zeros(n ** 3, np.uint32)
    for k in range(1, n - 1):
        for j in range(1, top + 1):
            for i in range(1, n - 1):
                occ[i + n * (j + n * k)] = 1
    ct, _ = orc.classify(P, occ)
    dv = face_dims(n, 1)
    v = np.zeros(dv[0] * dv[1] * dv[2], np.float32)
    for k in range(n):
        for j in range(n + 1):
            for i in range(n):
                lo = ct[i + n * ((j - 1) + n * k)] if j > 0 else 0
                hi = ct[i + n * (j + n * k)] if j < n else 0
                if lo != 0 and hi != 0 and (lo == 1 or hi == 1):
                    v[fidx(dv, i, j, k)] = np.float32(g) * np.float32(dt)
    vel = [np.zeros((n + 1) * n * n, np.float32), v, np.zeros((n + 1) * n * n, np.float32)]
    div = orc.divergence(P, vel, ct)
    p, st = orc.pressure_solve(P, ct, div)
    for j in range(1, top + 1):
        depth = top + 1 - j
        expect = rho * abs(g) * h * depth
        got = p[3 + n * (j + n * 4)]
        assert abs(got - expect) / expect < 1e-4, (j, got, expect)
    v2 = orc.gradient_enforce(P, p, ct, vel)
    # interior v faces of the column come to rest
    for j in range(2, top + 1):
        assert abs(v2[1][fidx(dv, 3, j, 4)]) < 1e-5


def _uniform_faces(n, val):
    return [np.full(((n + 1) * n * n), val[a], np.float32) for a in range(3)]


@pytest.mark.parametrize("alpha", [0.0, 1.0, 0.3])
def test_g2p_uniform_field_flip_pic_blend(orc, alpha):
    n = 16
    P = orc.make_params(n, n, n, 1.0 / n, flip_blend=alpha, wall_friction=0.0, cfl=6.0)
    field = (0.4, -0.7, 0.25)
    pos = np.array([[0.3, 0.5, 0.6], [0.2, 0.11, 0.85]], np.float32)
    pv = np.array([[1.0, 0.5, -0.2], [-0.3, 0.0, 0.1]], np.float32)
    f = _uniform_faces(n, field)
    newp, newv = orc.g2p(P, pos, pv, f, f)
    expect_v = np.float32(1 - alpha) * np.array(field, np.float32) + np.float32(alpha) * pv
    assert np.allclose(newv, expect_v, atol=1e-6)
    assert np.allclose(newp, pos + newv * np.float32(P.dt), atol=1e-6)


def test_g2p_cfl_clamp_and_wall_contact(orc):
    n = 16
    h, dt = 1.0 / n, 1.0 / 120.0
    P = orc.make_params(n, n, n, h, dt=dt, flip_blend=0.0, cfl=2.0)
    f = _uniform_faces(n, (100.0, 0.0, 0.0))
    pos = np.array([[0.5, 0.5, 0.5]], np.float32)
    _, v = orc.g2p(P, pos, np.zeros((1, 3), np.float32), f, f)
    maxs = np.float32(2.0) * np.float32(h) / np.float32(dt)
    assert abs(np.linalg.norm(v[0]) - maxs) < 1e-4 * maxs and v[0, 1] == 0
    # wall contact: moving into the low-x wall clamps to origin + 1.05h, v.x = 0
    f2 = _uniform_faces(n, (-10.0, 0.3, 0.0))
    pos2 = np.array([[0.07, 0.5, 0.5]], np.float32)
    p2, v2 = orc.g2p(P, pos2, np.zeros((1, 3), np.float32), f2, f2)
    assert p2[0, 0] == np.float32(np.float32(h) * np.float32(1.05)) and v2[0, 0] == 0
    assert abs(v2[0, 1] - 0.3) < 1e-6


def test_g2p_wall_friction_damps_tangential_components(orc):
    n = 16
    h = np.float32(1.0 / n)
    P = orc.make_params(n, n, n, float(h), flip_blend=0.0, wall_friction=0.5, cfl=6.0)
    lo = h * np.float32(1.05)
    f = _uniform_faces(n, (0.2, 0.4, -0.6))
    pos = np.array([[lo, 0.5, 0.5]], np.float32)  # on the low-x wall band edge: near_x = 1
    _, v = orc.g2p(P, pos, np.zeros((1, 3), np.float32), f, f)
    assert abs(v[0, 0] - 0.2) < 1e-6            # normal to the x walls: untouched
    assert abs(v[0, 1] - 0.4 * 0.5) < 1e-6      # tangential: x (1 - friction)
    assert abs(v[0, 2] + 0.6 * 0.5) < 1e-6
    pos_mid = np.array([[0.5, 0.5, 0.5]], np.float32)
    _, vm = orc.g2p(P, pos_mid, np.zeros((1, 3), np.float32), f, f)
    assert np.allclose(vm[0], [0.2, 0.4, -0.6], atol=1e-6)


def test_full_step_conserves_particles_and_keeps_them_inside(orc):
    import scenes
    from helpers import oracle_params
    sc = scenes.small_tank(n=16, seed=3)
    P = oracle_params(sc, 0)
    r = orc.run(P, sc["pos"], sc["vel"], steps=3)
    h = 1.0 / 16
    assert r.pos.shape == sc["pos"].shape
    assert r.pos.min() >= np.float32(1.05 * h) - 1e-7 and r.pos.max() <= 1 - 1.05 * h + 1e-7
    assert int(r.occupancy.sum()) == len(sc["pos"])
    assert np.all(np.isfinite(r.vel_particles))


# ---------------------------------------------------------------------------
# Face epilogue: normalize (sh/normalize.wgsl:15-27), save (sh/save_vel.wgsl:
# 14-19), forces (sh/forces.wgsl:30-84) and the pre-pressure enforce
# (sh/boundaries.wgsl:20-61), checked face by face against a brute-force
# statement built from the face's two neighbouring cells.

EPI_DIMS = (6, 7, 5)  # non-cubic: a transposed index or a swapped axis shows


def _epilogue_case(seed, dims=EPI_DIMS, interior_solids=3):
    """Cell types with a solid shell, random liquid / air inside and a few
    interior solid cells; num/den with zero, negative (wrapped) and
    positive denominators, num nonzero where den is zero."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    ct3 = np.zeros((nz, ny, nx), np.uint32)  # [k, j, i]
    ct3[1:-1, 1:-1, 1:-1] = rng.choice(np.array([1, 2], np.uint32), size=(nz - 2, ny - 2, nx - 2), p=[0.55, 0.45])
    inner = np.argwhere(ct3[1:-1, 1:-1, 1:-1] >= 0) + 1
    for k, j, i in inner[rng.choice(len(inner), interior_solids, replace=False)]:
        ct3[k, j, i] = 0
    num, den = [], []
    for a in range(3):
        d = [nx, ny, nz]
        d[a] += 1
        cnt = d[0] * d[1] * d[2]
        dd = rng.integers(1, 200_000, cnt).astype(np.int64)
        kind = rng.random(cnt)
        dd[kind < 0.25] = 0
        dd[(kind >= 0.25) & (kind < 0.3)] = -rng.integers(1, 1000, int(np.sum((kind >= 0.25) & (kind < 0.3))))
        nn = rng.integers(-3_000_000, 3_000_000, cnt).astype(np.int64)
        num.append(nn.astype(np.int32))
        den.append(dd.astype(np.int32))
    return ct3, num, den


def _epilogue_expected(dims, ct3, num, den, grav, dt, real=np.float32, mutate=None):
    """Per axis a, on the face array [k, j, i] (n_a + 1 along a): the lo cell
    is the face's coordinate minus one and the hi cell its coordinate along
    a, a cell outside the grid counting as solid.  saved = num/den where
    den > 0 else 0; vel = saved + g_a*dt where the lo or hi cell is liquid;
    then vel = 0 where the lo or hi cell is solid.  `mutate` plants one
    plausible mistake in this statement (test of the test's reach)."""
    nx, ny, nz = dims
    outs_v, outs_s = [], []
    for a in range(3):
        d = [nx, ny, nz]
        d[a] += 1
        shape = (d[2], d[1], d[0])
        n3 = num[a].reshape(shape).astype(np.int64)
        d3 = den[a].reshape(shape).astype(np.int64)
        ax = 2 - a  # numpy axis of the face's own direction
        keep = d3 > 0 if mutate != "den_nonzero" else d3 != 0
        safe = np.where(keep, d3, 1)
        saved = np.where(keep, real(1) * n3.astype(real) / safe.astype(real), real(0)).astype(real)
        if mutate == "den_zero_keeps_num":
            saved = np.where(d3 == 0, n3.astype(real), saved).astype(real)
        pad_w = [(0, 0)] * 3
        pad_w[ax] = (1, 1)
        ctp = np.pad(ct3, pad_w, constant_values=0)  # outside the grid: solid
        lo = np.take(ctp, np.arange(0, shape[ax]), axis=ax)
        hi = np.take(ctp, np.arange(1, shape[ax] + 1), axis=ax)
        if mutate == "hi_cell_shifted":
            hi = np.take(np.pad(ctp, pad_w, constant_values=0), np.arange(3, shape[ax] + 3), axis=ax)
        wet = (lo == 1) | (hi == 1)
        if mutate == "gravity_on_dry_faces":
            wet = np.ones_like(wet)
        g_a = grav[(a + 1) % 3] if mutate == "gravity_axis_rotated" else grav[a]
        gdt = real(g_a) * real(dt)
        vel = np.where(wet, saved + gdt, saved).astype(real)
        solid = (lo == 0) | (hi == 0)
        if mutate == "interior_solids_ignored":
            coord = np.indices(shape)[ax]
            solid = (coord == 0) | (coord == shape[ax] - 1)
        vel = np.where(solid, real(0), vel).astype(real)
        outs_v.append(vel.reshape(-1))
        outs_s.append(saved.reshape(-1))
    return outs_v, outs_s


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("f64", [False, True])
def test_face_epilogue_matches_the_per_face_closed_form(orc, seed, f64):
    nx, ny, nz = EPI_DIMS
    grav, dt = (0.7, -9.81, 2.3), 1.0 / 120.0
    P = orc.make_params(nx, ny, nz, 0.125, dt=dt, grav=grav)
    ct3, num, den = _epilogue_case(seed)
    vel, saved = orc.p2g_epilogue(P, ct3.reshape(-1), num, den, f64=f64)
    real = np.float64 if f64 else np.float32
    ev, es = _epilogue_expected(EPI_DIMS, ct3, num, den, [real(np.float32(g)) for g in grav],
                                real(np.float32(dt)), real=real)
    for a in range(3):
        assert np.array_equal(saved[a], es[a]), f"saved axis {a}"
        assert np.array_equal(vel[a], ev[a]), f"vel axis {a}"
    # the case exercises every branch: empty, negative and positive den;
    # wet and dry faces; faces zeroed by interior solids
    assert any(np.any(d == 0) and np.any(d < 0) for d in den)
    assert any(np.any((v != 0) & (s == 0)) for v, s in zip(vel, saved))  # gravity on a wet empty face


@pytest.mark.parametrize("mutation", ["gravity_on_dry_faces", "den_nonzero", "den_zero_keeps_num",
                                      "hi_cell_shifted", "interior_solids_ignored", "gravity_axis_rotated"])
def test_face_epilogue_closed_form_rejects_planted_mistakes(orc, mutation):
    """Each plausible mistake, planted in the closed form, disagrees with the
    restatement on this case: the per-face check above would catch it."""
    nx, ny, nz = EPI_DIMS
    grav, dt = (0.7, -9.81, 2.3), 1.0 / 120.0
    P = orc.make_params(nx, ny, nz, 0.125, dt=dt, grav=grav)
    ct3, num, den = _epilogue_case(0)
    vel, saved = orc.p2g_epilogue(P, ct3.reshape(-1), num, den)
    g32 = [np.float32(g) for g in grav]
    ev, es = _epilogue_expected(EPI_DIMS, ct3, num, den, g32, np.float32(dt), mutate=mutation)
    with np.errstate(invalid="ignore"):
        same = all(np.array_equal(vel[a], ev[a]) and np.array_equal(saved[a], es[a]) for a in range(3))
    assert not same


def test_face_epilogue_of_the_step_matches_the_closed_form(orc):
    """Inside a full substep: the restatement's vel/saved after P2G equal the
    closed form applied to its own num/den and cell types."""
    import scenes
    from helpers import oracle_params
    sc = scenes.small_tank(n=12, seed=8, fill=0.6)
    P = oracle_params(sc, 0)
    r = orc.run(P, sc["pos"], sc["vel"], steps=1)
    ct3 = r.cell_type.reshape(12, 12, 12)
    ev, es = _epilogue_expected((12, 12, 12), ct3, r.num, r.den, [np.float32(g) for g in P.grav],
                                np.float32(P.dt))
    for a in range(3):
        assert np.array_equal(r.saved[a], es[a]) and np.array_equal(r.vel[a], ev[a])
