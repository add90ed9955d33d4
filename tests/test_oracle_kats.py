"""The reference's own known-answer tests, run against the CPU restatement.

Each test names the reference test it ports (app/crates/fluid-lab/src/...).
Inputs are generated exactly as the reference tests generate them (the same
LCGs in f32); expected values are the reference's asserted numbers.
"""
import numpy as np
import pytest


def lcg_stream(seed):
    st = np.uint32(seed)
    a, c = np.uint32(1664525), np.uint32(1013904223)
    while True:
        with np.errstate(over="ignore"):
            st = np.uint32(st * a + c)
        yield np.float32(int(st) >> 8) / np.float32(1 << 24)


def sample_positions(n, origin, h, dims):
    """src/lib.rs:1214-1229 (LCG seed 0x9E3779B9, f32 arithmetic)."""
    h = np.float32(h)
    ext = [np.float32(d) * h for d in dims]
    org = [np.float32(o) for o in origin]
    g = lcg_stream(0x9E3779B9)
    out = np.zeros((n, 3), np.float32)
    for i in range(n):
        for a in range(3):
            out[i, a] = org[a] + next(g) * ext[a]
    return out


def keys_of(orc, P, pos):
    return np.array([orc.cell_key(P, *map(float, p)) for p in pos], np.uint32)


# ---- src/lib.rs particle_sort_tests ----
def test_exclusive_scan_matches_definition(orc):  # src/lib.rs:1232-1238
    hist = np.array([3, 0, 5, 1, 0, 4], np.uint32)
    off = orc.exclusive_scan(hist)
    assert off.tolist() == [0, 3, 3, 8, 9, 9]
    assert off[5] + hist[5] == hist.sum()


def test_permutation_is_a_bijection(orc):  # src/lib.rs:1240-1254
    dims, h, origin = (16, 8, 16), 0.1, (-0.8, -0.4, -0.8)
    P = orc.make_params(*dims, h, origin=origin)
    pos = sample_positions(5000, origin, h, dims)
    keys = keys_of(orc, P, pos)
    perm = orc.stable_permutation(keys, 16 * 8 * 16)
    assert len(perm) == len(pos)
    assert np.array_equal(np.sort(perm), np.arange(len(pos)))
    # Stable order = brute-force stable argsort of the keys.
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))


def test_each_bucket_shares_one_key_and_is_monotone(orc):  # src/lib.rs:1256-1284
    dims, h, origin = (12, 6, 10), 0.1, (-0.6, -0.3, -0.5)
    C = 12 * 6 * 10
    P = orc.make_params(*dims, h, origin=origin)
    pos = sample_positions(4000, origin, h, dims)
    keys = keys_of(orc, P, pos)
    perm = orc.stable_permutation(keys, C)
    sk = keys[perm]
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    hist = orc.histogram(P, pos)
    assert np.array_equal(hist, np.bincount(keys, minlength=C).astype(np.uint32))
    starts = orc.exclusive_scan(hist)
    for c in range(C):
        s, e = int(starts[c]), int(starts[c]) + int(hist[c])
        assert np.all(keys[perm[s:e]] == c)


def test_clamps_out_of_bounds_positions_into_edge_cells(orc):  # src/lib.rs:1286-1300
    P = orc.make_params(4, 4, 4, 1.0)
    pos = np.array([[-100, -100, -100], [100, 100, 100]], np.float32)
    keys = keys_of(orc, P, pos)
    assert keys.min() == 0 and keys.max() == 63
    perm = orc.stable_permutation(keys, 64)
    assert sorted(perm.tolist()) == [0, 1]


# ---- src/sim/mod.rs tests ----
def test_world_to_cell_roundtrips_centers_and_clamps_escapes(orc):  # src/sim/mod.rs:245-260
    P = orc.make_params(8, 8, 8, 0.25, origin=(-1, -1, -1))
    h = np.float32(0.25)
    c = [np.float32(-1) + (np.float32(i) + np.float32(0.5)) * h for i in (3, 4, 5)]
    assert orc.cell_key(P, *map(float, c)) == 3 + 8 * (4 + 8 * 5)
    assert orc.cell_key(P, -100, -100, -100) == 0
    assert orc.cell_key(P, 100, 100, 100) == 7 + 8 * (7 + 8 * 7)


def test_world_to_cell_roundtrips_on_rectangular_tank(orc):  # src/sim/mod.rs:304-330
    nx, ny, nz, h = 5, 9, 7, np.float32(0.3)
    origin = (np.float32(-0.75), np.float32(1.35), np.float32(-2.1))
    P = orc.make_params(nx, ny, nz, float(h), origin=origin)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = [origin[a] + (np.float32(v) + np.float32(0.5)) * h for a, v in enumerate((i, j, k))]
                assert orc.cell_key(P, *map(float, c)) == i + nx * (j + ny * k)
    assert orc.cell_key(P, -1e3, -1e3, -1e3) == 0
    assert orc.cell_key(P, 1e3, 1e3, 1e3) == (nx - 1) + nx * ((ny - 1) + ny * (nz - 1))


def test_boundary_cells_are_solid_interior_marked_by_occupancy(orc):  # src/sim/mod.rs:262-276
    n = 6
    P = orc.make_params(n, n, n, 1.0, liquid_threshold=1, surface_dilation=0)
    idx = lambda i, j, k: i + n * (j + n * k)
    occ = np.zeros(n ** 3, np.uint32)
    occ[idx(2, 2, 2)] = 1
    ct, nliq = orc.classify(P, occ)
    assert ct[idx(0, 3, 3)] == 0 and ct[idx(5, 3, 3)] == 0 and ct[idx(3, 0, 3)] == 0
    assert ct[idx(2, 2, 2)] == 1 and ct[idx(3, 3, 3)] == 2
    assert nliq == 1


def test_particles_mark_their_containing_cells(orc):  # src/sim/mod.rs:278-288
    n = 8
    P = orc.make_params(n, n, n, 1.0, liquid_threshold=1, surface_dilation=0)
    pos = np.array([[2.5, 3.5, 4.5], [5.5, 5.5, 5.5]], np.float32)
    occ = orc.histogram(P, pos)
    idx = lambda i, j, k: i + n * (j + n * k)
    assert occ[idx(2, 3, 4)] == 1 and occ[idx(5, 5, 5)] == 1 and occ.sum() == 2
    ct, _ = orc.classify(P, occ)
    assert ct[idx(2, 3, 4)] == 1 and ct[idx(5, 5, 5)] == 1


def _wall_zeroed_v(n, value):
    """fill_wall_zeroed_faces for V (src/sim/mod.rs:427-442): faces passing the
    static-solid test keep `value`."""
    out = np.zeros(n * (n + 1) * n, np.float32)
    for k in range(n):
        for j in range(n + 1):
            for i in range(n):
                solid = i == 0 or i >= n - 1 or j <= 1 or j >= n - 1 or k == 0 or k >= n - 1
                if not solid:
                    out[i + n * (j + (n + 1) * k)] = value
    return out


@pytest.mark.parametrize("p,vals,old,fixed,tol", [
    # src/sim/mod.rs:527-555
    ((1.05, 3.5, 3.5), (10.0, 7.0), (5.5, 3.85), (10.0, 7.0), 1e-6),
    # src/sim/mod.rs:557-585
    ((3.5, 6.95, 3.5), (-10.0, -6.0), (-0.5, -0.3), (-10.0, -6.0), 1e-5),
])
def test_wall_aware_mac_gather(orc, p, vals, old, fixed, tol):
    n = 8
    P = orc.make_params(n, n, n, 1.0)
    fin = _wall_zeroed_v(n, vals[0])
    sav = _wall_zeroed_v(n, vals[1])
    o = orc.sample_axis(P, 1, p, fin, sav, wall_aware=False)
    f = orc.sample_axis(P, 1, p, fin, sav, wall_aware=True)
    assert abs(o[0] - old[0]) < tol and abs(o[1] - old[1]) < tol
    assert abs(f[0] - fixed[0]) < 1e-6 and abs(f[1] - fixed[1]) < 1e-6


# ---- src/sim/pressure.rs tests ----
def _divergent_tank(orc, reduce_mode=1, **kw):
    """divergent_16cubed_tank (src/sim/pressure.rs:462-505): 16^3, h = rho = dt
    = 1, interior liquid except the top interior layer (air), LCG(12345) faces
    in [-1,1), solid-touching faces zeroed."""
    n = 16
    P = orc.make_params(n, n, n, 1.0, dt=1.0, rho=1.0, stiff=0.0, liquid_threshold=1, surface_dilation=0,
                        reduce_mode=reduce_mode, warm_start=kw.pop("warm_start", 0), **kw)
    idx = lambda i, j, k: i + n * (j + n * k)
    occ = np.zeros(n ** 3, np.uint32)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                if j != n - 2:
                    occ[idx(i, j, k)] = 1
    ct, _ = orc.classify(P, occ)
    assert ct[idx(8, n - 2, 8)] == 2 and ct[idx(8, n - 3, 8)] == 1
    g = lcg_stream(12345)
    vel = []
    for a in range(3):
        cnt = orc.face_count(P, a)
        vel.append(np.array([next(g) * np.float32(2) - np.float32(1) for _ in range(cnt)], np.float32))
    vel = zero_solid_faces(n, ct, vel)
    return P, ct, vel


def zero_solid_faces(n, ct, vel):
    """src/sim/pressure.rs:585-627."""
    vel = [v.copy() for v in vel]
    cidx = lambda i, j, k: i + n * (j + n * k)
    for a in range(3):
        d = [n, n, n]
        d[a] += 1
        for k in range(d[2]):
            for j in range(d[1]):
                for i in range(d[0]):
                    comp = (i, j, k)[a]
                    lo = [i, j, k]
                    lo[a] -= 1
                    lo_solid = comp == 0 or ct[cidx(*lo)] == 0
                    hi_solid = comp == n or ct[cidx(i, j, k)] == 0
                    if lo_solid or hi_solid:
                        vel[a][i + d[0] * (j + d[1] * k)] = 0
    return vel


def l2_div(div, ct):
    m = ct == 1
    return float(np.sqrt(np.sum(div[m].astype(np.float64) ** 2)))


def maxabs_div(div, ct):
    return float(np.max(np.abs(div[ct == 1])))


def test_jacobi_reduces_divergence_16cubed(orc):  # src/sim/pressure.rs:513-580
    P, ct, vel = _divergent_tank(orc)
    div0 = orc.divergence(P, vel, ct)
    l2b, mb = l2_div(div0, ct), maxabs_div(div0, ct)
    assert l2b > 1.0
    p = orc.jacobi_solve(P, ct, div0, 200)
    v1 = orc.gradient_enforce(P, p, ct, vel)
    div1 = orc.divergence(P, v1, ct)
    assert l2_div(div1, ct) < 0.1 * l2b
    assert maxabs_div(div1, ct) < mb


@pytest.mark.parametrize("reduce_mode", [1, 0])
def test_cg_beats_jacobi_16cubed(orc, reduce_mode):  # src/sim/pressure.rs:629-698
    P, ct, vel = _divergent_tank(orc, reduce_mode=reduce_mode, pressure_iters=40, tol=0.0)
    div0 = orc.divergence(P, vel, ct)
    l2b = l2_div(div0, ct)
    pj = orc.jacobi_solve(P, ct, div0, 200)
    l2j = l2_div(orc.divergence(P, orc.gradient_enforce(P, pj, ct, vel), ct), ct)
    pc, st = orc.pressure_solve(P, ct, div0)
    assert st.cg_iterations == 40
    l2c = l2_div(orc.divergence(P, orc.gradient_enforce(P, pc, ct, vel), ct), ct)
    assert l2c < 0.01 * l2b
    assert l2c <= l2j


@pytest.mark.parametrize("reduce_mode", [1, 0])
def test_cg_zero_initial_guess_matches_fixed_iteration_output(orc, reduce_mode):  # :700-721
    P, ct, vel = _divergent_tank(orc, reduce_mode=reduce_mode, pressure_iters=40, tol=0.0)
    div0 = orc.divergence(P, vel, ct)
    fixed, _ = orc.pressure_solve(P, ct, div0)
    Pw = orc.make_params(16, 16, 16, 1.0, dt=1.0, rho=1.0, stiff=0.0, reduce_mode=reduce_mode,
                         warm_start=1, pressure_iters=40, tol=0.0)
    zero_init, st = orc.pressure_solve(Pw, ct, div0, p0=np.zeros(16 ** 3, np.float32))
    assert st.cg_iterations == 40
    assert np.array_equal(zero_init.view(np.uint32), fixed.view(np.uint32))
    assert st.rs_initial > st.rs_final


def test_cg_warm_initial_guess_reaches_lower_residual_with_fewer_iterations(orc):  # :723-763
    P, ct, vel = _divergent_tank(orc, reduce_mode=1, pressure_iters=12, tol=0.0)
    div0 = orc.divergence(P, vel, ct)
    zero_ref, zst = orc.pressure_solve(P, ct, div0)
    Pw = orc.make_params(16, 16, 16, 1.0, dt=1.0, rho=1.0, stiff=0.0, reduce_mode=1, warm_start=1,
                         pressure_iters=40, tol=1.0, clamp_tol=False)
    _, wst = orc.pressure_solve(Pw, ct, div0, p0=zero_ref)
    assert wst.cg_iterations < zst.cg_iterations
    assert wst.rs_final <= zst.rs_final
    assert wst.rs_initial < zst.rs_initial


def test_divergence_free_field_has_zero_divergence(orc):  # src/sim/pressure.rs:765-783
    n = 8
    P = orc.make_params(n, n, n, 0.5, stiff=0.0)
    occ = np.zeros(n ** 3, np.uint32)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                occ[i + n * (j + n * k)] = 1
    ct, _ = orc.classify(P, occ)
    vel = [np.full(orc.face_count(P, 0), 3.0, np.float32), np.full(orc.face_count(P, 1), -2.0, np.float32),
           np.full(orc.face_count(P, 2), 1.5, np.float32)]
    div = orc.divergence(P, vel, ct)
    assert maxabs_div(div, ct) < 1e-5
