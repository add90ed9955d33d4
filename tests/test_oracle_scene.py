"""Scene presets, spawn counts and the initial lattice: the CPU restatement
(oracle/flip_oracle_scene.c) and the product's host functions
(fs_scene_preset_blocks / fs_scene_from_registry / fs_scene_seed's count),
against the reference's own scene tests (src/scene/mod.rs:470-1000), exact
integer arithmetic for the generator, and each other bit for bit.  No GPU.
"""
import itertools

import numpy as np
import pytest

import fluidstep as fsm
import oracle

H = np.float32(2.0 / 64.0)
PRESETS = (fsm.SCENE_FALLING_BLOB, fsm.SCENE_DAM_BREAK, fsm.SCENE_DOUBLE_SPLASH)
DEFAULT_DROP = 0.72


class OracleImpl:
    name = "oracle"

    def blocks(self, preset, drop, fill):
        return oracle.scene_preset_blocks(preset, drop, fill)

    def counts(self, preset, res, fill, density, override=0, drop=1.0):
        """(blocks, resolved count, effective density, generated count) of
        SceneConfig::from_settings for these registry values."""
        b = self.blocks(preset, drop, np.float32(np.float32(fill * 100.0) / np.float32(100.0)))
        c = oracle.scene_resolved_count(res, b, density, override)
        eff = oracle.scene_effective_density(res, b, c)
        origin = [-(np.float32(r) * H) / np.float32(2.0) for r in res]
        gen = oracle.scene_generate(b, c, H, origin, res, want=False)
        return b, c, eff, gen


class ProductImpl:
    name = "product"

    def blocks(self, preset, drop, fill):
        return fsm.scene_preset_blocks(preset, drop, fill)

    def counts(self, preset, res, fill, density, override=0, drop=1.0):
        r = fsm.Registry()
        r.set("scene.preset", preset)
        r.set("grid.res_x", res[0])
        r.set("grid.res_y", res[1])
        r.set("grid.res_z", res[2])
        r.set("scene.fill_level", fill * 100.0)
        r.set("scene.drop_height", drop)
        r.set("particles.density", density)
        r.set("particles.count", override)
        sc = fsm.scene_from_registry(r)
        return sc.blocks_array(), sc.particle_count, np.float32(sc.effective_density), sc.generated_count


IMPLS = [OracleImpl(), ProductImpl()]


def frac(b):
    return float(oracle.scene_seeded_fraction(b))


def lcg_draws(n):
    """The generator of generate_particles in exact integers: rand01 of
    draws 1..n from seed 0x12345678."""
    st = 0x12345678
    out = np.empty(n, np.float64)
    for i in range(n):
        st = (st * 1664525 + 1013904223) % (1 << 32)
        out[i] = (st >> 8) / float(1 << 24)
    return out


# ---- the reference's scene tests (src/scene/mod.rs), on both implementations ----

@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_default_fill_level_is_falling_blob(impl):  # :470-486
    b = impl.blocks(fsm.SCENE_FALLING_BLOB, DEFAULT_DROP, 0.2)[0]
    assert b[1] > 0.0 and b[4] < 1.0
    assert abs(frac(b[None]) - 0.2) < 0.01


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_falling_blob_uses_drop_height(impl):  # :488-499
    low = impl.blocks(fsm.SCENE_FALLING_BLOB, 0.0, 0.2)[0]
    high = impl.blocks(fsm.SCENE_FALLING_BLOB, 1.0, 0.2)[0]
    assert abs(low[1]) < 1e-6
    assert high[1] > low[1] + 0.25
    assert abs(high[4] - 1.0) < 1e-6


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_double_splash_drop_height_clamps_inside_tank(impl):  # :501-515
    for drop in (0.0, 1.0):
        b = impl.blocks(fsm.SCENE_DOUBLE_SPLASH, drop, 0.2)[0]
        assert b[1] >= -1e-6 and b[4] <= 1.0 + 1e-6 and b[4] > b[1]


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_default_density_derives_count_from_seeded_cells(impl):  # :517-527
    _, c, _, _ = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.2, 10.0)
    assert 511_000 <= c <= 513_000


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_density_scales_count_with_grid_resolution(impl):  # :529-540
    _, c, _, _ = impl.counts(fsm.SCENE_FALLING_BLOB, (128, 40, 128), 0.2, 10.0)
    assert 1_305_000 <= c <= 1_315_000


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_nonzero_particle_count_override_wins_over_density(impl):  # :542-550
    _, c, _, _ = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.2, 10.0, override=500_000)
    assert c == 500_000


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_dam_break_ignores_drop_height(impl):  # :552-561
    low = impl.blocks(fsm.SCENE_DAM_BREAK, 0.0, 0.2)[0]
    high = impl.blocks(fsm.SCENE_DAM_BREAK, 1.0, 0.2)[0]
    assert abs(low[1]) < 1e-6 and abs(high[1] - low[1]) < 1e-6 and abs(high[4] - low[4]) < 1e-6


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_full_fill_uses_preset_authored_max_not_whole_tank(impl):  # :563-572
    b, _, _, _ = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 1.0, 10.0)
    assert 0.55 <= frac(b) <= 0.75


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_preset_seeded_fraction_keeps_authored_scale(impl):  # :605-621
    cases = [(fsm.SCENE_FALLING_BLOB, 0.2, 0.199888), (fsm.SCENE_FALLING_BLOB, 0.5, 0.499720),
             (fsm.SCENE_DAM_BREAK, 0.5, 0.163170), (fsm.SCENE_DOUBLE_SPLASH, 0.5, 0.173280)]
    for preset, fill, expected in cases:
        assert abs(frac(impl.blocks(preset, DEFAULT_DROP, fill)) - expected) < 0.005


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_zero_fill_keeps_minimum_viable_seed(impl):  # :623-637
    for preset in PRESETS:
        b = impl.blocks(preset, DEFAULT_DROP, 0.0)
        assert len(b) > 0 and frac(b) > 0.0
        _, c, _, _ = impl.counts(preset, (80, 40, 80), 0.0, 8.0, drop=DEFAULT_DROP)
        assert c == 1024


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_fill_level_is_monotone_in_seeded_fraction_and_count(impl):  # :639-663
    for preset in PRESETS:
        last_f, last_c = -1.0, 0
        for fill in (0.1, 0.2, 0.5, 1.0):
            b, c, _, _ = impl.counts(preset, (64, 64, 64), fill, 8.0, drop=DEFAULT_DROP)
            assert frac(b) > last_f + 1e-6 and c >= last_c
            last_f, last_c = frac(b), c


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_default_scene_fill_is_roughly_linear_before_clamping(impl):  # :665-683
    f20 = frac(impl.blocks(fsm.SCENE_FALLING_BLOB, DEFAULT_DROP, 0.2))
    f50 = frac(impl.blocks(fsm.SCENE_FALLING_BLOB, DEFAULT_DROP, 0.5))
    assert abs(f20 - 0.2) < 0.001 and 2.4 <= f50 / f20 <= 2.6


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_count_scales_with_density_at_fixed_fill_level(impl):  # :685-711
    for preset in PRESETS:
        _, c8, _, _ = impl.counts(preset, (64, 64, 64), 0.5, 8.0, drop=DEFAULT_DROP)
        _, c2, _, _ = impl.counts(preset, (64, 64, 64), 0.5, 2.0, drop=DEFAULT_DROP)
        assert 3.6 <= c8 / c2 <= 4.4


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_requested_effective_density_tracks_slider_density(impl):  # :738-753
    for d in (1.0, 2.0, 4.0, 8.0, 16.0):
        _, _, eff, _ = impl.counts(fsm.SCENE_DAM_BREAK, (64, 64, 64), 0.5, d, drop=DEFAULT_DROP)
        assert abs(eff - d) / d < 0.02


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_actual_effective_density_uses_generated_count(impl):  # :755-771
    b, c, eff, gen = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 8.0)
    actual = gen / (80 * 40 * 80 * float(frac(b)))
    assert abs(eff - 8.0) < 0.001 and actual < eff and abs(actual - 7.920) < 0.02


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_fill_level_clamps_keep_blocks_valid(impl):  # :806-820
    for preset, fill in itertools.product(PRESETS, (0.0, 0.1, 1.0)):
        for b in impl.blocks(preset, DEFAULT_DROP, fill):
            assert b[4] > b[1] and b[1] >= -1e-6 and b[4] <= 1.0 + 1e-6


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_measurement_matrix_bounds(impl):  # :852-903
    cases = [(fsm.SCENE_FALLING_BLOB, (32, 32, 32), 0.2, 8.0), (fsm.SCENE_FALLING_BLOB, (64, 64, 64), 0.5, 8.0),
             (fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 1.0), (fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 8.0),
             (fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 32.0), (fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.8, 8.0),
             (fsm.SCENE_DAM_BREAK, (80, 40, 80), 0.5, 8.0), (fsm.SCENE_DAM_BREAK, (80, 40, 80), 1.0, 8.0),
             (fsm.SCENE_DOUBLE_SPLASH, (80, 40, 80), 0.5, 8.0), (fsm.SCENE_FALLING_BLOB, (96, 24, 96), 0.8, 8.0)]
    for preset, res, fill, d in cases:
        b, c, eff, gen = impl.counts(preset, res, fill, d)
        actual = gen / (res[0] * res[1] * res[2] * float(frac(b)))
        assert frac(b) > 0 and c >= 1024 and 0 < gen <= c and actual <= eff + 1e-6


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_density_changes_count_not_seed_geometry(impl):  # :905-936
    low = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 1.0)
    ref = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 8.0)
    high = impl.counts(fsm.SCENE_FALLING_BLOB, (80, 40, 80), 0.5, 32.0)
    assert np.array_equal(low[0], ref[0]) and np.array_equal(high[0], ref[0])
    assert low[3] < ref[3] < high[3]
    cells = 80 * 40 * 80 * float(frac(ref[0]))
    for (b, c, eff, gen), d, tol in ((low, 1.0, 0.10), (ref, 8.0, 0.05), (high, 32.0, 0.05)):
        assert abs(gen / cells - d) / d < tol


@pytest.mark.parametrize("impl", IMPLS, ids=lambda i: i.name)
def test_behavior_formula_supports_preset_authored_amount(impl):  # :938-957 (registry drop height 1.0)
    for preset, expected in ((fsm.SCENE_FALLING_BLOB, 0.434), (fsm.SCENE_DAM_BREAK, 0.163),
                             (fsm.SCENE_DOUBLE_SPLASH, 0.158)):
        b, _, _, _ = impl.counts(preset, (80, 40, 80), 0.5, 8.0)
        assert abs(frac(b) - expected) < 0.005


# ---- the generator ----

@pytest.mark.parametrize("preset,res,fill,h,origin", [
    (fsm.SCENE_DAM_BREAK, (32, 32, 32), 0.5, 1.0 / 32.0, (0.0, 0.0, 0.0)),
    (fsm.SCENE_DOUBLE_SPLASH, (40, 24, 40), 0.3, 2.0 / 64.0, (-0.625, -0.375, -0.625)),
])
def test_generator_draws_are_the_lcg_stream_in_lattice_order(preset, res, fill, h, origin):
    """Recover each particle's three draws from its position, (p - site) /
    (spacing / 2) + 1/2 at lattice site wmin + (i + 1/2) spacing, and compare
    them with the generator computed in exact integers: this fixes the seed,
    the constants, three draws per particle in x, y, z order, the lattice
    order (ix outermost, iz innermost, blocks in order) and the jitter scale.
    Positions on the clamp box are skipped."""
    b = oracle.scene_preset_blocks(preset, DEFAULT_DROP, fill)
    count = oracle.scene_resolved_count(res, b, 8.0)
    pos, info = oracle.scene_generate(b, count, h, origin, res, info=True)
    draws = lcg_draws(3 * len(pos)).reshape(-1, 3)
    o = np.asarray(origin, np.float32)
    ext = np.asarray(res, np.float32) * np.float32(h)
    lo, hi = o + np.float32(h) * np.float32(1.05), o + ext - np.float32(h) * np.float32(1.05)
    g = 0
    checked = 0
    for blk in info:
        wmin, sp, cnt = blk[:3].astype(np.float64), float(blk[3]), blk[4:].astype(np.int64)
        n = int(np.prod(cnt))
        ix, iy, iz = np.unravel_index(np.arange(n), tuple(cnt))
        site = wmin + (np.stack([ix, iy, iz], 1) + 0.5) * sp
        p = pos[g:g + n].astype(np.float64)
        r = (p - site) / (sp / 2.0) + 0.5
        free = np.all((pos[g:g + n] > lo) & (pos[g:g + n] < hi), axis=1)
        assert np.max(np.abs(r[free] - draws[g:g + n][free])) < 1e-3
        checked += int(free.sum())
        g += n
    assert g == len(pos) and checked > 0.9 * len(pos)


def test_generator_lattice_counts_and_clamp_box():
    """The lattice holds floor(extent / spacing) points per axis (at least 1);
    all positions lie in the clamp box; a quarter spacing bounds the jitter."""
    for preset, fill in itertools.product(PRESETS, (0.0, 0.2, 1.0)):
        res = (48, 32, 40)
        b = oracle.scene_preset_blocks(preset, DEFAULT_DROP, fill)
        count = oracle.scene_resolved_count(res, b, 6.0)
        origin = np.array([-0.75, -0.5, -0.625], np.float32)
        pos, info = oracle.scene_generate(b, count, H, origin, res, info=True)
        ext = np.asarray(res, np.float32) * H
        lo, hi = origin + H * np.float32(1.05), origin + ext - H * np.float32(1.05)
        assert np.all(pos >= lo) and np.all(pos <= hi)
        g = 0
        for k, blk in enumerate(info):
            bext = (b[k, 3:] - b[k, :3]) * ext
            expect = np.maximum(np.floor(bext / blk[3]), 1)
            # floor at the ulp of the f32 quotient may differ from the f64 one
            assert np.all(np.abs(blk[4:] - expect) <= 1)
            n = int(np.prod(blk[4:]))
            ix, iy, iz = np.unravel_index(np.arange(n), tuple(blk[4:].astype(int)))
            site = blk[:3] + (np.stack([ix, iy, iz], 1) + 0.5) * blk[3]
            d = np.abs(pos[g:g + n] - site)
            inside = np.all((pos[g:g + n] > lo) & (pos[g:g + n] < hi), axis=1)
            assert np.all(d[inside] <= blk[3] / 4 * (1 + 1e-5))
            g += n
        assert g == len(pos)


# ---- product host functions equal the restatement bit for bit ----

def test_product_blocks_equal_oracle():
    for preset, drop, fill in itertools.product((0, 1, 2, 7), (0.0, 0.3, 0.72, 1.0),
                                                (0.0, 0.05, 0.2, 0.37, 0.5, 0.8, 1.0)):
        a = fsm.scene_preset_blocks(preset, drop, fill)
        b = oracle.scene_preset_blocks(preset, drop, fill)
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), (preset, drop, fill)


def test_product_counts_equal_oracle():
    P, O = ProductImpl(), OracleImpl()
    for preset, res, fill, d, ov, drop in itertools.product(
            PRESETS, ((80, 40, 80), (128, 128, 128), (16, 16, 16), (96, 24, 112)), (0.0, 0.2, 0.55, 1.0),
            (1.0, 7.5, 10.0, 32.0), (0, 123_457), (0.0, 1.0)):
        a = P.counts(preset, res, fill, d, ov, drop)
        b = O.counts(preset, res, fill, d, ov, drop)
        assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[3] == b[3], (preset, res, fill, d, ov, drop)
        assert np.float32(a[2]).view(np.uint32) == np.float32(b[2]).view(np.uint32)


def test_product_seed_count_query_needs_no_gpu():
    r = fsm.Registry()
    sc = fsm.scene_from_registry(r)
    for h, origin in ((H, (-1.25, -0.625, -1.25)), (np.float32(0.01), (0.0, 0.0, 0.0))):
        k = fsm.scene_seed(sc, h, origin, want=False)
        assert k == oracle.scene_generate(sc.blocks_array(), sc.particle_count, h, origin, tuple(sc.res), want=False)


def test_registry_has_the_scene_rows():  # src/settings/mod.rs:407-496
    r = fsm.Registry()
    assert r.get("scene.preset") == 0 and r.get("scene.fill_level") == 20.0 and r.get("scene.drop_height") == 1.0
    assert r.get("particles.count") == 0 and r.get("particles.density") == 10.0
    res = r.set("scene.preset", 5)
    assert res["clamped"] and r.get("scene.preset") == 2 and res["needs_reset"]
    assert r.set("particles.count", 2e9)["clamped"] and r.get("particles.count") == 134_217_728
