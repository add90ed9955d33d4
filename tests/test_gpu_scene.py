"""Scene seeding on the GPU (fs_scene_seed, fs_create_from_settings) against
the CPU restatement of generate_particles (src/gpu/fluid.rs:1838-1913):
bit-exact positions; a tank created from the registry holds exactly those
particles at rest, with the Auto rest density and surface dilation of its
scene, and its first step matches the restatement's step."""
import sys

import numpy as np
import pytest

import fluidstep as fsm
import oracle
from helpers import bits_equal, rel_l2

pytestmark = pytest.mark.gpu
H = np.float32(2.0 / 64.0)
TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def registry(preset=0, res=(80, 40, 80), fill=20.0, drop=1.0, density=10.0, override=0, **extra):
    r = fsm.Registry()
    for k, v in (("scene.preset", preset), ("grid.res_x", res[0]), ("grid.res_y", res[1]), ("grid.res_z", res[2]),
                 ("scene.fill_level", fill), ("scene.drop_height", drop), ("particles.density", density),
                 ("particles.count", override), *extra.items()):
        assert r.set(k, v)["ok"]
    return r


def centred(res, h=H):
    return [-(np.float32(n) * np.float32(h)) / np.float32(2.0) for n in res]


@pytest.mark.parametrize("case", [
    dict(),  # the reference's default scene: falling blob, 80x40x80, 10 per seeded cell
    dict(preset=1, fill=50.0, density=8.0),
    dict(preset=2, fill=35.0, drop=0.3, density=6.0),
    dict(preset=2, fill=100.0, drop=0.0, res=(96, 24, 112), density=3.0),
    dict(preset=0, fill=0.0, density=1.0),
    dict(preset=1, fill=80.0, override=123_457, res=(16, 64, 16)),
])
def test_seed_is_bit_exact(case):
    r = registry(**case)
    sc = fsm.scene_from_registry(r)
    origin = centred(sc.res)
    say("seed", case, "target", sc.particle_count, "lattice", sc.generated_count)
    pos = fsm.scene_seed(sc, H, origin)
    ref = oracle.scene_generate(sc.blocks_array(), sc.particle_count, H, origin, tuple(sc.res))
    assert len(pos) == sc.generated_count == len(ref)
    assert bits_equal(pos, ref)


def test_seed_other_cell_size_and_origin():
    sc = fsm.scene_from_registry(registry(preset=2, fill=60.0, drop=0.5, density=12.0, res=(64, 48, 80)))
    for h, origin in ((np.float32(1.0 / 64.0), (0.0, 0.0, 0.0)), (np.float32(0.037), (3.5, -2.25, 0.125))):
        say("seed h", h, "origin", origin)
        pos = fsm.scene_seed(sc, h, origin)
        ref = oracle.scene_generate(sc.blocks_array(), sc.particle_count, h, origin, tuple(sc.res))
        assert bits_equal(pos, ref)


def test_seed_large_lattice():
    """128^3 at 16 per seeded cell, half fill: 14.5M particles (the lattice
    floors the 16.8M target), i.e. LCG states up to ~4.3e7 reached by each
    thread's jump."""
    sc = fsm.scene_from_registry(registry(preset=0, res=(128, 128, 128), fill=50.0, density=16.0))
    say("seed large", sc.generated_count)
    pos = fsm.scene_seed(sc, H, centred(sc.res))
    ref = oracle.scene_generate(sc.blocks_array(), sc.particle_count, H, centred(sc.res), tuple(sc.res))
    assert len(pos) == len(ref) > 14_000_000
    assert bits_equal(pos, ref)


def oracle_params_of(p):
    """The restatement's Params for a tank's fs_params (reduce mode of the
    fused engine)."""
    return oracle.make_params(
        p.nx, p.ny, p.nz, p.h, origin=tuple(p.origin), dt=p.dt, rho=p.rho, flip_blend=p.flip_blend,
        wall_friction=p.wall_friction, grav=tuple(p.gravity), rest=p.rest_density, stiff=p.volume_stiffness,
        drift_clamp=p.drift_clamp, liquid_threshold=p.liquid_threshold, surface_dilation=p.surface_dilation,
        cfl=p.cfl, pressure_iters=p.pressure_iterations, warm_start=p.pressure_warm_start,
        tol=p.pressure_residual_tolerance, reduce_mode=1)


@pytest.mark.parametrize("case,dilation", [(dict(), 0), (dict(preset=1, fill=40.0, density=4.0), 1)])
def test_create_from_settings(case, dilation):
    r = registry(**case)
    sc = fsm.scene_from_registry(r)
    say("create from settings", case)
    t = fsm.FluidStep.from_settings(r, sc)
    p = t.params
    origin = centred((p.nx, p.ny, p.nz))
    ref_pos = oracle.scene_generate(sc.blocks_array(), sc.particle_count, H, origin, (p.nx, p.ny, p.nz))
    pos, vel = t.particles()
    assert t.n == sc.generated_count and bits_equal(pos, ref_pos) and not np.any(vel)
    # Auto rest density = the scene's effective density; Auto dilation below 8 per cell.
    assert np.float32(p.rest_density) == np.float32(sc.effective_density)
    assert p.surface_dilation == dilation and p.h == H and list(p.origin) == [float(o) for o in origin]
    t.step(1)
    t.synchronize()
    ref = oracle.run(oracle_params_of(p), ref_pos, np.zeros_like(ref_pos), steps=1, sort_period=1)
    gp, gv = t.particles()
    assert rel_l2(gp, ref.pos) <= TOL and rel_l2(gv, ref.vel_particles) <= TOL
    t.reset()
    assert bits_equal(t.particles()[0], ref_pos)
    t.close()
