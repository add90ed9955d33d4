"""fs_scene_seed_noalloc: the seeding of fs_scene_seed with the lattice staged
in the library's fixed per-GPU array instead of a hipMalloc'd buffer.  Its
output must equal fs_scene_seed's bit for bit, on two calls in a row (the
second lattice smaller, so a stale tail of the array must not show), and a
lattice above the array's size is rejected before any GPU work."""
import sys

import numpy as np
import pytest

import fluidstep as fsm
from helpers import bits_equal

pytestmark = pytest.mark.gpu
H = np.float32(2.0 / 64.0)
FS_ERR_INVALID_ARGUMENT = -1
MAX_PARTICLES = (128 * 1024 * 1024) // 12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def registry(preset=0, res=(80, 40, 80), fill=20.0, drop=1.0, density=10.0):
    r = fsm.Registry()
    for k, v in (("scene.preset", preset), ("grid.res_x", res[0]), ("grid.res_y", res[1]), ("grid.res_z", res[2]),
                 ("scene.fill_level", fill), ("scene.drop_height", drop), ("particles.density", density)):
        assert r.set(k, v)["ok"]
    return r


def centred(res):
    return [-(np.float32(n) * H) / np.float32(2.0) for n in res]


def test_noalloc_seed_equals_the_allocating_seed_twice_then_smaller():
    cases = [
        dict(preset=1, res=(128, 96, 128), fill=60.0, density=12.0),  # 3,664,116 particles
        dict(),                                                         # the default scene, 494,384
    ]
    counts = []
    for case in cases:
        sc = fsm.scene_from_registry(registry(**case))
        origin = centred(tuple(sc.res))
        say("noalloc seed", case, "lattice", sc.generated_count)
        got = fsm.scene_seed_noalloc(sc, H, origin)
        ref = fsm.scene_seed(sc, H, origin)
        assert len(got) == len(ref) == sc.generated_count
        assert bits_equal(got, ref)
        counts.append(len(got))
    assert counts[0] > counts[1] > 0 and counts[0] <= MAX_PARTICLES


def test_noalloc_seed_rejects_a_lattice_above_the_array():
    sc = fsm.scene_from_registry(registry(preset=0, res=(128, 128, 128), fill=50.0, density=16.0))
    origin = centred(tuple(sc.res))
    n = fsm.scene_seed(sc, H, origin, want=False)
    assert n > MAX_PARTICLES
    out = np.zeros((n, 3), np.float32)
    k = fsm.C.c_uint32()
    o = (fsm.C.c_float * 3)(*origin)
    rc = fsm.lib().fs_scene_seed_noalloc(fsm.C.byref(sc), H, o, fsm._ptr(out), n, fsm.C.byref(k))
    assert rc == FS_ERR_INVALID_ARGUMENT and k.value == n
    assert b"staging array" in fsm.lib().fs_last_error()
    assert not np.any(out)
