"""The live-buffer hand-off (fs_get_device_view, fs_take_sort_swapped,
fs_pack_particles_vec4, fs_pack_cell_types_u32): what a consumer on the GPU
reads through the lent device pointers equals the tank's read-outs bit for
bit; the live side and the swapped flag follow particle_buffer() /
take_sort_swapped() (src/gpu/fluid.rs:1205-1226, reset :1408-1418); the packed
layouts are the reference's 32-byte record and u32 cell types."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

import fluidstep as fsm
import scenes
from helpers import bits_equal, gpu_params

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        fsm.lib()  # loads the HIP runtime the library links
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _hip


def fetch(ptr, count, dtype):
    """Device -> host copy of `count` elements at a lent device pointer."""
    out = np.empty(count, dtype)
    assert hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return out


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tank(velocity=fsm.VELOCITY_F32, period=1):
    sc = scenes.small_tank(n=24, seed=3)
    p = gpu_params(sc, fsm.ENGINE_FUSED, velocity=velocity, particle_sort_period=period)
    return sc, fsm.FluidStep(p, sc["pos"], sc["vel"])


def check_view_matches_readouts(t):
    t.synchronize()
    v = t.device_view()
    n, (nx, ny, nz) = v.n, tuple(v.dims)
    C_ = nx * ny * nz
    pos, vel = t.particles()
    for a in range(3):
        assert bits_equal(fetch(v.pos[a], n, np.float32), pos[:, a].copy())
        if v.velocity_storage == fsm.VELOCITY_BF16:
            raw = fetch(v.vel[a], n, np.uint16).astype(np.uint32) << 16
            assert np.array_equal(raw.view(np.float32), vel[:, a])
        else:
            assert bits_equal(fetch(v.vel[a], n, np.float32), vel[:, a].copy())
    faces = t.faces(fsm.FACE_VELOCITY)
    dims = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
    for a in range(3):
        assert bits_equal(fetch(v.face_vel[a], dims[a], np.float32), faces[a])
    assert bits_equal(fetch(v.pressure, C_, np.float32), t.pressure())
    assert np.array_equal(fetch(v.cell_type, C_, np.uint8).astype(np.uint32), t.cell_type())
    assert np.array_equal(fetch(v.occupancy, C_, np.uint32), t.occupancy())
    assert fetch(v.stats, 1, np.uint32)[0] == t.stats()["liquid_cells"]
    return v


@pytest.mark.parametrize("velocity", [fsm.VELOCITY_F32, fsm.VELOCITY_BF16])
def test_view_equals_readouts_and_follows_the_live_side(velocity):
    sc, t = tank(velocity)
    say("device view", velocity)
    v0 = check_view_matches_readouts(t)
    assert v0.live_side == 0 and v0.n == len(sc["pos"])
    assert not t.take_sort_swapped()
    t.step(1)  # sorts (period 1): the live side flips
    v1 = check_view_matches_readouts(t)
    assert v1.live_side == 1 and v1.pos[0] != v0.pos[0]
    assert t.take_sort_swapped() and not t.take_sort_swapped()
    t.step(1)
    assert check_view_matches_readouts(t).live_side == 0 and t.take_sort_swapped()
    t.step(1)
    assert t.take_sort_swapped()
    t.reset()  # rewrites side 0 and snaps back to it (src/gpu/fluid.rs:1408-1418)
    assert check_view_matches_readouts(t).live_side == 0 and t.take_sort_swapped()
    t.close()


def test_no_flip_without_a_sort():
    sc, t = tank(period=4)
    t.step(1)  # tick 0 sorts
    assert t.take_sort_swapped()
    t.step(3)  # ticks 1-3 do not
    assert not t.take_sort_swapped()
    check_view_matches_readouts(t)
    t.close()


def test_consumer_reads_ordered_on_the_tanks_stream():
    """A read enqueued on the lent stream right after fs_step (no host wait)
    sees the step's result."""
    sc, t = tank()
    t.step(1)
    v = t.device_view()
    n = v.n
    out = np.empty(n, np.float32)
    t.step(2)
    v = t.device_view()
    assert hip().hipMemcpyAsync(out.ctypes.data, v.pos[1], out.nbytes, 2, v.stream) == 0
    assert hip().hipStreamSynchronize(v.stream) == 0
    assert bits_equal(out, t.particles()[0][:, 1].copy())
    t.close()


@pytest.mark.parametrize("velocity", [fsm.VELOCITY_F32, fsm.VELOCITY_BF16])
def test_pack_particles_vec4_is_the_reference_record(velocity):
    sc, t = tank(velocity)
    t.step(2)
    n = t.n
    dst = torch.full((8 * n,), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t.pack_particles_vec4(dst.data_ptr())
    t.synchronize()
    rec = dst.cpu().numpy().reshape(n, 8)
    pos, vel = t.particles()
    exp = np.zeros((n, 8), np.float32)
    exp[:, 0:3], exp[:, 4:7] = pos, vel
    assert bits_equal(rec, exp)
    t.close()


def test_pack_cell_types_u32():
    sc, t = tank()
    t.step(1)
    nx, ny, nz = t.dims
    dst = torch.full((nx * ny * nz,), 7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t.pack_cell_types_u32(dst.data_ptr())
    t.synchronize()
    assert np.array_equal(dst.cpu().numpy().view(np.uint32), t.cell_type())
    t.close()
