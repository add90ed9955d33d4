"""Shared helpers of the parity tests: the same workload on the HIP step and on
the CPU restatement (oracle/), and the comparison metrics."""
import numpy as np

import fluidstep
import oracle


def gpu_params(scene, engine=fluidstep.ENGINE_FUSED, precision=None, velocity=fluidstep.VELOCITY_F32, **over):
    n = scene["n"]
    ph = dict(scene["physics"])
    ph.update(over)
    h = ph.pop("h")
    if precision is None:
        precision = fluidstep.PRECISION_MIXED if engine == fluidstep.ENGINE_FUSED else fluidstep.PRECISION_F32
    return fluidstep.make_params(n, n, n, float(h), engine=engine, precision=precision,
                                 velocity_storage=velocity, **ph)


def oracle_params(scene, reduce_mode, **over):
    n = scene["n"]
    ph = dict(scene["physics"])
    ph.update(over)
    return oracle.make_params(
        n, n, n, float(ph["h"]), dt=ph["dt"], rho=ph["rho"], flip_blend=ph["flip_blend"],
        wall_friction=ph["wall_friction"], grav=ph["gravity"], rest=ph["rest_density"],
        stiff=ph["volume_stiffness"], drift_clamp=ph["drift_clamp"],
        liquid_threshold=ph["liquid_threshold"], surface_dilation=ph["surface_dilation"], cfl=ph["cfl"],
        pressure_iters=ph["pressure_iterations"], warm_start=ph["pressure_warm_start"],
        tol=ph["pressure_residual_tolerance"], reduce_mode=reduce_mode)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return num / den if den > 0 else num


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
