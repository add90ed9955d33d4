"""Error reporting of the persistent pressure solve.

A solve whose grid barrier times out sets a sticky failure word that later
successful solves do not clear: fs_synchronize reports it once (FS_ERR_SOLVER)
and clears it.  The failure is forced from the host by shortening the
kernel's own bounded waits (FLUIDSTEP_CG_SPIN_TICKS) with the counter form of
the memory-ordering barrier (FLUIDSTEP_CG_DATA_BARRIER=counter, whose wait
tests its limit on the first poll): a workgroup that is not the last to
arrive gives up at once and raises the abort flag, and every workgroup
leaves the kernel, so the GPU never waits on a barrier that cannot
complete.  The 64^3 tank solves on more than one workgroup."""
import sys

import numpy as np
import pytest

import fluidstep as fsm
import scenes
from helpers import gpu_params

pytestmark = pytest.mark.gpu
FS_ERR_SOLVER = -6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_solver_failure_is_sticky_across_later_solves(monkeypatch):
    sc = scenes.control_64()
    monkeypatch.setenv("FLUIDSTEP_CG_DATA_BARRIER", "counter")
    t = fsm.FluidStep(gpu_params(sc, fsm.ENGINE_FUSED), sc["pos"], sc["vel"])
    t.step(1)
    t.synchronize()  # a normal solve reports nothing
    monkeypatch.setenv("FLUIDSTEP_CG_SPIN_TICKS", "1")
    say("forced solver timeout: one substep with 1-tick waits, then two normal substeps")
    t.step(1)
    monkeypatch.delenv("FLUIDSTEP_CG_SPIN_TICKS")
    t.step(2)
    with pytest.raises(fsm.FluidStepError) as e:
        t.synchronize()
    assert e.value.code == FS_ERR_SOLVER
    t.synchronize()  # reported once, then cleared
    st = t.stats()  # the last solve ran normally
    assert st["cg_iterations"] > 0 and np.isfinite(st["rs_final"])
    pos, vel = t.particles()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    # a failing solve that is the last one: stats report it too
    monkeypatch.setenv("FLUIDSTEP_CG_SPIN_TICKS", "1")
    t.step(1)
    monkeypatch.delenv("FLUIDSTEP_CG_SPIN_TICKS")
    with pytest.raises(fsm.FluidStepError) as e:
        t.stats()
    assert e.value.code == FS_ERR_SOLVER
    t.reset()  # clears the sticky word
    t.step(1)
    t.synchronize()
    t.close()
