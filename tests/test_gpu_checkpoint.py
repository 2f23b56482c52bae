"""Restart from a checkpoint continues a run bit for bit (FP64 and mixed precision): 2 steps,
Checkpoint_Output, Checkpoint_Input, 2 more steps == 4 steps in one go."""
import numpy as np
import pytest
import torch

from cssx import checkpoint as ck
from cssx.model import Config, Model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_restart_is_bitwise(tmp_path, precision):
    from cssx.mhd import Solver
    print(f"checkpoint restart {precision}", flush=True)
    c = Config(nr=24, nth=32, nphi=16, sld_coef=0.1)
    m = Model(c)
    sol = Solver(m)
    y0 = torch.as_tensor(m.initial_state(), device="cuda")
    dt = sol.cfl(y0)
    straight = y0.clone()
    sol.rk4(straight, dt, 4, precision)
    part = y0.clone()
    sol.rk4(part, dt, 2, precision)
    d = ck.write(str(tmp_path) + "/Checkpoints/", 2, part, deltat=dt, sim_time=2 * dt, r1=c.r1, r2=c.r2,
                 th1=c.th1, th2=c.th2, phi1=c.phi1, phi2=c.phi2)
    hdr, st = ck.read(d)
    assert hdr.istep == 2 and hdr.nv == 8
    resumed = ck.to_device(st)
    # the header keeps deltat to 7 digits (e14.7), as the reference's restarts do; continue with
    # the exact step to compare bit for bit
    assert abs(hdr.deltat - dt) <= 5e-7 * dt
    sol.rk4(resumed, dt, 2, precision)
    torch.cuda.synchronize()
    assert torch.equal(resumed, straight)
    sol.close()
