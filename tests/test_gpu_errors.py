"""Error paths of the device API through the C ABI: every refusal raises with a message
(uc_last_error) and leaves the context usable; no CPU fallback is taken."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_gpu_refusals_leave_the_context_usable(torch_cuda):
    from unicorn_amd.chem import ChemError, ChemStep, Controls, DeviceCoupling, DeviceField
    from unicorn_amd.fields import flame_field, synthetic_coupling
    f0 = flame_field(21, 13, "CH4", seed=3)
    ctx = ChemStep(f0.n)
    print("error paths", flush=True)
    # no active interior cell: every cell skipped
    dead = f0.copy()
    dead.iskip[:] = 10
    with pytest.raises(ChemError, match="no active interior cell"):
        ctx.step(DeviceField(dead), Controls(dt=f0.dt))
    # a grid larger than the context
    big = flame_field(31, 21, "CH4", seed=3)
    with pytest.raises(ChemError, match="does not fit"):
        ctx.step(DeviceField(big), Controls(dt=big.dt))
    # controls out of range
    with pytest.raises(ChemError, match="sweep limits"):
        ctx.step(DeviceField(f0), Controls(dt=f0.dt, isor1=0))
    with pytest.raises(ChemError, match="relaxation factors"):
        ctx.step(DeviceField(f0), Controls(dt=f0.dt, relxsp=1.5))
    # the coupled step needs the global stops, and the flux inputs need the geometry
    cp = synthetic_coupling(f0, seed=4)
    with pytest.raises(ChemError, match="global stops"):
        ctx.step_coupled(DeviceField(f0), Controls(dt=f0.dt, mode=2), DeviceCoupling(cp))
    with pytest.raises(ChemError, match="geometry"):
        ctx.step_coupled(DeviceField(f0), Controls(dt=f0.dt), DeviceCoupling(cp, keep=("coef", "vdf")))
    # after all of that the context still steps, and agrees with a fresh one
    a, b = DeviceField(f0), DeviceField(f0)
    sa = ctx.step(a, Controls(dt=f0.dt))
    sb = ChemStep(f0.n).step(b, Controls(dt=f0.dt))
    assert (sa.isor1, sa.isor2) == (sb.isor1, sb.isor2)
    assert np.array_equal(a.to_host().y, b.to_host().y)
