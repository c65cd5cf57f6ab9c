"""Host logic of solve() and of checkpoints, on the CPU: the fixed-step plan
(full steps, a shortened last step, save times on the step grid) and the
checkpoint file round trip at the state's own precision."""
import os

import numpy as np
import pytest
import torch


def gw():
    import geothermalwells_jl_amd as g
    return g


def test_plan_of_the_examples():
    """examples/single_borehole_basic.jl: tspan (0, 800), dt 80, saveat
    range(0, 800, 11) -- 10 steps, every save time a step time."""
    n, last = gw().step_plan((0.0, 800.0), 80.0, np.linspace(0.0, 800.0, 11))
    assert (n, last) == (10, 0.0)


def test_unaligned_save_times_raise():
    """examples/well_array_basic.jl: saveat range(0, 800, 5) = 0, 200, ...
    with dt 80: 200 and 600 fall between steps; the reference interpolates
    there, this mirror raises rather than drop them."""
    with pytest.raises(ValueError, match="200"):
        gw().step_plan((0.0, 800.0), 80.0, np.linspace(0.0, 800.0, 5))
    # times outside tspan are not save times of the run
    assert gw().step_plan((0.0, 800.0), 80.0, [-5.0, 900.0]) == (10, 0.0)


def test_shortened_last_step():
    n, last = gw().step_plan((0.0, 1000.0), 300.0)
    assert n == 3 and last == pytest.approx(100.0, abs=1e-12)
    # the end of tspan is a save time even when it is not t0 + k dt
    assert gw().step_plan((0.0, 1000.0), 300.0, [0.0, 600.0, 1000.0])[0] == 3
    with pytest.raises(ValueError):
        gw().step_plan((0.0, 1000.0), 300.0, [950.0])
    # whole numbers of steps that are inexact in binary
    assert gw().step_plan((0.0, 0.3), 0.1) == (3, 0.0)
    assert gw().step_plan((0.0, 31536000.0), 600.0) == (52560, 0.0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_checkpoint_round_trip(tmp_path, dtype):
    g = gw()
    u = torch.randn(3 * 4 * 5, dtype=torch.float64).to(dtype)
    path = os.path.join(tmp_path, "ck.npz")
    g.save_checkpoint(path, u, 1234.5, naccept=7)
    ck = g.load_checkpoint(path)
    assert ck.u.dtype == dtype
    assert torch.equal(ck.u, u)
    assert ck.t == 1234.5 and ck.naccept == 7
    assert ck.T_outlet is None and ck.shapes is None
    with np.load(path, allow_pickle=False) as f:
        assert f["u"].dtype == (np.float64 if dtype == torch.float64 else np.float32)


def test_checkpoint_version_is_checked(tmp_path):
    path = os.path.join(tmp_path, "bad.npz")
    np.savez(path, version=np.int64(99), u=np.zeros(3), t=np.float64(0.0), naccept=np.int64(0))
    with pytest.raises(ValueError, match="version"):
        gw().load_checkpoint(path)
