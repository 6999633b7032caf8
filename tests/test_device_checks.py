"""Device-tensor checks of the Python actors, run before anything is launched.

Undersized or misplaced device tensors would make the kernels read or write
out of bounds, so they must raise on the host.  Asynchronous calls refuse
inputs that need a converted copy (the copy could be reused by the caching
allocator while the solver stream still reads it).  The checks take any
object with the tensor attributes they read, so they are exercised here with
stand-ins that need no GPU.
"""
import pytest

from espresso_amd.electrostatics import (check_device_tensor, device_input,
                                         device_output)


class _Dev:
    def __init__(self, index):
        self.index = index


class FakeTensor:
    def __init__(self, numel, dtype="torch.float64", contiguous=True, device=0,
                 cuda=True):
        self._numel = numel
        self.dtype = dtype
        self._contig = contiguous
        self.device = _Dev(device)
        self.is_cuda = cuda

    def numel(self):
        return self._numel

    def is_contiguous(self):
        return self._contig


def test_sizes_are_checked():
    assert check_device_tensor(FakeTensor(30), "pos", 30, 0)
    with pytest.raises(ValueError, match="'pos' has 29 elements, expected 30"):
        check_device_tensor(FakeTensor(29), "pos", 30, 0)
    with pytest.raises(ValueError, match="expected 30"):
        check_device_tensor(FakeTensor(31), "force", 30, 0)
    # energy/virial outputs: at least 1 / 9 elements
    device_output(FakeTensor(1), "energy", 1, 0, minimum=True)
    device_output(FakeTensor(12), "virial", 9, 0, minimum=True)
    with pytest.raises(ValueError, match="at least 9"):
        device_output(FakeTensor(8), "virial", 9, 0, minimum=True)
    with pytest.raises(ValueError, match="at least 1"):
        device_output(FakeTensor(0), "energy", 1, 0, minimum=True)


def test_device_and_kind_are_checked():
    with pytest.raises(ValueError, match="on device 1, the solver on device 0"):
        check_device_tensor(FakeTensor(3, device=1), "q", 3, 0)
    with pytest.raises(TypeError, match="must be a device tensor"):
        check_device_tensor(FakeTensor(3, cuda=False), "q", 3, 0)
    with pytest.raises(TypeError, match="contiguous float64"):
        device_output(FakeTensor(9, dtype="torch.float32"), "force", 9, 0)
    with pytest.raises(TypeError, match="contiguous float64"):
        device_output(FakeTensor(9, contiguous=False), "torque", 9, 0)


def test_asynchronous_calls_refuse_converted_inputs():
    ok = FakeTensor(9)
    assert device_input(ok, "pos", 9, 0, asynchronous=True) is ok
    for bad in (FakeTensor(9, dtype="torch.float32"), FakeTensor(9, contiguous=False)):
        with pytest.raises(TypeError, match="asynchronous call"):
            device_input(bad, "pos", 9, 0, asynchronous=True)
    # sizes are checked before the conversion question arises
    with pytest.raises(ValueError, match="expected 9"):
        device_input(FakeTensor(6, dtype="torch.float32"), "pos", 9, 0, asynchronous=False)


@pytest.mark.gpu
def test_actor_device_checks_on_the_gpu():
    import numpy as np
    import torch
    from espresso_amd.electrostatics import P3M
    from espresso_amd.magnetostatics import DipolarP3M
    rng = np.random.default_rng(3)
    pos = rng.uniform(0., 10., (200, 3))
    q = np.where(np.arange(200) % 2 == 0, 1., -1.)
    p3m = P3M(prefactor=1., accuracy=1e-3, mesh=16, cao=3, r_cut=3., alpha=0.9,
              tune=False)
    p3m.activate(pos, q, box_l=[10.] * 3)
    tp = torch.tensor(pos, device="cuda")
    tq = torch.tensor(q, device="cuda")
    f = torch.zeros_like(tp)
    with pytest.raises(ValueError, match="'force' has"):
        p3m.add_long_range_forces(tp, tq, f[:-1])
    with pytest.raises(ValueError, match="'pos' has"):
        p3m.add_long_range_forces(tp[:-1], tq, f)
    with pytest.raises(TypeError, match="asynchronous call"):
        p3m.add_long_range_forces(tp.float(), tq, f, sync=False)
    with pytest.raises(TypeError, match="asynchronous call"):
        p3m.add_long_range_forces(tp.t().contiguous().t(), tq, f, sync=False)
    with pytest.raises(ValueError, match="at least 9"):
        p3m.solve_to_device(tp, tq, f, virial=torch.zeros(8, dtype=torch.float64,
                                                          device="cuda"), sync=True)
    # a synchronous call converts, and gives the float64 result
    f.zero_()
    p3m.add_long_range_forces(tp, tq, f)
    g = torch.zeros_like(tp)
    p3m.add_long_range_forces(tp.t().contiguous().t(), tq.float(), g)
    assert float((f - g).norm() / f.norm()) <= 1e-12
    dp = DipolarP3M(prefactor=1., accuracy=1e-3, mesh=16, cao=3, r_cut=3., alpha=1.,
                    tune=False, verbose=False)
    dip = rng.normal(size=(200, 3))
    dp.activate(pos, dip, box_l=10.)
    td = torch.tensor(dip, device="cuda")
    tt = torch.zeros_like(tp)
    with pytest.raises(ValueError, match="'torque' has"):
        dp.add_long_range_forces(tp, td, f, tt[:-2])
    with pytest.raises(TypeError, match="asynchronous call"):
        dp.add_long_range_forces(tp, td.float(), f, tt, sync=False)
