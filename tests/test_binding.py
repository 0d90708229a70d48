"""Python binding argument checks, run without a GPU: every bad size, dtype,
layout or device is refused before a pointer reaches the C ABI (an
undersized output or a strided view would otherwise send the kernels past
an allocation)."""
import ctypes

import pytest
import torch

import apsep_amd as A


def _scratch_without_device(n_cap=100, device=0):
    # a Scratch whose handle is null: any call that reached the library
    # would fail with ERR_INVALID_ARG instead of the ValueError checked here
    s = A.Scratch.__new__(A.Scratch)
    s._h = ctypes.c_void_p(None)
    s.n_cap, s.incl, s.device = n_cap, False, device
    return s


def _plan_without_device(total_cap=1000, nseg_cap=4, device=0):
    p = A.BatchPlan.__new__(A.BatchPlan)
    p._h = ctypes.c_void_p(None)
    p.total_cap, p.nseg_cap, p.incl, p.device = total_cap, nseg_cap, False, device
    return p


def i32(n):
    return torch.zeros(n, dtype=torch.int32)


def test_single_dtype():
    with pytest.raises(TypeError):
        A.check_single_args(torch.zeros(8, dtype=torch.int64), i32(8), 8, 100, 0)
    with pytest.raises(TypeError):
        A.check_single_args(i32(8), torch.zeros(8, dtype=torch.float32), 8, 100, 0)


@pytest.mark.parametrize("n,nx,nout,cap", [
    (9, 8, 16, 100),    # n beyond the input
    (9, 16, 8, 100),    # n beyond the output
    (101, 200, 200, 100),  # n beyond the capacity
    (-1, 8, 8, 100),
])
def test_single_sizes(n, nx, nout, cap):
    with pytest.raises(ValueError):
        A.check_single_args(i32(nx), i32(nout), n, cap, 0)


def test_single_strided_and_host_tensors():
    with pytest.raises(ValueError, match="contiguous"):
        A.check_single_args(i32(16)[::2], i32(8), 8, 100, 0)
    with pytest.raises(ValueError, match="contiguous"):
        A.check_single_args(i32(8), i32(16)[::2], 8, 100, 0)
    with pytest.raises(ValueError, match="device tensors"):
        A.check_single_args(i32(8), i32(8), 8, 100, 0)


def test_scratch_run_checks_before_any_call():
    s = _scratch_without_device(n_cap=100)
    with pytest.raises(ValueError, match="exceeds the input"):
        s.run(i32(50), i32(100), n=60)
    with pytest.raises(ValueError, match="exceeds the input"):
        s.run(i32(100), i32(50), n=60)
    with pytest.raises(ValueError, match="capacity"):
        s.run(i32(200), i32(200))
    with pytest.raises(ValueError, match="device tensors"):
        s.run(i32(100), i32(100))
    with pytest.raises(ValueError, match="exceeds"):
        s.time_kernels(i32(10), i32(100), 1, n=20)


def _offs(*v):
    return torch.tensor(v, dtype=torch.int64)


def test_segmented_dtypes_and_shapes():
    with pytest.raises(TypeError):
        A.check_segmented_args(i32(8), _offs(0, 8).to(torch.int32), i32(8), 8, 100, 4, 0)
    with pytest.raises(ValueError, match="1-D"):
        A.check_segmented_args(i32(8), _offs(0, 8).reshape(1, 2), i32(8), 8, 100, 4, 0)
    with pytest.raises(ValueError, match="1-D"):
        A.check_segmented_args(i32(8), torch.zeros(0, dtype=torch.int64), i32(8), 8, 100, 4, 0)
    with pytest.raises(ValueError, match="segments exceed"):
        A.check_segmented_args(i32(8), _offs(0, 1, 2, 3, 4, 8), i32(8), 8, 100, 4, 0)


@pytest.mark.parametrize("total,nx,nout,cap", [
    (9, 8, 16, 100), (9, 16, 8, 100), (101, 200, 200, 100), (-1, 8, 8, 100)])
def test_segmented_sizes(total, nx, nout, cap):
    with pytest.raises(ValueError):
        A.check_segmented_args(i32(nx), _offs(0, 8), i32(nout), total, cap, 4, 0)


def test_segmented_layout_and_host_tensors():
    with pytest.raises(ValueError, match="contiguous"):
        A.check_segmented_args(i32(16)[::2], _offs(0, 8), i32(8), 8, 100, 4, 0)
    with pytest.raises(ValueError, match="contiguous"):
        A.check_segmented_args(i32(8), _offs(0, 4, 4, 8)[::2], i32(8), 8, 100, 4, 0)
    with pytest.raises(ValueError, match="device tensors"):
        A.check_segmented_args(i32(8), _offs(0, 8), i32(8), 8, 100, 4, 0)


def test_plan_run_and_tune_check_before_any_call():
    p = _plan_without_device(total_cap=1000, nseg_cap=2)
    with pytest.raises(ValueError, match="segments exceed"):
        p.run(i32(10), _offs(0, 3, 6, 10), i32(10))
    with pytest.raises(ValueError, match="exceeds the keys"):
        p.run(i32(10), _offs(0, 20), i32(30), total=20)
    with pytest.raises(ValueError, match="capacity"):
        p.run(i32(2000), _offs(0, 2000), i32(2000))
    with pytest.raises(ValueError, match="device tensors"):
        p.tune(i32(10), _offs(0, 10), i32(10))
