"""Argument checks of the array-level calls that run before any GPU work:
device tensors are handed to the library as raw pointers, so a tensor of
the wrong dtype or shape must be refused (or, for integer step limits,
converted) on the host.  CPU only: the checks fire before a device is
touched."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_float_tensor_of_wrong_dtype_is_rejected():
    from pycpet_amd.ops import dev_tensor

    for dt in (torch.float64, torch.float16, torch.bfloat16, torch.int32):
        with pytest.raises(ValueError, match="float32"):
            dev_tensor(torch.zeros(6, dtype=dt), "f32", "x")


def test_step_limits_must_be_integers():
    from pycpet_amd.ops import dev_tensor

    with pytest.raises(ValueError, match="integer"):
        dev_tensor(torch.zeros(4, dtype=torch.float32), "i32", "n_iter")
    with pytest.raises(ValueError, match="integer"):
        dev_tensor(torch.zeros(4, dtype=torch.bool), "i32", "n_iter")


def test_host_tensors_are_rejected_after_the_dtype_check():
    from pycpet_amd.ops import dev_tensor

    with pytest.raises(ValueError, match="device tensor"):
        dev_tensor(torch.zeros(6, dtype=torch.float32), "f32", "x")
    with pytest.raises(ValueError, match="device tensor"):
        dev_tensor(torch.zeros(4, dtype=torch.int64), "i32", "n_iter")


def test_rows_of_three():
    from pycpet_amd.ops import _rows3

    assert _rows3(np.zeros((5, 3)), "x") == 5
    assert _rows3(np.zeros(12), "x") == 4
    assert _rows3(torch.zeros(4, 3), "x") == 4
    for bad in (np.zeros((4, 2)), np.zeros(7), np.zeros((2, 3, 3))):
        with pytest.raises(ValueError, match="rows of 3"):
            _rows3(bad, "x")


def test_batch_rejects_wrong_dtype_before_the_device():
    from pycpet_amd.batch import _cat

    parts = [torch.zeros(6, dtype=torch.float64), torch.zeros(3, dtype=torch.float64)]
    with pytest.raises(ValueError, match="float32"):
        _cat(parts, np.float32, True, "x")
    with pytest.raises(ValueError, match="mix"):
        _cat([torch.zeros(3), np.zeros(3)], np.float32, True, "x")
    got = _cat([np.arange(3), np.arange(2)], np.int32, False, "n_iter")
    assert got.dtype == np.int32 and list(got) == [0, 1, 2, 0, 1]
