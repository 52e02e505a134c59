"""Checks the Python layer makes on torch tensors before their data_ptr goes
to the C-ABI (CPU tensors here; the device path runs in the GPU tests)."""
import ctypes

import pytest
import torch

from diffcp_amd.engine import _is_tensor, _tensor_arg


def test_tensor_argument_checks():
    t = torch.zeros(4, 3, dtype=torch.float64)
    p = _tensor_arg(t, "float64", (4, 3), "x")
    assert ctypes.cast(p, ctypes.c_void_p).value == t.data_ptr()
    with pytest.raises(TypeError, match="float32 tensor expected"):
        _tensor_arg(t, "float32", (4, 3), "x")
    with pytest.raises(ValueError, match="contiguous"):
        _tensor_arg(torch.zeros(3, 4, dtype=torch.float64).t(), "float64", (4, 3), "x")
    with pytest.raises(ValueError, match="shape"):
        _tensor_arg(t, "float64", (3, 4), "x")
    assert _is_tensor(t) and not _is_tensor(t.numpy())
