"""Host-tensor arguments of the host-streamed sweep (api._host_ptr): the buffer is read as the
column-major tensor, so flat arrays and Fortran-ordered arrays pass and a C-ordered N-d array, a
wrong dtype or a device tensor is refused before any call into the library."""
import numpy as np
import pytest
import torch

from tensor_train_ring_decompositions_hpc_gpu_amd.api import _host_ptr


def test_accepts_flat_and_fortran_order():
    a = np.zeros(24, np.float32)
    assert _host_ptr(a) == a.ctypes.data
    b = np.zeros((2, 3, 4), np.float32, order="F")
    assert _host_ptr(b) == b.ctypes.data
    t = torch.zeros(24, dtype=torch.float32)
    assert _host_ptr(t) == t.data_ptr()


def test_refuses_what_would_be_misread():
    with pytest.raises(TypeError):
        _host_ptr(np.zeros((2, 3, 4), np.float32))           # C order: would be read transposed
    with pytest.raises(TypeError):
        _host_ptr(np.zeros(24, np.float64))
    with pytest.raises(TypeError):
        _host_ptr(torch.zeros(24, dtype=torch.float64))
    with pytest.raises(TypeError):
        _host_ptr(torch.zeros(4, 6, dtype=torch.float32).t())   # not contiguous


def test_refuses_nd_torch_tensors():
    """A C-contiguous N-d torch tensor would be read transposed as the column-major tensor: only flat
    tensors pass (numpy arrays can say Fortran order; torch tensors cannot)."""
    with pytest.raises(TypeError, match="flat"):
        _host_ptr(torch.zeros(2, 3, 4, dtype=torch.float32))
    with pytest.raises(TypeError, match="flat"):
        _host_ptr(torch.zeros(4, 6, dtype=torch.float32))
