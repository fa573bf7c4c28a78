"""Host-side argument checks of the Python binding, without a GPU: device
outputs that the library would write through are validated before any
library call, so a CPU tensor, another dtype, a strided view or a short
buffer raises ValueError instead of reaching a device copy."""
import types

import numpy as np
import pytest
import torch

import basilisk_amd as bmg


class FakeCuda:
    """Stands in for a CUDA tensor: the checks read only these attributes."""

    def __init__(self, numel, dtype=torch.float64, contiguous=True):
        self.is_cuda, self.dtype, self._n, self._c = True, dtype, numel, contiguous

    def data_ptr(self):
        return 0

    def is_contiguous(self):
        return self._c

    def numel(self):
        return self._n


def _grid_stub(dim=2, depth=3):
    n = 1 << depth
    return types.SimpleNamespace(dim=dim, depth=depth, lib=None, _g=None,
                                 shape=lambda level=None: (n,) * dim,
                                 face_shape=lambda c, level=None: tuple(n + (d == c) for d in range(dim)),
                                 nleaves=lambda: 10, _t=None)


BAD = [torch.zeros(64, dtype=torch.float64),           # host tensor
       FakeCuda(64, dtype=torch.float32),              # wrong dtype
       FakeCuda(64, contiguous=False),                 # strided view
       FakeCuda(63)]                                   # too short


@pytest.mark.parametrize("out", BAD)
def test_grid_get_rejects_bad_out(out):
    with pytest.raises(ValueError):
        bmg.Grid._get(_grid_stub(), bmg.F_A, None, (8, 8), out)


@pytest.mark.parametrize("out", [torch.zeros(10, dtype=torch.float64), FakeCuda(9), FakeCuda(10, torch.float32)])
def test_tree_get_rejects_bad_out(out):
    with pytest.raises(ValueError):
        bmg.Tree._get(_grid_stub(), bmg.F_A, out)


def test_device_ptr_accepts_matching_tensor():
    assert bmg._device_ptr(FakeCuda(64), 64, "x").value in (None, 0)


@pytest.mark.parametrize("which", ["psi", "alpha", "lambda"])
def test_coef_gradient_rejects_bad_tensors(which):
    g = _grid_stub()
    psi = FakeCuda(64)
    ga = [FakeCuda(72), FakeCuda(72)]
    gl = FakeCuda(64)
    if which == "psi":
        psi = FakeCuda(64, dtype=torch.float32)
    elif which == "alpha":
        ga[1] = FakeCuda(64)  # n x (n+1) faces expected
    else:
        gl = FakeCuda(64, contiguous=False)
    with pytest.raises(ValueError):
        bmg.Grid.coef_gradient(g, psi, ga, gl)
