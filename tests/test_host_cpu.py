"""Host-side argument checks that need no GPU: they raise before any library call."""
import os

import numpy as np
import pytest
import torch

from dmff_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "dmff_amd", "libadmp_hip.so")


def test_scale_vectors_must_have_equal_lengths():
    m, p, d = L.check_scales([0, 0, 0, 0, 1, 1], [0, 0, 0, 0, 1, 1], [1, 1, 1, 1, 1, 1])
    assert m.shape == p.shape == d.shape == (6,)
    with pytest.raises(ValueError, match="equal lengths"):
        L.check_scales([0, 0, 0, 0, 1, 1], [0, 0, 1])
    with pytest.raises(ValueError, match="equal lengths"):
        L.check_scales([0, 0, 0, 0, 1, 1], None, [1, 1])
    with pytest.raises(ValueError, match="1..8"):
        L.check_scales(np.ones(9))
    with pytest.raises(ValueError, match="1..8"):
        L.check_scales([])


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_type_reducer_refuses_host_tensors_for_the_hip_library():
    from dmff_amd.admp.pme import TypeReducer
    assert not L.HOST_MEMORY
    tr = TypeReducer(np.array([0, 1, 1, 0]))
    with pytest.raises(ValueError, match="CUDA tensor"):
        tr(torch.zeros(4, 3, dtype=torch.float64))
    with pytest.raises(ValueError, match="type map"):
        tr(torch.zeros(5, 3, dtype=torch.float64))
