"""Host side of the batched moment-map objective (no GPU): the data,
error and mask maps stacked in the data model's key order, masked pixels
given zero data and unit error, per-key weights and errors."""
import numpy as np
import pytest

from gbkfit_amd.batch import BatchEvaluator
from gbkfit_amd.dataset import Data, DatasetMMaps
from gbkfit_amd.dmodels import DModelMMaps


def test_mmap_data_stacking():
    dm = DModelMMaps((6, 4), (1, 1), orders=(0, 2))
    d0 = np.arange(24, dtype=np.float32).reshape(4, 6)
    d2 = -d0
    d2[1, 2] = np.nan
    m2 = np.ones_like(d2)
    m2[1, 2] = 0
    ds = DatasetMMaps(mmap0=Data(d0, error=np.full_like(d0, 2.0)),
                      mmap2=Data(d2, mask=m2, error=np.full_like(d2, 3.0)))
    orders, D, E, esc, M, W, cutoff, z0, dz = BatchEvaluator._mmap_data(
        None, dm, ds, None, None, {'mmap2': 0.5})
    assert list(orders) == [0, 2] and D.shape == (2, 4, 6)
    assert np.array_equal(D[0], d0) and D[1, 1, 2] == 0.0
    assert E[0, 0, 0] == 2.0 and E[1, 0, 0] == 3.0 and E[1, 1, 2] == 1.0
    assert M[1, 1, 2] == 0 and M[0].min() == 1
    assert list(W) == [1.0, 0.5] and cutoff == 1e-6
    assert dz == dm.dcube().step()[2] and z0 == dm.dcube().zero()[2]
    # plain arrays with a scalar error and no masks
    orders, D, E, esc, M, W, *_ = BatchEvaluator._mmap_data(
        None, dm, {'mmap0': d0, 'mmap2': -d0}, 0.25, None, 2.0)
    assert E is None and M is None and list(esc) == [0.25, 0.25]
    assert list(W) == [2.0, 2.0]
    with pytest.raises(RuntimeError, match='mmap2'):
        BatchEvaluator._mmap_data(None, dm, {'mmap0': d0}, 1.0, None, 1.0)
    with pytest.raises(RuntimeError, match='shape'):
        BatchEvaluator._mmap_data(None, dm, {'mmap0': d0, 'mmap2': d0.T},
                                  1.0, None, 1.0)
