"""Host-side checks of the PolyCommitmentScheme adapter (pcs/src/noop.rs:23-48): argument
validation happens before any device call, and verification is declined rather than accepted."""
from types import SimpleNamespace

import numpy as np
import pytest

from ligesis_amd import LigesisError, RandomOracle
from ligesis_amd.pcs import LigesisPcs, LigesisPcsData, PolyCommitmentScheme


def _scheme():
    return LigesisPcs(SimpleNamespace(log_m=2, log_n=6, ctx=None))


def test_is_a_poly_commitment_scheme():
    assert isinstance(_scheme(), PolyCommitmentScheme)
    assert _scheme().size == 256


@pytest.mark.parametrize("shape", [(255,), (256, 3), (128, 2), (2, 256)])
def test_commit_rejects_wrong_shapes(shape):
    with pytest.raises(LigesisError):
        _scheme().commit(np.zeros(shape, np.uint64))


def test_components_split_extension_values():
    s = _scheme()
    base, ext = s._components(np.arange(256, dtype=np.uint64))
    assert not ext and len(base) == 1
    poly = np.arange(512, dtype=np.uint64).reshape(256, 2)
    comps, ext = s._components(poly)
    assert ext
    np.testing.assert_array_equal(comps[0], poly[:, 0])
    np.testing.assert_array_equal(comps[1], poly[:, 1])


def test_batch_prove_rejects_foreign_data_and_bad_points():
    s = _scheme()
    ro = RandomOracle(np.zeros((8, 2), np.uint64), np.zeros(8, np.uint64))
    with pytest.raises(LigesisError):
        s.batch_prove([(object(), np.zeros((8, 2), np.uint64))], ro)
    with pytest.raises(LigesisError):
        s.batch_prove([(LigesisPcsData((), False), np.zeros((7, 2), np.uint64))], ro)
    assert ro.drawn() == (0, 0)


def test_batch_verify_is_declined():
    with pytest.raises(NotImplementedError):
        _scheme().batch_verify([], [], None)
