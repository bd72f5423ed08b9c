"""Argument checks of the standalone Basefold commitment (ligesis_amd.Basefold) that run before
any device call: device tensors must be contiguous 8-byte words, extension input is (N, 2)
pairs, commit_base takes base-field input only, lengths are powers of two
(pcs/src/basefold.rs:309-382 take MlPoly<BF> / MlPoly<EF> of 2^num_vars evaluations)."""
import numpy as np
import pytest
import torch

from ligesis_amd import Basefold, LigesisError


def test_host_shape_checks():
    with pytest.raises(LigesisError):
        Basefold._words(np.zeros((16, 3), np.uint64), None, False)
    with pytest.raises(LigesisError):
        Basefold._words(np.zeros(12, np.uint64), None, False)
    with pytest.raises(LigesisError):
        Basefold._words(np.zeros((16, 2), np.uint64), False, False)
    with pytest.raises(LigesisError):
        Basefold._words(np.zeros((2, 2, 2), np.uint64), None, False)


def test_host_ext_upcast():
    base = np.arange(16, dtype=np.uint64)
    _, nv, ext, keep = Basefold._words(base, True, False)
    assert (nv, ext) == (4, True)
    assert keep.shape == (16, 2) and (keep[:, 0] == base).all() and (keep[:, 1] == 0).all()
    _, nv, ext, _ = Basefold._words(base, None, False)
    assert (nv, ext) == (4, False)


def test_device_path_checks_dtype_layout_and_pairs():
    # CPU tensors stand in for device tensors: the checks run before the pointer is used
    x = torch.zeros((16, 2), dtype=torch.int64)
    with pytest.raises(LigesisError):
        Basefold._words(x[:, 0], None, True)  # strided column
    with pytest.raises(LigesisError):
        Basefold._words(torch.zeros(16, dtype=torch.int32), None, True)
    with pytest.raises(LigesisError):
        Basefold._words(torch.zeros((16, 3), dtype=torch.int64), None, True)
    with pytest.raises(LigesisError):
        Basefold._words(x, False, True)
    with pytest.raises(LigesisError):
        Basefold._words(torch.zeros(24, dtype=torch.int64), None, True)
    _, nv, ext, keep = Basefold._words(torch.arange(8, dtype=torch.int64), True, True)
    assert (nv, ext) == (3, True) and tuple(keep.shape) == (8, 2) and keep.is_contiguous()
    _, nv, ext, _ = Basefold._words(x, None, True)
    assert (nv, ext) == (4, True)
