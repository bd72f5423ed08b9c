"""Pins the oracle's subset-sum hash (oracle/ligesis.hpp `hash_column`, `add_noncanon`) against
Python big integers, restating the reference's own tests pcs/src/subset_sum.rs:89-124:

- `hash_column_matches_naive_matrix_vector_product` (width 128): H[r] = sum of A[r][i] over the
  set bits x[i], mod p;
- `hash_columns_is_columnwise_hash` (width 64, 10 columns): column j of hash_columns is the hash
  of column j alone.

The oracle keeps its running sums in [0, 2^64) and folds a carry out of 2^64 back as 2^32 - 1
(p's complement); rows of p - 1 and of values just below p make almost every addition carry.
Its second-carry branch cannot run for canonical addends: with a < 2^64 and b < p, a wrapped
sum is a + b - 2^64 <= p - 2, and adding 2^32 - 1 stays below 2^64. The big-integer check
covers the result either way.
"""
import numpy as np
import pytest

import oracle
from ligesis_amd import synth

P = 2**64 - 2**32 + 1
C = 32


def _naive(a: np.ndarray, bits: np.ndarray) -> list[list[int]]:
    """H[r][j] = sum_i a[r][i] * bits[j][i] mod p with Python integers."""
    A = [[int(v) for v in row] for row in a]
    out = [[0] * bits.shape[0] for _ in range(C)]
    for j in range(bits.shape[0]):
        sel = [i for i in range(bits.shape[1]) if bits[j, i]]
        for r in range(C):
            out[r][j] = sum(A[r][i] for i in sel) % P
    return out


def _matrix(width: int, seed: int) -> np.ndarray:
    a = synth.field_elems(seed, C * width).reshape(C, width)
    a[0, :] = P - 1                      # every add carries out of 2^64
    a[1, :] = P - 2
    a[2, :] = 2**63                      # two of these wrap exactly
    a[3, :] = 0
    a[4, :] = 1
    a[5, ::2] = P - 1                    # alternating large / small
    a[5, 1::2] = 2**32 - 1
    a[6, :] = 0xFFFFFFFF00000000          # p - 1 in the other canonical form's neighbourhood
    return a


@pytest.mark.parametrize("width,seed", [(128, 1), (128, 2), (1024, 3)])
def test_hash_column_matches_naive_matrix_vector_product(width, seed):
    a = _matrix(width, 100 + seed)
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2, size=(6, width), dtype=np.uint8)
    bits[0, :] = 1   # all bits set: W additions per row
    bits[1, :] = 0   # empty column hashes to zero
    bits[2, :] = 0
    bits[2, width - 1] = 1  # a single bit selects one column of A
    got = oracle.hash_columns(a, bits)
    want = _naive(a, bits)
    for r in range(C):
        assert [int(v) for v in got[r]] == want[r], f"row {r}"
    assert (got < np.uint64(P)).all(), "outputs must be canonical"
    # closed forms on the structured rows
    assert int(got[0, 0]) == (width * (P - 1)) % P
    assert int(got[3, 0]) == 0 and int(got[4, 0]) == width
    assert (got[:, 1] == 0).all()
    np.testing.assert_array_equal(got[:, 2], a[:, width - 1])


def test_hash_columns_is_columnwise_hash():
    width = 64
    a = _matrix(width, 7)
    rng = np.random.default_rng(11)
    cols = rng.integers(0, 2, size=(10, width), dtype=np.uint8)
    hm = oracle.hash_columns(a, cols)
    for j in range(10):
        d = oracle.hash_columns(a, cols[j:j + 1])
        np.testing.assert_array_equal(hm[:, j], d[:, 0])


def test_hash_is_linear_over_disjoint_supports():
    # H(x | y) = H(x) + H(y) mod p when x and y share no set bit
    width = 256
    a = _matrix(width, 9)
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2, size=width, dtype=np.uint8)
    y = (1 - x) * rng.integers(0, 2, size=width, dtype=np.uint8)
    h = oracle.hash_columns(a, np.stack([x, y, x | y]))
    for r in range(C):
        assert (int(h[r, 0]) + int(h[r, 1])) % P == int(h[r, 2])
