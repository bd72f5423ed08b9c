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


# ---- the Z_2^64 hash (ring64): the same sums mod 2^64, over any 64-bit A ---------------------
def _naive_ring(a: np.ndarray, bits: np.ndarray) -> list[list[int]]:
    A = [[int(v) for v in row] for row in a]
    out = [[0] * bits.shape[0] for _ in range(C)]
    for j in range(bits.shape[0]):
        sel = [i for i in range(bits.shape[1]) if bits[j, i]]
        for r in range(C):
            out[r][j] = sum(A[r][i] for i in sel) % 2**64
    return out


@pytest.mark.parametrize("width,seed", [(128, 11), (1024, 12)])
def test_ring64_hash_matches_big_integer_sums(width, seed):
    """Words at and above p (never canonical field elements), 2^64 - 1, 2^63: the sums wrap mod
    2^64 where the field hash would reduce mod p, so the two differ on such rows."""
    a = synth.splitmix64(seed, C * width).reshape(C, width)
    a[0, :] = 2**64 - 1
    a[1, :] = P
    a[2, :] = 2**63
    a[3, :] = 0x7F7F7F7F7F7F7F80  # the balanced-digit boundary of the MFMA kernels
    a[4, :] = 0x8080808080808080
    bits = (synth.splitmix64(seed + 1, 6 * width) & 1).astype(np.uint8).reshape(6, width)
    bits[0, :] = 1
    got = oracle.hash_columns(a, bits, ring64=True)
    want = _naive_ring(a, bits)
    for r in range(C):
        for j in range(bits.shape[0]):
            assert int(got[r, j]) == want[r][j], (r, j)
    assert int(got[0, 0]) == (width * (2**64 - 1)) % 2**64
    field = oracle.hash_columns(a % P, bits)
    assert any(int(field[r, 0]) != int(got[r, 0]) for r in range(C))


def test_ring64_hash_is_linear_mod_2_64():
    width = 256
    a = synth.splitmix64(21, C * width).reshape(C, width)
    x = (synth.splitmix64(22, width) & 1).astype(np.uint8)
    y = (synth.splitmix64(23, width) & 1).astype(np.uint8) * (1 - x)  # disjoint support
    hx, hy, hxy = (oracle.hash_columns(a, b[None, :], ring64=True)[:, 0] for b in (x, y, x + y))
    for r in range(C):
        assert (int(hx[r]) + int(hy[r])) % 2**64 == int(hxy[r])


def test_ring64_commit_root_is_tree_over_raw_words():
    """The ring64 commit hashes each H column's 32 raw words (not reduced mod p) into its leaf."""
    lm, ln = 2, 6
    inp = synth.Inputs(lm, ln, seed=31, n_fields=4, n_ints=4)
    a = synth.splitmix64(32, C * (64 << lm))
    pk = oracle.ProvingKey(lm, ln, a, commit_a=False, ring64=True)
    cd = oracle.CommitData(pk, inp.f, keep_enc=True)
    enc = cd.enc()
    bits = np.array([[(int(enc[i, j]) >> k) & 1 for i in range(1 << lm) for k in range(64)]
                     for j in range(2 << ln)], np.uint8)
    np.testing.assert_array_equal(cd.h(), oracle.hash_columns(a.reshape(C, -1), bits, ring64=True))
    cols = np.frombuffer(np.ascontiguousarray(cd.h().T).astype("<u8").tobytes(), np.uint8).reshape(2 << ln, 256)
    assert cd.root == oracle.merkle_root(cols)  # leaf j = blake3 of column j's 32 raw words
