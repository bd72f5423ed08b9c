"""Pins BLAKE3 (oracle/blake3.hpp) to published digests and the Merkle tree / multiproof
(oracle/merkle.hpp) to the reference tests utils/src/merkle.rs:126-156 and to hand-derived
proof layouts."""
import struct

import numpy as np
import pytest

import oracle

P = 2**64 - 2**32 + 1


def test_blake3_published_vectors():
    # BLAKE3 specification / official test vectors (input_len 0 and 1 of the 0,1,2,.. pattern)
    assert oracle.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert oracle.blake3(b"abc").hex() == "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"
    assert oracle.blake3(bytes([0])).hex() == "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213"


def test_blake3_multiblock_and_multichunk_are_prefix_sensitive():
    data = bytes(i % 251 for i in range(4096))
    seen = set()
    for n in (63, 64, 65, 127, 128, 256, 1023, 1024, 1025, 2048, 2049, 3072, 4096):
        h = oracle.blake3(data[:n])
        assert h not in seen
        seen.add(h)


def _ser(vals):
    # Serialize::serialize_fields: canonical little-endian u64 (merkle.rs:24-29)
    return b"".join(struct.pack("<Q", v % P) for v in vals)


def test_reference_serialize():
    # merkle.rs:146-156: u64::MAX reduces to its canonical value 2^32 - 2
    assert _ser([42, 2**64 - 1]) == struct.pack("<Q", 42) + struct.pack("<Q", 2**32 - 2)


def test_reference_commit_and_open():
    # merkle.rs:126-144: 8 leaves serialize([2x, 2x+1]); open {2, 3}; verify
    leaves = np.frombuffer(b"".join(_ser([2 * x, 2 * x + 1]) for x in range(8)), np.uint8).reshape(8, 16)
    root = oracle.merkle_root(leaves)
    proof = oracle.merkle_open(leaves, [2, 3])
    assert oracle.merkle_verify(root, 8, proof, [2, 3], leaves[[2, 3]])
    # layout: 2 and 3 are siblings; their parent (layer 1 node 1) needs node 0, then layer 2
    # node 0 needs node 1 -> two hashes, in that order
    lh = [oracle.blake3(leaves[i].tobytes()) for i in range(8)]
    l1 = [oracle.blake3(lh[2 * i] + lh[2 * i + 1]) for i in range(4)]
    l2 = [oracle.blake3(l1[2 * i] + l1[2 * i + 1]) for i in range(2)]
    assert root == oracle.blake3(l2[0] + l2[1])
    assert proof == l1[0] + l2[1]


@pytest.mark.parametrize("n,idx", [(16, [0]), (16, [15]), (16, [1, 2, 9]), (64, [3, 4, 5, 40, 63]),
                                   (32, list(range(32))), (1, [0]), (2, [1])])
def test_multiproof_round_trip_and_length(n, idx):
    rng = np.random.default_rng(n + len(idx))
    leaves = rng.integers(0, 256, size=(n, 48), dtype=np.uint8)
    root = oracle.merkle_root(leaves)
    proof = oracle.merkle_open(leaves, idx)
    assert oracle.merkle_verify(root, n, proof, idx, leaves[idx])
    # hashes in a multiproof = nodes outside the known set whose sibling is in it, per layer
    known, count, size = set(idx), 0, n
    while size > 1:
        count += sum(1 for i in known if (i ^ 1) not in known and (i ^ 1) < size)
        known = {i // 2 for i in known}
        size //= 2
    assert len(proof) == 32 * count
    if proof:
        bad = bytearray(proof)
        bad[5] ^= 1
        assert not oracle.merkle_verify(root, n, bytes(bad), idx, leaves[idx])
    wrong = leaves[idx].copy()
    wrong[0, 0] ^= 1
    assert not oracle.merkle_verify(root, n, proof, idx, wrong)
