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


# The BLAKE3 project's published test vectors (test_vectors/test_vectors.json in the BLAKE3
# repository): input byte i is i mod 251; the digest is the first 32 bytes of the extended
# output "hash". Lengths cover one block, block boundaries (63/64/65, 127/128/129), one chunk
# (1023/1024), two or more chunks (1025 .. 102400, where the chunk tree has several levels).
# Every H leaf of the commit is a 4-block message and every Basefold leaf up to 16 blocks.
BLAKE3_PUBLISHED = {
    0: "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262",
    1: "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213",
    2: "7b7015bb92cf0b318037702a6cdd81dee41224f734684c2c122cd6359cb1ee63",
    3: "e1be4d7a8ab5560aa4199eea339849ba8e293d55ca0a81006726d184519e647f",
    8: "2351207d04fc16ade43ccab08600939c7c1fa70a5c0aaca76063d04c3228eaeb",
    63: "e9bc37a594daad83be9470df7f7b3798297c3d834ce80ba85d6e207627b7db7b",
    64: "4eed7141ea4a5cd4b788606bd23f46e212af9cacebacdc7d1f4c6dc7f2511b98",
    65: "de1e5fa0be70df6d2be8fffd0e99ceaa8eb6e8c93a63f2d8d1c30ecb6b263dee",
    127: "d81293fda863f008c09e92fc382a81f5a0b4a1251cba1634016a0f86a6bd640d",
    128: "f17e570564b26578c33bb7f44643f539624b05df1a76c81f30acd548c44b45ef",
    129: "683aaae9f3c5ba37eaaf072aed0f9e30bac0865137bae68b1fde4ca2aebdcb12",
    1023: "10108970eeda3eb932baac1428c7a2163b0e924c9a9e25b35bba72b28f70bd11",
    1024: "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7",
    1025: "d00278ae47eb27b34faecf67b4fe263f82d5412916c1ffd97c8cb7fb814b8444",
    2048: "e776b6028c7cd22a4d0ba182a8bf62205d2ef576467e838ed6f2529b85fba24a",
    2049: "5f4d72f40d7a5f82b15ca2b2e44b1de3c2ef86c426c95c1af0b6879522563030",
    3072: "b98cb0ff3623be03326b373de6b9095218513e64f1ee2edd2525c7ad1e5cffd2",
    3073: "7124b49501012f81cc7f11ca069ec9226cecb8a2c850cfe644e327d22d3e1cd3",
    4096: "015094013f57a5277b59d8475c0501042c0b642e531b0a1c8f58d2163229e969",
    4097: "9b4052b38f1c5fc8b1f9ff7ac7b27cd242487b3d890d15c96a1c25b8aa0fb995",
    8192: "aae792484c8efe4f19e2ca7d371d8c467ffb10748d8a5a1ae579948f718a2a63",
    8193: "bab6c09cb8ce8cf459261398d2e7aef35700bf488116ceb94a36d0f5f1b7bc3b",
    16384: "f875d6646de28985646f34ee13be9a576fd515f76b5b0a26bb324735041ddde4",
    31744: "62b6960e1a44bcc1eb1a611a8d6235b6b4b78f32e7abc4fb4c6cdcce94895c47",
    102400: "bc3e3d41a1146b069abffad3c0d44860cf664390afce4d9661f7902e7943e085",
}


@pytest.mark.parametrize("n", sorted(BLAKE3_PUBLISHED))
def test_blake3_published_test_vectors(n):
    assert oracle.blake3(bytes(i % 251 for i in range(n))).hex() == BLAKE3_PUBLISHED[n]


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
