"""Versioned on-wire records for LigeSIS commitments and proofs.

The reference has no serializer for either (``LigesisCommit([u8; 32])``, pcs/src/ligesis.rs:62;
``LigesisProof`` has private fields and only ``size_bytes``, pcs/src/ligesis.rs:99-138). The
library's proof bytes are a canonical encoding of the proof struct: fields in declaration
order, u64 little-endian lengths and integers, canonical field elements. These records add what
a receiver in another language needs before it parses them:

commitment (40 bytes)::

    b"LGSC" | u16 version | u8 log_m | u8 log_n | root[32]

proof (20 + len bytes)::

    b"LGSP" | u16 version | u8 log_m | u8 log_n | u64 len | u32 crc32(payload) | payload[len]

All integers are little-endian. ``version`` is 1 for the reference's scheme and 2 for a
commitment made with the Z_2^64 subset-sum hash (``ring64`` keys, commit only): the same record,
but a different hash, so a reader that knows only version 1 refuses it instead of treating the
root as a version-1 commitment. The shape fixes every length inside a proof's payload
(pcs/src/ligesis.rs:239-273), so a receiver checks it against its verifying key before parsing.
The CRC guards against transport damage only; soundness is the verifier's.
"""
from __future__ import annotations

import struct
import zlib

from . import LigesisError, LGS_ERR_ARG

VERSION = 1
VERSION_RING64 = 2
COMMIT_MAGIC = b"LGSC"
PROOF_MAGIC = b"LGSP"
_COMMIT = struct.Struct("<4sHBB32s")
_PROOF_HEAD = struct.Struct("<4sHBBQI")


def _shape(log_m: int, log_n: int):
    if not (0 <= log_m < 256 and 0 <= log_n < 256):
        raise LigesisError(LGS_ERR_ARG, f"wire: shape ({log_m}, {log_n}) does not fit the record")


def encode_commitment(root: bytes, log_m: int, log_n: int, ring64: bool = False) -> bytes:
    """``ring64``: the root of a Z_2^64-hash commitment (record version 2)."""
    _shape(log_m, log_n)
    if len(root) != 32:
        raise LigesisError(LGS_ERR_ARG, "wire: a commitment root is 32 bytes")
    return _COMMIT.pack(COMMIT_MAGIC, VERSION_RING64 if ring64 else VERSION, log_m, log_n, bytes(root))


def decode_commitment(b: bytes, ring64: bool = False):
    """-> (root, log_m, log_n). The record's version must be the one of the expected scheme
    (``ring64``: version 2, else 1)."""
    if len(b) != _COMMIT.size:
        raise LigesisError(LGS_ERR_ARG, f"wire: a commitment record is {_COMMIT.size} bytes, got {len(b)}")
    magic, ver, log_m, log_n, root = _COMMIT.unpack(b)
    if magic != COMMIT_MAGIC:
        raise LigesisError(LGS_ERR_ARG, "wire: not a commitment record")
    want = VERSION_RING64 if ring64 else VERSION
    if ver != want:
        raise LigesisError(LGS_ERR_ARG, f"wire: commitment record version {ver}, expected {want}")
    return root, log_m, log_n


def encode_proof(proof: bytes, log_m: int, log_n: int) -> bytes:
    _shape(log_m, log_n)
    payload = bytes(proof)
    head = _PROOF_HEAD.pack(PROOF_MAGIC, VERSION, log_m, log_n, len(payload), zlib.crc32(payload))
    return head + payload


def decode_proof(b: bytes):
    """-> (proof bytes, log_m, log_n)"""
    if len(b) < _PROOF_HEAD.size:
        raise LigesisError(LGS_ERR_ARG, "wire: proof record shorter than its header")
    magic, ver, log_m, log_n, length, crc = _PROOF_HEAD.unpack_from(b)
    if magic != PROOF_MAGIC:
        raise LigesisError(LGS_ERR_ARG, "wire: not a proof record")
    if ver != VERSION:
        raise LigesisError(LGS_ERR_ARG, f"wire: proof record version {ver}, this reader knows {VERSION}")
    payload = bytes(b[_PROOF_HEAD.size:])
    if len(payload) != length:
        raise LigesisError(LGS_ERR_ARG, f"wire: proof record holds {len(payload)} payload bytes, header says {length}")
    if zlib.crc32(payload) != crc:
        raise LigesisError(LGS_ERR_ARG, "wire: proof payload checksum mismatch")
    return payload, log_m, log_n
