"""Versioned commitment and proof records: their byte layout, round trips of an oracle proof,
and rejection of damaged or foreign records."""
import struct
import zlib

import pytest

import oracle
from ligesis_amd import LigesisError, synth, wire

from . import proof_format as pfmt


def test_commitment_record_layout():
    root = bytes(range(32))
    rec = wire.encode_commitment(root, 3, 10)
    assert rec == b"LGSC" + b"\x01\x00" + b"\x03\x0a" + root
    assert wire.decode_commitment(rec) == (root, 3, 10)


def _oracle_proof():
    inp = synth.Inputs(3, 10, seed=8)
    pk = oracle.ProvingKey(3, 10, inp.a)
    cd = oracle.CommitData(pk, inp.f)
    proof, _ = oracle.prove(pk, cd, inp.z, inp.fields, inp.ints)
    return proof


def test_proof_record_round_trip():
    proof = _oracle_proof()
    rec = wire.encode_proof(proof, 3, 10)
    assert rec[:4] == b"LGSP" and struct.unpack_from("<HBBQI", rec, 4) == (1, 3, 10, len(proof), zlib.crc32(proof))
    got, lm, ln = wire.decode_proof(rec)
    assert (got, lm, ln) == (proof, 3, 10)
    assert pfmt.encode(pfmt.decode(got)) == proof  # the payload is the canonical proof encoding


@pytest.mark.parametrize("damage", ["payload", "magic", "version", "truncated", "length"])
def test_damaged_proof_records_rejected(damage):
    rec = bytearray(wire.encode_proof(_oracle_proof(), 3, 10))
    if damage == "payload":
        rec[-1] ^= 1
    elif damage == "magic":
        rec[0:4] = b"LGSC"
    elif damage == "version":
        rec[4] = 2
    elif damage == "truncated":
        rec = rec[:-8]
    else:
        struct.pack_into("<Q", rec, 8, struct.unpack_from("<Q", rec, 8)[0] + 1)
    with pytest.raises(LigesisError):
        wire.decode_proof(bytes(rec))


@pytest.mark.parametrize("rec", [b"LGSC\x01\x00\x03\x0a" + bytes(31), b"LGSP\x01\x00\x03\x0a" + bytes(32),
                                 b"LGSC\x02\x00\x03\x0a" + bytes(32)])
def test_bad_commitment_records_rejected(rec):
    with pytest.raises(LigesisError):
        wire.decode_commitment(rec)


def test_encoders_reject_what_does_not_fit():
    with pytest.raises(LigesisError):
        wire.encode_commitment(bytes(31), 3, 10)
    with pytest.raises(LigesisError):
        wire.encode_proof(b"", 256, 10)


def test_ring64_commitment_record_is_version_2_and_refused_as_version_1():
    root = bytes(range(32, 64))
    rec = wire.encode_commitment(root, 4, 9, ring64=True)
    assert rec == b"LGSC" + b"\x02\x00" + b"\x04\x09" + root
    assert wire.decode_commitment(rec, ring64=True) == (root, 4, 9)
    with pytest.raises(LigesisError):
        wire.decode_commitment(rec)
    with pytest.raises(LigesisError):
        wire.decode_commitment(wire.encode_commitment(root, 4, 9), ring64=True)
