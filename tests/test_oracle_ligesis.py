"""Pins the oracle's LigeSIS (oracle/ligesis.hpp) with the reference tests
pcs/src/ligesis.rs:807-868 (round trip with proof.y == f(z), tampered eval / sumcheck round /
H column rejected) on seeded inputs, plus transcript-length and encoding checks."""
import numpy as np
import pytest

import oracle
from ligesis_amd import synth

from . import proof_format as pfmt

P = synth.P


@pytest.fixture(scope="module")
def inst():
    log_m, log_n = 3, 10  # the reference test shape (ligesis.rs:810)
    inp = synth.Inputs(log_m, log_n, seed=7)
    pk = oracle.ProvingKey(log_m, log_n, inp.a)
    cd = oracle.CommitData(pk, inp.f)
    proof, used = oracle.prove(pk, cd, inp.z, inp.fields, inp.ints)
    return log_m, log_n, inp, pk, cd, proof, used


def _verify(inst, proof):
    log_m, log_n, inp, pk, cd, _, _ = inst
    return oracle.verify(log_m, log_n, pk.com_a(), cd.root, inp.z, proof, inp.fields, inp.ints)


def test_round_trip_and_claimed_value(inst):
    log_m, log_n, inp, pk, cd, proof, used = inst
    d = pfmt.decode(proof)
    f_ef = np.stack([inp.f, np.zeros_like(inp.f)], axis=1)
    assert d["y"] == tuple(int(x) for x in oracle.mle_eval(f_ef, inp.z))
    assert _verify(inst, proof)
    assert pfmt.encode(d) == proof
    assert oracle.proof_reencode(proof) == proof


def test_transcript_lengths(inst):
    log_m, log_n, _, _, _, _, used = inst
    log_w, log_s = 6 + log_m, 7
    rounds = max(11 + log_m, 13 + log_m, log_n) - 6
    fields = ((log_w + log_s) + 5 + (log_w + log_s) + log_s + 2 * log_w + (log_n + 1) + log_n
              + 1 + (log_w + log_s) + 1 + log_n + 3 + rounds)
    assert used == (fields, 4 * 128 + 100)


def test_tampered_eval_rejected(inst):
    d = pfmt.decode(inst[5])
    d["y"] = ((d["y"][0] + 1) % P, d["y"][1])
    assert not _verify(inst, pfmt.encode(d))


def test_tampered_sumcheck_rejected(inst):
    d = pfmt.decode(inst[5])
    e = d["sc_hash"]["round_polys"][0][0]
    d["sc_hash"]["round_polys"][0][0] = ((e[0] + 1) % P, e[1])
    assert not _verify(inst, pfmt.encode(d))


def test_tampered_h_column_rejected(inst):
    d = pfmt.decode(inst[5])
    d["h_cols"][0][0] = (d["h_cols"][0][0] + 1) % P
    assert not _verify(inst, pfmt.encode(d))


def test_tampered_batch_rejected(inst):
    d = pfmt.decode(inst[5])
    e = d["batch"]["reduction_msgs"][1][0]
    d["batch"]["reduction_msgs"][1][0] = ((e[0] + 1) % P, e[1])
    assert not _verify(inst, pfmt.encode(d))
    d = pfmt.decode(inst[5])
    k, v = d["batch"]["base_bf"][1]["values"][0]
    v[3] = 1 - v[3] if v[3] in (0, 1) else 0
    assert not _verify(inst, pfmt.encode(d))


def test_truncated_proof_rejected(inst):
    assert not _verify(inst, inst[5][:-8])


def test_shape_window():
    # ligesis.rs:248-251 and README.md:187-189: 7 + log_m <= log_n <= 17 + log_m
    for lm in range(0, 10):
        for ln in range(1, 30):
            assert oracle.shape_ok(lm, ln) == (7 + lm <= ln <= 17 + lm)
    with pytest.raises(RuntimeError):
        oracle.ProvingKey(12, 12, np.zeros(32 * (64 << 12), np.uint64))
