"""The oracle reproduces the committed fixtures (tests/golden/fixtures.json), and the GPU path
reproduces the LigeSIS ones without running the oracle."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from ligesis_amd import synth

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


def test_oracle_dft_fixtures():
    for d in FIX["dft"]:
        x = synth.field_elems(d["seed"], 1 << d["logn"])
        assert [str(v) for v in oracle.dft(x)] == d["out"]


def test_oracle_blake3_fixtures():
    for d in FIX["blake3"]:
        assert oracle.blake3(bytes.fromhex(d["hex_in"])).hex() == d["digest"]


def test_oracle_ligesis_fixtures():
    for d in FIX["ligesis"]:
        inp = synth.Inputs(d["log_m"], d["log_n"], seed=d["seed"])
        pk = oracle.ProvingKey(d["log_m"], d["log_n"], inp.a)
        cd = oracle.CommitData(pk, inp.f)
        assert pk.com_a().hex() == d["com_a"]
        assert cd.root.hex() == d["root"]
        proof, used = oracle.prove(pk, cd, inp.z, inp.fields, inp.ints)
        assert hashlib.sha256(proof).hexdigest() == d["proof_sha256"]
        assert list(used) == d["used"]


@pytest.mark.gpu
def test_gpu_ligesis_fixtures(gpu_ctx):
    from ligesis_amd import Ligesis, RandomOracle

    for d in FIX["ligesis"]:
        print(f"golden shape=({d['log_m']},{d['log_n']})", flush=True)
        inp = synth.Inputs(d["log_m"], d["log_n"], seed=d["seed"])
        pk, vk = Ligesis.setup(d["log_m"], d["log_n"], inp.a, ctx=gpu_ctx)
        root, cd = Ligesis.commit(pk, inp.f)
        assert vk.com_a.hex() == d["com_a"]
        assert root.hex() == d["root"]
        assert [str(v) for v in cd.digests()[:, 0]] == d["h_col0"]
        proof = Ligesis.prove(pk, cd, inp.z, RandomOracle(inp.fields, inp.ints))
        assert proof[:112].hex() == d["proof_head"]
        assert len(proof) == d["proof_len"]
        assert hashlib.sha256(proof).hexdigest() == d["proof_sha256"]
    for d in FIX["commit_only"]:
        inp = synth.Inputs(d["log_m"], d["log_n"], seed=d["seed"], n_fields=4, n_ints=4)
        pk, _ = Ligesis.setup(d["log_m"], d["log_n"], inp.a, ctx=gpu_ctx, commit_only=True)
        root, _ = Ligesis.commit(pk, inp.f)
        assert root.hex() == d["root"]
    _ = np
