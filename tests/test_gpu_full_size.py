"""Byte equality with the oracle at the benchmark sizes, through digests the oracle wrote on the CPU
(tests/golden/full_size.json, from tests/golden/make_full_size.py; the oracle needs minutes per
case there, so it is not run here):

- the mu = 24 square commit (2^12 x 2^12, the commit benchmark's seeded input) and the mu = 20
  square commit (2^10 x 2^10): root, sha256 of all of F' and of all of H
  (pcs/src/ligesis.rs:277-325);
- full eval-proofs at mu = 24 on the reference-valid shapes 2^8 x 2^16 and 2^7 x 2^17: com_a,
  root, proof length, sha256, leading bytes and oracle draws (pcs/src/ligesis.rs:347-549);
- the 64-polynomial batched commit (mu uniform in [16, 24]): every root, batched and one by one.
"""
import hashlib
import json
import os

import numpy as np
import pytest
import torch

from ligesis_amd import Ligesis, RandomOracle
from ligesis_amd import synth

pytestmark = pytest.mark.gpu

FIX_PATH = os.path.join(os.path.dirname(__file__), "golden", "full_size.json")
FIX = json.load(open(FIX_PATH)) if os.path.exists(FIX_PATH) else {"commit": [], "prove": [], "batch": []}


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


@pytest.mark.parametrize("case", FIX["commit"], ids=lambda d: f"{d['log_m']}x{d['log_n']}")
def test_square_commit_matches_oracle_digests(gpu_ctx, case):
    lm, ln = case["log_m"], case["log_n"]
    print(f"full-size commit shape=({lm},{ln}) seed={case['seed']}", flush=True)
    inp = synth.Inputs(lm, ln, seed=case["seed"], n_fields=4, n_ints=4)
    pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
    root, cd = Ligesis.commit(pk, inp.f)
    assert root.hex() == case["root"]
    assert _sha(cd.encoded()) == case["enc_sha256"]
    assert _sha(cd.digests()) == case["h_sha256"]
    cd.release()
    pk.release()


@pytest.mark.parametrize("case", FIX["prove"], ids=lambda d: f"{d['log_m']}x{d['log_n']}")
def test_mu24_proof_matches_oracle_bytes(gpu_ctx, case):
    lm, ln = case["log_m"], case["log_n"]
    print(f"full-size prove shape=({lm},{ln}) seed={case['seed']}", flush=True)
    inp = synth.Inputs(lm, ln, seed=case["seed"])
    pk, vk = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx)
    assert vk.com_a.hex() == case["com_a"]
    root, cd = Ligesis.commit(pk, inp.f)
    assert root.hex() == case["root"]
    ro = RandomOracle(inp.fields, inp.ints)
    proof = Ligesis.prove(pk, cd, inp.z, ro)
    assert len(proof) == case["proof_len"]
    assert proof[:112].hex() == case["proof_head"]
    assert hashlib.sha256(proof).hexdigest() == case["proof_sha256"]
    assert list(ro.drawn()) == case["used"]
    cd.release()
    pk.release()


def test_batched_commit_of_64_polys_matches_oracle_roots(gpu_ctx):
    if len(FIX["batch"]) != 64:
        pytest.skip("batch fixtures not written")
    # the batched workload's draw (tests/golden/make_full_size.py batch_shapes, bench.py bench_batch)
    mus = np.random.default_rng(2024).integers(16, 25, size=64)
    shapes = [(int(mu) // 2, int(mu) - int(mu) // 2) for mu in mus]
    assert [(d["log_m"], d["log_n"]) for d in FIX["batch"]] == shapes
    print("full-size batched commit of 64 polys", flush=True)
    keys = {}
    for lm, ln in sorted(set(shapes)):
        a = synth.field_elems(1000 + 64 * lm + ln, 32 * (64 << lm))
        keys[(lm, ln)] = Ligesis.setup(lm, ln, a, ctx=gpu_ctx, commit_only=True)[0]
    pool = synth.field_elems(77, sum(1 << (lm + ln) for lm, ln in shapes))
    fs, off = [], 0
    for lm, ln in shapes:
        sz = 1 << (lm + ln)
        fs.append(torch.from_numpy(pool[off:off + sz].view(np.int64)).to("cuda:0"))
        off += sz
    ks = [keys[s] for s in shapes]
    res = Ligesis.commit_batch(ks, fs, ctx=gpu_ctx, device=True)
    want = [d["root"] for d in FIX["batch"]]
    assert [r.hex() for r, _ in res] == want
    for _, cd in res:
        cd.release()
    for i in (0, 17, 63):  # members committed alone take the same schedule as the benchmark's
        r, cd = Ligesis.commit(ks[i], fs[i], device=True)
        assert r.hex() == want[i]
        cd.release()
    del fs
    for k in keys.values():
        k.release()
