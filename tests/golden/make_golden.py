"""Writes tests/golden/fixtures.json from the oracle (CPU restatement) on seeded inputs.

The reference has no fixed-value fixtures for this path (every test draws from OS entropy), so
these vectors are the oracle's own outputs, recorded so that later changes of the oracle or the
GPU path are caught without re-running the oracle. Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from ligesis_amd import synth  # noqa: E402


def main():
    out = {"note": "oracle outputs on SplitMix64-seeded inputs (ligesis_amd/synth.py)"}
    out["dft"] = []
    for logn, seed in [(3, 1), (6, 2)]:
        x = synth.field_elems(seed, 1 << logn)
        out["dft"].append({"seed": seed, "logn": logn, "out": [str(v) for v in oracle.dft(x)]})
    msgs = [b"", b"abc", bytes(range(64)), bytes(i % 251 for i in range(256)),
            bytes(i % 251 for i in range(1024)), bytes(i % 251 for i in range(1025))]
    out["blake3"] = [{"hex_in": m.hex(), "digest": oracle.blake3(m).hex()} for m in msgs]
    out["ligesis"] = []
    for (lm, ln, seed) in [(3, 10, 7), (1, 8, 3)]:
        inp = synth.Inputs(lm, ln, seed=seed)
        pk = oracle.ProvingKey(lm, ln, inp.a)
        cd = oracle.CommitData(pk, inp.f)
        proof, used = oracle.prove(pk, cd, inp.z, inp.fields, inp.ints)
        out["ligesis"].append({
            "log_m": lm, "log_n": ln, "seed": seed,
            "com_a": pk.com_a().hex(), "root": cd.root.hex(),
            "h_col0": [str(v) for v in cd.h()[:, 0]],
            "proof_len": len(proof), "proof_sha256": hashlib.sha256(proof).hexdigest(),
            "proof_head": proof[:112].hex(), "used": list(used),
        })
    sq = []
    for (lm, ln, seed) in [(5, 5, 11), (6, 6, 12)]:
        inp = synth.Inputs(lm, ln, seed=seed, n_fields=4, n_ints=4)
        pk = oracle.ProvingKey(lm, ln, inp.a, commit_a=False)
        cd = oracle.CommitData(pk, inp.f)
        sq.append({"log_m": lm, "log_n": ln, "seed": seed, "root": cd.root.hex()})
    out["commit_only"] = sq
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    _ = np


if __name__ == "__main__":
    main()
