"""Writes tests/golden/full_size.json: oracle (CPU restatement) outputs at the benchmark sizes.

The GPU tests compare the library with these digests at the sizes the benchmarks run, where the
oracle itself is too slow to run inside a test (a mu = 24 commit takes about 140 s on one core,
as the reference's own commit does, README.md:175). Inputs are the seeded ones of
``ligesis_amd/synth.py``; every case names its seeds, so the GPU side regenerates them.

Cases:
- ``commit``: mu = 24 at 2^12 x 2^12 (seed 1, the benchmark's commit) and mu = 20 at 2^10 x 2^10
  (seed 20): the root and the sha256 of F' (m x 2n canonical u64, row-major) and of H
  (32 x 2n, row-major), pcs/src/ligesis.rs:277-325;
- ``prove``: mu = 24 at 2^8 x 2^16 (seed 5) and 2^7 x 2^17 (seed 6), the reference-valid
  opening shapes: com_a, root, proof length, sha256 and first 112 bytes, oracle draws used,
  pcs/src/ligesis.rs:347-549;
- ``batch``: the 64-polynomial batched-commit workload (mu uniform in [16, 24], m = 2^floor(mu/2),
  n = 2^ceil(mu/2); A per shape from seed 1000 + 64 log_m + log_n, f cut from one stream of
  seed 77 in order): every root.

Run: python tests/golden/make_full_size.py [--procs 6]   (about 10 minutes on 6 cores)
"""
from __future__ import annotations

import argparse
import hashlib
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "full_size.json")


def batch_shapes(count: int = 64, seed: int = 2024):
    """The batched-commit workload's shapes, in order (the same draw as the benchmark's)."""
    rng = np.random.default_rng(seed)
    mus = rng.integers(16, 25, size=count)
    return [(int(mu) // 2, int(mu) - int(mu) // 2) for mu in mus]


def batch_inputs(i: int, shapes, pool=None):
    """A and f of polynomial i of the batched workload (``pool``: the whole f stream, if made)."""
    from ligesis_amd import synth

    lm, ln = shapes[i]
    a = synth.field_elems(1000 + 64 * lm + ln, 32 * (64 << lm))
    off = sum(1 << (x + y) for x, y in shapes[:i])
    end = off + (1 << (lm + ln))
    f = (pool if pool is not None else synth.field_elems(77, end))[off:end]
    return a, f


def _sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def run_case(case):
    import oracle
    from ligesis_amd import synth

    t0 = time.time()
    kind = case[0]
    if kind == "commit":
        _, lm, ln, seed = case
        inp = synth.Inputs(lm, ln, seed=seed, n_fields=4, n_ints=4)
        pk = oracle.ProvingKey(lm, ln, inp.a, commit_a=False)
        cd = oracle.CommitData(pk, inp.f, keep_enc=True)
        res = {"log_m": lm, "log_n": ln, "seed": seed, "root": cd.root.hex(),
               "enc_sha256": _sha(cd.enc()), "h_sha256": _sha(cd.h())}
    elif kind == "prove":
        _, lm, ln, seed = case
        inp = synth.Inputs(lm, ln, seed=seed)
        pk = oracle.ProvingKey(lm, ln, inp.a)
        cd = oracle.CommitData(pk, inp.f)
        proof, used = oracle.prove(pk, cd, inp.z, inp.fields, inp.ints)
        ok = oracle.verify(lm, ln, pk.com_a(), cd.root, inp.z, proof, inp.fields, inp.ints)
        assert ok, "oracle verifier rejected the oracle's own proof"
        res = {"log_m": lm, "log_n": ln, "seed": seed, "com_a": pk.com_a().hex(), "root": cd.root.hex(),
               "proof_len": len(proof), "proof_sha256": hashlib.sha256(proof).hexdigest(),
               "proof_head": proof[:112].hex(), "used": list(used)}
    elif kind == "batch":
        _, idxs = case
        shapes = batch_shapes()
        pool = synth.field_elems(77, sum(1 << (x + y) for x, y in shapes))
        res = []
        for i in idxs:
            lm, ln = shapes[i]
            a, f = batch_inputs(i, shapes, pool)
            pk = oracle.ProvingKey(lm, ln, a, commit_a=False)
            cd = oracle.CommitData(pk, f)
            res.append({"index": i, "log_m": lm, "log_n": ln, "root": cd.root.hex()})
    else:
        raise ValueError(kind)
    print(f"done {case!r:.60} in {time.time() - t0:.0f} s", flush=True)
    return case, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--procs", type=int, default=6)
    args = ap.parse_args()
    shapes = batch_shapes()
    # batch members spread over processes by size (largest first)
    order = sorted(range(len(shapes)), key=lambda i: -(shapes[i][0] + shapes[i][1]))
    bins = [[] for _ in range(8)]
    load = [0] * 8
    for i in order:
        b = load.index(min(load))
        bins[b].append(i)
        load[b] += 1 << (shapes[i][0] + shapes[i][1])
    cases = [("commit", 12, 12, 1), ("prove", 8, 16, 5), ("prove", 7, 17, 6), ("commit", 10, 10, 20)]
    cases += [("batch", tuple(sorted(b))) for b in bins if b]
    out = {"note": "oracle outputs at benchmark sizes on SplitMix64-seeded inputs (ligesis_amd/synth.py); "
                   "written by tests/golden/make_full_size.py",
           "commit": [], "prove": [], "batch": []}
    with mp.get_context("spawn").Pool(args.procs) as pool:
        for case, res in pool.imap_unordered(run_case, cases):
            if case[0] == "batch":
                out["batch"].extend(res)
            else:
                out[case[0]].append(res)
    out["batch"].sort(key=lambda r: r["index"])
    out["commit"].sort(key=lambda r: -r["log_m"])
    out["prove"].sort(key=lambda r: -r["log_m"])
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
