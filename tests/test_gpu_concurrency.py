"""Kernels of different contexts running at once on the device give the same results as alone.

The MFMA hash kernels that copy their operands straight into LDS (global_load_lds) corrupted the
LDS of workgroups of other kernels sharing a CU (about 6 % of concurrent Basefold commitments
came out wrong while a mu = 24 commit ran on another stream); those launches now hold a CU's
whole LDS. This runs commits with those kernels on two contexts in background threads while
other threads repeat Basefold commitments (row transform + Merkle tree, both LDS users) and a
commit with the register-staged hash kernel, and checks every result against the one computed
alone.
"""
import threading
import time

import numpy as np
import pytest

from ligesis_amd import Basefold, Context, Ligesis, synth

pytestmark = pytest.mark.gpu

SECONDS = 3.0


def _device(words):
    import torch

    return torch.from_numpy(np.ascontiguousarray(words).view(np.int64)).to("cuda:0")


def test_concurrent_contexts_give_results_computed_alone():
    import torch

    cases = [  # B_I-like (2^18 base values, interleave 6), a-bar-like (2^17 extension values)
        (_device(synth.field_elems(1, 1 << 18)), 6, 1, False),
        (_device(synth.ext_elems(2, 1 << 17).reshape(-1)).reshape(-1, 2), 0, 1, True),
        (_device(synth.field_elems(3, 1 << 16)), 4, 1, False),
    ]
    ref_ctx = Context(0)
    want = [Basefold.commit_root(p, cr, li, ctx=ref_ctx, device=True, ext=e) for p, li, cr, e in cases]
    big = synth.Inputs(12, 12, seed=1, n_fields=4, n_ints=4)
    small = synth.Inputs(8, 10, seed=5, n_fields=4, n_ints=4)
    stop = threading.Event()
    errors, counts = [], {}

    def committer(name, inp, options):
        ctx = Context(0)
        for k, v in options.items():
            ctx.set_option(k, v)
        pk, _ = Ligesis.setup(inp.log_m, inp.log_n, inp.a, ctx=ctx, commit_only=True)
        f = _device(inp.f)
        ref, cd = Ligesis.commit(pk, f, device=True)
        cd.release()
        n = 0
        while not stop.is_set():
            r, cd = Ligesis.commit(pk, f, device=True)
            cd.release()
            n += 1
            if r != ref:
                errors.append(f"{name}: commit {n} root differs")
        counts[name] = n

    def basefold(w):
        ctx = Context(0)
        n = 0
        while not stop.is_set():
            k = (n + w) % len(cases)
            p, li, cr, e = cases[k]
            if Basefold.commit_root(p, cr, li, ctx=ctx, device=True, ext=e) != want[k]:
                errors.append(f"basefold worker {w}: commitment {n} (case {k}) root differs")
            n += 1
        counts[f"basefold{w}"] = n

    threads = [threading.Thread(target=committer, args=("dma hash, 8 waves", big, {})),
               threading.Thread(target=committer, args=("dma hash, 4 waves", big, {"hash": 1, "hash_impl": 12})),
               threading.Thread(target=committer, args=("register-staged hash", small, {"hash": 1, "hash_impl": 0})),
               threading.Thread(target=basefold, args=(0,)),
               threading.Thread(target=basefold, args=(1,))]
    for t in threads:
        t.start()
    time.sleep(SECONDS)
    stop.set()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    print("operations per thread:", counts, flush=True)
    assert all(v > 0 for v in counts.values()) and len(counts) == len(threads), counts
    assert not errors, errors[:10]
