"""The Z_2^64 subset-sum commitment (``ring64`` keys, LGS_SETUP_RING64) against the oracle's.

A paper-level variant the reference leaves out (pcs/src/subset_sum.rs:10-12): the commit of
pcs/src/ligesis.rs:277-325 with H = A . B summed mod 2^64 over any 64-bit A. The oracle's
restatement is pinned against big-integer sums mod 2^64 (tests/test_oracle_subset_sum.py).
"""
import numpy as np
import pytest

import oracle
from ligesis_amd import Ligesis, LigesisError, RandomOracle, Server, synth

pytestmark = pytest.mark.gpu

C = 32
P = 2**64 - 2**32 + 1


def _a(lm: int, seed: int) -> np.ndarray:
    """64-bit words, most of them outside [0, p), with the balanced-digit boundaries."""
    w = 64 << lm
    a = synth.splitmix64(seed, C * w).reshape(C, w)
    a[0, :] = 2**64 - 1
    a[1, :] = P
    a[2, :] = 2**63
    a[3, ::2] = 0x7F7F7F7F7F7F7F7F
    a[3, 1::2] = 0x7F7F7F7F7F7F7F80
    a[4, ::2] = 0x8080808080808080
    a[4, 1::2] = 0x807F7F7F7F7F7F7F
    return a.reshape(-1)


@pytest.mark.parametrize("shape", [(1, 6), (3, 10), (5, 9), (8, 10), (2, 8)])
def test_ring64_commit_matches_oracle(gpu_ctx, shape):
    lm, ln = shape
    inp = synth.Inputs(lm, ln, seed=50 + lm + ln, n_fields=4, n_ints=4)
    a = _a(lm, 60 + lm)
    pk, _ = Ligesis.setup(lm, ln, a, ctx=gpu_ctx, ring64=True)
    ocd = oracle.CommitData(oracle.ProvingKey(lm, ln, a, commit_a=False, ring64=True), inp.f, keep_enc=True)
    field_root = oracle.CommitData(oracle.ProvingKey(lm, ln, a % P, commit_a=False), inp.f).root
    assert ocd.root != field_root
    try:
        for impl in (17, 16, 15, 14, 12, 11):
            for ks in (0, 1, 2, 4):
                if ks > (1 << lm) // 2:
                    continue
                print(f"ring64 commit shape={shape} hash_impl={impl} ks={ks}", flush=True)
                gpu_ctx.set_option("hash_impl", impl)
                gpu_ctx.set_option("hash_ksplit", ks)
                root, cd = Ligesis.commit(pk, inp.f)
                np.testing.assert_array_equal(cd.digests(), ocd.h())
                assert root == ocd.root, (impl, ks)
    finally:
        gpu_ctx.set_option("hash_impl", -1)
        gpu_ctx.set_option("hash_ksplit", 0)


def test_ring64_full_size_square(gpu_ctx):
    """mu = 24 (2^12 x 2^12): sampled H columns against the oracle's Z_2^64 hash of the GPU's F'
    columns (F' is the field-mode transform, checked elsewhere), and the root against the oracle's
    tree over the GPU's H (leaves = raw words)."""
    lm = ln = 12
    inp = synth.Inputs(lm, ln, seed=1, n_fields=4, n_ints=4)
    a = _a(lm, 71)
    pk, _ = Ligesis.setup(lm, ln, a, ctx=gpu_ctx, ring64=True)
    print("ring64 commit at 2^12 x 2^12", flush=True)
    root, cd = Ligesis.commit(pk, inp.f)
    enc, h = cd.encoded(), cd.digests()
    cols = [0, 1, 4095, 4096, 8191]
    bits = np.stack([np.unpackbits(enc[:, j].astype("<u8").view(np.uint8), bitorder="little") for j in cols])
    np.testing.assert_array_equal(h[:, cols], oracle.hash_columns(a.reshape(C, -1), bits, ring64=True))
    leaves = np.frombuffer(np.ascontiguousarray(h.T).astype("<u8").tobytes(), np.uint8).reshape(8192, 256)
    assert root == oracle.merkle_root(leaves)


def test_ring64_keys_commit_only_and_other_paths(gpu_ctx):
    lm, ln = 3, 9
    inp = synth.Inputs(lm, ln, seed=81)
    a = _a(lm, 82)
    pk, vk = Ligesis.setup(lm, ln, a, ctx=gpu_ctx, ring64=True)
    assert pk.commit_only and vk.com_a == b""
    root, cd = Ligesis.commit(pk, inp.f)
    want = oracle.CommitData(oracle.ProvingKey(lm, ln, a, commit_a=False, ring64=True), inp.f).root
    assert root == want
    with pytest.raises(LigesisError):  # no LigeSIS opening over Z_2^64
        Ligesis.prove(pk, cd, inp.z, RandomOracle(inp.fields, inp.ints))
    with pytest.raises(LigesisError):  # the MFMA kernels need 2n >= 128
        Ligesis.setup(1, 5, synth.splitmix64(1, C * 128), ctx=gpu_ctx, ring64=True)
    with pytest.raises(LigesisError):  # a field-mode key takes field elements only
        Ligesis.setup(lm, ln, a, ctx=gpu_ctx, commit_only=True)
    # batched commits (a ring key takes the per-polynomial path) and a served commit-only request
    pk2, _ = Ligesis.setup(lm, ln, a % P, ctx=gpu_ctx, commit_only=True)
    res = Ligesis.commit_batch([pk, pk2], [inp.f, inp.f], ctx=gpu_ctx)
    assert res[0][0] == want and res[1][0] == Ligesis.commit(pk2, inp.f)[0] != want
    srv = Server(0, hbm_budget=4 << 30, prove_workers=1, max_batch=2)
    try:
        t = srv.submit_commit(pk, inp.f)
        got, proof, _ = srv.poll(t)
        assert got == want and proof == b""
        srv.release(t)
    finally:
        srv.stop()


def test_ring64_noalloc_matches_commit(gpu_ctx):
    import torch

    lm, ln = 4, 10
    inp = synth.Inputs(lm, ln, seed=91, n_fields=4, n_ints=4)
    a = _a(lm, 92)
    pk, _ = Ligesis.setup(lm, ln, a, ctx=gpu_ctx, ring64=True)
    want, _ = Ligesis.commit(pk, inp.f)
    m, n = 1 << lm, 1 << ln
    f = torch.from_numpy(inp.f.view(np.int64)).to("cuda:0")
    enc = torch.empty(m * 2 * n, dtype=torch.int64, device="cuda:0")
    h = torch.empty(C * 2 * n, dtype=torch.int64, device="cuda:0")
    tree = torch.empty((4 * n - 1) * 32, dtype=torch.uint8, device="cuda:0")
    assert Ligesis.commit_noalloc(pk, f, enc, h, tree) == want
