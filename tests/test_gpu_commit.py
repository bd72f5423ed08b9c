"""GPU commit parity: F', H and the Merkle root against the oracle, bit for bit.

Restates the reference's commit (pcs/src/ligesis.rs:277-325) checks: the oracle runs the same
seeded inputs on the CPU; every stage the GPU keeps resident is read back and compared.
"""
import numpy as np
import pytest

import oracle
from ligesis_amd import Ligesis, LigesisError
from ligesis_amd import synth

pytestmark = pytest.mark.gpu

# (log_m, log_n): the reference test shape (3, 10), tiny and ragged-split shapes, a shape whose
# rows exceed one LDS tile (2n = 2^16: global NTT stages), and a larger square.
SHAPES = [(3, 10), (0, 7), (1, 5), (2, 6), (4, 8), (6, 9), (2, 15), (8, 10)]


def _commit_both(gpu_ctx, log_m, log_n, seed, variant):
    inp = synth.Inputs(log_m, log_n, seed=seed)
    print(f"commit parity shape=({log_m},{log_n}) seed={seed} hash={variant}", flush=True)
    gpu_ctx.set_option("hash", variant)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    root, cd = Ligesis.commit(pk, inp.f)
    opk = oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False)
    ocd = oracle.CommitData(opk, inp.f, keep_enc=True)
    return root, cd, ocd


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", [0, 2])
def test_commit_matches_oracle(gpu_ctx, shape, variant):
    log_m, log_n = shape
    root, cd, ocd = _commit_both(gpu_ctx, log_m, log_n, seed=11 + log_m + 3 * log_n, variant=variant)
    np.testing.assert_array_equal(cd.encoded(), ocd.enc())
    np.testing.assert_array_equal(cd.digests(), ocd.h())
    assert root == ocd.root
    gpu_ctx.set_option("hash", 0)


@pytest.mark.parametrize("shape", [(1, 3), (2, 4), (3, 5), (4, 11)])
def test_tree_and_split_sum_small_and_odd_splits(gpu_ctx, shape):
    """The K-split sum fused into the tree's first launch: trees of 16 and 32 leaves (one
    workgroup reaches the root), 64 and 4096 leaves (a second launch), every K-split count."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=90 + log_n)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    variants = (2,) if (2 << log_n) % 128 else (1, 2)
    try:
        for variant in variants:
            splits = [1 << k for k in range(log_m + 1)] if variant == 2 else [1 << k for k in range(log_m)]
            for ks in splits:
                print(f"tree shape={shape} hash={variant} ksplit={ks}", flush=True)
                gpu_ctx.set_option("hash", variant)
                gpu_ctx.set_option("hash_ksplit", ks)
                root, cd = Ligesis.commit(pk, inp.f)
                np.testing.assert_array_equal(cd.digests(), ocd.h())
                assert root == ocd.root, (variant, ks)
    finally:
        gpu_ctx.set_option("hash", 0)
        gpu_ctx.set_option("hash_ksplit", 0)


@pytest.mark.parametrize("shape", [(1, 17), (0, 19)])
def test_long_rows_match_oracle(gpu_ctx, shape):
    """Rows of 2^18 and 2^20 points: LDS blocks of 2^13, then one and three global radix-2 stages
    before the four-stage tail pass (ntt_impl 0), or radix-2 stages throughout (ntt_impl 1)."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=40 + log_n)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    try:
        for impl in (0, 1):
            print(f"long-row commit shape={shape} ntt_impl={impl}", flush=True)
            gpu_ctx.set_option("ntt_impl", impl)
            root, cd = Ligesis.commit(pk, inp.f)
            np.testing.assert_array_equal(cd.encoded(), ocd.enc())
            assert root == ocd.root
    finally:
        gpu_ctx.set_option("ntt_impl", -1)


@pytest.mark.parametrize("shape", [(3, 12), (7, 12), (2, 11)])
def test_every_row_transform_matches_oracle(gpu_ctx, shape):
    """Every row-transform variant gives the oracle's F'; ntt_impl 4 and 5 (the compile-time 2^13
    radix-8 and radix-16 kernels with per-pass twiddle tables) run at rows of 2^12 values and fall
    back to 3 elsewhere."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=60 + log_m + log_n)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    want = ocd.enc()
    try:
        for impl in (0, 1, 2, 3, 4, 5):
            print(f"row transform ntt_impl={impl} shape={shape}", flush=True)
            gpu_ctx.set_option("ntt_impl", impl)
            root, cd = Ligesis.commit(pk, inp.f)
            np.testing.assert_array_equal(cd.encoded(), want)
            assert root == ocd.root, impl
    finally:
        gpu_ctx.set_option("ntt_impl", -1)


def test_row_transform_extreme_values(gpu_ctx):
    """Rows of boundary values through every row-transform variant at rows of 2^12 (the
    compile-time kernel with Montgomery-form pass twiddles and shift/division power-of-two
    factors): p - 1, p - 2, 2^32 - 1, 2^32, 2^63, 1, alternating 0 / p - 1, and a random row."""
    log_m, log_n = 3, 12
    p = synth.P
    n = 1 << log_n
    rows = [np.full(n, p - 1), np.full(n, p - 2), np.full(n, 0xFFFFFFFF), np.full(n, 1 << 32),
            np.full(n, 1 << 63), np.ones(n), np.tile([0, p - 1], n // 2),
            synth.field_elems(91, n)]
    f = np.concatenate([np.asarray(r, dtype=np.uint64) for r in rows])
    inp = synth.Inputs(log_m, log_n, seed=92)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), f, keep_enc=True)
    try:
        for impl in (3, 4, 5):
            print(f"extreme rows ntt_impl={impl}", flush=True)
            gpu_ctx.set_option("ntt_impl", impl)
            root, cd = Ligesis.commit(pk, f)
            np.testing.assert_array_equal(cd.encoded(), ocd.enc())
            assert root == ocd.root, impl
    finally:
        gpu_ctx.set_option("ntt_impl", -1)


def test_mfma_and_valu_hash_agree(gpu_ctx):
    inp = synth.Inputs(5, 9, seed=3)
    pk, _ = Ligesis.setup(5, 9, inp.a, ctx=gpu_ctx, commit_only=True)
    out = {}
    for variant in (1, 2):
        gpu_ctx.set_option("hash", variant)
        for ks in (0, 1, 2, 4):
            gpu_ctx.set_option("hash_ksplit", ks)
            root, cd = Ligesis.commit(pk, inp.f)
            out[(variant, ks)] = (root, cd.digests())
    gpu_ctx.set_option("hash", 0)
    gpu_ctx.set_option("hash_ksplit", 0)
    ref = out[(2, 1)]
    for k, v in out.items():
        assert v[0] == ref[0], k
        np.testing.assert_array_equal(v[1], ref[1])


@pytest.mark.parametrize("shape", [(5, 9), (3, 12), (1, 7)])
def test_every_mfma_kernel_matches_oracle(gpu_ctx, shape):
    """All MFMA hash kernels (pipelined, basic, 256-wide tiles, load rings) and K-splits give the
    oracle's H; shapes where a kernel's tiling does not apply take its fallback."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=21 + log_n)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    want = ocd.h()
    gpu_ctx.set_option("hash", 1)
    try:
        for impl in list(range(13)) + [14, 15, 16, 17]:
            for ks in (0, 1, 2, 4):
                if ks > (1 << log_m) // 2:
                    continue
                print(f"hash impl={impl} ks={ks} shape={shape}", flush=True)
                gpu_ctx.set_option("hash_impl", impl)
                gpu_ctx.set_option("hash_ksplit", ks)
                root, cd = Ligesis.commit(pk, inp.f)
                assert root == ocd.root, (impl, ks)
                np.testing.assert_array_equal(cd.digests(), want)
    finally:
        gpu_ctx.set_option("hash", 0)
        gpu_ctx.set_option("hash_impl", -1)
        gpu_ctx.set_option("hash_ksplit", 0)


def test_extreme_field_values(gpu_ctx):
    """All-(p-1) and all-zero rows and an A of p-1 entries (largest digits / carries)."""
    log_m, log_n = 2, 7
    p = synth.P
    f = np.full(1 << (log_m + log_n), p - 1, np.uint64)
    f[: 1 << log_n] = 0
    a = np.full(32 * (64 << log_m), p - 1, np.uint64)
    a[::3] = 0x7F7F7F7F7F7F7F7F
    a[1::3] = 0x7F7F7F7F7F7F7F80
    for variant in (1, 2):
        gpu_ctx.set_option("hash", variant)
        pk, _ = Ligesis.setup(log_m, log_n, a, ctx=gpu_ctx, commit_only=True)
        root, cd = Ligesis.commit(pk, f)
        opk = oracle.ProvingKey(log_m, log_n, a, commit_a=False)
        ocd = oracle.CommitData(opk, f, keep_enc=True)
        np.testing.assert_array_equal(cd.digests(), ocd.h())
        assert root == ocd.root
    gpu_ctx.set_option("hash", 0)


def test_commit_batch_matches_single(gpu_ctx):
    shapes = [(3, 10), (2, 6), (5, 5), (4, 9)]
    keys, fs, expect = [], [], []
    for i, (lm, ln) in enumerate(shapes):
        inp = synth.Inputs(lm, ln, seed=100 + i)
        pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
        keys.append(pk)
        fs.append(inp.f)
        expect.append(Ligesis.commit(pk, inp.f)[0])
    res = Ligesis.commit_batch(keys, fs, ctx=gpu_ctx)
    assert [r for r, _ in res] == expect


def test_commit_batch_of_wide_shapes_matches_single(gpu_ctx):
    """Every job with 2n a multiple of 256: the grouped hash runs 256-column tiles (8 waves); jobs
    whose K-splits have 2 or 4k slabs take the 3- or 4-slot ring."""
    shapes = [(1, 9), (2, 9), (3, 10), (1, 7), (5, 8), (0, 8)]
    keys, fs, expect = [], [], []
    for i, (lm, ln) in enumerate(shapes):
        inp = synth.Inputs(lm, ln, seed=130 + i, n_fields=4, n_ints=4)
        pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
        root, cd = Ligesis.commit(pk, inp.f)
        keys.append(pk)
        fs.append(inp.f)
        expect.append((root, cd.digests()))
    print("batched commit of wide shapes", shapes, flush=True)
    res = Ligesis.commit_batch(keys, fs, ctx=gpu_ctx)
    for (r, cd), (er, eh) in zip(res, expect):
        assert r == er
        np.testing.assert_array_equal(cd.digests(), eh)


def test_commit_rejects_bad_length(gpu_ctx):
    inp = synth.Inputs(2, 6, seed=5)
    pk, _ = Ligesis.setup(2, 6, inp.a, ctx=gpu_ctx, commit_only=True)
    with pytest.raises(LigesisError):
        Ligesis.commit(pk, inp.f[:-1])


def test_full_size_square_mu24_chain(gpu_ctx):
    """mu = 24 at 2^12 x 2^12 (the commit benchmark shape): sampled rows of F' equal the oracle's
    row encoding, sampled columns of H equal the oracle's subset-sum hash of the bits of F', and
    the root equals the oracle's Merkle root over the GPU's H. Both hash kernels agree."""
    log_m = log_n = 12
    inp = synth.Inputs(log_m, log_n, seed=24, n_fields=4, n_ints=4)
    print("full-size mu=24 square chain check", flush=True)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    gpu_ctx.set_option("hash", 1)
    root, cd = Ligesis.commit(pk, inp.f)
    n, m = 1 << log_n, 1 << log_m
    enc = cd.encoded()
    h = cd.digests()
    rng = np.random.default_rng(0)
    rows = rng.choice(m, 6, replace=False)
    f2 = inp.f.reshape(m, n)
    for i in rows:
        np.testing.assert_array_equal(enc[i], oracle.rs_encode_rows(f2[i], 1, n)[0])
    cols = rng.choice(2 * n, 4, replace=False)
    bits = ((enc[:, cols, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
    bits = bits.transpose(1, 0, 2).reshape(len(cols), 64 * m)
    np.testing.assert_array_equal(h[:, cols], oracle.hash_columns(inp.a.reshape(32, 64 * m), bits))
    leaves = h.T.copy().view(np.uint8).reshape(2 * n, 256)
    assert oracle.merkle_root(leaves) == root
    gpu_ctx.set_option("hash", 2)
    root2, cd2 = Ligesis.commit(pk, inp.f)
    assert root2 == root
    gpu_ctx.set_option("hash", 0)


def test_handles_outlive_context_shutdown():
    """Regression: keys and commitments released after lgs_shutdown must not touch freed state
    (the context is reference-counted by its keys and commitments)."""
    import ligesis_amd as L

    ctx = L.Context(0)
    inp = synth.Inputs(2, 6, seed=9)
    pk, _ = L.Ligesis.setup(2, 6, inp.a, ctx=ctx, commit_only=True)
    root, cd = L.Ligesis.commit(pk, inp.f)
    ctx.close()
    assert cd.digests().shape == (32, 128)  # still usable after the caller's shutdown
    pk.release()
    cd.release()


def test_batched_commit_mixed_shapes_matches_single(gpu_ctx):
    """Grouped launches over many shapes (including ones that take the per-polynomial path)
    give the same roots and digests as committing each polynomial alone."""
    shapes = [(8, 8), (7, 8), (8, 9), (3, 10), (1, 5), (6, 12), (9, 10), (2, 11), (0, 7), (5, 6),
              (10, 10), (4, 13)]
    keys, fs, expect = [], [], []
    for i, (lm, ln) in enumerate(shapes):
        inp = synth.Inputs(lm, ln, seed=300 + i, n_fields=4, n_ints=4)
        pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
        root, cd = Ligesis.commit(pk, inp.f)
        keys.append(pk)
        fs.append(inp.f)
        expect.append((root, cd.digests()))
    print("batched commit of", len(shapes), "shapes", flush=True)
    res = Ligesis.commit_batch(keys, fs, ctx=gpu_ctx)
    for (r, cd), (er, eh) in zip(res, expect):
        assert r == er
        np.testing.assert_array_equal(cd.digests(), eh)


@pytest.mark.parametrize("shape", [(13, 13), (9, 17)])
def test_full_size_mu26_chain(gpu_ctx, shape):
    """mu = 26, the largest shapes served: rows of 2^14 and 2^18 points (transforms longer than
    one LDS tile, at full height), W = 2^19 and 2^15 hash inputs per column. Sampled rows of F'
    equal the oracle's row encoding, sampled columns of H equal the oracle's subset-sum hash of
    the bits of F', and the root equals the oracle's Merkle root over the GPU's H."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=26 + log_m, n_fields=4, n_ints=4)
    print(f"full-size mu=26 chain check shape={shape}", flush=True)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    root, cd = Ligesis.commit(pk, inp.f)
    n, m = 1 << log_n, 1 << log_m
    enc = cd.encoded()
    h = cd.digests()
    rng = np.random.default_rng(1)
    rows = rng.choice(m, 4, replace=False)
    f2 = inp.f.reshape(m, n)
    for i in rows:
        np.testing.assert_array_equal(enc[i], oracle.rs_encode_rows(f2[i], 1, n)[0])
    cols = rng.choice(2 * n, 3, replace=False)
    bits = ((enc[:, cols, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).astype(np.uint8)
    bits = bits.transpose(1, 0, 2).reshape(len(cols), 64 * m)
    np.testing.assert_array_equal(h[:, cols], oracle.hash_columns(inp.a.reshape(32, 64 * m), bits))
    leaves = h.T.copy().view(np.uint8).reshape(2 * n, 256)
    assert oracle.merkle_root(leaves) == root
    del enc
    cd.release()
