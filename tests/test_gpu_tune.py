"""Autotuner on the GPU: the schedule it picks is valid, is what later commits use, survives a
save / load round trip, and leaves every output byte unchanged (root, F', H against the oracle)."""
import numpy as np
import pytest

import ligesis_amd
import oracle
from ligesis_amd import Ligesis, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean_table():
    ligesis_amd.tune_clear()
    yield
    ligesis_amd.tune_clear()


@pytest.mark.parametrize("shape", [(5, 9), (2, 6), (3, 12)])
def test_autotuned_commit_matches_oracle(gpu_ctx, shape, tmp_path):
    log_m, log_n = shape
    print(f"autotune shape={shape}", flush=True)
    inp = synth.Inputs(log_m, log_n, seed=40 + log_m)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    best = gpu_ctx.autotune(pk, reps=2)
    assert best["ms"] > 0
    assert best["hash_variant"] in (1, 2) and best["hash_impl"] in (*range(13), 14, 15, 16, 17)
    assert 0 <= best["ntt_impl"] <= 5 and 6 <= best["merkle_lg"] <= 8 and best["pipeline"] in (1, 2, 4, 8)
    cfg = gpu_ctx.tuning(log_m, log_n)
    assert cfg == {k: best[k] for k in ligesis_amd.TUNE_FIELDS}
    root, cd = Ligesis.commit(pk, inp.f)
    opk = oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False)
    ocd = oracle.CommitData(opk, inp.f, keep_enc=True)
    assert root == ocd.root
    np.testing.assert_array_equal(cd.encoded(), ocd.enc())
    np.testing.assert_array_equal(cd.digests(), ocd.h())
    # the table round-trips through a file
    path = str(tmp_path / "tune.txt")
    assert ligesis_amd.tune_save(path) == 1
    ligesis_amd.tune_clear()
    assert ligesis_amd.tune_load(path) == 1
    assert gpu_ctx.tuning(log_m, log_n) == cfg


def test_explicit_option_overrides_table(gpu_ctx):
    inp = synth.Inputs(4, 8, seed=5)
    pk, _ = Ligesis.setup(4, 8, inp.a, ctx=gpu_ctx, commit_only=True)
    gpu_ctx.autotune(pk, reps=1)
    try:
        gpu_ctx.set_option("ntt_impl", 1)
        gpu_ctx.set_option("merkle_lg", 6)
        cfg = gpu_ctx.tuning(4, 8)
        assert cfg["ntt_impl"] == 1 and cfg["merkle_lg"] == 6
        root, _ = Ligesis.commit(pk, inp.f)
    finally:
        gpu_ctx.set_option("ntt_impl", -1)
        gpu_ctx.set_option("merkle_lg", 0)
    root2, _ = Ligesis.commit(pk, inp.f)
    assert root == root2


@pytest.mark.parametrize("lg", [6, 7, 8])
def test_merkle_widths_agree(gpu_ctx, lg):
    """Every Merkle workgroup width gives the oracle's root (2n = 2^11 leaves: 2 launches at 6)."""
    inp = synth.Inputs(3, 10, seed=9)
    pk, _ = Ligesis.setup(3, 10, inp.a, ctx=gpu_ctx, commit_only=True)
    gpu_ctx.set_option("merkle_lg", lg)
    try:
        root, _ = Ligesis.commit(pk, inp.f)
    finally:
        gpu_ctx.set_option("merkle_lg", 0)
    opk = oracle.ProvingKey(3, 10, inp.a, commit_a=False)
    assert root == oracle.CommitData(opk, inp.f).root


@pytest.mark.parametrize("shape,pipeline", [((3, 12), 2), ((4, 9), 4), ((5, 9), 8), ((2, 7), 2)])
def test_pipelined_commit_matches_oracle(gpu_ctx, shape, pipeline):
    """Row chunks encoded on one stream and hashed on another give the oracle's F', H and root;
    (2, 7) has too few K-splits for two chunks and takes the plain path."""
    log_m, log_n = shape
    print(f"pipelined commit shape={shape} chunks={pipeline}", flush=True)
    inp = synth.Inputs(log_m, log_n, seed=70 + log_m)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    gpu_ctx.set_option("pipeline", pipeline)
    gpu_ctx.set_option("profile", 1)
    try:
        root, cd = Ligesis.commit(pk, inp.f)
        st = gpu_ctx.stage_times()
    finally:
        gpu_ctx.set_option("pipeline", 0)
        gpu_ctx.set_option("profile", 0)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    assert root == ocd.root
    np.testing.assert_array_equal(cd.encoded(), ocd.enc())
    np.testing.assert_array_equal(cd.digests(), ocd.h())
    assert all(v >= 0 for v in st.values()) and st["commit_ms"] > 0


@pytest.mark.parametrize("variant", [1, 2])
def test_hash_kernel_timing_matches_oracle(gpu_ctx, variant):
    """Profile level 2 (the timed region of the commit benchmark) times only the hash kernel,
    from the MFMA kernel dispatch's own start and stop times (the VALU hash from events around
    it); the commit is unchanged and the time is positive, the other stages zero."""
    log_m, log_n = 4, 10
    print(f"profile 2 commit hash_variant={variant}", flush=True)
    inp = synth.Inputs(log_m, log_n, seed=81)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx, commit_only=True)
    gpu_ctx.set_option("hash", variant)
    gpu_ctx.set_option("profile", 2)
    try:
        times = []
        for _ in range(3):
            root, cd = Ligesis.commit(pk, inp.f)
            times.append(gpu_ctx.stage_times())
    finally:
        gpu_ctx.set_option("hash", 0)
        gpu_ctx.set_option("profile", 0)
    ocd = oracle.CommitData(oracle.ProvingKey(log_m, log_n, inp.a, commit_a=False), inp.f, keep_enc=True)
    assert root == ocd.root
    np.testing.assert_array_equal(cd.digests(), ocd.h())
    for st in times:
        assert st["hash_ms"] > 0 and st["ntt_ms"] == 0 and st["merkle_ms"] == 0
