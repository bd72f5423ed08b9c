"""GPU eval-prove parity: the proof bytes equal the oracle's (every sumcheck round polynomial,
Merkle opening and Basefold message), the oracle's verifier accepts the GPU proof, and the
reference's tamper cases (pcs/src/ligesis.rs:824-868) are rejected."""
import numpy as np
import pytest

import oracle
from ligesis_amd import Ligesis, LigesisError, RandomOracle
from ligesis_amd import synth

from . import proof_format as pfmt

pytestmark = pytest.mark.gpu

P = synth.P

# The reference test shape, the window edges (log_n = 7 + log_m and 17 + log_m), and the
# config-1 opening shape (mu = 20 at 2^6 x 2^14).
SHAPES = [(3, 10), (0, 7), (1, 8), (2, 12), (0, 17), (4, 11), (6, 14)]


def _instance(gpu_ctx, log_m, log_n, seed):
    print(f"prove parity shape=({log_m},{log_n}) seed={seed}", flush=True)
    inp = synth.Inputs(log_m, log_n, seed=seed)
    pk, vk = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx)
    root, cd = Ligesis.commit(pk, inp.f)
    ro = RandomOracle(inp.fields, inp.ints)
    proof = Ligesis.prove(pk, cd, inp.z, ro)
    return inp, pk, vk, root, cd, ro, proof


@pytest.mark.parametrize("shape", SHAPES)
def test_proof_matches_oracle_and_verifies(gpu_ctx, shape):
    log_m, log_n = shape
    inp, pk, vk, root, cd, ro, proof = _instance(gpu_ctx, log_m, log_n, seed=31 + 5 * log_m + log_n)
    opk = oracle.ProvingKey(log_m, log_n, inp.a)
    ocd = oracle.CommitData(opk, inp.f)
    assert vk.com_a == opk.com_a()
    assert root == ocd.root
    oproof, oused = oracle.prove(opk, ocd, inp.z, inp.fields, inp.ints)
    assert ro.drawn() == oused
    if proof != oproof:  # name the first differing field before failing
        g, o = pfmt.decode(proof), pfmt.decode(oproof)
        diff = [k for k in o if g.get(k) != o[k]]
        if "batch" in diff:
            diff += ["batch." + k for k in o["batch"] if g["batch"][k] != o["batch"][k]]
        pytest.fail(f"proof differs from the oracle in {diff}")
    assert oracle.verify(log_m, log_n, vk.com_a, root, inp.z, proof, inp.fields, inp.ints)


def test_claimed_value_is_f_at_z(gpu_ctx):
    inp, pk, vk, root, cd, ro, proof = _instance(gpu_ctx, 2, 9, seed=5)
    f_ef = np.stack([inp.f, np.zeros_like(inp.f)], axis=1)
    assert pfmt.decode(proof)["y"] == tuple(int(x) for x in oracle.mle_eval(f_ef, inp.z))


def test_tampered_proofs_rejected(gpu_ctx):
    log_m, log_n = 3, 10
    inp, pk, vk, root, cd, ro, proof = _instance(gpu_ctx, log_m, log_n, seed=8)

    def verify(b):
        return oracle.verify(log_m, log_n, vk.com_a, root, inp.z, b, inp.fields, inp.ints)

    assert verify(proof)
    d = pfmt.decode(proof)
    d["y"] = ((d["y"][0] + 1) % P, d["y"][1])
    assert not verify(pfmt.encode(d))
    d = pfmt.decode(proof)
    e = d["sc_hash"]["round_polys"][0][0]
    d["sc_hash"]["round_polys"][0][0] = ((e[0] + 1) % P, e[1])
    assert not verify(pfmt.encode(d))
    d = pfmt.decode(proof)
    d["h_cols"][0][0] = (d["h_cols"][0][0] + 1) % P
    assert not verify(pfmt.encode(d))


def test_prove_is_repeatable_and_commitment_reusable(gpu_ctx):
    inp, pk, vk, root, cd, ro, proof = _instance(gpu_ctx, 1, 9, seed=12)
    ro.restart()
    assert Ligesis.prove(pk, cd, inp.z, ro) == proof
    # a second opening point on the same resident commitment
    z2 = synth.ext_elems(999, 10)
    ro.restart()
    p2 = Ligesis.prove(pk, cd, z2, ro)
    assert oracle.verify(1, 9, vk.com_a, root, z2, p2, inp.fields, inp.ints)


@pytest.mark.parametrize("shape", [(3, 10), (4, 13)])
def test_repeated_proofs_replay_a_graph(gpu_ctx, shape):
    """The second proof of a commitment captures its launches into a graph and later ones
    replay it with new constants: each of five proofs (own point and oracle streams) equals the
    oracle's, with other commits and proofs on the context between the replays, and again with
    the graph switched off."""
    log_m, log_n = shape
    inp = synth.Inputs(log_m, log_n, seed=91 + log_n)
    pk, _ = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx)
    _, cd = Ligesis.commit(pk, inp.f)
    opk = oracle.ProvingKey(log_m, log_n, inp.a)
    ocd = oracle.CommitData(opk, inp.f)
    other = synth.Inputs(2, 9, seed=17)
    opk2, _ = Ligesis.setup(2, 9, other.a, ctx=gpu_ctx)
    try:
        for k, graph in enumerate([1, 1, 1, 1, 1, 0]):
            gpu_ctx.set_option("prove_graph", graph)
            q = synth.Inputs(log_m, log_n, seed=500 + k, with_a=False)
            print(f"prove {k} shape={shape} graph={graph}", flush=True)
            proof = Ligesis.prove(pk, cd, q.z, RandomOracle(q.fields, q.ints))
            want, _ = oracle.prove(opk, ocd, q.z, q.fields, q.ints)
            assert proof == want, k
            # other work on the context between the proofs
            _, cd2 = Ligesis.commit(opk2, other.f)
            Ligesis.prove(opk2, cd2, other.z, RandomOracle(other.fields, other.ints))
    finally:
        gpu_ctx.set_option("prove_graph", 0)


def test_prove_errors(gpu_ctx):
    inp = synth.Inputs(1, 8, seed=3)
    pk, _ = Ligesis.setup(1, 8, inp.a, ctx=gpu_ctx)
    _, cd = Ligesis.commit(pk, inp.f)
    with pytest.raises(LigesisError) as e:
        Ligesis.prove(pk, cd, inp.z, RandomOracle(inp.fields[:10], inp.ints))
    assert e.value.code == 5
    with pytest.raises(LigesisError):
        Ligesis.prove(pk, cd, inp.z[:-1], RandomOracle(inp.fields, inp.ints))
    # all 512 sampled integers land on one column: the reference's assert fires
    same = np.zeros_like(inp.ints)
    with pytest.raises(LigesisError):
        Ligesis.prove(pk, cd, inp.z, RandomOracle(inp.fields, same))
    # commit-only keys cannot prove; shapes outside the window are refused at setup
    pk2, _ = Ligesis.setup(1, 8, inp.a, ctx=gpu_ctx, commit_only=True)
    _, cd2 = Ligesis.commit(pk2, inp.f)
    with pytest.raises(LigesisError) as e:
        Ligesis.prove(pk2, cd2, inp.z, RandomOracle(inp.fields, inp.ints))
    assert e.value.code == 7
    with pytest.raises(LigesisError) as e:
        Ligesis.setup(4, 4, synth.field_elems(1, 32 * 64 * 16), ctx=gpu_ctx)
    assert e.value.code == 2


@pytest.mark.parametrize("shape", [(8, 16), (7, 17)])
def test_full_size_proof_verifies(gpu_ctx, shape):
    """mu = 24 at the benchmark's opening shape and the reference README's: the oracle's prover is
    too slow here, so the size-independent properties are checked -- the oracle's verifier
    accepts the GPU proof, its claimed value is f(z), a second proof from a restarted oracle has
    the same bytes, and a proof with a changed opened H word or sumcheck round is rejected."""
    log_m, log_n = shape
    inp, pk, vk, root, cd, ro, proof = _instance(gpu_ctx, log_m, log_n, seed=24 + log_m)

    def verify(b):
        return oracle.verify(log_m, log_n, vk.com_a, root, inp.z, b, inp.fields, inp.ints)

    print(f"full-size verify shape={shape} bytes={len(proof)}", flush=True)
    assert verify(proof)
    f_ef = np.stack([inp.f, np.zeros_like(inp.f)], axis=1)
    assert pfmt.decode(proof)["y"] == tuple(int(x) for x in oracle.mle_eval(f_ef, inp.z))
    ro.restart()
    assert Ligesis.prove(pk, cd, inp.z, ro) == proof
    d = pfmt.decode(proof)
    d["h_cols"][-1][-1] = (d["h_cols"][-1][-1] + 1) % P
    assert not verify(pfmt.encode(d))
    d = pfmt.decode(proof)
    e = d["sc_binary"]["round_polys"][-1][0]
    d["sc_binary"]["round_polys"][-1][0] = ((e[0] + 1) % P, e[1])
    assert not verify(pfmt.encode(d))
