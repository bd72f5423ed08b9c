"""The PolyCommitmentScheme adapter on the GPU: its commitments are LigeSIS commitments, its
batched proofs are the proofs sequential eval-proves give on one shared oracle (each accepted by
the oracle's verifier), the component openings of an extension-field polynomial combine to its
evaluation, f(z) = f0(z) + X f1(z), and batch_verify with the oracle's verifier accepts the
batch and rejects a wrong claim. Ligesis.prove_batch over keys of different shapes equals
sequential proves and the oracle's proofs."""
import numpy as np
import pytest

import oracle
from ligesis_amd import Ligesis, RandomOracle
from ligesis_amd import synth
from ligesis_amd.pcs import LigesisPcs

from . import proof_format as pfmt

pytestmark = pytest.mark.gpu

P = synth.P


def _ef_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def _times_x(a):  # (a0 + a1 X) X = 7 a1 + a0 X
    return ((7 * a[1]) % P, a[0])


@pytest.mark.parametrize("shape", [(3, 10), (1, 9)])
def test_batch_prove_matches_sequential_proofs(gpu_ctx, shape):
    log_m, log_n = shape
    mu = log_m + log_n
    print(f"pcs adapter shape={shape}", flush=True)
    inp = synth.Inputs(log_m, log_n, seed=70 + mu)
    pk, vk = Ligesis.setup(log_m, log_n, inp.a, ctx=gpu_ctx)
    pcs = LigesisPcs(pk, com_a=vk.com_a, verifier=oracle.verify)
    f_base = inp.f
    f_ext = np.stack([synth.field_elems(900 + mu, 1 << mu), synth.field_elems(901 + mu, 1 << mu)], axis=1)
    z1, z2 = synth.ext_elems(902 + mu, mu), synth.ext_elems(903 + mu, mu)
    fields, ints = synth.ext_elems(904 + mu, 4096), synth.oracle_ints(905 + mu, 8192)

    d1, c1 = pcs.commit(f_base)
    d2, c2 = pcs.commit(f_ext)
    assert len(c1.roots) == 1 and len(c2.roots) == 2
    for root, f in [(c1.roots[0], f_base), (c2.roots[0], f_ext[:, 0]), (c2.roots[1], f_ext[:, 1])]:
        assert root == Ligesis.commit(pk, f)[0]

    ro = RandomOracle(fields, ints)
    proofs = pcs.batch_prove([(d1, z1), (d2, z2)], ro)
    assert [len(p) for p in proofs] == [1, 2]

    # the same transcript from sequential proofs, each continuing where the last stopped
    fi = ii = 0
    flat = [(d1.parts[0], c1.roots[0], z1, proofs[0][0]), (d2.parts[0], c2.roots[0], z2, proofs[1][0]),
            (d2.parts[1], c2.roots[1], z2, proofs[1][1])]
    for cd, root, z, proof in flat:
        sub_f, sub_i = fields[fi:], ints[ii:]
        assert oracle.verify(log_m, log_n, vk.com_a, root, z, proof, sub_f, sub_i)
        ref = RandomOracle(sub_f, sub_i)
        assert Ligesis.prove(pk, cd, z, ref) == proof
        fi, ii = fi + ref.fields_idx, ii + ref.ints_idx
    assert ro.drawn() == (fi, ii)

    y = [pfmt.decode(p)["y"] for p in proofs[1]]
    want = tuple(int(v) for v in oracle.mle_eval(f_ext, z2))
    assert _ef_add(y[0], _times_x(y[1])) == want

    # batch_verify: the claims are the polynomials' values at their points
    lifted = np.stack([f_base, np.zeros_like(f_base)], axis=1)
    y_base = oracle.mle_eval(lifted, z1)
    claims = [(c1, z1, y_base), (c2, z2, np.array(want, np.uint64))]
    assert pcs.batch_verify(claims, proofs, RandomOracle(fields, ints))
    wrong = [(c1, z1, y_base), (c2, z2, np.array([want[0], (want[1] + 1) % P], np.uint64))]
    assert not pcs.batch_verify(wrong, proofs, RandomOracle(fields, ints))
    swapped = [claims[1], claims[0]]
    assert not pcs.batch_verify(swapped, [proofs[1], proofs[0]], RandomOracle(fields, ints))
    d1.release()
    d2.release()


def test_prove_batch_mixed_shapes_matches_sequential_and_oracle(gpu_ctx):
    shapes = [(3, 10), (0, 7), (3, 10), (2, 12), (1, 9)]
    keys, cds, zs, inps = [], [], [], []
    made = {}
    for i, (lm, ln) in enumerate(shapes):
        inp = synth.Inputs(lm, ln, seed=310 + i)
        if (lm, ln) not in made:
            made[(lm, ln)] = (Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx)[0], inp.a)
        pk, a = made[(lm, ln)]
        _, cd = Ligesis.commit(pk, inp.f)
        keys.append(pk)
        cds.append(cd)
        zs.append(inp.z)
        inps.append((inp, a))
    fields, ints = synth.ext_elems(330, 4096), synth.oracle_ints(331, 8192)
    ro = RandomOracle(fields, ints)
    print("prove_batch of", len(shapes), "proofs", flush=True)
    proofs = Ligesis.prove_batch(keys, cds, zs, ro)
    fi = ii = 0
    for (lm, ln), pk, cd, z, (inp, a), proof in zip(shapes, keys, cds, zs, inps, proofs):
        ref = RandomOracle(fields[fi:], ints[ii:])
        assert Ligesis.prove(pk, cd, z, ref) == proof
        opk = oracle.ProvingKey(lm, ln, a)
        ocd = oracle.CommitData(opk, inp.f)
        oproof, oused = oracle.prove(opk, ocd, z, fields[fi:], ints[ii:])
        assert oproof == proof
        assert oused == ref.drawn()
        fi, ii = fi + ref.fields_idx, ii + ref.ints_idx
    assert ro.drawn() == (fi, ii)
    for cd in cds:
        cd.release()
