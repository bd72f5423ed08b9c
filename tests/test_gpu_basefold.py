"""Standalone Basefold commitment on the GPU (pcs/src/basefold.rs:309-382): roots equal the
oracle's for base and extension polynomials, interleave depths 0..6 (6 = the reference's
LOG_INTERLEAVE, basefold.rs:16), code rates 1..4 on both transform paths (rows held in LDS, and
rows longer than one LDS tile), host and device inputs; bad arguments are rejected."""
import numpy as np
import pytest
import torch

import oracle
from ligesis_amd import Basefold, LigesisError
from ligesis_amd import synth

pytestmark = pytest.mark.gpu

CASES = [(6, 0, 1, False), (10, 6, 1, False), (12, 4, 2, False), (16, 6, 1, False), (20, 6, 1, False),
         (9, 3, 1, True), (14, 6, 1, True), (13, 2, 3, False), (18, 5, 2, True),
         (9, 3, 3, False), (10, 6, 4, False), (16, 2, 4, False), (11, 4, 4, True),
         # rows of 2^16 - 2^19 points over enough rows for the register column pass (radix-8 over
         # 3 or 6 stages, radix-8 then radix-16 over 7); at 2^16 with rate 1 the compile-time-shaped
         # 2^13-point subsequence transforms before it
         (22, 6, 1, False), (20, 3, 1, True), (21, 3, 1, False), (21, 6, 1, False), (20, 5, 1, True),
         (19, 3, 2, False)]


@pytest.mark.parametrize("nv,li,code_rate,ext", CASES)
def test_root_matches_oracle(gpu_ctx, nv, li, code_rate, ext):
    print(f"basefold commit nv={nv} li={li} code_rate={code_rate} ext={ext}", flush=True)
    poly = synth.ext_elems(nv + 40, 1 << nv) if ext else synth.field_elems(nv + 40, 1 << nv)
    want = oracle.basefold_commit(poly, li, code_rate)
    assert Basefold.commit_root(poly, code_rate, li, ctx=gpu_ctx) == want
    dev = torch.from_numpy(np.ascontiguousarray(poly).view(np.int64)).to("cuda:0")
    assert Basefold.commit_root(dev, code_rate, li, ctx=gpu_ctx, device=True) == want


def test_commit_base_uses_log_interleave_6(gpu_ctx):
    poly = synth.field_elems(77, 1 << 12)
    st, root = Basefold.commit_base(poly, 1, ctx=gpu_ctx)
    assert root == oracle.basefold_commit(poly, 6, 1)
    st.release()


@pytest.mark.parametrize("nv,li,code_rate", [(5, 6, 1), (8, 7, 1), (8, 2, 0), (8, 2, 5), (31, 2, 1)])
def test_bad_arguments_rejected(gpu_ctx, nv, li, code_rate):
    poly = np.zeros(1 << min(nv, 8), np.uint64)
    with pytest.raises(LigesisError):
        if nv > 8:  # the size limit is checked before the polynomial is read
            from ligesis_amd import lib, _ptr
            gpu_ctx.check(lib().lgs_basefold_commit(gpu_ctx.h, _ptr(poly), nv, 0, li, code_rate,
                                                    _ptr(np.zeros(32, np.uint8))))
        else:
            Basefold.commit_root(poly, code_rate, li, ctx=gpu_ctx)


def test_commit_ext_upcasts_a_base_polynomial(gpu_ctx):
    # commit_ext takes MlPoly<EF>: a base polynomial enters as (c0, 0) pairs (16-byte leaves)
    poly = synth.field_elems(78, 1 << 10)
    pairs = np.stack([poly, np.zeros_like(poly)], axis=1)
    want = oracle.basefold_commit(pairs, 6, 1)
    assert want != oracle.basefold_commit(poly, 6, 1)
    assert Basefold.commit_ext(poly, 1, ctx=gpu_ctx)[1] == want
    assert Basefold.commit_ext(pairs, 1, ctx=gpu_ctx)[1] == want


def test_device_input_checks(gpu_ctx):
    poly = torch.from_numpy(synth.field_elems(79, 1 << 9).view(np.int64)).to("cuda:0")
    pairs = torch.from_numpy(synth.ext_elems(80, 1 << 9).view(np.int64)).to("cuda:0")
    with pytest.raises(LigesisError):
        Basefold.commit_root(pairs[:, 0], 1, 3, ctx=gpu_ctx, device=True)  # strided column
    with pytest.raises(LigesisError):
        Basefold.commit_root(poly.to(torch.int32), 1, 3, ctx=gpu_ctx, device=True)
    with pytest.raises(LigesisError):
        Basefold.commit_base_on_domain(pairs, 1, 3, ctx=gpu_ctx, device=True)
    want = oracle.basefold_commit(pairs.cpu().numpy().view(np.uint64), 3, 1)
    assert Basefold.commit_ext_on_domain(pairs, 1, 3, ctx=gpu_ctx, device=True)[1] == want


# ---- Basefold::prove (basefold.rs:384-516) on the GPU: bytes equal to the oracle's ----
PROVE_CASES = [(12, 6, 2, False), (12, 6, 2, True), (8, 0, 1, False), (9, 3, 3, True), (10, 2, 4, False),
               (7, 6, 1, True), (6, 6, 1, False), (14, 6, 1, False), (16, 4, 1, True), (20, 6, 1, False),
               (22, 6, 1, False)]


@pytest.mark.parametrize("nv,li,cr,ext", PROVE_CASES)
def test_basefold_proof_matches_oracle(gpu_ctx, nv, li, cr, ext):
    from ligesis_amd import RandomOracle

    print(f"basefold prove nv={nv} li={li} code_rate={cr} ext={ext}", flush=True)
    poly = synth.ext_elems(nv + 140, 1 << nv) if ext else synth.field_elems(nv + 140, 1 << nv)
    point = synth.ext_elems(nv + 141, nv)
    fields = synth.ext_elems(nv + 142, 64)
    ints = synth.oracle_ints(nv + 143, 128)
    st, root = Basefold.commit_ext_on_domain(poly, cr, li, ctx=gpu_ctx) if ext else \
        Basefold.commit_base_on_domain(poly, cr, li, ctx=gpu_ctx)
    ro = RandomOracle(fields, ints)
    proof = Basefold.prove(st, point, ro)
    oroot, oproof, oused = oracle.basefold_prove(poly, li, cr, point, fields, ints)
    assert root == oroot
    assert len(proof) == len(oproof)
    assert proof == oproof
    assert ro.drawn() == oused
    if nv > li:  # with no reduction round the reference's verifier has no fold chain to check
        assert oracle.basefold_verify(root, point, proof, ext, fields, ints)
    # a second proof at another point on the same resident state
    p2 = synth.ext_elems(nv + 144, nv)
    p2[0] = [0, 0]  # a zero coordinate is allowed (both reduction values are sent)
    proof2 = Basefold.prove(st, p2, RandomOracle(fields, ints))
    assert proof2 == oracle.basefold_prove(poly, li, cr, p2, fields, ints)[1]
    st.release()


def test_basefold_prove_device_input_and_oracle_errors(gpu_ctx):
    from ligesis_amd import RandomOracle

    nv = 12
    poly = synth.field_elems(300, 1 << nv)
    dev = torch.from_numpy(poly.view(np.int64)).to("cuda:0")
    st, root = Basefold.commit_base(dev, 1, ctx=gpu_ctx, device=True)
    point = synth.ext_elems(301, nv)
    fields, ints = synth.ext_elems(302, 16), synth.oracle_ints(303, 128)
    proof = Basefold.prove(st, point, RandomOracle(fields, ints))
    assert proof == oracle.basefold_prove(poly, 6, 1, point, fields, ints)[1]
    with pytest.raises(LigesisError) as e:  # 6 rounds need 6 fields
        Basefold.prove(st, point, RandomOracle(fields[:5], ints))
    assert e.value.code == 5
    with pytest.raises(LigesisError) as e:  # 100 queries at code rate 1
        Basefold.prove(st, point, RandomOracle(fields, ints[:99]))
    assert e.value.code == 5
    with pytest.raises(LigesisError):
        Basefold.prove(st, point[:-1], RandomOracle(fields, ints))
    st.release()
