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
         (9, 3, 3, False), (10, 6, 4, False), (16, 2, 4, False), (11, 4, 4, True)]


@pytest.mark.parametrize("nv,li,code_rate,ext", CASES)
def test_root_matches_oracle(gpu_ctx, nv, li, code_rate, ext):
    print(f"basefold commit nv={nv} li={li} code_rate={code_rate} ext={ext}", flush=True)
    poly = synth.ext_elems(nv + 40, 1 << nv) if ext else synth.field_elems(nv + 40, 1 << nv)
    want = oracle.basefold_commit(poly, li, code_rate)
    assert Basefold.commit_on_domain(poly, code_rate, li, ctx=gpu_ctx) == want
    dev = torch.from_numpy(np.ascontiguousarray(poly).view(np.int64)).to("cuda:0")
    assert Basefold.commit_on_domain(dev, code_rate, li, ctx=gpu_ctx, device=True) == want


def test_commit_base_uses_log_interleave_6(gpu_ctx):
    poly = synth.field_elems(77, 1 << 12)
    assert Basefold.commit_base(poly, 1, ctx=gpu_ctx) == oracle.basefold_commit(poly, 6, 1)


@pytest.mark.parametrize("nv,li,code_rate", [(5, 6, 1), (8, 7, 1), (8, 2, 0), (8, 2, 5), (31, 2, 1)])
def test_bad_arguments_rejected(gpu_ctx, nv, li, code_rate):
    poly = np.zeros(1 << min(nv, 8), np.uint64)
    with pytest.raises(LigesisError):
        if nv > 8:  # the size limit is checked before the polynomial is read
            from ligesis_amd import lib, _ptr
            gpu_ctx.check(lib().lgs_basefold_commit(gpu_ctx.h, _ptr(poly), nv, 0, li, code_rate,
                                                    _ptr(np.zeros(32, np.uint8))))
        else:
            Basefold.commit_on_domain(poly, code_rate, li, ctx=gpu_ctx)


def test_commit_ext_upcasts_a_base_polynomial(gpu_ctx):
    # commit_ext takes MlPoly<EF>: a base polynomial enters as (c0, 0) pairs (16-byte leaves)
    poly = synth.field_elems(78, 1 << 10)
    pairs = np.stack([poly, np.zeros_like(poly)], axis=1)
    want = oracle.basefold_commit(pairs, 6, 1)
    assert want != oracle.basefold_commit(poly, 6, 1)
    assert Basefold.commit_ext(poly, 1, ctx=gpu_ctx) == want
    assert Basefold.commit_ext(pairs, 1, ctx=gpu_ctx) == want


def test_device_input_checks(gpu_ctx):
    poly = torch.from_numpy(synth.field_elems(79, 1 << 9).view(np.int64)).to("cuda:0")
    pairs = torch.from_numpy(synth.ext_elems(80, 1 << 9).view(np.int64)).to("cuda:0")
    with pytest.raises(LigesisError):
        Basefold.commit_on_domain(pairs[:, 0], 1, 3, ctx=gpu_ctx, device=True)  # strided column
    with pytest.raises(LigesisError):
        Basefold.commit_on_domain(poly.to(torch.int32), 1, 3, ctx=gpu_ctx, device=True)
    with pytest.raises(LigesisError):
        Basefold.commit_base_on_domain(pairs, 1, 3, ctx=gpu_ctx, device=True)
    want = oracle.basefold_commit(pairs.cpu().numpy().view(np.uint64), 3, 1)
    assert Basefold.commit_ext_on_domain(pairs, 1, 3, ctx=gpu_ctx, device=True) == want
