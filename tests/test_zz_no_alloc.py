"""lgs_commit_noalloc (the commit into caller-owned device buffers, scratch in a static device
array) against lgs_commit (pcs/src/ligesis.rs:277-325 on the GPU) on the same inputs: the root,
F', H and every tree node are identical, on two calls in a row with the second smaller, and
shapes whose scratch exceeds the fixed size are rejected with LGS_ERR_ARG."""
import numpy as np
import pytest
import torch

from ligesis_amd import LGS_ERR_ARG, Ligesis, LigesisError
from ligesis_amd import synth

pytestmark = pytest.mark.gpu


def _outputs(lm, ln):
    m, n = 1 << lm, 1 << ln
    # filled with a pattern so that a word the entry point fails to write cannot match
    enc = torch.full((m * 2 * n,), -1, dtype=torch.int64, device="cuda:0")
    h = torch.full((32 * 2 * n,), -1, dtype=torch.int64, device="cuda:0")
    tree = torch.full(((4 * n - 1) * 32,), 0xA5, dtype=torch.uint8, device="cuda:0")
    return enc, h, tree


def _check(gpu_ctx, lm, ln, seed, stream=None):
    print(f"noalloc commit shape=({lm},{ln}) seed={seed} stream={'torch' if stream else 'ctx'}", flush=True)
    inp = synth.Inputs(lm, ln, seed=seed, n_fields=4, n_ints=4)
    pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
    root, cd = Ligesis.commit(pk, inp.f)
    f = torch.from_numpy(inp.f.view(np.int64)).to("cuda:0")
    enc, h, tree = _outputs(lm, ln)
    torch.cuda.synchronize()
    if stream is None:
        got = Ligesis.commit_noalloc(pk, f, enc, h, tree)
    else:
        assert Ligesis.commit_noalloc(pk, f, enc, h, tree, stream=stream, want_root=False) is None
        stream.synchronize()
        got = tree[-32:].cpu().numpy().tobytes()
    assert got == root
    np.testing.assert_array_equal(enc.cpu().numpy().view(np.uint64).reshape(1 << lm, 2 << ln), cd.encoded())
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64).reshape(32, 2 << ln), cd.digests())
    assert tree.cpu().numpy().tobytes() == cd.tree()
    cd.release()
    pk.release()


def test_noalloc_matches_commit_then_smaller(gpu_ctx):
    _check(gpu_ctx, 12, 12, 1)   # the mu = 24 square commit
    _check(gpu_ctx, 6, 9, 3)     # a smaller one right after: tables and partials are rebuilt


def test_noalloc_on_a_caller_stream(gpu_ctx):
    s = torch.cuda.Stream()
    _check(gpu_ctx, 8, 10, 4, stream=s)
    _check(gpu_ctx, 3, 7, 5, stream=s)


def test_noalloc_rejects_shapes_beyond_the_scratch(gpu_ctx):
    # rows of 2^17 values: one K-split of partials is 32 x 2^18 x 8 bytes = 64 MiB > 40 MiB
    lm, ln = 2, 17
    inp = synth.Inputs(lm, ln, seed=9, n_fields=4, n_ints=4)
    pk, _ = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx, commit_only=True)
    f = torch.from_numpy(inp.f.view(np.int64)).to("cuda:0")
    enc, h, tree = _outputs(lm, ln)
    with pytest.raises(LigesisError) as e:
        Ligesis.commit_noalloc(pk, f, enc, h, tree)
    assert e.value.code == LGS_ERR_ARG
    pk.release()


def test_noalloc_captured_in_a_graph(gpu_ctx):
    """lgs_commit_noalloc allocates nothing and synchronizes nothing without a root, so a caller can
    capture it into a HIP graph (here torch.cuda.CUDAGraph on a caller stream) and replay it with
    new polynomials written into the captured input buffer. Each of four replays (inputs f1, f2,
    f1, f2), with eager commits and synchronizations between them, must equal lgs_commit on the
    same input."""
    lm, ln = 8, 10
    m, n = 1 << lm, 1 << ln
    inp1 = synth.Inputs(lm, ln, seed=31, n_fields=4, n_ints=4)
    inp2 = synth.Inputs(lm, ln, seed=32, n_fields=4, n_ints=4)
    pk, _ = Ligesis.setup(lm, ln, inp1.a, ctx=gpu_ctx, commit_only=True)
    want = []
    for inp in (inp1, inp2):
        root, cd = Ligesis.commit(pk, inp.f)
        want.append((root, cd.encoded(), cd.digests(), cd.tree()))
        cd.release()
    f = torch.empty(m * n, dtype=torch.int64, device="cuda:0")
    f_src = [torch.from_numpy(inp.f.view(np.int64)).to("cuda:0") for inp in (inp1, inp2)]
    enc, h, tree = _outputs(lm, ln)
    s = torch.cuda.Stream()
    f.copy_(f_src[0])
    torch.cuda.synchronize()
    # one eager call on the stream first (it fills the library's per-kernel launch tables)
    Ligesis.commit_noalloc(pk, f, enc, h, tree, stream=s, want_root=False)
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        Ligesis.commit_noalloc(pk, f, enc, h, tree, stream=s, want_root=False)
    other = synth.Inputs(6, 9, seed=33, n_fields=4, n_ints=4)
    opk, _ = Ligesis.setup(6, 9, other.a, ctx=gpu_ctx, commit_only=True)
    for rep, which in enumerate((0, 1, 0, 1)):
        print(f"graph replay {rep} input {which}", flush=True)
        f.copy_(f_src[which])
        enc.fill_(-1)
        h.fill_(-1)
        tree.fill_(0xA5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        root, e, d, t = want[which]
        assert tree[-32:].cpu().numpy().tobytes() == root, rep
        np.testing.assert_array_equal(enc.cpu().numpy().view(np.uint64).reshape(m, 2 * n), e)
        np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64).reshape(32, 2 * n), d)
        assert tree.cpu().numpy().tobytes() == t
        # eager work between the replays: a commit of another shape through the same scratch
        e2, h2, t2 = _outputs(6, 9)
        r2 = Ligesis.commit_noalloc(opk, torch.from_numpy(other.f.view(np.int64)).to("cuda:0"), e2, h2, t2)
        r3, cd3 = Ligesis.commit(opk, other.f)
        assert r2 == r3
        cd3.release()
    del g
    opk.release()
    pk.release()
