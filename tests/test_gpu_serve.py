"""Serving loop: every request's root and proof are the same alone and in a mix of concurrent
requests of different sizes; the oracle verifier accepts them; admission respects the HBM
budget; stop cancels queued work cleanly."""
import numpy as np
import pytest

import oracle
from tests import proof_format
from ligesis_amd import Ligesis, LigesisError, RandomOracle, Server
from ligesis_amd import synth

pytestmark = pytest.mark.gpu

SHAPES = [(0, 7), (1, 9), (2, 10), (3, 10), (1, 12), (4, 12), (2, 14), (6, 14)]


def _requests(gpu_ctx, n_per_shape=2):
    keys, reqs = {}, []
    for i, (lm, ln) in enumerate(SHAPES):
        inp = synth.Inputs(lm, ln, seed=500 + i)
        keys[(lm, ln)] = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx)
        for k in range(n_per_shape):
            r = synth.Inputs(lm, ln, seed=700 + 10 * i + k)
            reqs.append(((lm, ln), r))
    return keys, reqs


def test_served_results_equal_alone(gpu_ctx):
    keys, reqs = _requests(gpu_ctx)
    alone = []
    for shape, r in reqs:
        pk, _ = keys[shape]
        root, cd = Ligesis.commit(pk, r.f)
        alone.append((root, Ligesis.prove(pk, cd, r.z, RandomOracle(r.fields, r.ints))))
        cd.release()
    srv = Server(0, hbm_budget=8 << 30, prove_workers=3, max_batch=8)
    print("serving", len(reqs), "requests", flush=True)
    tickets = [srv.submit(keys[shape][0], r.f, r.z, RandomOracle(r.fields, r.ints)) for shape, r in reqs]
    got = [srv.poll(t) for t in tickets]
    st = srv.stats()
    srv.stop()
    for (root, proof, times), (eroot, eproof) in zip(got, alone):
        assert root == eroot
        assert proof == eproof
        assert times["total_ms"] >= times["commit_ms"] >= 0
    assert st["requests"] == len(reqs)
    assert st["batches"] < len(reqs)  # requests were batched together
    # the oracle's verifier accepts a served proof
    (shape, r), (root, proof, _) = reqs[3], got[3]
    vk = keys[shape][1]
    assert oracle.verify(shape[0], shape[1], vk.com_a, root, r.z, proof, r.fields, r.ints)


def test_served_stream_at_full_sizes(gpu_ctx):
    # the served stream's mu range [14, 26] at its opening shapes log_m = floor((mu - 7) / 2),
    # two requests each of mu = 24 and 26 among smaller ones, all in flight together
    mus = [14, 24, 18, 26, 21, 24, 26, 16]
    keys, reqs = {}, []
    for i, mu in enumerate(mus):
        lm = (mu - 7) // 2
        ln = mu - lm
        if (lm, ln) not in keys:
            keys[(lm, ln)] = Ligesis.setup(lm, ln, synth.field_elems(3000 + mu, 32 * (64 << lm)), ctx=gpu_ctx)
        reqs.append(((lm, ln), synth.Inputs(lm, ln, seed=900 + i, with_a=False)))
    srv = Server(0, hbm_budget=64 << 30, prove_workers=2, max_batch=4)
    print("serving mu", mus, flush=True)
    tickets = [srv.submit(keys[shape][0], r.f, r.z, RandomOracle(r.fields, r.ints)) for shape, r in reqs]
    got = [srv.poll(t) for t in tickets]
    srv.stop()
    for k, ((shape, r), (root, proof, _)) in enumerate(zip(reqs, got)):
        print(f"lone commit + prove of request {k} shape={shape}", flush=True)
        pk, vk = keys[shape]
        eroot, cd = Ligesis.commit(pk, r.f)
        eproof = Ligesis.prove(pk, cd, r.z, RandomOracle(r.fields, r.ints))
        cd.release()
        served_ok = oracle.verify(shape[0], shape[1], vk.com_a, root, r.z, proof, r.fields, r.ints)
        lone_ok = oracle.verify(shape[0], shape[1], vk.com_a, eroot, r.z, eproof, r.fields, r.ints)
        if proof != eproof:
            a, b = proof_format.decode(proof), proof_format.decode(eproof)
            print(f"request {k}: served proof differs in {[f for f in a if a[f] != b[f]]}", flush=True)
            if shape[0] + shape[1] <= 21:  # the oracle's prove takes seconds at these sizes
                opk = oracle.ProvingKey(shape[0], shape[1], synth.field_elems(3000 + mus[k], 32 * (64 << shape[0])))
                oproof, _ = oracle.prove(opk, oracle.CommitData(opk, r.f), r.z, r.fields, r.ints)
                print(f"request {k}: served == oracle {proof == oproof}, lone == oracle {eproof == oproof}",
                      flush=True)
        assert (served_ok, lone_ok) == (True, True)
        assert root == eroot
        assert proof == eproof
    for pk, _ in keys.values():
        pk.release()


def test_server_budget_and_errors(gpu_ctx):
    inp = synth.Inputs(1, 9, seed=1)
    pk, _ = Ligesis.setup(1, 9, inp.a, ctx=gpu_ctx)
    tiny = Server(0, hbm_budget=1 << 20, prove_workers=1)
    with pytest.raises(LigesisError) as e:
        tiny.submit(pk, inp.f, inp.z, RandomOracle(inp.fields, inp.ints))
    assert e.value.code == 4
    tiny.stop()
    # a budget that fits one request at a time still completes everything, one batch each
    srv = Server(0, hbm_budget=140 << 20, prove_workers=2)
    ts = [srv.submit(pk, inp.f, inp.z, RandomOracle(inp.fields, inp.ints)) for _ in range(3)]
    res = [srv.poll(t) for t in ts]
    assert len({r[1] for r in res}) == 1
    # an exhausted oracle fails that request only
    bad = srv.submit(pk, inp.f, inp.z, RandomOracle(inp.fields[:3], inp.ints))
    good = srv.submit(pk, inp.f, inp.z, RandomOracle(inp.fields, inp.ints))
    with pytest.raises(LigesisError) as e:
        srv.poll(bad)
    assert e.value.code == 5
    assert srv.poll(good)[1] == res[0][1]
    srv.stop()
    _ = np


def test_resident_commitment_with_separate_opens(gpu_ctx):
    # commit-only requests keep their commitment resident; opens at different points, submitted
    # before and after the commit is done, equal lone proves; a released commitment refuses opens
    shapes = [(2, 12), (6, 14)]
    srv = Server(0, hbm_budget=8 << 30, prove_workers=2, max_batch=4)
    for k, (lm, ln) in enumerate(shapes):
        print(f"resident commitment shape=({lm},{ln})", flush=True)
        inp = synth.Inputs(lm, ln, seed=950 + k)
        pk, vk = Ligesis.setup(lm, ln, inp.a, ctx=gpu_ctx)
        c = srv.submit_commit(pk, inp.f)
        pts = [synth.ext_elems(960 + 10 * k + j, lm + ln) for j in range(3)]
        early = srv.submit_open(c, pts[0], RandomOracle(inp.fields, inp.ints))  # may wait for the commit
        root, proof, _ = srv.poll(c)
        assert proof == b""
        eroot, cd = Ligesis.commit(pk, inp.f)
        assert root == eroot
        late = [srv.submit_open(c, p, RandomOracle(inp.fields, inp.ints)) for p in pts[1:]]
        for t, p in zip([early] + late, pts):
            r, pf, times = srv.poll(t)
            assert r == eroot
            assert pf == Ligesis.prove(pk, cd, p, RandomOracle(inp.fields, inp.ints))
            assert oracle.verify(lm, ln, vk.com_a, r, p, pf, inp.fields, inp.ints)
        cd.release()
        srv.release(c)
        with pytest.raises(LigesisError):
            srv.submit_open(c, pts[0], RandomOracle(inp.fields, inp.ints))
        with pytest.raises(LigesisError):
            srv.release(c)
    st = srv.stats()
    assert st["reserved_bytes"] == 0
    srv.stop()
