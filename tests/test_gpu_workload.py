"""The batched and serving workloads at N=2^16, L=24, dnum=3 against the
oracle (needs a GPU): every output's SHA-256 digest must equal the digest of
the oracle's result, committed in tests/golden/workload_n65536_digests.json
by tests/golden/make_workload_digests.py.  Inputs are regenerated here from
the same seeds (tests/workload_n65536.py); none of the oracle's ops run on
the GPU box.

  * fides_eval_mult_batch over the 128 pairs at levels U{4..24}, with
    rescale, first run and replayed from the captured plan
  * fides_eval_rotate_batch over the 128 first ciphertexts by 2^U{0..15}
  * fides_eval_rotate_hoisted with the 27 indices at level 24
  * the request loop on a 6-job mix (HMult, rotations, logistic-regression
    scoring at level 24), each job's digest in the mix
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import workload_n65536 as W
from tests.gpu_util import log

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workload_n65536_digests.json")


@pytest.fixture(scope="module")
def env():
    import pyfideslib_amd as F

    with open(GOLDEN) as f:
        want = json.load(f)
    P = W.params()
    assert want["params"]["moduli"] == [str(q) for q in P.moduli]
    o = O.Oracle(P, seed=0)
    raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, 0,
                                 [2.0 ** P.scale_bits] * (P.L + 1))
    ctx = F.Context(raw)
    log("workload keys")
    ctx.load_eval_key(W.relin_key(o))
    for r in W.rotation_indices():
        ctx.load_rotation_key(r, W.rotation_key(o, r))
    yield P, o, ctx, want
    ctx.close()


def _digest(g):
    return W.digest(g.store())


def _load(ctx, P, c):
    return ctx.load(c[0], c[1], noise=2.0 ** P.scale_bits)


def test_batched_mult_128_pairs(env):
    P, o, ctx, want = env
    levels = W.pair_levels()
    assert levels == want["pair_levels"]
    A, B = [], []
    for i, lv in enumerate(levels):
        a, b = W.pair(o, i, lv)
        A.append(_load(ctx, P, a))
        B.append(_load(ctx, P, b))
    outs = [ctx.ciphertext() for _ in levels]
    for rep in range(2):  # the plan is built and run, then captured and replayed
        log(f"batched mult 128 pairs, run {rep}")
        ctx.mult_batch(A, B, outs, rescale=True)
        got = [_digest(x) for x in outs]
        bad = [i for i, (g, w) in enumerate(zip(got, want["mult_rescale"])) if g != w]
        assert not bad, f"pairs differing from the oracle: {bad[:10]} (levels {[levels[i] for i in bad[:10]]})"
        assert [x.level for x in outs] == [lv - 1 for lv in levels]


def test_batched_rotations_128(env):
    P, o, ctx, want = env
    levels, rots = W.pair_levels(), W.batch_rotations()
    assert rots == want["batch_rotations"]
    A = [_load(ctx, P, W.pair(o, i, lv)[0]) for i, lv in enumerate(levels)]
    outs = [ctx.ciphertext() for _ in levels]
    log("batched rotations 128")
    ctx.rotate_batch(A, rots, outs)
    bad = [i for i, (x, w) in enumerate(zip(outs, want["rotate"])) if _digest(x) != w]
    assert not bad, f"rotations differing from the oracle: {bad[:10]}"


def test_hoisted_27_indices(env):
    P, o, ctx, want = env
    assert W.HOIST_IDX == want["hoist_indices"]
    g = _load(ctx, P, W.hoist_input(o))
    outs = [ctx.ciphertext() for _ in W.HOIST_IDX]
    for rep in range(2):  # eager, then captured on the set's second use
        log(f"hoisted 27 indices, run {rep}")
        g.rotate_hoisted(W.HOIST_IDX, outs)
        bad = [r for r, x, w in zip(W.HOIST_IDX, outs, want["hoisted"]) if _digest(x) != w]
        assert not bad, f"hoisted rotations differing from the oracle: {bad}"


def test_serving_mix(env):
    import pyfideslib_amd as F

    P, o, ctx, want = env
    jobs = W.serve_jobs(o)
    assert [j[0] for j in jobs] == want["serve_kinds"]
    dev = []
    for j in jobs:
        if j[0] == "mult":
            dev.append((_load(ctx, P, j[1]), _load(ctx, P, j[2])))
        elif j[0] == "rot":
            dev.append((_load(ctx, P, j[1]),))
        else:
            dev.append((_load(ctx, P, j[1]), ctx.plaintext(j[2], noise=2.0 ** P.scale_bits)))
    outs = [ctx.ciphertext() for _ in jobs]
    srv = F.Server(ctx, max_slots=16)
    log("serving mix at N=2^16")
    ids = {}
    for k, (j, d) in enumerate(zip(jobs, dev)):
        if j[0] == "mult":
            jid = srv.submit_mult(d[0], d[1], outs[k])
        elif j[0] == "rot":
            jid = srv.submit_rotate(d[0], j[2], outs[k])
        else:
            jid = srv.submit_logreg(d[0], d[1], W.LOGREG_ROT, W.LOGREG_COEFFS, outs[k])
        ids[jid] = k
        if k % 2 == 1:
            srv.step()  # later jobs arrive while earlier ones are in flight
    assert sorted(srv.drain()) == sorted(ids)
    srv.close()
    bad = [(k, jobs[k][0]) for k in range(len(jobs)) if _digest(outs[k]) != want["serve"][k]]
    assert not bad, f"served jobs differing from the oracle: {bad}"
