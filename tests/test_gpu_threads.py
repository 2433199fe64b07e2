"""Thread safety at the handle level: C ABI calls on one context from several
threads are serialized by the context's lock, calls on different contexts run
concurrently; every result equals the single-threaded one bit for bit."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


def _pairs(o, P, seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        z1, z2 = rng.uniform(-1, 1, P.N // 2), rng.uniform(-1, 1, P.N // 2)
        out.append((o.encrypt_slots(z1, P.L, seed * 100 + 2 * i), o.encrypt_slots(z2, P.L, seed * 100 + 2 * i + 1)))
    return out


def _work(ctx, P, a, b):
    s = 2.0 ** P.scale_bits
    x = ctx.load(*a, noise=s).mult(ctx.load(*b, noise=s), rescale=True)
    return x.rotate(1).store()


def _run_threads(fns, timeout=120):
    errs = []

    def wrap(f):
        try:
            f()
        except Exception as e:  # reported below
            errs.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout)
        assert not t.is_alive(), "thread did not finish"
    assert not errs, errs


def test_threads_share_one_context():
    P = O.params_small()
    o, ctx, _ = make_env(P, seed=3, rotations=(1,))
    try:
        pairs = _pairs(o, P, 3, 8)
        ref = [_work(ctx, P, a, b) for a, b in pairs]
        got = [None] * len(pairs)

        def job(i):
            def f():
                for _ in range(3):
                    got[i] = _work(ctx, P, *pairs[i])
            return f

        log("8 threads on one context")
        _run_threads([job(i) for i in range(len(pairs))])
        for r, g in zip(ref, got):
            assert np.array_equal(r[0], g[0]) and np.array_equal(r[1], g[1])
    finally:
        ctx.close()


def test_threads_on_two_contexts():
    P = O.params_small()
    envs = [make_env(P, seed=s, rotations=(1,)) for s in (4, 5)]
    try:
        pairs = [_pairs(o, P, s, 4) for (o, _, _), s in zip(envs, (4, 5))]
        ref = [[_work(ctx, P, a, b) for a, b in pr] for (_, ctx, _), pr in zip(envs, pairs)]
        got = [[None] * 4 for _ in envs]

        def job(k):
            def f():
                ctx = envs[k][1]
                for _ in range(3):
                    for i, (a, b) in enumerate(pairs[k]):
                        got[k][i] = _work(ctx, P, a, b)
            return f

        log("2 threads on 2 contexts")
        _run_threads([job(0), job(1)])
        for rk, gk in zip(ref, got):
            for r, g in zip(rk, gk):
                assert np.array_equal(r[0], g[0]) and np.array_equal(r[1], g[1])
    finally:
        for _, ctx, _ in envs:
            ctx.close()
