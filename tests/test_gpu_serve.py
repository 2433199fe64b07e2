"""The request loop (fides_server_*): every job's result is the same alone
and in a mix of concurrently admitted jobs (bit-exact: all arithmetic is
exact modular arithmetic), matches the oracle's composition of the same ops,
and the logistic-regression job decrypts to the scored polynomial."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu

N_ROT = 8  # rotate-and-sum over 256 features
COEFFS = (0.5, 0.197, -0.004)  # sigma(y) ~ 0.5 + 0.197 y - 0.004 y^3


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=0, rotations=tuple(1 << i for i in range(N_ROT)) + (1 << 10,))
    yield P, o, ctx, keys
    ctx.close()


def _enc(o, P, z, level, seed):
    return o.encrypt_slots(z, level, seed)


def _jobs(P, o, ctx, seed=0):
    """A deterministic mix: HMults at levels 1..3, power-of-two rotations, scoring jobs."""
    rng = np.random.default_rng(seed)
    scale = 2.0 ** P.scale_bits
    jobs = []
    for i in range(12):
        kind = ("mult", "rot", "logreg")[i % 3]
        if kind == "mult":
            lv = int(rng.integers(1, P.L + 1))
            za, zb = rng.uniform(-1, 1, P.N // 2), rng.uniform(-1, 1, P.N // 2)
            a, b = _enc(o, P, za, lv, 10 + i), _enc(o, P, zb, lv, 50 + i)
            jobs.append(dict(kind=kind, a=a, b=b, za=za, zb=zb))
        elif kind == "rot":
            lv = int(rng.integers(0, P.L + 1))
            r = int(1 << rng.integers(0, N_ROT)) if i != 4 else (1 << 10)
            za = rng.uniform(-1, 1, P.N // 2)
            jobs.append(dict(kind=kind, a=_enc(o, P, za, lv, 90 + i), r=r, za=za))
        else:
            x = np.zeros(P.N // 2)
            w = np.zeros(P.N // 2)
            x[:256] = rng.uniform(-1, 1, 256)
            w[:256] = rng.normal(0, 0.1, 256)
            jobs.append(dict(kind=kind, x=_enc(o, P, x, P.L, 130 + i), w=o.encode(w, P.L), zx=x, zw=w))
    dev = []
    for j in jobs:
        if j["kind"] == "mult":
            d = dict(a=ctx.load(*j["a"], noise=scale), b=ctx.load(*j["b"], noise=scale))
        elif j["kind"] == "rot":
            d = dict(a=ctx.load(*j["a"], noise=scale))
        else:
            d = dict(x=ctx.load(*j["x"], noise=scale), w=ctx.plaintext(j["w"], noise=scale))
        d["out"] = ctx.ciphertext()
        dev.append(d)
    return jobs, dev


def _submit(srv, j, d):
    if j["kind"] == "mult":
        return srv.submit_mult(d["a"], d["b"], d["out"])
    if j["kind"] == "rot":
        return srv.submit_rotate(d["a"], j["r"], d["out"])
    return srv.submit_logreg(d["x"], d["w"], N_ROT, COEFFS, d["out"])


def _oracle_logreg(P, o, keys, j):
    sf = 2.0 ** P.scale_bits
    y = o.rescale(o.mult_plain(j["x"], j["w"]))
    for i in range(N_ROT):
        t = o.rotate(y, 1 << i, keys[1 << i])
        y = (o.add(y[0], t[0]), o.add(y[1], t[1]))
    ly = y[0].shape[0] - 1
    y2 = o.rescale(o.mult(y, y, keys["relin"]))
    u = o.rescale(o.mult_scalar(y, O.elem_for_mult(P.moduli, ly, COEFFS[2], sf)))
    y3 = o.rescale(o.mult(y2, u, keys["relin"]))
    v = o.rescale(o.mult_scalar(y, O.elem_for_mult(P.moduli, ly, COEFFS[1], sf)))
    n3 = y3[0].shape[0]
    out = (o.add(y3[0], np.ascontiguousarray(v[0][:n3])), o.add(y3[1], np.ascontiguousarray(v[1][:n3])))
    e = O.elem_for_add(P.moduli, n3 - 1, abs(COEFFS[0]), 1, sf)
    return o.add_scalar(out, e)


def test_serve_alone_equals_mix_and_oracle(env):
    import pyfideslib_amd as F

    P, o, ctx, keys = env
    log("serve: alone")
    jobs, dev = _jobs(P, o, ctx)
    alone = []
    srv = F.Server(ctx, max_slots=64)
    for j, d in zip(jobs, dev):
        jid = _submit(srv, j, d)
        assert srv.drain() == [jid]
        alone.append(d["out"].store())
        d["out"] = ctx.ciphertext()  # fresh output for the mixed run

    log("serve: mix")
    ids = {}
    # continuous batching: a third arrives up front, the rest while those are in flight
    for k, (j, d) in enumerate(zip(jobs, dev)):
        ids[_submit(srv, j, d)] = k
        if k % 4 == 3:
            srv.step()
    done = srv.drain()
    assert sorted(done) == sorted(ids)
    for k, (j, d) in enumerate(zip(jobs, dev)):
        c0, c1 = d["out"].store()
        assert np.array_equal(c0, alone[k][0]) and np.array_equal(c1, alone[k][1]), f"job {k} ({j['kind']})"

    log("serve: oracle")
    for j, d in zip(jobs, dev):
        if j["kind"] == "mult":
            ref = o.rescale(o.mult(j["a"], j["b"], keys["relin"]))
        elif j["kind"] == "rot":
            ref = o.rotate(j["a"], j["r"], keys[j["r"]])
        else:
            ref = _oracle_logreg(P, o, keys, j)
        c0, c1 = d["out"].store()
        assert np.array_equal(c0, ref[0]) and np.array_equal(c1, ref[1]), j["kind"]
        if j["kind"] == "logreg":
            dec = o.decrypt_slots(c0, c1, d["out"].info()["noise"])
            s = float(np.dot(j["zx"][:256], j["zw"][:256]))
            want = COEFFS[0] + COEFFS[1] * s + COEFFS[2] * s ** 3
            assert abs(dec[0] - want) < 1e-3, (dec[0], want)
    srv.close()


def test_serve_admission_by_bytes(env):
    """A budget that fits one job at a time still runs every job, with the same results."""
    import pyfideslib_amd as F

    P, o, ctx, keys = env
    log("serve: admission")
    jobs, dev = _jobs(P, o, ctx, seed=1)
    big = F.Server(ctx, max_slots=64)
    for j, d in zip(jobs, dev):
        _submit(big, j, d)
    big.drain()
    ref = [d["out"].store() for d in dev]
    for d in dev:
        d["out"] = ctx.ciphertext()
    ct_bytes = 2 * (P.L + 1) * P.N * 8
    one = F.Server(ctx, max_slots=64, budget_bytes=12 * ct_bytes)  # one scoring job at a time
    for j, d in zip(jobs, dev):
        _submit(one, j, d)
    peak = 0
    while one.step() > 0:
        peak = max(peak, one.counts()["used_bytes"])
    assert peak <= 12 * ct_bytes
    ctx.synchronize()
    assert len(one.poll()) == len(jobs)
    for (c0, c1), d in zip(ref, dev):
        g0, g1 = d["out"].store()
        assert np.array_equal(g0, c0) and np.array_equal(g1, c1)
    tiny = F.Server(ctx, max_slots=4, budget_bytes=1024)
    with pytest.raises(F.FidesError) as e:
        _submit(tiny, jobs[0], dev[0])
    assert e.value.code == 6
    for s in (big, one, tiny):
        s.close()
