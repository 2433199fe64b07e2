"""GPU parity of the CKKS-layer ops the reference's linear transforms and
EvalMod build on: subPt, addMultPt, addMultScalar, evalLinearWSumMutable,
EvalMult without its ModDown (the product left in P*Q) and the ciphertext dot
product with one key switch -- bit-exact against the oracle's composition of
the same steps."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=21)
    yield P, o, ctx, keys
    ctx.close()


def enc(o, P, level, seed):
    return o.encrypt_slots(np.random.default_rng(seed).uniform(-1, 1, P.N // 2), level, seed)


def neg(P, x):
    q = np.array(P.moduli[: x.shape[0]], dtype=np.uint64)[:, None]
    return np.where(x == 0, x, q - x)


def test_sub_plain(env):
    P, o, ctx, keys = env
    log("subPt")
    ct = enc(o, P, 2, 1)
    pt = o.encode(np.linspace(-1, 1, P.N // 2), P.L)
    got = ctx.load(*ct, noise=2.0 ** P.scale_bits).sub_plain(ctx.plaintext(pt, noise=2.0 ** P.scale_bits)).store()
    assert np.array_equal(got[0], o.add(ct[0], neg(P, pt[:3]))) and np.array_equal(got[1], ct[1])


@pytest.mark.parametrize("rescale", [False, True])
def test_add_mult_plain(env, rescale):
    P, o, ctx, keys = env
    log(f"addMultPt rescale {rescale}")
    scale = 2.0 ** P.scale_bits
    a, b = enc(o, P, 3, 2), enc(o, P, 3, 3)
    p1, p2 = o.random_poly(P.L + 1, 4), o.random_poly(P.L + 1, 5)
    acc = ctx.load(*a, noise=scale).mult_plain(ctx.plaintext(p1, noise=scale))
    acc.add_mult_plain_(ctx.load(*b, noise=scale), ctx.plaintext(p2, noise=scale), rescale=rescale)
    t1, t2 = o.mult_plain(a, p1), o.mult_plain(b, p2)
    want = (o.add(t1[0], t2[0]), o.add(t1[1], t2[1]))
    if rescale:
        want = o.rescale(want)
    got = acc.store()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert acc.info()["noise_level"] == (1 if rescale else 2)


def test_add_mult_scalar_and_linear_wsum(env):
    P, o, ctx, keys = env
    log("addMultScalar, evalLinearWSumMutable")
    scale = 2.0 ** P.scale_bits
    lv = 2
    xs = [enc(o, P, lv, 10 + i) for i in range(4)]
    ws = [0.75, -1.5, 3e-3, 2.0]
    gs = [ctx.load(*x, noise=scale) for x in xs]

    def ms(x, v, level):
        return o.mult_scalar(x, O.elem_for_mult(P.moduli, level, v, scale))

    want = None
    for x, v in zip(xs, ws):
        t = ms(x, v, lv)
        want = t if want is None else (o.add(want[0], t[0]), o.add(want[1], t[1]))
    out = ctx.linear_wsum(gs, ws)
    got = out.store()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    info = out.info()
    assert info["noise_level"] == 2 and info["noise"] == scale * scale and info["level"] == lv
    ctx.linear_wsum(gs, ws, out=gs[0])  # in place over the first term
    assert np.array_equal(gs[0].store()[0], want[0])
    # acc += b * d, b one level above acc
    b = enc(o, P, lv + 1, 30)
    acc = ctx.linear_wsum([ctx.load(*x, noise=scale) for x in xs], ws)
    acc.add_mult_scalar_(ctx.load(*b, noise=scale), -0.625)
    t = ms((np.ascontiguousarray(b[0][: lv + 1]), np.ascontiguousarray(b[1][: lv + 1])), -0.625, lv)
    got = acc.store()
    assert np.array_equal(got[0], o.add(want[0], t[0])) and np.array_equal(got[1], o.add(want[1], t[1]))


@pytest.mark.parametrize("level", [1, 3])
def test_mult_without_moddown(env, level):
    P, o, ctx, keys = env
    log(f"mult without ModDown at level {level}")
    E = O.ExtBasis(o)
    scale = 2.0 ** P.scale_bits
    a, b = enc(o, P, level, 40), enc(o, P, level, 41)
    x = ctx.load(*a, noise=scale).mult_ext(ctx.load(*b, noise=scale))
    assert x.is_ext and x.info()["noise_level"] == 2
    d0, d1, d2 = o.tensor(a[0], a[1], b[0], b[1])
    acc_a, acc_b = o.ksk_inner(o.modup(d2), keys["relin"])
    got = x.store_ext()
    assert np.array_equal(got[0], E.add(E.extend(d0), acc_b)) and np.array_equal(got[1], E.add(E.extend(d1), acc_a))
    md = x.mod_down().store()
    want = o.mult(a, b, keys["relin"])
    assert np.array_equal(md[0], want[0]) and np.array_equal(md[1], want[1])


@pytest.mark.parametrize("n, level", [(1, 3), (3, 2), (6, 3)])
def test_dot_product(env, n, level):
    """Ciphertext::dotProduct: the tensors of n pairs summed, ONE key switch of
    the summed d2 -- bit-exact against the oracle's tensor, add and key switch
    (Q form) or its ModUp and KSK with the extended (d0, d1) (P*Q form); n = 1
    equals EvalMult without rescale; the slots decrypt to sum a_i b_i; in place
    over an operand."""
    P, o, ctx, keys = env
    log(f"dot product of {n} pairs at level {level}")
    E = O.ExtBasis(o)
    scale = 2.0 ** P.scale_bits
    rng = np.random.default_rng(100 + n)
    za = [rng.uniform(-1, 1, P.N // 2) for _ in range(n)]
    zb = [rng.uniform(-1, 1, P.N // 2) for _ in range(n)]
    a = [o.encrypt_slots(z, level, 200 + i) for i, z in enumerate(za)]
    b = [o.encrypt_slots(z, level, 300 + i) for i, z in enumerate(zb)]
    ga = [ctx.load(*x, noise=scale) for x in a]
    gb = [ctx.load(*x, noise=scale) for x in b]
    d = None
    for x, y in zip(a, b):
        t = o.tensor(x[0], x[1], y[0], y[1])
        d = t if d is None else [o.add(u, v) for u, v in zip(d, t)]
    kb, ka = o.keyswitch(d[2], keys["relin"])
    want = (o.add(d[0], kb), o.add(d[1], ka))
    out = ctx.dot_product(ga, gb)
    info = out.info()
    assert info["noise_level"] == 2 and info["level"] == level and info["noise"] == scale * scale
    got = out.store()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    dec = o.decrypt_slots(*got, scale * scale)
    assert np.abs(dec - sum(x * y for x, y in zip(za, zb))).max() < 1e-4
    if n == 1:
        m = o.mult(a[0], b[0], keys["relin"])
        assert np.array_equal(got[0], m[0]) and np.array_equal(got[1], m[1])
    # P*Q form
    acc_a, acc_b = o.ksk_inner(o.modup(d[2]), keys["relin"])
    x = ctx.dot_product(ga, gb, ext=True)
    assert x.is_ext and x.info()["noise_level"] == 2
    ge = x.store_ext()
    assert np.array_equal(ge[0], E.add(E.extend(d[0]), acc_b)) and np.array_equal(ge[1], E.add(E.extend(d[1]), acc_a))
    # in place over the first operand
    ctx.dot_product(ga, gb, out=ga[0])
    g0 = ga[0].store()
    assert np.array_equal(g0[0], want[0]) and np.array_equal(g0[1], want[1])
