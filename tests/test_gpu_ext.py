"""GPU parity of the P*Q (extended-basis) ciphertext entry points -- extend,
ModDown, hoisted rotations left in P*Q, adds of P*Q operands and plaintext
dot products -- against the oracle's restatement of the same forms
(oracle.ExtBasis), bit-exact, and the Q-basis entry points' refusal of a P*Q
operand."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu

ROTS = (1, 2, 5)


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=3, rotations=ROTS)
    yield P, o, ctx, keys, O.ExtBasis(o)
    ctx.close()


def enc(o, P, level, seed):
    rng = np.random.default_rng(seed)
    return o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), level, seed)


@pytest.mark.parametrize("level", [0, 2, 3])
def test_extend_and_mod_down(env, level):
    P, o, ctx, keys, E = env
    log(f"extend / mod_down at level {level}")
    ct = enc(o, P, level, 10 + level)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    x = g.extend()
    assert x.is_ext and not g.is_ext
    got = x.store_ext()
    assert np.array_equal(got[0], E.extend(ct[0])) and np.array_equal(got[1], E.extend(ct[1]))
    y = x.mod_down()
    assert not y.is_ext
    c0, c1 = y.store()
    assert np.array_equal(c0, ct[0]) and np.array_equal(c1, ct[1])
    x.mod_down(out=x)  # in place
    c0, c1 = x.store()
    assert np.array_equal(c0, ct[0]) and np.array_equal(c1, ct[1])
    g.extend(out=g)  # in place: a Q buffer grows to the P*Q layout
    assert np.array_equal(g.store_ext()[0], E.extend(ct[0]))


@pytest.mark.parametrize("level", [1, 3])
def test_rotate_hoisted_ext(env, level):
    P, o, ctx, keys, E = env
    log(f"hoisted rotations in P*Q at level {level}")
    ct = enc(o, P, level, 20 + level)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    idx = [0] + list(ROTS)
    outs = g.rotate_hoisted(idx, ext=True)
    for r, out in zip(idx, outs):
        assert out.is_ext and out.level == level
        got = out.store_ext()
        want = (E.extend(ct[0]), E.extend(ct[1])) if r == 0 else \
            E.rotate(ct[0], ct[1], O.rot_index(P.logN, r), keys[r])
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), f"index {r}"
        md = out.mod_down().store()
        wm = E.moddown(want)
        assert np.array_equal(md[0], wm[0]) and np.array_equal(md[1], wm[1])
    with pytest.raises(Exception):
        g.rotate_hoisted([1], outs=[g], ext=True)  # an output aliasing the input


def test_add_ext_and_refusals(env):
    import pyfideslib_amd as F

    P, o, ctx, keys, E = env
    level = 2
    a, b = enc(o, P, level, 31), enc(o, P, level, 32)
    ga, gb = ctx.load(*a, noise=2.0 ** P.scale_bits), ctx.load(*b, noise=2.0 ** P.scale_bits)
    xa = ga.rotate_hoisted([1], ext=True)[0]
    xb = gb.extend()
    wa = E.rotate(a[0], a[1], O.rot_index(P.logN, 1), keys[1])
    wb = (E.extend(b[0]), E.extend(b[1]))
    s = xa.add_ext(xb).store_ext()
    assert np.array_equal(s[0], E.add(wa[0], wb[0])) and np.array_equal(s[1], E.add(wa[1], wb[1]))
    s2 = xa.add_ext(gb).store_ext()  # a Q operand is extended first
    assert np.array_equal(s2[0], s[0]) and np.array_equal(s2[1], s[1])
    xa.add_ext(xb, out=xa)  # in place
    assert np.array_equal(xa.store_ext()[0], s[0])
    # Q-basis entry points refuse a P*Q operand, as input and as output
    for call in (lambda: xb.store(), lambda: xb.rotate(1), lambda: ga.add(gb, out=xb), lambda: xb.rescale()):
        with pytest.raises(F.FidesError) as e:
            call()
        assert e.value.code == 3
    with pytest.raises(F.FidesError):
        ga.add_ext(gb)  # neither operand in P*Q form


@pytest.mark.parametrize("ext", [False, True])
def test_dot_product_plain(env, ext):
    P, o, ctx, keys, E = env
    level, n = 3, 5
    log(f"dot product of {n} terms, P*Q {ext}")
    rng = np.random.default_rng(40 + ext)
    cts = [enc(o, P, level, 50 + i) for i in range(n)]
    gs = [ctx.load(*c, noise=2.0 ** P.scale_bits) for c in cts]
    if ext:
        pq = E.ids(P.L)
        pts_h = [o.random_poly(len(pq), 60 + i, primeids=pq) for i in range(n)]
        pts = [ctx.plaintext_ext(p) for p in pts_h]
        xs = [g.extend() for g in gs]
        want = None
        for c, p in zip(cts, pts_h):
            pl = np.ascontiguousarray(np.concatenate([p[: level + 1], p[P.L + 1:]]))
            t = (E.mul(E.extend(c[0]), pl), E.mul(E.extend(c[1]), pl))
            want = t if want is None else (E.add(want[0], t[0]), E.add(want[1], t[1]))
        out = ctx.dot_product_plain(xs, pts)
        got = out.store_ext()
        assert out.info()["noise_level"] == 2
        # in place over the first operand
        ctx.dot_product_plain(xs, pts, out=xs[0])
        got2 = xs[0].store_ext()
    else:
        pts_h = [o.random_poly(P.L + 1, 60 + i) for i in range(n)]
        pts = [ctx.plaintext(p) for p in pts_h]
        want = None
        for c, p in zip(cts, pts_h):
            t = o.mult_plain(c, p)
            want = t if want is None else (o.add(want[0], t[0]), o.add(want[1], t[1]))
        out = ctx.dot_product_plain(gs, pts)
        got = out.store()
        ctx.dot_product_plain(gs, pts, out=gs[0])
        got2 = gs[0].store()
    for g_ in (got, got2):
        assert np.array_equal(g_[0], want[0]) and np.array_equal(g_[1], want[1])
