"""GPU parity of the bootstrapping building blocks against the oracle,
bit-exact: ModRaise, conjugation (automorphism 2N-1), multiplication by a
monomial and by an integer."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=0)
    k = 2 * P.N - 1
    keys["conj"] = o.automorph_key(k)
    ctx.load_automorph_key(k, keys["conj"])
    yield P, o, ctx, keys
    ctx.close()


def _eq(g, ref):
    c0, c1 = g.store()
    assert c0.shape == ref[0].shape
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])}"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])}"


def _ct(o, P, level, seed):
    z = np.random.default_rng(seed).uniform(-1, 1, P.N // 2)
    return z, o.encrypt_slots(z, level, seed)


@pytest.mark.parametrize("to", [1, 3])
def test_mod_raise(env, to):
    P, o, ctx, _ = env
    log(f"mod raise to {to}")
    _, ct = _ct(o, P, 0, 31)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    out = g.mod_raise(to)
    assert out.level == to
    _eq(out, o.mod_raise(ct, to))
    g.mod_raise(to, out=g)  # in place
    _eq(g, o.mod_raise(ct, to))


def test_conjugate(env):
    P, o, ctx, keys = env
    log("conjugate")
    for level in (P.L, 1):
        z, ct = _ct(o, P, level, 40 + level)
        g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
        ref = o.rotate_k(ct, 2 * P.N - 1, keys["conj"])
        _eq(g.conjugate(), ref)
        _eq(g.automorph(2 * P.N - 1), ref)
        dec = o.decrypt_slots(*g.conjugate().store(), 2.0 ** P.scale_bits)
        assert np.abs(dec - z).max() < 1e-5  # real slots


@pytest.mark.parametrize("m", [3, 8192 + 5, 3 * 8192 // 2, -7])
def test_mult_monomial(env, m):
    P, o, ctx, _ = env
    log(f"monomial {m}")
    _, ct = _ct(o, P, P.L, 50)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    _eq(g.mult_monomial(m), o.mult_monomial(ct, m))


def test_mult_int(env):
    P, o, ctx, _ = env
    log("mult int")
    z, ct = _ct(o, P, 2, 60)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    s = 12345
    out = g.mult_int(s)
    _eq(out, o.mult_scalar(ct, [s % q for q in P.moduli[:3]]))
    dec = o.decrypt_slots(*out.store(), 2.0 ** P.scale_bits)
    assert np.abs(dec - s * z).max() < 1e-5 * s
