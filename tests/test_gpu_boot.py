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


def _neg_elems(P, level, v, deg, sf):
    e = O.elem_for_add(P.moduli, level, abs(v), deg, sf)
    return [(q - x) % q for q, x in zip(P.moduli, e)] if v < 0 else e


def test_horner_series(env):
    """Degree-3 polynomial (the scoring sigmoid) by Horner's rule, FIDESlib's
    rescale placement, against the oracle's composition and the closed form."""
    P, o, ctx, keys = env
    log("horner series")
    sf = 2.0 ** P.scale_bits
    coeffs = [0.5, 0.197, 0.01, -0.004]
    z, ct = _ct(o, P, P.L, 70)
    g = ctx.load(*ct, noise=sf)
    out = g.horner_series(coeffs)
    # oracle: acc = x c3 + c2 (degree 2), then rescale / mult / rescale / add
    n = len(coeffs) - 1
    acc = o.mult_scalar(ct, O.elem_for_mult(P.moduli, P.L, coeffs[n], sf))
    acc = o.add_scalar(acc, _neg_elems(P, P.L, coeffs[n - 1], 2, sf))
    deg = 2
    for i in range(n - 2, -1, -1):
        if deg == 2:
            acc = o.rescale(acc)
        lv = acc[0].shape[0]
        acc = o.rescale(o.mult(acc, (np.ascontiguousarray(ct[0][:lv]), np.ascontiguousarray(ct[1][:lv])),
                               keys["relin"]))
        deg = 1
        acc = o.add_scalar(acc, _neg_elems(P, acc[0].shape[0] - 1, coeffs[i], 1, sf))
    _eq(out, acc)
    dec = o.decrypt_slots(*out.store(), out.info()["noise"])
    want = sum(c * z ** k for k, c in enumerate(coeffs))
    assert np.abs(dec - want).max() < 1e-4


def test_double_angle(env):
    P, o, ctx, keys = env
    log("double angle")
    sf = 2.0 ** P.scale_bits
    za, a = _ct(o, P, P.L, 80)
    zb, b = _ct(o, P, P.L, 81)
    ga, gb = ctx.load(*a, noise=sf), ctx.load(*b, noise=sf)
    m = ga.mult(gb)  # noise degree 2, as EvalMod's series leaves it
    r = 2
    out = m.double_angle(r)
    ref = o.mult(a, b, keys["relin"])
    y = za * zb
    for j in range(1, r + 1):
        ref = o.rescale(ref)
        ref = o.square(ref, keys["relin"])
        ref = (o.add(ref[0], ref[0]), o.add(ref[1], ref[1]))
        v = -1.0 / (2 * np.pi) ** (2.0 ** (j - r))
        ref = o.add_scalar(ref, _neg_elems(P, ref[0].shape[0] - 1, v, 2, sf))
        y = 2 * y * y + v
    _eq(out, ref)
    dec = o.decrypt_slots(*out.store(), out.info()["noise"])
    assert np.abs(dec - y).max() < 1e-4
