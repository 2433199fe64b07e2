"""GPU parity of the plaintext and scalar ops (FIDESlib Ciphertext::multPt,
addPt, multScalar, addScalar, dropToLevel) against the oracle, bit-exact, plus
the scalar encodings computed by the library against the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=0, rotations=(1,))
    yield P, o, ctx, keys
    ctx.close()


def _ct(o, P, level, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, P.N // 2)
    return z, o.encrypt_slots(z, level, seed)


def _eq(g, ref):
    c0, c1 = g.store()
    assert c0.shape == ref[0].shape
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])}"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])}"


def test_scalar_encodings_match_oracle(env):
    P, o, ctx, _ = env
    log("scalar encodings")
    sf = 2.0 ** P.scale_bits
    for level in (P.L, 1):
        for v in (0.3, -1.7, 2.0 ** 70, -3e-9, 0.0):
            got = ctx.scalar_encoding(level, v, "mult")
            assert got.tolist() == O.elem_for_mult(P.moduli, level, v, sf)
        for v, deg in ((0.25, 1), (1.5, 2), (2.0 ** 30, 1), (0.7, 3)):
            got = ctx.scalar_encoding(level, v, "add", deg)
            assert got.tolist() == O.elem_for_add(P.moduli, level, v, deg, sf)


@pytest.mark.parametrize("rescale", [False, True])
def test_mult_plain(env, rescale):
    P, o, ctx, _ = env
    log(f"mult plain rescale={rescale}")
    scale = 2.0 ** P.scale_bits
    for level in (P.L, 2):
        z, ct = _ct(o, P, level, 40 + level)
        w = np.random.default_rng(7).uniform(-1, 1, P.N // 2)
        pt_limbs = o.encode(w, P.L)  # more limbs than the ciphertext: its first level+1 are used
        g = ctx.load(ct[0], ct[1], noise=scale)
        pt = ctx.plaintext(pt_limbs, noise=scale)
        out = g.mult_plain(pt, rescale=rescale)
        ref = o.mult_plain(ct, pt_limbs)
        if rescale:
            ref = o.rescale(ref)
        _eq(out, ref)
        info = out.info()
        assert info["noise_level"] == (1 if rescale else 2)
        d = o.decrypt_slots(*out.store(), info["noise"])
        assert np.abs(d - z * w).max() < 1e-5


def test_add_plain_and_in_place(env):
    P, o, ctx, _ = env
    log("add plain")
    scale = 2.0 ** P.scale_bits
    z, ct = _ct(o, P, P.L, 61)
    w = np.random.default_rng(8).uniform(-1, 1, P.N // 2)
    pt_limbs = o.encode(w, P.L)
    pt = ctx.plaintext(pt_limbs, noise=scale)
    g = ctx.load(ct[0], ct[1], noise=scale)
    _eq(g.add_plain(pt), o.add_plain(ct, pt_limbs))
    g.add_plain(pt, out=g)  # in place
    _eq(g, o.add_plain(ct, pt_limbs))


@pytest.mark.parametrize("v", [0.3, -1.7])
def test_mult_scalar(env, v):
    P, o, ctx, _ = env
    log(f"mult scalar {v}")
    scale = 2.0 ** P.scale_bits
    for level in (P.L, 1):
        z, ct = _ct(o, P, level, 70 + level)
        g = ctx.load(ct[0], ct[1], noise=scale)
        elems = O.elem_for_mult(P.moduli, level, v, scale)
        ref = o.mult_scalar(ct, elems)
        _eq(g.mult_scalar(v), ref)
        out = g.mult_scalar(v, rescale=True)
        _eq(out, o.rescale(ref))
        d = o.decrypt_slots(*out.store(), out.info()["noise"])
        assert np.abs(d - v * z).max() < 1e-5


@pytest.mark.parametrize("v", [0.25, -0.6, 0.0])
def test_add_scalar(env, v):
    P, o, ctx, keys = env
    log(f"add scalar {v}")
    scale = 2.0 ** P.scale_bits
    z, ct = _ct(o, P, P.L, 80)
    g = ctx.load(ct[0], ct[1], noise=scale)

    def ref_add(c, deg):
        e = O.elem_for_add(P.moduli, c[0].shape[0] - 1, abs(v), deg, scale)
        if v < 0:
            e = [(q - x) % q for q, x in zip(P.moduli, e)]
        return o.add_scalar(c, e) if v != 0 else (c[0].copy(), c[1].copy())

    _eq(g.add_scalar(v), ref_add(ct, 1))
    # noise degree 2 (after a product): the constant carries scale^2
    m = g.mult(g)
    mref = o.mult(ct, ct, keys["relin"])
    _eq(m.add_scalar(v), ref_add(mref, 2))


def test_drop_to_level_then_add(env):
    P, o, ctx, _ = env
    log("drop to level")
    scale = 2.0 ** P.scale_bits
    _, hi = _ct(o, P, P.L, 90)
    _, lo = _ct(o, P, 1, 91)
    g_hi = ctx.load(hi[0], hi[1], noise=scale).drop_to_level(1)
    assert g_hi.level == 1
    g_lo = ctx.load(lo[0], lo[1], noise=scale)
    _eq(g_hi.add(g_lo), (o.add(hi[0][:2].copy(), lo[0]), o.add(hi[1][:2].copy(), lo[1])))
    with pytest.raises(Exception):
        g_hi.drop_to_level(2)


def test_negate(env):
    """EvalNegate: every residue q - x (0 stays 0), metadata unchanged, in place too."""
    P, o, ctx, _ = env
    for level in (P.L, 0):
        z = np.random.default_rng(level).uniform(-1, 1, P.N // 2)
        ct = o.encrypt_slots(z, level, 600 + level)
        g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
        out = g.negate()
        q = np.array(P.moduli[: level + 1], dtype=np.uint64)[:, None]
        for got, x in zip(out.store(), ct):
            assert np.array_equal(got, np.where(x == 0, x, q - x))
        assert out.info() == g.info()
        g.negate(out=g)
        assert np.array_equal(g.store()[0], out.store()[0])


def test_add_many(env):
    """EvalAddMany: equal to adds one after another (bit-exact), at the lowest
    level, also in place over the first input."""
    P, o, ctx, _ = env
    rng = np.random.default_rng(77)
    cts = [o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), lv, 700 + i) for i, lv in enumerate((P.L, 2, P.L, 1, 2))]
    # slot counts differ: the result carries the widest, as chained adds do
    slots = (64, 256, 128, 32, 16)
    gs = [ctx.load(*ct, noise=2.0 ** P.scale_bits, slots=sl) for ct, sl in zip(cts, slots)]
    out = ctx.add_many(gs)
    want = gs[0].copy()
    for g in gs[1:]:
        want.add(g, out=want)
    a, b = out.store(), want.store()
    assert out.level == 1 and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert out.info() == want.info() and out.info()["slots"] == 256
    ctx.add_many(gs, out=gs[0])
    c = gs[0].store()
    assert np.array_equal(c[0], b[0]) and np.array_equal(c[1], b[1])
