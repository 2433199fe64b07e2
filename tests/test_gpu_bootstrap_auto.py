"""Bootstrapping under FLEXIBLEAUTO and FIXEDAUTO on the GPU (FIDESlib
CKKS::Bootstrap with rescaleTechnique FLEXIBLEAUTO / FIXEDAUTO,
src/CKKS/Bootstrap.cu:169-300, :428-530; the reference's bootstrap tests run
both FLEXIBLEAUTO and FIXEDMANUAL parameter sets).

The input, at a level above 0 and at either noise degree, is first brought
onto the plan's scale by an integer product and a rescale (the reference's
adjustment product before ModRaise); SlotsToCoeffs is encoded at the scale of
each level it runs at, so under FLEXIBLEAUTO the result sits exactly on its
level's ScalingFactorReal and adds to a fresh ciphertext of that level with no
alignment.  The chain is numeric (EvalMod approximates), so every result is
checked against the message."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pyfideslib_amd import bootstrap as B
from tests.gpu_util import log

pytestmark = pytest.mark.gpu

LOGN, L = 12, 20


def _env(tech):
    import pyfideslib_amd as F

    P = O.make_params(LOGN, 60, 50, L, 60, 3)
    if tech == O.FLEXIBLEAUTO:
        sf, sfb = O.flexible_scaling_factors(P.q)
    else:
        sf = [2.0 ** P.scale_bits] * (P.L + 1)
        sfb = [x * x for x in sf]
    o = O.Oracle(P, seed=5, hamming=192)
    ctx = F.Context(F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, tech, sf, sfb))
    ctx.load_eval_key(o.relin_key())
    bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119)
    for r in bs.rotations:
        ctx.load_rotation_key(r, o.rotation_key(r))
    k = 2 * P.N - 1
    ctx.load_automorph_key(k, o.automorph_key(k))
    return P, o, ctx, bs, sf


@pytest.fixture(scope="module")
def flex():
    env = _env(O.FLEXIBLEAUTO)
    yield env
    env[3].close()
    env[2].close()


def _slots(rng, n):
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


@pytest.mark.parametrize("level_in", [3, 1, 7])
def test_flexibleauto_bootstrap(flex, level_in):
    P, o, ctx, bs, sf = flex
    log(f"FLEXIBLEAUTO bootstrap, input level {level_in}")
    rng = np.random.default_rng(level_in)
    z = _slots(rng, P.N // 2)
    ct = o.encrypt_slots(z, level_in, 40 + level_in, sf[level_in])
    g = ctx.load(*ct, noise=sf[level_in])
    out = bs.bootstrap(g)
    info = out.info()
    assert info["noise_level"] == 1 and 1 <= info["level"] < L - 12
    assert info["noise"] == sf[info["level"]]
    err = np.abs(o.decrypt_slots(*out.store(), info["noise"]) - z).max()
    log(f"FLEXIBLEAUTO bootstrap: output level {info['level']}, max error {err:.3e} ({-np.log2(err):.1f} bits)")
    assert err < 1e-3
    c0, c1 = g.store()  # the input is untouched
    assert np.array_equal(c0, ct[0]) and np.array_equal(c1, ct[1])


def test_flexibleauto_degree_two_input_and_use(flex):
    """A product left at noise degree 2 is rescaled first; the result adds to a
    fresh ciphertext of its level (same FLEXIBLEAUTO factor, no alignment) and
    multiplies on."""
    P, o, ctx, bs, sf = flex
    log("FLEXIBLEAUTO bootstrap of a degree-2 product, then add and mult")
    rng = np.random.default_rng(9)
    z1, z2 = rng.uniform(-1, 1, P.N // 2), rng.uniform(-1, 1, P.N // 2)
    lv = 4
    a = ctx.load(*o.encrypt_slots(z1, lv, 91, sf[lv]), noise=sf[lv])
    b = ctx.load(*o.encrypt_slots(z2, lv, 92, sf[lv]), noise=sf[lv])
    prod = a.mult(b)
    assert prod.info()["noise_level"] == 2
    out = bs.bootstrap(prod)
    info = out.info()
    err = np.abs(o.decrypt_slots(*out.store(), info["noise"]) - z1 * z2).max()
    log(f"degree-2 input: max error {err:.3e} ({-np.log2(err):.1f} bits)")
    assert err < 1e-3
    lo = info["level"]
    z3 = rng.uniform(-1, 1, P.N // 2)
    fresh = ctx.load(*o.encrypt_slots(z3, lo, 93, sf[lo]), noise=sf[lo])
    s = out.add(fresh)
    assert s.info()["level"] == lo and s.info()["noise"] == sf[lo]
    assert np.abs(o.decrypt_slots(*s.store(), sf[lo]) - (z1 * z2 + z3)).max() < 1e-3
    m = out.mult(fresh)  # FLEXIBLEAUTO: left at degree 2, rescaled by the next product
    mi = m.info()
    assert np.abs(o.decrypt_slots(*m.store(), mi["noise"]) - z1 * z2 * z3).max() < 1e-3


def test_flexibleauto_batch_equals_single(flex):
    """Two ciphertexts at different levels through one batched bootstrap equal
    their single bootstraps bit for bit."""
    P, o, ctx, bs, sf = flex
    log("FLEXIBLEAUTO bootstrap batch of 2")
    rng = np.random.default_rng(13)
    zs = [_slots(rng, P.N // 2) for _ in range(2)]
    gs = [ctx.load(*o.encrypt_slots(z, lv, 50 + lv, sf[lv]), noise=sf[lv]) for z, lv in zip(zs, (2, 5))]
    outs = bs.bootstrap_batch(gs)
    for g, z, out in zip(gs, zs, outs):
        single = bs.bootstrap(g)
        got, want = out.store(), single.store()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert out.info() == single.info()
        assert np.abs(o.decrypt_slots(*got, out.info()["noise"]) - z).max() < 1e-3


def test_flexibleauto_level_zero_refused(flex):
    """The scale adjustment needs a level to rescale into (the reference
    asserts level - noise degree + 1 >= 1)."""
    import pyfideslib_amd as F

    P, o, ctx, bs, sf = flex
    g = ctx.load(*o.encrypt_slots(np.zeros(P.N // 2), 0, 60, sf[0]), noise=sf[0])
    with pytest.raises(F.FidesError):
        bs.bootstrap(g)


def test_fixedauto_bootstrap():
    P, o, ctx, bs, sf = _env(O.FIXEDAUTO)
    try:
        log("FIXEDAUTO bootstrap, input level 2")
        z = _slots(np.random.default_rng(17), P.N // 2)
        out = bs.bootstrap(ctx.load(*o.encrypt_slots(z, 2, 71), noise=sf[2]))
        info = out.info()
        assert info["noise_level"] == 1
        err = np.abs(o.decrypt_slots(*out.store(), info["noise"]) - z).max()
        log(f"FIXEDAUTO bootstrap: output level {info['level']}, max error {err:.3e} ({-np.log2(err):.1f} bits)")
        assert err < 1e-3
    finally:
        bs.close()
        ctx.close()


def test_flexibleauto_bootstrap_production_size():
    """The bootstrap workload at full size under FLEXIBLEAUTO: N = 2^16, q0
    60-bit + 24 x 50-bit, dnum 3, a dense main secret with sparse-secret
    encapsulation (h = 192), 2^15 real slots, level budget (3, 3), degree-119
    EvalMod; the input at level 2 with its FLEXIBLEAUTO factor.  Same precision
    bounds as the FIXEDMANUAL workload test, and the result on its level's
    factor."""
    import time

    import pyfideslib_amd as F
    from tests.test_gpu_bootstrap import sparse_secret

    P = O.params_production()
    sf, sfb = O.flexible_scaling_factors(P.q)
    o = O.Oracle(P, seed=1)
    ctx = F.Context(F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, O.FLEXIBLEAUTO, sf, sfb))
    bs = None
    try:
        ctx.load_eval_key(o.relin_key())
        sp = sparse_secret(o, P.N, 192, 5)
        ctx.load_switch_key(F.SSE_DENSE_TO_SPARSE, o.ksk_gen(o.s, sp, 911))
        ctx.load_switch_key(F.SSE_SPARSE_TO_DENSE, o.ksk_gen(sp, o.s, 912))
        bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119,
                            sparse_encaps=True)
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        k = 2 * P.N - 1
        ctx.load_automorph_key(k, o.automorph_key(k))
        z = np.random.default_rng(4).uniform(-1, 1, P.N // 2)
        g = ctx.load(*o.encrypt_slots(z, 2, 4, sf[2]), noise=sf[2])
        out = bs.bootstrap(g)  # the first call also encodes SlotsToCoeffs
        ctx.synchronize()
        t = time.perf_counter()
        out = bs.bootstrap(g)
        ctx.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        info = out.info()
        dec = o.decrypt_slots(*out.store(), info["noise"])
        err = np.abs(dec - z).max()
        log(f"FLEXIBLEAUTO production bootstrap: output level {info['level']}, {ms:.1f} ms, max error {err:.3e} "
            f"({-np.log2(err):.1f} bits), mean error {np.abs(dec - z).mean():.1e}")
        assert info["level"] == 8 and info["noise"] == sf[8]
        assert err < 2.0 ** -5.7 and np.abs(dec - z).mean() < 2.0 ** -7.8
    finally:
        if bs is not None:
            bs.close()
        ctx.close()
