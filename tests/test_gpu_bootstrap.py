"""Bootstrapping of fully packed ciphertexts on the GPU (FIDESlib
CKKS::Bootstrap, src/CKKS/Bootstrap.cu:169-300): a level-0 ciphertext under a
sparse ternary secret (h = 192), or under a dense one with sparse-secret
encapsulation, is raised, taken through CoeffsToSlots, EvalMod and
SlotsToCoeffs, and must decrypt to the slots it started with.
Every op in the chain is bit-exact-tested against the oracle on its own; the
chain is checked against the message (the EvalMod approximation makes it
numeric).  The engine's encoder is checked against the oracle's encoding."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pyfideslib_amd import bootstrap as B
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu

LOGN, L = 12, 20


@pytest.fixture(scope="module")
def env():
    P = O.make_params(LOGN, 60, 50, L, 60, 3)
    o, ctx, keys = make_env(P, seed=0, hamming=192)
    bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119)
    for r in bs.rotations:
        keys[r] = o.rotation_key(r)
        ctx.load_rotation_key(r, keys[r])
    k = 2 * P.N - 1
    ctx.load_automorph_key(k, o.automorph_key(k))
    log(f"bootstrap setup: {len(bs.rotations)} rotation keys")
    yield P, o, ctx, bs
    bs.close()
    ctx.close()


def test_setup_layout_matches_model(env):
    """The C++ setup needs the rotations the FP64 layout model derives."""
    P, o, ctx, bs = env
    model = B.LayoutModel(P.N, P.moduli[0], 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3,
                          degree=119)
    assert bs.rotations == model.rotations


def test_encode(env):
    """The engine's encoder against the oracle's encoding: ct x pt decrypts the
    same with either plaintext, and to the slotwise product."""
    P, o, ctx, _ = env
    log("encode")
    rng = np.random.default_rng(1)
    z = rng.uniform(-1, 1, P.N // 2) + 1j * rng.uniform(-1, 1, P.N // 2)
    x = rng.uniform(-1, 1, P.N // 2)
    scale = 2.0 ** P.scale_bits
    for level in (1, 7, P.L):
        ct = o.encrypt_slots(x, level, 70 + level)
        g = ctx.load(*ct, noise=scale)
        mine = ctx.encode(z, level, scale)
        ref = ctx.plaintext(o.encode(z, level), noise=scale)
        a = o.decrypt_slots(*g.mult_plain(mine).store(), scale * scale)
        b = o.decrypt_slots(*g.mult_plain(ref).store(), scale * scale)
        assert np.abs(a - b).max() < 1e-9
        assert np.abs(a - x * z).max() < 1e-6
    import pyfideslib_amd as F

    with pytest.raises(ValueError):
        ctx.encode(z[:-1], 1, scale)
    with pytest.raises(F.FidesError):
        ctx.encode(z, P.L + 1, scale)
    with pytest.raises(F.FidesError):
        ctx.encode(z * 2.0 ** 30, 1, scale)  # coefficients beyond 62 bits
    real = ctx.encode(x, 2, scale)  # real slots (no imaginary part)
    ct = o.encrypt_slots(x, 2, 79)
    d = o.decrypt_slots(*ctx.load(*ct, noise=scale).mult_plain(real).store(), scale * scale)
    assert np.abs(d - x * x).max() < 1e-6


@pytest.mark.parametrize("seed, level_in", [(11, 0), (12, 3)])
def test_bootstrap_full_slots(env, seed, level_in):
    P, o, ctx, bs = env
    log(f"bootstrap seed {seed} input level {level_in}")
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, P.N // 2) + 1j * rng.uniform(-1, 1, P.N // 2)
    ct = o.encrypt_slots(z, level_in, seed)
    g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
    out = bs.bootstrap(g)
    info = out.info()
    assert info["noise_level"] == 1 and 1 <= info["level"] < L - 12
    dec = o.decrypt_slots(*out.store(), info["noise"])
    err = np.abs(dec - z).max()
    log(f"bootstrap: output level {info['level']}, max error {err:.3e} ({-np.log2(err):.1f} bits)")
    assert err < 1e-3
    # the input is untouched
    c0, c1 = g.store()
    assert np.array_equal(c0, ct[0]) and np.array_equal(c1, ct[1])


def test_bootstrap_batch_equals_single(env):
    """Three ciphertexts at different levels bootstrapped together: each output
    equals its own bootstrap bit for bit (the batched key switches are exact)
    and decrypts to its slots."""
    P, o, ctx, bs = env
    log("bootstrap batch of 3")
    rng = np.random.default_rng(21)
    zs = [rng.uniform(-1, 1, P.N // 2) + 1j * rng.uniform(-1, 1, P.N // 2) for _ in range(3)]
    gs = [ctx.load(*o.encrypt_slots(z, lv, 30 + lv), noise=2.0 ** P.scale_bits) for z, lv in zip(zs, (0, 2, 5))]
    outs = bs.bootstrap_batch(gs)
    for g, z, out in zip(gs, zs, outs):
        single = bs.bootstrap(g).store()
        got = out.store()
        assert np.array_equal(got[0], single[0]) and np.array_equal(got[1], single[1])
        dec = o.decrypt_slots(*got, out.info()["noise"])
        assert np.abs(dec - z).max() < 1e-3
    bs.bootstrap_batch(gs, outs=gs)  # in place
    for g, out in zip(gs, outs):
        a, b = g.store(), out.store()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_bootstrap_other_level_budgets(env):
    """Level budgets (2, 4): a six-stage CoeffsToSlots group whose offsets span
    every multiple of its stride, and a two-stage SlotsToCoeffs group; a
    lower-degree EvalMod (K = 28, r = 3, degree 89)."""
    P, o, ctx, _ = env
    log("bootstrap with level budgets (2, 4), degree 89")
    bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(2, 4), K=28, double_angle=3, degree=89)
    try:
        model = B.LayoutModel(P.N, P.moduli[0], 2.0 ** P.scale_bits, level_budget=(2, 4), K=28, double_angle=3,
                              degree=89)
        assert bs.rotations == model.rotations
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        z = np.random.default_rng(24).uniform(-1, 1, P.N // 2)
        out = bs.bootstrap(ctx.load(*o.encrypt_slots(z, 0, 24), noise=2.0 ** P.scale_bits))
        dec = o.decrypt_slots(*out.store(), out.info()["noise"])
        err = np.abs(dec - z).max()
        log(f"bootstrap (2, 4): output level {out.level}, max error {err:.3e} ({-np.log2(err):.1f} bits)")
        assert err < 1e-3
    finally:
        bs.close()


@pytest.mark.parametrize("slots", [256, 32])
def test_bootstrap_sparse_slots(env, slots):
    """Sparse packing (EvalBootstrapSetup's `slots` < N/2): the message repeats
    with period `slots`; the subring trace, the split into blocks, one EvalMod
    and the fold return it.  The rotation set matches the layout model."""
    P, o, ctx, _ = env
    log(f"bootstrap with {slots} slots")
    bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119,
                        slots=slots)
    try:
        model = B.LayoutModel(P.N, P.moduli[0], 2.0 ** P.scale_bits, level_budget=(3, 3), slots=slots)
        assert bs.rotations == model.rotations
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        rng = np.random.default_rng(slots)
        zp = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
        z = np.tile(zp, P.N // 2 // slots)
        out = bs.bootstrap(ctx.load(*o.encrypt_slots(z, 0, slots), noise=2.0 ** P.scale_bits))
        dec = o.decrypt_slots(*out.store(), out.info()["noise"])
        err = np.abs(dec - z).max()
        log(f"bootstrap {slots} slots: output level {out.level}, max error {err:.3e} ({-np.log2(err):.1f} bits)")
        assert err < 1e-3
    finally:
        bs.close()


def test_bootstrap_batched_giant_steps(env):
    """Giant steps of every transform as one batched rotation: the rotation set
    of the layout model, and the message back."""
    P, o, ctx, _ = env
    log("bootstrap with batched giant steps")
    bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), batched_giant=True)
    try:
        model = B.LayoutModel(P.N, P.moduli[0], 2.0 ** P.scale_bits, level_budget=(3, 3), batched_giant=True)
        assert bs.rotations == model.rotations
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        z = np.random.default_rng(31).uniform(-1, 1, P.N // 2)
        out = bs.bootstrap(ctx.load(*o.encrypt_slots(z, 0, 31), noise=2.0 ** P.scale_bits))
        err = np.abs(o.decrypt_slots(*out.store(), out.info()["noise"]) - z).max()
        log(f"bootstrap, batched giant steps: max error {err:.3e} ({-np.log2(err):.1f} bits)")
        assert err < 1e-3
    finally:
        bs.close()


def sparse_secret(o, N, h, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros(N, dtype=np.int64)
    s[rng.choice(N, h, replace=False)] = rng.choice(np.array([-1, 1]), h)
    return o.from_int(s, o.np)


def test_bootstrap_production_size():
    """The bootstrap workload at full size: N = 2^16, q0 60-bit + 24 x 50-bit,
    dnum 3, 60-bit special primes, a dense main secret with sparse-secret
    encapsulation (h = 192), 2^15 real slots uniform in [-1, 1], level budget
    (3, 3), degree-119 EvalMod."""
    import pyfideslib_amd as F

    P = O.params_production()
    o, ctx, keys = make_env(P, seed=1)
    bs = None
    try:
        sp = sparse_secret(o, P.N, 192, 5)
        ctx.load_switch_key(F.SSE_DENSE_TO_SPARSE, o.ksk_gen(o.s, sp, 911))
        ctx.load_switch_key(F.SSE_SPARSE_TO_DENSE, o.ksk_gen(sp, o.s, 912))
        bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119,
                            sparse_encaps=True)
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        k = 2 * P.N - 1
        ctx.load_automorph_key(k, o.automorph_key(k))
        log(f"production bootstrap: {len(bs.rotations)} rotation keys")
        z = np.random.default_rng(4).uniform(-1, 1, P.N // 2)
        ct = o.encrypt_slots(z, 0, 4)
        out = bs.bootstrap(ctx.load(*ct, noise=2.0 ** P.scale_bits))
        info = out.info()
        dec = o.decrypt_slots(*out.store(), info["noise"])
        err = np.abs(dec - z).max()
        log(f"production bootstrap: output level {info['level']}, max error {err:.3e} ({-np.log2(err):.1f} bits), "
            f"mean error {np.abs(dec - z).mean():.1e}, imaginary part {np.abs(dec.imag).max():.1e}")
        assert info["level"] == 8
        # rounding noise of every rescale / key switch grows with the dense
        # secret's norm and is amplified by K 2^10 sqrt(n) through EvalMod and
        # SlotsToCoeffs: measured 6.7-6.9 bits worst case and 8.8 bits mean
        # at a 2^50 scale (seeds 4 and 5); the bounds are 1 bit below that
        assert err < 2.0 ** -5.7 and np.abs(dec - z).mean() < 2.0 ** -7.8
    finally:
        if bs is not None:
            bs.close()
        ctx.close()


def test_bootstrap_sparse_secret_encapsulation():
    """Dense main secret (uniform ternary); the level-0 input is switched to an
    ephemeral sparse secret (h = 192) for ModRaise and back after it; the
    generic key switch is checked against the oracle on its own."""
    P = O.make_params(LOGN, 60, 50, L, 60, 3)
    o, ctx, keys = make_env(P, seed=2)  # dense secret
    bs = None
    try:
        import pyfideslib_amd as F

        rng = np.random.default_rng(77)
        sp = sparse_secret(o, P.N, 192, 77)
        down, up = o.ksk_gen(o.s, sp, 901), o.ksk_gen(sp, o.s, 902)
        ctx.load_switch_key(F.SSE_DENSE_TO_SPARSE, down)
        ctx.load_switch_key(F.SSE_SPARSE_TO_DENSE, up)
        with pytest.raises(F.FidesError) as e:  # no key under this id
            ctx.load(*o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), 1, 59), noise=1.0).key_switch(7)
        assert e.value.code == 4  # FIDES_E_NOKEY
        # the key switch alone: bit-exact against the oracle's, at level 0 and at the top
        z = rng.uniform(-1, 1, P.N // 2)
        for level, kid, key in ((0, F.SSE_DENSE_TO_SPARSE, down), (P.L, F.SSE_SPARSE_TO_DENSE, up)):
            ct = o.encrypt_slots(z, level, 60 + level)
            got = ctx.load(*ct, noise=2.0 ** P.scale_bits).key_switch(kid).store()
            kb, ka = o.keyswitch(ct[1], key)
            want = (o.add(ct[0], kb), ka)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        bs = B.Bootstrapper(ctx, 2.0 ** P.scale_bits, level_budget=(3, 3), K=28, double_angle=3, degree=119,
                            sparse_encaps=True)
        for r in bs.rotations:
            ctx.load_rotation_key(r, o.rotation_key(r))
        k = 2 * P.N - 1
        ctx.load_automorph_key(k, o.automorph_key(k))
        zc = z + 1j * rng.uniform(-1, 1, P.N // 2)
        ct = o.encrypt_slots(zc, 0, 78)
        out = bs.bootstrap(ctx.load(*ct, noise=2.0 ** P.scale_bits))
        info = out.info()
        dec = o.decrypt_slots(*out.store(), info["noise"])
        err = np.abs(dec - zc).max()
        log(f"bootstrap with sparse-secret encapsulation: output level {info['level']}, max error {err:.3e} "
            f"({-np.log2(err):.1f} bits)")
        assert err < 1e-3
    finally:
        if bs is not None:
            bs.close()
        ctx.close()
