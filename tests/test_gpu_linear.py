"""GPU parity of the baby-step / giant-step linear transform (FIDESlib
CKKS::LinearTransform) against the oracle's composition of the same ops
(bit-exact), and its decryption against the slotwise closed form."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu

B, STRIDE, OFFSET, NDIAG = 4, 2, 3, 10  # g = 3 giant steps, the last one ragged


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    rots = sorted({i * STRIDE for i in range(1, B)} | {j * B * STRIDE for j in (1, 2)} | {OFFSET})
    o, ctx, keys = make_env(P, seed=0, rotations=tuple(rots))
    yield P, o, ctx, keys
    ctx.close()


def oracle_lt(o, ct, diags, keys, b, s, offset, batched_giant=False):
    n = len(diags)
    g = -(-n // b)
    baby = [ct] + o.rotate_hoisted(ct, [i * s for i in range(1, b)], [keys[i * s] for i in range(1, b)])
    inner = []
    for j in range(g):
        acc = None
        for i in range(b):
            if b * j + i >= n:
                break
            t = o.mult_plain(baby[i], diags[b * j + i])
            acc = t if acc is None else (o.add(acc[0], t[0]), o.add(acc[1], t[1]))
        inner.append(acc)
    if batched_giant:  # sum_j rot_{j b s}(inner[j])
        acc = inner[0]
        for j in range(1, g):
            r = o.rotate(inner[j], j * b * s, keys[j * b * s])
            acc = (o.add(acc[0], r[0]), o.add(acc[1], r[1]))
    else:  # Horner, the reference's order
        acc = inner[g - 1]
        for j in range(g - 2, -1, -1):
            r = o.rotate(acc, b * s, keys[b * s])
            acc = (o.add(inner[j][0], r[0]), o.add(inner[j][1], r[1]))
    return o.rotate(acc, offset, keys[offset]) if offset else acc


@pytest.mark.parametrize("level, batched", [(3, False), (1, False), (3, True), (2, True)])
def test_linear_transform(env, level, batched):
    P, o, ctx, keys = env
    log(f"linear transform level {level} batched giant steps {batched}")
    rng = np.random.default_rng(level)
    scale = 2.0 ** P.scale_bits
    nsl = P.N // 2
    z = rng.uniform(-1, 1, nsl)
    ct = o.encrypt_slots(z, level, 300 + level)
    dz = [rng.uniform(-1, 1, nsl) for _ in range(NDIAG)]
    dl = [o.encode(d, P.L) for d in dz]
    g = ctx.load(ct[0], ct[1], noise=scale)
    pts = [ctx.plaintext(d, noise=scale) for d in dl]
    out = g.linear_transform(pts, B, STRIDE, OFFSET, batched_giant=batched)
    ref = oracle_lt(o, ct, dl, keys, B, STRIDE, OFFSET, batched)
    c0, c1 = out.store()
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])}"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])}"
    info = out.info()
    assert info["noise_level"] == 2 and info["level"] == level
    # closed form over the slots (left rotation by r = np.roll(., -r))
    want = np.zeros(nsl)
    for j in range(-(-NDIAG // B)):
        inner = sum(dz[B * j + i] * np.roll(z, -i * STRIDE) for i in range(B) if B * j + i < NDIAG)
        want += np.roll(inner, -j * B * STRIDE)
    want = np.roll(want, -OFFSET)
    dec = o.decrypt_slots(c0, c1, info["noise"])
    assert np.abs(dec - want).max() < 1e-4


@pytest.mark.parametrize("batched", [False, True])
def test_linear_transform_batch(env, batched):
    """Three ciphertexts through one batched transform: each output equals the
    single transform bit for bit."""
    P, o, ctx, keys = env
    log(f"linear transform batch of 3, batched giant steps {batched}")
    rng = np.random.default_rng(7)
    scale = 2.0 ** P.scale_bits
    dl = [o.encode(rng.uniform(-1, 1, P.N // 2), P.L) for _ in range(NDIAG)]
    pts = [ctx.plaintext(d, noise=scale) for d in dl]
    gs = [ctx.load(*o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), 2, 400 + i), noise=scale) for i in range(3)]
    outs = ctx.linear_transform_batch(gs, pts, B, STRIDE, OFFSET, batched_giant=batched)
    wants = []
    for g, out in zip(gs, outs):
        want = g.linear_transform(pts, B, STRIDE, OFFSET, batched_giant=batched).store()
        got = out.store()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        wants.append(want)
    ctx.linear_transform_batch(gs, pts, B, STRIDE, OFFSET, outs=gs, batched_giant=batched)  # in place
    for g, want in zip(gs, wants):
        got = g.store()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def pq_diags(o, E, P, zs, scale):
    """The diagonals' encodings in P*Q form (limbs q_0..q_L, then the special primes)."""
    pq = E.ids(P.L)
    return [o.from_int(O.encode_coeffs(z, P.logN, scale), len(pq), pq) for z in zs]


def at_level(P, d, level):
    """A P*Q plaintext's limbs as a transform at `level` reads them."""
    return np.ascontiguousarray(np.concatenate([d[: level + 1], d[P.L + 1:]]))


@pytest.mark.parametrize("level, batched", [(3, False), (1, False), (3, True), (2, True)])
def test_linear_transform_ext(env, level, batched):
    """Diagonals in P*Q form take the extended-basis path (baby steps and inner
    sums in P*Q, only c1 brought down per giant step, one final ModDown):
    bit-exact against the oracle's restatement of that path, and decrypting
    to the slotwise closed form."""
    P, o, ctx, keys = env
    E = O.ExtBasis(o)
    log(f"linear transform, P*Q diagonals, level {level} batched giant steps {batched}")
    rng = np.random.default_rng(50 + level)
    scale = 2.0 ** P.scale_bits
    nsl = P.N // 2
    z = rng.uniform(-1, 1, nsl)
    ct = o.encrypt_slots(z, level, 500 + level)
    dz = [rng.uniform(-1, 1, nsl) for _ in range(NDIAG)]
    dl = pq_diags(o, E, P, dz, scale)
    g = ctx.load(ct[0], ct[1], noise=scale)
    pts = [ctx.plaintext_ext(d, noise=scale) for d in dl]
    out = g.linear_transform(pts, B, STRIDE, OFFSET, batched_giant=batched)
    ref = E.linear_transform(ct, [at_level(P, d, level) for d in dl], keys, B, STRIDE, OFFSET, batched)
    c0, c1 = out.store()
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])}"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])}"
    info = out.info()
    assert info["noise_level"] == 2 and info["level"] == level
    want = np.zeros(nsl)
    for j in range(-(-NDIAG // B)):
        inner = sum(dz[B * j + i] * np.roll(z, -i * STRIDE) for i in range(B) if B * j + i < NDIAG)
        want += np.roll(inner, -j * B * STRIDE)
    want = np.roll(want, -OFFSET)
    assert np.abs(o.decrypt_slots(c0, c1, info["noise"]) - want).max() < 1e-4
    # the engine's own P*Q encoder (its FP64 special FFT rounds on its own, so
    # compared through the decryption)
    pts2 = [ctx.encode(zz, P.L, scale, ext=True) for zz in dz]
    c0b, c1b = g.linear_transform(pts2, B, STRIDE, OFFSET, batched_giant=batched).store()
    assert np.abs(o.decrypt_slots(c0b, c1b, info["noise"]) - want).max() < 1e-4


@pytest.mark.parametrize("batched", [False, True])
def test_linear_transform_ext_batch(env, batched):
    """Three ciphertexts through one batched extended-basis transform: each
    output equals the single transform bit for bit, also in place."""
    P, o, ctx, keys = env
    E = O.ExtBasis(o)
    log(f"linear transform batch of 3, P*Q diagonals, batched giant steps {batched}")
    rng = np.random.default_rng(8)
    scale = 2.0 ** P.scale_bits
    dl = pq_diags(o, E, P, [rng.uniform(-1, 1, P.N // 2) for _ in range(NDIAG)], scale)
    pts = [ctx.plaintext_ext(d, noise=scale) for d in dl]
    gs = [ctx.load(*o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), 2, 600 + i), noise=scale) for i in range(3)]
    outs = ctx.linear_transform_batch(gs, pts, B, STRIDE, OFFSET, batched_giant=batched)
    wants = []
    for g, out in zip(gs, outs):
        want = g.linear_transform(pts, B, STRIDE, OFFSET, batched_giant=batched).store()
        got = out.store()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        wants.append(want)
    ctx.linear_transform_batch(gs, pts, B, STRIDE, OFFSET, outs=gs, batched_giant=batched)  # in place
    for g, want in zip(gs, wants):
        got = g.store()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("size, stride, slots", [(8, 1, 4096), (16, 2, 4096), (512, 8, 4096)])
def test_accumulate_sum(size, stride, slots):
    """AccumulateSum against the oracle's composition of the same extended-basis
    hoisted sets, adds and ModDowns (bit-exact) and the slotwise closed form; a
    sum over every slot (size * stride == slots) leaves `stride` slots."""
    import pyfideslib_amd as F

    P = O.params_small()
    rots = F.accumulate_rotations(size, stride)
    o, ctx, keys = make_env(P, seed=9, rotations=tuple(rots))
    try:
        log(f"accumulate sum size {size} stride {stride}")
        rng = np.random.default_rng(size)
        z = rng.uniform(-1, 1, P.N // 2)
        ct = o.encrypt_slots(z, 2, 90 + size)
        g = ctx.load(*ct, noise=2.0 ** P.scale_bits, slots=slots)
        out = g.accumulate_sum(size, stride)
        assert out.info()["slots"] == (stride if size * stride == slots else slots)
        # each radix-4 step as the reference: hoisted rotations left in P*Q,
        # added to the extended running sum, one ModDown
        E = O.ExtBasis(o)
        acc = ct
        s = 1
        while s < size:
            idx = [i * s * stride for i in range(1, 4) if i * s < size]
            tot = (E.extend(acc[0]), E.extend(acc[1]))
            for r in idx:
                h = E.rotate(acc[0], acc[1], O.rot_index(P.logN, r), keys[r])
                tot = (E.add(tot[0], h[0]), E.add(tot[1], h[1]))
            acc = E.moddown(tot)
            s *= 4
        c0, c1 = out.store()
        assert np.array_equal(c0, acc[0]) and np.array_equal(c1, acc[1])
        want = sum(np.roll(z, -k * stride) for k in range(size))
        assert np.abs(o.decrypt_slots(c0, c1, out.info()["noise"]) - want).max() < 1e-4
    finally:
        ctx.close()


def test_linear_transform_ext_production():
    """The extended-basis transform at the production parameters (N = 2^16,
    L = 24, dnum = 3, K = 8): 4 P*Q diagonals, 2 baby steps, a giant step
    and an offset rotation, bit-exact against the oracle's restatement."""
    P = O.params_production()
    o, ctx, keys = make_env(P, seed=3, rotations=(1, 2, 3), relin=False)
    try:
        log("P*Q linear transform at N=2^16, L=24")
        E = O.ExtBasis(o)
        rng = np.random.default_rng(60)
        scale = 2.0 ** P.scale_bits
        level = P.L
        ct = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), level, 61)
        dl = pq_diags(o, E, P, [rng.uniform(-1, 1, P.N // 2) for _ in range(4)], scale)
        g = ctx.load(*ct, noise=scale)
        pts = [ctx.plaintext_ext(d, noise=scale) for d in dl]
        got = g.linear_transform(pts, 2, 1, 3).store()
        want = E.linear_transform(ct, [at_level(P, d, level) for d in dl], keys, 2, 1, 3)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    finally:
        ctx.close()
