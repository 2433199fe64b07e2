"""GPU parity of the convolution transform (FIDESlib CKKS::ConvolutionTransform,
src/CKKS/LinearTransform.cu:255-441, its Q-basis path) against the oracle's
composition of the same operations (bit-exact), and its decryption against the
slotwise closed form."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu

HALF = 2 ** 12  # slots of the small parameter set (N = 2^13)
# (bstep, gstep, n plaintexts, stride, baby-step indexes, input level)
CASES = [
    (3, 3, 9, 2, (0, 5, 11), 3),  # one block of 3
    (2, 10, 19, 1, (0, 1), 2),  # blocks of 8 and 2, the last giant step ragged
    (1, 17, 17, 1, (7,), 3),  # blocks of 8, 8 and 1
    (2, 3, 4, 3, (0, 2), 1),  # a giant step with no plaintext (zero)
    (1, 2, 2, HALF // 2, (0,), 2),  # the block's first rotation is by N/2: skipped
]


def _rotations():
    import pyfideslib_amd as F

    need = set()
    for b, g, n, s, idx, lv in CASES:
        need |= {r for r in idx if r % HALF}
        need |= {r for r in F.convolution_rotations(s, g) if r % HALF}
    return tuple(sorted(need))


@pytest.fixture(scope="module")
def env():
    P = O.params_small()
    assert P.N // 2 == HALF
    o, ctx, keys = make_env(P, seed=3, rotations=_rotations())
    yield P, o, ctx, keys
    ctx.close()


def oracle_conv(o, ct, pts, keys, b, g, s, idx):
    """Baby steps, inner sums, rotations within each block of 8 giant steps and of
    the block sums, in the order of LinearTransform.cu:287-431."""
    n = len(pts)
    nz = sorted({r for r in idx if r % HALF})
    hs = dict(zip(nz, o.rotate_hoisted(ct, nz, [keys[r] for r in nz]))) if nz else {}
    baby = [hs[r] if r % HALF else ct for r in idx]

    def add(x, y):
        return (o.add(x[0], y[0]), o.add(x[1], y[1]))

    def rot(x, r):
        return o.rotate(x, r, keys[r]) if r % HALF else x

    gl = -(-n // b)
    u = []
    for j in range(gl):
        acc = None
        for i in range(b):
            if b * j + i < n:
                t = o.mult_plain(baby[i], pts[b * j + i])
                acc = t if acc is None else add(acc, t)
        u.append(acc)
    nb = -(-g // 8)
    total = None
    for k in range(nb):
        G = min(8, g - 8 * k)
        blk = None
        for jj in range(G):
            j = 8 * k + jj
            if j >= gl:
                break
            v = rot(u[j], s * (G - jj))
            blk = v if blk is None else add(blk, v)
        if blk is None:
            continue
        v = rot(blk, 8 * s * (nb - 1 - k))
        total = v if total is None else add(total, v)
    return total


def closed_form(z, pz, b, g, s, idx):
    n = len(pz)
    nb = -(-g // 8)
    want = np.zeros_like(z)
    for k in range(nb):
        G = min(8, g - 8 * k)
        blk = np.zeros_like(z)
        for jj in range(G):
            j = 8 * k + jj
            inner = sum((pz[b * j + i] * np.roll(z, -idx[i]) for i in range(b) if b * j + i < n), np.zeros_like(z))
            blk += np.roll(inner, -s * (G - jj))
        want += np.roll(blk, -8 * s * (nb - 1 - k))
    return want


@pytest.mark.parametrize("case", range(len(CASES)))
def test_convolution_transform(env, case):
    P, o, ctx, keys = env
    b, g, n, s, idx, level = CASES[case]
    log(f"convolution transform bstep {b} gstep {g} n {n} stride {s} idx {idx} level {level}")
    rng = np.random.default_rng(case)
    scale = 2.0 ** P.scale_bits
    z = rng.uniform(-1, 1, HALF)
    ct = o.encrypt_slots(z, level, 500 + case)
    pz = [rng.uniform(-1, 1, HALF) for _ in range(n)]
    pl = [o.encode(d, P.L) for d in pz]
    gct = ctx.load(ct[0], ct[1], noise=scale)
    pts = [ctx.plaintext(d, noise=scale) for d in pl]
    out = gct.convolution_transform(pts, b, list(idx), g, stride=s)
    ref = oracle_conv(o, ct, pl, keys, b, g, s, idx)
    c0, c1 = out.store()
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])}"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])}"
    info = out.info()
    assert info["noise_level"] == 2 and info["level"] == level
    dec = o.decrypt_slots(c0, c1, info["noise"])
    err = np.abs(dec - closed_form(z, pz, b, g, s, idx)).max()
    assert err < 1e-4, err
    gct.convolution_transform(pts, b, list(idx), g, stride=s, out=gct)  # in place
    got = gct.store()
    assert np.array_equal(got[0], c0) and np.array_equal(got[1], c1)


def test_convolution_transform_degree_two_input(env):
    """A noise-degree-2 input is rescaled first (LinearTransform.cu:267-268)."""
    P, o, ctx, keys = env
    log("convolution transform of a degree-2 input")
    rng = np.random.default_rng(31)
    scale = 2.0 ** P.scale_bits
    z = rng.uniform(-1, 1, HALF)
    w = rng.uniform(-1, 1, HALF)
    a = ctx.load(*o.encrypt_slots(z, 3, 601), noise=scale)
    m = ctx.plaintext(o.encode(w, P.L), noise=scale)
    prod = a.mult_plain(m)
    assert prod.info()["noise_level"] == 2
    b, g, s, idx = 3, 3, 2, (0, 5, 11)
    pz = [rng.uniform(-1, 1, HALF) for _ in range(9)]
    pts = [ctx.plaintext(o.encode(d, P.L), noise=scale) for d in pz]
    out = prod.convolution_transform(pts, b, list(idx), g, stride=s)
    resc = prod.rescale()
    want = resc.convolution_transform(pts, b, list(idx), g, stride=s).store()
    got = out.store()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    info = out.info()
    dec = o.decrypt_slots(*got, info["noise"])
    assert np.abs(dec - closed_form(z * w, pz, b, g, s, idx)).max() < 1e-4


def test_convolution_transform_refusals(env):
    import pyfideslib_amd as F

    P, o, ctx, keys = env
    scale = 2.0 ** P.scale_bits
    a = ctx.load(*o.encrypt_slots(np.zeros(HALF), 2, 602), noise=scale)
    pts = [ctx.plaintext(o.encode(np.zeros(HALF), P.L), noise=scale) for _ in range(5)]
    with pytest.raises(F.FidesError):
        a.convolution_transform(pts, 2, [0, 1], 2)  # 5 plaintexts > bstep * gstep
    with pytest.raises(F.FidesError):
        a.convolution_transform(pts[:2], 1, [3], 2, stride=1000)  # no key for 1000 or 2000
