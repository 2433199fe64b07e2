"""Bit-exact parity of the HIP engine against the CPU oracle (needs a GPU).

Every residue of every limb must be identical: NTT/INTT, ModUp, ModDown, key
switch, tensor+relinearize, rescale, rotation, hoisted rotation sets and the
batched form.  Decrypted results are additionally checked against FP64 slot
arithmetic.  All calls go through the C ABI (pyfideslib_amd is its ctypes
binding).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import empty_dev, log, make_env, sync, to_dev, to_host

pytestmark = pytest.mark.gpu

ROT_SMALL = (1, 2, 5, -1, 1000)


@pytest.fixture(scope="module")
def small():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=0, rotations=ROT_SMALL)
    yield P, o, ctx, keys
    ctx.close()


def enc(o, P, level, seed, scale=None):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, P.N // 2)
    return z, o.encrypt_slots(z, level, seed, scale)


def load(ctx, ct, P, noise=None):
    noise = float(2 ** P.scale_bits) if noise is None else noise
    return ctx.load(ct[0], ct[1], noise=noise, noise_level=1)


def assert_ct_equal(gpu_ct, ref):
    c0, c1 = gpu_ct.store()
    assert c0.shape == ref[0].shape, (c0.shape, ref[0].shape)
    assert np.array_equal(c0, ref[0]), f"c0 differs in {np.count_nonzero(c0 != ref[0])} residues"
    assert np.array_equal(c1, ref[1]), f"c1 differs in {np.count_nonzero(c1 != ref[1])} residues"


# ------------------------------------------------------------------ NTT ----

def test_ntt_all_primes_small(small):
    P, o, ctx, _ = small
    log("ntt small")
    n = P.L + 1 + P.K
    pids = list(range(n))
    x = np.zeros((n, P.N), np.uint64)
    O.lib().orc_sample_uniform(o.ctx, 3, n, np.array(pids, np.int32), x)
    d = to_dev(ctx, x)
    sync(ctx)
    ctx.ntt(d, n, 0, inverse=False)
    sync(ctx)
    y = to_host(d)
    assert np.array_equal(y, o.ntt(x, pids, False))
    ctx.ntt(d, n, 0, inverse=True)
    sync(ctx)
    assert np.array_equal(to_host(d), x)


def test_ntt_limb_subranges(small):
    P, o, ctx, _ = small
    for first, n in ((0, 1), (2, 3), (P.L + 1, P.K), (P.L, 2)):
        pids = list(range(first, first + n))
        x = np.zeros((n, P.N), np.uint64)
        O.lib().orc_sample_uniform(o.ctx, 10 + first, n, np.array(pids, np.int32), x)
        d = to_dev(ctx, x)
        sync(ctx)
        ctx.ntt(d, n, first, inverse=True)
        sync(ctx)
        assert np.array_equal(to_host(d), o.ntt(x, pids, True))


# --------------------------------------------------------------- stages ----

@pytest.mark.parametrize("level", [3, 2, 1, 0])
def test_modup_moddown_keyswitch_stages(small, level):
    P, o, ctx, keys = small
    log(f"stages level {level}")
    nl = level + 1
    x = o.random_poly(nl, seed=40 + level)
    ext_ref = o.modup(x)
    dx = to_dev(ctx, x)
    dext = empty_dev(ctx, ext_ref.shape)
    sync(ctx)
    ctx.stage_modup(level, dx, dext)
    sync(ctx)
    assert np.array_equal(to_host(dext), ext_ref)
    # moddown of a random extended poly
    pids = list(range(nl)) + [P.L + 1 + k for k in range(P.K)]
    xe = np.zeros((nl + P.K, P.N), np.uint64)
    O.lib().orc_sample_uniform(o.ctx, 77 + level, nl + P.K, np.array(pids, np.int32), xe)
    dxe, dy = to_dev(ctx, xe), empty_dev(ctx, (nl, P.N))
    sync(ctx)
    ctx.stage_moddown(level, dxe, dy)
    sync(ctx)
    assert np.array_equal(to_host(dy), o.moddown(xe))
    # full key switch with the relinearization key
    kb, ka = o.keyswitch(x, keys["relin"])
    dkb, dka = empty_dev(ctx, (nl, P.N)), empty_dev(ctx, (nl, P.N))
    sync(ctx)
    ctx.stage_keyswitch(level, dx, dkb, dka)
    sync(ctx)
    assert np.array_equal(to_host(dkb), kb)
    assert np.array_equal(to_host(dka), ka)


# ------------------------------------------------------------------ ops ----

@pytest.mark.parametrize("level", [3, 2, 1])
def test_mult_relin_all_levels(small, level):
    P, o, ctx, keys = small
    log(f"mult level {level}")
    z1, c1 = enc(o, P, level, 100 + level)
    z2, c2 = enc(o, P, level, 200 + level)
    g1, g2 = load(ctx, c1, P), load(ctx, c2, P)
    ref = o.mult(c1, c2, keys["relin"])
    out = g1.mult(g2)
    assert_ct_equal(out, ref)
    info = out.info()
    assert info["noise_level"] == 2 and info["level"] == level
    # explicit rescale, and the fused mult+rescale form
    out.rescale()
    ref_r = o.rescale(ref)
    assert_ct_equal(out, ref_r)
    assert_ct_equal(g1.mult(g2, rescale=True), ref_r)
    scale = 2.0 ** P.scale_bits
    dec = o.decrypt_slots(*out.store(), scale * scale / P.moduli[level])
    assert np.abs(dec - z1 * z2).max() < 1e-6


def test_square_add_sub_inplace(small):
    P, o, ctx, keys = small
    z1, c1 = enc(o, P, P.L, 301)
    z2, c2 = enc(o, P, P.L, 302)
    g1, g2 = load(ctx, c1, P), load(ctx, c2, P)
    assert_ct_equal(g1.square(), o.square(c1, keys["relin"]))
    assert_ct_equal(g1.add(g2), (o.add(c1[0], c2[0]), o.add(c1[1], c2[1])))
    s0 = np.zeros_like(c1[0]), np.zeros_like(c1[1])
    O.lib().orc_sub(o.ctx, P.L + 1, O.ids(P.L + 1), c1[0], c2[0], s0[0])
    O.lib().orc_sub(o.ctx, P.L + 1, O.ids(P.L + 1), c1[1], c2[1], s0[1])
    assert_ct_equal(g1.sub(g2), s0)
    # in place: out aliases the first input
    ref = o.mult(c1, c2, keys["relin"])
    g1.mult(g2, out=g1)
    assert_ct_equal(g1, ref)


def test_mult_mixed_levels_reduces_to_lower(small):
    P, o, ctx, keys = small
    _, c1 = enc(o, P, 3, 401)
    _, c2 = enc(o, P, 2, 402)
    g = load(ctx, c1, P).mult(load(ctx, c2, P))
    ref = o.mult((c1[0][:3], c1[1][:3]), c2, keys["relin"])
    assert_ct_equal(g, ref)


@pytest.mark.parametrize("r", ROT_SMALL)
def test_rotate(small, r):
    P, o, ctx, keys = small
    log(f"rotate {r}")
    for level in (P.L, 1):
        z, c = enc(o, P, level, 500 + level)
        out = load(ctx, c, P).rotate(r)
        ref = o.rotate(c, r, keys[r])
        assert_ct_equal(out, ref)
    dec = o.decrypt_slots(*out.store(), 2.0 ** P.scale_bits)
    assert np.abs(dec - np.roll(z, -r)).max() < 1e-5


def test_rotate_in_place(small):
    """out aliasing the input: the permuting epilogue goes through a temporary."""
    P, o, ctx, keys = small
    log("rotate in place")
    for r in (1, 5):
        z, c = enc(o, P, P.L, 700 + r)
        g = load(ctx, c, P)
        assert g.rotate(r, out=g) is g
        assert_ct_equal(g, o.rotate(c, r, keys[r]))


def test_rotate_hoisted(small):
    P, o, ctx, keys = small
    z, c = enc(o, P, P.L, 600)
    g = load(ctx, c, P)
    rs = list(ROT_SMALL)
    outs = g.rotate_hoisted(rs)
    refs = o.rotate_hoisted(c, rs, [keys[r] for r in rs])
    for r, out, ref in zip(rs, outs, refs):
        assert_ct_equal(out, ref)


def test_batched_mult_uneven_levels(small):
    P, o, ctx, keys = small
    log("batch small")
    levels = [3, 1, 2, 3, 1, 2, 1, 3]
    A, B, refs = [], [], []
    for i, lv in enumerate(levels):
        _, c1 = enc(o, P, lv, 700 + i)
        _, c2 = enc(o, P, lv, 800 + i)
        A.append(load(ctx, c1, P))
        B.append(load(ctx, c2, P))
        refs.append(o.rescale(o.mult(c1, c2, keys["relin"])))
    outs = [ctx.ciphertext() for _ in levels]
    for _ in range(3):  # later calls replay the cached plan graph
        ctx.mult_batch(A, B, outs, rescale=True)
        for out, ref, lv in zip(outs, refs, levels):
            assert out.level == lv - 1
            assert_ct_equal(out, ref)


def test_batched_rotations_uneven_levels(small):
    """Many ciphertexts, each rotated by its own index (identity included)."""
    P, o, ctx, keys = small
    log("batch rotations")
    levels = [3, 1, 2, 3, 0, 2]
    rs = [1, 2, 5, -1, 1, P.N // 2]  # N/2 slots: the last one is the identity
    A, refs = [], []
    for i, (lv, r) in enumerate(zip(levels, rs)):
        _, c = enc(o, P, lv, 1100 + i)
        A.append(load(ctx, c, P))
        refs.append(c if r % (P.N // 2) == 0 else o.rotate(c, r, keys[r]))
    outs = [ctx.ciphertext() for _ in levels]
    for _ in range(3):  # eager, then captured, then replayed
        ctx.rotate_batch(A, rs, outs)
        for out, ref, lv in zip(outs, refs, levels):
            assert out.level == lv
            assert_ct_equal(out, ref)
    import pyfideslib_amd as F
    with pytest.raises(F.FidesError):
        ctx.rotate_batch([A[0]], [1], [A[0]])  # output aliasing its input


def test_batched_rotations_shared_keys(small):
    """Rotations whose keys repeat (3 keys over 9 ciphertexts at uneven levels)
    take the key-stationary KSK, the jobs grouped by key: each output equals
    the oracle's single rotation."""
    P, o, ctx, keys = small
    log("batch rotations, shared keys")
    levels = [3, 1, 2, 3, 0, 2, 1, 3, 2]
    rs = [1, 2, 1, 5, 2, 1, 5, 1, 2]
    A, refs = [], []
    for i, (lv, r) in enumerate(zip(levels, rs)):
        _, c = enc(o, P, lv, 1200 + i)
        A.append(load(ctx, c, P))
        refs.append(o.rotate(c, r, keys[r]))
    outs = [ctx.ciphertext() for _ in levels]
    for _ in range(3):  # eager, then captured, then replayed
        ctx.rotate_batch(A, rs, outs)
        for out, ref, lv in zip(outs, refs, levels):
            assert out.level == lv
            assert_ct_equal(out, ref)


def test_errors(small):
    import pyfideslib_amd as F

    P, o, ctx, _ = small
    _, c = enc(o, P, 0, 900)
    g = load(ctx, c, P)
    with pytest.raises(F.FidesError) as e:
        g.rescale()
    assert e.value.code == 5
    with pytest.raises(F.FidesError) as e:
        g.rotate(3)  # no key for 3
    assert e.value.code == 4
    with pytest.raises(F.FidesError) as e:
        ctx.ntt(0, 1, P.L + P.K, False)  # range beyond the prime list
    assert e.value.code == 1


# ------------------------------------------------------ production size ----

@pytest.fixture(scope="module")
def prod():
    P = O.params_production()
    o, ctx, keys = make_env(P, seed=0, rotations=(1, 2, 16))
    yield P, o, ctx, keys
    ctx.close()


def test_prod_ntt(prod):
    P, o, ctx, _ = prod
    log("prod ntt")
    n = P.L + 1 + P.K
    pids = list(range(n))
    x = np.zeros((n, P.N), np.uint64)
    O.lib().orc_sample_uniform(o.ctx, 3, n, np.array(pids, np.int32), x)
    d = to_dev(ctx, x)
    sync(ctx)
    ctx.ntt(d, n, 0, inverse=False)
    sync(ctx)
    assert np.array_equal(to_host(d), o.ntt(x, pids, False))
    ctx.ntt(d, n, 0, inverse=True)
    sync(ctx)
    assert np.array_equal(to_host(d), x)


def test_prod_mult_rescale(prod):
    P, o, ctx, keys = prod
    log("prod mult")
    for level in (P.L, 13, 4):
        z1, c1 = enc(o, P, level, 10 + level)
        z2, c2 = enc(o, P, level, 20 + level)
        ref = o.rescale(o.mult(c1, c2, keys["relin"]))
        out = load(ctx, c1, P).mult(load(ctx, c2, P), rescale=True)
        assert_ct_equal(out, ref)
    scale = 2.0 ** P.scale_bits
    dec = o.decrypt_slots(*out.store(), scale * scale / P.moduli[level])
    assert np.abs(dec - z1 * z2).max() < 1e-9


def test_prod_rotate_and_hoisted(prod):
    P, o, ctx, keys = prod
    log("prod rotate")
    z, c = enc(o, P, P.L, 31)
    g = load(ctx, c, P)
    assert_ct_equal(g.rotate(1), o.rotate(c, 1, keys[1]))
    rs = [1, 2, 16]
    outs = g.rotate_hoisted(rs)
    refs = o.rotate_hoisted(c, rs, [keys[r] for r in rs])
    for out, ref in zip(outs, refs):
        assert_ct_equal(out, ref)
    dec = o.decrypt_slots(*outs[2].store(), 2.0 ** P.scale_bits)
    assert np.abs(dec - np.roll(z, -16)).max() < 1e-6


def test_prod_batched_vs_plain(prod):
    P, o, ctx, keys = prod
    log("prod batch")
    levels = [24, 4, 17, 9, 24, 11]
    A, B, plain = [], [], []
    for i, lv in enumerate(levels):
        _, c1 = enc(o, P, lv, 40 + i)
        _, c2 = enc(o, P, lv, 60 + i)
        A.append(load(ctx, c1, P))
        B.append(load(ctx, c2, P))
        plain.append(A[-1].mult(B[-1], rescale=True).store())
    outs = [ctx.ciphertext() for _ in levels]
    ctx.mult_batch(A, B, outs, rescale=True)
    for out, ref in zip(outs, plain):
        assert_ct_equal(out, ref)
    # one pair against the oracle as well
    _, c1 = enc(o, P, levels[1], 40 + 1)
    _, c2 = enc(o, P, levels[1], 60 + 1)
    assert_ct_equal(outs[1], o.rescale(o.mult(c1, c2, keys["relin"])))


def test_graph_replays_match_eager(small):
    """Captured op graphs replayed several times, with eager work and
    synchronizations between replays, give the oracle's residues every time."""
    P, o, ctx, keys = small
    log("graph replays")
    ctx.set_op_graphs(True)
    _, c1 = enc(o, P, P.L, 1001)
    _, c2 = enc(o, P, P.L, 1002)
    g1, g2 = load(ctx, c1, P), load(ctx, c2, P)
    out, rot = ctx.ciphertext(), ctx.ciphertext()
    ref_m = o.rescale(o.mult(c1, c2, keys["relin"]))
    ref_r = o.rotate(c1, 1, keys[1])
    outs = [ctx.ciphertext() for _ in (1, 2)]
    ref_h = o.rotate_hoisted(c1, [1, 2], [keys[1], keys[2]])
    # call 1 runs eagerly, call 2 captures, calls 3.. replay
    for rep in range(5):
        g1.mult(g2, rescale=True, out=out)
        assert_ct_equal(out, ref_m)
        junk = g1.add(g2)  # eager work between replays
        ctx.synchronize()
        g1.rotate(1, out=rot)
        assert_ct_equal(rot, ref_r)
        g1.rotate_hoisted([1, 2], outs)
        for x, r in zip(outs, ref_h):
            assert_ct_equal(x, r)
        assert_ct_equal(junk, (o.add(c1[0], c2[0]), o.add(c1[1], c2[1])))
    ctx.set_op_graphs(False)


@pytest.mark.parametrize("logN", [14, 11, 17, 10])
def test_odd_row_split_mult_rotate(logN):
    """logN where the row pass's last phase is shorter than the others (K2 odd):
    the ModDown / rescale epilogues and the automorphism store at that split."""
    P = O.make_params(logN, 60, 40, 3, 60, 2)
    o, ctx, keys = make_env(P, seed=5, rotations=(1, 7))
    log(f"odd row split logN={logN}")
    try:
        _, c1 = enc(o, P, P.L, 1201)
        _, c2 = enc(o, P, P.L, 1202)
        g1, g2 = load(ctx, c1, P), load(ctx, c2, P)
        assert_ct_equal(g1.mult(g2, rescale=True), o.rescale(o.mult(c1, c2, keys["relin"])))
        assert_ct_equal(g1.mult(g2), o.mult(c1, c2, keys["relin"]))
        for r in (1, 7):
            assert_ct_equal(g1.rotate(r), o.rotate(c1, r, keys[r]))
        outs = g1.rotate_hoisted([1, 7])
        for x, ref in zip(outs, o.rotate_hoisted(c1, [1, 7], [keys[1], keys[7]])):
            assert_ct_equal(x, ref)
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def wide():
    """All-60-bit moduli with one digit of 10 primes and 11 special primes:
    the conversion sums exceed one Barrett reduction, so the engine takes its
    chunked-accumulation conversion kernels."""
    P = O.make_params(12, 60, 60, 9, 60, 1)
    assert P.K == 11
    o, ctx, keys = make_env(P, seed=3, rotations=(3,))
    yield P, o, ctx, keys
    ctx.close()


@pytest.mark.parametrize("level", [9, 5, 1])
def test_wide_digit_chunked_conversion(wide, level):
    P, o, ctx, keys = wide
    log(f"wide digit level {level}")
    _, c1 = enc(o, P, level, 70 + level, 2.0 ** 40)
    _, c2 = enc(o, P, level, 80 + level, 2.0 ** 40)
    ref = o.rescale(o.mult(c1, c2, keys["relin"]))
    g1 = load(ctx, c1, P, 2.0 ** 40)
    assert_ct_equal(g1.mult(load(ctx, c2, P, 2.0 ** 40), rescale=True), ref)
    assert_ct_equal(g1.rotate(3), o.rotate(c1, 3, keys[3]))
