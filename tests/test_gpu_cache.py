"""Launch-sequence caches across key replacement (needs a GPU).

Per-op graphs, batch plans and hoisted-rotation plans bake in device
addresses of the keys they were captured with.  After every key is replaced
and the pool's cached blocks are given back to the device, the same calls on
the same ciphertexts must give the oracle's residues under the NEW keys --
eagerly, captured and replayed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


def _eq(g, ref):
    a, b = g.store()
    return np.array_equal(a, ref[0]) and np.array_equal(b, ref[1])


def test_key_replacement_invalidates_cached_sequences():
    P = O.params_small()
    o, ctx, keys = make_env(P, seed=0, rotations=(1, 2))
    ctx.set_op_graphs(True)  # the per-op graphs are among the cached sequences
    try:
        rng = np.random.default_rng(5)
        c1 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 11)
        c2 = o.encrypt_slots(rng.uniform(-1, 1, P.N // 2), P.L, 12)
        g1 = ctx.load(*c1, noise=2.0 ** P.scale_bits)
        g2 = ctx.load(*c2, noise=2.0 ** P.scale_bits)
        out, rot = ctx.ciphertext(), ctx.ciphertext()
        hs = [ctx.ciphertext() for _ in range(2)]
        bm, br = [ctx.ciphertext()], [ctx.ciphertext(), ctx.ciphertext()]

        def run_all(K, tag):
            want_m = o.rescale(o.mult(c1, c2, K["relin"]))
            want_r = o.rotate(c1, 1, K[1])
            want_h = o.rotate_hoisted(c1, [1, 2], [K[1], K[2]])
            want_br = [o.rotate(c1, 1, K[1]), o.rotate(c2, 2, K[2])]
            for rep in range(4):  # eager, captured, replayed, replayed
                log(f"{tag} rep {rep}")
                g1.mult(g2, rescale=True, out=out)
                assert _eq(out, want_m), (tag, rep, "mult")
                g1.rotate(1, out=rot)
                assert _eq(rot, want_r), (tag, rep, "rotate")
                g1.rotate_hoisted([1, 2], hs)
                assert all(_eq(h, w) for h, w in zip(hs, want_h)), (tag, rep, "hoisted")
                ctx.mult_batch([g1], [g2], bm, rescale=True)
                assert _eq(bm[0], want_m), (tag, rep, "batched mult")
                ctx.rotate_batch([g1, g2], [1, 2], br)
                assert all(_eq(x, w) for x, w in zip(br, want_br)), (tag, rep, "batched rotate")

        run_all(keys, "first keys")
        # new keys of another secret, loaded over the old ones (same slots in
        # the context's key tables); then every cached block goes back
        o2 = O.Oracle(P, seed=7)
        keys2 = {"relin": o2.relin_key(), 1: o2.rotation_key(1), 2: o2.rotation_key(2)}
        ctx.load_eval_key(keys2["relin"])
        ctx.load_rotation_key(1, keys2[1])
        ctx.load_rotation_key(2, keys2[2])
        ctx.trim_pool(0)
        run_all(keys2, "replaced keys")
        # and once more, replacing only the B components' contents
        k3 = {"relin": (keys2["relin"][0], keys["relin"][1]), 1: (keys2[1][0], keys[1][1]),
              2: (keys2[2][0], keys[2][1])}
        ctx.load_eval_key(k3["relin"])
        ctx.load_rotation_key(1, k3[1])
        ctx.load_rotation_key(2, k3[2])
        run_all(k3, "mixed keys")
    finally:
        ctx.close()
