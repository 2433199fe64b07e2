"""Serialization of device objects on the GPU: a context's parameters, keys,
ciphertexts (both forms) and plaintexts saved, loaded into a context built
from the parameter file, and evaluated there bit-exactly against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import log, make_env

pytestmark = pytest.mark.gpu


def test_save_and_reload_everything(tmp_path):
    import pyfideslib_amd as F

    P = O.params_small()
    o, ctx, keys = make_env(P, seed=12, rotations=(1, 3))
    ctx2 = None
    try:
        log("serialization round trip")
        sw = o.ksk_gen(o.s, o.s, 77)
        ctx.load_switch_key(5, sw)
        ct = o.encrypt_slots(np.linspace(-1, 1, P.N // 2), 2, 7)
        ct2 = o.encrypt_slots(np.linspace(1, -1, P.N // 2), 2, 8)
        g = ctx.load(*ct, noise=2.0 ** P.scale_bits)
        x = g.extend()
        pt_q = ctx.plaintext(o.random_poly(P.L + 1, 9))
        E = O.ExtBasis(o)
        pq = E.ids(P.L)
        pt_x = ctx.plaintext_ext(o.random_poly(len(pq), 10, primeids=pq))
        d = tmp_path
        ctx.save_params(str(d / "params.bin"))
        ctx.save_key(F.KEY_KIND_EVAL, 0, str(d / "relin.key"))
        k1 = O.rot_index(P.logN, 1)
        ctx.save_key(F.KEY_KIND_AUTOMORPH, k1, str(d / "rot1.key"))
        ctx.save_key(F.KEY_KIND_SWITCH, 5, str(d / "sw5.key"))
        g.save(str(d / "ct.bin"))
        x.save(str(d / "ctx.bin"))
        pt_q.save(str(d / "ptq.bin"))
        pt_x.save(str(d / "ptx.bin"))
        with pytest.raises(F.FidesError):
            ctx.save_key(F.KEY_KIND_AUTOMORPH, O.rot_index(P.logN, 2), str(d / "none.key"))  # not loaded

        ctx2 = F.Context.from_file(str(d / "params.bin"))
        assert ctx2.raw == ctx.raw
        assert ctx2.load_key_file(str(d / "relin.key")) == (F.KEY_KIND_EVAL, 0)
        assert ctx2.load_key_file(str(d / "rot1.key")) == (F.KEY_KIND_AUTOMORPH, k1)
        assert ctx2.load_key_file(str(d / "sw5.key")) == (F.KEY_KIND_SWITCH, 5)
        h = ctx2.load_ciphertext_file(str(d / "ct.bin"))
        assert h.info() == g.info()
        c0, c1 = h.store()
        assert np.array_equal(c0, ct[0]) and np.array_equal(c1, ct[1])
        hx = ctx2.load_ciphertext_file(str(d / "ctx.bin"))
        assert hx.is_ext
        ex = hx.store_ext()
        assert np.array_equal(ex[0], E.extend(ct[0])) and np.array_equal(ex[1], E.extend(ct[1]))
        # evaluation in the reloaded context with the reloaded keys
        r = h.rotate(1).store()
        want = o.rotate(ct, 1, keys[1])
        assert np.array_equal(r[0], want[0]) and np.array_equal(r[1], want[1])
        h2 = ctx2.load(*ct2, noise=2.0 ** P.scale_bits)
        m = h.mult(h2, rescale=True).store()
        wm = o.rescale(o.mult(ct, ct2, keys["relin"]))
        assert np.array_equal(m[0], wm[0]) and np.array_equal(m[1], wm[1])
        s = h.key_switch(5).store()
        ws = o.keyswitch(ct[1], sw)
        assert np.array_equal(s[0], o.add(ct[0], ws[0])) and np.array_equal(s[1], ws[1])
        mq = h.mult_plain(ctx2.load_plaintext_file(str(d / "ptq.bin"))).store()
        wq = o.mult_plain(ct, o.random_poly(P.L + 1, 9))
        assert np.array_equal(mq[0], wq[0]) and np.array_equal(mq[1], wq[1])
        px = ctx2.load_plaintext_file(str(d / "ptx.bin"))
        dx = ctx2.dot_product_plain([hx], [px]).store_ext()
        pl = o.random_poly(len(pq), 10, primeids=pq)
        pl = np.ascontiguousarray(np.concatenate([pl[:3], pl[P.L + 1:]]))
        assert np.array_equal(dx[0], E.mul(E.extend(ct[0]), pl))
        # a file of other parameters does not load (other Q primes)
        P2 = O.make_params(13, 58, 42, 3, 60, 2)
        o3, ctx3, _ = make_env(P2, seed=1, relin=False)
        try:
            with pytest.raises(F.FidesError):
                ctx3.load_ciphertext_file(str(d / "ct.bin"))
            with pytest.raises(F.FidesError):
                ctx3.load_key_file(str(d / "rot1.key"))
        finally:
            ctx3.close()
    finally:
        if ctx2 is not None:
            ctx2.close()
        ctx.close()
