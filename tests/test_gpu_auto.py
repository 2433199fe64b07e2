"""Automatic rescaling on the GPU (needs a GPU): EvalAdd / EvalSub / EvalMult
of ciphertexts at different levels and noise degrees under FIXEDAUTO and
FLEXIBLEAUTO, through the C ABI, bit-exact against the oracle's restatement
of FIDESlib's adjustForAddOrSub / adjustForMult (src/CKKS/Ciphertext.cpp:
1472-1611, oracle.AutoScaling), with the same level, noise degree and
scaling factor, and decrypting to the combination of the messages."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import auto_cases as A
from tests.gpu_util import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import pyfideslib_amd as F

    P = A.params()
    o = O.Oracle(P, seed=0)
    rk = o.relin_key()
    ctxs = {}
    for tech in (O.FIXEDAUTO, O.FLEXIBLEAUTO):
        sf, sfb = A.scaling(P, tech)
        raw = F.RawParams.from_lists(P.logN, P.L, P.K, P.dnum, P.moduli, P.roots, tech, sf, sfb)
        ctx = F.Context(raw)
        ctx.load_eval_key(rk)
        ctxs[tech] = ctx
    yield P, o, rk, ctxs
    for c in ctxs.values():
        c.close()


@pytest.mark.parametrize("case", A.CASES, ids=[c[0] for c in A.CASES])
def test_auto_alignment_bitexact(env, case):
    P, o, rk, ctxs = env
    name, tech, op, (la, da), (lb, db) = case
    log(f"auto scaling case {name}")
    ctx = ctxs[tech]
    sf, sfb = A.scaling(P, tech)
    a, za = A.operand(o, sf, la, da, 11)
    b, zb = A.operand(o, sf, lb, db, 29)
    auto = O.AutoScaling(o, tech, sf, sfb)
    ga = ctx.load(a.c0, a.c1, noise=a.noise, noise_level=a.noise_level)
    gb = ctx.load(b.c0, b.c1, noise=b.noise, noise_level=b.noise_level)
    if op == "mult":
        want, got = auto.mult(a, b, rk), ga.mult(gb)
    elif op == "add":
        want, got = auto.add(a, b), ga.add(gb)
    else:
        want, got = auto.add(a, b, sub=True), ga.sub(gb)
    info = got.info()
    assert info["level"] == want.level and info["noise_level"] == want.noise_level
    assert info["noise"] == pytest.approx(want.noise, rel=1e-12)
    c0, c1 = got.store()
    assert np.array_equal(c0, want.c0) and np.array_equal(c1, want.c1), name
    dec = o.decrypt_slots(c0, c1, info["noise"])
    assert float(np.abs(dec - A.expected(op, za, zb)).max()) < 1e-4
    # the operands keep their residues and metadata
    assert ga.info()["level"] == la and ga.info()["noise_level"] == da
    assert np.array_equal(gb.store()[0], b.c0)
