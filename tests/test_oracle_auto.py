"""The oracle's restatement of automatic rescaling (oracle.AutoScaling, after
FIDESlib's adjustForAddOrSub / adjustForMult, src/CKKS/Ciphertext.cpp:1472-1611)
pinned by what the rules must achieve: every aligned sum, difference or
product decrypts to the same combination of the messages at the scale its
metadata records, and the result sits where the reference leaves it (level,
noise degree)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import auto_cases as A


@pytest.fixture(scope="module")
def env():
    P = A.params()
    o = O.Oracle(P, seed=0)
    return P, o, o.relin_key()


@pytest.mark.parametrize("case", A.CASES, ids=[c[0] for c in A.CASES])
def test_auto_alignment_decrypts(env, case):
    P, o, rk = env
    name, tech, op, (la, da), (lb, db) = case
    sf, sfb = A.scaling(P, tech)
    a, za = A.operand(o, sf, la, da, 11)
    b, zb = A.operand(o, sf, lb, db, 29)
    auto = O.AutoScaling(o, tech, sf, sfb)
    if op == "mult":
        r = auto.mult(a, b, rk)
        assert r.noise_level == 2
    else:
        r = auto.add(a, b, sub=(op == "sub"))
    assert r.level <= min(la, lb)
    got = o.decrypt_slots(r.c0, r.c1, r.noise)
    err = float(np.abs(got - A.expected(op, za, zb)).max())
    assert err < 1e-4, (name, err, r.level, r.noise_level)
    # the operands are untouched (the rules act on copies)
    assert a.level == la and a.noise_level == da and b.level == lb and b.noise_level == db


def test_flexible_factors_chain():
    P = A.params()
    sf, sfb = O.flexible_scaling_factors(P.q)
    assert sf[P.L] == float(P.q[P.L])
    for lv in range(1, P.L + 1):  # a level-lv product rescaled by q_lv lands on level lv-1's factor
        assert abs(sf[lv] * sf[lv] / P.q[lv] - sf[lv - 1]) < 1e-6 * sf[lv - 1]
    assert all(abs(b - s * s) < 1e-9 * b for s, b in zip(sf, sfb))


def test_fixedmanual_add_keeps_operands_as_they_are():
    """Under FIXEDMANUAL nothing is adjusted: the sum is taken at the lower
    level with the first operand's metadata."""
    P = A.params()
    o = O.Oracle(P, seed=0)
    sf, sfb = A.scaling(P, O.FIXEDAUTO)
    a, za = A.operand(o, sf, 3, 1, 5)
    b, zb = A.operand(o, sf, 2, 1, 6)
    r = O.AutoScaling(o, O.FIXEDMANUAL, sf, sfb).add(a, b)
    assert r.level == 2 and r.noise == a.noise
    assert np.array_equal(r.c0, o.add(np.ascontiguousarray(a.c0[:3]), b.c0))
