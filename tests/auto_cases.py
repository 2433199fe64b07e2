"""Operand sets of the automatic-rescaling tests (FIXEDAUTO / FLEXIBLEAUTO),
shared by the oracle's CPU check (tests/test_oracle_auto.py) and the GPU
parity test (tests/test_gpu_auto.py).  Each case names a branch of FIDESlib's
adjustForAddOrSub / adjustForMult (src/CKKS/Ciphertext.cpp:1472-1611)."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

# (name, technique, op, (level, noise degree) of a, (level, noise degree) of b)
CASES = [
    ("fixed_drop", O.FIXEDAUTO, "add", (3, 1), (2, 1)),
    ("fixed_rescale_a", O.FIXEDAUTO, "add", (3, 2), (2, 1)),
    ("fixed_raise_a", O.FIXEDAUTO, "add", (3, 1), (3, 2)),
    ("fixed_sub_raise_b", O.FIXEDAUTO, "sub", (3, 2), (3, 1)),
    ("fixed_mult_rescale_a", O.FIXEDAUTO, "mult", (3, 2), (2, 1)),
    ("fixed_mult_rescale_b", O.FIXEDAUTO, "mult", (2, 1), (3, 2)),
    ("flex_d11_two_levels", O.FLEXIBLEAUTO, "add", (3, 1), (1, 1)),
    ("flex_d21_one_level", O.FLEXIBLEAUTO, "add", (3, 2), (2, 1)),
    ("flex_d21_two_levels", O.FLEXIBLEAUTO, "add", (3, 2), (1, 1)),
    ("flex_d22", O.FLEXIBLEAUTO, "add", (3, 2), (2, 2)),
    ("flex_d12", O.FLEXIBLEAUTO, "add", (3, 1), (2, 2)),
    ("flex_b_higher", O.FLEXIBLEAUTO, "sub", (1, 1), (3, 1)),
    ("flex_same_level_raise", O.FLEXIBLEAUTO, "add", (2, 1), (2, 2)),
    ("flex_mult_levels", O.FLEXIBLEAUTO, "mult", (3, 1), (1, 1)),
    ("flex_mult_depth", O.FLEXIBLEAUTO, "mult", (3, 2), (3, 1)),
]


def params():
    return O.params_small()


def scaling(P, technique):
    """(ScalingFactorReal, ScalingFactorRealBig) by level: one factor 2^40 under
    FIXEDAUTO, the FLEXIBLEAUTO chain otherwise."""
    if technique == O.FIXEDAUTO:
        sf = [2.0 ** P.scale_bits] * (P.L + 1)
        return sf, [x * x for x in sf]
    return O.flexible_scaling_factors(P.q)


def operand(o, sf, level, degree, seed):
    """A MetaCt at `level` and noise degree 1 or 2 (a product of two fresh
    ciphertexts, not rescaled), with its message."""
    rng = np.random.default_rng(seed)
    if degree == 1:
        z = rng.uniform(-1, 1, o.N // 2)
        c = o.encrypt_slots(z, level, seed, sf[level])
        return O.MetaCt(c[0], c[1], sf[level], 1, o.N // 2), z
    z1, z2 = rng.uniform(-1, 1, o.N // 2), rng.uniform(-1, 1, o.N // 2)
    c1 = o.encrypt_slots(z1, level, seed, sf[level])
    c2 = o.encrypt_slots(z2, level, seed + 1, sf[level])
    c = o.mult(c1, c2, o.relin_key())
    return O.MetaCt(c[0], c[1], sf[level] * sf[level], 2, o.N // 2), z1 * z2


def expected(op, za, zb):
    return za + zb if op == "add" else za - zb if op == "sub" else za * zb
