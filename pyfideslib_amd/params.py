"""RNS-CKKS parameter generation (host side).

In FIDESlib the moduli and roots come from OpenFHE (GetRawParams,
src/CKKS/openfhe-interface/RawCiphertext.cu:219-467).  This module produces an
equivalent RawParams without OpenFHE, for synthetic workloads:
  - primes: the largest primes below 2^bits with q == 1 (mod 2N), q0 first,
    then the scaling primes, then the special primes, all distinct;
  - roots: the minimal primitive 2N-th root of unity of each prime (the rule
    OpenFHE's RootOfUnity applies);
  - K: the reference's ContextData::computeK rule (src/CKKS/Context.cu:286-302):
    special primes counted as (bits - 1) until the largest digit is covered.
"""
from __future__ import annotations

from . import RawParams, FIXEDMANUAL

_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in _BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_primes(logN: int, bits: int, count: int, exclude=()) -> list[int]:
    m = 2 << logN
    c = (1 << bits) - m + 1
    out = []
    ex = set(exclude)
    while len(out) < count:
        if c <= m:
            raise ValueError("not enough NTT-friendly primes")
        if c not in ex and is_prime(c):
            out.append(c)
        c -= m
    return out


def min_primitive_root(q: int, m: int) -> int:
    """Smallest primitive m-th root of unity mod q (m a power of two)."""
    e = (q - 1) // m
    w = None
    for g in range(2, q):
        x = pow(g, e, q)
        if pow(x, m // 2, q) == q - 1:
            w = x
            break
    w2 = w * w % q
    best = cur = w
    for _ in range(m // 2 - 1):
        cur = cur * w2 % q
        if cur < best:
            best = cur
    return best


def special_count(q_moduli: list[int], dnum: int, special_bits: int) -> int:
    alpha = -(-len(q_moduli) // dnum)
    digits = [q_moduli[i:i + alpha] for i in range(0, len(q_moduli), alpha)]
    maxbits = max(sum(q.bit_length() for q in d) for d in digits)
    k, bits = 0, 0
    while bits < maxbits:
        bits += special_bits - 1
        k += 1
    return k


def ckks_params(logN: int, first_bits: int, scale_bits: int, L: int, special_bits: int, dnum: int,
                K: int | None = None, scaling_technique: int = FIXEDMANUAL) -> RawParams:
    q0 = ntt_primes(logN, first_bits, 1)
    qs = ntt_primes(logN, scale_bits, L, exclude=q0)
    if K is None:
        K = special_count(q0 + qs, dnum, special_bits)
    ps = ntt_primes(logN, special_bits, K, exclude=q0 + qs)
    m = 2 << logN
    q = q0 + qs
    # ScalingFactorReal: under FIXEDMANUAL / FIXEDAUTO OpenFHE keeps 2^scale_bits at every level
    sfr = [2.0 ** scale_bits] * (L + 1) if scaling_technique in (0, 1) else None
    return RawParams(logN, L, K, dnum, q, [min_primitive_root(x, m) for x in q], ps,
                     [min_primitive_root(x, m) for x in ps], scaling_technique, None, sfr)


def production() -> RawParams:
    """N=2^16, q0 60-bit, 24 x 50-bit scaling primes, dnum=3, 60-bit special primes (K=8)."""
    return ckks_params(16, 60, 50, 24, 60, 3)


def small() -> RawParams:
    """N=2^13, q0 60-bit, 3 x 40-bit scaling primes, dnum=2, 60-bit special primes (K=2)."""
    return ckks_params(13, 60, 40, 3, 60, 2)
