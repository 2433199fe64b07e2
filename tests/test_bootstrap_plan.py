"""FP64 model of the bootstrap setup (no GPU): the factored special FFT against
the dense encoding matrix, the encoder against the oracle's 2N-point FFT
encoding, the baby-step / giant-step layout of every CoeffsToSlots /
SlotsToCoeffs group (evaluated the way the engine's linear transform composes
it) against the dense matrices, the EvalMod coefficients through the
double-angle recursion against sin, and the whole pipeline's scale
bookkeeping on plain complex vectors."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from pyfideslib_amd import bootstrap as B


def dense_u0(N: int) -> np.ndarray:
    n = N // 2
    e = np.array([pow(5, j, 2 * N) for j in range(n)])
    return np.exp(1j * np.pi * np.outer(e, np.arange(n)) / N)


def bitrev_perm(n: int) -> np.ndarray:
    bits = n.bit_length() - 1
    return np.array([int(format(i, f"0{bits}b")[::-1], 2) for i in range(n)])


@pytest.mark.parametrize("logN", [2, 5, 9])
def test_special_fft_closed_form(logN):
    N = 1 << logN
    n = N // 2
    rng = np.random.default_rng(logN)
    w = rng.normal(size=n) + 1j * rng.normal(size=n)
    U = dense_u0(N)
    assert np.abs(B.special_fft(w) - U @ w).max() < 1e-10 * n
    # U0^-1 = U0^H / n (the slot points are the n roots zeta^e, e = 1 mod 4)
    z = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.abs(B.special_fft_inverse(z) - U.conj().T @ z / n).max() < 1e-12 * n
    assert np.abs(B.special_fft(B.special_fft_inverse(z)) - z).max() < 1e-12 * n


def test_encode_matches_oracle_encoding():
    logN = 10
    N = 1 << logN
    rng = np.random.default_rng(3)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    scale = 2.0 ** 40
    mine = B.encode_coefficients(z, N, scale)
    ref = O.encode_coeffs(z, logN, scale)
    assert np.abs(mine - ref).max() <= 1
    assert np.abs(O.decode_coeffs(mine.astype(np.float64), logN, scale) - z).max() < 1e-9


def apply_lt(z: np.ndarray, g: dict) -> np.ndarray:
    """out = rot_offset(sum_j rot_(j b s)(sum_i A[b j + i] * rot_(i s)(z))), rot_r(x)[p] = x[p + r]."""
    rot = lambda x, r: np.roll(x, -r)  # noqa: E731
    b, s, off, A = g["bstep"], g["stride"], g["offset"], g["diag"]
    acc = np.zeros_like(z)
    for j in range(-(-len(A) // b)):
        inner = sum(A[b * j + i] * rot(z, i * s) for i in range(b) if b * j + i < len(A))
        acc = acc + rot(inner, j * b * s)
    return rot(acc, off)


@pytest.mark.parametrize("logN, budget", [(8, (3, 3)), (8, (1, 2)), (9, (2, 4)), (12, (3, 3))])
def test_bsgs_groups_equal_dense_transforms(logN, budget):
    N = 1 << logN
    n = N // 2
    q0 = (1 << 60) - 93
    bs = B.LayoutModel(N, q0, 2.0 ** 50, level_budget=budget, degree=15)
    rng = np.random.default_rng(logN)
    z = rng.normal(size=n) + 1j * rng.normal(size=n)
    P = bitrev_perm(n)
    # CoeffsToSlots: alpha * P U0^-1, alpha = S0 / (2 K q0)
    v = z
    for g in bs.cts:
        v = apply_lt(v, g)
    alpha = 2.0 ** 50 / (2 * bs.K * q0)
    want = alpha * B.special_fft_inverse(z)[P]
    assert np.abs(v - want).max() < 1e-9 * np.abs(want).max()
    # SlotsToCoeffs: beta * U0 P^-1, beta = q0 / (S0 2^deg)
    u = z
    for g in bs.stc:
        u = apply_lt(u, g)
    beta = q0 / (2.0 ** 50 * bs.post)
    w = np.empty(n, dtype=np.complex128)
    w[P] = z
    want = beta * B.special_fft(w)
    assert np.abs(u - want).max() < 1e-9 * np.abs(want).max()
    # every rotation the layout needs is listed
    for g in bs.cts + bs.stc:
        need = {(i * g["stride"]) % n for i in range(1, g["bstep"])} | {g["offset"]}
        if len(g["diag"]) > g["bstep"]:
            need.add((g["bstep"] * g["stride"]) % n)
        assert need - {0} <= set(bs.rotations)
    assert len(bs.cts) == budget[0] and len(bs.stc) == budget[1]


@pytest.mark.parametrize("K, r, deg", [(28, 3, 119), (12, 2, 59)])
def test_evalmod_coefficients_through_double_angle(K, r, deg):
    c = B.evalmod_coefficients(K, r, deg)
    x = np.linspace(-1, 1, 4001)
    y = np.polynomial.chebyshev.chebval(x, c)
    for j in range(1, r + 1):
        y = 2 * y * y - (2 * np.pi) ** (-(2.0 ** (j - r)))
    assert np.abs(y - np.sin(2 * np.pi * K * x) / (2 * np.pi)).max() < 1e-9


def test_pipeline_scale_bookkeeping():
    """ModRaise output c = S0 m + q0 I pushed through the CtS groups, the
    real / imaginary split, FP64 EvalMod, the integer factor and the StC
    groups returns m's slots (no encryption: checks the scales only)."""
    logN = 9
    N = 1 << logN
    n = N // 2
    q0 = (1 << 60) - 93
    S0 = 2.0 ** 50
    bs = B.LayoutModel(N, q0, S0, level_budget=(3, 2))
    rng = np.random.default_rng(5)
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    m = B.encode_coefficients(z, N, S0)  # S0 m as integers
    I = rng.integers(-20, 21, N)
    c = m.astype(np.float64) + float(q0) * I
    slots = B.special_fft(c[:n] + 1j * c[n:]) / S0  # what the raised ciphertext decodes to
    v = slots
    for g in bs.cts:
        v = apply_lt(v, g)
    re = v + np.conj(v)
    im = (v - np.conj(v)) * -1j
    assert np.abs(re.real).max() <= 1.0 and np.abs(im.real).max() <= 1.0

    def evalmod(x):
        y = np.polynomial.chebyshev.chebval(x.real, bs.coeffs)
        for j in range(1, bs.r + 1):
            y = 2 * y * y - (2 * np.pi) ** (-(2.0 ** (j - bs.r)))
        return y * bs.post

    u = evalmod(re) + 1j * evalmod(im)
    for g in bs.stc:
        u = apply_lt(u, g)
    assert np.abs(u - z).max() < 1e-4


@pytest.mark.parametrize("logN, slots, budget", [(9, 32, (3, 3)), (10, 64, (2, 3)), (9, 2, (1, 1))])
def test_sparse_slots_pipeline(logN, slots, budget):
    """Sparse packing (slots < N/2, the message periodic with period `slots`):
    the subring trace (rotations by slots 2^t), CoeffsToSlots with the
    real / imaginary split into blocks of `slots`, ONE FP64 EvalMod on the real
    part, SlotsToCoeffs and the fold by `slots` return the message."""
    N = 1 << logN
    n = N // 2
    q0 = (1 << 60) - 93
    S0 = 2.0 ** 50
    bs = B.LayoutModel(N, q0, S0, level_budget=budget, slots=slots)
    gap = n // slots
    assert set(slots << t for t in range(gap.bit_length() - 1)) <= set(bs.rotations)
    rng = np.random.default_rng(slots)
    z = np.tile(rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots), gap)
    m = B.encode_coefficients(z, N, S0)
    assert np.all(np.nonzero(m)[0] % gap == 0)  # the sparse encoding lives in the subring
    c = m.astype(np.float64) + float(q0) * rng.integers(-20, 21, N)
    slots_raised = B.special_fft(c[:n] + 1j * c[n:]) / S0
    v = sum(np.roll(slots_raised, -slots * k) for k in range(gap))  # the trace
    for g in bs.cts:
        v = apply_lt(v, g)
    re = v + np.conj(v)
    assert np.abs(re.real).max() <= 1.0
    y = np.polynomial.chebyshev.chebval(re.real, bs.coeffs)
    for j in range(1, bs.r + 1):
        y = 2 * y * y - (2 * np.pi) ** (-(2.0 ** (j - bs.r)))
    u = (y * bs.post).astype(np.complex128)
    for g in bs.stc:
        u = apply_lt(u, g)
    u = u + np.roll(u, -slots)
    assert np.abs(u - z).max() < 1e-4


def test_batched_giant_rotations():
    """With batched giant steps every giant step j of every group needs the key
    of j b stride; the Horner set is a subset."""
    N, q0 = 1 << 10, (1 << 60) - 93
    horner = B.LayoutModel(N, q0, 2.0 ** 50, level_budget=(3, 3))
    allg = B.LayoutModel(N, q0, 2.0 ** 50, level_budget=(3, 3), batched_giant=True)
    assert set(horner.rotations) <= set(allg.rotations)
    n = N // 2
    for g in allg.cts + allg.stc:
        ng = -(-len(g["diag"]) // g["bstep"])
        need = {(j * g["bstep"] * g["stride"]) % n for j in range(1, ng)} - {0}
        assert need <= set(allg.rotations)
