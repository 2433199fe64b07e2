"""CKKS bootstrapping of packed ciphertexts (N/2 slots or sparse packing): the Python
side of the engine's bootstrap (src/CKKS/Bootstrap.cu:169-300 in FIDESlib;
EvalBootstrapSetup / EvalBootstrap, api/CryptoContext.hpp:80, :180).

`Bootstrapper` wraps the engine's C++ setup and evaluation
(fides_bootstrap_setup / fides_eval_bootstrap).  The rest of this module is
an FP64 model of the layout the setup builds, used to check it on the CPU:

Encoding convention (the engine's and OpenFHE's): slot j of a polynomial m
is m(zeta^(5^j)), zeta = exp(i pi / N).  With n = N/2 and
w_k = m_k + i m_(k+n) this is z = U0 w, U0[j, k] = zeta^(5^j k), and U0
factors as U0 = B_n ... B_2 P: P the bit-reversal permutation and B_len one
radix-2 butterfly stage (pairs at distance len/2 inside blocks of len, twiddle
exp(2 pi i (5^j mod 4 len) / (4 len)) for pair j).  Each stage is a matrix
with three non-zero diagonals (offsets 0, +len/2, -len/2), so a group of s
consecutive stages has at most 2^(s+1) - 1 diagonals -- the level-budget
decomposition OpenFHE uses (EvalCoeffsToSlotsPrecompute).  CoeffsToSlots
applies B_2^-1 ... B_n^-1 (slots come out in bit-reversed coefficient order,
which the slotwise EvalMod does not care about); SlotsToCoeffs applies
B_n ... B_2.  Each group runs as one baby-step / giant-step linear transform
(CKKS::LinearTransform) with hoisted baby steps.

Scales.  The raised ciphertext holds c = S0 m + q0 I (S0 its scaling factor,
I the small integer overflow of the decryption over q0).  CoeffsToSlots
carries the factor S0 / (2 K q0), spread evenly over its groups, so that the
real and imaginary halves hold c / (K q0) in [-1, 1].  EvalMod evaluates
(2 pi)^(-1/2^r) cos(2 pi (K x - 1/4) / 2^r) as a Chebyshev series and r
double-angle steps, giving sin(2 pi c / q0) / (2 pi) ~= S0 m / q0, then
multiplies by the integer 2^deg, deg = round(log2(q0 / S0)) (the reference's
`post`).  SlotsToCoeffs carries q0 / (S0 2^deg) ~= 1.  The reference spends a
separate level on the scalar product by pre / (K N) (Bootstrap.cu:219-231);
here that factor lives in the CoeffsToSlots diagonals, which costs no level.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["encode", "encode_coefficients", "special_fft", "special_fft_inverse", "Bootstrapper", "LayoutModel",
           "evalmod_coefficients"]


# --------------------------------------------------------------------------
# the factored special FFT (host, FP64)
# --------------------------------------------------------------------------

def _bitrev(n: int) -> np.ndarray:
    bits = n.bit_length() - 1
    r = np.zeros(n, dtype=np.int64)
    for b in range(bits):
        r |= ((np.arange(n) >> b) & 1) << (bits - 1 - b)
    return r


def _stages(n: int) -> list:
    """[(h, tw)] for the butterfly stages B_2 .. B_n (h = len / 2, tw[j] the twiddle of pair j)."""
    out = []
    h = 1
    while h < n:
        m = 8 * h  # 4 len
        e = np.array([pow(5, j, m) for j in range(h)], dtype=np.float64)
        out.append((h, np.exp(2j * np.pi * e / m)))
        h *= 2
    return out


def special_fft(w: np.ndarray) -> np.ndarray:
    """z = U0 w (coefficients -> slots), O(n log n)."""
    n = len(w)
    v = np.asarray(w, dtype=np.complex128)[_bitrev(n)].copy()
    for h, tw in _stages(n):
        v = v.reshape(-1, 2, h)
        a, b = v[:, 0, :].copy(), v[:, 1, :] * tw
        v[:, 0, :], v[:, 1, :] = a + b, a - b
        v = v.reshape(n)
    return v


def special_fft_inverse(z: np.ndarray) -> np.ndarray:
    """w = U0^-1 z (slots -> coefficients)."""
    n = len(z)
    v = np.asarray(z, dtype=np.complex128).copy()
    for h, tw in reversed(_stages(n)):
        v = v.reshape(-1, 2, h)
        a, b = v[:, 0, :].copy(), v[:, 1, :].copy()
        v[:, 0, :], v[:, 1, :] = (a + b) / 2, (a - b) / (2 * tw)
        v = v.reshape(n)
    return v[_bitrev(n)]


def _stage_diagonals(n: int, h: int, tw: np.ndarray, inverse: bool) -> dict:
    """One butterfly stage as {offset d: diag} with out[p] = sum_d diag[p] in[(p + d) mod n]."""
    p = np.arange(n)
    r = p % (2 * h)
    low = r < h
    t = tw[np.where(low, r, r - h)]
    d0 = np.where(low, 0.5 if inverse else 1.0, -0.5 / t if inverse else -t)
    dp = np.where(low, 0.5 if inverse else t, 0.0)
    dm = np.where(low, 0.0, 0.5 / t if inverse else 1.0)
    out = {0: d0.astype(np.complex128)}
    for d, v in ((h % n, dp), ((n - h) % n, dm)):
        out[d] = out.get(d, 0) + v.astype(np.complex128)
    return out


def _compose(A: dict, B: dict, n: int) -> dict:
    """Diagonals of A @ B: (A B)_d[p] = sum_(a + b = d) A_a[p] B_b[p + a]."""
    out = {}
    for a, va in A.items():
        for b, vb in B.items():
            d = (a + b) % n
            out[d] = out.get(d, 0) + va * np.roll(vb, -a)
    return out


def _groups(n: int, budget: int, inverse: bool) -> list:
    """Stage groups in the order they are applied: [(min h, number of stages, diagonals)]."""
    st = _stages(n)
    if inverse:
        st = st[::-1]  # B_n^-1 first
    logn = len(st)
    budget = max(1, min(budget, logn))
    sizes = [logn // budget + (1 if i < logn % budget else 0) for i in range(budget)]
    if not inverse:
        sizes = sizes[::-1]
    groups, at = [], 0
    for s in sizes:
        M = None
        for h, tw in st[at:at + s]:
            D = _stage_diagonals(n, h, tw, inverse)
            M = D if M is None else _compose(D, M, n)
        hs = [h for h, _ in st[at:at + s]]
        groups.append((min(hs), s, M))
        at += s
    return groups


def _bsgs(n: int, stride: int, s: int, diags: dict, factor: complex, ns: int | None = None,
          left_mask=None, right_mask=None) -> dict:
    """Lay one stage group out for the linear transform over n slots: offsets
    (lo + t) stride, t < cnt, baby step b, giant step b stride; term b j + i
    carries its diagonal pre-rotated by lo stride + j b stride.  For ns < n
    slots the group's diagonals (period ns) are tiled over n; left_mask
    multiplies the output slots, right_mask the input slots (taken at the
    offset actually rotated by)."""
    ns = n if ns is None else ns
    lo, cnt = -((1 << s) - 1), (1 << (s + 1)) - 1
    if cnt > ns // stride:  # the group spans every multiple of the stride
        lo, cnt = 0, ns // stride
    b = 1 << max(0, math.ceil(math.log2(math.sqrt(cnt))))
    b = min(b, cnt)
    off = (lo * stride) % n
    vals = []
    for t in range(cnt):
        d_act = (lo + t) * stride
        v = diags.get(d_act % ns)
        v = np.zeros(n, dtype=np.complex128) if v is None else np.tile(v, n // ns) * factor
        if left_mask is not None:
            v = v * left_mask
        if right_mask is not None:
            v = v * np.roll(right_mask, -d_act)
        j = t // b
        vals.append(np.roll(v, lo * stride + j * b * stride))
    rots = {(i * stride) % n for i in range(1, b)}
    if cnt > b:
        rots.add((b * stride) % n)
    rots.add(off)
    rots.discard(0)
    return {"diag": vals, "bstep": b, "stride": stride, "offset": off, "rotations": rots}


# --------------------------------------------------------------------------
# encoding (MakeCKKSPackedPlaintext, api/CryptoContext.cpp:475-514)
# --------------------------------------------------------------------------

def encode_coefficients(z, N: int, scale: float) -> np.ndarray:
    """Integer coefficients round(scale * [Re w, Im w]) of the N/2 slot values z, w = U0^-1 z."""
    z = np.asarray(z, dtype=np.complex128)
    if z.shape != (N // 2,):
        raise ValueError(f"need {N // 2} slot values")
    w = special_fft_inverse(z)
    m = np.concatenate([w.real, w.imag]) * scale
    if np.abs(m).max() >= 2.0 ** 62:
        raise ValueError("encoded coefficients exceed 62 bits")
    return np.rint(m).astype(np.int64)


def encode(ctx, z, level: int, scale: float):
    """Plaintext (MakeCKKSPackedPlaintext) of the N/2 slot values z: the engine's encoder (fides_encode)."""
    return ctx.encode(z, level, scale)


# --------------------------------------------------------------------------
# EvalMod coefficients
# --------------------------------------------------------------------------

def evalmod_coefficients(K: float, double_angle: int, degree: int) -> np.ndarray:
    """Chebyshev coefficients (sum_k c_k T_k, c_0 not halved) on [-1, 1] of
    (2 pi)^(-1/2^r) cos(2 pi (K x - 1/4) / 2^r), so that r double-angle steps
    y <- 2 y^2 - (2 pi)^(-2^(j-r)) give sin(2 pi K x) / (2 pi)."""
    r = double_angle
    a = (2 * np.pi) ** (-1.0 / (1 << r))

    def f(x):
        return a * np.cos(2 * np.pi * (K * x - 0.25) / (1 << r))

    return np.polynomial.chebyshev.chebinterpolate(f, degree)


# --------------------------------------------------------------------------
# FP64 model of the setup's layout, and the bootstrapper
# --------------------------------------------------------------------------

class LayoutModel:
    """What fides_bootstrap_setup builds, in FP64: per CoeffsToSlots /
    SlotsToCoeffs group the slot values of every diagonal (rotated for the
    linear transform), baby step, stride and offset; the rotation indices; the
    EvalMod coefficients and the integer factor post."""

    def __init__(self, N: int, q0: int, scale: float, level_budget=(3, 3), K: float = 28.0, double_angle: int = 3,
                 degree: int = 119, slots: int | None = None, batched_giant: bool = False):
        self.N, self.n = N, N // 2
        self.slots = self.n if slots is None else int(slots)
        ns, n = self.slots, self.n
        if ns < 2 or ns > n or ns & (ns - 1):
            raise ValueError("slots must be a power of two in [2, N/2]")
        self.S0, self.K, self.r = float(scale), float(K), int(double_angle)
        self.deg = int(round(math.log2(float(q0) / self.S0)))
        self.post = 1 << self.deg
        self.coeffs = evalmod_coefficients(self.K, self.r, degree)
        cts = _groups(ns, level_budget[0], inverse=True)
        stc = _groups(ns, level_budget[1], inverse=False)
        gap = n // ns  # the subring trace multiplies by gap
        fa = (self.S0 / (2.0 * self.K * float(q0) * gap)) ** (1.0 / len(cts))
        fb = float(q0) / (self.S0 * self.post)
        lm = rm = None
        if ns < n:  # real / imaginary halves side by side in blocks of ns (period 2 ns)
            lo = ((np.arange(n) % (2 * ns)) < ns).astype(np.float64)
            lm, rm = lo - 1j * (1 - lo), lo + 1j * (1 - lo)
        self.cts = [_bsgs(n, h, s, D, fa, ns, left_mask=lm if g == len(cts) - 1 else None)
                    for g, (h, s, D) in enumerate(cts)]
        self.stc = [_bsgs(n, h, s, D, fb if g == 0 else 1.0, ns, right_mask=rm if g == 0 else None)
                    for g, (h, s, D) in enumerate(stc)]
        rots = set().union(*[g["rotations"] for g in self.cts + self.stc])
        if batched_giant:
            for g in self.cts + self.stc:
                ng = -(-len(g["diag"]) // g["bstep"])
                rots |= {(j * g["bstep"] * g["stride"]) % n for j in range(2, ng)}
        if ns < n:
            rots |= {ns << t for t in range(gap.bit_length() - 1)}  # trace, and the final fold by ns
        self.rotations = sorted(rots)


class Bootstrapper:
    """EvalBootstrapSetup + EvalBootstrap (the engine's C++ setup and pipeline).

    level_budget  (CoeffsToSlots levels, SlotsToCoeffs levels), as OpenFHE's levelBudget
    K             bound on the raised overflow |I| (OpenFHE's K_SPARSE = 28 for sparse secrets)
    double_angle  double-angle iterations r of EvalMod
    degree        Chebyshev degree of EvalMod
    scale         the scaling factor S0 of the ciphertexts to bootstrap (2^scale_bits)
    top_level     the level ModRaise lifts to (default L)

    batched_giant each transform's giant steps as one batched rotation (needs the keys of
                  every giant step; they are in `rotations`)
    slots         packed slots (a power of two <= N/2, default N/2): fewer is sparse
                  packing -- the message repeats with period `slots`
    q_basis_diagonals  the transforms' diagonals in the Q basis (every rotation
                  ModDowns at once) instead of P*Q, the extended-basis
                  transform OpenFHE and FIDESlib run
    sparse_encaps sparse-secret encapsulation: the level-0 input is switched to a
                  sparse secret before ModRaise and back after it (switching keys
                  SSE_DENSE_TO_SPARSE / SSE_SPARSE_TO_DENSE loaded in the context)

    `rotations` lists the rotation indices whose keys the context must hold
    (EvalBootstrapKeyGen's); the conjugation key (automorphism 2N - 1) is
    needed too."""

    def __init__(self, ctx, scale: float, level_budget=(3, 3), K: float = 28.0, double_angle: int = 3,
                 degree: int = 119, top_level: int | None = None, sparse_encaps: bool = False,
                 slots: int | None = None, batched_giant: bool = False, q_basis_diagonals: bool = False):
        import ctypes as C

        from . import BootstrapParams, _check, lib

        self.ctx = ctx
        p = BootstrapParams(int(level_budget[0]), int(level_budget[1]), float(K), int(double_angle), int(degree),
                            float(scale), -1 if top_level is None else int(top_level), 1 if sparse_encaps else 0,
                            0 if slots is None else int(slots), 1 if batched_giant else 0,
                            1 if q_basis_diagonals else 0)
        h = C.c_void_p()
        _check(lib().fides_bootstrap_setup(ctx.h, C.byref(p), C.byref(h)))
        self.h = h
        n = C.c_int32()
        _check(lib().fides_bootstrap_rotations(self.h, None, 0, C.byref(n)))
        idx = (C.c_int32 * max(1, n.value))()
        _check(lib().fides_bootstrap_rotations(self.h, idx, n.value, C.byref(n)))
        self.rotations = [int(v) for v in idx[: n.value]]

    def bootstrap(self, ct, out=None):
        """EvalBootstrap: a ciphertext at any level -> the same slots near the top of the chain."""
        from . import Ciphertext, _check, lib

        o = Ciphertext(self.ctx) if out is None else out
        _check(lib().fides_eval_bootstrap(self.ctx.h, self.h, ct.h, o.h))
        return o

    def bootstrap_batch(self, cts, outs=None):
        """Bootstrap several ciphertexts together (their EvalMods in lockstep, one batched key switch per
        product); returns the outputs."""
        import ctypes as C

        from . import Ciphertext, _check, lib

        n = len(cts)
        outs = [Ciphertext(self.ctx) for _ in cts] if outs is None else outs
        hin = (C.c_void_p * n)(*[c.h for c in cts])
        hout = (C.c_void_p * n)(*[o.h for o in outs])
        _check(lib().fides_eval_bootstrap_batch(self.ctx.h, self.h, hin, hout, n))
        return outs

    def close(self):
        """Free the plan's diagonals (before the context: a closed context has freed them already)."""
        from . import lib

        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().fides_bootstrap_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
