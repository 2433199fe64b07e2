"""CKKS bootstrapping for fully packed ciphertexts (N/2 slots), assembled from
the engine's entry points -- FIDESlib CKKS::Bootstrap (src/CKKS/Bootstrap.cu:169-300)
and its setup (EvalBootstrapSetup, api/CryptoContext.hpp:80; precomputation
src/CKKS/BootstrapPrecomputation.cuh), OpenFHE EvalBootstrap:

  ModRaise -> CoeffsToSlots -> (conjugate; real / imaginary split) ->
  EvalMod on both halves -> recombine -> SlotsToCoeffs

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

__all__ = ["encode", "encode_limbs", "encode_coefficients", "special_fft", "special_fft_inverse", "Bootstrapper", "evalmod_coefficients"]


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


def _bsgs(n: int, stride: int, s: int, diags: dict, factor: complex) -> dict:
    """Lay one stage group out for the linear transform: offsets (lo + t) stride,
    t < cnt, baby step b, giant step b stride; term b j + i carries its diagonal
    pre-rotated by lo stride + j b stride."""
    lo, cnt = -((1 << s) - 1), (1 << (s + 1)) - 1
    if cnt > n // stride:  # the group spans every multiple of the stride
        lo, cnt = 0, n // stride
    b = 1 << max(0, math.ceil(math.log2(math.sqrt(cnt))))
    b = min(b, cnt)
    off = (lo * stride) % n
    vals = []
    for t in range(cnt):
        d = ((lo + t) * stride) % n
        v = diags.get(d)
        v = np.zeros(n, dtype=np.complex128) if v is None else v * factor
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


def encode_limbs(ctx, z, level: int, scale: float) -> np.ndarray:
    """Evaluation-form residues [level + 1, N] of the N/2 slot values z:
    coefficients round(scale * [Re w, Im w]), w = U0^-1 z, reduced mod each
    q_i on the host and taken to evaluation form by the engine's NTT."""
    c = encode_coefficients(z, ctx.N, scale)
    q = ctx.raw.moduli[: level + 1]
    res = np.empty((level + 1, ctx.N), dtype=np.uint64)
    for i, qi in enumerate(q):
        res[i] = np.mod(c, np.int64(qi)).astype(np.uint64)
    d = ctx.device_array(res.shape, res)
    ctx.ntt(d, level + 1, 0, False)
    limbs = d.numpy()
    d.free()
    return limbs


def encode(ctx, z, level: int, scale: float, noise_level: int = 1):
    """Plaintext (MakeCKKSPackedPlaintext) of the N/2 slot values z at `level` and scaling factor `scale`."""
    return ctx.plaintext(encode_limbs(ctx, z, level, scale), noise=scale, noise_level=noise_level)


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
# the bootstrapper
# --------------------------------------------------------------------------

class Bootstrapper:
    """EvalBootstrapSetup + EvalBootstrap for fully packed ciphertexts.

    level_budget  (CoeffsToSlots groups, SlotsToCoeffs groups), as OpenFHE's levelBudget
    K             bound on |I| (OpenFHE K_SPARSE = 28 for sparse secrets)
    double_angle  double-angle iterations r
    degree        Chebyshev degree of EvalMod
    scale         the scaling factor S0 of the ciphertexts to bootstrap (2^scale_bits)

    `rotations` lists the rotation indices whose keys the context must hold;
    the conjugation key (automorphism 2N - 1) is needed too."""

    def __init__(self, ctx, scale: float, level_budget=(3, 3), K: float = 28.0, double_angle: int = 3,
                 degree: int = 119, top_level: int | None = None):
        self.ctx = ctx
        self.N = ctx.N
        self.n = ctx.N // 2
        self.top = ctx.L if top_level is None else top_level
        self.S0 = float(scale)
        self.K = float(K)
        self.r = int(double_angle)
        q0 = float(ctx.raw.moduli[0])
        self.deg = int(round(math.log2(q0 / self.S0)))
        self.post = 1 << self.deg
        self.coeffs = evalmod_coefficients(self.K, self.r, degree)
        cts = _groups(self.n, level_budget[0], inverse=True)
        stc = _groups(self.n, level_budget[1], inverse=False)
        a = self.S0 / (2.0 * self.K * q0)
        fa = a ** (1.0 / len(cts))
        fb = q0 / (self.S0 * self.post)
        self.cts = [_bsgs(self.n, h, s, D, fa) for h, s, D in cts]
        self.stc = [_bsgs(self.n, h, s, D, fb if g == 0 else 1.0) for g, (h, s, D) in enumerate(stc)]
        self.rotations = sorted(set().union(*[g["rotations"] for g in self.cts + self.stc]))
        self._pts = None

    def setup(self):
        """Encode every diagonal at the top level (the linear transform uses as
        many limbs as its input has)."""
        sc = self.S0
        self._pts = [[[encode(self.ctx, v, self.top, sc) for v in g["diag"]] for g in grp]
                     for grp in (self.cts, self.stc)]
        self.ctx.synchronize()
        return self

    def _lt(self, ct, g, pts):
        return ct.linear_transform(pts, g["bstep"], g["stride"], g["offset"])

    def bootstrap(self, ct):
        """EvalBootstrap: a ciphertext at any level -> the same slots at level top - depth()."""
        if self._pts is None:
            self.setup()
        x = ct.copy()
        if x.info()["noise_level"] == 2:
            x.rescale()
        if x.level > 0:
            x.drop_to_level(0)
        x = x.mod_raise(self.top)
        for g, pts in zip(self.cts, self._pts[0]):
            x = self._lt(x, g, pts)
        x.rescale()
        conj = x.conjugate()
        im = x.sub(conj)
        re = x.add(conj, out=x)
        im.mult_monomial(3 * self.N // 2, out=im)  # times -i
        re = re.mod_reduction(self.coeffs, self.r, self.post)
        im = im.mod_reduction(self.coeffs, self.r, self.post)
        im.mult_monomial(self.N // 2, out=im)  # times i
        u = re.add(im, out=re)
        for g, pts in zip(self.stc, self._pts[1]):
            u = self._lt(u, g, pts)
        u.rescale()
        return u
