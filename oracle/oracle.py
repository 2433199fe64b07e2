"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py, never by the product package.  It plays the part OpenFHE plays for
FIDESlib: parameter generation, key generation, CKKS encode/encrypt/decrypt and
a CPU evaluation of every op on the path, against which the HIP engine's
residues are compared bit-exactly.

Encoding follows OpenFHE's CKKS packed encoding (slot j <-> root
zeta^(5^j), zeta = exp(i pi / N)); it is FP64 host code.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
    lib = C.CDLL(_LIB_PATH)
    U64, I32, VP, D = C.c_uint64, C.c_int, C.c_void_p, C.c_double
    sig = {
        "orc_mulmod": (U64, [U64, U64, U64]),
        "orc_powmod": (U64, [U64, U64, U64]),
        "orc_invmod": (U64, [U64, U64]),
        "orc_is_prime": (I32, [U64]),
        "orc_gen_primes": (I32, [I32, I32, I32, u64p, I32, u64p]),
        "orc_min_root": (U64, [U64, U64]),
        "orc_ntt_tables": (None, [I32, U64, U64, u64p, u64p, C.POINTER(U64)]),
        "orc_ntt_fwd": (None, [u64p, I32, U64, u64p]),
        "orc_ntt_inv": (None, [u64p, I32, U64, u64p, U64]),
        "orc_ntt_bruteforce": (None, [u64p, u64p, I32, U64, U64]),
        "orc_ctx_create": (VP, [I32, I32, I32, I32, u64p, u64p]),
        "orc_ctx_free": (None, [VP]),
        "orc_ctx_alpha": (I32, [VP]),
        "orc_set_threads": (None, [I32]),
        "orc_get_threads": (I32, []),
        "orc_poly_ntt": (None, [VP, u64p, I32, i32p, I32]),
        "orc_sample_ternary": (None, [U64, I32, i8p]),
        "orc_poly_from_int": (None, [VP, i64p, I32, i32p, u64p]),
        "orc_sample_uniform": (None, [VP, U64, I32, i32p, u64p]),
        "orc_sample_gauss": (None, [VP, U64, D, I32, i32p, u64p]),
        "orc_ksk_gen": (None, [VP, U64, u64p, u64p, u64p, u64p]),
        "orc_encrypt_sk": (None, [VP, U64, I32, u64p, u64p, u64p, u64p]),
        "orc_decrypt_coef": (None, [VP, I32, u64p, u64p, u64p, u64p]),
        "orc_crt_to_double": (None, [VP, I32, u64p, f64p]),
        "orc_tensor": (None, [VP, I32, u64p, u64p, u64p, u64p, u64p, u64p, u64p]),
        "orc_num_digits": (I32, [VP, I32]),
        "orc_modup": (None, [VP, I32, u64p, u64p]),
        "orc_ksk_inner": (None, [VP, I32, u64p, u64p, u64p, u64p, u64p]),
        "orc_moddown": (None, [VP, I32, u64p, u64p]),
        "orc_keyswitch": (None, [VP, I32, u64p, u64p, u64p, u64p, u64p]),
        "orc_eval_mult": (None, [VP, I32, u64p, u64p, u64p, u64p, u64p, u64p, u64p, u64p]),
        "orc_eval_square": (None, [VP, I32, u64p, u64p, u64p, u64p, u64p, u64p]),
        "orc_rescale": (None, [VP, I32, u64p, u64p]),
        "orc_automorph": (None, [VP, I32, U64, u64p, u64p]),
        "orc_rot_index": (U64, [I32, I32]),
        "orc_rotate": (None, [VP, I32, U64, u64p, u64p, u64p, u64p, u64p, u64p]),
        "orc_rotate_hoisted": (None, [VP, I32, I32, u64p, u64p, u64p, VP, VP, VP, VP]),
        "orc_add": (None, [VP, I32, i32p, u64p, u64p, u64p]),
        "orc_sub": (None, [VP, I32, i32p, u64p, u64p, u64p]),
        "orc_mul": (None, [VP, I32, i32p, u64p, u64p, u64p]),
        "orc_scalar_mul": (None, [VP, I32, u64p, u64p, u64p]),
        "orc_scalar_add": (None, [VP, I32, u64p, u64p, u64p]),
        "orc_const_partqhatinv": (U64, [VP, I32, I32, I32]),
        "orc_const_partqhat_modp": (U64, [VP, I32, I32, I32, I32]),
        "orc_const_phatinv_modp": (U64, [VP, I32]),
        "orc_const_phat_modq": (U64, [VP, I32, I32]),
        "orc_const_pinv_modq": (U64, [VP, I32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def gen_primes(logN: int, bits: int, count: int, exclude=()) -> list[int]:
    out = np.zeros(max(count, 1), dtype=np.uint64)
    ex = np.array(list(exclude) or [0], dtype=np.uint64)
    n = lib().orc_gen_primes(logN, bits, count, ex, len(exclude), out)
    if n != count:
        raise ValueError("not enough primes")
    return [int(x) for x in out[:count]]


def compute_K(moduli_q: list[int], dnum: int, special_bits: int) -> int:
    """Number of special primes: the reference's ContextData::computeK
    (src/CKKS/Context.cu:286-302) counts each special prime as bits-1 until
    the largest digit's bit count is covered."""
    alpha = -(-len(moduli_q) // dnum)
    digits = [moduli_q[i:i + alpha] for i in range(0, len(moduli_q), alpha)]
    maxbits = max(sum(int(q).bit_length() for q in d) for d in digits)
    k, bits = 0, 0
    while bits < maxbits:
        bits += special_bits - 1
        k += 1
    return k


@dataclass
class Params:
    """RNS-CKKS parameter set (what FIDESlib reads as RawParams)."""

    logN: int
    L: int
    K: int
    dnum: int
    moduli: list  # Q primes (L+1) then P primes (K)
    roots: list
    scale_bits: int

    @property
    def N(self):
        return 1 << self.logN

    @property
    def q(self):
        return self.moduli[: self.L + 1]

    @property
    def p(self):
        return self.moduli[self.L + 1:]

    @property
    def alpha(self):
        return -(-(self.L + 1) // self.dnum)


def make_params(logN: int, first_bits: int, scale_bits: int, L: int, special_bits: int, dnum: int,
                K: int | None = None) -> Params:
    """Primes per the workload spec: the largest primes below 2^b with
    q == 1 mod 2N; q0, then L scaling primes, then the special primes (all
    distinct); psi = the minimal primitive 2N-th root of each prime."""
    q0 = gen_primes(logN, first_bits, 1)
    qs = gen_primes(logN, scale_bits, L, exclude=q0)
    if K is None:
        K = compute_K(q0 + qs, dnum, special_bits)
    ps = gen_primes(logN, special_bits, K, exclude=q0 + qs)
    moduli = q0 + qs + ps
    roots = [int(lib().orc_min_root(q, 2 << logN)) for q in moduli]
    return Params(logN, L, K, dnum, moduli, roots, scale_bits)


# workload parameter sets (see DESIGN.md for the sizes)
def params_small() -> Params:
    """N=2^13, q0 60-bit, 3 x 40-bit, dnum=2, K=2 x 60-bit, scale 2^40."""
    return make_params(13, 60, 40, 3, 60, 2)


def params_production() -> Params:
    """N=2^16, q0 60-bit, 24 x 50-bit, dnum=3, K=8 x 60-bit, scale 2^50."""
    return make_params(16, 60, 50, 24, 60, 3)


def ids(n: int) -> np.ndarray:
    return np.arange(n, dtype=np.int32)


def rot_index(logN: int, r: int) -> int:
    return int(lib().orc_rot_index(logN, r))


# --------------------------------------------------------------------------
# CKKS encoding (FP64, host)
# --------------------------------------------------------------------------

def slot_exponents(logN: int, nslots: int | None = None) -> np.ndarray:
    N = 1 << logN
    nslots = N // 2 if nslots is None else nslots
    e = np.empty(nslots, dtype=np.int64)
    x = 1
    for j in range(nslots):
        e[j] = x
        x = (x * 5) % (2 * N)
    return e


def encode_coeffs(z: np.ndarray, logN: int, scale: float) -> np.ndarray:
    """Slots (N/2 complex or real values) -> rounded integer coefficients."""
    N = 1 << logN
    z = np.asarray(z, dtype=np.complex128)
    assert z.shape == (N // 2,)
    e = slot_exponents(logN)
    w = np.zeros(2 * N, dtype=np.complex128)
    w[e] = z
    w[(2 * N - e) % (2 * N)] = np.conj(z)
    m = np.fft.fft(w)[:N].real / N
    return np.rint(m * scale).astype(np.int64)


def decode_coeffs(m: np.ndarray, logN: int, scale: float) -> np.ndarray:
    N = 1 << logN
    x = np.zeros(2 * N, dtype=np.float64)
    x[:N] = m
    ev = np.fft.ifft(x) * (2 * N)
    return ev[slot_exponents(logN)] / scale


class Oracle:
    """Oracle CKKS context: keys, ciphertexts and CPU evaluation of each op."""

    def __init__(self, params: Params, seed: int = 0, hamming: int | None = None):
        self.P = params
        self.hamming = hamming
        self.N = params.N
        m = np.array(params.moduli, dtype=np.uint64)
        r = np.array(params.roots, dtype=np.uint64)
        self.ctx = lib().orc_ctx_create(params.logN, params.L, params.K, params.dnum, m, r)
        self.np = params.L + 1 + params.K
        self.seed = seed
        self._s = None

    def __del__(self):
        try:
            if self.ctx:
                lib().orc_ctx_free(self.ctx)
                self.ctx = None
        except Exception:
            pass

    # ---- keys ----
    def keygen(self):
        s = np.zeros(self.N, dtype=np.int8)
        if self.hamming:  # sparse ternary secret with `hamming` non-zero coefficients
            rng = np.random.default_rng(self.seed * 7919 + 1)
            pos = rng.choice(self.N, self.hamming, replace=False)
            s[pos] = rng.choice(np.array([-1, 1], dtype=np.int8), self.hamming)
        else:
            lib().orc_sample_ternary(self.seed * 7919 + 1, self.N, s)
        self.s_coef = s
        self._s = self.from_int(s.astype(np.int64), self.np)
        return self._s

    @property
    def s(self):
        if self._s is None:
            self.keygen()
        return self._s

    def from_int(self, coef: np.ndarray, nlimbs: int, primeids=None) -> np.ndarray:
        pid = ids(nlimbs) if primeids is None else np.asarray(primeids, dtype=np.int32)
        out = np.zeros((nlimbs, self.N), dtype=np.uint64)
        lib().orc_poly_from_int(self.ctx, np.ascontiguousarray(coef, dtype=np.int64), nlimbs, pid, out)
        return out

    def ksk_gen(self, s_old: np.ndarray, s_new: np.ndarray, seed: int):
        shape = (self.P.dnum, self.np, self.N)
        a = np.zeros(shape, dtype=np.uint64)
        b = np.zeros(shape, dtype=np.uint64)
        lib().orc_ksk_gen(self.ctx, seed, np.ascontiguousarray(s_old), np.ascontiguousarray(s_new), a, b)
        return a, b

    def relin_key(self):
        s = self.s
        s2 = np.zeros_like(s)
        lib().orc_mul(self.ctx, self.np, ids(self.np), s, s, s2)
        return self.ksk_gen(s2, s, self.seed * 31 + 5)

    def rotation_key(self, r: int):
        """Key for a left rotation by r: switches from s to sigma_{k^-1}(s)
        (OpenFHE EvalAutomorphismKeyGen), k = 5^r mod 2N."""
        k = rot_index(self.P.logN, r)
        kinv = int(lib().orc_invmod(k, 2 * self.N))
        s_perm = self.automorph(self.s, kinv)
        return self.ksk_gen(self.s, s_perm, self.seed * 31 + 1000 + (k % 100003))

    # ---- encode / encrypt ----
    def encode(self, z: np.ndarray, level: int, scale: float | None = None) -> np.ndarray:
        scale = float(2 ** self.P.scale_bits) if scale is None else scale
        return self.from_int(encode_coeffs(z, self.P.logN, scale), level + 1)

    def encrypt(self, m: np.ndarray, level: int, seed: int):
        n = level + 1
        c0 = np.zeros((n, self.N), dtype=np.uint64)
        c1 = np.zeros((n, self.N), dtype=np.uint64)
        lib().orc_encrypt_sk(self.ctx, seed, level, np.ascontiguousarray(self.s[:n]), np.ascontiguousarray(m), c0,
                             c1)
        return c0, c1

    def encrypt_slots(self, z: np.ndarray, level: int, seed: int, scale: float | None = None):
        return self.encrypt(self.encode(z, level, scale), level, seed)

    def decrypt_slots(self, c0: np.ndarray, c1: np.ndarray, scale: float) -> np.ndarray:
        level = c0.shape[0] - 1
        m = np.zeros_like(c0)
        lib().orc_decrypt_coef(self.ctx, level, np.ascontiguousarray(self.s[: level + 1]), np.ascontiguousarray(c0),
                               np.ascontiguousarray(c1), m)
        d = np.zeros(self.N, dtype=np.float64)
        lib().orc_crt_to_double(self.ctx, level, m, d)
        return decode_coeffs(d, self.P.logN, scale)

    def random_poly(self, nlimbs: int, seed: int, primeids=None) -> np.ndarray:
        pid = ids(nlimbs) if primeids is None else np.asarray(primeids, dtype=np.int32)
        out = np.zeros((nlimbs, self.N), dtype=np.uint64)
        lib().orc_sample_uniform(self.ctx, seed, nlimbs, pid, out)
        return out

    # ---- ops ----
    def ntt(self, a: np.ndarray, primeids, inverse: bool) -> np.ndarray:
        out = np.ascontiguousarray(a.copy())
        pid = np.asarray(primeids, dtype=np.int32)
        lib().orc_poly_ntt(self.ctx, out, out.shape[0], pid, int(inverse))
        return out

    def num_digits(self, level: int) -> int:
        return lib().orc_num_digits(self.ctx, level)

    def tensor(self, c0, c1, d0, d1):
        level = c0.shape[0] - 1
        o = [np.zeros_like(c0) for _ in range(3)]
        lib().orc_tensor(self.ctx, level, c0, c1, d0, d1, *o)
        return o

    def modup(self, x: np.ndarray) -> np.ndarray:
        level = x.shape[0] - 1
        ext = np.zeros((self.num_digits(level), level + 1 + self.P.K, self.N), dtype=np.uint64)
        lib().orc_modup(self.ctx, level, np.ascontiguousarray(x), ext)
        return ext

    def ksk_inner(self, ext, key):
        level = ext.shape[1] - 1 - self.P.K
        oa = np.zeros(ext.shape[1:], dtype=np.uint64)
        ob = np.zeros_like(oa)
        lib().orc_ksk_inner(self.ctx, level, ext, key[0], key[1], oa, ob)
        return oa, ob

    def moddown(self, x: np.ndarray) -> np.ndarray:
        level = x.shape[0] - 1 - self.P.K
        y = np.zeros((level + 1, self.N), dtype=np.uint64)
        lib().orc_moddown(self.ctx, level, np.ascontiguousarray(x), y)
        return y

    def keyswitch(self, x, key):
        level = x.shape[0] - 1
        kb = np.zeros_like(x)
        ka = np.zeros_like(x)
        lib().orc_keyswitch(self.ctx, level, np.ascontiguousarray(x), key[0], key[1], kb, ka)
        return kb, ka

    def mult(self, ct1, ct2, key):
        level = ct1[0].shape[0] - 1
        o0 = np.zeros_like(ct1[0])
        o1 = np.zeros_like(ct1[0])
        lib().orc_eval_mult(self.ctx, level, ct1[0], ct1[1], ct2[0], ct2[1], key[0], key[1], o0, o1)
        return o0, o1

    def square(self, ct, key):
        level = ct[0].shape[0] - 1
        o0 = np.zeros_like(ct[0])
        o1 = np.zeros_like(ct[0])
        lib().orc_eval_square(self.ctx, level, ct[0], ct[1], key[0], key[1], o0, o1)
        return o0, o1

    def rescale_poly(self, x):
        level = x.shape[0] - 1
        y = np.zeros((level, self.N), dtype=np.uint64)
        lib().orc_rescale(self.ctx, level, np.ascontiguousarray(x), y)
        return y

    def rescale(self, ct):
        return self.rescale_poly(ct[0]), self.rescale_poly(ct[1])

    def automorph(self, x, k: int):
        out = np.zeros_like(x)
        lib().orc_automorph(self.ctx, x.shape[0], k, np.ascontiguousarray(x), out)
        return out

    def rotate(self, ct, r: int, key):
        level = ct[0].shape[0] - 1
        k = rot_index(self.P.logN, r)
        o0 = np.zeros_like(ct[0])
        o1 = np.zeros_like(ct[0])
        lib().orc_rotate(self.ctx, level, k, ct[0], ct[1], key[0], key[1], o0, o1)
        return o0, o1

    def rotate_hoisted(self, ct, rs, keys):
        level = ct[0].shape[0] - 1
        n = len(rs)
        ks = np.array([rot_index(self.P.logN, r) for r in rs], dtype=np.uint64)
        outs0 = [np.zeros_like(ct[0]) for _ in rs]
        outs1 = [np.zeros_like(ct[0]) for _ in rs]

        def ptrs(arrs):
            return (C.c_void_p * n)(*[a.ctypes.data for a in arrs])

        keep = [k[0] for k in keys] + [k[1] for k in keys]
        assert all(a.flags.c_contiguous for a in keep)
        lib().orc_rotate_hoisted(self.ctx, level, n, ks, ct[0], ct[1], ptrs([k[0] for k in keys]),
                                 ptrs([k[1] for k in keys]), ptrs(outs0), ptrs(outs1))
        return list(zip(outs0, outs1))

    def add(self, a, b):
        o = np.zeros_like(a)
        lib().orc_add(self.ctx, a.shape[0], ids(a.shape[0]), a, b, o)
        return o

    def automorph_key(self, k: int):
        """Key of the automorphism X -> X^k: switches from s to sigma_{k^-1}(s)."""
        kinv = int(lib().orc_invmod(k, 2 * self.N))
        return self.ksk_gen(self.s, self.automorph(self.s, kinv), self.seed * 31 + 2000 + (k % 100003))

    def rotate_k(self, ct, k: int, key):
        """Key switch, ModDown, add c0, then X -> X^k (OpenFHE order), for any odd k."""
        level = ct[0].shape[0] - 1
        o0 = np.zeros_like(ct[0])
        o1 = np.zeros_like(ct[0])
        lib().orc_rotate(self.ctx, level, k, ct[0], ct[1], key[0], key[1], o0, o1)
        return o0, o1

    def mod_raise(self, ct, to: int):
        """OpenFHE / FIDESlib ModRaise: INTT of the q0 limb, centered lift of the
        coefficients into q_0..q_to (SwitchModulus), NTT."""
        q0 = self.P.moduli[0]
        out = []
        for x in ct:
            coef = self.ntt(np.ascontiguousarray(x[:1]), [0], inverse=True)[0].astype(object)
            cen = np.where(coef > q0 // 2, coef - q0, coef)
            limbs = np.stack([np.array([int(v) % q for v in cen], dtype=np.uint64)
                              for q in self.P.moduli[: to + 1]])
            out.append(self.ntt(limbs, ids(to + 1), inverse=False))
        return tuple(out)

    def mult_monomial(self, ct, m: int):
        """Multiplication by X^m: the NTT of the monomial, limb by limb (negacyclic: X^N = -1)."""
        n = ct[0].shape[0]
        m %= 2 * self.N
        mono = np.zeros((n, self.N), dtype=np.uint64)
        for i in range(n):
            q = self.P.moduli[i]
            mono[i, m % self.N] = 1 if m < self.N else q - 1
        mono = self.ntt(mono, ids(n), inverse=False)
        return self.mul(ct[0], mono), self.mul(ct[1], mono)

    def mul(self, a, b):
        o = np.zeros_like(a)
        lib().orc_mul(self.ctx, a.shape[0], ids(a.shape[0]), np.ascontiguousarray(a), np.ascontiguousarray(b), o)
        return o

    # ---- plaintext / scalar ops (FIDESlib Ciphertext::multPt, addPt, multScalar, addScalar) ----
    def mult_plain(self, ct, pt):
        """(c0 pt, c1 pt) on the ciphertext's limbs (Ciphertext.cpp:369-370)."""
        n = ct[0].shape[0]
        return self.mul(ct[0], pt[:n]), self.mul(ct[1], pt[:n])

    def add_plain(self, ct, pt):
        n = ct[0].shape[0]
        return self.add(ct[0], np.ascontiguousarray(pt[:n])), ct[1].copy()

    def mult_scalar(self, ct, elems):
        n = ct[0].shape[0]
        k = np.ascontiguousarray(np.asarray(elems, dtype=np.uint64)[:n])
        outs = []
        for x in ct:
            o = np.zeros_like(x)
            lib().orc_scalar_mul(self.ctx, n, k, np.ascontiguousarray(x), o)
            outs.append(o)
        return tuple(outs)

    def add_scalar(self, ct, elems):
        n = ct[0].shape[0]
        k = np.ascontiguousarray(np.asarray(elems, dtype=np.uint64)[:n])
        o = np.zeros_like(ct[0])
        lib().orc_scalar_add(self.ctx, n, k, np.ascontiguousarray(ct[0]), o)
        return o, ct[1].copy()


class ExtBasis:
    """The extended-basis (P*Q) forms of FIDESlib's key-switching path, composed
    from the oracle's primitives: a P*Q polynomial holds the Q limbs
    q_0..q_level followed by the K special-prime limbs.

      extend          Ciphertext::extend (src/CKKS/Ciphertext.cpp:742-754):
                      P x on the Q limbs, zero special limbs
      rotate          one key switch of c1 kept in P*Q, then the automorphism
                      (rotate_hoisted with ext, Ciphertext.cpp:1100-1123, the
                      c0 addend scaled by P as LimbPartition::add's addScaleB_,
                      LimbPartition.cu:2216-2262; rotate(index, moddown =
                      false), Ciphertext.cpp:756-926, with a P*Q c0)
      moddown         RNSPoly::moddown of each part (Ciphertext::modDown)
      linear_transform  LinearTransform with ext = true, MODDOWN_HOIST and
                      ONLY_C1 (src/CKKS/LinearTransform.cu:82-193)
    """

    def __init__(self, o: Oracle):
        self.o = o
        self.L, self.K = o.P.L, o.P.K
        pp = 1
        for v in o.P.p:
            pp *= int(v)
        self.Pmod = [pp % int(q) for q in o.P.q]

    def ids(self, level: int) -> np.ndarray:
        return np.concatenate([ids(level + 1), np.arange(self.L + 1, self.L + 1 + self.K, dtype=np.int32)])

    def add(self, a, b):
        o = np.zeros_like(a)
        n = a.shape[0]
        lib().orc_add(self.o.ctx, n, self.ids(n - 1 - self.K), np.ascontiguousarray(a), np.ascontiguousarray(b), o)
        return o

    def mul(self, a, b):
        o = np.zeros_like(a)
        n = a.shape[0]
        lib().orc_mul(self.o.ctx, n, self.ids(n - 1 - self.K), np.ascontiguousarray(a), np.ascontiguousarray(b), o)
        return o

    def extend(self, x):
        n = x.shape[0]
        out = np.zeros((n + self.K, x.shape[1]), dtype=np.uint64)
        k = np.ascontiguousarray(np.array(self.Pmod[:n], dtype=np.uint64))
        lib().orc_scalar_mul(self.o.ctx, n, k, np.ascontiguousarray(x), out[:n])
        return out

    def rotate(self, c0, c1, k: int, key):
        """(sigma_k(acc_b + c0'), sigma_k(acc_a)) with (acc_a, acc_b) the key switch of
        c1 (Q basis) before ModDown; c0' = c0 when it is P*Q, else extend(c0)."""
        acc_a, acc_b = self.o.ksk_inner(self.o.modup(c1), key)
        c0p = c0 if c0.shape[0] == c1.shape[0] + self.K else self.extend(c0)
        return self.o.automorph(self.add(acc_b, c0p), k), self.o.automorph(acc_a, k)

    def moddown(self, ct):
        return self.o.moddown(ct[0]), self.o.moddown(ct[1])

    def linear_transform(self, ct, diags, keys, b: int, s: int, offset: int, batched_giant: bool = False):
        """diags: P*Q plaintexts (limbs [level+1+K]); keys by rotation amount.
        Baby steps ext-rotated (index 0: extend), inner sums in P*Q, then per
        giant step only c1 brought down (ModDown) before its key switch, the
        result kept in P*Q; the offset rotation likewise, then one ModDown.
        batched_giant: sum_j rot_{j b s}(inner[j]), each rotation from the
        moddowned c1 of inner[j]."""
        o = self.o
        N2 = o.N // 2
        n = len(diags)
        g = -(-n // b)
        lv = ct[0].shape[0] - 1
        baby = []
        for i in range(b):
            r = i * s
            if r % N2 == 0:
                baby.append((self.extend(ct[0]), self.extend(ct[1])))
            else:
                baby.append(self.rotate(ct[0], ct[1], rot_index(o.P.logN, r), keys[r]))
        inner = []
        for j in range(g):
            acc = None
            for i in range(b):
                if b * j + i >= n:
                    break
                d = diags[b * j + i]
                t = (self.mul(baby[i][0], d), self.mul(baby[i][1], d))
                acc = t if acc is None else (self.add(acc[0], t[0]), self.add(acc[1], t[1]))
            inner.append(acc)

        def rot(x, r):
            return self.rotate(x[0], o.moddown(x[1]), rot_index(o.P.logN, r), keys[r % N2])

        if batched_giant:
            acc = inner[0]
            for j in range(1, g):
                r = j * b * s
                y = inner[j] if r % N2 == 0 else rot(inner[j], r)
                acc = (self.add(acc[0], y[0]), self.add(acc[1], y[1]))
        else:
            acc = inner[g - 1]
            for j in range(g - 2, -1, -1):
                y = acc if (b * s) % N2 == 0 else rot(acc, b * s)
                acc = (self.add(inner[j][0], y[0]), self.add(inner[j][1], y[1]))
        if offset % N2:
            acc = rot(acc, offset)
        assert acc[0].shape[0] == lv + 1 + self.K
        return self.moddown(acc)


def _c_int_div(a: int, b: int) -> int:
    """C integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _c_mod(a: int, b: int) -> int:
    """C remainder (sign of the dividend)."""
    return a - _c_int_div(a, b) * b


def _trunc_double(x: float) -> int:
    """C conversion of a finite double to an integer type (truncation)."""
    return int(x)


def elem_for_mult(moduli, level: int, v: float, sf: float) -> list[int]:
    """FIDESlib ContextData::ElemForEvalMult(level, v) (src/CKKS/Context.cu:375-455):
    the 128-bit integer (int128)(v / 2^a * sf + 0.5), a > 0 only past 125 bits,
    reduced modulo each q_i with C remainder semantics, times 2^a mod q_i."""
    import math

    log_approx = 0
    res = abs(v * sf)
    if res > 0:
        log_sf = int(math.ceil(math.log2(res)))
        log_approx = log_sf - min(log_sf, 125)
    approx = 2.0 ** log_approx
    large = _trunc_double(v / approx * sf + 0.5)
    out = []
    for i in range(level + 1):
        q = moduli[i]
        r = _c_mod(large, q)  # both branches of the reference reduce with C's %
        r = r + q if r < 0 else r
        if log_approx > 0:
            r = r * pow(2, log_approx, q) % q
        out.append(r % q)
    return out


def elem_for_add(moduli, level: int, v_abs: float, noise_deg: int, sf: float) -> list[int]:
    """FIDESlib ContextData::ElemForEvalAddOrSub(level, |v|, noise_deg)
    (src/CKKS/Context.cu:471-564): (uint64)(|v| sf / 2^a + 0.5) (a > 0 only past
    61 bits), times 2^a and times round(sf)^(noise_deg-1), modulo each q_i."""
    import math

    log_approx = 0
    res = abs(v_abs * sf)
    if res > 0:
        log_sf = int(math.ceil(math.log2(res)))
        log_approx = log_sf - min(log_sf, 61)
    approx = 2.0 ** log_approx
    sc = _trunc_double(v_abs * sf / approx + 0.5)
    isf = _trunc_double(sf + 0.5)
    out = []
    for i in range(level + 1):
        q = moduli[i]
        r = sc % q
        if log_approx > 0:
            r = r * pow(2, log_approx, q) % q
        for _ in range(1, noise_deg):
            r = r * isf % q
        out.append(r)
    return out


# --------------------------------------------------------------------------
# Automatic rescaling (FIXEDAUTO / FLEXIBLEAUTO): FIDESlib's metadata rules
# restated over the oracle's ops.  Pinned in tests/test_oracle_auto.py by what
# the rules must achieve: every aligned sum or product decrypts to the sum or
# product of the messages at the scale the metadata records.
# --------------------------------------------------------------------------

@dataclass
class MetaCt:
    """A ciphertext with FIDESlib's metadata (Ciphertext.cuh: NoiseFactor,
    NoiseLevel; level = number of limbs - 1)."""

    c0: np.ndarray
    c1: np.ndarray
    noise: float
    noise_level: int = 1
    slots: int = 0

    @property
    def level(self) -> int:
        return self.c0.shape[0] - 1

    def copy(self) -> "MetaCt":
        return MetaCt(self.c0.copy(), self.c1.copy(), self.noise, self.noise_level, self.slots)


FIXEDMANUAL, FIXEDAUTO, FLEXIBLEAUTO, FLEXIBLEAUTOEXT = 0, 1, 2, 3


class AutoScaling:
    """adjustForAddOrSub / adjustForMult and the add / mult that call them
    (src/CKKS/Ciphertext.cpp:166-234, :409-421, :687-713, :1472-1611)."""

    def __init__(self, o: Oracle, technique: int, sf_real, sf_real_big=None, mod_reduce=None):
        self.o = o
        self.t = technique
        self.sfr = list(sf_real)
        self.sfrb = None if sf_real_big is None else list(sf_real_big)
        self.mrf = [float(q) for q in o.P.q] if mod_reduce is None else list(mod_reduce)

    # ---- primitives (Ciphertext::rescale, multScalarNoPrecheck, dropToLevel) ----
    def rescale(self, a: MetaCt):
        lv = a.level
        a.c0, a.c1 = self.o.rescale((a.c0, a.c1))
        a.noise /= self.mrf[lv]  # ModReduceFactor at the level rescaled from (Ciphertext.cpp:395)
        a.noise_level -= 1

    def mult_scalar_noprecheck(self, a: MetaCt, v: float):
        e = elem_for_mult(self.o.P.moduli, a.level, v, self.sfr[a.level])
        a.c0, a.c1 = self.o.mult_scalar((a.c0, a.c1), e)
        a.noise *= self.sfr[a.level]
        a.noise_level += 1

    @staticmethod
    def drop(a: MetaCt, level: int):
        a.c0 = np.ascontiguousarray(a.c0[: level + 1])
        a.c1 = np.ascontiguousarray(a.c1[: level + 1])

    # ---- Ciphertext::adjustForAddOrSub (Ciphertext.cpp:1472-1594) ----
    def adjust_for_add_or_sub(self, a: MetaCt, b: MetaCt) -> bool:
        if self.t in (FIXEDMANUAL, FIXEDAUTO):
            if a.level - a.noise_level > b.level - b.noise_level:
                if b.noise_level == 1 and a.noise_level == 2:
                    self.rescale(a)
                elif b.noise_level == 2 and a.noise_level == 1:
                    self.mult_scalar_noprecheck(a, 1.0)
                return True
            if b.noise_level == 1 and a.noise_level == 2:
                self.rescale(a)
                return True
            return not (a.noise_level == 1 and b.noise_level == 2)
        l1, l2, d1, d2 = a.level, b.level, a.noise_level, b.noise_level
        if l1 > l2:
            scf1, scf, q1 = a.noise, self.sfr[l1], self.mrf[l1]
            if d1 == 2:
                if d2 == 2:
                    self.mult_scalar_noprecheck(a, b.noise / scf1 * q1 / scf)
                    self.rescale(a)
                    if a.level > b.level:
                        self.drop(a, b.level)
                    a.noise = b.noise
                elif l1 - 1 == l2:
                    self.rescale(a)
                else:
                    self.mult_scalar_noprecheck(a, self.sfrb[l2 + 1] / scf1 * q1 / scf)
                    self.rescale(a)
                    if a.level - 1 > b.level:
                        self.drop(a, b.level + 1)
                    self.rescale(a)
                    a.noise = b.noise
            elif d2 == 2:
                self.mult_scalar_noprecheck(a, b.noise / scf1 / scf)
                self.drop(a, l2)
                a.noise = b.noise
            else:
                self.mult_scalar_noprecheck(a, self.sfrb[l2 + 1] / scf1 / scf)
                if l1 - 1 > l2:
                    self.drop(a, l2 + 1)
                self.rescale(a)
                a.noise = b.noise
            return True
        if l1 < l2:
            return False
        if d1 < d2:
            self.mult_scalar_noprecheck(a, 1.0)
            return True
        return d2 >= d1

    # ---- Ciphertext::adjustForMult (Ciphertext.cpp:1595-1611) ----
    def adjust_for_mult(self, a: MetaCt, b: MetaCt) -> bool:
        ok = self.adjust_for_add_or_sub(a, b)
        if a.noise_level == 2:
            self.rescale(a)
        return ok and b.noise_level != 2

    # ---- Ciphertext::add / sub (Ciphertext.cpp:166-234), on a copy of a ----
    def add(self, a: MetaCt, b: MetaCt, sub: bool = False) -> MetaCt:
        s = a.copy()
        if self.t != FIXEDMANUAL and not self.adjust_for_add_or_sub(s, b):
            b = b.copy()
            if not self.adjust_for_add_or_sub(b, s):
                raise ValueError("operands could not be aligned")
        if self.t != FIXEDMANUAL:
            assert s.noise_level == b.noise_level
        n = min(s.level, b.level) + 1
        f = (lambda x, y: self.o.add(x, self._neg(y))) if sub else self.o.add
        c0 = f(np.ascontiguousarray(s.c0[:n]), np.ascontiguousarray(b.c0[:n]))
        c1 = f(np.ascontiguousarray(s.c1[:n]), np.ascontiguousarray(b.c1[:n]))
        return MetaCt(c0, c1, s.noise, s.noise_level, max(s.slots, b.slots))

    def _neg(self, y):
        q = np.array(self.o.P.moduli[: y.shape[0]], dtype=np.uint64)[:, None]
        return np.where(y == 0, y, q - y).astype(np.uint64)

    # ---- Ciphertext::mult (Ciphertext.cpp:409-421), on a copy of a ----
    def mult(self, a: MetaCt, b: MetaCt, key, rescale: bool = False) -> MetaCt:
        s = a.copy()
        if self.t != FIXEDMANUAL and not self.adjust_for_mult(s, b):
            b = b.copy()
            if not self.adjust_for_mult(b, s):
                raise ValueError("operands could not be aligned")
        assert s.noise_level == 1 and b.noise_level == 1
        n = min(s.level, b.level) + 1
        c = self.o.mult((np.ascontiguousarray(s.c0[:n]), np.ascontiguousarray(s.c1[:n])),
                        (np.ascontiguousarray(b.c0[:n]), np.ascontiguousarray(b.c1[:n])), key)
        out = MetaCt(c[0], c[1], s.noise * b.noise, 2, max(s.slots, b.slots))
        if rescale:
            self.rescale(out)
        return out


def flexible_scaling_factors(moduli_q) -> tuple[list[float], list[float]]:
    """FLEXIBLEAUTO scaling factors indexed by level (FIDESlib's order, RawCiphertext.cu:274-282):
    the top level's factor is its prime, each lower one is the square of the one
    above divided by the prime rescaled away (so a product rescaled from level l
    lands on level l-1's factor); ScalingFactorRealBig[l] = ScalingFactorReal[l]^2,
    the scale of a noise-degree-2 ciphertext at level l."""
    L = len(moduli_q) - 1
    sf = [0.0] * (L + 1)
    sf[L] = float(moduli_q[L])
    for lv in range(L, 0, -1):
        sf[lv - 1] = sf[lv] * sf[lv] / float(moduli_q[lv])
    return sf, [x * x for x in sf]
