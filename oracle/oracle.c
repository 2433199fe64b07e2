/*
 * CPU oracle for the CKKS evaluation path -- TEST INFRASTRUCTURE ONLY (see
 * oracle.h).  Only tests/, smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function restates the algorithm that FIDESlib runs (or delegates to
 * OpenFHE v1.4.2, which is not vendored in the reference: deps/openfhe-src is
 * an empty submodule).  Citations are to the reference repository.
 *
 * Parity pins (tests/test_oracle.py):
 *   modpow/modinv        KATs of test/MathTest.cu:14-37
 *   NTT fwd/inv          cyclic KAT of test/NTTtests.cu:29-38 (p=17, w=9) via the
 *                        negacyclic twist; O(N^2) evaluation at psi^(2brev(j)+1);
 *                        round trip; negacyclic convolution theorem
 *   min root / primes    brute force for small q; primality of every prime
 *   base conversion      exact big-integer CRT sums (Python int) on random inputs
 *   KSK / ModDown        decryption identity <ks, (1,s)> ~ x * s_old within noise,
 *                        plus big-integer closed form of ApproxModDown
 *   rescale              big-integer closed form (x - [x_l]_centered) / q_l and the
 *                        QlQlInvModqlDivqlModq form, both must agree
 *   automorphism         permutation closed form X -> X^k on coefficients
 *   mult/rotate          decrypt-and-decode against FP64 slot arithmetic
 * The sampling functions are the oracle's own (not OpenFHE's streams): keys and
 * ciphertexts are generated here and fed identically to oracle and GPU.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* scalar number theory                                               */
/* ------------------------------------------------------------------ */

uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }

/* FIDESlib src/Math.cu modpow/modinv (square-and-multiply, Fermat inverse) */
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = orc_mulmod(r, a, p);
        a = orc_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

/* Extended Euclid: valid for any modulus coprime with a (MathTest uses 2^61-1). */
uint64_t orc_invmod(uint64_t a, uint64_t p) {
    __int128 t = 0, nt = 1;
    __int128 r = p, nr = a % p;
    while (nr != 0) {
        __int128 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += p;
    return (uint64_t)t;
}

int orc_is_prime(uint64_t n) {
    static const uint64_t bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n % bases[i] == 0) return n == bases[i];
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (int i = 0; i < 12; ++i) {
        uint64_t x = orc_powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; ++r) {
            x = orc_mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* Prime rule of the workload spec: largest primes below 2^bits with
 * q == 1 mod 2N (NTT-friendly), in decreasing order. */
int orc_gen_primes(int logN, int bits, int count, const uint64_t* exclude, int n_exclude, uint64_t* out) {
    uint64_t m = 2ull << logN;
    uint64_t cand = (bits >= 64 ? 0ull : (1ull << bits)) - m + 1;
    int found = 0;
    while (found < count && cand > m) {
        int skip = 0;
        for (int i = 0; i < n_exclude; ++i) skip |= (exclude[i] == cand);
        if (!skip && orc_is_prime(cand)) out[found++] = cand;
        cand -= m;
    }
    return found;
}

/* Minimal primitive m-th root of unity, the rule OpenFHE's RootOfUnity applies
 * after finding any primitive root (it cycles the coprime powers and keeps the
 * smallest).  m is a power of two here, so the coprime powers are the odd ones. */
uint64_t orc_min_root(uint64_t q, uint64_t m) {
    uint64_t e = (q - 1) / m, w = 0;
    for (uint64_t g = 2; g < q; ++g) {
        uint64_t x = orc_powmod(g, e, q);
        if (orc_powmod(x, m / 2, q) == q - 1) { w = x; break; }
    }
    uint64_t w2 = orc_mulmod(w, w, q), cur = w, best = w;
    for (uint64_t k = 3; k < m; k += 2) {
        cur = orc_mulmod(cur, w2, q);
        if (cur < best) best = cur;
    }
    return best;
}

/* ------------------------------------------------------------------ */
/* NTT                                                                */
/* ------------------------------------------------------------------ */

static inline uint32_t brev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

static inline uint64_t shoup_pre(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }
static inline uint64_t mul_shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t p) {
    uint64_t qt = (uint64_t)(((u128)x * wp) >> 64);
    uint64_t r = x * w - qt * p;
    return r >= p ? r - p : r;
}
static inline uint64_t addm(uint64_t a, uint64_t b, uint64_t p) { uint64_t s = a + b; return s >= p ? s - p : s; }
static inline uint64_t subm(uint64_t a, uint64_t b, uint64_t p) { return a >= b ? a - b : a + p - b; }

/* Tables as in ConstantsGPU.cu:224-323 (psi in bit-reversed order). */
void orc_ntt_tables(int logN, uint64_t q, uint64_t psi, uint64_t* psi_rev, uint64_t* psi_inv_rev, uint64_t* n_inv) {
    uint32_t N = 1u << logN;
    uint64_t psi_inv = orc_invmod(psi, q);
    uint64_t pw = 1, pwi = 1;
    for (uint32_t k = 0; k < N; ++k) {
        uint32_t r = brev(k, logN);
        psi_rev[r] = pw;
        psi_inv_rev[r] = pwi;
        pw = orc_mulmod(pw, psi, q);
        pwi = orc_mulmod(pwi, psi_inv, q);
    }
    *n_inv = orc_invmod(N, q);
}

/* Longa-Naehrig CT forward transform with bit-reversed output (OpenFHE
 * ForwardTransformToBitReverseInPlace; reference CPU form test/cpuNTT_nega.cpp:24-97
 * without its final swap loop). */
static void ntt_fwd_core(uint64_t* a, uint32_t N, uint64_t q, const uint64_t* w, const uint64_t* wp) {
    uint32_t t = N;
    for (uint32_t m = 1; m < N; m <<= 1) {
        t >>= 1;
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t j1 = 2 * i * t;
            uint64_t W = w[m + i], Wp = wp[m + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t U = a[j];
                uint64_t V = mul_shoup(a[j + t], W, Wp, q);
                a[j] = addm(U, V, q);
                a[j + t] = subm(U, V, q);
            }
        }
    }
}

/* Gentleman-Sande inverse from bit-reversed input, then multiply by N^-1
 * (OpenFHE InverseTransformFromBitReverseInPlace). */
static void ntt_inv_core(uint64_t* a, uint32_t N, uint64_t q, const uint64_t* w, const uint64_t* wp, uint64_t ninv,
                         uint64_t ninvp) {
    uint32_t t = 1;
    for (uint32_t m = N >> 1; m >= 1; m >>= 1) {
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t j1 = 2 * i * t;
            uint64_t W = w[m + i], Wp = wp[m + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t U = a[j], V = a[j + t];
                a[j] = addm(U, V, q);
                a[j + t] = mul_shoup(subm(U, V, q), W, Wp, q);
            }
        }
        t <<= 1;
    }
    for (uint32_t j = 0; j < N; ++j) a[j] = mul_shoup(a[j], ninv, ninvp, q);
}

void orc_ntt_fwd(uint64_t* a, int logN, uint64_t q, const uint64_t* psi_rev) {
    uint32_t N = 1u << logN;
    uint64_t* wp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint32_t k = 0; k < N; ++k) wp[k] = shoup_pre(psi_rev[k], q);
    ntt_fwd_core(a, N, q, psi_rev, wp);
    free(wp);
}

void orc_ntt_inv(uint64_t* a, int logN, uint64_t q, const uint64_t* psi_inv_rev, uint64_t n_inv) {
    uint32_t N = 1u << logN;
    uint64_t* wp = (uint64_t*)malloc(sizeof(uint64_t) * N);
    for (uint32_t k = 0; k < N; ++k) wp[k] = shoup_pre(psi_inv_rev[k], q);
    ntt_inv_core(a, N, q, psi_inv_rev, wp, n_inv, shoup_pre(n_inv, q));
    free(wp);
}

void orc_ntt_bruteforce(const uint64_t* a, uint64_t* out, int logN, uint64_t q, uint64_t psi) {
    uint32_t N = 1u << logN;
    for (uint32_t j = 0; j < N; ++j) {
        uint64_t x = orc_powmod(psi, 2ull * brev(j, logN) + 1, q);
        /* Horner evaluation of a at x */
        uint64_t acc = 0;
        for (int32_t k = (int32_t)N - 1; k >= 0; --k) acc = addm(orc_mulmod(acc, x, q), a[k] % q, q);
        out[j] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* context                                                            */
/* ------------------------------------------------------------------ */

struct orc_ctx {
    int logN, N, L, K, dnum, alpha, nprimes;
    uint64_t* q;
    uint64_t* root;
    uint64_t** tw;   /* psi_rev */
    uint64_t** twp;  /* shoup companions */
    uint64_t** itw;  /* psi_inv_rev */
    uint64_t** itwp;
    uint64_t* ninv;
    uint64_t* ninvp;
    /* base-conversion constants */
    uint64_t* partqhatinv;   /* [dnum][alpha (size-1)][alpha (i)] */
    uint64_t* partqhat_modp; /* [dnum][alpha][alpha][nprimes] */
    uint64_t* phatinv_modp;  /* [K] */
    uint64_t* phat_modq;     /* [K][nprimes] */
    uint64_t* pinv_modq;     /* [nprimes] */
    uint64_t* pmodq;         /* [nprimes] */
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_FOR _Pragma("omp parallel for schedule(dynamic,1) num_threads(orc_get_threads())")

int orc_ctx_alpha(const orc_ctx* c) { return c->alpha; }

static inline size_t PQ_IDX(const orc_ctx* c, int d, int s, int i) {
    return ((size_t)d * c->alpha + s) * c->alpha + i;
}

orc_ctx* orc_ctx_create(int logN, int L, int K, int dnum, const uint64_t* moduli, const uint64_t* roots) {
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->logN = logN;
    c->N = 1 << logN;
    c->L = L;
    c->K = K;
    c->dnum = dnum;
    c->alpha = (L + 1 + dnum - 1) / dnum; /* OpenFHE numPerPartQ = ceil((L+1)/dnum) */
    c->nprimes = L + 1 + K;
    int np = c->nprimes, N = c->N;
    c->q = (uint64_t*)malloc(sizeof(uint64_t) * np);
    c->root = (uint64_t*)malloc(sizeof(uint64_t) * np);
    memcpy(c->q, moduli, sizeof(uint64_t) * np);
    memcpy(c->root, roots, sizeof(uint64_t) * np);
    c->tw = (uint64_t**)malloc(sizeof(uint64_t*) * np);
    c->twp = (uint64_t**)malloc(sizeof(uint64_t*) * np);
    c->itw = (uint64_t**)malloc(sizeof(uint64_t*) * np);
    c->itwp = (uint64_t**)malloc(sizeof(uint64_t*) * np);
    c->ninv = (uint64_t*)malloc(sizeof(uint64_t) * np);
    c->ninvp = (uint64_t*)malloc(sizeof(uint64_t) * np);
    OMP_FOR
    for (int i = 0; i < np; ++i) {
        c->tw[i] = (uint64_t*)malloc(sizeof(uint64_t) * N);
        c->twp[i] = (uint64_t*)malloc(sizeof(uint64_t) * N);
        c->itw[i] = (uint64_t*)malloc(sizeof(uint64_t) * N);
        c->itwp[i] = (uint64_t*)malloc(sizeof(uint64_t) * N);
        orc_ntt_tables(logN, c->q[i], c->root[i], c->tw[i], c->itw[i], &c->ninv[i]);
        for (int k = 0; k < N; ++k) {
            c->twp[i][k] = shoup_pre(c->tw[i][k], c->q[i]);
            c->itwp[i][k] = shoup_pre(c->itw[i][k], c->q[i]);
        }
        c->ninvp[i] = shoup_pre(c->ninv[i], c->q[i]);
    }
    /* ModUp constants: RawCiphertext.cu:219-467 names PartQlHatInvModq and
     * PartQlHatModp; their values are products of the digit primes. */
    int A = c->alpha;
    c->partqhatinv = (uint64_t*)calloc((size_t)dnum * A * A, sizeof(uint64_t));
    c->partqhat_modp = (uint64_t*)calloc((size_t)dnum * A * A * np, sizeof(uint64_t));
    for (int d = 0; d < dnum; ++d) {
        int start = d * A;
        for (int s = 1; s <= A && start + s <= L + 1; ++s) {
            for (int i = 0; i < s; ++i) {
                uint64_t qi = c->q[start + i];
                uint64_t inv = 1;
                for (int j = 0; j < s; ++j)
                    if (j != i) inv = orc_mulmod(inv, orc_invmod(c->q[start + j] % qi, qi), qi);
                c->partqhatinv[PQ_IDX(c, d, s - 1, i)] = inv;
                for (int t = 0; t < np; ++t) {
                    uint64_t pt = c->q[t], prod = 1;
                    for (int j = 0; j < s; ++j)
                        if (j != i) prod = orc_mulmod(prod, c->q[start + j] % pt, pt);
                    c->partqhat_modp[PQ_IDX(c, d, s - 1, i) * np + t] = prod;
                }
            }
        }
    }
    /* ModDown constants: PHatInvModp, PHatModq, PInvModq. */
    c->phatinv_modp = (uint64_t*)malloc(sizeof(uint64_t) * (K > 0 ? K : 1));
    c->phat_modq = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(K > 0 ? K : 1) * np);
    c->pinv_modq = (uint64_t*)malloc(sizeof(uint64_t) * np);
    c->pmodq = (uint64_t*)malloc(sizeof(uint64_t) * np);
    for (int k = 0; k < K; ++k) {
        uint64_t pk = c->q[L + 1 + k], inv = 1;
        for (int j = 0; j < K; ++j)
            if (j != k) inv = orc_mulmod(inv, orc_invmod(c->q[L + 1 + j] % pk, pk), pk);
        c->phatinv_modp[k] = inv;
        for (int t = 0; t < np; ++t) {
            uint64_t qt = c->q[t], prod = 1;
            for (int j = 0; j < K; ++j)
                if (j != k) prod = orc_mulmod(prod, c->q[L + 1 + j] % qt, qt);
            c->phat_modq[(size_t)k * np + t] = prod;
        }
    }
    for (int t = 0; t < np; ++t) {
        uint64_t qt = c->q[t], prod = 1;
        for (int j = 0; j < K; ++j) prod = orc_mulmod(prod, c->q[L + 1 + j] % qt, qt);
        c->pmodq[t] = prod;
        c->pinv_modq[t] = t <= L ? orc_invmod(prod, qt) : 0;
    }
    return c;
}

void orc_ctx_free(orc_ctx* c) {
    if (!c) return;
    for (int i = 0; i < c->nprimes; ++i) {
        free(c->tw[i]); free(c->twp[i]); free(c->itw[i]); free(c->itwp[i]);
    }
    free(c->tw); free(c->twp); free(c->itw); free(c->itwp);
    free(c->ninv); free(c->ninvp); free(c->q); free(c->root);
    free(c->partqhatinv); free(c->partqhat_modp);
    free(c->phatinv_modp); free(c->phat_modq); free(c->pinv_modq); free(c->pmodq);
    free(c);
}

uint64_t orc_const_partqhatinv(const orc_ctx* c, int d, int size, int i) { return c->partqhatinv[PQ_IDX(c, d, size - 1, i)]; }
uint64_t orc_const_partqhat_modp(const orc_ctx* c, int d, int size, int i, int t) {
    return c->partqhat_modp[PQ_IDX(c, d, size - 1, i) * c->nprimes + t];
}
uint64_t orc_const_phatinv_modp(const orc_ctx* c, int k) { return c->phatinv_modp[k]; }
uint64_t orc_const_phat_modq(const orc_ctx* c, int k, int i) { return c->phat_modq[(size_t)k * c->nprimes + i]; }
uint64_t orc_const_pinv_modq(const orc_ctx* c, int i) { return c->pinv_modq[i]; }

static inline void ntt_limb(const orc_ctx* c, uint64_t* a, int pid, int inverse) {
    if (inverse)
        ntt_inv_core(a, c->N, c->q[pid], c->itw[pid], c->itwp[pid], c->ninv[pid], c->ninvp[pid]);
    else
        ntt_fwd_core(a, c->N, c->q[pid], c->tw[pid], c->twp[pid]);
}

void orc_poly_ntt(const orc_ctx* c, uint64_t* a, int nlimbs, const int* primeids, int inverse) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) ntt_limb(c, a + (size_t)i * c->N, primeids[i], inverse);
}

/* ------------------------------------------------------------------ */
/* sampling                                                           */
/* ------------------------------------------------------------------ */

typedef struct { uint64_t s[4]; } rng_t;
static uint64_t splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t* r, uint64_t seed, uint64_t stream) {
    uint64_t x = seed * 0x100000001B3ull ^ (stream + 0x632BE59BD9B4E019ull);
    for (int i = 0; i < 4; ++i) r->s[i] = splitmix(&x);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t rng_next(rng_t* r) { /* xoshiro256** */
    uint64_t* s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}
static uint64_t rng_mod(rng_t* r, uint64_t q) {
    uint64_t mask = q - 1;
    mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16; mask |= mask >> 32;
    for (;;) {
        uint64_t v = rng_next(r) & mask;
        if (v < q) return v;
    }
}
static double rng_unif01(rng_t* r) { return ((rng_next(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

void orc_sample_ternary(uint64_t seed, int N, int8_t* s) {
    rng_t r;
    rng_seed(&r, seed, 1);
    for (int i = 0; i < N; ++i) s[i] = (int8_t)((int)rng_mod(&r, 3) - 1);
}

void orc_poly_from_int(const orc_ctx* c, const int64_t* coef, int nlimbs, const int* primeids, uint64_t* out) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[primeids[i]];
        uint64_t* o = out + (size_t)i * c->N;
        for (int k = 0; k < c->N; ++k) {
            int64_t v = coef[k];
            uint64_t m = v >= 0 ? ((uint64_t)v) % q : (q - ((uint64_t)(-(v + 1)) + 1) % q) % q;
            o[k] = m;
        }
        ntt_limb(c, o, primeids[i], 0);
    }
}

void orc_sample_uniform(const orc_ctx* c, uint64_t seed, int nlimbs, const int* primeids, uint64_t* out) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        rng_t r;
        rng_seed(&r, seed, 1000 + (uint64_t)primeids[i]);
        uint64_t q = c->q[primeids[i]];
        uint64_t* o = out + (size_t)i * c->N;
        for (int k = 0; k < c->N; ++k) o[k] = rng_mod(&r, q);
    }
}

void orc_sample_gauss(const orc_ctx* c, uint64_t seed, double sigma, int nlimbs, const int* primeids, uint64_t* out) {
    int64_t* e = (int64_t*)malloc(sizeof(int64_t) * c->N);
    rng_t r;
    rng_seed(&r, seed, 2);
    for (int k = 0; k < c->N; k += 2) { /* Box-Muller, rounded */
        double u1 = rng_unif01(&r), u2 = rng_unif01(&r);
        double rad = sqrt(-2.0 * log(u1)) * sigma;
        e[k] = (int64_t)llround(rad * cos(2 * M_PI * u2));
        if (k + 1 < c->N) e[k + 1] = (int64_t)llround(rad * sin(2 * M_PI * u2));
    }
    orc_poly_from_int(c, e, nlimbs, primeids, out);
    free(e);
}

/* ------------------------------------------------------------------ */
/* elementwise                                                        */
/* ------------------------------------------------------------------ */

void orc_add(const orc_ctx* c, int nlimbs, const int* pid, const uint64_t* a, const uint64_t* b, uint64_t* o) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[pid[i]];
        size_t off = (size_t)i * c->N;
        for (int k = 0; k < c->N; ++k) o[off + k] = addm(a[off + k], b[off + k], q);
    }
}
void orc_sub(const orc_ctx* c, int nlimbs, const int* pid, const uint64_t* a, const uint64_t* b, uint64_t* o) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[pid[i]];
        size_t off = (size_t)i * c->N;
        for (int k = 0; k < c->N; ++k) o[off + k] = subm(a[off + k], b[off + k], q);
    }
}
void orc_mul(const orc_ctx* c, int nlimbs, const int* pid, const uint64_t* a, const uint64_t* b, uint64_t* o) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[pid[i]];
        size_t off = (size_t)i * c->N;
        for (int k = 0; k < c->N; ++k) o[off + k] = orc_mulmod(a[off + k], b[off + k], q);
    }
}

/* Per-limb scalar ops (FIDESlib RNSPoly::multScalar / addScalar with the
 * residues of ContextData::ElemForEvalMult / ElemForEvalAddOrSub):
 * o_i = a_i * k_i or a_i + k_i mod q_i on the Q limbs 0..nlimbs-1. */
void orc_scalar_mul(const orc_ctx* c, int nlimbs, const uint64_t* k, const uint64_t* a, uint64_t* o) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[i];
        size_t off = (size_t)i * c->N;
        for (int j = 0; j < c->N; ++j) o[off + j] = orc_mulmod(a[off + j], k[i], q);
    }
}
void orc_scalar_add(const orc_ctx* c, int nlimbs, const uint64_t* k, const uint64_t* a, uint64_t* o) {
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        uint64_t q = c->q[i];
        size_t off = (size_t)i * c->N;
        for (int j = 0; j < c->N; ++j) o[off + j] = addm(a[off + j], k[i] % q, q);
    }
}

static int* qids(int n) {
    int* p = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) p[i] = i;
    return p;
}

/* ------------------------------------------------------------------ */
/* keys, encryption                                                   */
/* ------------------------------------------------------------------ */

/* Hybrid key generation (OpenFHE KeySwitchHYBRID::KeySwitchGenInternal, read
 * through FIDESlib's loader RawCiphertext.cu:468-495): for digit d,
 *   a_d uniform over Q u P,  b_d = -a_d s_new + e_d + [i in digit d] (P mod q_i) s_old. */
void orc_ksk_gen(const orc_ctx* c, uint64_t seed, const uint64_t* s_old, const uint64_t* s_new, uint64_t* a_out,
                 uint64_t* b_out) {
    int np = c->nprimes, N = c->N;
    int* all = qids(np);
    uint64_t* e = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)np * N);
    for (int d = 0; d < c->dnum; ++d) {
        uint64_t* a = a_out + (size_t)d * np * N;
        uint64_t* b = b_out + (size_t)d * np * N;
        orc_sample_uniform(c, seed * 131 + 7 * d + 1, np, all, a);
        orc_sample_gauss(c, seed * 131 + 7 * d + 2, 3.2, np, all, e);
        int start = d * c->alpha, end = start + c->alpha;
        if (end > c->L + 1) end = c->L + 1;
        OMP_FOR
        for (int i = 0; i < np; ++i) {
            uint64_t q = c->q[i];
            size_t off = (size_t)i * N;
            int in_digit = (i >= start && i < end);
            for (int k = 0; k < N; ++k) {
                uint64_t v = subm(e[off + k], orc_mulmod(a[off + k], s_new[off + k], q), q);
                if (in_digit) v = addm(v, orc_mulmod(c->pmodq[i], s_old[off + k], q), q);
                b[off + k] = v;
            }
        }
    }
    free(e);
    free(all);
}

void orc_encrypt_sk(const orc_ctx* c, uint64_t seed, int level, const uint64_t* s, const uint64_t* m, uint64_t* c0,
                    uint64_t* c1) {
    int n = level + 1, N = c->N;
    int* ids = qids(n);
    uint64_t* e = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n * N);
    orc_sample_uniform(c, seed * 977 + 11, n, ids, c1);
    orc_sample_gauss(c, seed * 977 + 12, 3.2, n, ids, e);
    OMP_FOR
    for (int i = 0; i < n; ++i) {
        uint64_t q = c->q[i];
        size_t off = (size_t)i * N;
        for (int k = 0; k < N; ++k)
            c0[off + k] = addm(subm(e[off + k], orc_mulmod(c1[off + k], s[off + k], q), q), m[off + k], q);
    }
    free(e);
    free(ids);
}

void orc_decrypt_coef(const orc_ctx* c, int level, const uint64_t* s, const uint64_t* c0, const uint64_t* c1,
                      uint64_t* m) {
    int n = level + 1, N = c->N;
    OMP_FOR
    for (int i = 0; i < n; ++i) {
        uint64_t q = c->q[i];
        size_t off = (size_t)i * N;
        for (int k = 0; k < N; ++k) m[off + k] = addm(c0[off + k], orc_mulmod(c1[off + k], s[off + k], q), q);
        ntt_limb(c, m + off, i, 1);
    }
}

void orc_crt_to_double(const orc_ctx* c, int level, const uint64_t* m, double* out) {
    int N = c->N;
    uint64_t q0 = c->q[0];
    if (level == 0) {
        for (int k = 0; k < N; ++k) {
            uint64_t v = m[k];
            out[k] = v > q0 / 2 ? -(double)(q0 - v) : (double)v;
        }
        return;
    }
    uint64_t q1 = c->q[1];
    uint64_t q0inv = orc_invmod(q0 % q1, q1);
    u128 Q = (u128)q0 * q1, half = Q / 2;
    for (int k = 0; k < N; ++k) {
        uint64_t x0 = m[k], x1 = m[N + k];
        uint64_t h = orc_mulmod(subm(x1, x0 % q1, q1), q0inv, q1);
        u128 x = (u128)x0 + (u128)q0 * h;
        out[k] = x > half ? -(double)(Q - x) : (double)x;
    }
}

/* ------------------------------------------------------------------ */
/* evaluation ops                                                     */
/* ------------------------------------------------------------------ */

/* (c0,c1) x (d0,d1) -> (c0d0, c0d1+c1d0, c1d1): ElemenwiseBatchKernels.cu:1398-1438 */
void orc_tensor(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* d0,
                const uint64_t* d1, uint64_t* o0, uint64_t* o1, uint64_t* o2) {
    int n = level + 1, N = c->N;
    OMP_FOR
    for (int i = 0; i < n; ++i) {
        uint64_t q = c->q[i];
        size_t off = (size_t)i * N;
        for (int k = 0; k < N; ++k) {
            size_t x = off + k;
            uint64_t a0 = c0[x], a1 = c1[x], b0 = d0[x], b1 = d1[x];
            o0[x] = orc_mulmod(a0, b0, q);
            o1[x] = addm(orc_mulmod(a0, b1, q), orc_mulmod(a1, b0, q), q);
            o2[x] = orc_mulmod(a1, b1, q);
        }
    }
}

int orc_num_digits(const orc_ctx* c, int level) {
    int n = (level + 1 + c->alpha - 1) / c->alpha;
    return n > c->dnum ? c->dnum : n;
}

/* Index of position `pos` of an extended poly at level l in the full prime list. */
static inline int ext_pid(const orc_ctx* c, int level, int pos) { return pos <= level ? pos : c->L + 1 + (pos - level - 1); }

/* ModUp (OpenFHE KeySwitchHYBRID::EvalKeySwitchPrecomputeCore + ApproxSwitchCRTBasis;
 * FIDESlib Conv.cu:144-287 DecompAndModUpConv and LimbPartitionMGPU.cu:805-1153):
 * for digit d with present primes D,
 *   t_i = [x_i * (D/q_i)^-1]_{q_i},   y_p = sum_i t_i [D/q_i]_p  mod p  (no centering). */
void orc_modup(const orc_ctx* c, int level, const uint64_t* x, uint64_t* ext) {
    int N = c->N, nd = orc_num_digits(c, level), next = level + 1 + c->K;
    int n = level + 1;
    uint64_t* coef = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n * N);
    memcpy(coef, x, sizeof(uint64_t) * (size_t)n * N);
    int* ids = qids(n);
    orc_poly_ntt(c, coef, n, ids, 1);
    for (int d = 0; d < nd; ++d) {
        int start = d * c->alpha, size = n - start < c->alpha ? n - start : c->alpha;
        uint64_t* out = ext + (size_t)d * next * N;
        /* t_i in place per digit (separate buffer) */
        uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)size * N);
        for (int i = 0; i < size; ++i) {
            uint64_t q = c->q[start + i], w = c->partqhatinv[PQ_IDX(c, d, size - 1, i)];
            uint64_t wp = shoup_pre(w, q);
            for (int k = 0; k < N; ++k) t[(size_t)i * N + k] = mul_shoup(coef[(size_t)(start + i) * N + k], w, wp, q);
        }
        OMP_FOR
        for (int pos = 0; pos < next; ++pos) {
            uint64_t* o = out + (size_t)pos * N;
            if (pos >= start && pos < start + size) {
                memcpy(o, x + (size_t)pos * N, sizeof(uint64_t) * N);
                continue;
            }
            int pid = ext_pid(c, level, pos);
            uint64_t p = c->q[pid];
            for (int k = 0; k < N; ++k) {
                u128 acc = 0;
                for (int i = 0; i < size; ++i)
                    acc += (u128)t[(size_t)i * N + k] * c->partqhat_modp[PQ_IDX(c, d, size - 1, i) * c->nprimes + pid];
                o[k] = (uint64_t)(acc % p);
            }
            ntt_limb(c, o, pid, 0);
        }
        free(t);
    }
    free(ids);
    free(coef);
}

/* KSK inner product (OpenFHE EvalFastKeySwitchCoreExt; FIDESlib
 * ElemenwiseBatchKernels.cu:221-297 fusedDotKSK_2_): key limb of position pos
 * is pos for Q limbs and L+1+(pos-l-1) for P limbs. */
void orc_ksk_inner(const orc_ctx* c, int level, const uint64_t* ext, const uint64_t* ka, const uint64_t* kb,
                   uint64_t* out_a, uint64_t* out_b) {
    int N = c->N, nd = orc_num_digits(c, level), next = level + 1 + c->K, np = c->nprimes;
    OMP_FOR
    for (int pos = 0; pos < next; ++pos) {
        int pid = ext_pid(c, level, pos);
        uint64_t p = c->q[pid];
        for (int k = 0; k < N; ++k) {
            uint64_t sa = 0, sb = 0;
            for (int d = 0; d < nd; ++d) {
                uint64_t v = ext[((size_t)d * next + pos) * N + k];
                sa = addm(sa, orc_mulmod(v, ka[((size_t)d * np + pid) * N + k], p), p);
                sb = addm(sb, orc_mulmod(v, kb[((size_t)d * np + pid) * N + k], p), p);
            }
            out_a[(size_t)pos * N + k] = sa;
            out_b[(size_t)pos * N + k] = sb;
        }
    }
}

/* ApproxModDown with t = 0 (OpenFHE DCRTPoly::ApproxModDown; FIDESlib
 * Conv.cu:26-132 ModDown2 and NTTfusions.cuh:244-262):
 *   y_i = (x_i - NTT(sum_k [x_k (P/p_k)^-1]_{p_k} [P/p_k]_{q_i})) * P^-1 mod q_i. */
void orc_moddown(const orc_ctx* c, int level, const uint64_t* x, uint64_t* y) {
    int N = c->N, n = level + 1, K = c->K;
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)K * N);
    OMP_FOR
    for (int k = 0; k < K; ++k) {
        int pid = c->L + 1 + k;
        uint64_t p = c->q[pid], w = c->phatinv_modp[k], wp = shoup_pre(w, p);
        uint64_t* tk = t + (size_t)k * N;
        memcpy(tk, x + (size_t)(n + k) * N, sizeof(uint64_t) * N);
        ntt_limb(c, tk, pid, 1);
        for (int j = 0; j < N; ++j) tk[j] = mul_shoup(tk[j], w, wp, p);
    }
    OMP_FOR
    for (int i = 0; i < n; ++i) {
        uint64_t q = c->q[i];
        uint64_t* yi = y + (size_t)i * N;
        for (int j = 0; j < N; ++j) {
            u128 acc = 0;
            for (int k = 0; k < K; ++k) acc += (u128)t[(size_t)k * N + j] * c->phat_modq[(size_t)k * c->nprimes + i];
            yi[j] = (uint64_t)(acc % q);
        }
        ntt_limb(c, yi, i, 0);
        uint64_t pinv = c->pinv_modq[i], pinvp = shoup_pre(pinv, q);
        const uint64_t* xi = x + (size_t)i * N;
        for (int j = 0; j < N; ++j) yi[j] = mul_shoup(subm(xi[j], yi[j], q), pinv, pinvp, q);
    }
    free(t);
}

void orc_keyswitch(const orc_ctx* c, int level, const uint64_t* x, const uint64_t* ka, const uint64_t* kb,
                   uint64_t* ks_b, uint64_t* ks_a) {
    int N = c->N, nd = orc_num_digits(c, level), next = level + 1 + c->K;
    uint64_t* ext = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nd * next * N);
    uint64_t* oa = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)next * N);
    uint64_t* ob = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)next * N);
    orc_modup(c, level, x, ext);
    orc_ksk_inner(c, level, ext, ka, kb, oa, ob);
    orc_moddown(c, level, ob, ks_b);
    orc_moddown(c, level, oa, ks_a);
    free(ext); free(oa); free(ob);
}

/* EvalMult = tensor + relinearize (LeveledSHEBase::EvalMult; FIDESlib
 * Ciphertext.cpp:409-622): o0 = c0d0 + ks_b(c1d1), o1 = c0d1 + c1d0 + ks_a(c1d1). */
void orc_eval_mult(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* d0,
                   const uint64_t* d1, const uint64_t* ka, const uint64_t* kb, uint64_t* o0, uint64_t* o1) {
    int n = level + 1, N = c->N;
    size_t sz = sizeof(uint64_t) * (size_t)n * N;
    uint64_t *t2 = (uint64_t*)malloc(sz), *kb_ = (uint64_t*)malloc(sz), *ka_ = (uint64_t*)malloc(sz);
    int* ids = qids(n);
    orc_tensor(c, level, c0, c1, d0, d1, o0, o1, t2);
    orc_keyswitch(c, level, t2, ka, kb, kb_, ka_);
    orc_add(c, n, ids, o0, kb_, o0);
    orc_add(c, n, ids, o1, ka_, o1);
    free(t2); free(kb_); free(ka_); free(ids);
}

void orc_eval_square(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* ka,
                     const uint64_t* kb, uint64_t* o0, uint64_t* o1) {
    orc_eval_mult(c, level, c0, c1, c0, c1, ka, kb, o0, o1);
}

/* Rescale (OpenFHE DCRTPoly::DropLastElementAndScale; FIDESlib
 * LimbPartitionMGPU.cu:214-394, Rescale.cuh:45-65, NTTfusions.cuh:218-242):
 *   y_i = x_i q_l^-1 + NTT(Switch_{q_l->q_i}(INTT x_l)) * QlQlInvModqlDivqlModq_i
 * with Switch the centered lift (v > floor(q_l/2) -> v - q_l), and
 * QlQlInvModqlDivqlModq_i == -q_l^-1 mod q_i, so y_i = (x_i - NTT(Switch)) q_l^-1. */
void orc_rescale(const orc_ctx* c, int level, const uint64_t* x, uint64_t* y) {
    int N = c->N;
    uint64_t ql = c->q[level], half = ql >> 1;
    uint64_t* top = (uint64_t*)malloc(sizeof(uint64_t) * N);
    memcpy(top, x + (size_t)level * N, sizeof(uint64_t) * N);
    ntt_limb(c, top, level, 1);
    OMP_FOR
    for (int i = 0; i < level; ++i) {
        uint64_t q = c->q[i];
        uint64_t* yi = y + (size_t)i * N;
        uint64_t qlmod = ql % q;
        for (int j = 0; j < N; ++j) {
            uint64_t v = top[j];
            uint64_t r = v % q;
            if (v > half) r = subm(r, qlmod, q); /* v - q_l mod q */
            yi[j] = r;
        }
        ntt_limb(c, yi, i, 0);
        uint64_t qinv = orc_invmod(ql % q, q), qinvp = shoup_pre(qinv, q);
        const uint64_t* xi = x + (size_t)i * N;
        for (int j = 0; j < N; ++j) yi[j] = mul_shoup(subm(xi[j], yi[j], q), qinv, qinvp, q);
    }
    free(top);
}

/* Automorphism on evaluation form (OpenFHE PrecomputeAutoMap +
 * AutomorphismTransform; FIDESlib Rotation.cuh:14-36):
 *   out[brev(j)] = in[brev(((2j+1) k mod 2N) >> 1)]. */
void orc_automorph(const orc_ctx* c, int nlimbs, uint64_t k, const uint64_t* in, uint64_t* out) {
    int N = c->N, logN = c->logN;
    uint32_t* map = (uint32_t*)malloc(sizeof(uint32_t) * N);
    uint64_t mask = 2ull * N - 1;
    for (int j = 0; j < N; ++j) {
        uint64_t idx = (((2ull * j + 1) * k) & mask) >> 1;
        map[brev(j, logN)] = brev((uint32_t)idx, logN);
    }
    OMP_FOR
    for (int i = 0; i < nlimbs; ++i) {
        const uint64_t* a = in + (size_t)i * N;
        uint64_t* o = out + (size_t)i * N;
        for (int j = 0; j < N; ++j) o[j] = a[map[j]];
    }
    free(map);
}

/* FindAutomorphismIndex2nComplex: k = 5^r mod 2N, r reduced mod N/2 slots. */
uint64_t orc_rot_index(int logN, int r) {
    int64_t half = 1ll << (logN - 1);
    int64_t rr = r % half;
    if (rr < 0) rr += half;
    return orc_powmod(5, (uint64_t)rr, 2ull << logN);
}

/* EvalRotate = key switch of c1, add c0, then automorphism (OpenFHE
 * LeveledSHEBase::EvalAutomorphism order; FIDESlib's non-fused hoisted path
 * Ciphertext.cpp:1080-1128 uses the same order). */
void orc_rotate(const orc_ctx* c, int level, uint64_t k, const uint64_t* c0, const uint64_t* c1, const uint64_t* ka,
                const uint64_t* kb, uint64_t* o0, uint64_t* o1) {
    int n = level + 1, N = c->N;
    size_t sz = sizeof(uint64_t) * (size_t)n * N;
    uint64_t *kb_ = (uint64_t*)malloc(sz), *ka_ = (uint64_t*)malloc(sz);
    int* ids = qids(n);
    orc_keyswitch(c, level, c1, ka, kb, kb_, ka_);
    orc_add(c, n, ids, kb_, c0, kb_);
    orc_automorph(c, n, k, kb_, o0);
    orc_automorph(c, n, k, ka_, o1);
    free(kb_); free(ka_); free(ids);
}

/* EvalFastRotationPrecompute + EvalFastRotation: digits of c1 computed once. */
void orc_rotate_hoisted(const orc_ctx* c, int level, int nrot, const uint64_t* ks, const uint64_t* c0,
                        const uint64_t* c1, const uint64_t* const* ka, const uint64_t* const* kb, uint64_t* const* o0,
                        uint64_t* const* o1) {
    int n = level + 1, N = c->N, nd = orc_num_digits(c, level), next = n + c->K;
    uint64_t* ext = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nd * next * N);
    uint64_t* oa = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)next * N);
    uint64_t* ob = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)next * N);
    size_t sz = sizeof(uint64_t) * (size_t)n * N;
    uint64_t *kb_ = (uint64_t*)malloc(sz), *ka_ = (uint64_t*)malloc(sz);
    int* ids = qids(n);
    orc_modup(c, level, c1, ext);
    for (int r = 0; r < nrot; ++r) {
        orc_ksk_inner(c, level, ext, ka[r], kb[r], oa, ob);
        orc_moddown(c, level, ob, kb_);
        orc_moddown(c, level, oa, ka_);
        orc_add(c, n, ids, kb_, c0, kb_);
        orc_automorph(c, n, ks[r], kb_, o0[r]);
        orc_automorph(c, n, ks[r], ka_, o1[r]);
    }
    free(ext); free(oa); free(ob); free(kb_); free(ka_); free(ids);
}
