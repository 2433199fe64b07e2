/*
 * CPU oracle for the CKKS evaluation path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C restatement of the CPU algorithms that FIDESlib delegates to
 * OpenFHE v1.4.2 and that FIDESlib's GPU kernels re-implement.  It is the
 * checker for the HIP engine in pyfideslib_amd/: only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it.  The product never
 * links, calls or falls back to it.
 *
 * Conventions
 *   - A prime id indexes ctx->q: ids 0..L are the Q chain (q0 first), ids
 *     L+1..L+K are the special primes P.
 *   - A polynomial is a row-major uint64 array [limb][N].  "Evaluation form" is
 *     OpenFHE's bit-reversed NTT order: a_eval[j] = a(psi^(2*brev(j)+1)).
 *   - An extended polynomial at level l has l+1 Q limbs followed by K P limbs.
 *
 * Parity status of each function is recorded in the header comment of
 * oracle.c; every function there is pinned by a test in tests/ against the
 * reference's known-answer vectors or a closed mathematical form.
 */
#ifndef FIDES_ORACLE_H
#define FIDES_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------- scalar number theory ---------- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t orc_invmod(uint64_t a, uint64_t p);
int orc_is_prime(uint64_t n);
/* Largest `count` primes below 2^bits with p == 1 (mod 2^(logN+1)), in
 * decreasing order, skipping any listed in `exclude`.  Returns count found. */
int orc_gen_primes(int logN, int bits, int count, const uint64_t* exclude, int n_exclude, uint64_t* out);
/* Minimal primitive m-th root of unity mod q (m a power of two dividing q-1). */
uint64_t orc_min_root(uint64_t q, uint64_t m);

/* ---------- single-limb NTT ---------- */
/* psi_rev[k] = psi^brev(k), psi_inv_rev[k] = psi^-brev(k)  (k < N) */
void orc_ntt_tables(int logN, uint64_t q, uint64_t psi, uint64_t* psi_rev, uint64_t* psi_inv_rev,
                    uint64_t* n_inv);
void orc_ntt_fwd(uint64_t* a, int logN, uint64_t q, const uint64_t* psi_rev);
void orc_ntt_inv(uint64_t* a, int logN, uint64_t q, const uint64_t* psi_inv_rev, uint64_t n_inv);
/* O(N^2) evaluation a_eval[j] = sum_k a[k] psi^((2 brev(j)+1) k): pins the fast NTT. */
void orc_ntt_bruteforce(const uint64_t* a, uint64_t* out, int logN, uint64_t q, uint64_t psi);

/* ---------- CKKS context ---------- */
typedef struct orc_ctx orc_ctx;
/* moduli/roots: L+1 Q primes followed by K special primes.  roots[i] is the
 * primitive 2N-th root psi_i used by the NTT for prime i. */
orc_ctx* orc_ctx_create(int logN, int L, int K, int dnum, const uint64_t* moduli, const uint64_t* roots);
void orc_ctx_free(orc_ctx* c);
int orc_ctx_alpha(const orc_ctx* c);
void orc_set_threads(int n);
int orc_get_threads(void);

/* NTT over `nlimbs` limbs whose prime ids are primeids[i]. */
void orc_poly_ntt(const orc_ctx* c, uint64_t* a, int nlimbs, const int* primeids, int inverse);

/* ---------- sampling (own seeded PRNG; not OpenFHE's stream) ---------- */
/* Ternary secret s in {-1,0,1}^N (uniform), written as int8 coefficients. */
void orc_sample_ternary(uint64_t seed, int N, int8_t* s);
/* Evaluation form of an integer polynomial (int64 coefficients) over limbs
 * with the given prime ids. */
void orc_poly_from_int(const orc_ctx* c, const int64_t* coef, int nlimbs, const int* primeids, uint64_t* out);
void orc_sample_uniform(const orc_ctx* c, uint64_t seed, int nlimbs, const int* primeids, uint64_t* out);
/* Discrete Gaussian error (rounded normal, sigma), evaluation form. */
void orc_sample_gauss(const orc_ctx* c, uint64_t seed, double sigma, int nlimbs, const int* primeids, uint64_t* out);

/* ---------- keys ---------- */
/* Hybrid key-switching key from s_old to s_new (both evaluation form over all
 * L+1+K primes).  a_out/b_out: [dnum][L+1+K][N]. */
void orc_ksk_gen(const orc_ctx* c, uint64_t seed, const uint64_t* s_old, const uint64_t* s_new, uint64_t* a_out,
                 uint64_t* b_out);
/* Secret-key encryption of an evaluation-form plaintext at level l. */
void orc_encrypt_sk(const orc_ctx* c, uint64_t seed, int level, const uint64_t* s, const uint64_t* m,
                    uint64_t* c0, uint64_t* c1);
/* m = c0 + c1 s at level l, returned in coefficient form (l+1 limbs). */
void orc_decrypt_coef(const orc_ctx* c, int level, const uint64_t* s, const uint64_t* c0, const uint64_t* c1,
                      uint64_t* m_coef);
/* Centered CRT lift of coefficient-form residues over limbs 0..min(level,1),
 * as double.  Exact while |value| < q0*q1/2. */
void orc_crt_to_double(const orc_ctx* c, int level, const uint64_t* m_coef, double* out);

/* ---------- evaluation ops (all evaluation form) ---------- */
void orc_tensor(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* d0,
                const uint64_t* d1, uint64_t* o0, uint64_t* o1, uint64_t* o2);
int orc_num_digits(const orc_ctx* c, int level);
/* ModUp of x (level l): ext[d][l+1+K][N] for d < orc_num_digits. */
void orc_modup(const orc_ctx* c, int level, const uint64_t* x, uint64_t* ext);
/* KSK inner product: out_a/out_b [l+1+K][N]. */
void orc_ksk_inner(const orc_ctx* c, int level, const uint64_t* ext, const uint64_t* ka, const uint64_t* kb,
                   uint64_t* out_a, uint64_t* out_b);
/* ModDown of an extended poly at level l to l+1 limbs. */
void orc_moddown(const orc_ctx* c, int level, const uint64_t* x, uint64_t* y);
/* Full key switch of x: (ks_b, ks_a) = ModDown(<ModUp(x), (b, a)>). */
void orc_keyswitch(const orc_ctx* c, int level, const uint64_t* x, const uint64_t* ka, const uint64_t* kb,
                   uint64_t* ks_b, uint64_t* ks_a);
void orc_eval_mult(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* d0,
                   const uint64_t* d1, const uint64_t* ka, const uint64_t* kb, uint64_t* o0, uint64_t* o1);
void orc_eval_square(const orc_ctx* c, int level, const uint64_t* c0, const uint64_t* c1, const uint64_t* ka,
                     const uint64_t* kb, uint64_t* o0, uint64_t* o1);
/* Rescale one polynomial from level l to l-1 (output l limbs). */
void orc_rescale(const orc_ctx* c, int level, const uint64_t* x, uint64_t* y);
/* Automorphism X -> X^k on nlimbs evaluation-form limbs. */
void orc_automorph(const orc_ctx* c, int nlimbs, uint64_t k, const uint64_t* in, uint64_t* out);
/* Automorphism index for a left rotation by r slots (5^r mod 2N, r may be < 0). */
uint64_t orc_rot_index(int logN, int r);
void orc_rotate(const orc_ctx* c, int level, uint64_t k, const uint64_t* c0, const uint64_t* c1,
                const uint64_t* ka, const uint64_t* kb, uint64_t* o0, uint64_t* o1);
/* Hoisted rotation set: one ModUp of c1 shared by n rotations.
 * ka/kb: n pointers to keys; o0/o1: n output buffers. */
void orc_rotate_hoisted(const orc_ctx* c, int level, int n, const uint64_t* ks, const uint64_t* c0,
                        const uint64_t* c1, const uint64_t* const* ka, const uint64_t* const* kb, uint64_t* const* o0,
                        uint64_t* const* o1);

/* ---------- elementwise helpers ---------- */
void orc_add(const orc_ctx* c, int nlimbs, const int* primeids, const uint64_t* a, const uint64_t* b, uint64_t* o);
void orc_sub(const orc_ctx* c, int nlimbs, const int* primeids, const uint64_t* a, const uint64_t* b, uint64_t* o);
void orc_mul(const orc_ctx* c, int nlimbs, const int* primeids, const uint64_t* a, const uint64_t* b, uint64_t* o);
void orc_scalar_mul(const orc_ctx* c, int nlimbs, const uint64_t* k, const uint64_t* a, uint64_t* o);
void orc_scalar_add(const orc_ctx* c, int nlimbs, const uint64_t* k, const uint64_t* a, uint64_t* o);

/* ---------- precomputed constants (exposed so tests can pin them) ---------- */
/* [(Qpart/q_i)^-1]_{q_i} for digit d with `size` present primes. */
uint64_t orc_const_partqhatinv(const orc_ctx* c, int d, int size, int i);
/* [Qpart/q_i]_{p} for digit d, present size `size`, target prime id p. */
uint64_t orc_const_partqhat_modp(const orc_ctx* c, int d, int size, int i, int target_pid);
uint64_t orc_const_phatinv_modp(const orc_ctx* c, int k);
uint64_t orc_const_phat_modq(const orc_ctx* c, int k, int i);
uint64_t orc_const_pinv_modq(const orc_ctx* c, int i);

#ifdef __cplusplus
}
#endif
#endif
