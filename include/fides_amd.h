/*
 * fides_amd.h -- C ABI of the MI355X (gfx950) CKKS evaluation engine.
 *
 * This is the drop-in boundary for FIDESlib's GPU path: the engine takes the
 * raw RNS data FIDESlib extracts from OpenFHE (RawParams, RawCipherText,
 * RawKeySwitchKey -- src/CKKS/openfhe-interface/RawCiphertext.cuh:19-93) and
 * runs the evaluation ops of FIDESlib::CKKS::Ciphertext on one GPU.
 *
 * Conventions
 *   - Every function returns FIDES_OK (0) or a positive FIDES_E_* code; the
 *     message of the last failure on the calling thread is fides_last_error().
 *     The library never exits the process (FIDESlib's CudaCheckErrorMod does,
 *     src/CudaUtils.cuh:53-89).
 *   - Polynomials are uint64 arrays [limb][N] in OpenFHE EVALUATION format
 *     (bit-reversed NTT order), limb i reduced modulo prime i.
 *   - Prime ids: 0..L are the Q chain (moduli[]), L+1..L+K the special primes.
 *   - Keys are [dnum][L+1+K][N] per component (RawKeySwitchKey r_key[0] = A,
 *     r_key[1] = B).
 *   - All device work of a context runs in order on one HIP stream
 *     (fides_context_stream).  Host buffers passed to *_load / *_store are
 *     copied synchronously; ops are asynchronous until fides_synchronize or a
 *     store.
 *   - Threads: every call that takes a context (or a server) holds that
 *     context's lock for the call, so calls on one context from several
 *     threads run one at a time and calls on different contexts run
 *     concurrently.  A server is driven from one thread.
 *   - Destroy ciphertexts before the context that created them.
 */
#ifndef FIDES_AMD_H
#define FIDES_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIDES_OK 0
#define FIDES_E_ARG 1     /* invalid argument */
#define FIDES_E_HIP 2     /* HIP runtime error */
#define FIDES_E_STATE 3   /* object in the wrong state (e.g. not loaded, wrong noise level) */
#define FIDES_E_NOKEY 4   /* required key not loaded */
#define FIDES_E_LEVEL 5   /* level out of range */
#define FIDES_E_OOM 6     /* device allocation failed */

/* ScalingTechnique values (lbcrypto::ScalingTechnique order). */
#define FIDES_FIXEDMANUAL 0
#define FIDES_FIXEDAUTO 1
#define FIDES_FLEXIBLEAUTO 2
#define FIDES_FLEXIBLEAUTOEXT 3

typedef struct fides_context fides_context;
typedef struct fides_ciphertext fides_ciphertext;
typedef struct fides_plaintext fides_plaintext;
typedef struct fides_server fides_server;

/* The fields of FIDESlib::CKKS::RawParams (RawCiphertext.cuh:51-92) the engine
 * reads.  Twiddle tables and CRT constants are derived from moduli + roots. */
typedef struct {
    int32_t logN;                         /* ring dimension N = 2^logN, 10..17 */
    int32_t L;                            /* Q chain has L+1 primes */
    int32_t K;                            /* number of special primes */
    int32_t dnum;                         /* key-switching digits */
    const uint64_t* moduli;               /* [L+1] RawParams::moduli */
    const uint64_t* root_of_unity;        /* [L+1] primitive 2N-th roots (RawParams::root_of_unity) */
    const uint64_t* special_moduli;       /* [K]   RawParams::SPECIALmoduli */
    const uint64_t* special_root_of_unity;/* [K]   RawParams::SPECIALroot_of_unity */
    int32_t scaling_technique;            /* FIDES_FIXEDMANUAL, ... */
    const double* mod_reduce_factor;      /* [L+1] RawParams::ModReduceFactor, or NULL (= moduli) */
    const double* scaling_factor_real;    /* [L+1] RawParams::ScalingFactorReal, or NULL (no scalar ops) */
    const double* scaling_factor_real_big;/* [L+1] RawParams::ScalingFactorRealBig, or NULL (FLEXIBLEAUTO
                                           * alignment of a noise-degree-1 operand two or more levels up
                                           * then fails with FIDES_E_STATE) */
} fides_raw_params;

int32_t fides_abi_version(void);
const char* fides_last_error(void);

/* Context: replaces FIDESlib::CKKS::GenCryptoContextGPU(Parameters, devices)
 * (src/CKKS/Context.cu:969-987) for one device. */
int fides_context_create(const fides_raw_params* params, int32_t device, fides_context** out);
int fides_context_destroy(fides_context* ctx);
/* Per-op launch sequences captured into HIP graphs on a signature's second
 * use and replayed after (on != 0), or launched directly each time (0, the
 * default: replay measured 2-4% slower for single ops on MI355X; batched and
 * hoisted plans are captured either way; FIDES_OP_GRAPHS=1 sets it for every
 * context).  No reference counterpart (FIDESlib has no graph capture). */
int fides_context_set_op_graphs(fides_context* ctx, int32_t on);
int fides_synchronize(fides_context* ctx);
void* fides_context_stream(fides_context* ctx);
/* Bytes held by the context's device pool (ciphertexts, keys, scratch). */
int fides_context_pool_bytes(fides_context* ctx, uint64_t* bytes);
/* Gives cached, unused device blocks of the pool back to the device until at
 * most keep_bytes stay cached (waits for the device first).  No FIDESlib
 * counterpart: its limbs are allocated and freed one by one. */
int fides_context_trim_pool(fides_context* ctx, uint64_t keep_bytes);

/* Device memory on the context's device, for callers without a HIP runtime of
 * their own (tests, benchmarks).  Copies are ordered on the context stream and
 * complete before the call returns. */
int fides_device_alloc(fides_context* ctx, uint64_t bytes, void** ptr);
int fides_device_free(fides_context* ctx, void* ptr);
int fides_memcpy_htod(fides_context* ctx, void* dst, const void* src, uint64_t bytes);
int fides_memcpy_dtoh(fides_context* ctx, void* dst, const void* src, uint64_t bytes);

/* HIP events on the context stream, for timing the engine's own work. */
int fides_event_create(fides_context* ctx, void** ev);
int fides_event_destroy(void* ev);
int fides_event_record(fides_context* ctx, void* ev);
/* Milliseconds between two recorded events (waits for `stop`). */
int fides_event_elapsed_ms(void* start, void* stop, float* ms);

/* Ciphertexts: replace CKKS::Ciphertext::load / store with RawCipherText
 * (src/CKKS/Ciphertext.cpp:290-309, RawCiphertext.cu:48-132). */
int fides_ct_create(fides_context* ctx, fides_ciphertext** out);
int fides_ct_destroy(fides_ciphertext* ct);
int fides_ct_load(fides_context* ctx, fides_ciphertext* ct, const uint64_t* sub_0, const uint64_t* sub_1,
                  int32_t num_res, double noise, int32_t noise_level, int32_t slots);
/* Same as fides_ct_load with device-resident sub_0 / sub_1 (no host copy). */
int fides_ct_load_device(fides_context* ctx, fides_ciphertext* ct, const uint64_t* sub_0, const uint64_t* sub_1,
                         int32_t num_res, double noise, int32_t noise_level, int32_t slots);
/* sub_0 / sub_1 must hold (level+1)*N words each. */
int fides_ct_store(fides_context* ctx, const fides_ciphertext* ct, uint64_t* sub_0, uint64_t* sub_1);
int fides_ct_info(const fides_ciphertext* ct, int32_t* level, double* noise, int32_t* noise_level, int32_t* slots);
int fides_ct_copy(fides_context* ctx, const fides_ciphertext* src, fides_ciphertext* dst);
/* Device pointers of c0 and c1 (limb 0); valid while ct lives. */
int fides_ct_device_ptrs(const fides_ciphertext* ct, uint64_t** c0, uint64_t** c1);

/* Keys: replace KeySwitchingKey::Initialize(RawKeySwitchKey) and
 * Context::AddRotationKey (src/CKKS/KeySwitchingKey.cu:20-35, Context.cu:589-613). */
int fides_key_load_eval(fides_context* ctx, const uint64_t* a, const uint64_t* b);
/* Rotation key for a left rotation by `index` slots (automorphism 5^index mod 2N). */
int fides_key_load_rotation(fides_context* ctx, int32_t index, const uint64_t* a, const uint64_t* b);
/* Rotation key addressed by its automorphism index k (odd, < 2N). */
int fides_key_load_automorph(fides_context* ctx, uint32_t k, const uint64_t* a, const uint64_t* b);
int fides_rotation_automorph_index(fides_context* ctx, int32_t index, uint32_t* k);

/* Evaluation ops; `out` may alias an input unless stated. */
/* CKKS::Ciphertext::mult (src/CKKS/Ciphertext.cpp:409-622): tensor + relinearize,
 * then a rescale when rescale != 0. */
int fides_eval_mult(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out,
                    int32_t rescale);
int fides_eval_square(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out, int32_t rescale);
/* CKKS::Ciphertext::rescale (Ciphertext.cpp:379-398), in place. */
int fides_rescale(fides_context* ctx, fides_ciphertext* ct);
int fides_eval_add(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out);
int fides_eval_sub(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out);
/* CKKS::Ciphertext::rotate (Ciphertext.cpp:756-926), OpenFHE operation order. */
int fides_eval_rotate(fides_context* ctx, const fides_ciphertext* a, int32_t index, fides_ciphertext* out);
/* CKKS::Ciphertext::rotate_hoisted (Ciphertext.cpp:1053-1226): one ModUp shared
 * by n rotations; out[i] must not alias a. */
int fides_eval_rotate_hoisted(fides_context* ctx, const fides_ciphertext* a, const int32_t* indices, int32_t n,
                              fides_ciphertext* const* out);
/* Batched HMult+relin(+rescale) over n independent pairs at arbitrary levels,
 * one launch per stage for the whole batch (no counterpart in FIDESlib, whose
 * batched kernels, LimbPartitionBatch.cu:21-604, have no batched key switch). */
int fides_eval_mult_batch(fides_context* ctx, const fides_ciphertext* const* a, const fides_ciphertext* const* b,
                          fides_ciphertext* const* out, int32_t n, int32_t rescale);

/* Batched rotations: out[i] = rotate(a[i], indices[i]) for n independent
 * ciphertexts at arbitrary levels, each with its own rotation key, one launch
 * per stage (FIDESlib has no batched key switch); out[i] must not alias any
 * a[j]. */
int fides_eval_rotate_batch(fides_context* ctx, const fides_ciphertext* const* a, const int32_t* indices,
                            fides_ciphertext* const* out, int32_t n);

/* Plaintexts: replace CKKS::Plaintext(Context&, RawPlainText) (Plaintext.cuh:75-91)
 * -- one polynomial [num_res][N] in evaluation form, with its scale and
 * noise degree (RawPlainText::Noise / NoiseLevel). */
int fides_pt_create(fides_context* ctx, fides_plaintext** out);
int fides_pt_destroy(fides_plaintext* pt);
int fides_pt_load(fides_context* ctx, fides_plaintext* pt, const uint64_t* limbs, int32_t num_res, double noise,
                  int32_t noise_level, int32_t slots);
/* A plaintext in P*Q form: RawPlainText whose sub_0 also holds the special-
 * prime limbs (RNSPoly::loadConstant, RNSPoly.cpp:815-853, marks it modUp):
 * limbs [num_res + K][N], the Q limbs 0..num_res-1 then the K special limbs.
 * Linear transforms whose diagonals are all in this form (and bstep > 1) run
 * the extended-basis path, as LinearTransform.cu:93-104 decides. */
int fides_pt_load_ext(fides_context* ctx, fides_plaintext* pt, const uint64_t* limbs, int32_t num_res, double noise,
                      int32_t noise_level, int32_t slots);
/* CKKS::Ciphertext::multPt (Ciphertext.cpp:329-377): out = a * pt on a's
 * limbs (pt may hold more), then a rescale when rescale != 0 (FIXEDMANUAL). */
int fides_eval_mult_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                          fides_ciphertext* out, int32_t rescale);
/* CKKS::Ciphertext::addPt (Ciphertext.cpp:236-262): out = a + pt (c0 only). */
int fides_eval_add_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                         fides_ciphertext* out);
/* CKKS::Ciphertext::subPt (Ciphertext.cpp:252-275): out = a - pt (c0 only). */
int fides_eval_sub_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                         fides_ciphertext* out);
/* CKKS::Ciphertext::addMultPt (Ciphertext.cpp:1410-1426): acc += a * pt in
 * place (acc at noise level 2, a and pt at 1), then a rescale when rescale
 * != 0 (FIXEDMANUAL). */
int fides_eval_add_mult_plain(fides_context* ctx, fides_ciphertext* acc, const fides_ciphertext* a,
                              const fides_plaintext* pt, int32_t rescale);
/* CKKS::Ciphertext::addMultScalar (Ciphertext.cpp:1338-1358): acc += b * d in
 * place, d encoded by ElemForEvalMult at acc's level (acc at noise level 2, b
 * at 1). */
int fides_eval_add_mult_scalar(fides_context* ctx, fides_ciphertext* acc, const fides_ciphertext* b, double d);
/* CKKS::Ciphertext::evalLinearWSumMutable (Ciphertext.cpp:1272-1335): out =
 * sum_i w[i] a[i] in one launch, inputs at noise level 1 and one level; out at
 * noise level 2 with scale ScalingFactorReal[level]^2; out may alias an a[i]. */
int fides_eval_linear_wsum(fides_context* ctx, const fides_ciphertext* const* a, const double* w, int32_t n,
                           fides_ciphertext* out);
/* CKKS::Ciphertext::multScalar (Ciphertext.cpp:687-713) with the encoding of
 * ContextData::ElemForEvalMult (Context.cu:375-455); needs scaling_factor_real. */
int fides_eval_mult_scalar(fides_context* ctx, const fides_ciphertext* a, double v, fides_ciphertext* out,
                           int32_t rescale);
/* CKKS::Ciphertext::addScalar (Ciphertext.cpp:715-735) with the encoding of
 * ContextData::ElemForEvalAddOrSub (Context.cu:471-564). */
int fides_eval_add_scalar(fides_context* ctx, const fides_ciphertext* a, double v, fides_ciphertext* out);
/* CKKS::Ciphertext::dropToLevel (Ciphertext.cuh:401), in place. */
int fides_drop_to_level(fides_context* ctx, fides_ciphertext* ct, int32_t level);
/* The per-limb residues of those scalar encodings at `level` (kind 0:
 * ElemForEvalMult(v); kind 1: ElemForEvalAddOrSub(|v|, noise_deg)): out[level+1]. */
int fides_scalar_encoding(fides_context* ctx, int32_t level, double v, int32_t kind, int32_t noise_deg,
                          uint64_t* out);

/* Baby-step / giant-step linear transform, CKKS::LinearTransform(ct, rowSize = n,
 * bStep, pts, stride, offset) (src/CKKS/LinearTransform.cu:67-258):
 *   out = rot_offset( sum_{j<g} rot_{j*bstep*stride}( sum_{i<bstep} diag[bstep*j+i] * rot_{i*stride}(a) ) )
 * with g = ceil(n / bstep).  By default the giant steps are applied Horner
 * style with the rotation by bstep*stride, as the reference does; with
 * FIDES_LT_BATCHED_GIANT in flags the g-1 giant-step rotations run as one
 * batched key switch (needs the keys of every j*bstep*stride).  diag[] are
 * noise-level-1 plaintexts (pre-rotated by the caller, as OpenFHE's
 * precomputation does); needs the rotation keys of i*stride (i < bstep) and
 * offset.  out keeps noise level 2.  When bstep > 1 and every diagonal is a
 * P*Q plaintext (fides_pt_load_ext / fides_encode_ext) the transform runs the
 * reference's extended-basis path (LinearTransform.cu:82-193, ext, MODDOWN_HOIST,
 * ONLY_C1): baby steps and inner sums in P*Q, only c1 brought down per giant
 * step, one final ModDown -- different residues from the Q-basis path, the
 * same message. */
#define FIDES_LT_BATCHED_GIANT 1
int fides_eval_linear_transform(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* const* diag,
                                int32_t n, int32_t bstep, int32_t stride, int32_t offset, int32_t flags,
                                fides_ciphertext* out);
/* The same transform of m ciphertexts at one level (no FIDESlib counterpart;
 * its LTdotProductPtBatch, LimbPartitionBatch.cu, batches the inner products
 * only): hoisted baby steps per ciphertext, the inner sums of every giant step
 * of every ciphertext in one launch, every giant step (and the offset) one
 * batched rotation over all m; each output equals the single transform bit for
 * bit.  Outputs must be distinct and not alias another item's input. */
int fides_eval_linear_transform_batch(fides_context* ctx, const fides_ciphertext* const* a, int32_t m,
                                      const fides_plaintext* const* diag, int32_t n, int32_t bstep, int32_t stride,
                                      int32_t offset, int32_t flags, fides_ciphertext* const* out);

/* ConvolutionTransform(ctxt, rowSize = n, bStep, pts, stride, indexes, gStep)
 * (FIDESlib CKKS::ConvolutionTransform, src/CKKS/LinearTransform.cu:255-441;
 * OpenFHE-side ConvolutionTransformInPlace, api/CryptoContext.cpp:1702-1730):
 * baby steps v_i = rot_(idx[i])(a) (one hoisted set, bstep indexes), inner
 * sums u_j = sum_i pts[bstep j + i] * v_i, giant steps in blocks of 8 with
 * u_j of a block of G giant steps rotated by stride (G - j mod 8) and block k
 * of nb rotated by 8 stride (nb - 1 - k), all summed.  1 <= n <= bstep gstep;
 * a degree-2 input is rescaled first; out keeps noise level 2 and may alias a.
 * This is the reference's Q-basis path (its ext = false case: bStep = 1 or
 * diagonals not in P*Q); P*Q plaintexts contribute their Q limbs.
 * fides_convolution_rotations lists the rotations it needs besides idx
 * (GetConvolutionTransformRotationIndices, LinearTransform.cu:655-678, plus
 * stride min(gstep, 8), which the first giant step of a block needs): n gets
 * the count, out (if not NULL) the first min(count, cap). */
int fides_convolution_rotations(int32_t stride, int32_t gstep, int32_t* out, int32_t cap, int32_t* n);
int fides_eval_convolution_transform(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* const* pts,
                                     int32_t n, int32_t bstep, int32_t stride, const int32_t* idx, int32_t gstep,
                                     fides_ciphertext* out);

/* Ciphertexts in P*Q form (the extended basis; FIDESlib RNSPoly::isModUp on
 * both parts): the K special-prime limbs follow limb `level` in each part.
 * Only the entry points below, fides_ct_load (which resets the form),
 * fides_ct_copy, fides_ct_info and fides_ct_device_ptrs accept a P*Q
 * ciphertext; every other entry point returns FIDES_E_STATE for one, as
 * input or output, until fides_ct_mod_down brings it back to the Q basis.
 *   fides_ct_extend            Ciphertext::extend (Ciphertext.cuh:441,
 *                              Ciphertext.cpp:742-754): (P c0, P c1), special limbs 0
 *   fides_ct_mod_down          Ciphertext::modDown (Ciphertext.cuh:379-382,
 *                              Ciphertext.cpp:311-320); out may alias a
 *   fides_eval_rotate_hoisted_ext  Ciphertext::rotate_hoisted(indexes, results,
 *                              ext = true) (Ciphertext.cuh:453, Ciphertext.cpp:1053-1226;
 *                              OpenFHE EvalFastRotationExt, api/CryptoContext.hpp:166-169):
 *                              one ModUp of c1, each output (sigma(KS_b + P c0),
 *                              sigma(KS_a)) left in P*Q; index 0 gives extend(a);
 *                              outputs distinct and not aliasing a
 *   fides_eval_add_ext         Ciphertext::add of P*Q operands at one level (a Q
 *                              operand is extended first, RNSPoly.cpp:180-195)
 *   fides_eval_dot_product_plain  Ciphertext::dotProductPt (Ciphertext.cuh:596-617,
 *                              Ciphertext.cpp:1638-1663): sum_i a[i] * pt[i], all a[i]
 *                              at one level and in one form (P*Q ones need P*Q
 *                              plaintexts); out may alias an a[i]
 *   fides_ct_store_ext         the (level+1+K) limbs of each part
 *   fides_ct_is_ext            1 for the P*Q form */
int fides_ct_extend(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out);
int fides_ct_mod_down(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out);
int fides_eval_rotate_hoisted_ext(fides_context* ctx, const fides_ciphertext* a, const int32_t* indices, int32_t n,
                                  fides_ciphertext* const* out);
int fides_eval_add_ext(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b,
                       fides_ciphertext* out);
int fides_eval_dot_product_plain(fides_context* ctx, const fides_ciphertext* const* a,
                                 const fides_plaintext* const* pt, int32_t n, fides_ciphertext* out);
int fides_ct_store_ext(fides_context* ctx, const fides_ciphertext* ct, uint64_t* sub0, uint64_t* sub1);
/* Ciphertext::mult(b, rescale = false, moddown = false) (Ciphertext.cuh:264):
 * the relinearized product left in P*Q, (P d0 + KS_b(d2), P d1 + KS_a(d2));
 * fides_ct_mod_down of it equals fides_eval_mult without rescale. */
int fides_eval_mult_ext(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b,
                        fides_ciphertext* out);
int fides_ct_is_ext(const fides_ciphertext* ct, int32_t* ext);
/* Ciphertext::dotProduct(a, b, ext) (Ciphertext.cuh:617, Ciphertext.cpp:1665-1707,
 * RNSPoly::dotProduct RNSPoly.cpp:941-987): sum_i a[i] * b[i] with ONE key
 * switch of the summed d2; all operands in the Q basis, at one level, noise
 * degree 1; not rescaled (noise degree 2).  ext != 0 leaves the result in
 * P*Q as fides_eval_mult_ext does; out may alias an operand. */
int fides_eval_dot_product(fides_context* ctx, const fides_ciphertext* const* a, const fides_ciphertext* const* b,
                           int32_t n, int32_t ext, fides_ciphertext* out);

/* Serialization.  FIDESlib serializes the OpenFHE objects with cereal
 * (api/Serialize.cpp:14-65, :95-190) and its GPU state in a `<file>.dev` text
 * sidecar; this engine does not link OpenFHE, so it saves the raw RNS objects
 * that cross the boundary (RawParams, RawCipherText, RawPlainText,
 * RawKeySwitchKey) in its own little-endian binary files ("FIDESAMD" magic,
 * version, kind; each load checks the file's moduli against the context's),
 * and writes / parses the sidecar in the reference's exact text form:
 *   "<n> { <dev> ... }" / "AutoLoadCiphertexts: x" / "AutoLoadPlaintexts: x" /
 *   "RotationIndexes: { <i> ... }" / "KeyDist: x"
 * (missing lines keep the reference's defaults 1, 0, {}, 0).
 * Key kinds: FIDES_KEY_KIND_EVAL (relinearization), _AUTOMORPH (index = the
 * automorphism index k), _SWITCH (index = the caller id of
 * fides_key_load_switch); fides_key_load_file loads the key under the kind
 * and index the file records and returns them. */
#define FIDES_KEY_KIND_EVAL 0
#define FIDES_KEY_KIND_AUTOMORPH 1
#define FIDES_KEY_KIND_SWITCH 2
int fides_params_save(const fides_raw_params* params, const char* path);
int fides_context_save_params(fides_context* ctx, const char* path);
int fides_context_create_from_file(const char* path, int32_t device, fides_context** out);
int fides_ct_save(fides_context* ctx, const fides_ciphertext* ct, const char* path);
int fides_ct_load_file(fides_context* ctx, fides_ciphertext* ct, const char* path);
int fides_pt_save(fides_context* ctx, const fides_plaintext* pt, const char* path);
int fides_pt_load_file(fides_context* ctx, fides_plaintext* pt, const char* path);
int fides_key_save(fides_context* ctx, int32_t kind, int32_t index, const char* path);
int fides_key_load_file(fides_context* ctx, const char* path, int32_t* kind, int32_t* index);
int fides_dev_sidecar_write(const char* path, int32_t n_devices, const int32_t* devices, int32_t auto_load_ciphertexts,
                            int32_t auto_load_plaintexts, int32_t n_rotations, const int32_t* rotation_indexes,
                            int32_t key_dist);
/* n_devices / n_rotations get the counts; the arrays (if not NULL) the first
 * min(count, cap) entries. */
int fides_dev_sidecar_read(const char* path, int32_t* devices, int32_t dev_cap, int32_t* n_devices,
                           int32_t* auto_load_ciphertexts, int32_t* auto_load_plaintexts, int32_t* rotation_indexes,
                           int32_t rot_cap, int32_t* n_rotations, int32_t* key_dist);

/* Bootstrapping building blocks.
 * X -> X^k with the rotation key loaded for automorphism index k
 * (fides_key_load_automorph); key switch, ModDown, add c0, automorphism. */
int fides_eval_automorph(fides_context* ctx, const fides_ciphertext* a, uint32_t k, fides_ciphertext* out);
/* CKKS::Ciphertext::conjugate (Ciphertext.cpp:933-990): k = 2N - 1. */
int fides_eval_conjugate(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out);
/* ModRaise (Bootstrap.cu ModRaise: INTT, broadcastLimb0's centered SwitchModulus,
 * NTT): a level-0 ciphertext raised to `level`. */
int fides_mod_raise(fides_context* ctx, const fides_ciphertext* a, int32_t level, fides_ciphertext* out);
/* CKKS::Ciphertext::multMonomial (Ciphertext.cpp:1709-1760): out = a * X^power. */
int fides_eval_mult_monomial(fides_context* ctx, const fides_ciphertext* a, int64_t power, fides_ciphertext* out);
/* multIntScalar (Bootstrap.cu): out = a * s for an integer s (scale unchanged). */
int fides_eval_mult_int(fides_context* ctx, const fides_ciphertext* a, uint64_t s, fides_ciphertext* out);

/* evalHornerSeries (ApproxModEval.cu:710-745): out = sum_i coeffs[i] x^i by
 * Horner's rule with FIDESlib's rescale placement (FIXEDMANUAL). */
int fides_eval_horner_series(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                             fides_ciphertext* out);
/* applyDoubleAngleIterations (ApproxModEval.cu:747-761): `iterations` rounds of
 * rescale, square, double, add -1/(2 pi)^(2^(j - iterations)). */
int fides_eval_double_angle(fides_context* ctx, const fides_ciphertext* a, int32_t iterations,
                            fides_ciphertext* out);

/* evalChebyshevSeries (ApproxModEval.cu:370-700): out = sum_k coeffs[k] T_k(y),
 * y = (2x - (a + b)) / (b - a), by Paterson-Stockmeyer in the Chebyshev basis
 * (the engine's own operation order, with exact scales; under FIXEDAUTO and
 * FLEXIBLEAUTO too, a FLEXIBLEAUTO result on its level's ScalingFactorReal;
 * FLEXIBLEAUTOEXT is refused). */
int fides_eval_chebyshev_series(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                                double a, double b, fides_ciphertext* out);

/* approxModReductionSparse (ApproxModEval.cu:94-130): the Chebyshev series on
 * [-1, 1], double_angle_its double-angle iterations, then times `post`. */
int fides_eval_mod_reduction(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                             int32_t double_angle_its, uint64_t post, fides_ciphertext* out);

/* EvalNegate (api/CryptoContext.cpp:630-648): out = -a exactly (OpenFHE's
 * semantics, metadata unchanged; the reference's GPU branch multiplies by the
 * scalar -1.0, which under FIXEDMANUAL also raises the noise degree). */
int fides_eval_negate(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out);

/* EvalAddMany (api/CryptoContext.cpp:834-880): out = a[0] + .. + a[n-1] in one
 * launch at the lowest input level (inputs share the noise degree; out may
 * alias an input). */
int fides_eval_add_many(fides_context* ctx, const fides_ciphertext* const* a, int32_t n, fides_ciphertext* out);

/* AccumulateSum(ct, slots, stride) (api/CryptoContext.cpp:1649-1700; FIDESlib
 * CKKS::Accumulate, src/CKKS/AccumulateBroadcast.cu:41-67): out = sum over
 * k < size of the rotation of a by k stride, in radix-4 hoisted steps, each
 * summed in P*Q and brought down by one ModDown as the reference does
 * (rotate_hoisted(.., ext = true), extend, add, modDown); size a power of
 * two.  fides_accumulate_rotations lists the rotation keys it needs
 * (GetAccumulateRotationIndices with bStep 4): n gets the count, idx (if not
 * NULL) the first min(count, cap). */
int fides_accumulate_rotations(int32_t size, int32_t stride, int32_t* idx, int32_t cap, int32_t* n);
int fides_eval_accumulate_sum(fides_context* ctx, const fides_ciphertext* a, int32_t size, int32_t stride,
                              fides_ciphertext* out);

/* Secret-switching keys (GetSecretSwitchingKey, src/CKKS/openfhe-interface/
 * RawCiphertext.cu:388-459), stored under a caller id in the RawKeySwitchKey
 * layout of fides_key_load_eval; fides_eval_key_switch is
 * Ciphertext::keySwitch (Ciphertext.cpp:1450-1461): out = (c0 + KS_b(c1),
 * KS_a(c1)).  Bootstrapping's sparse-secret encapsulation uses the ids below. */
#define FIDES_KEY_SSE_DENSE_TO_SPARSE 1
#define FIDES_KEY_SSE_SPARSE_TO_DENSE 2
int fides_key_load_switch(fides_context* ctx, int32_t id, const uint64_t* A, const uint64_t* B);
int fides_eval_key_switch(fides_context* ctx, const fides_ciphertext* a, int32_t id, fides_ciphertext* out);

/* MakeCKKSPackedPlaintext (api/CryptoContext.cpp:475-514, OpenFHE's CKKS
 * encoder): the N/2 complex slot values re[j] + i im[j] (im may be NULL for
 * real slots) encoded at `level` with scaling factor `scale` into pt
 * (evaluation form, noise degree 1). */
int fides_encode(fides_context* ctx, const double* re, const double* im, int32_t level, double scale,
                 fides_plaintext* pt);
/* The same encoding in P*Q form (the coefficients also reduced mod the K
 * special primes), as OpenFHE encodes linear-transform diagonals in the
 * extended basis for EvalFastRotationExt (api/CryptoContext.hpp:166-169). */
int fides_encode_ext(fides_context* ctx, const double* re, const double* im, int32_t level, double scale,
                     fides_plaintext* pt);

/* Bootstrapping of packed ciphertexts (N/2 slots, or sparse packing of fewer).
 * fides_bootstrap_setup replaces EvalBootstrapSetup(levelBudget, dim1, slots,
 * correctionFactor) (api/CryptoContext.hpp:80): it builds the CoeffsToSlots /
 * SlotsToCoeffs diagonals (level budgets as OpenFHE's levelBudget, baby steps
 * chosen per level) and the EvalMod polynomial.  The keys EvalBootstrapKeyGen
 * (:82) would make are the rotation keys fides_bootstrap_rotations lists plus
 * the conjugation key (automorphism 2N - 1), loaded with fides_load_*_key.
 * fides_eval_bootstrap replaces EvalBootstrap(ct) (:180; FIDESlib
 * CKKS::Bootstrap, src/CKKS/Bootstrap.cu:169-300): any-level input, output at
 * top_level minus the levels the pipeline uses.  Under FIXEDAUTO and
 * FLEXIBLEAUTO (as the reference, Bootstrap.cu:428-446) the input needs a
 * level to spare (level - noise degree + 1 >= 1, else FIDES_E_LEVEL): it is
 * brought onto `scale` by an integer product and a rescale first, and the
 * result sits on ScalingFactorReal of its level; FLEXIBLEAUTOEXT is refused
 * (FIDES_E_STATE). */
typedef struct fides_bootstrap fides_bootstrap;
typedef struct {
    int32_t level_budget_cts;  /* CoeffsToSlots levels (OpenFHE levelBudget[0]) */
    int32_t level_budget_stc;  /* SlotsToCoeffs levels (levelBudget[1]) */
    double K;                  /* bound on the raised overflow |I| (28 for sparse secrets) */
    int32_t double_angle;      /* double-angle iterations of EvalMod */
    int32_t degree;            /* Chebyshev degree of EvalMod */
    double scale;              /* scaling factor of the ciphertexts to bootstrap */
    int32_t top_level;         /* level ModRaise lifts to (-1: L) */
    int32_t sparse_encaps;     /* 1: sparse-secret encapsulation around ModRaise
                                  (switching keys FIDES_KEY_SSE_* must be loaded) */
    int32_t slots;             /* packed slots (EvalBootstrapSetup's `slots`): a power of
                                  two <= N/2; 0 = N/2 */
    int32_t batched_giant;     /* 1: each transform's giant steps as one batched rotation
                                  (needs the keys of every giant step, listed) */
    int32_t q_basis_diagonals; /* 0: CoeffsToSlots / SlotsToCoeffs diagonals in P*Q form, the
                                  extended-basis transform (OpenFHE's and FIDESlib's default);
                                  1: Q-basis diagonals (every rotation ModDowns at once) */
} fides_bootstrap_params;
int fides_bootstrap_setup(fides_context* ctx, const fides_bootstrap_params* p, fides_bootstrap** out);
/* The rotation indices whose keys a bootstrap needs: n gets the count, idx (if
 * not NULL) the first min(count, cap) of them. */
int fides_bootstrap_rotations(const fides_bootstrap* b, int32_t* idx, int32_t cap, int32_t* n);
int fides_eval_bootstrap(fides_context* ctx, const fides_bootstrap* b, const fides_ciphertext* in,
                         fides_ciphertext* out);
/* Batched bootstrapping (no FIDESlib counterpart): n ciphertexts bootstrapped
 * together, their 2n EvalMod halves in lockstep (every product one batched
 * key switch); out[i] may alias in[i]. */
int fides_eval_bootstrap_batch(fides_context* ctx, const fides_bootstrap* b, const fides_ciphertext* const* in,
                               fides_ciphertext* const* out, int32_t n);
int fides_bootstrap_destroy(fides_bootstrap* b);

/* Request loop with continuous batching (no counterpart in FIDESlib, which has
 * no serving layer).  Jobs reference device-resident inputs and outputs that
 * stay alive until the job is reported done; in-flight jobs must be
 * independent (no job reads another in-flight job's output) and outputs
 * distinct.  A job is admitted when fewer than max_slots jobs are active and
 * its device-memory footprint fits budget_bytes beside the admitted ones;
 * fides_server_step advances every active job by one stage, running the
 * key-switched ops of all of them (HMult, rotations, any levels) as one batch.
 * Single-threaded: all calls on one server from one thread. */
int fides_server_create(fides_context* ctx, int32_t max_slots, uint64_t budget_bytes, fides_server** out);
int fides_server_destroy(fides_server* s);
/* out = a * b, relinearized and rescaled. */
int fides_server_submit_mult(fides_server* s, const fides_ciphertext* a, const fides_ciphertext* b,
                             fides_ciphertext* out, int64_t* job);
/* out = rotate(a, index). */
int fides_server_submit_rotate(fides_server* s, const fides_ciphertext* a, int32_t index, fides_ciphertext* out,
                               int64_t* job);
/* Encrypted logistic-regression scoring: y = x * w rescaled, n_rot rotate-and-sum
 * steps (rotations by 1, 2, 4, ...), out = coeffs[0] + coeffs[1] y + coeffs[2] y^3.
 * x at level >= 3 (FIXEDMANUAL); needs the rotation keys of 1 .. 2^(n_rot-1). */
int fides_server_submit_logreg(fides_server* s, const fides_ciphertext* x, const fides_plaintext* w, int32_t n_rot,
                               const double* coeffs, fides_ciphertext* out, int64_t* job);
/* One scheduling step; *pending = jobs still active or queued. */
int fides_server_step(fides_server* s, int32_t* pending);
/* Up to cap ids of jobs whose device work has completed since the last poll. */
int fides_server_poll(fides_server* s, int64_t* done, int32_t cap, int32_t* n_done);
int fides_server_counts(fides_server* s, int32_t* active, int32_t* queued, uint64_t* used_bytes);

/* Stage entry points on device memory (parity tests, benchmarks).
 * NTT of `num_limbs` limbs of prime ids first_prime.. (LimbPartition::NTT/INTT,
 * src/CKKS/LimbPartition.cu:365-478). */
int fides_poly_ntt(fides_context* ctx, uint64_t* data, int32_t num_limbs, int32_t first_prime, int32_t inverse);
/* ModUp of an evaluation-form poly at `level` into ext[nd][level+1+K][N]. */
int fides_stage_modup(fides_context* ctx, int32_t level, const uint64_t* x, uint64_t* ext);
/* Autotuned launch configurations (no counterpart in the reference, whose
 * NTT and key-switch launch shapes are compile-time constants: batch = 100 in
 * src/CKKS/Parameters.cuh:30, DEFAULT_ALGO in src/CKKS/forwardDefs.cuh:38).
 * The library loads its measured table (pyfideslib_amd/tune_gfx950.txt, or
 * FIDES_TUNE_FILE; empty = fixed rules) when the first context is created.
 * kind: FIDES_TUNE_NTT_FWD / _NTT_INV (size = limbs per launch; cfg = column
 * tile width, log2 values per column thread, row tile height, grid cap) or
 * FIDES_TUNE_MODUP / _MODDOWN (size = conversions per launch; cfg[0] =
 * target groups; FIDES_TUNE_MODUP cfg[1] = 1 selects the int8 matrix-core
 * conversion, built when the digit has at most 16 limbs).  fides_tune_force pins a kind's configuration for every
 * later launch in the process (cfg NULL releases it; kind < 0 releases all);
 * fides_tune_config reports what a launch of that size would use. */
#define FIDES_TUNE_NTT_FWD 0
#define FIDES_TUNE_NTT_INV 1
#define FIDES_TUNE_MODUP 2
#define FIDES_TUNE_MODDOWN 3
int fides_tune_load(const char* path, int32_t* entries); /* path NULL: fixed rules only */
/* NTT (inverse != 0: INTT) of data[limbs][N] in place, limb i over prime id
 * i mod (L+1+K), limbs <= 8 (L+1+K): the launch shapes batched stages make. */
int fides_ntt_limbs(fides_context* ctx, uint64_t* data, int32_t limbs, int32_t inverse);
int fides_tune_force(int32_t kind, const int32_t* cfg);
int fides_tune_config(fides_context* ctx, int32_t kind, int32_t size, int32_t* cfg, int32_t* from_table);
/* fides_stage_modup without any device allocation: its (level+1)*N-word
 * workspace is a static 16 MiB (2^21-word) device array, one per GPU, held
 * until the process ends.  No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under stream capture, serving
 * loops, hosts that forbid hipMalloc after start-up).  (level+1)*N above 2^21
 * fails with FIDES_E_ARG.  Every call on one GPU shares the array: calls that
 * overlap on one GPU (two contexts on one device, two threads) race. */
int fides_stage_modup_noalloc(fides_context* ctx, int32_t level, const uint64_t* x, uint64_t* ext);
/* ModDown of an extended poly [level+1+K][N] to [level+1][N]. */
int fides_stage_moddown(fides_context* ctx, int32_t level, const uint64_t* x_ext, uint64_t* y);
/* Full key switch of x with the relinearization key (rotation_k == 0) or the
 * rotation key of automorphism index rotation_k: ks_b, ks_a [level+1][N]. */
int fides_stage_keyswitch(fides_context* ctx, int32_t level, const uint64_t* x, uint32_t rotation_k, uint64_t* ks_b,
                          uint64_t* ks_a);

#ifdef __cplusplus
}
#endif
#endif
