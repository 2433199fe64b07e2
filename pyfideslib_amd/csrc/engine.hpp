// Internal declarations shared by the HIP kernels and the host runtime.
//
// Layout in HBM (all uint64, limb-major, each limb N contiguous words):
//   ciphertext   [2][L+1][N]      c0 then c1, level l uses limbs 0..l
//   ext digits   [nd][l+1+K][N]   ModUp output: Q positions 0..l then P
//   key (A or B) [dnum][L+1+K][N] RawKeySwitchKey r_key layout
// A "position" indexes the limbs of one polynomial; position pos maps to the
// prime id  pos            for pos <  nq  (Q chain, q0 = id 0)
//           pbase+(pos-nq) for pos >= nq  (special primes, pbase = L+1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.hpp"

namespace fides {

// Device constant tables (pointers into one device allocation), passed to
// every kernel by value.
struct DevTables {
    int logN, N, L, K, dnum, alpha, np;
    const PrimeInfo* primes;           // [np]
    const ShoupConst* tw_fwd;          // [np][N]  psi^brev(k) with Shoup companion
    const ShoupConst* tw_inv;          // [np][N]  psi^-brev(k)
    const ShoupConst* iscale_plain;    // [np][2]  (N^-1, psi_inv_rev[1] N^-1)
    const ShoupConst* iscale_modup;    // [L+1 level][L+1 pos][2]  N^-1 * PartQlHatInvModq
    const ShoupConst* iscale_moddown;  // [K][2]   N^-1 * PHatInvModp
    const uint64_t* modup_mat;         // [dnum][alpha size-1][np target][conv_bound(alpha) i]  PartQlHatModp
    const uint64_t* moddown_mat;       // [np target][conv_bound(K) k]  PHatModq
    const ShoupConst* pinv;            // [np]     P^-1 mod q_i
    const ShoupConst* pmodq;           // [np]     P mod q_i (0 for the special primes)
    const ShoupConst* qlinv;           // [L+1 l][L+1 i] q_l^-1 mod q_i
    const uint64_t* qlmodq;            // [L+1 l][L+1 i] q_l mod q_i
    // 1 when every base-conversion sum fits one Barrett reduction, i.e.
    // sum_i (q_i - 1)(p - 1) < 2^(bits(p) + 63) for every digit and target p
    int modup_one_reduce, moddown_one_reduce;
    // Matrix-core ModUp (k_modup_mfma): [dnum][alpha size-1][15 c][np target]
    // [4 ms i] bytes dm_(i, c-a) of the balanced digits of modup_mat, a = 0..7;
    // ms = ceil(alpha / 4) MFMA steps, 0 when not built (alpha > 16, or a
    // sum beyond one Barrett reduction)
    const uint64_t* modup_mfma;
    int modup_mfma_ms;
};

// Template bound of a base conversion with n input limbs (digit size or K):
// exact up to 10, then 12, 16, 32, 64.  Matrix rows are padded to it (with
// zeros) so a kernel reads a whole row with wide scalar loads.
__host__ __device__ constexpr int conv_bound(int n) {
    return n <= 10 ? n : n <= 12 ? 12 : n <= 16 ? 16 : n <= 32 ? 32 : 64;
}

__host__ __device__ inline int pos_pid(int pos, int nq, int q0, int pbase) {
    return pos < nq ? q0 + pos : pbase + (pos - nq);
}

// ---------------------------------------------------------------- jobs ----
// A job list is either a small inline array (plain per-op launches: no
// descriptor upload) or a device table built once for a batch.
// Global-memory views of device pointers.  Job tables hold generic pointers;
// accesses through them would compile to flat instructions, which also count
// against the LDS/scalar wait counter and so serialize LDS exchanges with
// outstanding HBM traffic.
typedef __attribute__((address_space(1))) uint64_t gu64;
__device__ __forceinline__ const gu64* gp(const uint64_t* p) { return (const gu64*)p; }
__device__ __forceinline__ gu64* gp(uint64_t* p) { return (gu64*)p; }

// Read of a constant table entry (DevTables: never written by a kernel)
// through the constant address space: with a wave-uniform address this is a
// scalar load.  A vector load here would have to be waited for with vmcnt,
// which on CDNA also waits for every store the wave has in flight.
template <class X>
__device__ __forceinline__ X cld(const X* p) {
    static_assert(sizeof(X) % 4 == 0, "table entries are whole dwords");
    union Words {
        X x;
        uint32_t w[sizeof(X) / 4];
    } u;
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)p;
#pragma unroll
    for (unsigned k = 0; k < sizeof(X) / 4; ++k) u.w[k] = q[k];
    return u.x;
}

template <class J, int CAP>
struct JobList {
    const J* dev;         // device table, or nullptr for the inline array
    const int* limbmap;   // device: flattened limb -> job index (NTT lists only)
    int n;
    J inl[CAP];
};

// Job i of a list (i wave-uniform): scalar loads from the device table, or a
// select over the inline array (constant offsets into the kernel arguments).
template <class J, int CAP>
__device__ __forceinline__ J get_job(const JobList<J, CAP>& L, int i) {
    static_assert(sizeof(J) % 4 == 0, "job records are whole dwords");
    constexpr int NW = sizeof(J) / 4;
    union Words {
        J j;
        uint32_t w[NW];
    } u;
    i = __builtin_amdgcn_readfirstlane(i);
    if (L.dev) {
        const __attribute__((address_space(1))) uint32_t* p =
            (const __attribute__((address_space(1))) uint32_t*)(L.dev + i);
#pragma unroll
        for (int k = 0; k < NW; ++k) u.w[k] = p[k];
    } else {
        const __attribute__((address_space(4))) uint32_t* p =
            (const __attribute__((address_space(4))) uint32_t*)(&L.inl[0]);
#pragma unroll
        for (int k = 0; k < NW; ++k) u.w[k] = p[i * NW + k];
    }
    return u.j;
}

enum NttLoad { LOAD_PLAIN = 0, LOAD_SWITCH = 1 };
enum NttStore {
    STORE_PLAIN = 0,
    STORE_MODDOWN = 1,
    STORE_RESCALE = 2,
    STORE_MODDOWN_RESCALE = 3,
    STORE_MODDOWN_AUTO = 4  // STORE_MODDOWN, then the automorphism of NttJob::autok
};
enum IScale { ISCALE_PLAIN = 0, ISCALE_MODUP = 1, ISCALE_MODDOWN = 2 };

// One NTT job: positions [0, npos) of a polynomial minus [skip_lo, skip_hi).
// Pass 1 reads src, writes mid; pass 2 reads mid, writes dst.
struct NttJob {
    const uint64_t* src;  // LOAD_SWITCH: the single coefficient-form limb of q_l
    uint64_t* mid;
    uint64_t* dst;
    const uint64_t* aux;  // STORE_MODDOWN(_RESCALE): key-switch accumulator limbs
    const uint64_t* add;  // STORE_MODDOWN(_RESCALE): optional addend limbs (may alias dst)
    int npos, nq, q0;
    int skip_lo, skip_hi;
    int level;            // ISCALE_MODUP level; LOAD_SWITCH / STORE_*RESCALE: l
    int scale_sel;        // IScale
    int first;            // flattened index of this job's first active limb
    uint32_t autok;       // STORE_MODDOWN_AUTO: k^-1 mod 2N of the automorphism X -> X^k
};
constexpr int NTT_INLINE = 8;
using NttJobs = JobList<NttJob, NTT_INLINE>;

struct TensorJob {
    const uint64_t *c0, *c1, *d0, *d1;
    uint64_t *o0, *o1, *o2;
    int nl;
};
constexpr int TENSOR_INLINE = 2;
using TensorJobs = JobList<TensorJob, TENSOR_INLINE>;

// ModUp base conversion of one polynomial: t = scaled coefficient limbs.
struct ModUpJob {
    const uint64_t* t;  // [l+1][N]
    uint64_t* ext;      // [nd][l+1+K][N]
    int level;
};
constexpr int MODUP_INLINE = 2;
using ModUpJobs = JobList<ModUpJob, MODUP_INLINE>;

// Key-switch inner product over the l+1+K positions of one polynomial.
struct KskJob {
    const uint64_t* ext;  // [nd][l+1+K][N]
    const uint64_t* own;  // [l+1][N] evaluation-form source (own-digit positions)
    const uint64_t* ka;   // [dnum][L+1+K][N]
    const uint64_t* kb;
    uint64_t* acc_a;      // [l+1+K][N]
    uint64_t* acc_b;
    int level;
};
constexpr int KSK_INLINE = 2;
using KskJobs = JobList<KskJob, KSK_INLINE>;

// ModDown base conversion P -> Q_l of one polynomial.  With top_acc set the
// job also folds the following rescale in (see k_moddown_conv): it reads the
// coefficient-form top limbs of the accumulator and of the addend and writes
// z_i = conv_i P^-1 + Switch_{q_l -> q_i}(top) for i < l.
struct ModDownJob {
    const uint64_t* t;        // [K][N] scaled coefficient limbs of the special primes
    uint64_t* y;              // [l+1][N] (or [l][N] z when rescaling)
    const uint64_t* top_acc;  // INTT of acc limb l, or nullptr
    const uint64_t* top_add;  // INTT of addend limb l
    int level;
};
constexpr int MODDOWN_INLINE = 4;
using ModDownJobs = JobList<ModDownJob, MODDOWN_INLINE>;

// Elementwise a (op) b over nl Q limbs.
struct EwJob {
    const uint64_t* a;
    const uint64_t* b;
    uint64_t* o;
    int nl;
};
constexpr int EW_INLINE = 4;
using EwJobs = JobList<EwJob, EW_INLINE>;
enum EwOp { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2, EW_MUL_SCALAR = 3, EW_ADD_SCALAR = 4 };

// Per-limb constants of a scalar op (limb i: c[i], with its Shoup companion
// for EW_MUL_SCALAR), passed by value in the kernel arguments.
constexpr int MAX_Q_LIMBS = 96;
struct LimbConsts {
    ShoupConst c[MAX_Q_LIMBS];
};

// Inner sums of a baby-step / giant-step linear transform, for one
// ciphertext and up to DOT_GMAX giant steps: o[j] = sum_t w[j][t] * x[t] on
// both parts, every baby step x[t] read once for all the job's giant steps.
// x0, x1 are device arrays of b polynomial pointers (limb 0 of each), w and wp
// [gc][b] (nullptr: no term, the ragged last giant step), o0 / o1 [gc].
// Limbs below nq are the Q primes 0 .. nq-1; limbs nq .. nl-1 the special
// primes (the P*Q form of the extended-basis transform), whose weights start
// at wp[j][t] (the special limbs of a P*Q plaintext, after its own Q limbs).
struct DotJob {
    const uint64_t* const* x0;
    const uint64_t* const* x1;
    const uint64_t* const* w;
    const uint64_t* const* wp;
    uint64_t* const* o0;
    uint64_t* const* o1;
    int b;
    int gc;
    int nl;
    int nq;
};
constexpr int DOT_GMAX = 8;
constexpr int DOT_INLINE = 2;
using DotJobs = JobList<DotJob, DOT_INLINE>;

// Outputs of extended-basis (P*Q) rotations, limbs [nl Q][K special] per part
// (FIDESlib Ciphertext::rotate_hoisted(.., ext = true), Ciphertext.cpp:1053-1226,
// and rotate(index, moddown = false), :756-926): per term t
//   o1 += sigma_k(acc_a),   o0 += sigma_k(acc_b + c0'),
// with c0' = P add0 on the Q limbs and 0 on the special limbs when scale_p is
// set (a Q-basis source, LimbPartition::add's addScaleB_), else add0 itself
// (a P*Q source); add1 enters o1 the same way (the identity rotation, whose
// acc is null: P c0, P c1).  base0 / base1 (P*Q, may alias the outputs) are
// added at the output index.  sigma_k gathers: output j reads input s with
// 2 brev(s) + 1 = (2 brev(j) + 1) k mod 2N.  No output may alias a gathered input.
struct ExtRotTerm {
    const uint64_t* acc_a;
    const uint64_t* acc_b;
    const uint64_t* add0;
    const uint64_t* add1;
    uint32_t k;
    int scale_p;
};
struct ExtRotJob {
    ExtRotTerm t0;            // the first term
    const ExtRotTerm* more;   // device array of the nt - 1 further terms
    const uint64_t* base0;
    const uint64_t* base1;
    uint64_t* o0;
    uint64_t* o1;
    int nt;
    int nl;  // Q limbs; every part has nl + K limbs
};
constexpr int EXTROT_INLINE = 2;
using ExtRotJobs = JobList<ExtRotJob, EXTROT_INLINE>;

// Linear combination of ciphertexts with integer constants over nl Q limbs:
// o0 = sum_t w[t][i] * x0[t], o1 = sum_t w[t][i] * x1[t] on limb i; x0, x1 are
// device arrays of n polynomial pointers, w a device table [n][nl] of
// per-limb constants with their Shoup companions.
// Sum of tensor products over n ciphertext pairs (Ciphertext::dotProduct):
// x = [a0 x n][a1 x n][b0 x n][b1 x n] limb pointers at one level.
struct TensorSumJob {
    const uint64_t* const* x;
    uint64_t *o0, *o1, *o2;
    int n;
    int nl;
};

struct LinCombJob {
    const uint64_t* const* x0;
    const uint64_t* const* x1;
    const ShoupConst* w;
    uint64_t* o0;
    uint64_t* o1;
    int n;
    int nl;
};

// ------------------------------------------------------------- tuning ----
// Launch configurations picked per (kind, logN, size) from a table measured
// on the device (pyfideslib_amd/tune.py writes it, tune.hip loads it once per
// process), falling back to the launchers' fixed rules when no entry covers
// the size.  NTT entries: {column tile width, log2 values per column thread,
// row tile height, grid cap (0 = one block per tile)} by limb count;
// conversion entries: {target groups} by conversions per launch.
enum TuneKind { TUNE_NTT_FWD = 0, TUNE_NTT_INV = 1, TUNE_MODUP = 2, TUNE_MODDOWN = 3, TUNE_KINDS = 4 };
constexpr int TUNE_NV = 4;
struct TuneCfg {
    int v[TUNE_NV];
};
// A forced configuration (tune_force), else the table's entry: true if found.
bool tune_get(int kind, int logN, int size, TuneCfg& out);
void tune_force(int kind, const TuneCfg* cfg);  // nullptr clears the kind; kind < 0 clears all
int tune_load(const char* path);                // entries read, -1 when the file cannot be read
void tune_load_default();                       // FIDES_TUNE_FILE, else the table next to the library
uint64_t tune_epoch();                          // changes whenever a lookup result may change
// The fixed rules (used when no entry applies), for reports.
TuneCfg ntt_rule_cfg(bool forward, int logN, int limbs);
int conv_rule_tg(int base_blocks, int max_targets);

// ------------------------------------------------------------ launchers ----
// All launchers enqueue on `stream` and return hipError_t.
hipError_t launch_ntt_fwd(const DevTables& T, const NttJobs& jobs, int total_limbs, int load_mode, int store_mode,
                          hipStream_t stream);
hipError_t launch_ntt_inv(const DevTables& T, const NttJobs& jobs, int total_limbs, hipStream_t stream);
hipError_t launch_tensor(const DevTables& T, const TensorJobs& jobs, int max_nl, hipStream_t stream);
hipError_t launch_modup_conv(const DevTables& T, const ModUpJobs& jobs, int max_nd, hipStream_t stream);
// shared: all jobs use the same ext / own / level (rotations of one hoisted set)
hipError_t launch_ksk_inner(const DevTables& T, const KskJobs& jobs, int max_npos, hipStream_t stream,
                            bool shared = false);
// The same for a list whose jobs all share one key (job 0's), key-stationary:
// each key value is read once per chunk of jobs (dnum <= 4).
// chunks: nchunks + 1 ascending job indices (device memory), each chunk's jobs
// sharing one key; ksk_keystat_chunk() is the largest chunk to cut.
hipError_t launch_ksk_keystat(const DevTables& T, const KskJobs& jobs, const int* chunks, int nchunks,
                              hipStream_t stream);
int ksk_keystat_chunk();
bool ksk_keystat_on();  // FIDES_KSK_KEYSTAT=0 turns the key-stationary form off
hipError_t launch_moddown_conv(const DevTables& T, const ModDownJobs& jobs, int max_nl, hipStream_t stream);
// rescale = every job carries top_acc/top_add (fused ModDown + rescale)
hipError_t launch_moddown_conv_mode(const DevTables& T, const ModDownJobs& jobs, bool rescale, hipStream_t stream);
hipError_t launch_elementwise(const DevTables& T, const EwJobs& jobs, int max_nl, int op, hipStream_t stream);
hipError_t launch_tensor_sum(const DevTables& T, const TensorSumJob& J, hipStream_t s);
hipError_t launch_lincomb(const DevTables& T, const LinCombJob& job, hipStream_t stream);
// max_gc: the largest gc of the list (selects the accumulator count)
hipError_t launch_dot_bsgs(const DevTables& T, const DotJobs& jobs, int max_nl, int max_gc, hipStream_t stream);
hipError_t launch_ext_rot(const DevTables& T, const ExtRotJobs& jobs, int max_nl, hipStream_t stream);
// o = a * X^m (evaluation form) on each limb; b unused.
hipError_t launch_monomial(const DevTables& T, const EwJobs& jobs, int max_nl, uint32_t m, hipStream_t stream);
// o = a * c[i] (EW_MUL_SCALAR) or a + c[i] (EW_ADD_SCALAR) on limb i; b unused.
hipError_t launch_scalar(const DevTables& T, const EwJobs& jobs, int max_nl, int op, const LimbConsts& k,
                         hipStream_t stream);

}  // namespace fides
