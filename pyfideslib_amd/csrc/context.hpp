// Host runtime of the CKKS engine: context, device memory, ciphertexts, keys.
//
// Mirrors FIDESlib's CKKS layer (src/CKKS/Context.cuh, Ciphertext.cuh,
// KeySwitchingKey.cuh): a Context is built from the RawParams an OpenFHE
// context exports (RawCiphertext.cu:219-467), ciphertexts and keys are loaded
// from the RawCipherText / RawKeySwitchKey limb arrays, and every evaluation op
// runs on the context's HIP stream.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <complex>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine.hpp"

namespace fides {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status { OK = 0, E_ARG = 1, E_HIP = 2, E_STATE = 3, E_NOKEY = 4, E_LEVEL = 5, E_OOM = 6 };

void hip_check(hipError_t e, const char* what);

// Scaling techniques, numbered as lbcrypto::ScalingTechnique.
enum ScalingTechnique { FIXEDMANUAL = 0, FIXEDAUTO = 1, FLEXIBLEAUTO = 2, FLEXIBLEAUTOEXT = 3 };

// Stream-ordered caching allocator: blocks are recycled by exact size.  All
// work of a context runs on its one stream, so a block released after its
// last use was enqueued may be handed to the next enqueued op.
// Device memory cache.  Blocks are rounded to size classes (256-byte granules
// up to 1 MiB, then 8 classes per power of two) and returned blocks are kept
// for reuse by later requests of the same class; all users order their work
// on one stream, so a block released while kernels still read it is only
// handed to work queued after them.  trim() gives cached blocks back to the
// device (hipFree waits for the device); a failed hipMalloc trims and retries.
class Pool {
   public:
    ~Pool();
    void* get(size_t bytes);
    void put(void* p, size_t bytes);
    // Frees cached (unused) blocks until at most keep_bytes stay cached.  Not
    // while a stream of this pool's users is being captured.
    void trim(size_t keep_bytes);
    size_t bytes_held() const { return held_; }
    size_t bytes_cached() const { return cached_; }
    static size_t size_class(size_t bytes);

   private:
    std::mutex mu_;
    std::unordered_map<size_t, std::vector<void*>> free_;
    std::unordered_map<void*, size_t> live_;  // every block this pool allocated
    size_t held_ = 0, cached_ = 0;
};

struct DevBuf {
    Pool* pool = nullptr;
    uint64_t* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(Pool* p, size_t words) : pool(p), ptr((uint64_t*)p->get(words * 8)), bytes(words * 8) {}
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        pool = o.pool; ptr = o.ptr; bytes = o.bytes;
        o.ptr = nullptr; o.bytes = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) pool->put(ptr, bytes);
        ptr = nullptr;
    }
};

class Context;

// KeySwitchingKey: RawKeySwitchKey r_key[0] (A) and r_key[1] (B), each
// [dnum][L+1+K][N] in evaluation form (FIDESlib KeySwitchingKey.cu:20-35).
struct KeySwitchingKey {
    DevBuf a, b;
};

struct Ciphertext {
    Context* ctx = nullptr;
    DevBuf data;      // [2][part words]: c0 then c1
    size_t part = 0;  // words per part: (L+1) N, or (L+1+K) N once it has held a P*Q form
    int level = -1;   // limbs 0..level are valid
    // 1: P*Q form (FIDESlib RNSPoly isModUp, both parts): the K special-prime
    // limbs follow limb `level` in each part.  Only the extended-basis entry
    // points accept it; ModDown returns it to the Q basis.
    int ext = 0;
    double noise = 1; // scaling factor (RawCipherText::Noise)
    int noise_level = 1;
    int slots = 0;
    uint64_t* c0() const { return data.ptr; }
    uint64_t* c1() const;
};

// Plaintext in evaluation form (FIDESlib CKKS::Plaintext, Plaintext.cuh:22-120;
// RawPlainText): one polynomial, limbs 0..level.
struct Plaintext {
    Context* ctx = nullptr;
    DevBuf data;      // [L+1][N], or [L+1+K][N] once loaded in P*Q form
    int level = -1;
    // 1: a P*Q plaintext (RNSPoly::loadConstant with special-prime limbs,
    // RNSPoly.cpp:815-853): the K special limbs follow limb `level`
    int ext = 0;
    double noise = 1; // scaling factor
    int noise_level = 1;
    int slots = 0;
    uint64_t* c0() const { return data.ptr; }
};

class Context {
   public:
    struct Params {
        int logN, L, K, dnum;
        std::vector<uint64_t> moduli, roots;  // Q then P
        int scaling_technique = FIXEDMANUAL;
        std::vector<double> mod_reduce_factor;  // [L+1] or empty
        std::vector<double> scaling_factor_real;  // [L+1] or empty (scalar ops then unavailable)
        std::vector<double> scaling_factor_real_big;  // [L+1] or empty (FLEXIBLEAUTO alignment then limited)
    };
    Context(const Params& p, int device);
    ~Context();

    int logN, N, L, K, dnum, alpha, np, device;
    Params params;  // as constructed (saved by params_save)
    int scaling_technique;
    std::vector<uint64_t> q;  // all primes
    std::vector<double> mod_reduce_factor;
    std::vector<double> scaling_factor_real;
    std::vector<double> scaling_factor_real_big;
    hipStream_t stream = nullptr;
    DevTables T{};
    Pool pool;
    std::unique_ptr<KeySwitchingKey> eval_key;
    std::map<uint32_t, std::unique_ptr<KeySwitchingKey>> rot_keys;  // by automorphism index
    std::map<int, std::unique_ptr<KeySwitchingKey>> switch_keys;     // secret-switching keys by caller id
    std::recursive_mutex api_mu;  // held by every C ABI call on this context
    // Caches of launch sequences.  A signature lists every device address a
    // cached sequence bakes in that is not owned by the entry itself (operand
    // buffers, both key components), so an entry only matches while those
    // buffers are the live ones at the same addresses.
    std::shared_ptr<void> plan_cache;  // last batched-op plan (job tables + scratch)
    std::shared_ptr<void> hoist_plan;  // last hoisted-rotation-set plan
    // Captured per-op launch sequences (hipGraph) keyed by operands, level and
    // key; each entry owns the scratch its graph was captured with.
    std::map<std::vector<uint64_t>, std::shared_ptr<void>> graph_cache;
    std::set<std::vector<uint64_t>> graph_seen;  // signatures run once (eagerly)
    std::shared_ptr<void> ks_workspace;          // key-switch buffers shared by all ops
    bool use_graphs = true;  // FIDES_GRAPHS=0 disables capture
    // Per-op launch sequences captured and replayed (run_cached): off by
    // default -- replaying an 8-launch op as a graph measured 2-4% slower than
    // launching it (HMult 3,780 vs 3,878 ops/s, rotation 4,431 vs 4,591); on
    // with FIDES_OP_GRAPHS=1 or fides_context_set_op_graphs.  Batch and
    // hoisted plans are captured either way (use_graphs).
    bool op_graphs = false;

    // Cache entries (captured graphs with their tables and scratch) leave a
    // cache through retire(): HIP does not document hipGraphExecDestroy as
    // safe while the graph still executes, so an entry is destroyed only once
    // an event recorded after its last launch has completed (checked by
    // reap() at the next retirement, or waited for when the context ends).
    void retire(std::vector<std::shared_ptr<void>> objs);
    void retire(std::shared_ptr<void> obj) { retire(std::vector<std::shared_ptr<void>>{std::move(obj)}); }
    void reap(bool wait);
    // Drops every cached plan and graph (a key was loaded or replaced).
    void drop_plans();

    int num_digits(int level) const {
        int n = (level + 1 + alpha - 1) / alpha;
        return n < dnum ? n : dnum;
    }
    int digit_size(int d, int level) const {
        int s = level + 1 - d * alpha;
        return s < alpha ? s : alpha;
    }
    size_t ct_words() const { return (size_t)2 * (L + 1) * N; }
    uint32_t rot_index(int r) const;  // 5^r mod 2N (r reduced mod N/2)
    uint32_t auto_inverse(uint32_t k) const;  // k^-1 mod 2N (k odd)
    void sync();
    // Copies host bytes to device memory in stream order through a ring of
    // pinned staging slots: the call returns once the bytes are staged,
    // without waiting for the device beyond the upload a full ring back
    // (which precedes every kernel queued since).
    void upload(void* dst, const void* src, size_t bytes);
    // Scratch shared by consecutive batch plans (grown, never shrunk): a plan
    // replaces the previous one in stream order, so one arena serves them all.
    uint64_t* batch_arena(size_t words);

   private:
    void* blob_ = nullptr;
    // pinned staging ring for upload(): a slot is reused once the copy from it
    // has run, so the host may queue up to STAGE_SLOTS uploads ahead of the
    // GPU.  Kept shallow: a deeper ring lets a request loop queue ticks far
    // ahead, and jobs that arrive meanwhile wait behind them (serving p50 at
    // 1000 jobs/s: 2.3 ms with 2 slots, 3.3 ms with 32), for no gain elsewhere.
    static constexpr int STAGE_SLOTS = 2;
    void* stage_[STAGE_SLOTS] = {};
    size_t stage_bytes_[STAGE_SLOTS] = {};
    hipEvent_t stage_ev_[STAGE_SLOTS] = {};
    bool stage_used_[STAGE_SLOTS] = {};
    int stage_next_ = 0;
    DevBuf arena_;
    std::vector<std::pair<hipEvent_t, std::vector<std::shared_ptr<void>>>> retired_;
    void build_tables(const std::vector<uint64_t>& roots);
};

// Captures the launches `body` enqueues on the context stream into an
// executable graph; a capture that throws is ended and its error re-raised.
template <class F>
hipGraphExec_t capture_graph(Context& c, F&& body) {
    hip_check(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipGraph_t graph = nullptr;
    try {
        body();
    } catch (...) {
        (void)hipStreamEndCapture(c.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    hip_check(hipStreamEndCapture(c.stream, &graph), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hip_check(e, "hipGraphInstantiate");
    return exec;
}

// ---- ops (ckks.hip) ----
std::unique_ptr<Ciphertext> ct_create(Context& c);
void ct_load(Context& c, Ciphertext& ct, const uint64_t* sub0, const uint64_t* sub1, int num_res, double noise,
             int noise_level, int slots, bool device_src);
void ct_store(Context& c, const Ciphertext& ct, uint64_t* sub0, uint64_t* sub1);
// Grow a ciphertext's buffer to hold a P*Q form ((L+1+K) N words per part);
// the contents are not kept.
void ct_reserve_ext(Context& c, Ciphertext& ct);
void ct_copy(Context& c, const Ciphertext& src, Ciphertext& dst);
void load_key(Context& c, KeySwitchingKey& k, const uint64_t* a, const uint64_t* b, bool device_src);

void eval_mult(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out, bool rescale);
void eval_square(Context& c, const Ciphertext& a, Ciphertext& out, bool rescale);
void rescale(Context& c, Ciphertext& a);
void eval_add(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out);
void eval_sub(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out);
void eval_rotate(Context& c, const Ciphertext& a, int index, Ciphertext& out);
std::unique_ptr<Plaintext> pt_create(Context& c);
// P*Q plaintext: limbs [num_res + K][N], the Q limbs 0 .. num_res-1 then the K special limbs.
void pt_load_ext(Context& c, Plaintext& pt, const uint64_t* limbs, int num_res, double noise, int noise_level,
                 int slots, bool device_src);
void pt_load(Context& c, Plaintext& pt, const uint64_t* limbs, int num_res, double noise, int noise_level, int slots,
             bool device_src);
void eval_mult_plain(Context& c, const Ciphertext& a, const Plaintext& pt, Ciphertext& out, bool rescale);
void eval_add_plain(Context& c, const Ciphertext& a, const Plaintext& pt, Ciphertext& out);
void eval_sub_plain(Context& c, const Ciphertext& a, const Plaintext& pt, Ciphertext& out);
void eval_add_mult_plain(Context& c, Ciphertext& acc, const Ciphertext& a, const Plaintext& pt, bool rescale);
void eval_add_mult_scalar(Context& c, Ciphertext& acc, const Ciphertext& b, double d);
void eval_linear_wsum(Context& c, const std::vector<const Ciphertext*>& x, const std::vector<double>& w,
                      Ciphertext& out);
// EvalMult without the ModDown (Ciphertext::mult(b, false, moddown = false)):
// the relinearized product left in P*Q.
void eval_mult_ext(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out);
void relin_add(Context& c, int level, const uint64_t* d2, uint64_t* o0, uint64_t* o1);
void eval_dot_product(Context& c, const std::vector<const Ciphertext*>& a, const std::vector<const Ciphertext*>& b,
                      bool ext, Ciphertext& out);
void eval_mult_scalar(Context& c, const Ciphertext& a, double v, Ciphertext& out, bool rescale);
void eval_add_scalar(Context& c, const Ciphertext& a, double v, Ciphertext& out);
void drop_to_level(Context& c, Ciphertext& ct, int level);
// Automatic scaling (FIXEDAUTO / FLEXIBLEAUTO): plain.hip
void mult_scalar_noprecheck(Context& c, Ciphertext& a, double v);
bool adjust_for_add_or_sub(Context& c, Ciphertext& a, const Ciphertext& b);
bool adjust_for_mult(Context& c, Ciphertext& a, const Ciphertext& b);
void eval_automorph(Context& c, const Ciphertext& a, uint32_t k, Ciphertext& out);
void eval_key_switch(Context& c, const Ciphertext& a, int key_id, Ciphertext& out);
// ids of the sparse-secret encapsulation keys among the switching keys
constexpr int KEY_SSE_DENSE_TO_SPARSE = 1, KEY_SSE_SPARSE_TO_DENSE = 2;
void mod_raise(Context& c, const Ciphertext& a, int to, Ciphertext& out);
void eval_mult_monomial(Context& c, const Ciphertext& a, int64_t m, Ciphertext& out);
void eval_negate(Context& c, const Ciphertext& a, Ciphertext& out);
void eval_add_many(Context& c, const std::vector<const Ciphertext*>& x, Ciphertext& out);
void eval_mult_int(Context& c, const Ciphertext& a, uint64_t s, Ciphertext& out);
void eval_horner_series(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, Ciphertext& out);
void eval_double_angle(Context& c, const Ciphertext& a, int iterations, Ciphertext& out);
// Bootstrapping (boot.hip): setup (diagonals, EvalMod coefficients) and evaluation.
// Compositions that keep every scale exact (the Chebyshev series, EvalMod,
// bootstrapping) run their steps with FIXEDMANUAL's explicit rescales
// whatever the context's technique: the technique is switched for the call
// (which holds the context's lock) and restored on exit.
struct ManualScaling {
    Context& c;
    int saved;
    explicit ManualScaling(Context& cc) : c(cc), saved(cc.scaling_technique) { c.scaling_technique = FIXEDMANUAL; }
    ~ManualScaling() { c.scaling_technique = saved; }
    ManualScaling(const ManualScaling&) = delete;
    ManualScaling& operator=(const ManualScaling&) = delete;
};

struct BootPlan;
struct BootPlanDeleter {
    void operator()(BootPlan* p) const;
};
using BootPlanPtr = std::unique_ptr<BootPlan, BootPlanDeleter>;
BootPlanPtr bootstrap_setup(Context& c, double scale, int budget_cts, int budget_stc, double K, int double_angle,
                            int degree, int top, bool sparse_encaps = false, int slots = 0,
                            bool batched_giant = false, bool ext_diagonals = true);
const std::vector<int>& bootstrap_rotations(const BootPlan& plan);
void eval_bootstrap(Context& c, const BootPlan& plan, const Ciphertext& in, Ciphertext& out);
void eval_bootstrap_batch(Context& c, const BootPlan& plan, const std::vector<const Ciphertext*>& ins,
                          const std::vector<Ciphertext*>& outs);
std::unique_ptr<Plaintext> encode_slots(Context& c, const std::vector<std::complex<double>>& z, int level,
                                        double scale, bool ext = false);
// Scale-exact forms (plain.hip): constants at the ciphertext's tracked scale.
void add_const_at_scale(Context& c, const Ciphertext& a, double v, Ciphertext& out);
void mult_const_to_scale(Context& c, const Ciphertext& a, int level, double v, double s_after, Ciphertext& out);
void double_angle_at_scale(Context& c, Ciphertext& out, int iterations);
void lincomb_to_scale(Context& c, const std::vector<const Ciphertext*>& x, const std::vector<double>& v, int level,
                      double s_after, Ciphertext& out);
void eval_chebyshev_series(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, double a, double b,
                           Ciphertext& out);
void eval_mod_reduction(Context& c, const Ciphertext& x, const std::vector<double>& coeffs, int double_angle_its,
                        uint64_t post, Ciphertext& out);
// The same over a batch in lockstep (every product one batched key switch).
void eval_chebyshev_series_batch(Context& c, const std::vector<const Ciphertext*>& xs,
                                 const std::vector<double>& coeffs, double a, double b,
                                 const std::vector<Ciphertext*>& outs);
void eval_mod_reduction_batch(Context& c, const std::vector<const Ciphertext*>& xs,
                              const std::vector<double>& coeffs, int double_angle_its, uint64_t post,
                              const std::vector<Ciphertext*>& outs);
void eval_linear_transform(Context& c, const Ciphertext& a, const std::vector<const Plaintext*>& diag, int bstep,
                           int stride, int offset, Ciphertext& out, bool batched_giant = false);
// ConvolutionTransform: giant steps go in blocks of CONV_BLOCK (the
// reference's INTERNAL_GSTEP, LinearTransform.cu:271).
constexpr int CONV_BLOCK = 8;
std::vector<int> convolution_rotations(int stride, int gstep);
void eval_convolution_transform(Context& c, const Ciphertext& a, const std::vector<const Plaintext*>& pts,
                                int bstep, int stride, const std::vector<int>& idx, int gstep, Ciphertext& out);
std::vector<int> accumulate_rotations(int size, int stride);
void eval_accumulate_sum(Context& c, const Ciphertext& a, int size, int stride, Ciphertext& out);
// P*Q ciphertexts (FIDESlib Ciphertext::extend / modDown / rotate_hoisted(ext) /
// add / dotProductPt, Ciphertext.cuh:379-382, 441-453, 596-617).
void eval_extend(Context& c, const Ciphertext& a, Ciphertext& out);
void eval_mod_down(Context& c, const Ciphertext& a, Ciphertext& out);
void eval_rotate_hoisted_ext(Context& c, const Ciphertext& a, const std::vector<int>& idx,
                             const std::vector<Ciphertext*>& outs);
void eval_add_ext(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out);
void eval_dot_product_plain(Context& c, const std::vector<const Ciphertext*>& a,
                            const std::vector<const Plaintext*>& pt, Ciphertext& out);
void eval_linear_transform_batch(Context& c, const std::vector<const Ciphertext*>& as,
                                 const std::vector<const Plaintext*>& diag, int bstep, int stride, int offset,
                                 const std::vector<Ciphertext*>& outs, bool batched_giant = false);
// Scalar encodings (FIDESlib ContextData::ElemForEvalMult / ElemForEvalAddOrSub).
std::vector<uint64_t> elem_for_mult(const Context& c, int level, double v);
std::vector<uint64_t> elem_for_add(const Context& c, int level, double v_abs, int noise_deg);
void eval_rotate_hoisted(Context& c, const Ciphertext& a, const std::vector<int>& idx,
                         const std::vector<Ciphertext*>& outs);
void eval_mult_batch(Context& c, const std::vector<const Ciphertext*>& a, const std::vector<const Ciphertext*>& b,
                     const std::vector<Ciphertext*>& out, bool rescale);
// One key-switched op of a batch (batch.hip): HMult+relin (a * b, rescaled when
// `rescale`) or a rotation of a (key = its rotation key, kinv = k^-1 mod 2N).
struct KsItem {
    bool mult;
    const Ciphertext* a;
    const Ciphertext* b;
    Ciphertext* out;
    const KeySwitchingKey* key;
    uint32_t kinv;
    bool rescale;
};
void ks_batch(Context& c, const std::vector<KsItem>& items);
void eval_rotate_batch(Context& c, const std::vector<const Ciphertext*>& a, const std::vector<int>& idx,
                       const std::vector<Ciphertext*>& out);

// ---- request loop (serve.hip) ----
class Server;
Server* server_create(Context& c, int max_slots, uint64_t budget_bytes);
void server_destroy(Server* s);
int64_t server_submit_mult(Server& s, const Ciphertext& a, const Ciphertext& b, Ciphertext& out);
int64_t server_submit_rotate(Server& s, const Ciphertext& a, int index, Ciphertext& out);
int64_t server_submit_logreg(Server& s, const Ciphertext& x, const Plaintext& w, int n_rot, double c0, double c1,
                             double c3, Ciphertext& out);
int server_step(Server& s);
std::vector<int64_t> server_poll(Server& s);
void server_counts(const Server& s, int& active, int& queued, uint64_t& used_bytes);
// Device work of a hoisted rotation set (batch.hip): ks = automorphism indices.
void rotate_hoisted_batch(Context& c, const Ciphertext& a, const std::vector<uint32_t>& ks,
                          const std::vector<Ciphertext*>& outs);
// NTT jobs, NTT_INLINE per launch (ckks.hip).
void run_ntt_jobs(Context& c, std::vector<NttJob> jobs, bool inverse, int load = LOAD_PLAIN, int store = STORE_PLAIN);
// ModUp of x (level l, evaluation form) into ext via the scaled coefficient limbs t.
void modup_to(Context& c, int level, const uint64_t* x, uint64_t* t, uint64_t* ext);

// Stage entry points for parity tests (device pointers).
void poly_ntt(Context& c, uint64_t* data, int nlimbs, int first_pid, bool inverse);
void ntt_limbs(Context& c, uint64_t* data, int limbs, bool inverse);
// ModUp of x into ext with the caller's (level+1)*N-word workspace t, which
// every call overwrites before reading (stage_modup allocates it).
void stage_modup_into(Context& c, int level, const uint64_t* x, uint64_t* t, uint64_t* ext);
// Static-scratch variants (noalloc.hip): noalloc_init runs once per device
// from the context constructor with the device current.
void noalloc_init(int device);
void stage_modup_noalloc(Context& c, int level, const uint64_t* x, uint64_t* ext);
void stage_modup(Context& c, int level, const uint64_t* x, uint64_t* ext);
void stage_keyswitch(Context& c, int level, const uint64_t* x, const KeySwitchingKey& k, uint64_t* ks_b,
                     uint64_t* ks_a);
void stage_moddown(Context& c, int level, const uint64_t* x_ext, uint64_t* y);

// Serialization (serial.hip): the engine's binary files of RawParams,
// RawCipherText, RawPlainText and RawKeySwitchKey data, and FIDESlib's
// `.dev` sidecar text (api/Serialize.cpp).
struct DevSidecar {
    std::vector<int> devices;
    bool auto_load_ciphertexts = true;
    bool auto_load_plaintexts = false;
    std::vector<int> rotation_indexes;
    int key_dist = 0;
};
void params_save(const Context::Params& p, const std::string& path);
Context::Params params_load(const std::string& path);
void ct_save(Context& c, const Ciphertext& ct, const std::string& path);
void ct_load_file(Context& c, Ciphertext& ct, const std::string& path);
void pt_save(Context& c, const Plaintext& pt, const std::string& path);
void pt_load_file(Context& c, Plaintext& pt, const std::string& path);
void key_save(Context& c, int kind, int index, const std::string& path);
void key_load_file(Context& c, const std::string& path, int& kind, int& index);
void dev_sidecar_write(const std::string& path, const DevSidecar& s);
DevSidecar dev_sidecar_read(const std::string& path);

}  // namespace fides
