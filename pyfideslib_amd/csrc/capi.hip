// extern "C" entry points of libfides_amd.so (declared in include/fides_amd.h).
#include <cstring>
#include <string>

#include "../../include/fides_amd.h"
#include "context.hpp"

using namespace fides;

struct fides_context {
    Context* c;
};
struct fides_ciphertext {
    std::unique_ptr<Ciphertext> ct;
    fides_context* owner;
};
struct fides_plaintext {
    std::unique_ptr<Plaintext> pt;
    fides_context* owner;
};

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
    try {
        f();
        return FIDES_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return FIDES_E_ARG;
    }
}

// Every entry point that takes a context (or a server, through its context)
// holds that context's API lock for the call: calls on one context are
// serialized, calls on different contexts run concurrently.
template <class F>
static int guard_ctx(fides_context* ctx, F&& f) {
    if (!ctx || !ctx->c) {
        g_err = "null context";
        return FIDES_E_ARG;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->c->api_mu);
    return guard(std::forward<F>(f));
}

static Context& C(fides_context* ctx) {
    if (!ctx || !ctx->c) throw Error(E_ARG, "null context");
    return *ctx->c;
}
// Any form (the extended-basis entry points, load / copy / info).
static Ciphertext& CTX(const fides_ciphertext* ct) {
    if (!ct || !ct->ct) throw Error(E_ARG, "null ciphertext");
    return *ct->ct;
}
// The Q basis: every other entry point, for inputs and outputs alike.
static Ciphertext& CT(const fides_ciphertext* ct) {
    Ciphertext& c = CTX(ct);
    if (c.ext) throw Error(E_STATE, "ciphertext in P*Q form: fides_ct_mod_down it first");
    return c;
}
static void same_ctx(fides_context* ctx, const fides_ciphertext* ct) {
    if (ct && ct->owner != ctx) throw Error(E_ARG, "ciphertext belongs to another context");
}

extern "C" {

int32_t fides_abi_version(void) { return 4; }
const char* fides_last_error(void) { return g_err.c_str(); }

static Context::Params params_of(const fides_raw_params* p) {
    if (!p || !p->moduli || !p->root_of_unity) throw Error(E_ARG, "null argument");
    if (p->K > 0 && (!p->special_moduli || !p->special_root_of_unity)) throw Error(E_ARG, "null special primes");
    Context::Params P;
    P.logN = p->logN;
    P.L = p->L;
    P.K = p->K;
    P.dnum = p->dnum;
    P.scaling_technique = p->scaling_technique;
    for (int i = 0; i <= p->L; ++i) {
        P.moduli.push_back(p->moduli[i]);
        P.roots.push_back(p->root_of_unity[i]);
    }
    for (int i = 0; i < p->K; ++i) {
        P.moduli.push_back(p->special_moduli[i]);
        P.roots.push_back(p->special_root_of_unity[i]);
    }
    if (p->mod_reduce_factor)
        for (int i = 0; i <= p->L; ++i) P.mod_reduce_factor.push_back(p->mod_reduce_factor[i]);
    if (p->scaling_factor_real)
        for (int i = 0; i <= p->L; ++i) P.scaling_factor_real.push_back(p->scaling_factor_real[i]);
    if (p->scaling_factor_real_big)
        for (int i = 0; i <= p->L; ++i) P.scaling_factor_real_big.push_back(p->scaling_factor_real_big[i]);
    return P;
}

static fides_context* make_context(const Context::Params& P, int32_t device) {
    auto* h = new fides_context{nullptr};
    try {
        h->c = new Context(P, device);
    } catch (...) {
        delete h;
        throw;
    }
    return h;
}

int fides_context_create(const fides_raw_params* p, int32_t device, fides_context** out) {
    return guard([&] {
        if (!out) throw Error(E_ARG, "null argument");
        Context::Params P = params_of(p);
        *out = make_context(P, device);
    });
}
int fides_params_save(const fides_raw_params* p, const char* path) {
    return guard([&] {
        if (!path) throw Error(E_ARG, "null path");
        params_save(params_of(p), path);
    });
}
int fides_context_save_params(fides_context* ctx, const char* path) {
    return guard_ctx(ctx, [&] {
        if (!path) throw Error(E_ARG, "null path");
        params_save(C(ctx).params, path);
    });
}
int fides_context_create_from_file(const char* path, int32_t device, fides_context** out) {
    return guard([&] {
        if (!path || !out) throw Error(E_ARG, "null argument");
        *out = make_context(params_load(path), device);
    });
}
int fides_context_destroy(fides_context* ctx) {
    return guard([&] {
        if (!ctx) return;
        delete ctx->c;
        delete ctx;
    });
}

int fides_synchronize(fides_context* ctx) { return guard_ctx(ctx, [&] { C(ctx).sync(); }); }
void* fides_context_stream(fides_context* ctx) { return ctx && ctx->c ? (void*)ctx->c->stream : nullptr; }
int fides_context_pool_bytes(fides_context* ctx, uint64_t* bytes) {
    return guard_ctx(ctx, [&] { *bytes = C(ctx).pool.bytes_held(); });
}
int fides_context_trim_pool(fides_context* ctx, uint64_t keep_bytes) {
    return guard_ctx(ctx, [&] { C(ctx).pool.trim(keep_bytes); });
}

int fides_device_alloc(fides_context* ctx, uint64_t bytes, void** ptr) {
    return guard_ctx(ctx, [&] {
        C(ctx);
        if (!ptr) throw Error(E_ARG, "null out");
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(E_OOM, "hipMalloc failed");
        }
    });
}
int fides_device_free(fides_context* ctx, void* ptr) {
    return guard_ctx(ctx, [&] {
        C(ctx).sync();
        hip_check(hipFree(ptr), "hipFree");
    });
}
int fides_memcpy_htod(fides_context* ctx, void* dst, const void* src, uint64_t bytes) {
    return guard_ctx(ctx, [&] {
        Context& c = C(ctx);
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream), "memcpy htod");
        c.sync();
    });
}
int fides_memcpy_dtoh(fides_context* ctx, void* dst, const void* src, uint64_t bytes) {
    return guard_ctx(ctx, [&] {
        Context& c = C(ctx);
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream), "memcpy dtoh");
        c.sync();
    });
}

int fides_event_create(fides_context* ctx, void** ev) {
    return guard_ctx(ctx, [&] {
        C(ctx);
        if (!ev) throw Error(E_ARG, "null out");
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        *ev = (void*)e;
    });
}
int fides_event_destroy(void* ev) {
    return guard([&] { hip_check(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy"); });
}
int fides_event_record(fides_context* ctx, void* ev) {
    return guard_ctx(ctx, [&] { hip_check(hipEventRecord((hipEvent_t)ev, C(ctx).stream), "hipEventRecord"); });
}
int fides_event_elapsed_ms(void* a, void* b, float* ms) {
    return guard([&] {
        hip_check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
    });
}

int fides_ct_create(fides_context* ctx, fides_ciphertext** out) {
    return guard_ctx(ctx, [&] {
        if (!out) throw Error(E_ARG, "null out");
        auto ct = ct_create(C(ctx));
        *out = new fides_ciphertext{std::move(ct), ctx};
    });
}
int fides_ct_destroy(fides_ciphertext* ct) {
    return guard([&] { delete ct; });
}

static int load_impl(fides_context* ctx, fides_ciphertext* ct, const uint64_t* s0, const uint64_t* s1, int32_t nr,
                     double noise, int32_t nl, int32_t slots, bool dev) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        if (!s0 || !s1) throw Error(E_ARG, "null limbs");
        ct_load(C(ctx), CTX(ct), s0, s1, nr, noise, nl, slots, dev);
    });
}
int fides_ct_load(fides_context* ctx, fides_ciphertext* ct, const uint64_t* s0, const uint64_t* s1, int32_t nr,
                  double noise, int32_t nl, int32_t slots) {
    return load_impl(ctx, ct, s0, s1, nr, noise, nl, slots, false);
}
int fides_ct_load_device(fides_context* ctx, fides_ciphertext* ct, const uint64_t* s0, const uint64_t* s1,
                         int32_t nr, double noise, int32_t nl, int32_t slots) {
    return load_impl(ctx, ct, s0, s1, nr, noise, nl, slots, true);
}
int fides_ct_store(fides_context* ctx, const fides_ciphertext* ct, uint64_t* s0, uint64_t* s1) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        if (!s0 || !s1) throw Error(E_ARG, "null output");
        if (CT(ct).level < 0) throw Error(E_STATE, "ciphertext not loaded");
        ct_store(C(ctx), CT(ct), s0, s1);
    });
}
int fides_ct_info(const fides_ciphertext* ct, int32_t* level, double* noise, int32_t* noise_level, int32_t* slots) {
    return guard([&] {
        const Ciphertext& c = CTX(ct);
        if (level) *level = c.level;
        if (noise) *noise = c.noise;
        if (noise_level) *noise_level = c.noise_level;
        if (slots) *slots = c.slots;
    });
}
int fides_ct_copy(fides_context* ctx, const fides_ciphertext* src, fides_ciphertext* dst) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, src);
        same_ctx(ctx, dst);
        ct_copy(C(ctx), CTX(src), CTX(dst));
    });
}
int fides_ct_device_ptrs(const fides_ciphertext* ct, uint64_t** c0, uint64_t** c1) {
    return guard([&] {
        if (c0) *c0 = CTX(ct).c0();
        if (c1) *c1 = CTX(ct).c1();
    });
}

static Plaintext& PT(const fides_plaintext* pt, fides_context* ctx) {
    if (!pt || !pt->pt) throw Error(E_ARG, "null plaintext");
    if (pt->owner != ctx) throw Error(E_ARG, "plaintext belongs to another context");
    return *pt->pt;
}
int fides_pt_create(fides_context* ctx, fides_plaintext** out) {
    return guard_ctx(ctx, [&] {
        if (!out) throw Error(E_ARG, "null out");
        auto pt = pt_create(C(ctx));
        *out = new fides_plaintext{std::move(pt), ctx};
    });
}
int fides_pt_destroy(fides_plaintext* pt) {
    return guard([&] { delete pt; });
}
int fides_pt_load(fides_context* ctx, fides_plaintext* pt, const uint64_t* limbs, int32_t nr, double noise,
                  int32_t nl, int32_t slots) {
    return guard_ctx(ctx, [&] { pt_load(C(ctx), PT(pt, ctx), limbs, nr, noise, nl, slots, false); });
}
int fides_pt_load_ext(fides_context* ctx, fides_plaintext* pt, const uint64_t* limbs, int32_t nr, double noise,
                      int32_t nl, int32_t slots) {
    return guard_ctx(ctx, [&] { pt_load_ext(C(ctx), PT(pt, ctx), limbs, nr, noise, nl, slots, false); });
}
int fides_eval_mult_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                          fides_ciphertext* out, int32_t rescale) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_mult_plain(C(ctx), CT(a), PT(pt, ctx), CT(out), rescale != 0);
    });
}
int fides_eval_add_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                         fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_add_plain(C(ctx), CT(a), PT(pt, ctx), CT(out));
    });
}
int fides_eval_mult_scalar(fides_context* ctx, const fides_ciphertext* a, double v, fides_ciphertext* out,
                           int32_t rescale) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_mult_scalar(C(ctx), CT(a), v, CT(out), rescale != 0);
    });
}
int fides_eval_add_scalar(fides_context* ctx, const fides_ciphertext* a, double v, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_add_scalar(C(ctx), CT(a), v, CT(out));
    });
}
int fides_drop_to_level(fides_context* ctx, fides_ciphertext* ct, int32_t level) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        drop_to_level(C(ctx), CT(ct), level);
    });
}
int fides_scalar_encoding(fides_context* ctx, int32_t level, double v, int32_t kind, int32_t noise_deg,
                          uint64_t* out) {
    return guard_ctx(ctx, [&] {
        if (!out) throw Error(E_ARG, "null out");
        Context& c = C(ctx);
        if (kind != 0 && kind != 1) throw Error(E_ARG, "kind must be 0 (mult) or 1 (add)");
        const std::vector<uint64_t> e =
            kind == 0 ? elem_for_mult(c, level, v) : elem_for_add(c, level, std::fabs(v), noise_deg);
        std::memcpy(out, e.data(), e.size() * 8);
    });
}

int fides_key_load_eval(fides_context* ctx, const uint64_t* a, const uint64_t* b) {
    return guard_ctx(ctx, [&] {
        if (!a || !b) throw Error(E_ARG, "null key");
        auto k = std::make_unique<KeySwitchingKey>();
        load_key(C(ctx), *k, a, b, false);
        C(ctx).drop_plans();
        C(ctx).eval_key = std::move(k);
    });
}
int fides_key_load_automorph(fides_context* ctx, uint32_t kidx, const uint64_t* a, const uint64_t* b) {
    return guard_ctx(ctx, [&] {
        Context& c = C(ctx);
        if (!a || !b) throw Error(E_ARG, "null key");
        if ((kidx & 1) == 0 || kidx >= 2u * c.N) throw Error(E_ARG, "automorphism index must be odd and < 2N");
        auto k = std::make_unique<KeySwitchingKey>();
        load_key(c, *k, a, b, false);
        c.drop_plans();
        c.rot_keys[kidx] = std::move(k);
    });
}
int fides_key_load_switch(fides_context* ctx, int32_t id, const uint64_t* a, const uint64_t* b) {
    return guard_ctx(ctx, [&] {
        Context& c = C(ctx);
        if (!a || !b) throw Error(E_ARG, "null key");
        auto k = std::make_unique<KeySwitchingKey>();
        load_key(c, *k, a, b, false);
        c.drop_plans();
        c.switch_keys[id] = std::move(k);
    });
}
int fides_eval_key_switch(fides_context* ctx, const fides_ciphertext* a, int32_t id, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_key_switch(C(ctx), CT(a), id, CT(out));
    });
}
int fides_key_load_rotation(fides_context* ctx, int32_t index, const uint64_t* a, const uint64_t* b) {
    int rc = guard([&] { C(ctx); });
    if (rc) return rc;
    return fides_key_load_automorph(ctx, ctx->c->rot_index(index), a, b);
}
int fides_rotation_automorph_index(fides_context* ctx, int32_t index, uint32_t* k) {
    return guard_ctx(ctx, [&] { *k = C(ctx).rot_index(index); });
}

int fides_eval_mult(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out,
                    int32_t rs) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a); same_ctx(ctx, b); same_ctx(ctx, out);
        eval_mult(C(ctx), CT(a), CT(b), CT(out), rs != 0);
    });
}
int fides_eval_square(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out, int32_t rs) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a); same_ctx(ctx, out);
        eval_square(C(ctx), CT(a), CT(out), rs != 0);
    });
}
int fides_rescale(fides_context* ctx, fides_ciphertext* ct) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        rescale(C(ctx), CT(ct));
    });
}
int fides_eval_add(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a); same_ctx(ctx, b); same_ctx(ctx, out);
        eval_add(C(ctx), CT(a), CT(b), CT(out));
    });
}
int fides_eval_sub(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a); same_ctx(ctx, b); same_ctx(ctx, out);
        eval_sub(C(ctx), CT(a), CT(b), CT(out));
    });
}
int fides_eval_rotate(fides_context* ctx, const fides_ciphertext* a, int32_t index, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a); same_ctx(ctx, out);
        eval_rotate(C(ctx), CT(a), index, CT(out));
    });
}
int fides_eval_rotate_hoisted(fides_context* ctx, const fides_ciphertext* a, const int32_t* idx, int32_t n,
                              fides_ciphertext* const* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        if (n < 0 || (n > 0 && (!idx || !out))) throw Error(E_ARG, "bad index list");
        std::vector<int> ix(idx, idx + n);
        std::vector<Ciphertext*> outs;
        for (int i = 0; i < n; ++i) {
            same_ctx(ctx, out[i]);
            outs.push_back(&CT(out[i]));
        }
        eval_rotate_hoisted(C(ctx), CT(a), ix, outs);
    });
}
int fides_eval_mult_batch(fides_context* ctx, const fides_ciphertext* const* a, const fides_ciphertext* const* b,
                          fides_ciphertext* const* out, int32_t n, int32_t rs) {
    return guard_ctx(ctx, [&] {
        if (n < 0 || (n > 0 && (!a || !b || !out))) throw Error(E_ARG, "bad batch");
        std::vector<const Ciphertext*> va, vb;
        std::vector<Ciphertext*> vo;
        for (int i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]); same_ctx(ctx, b[i]); same_ctx(ctx, out[i]);
            va.push_back(&CT(a[i]));
            vb.push_back(&CT(b[i]));
            vo.push_back(&CT(out[i]));
        }
        eval_mult_batch(C(ctx), va, vb, vo, rs != 0);
    });
}

int fides_eval_rotate_batch(fides_context* ctx, const fides_ciphertext* const* a, const int32_t* indices,
                            fides_ciphertext* const* out, int32_t n) {
    return guard_ctx(ctx, [&] {
        if (n < 0 || (n > 0 && (!a || !indices || !out))) throw Error(E_ARG, "bad batch");
        std::vector<const Ciphertext*> va;
        std::vector<int> vi;
        std::vector<Ciphertext*> vo;
        for (int i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]); same_ctx(ctx, out[i]);
            va.push_back(&CT(a[i]));
            vi.push_back(indices[i]);
            vo.push_back(&CT(out[i]));
        }
        eval_rotate_batch(C(ctx), va, vi, vo);
    });
}

int fides_eval_linear_transform(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* const* diag,
                                int32_t n, int32_t bstep, int32_t stride, int32_t offset, int32_t flags,
                                fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        if (n < 1 || !diag) throw Error(E_ARG, "no diagonals");
        std::vector<const Plaintext*> d;
        for (int i = 0; i < n; ++i) d.push_back(&PT(diag[i], ctx));
        if (flags & ~FIDES_LT_BATCHED_GIANT) throw Error(E_ARG, "unknown linear transform flags");
        eval_linear_transform(C(ctx), CT(a), d, bstep, stride, offset, CT(out), (flags & FIDES_LT_BATCHED_GIANT) != 0);
    });
}

int fides_eval_linear_transform_batch(fides_context* ctx, const fides_ciphertext* const* a, int32_t m,
                                      const fides_plaintext* const* diag, int32_t n, int32_t bstep, int32_t stride,
                                      int32_t offset, int32_t flags, fides_ciphertext* const* out) {
    return guard_ctx(ctx, [&] {
        if (m < 1 || !a || !out) throw Error(E_ARG, "empty batch");
        if (n < 1 || !diag) throw Error(E_ARG, "no diagonals");
        if (flags & ~FIDES_LT_BATCHED_GIANT) throw Error(E_ARG, "unknown linear transform flags");
        std::vector<const Ciphertext*> A;
        std::vector<Ciphertext*> O;
        for (int32_t i = 0; i < m; ++i) {
            same_ctx(ctx, a[i]);
            same_ctx(ctx, out[i]);
            A.push_back(&CT(a[i]));
            O.push_back(&CT(out[i]));
        }
        for (int32_t i = 0; i < m; ++i)
            for (int32_t j = 0; j < m; ++j)
                if (i != j && (O[i] == O[j] || O[i] == A[j])) throw Error(E_ARG, "batch output aliases another item");
        std::vector<const Plaintext*> d;
        for (int i = 0; i < n; ++i) d.push_back(&PT(diag[i], ctx));
        eval_linear_transform_batch(C(ctx), A, d, bstep, stride, offset, O, (flags & FIDES_LT_BATCHED_GIANT) != 0);
    });
}

int fides_eval_add_many(fides_context* ctx, const fides_ciphertext* const* a, int32_t n, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        if (n < 1 || !a) throw Error(E_ARG, "no ciphertexts");
        std::vector<const Ciphertext*> x;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]);
            x.push_back(&CT(a[i]));
        }
        same_ctx(ctx, out);
        eval_add_many(C(ctx), x, CT(out));
    });
}
int fides_eval_negate(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_negate(C(ctx), CT(a), CT(out));
    });
}
int fides_accumulate_rotations(int32_t size, int32_t stride, int32_t* idx, int32_t cap, int32_t* n) {
    return guard([&] {
        if (!n) throw Error(E_ARG, "null count");
        const std::vector<int> r = accumulate_rotations(size, stride);
        *n = (int32_t)r.size();
        if (idx)
            for (int32_t i = 0; i < (int32_t)r.size() && i < cap; ++i) idx[i] = r[i];
    });
}
int fides_convolution_rotations(int32_t stride, int32_t gstep, int32_t* out, int32_t cap, int32_t* n) {
    return guard([&] {
        if (!n) throw Error(E_ARG, "null count");
        if (gstep < 1) throw Error(E_ARG, "gstep must be at least 1");
        const std::vector<int> r = convolution_rotations(stride, gstep);
        *n = (int32_t)r.size();
        if (out)
            for (int32_t i = 0; i < (int32_t)r.size() && i < cap; ++i) out[i] = r[i];
    });
}
int fides_eval_convolution_transform(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* const* pts,
                                     int32_t n, int32_t bstep, int32_t stride, const int32_t* idx, int32_t gstep,
                                     fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        if (n < 1 || !pts) throw Error(E_ARG, "no plaintexts");
        if (bstep < 1 || !idx) throw Error(E_ARG, "no baby-step indexes");
        std::vector<const Plaintext*> p;
        for (int32_t i = 0; i < n; ++i) p.push_back(&PT(pts[i], ctx));
        eval_convolution_transform(C(ctx), CT(a), p, bstep, stride, std::vector<int>(idx, idx + bstep), gstep,
                                   CT(out));
    });
}
int fides_eval_accumulate_sum(fides_context* ctx, const fides_ciphertext* a, int32_t size, int32_t stride,
                              fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_accumulate_sum(C(ctx), CT(a), size, stride, CT(out));
    });
}

int fides_ct_save(fides_context* ctx, const fides_ciphertext* ct, const char* path) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        if (!path) throw Error(E_ARG, "null path");
        ct_save(C(ctx), CTX(ct), path);
    });
}
int fides_ct_load_file(fides_context* ctx, fides_ciphertext* ct, const char* path) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        if (!path) throw Error(E_ARG, "null path");
        ct_load_file(C(ctx), CTX(ct), path);
    });
}
int fides_pt_save(fides_context* ctx, const fides_plaintext* pt, const char* path) {
    return guard_ctx(ctx, [&] {
        if (!path) throw Error(E_ARG, "null path");
        pt_save(C(ctx), PT(pt, ctx), path);
    });
}
int fides_pt_load_file(fides_context* ctx, fides_plaintext* pt, const char* path) {
    return guard_ctx(ctx, [&] {
        if (!path) throw Error(E_ARG, "null path");
        pt_load_file(C(ctx), PT(pt, ctx), path);
    });
}
int fides_key_save(fides_context* ctx, int32_t kind, int32_t index, const char* path) {
    return guard_ctx(ctx, [&] {
        if (!path) throw Error(E_ARG, "null path");
        key_save(C(ctx), kind, index, path);
    });
}
int fides_key_load_file(fides_context* ctx, const char* path, int32_t* kind, int32_t* index) {
    return guard_ctx(ctx, [&] {
        if (!path) throw Error(E_ARG, "null path");
        int k = 0, i = 0;
        key_load_file(C(ctx), path, k, i);
        if (kind) *kind = k;
        if (index) *index = i;
    });
}
int fides_dev_sidecar_write(const char* path, int32_t n_devices, const int32_t* devices, int32_t auto_load_ciphertexts,
                            int32_t auto_load_plaintexts, int32_t n_rotations, const int32_t* rotation_indexes,
                            int32_t key_dist) {
    return guard([&] {
        if (!path || n_devices < 0 || n_rotations < 0 || (n_devices && !devices) || (n_rotations && !rotation_indexes))
            throw Error(E_ARG, "bad argument");
        DevSidecar s;
        s.devices.assign(devices, devices + n_devices);
        s.auto_load_ciphertexts = auto_load_ciphertexts != 0;
        s.auto_load_plaintexts = auto_load_plaintexts != 0;
        s.rotation_indexes.assign(rotation_indexes, rotation_indexes + n_rotations);
        s.key_dist = key_dist;
        dev_sidecar_write(path, s);
    });
}
int fides_dev_sidecar_read(const char* path, int32_t* devices, int32_t dev_cap, int32_t* n_devices,
                           int32_t* auto_load_ciphertexts, int32_t* auto_load_plaintexts, int32_t* rotation_indexes,
                           int32_t rot_cap, int32_t* n_rotations, int32_t* key_dist) {
    return guard([&] {
        if (!path) throw Error(E_ARG, "null path");
        const DevSidecar s = dev_sidecar_read(path);
        if (n_devices) *n_devices = (int32_t)s.devices.size();
        for (int32_t i = 0; devices && i < dev_cap && i < (int32_t)s.devices.size(); ++i) devices[i] = s.devices[i];
        if (auto_load_ciphertexts) *auto_load_ciphertexts = s.auto_load_ciphertexts;
        if (auto_load_plaintexts) *auto_load_plaintexts = s.auto_load_plaintexts;
        if (n_rotations) *n_rotations = (int32_t)s.rotation_indexes.size();
        for (int32_t i = 0; rotation_indexes && i < rot_cap && i < (int32_t)s.rotation_indexes.size(); ++i)
            rotation_indexes[i] = s.rotation_indexes[i];
        if (key_dist) *key_dist = s.key_dist;
    });
}

int fides_context_set_op_graphs(fides_context* ctx, int32_t on) {
    return guard_ctx(ctx, [&] { C(ctx).op_graphs = on != 0; });
}
int fides_ct_is_ext(const fides_ciphertext* ct, int32_t* ext) {
    return guard([&] {
        if (!ext) throw Error(E_ARG, "null output");
        *ext = CTX(ct).ext;
    });
}
int fides_ct_store_ext(fides_context* ctx, const fides_ciphertext* ct, uint64_t* s0, uint64_t* s1) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, ct);
        if (!s0 || !s1) throw Error(E_ARG, "null output");
        const Ciphertext& x = CTX(ct);
        if (x.level < 0) throw Error(E_STATE, "ciphertext not loaded");
        if (!x.ext) throw Error(E_STATE, "ciphertext not in P*Q form");
        Context& c = C(ctx);
        const size_t bytes = (size_t)(x.level + 1 + c.K) * c.N * 8;
        hip_check(hipMemcpyAsync(s0, x.c0(), bytes, hipMemcpyDeviceToHost, c.stream), "store c0");
        hip_check(hipMemcpyAsync(s1, x.c1(), bytes, hipMemcpyDeviceToHost, c.stream), "store c1");
        c.sync();
    });
}
int fides_ct_extend(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_extend(C(ctx), CT(a), CTX(out));
    });
}
int fides_ct_mod_down(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_mod_down(C(ctx), CTX(a), CTX(out));
    });
}
int fides_eval_rotate_hoisted_ext(fides_context* ctx, const fides_ciphertext* a, const int32_t* indices, int32_t n,
                                  fides_ciphertext* const* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        if (n < 1 || !indices || !out) throw Error(E_ARG, "no indices");
        std::vector<int> idx(indices, indices + n);
        std::vector<Ciphertext*> outs;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, out[i]);
            outs.push_back(&CTX(out[i]));
        }
        eval_rotate_hoisted_ext(C(ctx), CT(a), idx, outs);
    });
}
int fides_eval_add_ext(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b,
                       fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, b);
        same_ctx(ctx, out);
        eval_add_ext(C(ctx), CTX(a), CTX(b), CTX(out));
    });
}
int fides_eval_dot_product_plain(fides_context* ctx, const fides_ciphertext* const* a,
                                 const fides_plaintext* const* pt, int32_t n, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        if (n < 1 || !a || !pt) throw Error(E_ARG, "no terms");
        std::vector<const Ciphertext*> x;
        std::vector<const Plaintext*> p;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]);
            x.push_back(&CTX(a[i]));
            p.push_back(&PT(pt[i], ctx));
        }
        same_ctx(ctx, out);
        eval_dot_product_plain(C(ctx), x, p, CTX(out));
    });
}

int fides_eval_dot_product(fides_context* ctx, const fides_ciphertext* const* a, const fides_ciphertext* const* b,
                           int32_t n, int32_t ext, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        if (n < 1 || !a || !b) throw Error(E_ARG, "no terms");
        std::vector<const Ciphertext*> x, y;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]);
            same_ctx(ctx, b[i]);
            x.push_back(&CTX(a[i]));
            y.push_back(&CTX(b[i]));
        }
        same_ctx(ctx, out);
        eval_dot_product(C(ctx), x, y, ext != 0, CTX(out));
    });
}

int fides_eval_sub_plain(fides_context* ctx, const fides_ciphertext* a, const fides_plaintext* pt,
                         fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_sub_plain(C(ctx), CT(a), PT(pt, ctx), CT(out));
    });
}
int fides_eval_add_mult_plain(fides_context* ctx, fides_ciphertext* acc, const fides_ciphertext* a,
                              const fides_plaintext* pt, int32_t rescale) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, acc);
        same_ctx(ctx, a);
        eval_add_mult_plain(C(ctx), CT(acc), CT(a), PT(pt, ctx), rescale != 0);
    });
}
int fides_eval_add_mult_scalar(fides_context* ctx, fides_ciphertext* acc, const fides_ciphertext* b, double d) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, acc);
        same_ctx(ctx, b);
        eval_add_mult_scalar(C(ctx), CT(acc), CT(b), d);
    });
}
int fides_eval_linear_wsum(fides_context* ctx, const fides_ciphertext* const* a, const double* w, int32_t n,
                           fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        if (n < 1 || !a || !w) throw Error(E_ARG, "no terms");
        std::vector<const Ciphertext*> x;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, a[i]);
            x.push_back(&CT(a[i]));
        }
        same_ctx(ctx, out);
        eval_linear_wsum(C(ctx), x, std::vector<double>(w, w + n), CT(out));
    });
}
int fides_eval_mult_ext(fides_context* ctx, const fides_ciphertext* a, const fides_ciphertext* b,
                        fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, b);
        same_ctx(ctx, out);
        eval_mult_ext(C(ctx), CT(a), CT(b), CTX(out));
    });
}

int fides_eval_automorph(fides_context* ctx, const fides_ciphertext* a, uint32_t k, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_automorph(C(ctx), CT(a), k, CT(out));
    });
}
int fides_eval_conjugate(fides_context* ctx, const fides_ciphertext* a, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_automorph(C(ctx), CT(a), 2u * C(ctx).N - 1, CT(out));
    });
}
int fides_mod_raise(fides_context* ctx, const fides_ciphertext* a, int32_t level, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        mod_raise(C(ctx), CT(a), level, CT(out));
    });
}
int fides_eval_mult_monomial(fides_context* ctx, const fides_ciphertext* a, int64_t power, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_mult_monomial(C(ctx), CT(a), power, CT(out));
    });
}
int fides_eval_mult_int(fides_context* ctx, const fides_ciphertext* a, uint64_t s, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_mult_int(C(ctx), CT(a), s, CT(out));
    });
}

int fides_eval_horner_series(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                             fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, x);
        same_ctx(ctx, out);
        if (n < 1 || !coeffs) throw Error(E_ARG, "no coefficients");
        eval_horner_series(C(ctx), CT(x), std::vector<double>(coeffs, coeffs + n), CT(out));
    });
}
int fides_eval_double_angle(fides_context* ctx, const fides_ciphertext* a, int32_t iterations,
                            fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, a);
        same_ctx(ctx, out);
        eval_double_angle(C(ctx), CT(a), iterations, CT(out));
    });
}

int fides_eval_chebyshev_series(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                                double a, double b, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, x);
        same_ctx(ctx, out);
        if (n < 1 || !coeffs) throw Error(E_ARG, "no coefficients");
        eval_chebyshev_series(C(ctx), CT(x), std::vector<double>(coeffs, coeffs + n), a, b, CT(out));
    });
}

int fides_eval_mod_reduction(fides_context* ctx, const fides_ciphertext* x, const double* coeffs, int32_t n,
                             int32_t double_angle_its, uint64_t post, fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        same_ctx(ctx, x);
        same_ctx(ctx, out);
        if (n < 1 || !coeffs) throw Error(E_ARG, "no coefficients");
        if (double_angle_its < 0) throw Error(E_ARG, "negative double-angle iterations");
        eval_mod_reduction(C(ctx), CT(x), std::vector<double>(coeffs, coeffs + n), double_angle_its, post, CT(out));
    });
}

static int encode_impl(fides_context* ctx, const double* re, const double* im, int32_t level, double scale,
                       fides_plaintext* pt, bool ext) {
    return guard_ctx(ctx, [&] {
        if (!re) throw Error(E_ARG, "null slot values");
        Context& c = C(ctx);
        std::vector<std::complex<double>> z(c.N / 2);
        for (size_t j = 0; j < z.size(); ++j) z[j] = {re[j], im ? im[j] : 0.0};
        PT(pt, ctx);
        pt->pt = encode_slots(c, z, level, scale, ext);
    });
}
int fides_encode(fides_context* ctx, const double* re, const double* im, int32_t level, double scale,
                 fides_plaintext* pt) {
    return encode_impl(ctx, re, im, level, scale, pt, false);
}
int fides_encode_ext(fides_context* ctx, const double* re, const double* im, int32_t level, double scale,
                     fides_plaintext* pt) {
    return encode_impl(ctx, re, im, level, scale, pt, true);
}

struct fides_bootstrap {
    BootPlanPtr plan;
    fides_context* owner;
};

int fides_bootstrap_setup(fides_context* ctx, const fides_bootstrap_params* p, fides_bootstrap** out) {
    return guard_ctx(ctx, [&] {
        if (!p || !out) throw Error(E_ARG, "null argument");
        BootPlanPtr plan = bootstrap_setup(C(ctx), p->scale, p->level_budget_cts, p->level_budget_stc, p->K,
                                           p->double_angle, p->degree, p->top_level, p->sparse_encaps != 0,
                                           p->slots, p->batched_giant != 0, p->q_basis_diagonals == 0);
        *out = new fides_bootstrap{std::move(plan), ctx};
    });
}
int fides_bootstrap_rotations(const fides_bootstrap* b, int32_t* idx, int32_t cap, int32_t* n) {
    return guard([&] {
        if (!b || !b->plan || !n) throw Error(E_ARG, "null argument");
        const std::vector<int>& r = bootstrap_rotations(*b->plan);
        *n = (int32_t)r.size();
        if (idx)
            for (int32_t i = 0; i < (int32_t)r.size() && i < cap; ++i) idx[i] = r[i];
    });
}
int fides_eval_bootstrap(fides_context* ctx, const fides_bootstrap* b, const fides_ciphertext* in,
                         fides_ciphertext* out) {
    return guard_ctx(ctx, [&] {
        if (!b || !b->plan) throw Error(E_ARG, "null bootstrap");
        if (b->owner != ctx) throw Error(E_ARG, "bootstrap belongs to another context");
        same_ctx(ctx, in);
        same_ctx(ctx, out);
        eval_bootstrap(C(ctx), *b->plan, CT(in), CT(out));
    });
}
int fides_eval_bootstrap_batch(fides_context* ctx, const fides_bootstrap* b, const fides_ciphertext* const* in,
                               fides_ciphertext* const* out, int32_t n) {
    return guard_ctx(ctx, [&] {
        if (!b || !b->plan) throw Error(E_ARG, "null bootstrap");
        if (b->owner != ctx) throw Error(E_ARG, "bootstrap belongs to another context");
        if (n < 1 || !in || !out) throw Error(E_ARG, "empty batch");
        std::vector<const Ciphertext*> A;
        std::vector<Ciphertext*> O;
        for (int32_t i = 0; i < n; ++i) {
            same_ctx(ctx, in[i]);
            same_ctx(ctx, out[i]);
            A.push_back(&CT(in[i]));
            O.push_back(&CT(out[i]));
        }
        for (int32_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < n; ++j)
                if (i != j && (O[i] == O[j] || O[i] == A[j])) throw Error(E_ARG, "batch output aliases another item");
        eval_bootstrap_batch(C(ctx), *b->plan, A, O);
    });
}
int fides_bootstrap_destroy(fides_bootstrap* b) {
    return guard([&] { delete b; });
}

struct fides_server {
    Server* s;
    fides_context* owner;
    std::vector<int64_t> carry;  // completions beyond `cap` of an earlier poll
};
static Server& SV(fides_server* s) {
    if (!s || !s->s) throw Error(E_ARG, "null server");
    return *s->s;
}
int fides_server_create(fides_context* ctx, int32_t max_slots, uint64_t budget_bytes, fides_server** out) {
    return guard_ctx(ctx, [&] {
        if (!out) throw Error(E_ARG, "null out");
        Server* s = server_create(C(ctx), max_slots, budget_bytes);
        *out = new fides_server{s, ctx, {}};
    });
}
int fides_server_destroy(fides_server* s) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        if (!s) return;
        server_destroy(s->s);
        delete s;
    });
}
int fides_server_submit_mult(fides_server* s, const fides_ciphertext* a, const fides_ciphertext* b,
                             fides_ciphertext* out, int64_t* job) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        Server& v = SV(s);
        same_ctx(s->owner, a); same_ctx(s->owner, b); same_ctx(s->owner, out);
        const int64_t id = server_submit_mult(v, CT(a), CT(b), CT(out));
        if (job) *job = id;
    });
}
int fides_server_submit_rotate(fides_server* s, const fides_ciphertext* a, int32_t index, fides_ciphertext* out,
                               int64_t* job) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        Server& v = SV(s);
        same_ctx(s->owner, a); same_ctx(s->owner, out);
        const int64_t id = server_submit_rotate(v, CT(a), index, CT(out));
        if (job) *job = id;
    });
}
int fides_server_submit_logreg(fides_server* s, const fides_ciphertext* x, const fides_plaintext* w, int32_t n_rot,
                               const double* coeffs, fides_ciphertext* out, int64_t* job) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        Server& v = SV(s);
        if (!coeffs) throw Error(E_ARG, "null coefficients");
        same_ctx(s->owner, x); same_ctx(s->owner, out);
        const int64_t id =
            server_submit_logreg(v, CT(x), PT(w, s->owner), n_rot, coeffs[0], coeffs[1], coeffs[2], CT(out));
        if (job) *job = id;
    });
}
int fides_server_step(fides_server* s, int32_t* pending) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        const int p = server_step(SV(s));
        if (pending) *pending = p;
    });
}
int fides_server_poll(fides_server* s, int64_t* done, int32_t cap, int32_t* n_done) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        Server& v = SV(s);
        if (!n_done || (cap > 0 && !done)) throw Error(E_ARG, "null output");
        std::vector<int64_t>& carry = s->carry;
        std::vector<int64_t> ids = server_poll(v);
        carry.insert(carry.end(), ids.begin(), ids.end());
        const int n = std::min<int>((int)carry.size(), cap);
        for (int i = 0; i < n; ++i) done[i] = carry[i];
        carry.erase(carry.begin(), carry.begin() + n);
        *n_done = n;
    });
}
int fides_server_counts(fides_server* s, int32_t* active, int32_t* queued, uint64_t* used_bytes) {
    return guard_ctx(s ? s->owner : nullptr, [&] {
        int a, q;
        uint64_t u;
        server_counts(SV(s), a, q, u);
        if (active) *active = a;
        if (queued) *queued = q;
        if (used_bytes) *used_bytes = u;
    });
}

int fides_poly_ntt(fides_context* ctx, uint64_t* data, int32_t nl, int32_t first, int32_t inverse) {
    return guard_ctx(ctx, [&] {
        if (!data) throw Error(E_ARG, "null data");
        poly_ntt(C(ctx), data, nl, first, inverse != 0);
    });
}
int fides_stage_modup(fides_context* ctx, int32_t level, const uint64_t* x, uint64_t* ext) {
    return guard_ctx(ctx, [&] {
        if (!x || !ext) throw Error(E_ARG, "null buffer");
        stage_modup(C(ctx), level, x, ext);
    });
}
int fides_ntt_limbs(fides_context* ctx, uint64_t* data, int32_t limbs, int32_t inverse) {
    return guard_ctx(ctx, [&] {
        if (!data) throw Error(E_ARG, "null data");
        ntt_limbs(C(ctx), data, limbs, inverse != 0);
    });
}
int fides_tune_load(const char* path, int32_t* entries) {
    return guard([&] {
        const int n = tune_load(path);
        if (n < 0) throw Error(E_ARG, std::string("cannot read tuning table ") + path);
        if (entries) *entries = n;
    });
}
int fides_tune_force(int32_t kind, const int32_t* cfg) {
    return guard([&] {
        if (kind >= TUNE_KINDS) throw Error(E_ARG, "unknown tuning kind");
        if (kind < 0 || !cfg) {
            tune_force(kind, nullptr);
            return;
        }
        TuneCfg c;
        for (int i = 0; i < TUNE_NV; ++i) c.v[i] = cfg[i];
        const bool ntt = kind == TUNE_NTT_FWD || kind == TUNE_NTT_INV;
        auto w = [](int x) { return x == 4 || x == 8 || x == 16; };
        const bool conv_ok = c.v[0] >= 1 && c.v[0] <= 64 && (c.v[1] == 0 || (kind == TUNE_MODUP && c.v[1] == 1));
        if (ntt ? !(w(c.v[0]) && (c.v[1] == 2 || c.v[1] == 3) && w(c.v[2]) && c.v[3] >= 0) : !conv_ok)
            throw Error(E_ARG, "invalid launch configuration");
        tune_force(kind, &c);
    });
}
int fides_tune_config(fides_context* ctx, int32_t kind, int32_t size, int32_t* cfg, int32_t* from_table) {
    return guard_ctx(ctx, [&] {
        const Context& c = C(ctx);
        if (!cfg || kind < 0 || kind >= TUNE_KINDS || size < 1) throw Error(E_ARG, "invalid query");
        TuneCfg t;
        const bool tab = tune_get(kind, c.logN, size, t);
        if (!tab) {
            if (kind == TUNE_NTT_FWD || kind == TUNE_NTT_INV) {
                t = ntt_rule_cfg(kind == TUNE_NTT_FWD, c.logN, size);
            } else {
                const int targets = kind == TUNE_MODUP ? c.L + 1 + c.K - c.alpha : c.L + 1;
                t = TuneCfg{{conv_rule_tg((c.N / 256) * size, targets), 0, 0, 0}};
            }
        }
        for (int i = 0; i < TUNE_NV; ++i) cfg[i] = t.v[i];
        if (from_table) *from_table = tab ? 1 : 0;
    });
}
int fides_stage_modup_noalloc(fides_context* ctx, int32_t level, const uint64_t* x, uint64_t* ext) {
    return guard_ctx(ctx, [&] {
        if (!x || !ext) throw Error(E_ARG, "null buffer");
        stage_modup_noalloc(C(ctx), level, x, ext);
    });
}
int fides_stage_moddown(fides_context* ctx, int32_t level, const uint64_t* x, uint64_t* y) {
    return guard_ctx(ctx, [&] {
        if (!x || !y) throw Error(E_ARG, "null buffer");
        stage_moddown(C(ctx), level, x, y);
    });
}
int fides_stage_keyswitch(fides_context* ctx, int32_t level, const uint64_t* x, uint32_t rk, uint64_t* kb,
                          uint64_t* ka) {
    return guard_ctx(ctx, [&] {
        Context& c = C(ctx);
        if (!x || !kb || !ka) throw Error(E_ARG, "null buffer");
        const KeySwitchingKey* key = nullptr;
        if (rk == 0) {
            if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
            key = c.eval_key.get();
        } else {
            auto it = c.rot_keys.find(rk);
            if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for that automorphism index");
            key = it->second.get();
        }
        stage_keyswitch(c, level, x, *key, kb, ka);
    });
}

}  // extern "C"
