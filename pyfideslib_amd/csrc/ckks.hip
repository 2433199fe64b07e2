// CKKS evaluation ops: launch sequences over the kernels in ntt.hip and
// kernels.hip.  Each op mirrors a FIDESlib CKKS::Ciphertext method and keeps
// OpenFHE's arithmetic order so that every residue is bit-exact:
//   mult      tensor -> ModUp(d2) -> KSK inner product -> ModDown -> add
//             (Ciphertext.cpp:409-622, LimbPartitionMGPU.cu:573-1746)
//   rescale   centered-lift rescale (Ciphertext.cpp:379-398, LimbPartitionMGPU.cu:214-394)
//   rotate    key switch of c1, ModDown, add c0, THEN the automorphism -- the
//             OpenFHE order of EvalAutomorphism / FIDESlib's non-fused hoisted
//             path (Ciphertext.cpp:1080-1128); FIDESlib's fused kernel applies
//             the automorphism before ModDown and is not bit-exact.
#include <cstring>

#include "context.hpp"

namespace fides {

static void ck(hipError_t e, const char* w) { hip_check(e, w); }

// ------------------------------------------------------------ ciphertexts ----
std::unique_ptr<Ciphertext> ct_create(Context& c) {
    auto ct = std::make_unique<Ciphertext>();
    ct->ctx = &c;
    ct->data = DevBuf(&c.pool, c.ct_words());
    ct->level = -1;
    return ct;
}

void ct_load(Context& c, Ciphertext& ct, const uint64_t* sub0, const uint64_t* sub1, int num_res, double noise,
             int noise_level, int slots, bool device_src) {
    if (num_res < 1 || num_res > c.L + 1) throw Error(E_LEVEL, "num_res out of range");
    const size_t bytes = (size_t)num_res * c.N * 8;
    const hipMemcpyKind kind = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    ck(hipMemcpyAsync(ct.c0(), sub0, bytes, kind, c.stream), "load c0");
    ck(hipMemcpyAsync(ct.c1(), sub1, bytes, kind, c.stream), "load c1");
    if (!device_src) c.sync();  // caller's host buffers may go away
    ct.level = num_res - 1;
    ct.ext = 0;
    ct.noise = noise;
    ct.noise_level = noise_level;
    ct.slots = slots;
}

void ct_store(Context& c, const Ciphertext& ct, uint64_t* sub0, uint64_t* sub1) {
    const size_t bytes = (size_t)(ct.level + 1) * c.N * 8;
    ck(hipMemcpyAsync(sub0, ct.c0(), bytes, hipMemcpyDeviceToHost, c.stream), "store c0");
    ck(hipMemcpyAsync(sub1, ct.c1(), bytes, hipMemcpyDeviceToHost, c.stream), "store c1");
    c.sync();
}

void ct_reserve_ext(Context& c, Ciphertext& ct) {
    const size_t want = (size_t)(c.L + 1 + c.K) * c.N;
    if (ct.part >= want) return;
    ct.data = DevBuf(&c.pool, 2 * want);
    ct.part = want;
}

void ct_copy(Context& c, const Ciphertext& src, Ciphertext& dst) {
    if (&src == &dst) return;
    if (src.ext) ct_reserve_ext(c, dst);
    const size_t bytes = (size_t)(src.level + 1 + (src.ext ? c.K : 0)) * c.N * 8;
    ck(hipMemcpyAsync(dst.c0(), src.c0(), bytes, hipMemcpyDeviceToDevice, c.stream), "copy c0");
    ck(hipMemcpyAsync(dst.c1(), src.c1(), bytes, hipMemcpyDeviceToDevice, c.stream), "copy c1");
    dst.level = src.level;
    dst.ext = src.ext;
    dst.noise = src.noise;
    dst.noise_level = src.noise_level;
    dst.slots = src.slots;
}

void load_key(Context& c, KeySwitchingKey& k, const uint64_t* a, const uint64_t* b, bool device_src) {
    const size_t words = (size_t)c.dnum * c.np * c.N;
    k.a = DevBuf(&c.pool, words);
    k.b = DevBuf(&c.pool, words);
    const hipMemcpyKind kind = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    ck(hipMemcpyAsync(k.a.ptr, a, words * 8, kind, c.stream), "load key A");
    ck(hipMemcpyAsync(k.b.ptr, b, words * 8, kind, c.stream), "load key B");
    if (!device_src) c.sync();
}

// --------------------------------------------------------- job helpers ----
template <class JL>
static JL inline_list() {
    JL l;
    std::memset(&l, 0, sizeof(l));
    return l;
}

static NttJob ntt_job(const uint64_t* src, uint64_t* mid, uint64_t* dst, int npos, int nq, int q0) {
    NttJob j;
    std::memset(&j, 0, sizeof(j));
    j.src = src; j.mid = mid; j.dst = dst;
    j.npos = npos; j.nq = nq; j.q0 = q0;
    return j;
}

static int active(const NttJob& j) { return j.npos - (j.skip_hi - j.skip_lo); }

// Launch NTT jobs, NTT_INLINE per launch.
static void run_ntt(Context& c, std::vector<NttJob> jobs, bool inverse, int load = LOAD_PLAIN,
                    int store = STORE_PLAIN) {
    for (size_t b = 0; b < jobs.size(); b += NTT_INLINE) {
        NttJobs L = inline_list<NttJobs>();
        int total = 0;
        for (size_t i = b; i < jobs.size() && i < b + NTT_INLINE; ++i) {
            jobs[i].first = total;
            total += active(jobs[i]);
            L.inl[L.n++] = jobs[i];
        }
        hipError_t e = inverse ? launch_ntt_inv(c.T, L, total, c.stream)
                               : launch_ntt_fwd(c.T, L, total, load, store, c.stream);
        ck(e, inverse ? "intt launch" : "ntt launch");
    }
}

void run_ntt_jobs(Context& c, std::vector<NttJob> jobs, bool inverse, int load, int store) {
    run_ntt(c, std::move(jobs), inverse, load, store);
}

// ------------------------------------------------------ key-switch core ----
struct KsScratch {
    DevBuf t, ext, acc_a, acc_b, tk, y, d2, top, tmp;
    // mult: adds d2 (tensor output c1d1) and top (4 top limbs of the merged
    // ModDown+rescale); rot: adds tmp (the result when the output aliases the input, 2 polys)
    KsScratch(Context& c, int level, bool mult = false, bool rot = false) {
        const size_t N = c.N, nl = level + 1, npos = nl + c.K;
        t = DevBuf(&c.pool, nl * N);
        ext = DevBuf(&c.pool, (size_t)c.num_digits(level) * npos * N);
        acc_a = DevBuf(&c.pool, npos * N);
        acc_b = DevBuf(&c.pool, npos * N);
        tk = DevBuf(&c.pool, (size_t)2 * c.K * N);
        y = DevBuf(&c.pool, 2 * nl * N);
        if (mult) {
            d2 = DevBuf(&c.pool, nl * N);
            top = DevBuf(&c.pool, 4 * N);
        }
        if (rot) tmp = DevBuf(&c.pool, 2 * nl * N);
    }
};

// One key-switch workspace per context, sized for the top level: ops of a
// context run in order on its stream, so they never use it concurrently, and
// every captured op graph can bake in the same buffers.
static KsScratch& workspace(Context& c) {
    if (!c.ks_workspace) c.ks_workspace = std::make_shared<KsScratch>(c, c.L, true, true);
    return *static_cast<KsScratch*>(c.ks_workspace.get());
}

// Per-op graph: the launch sequence of one op on fixed operands, captured
// once and replayed with hipGraphLaunch.
struct OpGraph {
    hipGraphExec_t exec = nullptr;
    ~OpGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// Run `body(workspace)` through the context's graph cache: the first call
// with a signature runs eagerly, the second captures the launches, later
// calls replay them.  A capture that throws is ended and its error re-raised.
template <class Body>
static void run_cached(Context& c, const std::vector<uint64_t>& sig_in, Body body) {
    KsScratch& ws = workspace(c);
    if (!c.use_graphs || !c.op_graphs) {
        body(ws);
        return;
    }
    std::vector<uint64_t> sig(sig_in);
    sig.push_back(tune_epoch());  // the launch configurations the capture bakes in
    auto it = c.graph_cache.find(sig);
    if (it == c.graph_cache.end()) {
        // capture only signatures seen before: one-off operand sets (a stream
        // of distinct ciphertexts) run eagerly instead of paying for a capture
        if (c.graph_seen.size() >= 8192) c.graph_seen.clear();
        if (c.graph_seen.insert(sig).second) {
            body(ws);
            return;
        }
        if (c.graph_cache.size() >= 1024) {  // bounded host memory
            std::vector<std::shared_ptr<void>> old;
            for (auto& kv : c.graph_cache) old.push_back(kv.second);
            c.graph_cache.clear();
            c.retire(std::move(old));
        }
        auto g = std::make_shared<OpGraph>();
        g->exec = capture_graph(c, [&] { body(ws); });
        it = c.graph_cache.emplace(sig, g).first;
    }
    auto* g = (OpGraph*)it->second.get();
    hip_check(hipGraphLaunch(g->exec, c.stream), "hipGraphLaunch");
}

// ModUp of x (evaluation form, level l) into ext (evaluation form; the digit's
// own positions are left unwritten: consumers read them from x).
static void modup(Context& c, int level, const uint64_t* x, uint64_t* t, uint64_t* ext) {
    const int nl = level + 1, npos = nl + c.K, nd = c.num_digits(level);
    NttJob ji = ntt_job(x, t, t, nl, nl, 0);
    ji.level = level;
    ji.scale_sel = ISCALE_MODUP;
    run_ntt(c, {ji}, true);
    ModUpJobs mj = inline_list<ModUpJobs>();
    mj.inl[0] = ModUpJob{t, ext, level};
    mj.n = 1;
    ck(launch_modup_conv(c.T, mj, nd, c.stream), "modup conv");
    std::vector<NttJob> jobs;
    for (int d = 0; d < nd; ++d) {
        uint64_t* e = ext + (size_t)d * npos * c.N;
        NttJob j = ntt_job(e, e, e, npos, nl, 0);
        j.skip_lo = d * c.alpha;
        j.skip_hi = d * c.alpha + c.digit_size(d, level);
        jobs.push_back(j);
    }
    run_ntt(c, jobs, false);
}

// KSK inner product with `key`, then ModDown of both accumulators:
//   out_b = (acc_b - conv) P^-1 (+ add_b),  out_a likewise.
// With autok != 0 (k^-1 of an automorphism X -> X^k) both outputs are also
// permuted by the automorphism on the way out; then neither output may alias
// an addend.
static void ksk_moddown(Context& c, int level, const uint64_t* ext, const uint64_t* own, const KeySwitchingKey& key,
                        KsScratch& s, uint64_t* out_b, const uint64_t* add_b, uint64_t* out_a,
                        const uint64_t* add_a, uint32_t autok = 0) {
    const int nl = level + 1, npos = nl + c.K;
    const size_t N = c.N;
    KskJobs kj = inline_list<KskJobs>();
    kj.inl[0] = KskJob{ext, own, key.a.ptr, key.b.ptr, s.acc_a.ptr, s.acc_b.ptr, level};
    kj.n = 1;
    ck(launch_ksk_inner(c.T, kj, npos, c.stream), "ksk inner");
    // INTT of the special limbs, scaled by N^-1 (P/p_k)^-1.
    uint64_t* tka = s.tk.ptr;
    uint64_t* tkb = s.tk.ptr + (size_t)c.K * N;
    NttJob ja = ntt_job(s.acc_a.ptr + nl * N, tka, tka, c.K, 0, 0);
    NttJob jb = ntt_job(s.acc_b.ptr + nl * N, tkb, tkb, c.K, 0, 0);
    ja.scale_sel = jb.scale_sel = ISCALE_MODDOWN;
    run_ntt(c, {ja, jb}, true);
    uint64_t* ya = s.y.ptr;
    uint64_t* yb = s.y.ptr + nl * N;
    ModDownJobs mj = inline_list<ModDownJobs>();
    mj.inl[0] = ModDownJob{tka, ya, nullptr, nullptr, level};
    mj.inl[1] = ModDownJob{tkb, yb, nullptr, nullptr, level};
    mj.n = 2;
    ck(launch_moddown_conv(c.T, mj, nl, c.stream), "moddown conv");
    NttJob fa = ntt_job(ya, ya, out_a, nl, nl, 0);
    fa.aux = s.acc_a.ptr;
    fa.add = add_a;
    NttJob fb = ntt_job(yb, yb, out_b, nl, nl, 0);
    fb.aux = s.acc_b.ptr;
    fb.add = add_b;
    fa.autok = fb.autok = autok;
    run_ntt(c, {fa, fb}, false, LOAD_PLAIN, autok ? STORE_MODDOWN_AUTO : STORE_MODDOWN);
}

// KSK inner product, ModDown and rescale merged (HMult+relin+rescale): on
// entry c0/c1 hold the tensor terms (c0d0, c0d1+c1d0) at level l, on exit the
// rescaled result at level l-1.  The INTT launch of the special limbs also
// takes the four top limbs (accumulators and addends at q_l); the ModDown
// conversion then produces z = conv P^-1 + Switch(top) and a single NTT with
// the STORE_MODDOWN_RESCALE epilogue replaces ModDown's NTT, the rescale's
// INTT and the rescale's NTT (k_moddown_conv states the identity).
static void ksk_moddown_rescale(Context& c, int level, const uint64_t* ext, const uint64_t* own,
                                const KeySwitchingKey& key, KsScratch& s, uint64_t* c0, uint64_t* c1) {
    const int l = level, nl = l + 1, npos = nl + c.K;
    const size_t N = c.N;
    KskJobs kj = inline_list<KskJobs>();
    kj.inl[0] = KskJob{ext, own, key.a.ptr, key.b.ptr, s.acc_a.ptr, s.acc_b.ptr, level};
    kj.n = 1;
    ck(launch_ksk_inner(c.T, kj, npos, c.stream), "ksk inner");
    uint64_t* tka = s.tk.ptr;
    uint64_t* tkb = s.tk.ptr + (size_t)c.K * N;
    DevBuf& top = s.top;  // INTT of acc_a[l], acc_b[l], c1[l], c0[l]
    NttJob ja = ntt_job(s.acc_a.ptr + nl * N, tka, tka, c.K, 0, 0);
    NttJob jb = ntt_job(s.acc_b.ptr + nl * N, tkb, tkb, c.K, 0, 0);
    ja.scale_sel = jb.scale_sel = ISCALE_MODDOWN;
    NttJob la = ntt_job(s.acc_a.ptr + l * N, top.ptr, top.ptr, 1, 1, l);
    NttJob lb = ntt_job(s.acc_b.ptr + l * N, top.ptr + N, top.ptr + N, 1, 1, l);
    NttJob ca = ntt_job(c1 + l * N, top.ptr + 2 * N, top.ptr + 2 * N, 1, 1, l);
    NttJob cb = ntt_job(c0 + l * N, top.ptr + 3 * N, top.ptr + 3 * N, 1, 1, l);
    run_ntt(c, {ja, jb, la, lb, ca, cb}, true);
    uint64_t* za = s.y.ptr;
    uint64_t* zb = s.y.ptr + (size_t)l * N;
    ModDownJobs mj = inline_list<ModDownJobs>();
    mj.inl[0] = ModDownJob{tka, za, top.ptr, top.ptr + 2 * N, level};
    mj.inl[1] = ModDownJob{tkb, zb, top.ptr + N, top.ptr + 3 * N, level};
    mj.n = 2;
    ck(launch_moddown_conv_mode(c.T, mj, true, c.stream), "moddown+rescale conv");
    NttJob fa = ntt_job(za, za, c1, l, l, 0);
    fa.aux = s.acc_a.ptr;
    fa.add = c1;
    fa.level = l;
    NttJob fb = ntt_job(zb, zb, c0, l, l, 0);
    fb.aux = s.acc_b.ptr;
    fb.add = c0;
    fb.level = l;
    run_ntt(c, {fa, fb}, false, LOAD_PLAIN, STORE_MODDOWN_RESCALE);
}

// ------------------------------------------------------------ metadata ----
static void require_same_key_level(const Ciphertext& a, const Ciphertext& b) {
    if (a.level < 0 || b.level < 0) throw Error(E_STATE, "ciphertext not loaded");
}

// ----------------------------------------------------------------- ops ----
static void mult_core(Context& c, const Ciphertext* a, const Ciphertext* b, Ciphertext& out, bool do_rescale);

// Ciphertext::mult under FIXEDAUTO / FLEXIBLEAUTO (src/CKKS/Ciphertext.cpp:409-421):
// adjust `self` (a private copy of the first operand) towards b; when that
// is not enough, adjust a copy of b towards self and multiply again.
static void mult_auto(Context& c, Ciphertext& self, const Ciphertext& b, Ciphertext& out, bool do_rescale, int depth) {
    if (!adjust_for_mult(c, self, b)) {
        if (depth > 0) throw Error(E_STATE, "EvalMult: operands could not be aligned");
        std::unique_ptr<Ciphertext> tb = ct_create(c);
        ct_copy(c, b, *tb);
        if (!adjust_for_mult(c, *tb, self)) throw Error(E_STATE, "EvalMult: operands could not be aligned");
        mult_auto(c, self, *tb, out, do_rescale, depth + 1);
        return;
    }
    mult_core(c, &self, &b, out, do_rescale);
}

void eval_mult(Context& c, const Ciphertext& a_in, const Ciphertext& b_in, Ciphertext& out, bool do_rescale) {
    require_same_key_level(a_in, b_in);
    if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
    if (c.scaling_technique != FIXEDMANUAL) {
        std::unique_ptr<Ciphertext> ta = ct_create(c);
        ct_copy(c, a_in, *ta);
        mult_auto(c, *ta, &a_in == &b_in ? *ta : b_in, out, do_rescale, 0);
        return;
    }
    mult_core(c, &a_in, &b_in, out, do_rescale);
}

static void mult_core(Context& c, const Ciphertext* a, const Ciphertext* b, Ciphertext& out, bool do_rescale) {
    if (a->noise_level != 1 || b->noise_level != 1)
        throw Error(E_STATE, "EvalMult needs both inputs at noise level 1 (rescale first)");
    const int level = a->level < b->level ? a->level : b->level;  // LevelReduce the higher one
    const int nl = level + 1;
    if (do_rescale && level < 1) throw Error(E_LEVEL, "cannot rescale at level 0");
    const KeySwitchingKey& key = *c.eval_key;
    const uint64_t *a0 = a->c0(), *a1 = a->c1(), *b0 = b->c0(), *b1 = b->c1();
    uint64_t *o0 = out.c0(), *o1 = out.c1();
    auto body = [&](KsScratch& s) {
        TensorJobs tj = inline_list<TensorJobs>();
        tj.inl[0] = TensorJob{a0, a1, b0, b1, o0, o1, s.d2.ptr, nl};
        tj.n = 1;
        ck(launch_tensor(c.T, tj, nl, c.stream), "tensor");
        modup(c, level, s.d2.ptr, s.t.ptr, s.ext.ptr);
        if (do_rescale)
            ksk_moddown_rescale(c, level, s.ext.ptr, s.d2.ptr, key, s, o0, o1);
        else
            ksk_moddown(c, level, s.ext.ptr, s.d2.ptr, key, s, o0, o0, o1, o1);
    };
    // c1 addresses too: a buffer's c1 offset depends on its capacity
    const std::vector<uint64_t> sig{1, (uint64_t)a0, (uint64_t)b0, (uint64_t)o0, (uint64_t)level,
                                    (uint64_t)do_rescale, (uint64_t)key.a.ptr, (uint64_t)key.b.ptr,
                                    (uint64_t)a1, (uint64_t)b1, (uint64_t)o1};
    run_cached(c, sig, body);
    const double noise = a->noise * b->noise;
    const int slots = a->slots > b->slots ? a->slots : b->slots;
    if (do_rescale) {
        out.level = level - 1;
        out.noise = noise / c.mod_reduce_factor[level];
        out.noise_level = 1;
    } else {
        out.level = level;
        out.noise = noise;
        out.noise_level = 2;
    }
    out.slots = slots;
}

// Relinearization of a tensor's d2 into (o0, o1) in place: ModUp, KSK with
// the relinearization key, ModDown added to o0 / o1 -- the mult pipeline
// after its tensor product (Ciphertext::dotProduct's summed tensor).
void relin_add(Context& c, int level, const uint64_t* d2, uint64_t* o0, uint64_t* o1) {
    if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
    KsScratch& s = workspace(c);
    modup(c, level, d2, s.t.ptr, s.ext.ptr);
    ksk_moddown(c, level, s.ext.ptr, d2, *c.eval_key, s, o0, o0, o1, o1);
}

void eval_square(Context& c, const Ciphertext& a, Ciphertext& out, bool do_rescale) {
    eval_mult(c, a, a, out, do_rescale);
}

// Rescale with caller-provided scratch: top = the coefficient-form top limbs
// of both parts (2 limbs), mid = the forward NTT's intermediate (2 l limbs).
static void rescale_with(Context& c, Ciphertext& a, uint64_t* top, uint64_t* mid) {
    const int l = a.level;
    const size_t N = c.N;
    NttJob i0 = ntt_job(a.c0() + l * N, top, top, 1, 1, l);
    NttJob i1 = ntt_job(a.c1() + l * N, top + N, top + N, 1, 1, l);
    run_ntt(c, {i0, i1}, true);
    NttJob f0 = ntt_job(top, mid, a.c0(), l, l, 0);
    NttJob f1 = ntt_job(top + N, mid + l * N, a.c1(), l, l, 0);
    f0.level = f1.level = l;
    run_ntt(c, {f0, f1}, false, LOAD_SWITCH, STORE_RESCALE);
    a.level = l - 1;
    a.noise /= c.mod_reduce_factor[l];
    a.noise_level -= 1;
}

void rescale(Context& c, Ciphertext& a) {
    if (a.level < 1) throw Error(E_LEVEL, "cannot rescale at level 0");
    DevBuf top(&c.pool, 2 * (size_t)c.N), mid(&c.pool, 2 * (size_t)a.level * c.N);
    rescale_with(c, a, top.ptr, mid.ptr);
}

static void ew(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out, int op) {
    require_same_key_level(a, b);
    const int level = a.level < b.level ? a.level : b.level;
    EwJobs L = inline_list<EwJobs>();
    L.inl[0] = EwJob{a.c0(), b.c0(), out.c0(), level + 1};
    L.inl[1] = EwJob{a.c1(), b.c1(), out.c1(), level + 1};
    L.n = 2;
    ck(launch_elementwise(c.T, L, level + 1, op, c.stream), "elementwise");
    out.level = level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots > b.slots ? a.slots : b.slots;
}
// Ciphertext::add / sub under FIXEDAUTO / FLEXIBLEAUTO
// (src/CKKS/Ciphertext.cpp:166-234): align `self` (a private copy of the first
// operand) with b, or a copy of b with self, then add at the common level.
static void ew_auto(Context& c, Ciphertext& self, const Ciphertext& b, Ciphertext& out, int op, int depth) {
    if (!adjust_for_add_or_sub(c, self, b)) {
        if (depth > 0) throw Error(E_STATE, "EvalAdd: operands could not be aligned");
        std::unique_ptr<Ciphertext> tb = ct_create(c);
        ct_copy(c, b, *tb);
        if (!adjust_for_add_or_sub(c, *tb, self)) throw Error(E_STATE, "EvalAdd: operands could not be aligned");
        ew_auto(c, self, *tb, out, op, depth + 1);
        return;
    }
    if ((c.scaling_technique == FLEXIBLEAUTO || c.scaling_technique == FLEXIBLEAUTOEXT) && self.level != b.level)
        throw Error(E_STATE, "EvalAdd: FLEXIBLEAUTO operands left at different levels");
    if (self.noise_level != b.noise_level) throw Error(E_STATE, "EvalAdd: operands at different noise degrees");
    ew(c, self, b, out, op);
}

static void ew_dispatch(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out, int op) {
    if (c.scaling_technique == FIXEDMANUAL) {
        ew(c, a, b, out, op);
        return;
    }
    require_same_key_level(a, b);
    std::unique_ptr<Ciphertext> ta = ct_create(c);
    ct_copy(c, a, *ta);
    ew_auto(c, *ta, &a == &b ? *ta : b, out, op, 0);
}
void eval_add(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out) { ew_dispatch(c, a, b, out, EW_ADD); }
void eval_sub(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out) { ew_dispatch(c, a, b, out, EW_SUB); }

static const KeySwitchingKey& rot_key(Context& c, uint32_t k) {
    auto it = c.rot_keys.find(k);
    if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for automorphism index " + std::to_string(k));
    return *it->second;
}

static void finish_rotation(Context& c, const Ciphertext& a, int level, const uint64_t* ext, const uint64_t* own,
                            uint32_t k, KsScratch& s, Ciphertext& out) {
    const size_t N = c.N;
    const int nl = level + 1;
    // ModDown, + c0 and the automorphism in one epilogue; the automorphism
    // permutes, so an output that aliases the input goes through tmp
    const bool alias = out.c0() == a.c0() || out.c0() == a.c1() || out.c1() == a.c0() || out.c1() == a.c1();
    uint64_t* o0 = alias ? s.tmp.ptr : out.c0();
    uint64_t* o1 = alias ? s.tmp.ptr + (size_t)nl * N : out.c1();
    ksk_moddown(c, level, ext, own, rot_key(c, k), s, o0, a.c0(), o1, nullptr, c.auto_inverse(k));
    if (alias) {
        ck(hipMemcpyAsync(out.c0(), o0, (size_t)nl * N * 8, hipMemcpyDeviceToDevice, c.stream), "rotate copy c0");
        ck(hipMemcpyAsync(out.c1(), o1, (size_t)nl * N * 8, hipMemcpyDeviceToDevice, c.stream), "rotate copy c1");
    }
}

static void set_rot_meta(const Ciphertext& a, Ciphertext& out) {
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

void eval_rotate(Context& c, const Ciphertext& a, int index, Ciphertext& out) {
    eval_automorph(c, a, c.rot_index(index), out);
}

// X -> X^k with the key of automorphism index k (rotation: k = 5^r; conjugation:
// k = 2N - 1, FIDESlib Ciphertext::conjugate, Ciphertext.cpp:933-990).
void eval_automorph(Context& c, const Ciphertext& a, uint32_t k, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (!(k & 1) || k >= 2u * c.N) throw Error(E_ARG, "automorphism index must be odd and below 2N");
    if (k == 1) {
        ct_copy(c, a, out);
        return;
    }
    const int level = a.level;
    const KeySwitchingKey& key = rot_key(c, k);
    auto body = [&](KsScratch& s) {
        modup(c, level, a.c1(), s.t.ptr, s.ext.ptr);
        finish_rotation(c, a, level, s.ext.ptr, a.c1(), k, s, out);
    };
    const std::vector<uint64_t> sig{2, (uint64_t)a.c0(), (uint64_t)out.c0(), (uint64_t)level, k,
                                    (uint64_t)key.a.ptr, (uint64_t)key.b.ptr, (uint64_t)a.c1(), (uint64_t)out.c1()};
    run_cached(c, sig, body);
    set_rot_meta(a, out);
}

// Ciphertext::keySwitch (Ciphertext.cpp:1450-1461): out = (c0 + KS_b(c1),
// KS_a(c1)) with the secret-switching key loaded under `id`; c1 multiplies
// the key's source secret, the output decrypts under its target secret.
void eval_key_switch(Context& c, const Ciphertext& a, int id, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    const auto it = c.switch_keys.find(id);
    if (it == c.switch_keys.end()) throw Error(E_NOKEY, "no switching key " + std::to_string(id));
    const KeySwitchingKey& key = *it->second;
    const int level = a.level;
    auto body = [&](KsScratch& s) {
        modup(c, level, a.c1(), s.t.ptr, s.ext.ptr);
        ksk_moddown(c, level, s.ext.ptr, a.c1(), key, s, out.c0(), a.c0(), out.c1(), nullptr, 0);
    };
    const std::vector<uint64_t> sig{3, (uint64_t)a.c0(), (uint64_t)out.c0(), (uint64_t)level, (uint64_t)id,
                                    (uint64_t)key.a.ptr, (uint64_t)key.b.ptr, (uint64_t)a.c1(), (uint64_t)out.c1()};
    run_cached(c, sig, body);
    set_rot_meta(a, out);
}

void modup_to(Context& c, int level, const uint64_t* x, uint64_t* t, uint64_t* ext) { modup(c, level, x, t, ext); }

// ModRaise of bootstrapping (FIDESlib Bootstrap.cu ModRaise: INTT, then
// broadcastLimb0's SwitchModulus -- the centered lift of the q0 coefficients
// into every prime -- then NTT; OpenFHE's raise in EvalBootstrap): a level-0
// ciphertext becomes a level-`to` one whose coefficients are the centered
// representatives of a's mod q0.  Metadata other than the level is unchanged.
void mod_raise(Context& c, const Ciphertext& a, int to, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a.level != 0) throw Error(E_LEVEL, "ModRaise takes a level-0 ciphertext (rescale / drop first)");
    if (to < 1 || to > c.L) throw Error(E_LEVEL, "raise target out of range");
    const size_t N = c.N;
    const int n = to + 1;
    DevBuf top(&c.pool, 2 * N), mid(&c.pool, 2 * (size_t)n * N);
    NttJob i0 = ntt_job(a.c0(), top.ptr, top.ptr, 1, 1, 0);
    NttJob i1 = ntt_job(a.c1(), top.ptr + N, top.ptr + N, 1, 1, 0);
    run_ntt(c, {i0, i1}, true);
    NttJob f0 = ntt_job(top.ptr, mid.ptr, out.c0(), n, n, 0);
    NttJob f1 = ntt_job(top.ptr + N, mid.ptr + (size_t)n * N, out.c1(), n, n, 0);
    f0.level = f1.level = 0;  // lift from q_0
    run_ntt(c, {f0, f1}, false, LOAD_SWITCH, STORE_PLAIN);
    const double noise = a.noise;
    const int nlv = a.noise_level, slots = a.slots;
    out.level = to;
    out.noise = noise;
    out.noise_level = nlv;
    out.slots = slots;
}

void eval_rotate_hoisted(Context& c, const Ciphertext& a, const std::vector<int>& idx,
                         const std::vector<Ciphertext*>& outs) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (idx.size() != outs.size()) throw Error(E_ARG, "index / output count mismatch");
    for (size_t i = 0; i < outs.size(); ++i) {
        if (outs[i] == &a) throw Error(E_ARG, "hoisted rotation outputs must not alias the input");
        for (size_t j = 0; j < i; ++j)
            if (outs[j] == outs[i]) throw Error(E_ARG, "hoisted rotation outputs must be distinct");
    }
    std::vector<uint32_t> ks;
    for (int r : idx) {
        ks.push_back(c.rot_index(r));
        if (ks.back() != 1) rot_key(c, ks.back());  // E_NOKEY before any device work
    }
    rotate_hoisted_batch(c, a, ks, outs);
    for (auto* o : outs) set_rot_meta(a, *o);
}

// ------------------------------------------------------ stage entries ----
void poly_ntt(Context& c, uint64_t* data, int nlimbs, int first_pid, bool inverse) {
    if (nlimbs < 1 || first_pid < 0 || first_pid + nlimbs > c.np) throw Error(E_ARG, "limb range out of bounds");
    NttJob j = ntt_job(data, data, data, nlimbs, nlimbs, first_pid);
    run_ntt(c, {j}, inverse);
}

// NTT of `limbs` limbs whose prime ids cycle through all L+1+K primes (limb i
// has prime i mod (L+1+K)), one launch per NTT_INLINE polynomials: the launch
// shapes of batched stages, for the tuner and benchmarks.
void ntt_limbs(Context& c, uint64_t* data, int limbs, bool inverse) {
    if (limbs < 1 || limbs > NTT_INLINE * c.np) throw Error(E_ARG, "limb count out of range (1 .. 8 (L+1+K))");
    std::vector<NttJob> jobs;
    for (int first = 0; first < limbs; first += c.np) {
        const int n = limbs - first < c.np ? limbs - first : c.np;
        uint64_t* p = data + (size_t)first * c.N;
        jobs.push_back(ntt_job(p, p, p, n, c.L + 1, 0));
    }
    run_ntt(c, jobs, inverse);
}

void stage_modup(Context& c, int level, const uint64_t* x, uint64_t* ext) {
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    DevBuf t(&c.pool, (size_t)(level + 1) * c.N);
    stage_modup_into(c, level, x, t.ptr, ext);
}

void stage_modup_into(Context& c, int level, const uint64_t* x, uint64_t* t, uint64_t* ext) {
    modup(c, level, x, t, ext);
    // own-digit positions: copy x so the output matches the full extended digits
    const int nl = level + 1, npos = nl + c.K;
    for (int d = 0; d < c.num_digits(level); ++d) {
        const int st = d * c.alpha, sz = c.digit_size(d, level);
        ck(hipMemcpyAsync(ext + ((size_t)d * npos + st) * c.N, x + (size_t)st * c.N, (size_t)sz * c.N * 8,
                          hipMemcpyDeviceToDevice, c.stream),
           "modup own copy");
    }
}

void stage_keyswitch(Context& c, int level, const uint64_t* x, const KeySwitchingKey& k, uint64_t* ks_b,
                     uint64_t* ks_a) {
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    KsScratch s(c, level);
    modup(c, level, x, s.t.ptr, s.ext.ptr);
    ksk_moddown(c, level, s.ext.ptr, x, k, s, ks_b, nullptr, ks_a, nullptr);
}

void stage_moddown(Context& c, int level, const uint64_t* x_ext, uint64_t* y) {
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    const int nl = level + 1;
    const size_t N = c.N;
    DevBuf tk(&c.pool, (size_t)c.K * N), yy(&c.pool, (size_t)nl * N);
    NttJob j = ntt_job(x_ext + nl * N, tk.ptr, tk.ptr, c.K, 0, 0);
    j.scale_sel = ISCALE_MODDOWN;
    run_ntt(c, {j}, true);
    ModDownJobs mj = inline_list<ModDownJobs>();
    mj.inl[0] = ModDownJob{tk.ptr, yy.ptr, nullptr, nullptr, level};
    mj.n = 1;
    ck(launch_moddown_conv(c.T, mj, nl, c.stream), "moddown conv");
    NttJob f = ntt_job(yy.ptr, yy.ptr, y, nl, nl, 0);
    f.aux = x_ext;
    run_ntt(c, {f}, false, LOAD_PLAIN, STORE_MODDOWN);
}

}  // namespace fides
