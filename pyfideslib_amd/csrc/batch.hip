// Batched key-switched ops -- HMult + relinearize (+ rescale) pairs and
// rotations -- over many independent ciphertexts at uneven levels.
//
// Every stage of the per-op pipeline (tensor, INTT, ModUp conversion, NTT of
// the extended digits, KSK inner product, INTT of the special limbs, ModDown
// conversion, NTT+ModDown(+rescale) epilogue) becomes ONE launch covering all
// pairs: the grid walks a flattened list of (job, limb) tiles read from a
// device-resident job table, so ragged levels cost nothing and small levels
// still fill the 256 CUs.  Job tables and scratch are built once per batch
// signature and reused (the plan cache), so repeated calls upload nothing;
// the plan's launches are captured into a hipGraph and replayed.
//
// FIDESlib's batched kernels (LimbPartitionBatch.cu:21-604) batch elementwise
// and BSGS work but not the key switch; this is the build's own form.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "context.hpp"

namespace fides {

struct NttStage {
    const NttJob* jobs = nullptr;
    const int* map = nullptr;
    int njobs = 0, limbs = 0;
};

struct BatchPlan {
    std::vector<uint64_t> sig;
    hipGraphExec_t exec = nullptr;  // the plan's launches, captured on the plan's second use
    ~BatchPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
    DevBuf table;  // job tables; the scratch is the context's batch arena
    // device views
    const TensorJob* tensor = nullptr;
    int n_mult = 0, max_nl = 0, max_nd = 0, max_npos = 0;
    int n = 0;
    NttStage intt_up, ntt_ext, intt_sp;
    NttStage md_rescale, md_plain, md_auto;  // ModDown NTTs by epilogue
    const ModUpJob* modup = nullptr;
    const KskJob* ksk = nullptr;
    const ModDownJob* conv_rescale = nullptr;
    const ModDownJob* conv_plain = nullptr;
    int n_conv_rescale = 0, n_conv_plain = 0;
    // key-stationary KSK: the KSK jobs grouped by key, chunks of one key
    const int* ksk_chunks = nullptr;
    int n_ksk_chunks = 0;
    bool keystat = false;  // at least two jobs per key on average
};

// Key-stationary KSK for single-key batches (FIDES_KSK_KEYSTAT=0 turns it off).
bool ksk_keystat_on() {
    static const bool on = [] {
        const char* e = std::getenv("FIDES_KSK_KEYSTAT");
        return !e || std::atoi(e) != 0;
    }();
    return on;
}

static void ckb(hipError_t e, const char* w) { hip_check(e, w); }

// When a new batch plan is built and the pool holds more than
// POOL_CACHE_HIGH bytes of cached blocks, it gives blocks back down to
// POOL_CACHE_KEEP: the gap keeps a workload that frees and reallocates its
// temporaries (a bootstrap after a large batch) from trimming -- a device
// synchronization and hipFree / hipMalloc round trips -- at every plan.
constexpr size_t POOL_CACHE_HIGH = (size_t)64 << 30;
constexpr size_t POOL_CACHE_KEEP = (size_t)32 << 30;

static std::vector<int> limbmap_of(std::vector<NttJob>& jobs, int& total) {
    std::vector<int> map;
    total = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        jobs[j].first = total;
        const int a = jobs[j].npos - (jobs[j].skip_hi - jobs[j].skip_lo);
        for (int i = 0; i < a; ++i) map.push_back((int)j);
        total += a;
    }
    return map;
}

static int item_level(const KsItem& it) {
    return it.mult ? std::min(it.a->level, it.b->level) : it.a->level;
}

static std::shared_ptr<BatchPlan> build_plan(Context& c, const std::vector<KsItem>& items, std::vector<uint64_t> sig) {
    auto P = std::make_shared<BatchPlan>();
    P->sig = std::move(sig);
    P->n = (int)items.size();
    const size_t N = c.N;
    const int K = c.K;
    // scratch layout per item: [d2 (mult)], t, ext, acc_a, acc_b, tk(2K), y(2nl), [top(4) (rescale)]
    std::vector<size_t> off(P->n + 1, 0);
    for (int i = 0; i < P->n; ++i) {
        const KsItem& it = items[i];
        const int l = item_level(it);
        const size_t nl = l + 1, npos = nl + K, nd = c.num_digits(l);
        const size_t words = (it.mult ? nl : 0) + nl + nd * npos + 2 * npos + 2 * K + 2 * nl + (it.rescale ? 4 : 0);
        off[i + 1] = off[i] + words * N;
        P->max_nl = std::max(P->max_nl, (int)nl);
        P->max_nd = std::max(P->max_nd, (int)nd);
        P->max_npos = std::max(P->max_npos, (int)npos);
    }
    uint64_t* base = c.batch_arena(off[P->n]);

    std::vector<TensorJob> tj;
    std::vector<NttJob> iu, ne, isp, nmd_r, nmd_p, nmd_a;
    std::vector<ModUpJob> mu;
    std::vector<KskJob> kj;
    std::vector<ModDownJob> md_r, md_p;
    for (int i = 0; i < P->n; ++i) {
        const KsItem& it = items[i];
        const int l = item_level(it), nl = l + 1, npos = nl + K, nd = c.num_digits(l);
        uint64_t* s = base + off[i];
        uint64_t* d2 = it.mult ? s : nullptr;
        uint64_t* t = s + (it.mult ? nl * N : 0);
        uint64_t* ext = t + nl * N;
        uint64_t* acc_a = ext + (size_t)nd * npos * N;
        uint64_t* acc_b = acc_a + npos * N;
        uint64_t* tk = acc_b + npos * N;
        uint64_t* y = tk + 2 * K * N;
        uint64_t* top = y + 2 * nl * N;
        Ciphertext& O = *it.out;
        // the key-switch input: c1 d1 (mult, from the tensor) or c1 (rotation)
        const uint64_t* x = it.mult ? d2 : it.a->c1();
        if (it.mult)
            tj.push_back(TensorJob{it.a->c0(), it.a->c1(), it.b->c0(), it.b->c1(), O.c0(), O.c1(), d2, nl});
        NttJob j{};
        j.src = x; j.mid = t; j.dst = t; j.npos = nl; j.nq = nl; j.q0 = 0;
        j.level = l; j.scale_sel = ISCALE_MODUP;
        iu.push_back(j);
        mu.push_back(ModUpJob{t, ext, l});
        for (int d = 0; d < nd; ++d) {
            NttJob e{};
            uint64_t* ed = ext + (size_t)d * npos * N;
            e.src = ed; e.mid = ed; e.dst = ed; e.npos = npos; e.nq = nl; e.q0 = 0;
            e.skip_lo = d * c.alpha;
            e.skip_hi = d * c.alpha + c.digit_size(d, l);
            ne.push_back(e);
        }
        kj.push_back(KskJob{ext, x, it.key->a.ptr, it.key->b.ptr, acc_a, acc_b, l});
        // h = 0: a-part -> c1, h = 1: b-part -> c0 (ckks.hip ksk_moddown / ksk_moddown_rescale)
        for (int h = 0; h < 2; ++h) {
            uint64_t* acc = h ? acc_b : acc_a;
            uint64_t* tkh = tk + (size_t)h * K * N;
            uint64_t* yh = y + (size_t)h * nl * N;
            uint64_t* out = h ? O.c0() : O.c1();
            NttJob a{};
            a.src = acc + nl * N; a.mid = tkh; a.dst = tkh; a.npos = K; a.nq = 0; a.q0 = 0;
            a.scale_sel = ISCALE_MODDOWN;
            isp.push_back(a);
            NttJob f{};
            f.src = yh; f.mid = yh; f.dst = out; f.aux = acc; f.q0 = 0;
            if (it.mult && it.rescale) {
                f.add = out;  // the tensor terms (c0 d0, c0 d1 + c1 d0), in place
                for (int w = 0; w < 2; ++w) {  // top limbs of the accumulator and of the addend
                    NttJob t1{};
                    t1.src = (w ? out : acc) + (size_t)l * N;
                    t1.mid = t1.dst = top + (size_t)(h + 2 * w) * N;
                    t1.npos = 1; t1.nq = 1; t1.q0 = l;
                    isp.push_back(t1);
                }
                md_r.push_back(ModDownJob{tkh, yh, top + (size_t)h * N, top + (size_t)(h + 2) * N, l});
                f.npos = l; f.nq = l; f.level = l;
                nmd_r.push_back(f);
            } else if (it.mult) {
                f.add = out;
                md_p.push_back(ModDownJob{tkh, yh, nullptr, nullptr, l});
                f.npos = nl; f.nq = nl;
                nmd_p.push_back(f);
            } else {
                // rotation: ModDown, + c0 (b part), then the automorphism on the way out
                f.add = h ? it.a->c0() : nullptr;
                f.autok = it.kinv;
                md_p.push_back(ModDownJob{tkh, yh, nullptr, nullptr, l});
                f.npos = nl; f.nq = nl;
                nmd_a.push_back(f);
            }
        }
    }
    P->n_mult = (int)tj.size();
    // KSK jobs grouped by key (each job carries its own pointers, so their
    // order is free), cut into chunks of one key for the key-stationary form
    std::stable_sort(kj.begin(), kj.end(), [](const KskJob& x, const KskJob& y) { return x.ka < y.ka; });
    std::vector<int> chunks{0};
    int nkeys = 0;
    for (int j = 0; j < (int)kj.size(); ++j) {
        const bool new_key = j == 0 || kj[j].ka != kj[j - 1].ka;
        nkeys += new_key ? 1 : 0;
        if (j > 0 && (new_key || j - chunks.back() == ksk_keystat_chunk())) chunks.push_back(j);
    }
    chunks.push_back((int)kj.size());
    P->n_ksk_chunks = (int)chunks.size() - 1;
    P->keystat = !kj.empty() && 2 * nkeys <= (int)kj.size();
    int tot;
    std::vector<int> m_iu = limbmap_of(iu, tot); P->intt_up.limbs = tot;
    std::vector<int> m_ne = limbmap_of(ne, tot); P->ntt_ext.limbs = tot;
    std::vector<int> m_isp = limbmap_of(isp, tot); P->intt_sp.limbs = tot;
    std::vector<int> m_r = limbmap_of(nmd_r, tot); P->md_rescale.limbs = tot;
    std::vector<int> m_p = limbmap_of(nmd_p, tot); P->md_plain.limbs = tot;
    std::vector<int> m_a = limbmap_of(nmd_a, tot); P->md_auto.limbs = tot;

    // one table blob
    std::vector<char> host;
    auto append = [&](const void* src, size_t bytes) {
        size_t o = (host.size() + 255) & ~(size_t)255;
        host.resize(o + bytes);
        if (bytes) std::memcpy(host.data() + o, src, bytes);
        return o;
    };
    const size_t o_tj = append(tj.data(), tj.size() * sizeof(TensorJob));
    const size_t o_iu = append(iu.data(), iu.size() * sizeof(NttJob)), o_miu = append(m_iu.data(), m_iu.size() * 4);
    const size_t o_mu = append(mu.data(), mu.size() * sizeof(ModUpJob));
    const size_t o_ne = append(ne.data(), ne.size() * sizeof(NttJob)), o_mne = append(m_ne.data(), m_ne.size() * 4);
    const size_t o_kj = append(kj.data(), kj.size() * sizeof(KskJob));
    const size_t o_kc = append(chunks.data(), chunks.size() * sizeof(int));
    const size_t o_isp = append(isp.data(), isp.size() * sizeof(NttJob)),
                 o_misp = append(m_isp.data(), m_isp.size() * 4);
    const size_t o_cr = append(md_r.data(), md_r.size() * sizeof(ModDownJob));
    const size_t o_cp = append(md_p.data(), md_p.size() * sizeof(ModDownJob));
    const size_t o_r = append(nmd_r.data(), nmd_r.size() * sizeof(NttJob)), o_mr = append(m_r.data(), m_r.size() * 4);
    const size_t o_p = append(nmd_p.data(), nmd_p.size() * sizeof(NttJob)), o_mp = append(m_p.data(), m_p.size() * 4);
    const size_t o_a = append(nmd_a.data(), nmd_a.size() * sizeof(NttJob)), o_ma = append(m_a.data(), m_a.size() * 4);
    P->table = DevBuf(&c.pool, (host.size() + 7) / 8);
    c.upload(P->table.ptr, host.data(), host.size());
    char* d = (char*)P->table.ptr;
    P->tensor = (const TensorJob*)(d + o_tj);
    P->intt_up = {(const NttJob*)(d + o_iu), (const int*)(d + o_miu), (int)iu.size(), P->intt_up.limbs};
    P->modup = (const ModUpJob*)(d + o_mu);
    P->ntt_ext = {(const NttJob*)(d + o_ne), (const int*)(d + o_mne), (int)ne.size(), P->ntt_ext.limbs};
    P->ksk = (const KskJob*)(d + o_kj);
    P->ksk_chunks = (const int*)(d + o_kc);
    P->intt_sp = {(const NttJob*)(d + o_isp), (const int*)(d + o_misp), (int)isp.size(), P->intt_sp.limbs};
    P->conv_rescale = (const ModDownJob*)(d + o_cr);
    P->n_conv_rescale = (int)md_r.size();
    P->conv_plain = (const ModDownJob*)(d + o_cp);
    P->n_conv_plain = (int)md_p.size();
    P->md_rescale = {(const NttJob*)(d + o_r), (const int*)(d + o_mr), (int)nmd_r.size(), P->md_rescale.limbs};
    P->md_plain = {(const NttJob*)(d + o_p), (const int*)(d + o_mp), (int)nmd_p.size(), P->md_plain.limbs};
    P->md_auto = {(const NttJob*)(d + o_a), (const int*)(d + o_ma), (int)nmd_a.size(), P->md_auto.limbs};
    return P;
}

template <class JL, class J>
static JL dev_list(const J* jobs, int n, const int* map = nullptr) {
    JL l;
    std::memset(&l, 0, sizeof(l));
    l.dev = jobs;
    l.limbmap = map;
    l.n = n;
    return l;
}

static NttJobs ntt_list(const NttStage& st) { return dev_list<NttJobs>(st.jobs, st.njobs, st.map); }

static void run_plan(Context& c, const BatchPlan& P) {
    const DevTables& T = c.T;
    hipStream_t s = c.stream;
    if (P.n_mult) ckb(launch_tensor(T, dev_list<TensorJobs>(P.tensor, P.n_mult), P.max_nl, s), "batch tensor");
    ckb(launch_ntt_inv(T, ntt_list(P.intt_up), P.intt_up.limbs, s), "batch intt");
    ckb(launch_modup_conv(T, dev_list<ModUpJobs>(P.modup, P.n), P.max_nd, s), "batch modup");
    ckb(launch_ntt_fwd(T, ntt_list(P.ntt_ext), P.ntt_ext.limbs, LOAD_PLAIN, STORE_PLAIN, s), "batch ntt ext");
    if (P.keystat && T.dnum <= 4 && ksk_keystat_on())
        ckb(launch_ksk_keystat(T, dev_list<KskJobs>(P.ksk, P.n), P.ksk_chunks, P.n_ksk_chunks, s),
            "batch ksk (key-stationary)");
    else
        ckb(launch_ksk_inner(T, dev_list<KskJobs>(P.ksk, P.n), P.max_npos, s), "batch ksk");
    ckb(launch_ntt_inv(T, ntt_list(P.intt_sp), P.intt_sp.limbs, s), "batch intt sp");
    if (P.n_conv_rescale)
        ckb(launch_moddown_conv_mode(T, dev_list<ModDownJobs>(P.conv_rescale, P.n_conv_rescale), true, s),
            "batch moddown+rescale");
    if (P.n_conv_plain)
        ckb(launch_moddown_conv_mode(T, dev_list<ModDownJobs>(P.conv_plain, P.n_conv_plain), false, s),
            "batch moddown");
    if (P.md_rescale.njobs)
        ckb(launch_ntt_fwd(T, ntt_list(P.md_rescale), P.md_rescale.limbs, LOAD_PLAIN, STORE_MODDOWN_RESCALE, s),
            "batch ntt moddown+rescale");
    if (P.md_plain.njobs)
        ckb(launch_ntt_fwd(T, ntt_list(P.md_plain), P.md_plain.limbs, LOAD_PLAIN, STORE_MODDOWN, s),
            "batch ntt moddown");
    if (P.md_auto.njobs)
        ckb(launch_ntt_fwd(T, ntt_list(P.md_auto), P.md_auto.limbs, LOAD_PLAIN, STORE_MODDOWN_AUTO, s),
            "batch ntt moddown+automorphism");
}

// Runs a batch of key-switched ops (HMult+relin(+rescale) pairs and
// rotations, any levels) with one launch per stage and sets the outputs'
// metadata.  The plan (job tables + scratch) is kept for the next call and
// captured into a hipGraph when the same signature comes again.
void ks_batch(Context& c, const std::vector<KsItem>& items) {
    const size_t n = items.size();
    if (n == 0) return;
    std::vector<uint64_t> sig{tune_epoch()};
    for (size_t i = 0; i < n; ++i) {
        const KsItem& it = items[i];
        if (!it.a || !it.out || !it.key || (it.mult && !it.b)) throw Error(E_ARG, "incomplete batch item");
        if (it.a->level < 0 || (it.mult && it.b->level < 0)) throw Error(E_STATE, "ciphertext not loaded");
        if (it.mult && (it.a->noise_level != 1 || it.b->noise_level != 1))
            throw Error(E_STATE, "batched EvalMult needs inputs at noise level 1");
        if (it.mult && it.rescale && item_level(it) < 1) throw Error(E_LEVEL, "cannot rescale at level 0");
        if (!it.mult && it.out == it.a) throw Error(E_ARG, "batched rotation output must not alias its input");
        for (size_t j = 0; j < n; ++j)
            if (j != i && (items[j].out == it.a || (it.mult && items[j].out == it.b) || items[j].out == it.out))
                throw Error(E_ARG, "batch output aliases another item's input or output");
        sig.push_back(((uint64_t)it.mult << 1) | (uint64_t)it.rescale);
        sig.push_back((uint64_t)(uintptr_t)it.a->c0());
        sig.push_back(it.mult ? (uint64_t)(uintptr_t)it.b->c0() : it.kinv);
        sig.push_back((uint64_t)(uintptr_t)it.out->c0());
        sig.push_back((uint64_t)(uintptr_t)it.key->a.ptr);
        sig.push_back((uint64_t)(uintptr_t)it.key->b.ptr);
        sig.push_back(((uint64_t)it.a->level << 32) | (uint32_t)(it.mult ? it.b->level : 0));
        // c1 addresses too: a buffer's c1 offset depends on its capacity
        sig.push_back((uint64_t)(uintptr_t)it.a->c1());
        sig.push_back(it.mult ? (uint64_t)(uintptr_t)it.b->c1() : 0);
        sig.push_back((uint64_t)(uintptr_t)it.out->c1());
    }
    auto plan = std::static_pointer_cast<BatchPlan>(c.plan_cache);
    if (plan && plan->sig == sig) {
        if (!plan->exec && c.use_graphs) plan->exec = capture_graph(c, [&] { run_plan(c, *plan); });
        if (plan->exec)
            hip_check(hipGraphLaunch(plan->exec, c.stream), "hipGraphLaunch");
        else
            run_plan(c, *plan);
    } else {
        c.retire(std::move(c.plan_cache));
        c.plan_cache.reset();
        plan.reset();
        // a stream of differently sized batches (serving) would otherwise
        // leave one cached scratch block per size
        if (c.pool.bytes_cached() > POOL_CACHE_HIGH) c.pool.trim(POOL_CACHE_KEEP);
        plan = build_plan(c, items, sig);
        c.plan_cache = plan;
        run_plan(c, *plan);
    }
    for (const KsItem& it : items) {
        Ciphertext& o = *it.out;
        if (it.mult) {
            const int l = item_level(it);
            const double noise = it.a->noise * it.b->noise;
            const int slots = std::max(it.a->slots, it.b->slots);
            o.level = l;
            o.noise = noise;
            o.noise_level = 2;
            o.slots = slots;
            if (it.rescale) {
                o.level = l - 1;
                o.noise /= c.mod_reduce_factor[l];
                o.noise_level = 1;
            }
        } else {
            o.level = it.a->level;
            o.noise = it.a->noise;
            o.noise_level = it.a->noise_level;
            o.slots = it.a->slots;
        }
    }
}

void eval_mult_batch(Context& c, const std::vector<const Ciphertext*>& A, const std::vector<const Ciphertext*>& B,
                     const std::vector<Ciphertext*>& O, bool rescale) {
    const size_t n = A.size();
    if (B.size() != n || O.size() != n) throw Error(E_ARG, "batch size mismatch");
    if (n == 0) return;
    if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
    std::vector<KsItem> items;
    for (size_t i = 0; i < n; ++i) items.push_back(KsItem{true, A[i], B[i], O[i], c.eval_key.get(), 0, rescale});
    ks_batch(c, items);
}

void eval_rotate_batch(Context& c, const std::vector<const Ciphertext*>& A, const std::vector<int>& idx,
                       const std::vector<Ciphertext*>& O) {
    const size_t n = A.size();
    if (idx.size() != n || O.size() != n) throw Error(E_ARG, "batch size mismatch");
    std::vector<KsItem> items;
    std::vector<size_t> copies;
    for (size_t i = 0; i < n; ++i) {
        if (!A[i] || !O[i]) throw Error(E_ARG, "null ciphertext");
        if (A[i]->level < 0) throw Error(E_STATE, "ciphertext not loaded");
        const uint32_t k = c.rot_index(idx[i]);
        if (k == 1) {
            copies.push_back(i);
            continue;
        }
        auto it = c.rot_keys.find(k);
        if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for automorphism " + std::to_string(k));
        items.push_back(KsItem{false, A[i], nullptr, O[i], it->second.get(), c.auto_inverse(k), false});
    }
    for (size_t i : copies)
        for (size_t j = 0; j < n; ++j)
            if (j != i && (O[j] == A[i] || O[j] == O[i])) throw Error(E_ARG, "batch output aliases another item");
    ks_batch(c, items);
    for (size_t i : copies) ct_copy(c, *A[i], *O[i]);
}

// ------------------------------------------------- hoisted rotation set ----
// One ModUp of c1 shared by every rotation (FIDESlib Ciphertext::rotate_hoisted,
// Ciphertext.cpp:1053-1226), then each remaining stage as ONE launch over all
// rotations of the set: key-switch inner products with n keys, INTT of the 2n
// special-limb accumulators, ModDown conversions, the ModDown NTT epilogues
// (which add c0 to the b part and apply the automorphism on the way out).
// OpenFHE's order per rotation: key switch, ModDown, add c0, then the
// automorphism.
struct HoistPlan {
    std::vector<uint64_t> sig;
    hipGraphExec_t exec = nullptr;
    ~HoistPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
    DevBuf scratch, table;
    int level = 0, nk = 0, npos = 0;
    const uint64_t *c0 = nullptr, *c1 = nullptr;
    uint64_t *t = nullptr, *ext = nullptr;
    const KskJob* ksk = nullptr;
    NttStage intt_sp, ntt_md;
    const ModDownJob* moddown = nullptr;
    std::vector<std::pair<uint64_t*, uint64_t*>> copies;  // outputs of identity rotations (c0, c1)
};

static std::shared_ptr<HoistPlan> build_hoist_plan(Context& c, const Ciphertext& a, const std::vector<uint32_t>& ks,
                                                   const std::vector<Ciphertext*>& outs, std::vector<uint64_t> sig) {
    auto P = std::make_shared<HoistPlan>();
    P->sig = std::move(sig);
    const int l = a.level, nl = l + 1, K = c.K, nd = c.num_digits(l);
    const size_t N = c.N, npos = nl + K;
    P->level = l;
    P->npos = (int)npos;
    P->c0 = a.c0();
    P->c1 = a.c1();
    int nk = 0;
    for (uint32_t k : ks) nk += (k != 1);
    P->nk = nk;
    // scratch: t, ext, then per key-switched rotation acc_a, acc_b, tk(2K), y(2nl)
    const size_t per = 2 * npos + 2 * K + 2 * nl;
    P->scratch = DevBuf(&c.pool, (nl + (size_t)nd * npos + per * nk) * N);
    P->t = P->scratch.ptr;
    P->ext = P->t + nl * N;
    uint64_t* base = P->ext + (size_t)nd * npos * N;

    std::vector<KskJob> kj;
    std::vector<NttJob> isp, nmd;
    std::vector<ModDownJob> mdj;
    int r = 0;
    for (size_t i = 0; i < ks.size(); ++i) {
        if (ks[i] == 1) {
            P->copies.emplace_back(outs[i]->c0(), outs[i]->c1());
            continue;
        }
        auto it = c.rot_keys.find(ks[i]);
        if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for automorphism " + std::to_string(ks[i]));
        const KeySwitchingKey& key = *it->second;
        uint64_t* s = base + per * r * N;
        uint64_t* acc_a = s;
        uint64_t* acc_b = acc_a + npos * N;
        uint64_t* tka = acc_b + npos * N;
        uint64_t* tkb = tka + (size_t)K * N;
        uint64_t* ya = tkb + (size_t)K * N;
        uint64_t* yb = ya + nl * N;
        const uint32_t kinv = c.auto_inverse(ks[i]);
        kj.push_back(KskJob{P->ext, P->c1, key.a.ptr, key.b.ptr, acc_a, acc_b, l});
        for (int h = 0; h < 2; ++h) {
            NttJob j{};
            j.src = (h ? acc_b : acc_a) + nl * N;
            j.mid = j.dst = h ? tkb : tka;
            j.npos = K; j.nq = 0; j.q0 = 0;
            j.scale_sel = ISCALE_MODDOWN;
            isp.push_back(j);
            mdj.push_back(ModDownJob{h ? tkb : tka, h ? yb : ya, nullptr, nullptr, l});
            NttJob f{};
            f.src = f.mid = h ? yb : ya;
            // ModDown, + c0 (b part) and the automorphism, into the output
            f.dst = h ? outs[i]->c0() : outs[i]->c1();
            f.aux = h ? acc_b : acc_a;
            f.add = h ? P->c0 : nullptr;
            f.npos = nl; f.nq = nl; f.q0 = 0;
            f.autok = kinv;
            nmd.push_back(f);
        }
        ++r;
    }
    int tot;
    std::vector<int> m_isp = limbmap_of(isp, tot);
    P->intt_sp.limbs = tot;
    std::vector<int> m_nmd = limbmap_of(nmd, tot);
    P->ntt_md.limbs = tot;
    std::vector<char> host;
    auto append = [&](const void* src, size_t bytes) {
        size_t o = (host.size() + 255) & ~(size_t)255;
        host.resize(o + bytes);
        if (bytes) std::memcpy(host.data() + o, src, bytes);
        return o;
    };
    const size_t o_kj = append(kj.data(), kj.size() * sizeof(KskJob));
    const size_t o_isp = append(isp.data(), isp.size() * sizeof(NttJob)), o_misp = append(m_isp.data(), m_isp.size() * 4);
    const size_t o_md = append(mdj.data(), mdj.size() * sizeof(ModDownJob));
    const size_t o_nmd = append(nmd.data(), nmd.size() * sizeof(NttJob)), o_mnmd = append(m_nmd.data(), m_nmd.size() * 4);
    P->table = DevBuf(&c.pool, (host.size() + 7) / 8 + 1);
    c.upload(P->table.ptr, host.data(), host.size());
    char* d = (char*)P->table.ptr;
    P->ksk = (const KskJob*)(d + o_kj);
    P->intt_sp.jobs = (const NttJob*)(d + o_isp);
    P->intt_sp.map = (const int*)(d + o_misp);
    P->intt_sp.njobs = (int)isp.size();
    P->moddown = (const ModDownJob*)(d + o_md);
    P->ntt_md.jobs = (const NttJob*)(d + o_nmd);
    P->ntt_md.map = (const int*)(d + o_mnmd);
    P->ntt_md.njobs = (int)nmd.size();
    return P;
}

static void run_hoist_plan(Context& c, const HoistPlan& P) {
    const DevTables& T = c.T;
    hipStream_t s = c.stream;
    const int nl = P.level + 1;
    if (P.nk > 0) {
        modup_to(c, P.level, P.c1, P.t, P.ext);
        ckb(launch_ksk_inner(T, dev_list<KskJobs>(P.ksk, P.nk), P.npos, s, true), "hoist ksk");
        ckb(launch_ntt_inv(T, ntt_list(P.intt_sp), P.intt_sp.limbs, s), "hoist intt sp");
        ckb(launch_moddown_conv_mode(T, dev_list<ModDownJobs>(P.moddown, 2 * P.nk), false, s), "hoist moddown");
        ckb(launch_ntt_fwd(T, ntt_list(P.ntt_md), P.ntt_md.limbs, LOAD_PLAIN, STORE_MODDOWN_AUTO, s),
            "hoist ntt moddown automorph");
    }
    const size_t bytes = (size_t)nl * c.N * 8;
    for (const auto& cp : P.copies) {
        ckb(hipMemcpyAsync(cp.first, P.c0, bytes, hipMemcpyDeviceToDevice, s), "copy c0");
        ckb(hipMemcpyAsync(cp.second, P.c1, bytes, hipMemcpyDeviceToDevice, s), "copy c1");
    }
}

void rotate_hoisted_batch(Context& c, const Ciphertext& a, const std::vector<uint32_t>& ks,
                          const std::vector<Ciphertext*>& outs) {
    std::vector<uint64_t> sig{(uint64_t)(uintptr_t)a.c0(), (uint64_t)a.level, (uint64_t)(uintptr_t)a.c1(),
                              tune_epoch()};
    for (size_t i = 0; i < ks.size(); ++i) {
        sig.push_back(ks[i]);
        if (ks[i] != 1) {
            auto it = c.rot_keys.find(ks[i]);
            const bool have = it != c.rot_keys.end();
            sig.push_back(have ? (uint64_t)(uintptr_t)it->second->a.ptr : 0);
            sig.push_back(have ? (uint64_t)(uintptr_t)it->second->b.ptr : 0);
        }
        sig.push_back((uint64_t)(uintptr_t)outs[i]->c0());
        sig.push_back((uint64_t)(uintptr_t)outs[i]->c1());
    }
    auto plan = std::static_pointer_cast<HoistPlan>(c.hoist_plan);
    if (!plan || plan->sig != sig) {
        // a new set runs eagerly; a set that comes again is captured then
        // (one-off sets, e.g. each radix step of AccumulateSum, never pay for
        // a capture)
        c.retire(std::move(c.hoist_plan));
        c.hoist_plan.reset();
        plan.reset();
        plan = build_hoist_plan(c, a, ks, outs, sig);
        c.hoist_plan = plan;
    } else if (!plan->exec && c.use_graphs) {
        plan->exec = capture_graph(c, [&] { run_hoist_plan(c, *plan); });
    }
    if (plan->exec)
        hip_check(hipGraphLaunch(plan->exec, c.stream), "hipGraphLaunch");
    else
        run_hoist_plan(c, *plan);
}

}  // namespace fides
