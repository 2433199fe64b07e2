// Baby-step / giant-step linear transform (FIDESlib CKKS::LinearTransform,
// src/CKKS/LinearTransform.cu:67-258; OpenFHE EvalLinearTransform), the core
// of bootstrapping's CoeffsToSlots / SlotsToCoeffs:
//
//   out = rot_offset( sum_{j<g} rot_{j b s}( sum_{i<b} A[b j + i] * rot_{i s}(ct) ) )
//
// with b baby steps, g = ceil(n / b) giant steps and stride s; the diagonals
// A[b j + i] are plaintexts the caller encodes (pre-rotated by -j b s, as
// OpenFHE's precomputation does).  The b baby-step rotations share one ModUp
// (the hoisted rotation set), every inner sum of all g giant steps runs in ONE
// fused plaintext dot-product launch, and the giant steps are applied Horner
// style as the reference does: acc = inner[g-1]; acc = rot_{b s}(acc) +
// inner[j] for j = g-2 .. 0 -- or, with batched_giant, all g-1 giant-step
// rotations as one batched key switch (one launch per stage instead of g-1
// dependent key switches; needs the keys of every j b s).  The result keeps
// noise degree 2 (no rescale).
#include <algorithm>
#include <cstring>

#include "context.hpp"

namespace fides {

static void ck(hipError_t e, const char* w) { hip_check(e, w); }

// ------------------------------------------ extended-basis (P*Q) path ----
// The reference's LinearTransform when every diagonal is a P*Q plaintext
// (LinearTransform.cu:82-193, ext = true, MODDOWN_HOIST, ONLY_C1; OpenFHE's
// EvalFastRotationExt):
//   * baby steps: one ModUp of c1, each rotation's key switch left in P*Q
//     (no ModDown): (sigma(acc_b + P c0), sigma(acc_a)); the zero rotation is
//     (P c0, P c1) (extend());
//   * inner sums: the dot products over all nl + K limbs with the diagonals'
//     own special limbs;
//   * giant steps, Horner order: only c1 of the running sum is brought down
//     to Q (ModDown), switched, and the result kept in P*Q:
//     acc <- sigma(acc_b + acc.c0) + inner[j], sigma(acc_a) + inner[j].c1;
//   * the offset rotation the same way, then ONE ModDown of both parts.
// Per transform that is g + 2 polynomial ModDowns (g + 1 without an offset)
// instead of 2 (b + g) - 2 on the Q-basis path.  With batched_giant every
// giant step's c1 is brought down and switched together and their rotations
// summed in one gather launch.
namespace {

// Device copy of a host table; the buffer lives in `keep` until the caller
// returns (the pool recycles it in stream order, after the launches read it).
template <class J>
const J* upload_table(Context& c, std::vector<DevBuf>& keep, const std::vector<J>& v) {
    if (v.empty()) return nullptr;
    keep.emplace_back(&c.pool, (v.size() * sizeof(J) + 7) / 8);
    c.upload(keep.back().ptr, v.data(), v.size() * sizeof(J));
    return (const J*)keep.back().ptr;
}

template <class JL, class J>
JL table_list(const J* dev, int n) {
    JL l;
    std::memset(&l, 0, sizeof(l));
    l.dev = dev;
    l.n = n;
    return l;
}

// A job list of `jobs`: inline in the kernel arguments when it fits, else a
// device table.
template <class JL, class J>
JL job_list(Context& c, std::vector<DevBuf>& keep, const std::vector<J>& jobs) {
    JL l;
    std::memset(&l, 0, sizeof(l));
    l.n = (int)jobs.size();
    if (jobs.size() <= sizeof(l.inl) / sizeof(l.inl[0]))
        for (size_t i = 0; i < jobs.size(); ++i) l.inl[i] = jobs[i];
    else
        l.dev = upload_table(c, keep, jobs);
    return l;
}

// The inner sums of every giant step of every ciphertext in ONE launch: a
// job per ciphertext and chunk of up to DOT_GMAX giant steps (each baby step
// read once per job); job records and pointer tables go up in one upload.
// baby(k, t, part) is baby step t of ciphertext k, out(k, j, part) inner sum j;
// limbs nq .. nl-1 are special (P*Q diagonals).
template <class BabyF, class OutF>
void bsgs_inner_sums(Context& c, int B, int bstep, const std::vector<const Plaintext*>& diag, int nl, int nq,
                     BabyF baby, OutF out, std::vector<DevBuf>& keep) {
    const int n = (int)diag.size(), g = (n + bstep - 1) / bstep;
    struct Chunk {
        int k, j0, gc;
    };
    std::vector<Chunk> ch;
    size_t words = 0;
    for (int k = 0; k < B; ++k)
        for (int j0 = 0; j0 < g; j0 += DOT_GMAX) {
            const int gc = std::min(DOT_GMAX, g - j0);
            ch.push_back(Chunk{k, j0, gc});
            words += 2 * (size_t)bstep + 2 * (size_t)gc * bstep + 2 * (size_t)gc;
        }
    const size_t jb = (ch.size() * sizeof(DotJob) + 255) & ~(size_t)255;
    keep.emplace_back(&c.pool, (jb + words * 8 + 7) / 8);
    char* d = (char*)keep.back().ptr;
    std::vector<char> host(jb + words * 8);
    uint64_t* tab = (uint64_t*)(host.data() + jb);
    size_t at = 0;
    auto put = [&](const void* p) { tab[at++] = (uint64_t)(uintptr_t)p; };
    auto dev = [&]() { return (const uint64_t* const*)(d + jb + at * 8); };
    std::vector<DotJob> jobs;
    int max_gc = 0;
    for (const Chunk& q : ch) {
        DotJob J;
        std::memset(&J, 0, sizeof(J));
        J.x0 = dev();
        for (int t = 0; t < bstep; ++t) put(baby(q.k, t, 0));
        J.x1 = dev();
        for (int t = 0; t < bstep; ++t) put(baby(q.k, t, 1));
        J.w = dev();
        for (int j = q.j0; j < q.j0 + q.gc; ++j)
            for (int t = 0; t < bstep; ++t) put(bstep * j + t < n ? diag[bstep * j + t]->c0() : nullptr);
        J.wp = dev();
        for (int j = q.j0; j < q.j0 + q.gc; ++j)
            for (int t = 0; t < bstep; ++t) {
                const int idx = bstep * j + t;
                put(idx < n && nq < nl ? diag[idx]->c0() + (size_t)(diag[idx]->level + 1) * c.N : nullptr);
            }
        J.o0 = (uint64_t* const*)dev();
        for (int j = q.j0; j < q.j0 + q.gc; ++j) put(out(q.k, j, 0));
        J.o1 = (uint64_t* const*)dev();
        for (int j = q.j0; j < q.j0 + q.gc; ++j) put(out(q.k, j, 1));
        J.b = bstep;
        J.gc = q.gc;
        J.nl = nl;
        J.nq = nq;
        jobs.push_back(J);
        max_gc = std::max(max_gc, q.gc);
    }
    std::memcpy(host.data(), jobs.data(), jobs.size() * sizeof(DotJob));
    c.upload(d, host.data(), host.size());
    ck(launch_dot_bsgs(c.T, table_list<DotJobs>((const DotJob*)d, (int)jobs.size()), nl, max_gc, c.stream),
       "linear transform inner sums");
}

// One key switch of x (Q basis, evaluation form) into the P*Q accumulators.
struct KsReq {
    const uint64_t* x;
    const KeySwitchingKey* key;
    uint64_t* acc_a;
    uint64_t* acc_b;
};

// One ModUp per distinct source, then every key switch in one KSK launch.
void ext_keyswitch(Context& c, int level, const std::vector<KsReq>& reqs, std::vector<DevBuf>& keep) {
    if (reqs.empty()) return;
    const int nl = level + 1, nd = c.num_digits(level);
    const size_t N = c.N, npos = (size_t)nl + c.K;
    std::vector<const uint64_t*> xs;
    for (const KsReq& r : reqs)
        if (std::find(xs.begin(), xs.end(), r.x) == xs.end()) xs.push_back(r.x);
    keep.emplace_back(&c.pool, (size_t)nl * N);
    uint64_t* t = keep.back().ptr;
    keep.emplace_back(&c.pool, xs.size() * nd * npos * N);
    uint64_t* ext = keep.back().ptr;
    for (size_t i = 0; i < xs.size(); ++i) modup_to(c, level, xs[i], t, ext + i * nd * npos * N);
    std::vector<KskJob> kj;
    for (const KsReq& r : reqs) {
        const size_t i = std::find(xs.begin(), xs.end(), r.x) - xs.begin();
        kj.push_back(KskJob{ext + i * nd * npos * N, r.x, r.key->a.ptr, r.key->b.ptr, r.acc_a, r.acc_b, level});
    }
    // several sources switched by the same keys (the baby steps of a batch of
    // ciphertexts, a giant step of each): the jobs grouped by key and run
    // key-stationary, each key read once per chunk
    if (xs.size() > 1 && c.dnum <= 4 && ksk_keystat_on()) {
        std::stable_sort(kj.begin(), kj.end(), [](const KskJob& x, const KskJob& y) { return x.ka < y.ka; });
        std::vector<int> chunks{0};
        int nkeys = 0;
        for (int j = 0; j < (int)kj.size(); ++j) {
            const bool new_key = j == 0 || kj[j].ka != kj[j - 1].ka;
            nkeys += new_key ? 1 : 0;
            if (j > 0 && (new_key || j - chunks.back() == ksk_keystat_chunk())) chunks.push_back(j);
        }
        chunks.push_back((int)kj.size());
        if (2 * nkeys <= (int)kj.size()) {
            const int* dch = upload_table(c, keep, chunks);
            ck(launch_ksk_keystat(c.T, job_list<KskJobs>(c, keep, kj), dch, (int)chunks.size() - 1, c.stream),
               "ext key switch (key-stationary)");
            return;
        }
    }
    ck(launch_ksk_inner(c.T, job_list<KskJobs>(c, keep, kj), (int)npos, c.stream, xs.size() == 1),
       "ext key switch");
}

// ModDown of P*Q polynomials ([nl + K][N]) into Q-basis ones ([nl][N]).
void ext_moddown(Context& c, int level, const std::vector<std::pair<const uint64_t*, uint64_t*>>& polys,
                 std::vector<DevBuf>& keep) {
    if (polys.empty()) return;
    const int nl = level + 1, K = c.K;
    const size_t N = c.N, n = polys.size();
    keep.emplace_back(&c.pool, n * K * N);
    uint64_t* tk = keep.back().ptr;
    keep.emplace_back(&c.pool, n * nl * N);
    uint64_t* y = keep.back().ptr;
    std::vector<NttJob> isp, fwd;
    std::vector<ModDownJob> md;
    for (size_t i = 0; i < n; ++i) {
        uint64_t* ti = tk + i * K * N;
        uint64_t* yi = y + i * nl * N;
        NttJob j;
        std::memset(&j, 0, sizeof(j));
        j.src = polys[i].first + (size_t)nl * N;  // the special limbs
        j.mid = j.dst = ti;
        j.npos = K;
        j.scale_sel = ISCALE_MODDOWN;
        isp.push_back(j);
        md.push_back(ModDownJob{ti, yi, nullptr, nullptr, level});
        NttJob f;
        std::memset(&f, 0, sizeof(f));
        f.src = f.mid = yi;
        f.dst = polys[i].second;
        f.aux = polys[i].first;  // its Q limbs
        f.npos = nl;
        f.nq = nl;
        fwd.push_back(f);
    }
    run_ntt_jobs(c, isp, true);
    ck(launch_moddown_conv(c.T, job_list<ModDownJobs>(c, keep, md), nl, c.stream),
       "ext moddown conv");
    run_ntt_jobs(c, fwd, false, LOAD_PLAIN, STORE_MODDOWN);
}

struct RotOut {
    std::vector<ExtRotTerm> terms;
    const uint64_t* base0;
    const uint64_t* base1;
    uint64_t* o0;
    uint64_t* o1;
};

void ext_rot_outs(Context& c, int level, const std::vector<RotOut>& outs, std::vector<DevBuf>& keep) {
    if (outs.empty()) return;
    std::vector<ExtRotTerm> more;
    std::vector<size_t> first;
    for (const RotOut& o : outs) {
        if (o.terms.empty()) throw Error(E_STATE, "rotation output without terms");
        first.push_back(more.size());
        more.insert(more.end(), o.terms.begin() + 1, o.terms.end());
    }
    const ExtRotTerm* dm = upload_table(c, keep, more);
    std::vector<ExtRotJob> jobs;
    for (size_t i = 0; i < outs.size(); ++i)
        jobs.push_back(ExtRotJob{outs[i].terms[0], dm ? dm + first[i] : nullptr, outs[i].base0, outs[i].base1,
                                 outs[i].o0, outs[i].o1, (int)outs[i].terms.size(), level + 1});
    ck(launch_ext_rot(c.T, job_list<ExtRotJobs>(c, keep, jobs), level + 1, c.stream), "ext rotation outputs");
}

}  // namespace

static void lt_ext(Context& c, const std::vector<const Ciphertext*>& as, const std::vector<const Plaintext*>& diag,
                   int bstep, int stride, int offset, const std::vector<Ciphertext*>& outs, bool batched_giant) {
    const int B = (int)as.size(), n = (int)diag.size(), g = (n + bstep - 1) / bstep;
    const int level = as[0]->level, nl = level + 1;
    const size_t N = c.N, pw = ((size_t)nl + c.K) * N;  // words per P*Q part
    const int half = c.N / 2;
    std::vector<DevBuf> keep;
    auto alloc = [&](size_t words) {
        keep.emplace_back(&c.pool, words);
        return keep.back().ptr;
    };
    auto key_of = [&](int64_t r) -> const KeySwitchingKey* {
        const uint32_t k = c.rot_index((int)(r % half));
        auto it = c.rot_keys.find(k);
        if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for automorphism index " + std::to_string(k));
        return it->second.get();
    };

    // baby steps in P*Q: baby(k, i) = ext rot_{i s}(a_k)
    uint64_t* baby = alloc((size_t)B * bstep * 2 * pw);
    uint64_t* bacc = alloc((size_t)B * bstep * 2 * pw);
    auto bab = [&](int k, int i) { return baby + (size_t)(k * bstep + i) * 2 * pw; };
    {
        std::vector<KsReq> reqs;
        std::vector<RotOut> ro;
        for (int k = 0; k < B; ++k)
            for (int i = 0; i < bstep; ++i) {
                const int64_t r = (int64_t)i * stride;
                uint64_t* o0 = bab(k, i);
                if (r % half == 0) {
                    ro.push_back(RotOut{{ExtRotTerm{nullptr, nullptr, as[k]->c0(), as[k]->c1(), 1u, 1}}, nullptr,
                                        nullptr, o0, o0 + pw});
                    continue;
                }
                uint64_t* aa = bacc + (size_t)(k * bstep + i) * 2 * pw;
                reqs.push_back(KsReq{as[k]->c1(), key_of(r), aa, aa + pw});
                ro.push_back(RotOut{{ExtRotTerm{aa, aa + pw, as[k]->c0(), nullptr, c.rot_index((int)(r % half)), 1}},
                                    nullptr, nullptr, o0, o0 + pw});
            }
        ext_keyswitch(c, level, reqs, keep);
        ext_rot_outs(c, level, ro, keep);
    }

    // inner sums over P*Q: one launch for every giant step of every ciphertext
    uint64_t* inner = alloc((size_t)B * g * 2 * pw);
    auto inn = [&](int k, int j) { return inner + (size_t)(k * g + j) * 2 * pw; };
    bsgs_inner_sums(
        c, B, bstep, diag, nl + c.K, nl, [&](int k, int t, int h) { return bab(k, t) + h * pw; },
        [&](int k, int j, int h) { return inn(k, j) + h * pw; }, keep);

    // giant steps; c1 of a P*Q sum goes down to Q (ModDown) before its switch
    uint64_t* d1 = alloc((size_t)B * std::max(1, batched_giant ? g - 1 : 1) * nl * N);
    uint64_t* gacc = alloc((size_t)B * std::max(1, batched_giant ? g - 1 : 1) * 2 * pw);
    const int64_t gstep = (int64_t)bstep * stride;
    if (!batched_giant) {
        for (int j = g - 2; j >= 0; --j) {
            std::vector<RotOut> ro;
            if (gstep % half != 0) {
                std::vector<std::pair<const uint64_t*, uint64_t*>> md;
                std::vector<KsReq> reqs;
                for (int k = 0; k < B; ++k) {
                    md.emplace_back(inn(k, j + 1) + pw, d1 + (size_t)k * nl * N);
                    uint64_t* aa = gacc + (size_t)k * 2 * pw;
                    reqs.push_back(KsReq{d1 + (size_t)k * nl * N, key_of(gstep), aa, aa + pw});
                    ro.push_back(RotOut{{ExtRotTerm{aa, aa + pw, inn(k, j + 1), nullptr,
                                                    c.rot_index((int)(gstep % half)), 0}},
                                        inn(k, j), inn(k, j) + pw, inn(k, j), inn(k, j) + pw});
                }
                ext_moddown(c, level, md, keep);
                ext_keyswitch(c, level, reqs, keep);
            } else {
                for (int k = 0; k < B; ++k)
                    ro.push_back(RotOut{{ExtRotTerm{nullptr, nullptr, inn(k, j + 1), inn(k, j + 1) + pw, 1u, 0}},
                                        inn(k, j), inn(k, j) + pw, inn(k, j), inn(k, j) + pw});
            }
            ext_rot_outs(c, level, ro, keep);
        }
    } else if (g > 1) {
        std::vector<std::pair<const uint64_t*, uint64_t*>> md;
        std::vector<KsReq> reqs;
        std::vector<RotOut> ro(B);
        for (int k = 0; k < B; ++k) {
            ro[k] = RotOut{{}, inn(k, 0), inn(k, 0) + pw, inn(k, 0), inn(k, 0) + pw};
            for (int j = 1; j < g; ++j) {
                const int64_t r = (int64_t)j * gstep;
                if (r % half == 0) {
                    ro[k].terms.push_back(ExtRotTerm{nullptr, nullptr, inn(k, j), inn(k, j) + pw, 1u, 0});
                    continue;
                }
                uint64_t* dj = d1 + (size_t)(k * (g - 1) + j - 1) * nl * N;
                uint64_t* aa = gacc + (size_t)(k * (g - 1) + j - 1) * 2 * pw;
                md.emplace_back(inn(k, j) + pw, dj);
                reqs.push_back(KsReq{dj, key_of(r), aa, aa + pw});
                ro[k].terms.push_back(ExtRotTerm{aa, aa + pw, inn(k, j), nullptr, c.rot_index((int)(r % half)), 0});
            }
        }
        ext_moddown(c, level, md, keep);
        ext_keyswitch(c, level, reqs, keep);
        ext_rot_outs(c, level, ro, keep);
    }

    // offset rotation (kept in P*Q), then one ModDown of both parts
    std::vector<std::pair<const uint64_t*, uint64_t*>> fin;
    if (offset % half != 0) {
        uint64_t* rot = alloc((size_t)B * 2 * pw);
        std::vector<std::pair<const uint64_t*, uint64_t*>> md;
        std::vector<KsReq> reqs;
        std::vector<RotOut> ro;
        for (int k = 0; k < B; ++k) {
            md.emplace_back(inn(k, 0) + pw, d1 + (size_t)k * nl * N);
            uint64_t* aa = gacc + (size_t)k * 2 * pw;
            reqs.push_back(KsReq{d1 + (size_t)k * nl * N, key_of(offset), aa, aa + pw});
            uint64_t* o = rot + (size_t)k * 2 * pw;
            ro.push_back(RotOut{{ExtRotTerm{aa, aa + pw, inn(k, 0), nullptr, c.rot_index(offset % half), 0}},
                                nullptr, nullptr, o, o + pw});
            fin.emplace_back(o, outs[k]->c0());
            fin.emplace_back(o + pw, outs[k]->c1());
        }
        ext_moddown(c, level, md, keep);
        ext_keyswitch(c, level, reqs, keep);
        ext_rot_outs(c, level, ro, keep);
    } else {
        for (int k = 0; k < B; ++k) {
            fin.emplace_back(inn(k, 0), outs[k]->c0());
            fin.emplace_back(inn(k, 0) + pw, outs[k]->c1());
        }
    }
    ext_moddown(c, level, fin, keep);
    for (int k = 0; k < B; ++k) {
        Ciphertext& o = *outs[k];
        o.noise = as[k]->noise * diag[0]->noise;
        o.slots = as[k]->slots;
        o.level = level;
        o.noise_level = 2;
    }
}

// The transform of several ciphertexts with the same diagonals: baby steps
// per ciphertext (one hoisted set each), the inner sums of every giant step
// of every ciphertext in ONE dot launch, and each Horner giant step (and the
// offset) as one batched rotation over all of them.  B = 1 is
// eval_linear_transform.
void eval_linear_transform_batch(Context& c, const std::vector<const Ciphertext*>& as_in,
                                 const std::vector<const Plaintext*>& diag, int bstep, int stride, int offset,
                                 const std::vector<Ciphertext*>& outs, bool batched_giant) {
    const int B = (int)as_in.size();
    if (B < 1 || (int)outs.size() != B) throw Error(E_ARG, "batch size mismatch");
    const int n = (int)diag.size();
    if (n < 1 || bstep < 1 || bstep > n) throw Error(E_ARG, "need 1 <= bstep <= number of diagonals");
    const int g = (n + bstep - 1) / bstep;
    std::vector<std::unique_ptr<Ciphertext>> tmp(B);
    std::vector<const Ciphertext*> as(B);
    for (int k = 0; k < B; ++k) {
        const Ciphertext& a_in = *as_in[k];
        if (a_in.level < 0) throw Error(E_STATE, "ciphertext not loaded");
        as[k] = &a_in;
        if (a_in.noise_level == 2) {  // LinearTransform rescales a degree-2 input first
            tmp[k] = ct_create(c);
            ct_copy(c, a_in, *tmp[k]);
            rescale(c, *tmp[k]);
            as[k] = tmp[k].get();
        }
        if (as[k]->noise_level != 1) throw Error(E_STATE, "linear transform needs a noise level 1 input");
        if (as[k]->level != as[0]->level) throw Error(E_LEVEL, "a batched transform needs inputs at one level");
    }
    const int level = as[0]->level, nl = level + 1;
    bool ext = bstep > 1;
    for (const Plaintext* p : diag) {
        if (!p || p->level < level) throw Error(E_LEVEL, "diagonal missing or with fewer limbs than the input");
        if (p->noise_level != 1) throw Error(E_STATE, "diagonals must have noise level 1");
        ext = ext && p->ext;
    }
    if (ext) {  // every diagonal in P*Q form: the reference's extended-basis path
        lt_ext(c, as, diag, bstep, stride, offset, outs, batched_giant);
        return;
    }

    // baby steps: rot_{i s}(a), i < b, one shared ModUp per ciphertext
    std::vector<std::vector<std::unique_ptr<Ciphertext>>> baby(B);
    std::vector<int> idx;
    for (int i = 0; i < bstep; ++i) idx.push_back(i * stride);
    for (int k = 0; k < B; ++k) {
        std::vector<Ciphertext*> bp;
        for (int i = 0; i < bstep; ++i) {
            baby[k].push_back(ct_create(c));
            bp.push_back(baby[k].back().get());
        }
        eval_rotate_hoisted(c, *as[k], idx, bp);
    }

    // inner sums of every giant step of every ciphertext: one launch
    std::vector<std::vector<std::unique_ptr<Ciphertext>>> inner(B);
    for (int k = 0; k < B; ++k)
        for (int j = 0; j < g; ++j) inner[k].push_back(ct_create(c));
    std::vector<DevBuf> keep;
    bsgs_inner_sums(
        c, B, bstep, diag, nl, nl, [&](int k, int t, int h) { return h ? baby[k][t]->c1() : baby[k][t]->c0(); },
        [&](int k, int j, int h) { return h ? inner[k][j]->c1() : inner[k][j]->c0(); }, keep);
    for (int k = 0; k < B; ++k)
        for (int j = 0; j < g; ++j) {
            Ciphertext& o = *inner[k][j];
            o.level = level;
            o.noise = as[k]->noise * diag[bstep * j]->noise;
            o.noise_level = 2;
            o.slots = as[k]->slots;
        }

    // a rotation of every accumulator by r (one batched key switch when B > 1)
    auto rotate_all = [&](const std::vector<Ciphertext*>& src, int r, const std::vector<Ciphertext*>& dst) {
        if (B == 1)
            eval_rotate(c, *src[0], r, *dst[0]);
        else
            eval_rotate_batch(c, std::vector<const Ciphertext*>(src.begin(), src.end()), std::vector<int>(B, r),
                              dst);
    };
    const int half = c.N / 2;
    std::vector<Ciphertext*> acc(B);
    for (int k = 0; k < B; ++k) acc[k] = inner[k][0].get();
    if (batched_giant) {
        // every giant step rotated in ONE batched key switch (keys of all
        // j b s), then summed: sum_j rot_{j b s}(inner[j])
        std::vector<std::unique_ptr<Ciphertext>> rots;
        std::vector<const Ciphertext*> ins;
        std::vector<Ciphertext*> routs;
        std::vector<int> gi;
        for (int k = 0; k < B; ++k)
            for (int j = 1; j < g; ++j) {
                rots.push_back(ct_create(c));
                ins.push_back(inner[k][j].get());
                routs.push_back(rots.back().get());
                gi.push_back(j * bstep * stride);
            }
        eval_rotate_batch(c, ins, gi, routs);
        for (int k = 0; k < B; ++k)
            for (int j = 1; j < g; ++j) eval_add(c, *acc[k], *rots[(size_t)k * (g - 1) + j - 1], *acc[k]);
    } else {
        // Horner order, as the reference: acc = rot_{b s}(acc) + inner[j]
        std::vector<std::unique_ptr<Ciphertext>> rot(B);
        std::vector<Ciphertext*> rp(B);
        for (int k = 0; k < B; ++k) {
            rot[k] = ct_create(c);
            rp[k] = rot[k].get();
            acc[k] = inner[k][g - 1].get();
        }
        for (int j = g - 2; j >= 0; --j) {
            if (((int64_t)bstep * stride) % half != 0) {
                rotate_all(acc, bstep * stride, rp);
                for (int k = 0; k < B; ++k) eval_add(c, *inner[k][j], *rot[k], *inner[k][j]);
            } else {
                for (int k = 0; k < B; ++k) eval_add(c, *inner[k][j], *acc[k], *inner[k][j]);
            }
            for (int k = 0; k < B; ++k) acc[k] = inner[k][j].get();
        }
    }
    if (offset % half != 0)
        rotate_all(acc, offset, outs);
    else
        for (int k = 0; k < B; ++k) ct_copy(c, *acc[k], *outs[k]);
}

// AccumulateSum (api/CryptoContext.cpp:1649-1700; FIDESlib CKKS::Accumulate,
// src/CKKS/AccumulateBroadcast.cu:41-67): out = sum_(k < size) rot_(k stride)(a)
// in radix-4 steps -- for s = 1, 4, 16, .. < size the rotations by i s stride
// (i = 1..3, below size stride) as one extended-basis hoisted set, added in
// P*Q to the extended running sum, then ONE ModDown per step, as the
// reference does (rotate_hoisted(.., ext = true), extend(), add, modDown).
// Each step is one ModUp, one KSK launch, one gather launch summing the
// extended input and every rotation, and one ModDown.
std::vector<int> accumulate_rotations(int size, int stride) {
    if (size < 1 || (size & (size - 1)) || stride < 1) throw Error(E_ARG, "size must be a power of two, stride >= 1");
    std::vector<int> idx;
    for (int64_t s = 1; s < size; s *= 4)
        for (int64_t i = 1; i < 4 && i * s < size; ++i) idx.push_back((int)(i * s * stride));
    return idx;
}

void eval_accumulate_sum(Context& c, const Ciphertext& a, int size, int stride, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a.ext) throw Error(E_STATE, "ciphertext in P*Q form");
    (void)accumulate_rotations(size, stride);  // argument checks
    const int in_slots = a.slots;
    const int half = c.N / 2;
    std::unique_ptr<Ciphertext> acc = ct_create(c);
    ct_copy(c, a, *acc);
    const int level = acc->level, nl = level + 1;
    const size_t pw = ((size_t)nl + c.K) * c.N;
    for (int64_t s = 1; s < size; s *= 4) {
        std::vector<DevBuf> keep;
        keep.emplace_back(&c.pool, 3 * 2 * pw + 2 * pw);
        uint64_t* accs = keep.back().ptr;
        uint64_t* sum = accs + 3 * 2 * pw;
        std::vector<KsReq> reqs;
        RotOut ro{{ExtRotTerm{nullptr, nullptr, acc->c0(), acc->c1(), 1u, 1}}, nullptr, nullptr, sum, sum + pw};
        int r_i = 0;
        for (int64_t i = 1; i < 4 && i * s < size; ++i) {
            const int r = (int)((i * s * stride) % half);
            if (r == 0) {  // the identity: extend(acc) once more
                ro.terms.push_back(ExtRotTerm{nullptr, nullptr, acc->c0(), acc->c1(), 1u, 1});
                continue;
            }
            const uint32_t k = c.rot_index(r);
            auto it = c.rot_keys.find(k);
            if (it == c.rot_keys.end())
                throw Error(E_NOKEY, "no rotation key for automorphism index " + std::to_string(k));
            uint64_t* aa = accs + (size_t)r_i++ * 2 * pw;
            reqs.push_back(KsReq{acc->c1(), it->second.get(), aa, aa + pw});
            ro.terms.push_back(ExtRotTerm{aa, aa + pw, acc->c0(), nullptr, k, 1});
        }
        ext_keyswitch(c, level, reqs, keep);
        ext_rot_outs(c, level, {ro}, keep);
        ext_moddown(c, level, {{sum, acc->c0()}, {sum + pw, acc->c1()}}, keep);
    }
    ct_copy(c, *acc, out);
    // a sum over every slot leaves `stride` distinct slot values
    // (AccumulateBroadcast.cu:64-65)
    if ((int64_t)size * stride == in_slots) out.slots = stride;
}

// ------------------------------------------ P*Q ciphertexts at the API ----
// An output that aliases an input of a different layout is computed into a
// fresh P*Q buffer that then replaces the output's (pool blocks recycle in
// stream order, so the input stays valid for the launches already queued).
static std::unique_ptr<Ciphertext> ext_target(Context& c) {
    std::unique_ptr<Ciphertext> t = ct_create(c);
    ct_reserve_ext(c, *t);
    return t;
}
static void adopt(Ciphertext& out, std::unique_ptr<Ciphertext>& t) {
    out.data = std::move(t->data);
    out.part = t->part;
}
static bool same_buf(const Ciphertext& a, const Ciphertext& b) { return a.data.ptr == b.data.ptr; }

// Ciphertext::extend (Ciphertext.cpp:742-754): (P c0, P c1), special limbs 0.
void eval_extend(Context& c, const Ciphertext& a, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a.ext) throw Error(E_STATE, "ciphertext already in P*Q form");
    std::unique_ptr<Ciphertext> t = ext_target(c);
    std::vector<DevBuf> keep;
    ext_rot_outs(c, a.level,
                 {RotOut{{ExtRotTerm{nullptr, nullptr, a.c0(), a.c1(), 1u, 1}}, nullptr, nullptr, t->c0(), t->c1()}},
                 keep);
    const int level = a.level, nlev = a.noise_level, slots = a.slots;
    const double noise = a.noise;
    adopt(out, t);
    out.level = level;
    out.ext = 1;
    out.noise = noise;
    out.noise_level = nlev;
    out.slots = slots;
}

// Ciphertext::modDown (Ciphertext.cpp:311-320): both parts P*Q -> Q.
void eval_mod_down(Context& c, const Ciphertext& a, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (!a.ext) throw Error(E_STATE, "ciphertext not in P*Q form");
    std::vector<DevBuf> keep;
    // in place is safe: each ModDown output element is written by the thread
    // that read that element's Q residue
    ext_moddown(c, a.level, {{a.c0(), out.c0()}, {a.c1(), out.c1()}}, keep);
    if (&out != &a) {
        out.level = a.level;
        out.noise = a.noise;
        out.noise_level = a.noise_level;
        out.slots = a.slots;
    }
    out.ext = 0;
}

// Ciphertext::rotate_hoisted(indexes, results, ext = true)
// (Ciphertext.cpp:1053-1226): one ModUp of c1, each output left in P*Q
// (sigma(acc_b + P c0), sigma(acc_a)); index 0 gives extend(a).
void eval_rotate_hoisted_ext(Context& c, const Ciphertext& a, const std::vector<int>& idx,
                             const std::vector<Ciphertext*>& outs) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a.ext) throw Error(E_STATE, "ciphertext in P*Q form");
    if (idx.size() != outs.size() || idx.empty()) throw Error(E_ARG, "need one output per index");
    for (size_t i = 0; i < outs.size(); ++i) {
        if (same_buf(*outs[i], a)) throw Error(E_ARG, "an output aliases the input");
        for (size_t j = 0; j < i; ++j)
            if (same_buf(*outs[i], *outs[j])) throw Error(E_ARG, "outputs must be distinct");
    }
    const int level = a.level, half = c.N / 2;
    const size_t pw = ((size_t)level + 1 + c.K) * c.N;
    std::vector<DevBuf> keep;
    keep.emplace_back(&c.pool, std::max<size_t>(1, idx.size()) * 2 * pw);
    uint64_t* accs = keep.back().ptr;
    std::vector<KsReq> reqs;
    std::vector<RotOut> ro;
    for (size_t i = 0; i < idx.size(); ++i) {
        ct_reserve_ext(c, *outs[i]);
        uint64_t* o0 = outs[i]->c0();
        uint64_t* o1 = outs[i]->c1();
        const int r = ((idx[i] % half) + half) % half;
        if (r == 0) {
            ro.push_back(RotOut{{ExtRotTerm{nullptr, nullptr, a.c0(), a.c1(), 1u, 1}}, nullptr, nullptr, o0, o1});
            continue;
        }
        const uint32_t k = c.rot_index(r);
        auto it = c.rot_keys.find(k);
        if (it == c.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for automorphism index " + std::to_string(k));
        uint64_t* aa = accs + i * 2 * pw;
        reqs.push_back(KsReq{a.c1(), it->second.get(), aa, aa + pw});
        ro.push_back(RotOut{{ExtRotTerm{aa, aa + pw, a.c0(), nullptr, k, 1}}, nullptr, nullptr, o0, o1});
    }
    ext_keyswitch(c, level, reqs, keep);
    ext_rot_outs(c, level, ro, keep);
    for (Ciphertext* o : outs) {
        o->level = a.level;
        o->ext = 1;
        o->noise = a.noise;
        o->noise_level = a.noise_level;
        o->slots = a.slots;
    }
}

// Ciphertext::add with P*Q operands (RNSPoly::add, RNSPoly.cpp:180-195: a Q
// operand is extended, P x, first).
void eval_add_ext(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out) {
    if (a.level < 0 || b.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (!a.ext && !b.ext) throw Error(E_STATE, "neither operand in P*Q form (EvalAdd adds Q-basis ciphertexts)");
    if (a.level != b.level) throw Error(E_LEVEL, "P*Q operands must be at one level");
    std::unique_ptr<Ciphertext> t = ext_target(c);
    std::vector<DevBuf> keep;
    ext_rot_outs(c, a.level,
                 {RotOut{{ExtRotTerm{nullptr, nullptr, a.c0(), a.c1(), 1u, a.ext ? 0 : 1},
                          ExtRotTerm{nullptr, nullptr, b.c0(), b.c1(), 1u, b.ext ? 0 : 1}},
                         nullptr, nullptr, t->c0(), t->c1()}},
                 keep);
    const int level = a.level, nlev = std::max(a.noise_level, b.noise_level), slots = std::max(a.slots, b.slots);
    const double noise = a.noise;
    adopt(out, t);
    out.level = level;
    out.ext = 1;
    out.noise = noise;
    out.noise_level = nlev;
    out.slots = slots;
}

// Ciphertext::mult(b, rescale = false, moddown = false) (Ciphertext.cuh:264):
// the tensor product, then the key switch of d2 left in P*Q:
// (P d0 + KS_b(d2), P d1 + KS_a(d2)).  Its ModDown equals EvalMult's result
// (P d0 has no special part, so the conversion does not see it).
void eval_mult_ext(Context& c, const Ciphertext& a, const Ciphertext& b, Ciphertext& out) {
    if (a.level < 0 || b.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a.ext || b.ext) throw Error(E_STATE, "operands must be in the Q basis");
    if (a.noise_level != 1 || b.noise_level != 1)
        throw Error(E_STATE, "EvalMult needs both inputs at noise level 1 (rescale first)");
    if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key");
    const int level = std::min(a.level, b.level), nl = level + 1;
    const size_t N = c.N, pw = ((size_t)nl + c.K) * N;
    std::vector<DevBuf> keep;
    keep.emplace_back(&c.pool, 3 * (size_t)nl * N + 2 * pw);
    uint64_t* d0 = keep.back().ptr;
    uint64_t* d1 = d0 + (size_t)nl * N;
    uint64_t* d2 = d1 + (size_t)nl * N;
    uint64_t* acc = d2 + (size_t)nl * N;
    TensorJobs tj;
    std::memset(&tj, 0, sizeof(tj));
    tj.inl[0] = TensorJob{a.c0(), a.c1(), b.c0(), b.c1(), d0, d1, d2, nl};
    tj.n = 1;
    ck(launch_tensor(c.T, tj, nl, c.stream), "mult ext tensor");
    ext_keyswitch(c, level, {KsReq{d2, c.eval_key.get(), acc, acc + pw}}, keep);
    std::unique_ptr<Ciphertext> t = ext_target(c);
    ext_rot_outs(c, level, {RotOut{{ExtRotTerm{acc, acc + pw, d0, d1, 1u, 1}}, nullptr, nullptr, t->c0(), t->c1()}},
                 keep);
    const double noise = a.noise * b.noise;
    const int slots = std::max(a.slots, b.slots);
    adopt(out, t);
    out.level = level;
    out.ext = 1;
    out.noise = noise;
    out.noise_level = 2;
    out.slots = slots;
}

// Ciphertext::dotProduct (Ciphertext.cpp:1665-1707; RNSPoly::dotProduct,
// RNSPoly.cpp:941-987): out = sum_i a[i] b[i] with ONE key switch of the
// summed d2.  The tensor products of all pairs are summed in one launch
// (k_tensor_sum); then ModUp and KSK, and either ModDown added to the summed
// (d0, d1) (Q form) or, with ext, the P*Q form (P d0 + acc_b, P d1 + acc_a)
// that eval_mult_ext leaves a single product in.  Not rescaled (noise degree
// 2, scale a[0] b[0]), as the reference's multMetadata.
void eval_dot_product(Context& c, const std::vector<const Ciphertext*>& a, const std::vector<const Ciphertext*>& b,
                      bool ext, Ciphertext& out) {
    const int n = (int)a.size();
    if (n < 1 || (int)b.size() != n) throw Error(E_ARG, "need as many ciphertexts on both sides");
    if (!c.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
    if (!a[0] || a[0]->level < 0) throw Error(E_STATE, "ciphertext not loaded");
    const int level = a[0]->level;
    int slots = 0;
    for (int i = 0; i < n; ++i)
        for (const Ciphertext* x : {a[i], b[i]}) {
            if (!x || x->level < 0) throw Error(E_STATE, "ciphertext not loaded");
            if (x->ext) throw Error(E_STATE, "operands must be in the Q basis");
            if (x->noise_level != 1) throw Error(E_STATE, "dot product needs inputs at noise level 1");
            if (x->level != level) throw Error(E_LEVEL, "ciphertexts must be at one level");
            slots = std::max(slots, x->slots);
        }
    const double noise = a[0]->noise * b[0]->noise;
    const int nl = level + 1;
    const size_t N = c.N, pw = ((size_t)nl + c.K) * N;
    std::vector<const uint64_t*> host(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        host[i] = a[i]->c0();
        host[n + i] = a[i]->c1();
        host[2 * n + i] = b[i]->c0();
        host[3 * n + i] = b[i]->c1();
    }
    std::vector<DevBuf> keep;
    keep.emplace_back(&c.pool, host.size());
    c.upload(keep.back().ptr, host.data(), host.size() * sizeof(uint64_t*));
    const uint64_t* const* tab = (const uint64_t* const*)keep.back().ptr;
    if (!ext) {
        // in place is safe: a coefficient's terms are read by the thread that
        // writes it, and d2 has its own buffer
        keep.emplace_back(&c.pool, (size_t)nl * N);
        uint64_t* d2 = keep.back().ptr;
        ck(launch_tensor_sum(c.T, TensorSumJob{tab, out.c0(), out.c1(), d2, n, nl}, c.stream), "dot product tensor");
        relin_add(c, level, d2, out.c0(), out.c1());
        out.ext = 0;
    } else {
        keep.emplace_back(&c.pool, 3 * (size_t)nl * N + 2 * pw);
        uint64_t* d0 = keep.back().ptr;
        uint64_t* d1 = d0 + (size_t)nl * N;
        uint64_t* d2 = d1 + (size_t)nl * N;
        uint64_t* acc = d2 + (size_t)nl * N;
        ck(launch_tensor_sum(c.T, TensorSumJob{tab, d0, d1, d2, n, nl}, c.stream), "dot product tensor");
        ext_keyswitch(c, level, {KsReq{d2, c.eval_key.get(), acc, acc + pw}}, keep);
        std::unique_ptr<Ciphertext> t = ext_target(c);
        ext_rot_outs(c, level,
                     {RotOut{{ExtRotTerm{acc, acc + pw, d0, d1, 1u, 1}}, nullptr, nullptr, t->c0(), t->c1()}}, keep);
        adopt(out, t);
        out.ext = 1;
    }
    out.level = level;
    out.noise = noise;
    out.noise_level = 2;
    out.slots = slots;
}

// Ciphertext::dotProductPt (Ciphertext.cpp:1638-1663; RNSPoly::dotProductPt):
// out = sum_i a[i] * pt[i], in P*Q when the ciphertexts are (their
// plaintexts must be P*Q then too), else in Q.
void eval_dot_product_plain(Context& c, const std::vector<const Ciphertext*>& a,
                            const std::vector<const Plaintext*>& pt, Ciphertext& out) {
    const int n = (int)a.size();
    if (n < 1 || (int)pt.size() != n) throw Error(E_ARG, "need one plaintext per ciphertext");
    const int level = a[0]->level;
    const bool ext = a[0]->ext != 0;
    for (int i = 0; i < n; ++i) {
        if (!a[i] || !pt[i] || a[i]->level < 0 || pt[i]->level < 0) throw Error(E_STATE, "operand not loaded");
        if (a[i]->level != level) throw Error(E_LEVEL, "ciphertexts must be at one level");
        if ((a[i]->ext != 0) != ext) throw Error(E_STATE, "ciphertexts must all be in one form");
        if (pt[i]->level < level) throw Error(E_LEVEL, "plaintext has fewer limbs than the ciphertexts");
        if (ext && !pt[i]->ext) throw Error(E_STATE, "P*Q ciphertexts need P*Q plaintexts");
    }
    int slots = 0;
    for (int i = 0; i < n; ++i) slots = std::max(slots, std::max(a[i]->slots, pt[i]->slots));
    const double noise = a[0]->noise * pt[0]->noise;
    const int nlev = a[0]->noise_level + pt[0]->noise_level;
    if (ext) ct_reserve_ext(c, out);
    // in place is safe: every term of an output element is read by the
    // thread that writes it
    const int nl = level + 1;
    std::vector<DevBuf> keep;
    uint64_t* o0 = out.c0();
    uint64_t* o1 = out.c1();
    bsgs_inner_sums(
        c, 1, n, pt, ext ? nl + c.K : nl, nl, [&](int, int t, int h) { return h ? a[t]->c1() : a[t]->c0(); },
        [&](int, int, int h) { return h ? o1 : o0; }, keep);
    out.level = level;
    out.ext = ext ? 1 : 0;
    out.noise = noise;
    out.noise_level = nlev;
    out.slots = slots;
}

void eval_linear_transform(Context& c, const Ciphertext& a, const std::vector<const Plaintext*>& diag, int bstep,
                           int stride, int offset, Ciphertext& out, bool batched_giant) {
    eval_linear_transform_batch(c, {&a}, diag, bstep, stride, offset, {&out}, batched_giant);
}

// Rotations a convolution transform needs besides its baby-step indexes
// (FIDESlib GetConvolutionTransformRotationIndices, LinearTransform.cu:655-678,
// which lists s k for k < min(g, 8) only; the first giant step of a block of
// G also rotates by s G, so that one is listed too).
std::vector<int> convolution_rotations(int stride, int gstep) {
    std::vector<int> r;
    for (int k = 1; k <= std::min(gstep, CONV_BLOCK); ++k) r.push_back(stride * k);
    const int nb = (gstep + CONV_BLOCK - 1) / CONV_BLOCK;
    for (int k = 1; k < nb; ++k)
        if (k * CONV_BLOCK > std::min(gstep, CONV_BLOCK)) r.push_back(stride * CONV_BLOCK * k);
    return r;
}

// ConvolutionTransform (FIDESlib CKKS::ConvolutionTransform,
// LinearTransform.cu:255-441, with ext = false: bStep = 1 or the diagonals not
// in P*Q): baby steps v_i = rot_(idx_i)(a), one hoisted set; the inner sums
// u_j = sum_i pt[b j + i] (.) v_i of every giant step in one launch; giant
// steps in blocks of 8, u_j of a block of G rotated by s (G - j') (j' its place
// in the block), block k's sum by 8 s (nb - 1 - k), and everything summed.
// The rotations of all blocks go as one batched key switch, the block sums'
// as a second; rotations by a multiple of N/2 are skipped, and so are giant
// steps with no plaintext (they are zero).  A degree-2 input is rescaled
// first; out keeps noise level 2 and may alias a.  P*Q diagonals contribute
// their Q limbs.
void eval_convolution_transform(Context& c, const Ciphertext& a_in, const std::vector<const Plaintext*>& pts,
                                int bstep, int stride, const std::vector<int>& idx, int gstep, Ciphertext& out) {
    if (a_in.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (a_in.ext || out.ext) throw Error(E_STATE, "ciphertext in P*Q form");
    const int n = (int)pts.size();
    if (bstep < 1 || gstep < 1 || n < 1 || (int64_t)n > (int64_t)bstep * gstep)
        throw Error(E_ARG, "need 1 <= number of plaintexts <= bstep * gstep");
    if ((int)idx.size() != bstep) throw Error(E_ARG, "need one baby-step index per baby step");
    std::unique_ptr<Ciphertext> tmp;
    const Ciphertext* a = &a_in;
    if (a_in.noise_level == 2) {
        tmp = ct_create(c);
        ct_copy(c, a_in, *tmp);
        rescale(c, *tmp);
        a = tmp.get();
    }
    if (a->noise_level != 1) throw Error(E_STATE, "convolution transform needs a noise level 1 input");
    const int level = a->level, nl = level + 1, half = c.N / 2;
    for (const Plaintext* p : pts) {
        if (!p || p->level < level) throw Error(E_LEVEL, "plaintext missing or with fewer limbs than the input");
        if (p->noise_level != 1) throw Error(E_STATE, "plaintexts must have noise level 1");
    }
    auto zero_rot = [&](int64_t r) { return r % half == 0; };
    {  // every key up front: E_NOKEY before any device work
        std::vector<int> need(idx.begin(), idx.end());
        for (int r : convolution_rotations(stride, gstep)) need.push_back(r);
        for (int r : need)
            if (!zero_rot(r) && !c.rot_keys.count(c.rot_index(r)))
                throw Error(E_NOKEY, "no rotation key for index " + std::to_string(r));
    }

    std::vector<std::unique_ptr<Ciphertext>> baby;
    std::vector<Ciphertext*> bp;
    for (int i = 0; i < bstep; ++i) {
        baby.push_back(ct_create(c));
        bp.push_back(baby.back().get());
    }
    eval_rotate_hoisted(c, *a, idx, bp);

    // inner sums of the giant steps that hold a plaintext, in one launch
    const int gl = (n + bstep - 1) / bstep;
    std::vector<std::unique_ptr<Ciphertext>> inner;
    for (int j = 0; j < gl; ++j) inner.push_back(ct_create(c));
    std::vector<DevBuf> keep;
    bsgs_inner_sums(
        c, 1, bstep, pts, nl, nl, [&](int, int t, int h) { return h ? baby[t]->c1() : baby[t]->c0(); },
        [&](int, int j, int h) { return h ? inner[j]->c1() : inner[j]->c0(); }, keep);
    for (int j = 0; j < gl; ++j) {
        Ciphertext& o = *inner[j];
        o.level = level;
        o.noise = a->noise * pts[bstep * j]->noise;
        o.noise_level = 2;
        o.slots = a->slots;
    }

    // rotations inside the blocks: one batched key switch over all of them
    const int nb = (gstep + CONV_BLOCK - 1) / CONV_BLOCK;
    std::vector<Ciphertext*> u(gl);
    for (int j = 0; j < gl; ++j) u[j] = inner[j].get();
    {
        std::vector<std::unique_ptr<Ciphertext>> rots;
        std::vector<const Ciphertext*> ins;
        std::vector<Ciphertext*> routs;
        std::vector<int> ri;
        std::vector<int> at;
        for (int j = 0; j < gl; ++j) {
            const int blk = j / CONV_BLOCK, G = std::min(CONV_BLOCK, gstep - blk * CONV_BLOCK);
            const int r = stride * (G - (j - blk * CONV_BLOCK));
            if (zero_rot(r)) continue;
            rots.push_back(ct_create(c));
            ins.push_back(inner[j].get());
            routs.push_back(rots.back().get());
            ri.push_back(r);
            at.push_back(j);
        }
        if (!ins.empty()) eval_rotate_batch(c, ins, ri, routs);
        for (size_t m = 0; m < at.size(); ++m) u[at[m]] = rots[m].get();
        // block sums into the block's first inner-sum buffer
        std::vector<Ciphertext*> blk(nb, nullptr);
        for (int j = 0; j < gl; ++j) {
            const int b = j / CONV_BLOCK;
            if (!blk[b]) {
                blk[b] = inner[j].get();
                if (u[j] != blk[b]) ct_copy(c, *u[j], *blk[b]);
            } else {
                eval_add(c, *blk[b], *u[j], *blk[b]);
            }
        }
        // rotations of the block sums by 8 s (nb - 1 - k), one batched key switch
        std::vector<std::unique_ptr<Ciphertext>> brot;
        std::vector<const Ciphertext*> bin;
        std::vector<Ciphertext*> bout;
        std::vector<int> bri;
        std::vector<Ciphertext*> part;
        for (int k = 0; k < nb; ++k) {
            if (!blk[k]) continue;  // a block with no plaintext is zero
            const int64_t r = (int64_t)CONV_BLOCK * stride * (nb - 1 - k);
            if (zero_rot(r)) {
                part.push_back(blk[k]);
                continue;
            }
            brot.push_back(ct_create(c));
            bin.push_back(blk[k]);
            bout.push_back(brot.back().get());
            bri.push_back((int)r);
            part.push_back(brot.back().get());
        }
        if (!bin.empty()) eval_rotate_batch(c, bin, bri, bout);
        Ciphertext* acc = part[0];
        for (size_t k = 1; k < part.size(); ++k) eval_add(c, *acc, *part[k], *acc);
        ct_copy(c, *acc, out);
    }
}

}  // namespace fides
