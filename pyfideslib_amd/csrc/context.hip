// Context construction: parameter checks, host precomputation of every
// constant table the kernels read, one device allocation for all of them.
//
// Tables (FIDESlib computes them in src/ConstantsGPU.cu:127-708 from the
// OpenFHE values GetRawParams extracts; here they are derived from the moduli
// and roots alone, with exact 128-bit host arithmetic):
//   twiddles psi^brev(k), psi^-brev(k) with Shoup companions  (:224-323)
//   PartQlHatInvModq / PartQlHatModp  per (digit, present size)  (:546-644)
//   PHatInvModp, PHatModq, PInvModq                            (:482-544)
//   q_l^-1 mod q_i and q_l mod q_i for rescale                 (:426-478)
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "context.hpp"

namespace fides {

typedef unsigned __int128 u128;

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }
static uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    for (; e; e >>= 1) {
        if (e & 1) r = h_mulmod(r, a, p);
        a = h_mulmod(a, a, p);
    }
    return r;
}
static uint64_t h_invmod(uint64_t a, uint64_t p) {
    __int128 t = 0, nt = 1, r = p, nr = a % p;
    while (nr) {
        __int128 qq = r / nr, x;
        x = t - qq * nt; t = nt; nt = x;
        x = r - qq * nr; r = nr; nr = x;
    }
    if (r != 1) throw Error(E_ARG, "value not invertible modulo a prime (duplicate moduli?)");
    return (uint64_t)(t < 0 ? t + p : t);
}
static uint64_t h_shoup(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }
static ShoupConst sc(uint64_t w, uint64_t p) { return ShoupConst{w, h_shoup(w, p)}; }
static uint32_t h_brev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i, x >>= 1) r = (r << 1) | (x & 1);
    return r;
}

// ---------------------------------------------------------------- Pool ----
Pool::~Pool() {
    for (auto& kv : live_) (void)hipFree(kv.first);
}
size_t Pool::size_class(size_t b) {
    if (b == 0) b = 256;
    if (b <= ((size_t)1 << 20)) return (b + 255) & ~(size_t)255;
    const int lg = 63 - __builtin_clzll(b);
    const size_t step = (size_t)1 << (lg - 3);
    return (b + step - 1) & ~(step - 1);
}
void* Pool::get(size_t bytes) {
    bytes = size_class(bytes);
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = free_.find(bytes);
        if (it != free_.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            cached_ -= bytes;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim(0);
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    std::lock_guard<std::mutex> g(mu_);
    live_[p] = bytes;
    held_ += bytes;
    return p;
}
void Pool::put(void* p, size_t bytes) {
    bytes = size_class(bytes);
    std::lock_guard<std::mutex> g(mu_);
    free_[bytes].push_back(p);
    cached_ += bytes;
}
void Pool::trim(size_t keep_bytes) {
    std::lock_guard<std::mutex> g(mu_);
    if (cached_ <= keep_bytes) return;
    // a cached block may still be read by queued kernels (released blocks
    // are reused in stream order, never freed early)
    (void)hipDeviceSynchronize();
    // largest classes first
    std::vector<size_t> sizes;
    for (auto& kv : free_)
        if (!kv.second.empty()) sizes.push_back(kv.first);
    std::sort(sizes.rbegin(), sizes.rend());
    for (size_t sz : sizes) {
        auto& v = free_[sz];
        while (!v.empty() && cached_ > keep_bytes) {
            void* p = v.back();
            v.pop_back();
            (void)hipFree(p);
            live_.erase(p);
            cached_ -= sz;
            held_ -= sz;
        }
        if (cached_ <= keep_bytes) break;
    }
}

uint64_t* Ciphertext::c1() const { return data.ptr + (part ? part : (size_t)(ctx->L + 1) * ctx->N); }

// ------------------------------------------------------------- Context ----
Context::Context(const Params& p, int dev) {
    params = p;
    logN = p.logN;
    N = 1 << logN;
    L = p.L;
    K = p.K;
    dnum = p.dnum;
    device = dev;
    scaling_technique = p.scaling_technique;
    if (logN < 10 || logN > 17) throw Error(E_ARG, "logN must be in [10, 17]");
    if (L < 0 || K < 1 || dnum < 1 || dnum > L + 1) throw Error(E_ARG, "invalid L / K / dnum");
    np = L + 1 + K;
    if ((int)p.moduli.size() != np || (int)p.roots.size() != np) throw Error(E_ARG, "moduli/roots size mismatch");
    alpha = (L + 1 + dnum - 1) / dnum;  // OpenFHE numPerPartQ
    if (alpha > 64 || K > 32 || L + 1 + K > 96)
        throw Error(E_ARG, "unsupported sizes: digit above 64 primes, more than 32 special primes or L+1+K > 96");
    q = p.moduli;
    for (int i = 0; i < np; ++i) {
        if (q[i] < 3 || q[i] >= (1ull << 60)) throw Error(E_ARG, "moduli must be below 2^60");
        if ((q[i] - 1) % (2ull * N) != 0) throw Error(E_ARG, "modulus not 1 mod 2N");
        if (h_powmod(p.roots[i], N, q[i]) != q[i] - 1) throw Error(E_ARG, "root is not a primitive 2N-th root");
        for (int j = 0; j < i; ++j)
            if (q[j] == q[i]) throw Error(E_ARG, "duplicate modulus");
    }
    scaling_factor_real = p.scaling_factor_real;
    if (!scaling_factor_real.empty() && (int)scaling_factor_real.size() != L + 1)
        throw Error(E_ARG, "scaling_factor_real must have L+1 entries");
    scaling_factor_real_big = p.scaling_factor_real_big;
    if (!scaling_factor_real_big.empty() && (int)scaling_factor_real_big.size() != L + 1)
        throw Error(E_ARG, "scaling_factor_real_big must have L+1 entries");
    mod_reduce_factor = p.mod_reduce_factor;
    if (mod_reduce_factor.empty())
        for (int i = 0; i <= L; ++i) mod_reduce_factor.push_back((double)q[i]);
    if (const char* g = std::getenv("FIDES_GRAPHS")) use_graphs = std::atoi(g) != 0;
    if (const char* g = std::getenv("FIDES_OP_GRAPHS")) op_graphs = std::atoi(g) != 0;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    build_tables(p.roots);
    noalloc_init(device);
    tune_load_default();
}

Context::~Context() {
    if (stream) (void)hipStreamSynchronize(stream);
    reap(true);
    if (stream) (void)hipStreamDestroy(stream);
    for (int i = 0; i < STAGE_SLOTS; ++i) {
        if (stage_[i]) (void)hipHostFree(stage_[i]);
        if (stage_ev_[i]) (void)hipEventDestroy(stage_ev_[i]);
    }
    arena_.release();
    plan_cache.reset();
    hoist_plan.reset();
    graph_cache.clear();
    ks_workspace.reset();
    eval_key.reset();
    rot_keys.clear();
    if (blob_) (void)hipFree(blob_);
}

void Context::sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

void Context::retire(std::vector<std::shared_ptr<void>> objs) {
    reap(false);
    size_t live = 0;
    for (auto& o : objs) live += o != nullptr;
    if (live == 0) return;
    hipEvent_t ev = nullptr;
    hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    const hipError_t e = hipEventRecord(ev, stream);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        hip_check(e, "hipEventRecord");
    }
    retired_.emplace_back(ev, std::move(objs));
}

void Context::reap(bool wait) {
    size_t done = 0;
    for (; done < retired_.size(); ++done) {
        hipEvent_t ev = retired_[done].first;
        if (wait) {
            (void)hipEventSynchronize(ev);
        } else if (hipEventQuery(ev) != hipSuccess) {
            (void)hipGetLastError();  // "not ready" must not reach a later launch check
            break;                    // retired in stream order: the rest are later
        }
        (void)hipEventDestroy(ev);
        retired_[done].second.clear();
    }
    retired_.erase(retired_.begin(), retired_.begin() + done);
}

void Context::drop_plans() {
    std::vector<std::shared_ptr<void>> objs;
    for (auto& kv : graph_cache) objs.push_back(kv.second);
    graph_cache.clear();
    graph_seen.clear();
    objs.push_back(plan_cache);
    objs.push_back(hoist_plan);
    plan_cache.reset();
    hoist_plan.reset();
    retire(std::move(objs));
}

void Context::upload(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    const int i = stage_next_;
    stage_next_ = (stage_next_ + 1) % STAGE_SLOTS;
    if (stage_used_[i]) hip_check(hipEventSynchronize(stage_ev_[i]), "staging wait");
    if (stage_bytes_[i] < bytes) {
        if (stage_[i]) hip_check(hipHostFree(stage_[i]), "hipHostFree");
        stage_[i] = nullptr;
        stage_bytes_[i] = 0;
        // grow by powers of two from 1 MiB: freeing pinned memory synchronizes
        // the device, so a slot should rarely have to grow
        size_t want = (size_t)1 << 20;
        while (want < bytes) want <<= 1;
        hip_check(hipHostMalloc(&stage_[i], want, hipHostMallocDefault), "hipHostMalloc");
        stage_bytes_[i] = want;
    }
    if (!stage_ev_[i]) hip_check(hipEventCreateWithFlags(&stage_ev_[i], hipEventDisableTiming), "hipEventCreate");
    std::memcpy(stage_[i], src, bytes);
    hip_check(hipMemcpyAsync(dst, stage_[i], bytes, hipMemcpyHostToDevice, stream), "staged upload");
    hip_check(hipEventRecord(stage_ev_[i], stream), "hipEventRecord");
    stage_used_[i] = true;
}

uint64_t* Context::batch_arena(size_t words) {
    if (arena_.bytes < words * 8) {
        arena_.release();
        arena_ = DevBuf(&pool, words + words / 8);  // headroom for the next, slightly larger batch
    }
    return arena_.ptr;
}

uint32_t Context::rot_index(int r) const {
    const int64_t half = N / 2;
    int64_t rr = r % half;
    if (rr < 0) rr += half;
    return (uint32_t)h_powmod(5, (uint64_t)rr, 2ull * N);
}

uint32_t Context::auto_inverse(uint32_t k) const {
    if (!(k & 1)) throw Error(E_ARG, "automorphism index must be odd");
    uint32_t x = k;  // Newton: each step doubles the correct low bits (k x = 1 mod 8 at the start)
    for (int i = 0; i < 5; ++i) x *= 2u - k * x;
    return x & (uint32_t)(2 * N - 1);
}

void Context::build_tables(const std::vector<uint64_t>& roots) {
    const int A = alpha;
    // host staging, in the order of the blob
    std::vector<PrimeInfo> primes(np);
    std::vector<ShoupConst> tw_fwd((size_t)np * N), tw_inv((size_t)np * N);
    std::vector<ShoupConst> is_plain((size_t)np * 2), is_modup((size_t)(L + 1) * (L + 1) * 2), is_moddown((size_t)K * 2);
    const int AP = conv_bound(A), KP = conv_bound(K);
    std::vector<uint64_t> modup_mat((size_t)dnum * A * np * AP, 0), moddown_mat((size_t)np * KP, 0);
    std::vector<ShoupConst> pinv(np), pmodq(np), qlinv((size_t)(L + 1) * (L + 1));
    std::vector<uint64_t> qlmodq((size_t)(L + 1) * (L + 1), 0);

    for (int i = 0; i < np; ++i) {
        const uint64_t p = q[i];
        int n = 0;
        while ((p >> n) != 0) ++n;
        primes[i] = PrimeInfo{p, (uint64_t)(((u128)1 << (n + 63)) / p), (uint32_t)n, 0u, 2 * p};
        const uint64_t psi = roots[i], psi_inv = h_invmod(psi, p);
        uint64_t pw = 1, pwi = 1;
        for (int k = 0; k < N; ++k) {
            const uint32_t r = h_brev(k, logN);
            tw_fwd[(size_t)i * N + r] = sc(pw, p);
            tw_inv[(size_t)i * N + r] = sc(pwi, p);
            pw = h_mulmod(pw, psi, p);
            pwi = h_mulmod(pwi, psi_inv, p);
        }
        const uint64_t ninv = h_invmod(N % p, p);
        const uint64_t w1 = tw_inv[(size_t)i * N + 1].w;
        is_plain[i * 2] = sc(ninv, p);
        is_plain[i * 2 + 1] = sc(h_mulmod(w1, ninv, p), p);
    }
    // ModUp: digit d, present size s, element i.
    for (int d = 0; d < dnum; ++d) {
        const int start = d * A;
        for (int s = 1; s <= A && start + s <= L + 1; ++s) {
            for (int i = 0; i < s; ++i) {
                for (int t = 0; t < np; ++t) {
                    uint64_t prod = 1;
                    for (int j = 0; j < s; ++j)
                        if (j != i) prod = h_mulmod(prod, q[start + j] % q[t], q[t]);
                    modup_mat[(((size_t)d * A + (s - 1)) * np + t) * AP + i] = prod;
                }
            }
        }
    }
    for (int l = 0; l <= L; ++l) {
        for (int pos = 0; pos <= l; ++pos) {
            const int d = pos / A, i = pos - d * A;
            const int s = digit_size(d, l);
            const uint64_t p = q[pos];
            uint64_t inv = 1;
            for (int j = 0; j < s; ++j)
                if (j != i) inv = h_mulmod(inv, h_invmod(q[d * A + j] % p, p), p);
            const uint64_t scale = h_mulmod(inv, is_plain[pos * 2].w, p);
            const uint64_t w1 = tw_inv[(size_t)pos * N + 1].w;
            is_modup[((size_t)l * (L + 1) + pos) * 2] = sc(scale, p);
            is_modup[((size_t)l * (L + 1) + pos) * 2 + 1] = sc(h_mulmod(w1, scale, p), p);
        }
    }
    // ModDown.
    for (int k = 0; k < K; ++k) {
        const int pid = L + 1 + k;
        const uint64_t pk = q[pid];
        uint64_t inv = 1;
        for (int j = 0; j < K; ++j)
            if (j != k) inv = h_mulmod(inv, h_invmod(q[L + 1 + j] % pk, pk), pk);
        const uint64_t scale = h_mulmod(inv, is_plain[pid * 2].w, pk);
        const uint64_t w1 = tw_inv[(size_t)pid * N + 1].w;
        is_moddown[k * 2] = sc(scale, pk);
        is_moddown[k * 2 + 1] = sc(h_mulmod(w1, scale, pk), pk);
        for (int t = 0; t < np; ++t) {
            uint64_t prod = 1;
            for (int j = 0; j < K; ++j)
                if (j != k) prod = h_mulmod(prod, q[L + 1 + j] % q[t], q[t]);
            moddown_mat[(size_t)t * KP + k] = prod;
        }
    }
    for (int t = 0; t < np; ++t) {
        uint64_t prod = 1;
        for (int j = 0; j < K; ++j) prod = h_mulmod(prod, q[L + 1 + j] % q[t], q[t]);
        pinv[t] = t <= L ? sc(h_invmod(prod, q[t]), q[t]) : ShoupConst{0, 0};
        pmodq[t] = t <= L ? sc(prod, q[t]) : ShoupConst{0, 0};
    }
    // Rescale (q_l^-1 for i < l) and the centered lift of a q_l limb into every
    // prime (q_l mod q_i for all i: rescale's i < l, ModRaise's l = 0).
    for (int l = 0; l <= L; ++l)
        for (int i = 0; i <= L; ++i) {
            if (i < l) qlinv[(size_t)l * (L + 1) + i] = sc(h_invmod(q[l] % q[i], q[i]), q[i]);
            qlmodq[(size_t)l * (L + 1) + i] = q[l] % q[i];
        }

    // Bound checks for single-reduction base conversion (see DevTables).
    auto bits = [](uint64_t v) { int n = 0; while ((v >> n) != 0) ++n; return n; };
    auto fits = [&](const std::vector<int>& src, int tgt) {
        u128 sum = 0;
        for (int i : src) sum += (u128)(q[i] - 1) * (q[tgt] - 1);
        const int n = bits(q[tgt]);
        return n + 63 >= 128 || sum < ((u128)1 << (n + 63));
    };
    T.modup_one_reduce = 1;
    for (int d = 0; d < dnum; ++d) {
        std::vector<int> src;
        for (int i = d * A; i < (d + 1) * A && i <= L; ++i) src.push_back(i);
        for (int t = 0; t < np; ++t)
            if (!fits(src, t)) T.modup_one_reduce = 0;
    }
    {
        std::vector<int> src;
        for (int k = 0; k < K; ++k) src.push_back(L + 1 + k);
        T.moddown_one_reduce = 1;
        for (int t = 0; t <= L; ++t)
            if (!fits(src, t)) T.moddown_one_reduce = 0;
    }

    // Matrix-core ModUp tables: residues as 8 balanced signed bytes
    // (x = sum_a d_a 2^(8a), the bytes of x + 0x80..80 each less 128), and
    // for each c = a + b the A-operand word of (target, limb): byte a holds
    // digit c - a of the matrix entry (0 outside 0..7).
    const int MS = (A + 3) / 4, SP = 4 * MS;
    std::vector<uint64_t> modup_mfma;
    if (T.modup_one_reduce && A <= 16) {
        auto bal = [](uint64_t x) { const uint64_t C = 0x8080808080808080ull; return (x + C) ^ C; };
        modup_mfma.assign((size_t)dnum * A * 15 * np * SP, 0);
        for (int d = 0; d < dnum; ++d)
            for (int s = 1; s <= A && d * A + s <= L + 1; ++s)
                for (int t = 0; t < np; ++t)
                    for (int i = 0; i < s; ++i) {
                        const uint64_t mb = bal(modup_mat[(((size_t)d * A + (s - 1)) * np + t) * AP + i]);
                        for (int c = 0; c < 15; ++c) {
                            uint64_t w = 0;
                            for (int a = 0; a < 8; ++a) {
                                const int b = c - a;
                                if (b >= 0 && b < 8) w |= ((mb >> (8 * b)) & 0xff) << (8 * a);
                            }
                            modup_mfma[((((size_t)d * A + (s - 1)) * 15 + c) * np + t) * SP + i] = w;
                        }
                    }
    }

    // ---- one device blob ----
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_primes = place(primes.size() * sizeof(PrimeInfo));
    const size_t o_twf = place(tw_fwd.size() * sizeof(ShoupConst));
    const size_t o_twi = place(tw_inv.size() * sizeof(ShoupConst));
    const size_t o_isp = place(is_plain.size() * sizeof(ShoupConst));
    const size_t o_ism = place(is_modup.size() * sizeof(ShoupConst));
    const size_t o_isd = place(is_moddown.size() * sizeof(ShoupConst));
    const size_t o_mum = place(modup_mat.size() * 8);
    const size_t o_mdm = place(moddown_mat.size() * 8);
    const size_t o_pinv = place(pinv.size() * sizeof(ShoupConst));
    const size_t o_pmq = place(pmodq.size() * sizeof(ShoupConst));
    const size_t o_qli = place(qlinv.size() * sizeof(ShoupConst));
    const size_t o_qlm = place(qlmodq.size() * 8);
    const size_t o_mfu = place(modup_mfma.size() * 8);
    std::vector<char> host(off, 0);
    auto put = [&](size_t o, const void* src, size_t bytes) { if (bytes) std::memcpy(host.data() + o, src, bytes); };
    put(o_primes, primes.data(), primes.size() * sizeof(PrimeInfo));
    put(o_twf, tw_fwd.data(), tw_fwd.size() * sizeof(ShoupConst));
    put(o_twi, tw_inv.data(), tw_inv.size() * sizeof(ShoupConst));
    put(o_isp, is_plain.data(), is_plain.size() * sizeof(ShoupConst));
    put(o_ism, is_modup.data(), is_modup.size() * sizeof(ShoupConst));
    put(o_isd, is_moddown.data(), is_moddown.size() * sizeof(ShoupConst));
    put(o_mum, modup_mat.data(), modup_mat.size() * 8);
    put(o_mdm, moddown_mat.data(), moddown_mat.size() * 8);
    put(o_pinv, pinv.data(), pinv.size() * sizeof(ShoupConst));
    put(o_pmq, pmodq.data(), pmodq.size() * sizeof(ShoupConst));
    put(o_qli, qlinv.data(), qlinv.size() * sizeof(ShoupConst));
    put(o_qlm, qlmodq.data(), qlmodq.size() * 8);
    put(o_mfu, modup_mfma.data(), modup_mfma.size() * 8);
    hip_check(hipMalloc(&blob_, off), "hipMalloc(tables)");
    hip_check(hipMemcpy(blob_, host.data(), off, hipMemcpyHostToDevice), "hipMemcpy(tables)");
    char* b = (char*)blob_;
    T.logN = logN; T.N = N; T.L = L; T.K = K; T.dnum = dnum; T.alpha = alpha; T.np = np;
    T.primes = (const PrimeInfo*)(b + o_primes);
    T.tw_fwd = (const ShoupConst*)(b + o_twf);
    T.tw_inv = (const ShoupConst*)(b + o_twi);
    T.iscale_plain = (const ShoupConst*)(b + o_isp);
    T.iscale_modup = (const ShoupConst*)(b + o_ism);
    T.iscale_moddown = (const ShoupConst*)(b + o_isd);
    T.modup_mat = (const uint64_t*)(b + o_mum);
    T.moddown_mat = (const uint64_t*)(b + o_mdm);
    T.pinv = (const ShoupConst*)(b + o_pinv);
    T.pmodq = (const ShoupConst*)(b + o_pmq);
    T.qlinv = (const ShoupConst*)(b + o_qli);
    T.qlmodq = (const uint64_t*)(b + o_qlm);
    T.modup_mfma = modup_mfma.empty() ? nullptr : (const uint64_t*)(b + o_mfu);
    T.modup_mfma_ms = modup_mfma.empty() ? 0 : MS;
}

}  // namespace fides
