// CKKS bootstrapping of fully packed ciphertexts (N/2 slots) -- FIDESlib
// CKKS::Bootstrap (src/CKKS/Bootstrap.cu:169-300) and its setup
// (EvalBootstrapSetup, api/CryptoContext.hpp:80; BootstrapPrecomputation.cuh),
// OpenFHE EvalBootstrap:
//
//   ModRaise -> CoeffsToSlots -> conjugate, real / imaginary split ->
//   EvalMod on both halves -> recombine -> SlotsToCoeffs
//
// Slot j of a polynomial m is m(zeta^(5^j)), zeta = exp(i pi / N).  With
// n = N/2 and w_k = m_k + i m_(k+n) this is z = U0 w, U0[j, k] = zeta^(5^j k),
// and U0 = B_n ... B_2 P: P the bit-reversal permutation, B_len a radix-2
// butterfly stage (pairs at distance h = len/2 inside blocks of len, twiddle
// exp(2 pi i (5^j mod 4 len) / (4 len)) for pair j).  A stage has three
// non-zero diagonals (offsets 0, +h, -h), so s consecutive stages have at most
// 2^(s+1) - 1: OpenFHE's level-budget decomposition.  CoeffsToSlots applies
// B_2^-1 ... B_n^-1 (its slots come out in bit-reversed coefficient order,
// which the slotwise EvalMod ignores), SlotsToCoeffs B_n ... B_2; each group
// of stages is one baby-step / giant-step linear transform.
//
// Scales: the raised ciphertext holds c = S0 m + q0 I.  CoeffsToSlots
// carries S0 / (2 K q0), spread evenly over its groups, so the two halves hold
// c / (K q0) in [-1, 1]; EvalMod gives sin(2 pi c / q0) / (2 pi) ~ S0 m / q0
// and multiplies by 2^deg, deg = round(log2(q0 / S0)); SlotsToCoeffs carries
// q0 / (S0 2^deg).  The reference spends a level on a scalar product by
// pre / (K N) (Bootstrap.cu:219-231); here that factor sits in the
// CoeffsToSlots diagonals.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <set>

#include "context.hpp"

namespace fides {

using cd = std::complex<double>;
using Diags = std::map<int, std::vector<cd>>;

namespace {

const double PI = 3.14159265358979323846;

struct Stage {
    int h;
    std::vector<cd> tw;
};

// B_2 .. B_n: h = 1, 2, .., n/2, tw[j] = exp(2 pi i (5^j mod 8h) / (8h))
std::vector<Stage> stages(int n) {
    std::vector<Stage> out;
    for (int h = 1; h < n; h *= 2) {
        const int64_t m = 8 * (int64_t)h;
        Stage s{h, std::vector<cd>(h)};
        int64_t e = 1;
        for (int j = 0; j < h; ++j) {
            s.tw[j] = std::polar(1.0, 2.0 * PI * (double)e / (double)m);
            e = e * 5 % m;
        }
        out.push_back(std::move(s));
    }
    return out;
}

int bitrev(int x, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}

// w = U0^-1 z in place: the inverse stages from B_n down, then P
void special_fft_inverse(std::vector<cd>& v) {
    const int n = (int)v.size();
    const std::vector<Stage> st = stages(n);
    for (auto it = st.rbegin(); it != st.rend(); ++it) {
        const int h = it->h;
        for (int blk = 0; blk < n; blk += 2 * h)
            for (int j = 0; j < h; ++j) {
                const cd a = v[blk + j], b = v[blk + j + h];
                v[blk + j] = (a + b) * 0.5;
                v[blk + j + h] = (a - b) * 0.5 / it->tw[j];
            }
    }
    int bits = 0;
    while ((1 << bits) < n) ++bits;
    for (int i = 0; i < n; ++i) {
        const int r = bitrev(i, bits);
        if (r > i) std::swap(v[i], v[r]);
    }
}

// one stage as {offset d: diag}, out[p] = sum_d diag[p] in[(p + d) mod n]
Diags stage_diags(int n, const Stage& s, bool inverse) {
    const int h = s.h;
    std::vector<cd> d0(n), dp(n), dm(n);
    for (int p = 0; p < n; ++p) {
        const int r = p % (2 * h);
        const bool low = r < h;
        const cd t = s.tw[low ? r : r - h];
        if (low) {
            d0[p] = inverse ? 0.5 : 1.0;
            dp[p] = inverse ? cd(0.5) : t;
        } else {
            d0[p] = inverse ? -0.5 / t : -t;
            dm[p] = inverse ? 0.5 / t : cd(1.0);
        }
    }
    Diags out;
    out[0] = d0;
    auto addto = [&](int d, const std::vector<cd>& v) {
        auto& o = out[d];
        if (o.empty()) o.assign(n, 0.0);
        for (int p = 0; p < n; ++p) o[p] += v[p];
    };
    addto(h % n, dp);
    addto((n - h) % n, dm);
    return out;
}

// A B: (A B)_d[p] = sum_(a + b = d) A_a[p] B_b[p + a]
Diags compose(const Diags& A, const Diags& B, int n) {
    Diags out;
    for (const auto& [a, va] : A)
        for (const auto& [b, vb] : B) {
            auto& o = out[(a + b) % n];
            if (o.empty()) o.assign(n, 0.0);
            for (int p = 0; p < n; ++p) o[p] += va[p] * vb[(p + a) % n];
        }
    return out;
}

struct StageGroup {
    int stride, s;
    Diags D;
};

// the stage groups in the order they are applied
std::vector<StageGroup> groups(int n, int budget, bool inverse) {
    std::vector<Stage> st = stages(n);
    if (inverse) std::reverse(st.begin(), st.end());
    const int logn = (int)st.size();
    budget = std::max(1, std::min(budget, logn));
    std::vector<int> sizes;
    for (int i = 0; i < budget; ++i) sizes.push_back(logn / budget + (i < logn % budget ? 1 : 0));
    if (!inverse) std::reverse(sizes.begin(), sizes.end());
    std::vector<StageGroup> out;
    int at = 0;
    for (int s : sizes) {
        StageGroup g{n, s, {}};
        for (int k = at; k < at + s; ++k) {
            Diags D = stage_diags(n, st[k], inverse);
            g.D = g.D.empty() ? std::move(D) : compose(D, g.D, n);
            g.stride = std::min(g.stride, st[k].h);
        }
        out.push_back(std::move(g));
        at += s;
    }
    return out;
}

// Chebyshev interpolant (sum c_k T_k, c_0 not halved) on [-1, 1] of
// (2 pi)^(-1/2^r) cos(2 pi (K x - 1/4) / 2^r), at the degree+1 Chebyshev
// points of the first kind
std::vector<double> evalmod_coefficients(double K, int r, int degree) {
    const int n = degree + 1;
    const double a = std::pow(2.0 * PI, -1.0 / (double)(1 << r));
    std::vector<double> f(n), c(n, 0.0);
    for (int j = 0; j < n; ++j) {
        const double x = std::cos(PI * (j + 0.5) / n);
        f[j] = a * std::cos(2.0 * PI * (K * x - 0.25) / (double)(1 << r));
    }
    for (int k = 0; k < n; ++k) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += f[j] * std::cos(PI * k * (j + 0.5) / n);
        c[k] = s * (k == 0 ? 1.0 : 2.0) / n;
    }
    return c;
}

int64_t mod_n(int64_t x, int64_t n) { return ((x % n) + n) % n; }

}  // namespace

// MakeCKKSPackedPlaintext (api/CryptoContext.cpp:475-514): coefficients
// round(scale [Re w, Im w]), w = U0^-1 z, reduced mod q_0..q_level on the
// host, then the engine's forward NTT in place.  With ext the coefficients
// are also reduced mod the K special primes (a P*Q plaintext, as OpenFHE
// encodes the linear-transform diagonals in the extended basis for
// EvalFastRotationExt).
std::unique_ptr<Plaintext> encode_slots(Context& c, const std::vector<cd>& z, int level, double scale, bool ext) {
    const int n = c.N / 2;
    if ((int)z.size() != n) throw Error(E_ARG, "encode needs N/2 slot values");
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    std::vector<cd> w(z);
    special_fft_inverse(w);
    const int nl = level + 1, nr = nl + (ext ? c.K : 0);
    std::vector<uint64_t> res((size_t)nr * c.N);
    for (int k = 0; k < (int)c.N; ++k) {
        const double m = (k < n ? w[k].real() : w[k - n].imag()) * scale;
        if (!(std::fabs(m) < 0x1p62)) throw Error(E_ARG, "encoded coefficient exceeds 62 bits");
        const int64_t v = std::llround(m);
        for (int i = 0; i < nr; ++i) {
            const int64_t q = (int64_t)c.q[i < nl ? i : c.L + 1 + (i - nl)];
            res[(size_t)i * c.N + k] = (uint64_t)(((v % q) + q) % q);
        }
    }
    std::unique_ptr<Plaintext> pt = pt_create(c);
    if (ext) {
        pt_load_ext(c, *pt, res.data(), nl, scale, 1, n, false);
        poly_ntt(c, pt->c0() + (size_t)nl * c.N, c.K, c.L + 1, false);
    } else {
        pt_load(c, *pt, res.data(), nl, scale, 1, n, false);
    }
    poly_ntt(c, pt->c0(), nl, 0, false);
    return pt;
}

struct BootGroup {
    int bstep = 1, stride = 1, offset = 0;
    std::vector<std::unique_ptr<Plaintext>> diag;
};

struct BootPlan {
    bool sse = false;           // sparse-secret encapsulation around ModRaise
    bool batched_giant = false;  // giant steps of each transform as one batched rotation
    bool ext = false;            // P*Q diagonals
    int slots = 0;     // N/2, or fewer for sparse packing
    int top = 0, r = 0;
    double S0 = 0, K = 0, fb = 1;
    uint64_t post = 1;
    std::vector<double> coeffs;
    std::vector<BootGroup> cts, stc;
    std::vector<int> rotations;
    // FIXEDAUTO / FLEXIBLEAUTO: the SlotsToCoeffs groups are encoded on first
    // use, each at ScalingFactorReal of the level it runs at (a product
    // rescaled from level l then lands on level l-1's factor, as FLEXIBLEAUTO
    // requires); stc_level is the entry level they were encoded for.  Calls
    // hold the context's lock, so the cache is not raced.
    std::vector<StageGroup> stc_groups;
    mutable int stc_level = -1;
    mutable std::vector<BootGroup> stc_auto;
};

void BootPlanDeleter::operator()(BootPlan* p) const { delete p; }

// One stage group as a linear transform over n = N/2 slots: offsets
// (lo + t) stride for t < cnt, baby step b (the least power of two with
// b^2 >= cnt), giant step b stride; term b j + i carries its diagonal rotated
// right by lo stride + j b stride.  With ns < n slots the group's diagonals
// (period ns) are tiled over n; `left` multiplies the output slots by
// (1 | -i) and `right` the input slots by (1 | i) in alternating blocks of ns
// (the real / imaginary split of sparse packing), the input taken at the
// offset actually rotated by.
static BootGroup lay_out(Context& c, const StageGroup& g, int ns, cd factor, bool left, bool right, bool giant_all,
                         int top, double scale, std::set<int>& rots, bool ext, bool encode = true) {
    const int n = c.N / 2;
    int lo = -((1 << g.s) - 1), cnt = (1 << (g.s + 1)) - 1;
    if (cnt > ns / g.stride) {
        lo = 0;
        cnt = ns / g.stride;
    }
    int b = 1;
    while (b * b < cnt) b <<= 1;
    b = std::min(b, cnt);
    BootGroup out;
    out.bstep = b;
    out.stride = g.stride;
    out.offset = (int)mod_n((int64_t)lo * g.stride, n);
    auto block_hi = [&](int64_t q) { return mod_n(q, 2 * (int64_t)ns) >= ns; };
    std::vector<cd> v(n);
    for (int t = 0; t < cnt && encode; ++t) {
        const int64_t d_act = (int64_t)(lo + t) * g.stride;
        const auto it = g.D.find((int)mod_n(d_act, ns));
        const int64_t shift = (int64_t)lo * g.stride + (int64_t)(t / b) * b * g.stride;
        for (int p = 0; p < n; ++p) {
            if (it == g.D.end()) {
                v[p] = 0.0;
                continue;
            }
            const int64_t q = mod_n(p - shift, n);
            cd x = it->second[q % ns] * factor;
            if (left && block_hi(q)) x *= cd(0.0, -1.0);
            if (right && block_hi(q + d_act)) x *= cd(0.0, 1.0);
            v[p] = x;
        }
        out.diag.push_back(encode_slots(c, v, top, scale, ext && b > 1));
    }
    for (int i = 1; i < b; ++i) rots.insert((int)mod_n((int64_t)i * g.stride, n));
    if (cnt > b) rots.insert((int)mod_n((int64_t)b * g.stride, n));
    if (giant_all)  // batched giant steps rotate giant step j by j b stride, all at once
        for (int j = 2, ng = (cnt + b - 1) / b; j < ng; ++j) rots.insert((int)mod_n((int64_t)j * b * g.stride, n));
    rots.insert(out.offset);
    rots.erase(0);
    return out;
}

BootPlanPtr bootstrap_setup(Context& c, double scale, int budget_cts, int budget_stc, double K, int double_angle,
                            int degree, int top, bool sparse_encaps, int slots, bool batched_giant,
                            bool ext_diagonals) {
    if (!(scale > 0) || !(K > 0) || double_angle < 0 || double_angle > 16 || degree < 1)
        throw Error(E_ARG, "bootstrap parameters out of range");
    if (budget_cts < 1 || budget_stc < 1) throw Error(E_ARG, "level budgets must be positive");
    if (top < 0) top = c.L;
    if (top > c.L) throw Error(E_LEVEL, "top level beyond L");
    BootPlanPtr B(new BootPlan);
    const int n = c.N / 2;
    const int ns = slots <= 0 ? n : slots;
    if (ns < 2 || ns > n || (ns & (ns - 1))) throw Error(E_ARG, "slots must be a power of two in [2, N/2]");
    B->slots = ns;
    const double q0 = (double)c.q[0];
    B->sse = sparse_encaps;
    B->batched_giant = batched_giant;
    B->ext = ext_diagonals;
    B->top = top;
    B->S0 = scale;
    B->K = K;
    B->r = double_angle;
    const long deg = std::lround(std::log2(q0 / scale));
    if (deg < 0 || deg > 62) throw Error(E_ARG, "q0 / scale out of range");
    B->post = (uint64_t)1 << deg;
    B->coeffs = evalmod_coefficients(K, double_angle, degree);
    const std::vector<StageGroup> cts = groups(ns, budget_cts, true), stc = groups(ns, budget_stc, false);
    const int gap = n / ns;  // the subring trace of sparse packing multiplies by gap
    const double fa = std::pow(scale / (2.0 * K * q0 * gap), 1.0 / (double)cts.size());
    const double fb = q0 / (scale * (double)B->post);
    B->fb = fb;
    // the automatic modes encode SlotsToCoeffs at its levels' scales on first use
    const bool auto_scales = c.scaling_technique != FIXEDMANUAL;
    const bool sparse = ns < n;
    std::set<int> rots;
    for (size_t i = 0; i < cts.size(); ++i)
        B->cts.push_back(
            lay_out(c, cts[i], ns, fa, sparse && i + 1 == cts.size(), false, batched_giant, top, scale, rots,
                    ext_diagonals));
    for (size_t i = 0; i < stc.size(); ++i)
        B->stc.push_back(
            lay_out(c, stc[i], ns, i == 0 ? fb : 1.0, false, sparse && i == 0, batched_giant, top, scale, rots,
                    ext_diagonals, !auto_scales));
    if (auto_scales) B->stc_groups = stc;
    for (int r = ns; sparse && r < n; r <<= 1) rots.insert(r);  // the trace, and the final fold by ns
    B->rotations.assign(rots.begin(), rots.end());
    c.sync();
    return B;
}

const std::vector<int>& bootstrap_rotations(const BootPlan& B) { return B.rotations; }

// The SlotsToCoeffs groups for an entry at level ls, group j at
// ScalingFactorReal[ls - j] (automatic modes), encoded once per entry level.
static const std::vector<BootGroup>& stc_at_level(Context& c, const BootPlan& B, int ls) {
    if (B.stc_level == ls) return B.stc_auto;
    const int g = (int)B.stc_groups.size();
    if (ls - g < 0) throw Error(E_LEVEL, "SlotsToCoeffs has too few levels left");
    if ((int)c.scaling_factor_real.size() <= ls) throw Error(E_STATE, "bootstrapping needs ScalingFactorReal");
    const bool sparse = B.slots < (int)c.N / 2;
    std::set<int> rots;
    std::vector<BootGroup> out;
    for (int j = 0; j < g; ++j)
        out.push_back(lay_out(c, B.stc_groups[j], B.slots, j == 0 ? B.fb : 1.0, false, sparse && j == 0,
                              B.batched_giant, B.top, c.scaling_factor_real[ls - j], rots, B.ext));
    B.stc_auto = std::move(out);
    B.stc_level = ls;
    return B.stc_auto;
}


// one stage group applied to every ciphertext of xs (one batched transform)
static void run_lt(Context& c, std::vector<std::unique_ptr<Ciphertext>>& xs, const BootGroup& g,
                   bool batched_giant) {
    std::vector<const Plaintext*> d;
    for (const auto& p : g.diag) d.push_back(p.get());
    std::vector<std::unique_ptr<Ciphertext>> ys;
    std::vector<const Ciphertext*> in;
    std::vector<Ciphertext*> out;
    for (auto& x : xs) {
        ys.push_back(ct_create(c));
        in.push_back(x.get());
        out.push_back(ys.back().get());
    }
    eval_linear_transform_batch(c, in, d, g.bstep, g.stride, g.offset, out, batched_giant);
    xs = std::move(ys);
}

// EvalBootstrap (Bootstrap.cu:169-300) of k ciphertexts, full or sparse packing:
// each is raised on its own; CoeffsToSlots and SlotsToCoeffs run as batched
// linear transforms over all k (one dot launch per group, every giant step
// one batched rotation); the 2k real / imaginary halves go through EvalMod
// in lockstep (every product of the series one batched key switch).  k = 1
// is EvalBootstrap; its two halves still share the batched EvalMod (the
// reference's approxModReduction of ctxt and ctxtEncI).
void eval_bootstrap_batch(Context& c, const BootPlan& B, const std::vector<const Ciphertext*>& ins,
                          const std::vector<Ciphertext*>& outs) {
    const size_t k = ins.size();
    if (k == 0 || outs.size() != k) throw Error(E_ARG, "batch size mismatch");
    const int technique = c.scaling_technique;
    if (technique == FLEXIBLEAUTOEXT) throw Error(E_STATE, "bootstrapping under FLEXIBLEAUTOEXT is not supported");
    const bool autom = technique != FIXEDMANUAL;
    for (size_t i = 0; i < k; ++i) {
        if (!ins[i] || ins[i]->level < 0) throw Error(E_STATE, "ciphertext not loaded");
        if (autom && ins[i]->level - ins[i]->noise_level + 1 < 1)
            throw Error(E_LEVEL, "bootstrapping under FIXEDAUTO / FLEXIBLEAUTO needs a level above 0 to adjust the scale");
    }
    ManualScaling manual(c);
    std::vector<std::unique_ptr<Ciphertext>> raised(k), re(k), im(k);
    for (size_t i = 0; i < k; ++i) {
        std::unique_ptr<Ciphertext> x = ct_create(c);
        ct_copy(c, *ins[i], *x);
        if (x->noise_level == 2) rescale(c, *x);
        // automatic modes: the message onto the plan's scale S0 (the scale of
        // a ciphertext differs per level under FLEXIBLEAUTO): times
        // round(q_l S0 / scale), rescaled -- the reference's adjustment
        // product before ModRaise (Bootstrap.cu:428-446)
        if (autom) {
            mult_const_to_scale(c, *x, x->level, 1.0, B.S0, *x);
            rescale(c, *x);
        }
        if (x->level > 0) drop_to_level(c, *x, 0);
        std::unique_ptr<Ciphertext> y = ct_create(c);
        // sparse-secret encapsulation (ModRaise, Bootstrap.cu:575-584, :698-706):
        // at level 0 switch to the sparse secret, so the raise's overflow I is
        // that of a sparse key, and back to the dense secret after the raise
        if (B.sse) {
            eval_key_switch(c, *x, KEY_SSE_DENSE_TO_SPARSE, *y);
            std::swap(x, y);
        }
        mod_raise(c, *x, B.top, *y);
        if (B.sse) eval_key_switch(c, *y, KEY_SSE_SPARSE_TO_DENSE, *y);
        raised[i] = std::move(y);
    }
    const int n = (int)c.N / 2;
    const bool sparse = B.slots < n;
    // sparse packing: the trace onto the subring of the packing (Accumulate,
    // Bootstrap.cu:233): y += rot(y, slots 2^t) while slots 2^t < N/2
    std::unique_ptr<Ciphertext> tr = ct_create(c);
    for (int r = B.slots; sparse && r < n; r <<= 1)
        for (size_t i = 0; i < k; ++i) {
            eval_rotate(c, *raised[i], r, *tr);
            eval_add(c, *raised[i], *tr, *raised[i]);
        }
    for (const auto& g : B.cts) run_lt(c, raised, g, B.batched_giant);
    for (size_t i = 0; i < k; ++i) {
        std::unique_ptr<Ciphertext>& y = raised[i];
        rescale(c, *y);
        std::unique_ptr<Ciphertext> cj = ct_create(c);
        eval_automorph(c, *y, (uint32_t)(2 * c.N - 1), *cj);
        if (!sparse) {  // full packing: imaginary half (v - conj v) (-i)
            im[i] = ct_create(c);
            eval_sub(c, *y, *cj, *im[i]);
            eval_mult_monomial(c, *im[i], 3 * (int64_t)c.N / 2, *im[i]);  // times -i
        }
        eval_add(c, *y, *cj, *y);  // real part (sparse: both halves, in blocks of `slots`)
        re[i] = std::move(y);
    }
    const size_t halves = sparse ? 1 : 2;
    std::vector<std::unique_ptr<Ciphertext>> mod(halves * k);
    std::vector<const Ciphertext*> xs;
    std::vector<Ciphertext*> ms;
    for (size_t i = 0; i < k; ++i) {
        xs.push_back(re[i].get());
        if (!sparse) xs.push_back(im[i].get());
    }
    for (auto& m : mod) {
        m = ct_create(c);
        ms.push_back(m.get());
    }
    eval_mod_reduction_batch(c, xs, B.coeffs, B.r, B.post, ms);
    std::vector<std::unique_ptr<Ciphertext>> u(k);
    for (size_t i = 0; i < k; ++i) {
        if (!sparse) {
            Ciphertext& r = *mod[2 * i];
            Ciphertext& q = *mod[2 * i + 1];
            eval_mult_monomial(c, q, (int64_t)c.N / 2, q);  // times i
            eval_add(c, r, q, r);
        }
        u[i] = std::move(mod[halves * i]);
    }
    // SlotsToCoeffs (its first transform rescales a noise-degree-2 input on entry)
    const std::vector<BootGroup>& stc =
        autom ? stc_at_level(c, B, u[0]->level - (u[0]->noise_level == 2 ? 1 : 0)) : B.stc;
    for (const auto& g : stc) run_lt(c, u, g, B.batched_giant);
    for (size_t i = 0; i < k; ++i) {
        if (sparse) {  // fold the two blocks (Bootstrap.cu:289-292)
            eval_rotate(c, *u[i], B.slots, *tr);
            eval_add(c, *u[i], *tr, *u[i]);
        }
        rescale(c, *u[i]);
        // FLEXIBLEAUTO: the result sits on its level's factor (the products
        // above were encoded for it); the table value replaces the tracked
        // one, which differs from it only by FP64 rounding
        if (technique == FLEXIBLEAUTO) u[i]->noise = c.scaling_factor_real[u[i]->level];
        ct_copy(c, *u[i], *outs[i]);
    }
}

void eval_bootstrap(Context& c, const BootPlan& B, const Ciphertext& in, Ciphertext& out) {
    eval_bootstrap_batch(c, B, {&in}, {&out});
}

}  // namespace fides
