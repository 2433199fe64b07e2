// Plaintext and scalar operations: ct x pt, ct + pt, ct x scalar, ct + scalar,
// ct x integer, ct x monomial, drop to a level.
//
// FIDESlib counterparts: Ciphertext::multPt / addPt / multScalar / addScalar /
// multMonomial / dropToLevel (src/CKKS/Ciphertext.cpp:226-372, :687-735,
// :1709-1760; Ciphertext.cuh:190-401), multIntScalar (Bootstrap.cu)
// and the scalar encodings ContextData::ElemForEvalMult / ElemForEvalAddOrSub
// (src/CKKS/Context.cu:375-564), which restate OpenFHE's CKKS scalar
// encodings.  OpenFHE's big-integer CRT products become 128-bit integer
// products here; the results are the same residues.
#include <cmath>
#include <cstring>

#include "context.hpp"

namespace fides {

static void ck(hipError_t e, const char* w) { hip_check(e, w); }

static EwJobs ew_list() {
    EwJobs l;
    std::memset(&l, 0, sizeof(l));
    return l;
}

// OpenFHE LargeScalingFactorConstants: MAX_BITS_IN_WORD = 61, MAX_LOG_STEP = 60.
constexpr int MAX_BITS_IN_WORD = 61;
constexpr int MAX_BITS_IN_WORD_MULT = 125;  // ElemForEvalMult's 128-bit bound (Context.cu:399)

static uint64_t mulmod_h(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)((unsigned __int128)a * b % q);
}
// 2^e mod q
static uint64_t pow2mod(int e, uint64_t q) {
    uint64_t r = 1 % q, b = 2 % q;
    for (; e > 0; e >>= 1, b = mulmod_h(b, b, q))
        if (e & 1) r = mulmod_h(r, b, q);
    return r;
}

static double scaling_factor(const Context& c, int level) {
    if (c.scaling_factor_real.empty())
        throw Error(E_STATE, "scalar operations need RawParams::ScalingFactorReal (scaling_factor_real)");
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    return c.scaling_factor_real[level];
}

// ElemForEvalMult(level, v) for a ciphertext at `level` (Context.cu:375-455):
// round(v * sf) as a 128-bit integer (scaled down by 2^a when it would exceed
// 125 bits, the factor 2^a restored modulo each prime), reduced into [0, q_i).
std::vector<uint64_t> elem_for_mult(const Context& c, int level, double v) {
    const double sf = scaling_factor(c, level);
    int log_approx = 0;
    const double res = std::fabs(v * sf);
    if (res > 0) {
        const int log_sf = (int)std::ceil(std::log2(res));
        const int log_valid = log_sf <= MAX_BITS_IN_WORD_MULT ? log_sf : MAX_BITS_IN_WORD_MULT;
        log_approx = log_sf - log_valid;
    }
    const double approx = std::pow(2.0, log_approx);
    const __int128 large = (__int128)(v / approx * sf + 0.5);
    const __int128 mag = large < 0 ? -large : large;
    const __int128 bound = (__int128)1 << 63;
    std::vector<uint64_t> out(level + 1);
    for (int i = 0; i <= level; ++i) {
        const uint64_t q = c.q[i];
        uint64_t r;
        if (mag > bound) {
            __int128 t = large % (__int128)q;
            r = (uint64_t)(t < 0 ? t + q : t);
        } else {
            const int64_t t = (int64_t)large % (int64_t)q;
            r = (uint64_t)(t < 0 ? t + (int64_t)q : t);
        }
        if (log_approx > 0) r = mulmod_h(r, pow2mod(log_approx, q), q);
        out[i] = r % q;
    }
    return out;
}

// ElemForEvalAddOrSub(level, |v|, noise_deg) (Context.cu:471-564): the
// constant round(|v| sf) (same 2^a scaling above 61 bits), times the integer
// scaling factor round(sf) once per noise degree above 1.
std::vector<uint64_t> elem_for_add(const Context& c, int level, double v_abs, int noise_deg) {
    if (c.scaling_technique == FLEXIBLEAUTOEXT && level == c.L)
        throw Error(E_STATE, "FLEXIBLEAUTOEXT at the top level needs ScalingFactorRealBig (not supported)");
    const double sf = scaling_factor(c, level);
    int log_approx = 0;
    const double res = std::fabs(v_abs * sf);
    if (res > 0) {
        const int log_sf = (int)std::ceil(std::log2(res));
        const int log_valid = log_sf <= MAX_BITS_IN_WORD ? log_sf : MAX_BITS_IN_WORD;
        log_approx = log_sf - log_valid;
    }
    const double approx = std::pow(2.0, log_approx);
    const uint64_t sc = (uint64_t)(v_abs * sf / approx + 0.5);
    const uint64_t isf = (uint64_t)(sf + 0.5);
    std::vector<uint64_t> out(level + 1);
    for (int i = 0; i <= level; ++i) {
        const uint64_t q = c.q[i];
        uint64_t r = sc % q;
        if (log_approx > 0) r = mulmod_h(r, pow2mod(log_approx, q), q);
        for (int d = 1; d < noise_deg; ++d) r = mulmod_h(r, isf % q, q);
        out[i] = r;
    }
    return out;
}

static LimbConsts limb_consts(const Context& c, const std::vector<uint64_t>& v, bool shoup) {
    if ((int)v.size() > MAX_Q_LIMBS) throw Error(E_ARG, "too many limbs for a scalar op");
    LimbConsts k;
    std::memset(&k, 0, sizeof(k));
    for (size_t i = 0; i < v.size(); ++i) {
        k.c[i].w = v[i];
        k.c[i].wp = shoup ? (uint64_t)(((unsigned __int128)v[i] << 64) / c.q[i]) : 0;
    }
    return k;
}

static void copy_poly(Context& c, uint64_t* dst, const uint64_t* src, int nl) {
    if (dst != src)
        ck(hipMemcpyAsync(dst, src, (size_t)nl * c.N * 8, hipMemcpyDeviceToDevice, c.stream), "copy poly");
}

static void require_loaded(const Ciphertext& a) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
}

// The automatic scaling modes rescale a noise-level-2 input first
// (Ciphertext::multPt / multScalar, Ciphertext.cpp:329-338, :704-712).
static const Ciphertext& auto_rescaled(Context& c, const Ciphertext& a, std::unique_ptr<Ciphertext>& tmp) {
    if (c.scaling_technique == FIXEDMANUAL || a.noise_level != 2) return a;
    tmp = ct_create(c);
    ct_copy(c, a, *tmp);
    rescale(c, *tmp);
    return *tmp;
}

std::unique_ptr<Plaintext> pt_create(Context& c) {
    auto pt = std::make_unique<Plaintext>();
    pt->ctx = &c;
    pt->data = DevBuf(&c.pool, (size_t)(c.L + 1) * c.N);
    return pt;
}

void pt_load(Context& c, Plaintext& pt, const uint64_t* limbs, int num_res, double noise, int noise_level, int slots,
             bool device_src) {
    if (num_res < 1 || num_res > c.L + 1) throw Error(E_LEVEL, "num_res out of range");
    if (!limbs) throw Error(E_ARG, "null limbs");
    ck(hipMemcpyAsync(pt.c0(), limbs, (size_t)num_res * c.N * 8,
                      device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c.stream),
       "load plaintext");
    if (!device_src) c.sync();
    pt.level = num_res - 1;
    pt.ext = 0;
    pt.noise = noise;
    pt.noise_level = noise_level;
    pt.slots = slots;
}

void pt_load_ext(Context& c, Plaintext& pt, const uint64_t* limbs, int num_res, double noise, int noise_level,
                 int slots, bool device_src) {
    if (num_res < 1 || num_res > c.L + 1) throw Error(E_LEVEL, "num_res out of range");
    if (!limbs) throw Error(E_ARG, "null limbs");
    const size_t words = (size_t)(num_res + c.K) * c.N;
    if (pt.data.bytes < words * 8) pt.data = DevBuf(&c.pool, (size_t)(c.L + 1 + c.K) * c.N);
    ck(hipMemcpyAsync(pt.c0(), limbs, words * 8, device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                      c.stream),
       "load plaintext");
    if (!device_src) c.sync();
    pt.level = num_res - 1;
    pt.ext = 1;
    pt.noise = noise;
    pt.noise_level = noise_level;
    pt.slots = slots;
}

// out = a * pt on every limb of a's level (Ciphertext::multPt; the plaintext
// may hold more limbs than the ciphertext, as OpenFHE's LevelReduce of it).
void eval_mult_plain(Context& c, const Ciphertext& a_in, const Plaintext& pt, Ciphertext& out, bool do_rescale) {
    require_loaded(a_in);
    if (pt.level < 0) throw Error(E_STATE, "plaintext not loaded");
    std::unique_ptr<Ciphertext> tmp;
    const Ciphertext& a = auto_rescaled(c, a_in, tmp);
    if (pt.level < a.level) throw Error(E_LEVEL, "plaintext has fewer limbs than the ciphertext");
    if (a.noise_level >= 2 || pt.noise_level >= 2) throw Error(E_STATE, "multPt needs noise level 1 operands");
    if (do_rescale && c.scaling_technique == FIXEDMANUAL && a.level < 1)
        throw Error(E_LEVEL, "cannot rescale at level 0");
    const int nl = a.level + 1;
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), pt.c0(), out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), pt.c0(), out.c1(), nl};
    L.n = 2;
    ck(launch_elementwise(c.T, L, nl, EW_MUL, c.stream), "mult plain");
    const double noise = a.noise * pt.noise;
    const int nlev = a.noise_level + pt.noise_level, slots = std::max(a.slots, pt.slots);
    out.level = a.level;
    out.noise = noise;
    out.noise_level = nlev;
    out.slots = slots;
    if (do_rescale && c.scaling_technique == FIXEDMANUAL) rescale(c, out);
}

// out = a + pt (c0 only; Ciphertext::addPt).
void eval_add_plain(Context& c, const Ciphertext& a, const Plaintext& pt, Ciphertext& out) {
    require_loaded(a);
    if (pt.level < 0) throw Error(E_STATE, "plaintext not loaded");
    if (pt.level < a.level) throw Error(E_LEVEL, "plaintext has fewer limbs than the ciphertext");
    if (a.noise_level != pt.noise_level) throw Error(E_STATE, "addPt needs equal noise levels");
    const int nl = a.level + 1;
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), pt.c0(), out.c0(), nl};
    L.n = 1;
    ck(launch_elementwise(c.T, L, nl, EW_ADD, c.stream), "add plain");
    copy_poly(c, out.c1(), a.c1(), nl);
    const int slots = std::max(a.slots, pt.slots);
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = slots;
}

// out = a - pt (Ciphertext::subPt, Ciphertext.cpp:252-275): c0 minus the
// plaintext, c1 unchanged; operands at equal noise levels, the plaintext's
// limbs covering the ciphertext's.
void eval_sub_plain(Context& c, const Ciphertext& a, const Plaintext& pt, Ciphertext& out) {
    require_loaded(a);
    if (pt.level < 0) throw Error(E_STATE, "plaintext not loaded");
    if (pt.level < a.level) throw Error(E_LEVEL, "plaintext has fewer limbs than the ciphertext");
    if (a.noise_level != pt.noise_level) throw Error(E_STATE, "subPt needs equal noise levels");
    const int nl = a.level + 1;
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), pt.c0(), out.c0(), nl};
    L.n = 1;
    ck(launch_elementwise(c.T, L, nl, EW_SUB, c.stream), "sub plain");
    copy_poly(c, out.c1(), a.c1(), nl);
    const int slots = std::max(a.slots, pt.slots);
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = slots;
}

// acc += a * pt (Ciphertext::addMultPt, Ciphertext.cpp:1410-1426): acc at noise
// level 2, a and pt at 1, on acc's limbs; acc's metadata is kept; a rescale
// follows under FIXEDMANUAL when asked.
void eval_add_mult_plain(Context& c, Ciphertext& acc, const Ciphertext& a, const Plaintext& pt, bool do_rescale) {
    require_loaded(acc);
    require_loaded(a);
    if (pt.level < 0) throw Error(E_STATE, "plaintext not loaded");
    if (acc.noise_level != 2 || a.noise_level != 1 || pt.noise_level != 1)
        throw Error(E_STATE, "addMultPt needs an accumulator at noise level 2 and operands at 1");
    if (a.level < acc.level || pt.level < acc.level) throw Error(E_LEVEL, "operands have fewer limbs than acc");
    if (do_rescale && c.scaling_technique == FIXEDMANUAL && acc.level < 1)
        throw Error(E_LEVEL, "cannot rescale at level 0");
    const int nl = acc.level + 1;
    DevBuf t(&c.pool, (size_t)2 * nl * c.N);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), pt.c0(), t.ptr, nl};
    L.inl[1] = EwJob{a.c1(), pt.c0(), t.ptr + (size_t)nl * c.N, nl};
    L.n = 2;
    ck(launch_elementwise(c.T, L, nl, EW_MUL, c.stream), "add mult plain (product)");
    L.inl[0] = EwJob{acc.c0(), t.ptr, acc.c0(), nl};
    L.inl[1] = EwJob{acc.c1(), t.ptr + (size_t)nl * c.N, acc.c1(), nl};
    ck(launch_elementwise(c.T, L, nl, EW_ADD, c.stream), "add mult plain (sum)");
    if (do_rescale && c.scaling_technique == FIXEDMANUAL) rescale(c, acc);
}

// Sum of terms w[t] x[t] with per-limb constants (one k_lincomb launch); out
// may alias a term.
static void lincomb(Context& c, const std::vector<const Ciphertext*>& x, const std::vector<std::vector<uint64_t>>& w,
                    int nl, Ciphertext& out, const char* what) {
    const int n = (int)x.size();
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pb = (size_t)n * sizeof(uint64_t*), o_x1 = up(pb), o_w = up(o_x1 + pb);
    const size_t bytes = o_w + (size_t)n * nl * sizeof(ShoupConst);
    std::vector<char> host(bytes, 0);
    for (int t = 0; t < n; ++t) {
        const uint64_t* p0 = x[t]->c0();
        const uint64_t* p1 = x[t]->c1();
        std::memcpy(host.data() + t * sizeof(uint64_t*), &p0, sizeof(p0));
        std::memcpy(host.data() + o_x1 + t * sizeof(uint64_t*), &p1, sizeof(p1));
        ShoupConst* ws = (ShoupConst*)(host.data() + o_w) + (size_t)t * nl;
        for (int i = 0; i < nl; ++i)
            ws[i] = ShoupConst{w[t][i], (uint64_t)(((unsigned __int128)w[t][i] << 64) / c.q[i])};
    }
    DevBuf table(&c.pool, (bytes + 7) / 8);
    c.upload(table.ptr, host.data(), bytes);
    char* d = (char*)table.ptr;
    LinCombJob J{(const uint64_t* const*)d, (const uint64_t* const*)(d + o_x1), (const ShoupConst*)(d + o_w),
                 out.c0(), out.c1(), n, nl};
    ck(launch_lincomb(c.T, J, c.stream), what);
}

// acc += b * d (Ciphertext::addMultScalar, Ciphertext.cpp:1338-1358): b times
// ElemForEvalMult(d) at acc's level, added to acc (noise level 2; b at 1 and
// at least acc's level); acc's metadata is kept.
void eval_add_mult_scalar(Context& c, Ciphertext& acc, const Ciphertext& b, double d) {
    require_loaded(acc);
    require_loaded(b);
    if (acc.noise_level != 2 || b.noise_level != 1)
        throw Error(E_STATE, "addMultScalar needs an accumulator at noise level 2 and an operand at 1");
    if (b.level < acc.level) throw Error(E_LEVEL, "operand has fewer limbs than acc");
    const int nl = acc.level + 1;
    std::vector<uint64_t> one(nl, 1);
    lincomb(c, {&acc, &b}, {one, elem_for_mult(c, acc.level, d)}, nl, acc, "add mult scalar");
}

// out = sum_t w[t] x[t] (Ciphertext::evalLinearWSumMutable, Ciphertext.cpp:
// 1272-1335): each weight encoded by ElemForEvalMult at the common level, one
// launch; inputs at noise level 1 and one level; out at noise level 2 with
// scale ScalingFactorReal[l]^2, as the reference records it.
void eval_linear_wsum(Context& c, const std::vector<const Ciphertext*>& x, const std::vector<double>& w,
                      Ciphertext& out) {
    const int n = (int)x.size();
    if (n < 1 || (int)w.size() != n) throw Error(E_ARG, "need one weight per ciphertext");
    if (c.scaling_factor_real.empty()) throw Error(E_STATE, "needs RawParams::ScalingFactorReal");
    const int level = x[0]->level;
    int slots = 0;
    for (const Ciphertext* a : x) {
        require_loaded(*a);
        if (a->noise_level != 1) throw Error(E_STATE, "evalLinearWSum needs inputs at noise level 1");
        if (a->level != level) throw Error(E_LEVEL, "evalLinearWSum needs inputs at one level");
        slots = std::max(slots, a->slots);
    }
    std::vector<std::vector<uint64_t>> ws;
    for (double v : w) ws.push_back(elem_for_mult(c, level, v));
    lincomb(c, x, ws, level + 1, out, "linear weighted sum");
    const double sf = c.scaling_factor_real[level];
    out.level = level;
    out.noise = sf * sf;
    out.noise_level = 2;
    out.slots = slots;
}

// out = a * v (Ciphertext::multScalar): both parts times ElemForEvalMult(v);
// the noise degree rises by one and the scale by ScalingFactorReal[level].
// With rescale (FIXEDMANUAL) the product is rescaled, as multPt does; the
// reference's multScalar passes the flag on but its callee only rescales
// under FIXEDAUTO (Ciphertext.cpp:699-712).
void eval_mult_scalar(Context& c, const Ciphertext& a_in, double v, Ciphertext& out, bool do_rescale) {
    require_loaded(a_in);
    std::unique_ptr<Ciphertext> tmp;
    const Ciphertext& a = auto_rescaled(c, a_in, tmp);
    if (a.noise_level != 1) throw Error(E_STATE, "multScalar needs a noise level 1 input");
    if (do_rescale && c.scaling_technique == FIXEDMANUAL && a.level < 1)
        throw Error(E_LEVEL, "cannot rescale at level 0");
    const int nl = a.level + 1;
    const LimbConsts k = limb_consts(c, elem_for_mult(c, a.level, v), true);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, out.c1(), nl};
    L.n = 2;
    ck(launch_scalar(c.T, L, nl, EW_MUL_SCALAR, k, c.stream), "mult scalar");
    const double noise = a.noise * c.scaling_factor_real[a.level];
    out.level = a.level;
    out.noise = noise;
    out.noise_level = a.noise_level + 1;
    out.slots = a.slots;
    if (do_rescale && c.scaling_technique == FIXEDMANUAL) rescale(c, out);
}

// out = a + v (Ciphertext::addScalar): c0 plus ElemForEvalAddOrSub(|v|),
// negated for v < 0; c1 unchanged.
void eval_add_scalar(Context& c, const Ciphertext& a, double v, Ciphertext& out) {
    require_loaded(a);
    const int nl = a.level + 1;
    if (v != 0.0) {
        std::vector<uint64_t> e = elem_for_add(c, a.level, std::fabs(v), a.noise_level);
        if (v < 0)
            for (int i = 0; i < nl; ++i) e[i] = e[i] ? c.q[i] - e[i] : 0;
        const LimbConsts k = limb_consts(c, e, false);
        EwJobs L = ew_list();
        L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
        L.n = 1;
        ck(launch_scalar(c.T, L, nl, EW_ADD_SCALAR, k, c.stream), "add scalar");
    } else {
        copy_poly(c, out.c0(), a.c0(), nl);
    }
    copy_poly(c, out.c1(), a.c1(), nl);
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

// out = -a (EvalNegate, api/CryptoContext.cpp:630-648): both parts times
// q_i - 1, i.e. the exact negation of OpenFHE's EvalNegate; metadata
// unchanged.  (The reference's GPU branch calls multScalar(-1.0), which under
// FIXEDMANUAL also raises the noise degree and the scale.)
void eval_negate(Context& c, const Ciphertext& a, Ciphertext& out) {
    require_loaded(a);
    const int nl = a.level + 1;
    std::vector<uint64_t> v(nl);
    for (int i = 0; i < nl; ++i) v[i] = c.q[i] - 1;
    const LimbConsts k = limb_consts(c, v, true);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, out.c1(), nl};
    L.n = 2;
    ck(launch_scalar(c.T, L, nl, EW_MUL_SCALAR, k, c.stream), "negate");
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

// out = a * X^m (Ciphertext::multMonomial): metadata unchanged.
void eval_mult_monomial(Context& c, const Ciphertext& a, int64_t m, Ciphertext& out) {
    require_loaded(a);
    const int64_t two_n = 2 * (int64_t)c.N;
    const uint32_t mm = (uint32_t)(((m % two_n) + two_n) % two_n);
    const int nl = a.level + 1;
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, out.c1(), nl};
    L.n = 2;
    ck(launch_monomial(c.T, L, nl, mm, c.stream), "mult monomial");
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

// out = a * s for an integer s (FIDESlib multIntScalar, Bootstrap.cu): every
// limb times s mod q_i; scale and noise degree unchanged.
void eval_mult_int(Context& c, const Ciphertext& a, uint64_t sc, Ciphertext& out) {
    require_loaded(a);
    const int nl = a.level + 1;
    std::vector<uint64_t> v(nl);
    for (int i = 0; i < nl; ++i) v[i] = sc % c.q[i];
    const LimbConsts k = limb_consts(c, v, true);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, out.c1(), nl};
    L.n = 2;
    ck(launch_scalar(c.T, L, nl, EW_MUL_SCALAR, k, c.stream), "mult int");
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

// ---- scale-exact constants (used by the Chebyshev series and EvalMod) ----
// Under FIXEDMANUAL every rescale divides the tracked scale by the dropped
// prime, so ciphertexts reached along different paths carry scales that
// differ from ScalingFactorReal[level] (and from each other) by ~1e-9 per
// rescale, doubled by every squaring.  The two operations below work at the
// ciphertext's tracked scale instead: a constant is added as round(v * scale),
// and a product by a constant is steered onto a chosen scale.

// round(x) mod q_0 .. q_level for a double x: the rounded value written as
// m 2^e with m < 2^53, each residue (m mod q) 2^e mod q, negated for x < 0.
static std::vector<uint64_t> round_residues(const Context& c, int level, double x) {
    std::vector<uint64_t> out(level + 1, 0);
    double r = std::nearbyint(x);
    const bool neg = r < 0;
    r = std::fabs(r);
    if (r == 0.0) return out;
    if (!std::isfinite(r)) throw Error(E_ARG, "constant out of range");
    int e;
    const double f = std::frexp(r, &e);  // r = f 2^e, f in [0.5, 1)
    uint64_t m = (uint64_t)r;
    int sh = 0;
    if (e > 53) {
        m = (uint64_t)std::ldexp(f, 53);
        sh = e - 53;
    }
    for (int i = 0; i <= level; ++i) {
        const uint64_t q = c.q[i];
        const uint64_t v = mulmod_h(m % q, pow2mod(sh, q), q);
        out[i] = (neg && v) ? q - v : v;
    }
    return out;
}

// out = a + v at a's tracked scale: c0 plus round(v * a.noise); metadata unchanged.
void add_const_at_scale(Context& c, const Ciphertext& a, double v, Ciphertext& out) {
    require_loaded(a);
    const int nl = a.level + 1;
    if (v != 0.0) {
        const LimbConsts k = limb_consts(c, round_residues(c, a.level, v * a.noise), false);
        EwJobs L = ew_list();
        L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
        L.n = 1;
        ck(launch_scalar(c.T, L, nl, EW_ADD_SCALAR, k, c.stream), "add const at scale");
    } else {
        copy_poly(c, out.c0(), a.c0(), nl);
    }
    copy_poly(c, out.c1(), a.c1(), nl);
    out.level = a.level;
    out.noise = a.noise;
    out.noise_level = a.noise_level;
    out.slots = a.slots;
}

// out = v * (a at `level`), steered so that the rescale that follows leaves
// the scale s_after: both parts times the integer
// round(v * mrf[level] * s_after / a.noise); the tracked scale becomes
// s_after * mrf[level] and the noise degree rises by one (mrf = the factor
// the rescale divides the tracked scale by, the dropped prime by default).
void mult_const_to_scale(Context& c, const Ciphertext& a, int level, double v, double s_after, Ciphertext& out) {
    require_loaded(a);
    if (level < 1 || level > a.level) throw Error(E_LEVEL, "no level left to rescale into");
    const int nl = level + 1;
    const double mrf = c.mod_reduce_factor[level];
    const LimbConsts k = limb_consts(c, round_residues(c, level, v * mrf * (s_after / a.noise)), true);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, out.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, out.c1(), nl};
    L.n = 2;
    ck(launch_scalar(c.T, L, nl, EW_MUL_SCALAR, k, c.stream), "mult const to scale");
    out.level = level;
    out.noise = s_after * mrf;
    out.noise_level = a.noise_level + 1;
    out.slots = a.slots;
}

// out = sum_t v[t] * (x[t] at `level`), every term steered as in
// mult_const_to_scale onto s_after once rescaled, in ONE launch (each term
// read once): the same residues as the separate products added in turn.
void lincomb_to_scale(Context& c, const std::vector<const Ciphertext*>& x, const std::vector<double>& v, int level,
                      double s_after, Ciphertext& out) {
    const int n = (int)x.size();
    if (n < 1 || (int)v.size() != n) throw Error(E_ARG, "linear combination needs one constant per term");
    if (level < 1) throw Error(E_LEVEL, "no level left to rescale into");
    for (const Ciphertext* a : x) {
        require_loaded(*a);
        if (level > a->level) throw Error(E_LEVEL, "term below the combination's level");
        if (a->noise_level != x[0]->noise_level) throw Error(E_STATE, "terms at different noise levels");
    }
    const int nl = level + 1;
    const double mrf = c.mod_reduce_factor[level];
    // table: x0 pointers | x1 pointers | constants [n][nl]
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pb = (size_t)n * sizeof(uint64_t*), o_x1 = up(pb), o_w = up(o_x1 + pb);
    const size_t bytes = o_w + (size_t)n * nl * sizeof(ShoupConst);
    std::vector<char> host(bytes, 0);
    for (int t = 0; t < n; ++t) {
        const uint64_t* p0 = x[t]->c0();
        const uint64_t* p1 = x[t]->c1();
        std::memcpy(host.data() + t * sizeof(uint64_t*), &p0, sizeof(p0));
        std::memcpy(host.data() + o_x1 + t * sizeof(uint64_t*), &p1, sizeof(p1));
        const std::vector<uint64_t> r = round_residues(c, level, v[t] * mrf * (s_after / x[t]->noise));
        ShoupConst* w = (ShoupConst*)(host.data() + o_w) + (size_t)t * nl;
        for (int i = 0; i < nl; ++i) w[i] = ShoupConst{r[i], (uint64_t)(((unsigned __int128)r[i] << 64) / c.q[i])};
    }
    DevBuf table(&c.pool, (bytes + 7) / 8);
    c.upload(table.ptr, host.data(), bytes);
    char* d = (char*)table.ptr;
    LinCombJob J{(const uint64_t* const*)d, (const uint64_t* const*)(d + o_x1), (const ShoupConst*)(d + o_w),
                 out.c0(), out.c1(), n, nl};
    ck(launch_lincomb(c.T, J, c.stream), "linear combination");
    out.level = level;
    out.noise = s_after * mrf;
    out.noise_level = x[0]->noise_level + 1;
    out.slots = x[0]->slots;
}

// EvalAddMany (api/CryptoContext.cpp:834-880): out = sum of the ciphertexts
// in ONE launch (each read once; the linear-combination kernel with unit
// constants), at the lowest input level; metadata of the first input, as
// EvalAdd keeps its first operand's.  The inputs must share the noise degree.
void eval_add_many(Context& c, const std::vector<const Ciphertext*>& x, Ciphertext& out) {
    const int n = (int)x.size();
    if (n < 1) throw Error(E_ARG, "EvalAddMany needs at least one ciphertext");
    int level = x[0]->level;
    for (const Ciphertext* a : x) {
        require_loaded(*a);
        if (a->noise_level != x[0]->noise_level) throw Error(E_STATE, "EvalAddMany needs one noise degree");
        level = std::min(level, a->level);
    }
    const int nl = level + 1;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t pb = (size_t)n * sizeof(uint64_t*), o_x1 = up(pb), o_w = up(o_x1 + pb);
    const size_t bytes = o_w + (size_t)n * nl * sizeof(ShoupConst);
    std::vector<char> host(bytes, 0);
    for (int t = 0; t < n; ++t) {
        const uint64_t* p0 = x[t]->c0();
        const uint64_t* p1 = x[t]->c1();
        std::memcpy(host.data() + t * sizeof(uint64_t*), &p0, sizeof(p0));
        std::memcpy(host.data() + o_x1 + t * sizeof(uint64_t*), &p1, sizeof(p1));
        ShoupConst* w = (ShoupConst*)(host.data() + o_w) + (size_t)t * nl;
        for (int i = 0; i < nl; ++i) w[i] = ShoupConst{1, (uint64_t)(((unsigned __int128)1 << 64) / c.q[i])};
    }
    DevBuf table(&c.pool, (bytes + 7) / 8);
    c.upload(table.ptr, host.data(), bytes);
    char* d = (char*)table.ptr;
    // the kernel writes each coefficient after reading every term's, so out may alias an input
    LinCombJob J{(const uint64_t* const*)d, (const uint64_t* const*)(d + o_x1), (const ShoupConst*)(d + o_w),
                 out.c0(), out.c1(), n, nl};
    const Ciphertext& first = *x[0];
    const double noise = first.noise;
    const int noise_level = first.noise_level;
    int slots = 0;  // the widest input, as chained adds (Ciphertext.cpp:107)
    for (const Ciphertext* a : x) slots = std::max(slots, a->slots);
    ck(launch_lincomb(c.T, J, c.stream), "add many");
    out.level = level;
    out.noise = noise;
    out.noise_level = noise_level;
    out.slots = slots;
}

// The double-angle iterations with the constants added at the tracked scale
// (EvalMod's form; eval_double_angle keeps the reference's
// ScalingFactorReal encoding).
void double_angle_at_scale(Context& c, Ciphertext& out, int r) {
    require_loaded(out);
    if (r < 0) throw Error(E_ARG, "negative iteration count");
    const double two_pi = 2.0 * 3.14159265358979323846;
    for (int j = 1; j <= r; ++j) {
        if (out.noise_level == 2) rescale(c, out);
        eval_square(c, out, out, false);
        eval_add(c, out, out, out);
        add_const_at_scale(c, out, -1.0 / std::pow(two_pi, std::pow(2.0, j - r)), out);
    }
}

// Power-basis polynomial by Horner's rule (FIDESlib evalHornerSeries,
// ApproxModEval.cu:710-745, with its adjustForMult steps under FIXEDMANUAL):
// acc = x * c_n + c_(n-1) (scalar ops, noise degree 2), then for i = n-2 .. 0:
// rescale acc if at degree 2, acc = acc * x (relinearized, the lower level),
// rescale, acc += c_i.  coeffs[i] is the coefficient of x^i.
void eval_horner_series(Context& c, const Ciphertext& x_in, const std::vector<double>& coeffs, Ciphertext& out) {
    require_loaded(x_in);
    if (c.scaling_technique != FIXEDMANUAL) throw Error(E_STATE, "Horner series runs under FIXEDMANUAL");
    const int n = (int)coeffs.size() - 1;
    if (n < 0) throw Error(E_ARG, "no coefficients");
    std::unique_ptr<Ciphertext> xr;
    const Ciphertext* x = &x_in;
    if (x_in.noise_level == 2) {
        xr = ct_create(c);
        ct_copy(c, x_in, *xr);
        rescale(c, *xr);
        x = xr.get();
    }
    std::unique_ptr<Ciphertext> acc = ct_create(c);
    if (n == 0) {  // the reference: x * 0 + c_0
        eval_mult_scalar(c, *x, 0.0, *acc, false);
        eval_add_scalar(c, *acc, coeffs[0], out);
        return;
    }
    eval_mult_scalar(c, *x, coeffs[n], *acc, false);
    eval_add_scalar(c, *acc, coeffs[n - 1], *acc);
    for (int i = n - 2; i >= 0; --i) {
        if (acc->noise_level == 2) rescale(c, *acc);
        eval_mult(c, *acc, *x, *acc, false);
        rescale(c, *acc);
        eval_add_scalar(c, *acc, coeffs[i], *acc);
    }
    ct_copy(c, *acc, out);
}

// Double-angle iterations of EvalMod (FIDESlib applyDoubleAngleIterations,
// ApproxModEval.cu:747-761): for j = 1 .. r: rescale (FIXEDMANUAL; the
// reference's input is always at noise degree 2 there, a degree-1 input is
// not rescaled), square, double, add -1 / (2 pi)^(2^(j - r)).
void eval_double_angle(Context& c, const Ciphertext& a, int r, Ciphertext& out) {
    require_loaded(a);
    if (r < 0) throw Error(E_ARG, "negative iteration count");
    if (&out != &a) ct_copy(c, a, out);
    for (int j = 1; j <= r; ++j) {
        if (c.scaling_technique == FIXEDMANUAL && out.noise_level == 2) rescale(c, out);
        eval_square(c, out, out, false);
        eval_add(c, out, out, out);
        const double two_pi = 2.0 * 3.14159265358979323846;
        eval_add_scalar(c, out, -1.0 / std::pow(two_pi, std::pow(2.0, j - r)), out);
    }
}

// ------------------------------------------------- automatic rescaling ----
// Ciphertext::multScalarNoPrecheck (src/CKKS/Ciphertext.cpp:687-702), in
// place: both parts times ElemForEvalMult(level, v), one more noise degree,
// the scale times ScalingFactorReal[level].
void mult_scalar_noprecheck(Context& c, Ciphertext& a, double v) {
    require_loaded(a);
    const int nl = a.level + 1;
    const LimbConsts k = limb_consts(c, elem_for_mult(c, a.level, v), true);
    EwJobs L = ew_list();
    L.inl[0] = EwJob{a.c0(), nullptr, a.c0(), nl};
    L.inl[1] = EwJob{a.c1(), nullptr, a.c1(), nl};
    L.n = 2;
    ck(launch_scalar(c.T, L, nl, EW_MUL_SCALAR, k, c.stream), "mult scalar");
    a.noise *= c.scaling_factor_real[a.level];
    a.noise_level += 1;
}

static double sf_big(const Context& c, int level) {
    if (c.scaling_factor_real_big.empty())
        throw Error(E_STATE, "FLEXIBLEAUTO level alignment needs RawParams::ScalingFactorRealBig");
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    return c.scaling_factor_real_big[level];
}

// Ciphertext::adjustForAddOrSub (src/CKKS/Ciphertext.cpp:1472-1594), step for
// step: under FIXEDAUTO / FIXEDMANUAL a rescales or gains a noise degree
// (multScalar(1)) towards b's noise degree; under FLEXIBLEAUTO(EXT) a higher
// a is brought to b's level and scale (scalar multiple by the ratio of the
// scales, rescale, drop).  `a` is the caller's private copy.  Returns false
// when b must be adjusted towards a instead.
bool adjust_for_add_or_sub(Context& c, Ciphertext& a, const Ciphertext& b) {
    require_loaded(a);
    require_loaded(b);
    if (c.scaling_technique == FIXEDMANUAL || c.scaling_technique == FIXEDAUTO) {
        if (a.level - a.noise_level > b.level - b.noise_level) {
            if (b.noise_level == 1 && a.noise_level == 2)
                rescale(c, a);
            else if (b.noise_level == 2 && a.noise_level == 1)
                mult_scalar_noprecheck(c, a, 1.0);
            return true;
        }
        if (b.noise_level == 1 && a.noise_level == 2) {
            rescale(c, a);
            return true;
        }
        return !(a.noise_level == 1 && b.noise_level == 2);
    }
    const int l1 = a.level, l2 = b.level, d1 = a.noise_level, d2 = b.noise_level;
    if (l1 > l2) {
        const double scf1 = a.noise, scf = scaling_factor(c, l1), q1 = c.mod_reduce_factor[l1];
        if (d1 == 2) {
            if (d2 == 2) {
                mult_scalar_noprecheck(c, a, b.noise / scf1 * q1 / scf);
                rescale(c, a);
                if (a.level > b.level) drop_to_level(c, a, b.level);
                a.noise = b.noise;
            } else if (l1 - 1 == l2) {
                rescale(c, a);
            } else {
                mult_scalar_noprecheck(c, a, sf_big(c, l2 + 1) / scf1 * q1 / scf);
                rescale(c, a);
                if (a.level - 1 > b.level) drop_to_level(c, a, b.level + 1);
                rescale(c, a);
                a.noise = b.noise;
            }
        } else if (d2 == 2) {
            mult_scalar_noprecheck(c, a, b.noise / scf1 / scf);
            drop_to_level(c, a, l2);
            a.noise = b.noise;
        } else {
            mult_scalar_noprecheck(c, a, sf_big(c, l2 + 1) / scf1 / scf);
            if (l1 - 1 > l2) drop_to_level(c, a, l2 + 1);
            rescale(c, a);
            a.noise = b.noise;
        }
        return true;
    }
    if (l1 < l2) return false;
    if (d1 < d2) {
        mult_scalar_noprecheck(c, a, 1.0);  // multScalar(1.0, false) at noise degree 1
        return true;
    }
    return d2 >= d1;
}

// Ciphertext::adjustForMult (src/CKKS/Ciphertext.cpp:1595-1611): align as for
// an addition, then bring a to noise degree 1; false when b still has to be
// adjusted (or rescaled) towards a.
bool adjust_for_mult(Context& c, Ciphertext& a, const Ciphertext& b) {
    const bool ok = adjust_for_add_or_sub(c, a, b);
    if (a.noise_level == 2) rescale(c, a);
    return ok && b.noise_level != 2;
}

// Ciphertext::dropToLevel: keep limbs 0..level (metadata only -- a level-l
// ciphertext reads the first l+1 limbs of each part).
void drop_to_level(Context& c, Ciphertext& ct, int level) {
    (void)c;
    require_loaded(ct);
    if (level < 0 || level > ct.level) throw Error(E_LEVEL, "dropToLevel target out of range");
    ct.level = level;
}

}  // namespace fides
