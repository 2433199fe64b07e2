/*
 * fides_amd.hpp -- C++ binding of the MI355X CKKS engine (header only).
 *
 * Mirrors the GPU layer FIDESlib's api/ calls into -- FIDESlib::CKKS::Context,
 * Ciphertext and the raw OpenFHE interop records (src/CKKS/Ciphertext.cuh:34-626,
 * src/CKKS/openfhe-interface/RawCiphertext.cuh:19-93) -- over the C ABI in
 * fides_amd.h, so the GPU branch of api/CryptoContext.cpp can switch engines
 * by type.  Methods keep FIDESlib's in-place form (`a.mult(b)` sets a = a * b;
 * `a.mult(b, c)` sets a = b * c) and its argument meaning; failures throw
 * fides_amd::Error (with the FIDES_E_* code) instead of exiting the process.
 *
 * Link with -lfides_amd.  One GPU per Context.  Calls may come from several
 * threads: each call holds its Context's lock, so calls on one Context run one
 * at a time and calls on different Contexts run concurrently.  A Server is
 * driven from one thread.  Ciphertexts must not outlive their Context.
 */
#ifndef FIDES_AMD_HPP
#define FIDES_AMD_HPP

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fides_amd.h"

namespace fides_amd {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline void check(int rc) {
    if (rc != FIDES_OK) throw Error(rc, fides_last_error());
}

namespace CKKS {

// FIDESlib::CKKS::RawParams (RawCiphertext.cuh:51-92): the fields the engine reads.
struct RawParams {
    int logN = 0, L = 0, K = 0, dnum = 1;
    std::vector<uint64_t> moduli, root_of_unity;                // [L+1]
    std::vector<uint64_t> SPECIALmoduli, SPECIALroot_of_unity;  // [K]
    int scalingTechnique = FIDES_FIXEDMANUAL;                   // lbcrypto::ScalingTechnique order
    std::vector<double> ModReduceFactor;                        // [L+1] or empty (= moduli)
    std::vector<double> ScalingFactorReal;                      // [L+1] or empty (no scalar ops)
    std::vector<double> ScalingFactorRealBig;                   // [L+1] or empty (FLEXIBLEAUTO alignment)
};

// FIDESlib::CKKS::RawPlainText: one polynomial in EVALUATION format, sub_0[limb][N].
// sub_0 may hold the K special-prime limbs after the numRes Q limbs (a P*Q
// plaintext, as RNSPoly::loadConstant reads moduli beyond the Q primes).
struct RawPlainText {
    std::vector<std::vector<uint64_t>> sub_0;
    int numRes = 0;
    int N = 0;
    double Noise = 0.0;
    int NoiseLevel = 1;
    int slots = 0;
};

// FIDESlib::CKKS::RawCipherText (RawCiphertext.cuh:19-35): limbs in OpenFHE
// EVALUATION format, sub_x[limb][N].
struct RawCipherText {
    std::vector<std::vector<uint64_t>> sub_0, sub_1;
    int numRes = 0;
    int N = 0;
    double Noise = 0.0;
    int NoiseLevel = 1;
    int slots = 0;
};

// Key switching key of FIDESlib::CKKS::RawKeySwitchKey (RawCiphertext.cuh:88-93)
// flattened to [dnum][L+1+K][N] per component.
struct FlatKey {
    std::vector<uint64_t> a, b;
};

class Context {
   public:
    // GenCryptoContextGPU(Parameters, devices) for one device (Context.cu:969-987).
    explicit Context(const RawParams& p, int device = 0) {
        fides_raw_params r{};
        r.logN = p.logN;
        r.L = p.L;
        r.K = p.K;
        r.dnum = p.dnum;
        r.moduli = p.moduli.data();
        r.root_of_unity = p.root_of_unity.data();
        r.special_moduli = p.SPECIALmoduli.data();
        r.special_root_of_unity = p.SPECIALroot_of_unity.data();
        r.scaling_technique = p.scalingTechnique;
        r.mod_reduce_factor = p.ModReduceFactor.empty() ? nullptr : p.ModReduceFactor.data();
        r.scaling_factor_real = p.ScalingFactorReal.empty() ? nullptr : p.ScalingFactorReal.data();
        r.scaling_factor_real_big = p.ScalingFactorRealBig.empty() ? nullptr : p.ScalingFactorRealBig.data();
        if ((int)p.moduli.size() != p.L + 1 || (int)p.root_of_unity.size() != p.L + 1 ||
            (int)p.SPECIALmoduli.size() != p.K || (int)p.SPECIALroot_of_unity.size() != p.K ||
            (!p.ModReduceFactor.empty() && (int)p.ModReduceFactor.size() != p.L + 1) ||
            (!p.ScalingFactorReal.empty() && (int)p.ScalingFactorReal.size() != p.L + 1) ||
            (!p.ScalingFactorRealBig.empty() && (int)p.ScalingFactorRealBig.size() != p.L + 1))
            throw Error(FIDES_E_ARG, "RawParams vector sizes do not match L / K");
        check(fides_context_create(&r, device, &h_));
        scale_ = p.ScalingFactorReal.empty() ? 0.0 : p.ScalingFactorReal[0];
        N_ = 1 << p.logN;
        L_ = p.L;
        K_ = p.K;
        dnum_ = p.dnum;
    }
    ~Context() {
        if (boot_) fides_bootstrap_destroy(boot_);
        if (h_) fides_context_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    // KeySwitchingKey::Initialize + Context::AddEvalKey (KeySwitchingKey.cu:20-35).
    void AddEvalKey(const FlatKey& k) {
        key_size(k);
        check(fides_key_load_eval(h_, k.a.data(), k.b.data()));
    }
    // Context::AddRotationKey (Context.cu:589-613): key of a left rotation by `index`.
    void AddRotationKey(int index, const FlatKey& k) {
        key_size(k);
        check(fides_key_load_rotation(h_, index, k.a.data(), k.b.data()));
    }
    // Context::AddRotationKey by automorphism index k (e.g. 2N - 1: conjugation).
    void AddAutomorphKey(uint32_t k, const FlatKey& key) {
        key_size(key);
        check(fides_key_load_automorph(h_, k, key.a.data(), key.b.data()));
    }
    // GetSecretSwitchingKey (RawCiphertext.cu:388-459) under a caller id
    // (FIDES_KEY_SSE_* for bootstrapping's sparse-secret encapsulation).
    void AddSwitchKey(int id, const FlatKey& key) {
        key_size(key);
        check(fides_key_load_switch(h_, id, key.a.data(), key.b.data()));
    }
    void Synchronize() { check(fides_synchronize(h_)); }

    // EvalBootstrapSetup(levelBudget, dim1, slots, correctionFactor)
    // (api/CryptoContext.hpp:80): level budgets {CoeffsToSlots, SlotsToCoeffs}
    // and the slot count (0: N/2); dim1 (baby steps) is chosen per level
    // here, and the correction factor is not used.  EvalMod: K = 28, three
    // double-angle steps, degree 119, at ScalingFactorReal[0].
    void EvalBootstrapSetup(const std::vector<uint32_t>& levelBudget, const std::vector<uint32_t>& dim1 = {0, 0},
                            uint32_t slots = 0, uint32_t correctionFactor = 0, bool sparseEncaps = false) {
        (void)dim1;
        (void)correctionFactor;
        if (levelBudget.size() != 2) throw Error(FIDES_E_ARG, "levelBudget has two entries");
        if (scale_ <= 0) throw Error(FIDES_E_STATE, "bootstrapping needs RawParams::ScalingFactorReal");
        fides_bootstrap_params p{(int32_t)levelBudget[0], (int32_t)levelBudget[1], 28.0, 3, 119, scale_, -1,
                                 sparseEncaps ? 1 : 0, (int32_t)slots, 0, 0};
        fides_bootstrap* b = nullptr;
        check(fides_bootstrap_setup(h_, &p, &b));
        if (boot_) fides_bootstrap_destroy(boot_);
        boot_ = b;
    }
    // The rotation indices EvalBootstrapKeyGen would make keys for (plus the
    // conjugation key, automorphism 2N - 1).
    std::vector<int32_t> BootstrapRotations() const {
        if (!boot_) throw Error(FIDES_E_STATE, "EvalBootstrapSetup first");
        int32_t n = 0;
        check(fides_bootstrap_rotations(boot_, nullptr, 0, &n));
        std::vector<int32_t> r(n);
        check(fides_bootstrap_rotations(boot_, r.data(), n, &n));
        return r;
    }
    fides_bootstrap* bootstrap_handle() const { return boot_; }

    int N() const { return N_; }
    int L() const { return L_; }
    int K() const { return K_; }
    // Serialization of the raw objects (the reference serializes the OpenFHE
    // objects, api/Serialize.cpp): this context's parameters, a key of
    // FIDES_KEY_KIND_* by index, and a key file loaded under its recorded kind.
    void saveParams(const std::string& path) const { check(fides_context_save_params(h_, path.c_str())); }
    void saveKey(int kind, int index, const std::string& path) const {
        check(fides_key_save(h_, kind, index, path.c_str()));
    }
    void loadKeyFile(const std::string& path) { check(fides_key_load_file(h_, path.c_str(), nullptr, nullptr)); }
    fides_context* handle() const { return h_; }

   private:
    void key_size(const FlatKey& k) const {
        const size_t words = (size_t)dnum_ * (L_ + 1 + K_) * N_;
        if (k.a.size() != words || k.b.size() != words) throw Error(FIDES_E_ARG, "key size is not dnum*(L+1+K)*N");
    }
    fides_context* h_ = nullptr;
    fides_bootstrap* boot_ = nullptr;
    double scale_ = 0.0;
    int N_ = 0, L_ = 0, K_ = 0, dnum_ = 1;
};

// FIDESlib::CKKS::Plaintext (Plaintext.cuh:22-120) built from a RawPlainText.
class Plaintext {
   public:
    explicit Plaintext(Context& cc) : cc_(&cc) { check(fides_pt_create(cc.handle(), &h_)); }
    Plaintext(Context& cc, const RawPlainText& raw) : Plaintext(cc) { load(raw); }
    ~Plaintext() {
        if (h_) fides_pt_destroy(h_);
    }
    Plaintext(const Plaintext&) = delete;
    Plaintext& operator=(const Plaintext&) = delete;

    void load(const RawPlainText& raw) {
        const size_t N = cc_->N();
        if ((int)raw.sub_0.size() < raw.numRes) throw Error(FIDES_E_ARG, "RawPlainText has fewer limbs than numRes");
        std::vector<uint64_t> s0((size_t)raw.numRes * N);
        for (int i = 0; i < raw.numRes; ++i) {
            if (raw.sub_0[i].size() != N) throw Error(FIDES_E_ARG, "limb length is not N");
            std::copy(raw.sub_0[i].begin(), raw.sub_0[i].end(), s0.begin() + i * N);
        }
        if ((int)raw.sub_0.size() > raw.numRes) {  // the special limbs too: P*Q form
            if ((int)raw.sub_0.size() != raw.numRes + cc_->K()) throw Error(FIDES_E_ARG, "expected K special limbs");
            s0.resize((size_t)(raw.numRes + cc_->K()) * N);
            for (int i = raw.numRes; i < raw.numRes + cc_->K(); ++i) {
                if (raw.sub_0[i].size() != N) throw Error(FIDES_E_ARG, "limb length is not N");
                std::copy(raw.sub_0[i].begin(), raw.sub_0[i].end(), s0.begin() + i * N);
            }
            check(fides_pt_load_ext(cc_->handle(), h_, s0.data(), raw.numRes, raw.Noise, raw.NoiseLevel, raw.slots));
            return;
        }
        check(fides_pt_load(cc_->handle(), h_, s0.data(), raw.numRes, raw.Noise, raw.NoiseLevel, raw.slots));
    }
    // MakeCKKSPackedPlaintext (api/CryptoContext.cpp:475-514): N/2 complex slot
    // values (or real ones, im empty) encoded at `level` with scaling factor `scale`.
    void encode(const std::vector<double>& re, const std::vector<double>& im, int level, double scale) {
        const size_t n = (size_t)cc_->N() / 2;
        if (re.size() != n || (!im.empty() && im.size() != n)) throw Error(FIDES_E_ARG, "need N/2 slot values");
        check(fides_encode(cc_->handle(), re.data(), im.empty() ? nullptr : im.data(), level, scale, h_));
    }
    // The same encoding in P*Q form (linear-transform diagonals for the
    // extended-basis path).
    void encodeExt(const std::vector<double>& re, const std::vector<double>& im, int level, double scale) {
        const size_t n = (size_t)cc_->N() / 2;
        if (re.size() != n || (!im.empty() && im.size() != n)) throw Error(FIDES_E_ARG, "need N/2 slot values");
        check(fides_encode_ext(cc_->handle(), re.data(), im.empty() ? nullptr : im.data(), level, scale, h_));
    }
    void save(const std::string& path) const { check(fides_pt_save(cc_->handle(), h_, path.c_str())); }
    void loadFile(const std::string& path) { check(fides_pt_load_file(cc_->handle(), h_, path.c_str())); }
    fides_plaintext* handle() const { return h_; }

   private:
    Context* cc_;
    fides_plaintext* h_ = nullptr;
};

class Ciphertext {
   public:
    explicit Ciphertext(Context& cc) : cc_(&cc) { check(fides_ct_create(cc.handle(), &h_)); }
    Ciphertext(Context& cc, const RawCipherText& raw) : Ciphertext(cc) { load(raw); }
    ~Ciphertext() {
        if (h_) fides_ct_destroy(h_);
    }
    Ciphertext(const Ciphertext&) = delete;
    Ciphertext& operator=(const Ciphertext&) = delete;

    void load(const RawCipherText& raw) {
        const size_t N = cc_->N();
        if ((int)raw.sub_0.size() < raw.numRes || (int)raw.sub_1.size() < raw.numRes)
            throw Error(FIDES_E_ARG, "RawCipherText has fewer limbs than numRes");
        std::vector<uint64_t> s0((size_t)raw.numRes * N), s1((size_t)raw.numRes * N);
        for (int i = 0; i < raw.numRes; ++i) {
            if (raw.sub_0[i].size() != N || raw.sub_1[i].size() != N) throw Error(FIDES_E_ARG, "limb length is not N");
            std::copy(raw.sub_0[i].begin(), raw.sub_0[i].end(), s0.begin() + i * N);
            std::copy(raw.sub_1[i].begin(), raw.sub_1[i].end(), s1.begin() + i * N);
        }
        check(fides_ct_load(cc_->handle(), h_, s0.data(), s1.data(), raw.numRes, raw.Noise, raw.NoiseLevel,
                            raw.slots));
    }
    void store(RawCipherText& raw) const {
        int32_t level, nl, slots;
        double noise;
        check(fides_ct_info(h_, &level, &noise, &nl, &slots));
        const size_t N = cc_->N(), n = (size_t)level + 1;
        std::vector<uint64_t> s0(n * N), s1(n * N);
        check(fides_ct_store(cc_->handle(), h_, s0.data(), s1.data()));
        raw.numRes = (int)n;
        raw.N = (int)N;
        raw.Noise = noise;
        raw.NoiseLevel = nl;
        raw.slots = slots;
        raw.sub_0.assign(n, std::vector<uint64_t>(N));
        raw.sub_1.assign(n, std::vector<uint64_t>(N));
        for (size_t i = 0; i < n; ++i) {
            std::copy(s0.begin() + i * N, s0.begin() + (i + 1) * N, raw.sub_0[i].begin());
            std::copy(s1.begin() + i * N, s1.begin() + (i + 1) * N, raw.sub_1[i].begin());
        }
    }
    int getLevel() const {
        int32_t level, nl, slots;
        double noise;
        check(fides_ct_info(h_, &level, &noise, &nl, &slots));
        return level;
    }

    // this = this + b / this = a + b  (Ciphertext.cuh:167, :500)
    void add(const Ciphertext& b) { check(fides_eval_add(ctx(), h_, b.h_, h_)); }
    void add(const Ciphertext& a, const Ciphertext& b) { check(fides_eval_add(ctx(), a.h_, b.h_, h_)); }
    void sub(const Ciphertext& b) { check(fides_eval_sub(ctx(), h_, b.h_, h_)); }
    void sub(const Ciphertext& a, const Ciphertext& b) { check(fides_eval_sub(ctx(), a.h_, b.h_, h_)); }
    // this = this * b / this = b * c, relinearized, then rescaled when asked (:264, :277)
    void mult(const Ciphertext& b, bool rescale = false) { check(fides_eval_mult(ctx(), h_, b.h_, h_, rescale)); }
    void mult(const Ciphertext& b, const Ciphertext& c, bool rescale = false) {
        check(fides_eval_mult(ctx(), b.h_, c.h_, h_, rescale));
    }
    void square(bool rescale = false) { check(fides_eval_square(ctx(), h_, h_, rescale)); }
    void square(const Ciphertext& src, bool rescale = false) {
        check(fides_eval_square(ctx(), src.h_, h_, rescale));
    }
    void rescale() { check(fides_rescale(ctx(), h_)); }
    // this = rotate(this, index) / this = rotate(c, index)  (:348, :359)
    void rotate(int index) { check(fides_eval_rotate(ctx(), h_, index, h_)); }
    void rotate(const Ciphertext& c, int index) { check(fides_eval_rotate(ctx(), c.h_, index, h_)); }
    void copy(const Ciphertext& c) { check(fides_ct_copy(ctx(), c.h_, h_)); }
    // this = this * pt / this = c * pt, rescaled when asked (:226, :237)
    void multPt(const Plaintext& b, bool rescale = false) {
        check(fides_eval_mult_plain(ctx(), h_, b.handle(), h_, rescale));
    }
    void multPt(const Ciphertext& c, const Plaintext& b, bool rescale = false) {
        check(fides_eval_mult_plain(ctx(), c.h_, b.handle(), h_, rescale));
    }
    // EvalNegate, in place (exact negation, metadata unchanged)
    void negate() { check(fides_eval_negate(ctx(), h_, h_)); }
    // AccumulateSum(ct, slots, stride), in place
    void accumulateSum(int slots, int stride) { check(fides_eval_accumulate_sum(ctx(), h_, slots, stride, h_)); }
    // Ciphertext::keySwitch (Ciphertext.cpp:1450-1461) with a key of AddSwitchKey
    void keySwitch(int id) { check(fides_eval_key_switch(ctx(), h_, id, h_)); }
    void addPt(const Plaintext& b) { check(fides_eval_add_plain(ctx(), h_, b.handle(), h_)); }
    void addPt(const Ciphertext& c, const Plaintext& b) { check(fides_eval_add_plain(ctx(), c.h_, b.handle(), h_)); }
    // this = this * v / this = b * v (:303, :315); this = this + v / b + v (:211, :489)
    void multScalar(double v, bool rescale = false) { check(fides_eval_mult_scalar(ctx(), h_, v, h_, rescale)); }
    void multScalar(const Ciphertext& b, double v, bool rescale = false) {
        check(fides_eval_mult_scalar(ctx(), b.h_, v, h_, rescale));
    }
    void addScalar(double v) { check(fides_eval_add_scalar(ctx(), h_, v, h_)); }
    void addScalar(const Ciphertext& b, double v) { check(fides_eval_add_scalar(ctx(), b.h_, v, h_)); }
    void dropToLevel(int level) { check(fides_drop_to_level(ctx(), h_, level)); }
    // this = conj(c) (:369), this = this * X^power (:626), this = this * s (multIntScalar)
    void conjugate(const Ciphertext& c) { check(fides_eval_conjugate(ctx(), c.h_, h_)); }
    void multMonomial(int64_t power) { check(fides_eval_mult_monomial(ctx(), h_, power, h_)); }
    void multInt(uint64_t s) { check(fides_eval_mult_int(ctx(), h_, s, h_)); }
    // ModRaise of bootstrapping: this (level 0) raised to `level`
    void modRaise(int level) { check(fides_mod_raise(ctx(), h_, level, h_)); }
    // results[i] = rotate(this, indexes[i]) with one shared ModUp (:453);
    // with ext the results stay in P*Q (EvalFastRotationExt) until modDown()
    void rotate_hoisted(const std::vector<int>& indexes, const std::vector<Ciphertext*>& results,
                        bool ext = false) const {
        if (indexes.size() != results.size()) throw Error(FIDES_E_ARG, "index / result count mismatch");
        std::vector<int32_t> idx(indexes.begin(), indexes.end());
        std::vector<fides_ciphertext*> outs;
        for (Ciphertext* r : results) outs.push_back(r->h_);
        if (ext)
            check(fides_eval_rotate_hoisted_ext(ctx(), h_, idx.data(), (int32_t)idx.size(), outs.data()));
        else
            check(fides_eval_rotate_hoisted(ctx(), h_, idx.data(), (int32_t)idx.size(), outs.data()));
    }

    // P*Q form (Ciphertext.cuh:379-382, 441): extend() = (P c0, P c1) in P*Q,
    // modDown() back to the Q basis; isExt() reports the form.
    void extend() { check(fides_ct_extend(ctx(), h_, h_)); }
    void modDown() { check(fides_ct_mod_down(ctx(), h_, h_)); }
    bool isExt() const {
        int32_t e = 0;
        check(fides_ct_is_ext(h_, &e));
        return e != 0;
    }
    // mult(b, rescale, moddown) (:264): moddown = false leaves the
    // relinearized product in P*Q (no rescale then).
    void mult(const Ciphertext& b, bool rescale, bool moddown) {
        if (moddown) {
            mult(b, rescale);
            return;
        }
        if (rescale) throw Error(FIDES_E_ARG, "a product left in P*Q cannot be rescaled");
        check(fides_eval_mult_ext(ctx(), h_, b.h_, h_));
    }
    // P*Q addition (a Q operand is extended first)
    void addExt(const Ciphertext& b) { check(fides_eval_add_ext(ctx(), h_, b.h_, h_)); }
    // dotProductPt (:596-617): this = sum_i cts[i] * pts[i]
    void dotProductPt(const std::vector<const Ciphertext*>& cts, const std::vector<const Plaintext*>& pts) {
        if (cts.size() != pts.size()) throw Error(FIDES_E_ARG, "one plaintext per ciphertext");
        std::vector<const fides_ciphertext*> hc;
        std::vector<const fides_plaintext*> hp;
        for (const Ciphertext* c : cts) hc.push_back(c->h_);
        for (const Plaintext* p : pts) hp.push_back(p->handle());
        check(fides_eval_dot_product_plain(ctx(), hc.data(), hp.data(), (int32_t)hc.size(), h_));
    }
    // dotProduct (:617): this = sum_i a[i] * b[i], one key switch of the summed
    // d2, not rescaled; ext leaves the result in P*Q
    void dotProduct(const std::vector<Ciphertext*>& a, const std::vector<Ciphertext*>& b, bool ext) {
        if (a.size() != b.size()) throw Error(FIDES_E_ARG, "as many ciphertexts on both sides");
        std::vector<const fides_ciphertext*> ha, hb;
        for (const Ciphertext* c : a) ha.push_back(c->h_);
        for (const Ciphertext* c : b) hb.push_back(c->h_);
        check(fides_eval_dot_product(ctx(), ha.data(), hb.data(), (int32_t)ha.size(), ext ? 1 : 0, h_));
    }
    // subPt (:199), addMultPt (:249), addMultScalar (:479), evalLinearWSumMutable (:467)
    void subPt(const Plaintext& b) { check(fides_eval_sub_plain(ctx(), h_, b.handle(), h_)); }
    void addMultPt(const Ciphertext& c, const Plaintext& b, bool rescale = false) {
        check(fides_eval_add_mult_plain(ctx(), h_, c.h_, b.handle(), rescale));
    }
    void addMultScalar(const Ciphertext& b, double d) { check(fides_eval_add_mult_scalar(ctx(), h_, b.h_, d)); }
    void evalLinearWSumMutable(uint32_t n, const std::vector<Ciphertext*>& ctxs, const std::vector<double>& weights) {
        if (ctxs.size() < n || weights.size() < n) throw Error(FIDES_E_ARG, "fewer terms than n");
        std::vector<const fides_ciphertext*> hc;
        for (uint32_t i = 0; i < n; ++i) hc.push_back(ctxs[i]->h_);
        check(fides_eval_linear_wsum(ctx(), hc.data(), weights.data(), (int32_t)n, h_));
    }
    // the engine's ciphertext file (either form)
    void save(const std::string& path) const { check(fides_ct_save(ctx(), h_, path.c_str())); }
    void loadFile(const std::string& path) { check(fides_ct_load_file(ctx(), h_, path.c_str())); }

    fides_ciphertext* handle() const { return h_; }
    Context* context() const { return cc_; }

   private:
    fides_context* ctx() const { return cc_->handle(); }
    Context* cc_;
    fides_ciphertext* h_ = nullptr;
};

// Batched HMult+relin(+rescale): out[i] = a[i] * b[i] for every pair, at any
// mix of levels, one launch per stage (no FIDESlib counterpart).
inline void multBatch(Context& cc, const std::vector<const Ciphertext*>& a, const std::vector<const Ciphertext*>& b,
                      const std::vector<Ciphertext*>& out, bool rescale = true) {
    if (a.size() != b.size() || a.size() != out.size()) throw Error(FIDES_E_ARG, "batch size mismatch");
    std::vector<const fides_ciphertext*> ha, hb;
    std::vector<fides_ciphertext*> ho;
    for (size_t i = 0; i < a.size(); ++i) {
        ha.push_back(a[i]->handle());
        hb.push_back(b[i]->handle());
        ho.push_back(out[i]->handle());
    }
    check(fides_eval_mult_batch(cc.handle(), ha.data(), hb.data(), ho.data(), (int32_t)a.size(), rescale));
}

// Batched rotations: out[i] = rotate(a[i], indices[i]), one launch per stage.
inline void rotateBatch(Context& cc, const std::vector<const Ciphertext*>& a, const std::vector<int>& indices,
                        const std::vector<Ciphertext*>& out) {
    if (a.size() != indices.size() || a.size() != out.size()) throw Error(FIDES_E_ARG, "batch size mismatch");
    std::vector<const fides_ciphertext*> ha;
    std::vector<fides_ciphertext*> ho;
    for (size_t i = 0; i < a.size(); ++i) {
        ha.push_back(a[i]->handle());
        ho.push_back(out[i]->handle());
    }
    std::vector<int32_t> idx(indices.begin(), indices.end());
    check(fides_eval_rotate_batch(cc.handle(), ha.data(), idx.data(), ho.data(), (int32_t)a.size()));
}

// FIDESlib::CKKS::LinearTransform(ctxt, rowSize, bStep, pts, stride, offset)
// (LinearTransform.cu:67-258), in place; rowSize = pts.size().
inline void LinearTransform(Ciphertext& ctxt, int bStep, const std::vector<Plaintext*>& pts, int stride, int offset,
                            bool batchedGiantSteps = false) {
    std::vector<const fides_plaintext*> hp;
    for (Plaintext* p : pts) hp.push_back(p->handle());
    check(fides_eval_linear_transform(ctxt.context()->handle(), ctxt.handle(), hp.data(), (int32_t)hp.size(), bStep,
                                      stride, offset, batchedGiantSteps ? FIDES_LT_BATCHED_GIANT : 0, ctxt.handle()));
}

// The transform of several ciphertexts at one level with the same diagonals
// (fides_eval_linear_transform_batch), in place.
inline void LinearTransformBatch(const std::vector<Ciphertext*>& ctxts, int bStep, const std::vector<Plaintext*>& pts,
                                 int stride, int offset, bool batchedGiant = false) {
    if (ctxts.empty()) return;
    std::vector<fides_ciphertext*> hc;
    for (Ciphertext* c : ctxts) hc.push_back(c->handle());
    std::vector<const fides_plaintext*> hp;
    for (Plaintext* p : pts) hp.push_back(p->handle());
    check(fides_eval_linear_transform_batch(ctxts[0]->context()->handle(), (const fides_ciphertext* const*)hc.data(),
                                            (int32_t)hc.size(), hp.data(), (int32_t)hp.size(), bStep, stride, offset,
                                            batchedGiant ? FIDES_LT_BATCHED_GIANT : 0, hc.data()));
}

// FIDESlib::CKKS::ConvolutionTransform(ctxt, rowSize, bStep, pts, stride, indexes, gStep)
// (LinearTransform.cu:255-441), in place; rowSize = pts.size().
inline void ConvolutionTransform(Ciphertext& ctxt, int bStep, const std::vector<Plaintext*>& pts, int stride,
                                 const std::vector<int>& indexes, uint32_t gStep) {
    std::vector<const fides_plaintext*> hp;
    for (Plaintext* p : pts) hp.push_back(p->handle());
    std::vector<int32_t> ix(indexes.begin(), indexes.end());
    check(fides_eval_convolution_transform(ctxt.context()->handle(), ctxt.handle(), hp.data(), (int32_t)hp.size(),
                                           bStep, stride, ix.data(), (int32_t)gStep, ctxt.handle()));
}

// GetConvolutionTransformRotationIndices (LinearTransform.cu:655-678), with
// stride min(gStep, 8) included (the first giant step of a block rotates by it).
inline std::vector<int> GetConvolutionTransformRotationIndices(int stride, uint32_t gStep) {
    int32_t n = 0;
    check(fides_convolution_rotations(stride, (int32_t)gStep, nullptr, 0, &n));
    std::vector<int32_t> r((size_t)n);
    check(fides_convolution_rotations(stride, (int32_t)gStep, r.data(), n, &n));
    return std::vector<int>(r.begin(), r.end());
}

// evalHornerSeries / applyDoubleAngleIterations (ApproxModEval.cu:710-761), in place.
inline void evalHornerSeries(Ciphertext& ctxt, const std::vector<double>& coefficients) {
    check(fides_eval_horner_series(ctxt.context()->handle(), ctxt.handle(), coefficients.data(),
                                   (int32_t)coefficients.size(), ctxt.handle()));
}
inline void evalChebyshevSeries(Ciphertext& ctxt, const std::vector<double>& coefficients, double lower,
                                double upper) {
    check(fides_eval_chebyshev_series(ctxt.context()->handle(), ctxt.handle(), coefficients.data(),
                                      (int32_t)coefficients.size(), lower, upper, ctxt.handle()));
}
inline void applyDoubleAngleIterations(Ciphertext& ctxt, int its) {
    check(fides_eval_double_angle(ctxt.context()->handle(), ctxt.handle(), its, ctxt.handle()));
}

// EvalBootstrapSetup / EvalBootstrap (api/CryptoContext.hpp:80, :180; FIDESlib
// CKKS::Bootstrap, Bootstrap.cu:169-300) for fully packed ciphertexts.
// rotations() are the keys EvalBootstrapKeyGen would make, besides the
// conjugation key (automorphism 2N - 1).
class Bootstrapper {
   public:
    Bootstrapper(Context& cc, double scale, int levelBudgetCtS = 3, int levelBudgetStC = 3, double K = 28.0,
                 int doubleAngle = 3, int degree = 119, int topLevel = -1, bool sparseEncaps = false,
                 int slots = 0, bool batchedGiant = false, bool qBasisDiagonals = false)
        : cc_(&cc) {
        fides_bootstrap_params p{levelBudgetCtS, levelBudgetStC, K, doubleAngle, degree, scale, topLevel,
                                 sparseEncaps ? 1 : 0, slots, batchedGiant ? 1 : 0, qBasisDiagonals ? 1 : 0};
        check(fides_bootstrap_setup(cc.handle(), &p, &h_));
    }
    ~Bootstrapper() {
        if (h_) fides_bootstrap_destroy(h_);
    }
    Bootstrapper(const Bootstrapper&) = delete;
    Bootstrapper& operator=(const Bootstrapper&) = delete;

    std::vector<int32_t> rotations() const {
        int32_t n = 0;
        check(fides_bootstrap_rotations(h_, nullptr, 0, &n));
        std::vector<int32_t> r(n);
        check(fides_bootstrap_rotations(h_, r.data(), n, &n));
        return r;
    }
    // EvalBootstrapInPlace
    void bootstrap(Ciphertext& ctxt) const {
        check(fides_eval_bootstrap(cc_->handle(), h_, ctxt.handle(), ctxt.handle()));
    }
    // several ciphertexts in place, their EvalMods in lockstep
    void bootstrap(const std::vector<Ciphertext*>& cts) const {
        std::vector<fides_ciphertext*> h;
        for (Ciphertext* c : cts) h.push_back(c->handle());
        check(fides_eval_bootstrap_batch(cc_->handle(), h_, (const fides_ciphertext* const*)h.data(), h.data(),
                                         (int32_t)h.size()));
    }

   private:
    Context* cc_;
    fides_bootstrap* h_ = nullptr;
};

// EvalAddMany (api/CryptoContext.cpp:834-880): out = sum of cts, one launch.
inline void EvalAddMany(Ciphertext& out, const std::vector<const Ciphertext*>& cts) {
    std::vector<const fides_ciphertext*> h;
    for (const Ciphertext* c : cts) h.push_back(c->handle());
    check(fides_eval_add_many(out.context()->handle(), h.data(), (int32_t)h.size(), out.handle()));
}

// EvalBootstrapInPlace(ct) (api/CryptoContext.hpp:181) with the context's
// EvalBootstrapSetup plan.
inline void EvalBootstrapInPlace(Ciphertext& ctxt) {
    Context* cc = ctxt.context();
    if (!cc->bootstrap_handle()) throw Error(FIDES_E_STATE, "EvalBootstrapSetup first");
    check(fides_eval_bootstrap(cc->handle(), cc->bootstrap_handle(), ctxt.handle(), ctxt.handle()));
}

// Request loop with continuous batching (no FIDESlib counterpart): jobs on
// device-resident operands, advanced by step(); poll() returns finished ids.
class Server {
   public:
    Server(Context& cc, int maxSlots, uint64_t budgetBytes) {
        check(fides_server_create(cc.handle(), maxSlots, budgetBytes, &h_));
    }
    ~Server() {
        if (h_) fides_server_destroy(h_);
    }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    int64_t submitMult(const Ciphertext& a, const Ciphertext& b, Ciphertext& out) {
        int64_t id;
        check(fides_server_submit_mult(h_, a.handle(), b.handle(), out.handle(), &id));
        return id;
    }
    int64_t submitRotate(const Ciphertext& a, int index, Ciphertext& out) {
        int64_t id;
        check(fides_server_submit_rotate(h_, a.handle(), index, out.handle(), &id));
        return id;
    }
    int64_t submitLogReg(const Ciphertext& x, const Plaintext& w, int nRot, const double coeffs[3], Ciphertext& out) {
        int64_t id;
        check(fides_server_submit_logreg(h_, x.handle(), w.handle(), nRot, coeffs, out.handle(), &id));
        return id;
    }
    // one scheduling step; returns the jobs still active or queued
    int step() {
        int32_t p;
        check(fides_server_step(h_, &p));
        return p;
    }
    std::vector<int64_t> poll() {
        std::vector<int64_t> out;
        int64_t buf[256];
        int32_t n;
        do {
            check(fides_server_poll(h_, buf, 256, &n));
            out.insert(out.end(), buf, buf + n);
        } while (n == 256);
        return out;
    }

   private:
    fides_server* h_ = nullptr;
};

}  // namespace CKKS
}  // namespace fides_amd

#endif
