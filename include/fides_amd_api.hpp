/*
 * fides_amd_api.hpp -- the api/ surface of FIDESlib (api/CryptoContext.hpp:24-245,
 * api/Ciphertext.hpp) over the MI355X engine (header only, on fides_amd.hpp).
 *
 * CryptoContext and Ciphertext / Plaintext are shared_ptrs as in the reference
 * (api/Definitions.hpp:17-21); operations return new ciphertexts or work
 * *InPlace, with the reference's argument meaning: EvalMult(ct, ct) is tensor
 * + relinearization and leaves noise degree 2 (rescale explicit under
 * FIXEDMANUAL, lazy under FIXEDAUTO / FLEXIBLEAUTO), levels are OpenFHE's
 * (levels consumed: engine level = L - level).  The engine holds every
 * operand on the GPU from creation, so there is no LoadCiphertext /
 * auto-load step.  Key generation, encryption and decryption are OpenFHE's
 * CPU calls and stay outside: keys and ciphertexts enter in the interop
 * records (RawCipherText, RawPlainText, key A / B vectors) through
 * LoadCiphertext / LoadPlaintext / AddEvalKey / AddRotationKey, and leave
 * through StoreCiphertext.  Failures throw fides_amd::Error.
 */
#ifndef FIDES_AMD_API_HPP
#define FIDES_AMD_API_HPP

#include <cmath>
#include <complex>
#include <functional>
#include <memory>
#include <vector>

#include "fides_amd.hpp"

namespace fides_amd {
namespace api {

class CryptoContextImpl;
using CryptoContext = std::shared_ptr<CryptoContextImpl>;

// CiphertextImpl<DCRTPoly> (api/Ciphertext.hpp): a device ciphertext and the
// context that owns it.
class CiphertextImpl {
   public:
    explicit CiphertextImpl(const CryptoContext& cc);
    CKKS::Ciphertext& gpu() { return *ct_; }
    const CKKS::Ciphertext& gpu() const { return *ct_; }
    fides_ciphertext* handle() const { return ct_->handle(); }
    // OpenFHE's level (levels consumed) and noise scale degree
    uint32_t GetLevel() const;
    size_t GetNoiseScaleDeg() const { return (size_t)info().nl; }
    double GetScalingFactor() const { return info().noise; }
    uint32_t GetSlots() const { return (uint32_t)info().slots; }
    const CryptoContext& GetCryptoContext() const { return cc_; }

   private:
    struct Info {
        int32_t level, nl, slots;
        double noise;
    };
    Info info() const {
        Info i{};
        check(fides_ct_info(ct_->handle(), &i.level, &i.noise, &i.nl, &i.slots));
        return i;
    }
    CryptoContext cc_;
    std::unique_ptr<CKKS::Ciphertext> ct_;
};
using Ciphertext = std::shared_ptr<CiphertextImpl>;

// PlaintextImpl (OpenFHE Plaintext as the reference's GPU path holds it).
class PlaintextImpl {
   public:
    explicit PlaintextImpl(const CryptoContext& cc);
    CKKS::Plaintext& gpu() { return *pt_; }
    const CKKS::Plaintext& gpu() const { return *pt_; }
    fides_plaintext* handle() const { return pt_->handle(); }

   private:
    CryptoContext cc_;
    std::unique_ptr<CKKS::Plaintext> pt_;
};
using Plaintext = std::shared_ptr<PlaintextImpl>;

// CryptoContextImpl<DCRTPoly> (api/CryptoContext.hpp:24-245), GPU path only.
class CryptoContextImpl : public std::enable_shared_from_this<CryptoContextImpl> {
   public:
    CryptoContextImpl(const CKKS::RawParams& p, int device) : p_(p), cc_(p, device) {}
    CryptoContextImpl(const CryptoContextImpl&) = delete;
    CryptoContextImpl& operator=(const CryptoContextImpl&) = delete;

    // ---- getters ----
    uint32_t GetRingDimension() const { return (uint32_t)cc_.N(); }
    uint32_t GetCyclotomicOrder() const { return 2u * (uint32_t)cc_.N(); }
    uint32_t GetMultiplicativeDepth() const { return (uint32_t)cc_.L(); }
    CKKS::Context& gpu() { return cc_; }

    // ---- keys and operands from the OpenFHE interop records ----
    void AddEvalKey(const CKKS::FlatKey& k) { cc_.AddEvalKey(k); }
    void AddRotationKey(int index, const CKKS::FlatKey& k) { cc_.AddRotationKey(index, k); }
    void AddAutomorphKey(uint32_t k, const CKKS::FlatKey& key) { cc_.AddAutomorphKey(k, key); }
    void AddSwitchKey(int id, const CKKS::FlatKey& key) { cc_.AddSwitchKey(id, key); }
    Ciphertext LoadCiphertext(const CKKS::RawCipherText& raw) {
        Ciphertext c = fresh();
        c->gpu().load(raw);
        return c;
    }
    void StoreCiphertext(const Ciphertext& ct, CKKS::RawCipherText& raw) const { ct->gpu().store(raw); }
    Plaintext LoadPlaintext(const CKKS::RawPlainText& raw) {
        Plaintext p = std::make_shared<PlaintextImpl>(self());
        p->gpu().load(raw);
        return p;
    }

    // ---- encoding (api/CryptoContext.cpp:475-514) ----
    // Full packing (slots 0 or N/2; shorter inputs are zero-padded as OpenFHE
    // pads them), at OpenFHE level `level` with ScalingFactorReal of that level
    // raised to noiseScaleDeg.
    Plaintext MakeCKKSPackedPlaintext(const std::vector<std::complex<double>>& value, size_t noiseScaleDeg = 1,
                                      uint32_t level = 0, std::shared_ptr<void> params = nullptr,
                                      uint32_t slots = 0) {
        (void)params;
        const size_t n = (size_t)cc_.N() / 2;
        if (slots != 0 && slots != n) throw Error(FIDES_E_ARG, "only full packing (N/2 slots) is encoded here");
        if (value.size() > n) throw Error(FIDES_E_ARG, "more values than slots");
        if ((int)level > cc_.L()) throw Error(FIDES_E_LEVEL, "level beyond the multiplicative depth");
        if (noiseScaleDeg < 1 || noiseScaleDeg > 2) throw Error(FIDES_E_ARG, "noiseScaleDeg is 1 or 2");
        if (p_.ScalingFactorReal.empty()) throw Error(FIDES_E_STATE, "encoding needs RawParams::ScalingFactorReal");
        std::vector<double> re(n, 0.0), im(n, 0.0);
        for (size_t i = 0; i < value.size(); ++i) {
            re[i] = value[i].real();
            im[i] = value[i].imag();
        }
        const int el = cc_.L() - (int)level;
        const double sf = p_.ScalingFactorReal[el];
        Plaintext p = std::make_shared<PlaintextImpl>(self());
        p->gpu().encode(re, im, el, noiseScaleDeg == 2 ? sf * sf : sf);
        return p;
    }
    Plaintext MakeCKKSPackedPlaintext(const std::vector<double>& value, size_t noiseScaleDeg = 1, uint32_t level = 0,
                                      std::shared_ptr<void> params = nullptr, uint32_t slots = 0) {
        std::vector<std::complex<double>> v(value.begin(), value.end());
        return MakeCKKSPackedPlaintext(v, noiseScaleDeg, level, params, slots);
    }

    // ---- negation, addition, subtraction ----
    Ciphertext EvalNegate(const Ciphertext& ct) { return op1(ct, [&](auto o) { check(fides_eval_negate(h(), ct->handle(), o)); }); }
    void EvalNegateInPlace(Ciphertext& ct) { check(fides_eval_negate(h(), ct->handle(), ct->handle())); }

    Ciphertext EvalAdd(const Ciphertext& a, const Ciphertext& b) {
        return op1(a, [&](auto o) { check(fides_eval_add(h(), a->handle(), b->handle(), o)); });
    }
    Ciphertext EvalAdd(const Ciphertext& ct, Plaintext& pt) {
        return op1(ct, [&](auto o) { check(fides_eval_add_plain(h(), ct->handle(), pt->handle(), o)); });
    }
    Ciphertext EvalAdd(Plaintext& pt, const Ciphertext& ct) { return EvalAdd(ct, pt); }
    Ciphertext EvalAdd(const Ciphertext& ct, double s) {
        return op1(ct, [&](auto o) { check(fides_eval_add_scalar(h(), ct->handle(), s, o)); });
    }
    Ciphertext EvalAdd(double s, const Ciphertext& ct) { return EvalAdd(ct, s); }
    void EvalAddInPlace(Ciphertext& a, const Ciphertext& b) { check(fides_eval_add(h(), a->handle(), b->handle(), a->handle())); }
    void EvalAddInPlace(Ciphertext& ct, Plaintext& pt) { check(fides_eval_add_plain(h(), ct->handle(), pt->handle(), ct->handle())); }
    void EvalAddInPlace(Plaintext& pt, Ciphertext& ct) { EvalAddInPlace(ct, pt); }
    void EvalAddInPlace(Ciphertext& ct, double s) { check(fides_eval_add_scalar(h(), ct->handle(), s, ct->handle())); }
    void EvalAddInPlace(double s, Ciphertext& ct) { EvalAddInPlace(ct, s); }
    // The Mutable forms align their operands on private copies here, so they
    // equal the plain forms (the reference's may rescale the inputs themselves).
    Ciphertext EvalAddMutable(Ciphertext& a, Ciphertext& b) { return EvalAdd(a, b); }
    Ciphertext EvalAddMutable(Ciphertext& ct, Plaintext& pt) { return EvalAdd(ct, pt); }
    Ciphertext EvalAddMutable(Plaintext& pt, Ciphertext& ct) { return EvalAdd(ct, pt); }
    void EvalAddMutableInPlace(Ciphertext& a, Ciphertext& b) { EvalAddInPlace(a, b); }
    // EvalAddMany (api/CryptoContext.cpp:834-880): one launch
    Ciphertext EvalAddMany(const std::vector<Ciphertext>& cts) {
        if (cts.empty()) throw Error(FIDES_E_ARG, "no ciphertexts");
        Ciphertext o = fresh();
        add_many(cts, o->handle());
        return o;
    }
    void EvalAddManyInPlace(std::vector<Ciphertext>& cts) {  // the sum lands in cts[0]
        if (cts.empty()) throw Error(FIDES_E_ARG, "no ciphertexts");
        add_many(cts, cts[0]->handle());
    }

    Ciphertext EvalSub(const Ciphertext& a, const Ciphertext& b) {
        return op1(a, [&](auto o) { check(fides_eval_sub(h(), a->handle(), b->handle(), o)); });
    }
    Ciphertext EvalSub(const Ciphertext& ct, Plaintext& pt) {
        return op1(ct, [&](auto o) { check(fides_eval_sub_plain(h(), ct->handle(), pt->handle(), o)); });
    }
    Ciphertext EvalSub(Plaintext& pt, const Ciphertext& ct) {  // pt - ct = (-ct) + pt
        Ciphertext o = EvalNegate(ct);
        EvalAddInPlace(o, pt);
        return o;
    }
    Ciphertext EvalSub(const Ciphertext& ct, double s) { return EvalAdd(ct, -s); }
    Ciphertext EvalSub(double s, const Ciphertext& ct) {
        Ciphertext o = EvalNegate(ct);
        EvalAddInPlace(o, s);
        return o;
    }
    void EvalSubInPlace(Ciphertext& a, const Ciphertext& b) { check(fides_eval_sub(h(), a->handle(), b->handle(), a->handle())); }
    void EvalSubInPlace(Ciphertext& ct, double s) { EvalAddInPlace(ct, -s); }
    void EvalSubInPlace(double s, Ciphertext& ct) {
        EvalNegateInPlace(ct);
        EvalAddInPlace(ct, s);
    }
    Ciphertext EvalSubMutable(Ciphertext& a, Ciphertext& b) { return EvalSub(a, b); }
    Ciphertext EvalSubMutable(Ciphertext& ct, Plaintext& pt) { return EvalSub(ct, pt); }
    Ciphertext EvalSubMutable(Plaintext& pt, Ciphertext& ct) { return EvalSub(pt, ct); }
    void EvalSubMutableInPlace(Ciphertext& a, Ciphertext& b) { EvalSubInPlace(a, b); }

    // ---- multiplication (no rescale: the result keeps noise degree 2 under FIXEDMANUAL) ----
    Ciphertext EvalMult(const Ciphertext& a, const Ciphertext& b) {
        return op1(a, [&](auto o) { check(fides_eval_mult(h(), a->handle(), b->handle(), o, 0)); });
    }
    Ciphertext EvalMult(const Ciphertext& ct, Plaintext& pt) {
        return op1(ct, [&](auto o) { check(fides_eval_mult_plain(h(), ct->handle(), pt->handle(), o, 0)); });
    }
    Ciphertext EvalMult(Plaintext& pt, const Ciphertext& ct) { return EvalMult(ct, pt); }
    Ciphertext EvalMult(const Ciphertext& ct, double s) {
        return op1(ct, [&](auto o) { check(fides_eval_mult_scalar(h(), ct->handle(), s, o, 0)); });
    }
    Ciphertext EvalMult(double s, const Ciphertext& ct) { return EvalMult(ct, s); }
    void EvalMultInPlace(Ciphertext& ct, Plaintext& pt) {
        check(fides_eval_mult_plain(h(), ct->handle(), pt->handle(), ct->handle(), 0));
    }
    void EvalMultInPlace(Ciphertext& ct, double s) { check(fides_eval_mult_scalar(h(), ct->handle(), s, ct->handle(), 0)); }
    void EvalMultInPlace(double s, Ciphertext& ct) { EvalMultInPlace(ct, s); }
    Ciphertext EvalMultMutable(Ciphertext& a, Ciphertext& b) { return EvalMult(a, b); }
    Ciphertext EvalMultMutable(Ciphertext& ct, Plaintext& pt) { return EvalMult(ct, pt); }
    Ciphertext EvalMultMutable(Plaintext& pt, Ciphertext& ct) { return EvalMult(ct, pt); }
    void EvalMultMutableInPlace(Ciphertext& a, Ciphertext& b) {
        check(fides_eval_mult(h(), a->handle(), b->handle(), a->handle(), 0));
    }
    Ciphertext EvalSquare(const Ciphertext& ct) {
        return op1(ct, [&](auto o) { check(fides_eval_square(h(), ct->handle(), o, 0)); });
    }
    void EvalSquareInPlace(Ciphertext& ct) { check(fides_eval_square(h(), ct->handle(), ct->handle(), 0)); }
    Ciphertext EvalSquareMutable(Ciphertext& ct) { return EvalSquare(ct); }

    // ---- rotations (api/CryptoContext.cpp:1319-1509) ----
    Ciphertext EvalRotate(const Ciphertext& ct, int32_t index) {
        return op1(ct, [&](auto o) { check(fides_eval_rotate(h(), ct->handle(), index, o)); });
    }
    void EvalRotateInPlace(Ciphertext& ct, int32_t index) { check(fides_eval_rotate(h(), ct->handle(), index, ct->handle())); }
    // the GPU path has no precomputation object (api/CryptoContext.cpp:1360-1373)
    std::shared_ptr<void> EvalFastRotationPrecompute(const Ciphertext&) { return nullptr; }
    Ciphertext EvalFastRotation(const Ciphertext& ct, int32_t index, uint32_t m, const std::shared_ptr<void>& precomp) {
        (void)m;
        (void)precomp;
        return EvalRotate(ct, index);
    }
    std::vector<Ciphertext> EvalFastRotation(const Ciphertext& ct, const std::vector<int32_t>& indices, uint32_t m,
                                             const std::shared_ptr<void>& precomp) {
        (void)m;
        (void)precomp;
        return hoisted(ct, indices, false);
    }
    // results left in P*Q (rotate_hoisted(.., ext = true)); addFirst is not
    // consulted, as in the reference's GPU path (api/CryptoContext.cpp:1403-1428)
    Ciphertext EvalFastRotationExt(const Ciphertext& ct, int32_t index, const std::shared_ptr<void>& digits,
                                   bool addFirst) {
        (void)digits;
        (void)addFirst;
        return hoisted(ct, {index}, true)[0];
    }
    std::vector<Ciphertext> EvalFastRotationExt(const Ciphertext& ct, const std::vector<int32_t>& indices,
                                                const std::shared_ptr<void>& digits, bool addFirst) {
        (void)digits;
        (void)addFirst;
        return hoisted(ct, indices, true);
    }

    // ---- polynomials (OpenFHE's series p = c0/2 + sum_(k>=1) c_k T_k) ----
    Ciphertext EvalChebyshevSeries(const Ciphertext& ct, std::vector<double>& coeffs, double a, double b) {
        std::vector<double> c = half_c0(coeffs);
        return op1(ct, [&](auto o) {
            check(fides_eval_chebyshev_series(h(), ct->handle(), c.data(), (int32_t)c.size(), a, b, o));
        });
    }
    void EvalChebyshevSeriesInPlace(Ciphertext& ct, std::vector<double>& coeffs, double a, double b) {
        std::vector<double> c = half_c0(coeffs);
        check(fides_eval_chebyshev_series(h(), ct->handle(), c.data(), (int32_t)c.size(), a, b, ct->handle()));
    }
    // Chebyshev interpolation of func on [a, b] at the degree + 1 Chebyshev
    // nodes (the coefficients the series above expects; FIDESlib
    // get_chebyshev_coefficients, src/PolyApprox.cu:11-33).
    static std::vector<double> GetChebyshevCoefficients(std::function<double(double)>& func, double a, double b,
                                                        size_t degree) {
        if (degree == 0) throw Error(FIDES_E_ARG, "the degree of approximation cannot be zero");
        const size_t n = degree + 1;
        const double pi = std::acos(-1.0), half = 0.5 * (b - a), mid = 0.5 * (b + a);
        std::vector<double> f(n), c(n, 0.0);
        for (size_t j = 0; j < n; ++j) f[j] = func(std::cos(pi * (j + 0.5) / n) * half + mid);
        for (size_t k = 0; k < n; ++k) {
            double s = 0.0;
            for (size_t j = 0; j < n; ++j) s += f[j] * std::cos(pi * k * (j + 0.5) / n);
            c[k] = 2.0 * s / n;
        }
        return c;
    }

    // ---- levels ----
    Ciphertext Rescale(const Ciphertext& ct) {
        Ciphertext o = copy(ct);
        check(fides_rescale(h(), o->handle()));
        return o;
    }
    void RescaleInPlace(Ciphertext& ct) { check(fides_rescale(h(), ct->handle())); }
    // SetLevel (api/Ciphertext.cpp:99-130): dropToLevel(depth - level)
    static void SetLevel(Ciphertext& ct, size_t level) {
        CryptoContextImpl& cc = *ct->GetCryptoContext();
        if ((int)level > cc.cc_.L()) throw Error(FIDES_E_LEVEL, "level beyond the multiplicative depth");
        check(fides_drop_to_level(cc.h(), ct->handle(), cc.cc_.L() - (int)level));
    }

    // ---- bootstrapping (api/CryptoContext.cpp:1603-1647) ----
    void EvalBootstrapSetup(const std::vector<uint32_t>& levelBudget, std::vector<uint32_t> dim1 = {0, 0},
                            uint32_t slots = 0, uint32_t correctionFactor = 0, bool sparseEncaps = false) {
        cc_.EvalBootstrapSetup(levelBudget, dim1, slots, correctionFactor, sparseEncaps);
    }
    // the rotation keys EvalBootstrapKeyGen would make (plus the conjugation
    // key, automorphism 2N - 1)
    std::vector<int32_t> GetBootstrapRotationIndices() const { return cc_.BootstrapRotations(); }
    Ciphertext EvalBootstrap(const Ciphertext& ct, uint32_t numIterations = 1, uint32_t precision = 0,
                             bool prescaled = false) {
        Ciphertext o = copy(ct);
        EvalBootstrapInPlace(o, numIterations, precision, prescaled);
        return o;
    }
    void EvalBootstrapInPlace(Ciphertext& ct, uint32_t numIterations = 1, uint32_t precision = 0,
                              bool prescaled = false) {
        (void)precision;
        if (numIterations != 1) throw Error(FIDES_E_ARG, "one bootstrapping iteration only");
        if (prescaled) throw Error(FIDES_E_ARG, "prescaled inputs are not taken");
        CKKS::EvalBootstrapInPlace(ct->gpu());
    }

    // ---- sums and transforms ----
    Ciphertext AccumulateSum(const Ciphertext& ct, int slots, int stride = 1) {
        return op1(ct, [&](auto o) { check(fides_eval_accumulate_sum(h(), ct->handle(), slots, stride, o)); });
    }
    void AccumulateSumInPlace(Ciphertext& ct, int slots, int stride = 1) {
        check(fides_eval_accumulate_sum(h(), ct->handle(), slots, stride, ct->handle()));
    }
    // ConvolutionTransformInPlace (api/CryptoContext.cpp:1702-1730): rowSize 0
    // means bStep * gStep; the first rowSize plaintexts are the rows.
    void ConvolutionTransformInPlace(Ciphertext& ct, int gStep, int bStep, const std::vector<Plaintext>& pts,
                                     const std::vector<int>& indexes, int stride = 1, int rowSize = 0) {
        if (rowSize == 0) rowSize = bStep * gStep;
        if (rowSize < 1 || (size_t)rowSize > pts.size()) throw Error(FIDES_E_ARG, "fewer plaintexts than rowSize");
        if (indexes.size() != (size_t)bStep) throw Error(FIDES_E_ARG, "need one baby-step index per baby step");
        std::vector<const fides_plaintext*> hp;
        for (int i = 0; i < rowSize; ++i) hp.push_back(pts[i]->handle());
        std::vector<int32_t> ix(indexes.begin(), indexes.end());
        check(fides_eval_convolution_transform(h(), ct->handle(), hp.data(), rowSize, bStep, stride, ix.data(), gStep,
                                               ct->handle()));
    }
    static std::vector<int> GetConvolutionTransformRotationIndices(int rowSize, int bStep, int stride,
                                                                   uint32_t gStep) {
        (void)rowSize;
        (void)bStep;
        return CKKS::GetConvolutionTransformRotationIndices(stride, gStep);
    }

    void Synchronize() { cc_.Synchronize(); }

   private:
    CryptoContext self() { return shared_from_this(); }
    fides_context* h() const { return cc_.handle(); }
    Ciphertext fresh() { return std::make_shared<CiphertextImpl>(self()); }
    Ciphertext copy(const Ciphertext& ct) {
        Ciphertext o = fresh();
        o->gpu().copy(ct->gpu());
        return o;
    }
    template <class F>
    Ciphertext op1(const Ciphertext&, F f) {
        Ciphertext o = fresh();
        f(o->handle());
        return o;
    }
    void add_many(const std::vector<Ciphertext>& cts, fides_ciphertext* out) {
        std::vector<const fides_ciphertext*> hc;
        for (const Ciphertext& c : cts) hc.push_back(c->handle());
        check(fides_eval_add_many(h(), hc.data(), (int32_t)hc.size(), out));
    }
    std::vector<Ciphertext> hoisted(const Ciphertext& ct, const std::vector<int32_t>& indices, bool ext) {
        std::vector<Ciphertext> out;
        std::vector<fides_ciphertext*> ho;
        for (size_t i = 0; i < indices.size(); ++i) {
            out.push_back(fresh());
            ho.push_back(out.back()->handle());
        }
        if (ext)
            check(fides_eval_rotate_hoisted_ext(h(), ct->handle(), indices.data(), (int32_t)indices.size(), ho.data()));
        else
            check(fides_eval_rotate_hoisted(h(), ct->handle(), indices.data(), (int32_t)indices.size(), ho.data()));
        return out;
    }
    static std::vector<double> half_c0(const std::vector<double>& c) {
        std::vector<double> r(c);
        if (!r.empty()) r[0] *= 0.5;
        return r;
    }
    CKKS::RawParams p_;
    CKKS::Context cc_;
};

inline CiphertextImpl::CiphertextImpl(const CryptoContext& cc)
    : cc_(cc), ct_(std::make_unique<CKKS::Ciphertext>(cc->gpu())) {}
inline uint32_t CiphertextImpl::GetLevel() const { return (uint32_t)(cc_->GetMultiplicativeDepth() - info().level); }
inline PlaintextImpl::PlaintextImpl(const CryptoContext& cc) : cc_(cc), pt_(std::make_unique<CKKS::Plaintext>(cc->gpu())) {}

// GenCryptoContext for the GPU path: a context on one device from the raw
// parameters (GetRawParams of an OpenFHE context).
inline CryptoContext GenCryptoContext(const CKKS::RawParams& p, int device = 0) {
    return std::make_shared<CryptoContextImpl>(p, device);
}

}  // namespace api
}  // namespace fides_amd

#endif
