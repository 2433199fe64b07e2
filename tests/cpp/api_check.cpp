// Exercises the api-level binding (include/fides_amd_api.hpp) the way code
// written against FIDESlib's api/CryptoContext.hpp calls it: shared_ptr
// ciphertexts, new results or *InPlace, OpenFHE levels.
//   api_check run DIR   load DIR/{params,relin,rot1,rot2,rot3,ct1,ct2,pt}.bin,
//                       write DIR/api_{mult,resc,rot,fr2,fr3,add,sub,neg,sq,mpt,
//                       many,acc,setl,enc,csub}.bin (c0 then c1 limbs)
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "fides_amd_api.hpp"

using namespace fides_amd;
using namespace fides_amd::api;

static std::vector<uint64_t> read_u64(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    f.seekg(0, std::ios::end);
    const size_t n = (size_t)f.tellg() / 8;
    f.seekg(0);
    std::vector<uint64_t> v(n);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * 8));
    return v;
}

static void write(const CryptoContext& cc, const std::string& path, const Ciphertext& ct) {
    CKKS::RawCipherText r;
    cc->StoreCiphertext(ct, r);
    std::ofstream f(path, std::ios::binary);
    for (const auto* part : {&r.sub_0, &r.sub_1})
        for (const auto& limb : *part) f.write(reinterpret_cast<const char*>(limb.data()), (std::streamsize)(limb.size() * 8));
}

static void expect(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

static int run(const std::string& dir) {
    const std::vector<uint64_t> h = read_u64(dir + "/params.bin");
    CKKS::RawParams p;
    p.logN = (int)h[0];
    p.L = (int)h[1];
    p.K = (int)h[2];
    p.dnum = (int)h[3];
    const int np = p.L + 1 + p.K;
    p.moduli.assign(h.begin() + 4, h.begin() + 4 + p.L + 1);
    p.SPECIALmoduli.assign(h.begin() + 4 + p.L + 1, h.begin() + 4 + np);
    p.root_of_unity.assign(h.begin() + 4 + np, h.begin() + 4 + np + p.L + 1);
    p.SPECIALroot_of_unity.assign(h.begin() + 4 + np + p.L + 1, h.begin() + 4 + 2 * np);
    p.ScalingFactorReal.assign(p.L + 1, std::ldexp(1.0, 40));  // FIXEDMANUAL at 2^40
    CryptoContext cc = GenCryptoContext(p);
    const size_t N = cc->GetRingDimension();
    expect(cc->GetCyclotomicOrder() == 2 * N, "cyclotomic order");
    auto key = [&](const std::string& name) {
        const std::vector<uint64_t> k = read_u64(dir + "/" + name);
        CKKS::FlatKey fk;
        fk.a.assign(k.begin(), k.begin() + k.size() / 2);
        fk.b.assign(k.begin() + k.size() / 2, k.end());
        return fk;
    };
    cc->AddEvalKey(key("relin.bin"));
    cc->AddRotationKey(1, key("rot1.bin"));
    cc->AddRotationKey(2, key("rot2.bin"));
    cc->AddRotationKey(3, key("rot3.bin"));
    auto raw = [&](const std::string& name) {
        const std::vector<uint64_t> v = read_u64(dir + "/" + name);
        CKKS::RawCipherText r;
        r.numRes = (int)(v.size() / (2 * N));
        r.N = (int)N;
        r.Noise = std::ldexp(1.0, 40);
        r.NoiseLevel = 1;
        r.slots = (int)N / 2;
        for (int i = 0; i < r.numRes; ++i) {
            r.sub_0.emplace_back(v.begin() + i * N, v.begin() + (i + 1) * N);
            r.sub_1.emplace_back(v.begin() + (r.numRes + i) * N, v.begin() + (r.numRes + i + 1) * N);
        }
        return r;
    };
    Ciphertext a = cc->LoadCiphertext(raw("ct1.bin")), b = cc->LoadCiphertext(raw("ct2.bin"));
    expect(a->GetLevel() == 0 && a->GetNoiseScaleDeg() == 1, "fresh ciphertext level");

    Ciphertext m = cc->EvalMult(a, b);
    expect(m->GetNoiseScaleDeg() == 2 && m->GetLevel() == 0, "EvalMult leaves degree 2");
    write(cc, dir + "/api_mult.bin", m);
    Ciphertext rs = cc->Rescale(m);
    expect(rs->GetLevel() == 1 && rs->GetNoiseScaleDeg() == 1 && m->GetLevel() == 0, "Rescale returns a new one");
    write(cc, dir + "/api_resc.bin", rs);
    write(cc, dir + "/api_rot.bin", cc->EvalRotate(a, 1));
    std::vector<Ciphertext> fr = cc->EvalFastRotation(a, {2, 3}, 2 * (uint32_t)N, cc->EvalFastRotationPrecompute(a));
    write(cc, dir + "/api_fr2.bin", fr[0]);
    write(cc, dir + "/api_fr3.bin", fr[1]);
    write(cc, dir + "/api_add.bin", cc->EvalAdd(a, b));
    write(cc, dir + "/api_sub.bin", cc->EvalSub(a, b));
    write(cc, dir + "/api_neg.bin", cc->EvalNegate(a));
    write(cc, dir + "/api_sq.bin", cc->EvalSquare(a));

    const std::vector<uint64_t> pv = read_u64(dir + "/pt.bin");
    CKKS::RawPlainText rp;
    rp.numRes = (int)(pv.size() / N);
    rp.N = (int)N;
    rp.Noise = std::ldexp(1.0, 40);
    rp.NoiseLevel = 1;
    rp.slots = (int)N / 2;
    for (int i = 0; i < rp.numRes; ++i) rp.sub_0.emplace_back(pv.begin() + i * N, pv.begin() + (i + 1) * N);
    Plaintext pt = cc->LoadPlaintext(rp);
    write(cc, dir + "/api_mpt.bin", cc->EvalMult(a, pt));
    write(cc, dir + "/api_many.bin", cc->EvalAddMany({a, b, a}));
    write(cc, dir + "/api_acc.bin", cc->AccumulateSum(a, 4, 1));
    Ciphertext sl = cc->EvalAdd(a, b);
    CryptoContextImpl::SetLevel(sl, 2);
    expect(sl->GetLevel() == 2, "SetLevel");
    write(cc, dir + "/api_setl.bin", sl);

    // encoding: exact dyadic slot values, complex, OpenFHE level 0
    std::vector<std::complex<double>> v(N / 2);
    for (size_t i = 0; i < N / 2; ++i)
        v[i] = {(double)((long)((i * 37) % 201) - 100) / 128.0, (double)((long)((i * 11) % 97) - 48) / 64.0};
    Plaintext enc = cc->MakeCKKSPackedPlaintext(v);
    write(cc, dir + "/api_enc.bin", cc->EvalMult(a, enc));
    // scalar - ciphertext = negation plus the scalar
    Ciphertext cs = cc->EvalSub(0.25, a), want = cc->EvalNegate(a);
    cc->EvalAddInPlace(want, 0.25);
    write(cc, dir + "/api_csub.bin", cs);
    CKKS::RawCipherText x, y;
    cc->StoreCiphertext(cs, x);
    cc->StoreCiphertext(want, y);
    expect(x.sub_0 == y.sub_0 && x.sub_1 == y.sub_1, "EvalSub(double, ct)");

    // Chebyshev: x^2 on [-1, 1] is 1/2 T0 + 1/2 T2, so OpenFHE's coefficients
    // (c0 halved by the series) are {1, 0, 1/2}
    std::function<double(double)> sq = [](double t) { return t * t; };
    std::vector<double> c = CryptoContextImpl::GetChebyshevCoefficients(sq, -1.0, 1.0, 2);
    expect(c.size() == 3 && std::fabs(c[0] - 1.0) < 1e-12 && std::fabs(c[1]) < 1e-12 && std::fabs(c[2] - 0.5) < 1e-12,
           "Chebyshev coefficients of x^2");
    // the series with OpenFHE's convention equals the engine's series with c0 halved
    std::vector<double> co{0.3, -0.2, 0.1, 0.05};
    Ciphertext s1 = cc->EvalChebyshevSeries(a, co, -1.0, 1.0);
    Ciphertext s2 = cc->LoadCiphertext(raw("ct1.bin"));
    CKKS::evalChebyshevSeries(s2->gpu(), {0.15, -0.2, 0.1, 0.05}, -1.0, 1.0);
    cc->StoreCiphertext(s1, x);
    cc->StoreCiphertext(s2, y);
    expect(x.sub_0 == y.sub_0 && x.sub_1 == y.sub_1 && s1->GetLevel() == s2->GetLevel(), "Chebyshev c0 convention");

    // refusals keep their codes
    try {
        cc->EvalBootstrap(a, 2);
        throw std::runtime_error("two bootstrapping iterations did not fail");
    } catch (const Error& e) {
        if (e.code() != FIDES_E_ARG) throw;
    }
    try {
        cc->MakeCKKSPackedPlaintext(std::vector<double>(N), 1, 0);
        throw std::runtime_error("more values than slots did not fail");
    } catch (const Error& e) {
        if (e.code() != FIDES_E_ARG) throw;
    }
    cc->Synchronize();
    std::puts("api ok");
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 3 && std::strcmp(argv[1], "run") == 0) return run(argv[2]);
        std::puts("usage: api_check run DIR");
        return 2;
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
}
