// Exercises the C++ binding (include/fides_amd.hpp) the way FIDESlib's GPU
// branch would: RawParams -> Context, RawKeySwitchKey -> keys, RawCipherText
// -> Ciphertext, in-place ops, store.
//   facade_check errors      argument checks that need no GPU
//   facade_check run DIR     load DIR/{params,relin,rot1,rot2,rot3,ct1,ct2,pt}.bin,
//                            write DIR/{mult,rot,hoist,mpt,lt,srv,addmany,addalias,
//                            neg,acc,hoistext,multext,dot,subpt,conv}.bin (c0 then c1
//                            limbs)
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "fides_amd.hpp"

using namespace fides_amd;

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

static void write_ct(const std::string& path, const CKKS::RawCipherText& r) {
    std::ofstream f(path, std::ios::binary);
    for (const auto* sub : {&r.sub_0, &r.sub_1})
        for (const auto& limb : *sub) f.write(reinterpret_cast<const char*>(limb.data()), (std::streamsize)(limb.size() * 8));
}

static int errors() {
    CKKS::RawParams p;
    p.logN = 9;  // below the supported range: rejected before any device work
    p.L = 0;
    p.K = 1;
    p.moduli = {12289};
    p.root_of_unity = {3};
    p.SPECIALmoduli = {40961};
    p.SPECIALroot_of_unity = {3};
    try {
        CKKS::Context cc(p);
        std::puts("no error for logN=9");
        return 1;
    } catch (const Error& e) {
        if (e.code() != FIDES_E_ARG || std::string(e.what()).find("logN") == std::string::npos) {
            std::printf("unexpected error %d: %s\n", e.code(), e.what());
            return 1;
        }
    }
    p.logN = 13;
    p.root_of_unity.clear();  // size mismatch caught by the binding
    try {
        CKKS::Context cc(p);
        std::puts("no error for a short root_of_unity");
        return 1;
    } catch (const Error& e) {
        if (e.code() != FIDES_E_ARG) return 1;
    }
    std::puts("errors ok");
    return 0;
}

static int run(const std::string& dir) {
    // params.bin: logN, L, K, dnum, then L+1+K moduli and L+1+K roots
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
    CKKS::Context cc(p);
    const size_t N = (size_t)cc.N();
    auto key = [&](const std::string& name) {
        const std::vector<uint64_t> k = read_u64(dir + "/" + name);
        CKKS::FlatKey fk;
        fk.a.assign(k.begin(), k.begin() + k.size() / 2);
        fk.b.assign(k.begin() + k.size() / 2, k.end());
        return fk;
    };
    cc.AddEvalKey(key("relin.bin"));
    cc.AddRotationKey(1, key("rot1.bin"));
    cc.AddRotationKey(2, key("rot2.bin"));
    cc.AddRotationKey(3, key("rot3.bin"));
    auto raw = [&](const std::string& name) {
        const std::vector<uint64_t> v = read_u64(dir + "/" + name);
        CKKS::RawCipherText r;
        r.numRes = (int)(v.size() / (2 * N));
        r.N = (int)N;
        r.Noise = (double)(1ull << 40);
        r.NoiseLevel = 1;
        r.slots = (int)N / 2;
        for (int i = 0; i < r.numRes; ++i) {
            r.sub_0.emplace_back(v.begin() + i * N, v.begin() + (i + 1) * N);
            r.sub_1.emplace_back(v.begin() + (r.numRes + i) * N, v.begin() + (r.numRes + i + 1) * N);
        }
        return r;
    };
    CKKS::Ciphertext a(cc, raw("ct1.bin")), b(cc, raw("ct2.bin"));
    CKKS::Ciphertext m(cc), r(cc), h1(cc);
    m.mult(a, b, true);  // m = a * b, relinearized and rescaled
    r.rotate(a, 1);
    a.rotate_hoisted({1}, {&h1});
    CKKS::RawCipherText out;
    m.store(out);
    write_ct(dir + "/mult.bin", out);
    r.store(out);
    write_ct(dir + "/rot.bin", out);
    h1.store(out);
    write_ct(dir + "/hoist.bin", out);
    // plaintext product, a two-diagonal linear transform and a served HMult
    const std::vector<uint64_t> pv = read_u64(dir + "/pt.bin");
    CKKS::RawPlainText rp;
    rp.numRes = (int)(pv.size() / N);
    rp.N = (int)N;
    rp.Noise = (double)(1ull << 40);
    rp.NoiseLevel = 1;
    rp.slots = (int)N / 2;
    for (int i = 0; i < rp.numRes; ++i) rp.sub_0.emplace_back(pv.begin() + i * N, pv.begin() + (i + 1) * N);
    CKKS::Plaintext pt(cc, rp);
    CKKS::Ciphertext mp(cc), lt(cc), sv(cc);
    mp.multPt(a, pt);
    mp.store(out);
    write_ct(dir + "/mpt.bin", out);
    lt.copy(a);
    CKKS::LinearTransform(lt, 2, {&pt, &pt}, 1, 0);
    lt.store(out);
    write_ct(dir + "/lt.bin", out);
    CKKS::Server srv(cc, 4, 1ull << 34);
    const int64_t id = srv.submitMult(a, b, sv);
    while (srv.step() > 0) {
    }
    cc.Synchronize();
    const std::vector<int64_t> done = srv.poll();
    if (done.size() != 1 || done[0] != id) throw std::runtime_error("server did not report the job");
    sv.store(out);
    write_ct(dir + "/srv.bin", out);
    // EvalAddMany (also over an output that aliases its first input), negation
    // and AccumulateSum(4, 1) -- rotations by 1, 2, 3 hoisted
    CKKS::Ciphertext sum(cc), al(cc), ng(cc), ac(cc);
    EvalAddMany(sum, {&a, &b, &a});
    sum.store(out);
    write_ct(dir + "/addmany.bin", out);
    al.copy(a);
    EvalAddMany(al, {&al, &b});
    al.store(out);
    write_ct(dir + "/addalias.bin", out);
    ng.copy(a);
    ng.negate();
    ng.store(out);
    write_ct(dir + "/neg.bin", out);
    ac.copy(a);
    ac.accumulateSum(4, 1);
    ac.store(out);
    write_ct(dir + "/acc.bin", out);
    // the extended-basis and lazy forms: a hoisted rotation left in P*Q and
    // brought down, a product left in P*Q and brought down, subPt, and the
    // weighted sum 0.5 a - 2 b (this context has no ScalingFactorReal, so the
    // weighted sum must be refused as a state error)
    CKKS::Ciphertext x1(cc), mx(cc), sp(cc), ws(cc);
    a.rotate_hoisted({1}, {&x1}, true);
    if (!x1.isExt()) throw std::runtime_error("ext rotation output not in P*Q form");
    x1.modDown();
    x1.store(out);
    write_ct(dir + "/hoistext.bin", out);
    mx.copy(a);
    mx.mult(b, false, false);
    if (!mx.isExt()) throw std::runtime_error("mult without ModDown not in P*Q form");
    mx.modDown();
    mx.store(out);
    write_ct(dir + "/multext.bin", out);
    // ConvolutionTransform: bStep 2 (indexes 0, 1), gStep 2, three plaintexts
    if (CKKS::GetConvolutionTransformRotationIndices(1, 2) != std::vector<int>{1, 2})
        throw std::runtime_error("convolution rotation indices");
    CKKS::Ciphertext cv(cc);
    cv.copy(a);
    CKKS::ConvolutionTransform(cv, 2, {&pt, &pt, &pt}, 1, {0, 1}, 2);
    cv.store(out);
    write_ct(dir + "/conv.bin", out);
    // dotProduct: a*b + b*a with one key switch of the summed d2
    CKKS::Ciphertext dp(cc);
    dp.dotProduct({&a, &b}, {&b, &a}, false);
    dp.store(out);
    write_ct(dir + "/dot.bin", out);
    sp.copy(a);
    sp.subPt(pt);
    sp.store(out);
    write_ct(dir + "/subpt.bin", out);
    try {
        ws.evalLinearWSumMutable(2, {&a, &b}, {0.5, -2.0});
        throw std::runtime_error("weighted sum without ScalingFactorReal did not fail");
    } catch (const Error& e) {
        if (e.code() != FIDES_E_STATE) throw;
    }
    // bootstrapping entry points: this context has no ScalingFactorReal, so
    // EvalBootstrapSetup is a state error, and so is EvalBootstrapInPlace
    // before any setup
    for (int step = 0; step < 2; ++step) {
        try {
            if (step == 0)
                cc.EvalBootstrapSetup({2, 2});
            else
                CKKS::EvalBootstrapInPlace(a);
            throw std::runtime_error("bootstrap call without its prerequisite did not fail");
        } catch (const Error& e) {
            if (e.code() != FIDES_E_STATE) throw;
        }
    }
    std::printf("run ok level %d\n", m.getLevel());
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) return errors();
        if (argc >= 3 && std::strcmp(argv[1], "run") == 0) return run(argv[2]);
        std::puts("usage: facade_check errors | run DIR");
        return 2;
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
}
