// Serialization of the engine's device objects and of FIDESlib's device
// sidecar.
//
// FIDESlib serializes the OpenFHE objects themselves through cereal
// (api/Serialize.cpp:14-65, :95-190, OpenFHE's Serial::SerializeToFile) and
// writes its own GPU state beside them in a text sidecar `<file>.dev`:
//
//   <n> { <device> ... }
//   AutoLoadCiphertexts: <0|1>
//   AutoLoadPlaintexts: <0|1>
//   RotationIndexes: { <index> ... }
//   KeyDist: <SecretKeyDist>
//
// The sidecar is written and read here in exactly that text form.  The cereal
// streams belong to OpenFHE, which this engine does not link; what crosses the
// drop-in boundary are the raw RNS objects FIDESlib extracts from them
// (RawParams, RawCipherText, RawPlainText, RawKeySwitchKey), and those are
// saved in the engine's own little-endian binary form:
//
//   "FIDESAMD" u32 version u32 kind, then per kind
//   params      i32 logN L K dnum technique; u64 moduli[L+1+K], roots[L+1+K];
//               u32 present flags; f64 scaling_factor_real[L+1],
//               scaling_factor_real_big[L+1], mod_reduce_factor[L+1] when present
//   ciphertext  i32 level noise_level slots ext; f64 noise; u32 N nlimbs;
//               u64 moduli[nlimbs]; c0 [nlimbs][N]; c1 [nlimbs][N]
//   plaintext   the same with one part
//   key         i32 key kind, index; u32 N dnum npos; u64 moduli[npos];
//               A [dnum][npos][N]; B [dnum][npos][N]
//
// Every load checks the moduli against the context's, so a file only loads
// into a context of the same parameters.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "context.hpp"

namespace fides {

namespace {

constexpr char MAGIC[8] = {'F', 'I', 'D', 'E', 'S', 'A', 'M', 'D'};
constexpr uint32_t VERSION = 1;

struct Out {
    std::ofstream f;
    explicit Out(const std::string& path) : f(path, std::ios::binary | std::ios::trunc) {
        if (!f) throw Error(E_ARG, "cannot open " + path + " for writing");
    }
    void raw(const void* p, size_t n) {
        f.write((const char*)p, (std::streamsize)n);
        if (!f) throw Error(E_ARG, "write failed");
    }
    template <class T>
    void put(T v) {
        raw(&v, sizeof(v));
    }
    void header(uint32_t kind) {
        raw(MAGIC, 8);
        put<uint32_t>(VERSION);
        put<uint32_t>(kind);
    }
};

struct In {
    std::ifstream f;
    explicit In(const std::string& path) : f(path, std::ios::binary) {
        if (!f) throw Error(E_ARG, "cannot open " + path);
    }
    void raw(void* p, size_t n) {
        f.read((char*)p, (std::streamsize)n);
        if (!f) throw Error(E_ARG, "truncated file");
    }
    template <class T>
    T get() {
        T v;
        raw(&v, sizeof(v));
        return v;
    }
    void header(uint32_t kind) {
        char m[8];
        raw(m, 8);
        if (std::memcmp(m, MAGIC, 8) != 0) throw Error(E_ARG, "not an engine object file");
        if (get<uint32_t>() != VERSION) throw Error(E_ARG, "unsupported file version");
        if (get<uint32_t>() != kind) throw Error(E_ARG, "file holds another kind of object");
    }
};

// prime of limb i of a polynomial with nq Q limbs (then the special limbs)
uint64_t limb_prime(const Context& c, int i, int nq) { return c.q[i < nq ? i : c.L + 1 + (i - nq)]; }

void check_moduli(In& in, const Context& c, int nlimbs, int nq) {
    for (int i = 0; i < nlimbs; ++i)
        if (in.get<uint64_t>() != limb_prime(c, i, nq))
            throw Error(E_ARG, "the file's moduli differ from the context's");
}

void d2h(Context& c, void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream), "device to host");
    c.sync();
}

}  // namespace

void params_save(const Context::Params& p, const std::string& path) {
    Out o(path);
    o.header(1);
    for (int v : {p.logN, p.L, p.K, p.dnum, p.scaling_technique}) o.put<int32_t>(v);
    const size_t np = (size_t)p.L + 1 + p.K;
    if (p.moduli.size() != np || p.roots.size() != np) throw Error(E_ARG, "moduli / roots size mismatch");
    o.raw(p.moduli.data(), np * 8);
    o.raw(p.roots.data(), np * 8);
    const uint32_t flags = (!p.scaling_factor_real.empty() ? 1u : 0u) | (!p.scaling_factor_real_big.empty() ? 2u : 0u) |
                           (!p.mod_reduce_factor.empty() ? 4u : 0u);
    o.put<uint32_t>(flags);
    for (const auto* v : {&p.scaling_factor_real, &p.scaling_factor_real_big, &p.mod_reduce_factor})
        if (!v->empty()) {
            if (v->size() != (size_t)p.L + 1) throw Error(E_ARG, "per-level table size mismatch");
            o.raw(v->data(), v->size() * 8);
        }
}

Context::Params params_load(const std::string& path) {
    In in(path);
    in.header(1);
    Context::Params p;
    p.logN = in.get<int32_t>();
    p.L = in.get<int32_t>();
    p.K = in.get<int32_t>();
    p.dnum = in.get<int32_t>();
    p.scaling_technique = in.get<int32_t>();
    if (p.logN < 1 || p.logN > 20 || p.L < 0 || p.L > 1000 || p.K < 1 || p.K > 1000)
        throw Error(E_ARG, "implausible parameters in file");
    const size_t np = (size_t)p.L + 1 + p.K;
    p.moduli.resize(np);
    p.roots.resize(np);
    in.raw(p.moduli.data(), np * 8);
    in.raw(p.roots.data(), np * 8);
    const uint32_t flags = in.get<uint32_t>();
    auto vec = [&](std::vector<double>& v, uint32_t bit) {
        if (flags & bit) {
            v.resize(p.L + 1);
            in.raw(v.data(), v.size() * 8);
        }
    };
    vec(p.scaling_factor_real, 1);
    vec(p.scaling_factor_real_big, 2);
    vec(p.mod_reduce_factor, 4);
    return p;
}

void ct_save(Context& c, const Ciphertext& ct, const std::string& path) {
    if (ct.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    const int nq = ct.level + 1, nl = nq + (ct.ext ? c.K : 0);
    std::vector<uint64_t> h((size_t)nl * c.N);
    Out o(path);
    o.header(2);
    for (int v : {ct.level, ct.noise_level, ct.slots, ct.ext}) o.put<int32_t>(v);
    o.put<double>(ct.noise);
    o.put<uint32_t>(c.N);
    o.put<uint32_t>(nl);
    for (int i = 0; i < nl; ++i) o.put<uint64_t>(limb_prime(c, i, nq));
    for (const uint64_t* part : {ct.c0(), ct.c1()}) {
        d2h(c, h.data(), part, h.size() * 8);
        o.raw(h.data(), h.size() * 8);
    }
}

void ct_load_file(Context& c, Ciphertext& ct, const std::string& path) {
    In in(path);
    in.header(2);
    const int level = in.get<int32_t>(), noise_level = in.get<int32_t>(), slots = in.get<int32_t>(),
              ext = in.get<int32_t>();
    const double noise = in.get<double>();
    const uint32_t N = in.get<uint32_t>(), nl = in.get<uint32_t>();
    if (N != c.N) throw Error(E_ARG, "ring dimension differs from the context's");
    if (level < 0 || level > c.L || (ext != 0 && ext != 1)) throw Error(E_ARG, "bad ciphertext header");
    const int nq = level + 1;
    if ((int)nl != nq + (ext ? c.K : 0)) throw Error(E_ARG, "limb count does not match the level");
    check_moduli(in, c, (int)nl, nq);
    std::vector<uint64_t> h0((size_t)nl * N), h1((size_t)nl * N);
    in.raw(h0.data(), h0.size() * 8);
    in.raw(h1.data(), h1.size() * 8);
    if (ext) ct_reserve_ext(c, ct);
    hip_check(hipMemcpyAsync(ct.c0(), h0.data(), h0.size() * 8, hipMemcpyHostToDevice, c.stream), "load c0");
    hip_check(hipMemcpyAsync(ct.c1(), h1.data(), h1.size() * 8, hipMemcpyHostToDevice, c.stream), "load c1");
    c.sync();
    ct.level = level;
    ct.ext = ext;
    ct.noise = noise;
    ct.noise_level = noise_level;
    ct.slots = slots;
}

void pt_save(Context& c, const Plaintext& pt, const std::string& path) {
    if (pt.level < 0) throw Error(E_STATE, "plaintext not loaded");
    const int nq = pt.level + 1, nl = nq + (pt.ext ? c.K : 0);
    std::vector<uint64_t> h((size_t)nl * c.N);
    d2h(c, h.data(), pt.c0(), h.size() * 8);
    Out o(path);
    o.header(3);
    for (int v : {pt.level, pt.noise_level, pt.slots, pt.ext}) o.put<int32_t>(v);
    o.put<double>(pt.noise);
    o.put<uint32_t>(c.N);
    o.put<uint32_t>(nl);
    for (int i = 0; i < nl; ++i) o.put<uint64_t>(limb_prime(c, i, nq));
    o.raw(h.data(), h.size() * 8);
}

void pt_load_file(Context& c, Plaintext& pt, const std::string& path) {
    In in(path);
    in.header(3);
    const int level = in.get<int32_t>(), noise_level = in.get<int32_t>(), slots = in.get<int32_t>(),
              ext = in.get<int32_t>();
    const double noise = in.get<double>();
    const uint32_t N = in.get<uint32_t>(), nl = in.get<uint32_t>();
    if (N != c.N) throw Error(E_ARG, "ring dimension differs from the context's");
    if (level < 0 || level > c.L || (ext != 0 && ext != 1)) throw Error(E_ARG, "bad plaintext header");
    const int nq = level + 1;
    if ((int)nl != nq + (ext ? c.K : 0)) throw Error(E_ARG, "limb count does not match the level");
    check_moduli(in, c, (int)nl, nq);
    std::vector<uint64_t> h((size_t)nl * N);
    in.raw(h.data(), h.size() * 8);
    if (ext)
        pt_load_ext(c, pt, h.data(), nq, noise, noise_level, slots, false);
    else
        pt_load(c, pt, h.data(), nq, noise, noise_level, slots, false);
}

static const KeySwitchingKey& key_of(Context& c, int kind, int index) {
    const KeySwitchingKey* k = nullptr;
    if (kind == 0) {
        k = c.eval_key.get();
    } else if (kind == 1) {
        auto it = c.rot_keys.find((uint32_t)index);
        if (it != c.rot_keys.end()) k = it->second.get();
    } else if (kind == 2) {
        auto it = c.switch_keys.find(index);
        if (it != c.switch_keys.end()) k = it->second.get();
    } else {
        throw Error(E_ARG, "key kind must be 0 (relinearization), 1 (automorphism) or 2 (switching)");
    }
    if (!k) throw Error(E_NOKEY, "no such key loaded");
    return *k;
}

void key_save(Context& c, int kind, int index, const std::string& path) {
    const KeySwitchingKey& k = key_of(c, kind, index);
    const int npos = c.L + 1 + c.K;
    std::vector<uint64_t> h((size_t)c.dnum * npos * c.N);
    Out o(path);
    o.header(4);
    o.put<int32_t>(kind);
    o.put<int32_t>(index);
    o.put<uint32_t>(c.N);
    o.put<uint32_t>(c.dnum);
    o.put<uint32_t>(npos);
    for (int i = 0; i < npos; ++i) o.put<uint64_t>(c.q[i]);
    for (const uint64_t* part : {k.a.ptr, k.b.ptr}) {
        d2h(c, h.data(), part, h.size() * 8);
        o.raw(h.data(), h.size() * 8);
    }
}

// Loads the key into the context under the kind and index the file records;
// returns them through kind / index.
void key_load_file(Context& c, const std::string& path, int& kind, int& index) {
    In in(path);
    in.header(4);
    kind = in.get<int32_t>();
    index = in.get<int32_t>();
    const uint32_t N = in.get<uint32_t>(), dnum = in.get<uint32_t>(), npos = in.get<uint32_t>();
    if (N != c.N || (int)dnum != c.dnum || (int)npos != c.L + 1 + c.K) throw Error(E_ARG, "key shape differs");
    check_moduli(in, c, (int)npos, (int)npos);
    if (kind < 0 || kind > 2) throw Error(E_ARG, "bad key kind");
    if (kind == 1 && ((index & 1) == 0 || (uint32_t)index >= 2u * c.N)) throw Error(E_ARG, "bad automorphism index");
    std::vector<uint64_t> a((size_t)dnum * npos * N), b(a.size());
    in.raw(a.data(), a.size() * 8);
    in.raw(b.data(), b.size() * 8);
    auto k = std::make_unique<KeySwitchingKey>();
    load_key(c, *k, a.data(), b.data(), false);
    c.drop_plans();  // captured launch sequences may hold the replaced key
    if (kind == 0)
        c.eval_key = std::move(k);
    else if (kind == 1)
        c.rot_keys[(uint32_t)index] = std::move(k);
    else
        c.switch_keys[index] = std::move(k);
}

// ------------------------------------------------------ the .dev sidecar ----
// SerializeToFile / DeserializeFromFile(CryptoContext) (api/Serialize.cpp:35-65,
// :117-190): the same five lines, written and parsed the same way.
void dev_sidecar_write(const std::string& path, const DevSidecar& s) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw Error(E_ARG, "cannot open " + path + " for writing");
    std::string t = std::to_string(s.devices.size()) + " { ";
    for (int d : s.devices) t += std::to_string(d) + " ";
    t += "}\n";
    t += "AutoLoadCiphertexts: " + std::to_string(s.auto_load_ciphertexts ? 1 : 0) + "\n";
    t += "AutoLoadPlaintexts: " + std::to_string(s.auto_load_plaintexts ? 1 : 0) + "\n";
    t += "RotationIndexes: { ";
    for (int r : s.rotation_indexes) t += std::to_string(r) + " ";
    t += "}\n";
    t += "KeyDist: " + std::to_string(s.key_dist) + "\n";
    f.write(t.data(), (std::streamsize)t.size());
    if (!f) throw Error(E_ARG, "write failed");
}

// Missing lines keep the defaults, as the reference's reader does (a
// context's auto_load_ciphertexts = true, auto_load_plaintexts = false,
// keyDist = UNIFORM_TERNARY = 0; api/CryptoContext.hpp:208-218).
DevSidecar dev_sidecar_read(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(E_ARG, "cannot open " + path);
    DevSidecar s;
    std::string line;
    if (std::getline(f, line)) {
        std::istringstream is(line);
        size_t n = 0;
        char brace;
        is >> n >> brace;
        if (n > 4096) throw Error(E_ARG, "implausible device count");
        for (size_t i = 0; i < n; ++i) {
            int d;
            if (!(is >> d)) throw Error(E_ARG, "truncated device list");
            s.devices.push_back(d);
        }
    }
    auto flag = [&](bool& dst) {
        if (std::getline(f, line)) {
            std::istringstream is(line);
            std::string label;
            int v = 0;
            is >> label >> v;
            dst = v != 0;
        }
    };
    flag(s.auto_load_ciphertexts);
    flag(s.auto_load_plaintexts);
    if (std::getline(f, line)) {
        std::istringstream is(line);
        std::string label;
        char brace;
        is >> label >> brace;
        int r;
        while (is >> r) s.rotation_indexes.push_back(r);
    }
    if (std::getline(f, line)) {
        std::istringstream is(line);
        std::string label;
        int v = 0;
        is >> label >> v;
        s.key_dist = v;
    }
    return s;
}

}  // namespace fides
