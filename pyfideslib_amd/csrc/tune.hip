// Autotuned launch configurations: a process-wide table of measured choices
// per (kind, logN, size range), loaded from a text file, plus forced
// configurations for the tuner itself.
//
// File format, one entry per line ('#' starts a comment):
//   <kind> <logN> <size_lo> <size_hi> <v0> <v1> <v2> <v3>
// kind: ntt_fwd | ntt_inv (size = limbs per launch; v = column tile width,
// log2 values per column thread, row tile height, grid cap) or modup |
// moddown (size = conversions per launch; v0 = target groups; modup v1 = 1:
// the matrix-core conversion, k_modup_mfma).
#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <atomic>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace fides {

namespace {
struct Entry {
    int kind, logN, lo, hi;
    TuneCfg cfg;
};
std::mutex g_mu;
std::vector<Entry> g_table;
bool g_forced[TUNE_KINDS] = {};
TuneCfg g_force[TUNE_KINDS];
std::once_flag g_default_once;
// bumped on every change of the table or of a forced configuration: captured
// launch sequences key on it, so none replays a superseded configuration
std::atomic<uint64_t> g_epoch{1};

int kind_of(const std::string& k) {
    if (k == "ntt_fwd") return TUNE_NTT_FWD;
    if (k == "ntt_inv") return TUNE_NTT_INV;
    if (k == "modup") return TUNE_MODUP;
    if (k == "moddown") return TUNE_MODDOWN;
    return -1;
}

bool valid(int kind, const TuneCfg& c) {
    if (kind == TUNE_NTT_FWD || kind == TUNE_NTT_INV) {
        auto w = [](int x) { return x == 4 || x == 8 || x == 16; };
        return w(c.v[0]) && (c.v[1] == 2 || c.v[1] == 3) && w(c.v[2]) && c.v[3] >= 0;
    }
    // conversions: target groups, and 1 in v1 for the matrix-core ModUp
    return c.v[0] >= 1 && c.v[0] <= 64 && (c.v[1] == 0 || (kind == TUNE_MODUP && c.v[1] == 1));
}
}  // namespace

uint64_t tune_epoch() { return g_epoch.load(std::memory_order_relaxed); }

bool tune_get(int kind, int logN, int size, TuneCfg& out) {
    if (kind < 0 || kind >= TUNE_KINDS) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_forced[kind]) {
        out = g_force[kind];
        return true;
    }
    for (const Entry& e : g_table)
        if (e.kind == kind && e.logN == logN && size >= e.lo && size <= e.hi) {
            out = e.cfg;
            return true;
        }
    return false;
}

void tune_force(int kind, const TuneCfg* cfg) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_epoch.fetch_add(1, std::memory_order_relaxed);
    if (kind < 0) {
        for (bool& f : g_forced) f = false;
        return;
    }
    if (kind >= TUNE_KINDS) return;
    g_forced[kind] = cfg != nullptr;
    if (cfg) g_force[kind] = *cfg;
}

int tune_load(const char* path) {
    if (!path) {  // back to the fixed rules
        std::lock_guard<std::mutex> lk(g_mu);
        g_table.clear();
        g_epoch.fetch_add(1, std::memory_order_relaxed);
        return 0;
    }
    std::ifstream f(path);
    if (!f) return -1;
    std::vector<Entry> t;
    std::string line;
    while (std::getline(f, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string k;
        Entry e;
        if (!(is >> k)) continue;
        e.kind = kind_of(k);
        if (e.kind < 0 || !(is >> e.logN >> e.lo >> e.hi)) continue;
        for (int i = 0; i < TUNE_NV; ++i)
            if (!(is >> e.cfg.v[i])) e.cfg.v[i] = 0;
        if (valid(e.kind, e.cfg) && e.lo <= e.hi) t.push_back(e);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_table = std::move(t);
    g_epoch.fetch_add(1, std::memory_order_relaxed);
    return (int)g_table.size();
}

void tune_load_default() {
    std::call_once(g_default_once, [] {
        if (const char* e = std::getenv("FIDES_TUNE_FILE")) {
            if (*e) tune_load(e);  // empty: run on the fixed rules
            return;
        }
        Dl_info info;
        if (dladdr((void*)&tune_load_default, &info) && info.dli_fname) {
            std::string p(info.dli_fname);
            const size_t s = p.rfind('/');
            p = (s == std::string::npos ? std::string(".") : p.substr(0, s)) + "/tune_gfx950.txt";
            tune_load(p.c_str());
        }
    });
}

}  // namespace fides
