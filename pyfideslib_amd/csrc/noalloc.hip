// ModUp with a static device workspace: no allocation on any call.
//
// stage_modup (ckks.hip) allocates its (level+1)*N-word workspace t -- the
// digit limbs brought to coefficient form by the INTT, which the conversion
// reads -- from the context's pool.  Callers that cannot allocate device
// memory (code under stream capture, tight serving loops, hosts that forbid
// hipMalloc after start-up) use stage_modup_noalloc instead: t lives in the
// file-scope array `scratch` below, one copy per GPU, held from the library's
// first use on that GPU until the process ends.
//
// The array's host-side name is only a stand-in for the symbol, never a
// device address: noalloc_init resolves the device address of this GPU's copy
// once, with the device current (from the Context constructor), and every
// later use goes through that pointer.  All calls on one GPU share the copy,
// so calls that overlap on one GPU (two contexts on one device, used from two
// threads at once) race; calls on one context are serialised by its API lock.
// Every limb of t is written by the INTT before the conversion reads it, so
// nothing has to be cleared between calls.
#include <mutex>

#include "context.hpp"

namespace fides {

// 2^21 words = 16 MiB: level+1 <= 32 limbs at N = 2^16, 16 at N = 2^17.
static constexpr size_t NOALLOC_SCRATCH_WORDS = size_t(1) << 21;
static constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_WORDS];

static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

void noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Error(E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Error(E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = static_cast<uint64_t*>(p);
}

void stage_modup_noalloc(Context& c, int level, const uint64_t* x, uint64_t* ext) {
    if (level < 0 || level > c.L) throw Error(E_LEVEL, "level out of range");
    const size_t need = (size_t)(level + 1) * c.N;
    if (need > NOALLOC_SCRATCH_WORDS)
        throw Error(E_ARG, "ModUp workspace of (level+1)*N = " + std::to_string(need) +
                               " words exceeds the static scratch of 2^21 words (16 MiB)");
    uint64_t* t = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c.device >= 0 && c.device < NOALLOC_MAX_DEVICES) t = g_scratch_ptr[c.device];
    }
    if (!t) throw Error(E_STATE, "static scratch not initialised on this device");
    stage_modup_into(c, level, x, t, ext);
}

}  // namespace fides
