// Baby-step / giant-step linear transform (FIDESlib CKKS::LinearTransform,
// src/CKKS/LinearTransform.cu:67-258; OpenFHE EvalLinearTransform), the core
// of bootstrapping's CoeffsToSlots / SlotsToCoeffs:
//
//   out = rot_offset( sum_{j<g} rot_{j b s}( sum_{i<b} A[b j + i] * rot_{i s}(ct) ) )
//
// with b baby steps, g = ceil(n / b) giant steps and stride s; the diagonals
// A[b j + i] are plaintexts the caller encodes (pre-rotated by -j b s, as
// OpenFHE's precomputation does).  The b baby-step rotations share one ModUp
// (the hoisted rotation set), every inner sum of all g giant steps runs in ONE
// fused plaintext dot-product launch, and the giant steps are applied Horner
// style as the reference does: acc = inner[g-1]; acc = rot_{b s}(acc) +
// inner[j] for j = g-2 .. 0 -- or, with batched_giant, all g-1 giant-step
// rotations as one batched key switch (one launch per stage instead of g-1
// dependent key switches; needs the keys of every j b s).  The result keeps
// noise degree 2 (no rescale).
#include <cstring>

#include "context.hpp"

namespace fides {

static void ck(hipError_t e, const char* w) { hip_check(e, w); }

// The transform of several ciphertexts with the same diagonals: baby steps
// per ciphertext (one hoisted set each), the inner sums of every giant step
// of every ciphertext in ONE dot launch, and each Horner giant step (and the
// offset) as one batched rotation over all of them.  B = 1 is
// eval_linear_transform.
void eval_linear_transform_batch(Context& c, const std::vector<const Ciphertext*>& as_in,
                                 const std::vector<const Plaintext*>& diag, int bstep, int stride, int offset,
                                 const std::vector<Ciphertext*>& outs, bool batched_giant) {
    const int B = (int)as_in.size();
    if (B < 1 || (int)outs.size() != B) throw Error(E_ARG, "batch size mismatch");
    const int n = (int)diag.size();
    if (n < 1 || bstep < 1 || bstep > n) throw Error(E_ARG, "need 1 <= bstep <= number of diagonals");
    const int g = (n + bstep - 1) / bstep;
    std::vector<std::unique_ptr<Ciphertext>> tmp(B);
    std::vector<const Ciphertext*> as(B);
    for (int k = 0; k < B; ++k) {
        const Ciphertext& a_in = *as_in[k];
        if (a_in.level < 0) throw Error(E_STATE, "ciphertext not loaded");
        as[k] = &a_in;
        if (a_in.noise_level == 2) {  // LinearTransform rescales a degree-2 input first
            tmp[k] = ct_create(c);
            ct_copy(c, a_in, *tmp[k]);
            rescale(c, *tmp[k]);
            as[k] = tmp[k].get();
        }
        if (as[k]->noise_level != 1) throw Error(E_STATE, "linear transform needs a noise level 1 input");
        if (as[k]->level != as[0]->level) throw Error(E_LEVEL, "a batched transform needs inputs at one level");
    }
    const int level = as[0]->level, nl = level + 1;
    for (const Plaintext* p : diag) {
        if (!p || p->level < level) throw Error(E_LEVEL, "diagonal missing or with fewer limbs than the input");
        if (p->noise_level != 1) throw Error(E_STATE, "diagonals must have noise level 1");
    }

    // baby steps: rot_{i s}(a), i < b, one shared ModUp per ciphertext
    std::vector<std::vector<std::unique_ptr<Ciphertext>>> baby(B);
    std::vector<int> idx;
    for (int i = 0; i < bstep; ++i) idx.push_back(i * stride);
    for (int k = 0; k < B; ++k) {
        std::vector<Ciphertext*> bp;
        for (int i = 0; i < bstep; ++i) {
            baby[k].push_back(ct_create(c));
            bp.push_back(baby[k].back().get());
        }
        eval_rotate_hoisted(c, *as[k], idx, bp);
    }

    // inner sums of every giant step of every ciphertext: one launch
    std::vector<std::vector<std::unique_ptr<Ciphertext>>> inner(B);
    std::vector<DotJob> jobs;
    std::vector<const uint64_t*> x0s, x1s, ws;
    std::vector<size_t> first;
    for (int k = 0; k < B; ++k)
        for (int j = 0; j < g; ++j) {
            inner[k].push_back(ct_create(c));
            first.push_back(ws.size());
            int terms = 0;
            for (int i = 0; i < bstep && bstep * j + i < n; ++i, ++terms) {
                x0s.push_back(baby[k][i]->c0());
                x1s.push_back(baby[k][i]->c1());
                ws.push_back(diag[bstep * j + i]->c0());
            }
            jobs.push_back(DotJob{nullptr, nullptr, nullptr, inner[k][j]->c0(), inner[k][j]->c1(), terms, nl});
        }
    const size_t jb = jobs.size() * sizeof(DotJob), pb = ws.size() * sizeof(uint64_t*);
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_x0 = up(jb), o_x1 = up(o_x0 + pb), o_w = up(o_x1 + pb);
    DevBuf table(&c.pool, (o_w + pb + 7) / 8);
    char* d = (char*)table.ptr;
    for (size_t q = 0; q < jobs.size(); ++q) {
        jobs[q].x0 = (const uint64_t* const*)(d + o_x0) + first[q];
        jobs[q].x1 = (const uint64_t* const*)(d + o_x1) + first[q];
        jobs[q].w = (const uint64_t* const*)(d + o_w) + first[q];
    }
    std::vector<char> host(o_w + pb);
    std::memcpy(host.data(), jobs.data(), jb);
    std::memcpy(host.data() + o_x0, x0s.data(), pb);
    std::memcpy(host.data() + o_x1, x1s.data(), pb);
    std::memcpy(host.data() + o_w, ws.data(), pb);
    c.upload(table.ptr, host.data(), host.size());
    DotJobs L;
    std::memset(&L, 0, sizeof(L));
    L.dev = (const DotJob*)d;
    L.n = (int)jobs.size();
    ck(launch_dot_plain(c.T, L, nl, c.stream), "linear transform inner sums");
    for (int k = 0; k < B; ++k)
        for (int j = 0; j < g; ++j) {
            Ciphertext& o = *inner[k][j];
            o.level = level;
            o.noise = as[k]->noise * diag[bstep * j]->noise;
            o.noise_level = 2;
            o.slots = as[k]->slots;
        }

    // a rotation of every accumulator by r (one batched key switch when B > 1)
    auto rotate_all = [&](const std::vector<Ciphertext*>& src, int r, const std::vector<Ciphertext*>& dst) {
        if (B == 1)
            eval_rotate(c, *src[0], r, *dst[0]);
        else
            eval_rotate_batch(c, std::vector<const Ciphertext*>(src.begin(), src.end()), std::vector<int>(B, r),
                              dst);
    };
    const int half = c.N / 2;
    std::vector<Ciphertext*> acc(B);
    for (int k = 0; k < B; ++k) acc[k] = inner[k][0].get();
    if (batched_giant) {
        // every giant step rotated in ONE batched key switch (keys of all
        // j b s), then summed: sum_j rot_{j b s}(inner[j])
        std::vector<std::unique_ptr<Ciphertext>> rots;
        std::vector<const Ciphertext*> ins;
        std::vector<Ciphertext*> routs;
        std::vector<int> gi;
        for (int k = 0; k < B; ++k)
            for (int j = 1; j < g; ++j) {
                rots.push_back(ct_create(c));
                ins.push_back(inner[k][j].get());
                routs.push_back(rots.back().get());
                gi.push_back(j * bstep * stride);
            }
        eval_rotate_batch(c, ins, gi, routs);
        for (int k = 0; k < B; ++k)
            for (int j = 1; j < g; ++j) eval_add(c, *acc[k], *rots[(size_t)k * (g - 1) + j - 1], *acc[k]);
    } else {
        // Horner order, as the reference: acc = rot_{b s}(acc) + inner[j]
        std::vector<std::unique_ptr<Ciphertext>> rot(B);
        std::vector<Ciphertext*> rp(B);
        for (int k = 0; k < B; ++k) {
            rot[k] = ct_create(c);
            rp[k] = rot[k].get();
            acc[k] = inner[k][g - 1].get();
        }
        for (int j = g - 2; j >= 0; --j) {
            if (((int64_t)bstep * stride) % half != 0) {
                rotate_all(acc, bstep * stride, rp);
                for (int k = 0; k < B; ++k) eval_add(c, *inner[k][j], *rot[k], *inner[k][j]);
            } else {
                for (int k = 0; k < B; ++k) eval_add(c, *inner[k][j], *acc[k], *inner[k][j]);
            }
            for (int k = 0; k < B; ++k) acc[k] = inner[k][j].get();
        }
    }
    if (offset % half != 0)
        rotate_all(acc, offset, outs);
    else
        for (int k = 0; k < B; ++k) ct_copy(c, *acc[k], *outs[k]);
}

// AccumulateSum (api/CryptoContext.cpp:1649-1700; FIDESlib CKKS::Accumulate,
// src/CKKS/AccumulateBroadcast.cu:41-67): out = sum_(k < size) rot_(k stride)(a)
// in radix-4 steps -- for s = 1, 4, 16, .. < size one hoisted set of the
// rotations by i s stride (i = 1..3, below size stride) is added into the
// running sum.  The reference adds the set in the extended basis before one
// ModDown; here each rotation of the set is finished (ModDown, + c0) first.
std::vector<int> accumulate_rotations(int size, int stride) {
    if (size < 1 || (size & (size - 1)) || stride < 1) throw Error(E_ARG, "size must be a power of two, stride >= 1");
    std::vector<int> idx;
    for (int64_t s = 1; s < size; s *= 4)
        for (int64_t i = 1; i < 4 && i * s < size; ++i) idx.push_back((int)(i * s * stride));
    return idx;
}

void eval_accumulate_sum(Context& c, const Ciphertext& a, int size, int stride, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    (void)accumulate_rotations(size, stride);  // argument checks
    const int in_slots = a.slots;
    std::unique_ptr<Ciphertext> acc = ct_create(c);
    ct_copy(c, a, *acc);
    for (int64_t s = 1; s < size; s *= 4) {
        std::vector<int> idx;
        for (int64_t i = 1; i < 4 && i * s < size; ++i) idx.push_back((int)(i * s * stride));
        std::vector<std::unique_ptr<Ciphertext>> r;
        std::vector<Ciphertext*> rp;
        for (size_t i = 0; i < idx.size(); ++i) {
            r.push_back(ct_create(c));
            rp.push_back(r.back().get());
        }
        eval_rotate_hoisted(c, *acc, idx, rp);
        for (auto& x : r) eval_add(c, *acc, *x, *acc);
    }
    ct_copy(c, *acc, out);
    // a sum over every slot leaves `stride` distinct slot values
    // (AccumulateBroadcast.cu:64-65)
    if ((int64_t)size * stride == in_slots) out.slots = stride;
}

void eval_linear_transform(Context& c, const Ciphertext& a, const std::vector<const Plaintext*>& diag, int bstep,
                           int stride, int offset, Ciphertext& out, bool batched_giant) {
    eval_linear_transform_batch(c, {&a}, diag, bstep, stride, offset, {&out}, batched_giant);
}

}  // namespace fides
