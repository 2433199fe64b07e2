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

void eval_linear_transform(Context& c, const Ciphertext& a_in, const std::vector<const Plaintext*>& diag, int bstep,
                           int stride, int offset, Ciphertext& out, bool batched_giant) {
    if (a_in.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    const int n = (int)diag.size();
    if (n < 1 || bstep < 1 || bstep > n) throw Error(E_ARG, "need 1 <= bstep <= number of diagonals");
    const int g = (n + bstep - 1) / bstep;
    std::unique_ptr<Ciphertext> tmp;
    const Ciphertext* a = &a_in;
    if (a_in.noise_level == 2) {  // LinearTransform rescales a degree-2 input first
        tmp = ct_create(c);
        ct_copy(c, a_in, *tmp);
        rescale(c, *tmp);
        a = tmp.get();
    }
    if (a->noise_level != 1) throw Error(E_STATE, "linear transform needs a noise level 1 input");
    const int level = a->level, nl = level + 1;
    for (const Plaintext* p : diag) {
        if (!p || p->level < level) throw Error(E_LEVEL, "diagonal missing or with fewer limbs than the input");
        if (p->noise_level != 1) throw Error(E_STATE, "diagonals must have noise level 1");
    }

    // baby steps: rot_{i s}(a), i < b, one shared ModUp
    std::vector<std::unique_ptr<Ciphertext>> baby;
    std::vector<Ciphertext*> bp;
    std::vector<int> idx;
    for (int i = 0; i < bstep; ++i) {
        baby.push_back(ct_create(c));
        bp.push_back(baby.back().get());
        idx.push_back(i * stride);
    }
    eval_rotate_hoisted(c, *a, idx, bp);

    // inner sums of every giant step: one launch, one job per giant step
    std::vector<std::unique_ptr<Ciphertext>> inner;
    for (int j = 0; j < g; ++j) inner.push_back(ct_create(c));
    std::vector<DotJob> jobs;
    std::vector<const uint64_t*> x0s, x1s, ws;
    std::vector<size_t> first;
    for (int j = 0; j < g; ++j) {
        first.push_back(ws.size());
        int terms = 0;
        for (int i = 0; i < bstep && bstep * j + i < n; ++i, ++terms) {
            x0s.push_back(baby[i]->c0());
            x1s.push_back(baby[i]->c1());
            ws.push_back(diag[bstep * j + i]->c0());
        }
        jobs.push_back(DotJob{nullptr, nullptr, nullptr, inner[j]->c0(), inner[j]->c1(), terms, nl});
    }
    const size_t jb = jobs.size() * sizeof(DotJob), pb = ws.size() * sizeof(uint64_t*);
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_x0 = up(jb), o_x1 = up(o_x0 + pb), o_w = up(o_x1 + pb);
    DevBuf table(&c.pool, (o_w + pb + 7) / 8);
    char* d = (char*)table.ptr;
    for (size_t k = 0; k < jobs.size(); ++k) {
        jobs[k].x0 = (const uint64_t* const*)(d + o_x0) + first[k];
        jobs[k].x1 = (const uint64_t* const*)(d + o_x1) + first[k];
        jobs[k].w = (const uint64_t* const*)(d + o_w) + first[k];
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
    for (int j = 0; j < g; ++j) {
        Ciphertext& o = *inner[j];
        o.level = level;
        o.noise = a->noise * diag[bstep * j]->noise;
        o.noise_level = 2;
        o.slots = a->slots;
    }

    const int half = c.N / 2;
    Ciphertext* acc = inner[0].get();
    if (batched_giant) {
        // every giant step rotated in ONE batched key switch (keys of all
        // j b s), then summed: sum_j rot_{j b s}(inner[j])
        std::vector<std::unique_ptr<Ciphertext>> rots;
        std::vector<const Ciphertext*> ins;
        std::vector<Ciphertext*> outs;
        std::vector<int> gi;
        for (int j = 1; j < g; ++j) {
            rots.push_back(ct_create(c));
            ins.push_back(inner[j].get());
            outs.push_back(rots.back().get());
            gi.push_back(j * bstep * stride);
        }
        eval_rotate_batch(c, ins, gi, outs);
        for (auto& r : rots) eval_add(c, *acc, *r, *acc);
    } else {
        // Horner order, as the reference: acc = rot_{b s}(acc) + inner[j]
        std::unique_ptr<Ciphertext> rot = ct_create(c);
        acc = inner[g - 1].get();
        for (int j = g - 2; j >= 0; --j) {
            if (((int64_t)bstep * stride) % half != 0) {
                eval_rotate(c, *acc, bstep * stride, *rot);
                eval_add(c, *inner[j], *rot, *inner[j]);
            } else {
                eval_add(c, *inner[j], *acc, *inner[j]);
            }
            acc = inner[j].get();
        }
    }
    if (offset % half != 0)
        eval_rotate(c, *acc, offset, out);
    else
        ct_copy(c, *acc, out);
}

}  // namespace fides
