// Negacyclic NTT / INTT for gfx950.
//
// Transform (OpenFHE ForwardTransformToBitReverseInPlace, FIDESlib
// src/NTT.cu:292-659): a_eval[j] = a(psi^(2 brev(j) + 1)) mod q, bit-reversed
// output; the inverse (src/NTT.cu:43-233) includes N^-1.
//
// Structure: N = N1 * N2 (N1 = 2^ceil(logN/2)).  The Cooley-Tukey stages that
// pair elements more than N2 apart act along "columns" (stride N2), the rest
// along contiguous "rows".  Forward = column pass then row pass; inverse
// (Gentleman-Sande) = row pass then column pass.  Each pass is one launch: a
// workgroup stages a tile of W columns or W rows in LDS (16 columns or 4
// rows; see rule_pick), each
// thread keeps 4 values in registers and runs two radix-2 stages on them
// between LDS exchanges.  Short per-wave instruction streams and many waves
// matter more than register blocking here: a pass is latency-bound (a tile's
// load -> phases -> store chain), not VALU-bound.
//
// Twiddles psi^brev(k) come from the bit-reversed table with their Shoup
// companions, so each butterfly costs one Shoup product.  Column-pass
// twiddles (the first N1 table entries, shared by all columns) are staged in
// LDS; row-pass twiddles (distinct per row) are prefetched into registers
// before the data load so their latency overlaps it.  Butterflies are
// Harvey-lazy: forward values live in [0, 4p) and inverse values in [0, 2p)
// between stages, reduced once at the final store.
//
// Fusions (FIDESlib src/NTTfusions.cuh:218-262):
//   LOAD_SWITCH   (forward, pass 1): read the q_l coefficient limb and lift it
//                 centered into q_i (OpenFHE SwitchModulus) -- rescale.
//   STORE_MODDOWN (forward, pass 2): y = (acc - v) P^-1 (+ addend) -- ModDown.
//   STORE_RESCALE (forward, pass 2): y = (x - v) q_l^-1 -- rescale.
//   STORE_MODDOWN_RESCALE (forward, pass 2): y = (acc P^-1 + add - v) q_l^-1 --
//                 ModDown and rescale merged (k_moddown_conv explains the form).
//   STORE_MODDOWN_AUTO (forward, pass 2): STORE_MODDOWN, then the automorphism
//                 X -> X^k (rotation: ModDown, add c0, automorphism -- the
//                 k_automorph gather, FIDESlib src/Rotation.cuh:14-36) as a
//                 scatter through LDS; every row of the output is one row of
//                 the input, permuted.
//   inverse pass 2 folds N^-1 (times an optional per-limb base-conversion
//   factor) into its last stage.
#include <cstdlib>
#include <type_traits>

#include "engine.hpp"

namespace fides {

// log2 of the values per thread: 2, or 3 when a 16-wide column tile of
// 4-value threads would exceed 1024 threads (logN = 17).
template <int LOGN>
constexpr int ntt_lr() { return ((1 << ((LOGN + 1) / 2)) / 4) * 16 > 1024 ? 3 : 2; }

// W: tile width (columns or rows per block); LRC: log2 of the values per thread
// in the column pass (2, or 3 for radix-8 phases); the row pass always uses
// ntt_lr (its twiddles are register-prefetched, which needs short phases).
template <int LOGN, int W, int LRC = ntt_lr<LOGN>()>
struct NttGeom {
    static constexpr int K1 = (LOGN + 1) / 2, K2 = LOGN / 2;
    static constexpr int N1 = 1 << K1, N2 = 1 << K2, N = 1 << LOGN;
    static constexpr int LR = ntt_lr<LOGN>(), R = 1 << LR;
    static constexpr int LRC_ = LRC, RC = 1 << LRC;
    // column pass: CC columns of N1 rows per block, TC threads per column
    static constexpr int TC = N1 / RC;
    static constexpr int CC = N2 < W ? N2 : W;
    static constexpr int NTC = TC * CC;
    static constexpr int SC = N1 + N1 / 16 + 1;
    // row pass: RR rows of N2 elements per block, TR threads per row
    static constexpr int TR = N2 / R;
    static constexpr int RR = N1 < W ? N1 : W;
    static constexpr int NTR = TR * RR;
    static constexpr int SR = N2 + N2 / 16 + 1;
    // twiddle slots per thread in the row pass: K2 stages x R/2 butterflies
    static constexpr int NS = K2 * (R / 2);
};

__device__ __forceinline__ int find_ntt_job(const NttJobs& L, int limb) {
    int j = 0;
#pragma unroll
    for (int i = 1; i < NTT_INLINE; ++i)
        if (i < L.n && limb >= L.inl[i].first) j = i;
    return j;
}

__device__ __forceinline__ NttJob load_ntt_job(const NttJobs& L, int limb, int& pos) {
    limb = __builtin_amdgcn_readfirstlane(limb);
    const int ji = L.dev ? ((const __attribute__((address_space(1))) int*)L.limbmap)[limb] : find_ntt_job(L, limb);
    const NttJob J = get_job(L, ji);
    const int a = limb - J.first;
    pos = a < J.skip_lo ? a : a + (J.skip_hi - J.skip_lo);
    return J;
}

__device__ __forceinline__ uint32_t lds_off(uint32_t e) { return e + (e >> 4); }

// [0, 4p) -> [0, p)
__device__ __forceinline__ uint64_t reduce4p(uint64_t x, uint64_t p) {
    x = x >= 2 * p ? x - 2 * p : x;
    return x >= p ? x - p : x;
}
// [0, 2p) -> [0, p)
__device__ __forceinline__ uint64_t reduce2p(uint64_t x, uint64_t p) { return x >= p ? x - p : x; }

// Element index (within a 2^K sub-transform) of register j of thread t in a
// phase over bits [LO, LO+NB); the LR-NB spare register bits and the thread
// bits fill the remaining positions from the bottom up.
template <int LR, int LO, int NB>
__device__ __forceinline__ uint32_t elem(uint32_t t, uint32_t j) {
    constexpr uint32_t EX = LR - NB;
    const uint32_t jlo = j & ((1u << NB) - 1);
    const uint32_t other = (t << EX) | (j >> NB);
    return (other & ((1u << LO) - 1)) | (jlo << LO) | ((other >> LO) << (LO + NB));
}

// Index of butterfly (j, j | 1<<bb) among the R/2 butterflies of a stage.
__device__ __forceinline__ constexpr int bfly(int j, int bb) { return ((j >> (bb + 1)) << bb) | (j & ((1 << bb) - 1)); }

// Twiddle slot of a butterfly: phases visit the K bits in a fixed order (top
// down for the forward transform, bottom up for the inverse); `done` bits
// precede the phase, s is the stage's rank inside the phase.
template <int LR>
__device__ __forceinline__ constexpr int slot(int done, int s, int j, int bb) {
    return (done + s) * (1 << (LR - 1)) + bfly(j, bb);
}

// ---- LDS exchange between phases ----
template <int LR, int LO, int NB>
__device__ __forceinline__ void lds_get(const uint64_t* sub, uint32_t t, uint64_t* v) {
#pragma unroll
    for (int j = 0; j < (1 << LR); ++j) v[j] = sub[lds_off(elem<LR, LO, NB>(t, j))];
}
template <int LR, int LO, int NB>
__device__ __forceinline__ void lds_put(uint64_t* sub, uint32_t t, const uint64_t* v) {
#pragma unroll
    for (int j = 0; j < (1 << LR); ++j) sub[lds_off(elem<LR, LO, NB>(t, j))] = v[j];
}

// Phase boundaries of a K-bit sub-transform with up to LR bits per phase.
// Forward phases run from the top bit down: the first covers bits
// [K - NB, K), the last [0, fwd_last_nb).  Inverse phases run from bit 0 up:
// the first covers [0, NB), the last [inv_last_lo, K).
template <int LR, int K>
constexpr int first_nb() { return K < LR ? K : LR; }
template <int LR, int K>
constexpr int fwd_last_nb() { return K % LR == 0 ? LR : K % LR; }
template <int LR, int K>
constexpr int inv_last_lo() { return LR * ((K - 1) / LR); }

// ---- forward (Cooley-Tukey), bits processed from the top down ----
template <int LR, int K, int LO, int NB, class TW>
__device__ __forceinline__ void fwd_bfly(uint64_t* v, uint32_t t, const TW& tw, uint64_t p) {
    constexpr int R = 1 << LR;
#pragma unroll
    for (int bb = NB - 1; bb >= 0; --bb) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bb)) continue;
            const int j2 = j | (1 << bb);
            // Harvey: inputs in [0, 4p), outputs in [0, 4p)
            const ShoupConst w = tw(slot<LR>(K - LO - NB, NB - 1 - bb, j, bb), LO + bb, elem<LR, LO, NB>(t, j));
            uint64_t U = v[j];
            U = U >= 2 * p ? U - 2 * p : U;
            const uint64_t V = mul_shoup_lazy(v[j2], w.w, w.wp, p);
            v[j] = U + V;
            v[j2] = U - V + 2 * p;
        }
    }
}

// Forward phases from bit HI down to 0.  On entry v holds the elements of the
// phase that starts at HI (the caller loads them straight from global memory
// for the first phase); phases exchange through LDS; on exit v holds the
// elements of the last phase (bits [0, fwd_last_nb)), which the caller stores
// straight to global memory.  One barrier per exchange: a thread writes back
// exactly the elements it read, so reads and writes of one phase never race.
// hook(integral_constant<HI>) runs at the start of each phase (the row pass
// issues later phases' twiddle loads there).
struct NoHook {
    template <class X>
    __device__ __forceinline__ void operator()(X) const {}
};
template <int LR, int K, int HI, class TW, class HOOK = NoHook>
__device__ __forceinline__ void fwd_run(uint64_t* sub, uint64_t* v, uint32_t t, const TW& tw, uint64_t p,
                                        const HOOK& hook = HOOK{}) {
    constexpr int NB = HI < LR ? HI : LR;
    constexpr int LO = HI - NB;
    hook(std::integral_constant<int, HI>{});
    fwd_bfly<LR, K, LO, NB>(v, t, tw, p);
    if constexpr (LO > 0) {
        constexpr int NB2 = LO < LR ? LO : LR;
        lds_put<LR, LO, NB>(sub, t, v);
        __syncthreads();
        lds_get<LR, LO - NB2, NB2>(sub, t, v);
        fwd_run<LR, K, LO>(sub, v, t, tw, p, hook);
    }
}

// Prefetch of the forward twiddles into twr[] in slot order; IDX(b, e) gives
// the table index of the twiddle of stage bit b at lower element e.
// The twiddles of the one phase that starts at bit HI (none when HI <= 0).
template <int LR, int K, int HI, class IDX>
__device__ __forceinline__ void fwd_prefetch(ShoupConst* twr, const ShoupConst* tab, uint32_t t, const IDX& idx) {
    if constexpr (HI > 0) {
        constexpr int NB = HI < LR ? HI : LR;
        constexpr int LO = HI - NB;
#pragma unroll
        for (int bb = NB - 1; bb >= 0; --bb)
#pragma unroll
            for (int j = 0; j < (1 << LR); ++j)
                if (!(j & (1 << bb))) twr[slot<LR>(K - HI, NB - 1 - bb, j, bb)] = tab[idx(LO + bb, elem<LR, LO, NB>(t, j))];
    }
}

// ---- inverse (Gentleman-Sande), bits processed from the bottom up ----
// When SCALE, the stage on bit K-1 multiplies both outputs by the limb scale s
// (sum by s, difference by w*s).
template <int LR, int K, int LO, int NB, bool SCALE, class TW>
__device__ __forceinline__ void inv_bfly(uint64_t* v, uint32_t t, const TW& tw, uint64_t p, ShoupConst s,
                                         ShoupConst ws) {
    constexpr int R = 1 << LR;
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bb)) continue;
            const int j2 = j | (1 << bb);
            // Harvey: inputs in [0, 2p), outputs in [0, 2p)
            const uint64_t U = v[j], V = v[j2];
            if (SCALE && LO + bb == K - 1) {
                v[j] = mul_shoup_lazy(U + V, s.w, s.wp, p);
                v[j2] = mul_shoup_lazy(U - V + 2 * p, ws.w, ws.wp, p);
            } else {
                const ShoupConst w = tw(slot<LR>(LO, bb, j, bb), LO + bb, elem<LR, LO, NB>(t, j));
                const uint64_t X = U + V;
                v[j] = X >= 2 * p ? X - 2 * p : X;
                v[j2] = mul_shoup_lazy(U - V + 2 * p, w.w, w.wp, p);
            }
        }
    }
}

// Inverse phases from bit LO up to K, register/LDS discipline as fwd_run: v
// enters with the elements of the phase starting at LO and leaves with those
// of the last phase (bits [inv_last_lo, K)).
template <int LR, int K, int LO, bool SCALE, class TW, class HOOK = NoHook>
__device__ __forceinline__ void inv_run(uint64_t* sub, uint64_t* v, uint32_t t, const TW& tw, uint64_t p,
                                        ShoupConst s, ShoupConst ws, const HOOK& hook = HOOK{}) {
    constexpr int NB = (K - LO) < LR ? (K - LO) : LR;
    hook(std::integral_constant<int, LO>{});
    inv_bfly<LR, K, LO, NB, SCALE>(v, t, tw, p, s, ws);
    if constexpr (LO + NB < K) {
        constexpr int LO2 = LO + NB;
        constexpr int NB2 = (K - LO2) < LR ? (K - LO2) : LR;
        lds_put<LR, LO, NB>(sub, t, v);
        __syncthreads();
        lds_get<LR, LO2, NB2>(sub, t, v);
        inv_run<LR, K, LO2, SCALE>(sub, v, t, tw, p, s, ws, hook);
    }
}

// The twiddles of the one phase that starts at bit LO (none when LO >= K).
template <int LR, int K, int LO, class IDX>
__device__ __forceinline__ void inv_prefetch(ShoupConst* twr, const ShoupConst* tab, uint32_t t, const IDX& idx) {
    if constexpr (LO < K) {
        constexpr int NB = (K - LO) < LR ? (K - LO) : LR;
#pragma unroll
        for (int bb = 0; bb < NB; ++bb)
#pragma unroll
            for (int j = 0; j < (1 << LR); ++j)
                if (!(j & (1 << bb))) twr[slot<LR>(LO, bb, j, bb)] = tab[idx(LO + bb, elem<LR, LO, NB>(t, j))];
    }
}

// ------------------------------------------------------------ kernels ----
// Thread layout: in the column pass thread (cc, t) owns column c0 + cc and the
// rows elem(t, j) of the current phase; in the row pass thread (rr, t) owns
// row r0 + rr.  The first phase reads its elements straight from global
// memory and the last phase writes straight back (no LDS staging at either
// end): three LDS exchanges per 8-stage pass.

// Forward pass 1: column sub-transforms (global stages 0..K1-1).
template <int LOGN, int W, int LRC, int LOAD>
__global__ void __launch_bounds__((NttGeom<LOGN, W, LRC>::NTC)) k_ntt_fwd_cols(DevTables T, NttJobs L, int limbs) {
    using G = NttGeom<LOGN, W, LRC>;
    constexpr int R = G::RC, NB1 = first_nb<LRC, G::K1>(), LO1 = G::K1 - NB1, NBL = fwd_last_nb<LRC, G::K1>();
    __shared__ uint64_t lds[G::CC * G::SC];
    __shared__ ShoupConst twl[G::N1];
    constexpr int TPL = G::N2 / G::CC;  // tiles per limb
    // grid-stride loop over (limb, tile); a launch with one block per tile
    // runs one iteration per block
    for (int tile = blockIdx.x; tile < TPL * limbs; tile += gridDim.x) {
        const int limb_ = tile / TPL, tx_ = tile % TPL;
        if (tile != (int)blockIdx.x) __syncthreads();  // LDS reuse across tiles
        int pos;
        const NttJob J = load_ntt_job(L, limb_, pos);
        const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
        const PrimeInfo P = cld(T.primes + pid);
        const int c0 = tx_ * G::CC;
        const int tid = threadIdx.x;
        const int cc = tid % G::CC, t = tid / G::CC;

        {
            const ShoupConst* tw = T.tw_fwd + (size_t)pid * G::N;
            for (int i = tid; i < G::N1; i += G::NTC) twl[i] = tw[i];
        }
        uint64_t v[R];
        if constexpr (LOAD == LOAD_SWITCH) {
            const uint64_t ql = cld(T.primes + J.level).p, half = ql >> 1;
            const uint64_t qlmod = cld(T.qlmodq + (size_t)J.level * (T.L + 1) + pid);
            const gu64* src = gp(J.src) + c0 + cc;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t x = src[(size_t)elem<LRC, LO1, NB1>(t, j) * G::N2];
                uint64_t m = reduce64(x, P);
                if (x > half) m = submod(m, qlmod, P.p);
                v[j] = m;
            }
        } else {
            const gu64* src = gp(J.src) + (size_t)pos * G::N + c0 + cc;
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = src[(size_t)elem<LRC, LO1, NB1>(t, j) * G::N2];
        }
        __syncthreads();
        auto twf = [&](int, int b, uint32_t e) { return twl[(1u << (G::K1 - 1 - b)) + (e >> (b + 1))]; };
        fwd_run<LRC, G::K1, G::K1>(lds + cc * G::SC, v, t, twf, P.p);
        gu64* mid = gp(J.mid) + (size_t)pos * G::N + c0 + cc;
#pragma unroll
        for (int j = 0; j < R; ++j) mid[(size_t)elem<LRC, 0, NBL>(t, j) * G::N2] = v[j];
    }
}

// Forward pass 2: row sub-transforms (global stages K1..logN-1) + epilogue.
template <int LOGN, int W, int STORE>
__global__ void __launch_bounds__((NttGeom<LOGN, W>::NTR)) k_ntt_fwd_rows(DevTables T, NttJobs L, int limbs) {
    using G = NttGeom<LOGN, W>;
    constexpr int R = G::R, NB1 = first_nb<G::LR, G::K2>(), LO1 = G::K2 - NB1, NBL = fwd_last_nb<G::LR, G::K2>();
    __shared__ uint64_t lds[G::RR * G::SR];
    constexpr int TPL = G::N1 / G::RR;  // tiles per limb
    // grid-stride loop over (limb, tile); a launch with one block per tile
    // runs one iteration per block
    for (int tile = blockIdx.x; tile < TPL * limbs; tile += gridDim.x) {
        const int limb_ = tile / TPL, tx_ = tile % TPL;
        if (tile != (int)blockIdx.x) __syncthreads();  // LDS reuse across tiles
        int pos;
        const NttJob J = load_ntt_job(L, limb_, pos);
        const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
        const PrimeInfo P = cld(T.primes + pid);
        const int r0 = tx_ * G::RR;
        const int tid = threadIdx.x;
        const int t = tid % G::TR, rr = tid / G::TR;
        const size_t base = (size_t)pos * G::N + (size_t)(r0 + rr) * G::N2;

        // Loads run ahead of use without holding every phase's operands in
        // registers at once: each phase's twiddles are issued one phase
        // early (the first phase's with the tile), the epilogue operands of
        // the elements this thread stores two phases before the store.
        ShoupConst twr[G::NS];
        const uint32_t row = G::N1 + r0 + rr;
        const ShoupConst* twtab = T.tw_fwd + (size_t)pid * G::N;
        auto idx = [&](int b, uint32_t e) { return (row << (G::K2 - 1 - b)) + (e >> (b + 1)); };
        fwd_prefetch<G::LR, G::K2, G::K2>(twr, twtab, t, idx);
        uint64_t v[R];
        {
            const gu64* mid = gp(J.mid) + base;
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = mid[elem<G::LR, LO1, NB1>(t, j)];
        }
        uint64_t e0[R], e1[R];
        gu64* dst = gp(J.dst) + base;
        auto epi_load = [&]() {
            if constexpr (STORE == STORE_MODDOWN || STORE == STORE_MODDOWN_RESCALE || STORE == STORE_MODDOWN_AUTO) {
                const gu64* aux = gp(J.aux) + base;
                const gu64* add = J.add ? gp(J.add) + base : nullptr;
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    const uint32_t e = elem<G::LR, 0, NBL>(t, j);
                    e0[j] = aux[e];
                    e1[j] = add ? add[e] : 0;
                }
            } else if constexpr (STORE == STORE_RESCALE) {
#pragma unroll
                for (int j = 0; j < R; ++j) e0[j] = dst[elem<G::LR, 0, NBL>(t, j)];
            }
        };
        // the epilogue loads are issued at the start of the phase before the
        // last one (phases start at K2, K2 - LR, ..., fwd_last_nb)
        constexpr int EPI_HI = fwd_last_nb<G::LR, G::K2>() + G::LR < G::K2 ? fwd_last_nb<G::LR, G::K2>() + G::LR : G::K2;
        auto hook = [&](auto hi) {
            constexpr int H = decltype(hi)::value;
            fwd_prefetch<G::LR, G::K2, H - G::LR>(twr, twtab, t, idx);
            if constexpr (H == EPI_HI) epi_load();
        };
        auto twf = [&](int s, int, uint32_t) { return twr[s]; };
        fwd_run<G::LR, G::K2, G::K2>(lds + rr * G::SR, v, t, twf, P.p, hook);
        if constexpr (STORE == STORE_PLAIN) {
#pragma unroll
            for (int j = 0; j < R; ++j) dst[elem<G::LR, 0, NBL>(t, j)] = reduce4p(v[j], P.p);
        } else if constexpr (STORE == STORE_MODDOWN) {
            const ShoupConst pinv = cld(T.pinv + pid);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t x = reduce4p(v[j], P.p);
                // the addend is 0 when absent, so the add is a no-op then
                dst[elem<G::LR, 0, NBL>(t, j)] = addmod(mul_shoup(submod(e0[j], x, P.p), pinv.w, pinv.wp, P.p), e1[j], P.p);
            }
        } else if constexpr (STORE == STORE_MODDOWN_AUTO) {
            // Element jg = r N2 + e sits at exponent 2 brev(jg) + 1 =
            // 2 N1 brev(e) + 2 brev(r) + 1; times k (mod 2N) it lands at
            // row brev(((c mod 2N1) - 1) / 2), column brev((brev(e) k + (c >> (K1+1))) mod N2)
            // with c = (2 brev(r) + 1) k mod 2N: one destination row per row.
            const ShoupConst pinv = cld(T.pinv + pid);
            const uint32_t k = J.autok;
            const uint32_t cx = (uint32_t)(((2ull * brev_bits(r0 + rr, G::K1) + 1) * k) & (2ull * G::N - 1));
            const uint32_t rdst = brev_bits((cx & (2u * G::N1 - 1)) >> 1, G::K1);
            const uint32_t coff = cx >> (G::K1 + 1);
            uint64_t* row = lds + rr * G::SR;
            __syncthreads();  // the last phase's LDS reads are done
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t x = reduce4p(v[j], P.p);
                const uint32_t e = elem<G::LR, 0, NBL>(t, j);
                const uint32_t e2 = brev_bits((brev_bits(e, G::K2) * k + coff) & (G::N2 - 1), G::K2);
                row[lds_off(e2)] = addmod(mul_shoup(submod(e0[j], x, P.p), pinv.w, pinv.wp, P.p), e1[j], P.p);
            }
            __syncthreads();
            gu64* out = gp(J.dst) + (size_t)pos * G::N + (size_t)rdst * G::N2;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint32_t e = elem<G::LR, 0, NBL>(t, j);
                out[e] = row[lds_off(e)];
            }
        } else if constexpr (STORE == STORE_MODDOWN_RESCALE) {
            // r = (acc P^-1 + add - v) q_l^-1  (see k_moddown_conv)
            const ShoupConst pinv = cld(T.pinv + pid);
            const ShoupConst qi = cld(T.qlinv + (size_t)J.level * (T.L + 1) + pid);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t x = reduce4p(v[j], P.p);
                const uint64_t y = addmod(mul_shoup(e0[j], pinv.w, pinv.wp, P.p), e1[j], P.p);
                dst[elem<G::LR, 0, NBL>(t, j)] = mul_shoup(submod(y, x, P.p), qi.w, qi.wp, P.p);
            }
        } else {  // STORE_RESCALE
            const ShoupConst qi = cld(T.qlinv + (size_t)J.level * (T.L + 1) + pid);
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const uint64_t x = reduce4p(v[j], P.p);
                dst[elem<G::LR, 0, NBL>(t, j)] = mul_shoup(submod(e0[j], x, P.p), qi.w, qi.wp, P.p);
            }
        }
    }
}

// Inverse pass 1: row sub-transforms (global stages on bits 0..K2-1).
template <int LOGN, int W>
__global__ void __launch_bounds__((NttGeom<LOGN, W>::NTR)) k_ntt_inv_rows(DevTables T, NttJobs L, int limbs) {
    using G = NttGeom<LOGN, W>;
    constexpr int R = G::R, NB1 = first_nb<G::LR, G::K2>(), LOL = inv_last_lo<G::LR, G::K2>(), NBL = G::K2 - LOL;
    __shared__ uint64_t lds[G::RR * G::SR];
    constexpr int TPL = G::N1 / G::RR;  // tiles per limb
    // grid-stride loop over (limb, tile); a launch with one block per tile
    // runs one iteration per block
    for (int tile = blockIdx.x; tile < TPL * limbs; tile += gridDim.x) {
        const int limb_ = tile / TPL, tx_ = tile % TPL;
        if (tile != (int)blockIdx.x) __syncthreads();  // LDS reuse across tiles
        int pos;
        const NttJob J = load_ntt_job(L, limb_, pos);
        const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
        const PrimeInfo P = cld(T.primes + pid);
        const int r0 = tx_ * G::RR;
        const int tid = threadIdx.x;
        const int t = tid % G::TR, rr = tid / G::TR;
        const size_t base = (size_t)pos * G::N + (size_t)(r0 + rr) * G::N2;

        ShoupConst twr[G::NS];
        const uint32_t row = G::N1 + r0 + rr;
        const ShoupConst* twtab = T.tw_inv + (size_t)pid * G::N;
        auto idx = [&](int b, uint32_t e) { return (row << (G::K2 - 1 - b)) + (e >> (b + 1)); };
        // each phase's twiddles are issued one phase early (the first's with the tile)
        inv_prefetch<G::LR, G::K2, 0>(twr, twtab, t, idx);
        auto hook = [&](auto lo) { inv_prefetch<G::LR, G::K2, decltype(lo)::value + G::LR>(twr, twtab, t, idx); };
        uint64_t v[R];
        {
            const gu64* src = gp(J.src) + base;
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = src[elem<G::LR, 0, NB1>(t, j)];
        }
        auto twf = [&](int s, int, uint32_t) { return twr[s]; };
        inv_run<G::LR, G::K2, 0, false>(lds + rr * G::SR, v, t, twf, P.p, ShoupConst{0, 0}, ShoupConst{0, 0}, hook);
        gu64* mid = gp(J.mid) + base;
#pragma unroll
        for (int j = 0; j < R; ++j) mid[elem<G::LR, LOL, NBL>(t, j)] = v[j];
    }
}

// Inverse pass 2: column sub-transforms, last stage scaled by N^-1 (x factor).
template <int LOGN, int W, int LRC>
__global__ void __launch_bounds__((NttGeom<LOGN, W, LRC>::NTC)) k_ntt_inv_cols(DevTables T, NttJobs L, int limbs) {
    using G = NttGeom<LOGN, W, LRC>;
    constexpr int R = G::RC, NB1 = first_nb<LRC, G::K1>(), LOL = inv_last_lo<LRC, G::K1>(), NBL = G::K1 - LOL;
    __shared__ uint64_t lds[G::CC * G::SC];
    __shared__ ShoupConst twl[G::N1];
    constexpr int TPL = G::N2 / G::CC;  // tiles per limb
    // grid-stride loop over (limb, tile); a launch with one block per tile
    // runs one iteration per block
    for (int tile = blockIdx.x; tile < TPL * limbs; tile += gridDim.x) {
        const int limb_ = tile / TPL, tx_ = tile % TPL;
        if (tile != (int)blockIdx.x) __syncthreads();  // LDS reuse across tiles
        int pos;
        const NttJob J = load_ntt_job(L, limb_, pos);
        const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
        const PrimeInfo P = cld(T.primes + pid);
        const int c0 = tx_ * G::CC;
        const int tid = threadIdx.x;
        const int cc = tid % G::CC, t = tid / G::CC;

        const ShoupConst* sc;
        if (J.scale_sel == ISCALE_MODUP)
            sc = T.iscale_modup + ((size_t)J.level * (T.L + 1) + pos) * 2;
        else if (J.scale_sel == ISCALE_MODDOWN)
            sc = T.iscale_moddown + (size_t)(pos - J.nq) * 2;
        else
            sc = T.iscale_plain + (size_t)pid * 2;
        const ShoupConst s = cld(sc), ws = cld(sc + 1);

        {
            const ShoupConst* tw = T.tw_inv + (size_t)pid * G::N;
            for (int i = tid; i < G::N1; i += G::NTC) twl[i] = tw[i];
        }
        uint64_t v[R];
        {
            const gu64* mid = gp(J.mid) + (size_t)pos * G::N + c0 + cc;
#pragma unroll
            for (int j = 0; j < R; ++j) v[j] = mid[(size_t)elem<LRC, 0, NB1>(t, j) * G::N2];
        }
        __syncthreads();
        auto twf = [&](int, int b, uint32_t e) { return twl[(1u << (G::K1 - 1 - b)) + (e >> (b + 1))]; };
        inv_run<LRC, G::K1, 0, true>(lds + cc * G::SC, v, t, twf, P.p, s, ws);
        gu64* dst = gp(J.dst) + (size_t)pos * G::N + c0 + cc;
#pragma unroll
        for (int j = 0; j < R; ++j) dst[(size_t)elem<LRC, LOL, NBL>(t, j) * G::N2] = reduce2p(v[j], P.p);
    }
}

// ---------------------------------------------------------- launchers ----

// Launch configuration of a transform: column tile width, log2 values per
// column thread, row tile height, grid cap (0 = one block per tile).
struct NttPick {
    int cc, lrc, rr, pb;
};

// Fixed rules (measured on MI355X at N=2^16 over 2..25 limbs): column tiles
// are 16 wide (128-byte row segments; 4-wide tiles lose more to partial lines
// than they gain in parallelism) except for launches of under 64 tiles; row
// tiles are 4 rows (256 threads), which keeps register-prefetched twiddles at
// full occupancy; 8 values per column thread (radix-8 phases, one LDS
// exchange fewer) from 40 limbs up, where it measured 3-15% faster.
// FIDES_NTT_COLS / FIDES_NTT_ROWS (4, 8 or 16), FIDES_NTT_LRC (2 or 3) and
// FIDES_NTT_PB force a value.
static int env_int(const char* name) {
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : -1;
}
static NttPick rule_pick(int logN, int limbs) {
    static const int ov_c = env_int("FIDES_NTT_COLS"), ov_r = env_int("FIDES_NTT_ROWS");
    static const int ov_l = env_int("FIDES_NTT_LRC"), ov_pb = env_int("FIDES_NTT_PB");
    auto width = [](int w) { return w == 4 || w == 8 ? w : 16; };
    const int n2 = 1 << (logN / 2);
    const int tiles16 = limbs * (n2 < 16 ? 1 : n2 / 16);
    NttPick p;
    p.cc = ov_c > 0 ? width(ov_c) : (tiles16 < 64 ? 4 : 16);
    p.rr = ov_r > 0 ? width(ov_r) : 4;
    p.lrc = ov_l > 0 ? (ov_l == 3 ? 3 : 2) : (limbs >= 40 ? 3 : 2);
    p.pb = ov_pb > 0 ? ov_pb : 0;
    return p;
}

// The measured table's entry for this launch (pyfideslib_amd/tune.py), else
// the fixed rules.
static NttPick ntt_pick(bool forward, int logN, int limbs) {
    TuneCfg c;
    if (tune_get(forward ? TUNE_NTT_FWD : TUNE_NTT_INV, logN, limbs, c)) return NttPick{c.v[0], c.v[1], c.v[2], c.v[3]};
    return rule_pick(logN, limbs);
}

TuneCfg ntt_rule_cfg(bool, int logN, int limbs) {
    const NttPick p = rule_pick(logN, limbs);
    return TuneCfg{{p.cc, p.lrc, p.rr, p.pb}};
}

// Blocks of a pass: one per tile, or at most pb blocks looping over the tiles.
static dim3 pass_grid(int tiles, int pb) { return dim3(pb > 0 && pb < tiles ? pb : tiles); }

template <int LOGN, int W, int LRC>
static void fwd_cols(const DevTables& T, const NttJobs& L, int limbs, int load, int pb, hipStream_t s) {
    using G = NttGeom<LOGN, W, LRC>;
    const dim3 g = pass_grid(G::N2 / G::CC * limbs, pb);
    if (load == LOAD_SWITCH)
        hipLaunchKernelGGL((k_ntt_fwd_cols<LOGN, W, LRC, LOAD_SWITCH>), g, dim3(G::NTC), 0, s, T, L, limbs);
    else
        hipLaunchKernelGGL((k_ntt_fwd_cols<LOGN, W, LRC, LOAD_PLAIN>), g, dim3(G::NTC), 0, s, T, L, limbs);
}

template <int LOGN, int W, int LRC>
static void inv_cols(const DevTables& T, const NttJobs& L, int limbs, int pb, hipStream_t s) {
    using G = NttGeom<LOGN, W, LRC>;
    hipLaunchKernelGGL((k_ntt_inv_cols<LOGN, W, LRC>), pass_grid(G::N2 / G::CC * limbs, pb), dim3(G::NTC), 0, s, T,
                       L, limbs);
}

template <int LOGN, int W>
static void fwd_rows(const DevTables& T, const NttJobs& L, int limbs, int store, int pb, hipStream_t s) {
    using G = NttGeom<LOGN, W>;
    const dim3 g = pass_grid(G::N1 / G::RR * limbs, pb);
    if (store == STORE_MODDOWN)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, W, STORE_MODDOWN>), g, dim3(G::NTR), 0, s, T, L, limbs);
    else if (store == STORE_MODDOWN_AUTO)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, W, STORE_MODDOWN_AUTO>), g, dim3(G::NTR), 0, s, T, L, limbs);
    else if (store == STORE_MODDOWN_RESCALE)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, W, STORE_MODDOWN_RESCALE>), g, dim3(G::NTR), 0, s, T, L, limbs);
    else if (store == STORE_RESCALE)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, W, STORE_RESCALE>), g, dim3(G::NTR), 0, s, T, L, limbs);
    else
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, W, STORE_PLAIN>), g, dim3(G::NTR), 0, s, T, L, limbs);
}

template <int LOGN, int W>
static void inv_rows(const DevTables& T, const NttJobs& L, int limbs, int pb, hipStream_t s) {
    using G = NttGeom<LOGN, W>;
    hipLaunchKernelGGL((k_ntt_inv_rows<LOGN, W>), pass_grid(G::N1 / G::RR * limbs, pb), dim3(G::NTR), 0, s, T, L,
                       limbs);
}

template <class F>
static void by_width(int w, F&& f) {
    if (w == 4)
        f(std::integral_constant<int, 4>{});
    else if (w == 8)
        f(std::integral_constant<int, 8>{});
    else
        f(std::integral_constant<int, 16>{});
}

// Column-pass dispatch over (tile width, values per thread).  logN = 17 has
// only the 8-value form (a 16-wide tile of 4-value threads exceeds 1024).
template <int LOGN, bool FWD>
static void cols_dispatch(const DevTables& T, const NttJobs& L, int limbs, int load, const NttPick& pk,
                          hipStream_t s) {
    const int lrc = ntt_lr<LOGN>() == 3 ? 3 : pk.lrc;
    by_width(pk.cc, [&](auto wv) {
        constexpr int WW = decltype(wv)::value;
        auto go = [&](auto lv) {
            constexpr int LL = decltype(lv)::value;
            if constexpr (FWD)
                fwd_cols<LOGN, WW, LL>(T, L, limbs, load, pk.pb, s);
            else
                inv_cols<LOGN, WW, LL>(T, L, limbs, pk.pb, s);
        };
        if (lrc == 3)
            go(std::integral_constant<int, 3>{});
        else
            go(std::integral_constant<int, ntt_lr<LOGN>()>{});
    });
}

template <int LOGN>
static hipError_t fwd_impl(const DevTables& T, const NttJobs& L, int limbs, int load, int store, hipStream_t s) {
    const NttPick pk = ntt_pick(true, LOGN, limbs);
    cols_dispatch<LOGN, true>(T, L, limbs, load, pk, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    by_width(pk.rr, [&](auto wr) { fwd_rows<LOGN, decltype(wr)::value>(T, L, limbs, store, pk.pb, s); });
    return hipGetLastError();
}

template <int LOGN>
static hipError_t inv_impl(const DevTables& T, const NttJobs& L, int limbs, hipStream_t s) {
    const NttPick pk = ntt_pick(false, LOGN, limbs);
    by_width(pk.rr, [&](auto wr) { inv_rows<LOGN, decltype(wr)::value>(T, L, limbs, pk.pb, s); });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    cols_dispatch<LOGN, false>(T, L, limbs, 0, pk, s);
    return hipGetLastError();
}

hipError_t launch_ntt_fwd(const DevTables& T, const NttJobs& L, int limbs, int load, int store, hipStream_t s) {
    if (limbs <= 0) return hipSuccess;
    switch (T.logN) {
        case 10: return fwd_impl<10>(T, L, limbs, load, store, s);
        case 11: return fwd_impl<11>(T, L, limbs, load, store, s);
        case 12: return fwd_impl<12>(T, L, limbs, load, store, s);
        case 13: return fwd_impl<13>(T, L, limbs, load, store, s);
        case 14: return fwd_impl<14>(T, L, limbs, load, store, s);
        case 15: return fwd_impl<15>(T, L, limbs, load, store, s);
        case 16: return fwd_impl<16>(T, L, limbs, load, store, s);
        case 17: return fwd_impl<17>(T, L, limbs, load, store, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ntt_inv(const DevTables& T, const NttJobs& L, int limbs, hipStream_t s) {
    if (limbs <= 0) return hipSuccess;
    switch (T.logN) {
        case 10: return inv_impl<10>(T, L, limbs, s);
        case 11: return inv_impl<11>(T, L, limbs, s);
        case 12: return inv_impl<12>(T, L, limbs, s);
        case 13: return inv_impl<13>(T, L, limbs, s);
        case 14: return inv_impl<14>(T, L, limbs, s);
        case 15: return inv_impl<15>(T, L, limbs, s);
        case 16: return inv_impl<16>(T, L, limbs, s);
        case 17: return inv_impl<17>(T, L, limbs, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fides
