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
// workgroup stages a tile of columns or rows in LDS, each thread keeps 16
// values in registers and runs up to four radix-2 stages on them between LDS
// exchanges.  Twiddles psi^brev(k) come from the bit-reversed table with
// their Shoup companions, so each butterfly costs one Shoup product.
//
// Fusions (FIDESlib src/NTTfusions.cuh:218-262):
//   LOAD_SWITCH   (forward, pass 1): read the q_l coefficient limb and lift it
//                 centered into q_i (OpenFHE SwitchModulus) -- rescale.
//   STORE_MODDOWN (forward, pass 2): y = (acc - v) P^-1 (+ addend) -- ModDown.
//   STORE_RESCALE (forward, pass 2): y = (x - v) q_l^-1 -- rescale.
//   inverse pass 2 folds N^-1 (times an optional per-limb base-conversion
//   factor) into its last stage.
#include "engine.hpp"

namespace fides {

template <int LOGN>
struct NttGeom {
    static constexpr int K1 = (LOGN + 1) / 2, K2 = LOGN / 2;
    static constexpr int N1 = 1 << K1, N2 = 1 << K2, N = 1 << LOGN;
    static constexpr int LR = 4, R = 1 << LR;
    // column pass: CC columns of N1 rows per block, TC threads per column
    static constexpr int TC = N1 / R;
    static constexpr int CC = N2 < 16 ? N2 : 16;
    static constexpr int NTC = TC * CC;
    static constexpr int SC = N1 + N1 / 16 + 1;
    // row pass: RR rows of N2 elements per block, TR threads per row
    static constexpr int TR = N2 / R;
    static constexpr int RR0 = 256 / TR;
    static constexpr int RR = RR0 < N1 ? RR0 : N1;
    static constexpr int NTR = TR * RR;
    static constexpr int SR = N2 + N2 / 16 + 1;
};

__device__ __forceinline__ int find_ntt_job(const NttJobs& L, int limb) {
    if (L.dev) return L.limbmap[limb];
    int j = 0;
#pragma unroll
    for (int i = 1; i < NTT_INLINE; ++i)
        if (i < L.n && limb >= L.inl[i].first) j = i;
    return j;
}

__device__ __forceinline__ NttJob load_ntt_job(const NttJobs& L, int limb, int& pos) {
    const int ji = find_ntt_job(L, limb);
    NttJob J = L.dev ? L.dev[ji] : L.inl[ji];
    const int a = limb - J.first;
    pos = a < J.skip_lo ? a : a + (J.skip_hi - J.skip_lo);
    return J;
}

__device__ __forceinline__ uint32_t lds_off(uint32_t e) { return e + (e >> 4); }

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

// ---- forward (Cooley-Tukey), bits processed from the top down ----
template <int LR, int LO, int NB, class TW>
__device__ __forceinline__ void fwd_phase(uint64_t* sub, uint32_t t, const TW& tw, uint64_t p) {
    constexpr int R = 1 << LR;
    uint64_t v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = sub[lds_off(elem<LR, LO, NB>(t, j))];
#pragma unroll
    for (int bb = NB - 1; bb >= 0; --bb) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bb)) continue;
            const int j2 = j | (1 << bb);
            const ShoupConst w = tw(LO + bb, elem<LR, LO, NB>(t, j));
            const uint64_t U = v[j];
            const uint64_t V = mul_shoup(v[j2], w.w, w.wp, p);
            v[j] = addmod(U, V, p);
            v[j2] = submod(U, V, p);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) sub[lds_off(elem<LR, LO, NB>(t, j))] = v[j];
}

template <int LR, int HI, class TW>
__device__ __forceinline__ void fwd_phases(uint64_t* sub, uint32_t t, const TW& tw, uint64_t p) {
    if constexpr (HI > 0) {
        constexpr int NB = HI < LR ? HI : LR;
        constexpr int LO = HI - NB;
        fwd_phase<LR, LO, NB>(sub, t, tw, p);
        __syncthreads();
        fwd_phases<LR, LO>(sub, t, tw, p);
    }
}

// ---- inverse (Gentleman-Sande), bits processed from the bottom up ----
// When SCALE, the stage on bit K-1 multiplies both outputs by the limb scale s
// (sum by s, difference by w*s).
template <int LR, int K, int LO, int NB, bool SCALE, class TW>
__device__ __forceinline__ void inv_phase(uint64_t* sub, uint32_t t, const TW& tw, uint64_t p, ShoupConst s,
                                          ShoupConst ws) {
    constexpr int R = 1 << LR;
    uint64_t v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = sub[lds_off(elem<LR, LO, NB>(t, j))];
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (j & (1 << bb)) continue;
            const int j2 = j | (1 << bb);
            const uint64_t U = v[j], V = v[j2];
            if (SCALE && LO + bb == K - 1) {
                v[j] = mul_shoup(addmod(U, V, p), s.w, s.wp, p);
                v[j2] = mul_shoup(submod(U, V, p), ws.w, ws.wp, p);
            } else {
                const ShoupConst w = tw(LO + bb, elem<LR, LO, NB>(t, j));
                v[j] = addmod(U, V, p);
                v[j2] = mul_shoup(submod(U, V, p), w.w, w.wp, p);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) sub[lds_off(elem<LR, LO, NB>(t, j))] = v[j];
}

template <int LR, int K, int LO, bool SCALE, class TW>
__device__ __forceinline__ void inv_phases(uint64_t* sub, uint32_t t, const TW& tw, uint64_t p, ShoupConst s,
                                           ShoupConst ws) {
    if constexpr (LO < K) {
        constexpr int NB = (K - LO) < LR ? (K - LO) : LR;
        inv_phase<LR, K, LO, NB, SCALE>(sub, t, tw, p, s, ws);
        __syncthreads();
        inv_phases<LR, K, LO + NB, SCALE>(sub, t, tw, p, s, ws);
    }
}

// ------------------------------------------------------------ kernels ----

// Forward pass 1: column sub-transforms (global stages 0..K1-1).
template <int LOGN, int LOAD>
__global__ void __launch_bounds__(NttGeom<LOGN>::NTC) k_ntt_fwd_cols(DevTables T, NttJobs L) {
    using G = NttGeom<LOGN>;
    __shared__ uint64_t lds[G::CC * G::SC];
    int pos;
    const NttJob J = load_ntt_job(L, blockIdx.y, pos);
    const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
    const PrimeInfo P = T.primes[pid];
    const int c0 = blockIdx.x * G::CC;
    const int tid = threadIdx.x;

    if constexpr (LOAD == LOAD_SWITCH) {
        const uint64_t ql = T.primes[J.level].p, half = ql >> 1;
        const uint64_t qlmod = T.qlmodq[(size_t)J.level * (T.L + 1) + pid];
        const uint64_t* src = J.src;
#pragma unroll
        for (int it = 0; it < G::R; ++it) {
            const int f = tid + it * G::NTC, cc = f % G::CC, r = f / G::CC;
            const uint64_t v = src[r * G::N2 + c0 + cc];
            uint64_t m = reduce64(v, P);
            if (v > half) m = submod(m, qlmod, P.p);
            lds[cc * G::SC + lds_off(r)] = m;
        }
    } else {
        const uint64_t* src = J.src + (size_t)pos * G::N;
#pragma unroll
        for (int it = 0; it < G::R; ++it) {
            const int f = tid + it * G::NTC, cc = f % G::CC, r = f / G::CC;
            lds[cc * G::SC + lds_off(r)] = src[r * G::N2 + c0 + cc];
        }
    }
    __syncthreads();
    {
        const int cc = tid % G::CC, t = tid / G::CC;
        const ShoupConst* tw = T.tw_fwd + (size_t)pid * G::N;
        auto twf = [&](int b, uint32_t e) { return tw[(1u << (G::K1 - 1 - b)) + (e >> (b + 1))]; };
        fwd_phases<G::LR, G::K1>(lds + cc * G::SC, t, twf, P.p);
    }
    uint64_t* mid = J.mid + (size_t)pos * G::N;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTC, cc = f % G::CC, r = f / G::CC;
        mid[r * G::N2 + c0 + cc] = lds[cc * G::SC + lds_off(r)];
    }
}

// Forward pass 2: row sub-transforms (global stages K1..logN-1) + epilogue.
template <int LOGN, int STORE>
__global__ void __launch_bounds__(NttGeom<LOGN>::NTR) k_ntt_fwd_rows(DevTables T, NttJobs L) {
    using G = NttGeom<LOGN>;
    __shared__ uint64_t lds[G::RR * G::SR];
    int pos;
    const NttJob J = load_ntt_job(L, blockIdx.y, pos);
    const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
    const PrimeInfo P = T.primes[pid];
    const int r0 = blockIdx.x * G::RR;
    const int tid = threadIdx.x;
    const size_t base = (size_t)pos * G::N + (size_t)r0 * G::N2;

    const uint64_t* mid = J.mid + base;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
        lds[rr * G::SR + lds_off(c)] = mid[f];
    }
    __syncthreads();
    {
        const int t = tid % G::TR, rr = tid / G::TR;
        const uint32_t row = G::N1 + r0 + rr;
        const ShoupConst* tw = T.tw_fwd + (size_t)pid * G::N;
        auto twf = [&](int b, uint32_t e) { return tw[(row << (G::K2 - 1 - b)) + (e >> (b + 1))]; };
        fwd_phases<G::LR, G::K2>(lds + rr * G::SR, t, twf, P.p);
    }
    uint64_t* dst = J.dst + base;
    if constexpr (STORE == STORE_PLAIN) {
#pragma unroll
        for (int it = 0; it < G::R; ++it) {
            const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
            dst[f] = lds[rr * G::SR + lds_off(c)];
        }
    } else if constexpr (STORE == STORE_MODDOWN) {
        const ShoupConst pinv = T.pinv[pid];
        const uint64_t* aux = J.aux + base;
        const uint64_t* add = J.add ? J.add + base : nullptr;
#pragma unroll
        for (int it = 0; it < G::R; ++it) {
            const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
            const uint64_t v = lds[rr * G::SR + lds_off(c)];
            uint64_t o = mul_shoup(submod(aux[f], v, P.p), pinv.w, pinv.wp, P.p);
            if (add) o = addmod(o, add[f], P.p);
            dst[f] = o;
        }
    } else {  // STORE_RESCALE
        const ShoupConst qi = T.qlinv[(size_t)J.level * (T.L + 1) + pid];
#pragma unroll
        for (int it = 0; it < G::R; ++it) {
            const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
            const uint64_t v = lds[rr * G::SR + lds_off(c)];
            dst[f] = mul_shoup(submod(dst[f], v, P.p), qi.w, qi.wp, P.p);
        }
    }
}

// Inverse pass 1: row sub-transforms (global stages on bits 0..K2-1).
template <int LOGN>
__global__ void __launch_bounds__(NttGeom<LOGN>::NTR) k_ntt_inv_rows(DevTables T, NttJobs L) {
    using G = NttGeom<LOGN>;
    __shared__ uint64_t lds[G::RR * G::SR];
    int pos;
    const NttJob J = load_ntt_job(L, blockIdx.y, pos);
    const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
    const PrimeInfo P = T.primes[pid];
    const int r0 = blockIdx.x * G::RR;
    const int tid = threadIdx.x;
    const size_t base = (size_t)pos * G::N + (size_t)r0 * G::N2;

    const uint64_t* src = J.src + base;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
        lds[rr * G::SR + lds_off(c)] = src[f];
    }
    __syncthreads();
    {
        const int t = tid % G::TR, rr = tid / G::TR;
        const uint32_t row = G::N1 + r0 + rr;
        const ShoupConst* tw = T.tw_inv + (size_t)pid * G::N;
        auto twf = [&](int b, uint32_t e) { return tw[(row << (G::K2 - 1 - b)) + (e >> (b + 1))]; };
        inv_phases<G::LR, G::K2, 0, false>(lds + rr * G::SR, t, twf, P.p, ShoupConst{0, 0}, ShoupConst{0, 0});
    }
    uint64_t* mid = J.mid + base;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTR, rr = f / G::N2, c = f % G::N2;
        mid[f] = lds[rr * G::SR + lds_off(c)];
    }
}

// Inverse pass 2: column sub-transforms, last stage scaled by N^-1 (x factor).
template <int LOGN>
__global__ void __launch_bounds__(NttGeom<LOGN>::NTC) k_ntt_inv_cols(DevTables T, NttJobs L) {
    using G = NttGeom<LOGN>;
    __shared__ uint64_t lds[G::CC * G::SC];
    int pos;
    const NttJob J = load_ntt_job(L, blockIdx.y, pos);
    const int pid = pos_pid(pos, J.nq, J.q0, T.L + 1);
    const PrimeInfo P = T.primes[pid];
    const int c0 = blockIdx.x * G::CC;
    const int tid = threadIdx.x;

    const ShoupConst* sc;
    if (J.scale_sel == ISCALE_MODUP)
        sc = T.iscale_modup + ((size_t)J.level * (T.L + 1) + pos) * 2;
    else if (J.scale_sel == ISCALE_MODDOWN)
        sc = T.iscale_moddown + (size_t)(pos - J.nq) * 2;
    else
        sc = T.iscale_plain + (size_t)pid * 2;
    const ShoupConst s = sc[0], ws = sc[1];

    const uint64_t* mid = J.mid + (size_t)pos * G::N;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTC, cc = f % G::CC, r = f / G::CC;
        lds[cc * G::SC + lds_off(r)] = mid[r * G::N2 + c0 + cc];
    }
    __syncthreads();
    {
        const int cc = tid % G::CC, t = tid / G::CC;
        const ShoupConst* tw = T.tw_inv + (size_t)pid * G::N;
        auto twf = [&](int b, uint32_t e) { return tw[(1u << (G::K1 - 1 - b)) + (e >> (b + 1))]; };
        inv_phases<G::LR, G::K1, 0, true>(lds + cc * G::SC, t, twf, P.p, s, ws);
    }
    uint64_t* dst = J.dst + (size_t)pos * G::N;
#pragma unroll
    for (int it = 0; it < G::R; ++it) {
        const int f = tid + it * G::NTC, cc = f % G::CC, r = f / G::CC;
        dst[r * G::N2 + c0 + cc] = lds[cc * G::SC + lds_off(r)];
    }
}

// ---------------------------------------------------------- launchers ----

template <int LOGN>
static hipError_t fwd_impl(const DevTables& T, const NttJobs& L, int limbs, int load, int store, hipStream_t s) {
    using G = NttGeom<LOGN>;
    dim3 g1(G::N2 / G::CC, limbs), g2(G::N1 / G::RR, limbs);
    if (load == LOAD_SWITCH)
        hipLaunchKernelGGL((k_ntt_fwd_cols<LOGN, LOAD_SWITCH>), g1, dim3(G::NTC), 0, s, T, L);
    else
        hipLaunchKernelGGL((k_ntt_fwd_cols<LOGN, LOAD_PLAIN>), g1, dim3(G::NTC), 0, s, T, L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (store == STORE_MODDOWN)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, STORE_MODDOWN>), g2, dim3(G::NTR), 0, s, T, L);
    else if (store == STORE_RESCALE)
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, STORE_RESCALE>), g2, dim3(G::NTR), 0, s, T, L);
    else
        hipLaunchKernelGGL((k_ntt_fwd_rows<LOGN, STORE_PLAIN>), g2, dim3(G::NTR), 0, s, T, L);
    return hipGetLastError();
}

template <int LOGN>
static hipError_t inv_impl(const DevTables& T, const NttJobs& L, int limbs, hipStream_t s) {
    using G = NttGeom<LOGN>;
    dim3 g1(G::N1 / G::RR, limbs), g2(G::N2 / G::CC, limbs);
    hipLaunchKernelGGL((k_ntt_inv_rows<LOGN>), g1, dim3(G::NTR), 0, s, T, L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ntt_inv_cols<LOGN>), g2, dim3(G::NTC), 0, s, T, L);
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
