// Elementwise, base-conversion, key-switch and automorphism kernels (gfx950).
//
// All are HBM streams over limb-major polynomials: 256-thread blocks along the
// coefficient axis, the limb (position) along blockIdx.y and the job along
// blockIdx.z, so every access is a fully coalesced stream (16-byte lanes in
// the tensor and KSK kernels, which handle two coefficients per thread).  Sums of products are accumulated exactly
// in 128 bits and reduced once (Barrett, valid below 2^(n+63); sums of more
// than 8 products are reduced in chunks of 8).
#include "engine.hpp"

namespace fides {

constexpr int EW_BLOCK = 256;

template <class J, int CAP>
__device__ __forceinline__ J get_job(const JobList<J, CAP>& L, int i) {
    return L.dev ? L.dev[i] : L.inl[i];
}

__device__ __forceinline__ int num_digits(const DevTables& T, int nl) {
    const int n = (nl + T.alpha - 1) / T.alpha;
    return n < T.dnum ? n : T.dnum;
}

// Two adjacent coefficients per thread through 16-byte accesses.
struct u64x2 {
    uint64_t x, y;
};
__device__ __forceinline__ u64x2 ld2(const uint64_t* p) { return *(const u64x2*)p; }
__device__ __forceinline__ void st2(uint64_t* p, uint64_t a, uint64_t b) { *(u64x2*)p = u64x2{a, b}; }

// (c0,c1) x (d0,d1) -> (c0d0, c0d1 + c1d0, c1d1)
// FIDESlib ElemenwiseBatchKernels.cu:1398-1438 binomialMult_.
__device__ __forceinline__ void tensor1(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, const PrimeInfo& P,
                                        uint64_t& r0, uint64_t& r1, uint64_t& r2) {
    uint64_t lo, hi;
    mul128(a0, b0, lo, hi);
    r0 = reduce128(lo, hi, P);
    mul128(a0, b1, lo, hi);
    mac128(lo, hi, a1, b0);
    r1 = reduce128(lo, hi, P);
    mul128(a1, b1, lo, hi);
    r2 = reduce128(lo, hi, P);
}

__global__ void __launch_bounds__(EW_BLOCK) k_tensor(DevTables T, TensorJobs L) {
    const TensorJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const PrimeInfo P = T.primes[i];
    const size_t x = (size_t)i * T.N + 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const u64x2 a0 = ld2(J.c0 + x), a1 = ld2(J.c1 + x), b0 = ld2(J.d0 + x), b1 = ld2(J.d1 + x);
    uint64_t r0, r1, r2, s0, s1, s2;
    tensor1(a0.x, a1.x, b0.x, b1.x, P, r0, r1, r2);
    tensor1(a0.y, a1.y, b0.y, b1.y, P, s0, s1, s2);
    st2(J.o0 + x, r0, s0);
    st2(J.o1 + x, r1, s1);
    st2(J.o2 + x, r2, s2);
}

// ModUp base conversion (OpenFHE ApproxSwitchCRTBasis; FIDESlib
// src/CKKS/Conv.cu:144-287 DecompAndModUpConv).  Input t_i are the digit's
// coefficient limbs already multiplied by [(D/q_i)^-1]_{q_i} (folded into the
// INTT).  For every position outside the digit:
//   y = sum_i t_i [D/q_i]_p mod p   (exact, no centering).
// A block covers CONV_X coefficients of one digit: the digit's t_i for those
// coefficients are staged once in LDS, then CONV_G thread groups split the
// digit's target positions (Q outside the digit, then P) round-robin, so each
// input is read from HBM once and each output written once, with enough waves
// per launch to hide latency.
constexpr int CONV_X = 64;
constexpr int CONV_G = EW_BLOCK / CONV_X;

constexpr int CONV_MAXT = 96;  // targets per digit staged in LDS (L+1+K <= 96)

template <int MAXA>
__global__ void __launch_bounds__(EW_BLOCK) k_modup_conv(DevTables T, ModUpJobs L) {
    __shared__ uint64_t ts[MAXA * CONV_X];
    __shared__ uint64_t ms[MAXA * CONV_MAXT];  // conversion matrix [i][target]
    const ModUpJob J = get_job(L, blockIdx.z);
    const int d = blockIdx.y;
    const int nl = J.level + 1;
    if (d >= num_digits(T, nl)) return;
    const int start = d * T.alpha;
    const int size = (nl - start) < T.alpha ? (nl - start) : T.alpha;
    const int ntarget = nl - size + T.K;
    const int N = T.N;
    const int x0 = blockIdx.x * CONV_X;
    for (int i = threadIdx.x; i < size * CONV_X; i += EW_BLOCK)
        ts[i] = J.t[(size_t)(start + i / CONV_X) * N + x0 + i % CONV_X];
    const uint64_t* M = T.modup_mat + ((size_t)(d * T.alpha + size - 1) * T.alpha) * T.np;
    for (int i = threadIdx.x; i < size * ntarget; i += EW_BLOCK) {
        const int k = i % ntarget, pos = k < start ? k : k + size;
        ms[(i / ntarget) * CONV_MAXT + k] = M[(size_t)(i / ntarget) * T.np + pos_pid(pos, nl, 0, T.L + 1)];
    }
    __syncthreads();
    const int xl = threadIdx.x % CONV_X, g = threadIdx.x / CONV_X;
    uint64_t tv[MAXA];
#pragma unroll
    for (int i = 0; i < MAXA; ++i)
        if (i < size) tv[i] = ts[i * CONV_X + xl];
    const int npos = nl + T.K;
    uint64_t* out = J.ext + (size_t)d * npos * N + x0 + xl;
    for (int k = g; k < ntarget; k += CONV_G) {
        const int pos = k < start ? k : k + size;  // skips the digit's own positions
        const int pid = pos_pid(pos, nl, 0, T.L + 1);
        const PrimeInfo P = T.primes[pid];
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (int i = 0; i < MAXA; ++i) {
            if (i < size) {
                mac128(lo, hi, tv[i], ms[i * CONV_MAXT + k]);
                if ((i & 7) == 7 && i + 1 < size) {
                    lo = reduce128(lo, hi, P);
                    hi = 0;
                }
            }
        }
        out[(size_t)pos * N] = reduce128(lo, hi, P);
    }
}

// Key-switch inner product (OpenFHE EvalFastKeySwitchCoreExt; FIDESlib
// ElemenwiseBatchKernels.cu:221-297 fusedDotKSK_2_):
//   acc_a[pos] = sum_d D_d[pos] A_d[pid],  acc_b[pos] = sum_d D_d[pos] B_d[pid].
// D_d at the digit's own positions is the evaluation-form input itself.
__global__ void __launch_bounds__(EW_BLOCK) k_ksk_inner(DevTables T, KskJobs L) {
    const KskJob J = get_job(L, blockIdx.z);
    const int nl = J.level + 1;
    const int npos = nl + T.K;
    const int pos = blockIdx.y;
    if (pos >= npos) return;
    const int pid = pos_pid(pos, nl, 0, T.L + 1);
    const PrimeInfo P = T.primes[pid];
    const int nd = num_digits(T, nl);
    const int N = T.N;
    const size_t x = 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const int own_d = pos < nl ? pos / T.alpha : -1;
    uint64_t la0 = 0, ha0 = 0, lb0 = 0, hb0 = 0, la1 = 0, ha1 = 0, lb1 = 0, hb1 = 0;
    for (int d = 0; d < nd; ++d) {
        const u64x2 v = (d == own_d) ? ld2(J.own + (size_t)pos * N + x) : ld2(J.ext + ((size_t)d * npos + pos) * N + x);
        const size_t kx = ((size_t)d * T.np + pid) * N + x;
        const u64x2 ka = ld2(J.ka + kx), kb = ld2(J.kb + kx);
        mac128(la0, ha0, v.x, ka.x);
        mac128(lb0, hb0, v.x, kb.x);
        mac128(la1, ha1, v.y, ka.y);
        mac128(lb1, hb1, v.y, kb.y);
        if ((d & 7) == 7 && d + 1 < nd) {
            la0 = reduce128(la0, ha0, P); ha0 = 0;
            lb0 = reduce128(lb0, hb0, P); hb0 = 0;
            la1 = reduce128(la1, ha1, P); ha1 = 0;
            lb1 = reduce128(lb1, hb1, P); hb1 = 0;
        }
    }
    st2(J.acc_a + (size_t)pos * N + x, reduce128(la0, ha0, P), reduce128(la1, ha1, P));
    st2(J.acc_b + (size_t)pos * N + x, reduce128(lb0, hb0, P), reduce128(lb1, hb1, P));
}

// ModDown base conversion P -> Q_l (OpenFHE ApproxModDown's ApproxSwitchCRTBasis;
// FIDESlib src/CKKS/Conv.cu:26-132 ModDown2).  t_k already carry
// [(P/p_k)^-1]_{p_k} (folded into the INTT):  y_i = sum_k t_k [P/p_k]_{q_i}.
// Same blocking as the ModUp conversion: CONV_X coefficients per block, the K
// inputs and the matrix staged in LDS, CONV_G groups splitting the outputs.
//
// RESCALE: ModDown followed by rescale, merged by linearity.  With
//   x_i = (acc_i - NTT(y_i)) P^-1 + add_i               (ModDown, then add)
//   r_i = (x_i - NTT(Switch_i(INTT(x_l)))) q_l^-1        (rescale)
// the top limb in coefficient form is INTT(x_l) = (A_l - y_l) P^-1 + B_l with
// A_l = INTT(acc_l), B_l = INTT(add_l), and
//   r_i = (acc_i P^-1 + add_i - NTT(z_i)) q_l^-1,  z_i = y_i P^-1 + Switch_i(top).
// This kernel writes z_i (i < l); one NTT pass pair with the
// STORE_MODDOWN_RESCALE epilogue finishes.  Every step is an exact identity
// mod q_i, so the residues equal the two-step result.
template <int MAXK, bool RESCALE>
__global__ void __launch_bounds__(EW_BLOCK) k_moddown_conv(DevTables T, ModDownJobs L) {
    __shared__ uint64_t ts[MAXK * CONV_X];
    __shared__ uint64_t ms[MAXK * CONV_MAXT];  // [k][i]
    const ModDownJob J = get_job(L, blockIdx.y);
    const int nl = J.level + 1;
    const int N = T.N;
    const int x0 = blockIdx.x * CONV_X;
    for (int i = threadIdx.x; i < T.K * CONV_X; i += EW_BLOCK)
        ts[i] = J.t[(size_t)(i / CONV_X) * N + x0 + i % CONV_X];
    for (int i = threadIdx.x; i < T.K * nl; i += EW_BLOCK)
        ms[(i / nl) * CONV_MAXT + i % nl] = T.moddown_mat[(size_t)(i / nl) * T.np + i % nl];
    __syncthreads();
    const int xl = threadIdx.x % CONV_X, g = threadIdx.x / CONV_X;
    uint64_t tv[MAXK];
#pragma unroll
    for (int k = 0; k < MAXK; ++k)
        if (k < T.K) tv[k] = ts[k * CONV_X + xl];
    auto conv = [&](int i, const PrimeInfo& P) {
        uint64_t lo = 0, hi = 0;
#pragma unroll
        for (int k = 0; k < MAXK; ++k) {
            if (k < T.K) {
                mac128(lo, hi, tv[k], ms[k * CONV_MAXT + i]);
                if ((k & 7) == 7 && k + 1 < T.K) {
                    lo = reduce128(lo, hi, P);
                    hi = 0;
                }
            }
        }
        return reduce128(lo, hi, P);
    };
    if constexpr (!RESCALE) {
        for (int i = g; i < nl; i += CONV_G) J.y[(size_t)i * N + x0 + xl] = conv(i, T.primes[i]);
    } else {
        const int l = J.level;
        const PrimeInfo Pl = T.primes[l];
        const ShoupConst pl = T.pinv[l];
        // top limb (every group recomputes it for its own coefficient)
        const uint64_t yl = conv(l, Pl);
        const uint64_t top = addmod(mul_shoup(submod(J.top_acc[x0 + xl], yl, Pl.p), pl.w, pl.wp, Pl.p),
                                    J.top_add[x0 + xl], Pl.p);
        const bool neg = top > (Pl.p >> 1);
        for (int i = g; i < l; i += CONV_G) {
            const PrimeInfo P = T.primes[i];
            const ShoupConst pi = T.pinv[i];
            uint64_t sw = reduce64(top, P);  // centered lift of top into q_i
            if (neg) sw = submod(sw, T.qlmodq[(size_t)l * (T.L + 1) + i], P.p);
            J.y[(size_t)i * N + x0 + xl] = addmod(mul_shoup(conv(i, P), pi.w, pi.wp, P.p), sw, P.p);
        }
    }
}

// Automorphism X -> X^k on evaluation-form limbs, as a gather:
// out[brev(j)] = in[brev(((2j+1) k mod 2N) >> 1)]  (OpenFHE PrecomputeAutoMap;
// FIDESlib src/Rotation.cuh:14-36 is the equivalent scatter with k^-1).
__global__ void __launch_bounds__(EW_BLOCK) k_automorph(DevTables T, AutoJobs L) {
    const AutoJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const uint32_t y = blockIdx.x * EW_BLOCK + threadIdx.x;
    const uint32_t j = brev_bits(y, T.logN);
    const uint64_t mask = 2ull * T.N - 1;
    const uint32_t idx = (uint32_t)((((2ull * j + 1) * J.k) & mask) >> 1);
    const uint32_t src = brev_bits(idx, T.logN);
    J.out[(size_t)i * T.N + y] = J.in[(size_t)i * T.N + src];
}

template <int OP>
__global__ void __launch_bounds__(EW_BLOCK) k_elementwise(DevTables T, EwJobs L) {
    const EwJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const PrimeInfo P = T.primes[i];
    const size_t x = (size_t)i * T.N + blockIdx.x * EW_BLOCK + threadIdx.x;
    const uint64_t a = J.a[x], b = J.b[x];
    uint64_t r;
    if constexpr (OP == EW_ADD)
        r = addmod(a, b, P.p);
    else if constexpr (OP == EW_SUB)
        r = submod(a, b, P.p);
    else
        r = mulmod(a, b, P);
    J.o[x] = r;
}

// ---------------------------------------------------------- launchers ----

hipError_t launch_tensor(const DevTables& T, const TensorJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tensor, dim3(T.N / (2 * EW_BLOCK), max_nl, L.n), dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

hipError_t launch_modup_conv(const DevTables& T, const ModUpJobs& L, int max_nd, hipStream_t s) {
    if (L.n <= 0 || max_nd <= 0) return hipSuccess;
    if (T.L + 1 + T.K > CONV_MAXT) return hipErrorInvalidValue;
    dim3 g(T.N / CONV_X, max_nd, L.n), b(EW_BLOCK);
    if (T.alpha <= 4)
        hipLaunchKernelGGL(k_modup_conv<4>, g, b, 0, s, T, L);
    else if (T.alpha <= 8)
        hipLaunchKernelGGL(k_modup_conv<8>, g, b, 0, s, T, L);
    else if (T.alpha <= 16)
        hipLaunchKernelGGL(k_modup_conv<16>, g, b, 0, s, T, L);
    else if (T.alpha <= 32)
        hipLaunchKernelGGL(k_modup_conv<32>, g, b, 0, s, T, L);
    else if (T.alpha <= 64)
        hipLaunchKernelGGL(k_modup_conv<64>, g, b, 0, s, T, L);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ksk_inner(const DevTables& T, const KskJobs& L, int max_npos, hipStream_t s) {
    if (L.n <= 0 || max_npos <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ksk_inner, dim3(T.N / (2 * EW_BLOCK), max_npos, L.n), dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

hipError_t launch_moddown_conv(const DevTables& T, const ModDownJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0) return hipSuccess;
    if (max_nl > CONV_MAXT) return hipErrorInvalidValue;
    const bool rs = (L.dev == nullptr) ? (L.inl[0].top_acc != nullptr) : false;
    return launch_moddown_conv_mode(T, L, rs, s);
}

hipError_t launch_moddown_conv_mode(const DevTables& T, const ModDownJobs& L, bool rescale, hipStream_t s) {
    if (L.n <= 0) return hipSuccess;
    dim3 g(T.N / CONV_X, L.n), b(EW_BLOCK);
#define MD_LAUNCH(KK)                                                      \
    if (rescale)                                                           \
        hipLaunchKernelGGL((k_moddown_conv<KK, true>), g, b, 0, s, T, L);  \
    else                                                                   \
        hipLaunchKernelGGL((k_moddown_conv<KK, false>), g, b, 0, s, T, L);
    if (T.K <= 4) {
        MD_LAUNCH(4)
    } else if (T.K <= 8) {
        MD_LAUNCH(8)
    } else if (T.K <= 16) {
        MD_LAUNCH(16)
    } else if (T.K <= 32) {
        MD_LAUNCH(32)
    } else {
        return hipErrorInvalidValue;
    }
#undef MD_LAUNCH
    return hipGetLastError();
}

hipError_t launch_automorph(const DevTables& T, const AutoJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_automorph, dim3(T.N / EW_BLOCK, max_nl, L.n), dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

hipError_t launch_elementwise(const DevTables& T, const EwJobs& L, int max_nl, int op, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    dim3 g(T.N / EW_BLOCK, max_nl, L.n), b(EW_BLOCK);
    if (op == EW_ADD)
        hipLaunchKernelGGL(k_elementwise<EW_ADD>, g, b, 0, s, T, L);
    else if (op == EW_SUB)
        hipLaunchKernelGGL(k_elementwise<EW_SUB>, g, b, 0, s, T, L);
    else
        hipLaunchKernelGGL(k_elementwise<EW_MUL>, g, b, 0, s, T, L);
    return hipGetLastError();
}

}  // namespace fides
