// Elementwise, base-conversion, key-switch and automorphism kernels (gfx950).
//
// All are HBM streams over limb-major polynomials: 256-thread blocks along the
// coefficient axis, the limb (position) along blockIdx.y and the job along
// blockIdx.z, so every access is a fully coalesced stream (16-byte lanes in
// the tensor and KSK kernels, which handle two coefficients per thread).  Sums of products are accumulated exactly
// in 128 bits and reduced once (Barrett, valid below 2^(n+63); sums of more
// than 8 products are reduced in chunks of 8).
#include <cstdlib>

#include "engine.hpp"

namespace fides {

constexpr int EW_BLOCK = 256;

__device__ __forceinline__ int num_digits(const DevTables& T, int nl) {
    const int n = (nl + T.alpha - 1) / T.alpha;
    return n < T.dnum ? n : T.dnum;
}

// Two adjacent coefficients per thread through 16-byte accesses.
struct u64x2 {
    uint64_t x, y;
};
typedef uint64_t v2u64 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) v2u64 gv2u64;
__device__ __forceinline__ u64x2 ld2(const uint64_t* p) {
    const v2u64 v = *(const gv2u64*)p;
    return u64x2{v.x, v.y};
}
__device__ __forceinline__ void st2(uint64_t* p, uint64_t a, uint64_t b) { *(gv2u64*)p = v2u64{a, b}; }

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
    const PrimeInfo P = cld(T.primes + i);
    const size_t x = (size_t)i * T.N + 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const u64x2 a0 = ld2(J.c0 + x), a1 = ld2(J.c1 + x), b0 = ld2(J.d0 + x), b1 = ld2(J.d1 + x);
    uint64_t r0, r1, r2, s0, s1, s2;
    tensor1(a0.x, a1.x, b0.x, b1.x, P, r0, r1, r2);
    tensor1(a0.y, a1.y, b0.y, b1.y, P, s0, s1, s2);
    st2(J.o0 + x, r0, s0);
    st2(J.o1 + x, r1, s1);
    st2(J.o2 + x, r2, s2);
}

// ---------------------------------------------------------------------------
// Base conversions.  Layout: one lane per coefficient (CV_B consecutive
// coefficients per block), so every input and output access is a coalesced
// 512-byte row per wave; each lane loads its digit inputs once and produces
// its coefficient of every target limb the block owns.  Blocks of one
// coefficient range split the targets round-robin in TG groups (TG is chosen
// at launch so that small launches still fill the GPU; the inputs are then
// re-read TG times, from L2).  Matrix entries are wave-uniform and come in
// through scalar loads.
constexpr int CV_B = 256;
constexpr int CV_MAXT = 96;  // positions per polynomial (L+1+K <= 96, checked at context creation)
constexpr int CV_MIN_BLOCKS = 2048;

// Carry-free sum of products: residues (< 2^60) are split into 30-bit halves,
// so every partial product is below 2^60 and up to 16 terms accumulate in four
// 64-bit registers with one v_mad_u64_u32 each and no carry handling;
// combine() forms the exact 128-bit sum once.
struct Acc4 {
    uint64_t s00 = 0, s01 = 0, s10 = 0, s11 = 0;
};
struct Half2 {
    uint32_t lo, hi;  // x = hi * 2^30 + lo
};
__device__ __forceinline__ Half2 split30(uint64_t x) { return Half2{(uint32_t)(x & 0x3fffffffu), (uint32_t)(x >> 30)}; }
__device__ __forceinline__ void mac_split(Acc4& a, Half2 x, Half2 y) {
    a.s00 = (uint64_t)x.lo * y.lo + a.s00;
    a.s01 = (uint64_t)x.lo * y.hi + a.s01;
    a.s10 = (uint64_t)x.hi * y.lo + a.s10;
    a.s11 = (uint64_t)x.hi * y.hi + a.s11;
}
// X = s11 2^60 + (s01 + s10) 2^30 + s00 as (lo, hi)
__device__ __forceinline__ void combine(const Acc4& a, uint64_t& lo, uint64_t& hi) {
    lo = a.s00;
    hi = 0;
    uint64_t t = a.s01 << 30;
    lo += t;
    hi += (a.s01 >> 34) + (lo < t);
    t = a.s10 << 30;
    lo += t;
    hi += (a.s10 >> 34) + (lo < t);
    t = a.s11 << 60;
    lo += t;
    hi += (a.s11 >> 4) + (lo < t);
}

// Sum over i < n (n <= A, wave-uniform) of x_i * m[i * stride], reduced mod
// P.  The matrix column is read with scalar loads (uniform addresses).  ONE:
// the host proved the exact sum fits one Barrett reduction, so split 30-bit
// products accumulate carry-free; otherwise 128-bit MACs are reduced every 8
// terms.

template <int A>
struct MatRow {
    uint64_t v[A];
};

// Exact 128-bit MACs of terms [C0, A) in chunks of at most 8 (16 scalar
// registers of matrix at a time), reduced between chunks: 8 products of a
// residue < 2^60 by one < p stay below 2^(bits(p) + 63).
template <int A, int C0>
__device__ __forceinline__ void mac_rows(uint64_t& lo, uint64_t& hi, const uint64_t* x, const uint64_t* m,
                                         const PrimeInfo& P) {
    if constexpr (C0 < A) {
        constexpr int C = (A - C0) < 8 ? (A - C0) : 8;
        const MatRow<C> row = cld(reinterpret_cast<const MatRow<C>*>(m + C0));
#pragma unroll
        for (int i = 0; i < C; ++i) mac128(lo, hi, x[C0 + i], row.v[i]);
        if constexpr (C0 + C < A) {
            lo = reduce128(lo, hi, P);
            hi = 0;
            mac_rows<A, C0 + C>(lo, hi, x, m, P);
        }
    }
}

// Sum over i < A of x_i * m[i], reduced mod P, where m is one zero-padded
// matrix row (A entries, read with wide scalar loads).  Inputs beyond the
// live count are zero (see load_inputs).
template <int A, bool ONE>
__device__ __forceinline__ uint64_t conv_sum(const uint64_t* x, const Half2* xh, const uint64_t* m,
                                             const PrimeInfo& P) {
    uint64_t lo = 0, hi = 0;
    if constexpr (ONE) {
        const MatRow<A> row = cld(reinterpret_cast<const MatRow<A>*>(m));
        Acc4 acc;
#pragma unroll
        for (int i = 0; i < A; ++i) mac_split(acc, xh[i], split30(row.v[i]));
        combine(acc, lo, hi);
    } else {
        mac_rows<A, 0>(lo, hi, x, m, P);
    }
    return reduce128(lo, hi, P);
}

// The n (<= A, wave-uniform) inputs of this lane's coefficient, limbs
// src[i * N + x]; entries i >= n are zero.  Loads are unconditional (the row
// index is clamped) so they issue back to back.
template <int A, bool ONE>
__device__ __forceinline__ void load_inputs(const uint64_t* src, size_t N, size_t x, int n, uint64_t* tv, Half2* th) {
#pragma unroll
    for (int i = 0; i < A; ++i) tv[i] = gp(src)[(size_t)(i < n ? i : n - 1) * N + x];
#pragma unroll
    for (int i = 0; i < A; ++i) {
        tv[i] = i < n ? tv[i] : 0;
        th[i] = split30(tv[i]);
    }
}

// ModUp base conversion (OpenFHE ApproxSwitchCRTBasis; FIDESlib
// src/CKKS/Conv.cu:144-287 DecompAndModUpConv).  Input t_i are the digit's
// coefficient limbs already multiplied by [(D/q_i)^-1]_{q_i} (folded into the
// INTT).  For every position outside the digit:
//   y = sum_i t_i [D/q_i]_p mod p   (exact, no centering).
// Grid: (N / CV_B, digits x TG, jobs).  A is the template bound of the digit
// size (exact for the common sizes, so the sums unroll without guards).
template <int A, bool ONE>
__global__ void __launch_bounds__(CV_B) k_modup_conv(DevTables T, ModUpJobs L, int TG) {
    const ModUpJob J = get_job(L, blockIdx.z);
    const int d = blockIdx.y / TG, tg = blockIdx.y % TG;
    const int nl = __builtin_amdgcn_readfirstlane(J.level) + 1;
    if (d >= num_digits(T, nl)) return;
    const int start = d * T.alpha;
    const int size = __builtin_amdgcn_readfirstlane((nl - start) < T.alpha ? (nl - start) : T.alpha);
    const int ntarget = nl - size + T.K;
    const int N = T.N;
    const size_t x = (size_t)blockIdx.x * CV_B + threadIdx.x;
    uint64_t tv[A];
    Half2 th[A];
    load_inputs<A, ONE>(J.t + (size_t)start * N, N, x, size, tv, th);
    // conversion matrix of this digit size: rows [prime id][A]
    const uint64_t* M = T.modup_mat + (size_t)(d * T.alpha + size - 1) * T.np * A;
    const int npos = nl + T.K;
    gu64* out = gp(J.ext) + (size_t)d * npos * N + x;
    // two targets per iteration: independent MAC chains for the scheduler to
    // interleave (positions skip the digit's own ones)
    int k = tg;
    for (; k + TG < ntarget; k += 2 * TG) {
        const int pos0 = k < start ? k : k + size, pos1 = k + TG < start ? k + TG : k + TG + size;
        const int pid0 = pos_pid(pos0, nl, 0, T.L + 1), pid1 = pos_pid(pos1, nl, 0, T.L + 1);
        const uint64_t r0 = conv_sum<A, ONE>(tv, th, M + (size_t)pid0 * A, cld(T.primes + pid0));
        const uint64_t r1 = conv_sum<A, ONE>(tv, th, M + (size_t)pid1 * A, cld(T.primes + pid1));
        out[(size_t)pos0 * N] = r0;
        out[(size_t)pos1 * N] = r1;
    }
    if (k < ntarget) {
        const int pos = k < start ? k : k + size;
        const int pid = pos_pid(pos, nl, 0, T.L + 1);
        out[(size_t)pos * N] = conv_sum<A, ONE>(tv, th, M + (size_t)pid * A, cld(T.primes + pid));
    }
}

// ModUp base conversion on the matrix cores (same sums as k_modup_conv).
// Residues are split into 8 balanced signed bytes, x = sum_a d_a 2^(8a) with
// d_a in [-128, 127] (the bytes of x + 0x80..80, each less 128: two 64-bit
// operations per residue), so
//   sum_i x_i m_i = sum_(c=0..14) 2^(8c) S_c,   S_c = sum_i sum_(a+b=c) dx_(i,a) dm_(i,b),
// |S_c| < 2^21.  S_c over a tile of 16 targets x 16 coefficients is an int8
// product over K = (limb, byte): MS v_mfma_i32_16x16x32_i8 of 4 limbs each.
// Lane l supplies row / column l mod 16 and K entries 8 (l / 16) .. +7, so its
// B operand is the balanced word of one input residue and its A operand the
// host-built word of dm_(i, c-a), a = 0..7 (DevTables::modup_mfma), staged in
// LDS once per block; lane l receives rows 4 (l / 16) + v, column l mod 16.
// The S_c of each group of four c fold into a 64-bit partial sum (one
// multiply-add by 2^(8 (c mod 4))), the groups at 2^0, 2^32, 2^64, 2^96 give
// the exact 128-bit sum, reduced once (the host proved it below
// 2^(bits(p) + 63), as for k_modup_conv).  Grid: (coefficient ranges, digits x
// target tiles TT, jobs); each wave runs CT tiles of 16 coefficients.
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int MF_WAVES = 4;

__device__ __forceinline__ uint64_t balanced8(uint64_t x) {
    constexpr uint64_t C = 0x8080808080808080ull;
    return (x + C) ^ C;
}

template <int MS>
__global__ void __launch_bounds__(64 * MF_WAVES) k_modup_mfma(DevTables T, ModUpJobs L, int TT, int CT) {
    constexpr int SP = 4 * MS;
    __shared__ uint64_t As[15 * MS * 64];
    const ModUpJob J = get_job(L, blockIdx.z);
    const int d = blockIdx.y / TT, tt = blockIdx.y % TT;
    const int nl = __builtin_amdgcn_readfirstlane(J.level) + 1;
    if (d >= num_digits(T, nl)) return;
    const int start = d * T.alpha;
    const int size = __builtin_amdgcn_readfirstlane((nl - start) < T.alpha ? (nl - start) : T.alpha);
    const int ntarget = nl - size + T.K;
    if (tt * 16 >= ntarget) return;
    const int tid = threadIdx.x;
    {
        const gu64* TM = gp(T.modup_mfma) + (size_t)(d * T.alpha + size - 1) * 15 * T.np * SP;
        for (int e = tid; e < 15 * MS * 64; e += 64 * MF_WAVES) {
            const int ln = e & 63, cs = e >> 6, c = cs / MS, st = cs % MS;
            const int i = 4 * st + (ln >> 4), j = tt * 16 + (ln & 15);
            uint64_t w = 0;
            if (i < size && j < ntarget) {
                const int pos = j < start ? j : j + size;
                w = TM[((size_t)c * T.np + pos_pid(pos, nl, 0, T.L + 1)) * SP + i];
            }
            As[e] = w;
        }
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int N = T.N;
    PrimeInfo P[4];
    int opos[4];
    bool ok[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int j = tt * 16 + 4 * kg + v;
        ok[v] = j < ntarget;
        const int pos = j < start ? j : j + size;
        opos[v] = pos;
        P[v] = T.primes[ok[v] ? pos_pid(pos, nl, 0, T.L + 1) : 0];
    }
    const gu64* src = gp(J.t) + (size_t)start * N;
    gu64* out = gp(J.ext) + (size_t)d * (nl + T.K) * N;
    const int ntile = N / 16;
    const int ct0 = (blockIdx.x * MF_WAVES + wave) * CT;
    for (int ct = ct0; ct < ct0 + CT && ct < ntile; ++ct) {
        const int x = ct * 16 + m;
        uint64_t b[MS];
#pragma unroll
        for (int st = 0; st < MS; ++st) {
            const int i = 4 * st + kg;
            const uint64_t r = src[(size_t)(i < size ? i : size - 1) * N + x];
            b[st] = i < size ? balanced8(r) : 0;
        }
        // c in pairs (2k, 2k+1): two independent accumulator chains, folded in
        // 32 bits (S_2k + 2^8 S_2k+1 < 2^30) and into u_(k/2) at 2^(16 (k mod 2))
        // (|u| < 2^47); the four u at 2^0, 2^32, 2^64, 2^96 give the sum
        int64_t u[4][4] = {};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            asm volatile("" ::: "memory");  // keep each pair's A words loaded just in time
            v4i S0 = {0, 0, 0, 0}, S1 = {0, 0, 0, 0};
#pragma unroll
            for (int st = 0; st < MS; ++st) {
                S0 = __builtin_amdgcn_mfma_i32_16x16x32_i8((long)As[(2 * k * MS + st) * 64 + lane], (long)b[st], S0,
                                                           0, 0, 0);
                if (2 * k + 1 < 15)
                    S1 = __builtin_amdgcn_mfma_i32_16x16x32_i8((long)As[((2 * k + 1) * MS + st) * 64 + lane],
                                                               (long)b[st], S1, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
                u[k / 2][v] += (int64_t)(S0[v] + S1[v] * 256) * (int64_t)(k % 2 ? 65536 : 1);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const __int128 y = (__int128)u[0][v] + ((__int128)u[1][v] << 32) + ((__int128)u[2][v] << 64) +
                               ((__int128)u[3][v] << 96);
            if (ok[v]) out[(size_t)opos[v] * N + x] = reduce128((uint64_t)y, (uint64_t)(y >> 64), P[v]);
        }
    }
}

// Key-switch inner product (OpenFHE EvalFastKeySwitchCoreExt; FIDESlib
// ElemenwiseBatchKernels.cu:221-297 fusedDotKSK_2_):
//   acc_a[pos] = sum_d D_d[pos] A_d[pid],  acc_b[pos] = sum_d D_d[pos] B_d[pid].
// D_d at the digit's own positions is the evaluation-form input itself.
// Generic form (any number of digits): 128-bit MACs reduced every 8 digits.
__global__ void __launch_bounds__(EW_BLOCK) k_ksk_inner(DevTables T, KskJobs L) {
    const KskJob J = get_job(L, blockIdx.z);
    const int nl = J.level + 1;
    const int npos = nl + T.K;
    const int pos = blockIdx.y;
    if (pos >= npos) return;
    const int pid = pos_pid(pos, nl, 0, T.L + 1);
    const PrimeInfo P = cld(T.primes + pid);
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

// Up to D digits (the usual dnum <= 4): all digit and key loads issue back to
// back (indices clamped, digits beyond nd zeroed), and a sum of D products of
// residues below p stays below 2^(bits(p) + 63), so one Barrett step per
// output.  SHARED: every job of the list has the same ext / own / level (the
// rotations of a hoisted set) -- a block reads the digits once and runs all
// jobs' keys against them.
template <int D, bool SHARED>
__global__ void __launch_bounds__(EW_BLOCK) k_ksk_inner_d(DevTables T, KskJobs L) {
    const KskJob J0 = get_job(L, SHARED ? 0 : blockIdx.z);
    const int nl = __builtin_amdgcn_readfirstlane(J0.level) + 1;
    const int npos = nl + T.K;
    const int pos = blockIdx.y;
    if (pos >= npos) return;
    const int pid = pos_pid(pos, nl, 0, T.L + 1);
    const PrimeInfo P = cld(T.primes + pid);
    const int nd = num_digits(T, nl);
    const int N = T.N;
    const size_t x = 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const int own_d = pos < nl ? pos / T.alpha : -1;
    u64x2 v[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int dd = d < nd ? d : nd - 1;
        v[d] = ld2(dd == own_d ? J0.own + (size_t)pos * N + x : J0.ext + ((size_t)dd * npos + pos) * N + x);
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (d >= nd) v[d] = u64x2{0, 0};
    const int nj = SHARED ? L.n : 1;
    auto load_keys = [&](const KskJob& J, u64x2* ka, u64x2* kb) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int dd = d < nd ? d : nd - 1;
            const size_t kx = ((size_t)dd * T.np + pid) * N + x;
            ka[d] = ld2(J.ka + kx);
            kb[d] = ld2(J.kb + kx);
        }
    };
    // keys of the next job are loaded before this job's MACs (SHARED loops
    // over many keys; the loads of one overlap the arithmetic of the other)
    KskJob J = J0;
    u64x2 ka[D], kb[D];
    load_keys(J, ka, kb);
    for (int j = 0; j < nj; ++j) {
        KskJob Jn = J;
        u64x2 na[D], nb[D];
        if (SHARED && j + 1 < nj) {
            Jn = get_job(L, j + 1);
            load_keys(Jn, na, nb);
        }
        uint64_t la0 = 0, ha0 = 0, lb0 = 0, hb0 = 0, la1 = 0, ha1 = 0, lb1 = 0, hb1 = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            mac128(la0, ha0, v[d].x, ka[d].x);
            mac128(lb0, hb0, v[d].x, kb[d].x);
            mac128(la1, ha1, v[d].y, ka[d].y);
            mac128(lb1, hb1, v[d].y, kb[d].y);
        }
        st2(J.acc_a + (size_t)pos * N + x, reduce128(la0, ha0, P), reduce128(la1, ha1, P));
        st2(J.acc_b + (size_t)pos * N + x, reduce128(lb0, hb0, P), reduce128(lb1, hb1, P));
        if (SHARED && j + 1 < nj) {
            J = Jn;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ka[d] = na[d];
                kb[d] = nb[d];
            }
        }
    }
}

// Key-stationary form for a batch whose jobs share keys (batched HMult: the
// relinearization key; batched rotations: a few rotation keys): the jobs are
// grouped by key and cut into chunks of one key, chunk z being jobs
// chunks[z] .. chunks[z+1]-1.  A block owns one prime id, a tile of
// coefficients and a chunk, loads the key values there once and runs the
// chunk's jobs against them -- each key is read once per chunk instead of
// once per job.  A job whose level does not reach a Q prime skips it; the
// special primes sit at positions l+1.. of each job.
template <int D>
__global__ void __launch_bounds__(EW_BLOCK) k_ksk_keystat(DevTables T, KskJobs L, const int* chunks) {
    const int pid = blockIdx.y;
    const int j0 = chunks[blockIdx.z], j1 = chunks[blockIdx.z + 1];
    const KskJob K0 = get_job(L, j0);
    const PrimeInfo P = cld(T.primes + pid);
    const int N = T.N;
    const size_t x = 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    u64x2 ka[D], kb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int dk = d < T.dnum ? d : T.dnum - 1;  // the key holds dnum digits (D is a bound)
        const size_t kx = ((size_t)dk * T.np + pid) * N + x;
        ka[d] = ld2(K0.ka + kx);
        kb[d] = ld2(K0.kb + kx);
    }
    // jobs in pairs: both jobs' digit loads are issued before either's
    // products, so a block keeps two jobs' reads in flight
    struct Slot {
        bool on;
        int pos, nd;
        uint64_t *a, *b;
        u64x2 v[D];
    };
    auto fetch = [&](int j, Slot& s) {
        const KskJob J = get_job(L, j);
        const int nl = __builtin_amdgcn_readfirstlane(J.level) + 1;
        s.on = !(pid < T.L + 1 && pid >= nl);  // off: a Q prime above this job's level
        if (!s.on) return;
        s.pos = pid < T.L + 1 ? pid : nl + (pid - (T.L + 1));
        const int npos = nl + T.K;
        s.nd = num_digits(T, nl);
        const int own_d = s.pos < nl ? s.pos / T.alpha : -1;
        s.a = J.acc_a;
        s.b = J.acc_b;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int dd = d < s.nd ? d : s.nd - 1;
            s.v[d] = ld2(dd == own_d ? J.own + (size_t)s.pos * N + x : J.ext + ((size_t)dd * npos + s.pos) * N + x);
        }
    };
    auto finish = [&](const Slot& s) {
        if (!s.on) return;
        uint64_t la0 = 0, ha0 = 0, lb0 = 0, hb0 = 0, la1 = 0, ha1 = 0, lb1 = 0, hb1 = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (d >= s.nd) continue;
            mac128(la0, ha0, s.v[d].x, ka[d].x);
            mac128(lb0, hb0, s.v[d].x, kb[d].x);
            mac128(la1, ha1, s.v[d].y, ka[d].y);
            mac128(lb1, hb1, s.v[d].y, kb[d].y);
        }
        st2(s.a + (size_t)s.pos * N + x, reduce128(la0, ha0, P), reduce128(la1, ha1, P));
        st2(s.b + (size_t)s.pos * N + x, reduce128(lb0, hb0, P), reduce128(lb1, hb1, P));
    };
    int j = j0;
    for (; j + 1 < j1; j += 2) {
        Slot s0, s1;
        fetch(j, s0);
        fetch(j + 1, s1);
        finish(s0);
        finish(s1);
    }
    if (j < j1) {
        Slot s0;
        fetch(j, s0);
        finish(s0);
    }
}

// Jobs per key-stationary chunk (FIDES_KSK_JC overrides).
int ksk_keystat_chunk() {
    static const int v = [] {
        const char* e = std::getenv("FIDES_KSK_JC");
        const int x = e ? std::atoi(e) : 16;
        return x > 0 ? x : 16;
    }();
    return v;
}

hipError_t launch_ksk_keystat(const DevTables& T, const KskJobs& L, const int* chunks, int nchunks, hipStream_t s) {
    if (L.n <= 0 || nchunks <= 0) return hipSuccess;
    const dim3 g(T.N / (2 * EW_BLOCK), T.np, nchunks);
    if (T.dnum <= 2)
        hipLaunchKernelGGL((k_ksk_keystat<2>), g, dim3(EW_BLOCK), 0, s, T, L, chunks);
    else if (T.dnum == 3)
        hipLaunchKernelGGL((k_ksk_keystat<3>), g, dim3(EW_BLOCK), 0, s, T, L, chunks);
    else if (T.dnum == 4)
        hipLaunchKernelGGL((k_ksk_keystat<4>), g, dim3(EW_BLOCK), 0, s, T, L, chunks);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ModDown base conversion P -> Q_l (OpenFHE ApproxModDown's ApproxSwitchCRTBasis;
// FIDESlib src/CKKS/Conv.cu:26-132 ModDown2).  t_k already carry
// [(P/p_k)^-1]_{p_k} (folded into the INTT):  y_i = sum_k t_k [P/p_k]_{q_i}.
// Same layout as the ModUp conversion; grid (N / CV_B, TG, jobs).
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
template <int KK, bool ONE, bool RESCALE>
__global__ void __launch_bounds__(CV_B) k_moddown_conv(DevTables T, ModDownJobs L, int TG) {
    const ModDownJob J = get_job(L, blockIdx.z);
    const int tg = blockIdx.y;
    const int nl = __builtin_amdgcn_readfirstlane(J.level) + 1;
    const int N = T.N;
    const int K = T.K;
    const size_t x = (size_t)blockIdx.x * CV_B + threadIdx.x;
    uint64_t tv[KK];
    Half2 th[KK];
    load_inputs<KK, ONE>(J.t, N, x, K, tv, th);
    // conversion matrix: rows [prime id][KK]
    const uint64_t* M = T.moddown_mat;
    if constexpr (!RESCALE) {
        int i = tg;
        for (; i + TG < nl; i += 2 * TG) {
            const uint64_t r0 = conv_sum<KK, ONE>(tv, th, M + (size_t)i * KK, cld(T.primes + i));
            const uint64_t r1 = conv_sum<KK, ONE>(tv, th, M + (size_t)(i + TG) * KK, cld(T.primes + i + TG));
            gp(J.y)[(size_t)i * N + x] = r0;
            gp(J.y)[(size_t)(i + TG) * N + x] = r1;
        }
        if (i < nl) gp(J.y)[(size_t)i * N + x] = conv_sum<KK, ONE>(tv, th, M + (size_t)i * KK, cld(T.primes + i));
    } else {
        const int l = nl - 1;
        const PrimeInfo Pl = cld(T.primes + l);
        const ShoupConst pl = cld(T.pinv + l);
        const uint64_t yl = conv_sum<KK, ONE>(tv, th, M + (size_t)l * KK, Pl);
        const uint64_t top =
            addmod(mul_shoup(submod(gp(J.top_acc)[x], yl, Pl.p), pl.w, pl.wp, Pl.p), gp(J.top_add)[x], Pl.p);
        const bool neg = top > (Pl.p >> 1);
        auto z = [&](int i) {
            const PrimeInfo P = cld(T.primes + i);
            const ShoupConst pi = cld(T.pinv + i);
            uint64_t sw = reduce64(top, P);  // centered lift of top into q_i
            if (neg) sw = submod(sw, cld(T.qlmodq + (size_t)l * (T.L + 1) + i), P.p);
            const uint64_t y = conv_sum<KK, ONE>(tv, th, M + (size_t)i * KK, P);
            return addmod(mul_shoup(y, pi.w, pi.wp, P.p), sw, P.p);
        };
        int i = tg;
        for (; i + TG < l; i += 2 * TG) {
            const uint64_t r0 = z(i), r1 = z(i + TG);
            gp(J.y)[(size_t)i * N + x] = r0;
            gp(J.y)[(size_t)(i + TG) * N + x] = r1;
        }
        if (i < l) gp(J.y)[(size_t)i * N + x] = z(i);
    }
}

template <int OP>
__global__ void __launch_bounds__(EW_BLOCK) k_elementwise(DevTables T, EwJobs L) {
    const EwJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const PrimeInfo P = cld(T.primes + i);
    const size_t x = (size_t)i * T.N + blockIdx.x * EW_BLOCK + threadIdx.x;
    const uint64_t a = gp(J.a)[x], b = gp(J.b)[x];
    uint64_t r;
    if constexpr (OP == EW_ADD)
        r = addmod(a, b, P.p);
    else if constexpr (OP == EW_SUB)
        r = submod(a, b, P.p);
    else
        r = mulmod(a, b, P);
    gp(J.o)[x] = r;
}

// Scalar ops: the constant of limb i comes from the kernel arguments (a
// wave-uniform index, so a scalar load).
template <int OP>
__global__ void __launch_bounds__(EW_BLOCK) k_scalar(DevTables T, EwJobs L, LimbConsts K) {
    const EwJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const uint64_t p = cld(T.primes + i).p;
    const ShoupConst w = K.c[i];
    const size_t x = (size_t)i * T.N + 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const u64x2 a = ld2(J.a + x);
    if constexpr (OP == EW_MUL_SCALAR)
        st2(J.o + x, mul_shoup(a.x, w.w, w.wp, p), mul_shoup(a.y, w.w, w.wp, p));
    else
        st2(J.o + x, addmod(a.x, w.w, p), addmod(a.y, w.w, p));
}

// Multiplication by the monomial X^m on evaluation-form limbs (FIDESlib
// Ciphertext::multMonomial, Ciphertext.cpp:1709-1760, multiplies by the NTT of
// X^m): slot j of limb i holds a(psi^e_j) with e_j = 2 brev(j) + 1, so it is
// multiplied by psi^(m e_j mod 2N), read from the bit-reversed twiddle table
// (psi^N = -1 gives the upper half).  J.b is unused; the exponent m comes in
// as a kernel argument.
__global__ void __launch_bounds__(EW_BLOCK) k_monomial(DevTables T, EwJobs L, uint32_t m) {
    const EwJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const uint64_t p = cld(T.primes + i).p;
    const uint32_t y = blockIdx.x * EW_BLOCK + threadIdx.x;
    const uint32_t mask = 2u * T.N - 1;
    const uint32_t e = (uint32_t)(((uint64_t)m * (2u * brev_bits(y, T.logN) + 1)) & mask);
    const uint32_t h = e & (T.N - 1);
    const ShoupConst w = T.tw_fwd[(size_t)i * T.N + brev_bits(h, T.logN)];
    const size_t x = (size_t)i * T.N + y;
    uint64_t r = mul_shoup(gp(J.a)[x], w.w, w.wp, p);
    if (e >= (uint32_t)T.N && r) r = p - r;
    gp(J.o)[x] = r;
}

// Inner sums sum_t A[b j + t] * rot_t(ct) of a BSGS linear transform (FIDESlib
// DotProductPtInternal, LinearTransform.cu): one coefficient per thread, G
// giant steps per job, so each baby step is read once per job instead of
// once per giant step; both parts in one pass, so every diagonal is read
// once.  Exact 128-bit sums, reduced every 8 terms (8 products of residues
// below p < 2^60 stay below 2^(bits(p)+63), the Barrett bound).  Pointer
// tables are read with scalar loads (uniform index).
template <int G>
__global__ void __launch_bounds__(EW_BLOCK) k_dot_bsgs(DevTables T, DotJobs L) {
    const DotJob J = get_job(L, blockIdx.z);
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const bool special = i >= J.nq;
    const PrimeInfo P = cld(T.primes + (special ? T.L + 1 + (i - J.nq) : i));
    const size_t x = (size_t)i * T.N + blockIdx.x * EW_BLOCK + threadIdx.x;
    const size_t xw = special ? x - (size_t)J.nq * T.N : x;
    const uint64_t* const* wt = special ? J.wp : J.w;
    uint64_t l0[G], h0[G], l1[G], h1[G];
#pragma unroll
    for (int j = 0; j < G; ++j) l0[j] = h0[j] = l1[j] = h1[j] = 0;
    // baby steps in pairs: the loads of both issue before their MACs, so
    // twice the reads are in flight (b is even except for a ragged last one)
    for (int t = 0; t < J.b; t += 2) {
        const bool two = t + 1 < J.b;
        const int t1 = two ? t + 1 : t;
        const uint64_t a = gp(cld(J.x0 + t))[x], a1 = gp(cld(J.x0 + t1))[x];
        const uint64_t c = gp(cld(J.x1 + t))[x], c1 = gp(cld(J.x1 + t1))[x];
        uint64_t w0[G], w1[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const uint64_t* wj = j < J.gc ? cld(wt + j * J.b + t) : nullptr;
            const uint64_t* wk = j < J.gc && two ? cld(wt + j * J.b + t1) : nullptr;
            w0[j] = wj ? gp(wj)[xw] : 0;
            w1[j] = wk ? gp(wk)[xw] : 0;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (j < J.gc) {
                mac128(l0[j], h0[j], a, w0[j]);
                mac128(l1[j], h1[j], c, w0[j]);
                mac128(l0[j], h0[j], a1, w1[j]);
                mac128(l1[j], h1[j], c1, w1[j]);
            }
        }
        // reduce every 8 baby steps (t + 1 and t + 2 complete a multiple of 8 together)
        if (((t + 2) & 7) == 0 && t + 2 < J.b) {
#pragma unroll
            for (int j = 0; j < G; ++j) {
                l0[j] = reduce128(l0[j], h0[j], P);
                h0[j] = 0;
                l1[j] = reduce128(l1[j], h1[j], P);
                h1[j] = 0;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (j < J.gc) {
            gp(cld(J.o0 + j))[x] = reduce128(l0[j], h0[j], P);
            gp(cld(J.o1 + j))[x] = reduce128(l1[j], h1[j], P);
        }
    }
}

// Outputs of extended-basis rotations (ExtRotJob): one coefficient per
// thread, the output written in order and each term gathered through the
// automorphism's index map (a permutation within the limb, so the gathered
// limb stays in L2).
__global__ void __launch_bounds__(EW_BLOCK) k_ext_rot(DevTables T, ExtRotJobs L) {
    const ExtRotJob J = get_job(L, blockIdx.z);
    const int pos = blockIdx.y;
    if (pos >= J.nl + T.K) return;
    const bool q = pos < J.nl;
    const int pid = q ? pos : T.L + 1 + (pos - J.nl);
    const uint64_t p = cld(T.primes + pid).p;
    const ShoupConst pm = cld(T.pmodq + pid);
    const uint32_t j = blockIdx.x * EW_BLOCK + threadIdx.x;
    const size_t off = (size_t)pos * T.N;
    const uint32_t mask = 2u * (uint32_t)T.N - 1u;
    const uint32_t ej = 2u * brev_bits(j, T.logN) + 1u;
    uint64_t s0 = J.base0 ? J.base0[off + j] : 0;
    uint64_t s1 = J.base1 ? J.base1[off + j] : 0;
    for (int t = 0; t < J.nt; ++t) {
        const ExtRotTerm R = t == 0 ? J.t0 : cld(J.more + (t - 1));
        const uint32_t e = (uint32_t)(((uint64_t)ej * R.k) & mask);
        const size_t x = off + brev_bits(e >> 1, T.logN);
        if (R.acc_b) s0 = addmod(s0, R.acc_b[x], p);
        if (R.acc_a) s1 = addmod(s1, R.acc_a[x], p);
        if (R.scale_p) {
            // a Q-basis addend enters as P * add on the Q limbs, 0 on the special limbs
            if (q && R.add0) s0 = addmod(s0, mul_shoup(R.add0[x], pm.w, pm.wp, p), p);
            if (q && R.add1) s1 = addmod(s1, mul_shoup(R.add1[x], pm.w, pm.wp, p), p);
        } else {
            if (R.add0) s0 = addmod(s0, R.add0[x], p);
            if (R.add1) s1 = addmod(s1, R.add1[x], p);
        }
    }
    J.o0[off + j] = s0;
    J.o1[off + j] = s1;
}

// Linear combination sum_t w[t][i] x[t] of both ciphertext parts (the base
// sums of the Chebyshev series): each term read once, one Shoup product and
// one modular add per term and coefficient.
__global__ void __launch_bounds__(EW_BLOCK) k_lincomb(DevTables T, LinCombJob J) {
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const uint64_t p = cld(T.primes + i).p;
    const size_t x = (size_t)i * T.N + 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    uint64_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int t = 0; t < J.n; ++t) {
        const ShoupConst w = cld(J.w + (size_t)t * J.nl + i);
        const u64x2 a = ld2(cld(J.x0 + t) + x), b = ld2(cld(J.x1 + t) + x);
        a0 = addmod(a0, mul_shoup(a.x, w.w, w.wp, p), p);
        a1 = addmod(a1, mul_shoup(a.y, w.w, w.wp, p), p);
        b0 = addmod(b0, mul_shoup(b.x, w.w, w.wp, p), p);
        b1 = addmod(b1, mul_shoup(b.y, w.w, w.wp, p), p);
    }
    st2(J.o0 + x, a0, a1);
    st2(J.o1 + x, b0, b1);
}

// The tensor part of Ciphertext::dotProduct (RNSPoly::binomialDotProduct,
// src/CKKS/RNSPoly.cpp:941-987): over n pairs (a, b),
//   o0 = sum a0 b0,  o1 = sum (a0 b1 + a1 b0),  o2 = sum a1 b1,
// with exact 128-bit sums reduced every 4 pairs (at most 8 products of
// residues below p, below 2^(bits(p) + 63)) and once at the end.  Each
// coefficient's inputs are read before its outputs are written, so an output
// may alias an input.
__global__ void __launch_bounds__(EW_BLOCK) k_tensor_sum(DevTables T, TensorSumJob J) {
    const int i = blockIdx.y;
    if (i >= J.nl) return;
    const PrimeInfo P = cld(T.primes + i);
    const size_t x = (size_t)i * T.N + 2 * (blockIdx.x * EW_BLOCK + threadIdx.x);
    const int n = J.n;
    uint64_t l0[2] = {0, 0}, h0[2] = {0, 0}, l1[2] = {0, 0}, h1[2] = {0, 0}, l2[2] = {0, 0}, h2[2] = {0, 0};
    for (int t = 0; t < n; ++t) {
        const u64x2 a0 = ld2(cld(J.x + t) + x), a1 = ld2(cld(J.x + n + t) + x);
        const u64x2 b0 = ld2(cld(J.x + 2 * n + t) + x), b1 = ld2(cld(J.x + 3 * n + t) + x);
        mac128(l0[0], h0[0], a0.x, b0.x);
        mac128(l0[1], h0[1], a0.y, b0.y);
        mac128(l1[0], h1[0], a0.x, b1.x);
        mac128(l1[0], h1[0], a1.x, b0.x);
        mac128(l1[1], h1[1], a0.y, b1.y);
        mac128(l1[1], h1[1], a1.y, b0.y);
        mac128(l2[0], h2[0], a1.x, b1.x);
        mac128(l2[1], h2[1], a1.y, b1.y);
        if ((t & 3) == 3 && t + 1 < n) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                l0[k] = reduce128(l0[k], h0[k], P);
                l1[k] = reduce128(l1[k], h1[k], P);
                l2[k] = reduce128(l2[k], h2[k], P);
                h0[k] = h1[k] = h2[k] = 0;
            }
        }
    }
    st2(J.o0 + x, reduce128(l0[0], h0[0], P), reduce128(l0[1], h0[1], P));
    st2(J.o1 + x, reduce128(l1[0], h1[0], P), reduce128(l1[1], h1[1], P));
    st2(J.o2 + x, reduce128(l2[0], h2[0], P), reduce128(l2[1], h2[1], P));
}

// ---------------------------------------------------------- launchers ----

hipError_t launch_tensor_sum(const DevTables& T, const TensorSumJob& J, hipStream_t s) {
    if (J.n <= 0 || J.nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tensor_sum, dim3(T.N / (2 * EW_BLOCK), J.nl, 1), dim3(EW_BLOCK), 0, s, T, J);
    return hipGetLastError();
}

hipError_t launch_tensor(const DevTables& T, const TensorJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tensor, dim3(T.N / (2 * EW_BLOCK), max_nl, L.n), dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

// Target groups per coefficient range: enough blocks to fill the GPU
// (FIDES_CONV_TG forces a count).
int conv_rule_tg(int base_blocks, int max_targets) {
    static const int forced = [] {
        const char* e = std::getenv("FIDES_CONV_TG");
        return e ? std::atoi(e) : 0;
    }();
    if (forced > 0) return forced < max_targets ? forced : (max_targets > 0 ? max_targets : 1);
    int tg = (CV_MIN_BLOCKS + base_blocks - 1) / base_blocks;
    tg = tg < 1 ? 1 : (tg > 4 ? 4 : tg);
    return tg > max_targets ? (max_targets > 0 ? max_targets : 1) : tg;
}

// The measured table's target groups for `convs` conversions per launch
// (pyfideslib_amd/tune.py), else the rule above; at most max_targets.
static int conv_groups(int kind, int logN, int convs, int base_blocks, int max_targets) {
    TuneCfg c;
    if (tune_get(kind, logN, convs, c)) {
        const int tg = c.v[0];
        return tg < max_targets ? tg : (max_targets > 0 ? max_targets : 1);
    }
    return conv_rule_tg(base_blocks, max_targets);
}

template <int A>
static void modup_go(const DevTables& T, const ModUpJobs& L, dim3 g, int tg, hipStream_t s) {
    if (T.modup_one_reduce && A <= 16)
        hipLaunchKernelGGL((k_modup_conv<A, (A <= 16)>), g, dim3(CV_B), 0, s, T, L, tg);
    else
        hipLaunchKernelGGL((k_modup_conv<A, false>), g, dim3(CV_B), 0, s, T, L, tg);
}


static hipError_t launch_modup_mfma(const DevTables& T, const ModUpJobs& L, int max_nd, hipStream_t s) {
    const int TT = (T.L + T.K + 15) / 16;  // target tiles: at most L + K targets per digit
    const int tiles = T.N / 16;
    // about 4096 waves in all, each over CT coefficient tiles
    const long want = (long)tiles * max_nd * TT * L.n / 4096;
    const int CT = want < 1 ? 1 : (want > 64 ? 64 : (int)want);
    const dim3 g((tiles + MF_WAVES * CT - 1) / (MF_WAVES * CT), max_nd * TT, L.n);
    switch (T.modup_mfma_ms) {
        case 1: hipLaunchKernelGGL((k_modup_mfma<1>), g, dim3(64 * MF_WAVES), 0, s, T, L, TT, CT); break;
        case 2: hipLaunchKernelGGL((k_modup_mfma<2>), g, dim3(64 * MF_WAVES), 0, s, T, L, TT, CT); break;
        case 3: hipLaunchKernelGGL((k_modup_mfma<3>), g, dim3(64 * MF_WAVES), 0, s, T, L, TT, CT); break;
        case 4: hipLaunchKernelGGL((k_modup_mfma<4>), g, dim3(64 * MF_WAVES), 0, s, T, L, TT, CT); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_modup_conv(const DevTables& T, const ModUpJobs& L, int max_nd, hipStream_t s) {
    if (L.n <= 0 || max_nd <= 0) return hipSuccess;
    if (T.L + 1 + T.K > CV_MAXT || T.N % CV_B != 0) return hipErrorInvalidValue;
    // the matrix-core conversion where the tuning table (or a forced
    // configuration) selects it: v1 = 1
    TuneCfg tc;
    if (T.modup_mfma_ms > 0 && T.N % (16 * MF_WAVES) == 0 && tune_get(TUNE_MODUP, T.logN, max_nd * L.n, tc) &&
        tc.v[1] == 1)
        return launch_modup_mfma(T, L, max_nd, s);
    const int tg = conv_groups(TUNE_MODUP, T.logN, max_nd * L.n, (T.N / CV_B) * max_nd * L.n, T.L + 1 + T.K - T.alpha);
    dim3 g(T.N / CV_B, max_nd * tg, L.n);
    switch (conv_bound(T.alpha)) {  // the template bound the matrix rows are padded to
        case 1: modup_go<1>(T, L, g, tg, s); break;
        case 2: modup_go<2>(T, L, g, tg, s); break;
        case 3: modup_go<3>(T, L, g, tg, s); break;
        case 4: modup_go<4>(T, L, g, tg, s); break;
        case 5: modup_go<5>(T, L, g, tg, s); break;
        case 6: modup_go<6>(T, L, g, tg, s); break;
        case 7: modup_go<7>(T, L, g, tg, s); break;
        case 8: modup_go<8>(T, L, g, tg, s); break;
        case 9: modup_go<9>(T, L, g, tg, s); break;
        case 10: modup_go<10>(T, L, g, tg, s); break;
        case 12: modup_go<12>(T, L, g, tg, s); break;
        case 16: modup_go<16>(T, L, g, tg, s); break;
        case 32: modup_go<32>(T, L, g, tg, s); break;
        case 64: modup_go<64>(T, L, g, tg, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ksk_inner(const DevTables& T, const KskJobs& L, int max_npos, hipStream_t s, bool shared) {
    if (L.n <= 0 || max_npos <= 0) return hipSuccess;
    const unsigned gx = T.N / (2 * EW_BLOCK);
    if (T.dnum <= 2) {
        if (shared)
            hipLaunchKernelGGL((k_ksk_inner_d<2, true>), dim3(gx, max_npos, 1), dim3(EW_BLOCK), 0, s, T, L);
        else
            hipLaunchKernelGGL((k_ksk_inner_d<2, false>), dim3(gx, max_npos, L.n), dim3(EW_BLOCK), 0, s, T, L);
    } else if (T.dnum == 3) {
        if (shared)
            hipLaunchKernelGGL((k_ksk_inner_d<3, true>), dim3(gx, max_npos, 1), dim3(EW_BLOCK), 0, s, T, L);
        else
            hipLaunchKernelGGL((k_ksk_inner_d<3, false>), dim3(gx, max_npos, L.n), dim3(EW_BLOCK), 0, s, T, L);
    } else if (T.dnum == 4) {
        if (shared)
            hipLaunchKernelGGL((k_ksk_inner_d<4, true>), dim3(gx, max_npos, 1), dim3(EW_BLOCK), 0, s, T, L);
        else
            hipLaunchKernelGGL((k_ksk_inner_d<4, false>), dim3(gx, max_npos, L.n), dim3(EW_BLOCK), 0, s, T, L);
    } else {
        hipLaunchKernelGGL(k_ksk_inner, dim3(gx, max_npos, L.n), dim3(EW_BLOCK), 0, s, T, L);
    }
    return hipGetLastError();
}

hipError_t launch_moddown_conv(const DevTables& T, const ModDownJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0) return hipSuccess;
    if (max_nl > CV_MAXT) return hipErrorInvalidValue;
    const bool rs = (L.dev == nullptr) ? (L.inl[0].top_acc != nullptr) : false;
    return launch_moddown_conv_mode(T, L, rs, s);
}

template <int KK, bool ONE>
static void moddown_go(const DevTables& T, const ModDownJobs& L, bool rescale, dim3 g, int tg, hipStream_t s) {
    if (rescale)
        hipLaunchKernelGGL((k_moddown_conv<KK, ONE, true>), g, dim3(CV_B), 0, s, T, L, tg);
    else
        hipLaunchKernelGGL((k_moddown_conv<KK, ONE, false>), g, dim3(CV_B), 0, s, T, L, tg);
}

hipError_t launch_moddown_conv_mode(const DevTables& T, const ModDownJobs& L, bool rescale, hipStream_t s) {
    if (L.n <= 0) return hipSuccess;
    if (T.L + 1 > CV_MAXT || T.N % CV_B != 0) return hipErrorInvalidValue;
    const int tg = conv_groups(TUNE_MODDOWN, T.logN, L.n, (T.N / CV_B) * L.n, T.L + 1);
    dim3 g(T.N / CV_B, tg, L.n);
    const bool one = T.moddown_one_reduce != 0;
    switch (conv_bound(T.K)) {  // the template bound the matrix rows are padded to
        case 1: one ? moddown_go<1, true>(T, L, rescale, g, tg, s) : moddown_go<1, false>(T, L, rescale, g, tg, s); break;
        case 2: one ? moddown_go<2, true>(T, L, rescale, g, tg, s) : moddown_go<2, false>(T, L, rescale, g, tg, s); break;
        case 3: one ? moddown_go<3, true>(T, L, rescale, g, tg, s) : moddown_go<3, false>(T, L, rescale, g, tg, s); break;
        case 4: one ? moddown_go<4, true>(T, L, rescale, g, tg, s) : moddown_go<4, false>(T, L, rescale, g, tg, s); break;
        case 5: one ? moddown_go<5, true>(T, L, rescale, g, tg, s) : moddown_go<5, false>(T, L, rescale, g, tg, s); break;
        case 6: one ? moddown_go<6, true>(T, L, rescale, g, tg, s) : moddown_go<6, false>(T, L, rescale, g, tg, s); break;
        case 7: one ? moddown_go<7, true>(T, L, rescale, g, tg, s) : moddown_go<7, false>(T, L, rescale, g, tg, s); break;
        case 8: one ? moddown_go<8, true>(T, L, rescale, g, tg, s) : moddown_go<8, false>(T, L, rescale, g, tg, s); break;
        case 9: one ? moddown_go<9, true>(T, L, rescale, g, tg, s) : moddown_go<9, false>(T, L, rescale, g, tg, s); break;
        case 10: one ? moddown_go<10, true>(T, L, rescale, g, tg, s) : moddown_go<10, false>(T, L, rescale, g, tg, s); break;
        case 12: one ? moddown_go<12, true>(T, L, rescale, g, tg, s) : moddown_go<12, false>(T, L, rescale, g, tg, s); break;
        case 16: one ? moddown_go<16, true>(T, L, rescale, g, tg, s) : moddown_go<16, false>(T, L, rescale, g, tg, s); break;
        case 32: moddown_go<32, false>(T, L, rescale, g, tg, s); break;
        default: return hipErrorInvalidValue;
    }
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

hipError_t launch_scalar(const DevTables& T, const EwJobs& L, int max_nl, int op, const LimbConsts& k, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    if (max_nl > MAX_Q_LIMBS) return hipErrorInvalidValue;
    dim3 g(T.N / (2 * EW_BLOCK), max_nl, L.n), b(EW_BLOCK);
    if (op == EW_MUL_SCALAR)
        hipLaunchKernelGGL(k_scalar<EW_MUL_SCALAR>, g, b, 0, s, T, L, k);
    else if (op == EW_ADD_SCALAR)
        hipLaunchKernelGGL(k_scalar<EW_ADD_SCALAR>, g, b, 0, s, T, L, k);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_lincomb(const DevTables& T, const LinCombJob& J, hipStream_t s) {
    if (J.n <= 0 || J.nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lincomb, dim3(T.N / (2 * EW_BLOCK), J.nl, 1), dim3(EW_BLOCK), 0, s, T, J);
    return hipGetLastError();
}

hipError_t launch_dot_bsgs(const DevTables& T, const DotJobs& L, int max_nl, int max_gc, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0 || max_gc <= 0) return hipSuccess;
    if (max_gc > DOT_GMAX) return hipErrorInvalidValue;
    const dim3 grid(T.N / EW_BLOCK, max_nl, L.n);
    if (max_gc <= 1)
        hipLaunchKernelGGL(k_dot_bsgs<1>, grid, dim3(EW_BLOCK), 0, s, T, L);
    else if (max_gc <= 2)
        hipLaunchKernelGGL(k_dot_bsgs<2>, grid, dim3(EW_BLOCK), 0, s, T, L);
    else if (max_gc <= 4)
        hipLaunchKernelGGL(k_dot_bsgs<4>, grid, dim3(EW_BLOCK), 0, s, T, L);
    else
        hipLaunchKernelGGL(k_dot_bsgs<8>, grid, dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

hipError_t launch_ext_rot(const DevTables& T, const ExtRotJobs& L, int max_nl, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ext_rot, dim3(T.N / EW_BLOCK, max_nl + T.K, L.n), dim3(EW_BLOCK), 0, s, T, L);
    return hipGetLastError();
}

hipError_t launch_monomial(const DevTables& T, const EwJobs& L, int max_nl, uint32_t m, hipStream_t s) {
    if (L.n <= 0 || max_nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_monomial, dim3(T.N / EW_BLOCK, max_nl, L.n), dim3(EW_BLOCK), 0, s, T, L, m);
    return hipGetLastError();
}

}  // namespace fides
