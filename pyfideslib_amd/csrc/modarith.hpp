// 64-bit modular arithmetic for CDNA4 (gfx950), built from 32-bit multiplies.
//
// CDNA4 has no 64x64-bit integer multiply; the compiler lowers the pieces
// below to v_mul_hi_u32 / v_mul_lo_u32 / v_mad_u64_u32.  Every routine returns
// the exact residue (or an exact residue plus a stated multiple of p when
// named *_lazy); no floating point enters the ring.
//
// Reference counterparts: FIDESlib src/ModMult.cuh:46-74 (Barrett),
// :96-102 (Shoup), :288-313 (128-bit reduce), src/AddSub.cuh:37-50.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fides {

// Per-prime constants used by the kernels.
struct PrimeInfo {
    uint64_t p;      // modulus, < 2^60
    uint64_t mu;     // floor(2^(n+63) / p), n = bit length of p
    uint32_t nbits;  // n
    uint32_t pad;
    uint64_t p2;     // 2p
};

// Pair (w, floor(w * 2^64 / p)) for Shoup multiplication by the constant w.
struct ShoupConst {
    uint64_t w, wp;
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// Full 128-bit product with four 32x32->64 multiply-adds.
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t p01 = (uint64_t)a0 * b1 + (p00 >> 32);
    const uint64_t p10 = (uint64_t)a1 * b0 + (uint32_t)p01;
    hi = (uint64_t)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (uint32_t)p00;
}

// (lo, hi) += a * b   (caller guarantees no overflow of the 128-bit sum)
__device__ __forceinline__ void mac128(uint64_t& lo, uint64_t& hi, uint64_t a, uint64_t b) {
    uint64_t pl, ph;
    mul128(a, b, pl, ph);
    lo += pl;
    hi += ph + (lo < pl);
}

// Barrett reduction of X = hi*2^64 + lo, valid for X < 2^(n+63):
// qhat = floor((X >> (n-1)) * mu / 2^64) is within 2 of floor(X/p).
__device__ __forceinline__ uint64_t reduce128(uint64_t lo, uint64_t hi, const PrimeInfo& P) {
    const uint32_t s = P.nbits - 1;
    const uint64_t xs = (lo >> s) | (hi << (64 - s));
    const uint64_t q = mulhi64(xs, P.mu);
    uint64_t r = lo - q * P.p;
    r = r >= P.p2 ? r - P.p2 : r;
    return r >= P.p ? r - P.p : r;
}

// Reduce a single 64-bit value (< 2^64) modulo p.
__device__ __forceinline__ uint64_t reduce64(uint64_t x, const PrimeInfo& P) { return reduce128(x, 0, P); }

__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b, const PrimeInfo& P) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    return reduce128(lo, hi, P);
}

// Shoup: x * w mod p with x < 2^64, w < p; lazy result in [0, 2p).
__device__ __forceinline__ uint64_t mul_shoup_lazy(uint64_t x, uint64_t w, uint64_t wp, uint64_t p) {
    const uint64_t q = mulhi64(x, wp);
    return x * w - q * p;
}
__device__ __forceinline__ uint64_t mul_shoup(uint64_t x, uint64_t w, uint64_t wp, uint64_t p) {
    const uint64_t r = mul_shoup_lazy(x, w, wp, p);
    return r >= p ? r - p : r;
}

__device__ __forceinline__ uint64_t addmod(uint64_t a, uint64_t b, uint64_t p) {
    const uint64_t s = a + b;
    return s >= p ? s - p : s;
}
__device__ __forceinline__ uint64_t submod(uint64_t a, uint64_t b, uint64_t p) {
    return a >= b ? a - b : a + p - b;
}

__device__ __forceinline__ uint32_t brev_bits(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

}  // namespace fides
