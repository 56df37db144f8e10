// exact_sum.h -- order-independent (bit-reproducible) sums of FP32 values
// through 64-bit integer limbs; host and device.
//
// A value c (|c| < 2^31) is the integer v = trunc(|c| * 2^64) < 2^95 with c's
// sign, split into 32-bit limbs v = l2 * 2^64 + l1 * 2^32 + l0; accumulator k
// of parameter i holds the sum of the limbs k of its values (three int64 per
// parameter, limb-major: [k * n + i]).  Each limb is below 2^32, so 2^31 values
// per parameter cannot overflow one; integer addition is associative, so the
// sum does not depend on the order the values arrive in.  The result is
// l2 + l1 * 2^-32 + l0 * 2^-64 after the carries are moved up (limb_value).
// Bits below 2^-64 (values under about 2^-41) are dropped, the same way every
// time.
#pragma once
#include <cstdint>
#include <cstring>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace eos {

// false: c is not finite or |c| >= 2^31 (nothing to add).
__host__ __device__ __forceinline__ bool to_limbs(float c, long long& l0, long long& l1, long long& l2) {
    uint32_t b;
    std::memcpy(&b, &c, sizeof b);
    const int e = (int)((b >> 23) & 0xFFu);
    l0 = l1 = l2 = 0;
    if (e == 0) return true;                 // zero or below 2^-126: nothing at 2^-64 resolution
    if (e >= 127 + 31) return false;         // |c| >= 2^31, inf or NaN
    const unsigned long long m = (unsigned long long)((b & 0x7FFFFFu) | 0x800000u);   // |c| = m * 2^(e - 150)
    const int k = e - 150 + 64;                                                       // v = m * 2^k, k < 70
    unsigned long long lo = 0, hi = 0;
    if (k < 0) {
        lo = k <= -24 ? 0ull : m >> (-k);
    } else if (k < 64) {
        lo = m << k;
        hi = k > 40 ? m >> (64 - k) : 0ull;  // m < 2^24: nothing spills below k = 41
    } else {
        hi = m << (k - 64);
    }
    l0 = (long long)(lo & 0xFFFFFFFFull);
    l1 = (long long)(lo >> 32);
    l2 = (long long)hi;
    if (b >> 31) {
        l0 = -l0;
        l1 = -l1;
        l2 = -l2;
    }
    return true;
}

// Parameter i of the limbs as a double: carries moved up so the two lower
// limbs are in [0, 2^32) (exact in FP64), then l2 + l1 * 2^-32 + l0 * 2^-64.
__host__ __device__ inline double limb_value(const unsigned long long* limbs, int n, int i) {
    long long l0 = (long long)limbs[i], l1 = (long long)limbs[n + i], l2 = (long long)limbs[2 * (size_t)n + i];
    l1 += l0 >> 32;   // arithmetic shifts: floor division by 2^32
    l0 &= 0xFFFFFFFFll;
    l2 += l1 >> 32;
    l1 &= 0xFFFFFFFFll;
    return (double)l2 + ((double)l1 * 0x1p-32 + (double)l0 * 0x1p-64);
}

}  // namespace eos
