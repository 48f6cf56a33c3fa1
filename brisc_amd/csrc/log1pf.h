// log1p in single precision, bit-identical to the host C library the
// reference links (glibc's log1pf): the published fdlibm algorithm for
// log1pf (Sun Microsystems, 1993; single-precision version by Ian Lance
// Taylor, Cygnus Support), which glibc uses, restated.  Argument reduction
// 1 + x = 2^k (1 + f) with sqrt(2)/2 < 1 + f < sqrt(2) and a correction term
// c for the rounding of 1 + x, then log(1 + f) = f - hfsq + s (hfsq + R(z))
// with s = f / (2 + f), z = s^2 and the minimax polynomial R of fdlibm
// (coefficients Lp1..Lp7), summed as fdlibm orders it.
//
// Every operation must be a single correctly rounded IEEE step: files that
// include this header are compiled with -ffp-contract=off (no fused
// multiply-adds) and correctly rounded division.  Checked against the host's
// log1pf on every positive float and every float in [-1, 0].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bpca {

__host__ __device__ inline float log1pf_ref(float x) {
    const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f;
    const float Lp1 = 6.6666668653e-01f, Lp2 = 4.0000000596e-01f, Lp3 = 2.8571429849e-01f,
                Lp4 = 2.2222198546e-01f, Lp5 = 1.8183572590e-01f, Lp6 = 1.5313838422e-01f,
                Lp7 = 1.4798198640e-01f;
    float f = 0.0f, c = 0.0f;
    int32_t hu = 0;
    const int32_t hx = (int32_t)__builtin_bit_cast(uint32_t, x);
    const int32_t ax = hx & 0x7fffffff;
    int32_t k = 1;
    if (hx < 0x3ed413d7) {                 // x < 0.41422
        if (ax >= 0x3f800000) {            // x <= -1
            if (x == -1.0f) return -__builtin_inff();
            return __builtin_nanf("");
        }
        if (ax < 0x31000000) {             // |x| < 2^-29
            if (ax < 0x24800000) return x;  // |x| < 2^-54
            return x - x * x * 0.5f;
        }
        if (hx > 0 || hx <= (int32_t)0xbe95f61f) {  // -0.2929 < x < 0.41422
            k = 0;
            f = x;
            hu = 1;
        }
    } else if (hx >= 0x7f800000) {
        return x + x;
    }
    if (k != 0) {
        float u;
        if (hx < 0x5a000000) {
            u = 1.0f + x;
            hu = (int32_t)__builtin_bit_cast(uint32_t, u);
            k = (hu >> 23) - 127;
            c = (k > 0) ? 1.0f - (u - x) : x - (u - 1.0f);  // rounding of 1 + x
            c /= u;
        } else {
            u = x;
            hu = (int32_t)__builtin_bit_cast(uint32_t, u);
            k = (hu >> 23) - 127;
            c = 0.0f;
        }
        hu &= 0x007fffff;
        if (hu < 0x3504f7) {
            u = __builtin_bit_cast(float, (uint32_t)hu | 0x3f800000u);   // u in [1, sqrt 2)
        } else {
            k += 1;
            u = __builtin_bit_cast(float, (uint32_t)hu | 0x3f000000u);   // u / 2
            hu = (0x00800000 - hu) >> 2;
        }
        f = u - 1.0f;
    }
    const float hfsq = 0.5f * f * f;
    const float fk = (float)k;
    if (hu == 0) {  // |f| < 2^-20
        if (f == 0.0f) {
            if (k == 0) return 0.0f;
            c += fk * ln2_lo;
            return fk * ln2_hi + c;
        }
        const float R = hfsq * (1.0f - 0.66666666666666666f * f);
        if (k == 0) return f - R;
        return fk * ln2_hi - ((R - (fk * ln2_lo + c)) - f);
    }
    const float s = f / (2.0f + f);
    const float z = s * s;
    const float R = z * (Lp1 + z * (Lp2 + z * (Lp3 + z * (Lp4 + z * (Lp5 + z * (Lp6 + z * Lp7))))));
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return fk * ln2_hi - ((hfsq - (s * (hfsq + R) + (fk * ln2_lo + c))) - f);
}

}  // namespace bpca
