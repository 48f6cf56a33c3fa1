// Host-side random draws that must reproduce the reference's streams bit for
// bit: PCG-XSH-RR 64/32 (O'Neill 2014) seeded as brisc's srand does
// (brisc/cyutils.pyx:1153-1158, 1228-1232), Lemire's bounded draw
// (cyutils.pyx:1161-1171) and the FP32 Box-Muller normal draw
// (cyutils.pyx:1174-1191).  Used for the IRLBA start vector
// (brisc/pca.pyx:801-806), for kmeans()'s random initial cells
// (brisc/kmeans.pyx:13-39) and for the k-means|| draws (kmeans.pyx:41-360),
// whose per-cell uniforms are also drawn on the device.
#pragma once

#include <climits>
#include <cmath>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace bpca {

struct Pcg32 {
    uint64_t state;
    __host__ __device__ explicit Pcg32(uint64_t seed) : state(seed + 1442695040888963407ULL) { next(); }
    __host__ __device__ uint32_t next() {
        const uint64_t x = state;
        state = x * 6364136223846793005ULL + 1442695040888963407ULL;
        const uint32_t s = (uint32_t)((x ^ (x >> 18)) >> 27);
        const uint32_t rot = (uint32_t)(x >> 59);
        return (s >> rot) | (s << ((-rot) & 31u));
    }
    // Uniform integer in [0, bound).
    __host__ __device__ uint32_t bounded(uint32_t bound) {
        uint64_t m = (uint64_t)next() * bound;
        uint32_t l = (uint32_t)m;
        if (l < bound) {
            const uint32_t threshold = (0u - bound) % bound;
            while (l < threshold) {
                m = (uint64_t)next() * bound;
                l = (uint32_t)m;
            }
        }
        return (uint32_t)(m >> 32);
    }
    // FP32 uniform in [0, 1]: the draw divided by UINT_MAX in single
    // precision.
    __host__ __device__ float uniform() { return (float)next() / (float)UINT_MAX; }
    // FP32 standard normal (Box-Muller): r = sqrt(-2 log u1) and
    // theta = 2 pi u2 (formed in double, rounded to float), result r cos theta,
    // every function the single-precision overload.
    float normal() {
        float u1;
        do { u1 = uniform(); } while (u1 == 0.0f);
        const float u2 = uniform();
        const float r = std::sqrt(-2.0f * std::log(u1));
        const float theta = (float)(2.0 * M_PI * (double)u2);
        return r * std::cos(theta);
    }
};

}  // namespace bpca
