// Sequential 24-bit uniforms of one Philox stream: draw k is word k % 4 of
// block (k / 4, stream, id_lo, id_hi) under the key (seed_lo, seed_hi), as
// the reference's generator is driven through random::Generator::uniform.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "propagate_core.hpp"

namespace mlgpu {

struct StreamRng {
    uint32_t k0, k1, id_lo, id_hi, stream, next;
    uint32_t b[4];
    __device__ void init(uint64_t seed, uint64_t id, uint32_t s) {
        k0 = static_cast<uint32_t>(seed);
        k1 = static_cast<uint32_t>(seed >> 32);
        id_lo = static_cast<uint32_t>(id);
        id_hi = static_cast<uint32_t>(id >> 32);
        stream = s;
        next = 0;
    }
    __device__ double uniform() {
        if ((next & 3u) == 0u) philox4x32_10(next >> 2, stream, id_lo, id_hi, k0, k1, b[0], b[1], b[2], b[3]);
        const uint32_t w = b[next & 3u];
        next++;
        return static_cast<double>(w >> 8) * (1.0 / 16777216.0);
    }
    // Generator::expovariate (merlict/random/Generator.cpp:36-39)
    __device__ double expovariate(double rate) { return -log(uniform()) / rate; }
    // normal(mean, std) as the reference's Philox generator draws it:
    // Box-Muller on 1 - u1 and u2
    __device__ double normal(double mean, double std_dev) {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return mean + std_dev * sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
    }
};

}  // namespace mlgpu
