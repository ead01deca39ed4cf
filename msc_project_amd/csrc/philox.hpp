// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11),
// identical on host and device, so that seeded inputs do not depend on
// thread counts or launch shapes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lgf {

__host__ __device__ __forceinline__ void philox4x32(const uint32_t in[4], uint64_t seed, uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// two uniforms in (0, 1] from one block
__host__ __device__ __forceinline__ void uniform2(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                                  double* u) {
    const uint32_t in[4] = {a, b, c, d};
    uint32_t o[4];
    philox4x32(in, seed, o);
    const uint64_t x = ((uint64_t)o[0] << 32) | o[1];
    const uint64_t y = ((uint64_t)o[2] << 32) | o[3];
    u[0] = ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    u[1] = ((double)(y >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

}  // namespace lgf
