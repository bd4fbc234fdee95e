// Counter-based random numbers for the collision kernels: Philox4x32-10 keyed by a 64-bit seed,
// counted by (block, stream id, step, tag), so a stream's draws do not depend on the order in
// which streams are processed.
#pragma once
#include "sf_internal.h"

namespace sf {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of two 32x32->64 multiplies
// with the key bumped by the Weyl constants between rounds.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// A particle's stream: draw k uses half k&1 of block k>>1 of the counter (block, id, step, tag),
// 53 bits in [0, 1).
struct ParticleRng {   // stream id: a particle id (MCC) or a cell index (DSMC)
    uint32_t k0, k1, id, step, tag;
    __device__ __forceinline__ double draw(uint32_t k) const {
        uint4 o = philox4x32_10(make_uint4(k >> 1, id, step, tag), k0, k1);
        uint64_t bits = (k & 1) ? (((uint64_t)o.w << 32) | o.z) : (((uint64_t)o.y << 32) | o.x);
        return (double)(bits >> 11) * 0x1.0p-53;
    }
    // twelve consecutive draws summed in order (Utils.SampleMaxw1D), six Philox blocks
    __device__ __forceinline__ double sum12(uint32_t k) const {
        double s = 0;
        for (uint32_t b = 0; b < 6; b++) {
            uint4 o = philox4x32_10(make_uint4((k >> 1) + b, id, step, tag), k0, k1);
            s += (double)(((((uint64_t)o.y << 32) | o.x)) >> 11) * 0x1.0p-53;
            s += (double)(((((uint64_t)o.w << 32) | o.z)) >> 11) * 0x1.0p-53;
        }
        return s;
    }
};

}  // namespace sf
