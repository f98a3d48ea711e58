// Counter-based normal deviates for the Langevin thermostat.
//
// Philox4x32-10 as published by Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64
// multiplications with the Weyl key schedule.  A deviate depends only on
// (seed, counter), so every atom draws its own numbers in any launch order
// and a step re-run after a rebuild draws the same ones.  Uniforms take 53
// bits from two words and lie in (0, 1]; Box-Muller turns two of them into
// two standard normal deviates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdc {

struct U32x4 {
    uint32_t v[4];
};

__host__ __device__ inline void philox_round(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

__host__ __device__ inline U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
    uint32_t c[4] = {c0, c1, c2, c3};
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        philox_round(c, k0, k1);
    }
    return U32x4{{c[0], c[1], c[2], c[3]}};
}

// 53-bit uniform in (0, 1] from two words
__host__ __device__ inline double u01(uint32_t hi, uint32_t lo) {
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    return (double)((w >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// Two standard normal deviates from one Philox block
__host__ __device__ inline void normal2(const U32x4& x, double& a, double& b) {
    const double u1 = u01(x.v[0], x.v[1]), u2 = u01(x.v[2], x.v[3]);
    const double r = sqrt(-2.0 * log(u1));
    const double t = 6.283185307179586 * u2;
    a = r * cos(t);
    b = r * sin(t);
}

// Three deviates for atom `atom` at thermostat application `call` of
// thermostat `k` under `seed`: counters (atom, call lo, call hi, 2k + q).
__host__ __device__ inline void normal3(uint64_t seed, uint64_t call, uint32_t atom, int k, double out[3]) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t clo = (uint32_t)call, chi = (uint32_t)(call >> 32);
    double d;
    normal2(philox4x32_10(atom, clo, chi, 2u * k, k0, k1), out[0], out[1]);
    normal2(philox4x32_10(atom, clo, chi, 2u * k + 1u, k0, k1), out[2], d);
}

}  // namespace mdc
