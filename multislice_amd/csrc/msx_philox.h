// msx_philox.h -- counter-based random numbers for frozen-phonon displacements,
// identical on the host and the GPU.
//
// Philox-4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC11, 2011): ten rounds of
// two 32x32->64-bit multiplies with the published multipliers and Weyl key
// increments.  The stream of a frozen-phonon displacement is addressed by
// counter = (atom index lo, atom index hi, configuration, 0) and
// key = (seed lo, seed hi), so every (seed, configuration, atom) triple has its
// own draw and the result does not depend on thread or batch order.
//
// The Gaussian transform is Box-Muller in FP64 on uniforms built from the top
// 53 bits of two 32-bit words, (u + 0.5) * 2^-53 (never 0 or 1).  The isotropic
// displacement follows the reference's frozen-phonon models (Kirkland
// ReadXYZcoord wobble, slicelib.cpp:2197-2320; Musli Vobble,
// MusliRandom.cs:43-55): independent N(0, u^2) offsets per axis.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define MSX_HD __host__ __device__ __forceinline__
#else
#define MSX_HD inline
#endif

namespace msx {

struct Philox4 {
    uint32_t v[4];
};

MSX_HD void philox_round(uint32_t (&c)[4], const uint32_t (&k)[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    const uint32_t n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
}

MSX_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    uint32_t c[4] = {c0, c1, c2, c3};
    uint32_t k[2] = {k0, k1};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        if (r < 9) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
    }
    Philox4 out;
    for (int i = 0; i < 4; ++i) out.v[i] = c[i];
    return out;
}

// uniform in (0, 1) from two words, 53 bits
MSX_HD double uniform53(uint32_t hi, uint32_t lo) {
    const uint64_t u = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)u + 0.5) * (1.0 / 9007199254740992.0);
}

}  // namespace msx
