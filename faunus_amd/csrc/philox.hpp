// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11) and the
// translation-proposal scheme of the device-resident Metropolis loop.  Host and
// device run the same integer arithmetic, so a proposal stream is reproducible on
// the CPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fmc {

struct Philox4 {
    uint32_t v[4];
};

__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                  uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        c1 = static_cast<uint32_t>(p1);
        c3 = static_cast<uint32_t>(p0);
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{{c0, c1, c2, c3}};
}

// One translation proposal (AtomicTranslateRotate::translateParticle semantics,
// move.cpp:225-240: uniform random particle, isotropic unit vector times dp*U[0,1)).
struct Proposal {
    int index;       // particle within the moved set [0, n)
    double dir[3];   // unit vector
    double length_u; // U[0,1) scaling dp
    double accept_u; // U[0,1) of the acceptance test
};

__host__ __device__ inline double u01_from32(uint32_t x) { return (static_cast<double>(x) + 0.5) * 2.3283064365386963e-10; }

__host__ __device__ inline Proposal make_proposal(uint64_t seed, uint32_t replica, uint64_t move, int n) {
    const Philox4 a = philox4x32_10(static_cast<uint32_t>(move), static_cast<uint32_t>(move >> 32), replica, 0u,
                                    static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    const Philox4 b = philox4x32_10(static_cast<uint32_t>(move), static_cast<uint32_t>(move >> 32), replica, 1u,
                                    static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    Proposal p;
    p.index = static_cast<int>((static_cast<uint64_t>(a.v[0]) * static_cast<uint64_t>(n)) >> 32);
    const double z = 2.0 * u01_from32(a.v[1]) - 1.0;
    const double phi = 6.283185307179586 * u01_from32(a.v[2]);
    const double s = sqrt(1.0 - z * z);
    p.dir[0] = s * cos(phi);
    p.dir[1] = s * sin(phi);
    p.dir[2] = z;
    p.length_u = static_cast<double>(a.v[3]) * 2.3283064365386963e-10;
    p.accept_u = (static_cast<double>(b.v[0] >> 5) * 67108864.0 + static_cast<double>(b.v[1] >> 6)) *
                 1.1102230246251565e-16;
    return p;
}

} // namespace fmc
