// Device helpers shared by the device-resident Metropolis kernels: the FP32 pair
// form, minimum-image and boundary arithmetic, cell coordinates and the Metropolis
// decision.
#pragma once
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine.hpp"

namespace fmc {

constexpr int kMaxGk = 1025; // |n|^2 table (ncut <= 32)

struct DevPairF {
    float lj_s2, lj_e4, lj_cut2, wca_s2, wca_e4, hs_s2, coul, pad;
};

__device__ __forceinline__ float dist2f(float L, float half, float ax, float ay, float az, float bx, float by,
                                        float bz) {
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    dx = dx > half ? dx - L : (dx < -half ? dx + L : dx);
    dy = dy > half ? dy - L : (dy < -half ? dy + L : dy);
    dz = dz > half ? dz - L : (dz < -half ? dz + L : dz);
    return dx * dx + dy * dy + dz * dz;
}

// FP32 pair energy: the algebra of pair_energy in single precision, written without
// per-lane branches (every term is evaluated and selected) and with the hardware
// reciprocal and reciprocal square root (1 ulp), which the FP32 form's tolerance covers.
__device__ __forceinline__ float pair_energy_f(const DevPairF& p, const DevCoulomb& c, const float* knots,
                                               const float* coeff, float qa, float qb, float r2) {
    const float ir2 = __builtin_amdgcn_rcpf(r2);
    float x = p.lj_s2 * ir2;
    x = x * x * x;
    float u = (r2 < p.lj_cut2) ? p.lj_e4 * (x * x - x) : 0.0f;
    float w = p.wca_s2 * ir2;
    w = w * w * w;
    u += (r2 > p.wca_s2 * 1.2599210498948732f) ? 0.0f : p.wca_e4 * (w * w - w + 0.25f);
    u += (r2 < p.hs_s2) ? __builtin_huge_valf() : 0.0f;
    if (c.scheme != kCoulNone) { // the same scheme for every lane
        const float lB = static_cast<float>(c.lB);
        const float ir = __builtin_amdgcn_rsqf(r2);
        float e;
        if (c.scheme == kCoulPM || c.scheme == kCoulPlain) {
            e = lB * (qa * qb * ir);
        } else {
            const float r = r2 * ir;
            const float qq = r * static_cast<float>(c.rc_inv);
            float s;
            if (c.splined) s = spline_eval<float>(knots, coeff, c.nknots, qq, c.bucket, c.spline_steps);
            else if (c.scheme == kCoulEwald) s = erfcf(static_cast<float>(c.alpha * c.rc) * qq);
            else {
                const float q5 = qq * qq * qq * qq * qq;
                s = 1.0f - 1.75f * qq + 5.25f * q5 - 7.0f * q5 * qq + 2.5f * q5 * qq * qq;
            }
            e = (r < static_cast<float>(c.rc)) ? lB * (qa * qb * ir * s) : 0.0f;
        }
        u += (p.coul != 0.0f) ? e : 0.0f;
    }
    return u;
}

__device__ __forceinline__ int cell_of_coord(double x, double half, double inv_cell, int nc) {
    int c = static_cast<int>(floor((x + half) * inv_cell));
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

__device__ __forceinline__ double wrap(double x, double L, double half, double inv) {
    // Cuboid::boundary (geometry.cpp:91-104): x -= L * anint(x / L) when |x| > L/2
    if (fabs(x) > half) {
        const double s = x * inv;
        x -= L * static_cast<double>(static_cast<int>(s > 0.0 ? s + 0.5 : s - 0.5));
    }
    return x;
}

// getEnergyChange (montecarlo.cpp:193-209) on the split totals
__device__ __forceinline__ double energy_change(double pair_new, double pair_old, double drec) {
    const double e_new = pair_new + drec; // the reciprocal term enters as a difference
    if (isnan(pair_old) && !isnan(e_new)) return -__builtin_huge_val();
    if (isnan(e_new)) return __builtin_huge_val();
    if (e_new > 0.0 && isinf(e_new)) return __builtin_huge_val();
    const double du = (pair_new - pair_old) + drec;
    if (isnan(du)) return 0.0;
    return du;
}

// metropolisCriterion (montecarlo.cpp:17-34)
__device__ __forceinline__ bool metropolis(double du, double u) {
    if (isnan(du)) return false;
    if (isinf(du) && du < 0.0) return true;
    if (-du > 709.782712893384) return true;
    return u <= exp(-du);
}


// Exclusive prefix sum of one int per thread over a block of kThreads threads
// (wavefront shuffles, then the wave totals in LDS); *total receives the block sum.
template <int kThreads>
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wave) base += s_wave[w];
        tot += s_wave[w];
    }
    __syncthreads(); // s_wave may be reused by the caller
    *total = tot;
    return base + x - v;
}

} // namespace fmc
