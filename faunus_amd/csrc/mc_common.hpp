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

// FP32 pair energy (same algebra as pair_energy, single precision).
__device__ __forceinline__ float pair_energy_f(const DevPairF& p, const DevCoulomb& c, const float* knots,
                                               const float* coeff, float qa, float qb, float r2) {
    float u = 0.0f;
    if (p.lj_e4 != 0.0f && r2 < p.lj_cut2) {
        float x = p.lj_s2 / r2;
        x = x * x * x;
        u = u + p.lj_e4 * (x * x - x);
    }
    if (p.wca_e4 != 0.0f && !(r2 > p.wca_s2 * 1.2599210498948732f)) {
        float x = p.wca_s2 / r2;
        x = x * x * x;
        u = u + p.wca_e4 * (x * x - x + 0.25f);
    }
    if (r2 < p.hs_s2) u = u + __builtin_huge_valf();
    if (p.coul != 0.0f) {
        const float lB = static_cast<float>(c.lB);
        if (c.scheme == kCoulPM) {
            u = u + lB * qa * qb / sqrtf(r2);
        } else {
            const float r = sqrtf(r2);
            if (c.scheme == kCoulPlain) u = u + lB * (qa * qb / r);
            else if (r < static_cast<float>(c.rc)) {
                const float qq = r * static_cast<float>(c.rc_inv);
                float s;
                if (c.splined) s = spline_eval<float>(knots, coeff, c.nknots, qq);
                else if (c.scheme == kCoulEwald) s = erfcf(static_cast<float>(c.alpha * c.rc) * qq);
                else {
                    const float q5 = qq * qq * qq * qq * qq;
                    s = 1.0f - 1.75f * qq + 5.25f * q5 - 7.0f * q5 * qq + 2.5f * q5 * qq * qq;
                }
                u = u + lB * (qa * qb / r * s);
            }
        }
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


} // namespace fmc
