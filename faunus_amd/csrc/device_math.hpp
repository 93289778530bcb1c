// Device-side pair arithmetic shared by the per-move kernels and the
// device-resident Metropolis kernels.  Formulas follow the reference's pair
// potentials (potentials.h:42-49 LJ, 151-160 WCA, 203-207 hard sphere, 472-476
// plain Coulomb, 591-598 CoulombGalore wrapper) and the orthogonal minimum image
// (geometry.h:429-485).
#pragma once
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace fmc {

struct Box {
    double L[3], half[3];
};

__device__ __forceinline__ double min_image(double d, double L, double half) {
    if (d > half) d -= L;
    else if (d < -half) d += L;
    return d;
}

__device__ __forceinline__ double dist2(const Box& b, double ax, double ay, double az, double bx, double by,
                                        double bz) {
    const double dx = min_image(ax - bx, b.L[0], b.half[0]);
    const double dy = min_image(ay - by, b.L[1], b.half[1]);
    const double dz = min_image(az - bz, b.L[2], b.half[2]);
    return dx * dx + dy * dy + dz * dz;
}

// tabulate.h:184-202: interval by lower_bound, Horner from the top coefficient.  The
// lower_bound starts one knot before the first knot of x's uniform bucket and takes a
// fixed number of compare-and-step rounds (enough for the fullest bucket), which finds
// the same index as a binary search without data-dependent branches.
template <class R>
__device__ __forceinline__ R spline_eval(const R* __restrict__ knots, const R* __restrict__ coeff, int n, R x,
                                         const unsigned short* __restrict__ bucket, int steps) {
    int b = static_cast<int>(x * static_cast<R>(kSplineBuckets));
    b = b < 0 ? 0 : (b > kSplineBuckets - 1 ? kSplineBuckets - 1 : b);
    int lo = static_cast<int>(bucket[b]) - 1;
    lo = lo < 0 ? 0 : lo;
    for (int s = 0; s < steps; ++s) {
        const R k = knots[lo < n ? lo : n - 1];
        lo += (lo < n && k < x) ? 1 : 0;
    }
    int pos = lo - 1;
    pos = pos < 0 ? 0 : (pos > n - 2 ? n - 2 : pos);
    const R* c = coeff + 6 * pos;
    const R z = x - knots[pos];
    R s = R(0);
    s = z * (s + c[5]);
    s = z * (s + c[4]);
    s = z * (s + c[3]);
    s = z * (s + c[2]);
    s = z * (s + c[1]);
    return s + c[0];
}

// Coulomb short-range part of one pair (kT).
__device__ __forceinline__ double coulomb_pair(const DevCoulomb& c, double qa, double qb, double r2) {
    if (c.scheme == kCoulPM) return c.lB * qa * qb / sqrt(r2);
    const double r = sqrt(r2) + 2.220446049250313e-16;
    if (c.scheme == kCoulPlain) return c.lB * (qa * qb / r);
    if (!(r < c.rc)) return 0.0;
    const double q = r * c.rc_inv;
    double s;
    if (c.splined) s = spline_eval<double>(c.knots, c.coeff, c.nknots, q, c.bucket, c.spline_steps);
    else if (c.scheme == kCoulEwald) s = erfc(c.alpha * c.rc * q);
    else {
        const double q5 = q * q * q * q * q;
        s = 1.0 - 1.75 * q + 5.25 * q5 - 7.0 * q5 * q + 2.5 * q5 * q * q;
    }
    return c.lB * (qa * qb / r * s);
}

// Additive pair energy u = LJ + WCA + HS + Coulomb (absent terms have zero parameters).
__device__ __forceinline__ double pair_energy_exact(const DevPair& p, const DevCoulomb& c, double qa, double qb,
                                                    double r2) {
    double u = 0.0;
    if (p.lj_e4 != 0.0) {
        double x = p.lj_s2 / r2;
        x = x * x * x;
        u = u + p.lj_e4 * (x * x - x);
    }
    if (p.wca_e4 != 0.0 && !(r2 > p.wca_s2 * 1.2599210498948732)) {
        double x = p.wca_s2 / r2;
        x = x * x * x;
        u = u + p.wca_e4 * (x * x - x + 0.25);
    }
    if (r2 < p.hs_s2) u = u + __builtin_huge_val();
    if (p.coul != 0.0) u = u + coulomb_pair(c, qa, qb, r2);
    return u;
}

// Andrea spline of a nonbonded_splined table at r2 in (rmin2, rmax2) (tabulate.h:184-202:
// interval by lower_bound, Horner from the top coefficient).  The search starts one
// knot before the first knot of r2's bucket; within the table's `steps` knots from there
// lower_bound's index advances once per knot below r2, and the knots are increasing, so
// those knots are read independently and counted (one LDS round, not `steps` dependent ones).
__device__ __forceinline__ double spline_table_eval(const DevSplineTable& T, double r2) {
    int b = static_cast<int>((r2 - T.x0) * T.inv_w);
    b = b < 0 ? 0 : (b > kSplBuckets - 1 ? kSplBuckets - 1 : b);
    int lo = static_cast<int>(T.bucket[b]) - 1;
    lo = lo < 0 ? 0 : lo;
    const int n = T.n, steps = T.steps;
    int below = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int k = lo + s;
        const double v = T.knots[k < n ? k : n - 1];
        below += (s < steps && k < n && v < r2) ? 1 : 0;
    }
    for (int s = 8; s < steps; ++s) {
        const int k = lo + s;
        below += (k < n && T.knots[k < n ? k : n - 1] < r2) ? 1 : 0;
    }
    lo += below;
    int pos = lo - 1;
    pos = pos < 0 ? 0 : (pos > n - 2 ? n - 2 : pos);
    const double* c = T.coeff + 6 * pos;
    const double z = r2 - T.knots[pos];
    double s = 0.0;
    s = z * (s + c[5]);
    s = z * (s + c[4]);
    s = z * (s + c[3]);
    s = z * (s + c[2]);
    s = z * (s + c[1]);
    return s + c[0];
}

// SplinedPotential::operator() (potentials.h:807-823): zero from rmax on, the spline
// above rmin, +inf below rmin with hard-sphere repulsion, else the exact energy.
// `below` is true when the exact energy is needed (r2 <= rmin2, no hard sphere).
__device__ __forceinline__ double splined_pair(const DevSplineTable& T, double r2, bool& below) {
    below = false;
    if (r2 >= T.rmax2) return 0.0;
    if (r2 > T.rmin2) return spline_table_eval(T, r2);
    if (T.hs) return __builtin_huge_val();
    below = true;
    return 0.0;
}

// The nonbonded term's pair energy: splined when the term is nonbonded_splined.
__device__ __forceinline__ double pair_energy(const DevPair& p, const DevCoulomb& c, double qa, double qb,
                                              double r2) {
    if (c.nspl > 0) {
        bool below;
        const double u = splined_pair(c.spl[p.spl], r2, below);
        if (!below) return u;
    }
    return pair_energy_exact(p, c, qa, qb, r2);
}

// Copy of a struct into LDS by all threads of the block, 16 bytes per thread and step.
template <class T>
__device__ __forceinline__ void lds_copy(T& dst, const T* __restrict__ src, int tid, int nthreads) {
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "16-byte granules");
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(&dst);
    for (int i = tid; i < static_cast<int>(sizeof(T) / 16); i += nthreads) d[i] = s[i];
}

// Lane `lane` (uniform) of x, in every lane.
__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Sum of x over the 64 lanes of a wave, the same value in every lane and the same order
// of additions on every call: DPP moves within rows of 16 lanes (pair and quad swaps,
// half-row and row mirrors; no LDS round trip), then the four row sums read from lanes
// 0, 16, 32 and 48.  Every lane of the wave must be active.
__device__ __forceinline__ double wave_sum(double x) {
    x += __builtin_amdgcn_update_dpp(0.0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    x += __builtin_amdgcn_update_dpp(0.0, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    x += __builtin_amdgcn_update_dpp(0.0, x, 0x141, 0xF, 0xF, false); // row_half_mirror
    x += __builtin_amdgcn_update_dpp(0.0, x, 0x140, 0xF, 0xF, false); // row_mirror
    const double r0 = readlane_f64(x, 0), r1 = readlane_f64(x, 16);
    const double r2 = readlane_f64(x, 32), r3 = readlane_f64(x, 48);
    return (r0 + r1) + (r2 + r3);
}

// Wavefront (64 lanes) sum, then across the block's waves.
template <int kThreads>
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kThreads / 64; ++w) s += lds[w];
    }
    return s; // valid in thread 0
}

// Intra-molecular pair exclusion (MoleculeData::isPairExcluded) from per-particle
// sorted partner lists; no lists means no exclusions.
__device__ __forceinline__ bool pair_excluded(const int* __restrict__ off, const int* __restrict__ list, int a, int b) {
    if (off == nullptr) return false;
    int lo = off[a], hi = off[a + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == b) return true;
        if (v < b) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

// GroupCutoff::cut's squared cutoff (energy.h:761-768) for two groups' flag words
// (molecule type from bit 8) from the per-pair table of molecule types.
__device__ __forceinline__ double g2g_cutoff2(const double* __restrict__ table, int nmol, int flags_a, int flags_b) {
    return table[(flags_a >> 8) * nmol + (flags_b >> 8)];
}

__device__ __forceinline__ int unpack_n(int packed, int which) {
    return static_cast<int>(static_cast<signed char>((packed >> (8 * which)) & 0xff));
}

} // namespace fmc
