// Device helpers shared by the device-resident Metropolis kernels: the FP32 pair
// form, minimum-image and boundary arithmetic, cell coordinates and the Metropolis
// decision.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "device_math.hpp"
#include "engine.hpp"

namespace fmc {

constexpr int kMaxGk = 1025; // |n|^2 table (ncut <= 32)

namespace {
// Zero fill of device words by one block (a kernel rather than a memset, so that a
// captured launch sequence holds no memset node).
__global__ void zero_words_kernel(unsigned long long* __restrict__ p, int words) {
    for (int i = static_cast<int>(threadIdx.x); i < words; i += static_cast<int>(blockDim.x)) p[i] = 0ull;
}
} // namespace

struct alignas(16) DevPairF {
    float lj_s2, lj_e4, wca_s2, wca_e4, hs_s2, coul;
    int spl;
    float hot_r2; // below this r^2 the pair is evaluated again in FP64 (hot_radius2)
};

constexpr double kHotRepulsion = 8.0; // kT

// r^2 below which the pair is evaluated again in FP64: where its LJ or WCA repulsion
// e4 (s2/r2)^6 exceeds kHotRepulsion (FP32 rounding of r^2, grown sixfold by r^-12,
// would dominate the move's error), and inside hard-sphere contact (the FP32 form marks
// an overlap only below 1 - 2^-20 of the contact r^2, so that the FP64 form decides the
// pairs the two distances could place on different sides).
__device__ __forceinline__ double hot_radius2(const DevPair& p) {
    double r2 = p.hs_s2;
    if (p.lj_e4 > 0.0) r2 = fmax(r2, p.lj_s2 * sqrt(cbrt(p.lj_e4 / kHotRepulsion)));
    if (p.wca_e4 > 0.0) r2 = fmax(r2, p.wca_s2 * sqrt(cbrt(p.wca_e4 / kHotRepulsion)));
    return r2;
}

__device__ __forceinline__ DevPairF to_f32(const DevPair& p) {
    return DevPairF{static_cast<float>(p.lj_s2), static_cast<float>(p.lj_e4), static_cast<float>(p.wca_s2),
                    static_cast<float>(p.wca_e4), static_cast<float>(p.hs_s2), static_cast<float>(p.coul), p.spl,
                    static_cast<float>(hot_radius2(p))};
}

// An empty asm statement that takes and returns the values: the loads that produce
// them happen here, unconditionally, and are not moved into the branches the compiler
// would otherwise make of the selects that use them (a branch per load serialises the
// LDS latency of every pair term).
__device__ __forceinline__ void materialize(float4& v) { asm("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }

// Pair terms present anywhere in a replica's pair table (a uniform switch per kernel).
constexpr int kTermLJ = 1, kTermWCA = 2, kTermHS = 4, kTermAll = 7;

// FP64 pair energy with the same term switch (exact arithmetic of pair_energy).
__device__ __forceinline__ double pair_energy_t(const DevPair& p, const DevCoulomb& c, double qa, double qb, double r2,
                                                int terms) {
    if (c.nspl > 0) {
        bool below;
        const double us = splined_pair(c.spl[p.spl], r2, below);
        if (!below) return us;
    }
    double u = 0.0;
    if (terms & kTermLJ) {
        double x = p.lj_s2 / r2;
        x = x * x * x;
        u = u + ((p.lj_e4 != 0.0) ? p.lj_e4 * (x * x - x) : 0.0);
    }
    if (terms & kTermWCA) {
        double x = p.wca_s2 / r2;
        x = x * x * x;
        u = u + ((p.wca_e4 != 0.0 && !(r2 > p.wca_s2 * 1.2599210498948732)) ? p.wca_e4 * (x * x - x + 0.25) : 0.0);
    }
    if (terms & kTermHS) u = u + ((r2 < p.hs_s2) ? __builtin_huge_val() : 0.0);
    if (c.scheme != kCoulNone) {
        const double e = coulomb_pair(c, qa, qb, r2);
        u = u + ((p.coul != 0.0) ? e : 0.0);
    }
    return u;
}

// FP32 forms: the Coulomb short-range function S(q) of the splined (or exact)
// Ewald/Fanourgakis schemes as a uniform cubic Hermite table on q in [0, 1], built on
// the host from the FP64 form's S(q) and S'(q) (interpolation error ~1e-13, far below
// FP32 rounding), and the scheme constants held in registers.
constexpr int kSTable = 1024;

struct CoulF {
    int scheme;
    float lB, rc, rc_inv;
    const DevSplineTable* spl; // nonbonded_splined tables (in LDS), or null
};

__device__ __forceinline__ CoulF coul_f32(const DevCoulomb& c) {
    return CoulF{c.scheme, static_cast<float>(c.lB), static_cast<float>(c.rc), static_cast<float>(c.rc_inv),
                 c.nspl > 0 ? c.spl : nullptr};
}

// Coordinates of the FP32-pair form: each axis in 32-bit fixed point, the cubic box
// [-L/2, L/2) mapped onto [0, 2^32) in steps of h = L / 2^32, and the charge as FP32
// bits.  The difference of two fixed-point words taken modulo 2^32 and read as a signed
// integer is the minimum-image displacement in units of h, with the same resolution
// everywhere in the box (h = 6e-8 A at L = 255 A, where an FP32 coordinate near the box
// edge resolves 8e-6 A and LJ's r^-12 would turn that into 1e-5 relative errors of
// contact pairs).  Integer displacements also convert exactly to FP64 for the pairs
// re-evaluated there.
struct alignas(16) FixPos {
    uint32_t x, y, z;
    float q;
};

constexpr double kFix = 4294967296.0;        // 2^32
constexpr double kFixInv = 2.3283064365386963e-10; // 2^-32

// round(x 2^32 / L) + 2^31 modulo 2^32 for x in [-L/2, L/2]; `s` = 2^32 / L.  One
// multiply and a round to nearest even: the same word on the host and the device.
__host__ __device__ inline uint32_t fixed_coord(double x, double s) {
    return static_cast<uint32_t>(static_cast<long long>(rint(x * s)) + 2147483648ll);
}

__host__ __device__ inline FixPos fixed_pos(double x, double y, double z, double q, double inv) {
    const double s = kFix * inv;
    return FixPos{fixed_coord(x, s), fixed_coord(y, s), fixed_coord(z, s), static_cast<float>(q)};
}

// GroupCutoff::cut (energy.h:761-768) between a moved molecular group and static group
// gj: the groups' flag words (bit 0 molecular, molecule type from bit 8), the accepted
// mass centres and the squared cutoff per pair of molecule types.
struct GroupCut {
    const int* flags;
    const double* cm;
    const double* table;
    int nmol;
    int fg; // flags of the moved group
};

// True when gj is molecular and its centre is at least the cutoff away from c (the moved
// group's centre in one state).
__device__ __forceinline__ bool group_cut(const GroupCut& gc, int gj, const double* c, const Box& b) {
    const int fj = gc.flags[gj];
    if (!(fj & 1)) return false;
    const double* cj = gc.cm + 3 * gj;
    return !(dist2(b, c[0], c[1], c[2], cj[0], cj[1], cj[2]) < g2g_cutoff2(gc.table, gc.nmol, gc.fg, fj));
}

// The FP32-pair form's coordinates of n particles (xyz: 3 per particle, inside the box).
inline std::vector<FixPos> fixed_positions(const double* xyz, const std::vector<double>& q, int n, double inv) {
    std::vector<FixPos> out(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        out[static_cast<size_t>(i)] = fixed_pos(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], q[static_cast<size_t>(i)], inv);
    return out;
}

// Squared minimum-image distance in FP32 (h = L / 2^32 as FP32).
__device__ __forceinline__ float dist2_fix(const FixPos& a, const FixPos& b, float h) {
    const float dx = static_cast<float>(static_cast<int>(a.x - b.x)) * h;
    const float dy = static_cast<float>(static_cast<int>(a.y - b.y)) * h;
    const float dz = static_cast<float>(static_cast<int>(a.z - b.z)) * h;
    return dx * dx + dy * dy + dz * dz;
}

// Hot pairs (inside their type pair's hot radius, hot_radius2) are evaluated again in
// FP64; an FP32 hard-sphere overlap (+inf) stands, except within 2^-20 of contact where
// the FP32 form adds nothing and the FP64 form decides.  The pair loops queue hot pairs in LDS (the neighbour, and moved atom
// << 1 | state with state 0 new and 1 old); after the loops each queued pair is
// evaluated from the FP64 coordinates, so the FP64 pair code stays out of the loops'
// registers.  Loops that sum blocks of neighbours per lane queue all pairs of a block
// that holds a hot pair, or none; a pair or block that finds the queue full keeps its
// FP32 energy.
constexpr int kHotCap = 256;
constexpr float kHsSure = 0.99999904632568359375f; // 1 - 2^-20

// Queue one hot pair; false when the queue is full.
__device__ __forceinline__ bool push_hot(int* n, int2* q, int j, int tag) {
    const int s = atomicAdd(n, 1);
    if (s >= kHotCap) return false;
    q[s] = make_int2(j, tag);
    return true;
}

// Reserve cnt consecutive queue entries, all or none (the count never passes the
// capacity, so every entry below it has been written once the loops end); -1 when full.
__device__ __forceinline__ int reserve_hot(int* n, int cnt) {
    int cur = __hip_atomic_load(n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (true) {
        if (cur + cnt > kHotCap) return -1;
        const int prev = atomicCAS(n, cur, cur + cnt);
        if (prev == cur) return cur;
        cur = prev;
    }
}

// FP32 pair energy; `hot` tells whether the pair must be evaluated again in FP64.
__device__ __forceinline__ float pair_energy_ft(const DevPairF& p, const CoulF& c, const float4* __restrict__ tab,
                                                int terms, float qa, float qb, float r2, bool& hot) {
    // both halves of the pair parameters in two vector loads, then register arithmetic
    float4 pa = reinterpret_cast<const float4*>(&p)[0]; // lj_s2, lj_e4, wca_s2, wca_e4
    float4 pb = reinterpret_cast<const float4*>(&p)[1]; // hs_s2, coul, spl (bits), hot_r2
    materialize(pa);
    materialize(pb);
    hot = r2 < pb.w;
    if (c.spl != nullptr) {
        // nonbonded_splined: the table is evaluated in FP64 from the FP32 distance
        bool below;
        const double us = splined_pair(c.spl[__float_as_int(pb.z)], static_cast<double>(r2), below);
        if (!below) return static_cast<float>(us);
    }
    float u = 0.0f;
    if (terms & (kTermLJ | kTermWCA)) {
        const float ir2 = __builtin_amdgcn_rcpf(r2);
        if (terms & kTermLJ) {
            float x = pa.x * ir2;
            x = x * x * x;
            u += pa.y * (x * x - x);
        }
        if (terms & kTermWCA) {
            float w = pa.z * ir2;
            w = w * w * w;
            u += (r2 > pa.z * 1.2599210498948732f) ? 0.0f : pa.w * (w * w - w + 0.25f);
        }
    }
    if (terms & kTermHS) u += (r2 < pb.x * kHsSure) ? __builtin_huge_valf() : 0.0f;
    if (c.scheme != kCoulNone) {
        const float ir = __builtin_amdgcn_rsqf(r2);
        float e;
        if (c.scheme == kCoulPM || c.scheme == kCoulPlain) {
            e = c.lB * (qa * qb * ir);
        } else {
            const float r = r2 * ir;
            const float t = r * c.rc_inv * static_cast<float>(kSTable);
            int i = static_cast<int>(t);
            i = i > kSTable - 1 ? kSTable - 1 : i;
            i = i < 0 ? 0 : i;
            const float f = t - static_cast<float>(i);
            float4 k = tab[i];
            materialize(k);
            const float sv = ((k.w * f + k.z) * f + k.y) * f + k.x;
            e = (r < c.rc) ? c.lB * (qa * qb * ir * sv) : 0.0f;
        }
        u += (pb.y != 0.0f) ? e : 0.0f;
    }
    return u;
}

// Pair sums of one moved particle, at its new (a) and old (b) position, against every
// particle j in [jb, n) except `skip`; pairs inside group g count only if same_ok.
// Lanes stride over j; each lane loads kU neighbours before evaluating them.  FP32
// terms are summed in FP32 per block of kU and then in FP64; hot pairs go to the queue
// (hot_n, hot_q) with tag = moved atom << 1.
template <int kThreads>
__device__ __forceinline__ void pair_row_f32(const FixPos* __restrict__ pos32, const uint32_t* __restrict__ meta, int jb,
                                             int n, int skip, int g, bool same_ok, const DevPairF* row, const CoulF& c,
                                             const float4* tab, int terms, float h, const FixPos& a, const FixPos& b,
                                             int* hot_n, int2* hot_q, int tag, double& e_new, double& e_old) {
    constexpr int kU = 4;
    for (int j0 = jb + static_cast<int>(threadIdx.x); j0 < n; j0 += kThreads * kU) {
        FixPos pj[kU];
        uint32_t mj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int j = j0 + u * kThreads;
            pj[u] = j < n ? pos32[j] : FixPos{0u, 0u, 0u, 0.0f};
            mj[u] = j < n ? meta[j] : 0u;
        }
        float sn = 0.0f, so = 0.0f;
        bool anyhot = false;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int j = j0 + u * kThreads;
            const bool ok = j < n && j != skip && (static_cast<int>(mj[u] >> 8) != g || same_ok);
            const DevPairF& p = row[ok ? static_cast<int>(mj[u] & 0xffu) : 0];
            bool hn, ho;
            const float un = pair_energy_ft(p, c, tab, terms, a.q, pj[u].q, dist2_fix(a, pj[u], h), hn);
            const float uo = pair_energy_ft(p, c, tab, terms, a.q, pj[u].q, dist2_fix(b, pj[u], h), ho);
            sn += ok ? un : 0.0f;
            so += ok ? uo : 0.0f;
            anyhot = anyhot || (ok && (hn || ho));
        }
        if (anyhot) {
            // a hot pair in this block: every pair of the block goes to the queue for
            // FP64 (a state whose block sum is +inf, an overlap, keeps it)
            const bool fin_n = sn != __builtin_huge_valf(), fin_o = so != __builtin_huge_valf();
            int cnt = 0;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int j = j0 + u * kThreads;
                const bool ok = j < n && j != skip && (static_cast<int>(mj[u] >> 8) != g || same_ok);
                cnt += ok ? (fin_n ? 1 : 0) + (fin_o ? 1 : 0) : 0;
            }
            const int base = cnt > 0 ? reserve_hot(hot_n, cnt) : -1;
            if (base >= 0) {
                int k = base;
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int j = j0 + u * kThreads;
                    const bool ok = j < n && j != skip && (static_cast<int>(mj[u] >> 8) != g || same_ok);
                    if (ok && fin_n) hot_q[k++] = make_int2(j, tag);
                    if (ok && fin_o) hot_q[k++] = make_int2(j, tag | 1);
                }
                if (fin_n) sn = 0.0f;
                if (fin_o) so = 0.0f;
            }
        }
        e_new += static_cast<double>(sn);
        e_old += static_cast<double>(so);
    }
}

template <int kThreads>
__device__ __forceinline__ void pair_row_f64(const double4* __restrict__ pos, const uint32_t* __restrict__ meta, int jb,
                                             int n, int skip, int g, bool same_ok, const DevPair* row,
                                             const DevCoulomb& c, int terms, const Box& box, double ax, double ay,
                                             double az, double bx, double by, double bz, double qa, double& e_new,
                                             double& e_old) {
    constexpr int kU = 4;
    for (int j0 = jb + static_cast<int>(threadIdx.x); j0 < n; j0 += kThreads * kU) {
        double4 pj[kU];
        uint32_t mj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int j = j0 + u * kThreads;
            pj[u] = j < n ? pos[j] : make_double4(0.0, 0.0, 0.0, 0.0);
            mj[u] = j < n ? meta[j] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int j = j0 + u * kThreads;
            const bool ok = j < n && j != skip && (static_cast<int>(mj[u] >> 8) != g || same_ok);
            const DevPair& p = row[ok ? static_cast<int>(mj[u] & 0xffu) : 0];
            const double un = pair_energy_t(p, c, qa, pj[u].w, dist2(box, ax, ay, az, pj[u].x, pj[u].y, pj[u].z), terms);
            const double uo = pair_energy_t(p, c, qa, pj[u].w, dist2(box, bx, by, bz, pj[u].x, pj[u].y, pj[u].z), terms);
            e_new += ok ? un : 0.0;
            e_old += ok ? uo : 0.0;
        }
    }
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
