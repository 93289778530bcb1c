// Device-resident Metropolis over a general move set: translations of single atoms
// inside any molecule type, rigid translate+rotate of molecules and pivot rotations
// of chains.  One workgroup advances one replica; each move's nonbonded pair dU, Ewald
// reciprocal dU and harmonic bonded dU are evaluated in the same pass (the reference's
// MoveCollection::sample + performMove loop, move.cpp:513-525 and
// montecarlo.cpp:139-187, with the proposals of move.cpp:225-345 (atoms),
// move.cpp:1618-1689 (molecules) and chainmove.cpp:76-104, 216-268 (pivot)).
//
// Split form: a large replica is shared by `parts` workgroups.  Part p owns the p-th
// slice of the k-vectors (its Q(k) stays in registers) and the p-th slice of the
// static particles of the pair sums.  Every part draws the same proposal, publishes
// its five partial sums in an exchange slot (agent-scope release store of the move's
// tag), waits for the other parts' slots (acquire loads, bounded spin), adds the
// partials in part order and so takes the same decision as every other part.  Every
// part writes accepted coordinates and mass centres itself, so a part only ever reads
// particle data it wrote; the slot exchange orders each part's writes of one move
// before any part's writes of a later move.  The launch is cooperative.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <type_traits>

#include "device_math.hpp"
#include "mc_batch.hpp"
#include "mc_common.hpp"
#include "philox.hpp"

namespace fmc {
namespace {

constexpr int kMaxKinds = 8;   // move kinds per replica
constexpr int kChunk = 8;      // moved atoms per Ewald phase-table chunk
constexpr int kMaxComb = 512;  // (moved atom, state, cell) combinations of the flattened stencil walk
constexpr int kMaxParts = 64;  // workgroups per replica in the split form (one waiting lane each)
constexpr int kMolStencil = 16; // moved atoms that may share one stencil (each member against all of them)

// Exchange slot of one part for one move parity: the partial sums (pair new/old,
// Ewald, bonds new/old) and the tag of the move they belong to (move index + 1).
struct alignas(64) GeneralSlot {
    double v[5];
    unsigned long long seq;
};

__device__ __forceinline__ int pick(uint32_t x, int n) {
    return static_cast<int>((static_cast<uint64_t>(x) * static_cast<uint64_t>(n)) >> 32);
}

// Isotropic unit vector from z = 2u-1 and phi = 2 pi u'.
__device__ __forceinline__ void unit_vector(uint32_t zw, uint32_t pw, double* v) {
    const double z = 2.0 * u01_from32(zw) - 1.0;
    const double phi = 6.283185307179586 * u01_from32(pw);
    const double s = sqrt(1.0 - z * z);
    v[0] = s * cos(phi);
    v[1] = s * sin(phi);
    v[2] = z;
}

// Rotation by `angle` about the unit vector u (Rodrigues' formula), row-major.
__device__ __forceinline__ void rotation_matrix_sc(const double* u, double s, double c, double* R) {
    const double sa0 = u[0] * s, sa1 = u[1] * s, sa2 = u[2] * s;
    const double ca0 = u[0] * (1.0 - c), ca1 = u[1] * (1.0 - c), ca2 = u[2] * (1.0 - c);
    R[0] = ca0 * u[0] + c;
    R[1] = ca0 * u[1] - sa2;
    R[2] = ca0 * u[2] + sa1;
    R[3] = ca0 * u[1] + sa2;
    R[4] = ca1 * u[1] + c;
    R[5] = ca1 * u[2] - sa0;
    R[6] = ca0 * u[2] - sa1;
    R[7] = ca1 * u[2] + sa0;
    R[8] = ca2 * u[2] + c;
}

__device__ __forceinline__ void rotation_matrix(const double* u, double angle, double* R) {
    rotation_matrix_sc(u, sin(angle), cos(angle), R);
}

// Random part of one proposal: everything that does not depend on the current
// coordinates (move kind, target group and atoms, displacement, rotation, acceptance
// deviate).  Drawn kDraw moves ahead, one lane of wave 0 per move, so that the
// Philox blocks, the transcendental functions and the table lookups of 64 proposals
// overlap instead of forming one serial chain per move.
constexpr int kDraw = 64;

struct MoveDraw {
    int kind;                // -1: nothing moves (m = 0)
    int group, mlo, mhi;
    int internal, flags;     // flags: 1 translate, 2 rotate (kind 1); 4 the moved atom has bonds (kind 0)
    int atom0, atom1;        // kind 2: pivot origin and the bond partner
    double accept_u, scale;
    double dir[3];           // kind 0: unit direction (times scale at the trial); kind 1: displacement
    double R[9];             // kind 1 rotation; kind 2: R[0], R[1] = sin, cos of the angle
};

// move.cpp:225-345 (atoms), move.cpp:1618-1689 (molecules), chainmove.cpp:216-268 (pivot)
__device__ void draw_move(MoveDraw& D, unsigned long long mvi, uint32_t replica, unsigned long long seed,
                          const MoveKindDev* kinds, int nkinds, unsigned long long total_weight,
                          const int* __restrict__ kind_groups, const int* __restrict__ gbegin,
                          const int* __restrict__ gsize, const int2* __restrict__ pivot_bonds,
                          const uint32_t* __restrict__ meta, const double* dpa, const int* __restrict__ bond_off) {
    const uint32_t c0 = static_cast<uint32_t>(mvi), c1 = static_cast<uint32_t>(mvi >> 32);
    const uint32_t key0 = static_cast<uint32_t>(seed), key1 = static_cast<uint32_t>(seed >> 32);
    const Philox4 ra = philox4x32_10(c0, c1, replica, 0u, key0, key1);
    const Philox4 rb = philox4x32_10(c0, c1, replica, 1u, key0, key1);
    const Philox4 rc = philox4x32_10(c0, c1, replica, 2u, key0, key1);
    D.accept_u = (static_cast<double>(rb.v[0] >> 5) * 67108864.0 + static_cast<double>(rb.v[1] >> 6)) *
                 1.1102230246251565e-16;
    D.kind = -1;
    D.group = -1;
    D.mlo = D.mhi = 0;
    D.internal = 0;
    D.flags = 0;
    D.atom0 = D.atom1 = 0;
    D.scale = 0.0;
    const uint64_t sel = (static_cast<uint64_t>(rb.v[2]) * total_weight) >> 32;
    int k = 0;
    while (k + 1 < nkinds && sel >= kinds[k].cum) ++k;
    const MoveKindDev K = kinds[k];
    if (K.kind == 0) {
        const int g = kind_groups[K.group_first + (K.group_count == 1 ? 0 : pick(rc.v[3], K.group_count))];
        D.group = g;
        const int size = gsize[g];
        if (size > 0) {
            const int i = gbegin[g] + pick(ra.v[0], size);
            const int ti = static_cast<int>(meta[i] & 0xffu);
            const double step = dpa[ti] >= 0.0 ? dpa[ti] : K.dp;
            if (step > 0.0) {
                unit_vector(ra.v[1], ra.v[2], D.dir);
                D.scale = step * (static_cast<double>(ra.v[3]) * 2.3283064365386963e-10);
                D.kind = 0;
                D.mlo = i;
                D.mhi = i + 1;
                D.internal = 1;
                if (bond_off != nullptr && bond_off[i + 1] > bond_off[i]) D.flags = 4;
            }
        }
    } else if (K.kind == 1) {
        const int g = kind_groups[K.group_first + pick(ra.v[0], K.group_count)];
        D.group = g;
        const int size = gsize[g];
        if (size > 0) {
            if (K.dp > 0.0) {
                unit_vector(ra.v[1], ra.v[2], D.dir);
                D.scale = K.dp * (static_cast<double>(ra.v[3]) * 2.3283064365386963e-10);
                D.dir[0] *= D.scale;
                D.dir[1] *= D.scale;
                D.dir[2] *= D.scale;
                D.flags |= 1;
            }
            if (K.dprot > DBL_EPSILON) {
                double axis[3];
                unit_vector(rc.v[0], rc.v[1], axis);
                rotation_matrix(axis, K.dprot * (u01_from32(rc.v[2]) - 0.5), D.R);
                D.flags |= 2;
            }
            if (D.flags) {
                D.kind = 1;
                D.mlo = gbegin[g];
                D.mhi = D.mlo + size;
            }
        }
    } else {
        const int g = kind_groups[K.group_first + pick(ra.v[0], K.group_count)];
        D.group = g;
        const int size = gsize[g], b0 = gbegin[g];
        if (size > 2 && K.bond_count > 0 && fabs(K.dprot) > 1e-9) {
            const int2 bd = pivot_bonds[K.bond_first + pick(rc.v[3], K.bond_count)];
            int atom0 = b0 + bd.x, atom1 = b0 + bd.y;
            int slo, shi;
            if (u01_from32(rb.v[3]) > 0.5) {
                slo = atom1 + 1;
                shi = b0 + size;
                const int tmp = atom0;
                atom0 = atom1;
                atom1 = tmp;
            } else {
                slo = b0;
                shi = atom0;
            }
            if (shi > slo) {
                const double angle = K.dprot * (u01_from32(rc.v[2]) - 0.5);
                D.R[0] = sin(angle);
                D.R[1] = cos(angle);
                D.atom0 = atom0;
                D.atom1 = atom1;
                D.kind = 2;
                D.mlo = slo;
                D.mhi = shi;
                D.internal = 1;
            }
        }
    }
}

// Proposal of one move, identical on every lane of the workgroup.
struct GeneralMove {
    int kind, group, mlo, mhi, internal;
    double scale, dir[3];       // kind 0 / kind 1 translation
    bool translate, rotate;     // kind 1
    double R[9];                // kind 1 and 2 rotation
    double cm_new[3];           // kind 1 rotation centre after the translation
    double shift[3], origin[3]; // kind 2 frame
    double accept_u;
};

// The atoms of one move: [mlo, mhi) in the moved group g, their new and old
// coordinates and types in LDS; pairs inside g count only if same_ok, minus exclusions.
struct MovedSet {
    int mlo, mhi, m, g;
    bool same_ok;
    const double4* s_new;
    const double4* s_old;
    const FixPos* s_fnew; // FP32 form: fixed-point copies of s_new and s_old in LDS, or null
    const FixPos* s_fold; //   (then converted here from s_new and s_old with inv = 1 / L)
    double inv;
    const int* s_mt;
    const int* excl_off;
    const int* excl;
    GroupCut cut;         // cutoff_g2g against the moved molecular group (kCut)
    const double* cm_new; // the moved group's mass centre (new, old state)
    const double* cm_old;
};

// Pair sums of the moved atoms (new and old coordinates) against the particles
// base + tid + u * kThreads (u < kB) below je: the kB neighbours sit in registers and
// the moved atoms stream from LDS, kB independent pair terms per atom and state.  In
// the FP32 form, hot pairs go to the queue (hot_n, hot_q).
// kExcl: the system has pair exclusions (looked up per pair inside the moved group);
// without them the loop over the neighbours is straight-line code.
template <int kThreads, int kB, bool kF32, bool kExcl, bool kCut>
__device__ __forceinline__ void moved_block(int base, int je, const MovedSet& ms, const uint32_t* __restrict__ meta,
                                            const FixPos* __restrict__ pos32, const double4* __restrict__ pos,
                                            const DevPairF* s_pairs_f, const DevPair* s_pairs, int nt, const CoulF& cf,
                                            const float4* s_stab, const DevCoulomb& s_coul, int terms, const Box& box,
                                            float hf, int* hot_n, int2* hot_q, double& e_new, double& e_old) {
    const int j0 = base + static_cast<int>(threadIdx.x);
    uint32_t mj[kB];
    double4 pj[kB];
    FixPos pf[kB];
    // bits 0-1: 0 no pair, 1 pair, 2 pair inside the moved group (exclusions apply);
    // bit 2 / bit 3: the group pair is cut in the new / old state (cutoff_g2g)
    int use[kB];
    int any = 0;
#pragma unroll
    for (int u = 0; u < kB; ++u) {
        const int j = j0 + u * kThreads;
        const bool in = j < je && !(j >= ms.mlo && j < ms.mhi);
        mj[u] = in ? meta[j] : 0u;
        if (kF32) pf[u] = in ? pos32[j] : FixPos{0u, 0u, 0u, 0.0f};
        else pj[u] = in ? pos[j] : make_double4(0.0, 0.0, 0.0, 0.0);
        const bool same = static_cast<int>(mj[u] >> 8) == ms.g;
        use[u] = !in ? 0 : (same ? (ms.same_ok ? 2 : 0) : 1);
        if (kCut && use[u] == 1) {
            const int gj = static_cast<int>(mj[u] >> 8);
            use[u] |= (group_cut(ms.cut, gj, ms.cm_new, box) ? 4 : 0) | (group_cut(ms.cut, gj, ms.cm_old, box) ? 8 : 0);
        }
        any |= use[u];
    }
    if (!any) return;
    for (int a = 0; a < ms.m; ++a) {
        const int ta = ms.s_mt[a];
        if (kF32) {
            const DevPairF* row = s_pairs_f + ta * nt;
            FixPos fa, fb;
            if (ms.s_fnew != nullptr) {
                fa = ms.s_fnew[a];
                fb = ms.s_fold[a];
            } else {
                const double4 A = ms.s_new[a], B = ms.s_old[a];
                fa = fixed_pos(A.x, A.y, A.z, A.w, ms.inv);
                fb = fixed_pos(B.x, B.y, B.z, B.w, ms.inv);
            }
            auto pair_ok = [&](int u, bool& okn, bool& oko) {
                const int uu = use[u] & 3;
                const bool ok = kExcl ? (uu == 1 || (uu == 2 && !pair_excluded(ms.excl_off, ms.excl, ms.mlo + a,
                                                                                j0 + u * kThreads)))
                                      : uu != 0;
                okn = ok && !(use[u] & 4);
                oko = ok && !(use[u] & 8);
                return ok;
            };
            float sn = 0.0f, so = 0.0f;
            bool anyhot = false;
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                bool okn, oko;
                const bool ok = pair_ok(u, okn, oko);
                const DevPairF& p = row[ok ? static_cast<int>(mj[u] & 0xffu) : 0];
                bool hn, ho;
                const float un = pair_energy_ft(p, cf, s_stab, terms, fa.q, pf[u].q, dist2_fix(fa, pf[u], hf), hn);
                const float uo = pair_energy_ft(p, cf, s_stab, terms, fa.q, pf[u].q, dist2_fix(fb, pf[u], hf), ho);
                sn += okn ? un : 0.0f;
                so += oko ? uo : 0.0f;
                anyhot = anyhot || (okn && hn) || (oko && ho);
            }
            if (anyhot) {
                // a hot pair in this block: every pair of the block goes to the queue for
                // FP64 (a state whose block sum is +inf, an overlap, keeps it)
                const bool fin_n = sn != __builtin_huge_valf(), fin_o = so != __builtin_huge_valf();
                int cnt = 0;
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    bool okn, oko;
                    pair_ok(u, okn, oko);
                    cnt += (okn && fin_n ? 1 : 0) + (oko && fin_o ? 1 : 0);
                }
                const int base = cnt > 0 ? reserve_hot(hot_n, cnt) : -1;
                if (base >= 0) {
                    int k = base;
#pragma unroll
                    for (int u = 0; u < kB; ++u) {
                        bool okn, oko;
                        pair_ok(u, okn, oko);
                        const int j = j0 + u * kThreads;
                        if (okn && fin_n) hot_q[k++] = make_int2(j, a << 1);
                        if (oko && fin_o) hot_q[k++] = make_int2(j, (a << 1) | 1);
                    }
                    if (fin_n) sn = 0.0f;
                    if (fin_o) so = 0.0f;
                }
            }
            e_new += static_cast<double>(sn);
            e_old += static_cast<double>(so);
        } else {
            const double4 A = ms.s_new[a], B = ms.s_old[a];
            const DevPair* row = s_pairs + ta * nt;
            double sn = 0.0, so = 0.0;
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int uu = use[u] & 3;
                const bool ok = kExcl ? (uu == 1 || (uu == 2 && !pair_excluded(ms.excl_off, ms.excl, ms.mlo + a,
                                                                                j0 + u * kThreads)))
                                      : uu != 0;
                const DevPair& p = row[ok ? static_cast<int>(mj[u] & 0xffu) : 0];
                const double un =
                    pair_energy_t(p, s_coul, A.w, pj[u].w, dist2(box, A.x, A.y, A.z, pj[u].x, pj[u].y, pj[u].z), terms);
                const double uo =
                    pair_energy_t(p, s_coul, B.w, pj[u].w, dist2(box, B.x, B.y, B.z, pj[u].x, pj[u].y, pj[u].z), terms);
                sn += (ok && !(use[u] & 4)) ? un : 0.0;
                so += (ok && !(use[u] & 8)) ? uo : 0.0;
            }
            e_new += sn;
            e_old += so;
        }
    }
}

template <int kThreads, int kMaxKPT, bool kCells, bool kF32, bool kSplit>
__global__ __launch_bounds__(kThreads) void mc_general_kernel(ReplicaState* __restrict__ states,
                                                               const long long* __restrict__ moves,
                                                               TraceEntry* __restrict__ trace, int trace_cap,
                                                               const int2* __restrict__ blockmap,
                                                               long long spin_limit, int fix_lds) {
    static_assert(!(kSplit && kCells), "the split form walks all particles");
    // block -> (replica, part); without split every block is part 0 of its own replica
    const int2 bm = kSplit ? blockmap[blockIdx.x] : make_int2(static_cast<int>(blockIdx.x), 0);
    const int rep = bm.x, part = bm.y;
    if (kSplit && rep < 0) return; // an idle block of a placement map
    ReplicaState& S = states[rep];
    const long long M = moves[rep];
    if (M <= 0 || S.error) return;
    const int G = kSplit ? max(1, S.parts) : 1; // replicas of one part run unsplit inside a split launch
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevPairF s_pairs_f[kF32 ? kMaxTypes * kMaxTypes : 1];
    __shared__ float4 s_stab[kF32 ? kSTable : 1];
    __shared__ DevCoulomb s_coul;
    __shared__ double s_gk[kMaxGk];
    __shared__ double s_dpa[kMaxTypes];
    __shared__ double s_mw[kMaxTypes];          // atom masses (cutoff_g2g mass centres)
    __shared__ double s_cm[2][3];               // the moved group's mass centre, new and old state
    __shared__ double s_cmred[4][kThreads / 64];
    __shared__ GeneralMove s_P;                 // the current move's parameters
    __shared__ MoveKindDev s_kinds[kMaxKinds];
    __shared__ double s_red[5][kThreads / 64];
    __shared__ int s_accept, s_abort;
    __shared__ int s_nhot;                        // hot pairs queued this move (FP32 form)
    __shared__ int2 s_hot[kF32 ? kHotCap : 1];
    __shared__ int s_ccell[kCells ? kMaxComb : 1];
    __shared__ int s_cpre[kCells ? kMaxComb + 1 : 1];
    __shared__ int s_wsum[kThreads / 64];
    __shared__ MoveDraw s_draw[kDraw];
    // kn | phase chunk tables | new coordinates (two buffers in the split form) | old
    // coordinates | (FP32 form) fixed-point new and old coordinates | moved types
    extern __shared__ int s_dyn[];

    const int tid = threadIdx.x;
    const int n = S.n, nk = S.nk, ncut = S.ncut, nt = S.ntypes;
    const int width = 2 * ncut + 1;
    // this part's k-vectors [kb, ke) (kpt consecutive ones per thread) and static
    // particles [jb, je) of the pair sums
    const int kw = kSplit ? (nk + G - 1) / G : nk;
    const int kb = kSplit ? part * kw : 0;
    const int ke = kSplit ? min(nk, kb + kw) : nk;
    const int kpt = (kw + kThreads - 1) / kThreads;
    const int jw = kSplit ? (n + G - 1) / G : n;
    const int jb = kSplit ? min(n, part * jw) : 0;
    const int je = kSplit ? min(n, jb + jw) : n;
    const double L = S.L, half = S.half, inv = S.inv;
    const int nk8 = (nk + 7) & ~7;
    int* s_kn = s_dyn;
    double2* s_phase = reinterpret_cast<double2*>(s_dyn + nk8);
    const int max_moved = S.max_moved;
    double4* s_nbuf = reinterpret_cast<double4*>(s_phase + kChunk * 6 * width);
    double4* s_old = s_nbuf + (kSplit ? 2 : 1) * max_moved;
    // fixed-point copies of the moved atoms for the FP32 pair paths when the launch's
    // LDS holds them (fix_lds); otherwise they are converted where they are used
    const bool fixlds = kF32 && fix_lds != 0;
    FixPos* s_fnew = reinterpret_cast<FixPos*>(s_old + max_moved);
    FixPos* s_fold = s_fnew + (fixlds ? max_moved : 0);
    int* s_mt = reinterpret_cast<int*>(s_fold + (fixlds ? max_moved : 0));

    for (int i = tid; i < nt * nt; i += kThreads) {
        s_pairs[i] = S.pairs[i];
        if (kF32) {
            const DevPair& p = S.pairs[i];
            s_pairs_f[i] = to_f32(p);
        }
    }
    lds_copy(s_coul, S.coul, tid, kThreads);
    if (tid == 0) {
        s_abort = 0;
        s_nhot = 0;
    }
    for (int i = tid; i < nt; i += kThreads) {
        s_dpa[i] = S.dp_atom[i];
        s_mw[i] = S.mw != nullptr ? S.mw[i] : 1.0;
    }
    for (int i = tid; i < S.nkinds; i += kThreads) s_kinds[i] = S.kinds[i];
    for (int i = tid; i <= ncut * ncut && i < kMaxGk; i += kThreads) s_gk[i] = S.gk[i];
    for (int k = tid; k < nk; k += kThreads) s_kn[k] = S.kn[k];
    __syncthreads();
    if (kF32) {
        if (S.stab)
            for (int i = tid; i < kSTable; i += kThreads) s_stab[i] = S.stab[i];
    }
    const CoulF cf = coul_f32(s_coul);

    const int k0 = kb + tid * kpt;
    double2 q[kMaxKPT];
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) q[s] = (s < kpt && k0 + s < ke) ? S.q[k0 + s] : make_double2(0.0, 0.0);

    double4* __restrict__ pos = S.pos;
    FixPos* __restrict__ pos32 = S.pos32;
    const uint32_t* __restrict__ meta = S.meta;
    const float hf = static_cast<float>(L * kFixInv);
    const unsigned long long move0 = S.moves;
    const double beta = S.beta;
    const double pref = S.rec_prefactor;
    const double surf = S.surf;
    double dip0 = S.dip[0], dip1 = S.dip[1], dip2 = S.dip[2]; // thread 0's running dipole (every part)
    double energy = S.energy, sum_du = S.sum_du;
    unsigned long long accepted = S.accepted;
    const uint32_t replica = static_cast<uint32_t>(S.replica_id);
    const unsigned long long seed = S.seed;
    const Box box{{L, L, L}, {half, half, half}};
    const int nkinds = S.nkinds;
    const unsigned long long total_weight = S.total_weight;
    const int* __restrict__ kind_groups = S.kind_groups;
    const int* __restrict__ gbegin = S.gbegin;
    const int* __restrict__ gsize = S.gsize;
    const int* __restrict__ gflags = S.gflags;
    const int2* __restrict__ pivot_bonds = S.pivot_bonds;
    double* __restrict__ gcm = S.gcm;
    const int* __restrict__ bond_off = S.bond_off;
    const DevBond* __restrict__ bonds = S.bonds;
    const int* __restrict__ excl_off = S.excl_off;
    const int* __restrict__ excl = S.excl;
    const int terms = S.terms;
    const double* __restrict__ g2g2 = S.g2g2; // cutoff_g2g per pair of molecule types, or null
    const long long sample_every = S.sample_every;
    unsigned long long e_samples = S.e_samples;
    double e_sum = S.e_sum, e_sum2 = S.e_sum2;
    // split form: a part sees the coordinates (and mass centre) the previous move
    // committed in other parts only after that move's successor has been exchanged, so
    // reads of the previous move's accepted atoms take this part's own copy
    int prev_lo = 0, prev_hi = 0, prev_cg = -1;
    double prev_cm[3] = {0.0, 0.0, 0.0};
    auto load_pos = [&](int idx, const double4* s_prev) -> double4 {
        if (kSplit && idx >= prev_lo && idx < prev_hi) return s_prev[idx - prev_lo];
        return pos[idx];
    };
    long long t_done = M; // moves decided by this launch
    __syncthreads();

    for (long long t = 0; t < M; ++t) {
        // ---- proposal: the drawn part, completed with the current coordinates
        const unsigned long long mvi = move0 + static_cast<unsigned long long>(t);
        const int slot = static_cast<int>(t % kDraw);
        if (slot == 0) {
            // the previous draws were last read before this move's barriers
            if (tid < kDraw)
                draw_move(s_draw[tid], mvi + static_cast<unsigned long long>(tid), replica, seed, s_kinds, nkinds,
                          total_weight, kind_groups, gbegin, gsize, pivot_bonds, meta, s_dpa, bond_off);
            __syncthreads();
        }
        double4* s_new = s_nbuf + (kSplit ? (t & 1) * max_moved : 0);
        const double4* s_prev = s_nbuf + (kSplit ? ((t + 1) & 1) * max_moved : 0);
        const MoveDraw& D = s_draw[slot];
        GeneralMove Pm;
        Pm.accept_u = D.accept_u;
        Pm.kind = D.kind;
        Pm.group = D.group;
        Pm.mlo = D.mlo;
        Pm.mhi = D.mhi;
        Pm.internal = D.internal;
        Pm.translate = (D.flags & 1) != 0;
        Pm.rotate = (D.flags & 2) != 0;
        Pm.scale = D.scale;
        Pm.dir[0] = D.dir[0];
        Pm.dir[1] = D.dir[1];
        Pm.dir[2] = D.dir[2];
        if (Pm.kind == 1) {
            const int g = Pm.group;
            if (kSplit && g == prev_cg) {
                Pm.cm_new[0] = prev_cm[0];
                Pm.cm_new[1] = prev_cm[1];
                Pm.cm_new[2] = prev_cm[2];
            } else {
                Pm.cm_new[0] = gcm[3 * g];
                Pm.cm_new[1] = gcm[3 * g + 1];
                Pm.cm_new[2] = gcm[3 * g + 2];
            }
            if (Pm.translate) {
                Pm.cm_new[0] = wrap(Pm.cm_new[0] + Pm.dir[0], L, half, inv);
                Pm.cm_new[1] = wrap(Pm.cm_new[1] + Pm.dir[1], L, half, inv);
                Pm.cm_new[2] = wrap(Pm.cm_new[2] + Pm.dir[2], L, half, inv);
            }
            if (Pm.rotate)
                for (int e = 0; e < 9; ++e) Pm.R[e] = D.R[e];
        } else if (Pm.kind == 2) {
            const double4 pa = load_pos(D.atom0, s_prev), pb = load_pos(D.atom1, s_prev);
            Pm.shift[0] = pa.x;
            Pm.shift[1] = pa.y;
            Pm.shift[2] = pa.z;
            Pm.origin[0] = wrap(pa.x - Pm.shift[0], L, half, inv);
            Pm.origin[1] = wrap(pa.y - Pm.shift[1], L, half, inv);
            Pm.origin[2] = wrap(pa.z - Pm.shift[2], L, half, inv);
            double v[3] = {Pm.origin[0] - wrap(pb.x - Pm.shift[0], L, half, inv),
                           Pm.origin[1] - wrap(pb.y - Pm.shift[1], L, half, inv),
                           Pm.origin[2] - wrap(pb.z - Pm.shift[2], L, half, inv)};
            for (int d = 0; d < 3; ++d) v[d] = min_image(v[d], L, half);
            const double norm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const double u[3] = {v[0] / norm, v[1] / norm, v[2] / norm};
            rotation_matrix_sc(u, D.R[0], D.R[1], Pm.R);
        }
        const int m = Pm.mhi - Pm.mlo;
        const int g = Pm.group;
        // the move's parameters in LDS for the rest of the move (uniform values that
        // would otherwise hold registers through the pair loops); read after the barrier
        // that ends the trial coordinates
        if (tid == 0) s_P = Pm;

        // ---- trial coordinates of the moved atoms
        for (int a = tid; a < m; a += kThreads) {
            const int idx = Pm.mlo + a;
            const double4 p = load_pos(idx, s_prev);
            s_old[a] = p;
            s_mt[a] = static_cast<int>(meta[idx] & 0xffu);
            double x = p.x, y = p.y, z = p.z;
            if (Pm.kind == 0) {
                x = wrap(x + Pm.dir[0] * Pm.scale, L, half, inv);
                y = wrap(y + Pm.dir[1] * Pm.scale, L, half, inv);
                z = wrap(z + Pm.dir[2] * Pm.scale, L, half, inv);
            } else if (Pm.kind == 1) {
                if (Pm.translate) { // Group::translate (group.cpp:115-123)
                    x = wrap(x + Pm.dir[0], L, half, inv);
                    y = wrap(y + Pm.dir[1], L, half, inv);
                    z = wrap(z + Pm.dir[2], L, half, inv);
                }
                if (Pm.rotate) { // Geometry::rotate about the mass centre (geometry.h:613-626)
                    x = wrap(x - Pm.cm_new[0], L, half, inv);
                    y = wrap(y - Pm.cm_new[1], L, half, inv);
                    z = wrap(z - Pm.cm_new[2], L, half, inv);
                    const double rx = Pm.R[0] * x + Pm.R[1] * y + Pm.R[2] * z;
                    const double ry = Pm.R[3] * x + Pm.R[4] * y + Pm.R[5] * z;
                    const double rz = Pm.R[6] * x + Pm.R[7] * y + Pm.R[8] * z;
                    x = wrap(rx + Pm.cm_new[0], L, half, inv);
                    y = wrap(ry + Pm.cm_new[1], L, half, inv);
                    z = wrap(rz + Pm.cm_new[2], L, half, inv);
                }
            } else { // ChainRotationMove::rotate_segment (chainmove.cpp:76-104)
                const double fx = wrap(x - Pm.shift[0], L, half, inv) - Pm.origin[0];
                const double fy = wrap(y - Pm.shift[1], L, half, inv) - Pm.origin[1];
                const double fz = wrap(z - Pm.shift[2], L, half, inv) - Pm.origin[2];
                const double rx = Pm.R[0] * fx + Pm.R[1] * fy + Pm.R[2] * fz + Pm.origin[0];
                const double ry = Pm.R[3] * fx + Pm.R[4] * fy + Pm.R[5] * fz + Pm.origin[1];
                const double rz = Pm.R[6] * fx + Pm.R[7] * fy + Pm.R[8] * fz + Pm.origin[2];
                x = wrap(rx + Pm.shift[0], L, half, inv);
                y = wrap(ry + Pm.shift[1], L, half, inv);
                z = wrap(rz + Pm.shift[2], L, half, inv);
            }
            s_new[a] = make_double4(x, y, z, p.w);
            if (fixlds) {
                s_fnew[a] = fixed_pos(x, y, z, p.w, inv);
                s_fold[a] = fixed_pos(p.x, p.y, p.z, p.w, inv);
            }
        }
        __syncthreads();
        const GeneralMove& P = s_P;

        // ---- cutoff_g2g: the moved molecular group's mass centre in both states
        // (GroupCutoff::cut, energy.h:761-768).  A translate+rotate carries it along; atom
        // translations and pivots recompute it from the group's trial positions
        // (Geometry::massCenter as translateParticle, move.cpp:225-240, and
        // rotate_segment, chainmove.cpp:76-104, call it)
        const int gflag = g >= 0 ? gflags[g] : 0;
        const bool cutg = g2g2 != nullptr && m > 0 && (gflag & 1);
        if (cutg) {
            const double* co = (kSplit && g == prev_cg) ? prev_cm : gcm + 3 * g;
            if (P.kind == 1) {
                if (tid == 0)
                    for (int d = 0; d < 3; ++d) {
                        s_cm[0][d] = P.cm_new[d];
                        s_cm[1][d] = co[d];
                    }
            } else {
                // positions relative to the old centre (atom moves) or to the pivot's shift
                // atom, the rotated segment in the unwrapped frame (pivots)
                const double sx = P.kind == 0 ? co[0] : P.shift[0], sy = P.kind == 0 ? co[1] : P.shift[1],
                             sz = P.kind == 0 ? co[2] : P.shift[2];
                double v[4] = {0.0, 0.0, 0.0, 0.0};
                const int gb0 = gbegin[g], gn = gsize[g];
                for (int k = tid; k < gn; k += kThreads) {
                    const int idx = gb0 + k;
                    double fx, fy, fz;
                    if (idx >= P.mlo && idx < P.mhi && P.kind == 2) {
                        const double4 p = s_old[idx - P.mlo];
                        const double ox = wrap(p.x - P.shift[0], L, half, inv) - P.origin[0];
                        const double oy = wrap(p.y - P.shift[1], L, half, inv) - P.origin[1];
                        const double oz = wrap(p.z - P.shift[2], L, half, inv) - P.origin[2];
                        fx = P.R[0] * ox + P.R[1] * oy + P.R[2] * oz + P.origin[0];
                        fy = P.R[3] * ox + P.R[4] * oy + P.R[5] * oz + P.origin[1];
                        fz = P.R[6] * ox + P.R[7] * oy + P.R[8] * oz + P.origin[2];
                    } else {
                        const double4 p = (idx >= P.mlo && idx < P.mhi) ? s_new[idx - P.mlo] : load_pos(idx, s_prev);
                        fx = wrap(p.x - sx, L, half, inv);
                        fy = wrap(p.y - sy, L, half, inv);
                        fz = wrap(p.z - sz, L, half, inv);
                    }
                    const double w = s_mw[meta[idx] & 0xffu];
                    v[0] += w * fx;
                    v[1] += w * fy;
                    v[2] += w * fz;
                    v[3] += w;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    v[r] = wave_sum(v[r]);
                if ((tid & 63) == 0)
#pragma unroll
                    for (int r = 0; r < 4; ++r) s_cmred[r][tid >> 6] = v[r];
                __syncthreads();
                if (tid == 0) {
                    double c[4] = {0.0, 0.0, 0.0, 0.0};
                    for (int w = 0; w < kThreads / 64; ++w)
                        for (int r = 0; r < 4; ++r) c[r] += s_cmred[r][w];
                    s_cm[0][0] = wrap(c[0] / c[3] + sx, L, half, inv);
                    s_cm[0][1] = wrap(c[1] / c[3] + sy, L, half, inv);
                    s_cm[0][2] = wrap(c[2] / c[3] + sz, L, half, inv);
                    for (int d = 0; d < 3; ++d) s_cm[1][d] = co[d];
                }
            }
            __syncthreads();
        }
        const GroupCut gcut{gflags, gcm, g2g2, S.nmol, gflag};

        // ---- surface term: the moved atoms' change of the dipole sum q r (thread 0)
        double dmx = 0.0, dmy = 0.0, dmz = 0.0;
        if (surf != 0.0 && m > 0) {
            double v[3] = {0.0, 0.0, 0.0};
            for (int a = tid; a < m; a += kThreads) {
                const double4 An = s_new[a], Ao = s_old[a];
                v[0] += An.w * (An.x - Ao.x);
                v[1] += An.w * (An.y - Ao.y);
                v[2] += An.w * (An.z - Ao.z);
            }
#pragma unroll
            for (int r = 0; r < 3; ++r)
                v[r] = wave_sum(v[r]);
            if ((tid & 63) == 0)
#pragma unroll
                for (int r = 0; r < 3; ++r) s_cmred[r][tid >> 6] = v[r];
            __syncthreads();
            if (tid == 0)
                for (int w = 0; w < kThreads / 64; ++w) {
                    dmx += s_cmred[0][w];
                    dmy += s_cmred[1][w];
                    dmz += s_cmred[2][w];
                }
        }

        // ---- nonbonded pairs: moved x static (GroupPairing, energy.h:885-1027, 1155-1224)
        double e_new = 0.0, e_old = 0.0;
        const bool same_ok = P.internal && !(gflag & 2); // pairs inside the moved group
        if (m == 1 && !kCells && excl_off == nullptr && !cutg) {
            // one moved atom: the row sums with the moved coordinates in registers
            const double4 A = s_new[0], B = s_old[0];
            const int ta = s_mt[0];
            if (kF32) {
                pair_row_f32<kThreads>(pos32, meta, jb, je, P.mlo, g, same_ok, s_pairs_f + ta * nt, cf, s_stab, terms, hf,
                                       fixed_pos(A.x, A.y, A.z, A.w, inv), fixed_pos(B.x, B.y, B.z, B.w, inv), &s_nhot,
                                       s_hot, 0, e_new, e_old);
            } else {
                pair_row_f64<kThreads>(pos, meta, jb, je, P.mlo, g, same_ok, s_pairs + ta * nt, s_coul, terms, box, A.x,
                                       A.y, A.z, B.x, B.y, B.z, A.w, e_new, e_old);
            }
        } else if (m > 0) {
            if (!kCells) {
                // blocks of kB neighbours per lane (kB from 8 down to 1 as the slice
                // runs out, so that few lanes sit idle on short slices)
                const MovedSet ms{P.mlo,    P.mhi, m,    g,    same_ok, s_new, s_old,   fixlds ? s_fnew : nullptr,
                                  s_fold,   inv,   s_mt, excl_off, excl, gcut,  s_cm[0], s_cm[1]};
                auto blocks = [&](auto excl_tag, auto cut_tag) {
                    constexpr bool kX = decltype(excl_tag)::value;
                    constexpr bool kC = decltype(cut_tag)::value;
                    int base = jb;
                    while (base < je) {
                        const int rem = je - base;
                        if (kF32 && rem >= 8 * kThreads) {
                            moved_block<kThreads, 8, kF32, kX, kC>(base, je, ms, meta, pos32, pos, s_pairs_f, s_pairs, nt, cf,
                                                               s_stab, s_coul, terms, box, hf, &s_nhot, s_hot, e_new, e_old);
                            base += 8 * kThreads;
                        } else if (rem >= 4 * kThreads) {
                            moved_block<kThreads, 4, kF32, kX, kC>(base, je, ms, meta, pos32, pos, s_pairs_f, s_pairs, nt, cf,
                                                               s_stab, s_coul, terms, box, hf, &s_nhot, s_hot, e_new, e_old);
                            base += 4 * kThreads;
                        } else if (rem > kThreads) {
                            moved_block<kThreads, 2, kF32, kX, kC>(base, je, ms, meta, pos32, pos, s_pairs_f, s_pairs, nt, cf,
                                                               s_stab, s_coul, terms, box, hf, &s_nhot, s_hot, e_new, e_old);
                            base += 2 * kThreads;
                        } else {
                            moved_block<kThreads, 1, kF32, kX, kC>(base, je, ms, meta, pos32, pos, s_pairs_f, s_pairs, nt, cf,
                                                               s_stab, s_coul, terms, box, hf, &s_nhot, s_hot, e_new, e_old);
                            base += kThreads;
                        }
                    }
                };
                if (excl_off != nullptr) {
                    if (cutg) blocks(std::true_type{}, std::true_type{});
                    else blocks(std::true_type{}, std::false_type{});
                } else {
                    if (cutg) blocks(std::false_type{}, std::true_type{});
                    else blocks(std::false_type{}, std::false_type{});
                }
            } else {
              // stencils of (2r+1)^3 cells (cells of at least range / r) around each moved
              // atom's new (st=0) and old (st=1) position.  When the pair range plus the
              // moved atoms' largest distance from the first one fits in r cells (a small
              // molecule), the first atom's stencils hold every partner of every moved
              // atom: one stencil per state is walked and each member is paired with all
              // moved atoms (pairs beyond the range add exactly zero)
              const int cr = S.cell_radius, cw = 2 * cr + 1, ns = cw * cw * cw;
              bool mol = false;
              if (m > 1 && m <= kMolStencil) {
                  double ext2 = 0.0;
                  for (int a = 1; a < m; ++a) {
                      ext2 = fmax(ext2, dist2(box, s_new[a].x, s_new[a].y, s_new[a].z, s_new[0].x, s_new[0].y, s_new[0].z));
                      ext2 = fmax(ext2, dist2(box, s_old[a].x, s_old[a].y, s_old[a].z, s_old[0].x, s_old[0].y, s_old[0].z));
                  }
                  mol = S.pair_range + sqrt(ext2) <= static_cast<double>(cr) / S.inv_cell;
              }
              const int na = mol ? 1 : m; // stencil anchors per state
              if (na * 2 * ns <= kMaxComb) {
                // flattened: the members of all (anchor, state, cell) combinations are
                // numbered by a prefix sum and dealt to the lanes in strides, with the
                // loads of several members issued together
                const int ncomb = na * 2 * ns;
                int carry = 0;
                for (int base = 0; base < ncomb; base += kThreads) {
                    const int cb = base + tid;
                    int cnt = 0;
                    if (cb < ncomb) {
                        const int a = cb / (2 * ns), st = (cb / ns) & 1, c = cb % ns;
                        const double4 A = st == 0 ? s_new[a] : s_old[a];
                        const int cx = (cell_of_coord(A.x, half, S.inv_cell, S.nc0) + (c % cw) - cr + S.nc0) % S.nc0;
                        const int cy = (cell_of_coord(A.y, half, S.inv_cell, S.nc1) + ((c / cw) % cw) - cr + S.nc1) % S.nc1;
                        const int cz = (cell_of_coord(A.z, half, S.inv_cell, S.nc2) + (c / (cw * cw)) - cr + S.nc2) % S.nc2;
                        const int cell = (cz * S.nc1 + cy) * S.nc0 + cx;
                        s_ccell[cb] = cell;
                        cnt = S.cell_count[cell];
                    }
                    int tot;
                    const int ex = block_exclusive_scan<kThreads>(cnt, s_wsum, &tot);
                    if (cb < ncomb) s_cpre[cb] = carry + ex;
                    carry += tot;
                }
                if (tid == 0) s_cpre[ncomb] = carry;
                __syncthreads();
                const int W = carry;
                constexpr int kU = 4;
                int c = 0;
                for (int w0 = tid; w0 < W; w0 += kThreads * kU) {
                    int jj[kU], cc[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int w = w0 + u * kThreads;
                        jj[u] = -1;
                        cc[u] = 0;
                        if (w < W) {
                            while (s_cpre[c + 1] <= w) ++c;
                            cc[u] = c;
                            jj[u] = S.cell_members[s_ccell[c] * S.cap + (w - s_cpre[c])];
                        }
                    }
                    uint32_t mj[kU];
                    double4 pj[kU];
                    FixPos pf[kU];
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (jj[u] >= 0) {
                            mj[u] = meta[jj[u]];
                            if (kF32) pf[u] = pos32[jj[u]];
                            else pj[u] = pos[jj[u]];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const int j = jj[u];
                        if (j < 0 || (j >= P.mlo && j < P.mhi)) continue;
                        const int st = (cc[u] / ns) & 1;
                        const int alo = mol ? 0 : cc[u] / (2 * ns), ahi = mol ? m : alo + 1;
                        const bool same = static_cast<int>(mj[u] >> 8) == g;
                        if (same && !same_ok) continue;
                        if (cutg && !same && group_cut(gcut, static_cast<int>(mj[u] >> 8), s_cm[st], box)) continue;
                        const int tj = static_cast<int>(mj[u] & 0xffu);
                        double e = 0.0;
                        for (int a = alo; a < ahi; ++a) {
                            if (same && pair_excluded(excl_off, excl, P.mlo + a, j)) continue;
                            const double4 A = st == 0 ? s_new[a] : s_old[a];
                            const int ta = s_mt[a];
                            if (kF32) {
                                const FixPos& FA = st == 0 ? s_fnew[a] : s_fold[a];
                                bool hot;
                                const float uf = pair_energy_ft(s_pairs_f[ta * nt + tj], cf, s_stab, terms, FA.q,
                                                                pf[u].q, dist2_fix(FA, pf[u], hf), hot);
                                e += (hot && push_hot(&s_nhot, s_hot, j, (a << 1) | st)) ? 0.0 : static_cast<double>(uf);
                            } else {
                                e += pair_energy(s_pairs[ta * nt + tj], s_coul, A.w, pj[u].w,
                                                 dist2(box, A.x, A.y, A.z, pj[u].x, pj[u].y, pj[u].z));
                            }
                        }
                        if (st == 0) e_new += e;
                        else e_old += e;
                    }
                }
              } else {
                // many moved atoms: one (atom, state, cell) combination per lane
                for (int cb = tid; cb < m * 2 * ns; cb += kThreads) {
                    const int a = cb / (2 * ns), st = (cb / ns) & 1, c = cb % ns;
                    const double4 A = st == 0 ? s_new[a] : s_old[a];
                    const int ta = s_mt[a];
                    const int cx = (cell_of_coord(A.x, half, S.inv_cell, S.nc0) + (c % cw) - cr + S.nc0) % S.nc0;
                    const int cy = (cell_of_coord(A.y, half, S.inv_cell, S.nc1) + ((c / cw) % cw) - cr + S.nc1) % S.nc1;
                    const int cz = (cell_of_coord(A.z, half, S.inv_cell, S.nc2) + (c / (cw * cw)) - cr + S.nc2) % S.nc2;
                    const int cell = (cz * S.nc1 + cy) * S.nc0 + cx;
                    const int cnt = S.cell_count[cell];
                    for (int kk = 0; kk < cnt; ++kk) {
                        const int j = S.cell_members[cell * S.cap + kk];
                        if (j >= P.mlo && j < P.mhi) continue;
                        const uint32_t mj = meta[j];
                        if (static_cast<int>(mj >> 8) == g &&
                            (!same_ok || pair_excluded(excl_off, excl, P.mlo + a, j)))
                            continue;
                        if (cutg && static_cast<int>(mj >> 8) != g && group_cut(gcut, static_cast<int>(mj >> 8), s_cm[st], box))
                            continue;
                        const int tj = static_cast<int>(mj & 0xffu);
                        double u;
                        if (kF32) {
                            const FixPos pj = pos32[j];
                            const FixPos& FA = st == 0 ? s_fnew[a] : s_fold[a];
                            bool hot;
                            const float uf = pair_energy_ft(s_pairs_f[ta * nt + tj], cf, s_stab, terms, FA.q, pj.q,
                                                            dist2_fix(FA, pj, hf), hot);
                            u = (hot && push_hot(&s_nhot, s_hot, j, (a << 1) | st)) ? 0.0 : static_cast<double>(uf);
                        } else {
                            const double4 pj = pos[j];
                            u = pair_energy(s_pairs[ta * nt + tj], s_coul, A.w, pj.w, dist2(box, A.x, A.y, A.z, pj.x, pj.y, pj.z));
                        }
                        if (st == 0) e_new += u;
                        else e_old += u;
                    }
                }
              }
            }
            // moved x moved inside the group (internal changes of flexible groups)
            if (same_ok && m > 1) {
                for (int idx = part * kThreads + tid; idx < m * m; idx += G * kThreads) {
                    const int a = idx / m, b = idx % m;
                    if (a >= b || pair_excluded(excl_off, excl, P.mlo + a, P.mlo + b)) continue;
                    const DevPair& p = s_pairs[s_mt[a] * nt + s_mt[b]];
                    const double4 An = s_new[a], Bn = s_new[b], Ao = s_old[a], Bo = s_old[b];
                    e_new += pair_energy(p, s_coul, An.w, Bn.w, dist2(box, An.x, An.y, An.z, Bn.x, Bn.y, Bn.z));
                    e_old += pair_energy(p, s_coul, Ao.w, Bo.w, dist2(box, Ao.x, Ao.y, Ao.z, Bo.x, Bo.y, Bo.z));
                }
            }
        }

        // ---- harmonic bonds with a moved atom (Bonded::internalGroupEnergy, energy.cpp:1068-1088)
        double b_new = 0.0, b_old = 0.0;
        if (P.internal && bond_off != nullptr && part == 0 && (P.kind != 0 || (D.flags & 4))) {
            for (int a = tid; a < m; a += kThreads) {
                const int ia = P.mlo + a;
                for (int e = bond_off[ia]; e < bond_off[ia + 1]; ++e) {
                    const DevBond B = bonds[e];
                    const bool om = B.other >= P.mlo && B.other < P.mhi;
                    if (om && B.other < ia) continue; // visited from the lower moved atom
                    const double4 on = om ? s_new[B.other - P.mlo] : load_pos(B.other, s_prev);
                    const double4 oo = om ? s_old[B.other - P.mlo] : on;
                    const double4 An = s_new[a], Ao = s_old[a];
                    const double xn = B.req - sqrt(dist2(box, An.x, An.y, An.z, on.x, on.y, on.z));
                    const double xo = B.req - sqrt(dist2(box, Ao.x, Ao.y, Ao.z, oo.x, oo.y, oo.z));
                    b_new += B.half_k * xn * xn;
                    b_old += B.half_k * xo * xo;
                }
            }
        }

        // ---- Ewald: dQ(k) = sum_a q_a (e^{ik.r_new} - e^{ik.r_old}) over chunks of moved atoms
        double dqr[kMaxKPT], dqi[kMaxKPT];
#pragma unroll
        for (int s = 0; s < kMaxKPT; ++s) dqr[s] = dqi[s] = 0.0;
        for (int cst = 0; cst < m; cst += kChunk) {
            const int cm_ = min(kChunk, m - cst);
            if (cst > 0) __syncthreads(); // previous chunk's tables consumed
            for (int e = tid; e < cm_ * 6 * width; e += kThreads) {
                const int nn = e % width - ncut;
                const int axis = (e / width) % 3;
                const int st = (e / (3 * width)) & 1;
                const int a = e / (6 * width);
                const double4 A = st == 0 ? s_new[cst + a] : s_old[cst + a];
                const double x = axis == 0 ? A.x : (axis == 1 ? A.y : A.z);
                double sn, cs;
                sincos(6.283185307179586 * (static_cast<double>(nn) / L) * x, &sn, &cs);
                s_phase[e] = make_double2(cs, sn);
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                const int k = k0 + s;
                if (s < kpt && k < ke) {
                    const int packed = s_kn[k];
                    const int nx = unpack_n(packed, 0) + ncut, ny = unpack_n(packed, 1) + ncut,
                              nz = unpack_n(packed, 2) + ncut;
                    for (int a = 0; a < cm_; ++a) {
                        const double2* tb = s_phase + a * 6 * width;
                        const double2 xn = tb[nx], yn = tb[width + ny], zn = tb[2 * width + nz];
                        const double2 xo = tb[3 * width + nx], yo = tb[4 * width + ny], zo = tb[5 * width + nz];
                        const double rn = xn.x * yn.x - xn.y * yn.y, in_ = xn.x * yn.y + xn.y * yn.x;
                        const double ro = xo.x * yo.x - xo.y * yo.y, io = xo.x * yo.y + xo.y * yo.x;
                        const double qa = s_new[cst + a].w;
                        dqr[s] += qa * ((rn * zn.x - in_ * zn.y) - (ro * zo.x - io * zo.y));
                        dqi[s] += qa * ((rn * zn.y + in_ * zn.x) - (ro * zo.y + io * zo.x));
                    }
                }
            }
        }
        if (kF32 && m > 0) {
            // hot pairs in FP64 from the FP64 coordinates (the Ewald chunk loop's barrier
            // follows every push)
            const int nh = min(s_nhot, kHotCap);
            for (int h = tid; h < nh; h += kThreads) {
                const int2 e = s_hot[h];
                const int a = e.y >> 1;
                const bool nw = (e.y & 1) == 0;
                const double4 A = nw ? s_new[a] : s_old[a];
                const double4 pj = load_pos(e.x, s_prev);
                const int tj = static_cast<int>(meta[e.x] & 0xffu);
                const double u = pair_energy_t(s_pairs[s_mt[a] * nt + tj], s_coul, A.w, pj.w,
                                               dist2(box, A.x, A.y, A.z, pj.x, pj.y, pj.z), terms);
                if (nw) e_new += u;
                else e_old += u;
            }
        }
        double de = 0.0;
#pragma unroll
        for (int s = 0; s < kMaxKPT; ++s) {
            const int k = k0 + s;
            if (s < kpt && k < ke) {
                const int packed = s_kn[k];
                const int nx = unpack_n(packed, 0), ny = unpack_n(packed, 1), nz = unpack_n(packed, 2);
                const double A = (nx > 0 ? 2.0 : 1.0) * s_gk[nx * nx + ny * ny + nz * nz];
                de += A * (2.0 * (q[s].x * dqr[s] + q[s].y * dqi[s]) + (dqr[s] * dqr[s] + dqi[s] * dqi[s]));
            }
        }

        // ---- reduction (fixed order) and the decision on lane 0
        {
            double v[5] = {e_new, e_old, de, b_new, b_old};
#pragma unroll
            for (int r = 0; r < 5; ++r)
                v[r] = wave_sum(v[r]);
            if ((tid & 63) == 0)
#pragma unroll
                for (int r = 0; r < 5; ++r) s_red[r][tid >> 6] = v[r];
        }
        __syncthreads();
        double tot[5] = {0.0, 0.0, 0.0, 0.0, 0.0}; // totals, valid in thread 0
        if (kSplit && G > 1) {
            // publish this part's partials, then gather every part's (lane p of wave 0
            // waits for part p) and add them with the same butterfly in every part
            GeneralSlot* slots = static_cast<GeneralSlot*>(S.gslots);
            const int parity = static_cast<int>(mvi & 1ull);
            const unsigned long long tag = mvi + 1ull;
            if (tid == 0) {
                GeneralSlot* sl = slots + parity * G + part;
                for (int r = 0; r < 5; ++r) {
                    double x = 0.0;
                    for (int w = 0; w < kThreads / 64; ++w) x += s_red[r][w];
                    sl->v[r] = x;
                }
                __hip_atomic_store(&sl->seq, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (tid < 64) {
                double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                int ok = 1;
                if (tid < G) {
                    GeneralSlot* sl = slots + parity * G + tid;
                    long long spins = 0;
                    // relaxed polling, then one acquire fence (an acquire load per poll
                    // would invalidate the caches on every iteration)
                    while (__hip_atomic_load(&sl->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != tag) {
                        if (++spins > spin_limit) {
                            ok = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    if (ok)
                        for (int r = 0; r < 5; ++r) v[r] = sl->v[r];
                }
#pragma unroll
                for (int r = 0; r < 5; ++r)
                    v[r] = wave_sum(v[r]);
                const int all_ok = __all(ok);
                if (tid == 0) {
                    for (int r = 0; r < 5; ++r) tot[r] = v[r];
                    if (!all_ok) s_abort = 1; // a part did not arrive: the replica stops with an error
                }
            }
        } else if (tid == 0) {
            for (int w = 0; w < kThreads / 64; ++w)
                for (int r = 0; r < 5; ++r) tot[r] += s_red[r][w];
        }
        if (tid == 0) {
            double du = 0.0;
            bool acc = true;
            if (m > 0) {
                // the surface term enters with the reciprocal one (Ewald::energy, energy.cpp:581-587)
                double dsurf = 0.0;
                if (surf != 0.0) {
                    const double mx = dip0 + dmx, my = dip1 + dmy, mz = dip2 + dmz;
                    dsurf = surf * ((mx * mx + my * my + mz * mz) - (dip0 * dip0 + dip1 * dip1 + dip2 * dip2));
                }
                du = energy_change(tot[0] + tot[3], tot[1] + tot[4], pref * tot[2] + dsurf);
                acc = metropolis(beta * du, P.accept_u);
                if (acc) {
                    dip0 += dmx;
                    dip1 += dmy;
                    dip2 += dmz;
                }
            }
            s_accept = acc ? 1 : 0;
            if (kF32) s_nhot = 0; // every thread has read the queue before the reduction
            if (trace && t < trace_cap && part == 0) {
                TraceEntry& T = trace[static_cast<size_t>(rep) * trace_cap + t];
                T.index = m > 0 ? P.mlo : -1;
                T.accepted = acc ? 1 : 0;
                T.du = du;
                T.x = m > 0 ? s_new[0].x : 0.0;
                T.y = m > 0 ? s_new[0].y : 0.0;
                T.z = m > 0 ? s_new[0].z : 0.0;
            }
            ++accepted;
            if (!acc) --accepted;
            if (acc && m > 0) {
                energy += du;
                sum_du += du;
            }
            if (sample_every > 0 && isfinite(energy) &&
                (mvi + 1ull) % static_cast<unsigned long long>(sample_every) == 0ull) {
                e_sum += energy;
                e_sum2 += energy * energy;
                ++e_samples;
            }
            if (acc && m > 0) {
                // every part writes the (identical) centre, so each part's later reads see
                // its own write; reads of this move's old centre precede every part's commit
                if (cutg && P.kind != 1)
                    for (int d = 0; d < 3; ++d) gcm[3 * g + d] = s_cm[0][d];
                else if (P.kind == 1 && P.translate)
                    for (int d = 0; d < 3; ++d) gcm[3 * g + d] = P.cm_new[d];
                if (kCells) {
                    auto cell_at = [&](const double4& A) {
                        return (cell_of_coord(A.z, half, S.inv_cell, S.nc2) * S.nc1 +
                                cell_of_coord(A.y, half, S.inv_cell, S.nc1)) * S.nc0 +
                               cell_of_coord(A.x, half, S.inv_cell, S.nc0);
                    };
                    // every cell that gains atoms must hold them all (counting the atoms of
                    // this move that enter it); otherwise the lists are left as they are and
                    // the chain stops after this move
                    bool fits = true;
                    for (int a = 0; a < m && fits; ++a) {
                        const int cnew = cell_at(s_new[a]);
                        if (cnew == S.cell_of[P.mlo + a]) continue;
                        if (S.cell_count[cnew] + m <= S.cap) continue;
                        int entering = 0;
                        for (int b = 0; b < m; ++b)
                            if (cell_at(s_new[b]) == cnew && S.cell_of[P.mlo + b] != cnew) ++entering;
                        if (S.cell_count[cnew] + entering > S.cap) fits = false;
                    }
                    if (!fits) S.error = 1;
                    for (int a = 0; a < m && fits; ++a) {
                        const int i = P.mlo + a;
                        const int cnew = cell_at(s_new[a]);
                        const int cold = S.cell_of[i];
                        if (cnew == cold) continue;
                        const int nold = S.cell_count[cold];
                        for (int kk = 0; kk < nold; ++kk)
                            if (S.cell_members[cold * S.cap + kk] == i) {
                                S.cell_members[cold * S.cap + kk] = S.cell_members[cold * S.cap + nold - 1];
                                break;
                            }
                        S.cell_count[cold] = nold - 1;
                        const int nn = S.cell_count[cnew];
                        S.cell_members[cnew * S.cap + nn] = i;
                        S.cell_count[cnew] = nn + 1;
                        S.cell_of[i] = cnew;
                    }
                }
            }
        }
        __syncthreads();
        if (kSplit && s_abort) {
            if (tid == 0) atomicExch(&S.error, 2);
            t_done = t;
            break;
        }
        const bool committed = s_accept && m > 0;
        if (committed) {
            // each particle is written by the part whose pair slice holds it (all of them
            // without split); the next exchange orders these writes before later reads
            for (int a = tid; a < m; a += kThreads) {
                const int idx = P.mlo + a;
                if (kSplit && (idx < jb || idx >= je)) continue;
                const double4 A = s_new[a];
                pos[idx] = A;
                if (kF32) pos32[idx] = fixlds ? s_fnew[a] : fixed_pos(A.x, A.y, A.z, A.w, inv);
            }
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                q[s].x += dqr[s];
                q[s].y += dqi[s];
            }
        }
        if (kSplit) {
            prev_lo = committed ? P.mlo : 0;
            prev_hi = committed ? P.mhi : 0;
            prev_cg = (committed && P.kind == 1 && P.translate) ? g : -1;
            prev_cm[0] = P.cm_new[0];
            prev_cm[1] = P.cm_new[1];
            prev_cm[2] = P.cm_new[2];
        }
        if (!kSplit && S.error) {
            t_done = t + 1; // stopped after committing this move (cell capacity)
            break;
        }
        __syncthreads(); // moved coordinates and tables consumed before the next move
    }

#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s)
        if (s < kpt && k0 + s < ke) S.q[k0 + s] = q[s];
    if (tid == 0 && part == 0) {
        S.moves = move0 + static_cast<unsigned long long>(t_done);
        S.accepted = accepted;
        S.energy = energy;
        S.sum_du = sum_du;
        S.dip[0] = dip0;
        S.dip[1] = dip1;
        S.dip[2] = dip2;
        S.e_samples = e_samples;
        S.e_sum = e_sum;
        S.e_sum2 = e_sum2;
    }
}

// Workgroup size: 512 threads, two waves per SIMD (the second wave hides the latency
// of the first; some cold paths spill to scratch), up to 8 k-vectors per thread.
constexpr int kWide = 512;

int threads_of(bool) { return kWide; }
int kpt_cap(bool) { return 8; } // 4096 k-vectors per workgroup

template <int kThreads, int kMaxKPT, bool kCells, bool kF32, bool kSplit>
const void* kernel_of() {
    return reinterpret_cast<const void*>(mc_general_kernel<kThreads, kMaxKPT, kCells, kF32, kSplit>);
}

const void* select_kernel(int maxkpt, bool cells, bool f32, bool split) {
    if (cells) {
        if (maxkpt <= 4) return f32 ? kernel_of<kWide, 4, true, true, false>() : kernel_of<kWide, 4, true, false, false>();
        return f32 ? kernel_of<kWide, 8, true, true, false>() : kernel_of<kWide, 8, true, false, false>();
    }
    if (maxkpt <= 4) {
        if (split) return f32 ? kernel_of<kWide, 4, false, true, true>() : kernel_of<kWide, 4, false, false, true>();
        return f32 ? kernel_of<kWide, 4, false, true, false>() : kernel_of<kWide, 4, false, false, false>();
    }
    if (split) return f32 ? kernel_of<kWide, 8, false, true, true>() : kernel_of<kWide, 8, false, false, true>();
    return f32 ? kernel_of<kWide, 8, false, true, false>() : kernel_of<kWide, 8, false, false, false>();
}

constexpr size_t kLdsPerCU = 160 * 1024;

// Dynamic LDS a kernel may take: what its static LDS leaves of the CU's 160 KB, at
// most 96 KB (set once per kernel).
size_t prepare(const void* fn) {
    static std::vector<std::pair<const void*, size_t>> done;
    for (const auto& d : done)
        if (d.first == fn) return d.second;
    hipFuncAttributes attr;
    check_hip(hipFuncGetAttributes(&attr, fn), "hipFuncGetAttributes");
    const size_t limit = std::min<size_t>(96 * 1024, kLdsPerCU - attr.sharedSizeBytes);
    check_hip(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(limit)),
              "hipFuncSetAttribute");
    done.emplace_back(fn, limit);
    return limit;
}

// k-vectors per thread of a replica's part (its slice of ceil(nk / parts) k-vectors).
int part_kpt(const ReplicaState& s, int parts) {
    parts = parts > 1 ? parts : 1;
    const int kw = (s.nk + parts - 1) / parts;
    const int threads = threads_of(s.cells != 0);
    return (kw + threads - 1) / threads;
}

struct GeneralLaunch {
    int maxkpt;
    bool cells, split;
    size_t lds;
    int grid;
    size_t fix_lds_bytes; // FP32 form: fixed-point copies of the moved atoms
};

GeneralLaunch plan(const ReplicaState* h_states, int count, bool f32) {
    GeneralLaunch g{0, h_states[0].cells != 0, false, 0, 0};
    int maxnk = 0, maxw = 0, maxm = 0;
    for (int r = 0; r < count; ++r) {
        const ReplicaState& s = h_states[r];
        if (s.nkinds < 1 || s.nkinds > kMaxKinds) throw std::runtime_error("general move set: 1 to 8 move kinds");
        if (s.cells != h_states[0].cells) throw std::runtime_error("replicas of one launch must agree on the cell list");
        if (s.ntypes > kMaxTypes) throw std::runtime_error("too many atom types");
        if (s.parts > 1) {
            if (s.parts > kMaxParts) throw std::runtime_error("split general launch: at most 64 workgroups per replica");
            if (s.cells) throw std::runtime_error("split general launch: replicas without a cell list only");
            if (s.gslots == nullptr) throw std::runtime_error("split general launch: exchange slots missing");
            g.split = true;
        }
        g.grid += s.parts > 1 ? s.parts : 1;
        g.maxkpt = std::max(g.maxkpt, part_kpt(s, s.parts));
        maxnk = std::max(maxnk, s.nk);
        maxw = std::max(maxw, 2 * s.ncut + 1);
        maxm = std::max(maxm, s.max_moved);
    }
    if (g.maxkpt > kpt_cap(g.cells))
        throw std::runtime_error("too many k-vectors for the general device MC (max 4096 per workgroup)");
    // the split form keeps the previous move's coordinates in a second buffer
    g.lds = sizeof(int) * static_cast<size_t>((maxnk + 7) & ~7) + sizeof(double2) * kChunk * 6 * static_cast<size_t>(maxw) +
            ((g.split ? 3 : 2) * sizeof(double4) + sizeof(int)) * static_cast<size_t>(maxm);
    g.fix_lds_bytes = 2 * sizeof(FixPos) * static_cast<size_t>(maxm);
    return g;
}

} // namespace

int general_coresident_blocks(bool f32) {
    const void* fn = select_kernel(8, false, f32, true);
    const size_t lds = prepare(fn);
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWide, lds), "occupancy");
    return per_cu * prop.multiProcessorCount;
}

void set_replica_parts(ReplicaState& s, int parts, std::vector<void*>& owned) {
    if (parts < 1 || parts > kMaxParts) throw std::runtime_error("split general launch: 1 to 64 workgroups per replica");
    if (parts > 1 && s.cells) throw std::runtime_error("split general launch: replicas without a cell list only");
    if (parts > 1 && part_kpt(s, parts) > kpt_cap(false))
        throw std::runtime_error("too many k-vectors per workgroup for the split general launch");
    s.parts = parts;
    s.gslots = nullptr;
    if (parts == 1) return;
    void* p = nullptr;
    const size_t bytes = sizeof(GeneralSlot) * 2 * static_cast<size_t>(parts);
    check_hip(hipMalloc(&p, bytes), "hipMalloc exchange slots");
    owned.push_back(p);
    // tags are move indices + 1, so zeroed slots never match a move
    check_hip(hipMemset(p, 0, bytes), "clear exchange slots");
    s.gslots = p;
}

int general_blockmap(const ReplicaState* h_states, int count, std::vector<int2>& map, int stride) {
    if (stride < 1) throw std::runtime_error("block map: stride >= 1");
    map.clear();
    for (int r = 0; r < count; ++r) {
        const int parts = h_states[r].parts > 1 ? h_states[r].parts : 1;
        for (int p = 0; p < parts; ++p) {
            map.push_back(make_int2(r, p));
            for (int i = 1; i < stride; ++i) map.push_back(make_int2(-1, -1));
        }
    }
    return static_cast<int>(map.size());
}

namespace {
// Zero the exchange slots of every split replica of a launch before it starts: a slot
// left by an earlier launch carries a tag that a resumed (rewound) chain would take for
// its own.  Block r clears replica r's 2 x parts slots.
__global__ void clear_general_slots_kernel(ReplicaState* __restrict__ states, int count) {
    const int r = static_cast<int>(blockIdx.x);
    if (r >= count) return;
    const ReplicaState& s = states[r];
    if (s.parts <= 1 || s.gslots == nullptr) return;
    unsigned long long* w = static_cast<unsigned long long*>(s.gslots);
    const int words = static_cast<int>(sizeof(GeneralSlot) / 8) * 2 * s.parts;
    for (int i = static_cast<int>(threadIdx.x); i < words; i += static_cast<int>(blockDim.x)) w[i] = 0ull;
}
} // namespace

void launch_general(const ReplicaState* h_states, int count, ReplicaState* d_states, const long long* d_moves, bool f32,
                    TraceEntry* trace, int trace_cap, hipStream_t stream, const int2* d_blockmap, int map_blocks) {
    GeneralLaunch g = plan(h_states, count, f32);
    const void* fn = select_kernel(g.maxkpt, g.cells, f32, g.split);
    const size_t limit = prepare(fn);
    // the FP32 form keeps fixed-point copies of the moved atoms in LDS where they fit
    int fix_lds = f32 && g.lds + g.fix_lds_bytes <= limit ? 1 : 0;
    if (fix_lds) g.lds += g.fix_lds_bytes;
    if (f32 && g.cells && !fix_lds) throw std::runtime_error("molecules too large for the FP32 cell-list kernel's LDS");
    if (g.lds > limit) throw std::runtime_error("k-vector and molecule tables too large for LDS");
    // each workgroup lays out its own replica's tables inside the dynamic LDS
    long long spin = 1LL << 24;
    if (!g.split) {
        const int2* none = nullptr;
        void* args[] = {&d_states, &d_moves, &trace, &trace_cap, &none, &spin, &fix_lds};
        check_hip(hipLaunchKernel(fn, dim3(count), dim3(threads_of(g.cells)), args, g.lds, stream),
                  "mc_general_kernel launch");
        return;
    }
    if (d_blockmap == nullptr) throw std::runtime_error("split general launch: block map missing");
    const int grid = map_blocks > 0 ? map_blocks : g.grid;
    if (grid < g.grid) throw std::runtime_error("split general launch: block map shorter than the parts");
    if (grid > general_coresident_blocks(f32)) throw std::runtime_error("split general launch: the grid cannot be co-resident");
    static_assert(sizeof(GeneralSlot) % 8 == 0, "whole words");
    hipLaunchKernelGGL(clear_general_slots_kernel, dim3(count), dim3(64), 0, stream, d_states, count);
    check_hip(hipGetLastError(), "clear_general_slots_kernel");
    void* args[] = {&d_states, &d_moves, &trace, &trace_cap, &d_blockmap, &spin, &fix_lds};
    check_hip(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kWide), args, static_cast<unsigned>(g.lds), stream),
              "mc_general_kernel cooperative launch");
}

} // namespace fmc
