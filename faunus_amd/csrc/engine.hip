// GPU energy engine: kernels and host orchestration for the plain path
// (one trial move per launch) — the drop-in for the reference's
// Nonbonded::energy(Change) (energy.h:1561-1577) and Ewald::updateState/energy
// (energy.cpp:564-587).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

#include "device_math.hpp"
#include "engine.hpp"

namespace fmc {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

struct GroupTable {
    const int* flags;     // bit0 molecular (not atomic), bit1 rigid, molecule type from bit 8
    const double* cm;     // 3 per group (accepted state)
    const double* g2g2;   // squared cutoff per pair of molecule types (null: none)
    int nmol;
    const int* excl_off;  // per-particle excluded partners (null: none)
    const int* excl;
};

struct CellArgs {
    const int* count;
    const int* members;
    int nc0, nc1, nc2, cap;
    double inv0, inv1, inv2;
};

__device__ __forceinline__ int cell_coord(double x, double half, double inv_cell, int nc) {
    int c = static_cast<int>(floor((x + half) * inv_cell));
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

// ------------------------------------------------------------------------------
// Fused per-move kernel.  Blocks [0, pair_blocks) sum the pair energies of the
// affected pairs in the new and old state; blocks [pair_blocks, grid) update the
// trial structure factors Q_trial(k) = Q_acc(k) + dQ(k) and sum A_k|Q|^2 for both
// states.  The last block to finish reduces all block partials in a fixed order
// and writes the results to host-visible memory.
// ------------------------------------------------------------------------------
template <bool kGroups, bool kCells>
__global__ __launch_bounds__(kBlock) void delta_kernel(
    const double4* __restrict__ pos, const uint32_t* __restrict__ meta, int n, const DevPair* __restrict__ pairs_g,
    int ntypes, const DevCoulomb* __restrict__ coul_g, Box box, GroupTable groups, CellArgs cells, MoveArgs mv,
    int pair_blocks, const int* __restrict__ kn, const double* __restrict__ ak, const double2* __restrict__ q_acc,
    double2* __restrict__ q_trial, int nk, int ncut, double2* __restrict__ partials, unsigned* __restrict__ ticket,
    DeltaResult* __restrict__ out, unsigned long long seq, double rec_prefactor, double cm_new_x, double cm_new_y,
    double cm_new_z, double cm_old_x, double cm_old_y, double cm_old_z, int ewald_end,
    const int* __restrict__ bond_off, const DevBond* __restrict__ bonds) {
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevCoulomb s_coul;
    __shared__ MovedParticle s_moved[kPhaseTableMoved];
    __shared__ double s_red[2][kBlock / 64];
    __shared__ bool s_last;
    extern __shared__ double2 s_phase[]; // Ewald phase tables

    const int tid = threadIdx.x;
    const int m = mv.m;
    const MovedParticle* moved = (m <= kMaxInlineMoved) ? mv.inl : mv.ext;
    const int m_lds = m < kPhaseTableMoved ? m : kPhaseTableMoved;
    for (int i = tid; i < m_lds; i += kBlock) s_moved[i] = moved[i];
    const MovedParticle* mp = (m <= kPhaseTableMoved) ? s_moved : moved;

    double acc_new = 0.0, acc_old = 0.0;
    const bool is_pair_block = static_cast<int>(blockIdx.x) < pair_blocks;
    if (is_pair_block) {
        for (int i = tid; i < ntypes * ntypes; i += kBlock) s_pairs[i] = pairs_g[i];
        lds_copy(s_coul, coul_g, tid, kBlock);
        __syncthreads();
        const bool single = mv.change_group >= 0;
        auto allowed_other = [&](int gi, int gj, bool new_state) -> bool {
            // pair between a moved particle in group gi and a static particle in group gj
            if (gi == gj) return single && mv.internal && (!kGroups || !(groups.flags[gi] & 2));
            if (!kGroups) return true;
            if (groups.g2g2 && (groups.flags[gi] & 1) && (groups.flags[gj] & 1)) { // GroupCutoff::cut (energy.h:761-768)
                const double* cj = groups.cm + 3 * gj;
                const double r2 = new_state ? dist2(box, cm_new_x, cm_new_y, cm_new_z, cj[0], cj[1], cj[2])
                                            : dist2(box, cm_old_x, cm_old_y, cm_old_z, cj[0], cj[1], cj[2]);
                return r2 < g2g_cutoff2(groups.g2g2, groups.nmol, groups.flags[gi], groups.flags[gj]);
            }
            return true;
        };
        auto visit = [&](int j, const double4 pj, uint32_t mj) {
            const int tj = static_cast<int>(mj & 0xffu);
            const int gj = static_cast<int>(mj >> 8);
            // is j one of the moved particles (sorted by index)?
            int jm = -1;
            {
                int lo = 0, hi = m;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (mp[mid].idx < j) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < m && mp[lo].idx == j) jm = lo;
            }
            if (jm < 0) {
                for (int a = 0; a < m; ++a) {
                    const MovedParticle& M = mp[a];
                    if (kGroups && M.gid == gj && pair_excluded(groups.excl_off, groups.excl, M.idx, j)) continue;
                    const DevPair& p = s_pairs[M.type * ntypes + tj];
                    if (allowed_other(M.gid, gj, true))
                        acc_new += pair_energy(p, s_coul, M.newp.w, pj.w,
                                               dist2(box, M.newp.x, M.newp.y, M.newp.z, pj.x, pj.y, pj.z));
                    if (allowed_other(M.gid, gj, false))
                        acc_old += pair_energy(p, s_coul, M.oldp.w, pj.w,
                                               dist2(box, M.oldp.x, M.oldp.y, M.oldp.z, pj.x, pj.y, pj.z));
                }
            } else {
                // moved <-> moved pairs (a < jm), both particles at their state's coordinates
                const MovedParticle& B = mp[jm];
                for (int a = 0; a < jm; ++a) {
                    const MovedParticle& A = mp[a];
                    bool ok;
                    if (A.gid == B.gid)
                        ok = single && mv.internal && (!kGroups || (!(groups.flags[A.gid] & 2) &&
                                                                    !pair_excluded(groups.excl_off, groups.excl, A.idx, B.idx)));
                    else ok = true; // several moved groups: atomic (molecular multi-group changes are rejected on the host)
                    if (!ok) continue;
                    const DevPair& p = s_pairs[A.type * ntypes + B.type];
                    acc_new += pair_energy(p, s_coul, A.newp.w, B.newp.w,
                                           dist2(box, A.newp.x, A.newp.y, A.newp.z, B.newp.x, B.newp.y, B.newp.z));
                    acc_old += pair_energy(p, s_coul, A.oldp.w, B.oldp.w,
                                           dist2(box, A.oldp.x, A.oldp.y, A.oldp.z, B.oldp.x, B.oldp.y, B.oldp.z));
                }
            }
        };
        if (!kCells) {
            for (int j = blockIdx.x * kBlock + tid; j < n; j += pair_blocks * kBlock) visit(j, pos[j], meta[j]);
        } else {
            // Stencils of 27 cells around each moved particle's new and old position;
            // candidates are evaluated only for the state whose stencil holds them.
            const int stencil_jobs = m * 2 * 27;
            for (int job = blockIdx.x; job < stencil_jobs; job += pair_blocks) {
                const int a = job / 54;
                const int s = (job / 27) & 1; // 0: new, 1: old
                const int c = job % 27;
                const MovedParticle& M = mp[a];
                const double4 P = s == 0 ? M.newp : M.oldp;
                const int cx = (cell_coord(P.x, box.half[0], cells.inv0, cells.nc0) + (c % 3) - 1 + cells.nc0) % cells.nc0;
                const int cy = (cell_coord(P.y, box.half[1], cells.inv1, cells.nc1) + ((c / 3) % 3) - 1 + cells.nc1) % cells.nc1;
                const int cz = (cell_coord(P.z, box.half[2], cells.inv2, cells.nc2) + (c / 9) - 1 + cells.nc2) % cells.nc2;
                const int cell = (cz * cells.nc1 + cy) * cells.nc0 + cx;
                const int cnt = cells.count[cell];
                for (int k = tid; k < cnt; k += kBlock) {
                    const int j = cells.members[cell * cells.cap + k];
                    const double4 pj = pos[j];
                    const uint32_t mj = meta[j];
                    int jm = -1;
                    for (int b = 0; b < m; ++b)
                        if (mp[b].idx == j) jm = b;
                    if (jm >= 0) continue; // moved <-> moved handled below
                    const int tj = static_cast<int>(mj & 0xffu);
                    const int gj = static_cast<int>(mj >> 8);
                    if (!allowed_other(M.gid, gj, s == 0)) continue;
                    if (kGroups && M.gid == gj && pair_excluded(groups.excl_off, groups.excl, M.idx, j)) continue;
                    const DevPair& p = s_pairs[M.type * ntypes + tj];
                    const double u = pair_energy(p, s_coul, P.w, pj.w, dist2(box, P.x, P.y, P.z, pj.x, pj.y, pj.z));
                    if (s == 0) acc_new += u;
                    else acc_old += u;
                }
            }
            if (blockIdx.x == 0) {
                for (int pairidx = tid; pairidx < m * m; pairidx += kBlock) {
                    const int a = pairidx / m, b = pairidx % m;
                    if (a >= b) continue;
                    const MovedParticle& A = mp[a];
                    const MovedParticle& B = mp[b];
                    bool ok = (A.gid == B.gid)
                                  ? (single && mv.internal &&
                                     (!kGroups || (!(groups.flags[A.gid] & 2) &&
                                                   !pair_excluded(groups.excl_off, groups.excl, A.idx, B.idx))))
                                  : true;
                    if (!ok) continue;
                    const DevPair& p = s_pairs[A.type * ntypes + B.type];
                    acc_new += pair_energy(p, s_coul, A.newp.w, B.newp.w,
                                           dist2(box, A.newp.x, A.newp.y, A.newp.z, B.newp.x, B.newp.y, B.newp.z));
                    acc_old += pair_energy(p, s_coul, A.oldp.w, B.oldp.w,
                                           dist2(box, A.oldp.x, A.oldp.y, A.oldp.z, B.oldp.x, B.oldp.y, B.oldp.z));
                }
            }
        }
    } else if (static_cast<int>(blockIdx.x) >= ewald_end) {
        // Bonded::internalGroupEnergy (energy.cpp:1068-1088): bonds with at least one
        // moved atom, each counted once, in both states; only for internal changes.
        __syncthreads();
        if (mv.internal) {
            for (int a = tid; a < m; a += kBlock) {
                const MovedParticle& A = mp[a];
                for (int e = bond_off[A.idx]; e < bond_off[A.idx + 1]; ++e) {
                    const DevBond B = bonds[e];
                    int om = -1;
                    {
                        int lo = 0, hi = m;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (mp[mid].idx < B.other) lo = mid + 1;
                            else hi = mid;
                        }
                        if (lo < m && mp[lo].idx == B.other) om = lo;
                    }
                    if (om >= 0 && om < a) continue; // the pair is visited from the lower moved atom
                    const double4 on = om >= 0 ? mp[om].newp : pos[B.other];
                    const double4 oo = om >= 0 ? mp[om].oldp : on;
                    const double xn = B.req - sqrt(dist2(box, A.newp.x, A.newp.y, A.newp.z, on.x, on.y, on.z));
                    const double xo = B.req - sqrt(dist2(box, A.oldp.x, A.oldp.y, A.oldp.z, oo.x, oo.y, oo.z));
                    acc_new += B.half_k * xn * xn;
                    acc_old += B.half_k * xo * xo;
                }
            }
        }
    } else if (nk > 0) {
        __syncthreads();
        // Ewald partial update (energy.cpp:219-247): dQ(k) = sum_i q_i (e^{ik.r_new} - e^{ik.r_old}).
        const int eb = blockIdx.x - pair_blocks;
        const int e_blocks = ewald_end - pair_blocks;
        const int width = 2 * ncut + 1;
        const double twopi = 6.283185307179586;
        if (m <= kPhaseTableMoved) {
            // per moved particle, state and axis: e^{i k_a(n) r_a} for n in [-ncut, ncut]
            const int entries = m * 2 * 3 * width;
            for (int e = tid; e < entries; e += kBlock) {
                const int nn = e % width - ncut;
                const int axis = (e / width) % 3;
                const int st = (e / (3 * width)) & 1;
                const int a = e / (6 * width);
                const double4 P = st == 0 ? mp[a].newp : mp[a].oldp;
                const double x = axis == 0 ? P.x : (axis == 1 ? P.y : P.z);
                const double kk = twopi * (static_cast<double>(nn) / box.L[axis]);
                double sn, cs;
                sincos(kk * x, &sn, &cs);
                s_phase[e] = make_double2(cs, sn);
            }
            __syncthreads();
            for (int k = eb * kBlock + tid; k < nk; k += e_blocks * kBlock) {
                const int packed = kn[k];
                const int nx = unpack_n(packed, 0) + ncut, ny = unpack_n(packed, 1) + ncut,
                          nz = unpack_n(packed, 2) + ncut;
                double dre = 0.0, dim = 0.0;
                for (int a = 0; a < m; ++a) {
                    for (int st = 0; st < 2; ++st) {
                        const double2* t = s_phase + (a * 2 + st) * 3 * width;
                        const double2 ex = t[nx], ey = t[width + ny], ez = t[2 * width + nz];
                        const double re1 = ex.x * ey.x - ex.y * ey.y;
                        const double im1 = ex.x * ey.y + ex.y * ey.x;
                        const double re = re1 * ez.x - im1 * ez.y;
                        const double im = re1 * ez.y + im1 * ez.x;
                        const double q = st == 0 ? mp[a].newp.w : -mp[a].oldp.w;
                        dre += q * re;
                        dim += q * im;
                    }
                }
                const double2 qa = q_acc[k];
                const double2 qt = make_double2(qa.x + dre, qa.y + dim);
                q_trial[k] = qt;
                const double A = ak[k];
                acc_new += A * (qt.x * qt.x + qt.y * qt.y);
                acc_old += A * (qa.x * qa.x + qa.y * qa.y);
            }
        } else {
            for (int k = eb * kBlock + tid; k < nk; k += e_blocks * kBlock) {
                const int packed = kn[k];
                const double kx = twopi * (static_cast<double>(unpack_n(packed, 0)) / box.L[0]);
                const double ky = twopi * (static_cast<double>(unpack_n(packed, 1)) / box.L[1]);
                const double kz = twopi * (static_cast<double>(unpack_n(packed, 2)) / box.L[2]);
                double dre = 0.0, dim = 0.0;
                for (int a = 0; a < m; ++a) {
                    const MovedParticle& M = mp[a];
                    double sn, cs;
                    sincos(kx * M.newp.x + ky * M.newp.y + kz * M.newp.z, &sn, &cs);
                    dre += M.newp.w * cs;
                    dim += M.newp.w * sn;
                    sincos(kx * M.oldp.x + ky * M.oldp.y + kz * M.oldp.z, &sn, &cs);
                    dre -= M.oldp.w * cs;
                    dim -= M.oldp.w * sn;
                }
                const double2 qa = q_acc[k];
                const double2 qt = make_double2(qa.x + dre, qa.y + dim);
                q_trial[k] = qt;
                const double A = ak[k];
                acc_new += A * (qt.x * qt.x + qt.y * qt.y);
                acc_old += A * (qa.x * qa.x + qa.y * qa.y);
            }
        }
    } else {
        __syncthreads();
    }

    // block partials
    const double bn = block_sum<kBlock>(acc_new, s_red[0]);
    const double bo = block_sum<kBlock>(acc_old, s_red[1]);
    if (tid == 0) {
        partials[blockIdx.x] = make_double2(bn, bo);
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    // last block: fixed-order reduction of the pair partials and of the Ewald partials
    double pn = 0.0, po = 0.0, en = 0.0, eo = 0.0, bn_ = 0.0, bo_ = 0.0;
    if (tid == 0) {
        for (int b = 0; b < static_cast<int>(gridDim.x); ++b) {
            const double* pb = reinterpret_cast<const double*>(partials) + 2 * b;
            const double2 v = make_double2(__hip_atomic_load(pb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                                           __hip_atomic_load(pb + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (b < pair_blocks) {
                pn += v.x;
                po += v.y;
            } else if (b < ewald_end) {
                en += v.x;
                eo += v.y;
            } else {
                bn_ += v.x;
                bo_ += v.y;
            }
        }
        out->pair_new = pn;
        out->pair_old = po;
        out->rec_new = rec_prefactor * en;
        out->rec_old = rec_prefactor * eo;
        out->bond_new = bn_;
        out->bond_old = bo_;
        __threadfence_system();
        out->seq = seq;
        __threadfence_system();
        *ticket = 0u;
    }
}

// Accept: write the moved particles' trial coordinates into the accepted state and
// move them between cells of the cell list (one thread, deterministic order).
__global__ void commit_kernel(double4* pos, MoveArgs mv, CellArgs cells, int* cell_of, int* overflow, Box box,
                              int use_cells) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const MovedParticle* moved = (mv.m <= kMaxInlineMoved) ? mv.inl : mv.ext;
    for (int a = 0; a < mv.m; ++a) {
        const MovedParticle& M = moved[a];
        pos[M.idx] = M.newp;
        if (!use_cells) continue;
        const int cx = cell_coord(M.newp.x, box.half[0], cells.inv0, cells.nc0);
        const int cy = cell_coord(M.newp.y, box.half[1], cells.inv1, cells.nc1);
        const int cz = cell_coord(M.newp.z, box.half[2], cells.inv2, cells.nc2);
        const int cnew = (cz * cells.nc1 + cy) * cells.nc0 + cx;
        const int cold = cell_of[M.idx];
        if (cnew == cold) continue;
        int* cnt = const_cast<int*>(cells.count);
        int* mem = const_cast<int*>(cells.members);
        const int nold = cnt[cold];
        for (int k = 0; k < nold; ++k)
            if (mem[cold * cells.cap + k] == M.idx) {
                mem[cold * cells.cap + k] = mem[cold * cells.cap + nold - 1];
                break;
            }
        cnt[cold] = nold - 1;
        const int nn = cnt[cnew];
        if (nn >= cells.cap) {
            *overflow = 1;
            continue;
        }
        mem[cnew * cells.cap + nn] = M.idx;
        cnt[cnew] = nn + 1;
        cell_of[M.idx] = cnew;
    }
}

// Full structure factors Q(k) = sum_j q_j e^{ik.r_j} (energy.cpp:191-206); particle
// order as in the reference (groups in order, particles in order).
__global__ __launch_bounds__(kBlock) void full_q_kernel(const double4* __restrict__ pos, int n,
                                                         const int* __restrict__ kn, int nk, Box box,
                                                         double2* __restrict__ q_out) {
    __shared__ double4 tile[kBlock];
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const double twopi = 6.283185307179586;
    double kx = 0, ky = 0, kz = 0;
    if (k < nk) {
        const int packed = kn[k];
        kx = twopi * (static_cast<double>(unpack_n(packed, 0)) / box.L[0]);
        ky = twopi * (static_cast<double>(unpack_n(packed, 1)) / box.L[1]);
        kz = twopi * (static_cast<double>(unpack_n(packed, 2)) / box.L[2]);
    }
    double re = 0.0, im = 0.0;
    for (int base = 0; base < n; base += kBlock) {
        __syncthreads();
        if (base + threadIdx.x < static_cast<unsigned>(n)) tile[threadIdx.x] = pos[base + threadIdx.x];
        __syncthreads();
        const int cnt = min(kBlock, n - base);
        if (k < nk) {
            for (int j = 0; j < cnt; ++j) {
                const double4 p = tile[j];
                double sn, cs;
                sincos(kx * p.x + ky * p.y + kz * p.z, &sn, &cs);
                re += p.w * cs;
                im += p.w * sn;
            }
        }
    }
    if (k < nk) q_out[k] = make_double2(re, im);
}

// All pairs i<j (Change::everything, energy.h:1290-1299): one thread per i, j>i.
template <bool kGroups>
__global__ __launch_bounds__(kBlock) void all_pairs_kernel(const double4* __restrict__ pos,
                                                            const uint32_t* __restrict__ meta, int n,
                                                            const DevPair* __restrict__ pairs_g, int ntypes,
                                                            const DevCoulomb* __restrict__ coul_g, Box box,
                                                            GroupTable groups, double* __restrict__ partials) {
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevCoulomb s_coul;
    __shared__ double4 tile[kBlock];
    __shared__ uint32_t tmeta[kBlock];
    __shared__ double s_red[kBlock / 64];
    for (int i = threadIdx.x; i < ntypes * ntypes; i += kBlock) s_pairs[i] = pairs_g[i];
    lds_copy(s_coul, coul_g, static_cast<int>(threadIdx.x), kBlock);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double4 pi = make_double4(0, 0, 0, 0);
    uint32_t mi = 0;
    if (i < n) {
        pi = pos[i];
        mi = meta[i];
    }
    const int ti = static_cast<int>(mi & 0xffu), gi = static_cast<int>(mi >> 8);
    double acc = 0.0;
    for (int base = blockIdx.x * kBlock; base < n; base += kBlock) {
        __syncthreads();
        if (base + threadIdx.x < static_cast<unsigned>(n)) {
            tile[threadIdx.x] = pos[base + threadIdx.x];
            tmeta[threadIdx.x] = meta[base + threadIdx.x];
        }
        __syncthreads();
        const int cnt = min(kBlock, n - base);
        if (i < n) {
            for (int jj = 0; jj < cnt; ++jj) {
                const int j = base + jj;
                if (j <= i) continue;
                const uint32_t mj = tmeta[jj];
                const int gj = static_cast<int>(mj >> 8);
                if (kGroups) {
                    if (gi == gj) {
                        if (groups.flags[gi] & 2) continue; // rigid: no internal pairs
                        if (pair_excluded(groups.excl_off, groups.excl, i, j)) continue;
                    } else if (groups.g2g2 && (groups.flags[gi] & 1) && (groups.flags[gj] & 1)) {
                        const double* a = groups.cm + 3 * gi;
                        const double* b = groups.cm + 3 * gj;
                        if (!(dist2(box, a[0], a[1], a[2], b[0], b[1], b[2]) <
                              g2g_cutoff2(groups.g2g2, groups.nmol, groups.flags[gi], groups.flags[gj])))
                            continue;
                    }
                }
                const double4 pj = tile[jj];
                const DevPair& p = s_pairs[ti * ntypes + static_cast<int>(mj & 0xffu)];
                acc += pair_energy(p, s_coul, pi.w, pj.w, dist2(box, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z));
            }
        }
    }
    const double s = block_sum<kBlock>(acc, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Plain streaming copy of a known byte count; its HBM counter readings calibrate
// the FETCH_SIZE / WRITE_SIZE scale of the profiles.
// Sum of all intra-molecular bond energies (Bonded::energy on Change::everything,
// energy.cpp:1047-1066); one partial per block.
__global__ __launch_bounds__(kBlock) void bond_total_kernel(const double4* __restrict__ pos,
                                                             const DevBondPair* __restrict__ bonds, int nb, Box box,
                                                             double* __restrict__ partials) {
    __shared__ double s_red[kBlock / 64];
    double acc = 0.0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < nb; e += gridDim.x * kBlock) {
        const DevBondPair B = bonds[e];
        const double4 a = pos[B.i], b = pos[B.j];
        const double x = B.req - sqrt(dist2(box, a.x, a.y, a.z, b.x, b.y, b.z));
        acc += B.half_k * x * x;
    }
    const double s = block_sum<kBlock>(acc, s_red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void copy_probe_kernel(const double4* __restrict__ src, double4* __restrict__ dst,
                                                            long long n) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * kBlock)
        dst[i] = src[i];
}

} // namespace

double copy_probe(long long bytes, int repeats) {
    const long long n = bytes / static_cast<long long>(sizeof(double4));
    if (n <= 0 || repeats <= 0) throw std::runtime_error("copy_probe: bad size");
    double4 *src = nullptr, *dst = nullptr;
    check_hip(hipMalloc(&src, sizeof(double4) * n), "hipMalloc probe");
    check_hip(hipMalloc(&dst, sizeof(double4) * n), "hipMalloc probe");
    check_hip(hipMemset(src, 0, sizeof(double4) * n), "memset probe");
    hipEvent_t a, b;
    check_hip(hipEventCreate(&a), "event");
    check_hip(hipEventCreate(&b), "event");
    hipLaunchKernelGGL(copy_probe_kernel, dim3(2048), dim3(kBlock), 0, nullptr, src, dst, n);
    check_hip(hipEventRecord(a, nullptr), "event");
    for (int r = 0; r < repeats; ++r) hipLaunchKernelGGL(copy_probe_kernel, dim3(2048), dim3(kBlock), 0, nullptr, src, dst, n);
    check_hip(hipEventRecord(b, nullptr), "event");
    check_hip(hipEventSynchronize(b), "event sync");
    float ms = 0;
    check_hip(hipEventElapsedTime(&ms, a, b), "elapsed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    return ms / repeats;
}

// ==============================================================================
Engine::Engine(const std::string& config_json) : model_(Model::from_json(config_json)) {
    int ndev = 0;
    check_hip(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (ndev < 1) throw std::runtime_error("no HIP device visible: the GPU engine needs an MI355X");
    check_hip(hipGetDevice(&device_), "hipGetDevice");
    hipDeviceProp_t prop;
    check_hip(hipGetDeviceProperties(&prop, device_), "hipGetDeviceProperties");
    num_cu_ = prop.multiProcessorCount;
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    if (model_.ntypes > kMaxTypes) throw std::runtime_error("too many atom types for the GPU engine");
    max_blocks_ = 4096;
    check_hip(hipMalloc(&d_partials_, sizeof(double2) * max_blocks_), "hipMalloc partials");
    check_hip(hipMalloc(&d_ticket_, sizeof(unsigned)), "hipMalloc ticket");
    check_hip(hipMemset(d_ticket_, 0, sizeof(unsigned)), "hipMemset ticket");
    check_hip(hipHostMalloc(&h_result_, sizeof(DeltaResult), hipHostMallocCoherent), "hipHostMalloc result");
    std::memset(h_result_, 0, sizeof(DeltaResult));
    check_hip(hipMalloc(&d_moved_, sizeof(MovedParticle) * kMaxMoved), "hipMalloc moved");
    upload_potentials();
}

Engine::~Engine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    (void)hipFree(d_pos_);
    (void)hipFree(d_meta_);
    (void)hipFree(d_pairs_);
    (void)hipFree(d_coul_);
    (void)hipFree(d_kn_);
    (void)hipFree(d_ak_);
    (void)hipFree(d_q_acc_);
    (void)hipFree(d_q_trial_);
    (void)hipFree(d_partials_);
    (void)hipFree(d_ticket_);
    (void)hipFree(d_moved_);
    (void)hipFree(cells_.count);
    (void)hipFree(cells_.members);
    (void)hipFree(cells_.cell_of);
    (void)hipFree(cells_.overflow);
    (void)hipFree(d_gflags_);
    (void)hipFree(d_gcm_);
    (void)hipFree(d_g2g_);
    (void)hipFree(d_bond_off_);
    (void)hipFree(d_bond_);
    (void)hipFree(d_bond_all_);
    (void)hipFree(d_excl_off_);
    (void)hipFree(d_excl_);
    if (h_result_) (void)hipHostFree(h_result_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::upload_potentials() {
    const int nt = model_.ntypes;
    h_pairs_.assign(static_cast<size_t>(nt * nt), DevPair{});
    for (int i = 0; i < nt * nt; ++i) {
        const PairParam& p = model_.pairs[static_cast<size_t>(i)];
        h_pairs_[static_cast<size_t>(i)] = DevPair{p.lj_s2, p.lj_e4, p.wca_s2, p.wca_e4, p.hs_s2, p.coulomb, 0, 0, 0.0};
    }
    std::memset(&h_coul_, 0, sizeof(h_coul_));
    if (model_.splined) {
        // one table per unordered type pair
        std::vector<int> slot(static_cast<size_t>(nt * nt), -1);
        int ns = 0;
        for (int a = 0; a < nt; ++a)
            for (int b = 0; b <= a; ++b) {
                if (ns == kSplMaxTables)
                    throw std::runtime_error("nonbonded_splined: at most " + std::to_string(kSplMaxTables) +
                                             " type-pair tables on the device (four atom types)");
                const SplinedPairCfg& s = model_.spl[static_cast<size_t>(a * nt + b)];
                const int n = static_cast<int>(s.table.knots.size());
                if (n < 2 || n > kSplMaxKnots)
                    throw std::runtime_error("nonbonded_splined: a table has " + std::to_string(n) +
                                             " knots; the device holds at most " + std::to_string(kSplMaxKnots));
                DevSplineTable& T = h_coul_.spl[ns];
                T.rmin2 = s.rmin2;
                T.rmax2 = s.rmax2;
                T.n = n;
                T.hs = s.hardsphere ? 1 : 0;
                std::copy(s.table.knots.begin(), s.table.knots.end(), T.knots);
                std::copy(s.table.coeff.begin(), s.table.coeff.end(), T.coeff);
                T.x0 = s.table.knots.front();
                const double span = s.table.knots.back() - T.x0;
                T.inv_w = kSplBuckets / span;
                for (int k = 0; k <= kSplBuckets; ++k) {
                    const double x = T.x0 + k / T.inv_w;
                    T.bucket[k] = static_cast<unsigned char>(
                        std::lower_bound(s.table.knots.begin(), s.table.knots.end(), x) - s.table.knots.begin());
                }
                // lower_bound(x) for x in bucket k lies in [bucket[k], bucket[k+1]] (one more
                // for rounding of the bucket edges); the search starts one knot early
                int steps = 1;
                for (int k = 0; k < kSplBuckets; ++k) {
                    const int start = std::max(0, static_cast<int>(T.bucket[k]) - 1);
                    steps = std::max(steps, std::min(n, static_cast<int>(T.bucket[k + 1]) + 1) - start + 1);
                }
                T.steps = steps;
                slot[static_cast<size_t>(a * nt + b)] = slot[static_cast<size_t>(b * nt + a)] = ns++;
            }
        h_coul_.nspl = ns;
        for (int i = 0; i < nt * nt; ++i) h_pairs_[static_cast<size_t>(i)].spl = slot[static_cast<size_t>(i)];
    }
    const CoulombModel& c = model_.coulomb;
    h_coul_.scheme = c.scheme;
    h_coul_.splined = c.splined ? 1 : 0;
    h_coul_.lB = c.lB;
    h_coul_.rc = c.rc;
    h_coul_.rc_inv = c.rc_inv;
    h_coul_.alpha = c.alpha;
    if (c.splined && (c.scheme == kCoulEwald || c.scheme == kCoulFanourgakis)) {
        if (static_cast<int>(c.S.knots.size()) > kMaxKnots) throw std::runtime_error("spline has too many knots");
        h_coul_.nknots = static_cast<int>(c.S.knots.size());
        std::copy(c.S.knots.begin(), c.S.knots.end(), h_coul_.knots);
        std::copy(c.S.coeff.begin(), c.S.coeff.end(), h_coul_.coeff);
        for (int b = 0; b <= kSplineBuckets; ++b) {
            const double x = static_cast<double>(b) / kSplineBuckets;
            h_coul_.bucket[b] = static_cast<unsigned short>(
                std::lower_bound(c.S.knots.begin(), c.S.knots.end(), x) - c.S.knots.begin());
        }
        // lower_bound(x) for x in bucket b lies in [bucket[b], bucket[b+1]]; the search
        // starts one knot early and gets one extra step for FP32 rounding of the knots
        int steps = 1;
        for (int b = 0; b < kSplineBuckets; ++b) {
            const int start = std::max(0, static_cast<int>(h_coul_.bucket[b]) - 1);
            steps = std::max(steps, static_cast<int>(h_coul_.bucket[b + 1]) - start + 1);
        }
        h_coul_.spline_steps = steps;
    }
    check_hip(hipMalloc(&d_pairs_, sizeof(DevPair) * h_pairs_.size()), "hipMalloc pairs");
    check_hip(hipMalloc(&d_coul_, sizeof(DevCoulomb)), "hipMalloc coulomb");
    check_hip(hipMemcpy(d_pairs_, h_pairs_.data(), sizeof(DevPair) * h_pairs_.size(), hipMemcpyHostToDevice), "copy pairs");
    check_hip(hipMemcpy(d_coul_, &h_coul_, sizeof(DevCoulomb), hipMemcpyHostToDevice), "copy coulomb");
}

void Engine::set_particles(int n, const double* xyz, const double* charge, const int* atom_id) {
    if (n <= 0) throw std::runtime_error("set_particles: n must be positive");
    n_ = n;
    h_xyz_.assign(xyz, xyz + 3 * static_cast<size_t>(n));
    h_q_.assign(charge, charge + n);
    h_type_.assign(atom_id, atom_id + n);
    for (int t : h_type_)
        if (t < 0 || t >= model_.ntypes) throw std::runtime_error("set_particles: atom id out of range");
    h_gid_.assign(static_cast<size_t>(n), 0);
}

void Engine::set_groups(int ng, const int* begin, const int* size, const int* molid, const double* mass_center) {
    if (n_ == 0) throw std::runtime_error("set_groups: call set_particles first");
    if (ng <= 0 || ng >= (1 << 24)) throw std::runtime_error("set_groups: bad number of groups");
    g_begin_.assign(begin, begin + ng);
    g_size_.assign(size, size + ng);
    g_molid_.assign(molid, molid + ng);
    g_cm_.assign(3 * static_cast<size_t>(ng), 0.0);
    if (mass_center) g_cm_.assign(mass_center, mass_center + 3 * static_cast<size_t>(ng));
    int expect = 0;
    any_molecular_ = false;
    for (int g = 0; g < ng; ++g) {
        if (g_begin_[g] != expect) throw std::runtime_error("set_groups: groups must tile the particles in order");
        if (g_size_[g] < 0 || g_size_[g] > n_ - expect)
            throw std::runtime_error("set_groups: group extends past the last particle");
        if (g_molid_[g] < 0 || g_molid_[g] >= static_cast<int>(model_.molecules.size()))
            throw std::runtime_error("set_groups: molecule id out of range");
        for (int i = 0; i < g_size_[g]; ++i) h_gid_[static_cast<size_t>(g_begin_[g] + i)] = g;
        expect += g_size_[g];
        const auto& mol = model_.molecules[static_cast<size_t>(g_molid_[g])];
        if (!mol.atomic) any_molecular_ = true;
    }
    if (expect != n_) throw std::runtime_error("set_groups: groups must cover all particles");
    if (!mass_center)
        for (int g = 0; g < ng; ++g)
            if (!model_.molecules[static_cast<size_t>(g_molid_[g])].atomic && g_size_[g] > 0)
                group_mass_center_host(g, &g_cm_[static_cast<size_t>(3 * g)]);
}

// Geometry::weightedCenter with masses (geometry.h:505-525): shift by minus the first
// atom, wrap, mass-weighted mean, shift back, wrap.
void Engine::group_mass_center_host(int g, double* cm) const {
    const int b = g_begin_[static_cast<size_t>(g)];
    const double shift[3] = {-h_xyz_[3 * b], -h_xyz_[3 * b + 1], -h_xyz_[3 * b + 2]};
    auto wrap = [&](double& x, int d) {
        const double L = model_.box[d];
        if (std::fabs(x) > 0.5 * L) {
            const double t = x / L;
            x -= L * static_cast<double>(static_cast<int>(t > 0.0 ? t + 0.5 : t - 0.5));
        }
    };
    double c[3] = {0, 0, 0}, wsum = 0;
    for (int k = 0; k < g_size_[static_cast<size_t>(g)]; ++k) {
        const int i = b + k;
        const double w = model_.atoms[static_cast<size_t>(h_type_[static_cast<size_t>(i)])].mw;
        for (int d = 0; d < 3; ++d) {
            double x = h_xyz_[3 * i + d] + shift[d];
            wrap(x, d);
            c[d] += w * x;
        }
        wsum += w;
    }
    for (int d = 0; d < 3; ++d) {
        double x = c[d] / wsum - shift[d];
        wrap(x, d);
        cm[d] = x;
    }
}

void Engine::stage_mass_center(int group, const double* cm) {
    if (group < 0 || group >= static_cast<int>(g_begin_.size())) throw std::runtime_error("stage_mass_center: group out of range");
    staged_cm_group_ = group;
    for (int d = 0; d < 3; ++d) staged_cm_[d] = cm[d];
}

// Group flags, mass centres and the bond tables on the device.
void Engine::build_group_tables() {
    const int ng = static_cast<int>(g_begin_.size());
    std::vector<int> flags(static_cast<size_t>(ng));
    for (int g = 0; g < ng; ++g) {
        const MoleculeCfg& mol = model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(g)])];
        flags[static_cast<size_t>(g)] = (mol.atomic ? 0 : 1) | (mol.rigid ? 2 : 0) | (g_molid_[static_cast<size_t>(g)] << 8);
    }
    (void)hipFree(d_g2g_);
    d_g2g_ = nullptr;
    if (model_.has_g2g()) {
        check_hip(hipMalloc(&d_g2g_, sizeof(double) * model_.g2g2.size()), "hipMalloc g2g cutoffs");
        check_hip(hipMemcpy(d_g2g_, model_.g2g2.data(), sizeof(double) * model_.g2g2.size(), hipMemcpyHostToDevice),
                  "copy g2g cutoffs");
    }
    (void)hipFree(d_gflags_);
    (void)hipFree(d_gcm_);
    check_hip(hipMalloc(&d_gflags_, sizeof(int) * std::max(1, ng)), "hipMalloc group flags");
    check_hip(hipMalloc(&d_gcm_, sizeof(double) * 3 * std::max(1, ng)), "hipMalloc group centres");
    check_hip(hipMemcpy(d_gflags_, flags.data(), sizeof(int) * ng, hipMemcpyHostToDevice), "copy group flags");
    check_hip(hipMemcpy(d_gcm_, g_cm_.data(), sizeof(double) * 3 * ng, hipMemcpyHostToDevice), "copy group centres");
    // bonds (Bonded::updateGroupBonds, energy.cpp:974-984: molecule bonds shifted to absolute indices)
    std::vector<std::vector<DevBond>> per(static_cast<size_t>(n_));
    std::vector<DevBondPair> flat;
    if (model_.bonded_term >= 0) {
        for (int g = 0; g < ng; ++g) {
            const MoleculeCfg& mol = model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(g)])];
            const int b0 = g_begin_[static_cast<size_t>(g)], sz = g_size_[static_cast<size_t>(g)];
            for (const BondCfg& bd : mol.bonds) {
                if (bd.i >= sz || bd.j >= sz) throw std::runtime_error("bonded index out of range for group size");
                per[static_cast<size_t>(b0 + bd.i)].push_back(DevBond{b0 + bd.j, 0, bd.half_k, bd.req});
                per[static_cast<size_t>(b0 + bd.j)].push_back(DevBond{b0 + bd.i, 0, bd.half_k, bd.req});
                flat.push_back(DevBondPair{b0 + bd.i, b0 + bd.j, bd.half_k, bd.req});
            }
        }
    }
    // exclusions (MoleculeData::isPairExcluded): per-particle sorted absolute partners
    std::vector<std::vector<int>> ex(static_cast<size_t>(n_));
    bool any_excl = false;
    for (int g = 0; g < ng; ++g) {
        const MoleculeCfg& mol = model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(g)])];
        if (mol.atomic) continue;
        const int b0 = g_begin_[static_cast<size_t>(g)], sz = g_size_[static_cast<size_t>(g)];
        for (auto [i, j] : mol.excluded) {
            if (i >= sz || j >= sz) continue; // pairs beyond the group's atoms do not apply
            ex[static_cast<size_t>(b0 + i)].push_back(b0 + j);
            ex[static_cast<size_t>(b0 + j)].push_back(b0 + i);
            any_excl = true;
        }
    }
    (void)hipFree(d_excl_off_);
    (void)hipFree(d_excl_);
    d_excl_off_ = d_excl_ = nullptr;
    if (any_excl) {
        std::vector<int> eoff(static_cast<size_t>(n_) + 1, 0), elist;
        for (int i = 0; i < n_; ++i) {
            auto& v = ex[static_cast<size_t>(i)];
            std::sort(v.begin(), v.end());
            elist.insert(elist.end(), v.begin(), v.end());
            eoff[static_cast<size_t>(i) + 1] = static_cast<int>(elist.size());
        }
        check_hip(hipMalloc(&d_excl_off_, sizeof(int) * eoff.size()), "hipMalloc exclusion offsets");
        check_hip(hipMalloc(&d_excl_, sizeof(int) * elist.size()), "hipMalloc exclusions");
        check_hip(hipMemcpy(d_excl_off_, eoff.data(), sizeof(int) * eoff.size(), hipMemcpyHostToDevice), "copy exclusions");
        check_hip(hipMemcpy(d_excl_, elist.data(), sizeof(int) * elist.size(), hipMemcpyHostToDevice), "copy exclusions");
    }
    std::vector<int> off(static_cast<size_t>(n_) + 1, 0);
    std::vector<DevBond> list;
    for (int i = 0; i < n_; ++i) {
        off[static_cast<size_t>(i) + 1] = off[static_cast<size_t>(i)] + static_cast<int>(per[static_cast<size_t>(i)].size());
        list.insert(list.end(), per[static_cast<size_t>(i)].begin(), per[static_cast<size_t>(i)].end());
    }
    nbonds_ = static_cast<int>(flat.size());
    (void)hipFree(d_bond_off_);
    (void)hipFree(d_bond_);
    (void)hipFree(d_bond_all_);
    check_hip(hipMalloc(&d_bond_off_, sizeof(int) * off.size()), "hipMalloc bond offsets");
    check_hip(hipMalloc(&d_bond_, sizeof(DevBond) * std::max<size_t>(1, list.size())), "hipMalloc bonds");
    check_hip(hipMalloc(&d_bond_all_, sizeof(DevBondPair) * std::max<size_t>(1, flat.size())), "hipMalloc bond list");
    check_hip(hipMemcpy(d_bond_off_, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice), "copy bond offsets");
    if (!list.empty())
        check_hip(hipMemcpy(d_bond_, list.data(), sizeof(DevBond) * list.size(), hipMemcpyHostToDevice), "copy bonds");
    if (!flat.empty())
        check_hip(hipMemcpy(d_bond_all_, flat.data(), sizeof(DevBondPair) * flat.size(), hipMemcpyHostToDevice),
                  "copy bond list");
}

void Engine::build_kvectors() {
    // PolicyIonIon::updateBox (energy.cpp:133-186): n_x >= 0, |n|^2 <= n_c^2, n != 0.
    const EwaldModel& e = model_.ewald;
    const int nc = static_cast<int>(std::ceil(e.n_cutoff));
    if (nc > 127) throw std::runtime_error("ncutoff too large");
    const double maxL = std::max(model_.box[0], std::max(model_.box[1], model_.box[2]));
    const double k2zero = 0.1 * std::pow(2 * kPi / maxL, 2);
    const double nc2 = e.n_cutoff * e.n_cutoff;
    std::vector<int> kn;
    std::vector<double> ak;
    for (int nx = 0; nx <= nc; nx++)
        for (int ny = -nc; ny <= nc; ny++)
            for (int nz = -nc; nz <= nc; nz++) {
                const double kx = 2 * kPi * (double(nx) / model_.box[0]);
                const double ky = 2 * kPi * (double(ny) / model_.box[1]);
                const double kz = 2 * kPi * (double(nz) / model_.box[2]);
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < k2zero) continue;
                if (e.spherical_sum && (double(nx * nx) + double(ny * ny) + double(nz * nz)) / nc2 > 1) continue;
                kn.push_back((nx & 0xff) | ((ny & 0xff) << 8) | ((nz & 0xff) << 16));
                ak.push_back((nx > 0 ? 2.0 : 1.0) * std::exp(-k2 / (4 * e.alpha * e.alpha)) / k2);
            }
    nk_ = static_cast<int>(kn.size());
    check_hip(hipMalloc(&d_kn_, sizeof(int) * std::max(1, nk_)), "hipMalloc kn");
    check_hip(hipMalloc(&d_ak_, sizeof(double) * std::max(1, nk_)), "hipMalloc ak");
    check_hip(hipMalloc(&d_q_acc_, sizeof(double2) * std::max(1, nk_)), "hipMalloc q");
    check_hip(hipMalloc(&d_q_trial_, sizeof(double2) * std::max(1, nk_)), "hipMalloc q trial");
    if (nk_ > 0) {
        check_hip(hipMemcpy(d_kn_, kn.data(), sizeof(int) * nk_, hipMemcpyHostToDevice), "copy kn");
        check_hip(hipMemcpy(d_ak_, ak.data(), sizeof(double) * nk_, hipMemcpyHostToDevice), "copy ak");
    }
}

static Box make_box(const Model& m) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.L[d] = m.box[d];
        b.half[d] = 0.5 * m.box[d];
    }
    return b;
}

void Engine::full_q(double2* out) {
    if (nk_ == 0) return;
    const Box box = make_box(model_);
    hipLaunchKernelGGL(full_q_kernel, dim3((nk_ + kBlock - 1) / kBlock), dim3(kBlock), 0, stream_, d_pos_, n_, d_kn_,
                       nk_, box, out);
    check_hip(hipGetLastError(), "full_q_kernel");
}

void Engine::set_cell_capacity(int cap) {
    if (cap < 0) throw std::runtime_error("cell capacity must be >= 0");
    cell_capacity_ = cap;
}

void Engine::set_cell_divisions(int d) {
    if (d < 1 || d > 2) throw std::runtime_error("cell divisions: 1 or 2");
    if (d_pos_ != nullptr) throw std::runtime_error("cell divisions must be set before init");
    cell_divisions_ = d;
}

void Engine::regrid(int d) {
    if (d < 1 || d > 2) throw std::runtime_error("cell divisions: 1 or 2");
    cell_divisions_ = d; // the next reset_positions builds the grid
}

// Distance beyond which no pair of atoms interacts: the pair potentials' range, or with
// cutoff_g2g in a system made only of rigid molecules the largest group cutoff plus the
// two largest distances of an atom from its group's mass centre (GroupCutoff::cut drops
// every pair of groups whose centres are farther apart).
double Engine::interaction_range() const {
    double range = model_.pair_range;
    if (!model_.has_g2g() || g_begin_.empty()) return range;
    double rmax = 0.0;
    for (size_t g = 0; g < g_begin_.size(); ++g) {
        const MoleculeCfg& mol = model_.molecules[static_cast<size_t>(g_molid_[g])];
        if (mol.atomic || !mol.rigid) return range;
        for (int k = 0; k < g_size_[g]; ++k) {
            const int i = g_begin_[g] + k;
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                double x = h_xyz_[3 * static_cast<size_t>(i) + d] - g_cm_[3 * g + d];
                x -= model_.box[d] * std::round(x / model_.box[d]);
                r2 += x * x;
            }
            rmax = std::max(rmax, std::sqrt(r2));
        }
    }
    double c2 = 0.0;
    for (double v : model_.g2g2) c2 = std::max(c2, v);
    // 1e-6 A above the bound: rigid rotations keep the atom distances up to rounding
    return std::min(range, std::sqrt(c2) + 2.0 * rmax + 1e-6);
}

void Engine::build_cell_list() {
    cells_.enabled = false;
    const double range = interaction_range();
    cell_range_ = range;
    if (!(range > 0) || !std::isfinite(range)) return;
    int nc[3];
    // cells of at least range / r with a (2r+1)^3 stencil; the finer grid needs 2r+1
    // cells per side, else the 27-cell grid is used
    int r = cell_divisions_;
    for (int d = 0; d < 3; ++d)
        if (std::floor(model_.box[d] * r / range) < 2 * r + 1) r = 1;
    for (int d = 0; d < 3; ++d) {
        if (range > 0.5 * model_.box[d]) return;
        nc[d] = static_cast<int>(std::floor(model_.box[d] * r / range));
        if (nc[d] < 2 * r + 1) return;
    }
    cells_.radius = r;
    const long ncell = static_cast<long>(nc[0]) * nc[1] * nc[2];
    if (ncell > (1L << 24)) return;
    std::vector<int> count(static_cast<size_t>(ncell), 0), cell_of(static_cast<size_t>(n_));
    auto coord = [&](double x, int d) {
        int c = static_cast<int>(std::floor((x + 0.5 * model_.box[d]) * (nc[d] / model_.box[d])));
        c %= nc[d];
        if (c < 0) c += nc[d];
        return c;
    };
    for (int i = 0; i < n_; ++i) {
        const int c = (coord(h_xyz_[3 * i + 2], 2) * nc[1] + coord(h_xyz_[3 * i + 1], 1)) * nc[0] + coord(h_xyz_[3 * i], 0);
        cell_of[static_cast<size_t>(i)] = c;
        count[static_cast<size_t>(c)]++;
    }
    const int maxcount = *std::max_element(count.begin(), count.end());
    const double mean = static_cast<double>(n_) / static_cast<double>(ncell);
    int cap = std::max(maxcount * 2, static_cast<int>(4 * mean + 32));
    cap = (cap + 7) & ~7;
    if (cell_capacity_ > 0) {
        if (cell_capacity_ < maxcount) throw std::runtime_error("cell capacity below the fullest initial cell");
        cap = cell_capacity_;
    }
    std::vector<int> members(static_cast<size_t>(ncell * cap), -1);
    std::fill(count.begin(), count.end(), 0);
    for (int i = 0; i < n_; ++i) {
        const int c = cell_of[static_cast<size_t>(i)];
        members[static_cast<size_t>(c) * cap + count[static_cast<size_t>(c)]++] = i;
    }
    cells_.enabled = true;
    cells_.cap = cap;
    for (int d = 0; d < 3; ++d) {
        cells_.nc[d] = nc[d];
        cells_.inv_cell[d] = nc[d] / model_.box[d];
    }
    check_hip(hipMalloc(&cells_.count, sizeof(int) * ncell), "hipMalloc cells");
    check_hip(hipMalloc(&cells_.members, sizeof(int) * ncell * cap), "hipMalloc cells");
    check_hip(hipMalloc(&cells_.cell_of, sizeof(int) * n_), "hipMalloc cells");
    check_hip(hipMalloc(&cells_.overflow, sizeof(int)), "hipMalloc cells");
    check_hip(hipMemcpy(cells_.count, count.data(), sizeof(int) * ncell, hipMemcpyHostToDevice), "copy cells");
    check_hip(hipMemcpy(cells_.members, members.data(), sizeof(int) * ncell * cap, hipMemcpyHostToDevice), "copy cells");
    check_hip(hipMemcpy(cells_.cell_of, cell_of.data(), sizeof(int) * n_, hipMemcpyHostToDevice), "copy cells");
    check_hip(hipMemset(cells_.overflow, 0, sizeof(int)), "memset cells");
}

void Engine::init() {
    if (n_ == 0 || g_begin_.empty()) throw std::runtime_error("init: particles and groups must be set");
    // device particle arrays
    std::vector<double4> p(static_cast<size_t>(n_));
    std::vector<uint32_t> meta(static_cast<size_t>(n_));
    for (int i = 0; i < n_; ++i) {
        p[static_cast<size_t>(i)] = make_double4(h_xyz_[3 * i], h_xyz_[3 * i + 1], h_xyz_[3 * i + 2], h_q_[static_cast<size_t>(i)]);
        meta[static_cast<size_t>(i)] = static_cast<uint32_t>(h_type_[static_cast<size_t>(i)]) |
                                       (static_cast<uint32_t>(h_gid_[static_cast<size_t>(i)]) << 8);
    }
    (void)hipFree(d_pos_);
    (void)hipFree(d_meta_);
    check_hip(hipMalloc(&d_pos_, sizeof(double4) * n_), "hipMalloc pos");
    check_hip(hipMalloc(&d_meta_, sizeof(uint32_t) * n_), "hipMalloc meta");
    check_hip(hipMemcpy(d_pos_, p.data(), sizeof(double4) * n_, hipMemcpyHostToDevice), "copy pos");
    check_hip(hipMemcpy(d_meta_, meta.data(), sizeof(uint32_t) * n_, hipMemcpyHostToDevice), "copy meta");
    if (model_.ewald.present) {
        build_kvectors();
        full_q(d_q_acc_);
        check_hip(hipStreamSynchronize(stream_), "init full Q");
    }
    for (int d = 0; d < 3; ++d) {
        dipole_[d] = 0;
        for (int i = 0; i < n_; ++i) dipole_[d] += h_xyz_[3 * i + d] * h_q_[static_cast<size_t>(i)];
    }
    build_cell_list();
    build_group_tables();
}

void Engine::stage_trial(int m, const int* idx, const double* xyz) {
    if (m < 0 || m > kMaxMoved) throw std::runtime_error("stage_trial: too many moved particles");
    staged_.resize(static_cast<size_t>(m));
    for (int a = 0; a < m; ++a) {
        const int i = idx[a];
        if (i < 0 || i >= n_) throw std::runtime_error("stage_trial: particle index out of range");
        MovedParticle& M = staged_[static_cast<size_t>(a)];
        M.idx = i;
        M.type = h_type_[static_cast<size_t>(i)];
        M.gid = h_gid_[static_cast<size_t>(i)];
        M.pad = 0;
        M.newp = make_double4(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], h_q_[static_cast<size_t>(i)]);
        M.oldp = make_double4(h_xyz_[3 * i], h_xyz_[3 * i + 1], h_xyz_[3 * i + 2], h_q_[static_cast<size_t>(i)]);
    }
}

double Engine::self_energy_of(int idx) const {
    const double q = h_q_[static_cast<size_t>(idx)];
    return model_.coulomb.self_prefactor * q * q;
}

double Engine::surface_energy(const double* dip) const {
    const EwaldModel& e = model_.ewald;
    if (e.const_inf < 0.5) return 0.0;
    const double V = model_.box[0] * model_.box[1] * model_.box[2];
    const double m2 = dip[0] * dip[0] + dip[1] * dip[1] + dip[2] * dip[2];
    return e.const_inf * 2.0 * kPi / ((2.0 * e.epss + 1.0) * V) * m2 * e.lB;
}

void Engine::set_timing(bool on) {
    timing_ = on;
    ev_pairs_.clear();
}

void Engine::timing(double* total_ms, long* launches) {
    check_hip(hipStreamSynchronize(stream_), "timing sync");
    double t = 0;
    for (auto& pr : ev_pairs_) {
        float ms = 0;
        check_hip(hipEventElapsedTime(&ms, ev_pool_[static_cast<size_t>(pr.first)], ev_pool_[static_cast<size_t>(pr.second)]), "event elapsed");
        t += ms;
    }
    *total_ms = t;
    *launches = static_cast<long>(ev_pairs_.size());
}

void Engine::launch_delta(const MoveArgs& args, int pair_blocks, int ewald_blocks, bool bond_block) {
    const Box box = make_box(model_);
    GroupTable gt{d_gflags_, d_gcm_, d_g2g_, static_cast<int>(model_.molecules.size()), d_excl_off_, d_excl_};
    CellArgs ca{cells_.count, cells_.members, cells_.nc[0], cells_.nc[1], cells_.nc[2], cells_.cap,
                cells_.inv_cell[0], cells_.inv_cell[1], cells_.inv_cell[2]};
    const int ewald_end = pair_blocks + ewald_blocks;
    const int grid = ewald_end + (bond_block ? 1 : 0);
    if (grid > max_blocks_) throw std::runtime_error("too many blocks");
    const int ncut = model_.ewald.present ? static_cast<int>(std::ceil(model_.ewald.n_cutoff)) : 0;
    const size_t lds = (args.m <= kPhaseTableMoved && nk_ > 0)
                           ? sizeof(double2) * static_cast<size_t>(args.m) * 2 * 3 * (2 * ncut + 1)
                           : 0;
    const double V = model_.box[0] * model_.box[1] * model_.box[2];
    const double pref = model_.ewald.present ? 2 * kPi * model_.ewald.lB / V : 0.0;
    // reference order: 2*pi*energy*lB/V (energy.cpp:530); the prefactor is applied to
    // the sum, a difference in the last bit at most.
    ++seq_;
    int e0 = -1, e1 = -1;
    if (timing_) {
        while (ev_pool_.size() < ev_pairs_.size() * 2 + 2) {
            hipEvent_t e;
            check_hip(hipEventCreate(&e), "hipEventCreate");
            ev_pool_.push_back(e);
        }
        e0 = static_cast<int>(ev_pairs_.size() * 2);
        e1 = e0 + 1;
        check_hip(hipEventRecord(ev_pool_[static_cast<size_t>(e0)], stream_), "event record");
    }
    // mass centres of the moved group (group cutoff between molecular groups)
    double cn[3] = {0, 0, 0}, co[3] = {0, 0, 0};
    if (args.change_group >= 0) {
        for (int d = 0; d < 3; ++d) co[d] = cn[d] = g_cm_[static_cast<size_t>(3 * args.change_group + d)];
        if (staged_cm_group_ == args.change_group)
            for (int d = 0; d < 3; ++d) cn[d] = staged_cm_[d];
    }
    if (cells_.enabled && cells_.radius != 1)
        throw std::runtime_error("the per-move path walks 27-cell stencils (cell divisions 1)");
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), lds, stream_, d_pos_, d_meta_, n_, d_pairs_,
                           model_.ntypes, d_coul_, box, gt, ca, args, pair_blocks, d_kn_, d_ak_, d_q_acc_, d_q_trial_,
                           nk_, ncut, d_partials_, d_ticket_, h_result_, seq_, pref, cn[0], cn[1], cn[2], co[0], co[1],
                           co[2], ewald_end, d_bond_off_, d_bond_);
    };
    if (any_molecular_) {
        if (cells_.enabled) go(delta_kernel<true, true>);
        else go(delta_kernel<true, false>);
    } else {
        if (cells_.enabled) go(delta_kernel<false, true>);
        else go(delta_kernel<false, false>);
    }
    check_hip(hipGetLastError(), "delta_kernel launch");
    if (timing_) {
        check_hip(hipEventRecord(ev_pool_[static_cast<size_t>(e1)], stream_), "event record");
        ev_pairs_.emplace_back(e0, e1);
    }
    ++launches_;
}

void Engine::delta(const std::vector<int>& groups, const std::vector<char>& all, const std::vector<char>& internal,
                   const std::vector<std::vector<int>>& rel, double* e_new, double* e_old) {
    if (d_pos_ == nullptr) throw std::runtime_error("delta: engine not initialised");
    if (groups.empty()) throw std::runtime_error("delta: empty change");
    // Moved set of the pair loop (GroupPairing::accumulate, energy.h:1447-1475) and of
    // the Ewald update (energy.cpp:219-247) for this Change.
    std::vector<int> moved;
    const bool single = groups.size() == 1;
    for (size_t g = 0; g < groups.size(); ++g) {
        const int gi = groups[g];
        if (gi < 0 || gi >= static_cast<int>(g_begin_.size())) throw std::runtime_error("delta: group index out of range");
        const bool whole = !single || all[g] || rel[g].empty();
        if (!single && !all[g] && !rel[g].empty() && static_cast<int>(rel[g].size()) != g_size_[static_cast<size_t>(gi)])
            throw std::runtime_error("delta: partial moves of several groups are not supported");
        if (whole) {
            if (!all[g] && rel[g].empty()) throw std::runtime_error("delta: a group change without indices is not supported");
            for (int i = 0; i < g_size_[static_cast<size_t>(gi)]; ++i) moved.push_back(g_begin_[static_cast<size_t>(gi)] + i);
        } else {
            for (int r : rel[g]) {
                if (r < 0 || r >= g_size_[static_cast<size_t>(gi)]) throw std::runtime_error("delta: relative index out of range");
                moved.push_back(g_begin_[static_cast<size_t>(gi)] + r);
            }
        }
    }
    if (!single && model_.has_g2g())
        for (int gi : groups)
            if (!model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(gi)])].atomic)
                throw std::runtime_error("delta: changes of several molecular groups with cutoff_g2g are not supported");
    std::sort(moved.begin(), moved.end());
    if (std::adjacent_find(moved.begin(), moved.end()) != moved.end()) throw std::runtime_error("delta: duplicate moved particle");
    if (static_cast<int>(moved.size()) > kMaxMoved) throw std::runtime_error("delta: too many moved particles");
    // coordinates: staged trial coordinates where given, accepted ones otherwise
    std::vector<MovedParticle> mv(moved.size());
    for (size_t a = 0; a < moved.size(); ++a) {
        const int i = moved[a];
        MovedParticle M;
        M.idx = i;
        M.type = h_type_[static_cast<size_t>(i)];
        M.gid = h_gid_[static_cast<size_t>(i)];
        M.pad = 0;
        M.oldp = make_double4(h_xyz_[3 * i], h_xyz_[3 * i + 1], h_xyz_[3 * i + 2], h_q_[static_cast<size_t>(i)]);
        M.newp = M.oldp;
        for (const auto& s : staged_)
            if (s.idx == i) M.newp = s.newp;
        mv[a] = M;
    }
    MoveArgs args{};
    args.m = static_cast<int>(mv.size());
    args.change_group = single ? groups[0] : -1;
    args.internal = single ? (internal[0] ? 1 : 0) : 0;
    if (args.m <= kMaxInlineMoved) {
        for (int a = 0; a < args.m; ++a) args.inl[a] = mv[static_cast<size_t>(a)];
        args.ext = nullptr;
    } else {
        check_hip(hipMemcpyAsync(d_moved_, mv.data(), sizeof(MovedParticle) * mv.size(), hipMemcpyHostToDevice, stream_),
                  "copy moved");
        args.ext = d_moved_;
    }
    int pair_blocks;
    if (cells_.enabled) pair_blocks = std::min(args.m * 54, 256);
    else pair_blocks = std::max(1, std::min((n_ + kBlock * 4 - 1) / (kBlock * 4), 2 * num_cu_));
    int ewald_blocks = 0;
    if (nk_ > 0) ewald_blocks = std::max(1, std::min((nk_ + 2 * kBlock - 1) / (2 * kBlock), 2 * num_cu_));
    if (single && any_molecular_ && model_.has_g2g() &&
        !model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(groups[0])])].atomic &&
        staged_cm_group_ != groups[0])
        throw std::runtime_error("delta: stage the trial mass centre of a moved molecular group when cutoff_g2g is set");
    launch_delta(args, pair_blocks, ewald_blocks, nbonds_ > 0 && model_.bonded_term >= 0);
    check_hip(hipStreamSynchronize(stream_), "delta sync");
    volatile DeltaResult* r = h_result_;
    if (r->seq != seq_) throw std::runtime_error("delta: result sequence mismatch");
    // dipole of the trial state (surface term, energy.cpp:466-482)
    for (int d = 0; d < 3; ++d) dipole_trial_[d] = dipole_[d];
    for (const auto& M : mv) {
        dipole_trial_[0] += M.newp.w * M.newp.x - M.oldp.w * M.oldp.x;
        dipole_trial_[1] += M.newp.w * M.newp.y - M.oldp.w * M.oldp.y;
        dipole_trial_[2] += M.newp.w * M.newp.z - M.oldp.w * M.oldp.z;
    }
    // Hamiltonian::energy order and early break (energy.cpp:1227-1241)
    double self = 0;
    if (model_.has_self_term) {
        for (size_t g = 0; g < groups.size(); ++g) {
            const int gi = groups[g];
            if (all[g]) {
                for (int i = 0; i < g_size_[static_cast<size_t>(gi)]; ++i) self += self_energy_of(g_begin_[static_cast<size_t>(gi)] + i);
            } else {
                for (int r2 : rel[g]) self += self_energy_of(g_begin_[static_cast<size_t>(gi)] + r2);
            }
        }
    }
    auto fill = [&](double* out, double pair, double rec, double bond, const double* dip) {
        bool stop = false;
        for (size_t t = 0; t < model_.term_names.size(); ++t) {
            const std::string& name = model_.term_names[t];
            double v = 0.0;
            if (name == "selfenergy") v = self;
            else if (name == "nonbonded") v = pair;
            else if (name == "ewald") v = surface_energy(dip) + rec;
            else if (name == "bonded") v = bond;
            out[t] = stop ? 0.0 : v;
            if (!stop && (v >= model_.maximum_allowed_energy || std::isnan(v))) stop = true;
        }
    };
    fill(e_new, r->pair_new, r->rec_new, r->bond_new, dipole_trial_);
    fill(e_old, r->pair_old, r->rec_old, r->bond_old, dipole_);
    pending_moved_ = mv;
    pending_cm_group_ = (single && staged_cm_group_ == groups[0]) ? staged_cm_group_ : -1;
    for (int d = 0; d < 3; ++d) pending_cm_[d] = staged_cm_[d];
    pending_ = true;
}

void Engine::sync(bool accept) {
    if (!pending_) return;
    if (accept) {
        MoveArgs args{};
        args.m = static_cast<int>(pending_moved_.size());
        if (args.m <= kMaxInlineMoved) {
            for (int a = 0; a < args.m; ++a) args.inl[a] = pending_moved_[static_cast<size_t>(a)];
        } else {
            check_hip(hipMemcpyAsync(d_moved_, pending_moved_.data(), sizeof(MovedParticle) * pending_moved_.size(),
                                     hipMemcpyHostToDevice, stream_), "copy moved");
            args.ext = d_moved_;
        }
        CellArgs ca{cells_.count, cells_.members, cells_.nc[0], cells_.nc[1], cells_.nc[2], cells_.cap,
                    cells_.inv_cell[0], cells_.inv_cell[1], cells_.inv_cell[2]};
        hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(64), 0, stream_, d_pos_, args, ca, cells_.cell_of,
                           cells_.overflow, make_box(model_), cells_.enabled ? 1 : 0);
        check_hip(hipGetLastError(), "commit_kernel");
        for (const auto& M : pending_moved_) {
            h_xyz_[3 * M.idx] = M.newp.x;
            h_xyz_[3 * M.idx + 1] = M.newp.y;
            h_xyz_[3 * M.idx + 2] = M.newp.z;
        }
        std::swap(d_q_acc_, d_q_trial_);
        for (int d = 0; d < 3; ++d) dipole_[d] = dipole_trial_[d];
        if (pending_cm_group_ >= 0) {
            for (int d = 0; d < 3; ++d) g_cm_[static_cast<size_t>(3 * pending_cm_group_ + d)] = pending_cm_[d];
            check_hip(hipMemcpy(d_gcm_ + 3 * pending_cm_group_, pending_cm_, sizeof(double) * 3, hipMemcpyHostToDevice),
                      "copy mass centre");
        }
    }
    pending_ = false;
    pending_cm_group_ = -1;
    staged_cm_group_ = -1;
    staged_.clear();
}

double Engine::reciprocal_energy_host(const std::vector<double2>& q) const {
    std::vector<double> ak(static_cast<size_t>(nk_));
    check_hip(hipMemcpy(ak.data(), d_ak_, sizeof(double) * nk_, hipMemcpyDeviceToHost), "copy ak");
    double e = 0;
    for (int k = 0; k < nk_; ++k) e += ak[static_cast<size_t>(k)] * (q[static_cast<size_t>(k)].x * q[static_cast<size_t>(k)].x + q[static_cast<size_t>(k)].y * q[static_cast<size_t>(k)].y);
    const double V = model_.box[0] * model_.box[1] * model_.box[2];
    return 2 * kPi * e * model_.ewald.lB / V;
}

void Engine::total_energy(double* terms) {
    check_hip(hipStreamSynchronize(stream_), "sync");
    int cells_overflow = 0;
    if (cells_.enabled) {
        check_hip(hipMemcpy(&cells_overflow, cells_.overflow, sizeof(int), hipMemcpyDeviceToHost), "copy overflow");
        if (cells_overflow) throw std::runtime_error("cell list overflow: a cell exceeded its capacity");
    }
    double self = 0;
    if (model_.has_self_term) {
        for (size_t g = 0; g < g_begin_.size(); ++g) {
            double gs = 0;
            for (int i = 0; i < g_size_[g]; ++i) gs += self_energy_of(g_begin_[g] + i);
            self += gs;
        }
    }
    // nonbonded: all pairs (GroupPairing::all, energy.h:1290-1299)
    const int blocks = (n_ + kBlock - 1) / kBlock;
    const int bond_blocks = nbonds_ > 0 ? std::min(64, (nbonds_ + kBlock - 1) / kBlock) : 0;
    double* d_part = nullptr;
    check_hip(hipMalloc(&d_part, sizeof(double) * (blocks + bond_blocks)), "hipMalloc partials");
    GroupTable gt{d_gflags_, d_gcm_, d_g2g_, static_cast<int>(model_.molecules.size()), d_excl_off_, d_excl_};
    if (any_molecular_)
        hipLaunchKernelGGL((all_pairs_kernel<true>), dim3(blocks), dim3(kBlock), 0, stream_, d_pos_, d_meta_, n_,
                           d_pairs_, model_.ntypes, d_coul_, make_box(model_), gt, d_part);
    else
        hipLaunchKernelGGL((all_pairs_kernel<false>), dim3(blocks), dim3(kBlock), 0, stream_, d_pos_, d_meta_, n_,
                           d_pairs_, model_.ntypes, d_coul_, make_box(model_), gt, d_part);
    check_hip(hipGetLastError(), "all_pairs_kernel");
    if (bond_blocks > 0) {
        hipLaunchKernelGGL(bond_total_kernel, dim3(bond_blocks), dim3(kBlock), 0, stream_, d_pos_, d_bond_all_, nbonds_,
                           make_box(model_), d_part + blocks);
        check_hip(hipGetLastError(), "bond_total_kernel");
    }
    std::vector<double> part(static_cast<size_t>(blocks + bond_blocks));
    check_hip(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, stream_),
              "copy partials");
    check_hip(hipStreamSynchronize(stream_), "sync");
    (void)hipFree(d_part);
    double pair = 0, bond = 0;
    for (int b = 0; b < blocks; ++b) pair += part[static_cast<size_t>(b)];
    for (int b = 0; b < bond_blocks; ++b) bond += part[static_cast<size_t>(blocks + b)];
    double ewald = 0;
    if (model_.ewald.present) {
        std::vector<double2> q(static_cast<size_t>(nk_));
        check_hip(hipMemcpy(q.data(), d_q_acc_, sizeof(double2) * nk_, hipMemcpyDeviceToHost), "copy q");
        ewald = surface_energy(dipole_) + reciprocal_energy_host(q);
    }
    bool stop = false;
    for (size_t t = 0; t < model_.term_names.size(); ++t) {
        const std::string& name = model_.term_names[t];
        double v = 0.0;
        if (name == "selfenergy") v = self;
        else if (name == "nonbonded") v = pair;
        else if (name == "ewald") v = ewald;
        else if (name == "bonded") v = bond;
        terms[t] = stop ? 0.0 : v;
        if (!stop && (v >= model_.maximum_allowed_energy || std::isnan(v))) stop = true;
    }
}

void Engine::check_q(double* max_abs_dq, double* rel_rec) {
    *max_abs_dq = 0;
    *rel_rec = 0;
    if (!model_.ewald.present) return;
    double2* d_tmp = nullptr;
    check_hip(hipMalloc(&d_tmp, sizeof(double2) * nk_), "hipMalloc q tmp");
    full_q(d_tmp);
    std::vector<double2> qf(static_cast<size_t>(nk_)), qr(static_cast<size_t>(nk_));
    check_hip(hipMemcpyAsync(qf.data(), d_tmp, sizeof(double2) * nk_, hipMemcpyDeviceToHost, stream_), "copy");
    check_hip(hipMemcpyAsync(qr.data(), d_q_acc_, sizeof(double2) * nk_, hipMemcpyDeviceToHost, stream_), "copy");
    check_hip(hipStreamSynchronize(stream_), "sync");
    (void)hipFree(d_tmp);
    double mx = 0;
    for (int k = 0; k < nk_; ++k)
        mx = std::max(mx, std::max(std::fabs(qf[static_cast<size_t>(k)].x - qr[static_cast<size_t>(k)].x),
                                   std::fabs(qf[static_cast<size_t>(k)].y - qr[static_cast<size_t>(k)].y)));
    const double ef = reciprocal_energy_host(qf), er = reciprocal_energy_host(qr);
    *max_abs_dq = mx;
    *rel_rec = (er - ef) / (std::fabs(ef) > 0 ? std::fabs(ef) : 1.0);
}

void Engine::get_positions(double* xyz) const { std::copy(h_xyz_.begin(), h_xyz_.end(), xyz); }

void Engine::reset_positions(const double* xyz, const double* mass_center) {
    if (!d_pos_) throw std::runtime_error("reset_positions: engine not initialised");
    if (pending_) throw std::runtime_error("reset_positions: a move is pending (sync first)");
    check_hip(hipStreamSynchronize(stream_), "sync");
    h_xyz_.assign(xyz, xyz + 3 * static_cast<size_t>(n_));
    std::vector<double4> p(static_cast<size_t>(n_));
    for (int i = 0; i < n_; ++i)
        p[static_cast<size_t>(i)] = make_double4(h_xyz_[3 * i], h_xyz_[3 * i + 1], h_xyz_[3 * i + 2], h_q_[static_cast<size_t>(i)]);
    check_hip(hipMemcpy(d_pos_, p.data(), sizeof(double4) * n_, hipMemcpyHostToDevice), "copy pos");
    const int ng = static_cast<int>(g_begin_.size());
    if (mass_center) g_cm_.assign(mass_center, mass_center + 3 * static_cast<size_t>(ng));
    else
        for (int g = 0; g < ng; ++g)
            if (!model_.molecules[static_cast<size_t>(g_molid_[static_cast<size_t>(g)])].atomic && g_size_[static_cast<size_t>(g)] > 0)
                group_mass_center_host(g, &g_cm_[static_cast<size_t>(3 * g)]);
    check_hip(hipMemcpy(d_gcm_, g_cm_.data(), sizeof(double) * 3 * ng, hipMemcpyHostToDevice), "copy group centres");
    for (int d = 0; d < 3; ++d) {
        dipole_[d] = 0;
        for (int i = 0; i < n_; ++i) dipole_[d] += h_xyz_[3 * i + d] * h_q_[static_cast<size_t>(i)];
    }
    if (model_.ewald.present) {
        full_q(d_q_acc_);
        check_hip(hipStreamSynchronize(stream_), "full Q");
    }
    (void)hipFree(cells_.count);
    (void)hipFree(cells_.members);
    (void)hipFree(cells_.cell_of);
    (void)hipFree(cells_.overflow);
    cells_ = CellList{};
    cell_capacity_ = 0; // a fixed capacity applies to the first build only
    build_cell_list();
}

void Engine::refresh_host_positions() {
    std::vector<double4> p(static_cast<size_t>(n_));
    check_hip(hipStreamSynchronize(stream_), "sync");
    check_hip(hipMemcpy(p.data(), d_pos_, sizeof(double4) * n_, hipMemcpyDeviceToHost), "copy pos");
    for (int i = 0; i < n_; ++i) {
        h_xyz_[3 * i] = p[static_cast<size_t>(i)].x;
        h_xyz_[3 * i + 1] = p[static_cast<size_t>(i)].y;
        h_xyz_[3 * i + 2] = p[static_cast<size_t>(i)].z;
    }
    for (int d = 0; d < 3; ++d) {
        dipole_[d] = 0;
        for (int i = 0; i < n_; ++i) dipole_[d] += h_xyz_[3 * i + d] * h_q_[static_cast<size_t>(i)];
    }
}

} // namespace fmc
