// One large replica over many workgroups with the per-move exchange taken off the
// critical path.  Move t is evaluated while the decisions of up to D earlier moves
// ("pending" moves) are still unknown, and published with the terms that make the
// later decision exact:
//   Ewald:  dE_t = a_t + sum_d acc_{t-d} c_{t,d},   a_t = sum_k A_k(2Re(conj(Q) dQ_t) + |dQ_t|^2)
//           with Q holding every decided move, c_{t,d} = sum_k A_k 2Re(conj(dQ_{t-d}) dQ_t);
//   pairs:  sums over the committed configuration without the pending particles, plus
//           the pair with each pending particle at its old and at its new position.
// Decisions are taken strictly in move order, identically in every workgroup.  A
// proposal that picks a pending particle first drains the pipeline.  The Markov
// chain is exactly that of one-move-at-a-time Metropolis.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "device_math.hpp"
#include "mc_batch.hpp"
#include "mc_common.hpp"
#include "philox.hpp"

namespace fmc {
namespace {

constexpr int kPipeMaxKPT = 6;
constexpr int kPipeMaxG = 64;
constexpr int kPipeGk = 1025;
constexpr int kPipeMaxD = 4;
constexpr int kPipeVals = 3 + 5 * kPipeMaxD; // a, c[D], pn0, po0, alt[4D]
constexpr int kHist = 4 * kPipeMaxD + 4;

struct alignas(64) PipeSlot {
    double v[kPipeVals];
    unsigned long long seq;
};

struct Pending {
    int idx, type, depth, pad; // depth = pending moves when it was evaluated
    double ox, oy, oz, q;
    double nx, ny, nz, accept_u;
    unsigned long long mid;
};

__device__ __forceinline__ int pcell(double x, double half, double inv_cell, int nc) {
    int c = static_cast<int>(floor((x + half) * inv_cell));
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

__device__ __forceinline__ double pwrap(double x, double L, double half, double inv) {
    if (fabs(x) > half) {
        const double s = x * inv;
        x -= L * static_cast<double>(static_cast<int>(s > 0.0 ? s + 0.5 : s - 0.5));
    }
    return x;
}

template <int D, bool kCells>
__global__ __launch_bounds__(256) void mc_pipe_kernel(ReplicaState* __restrict__ states, long long M, int G,
                                                      PipeSlot* __restrict__ slots, int R, int* __restrict__ err,
                                                      TraceEntry* __restrict__ trace, int trace_cap,
                                                      long long spin_limit) {
    const int g = blockIdx.x;
    if (g >= G) return;
    ReplicaState& S = states[0];
    if (S.error) return; // a stopped replica (read before any workgroup of this launch can set it)
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevCoulomb s_coul;
    __shared__ double s_dp[kMaxTypes];
    __shared__ double s_gk[kPipeGk];
    __shared__ double s_red[3 + kPipeMaxD][4];
    __shared__ double s_alt[4 * kPipeMaxD];
    __shared__ double s_vals[kPipeMaxG][kPipeVals];
    __shared__ double2 s_phase[6 * 65];
    __shared__ Pending s_pend[D + 1];
    __shared__ int s_hist_idx[kHist];
    __shared__ double s_hist_pos[kHist][3];
    __shared__ int s_hist_n;
    __shared__ unsigned char s_acc_ring[64];
    __shared__ int s_accept, s_abort, s_p;

    const int tid = threadIdx.x;
    const int n = S.n, nk = S.nk, ncut = S.ncut, nt = S.ntypes;
    const int width = 2 * ncut + 1;
    const double L = S.L, half = S.half, inv = S.inv;
    const Box box{{L, L, L}, {half, half, half}};
    for (int i = tid; i < nt * nt; i += 256) s_pairs[i] = S.pairs[i];
    lds_copy(s_coul, S.coul, tid, 256);
    if (tid == 0) {
        s_abort = 0;
        s_p = 0;
        s_hist_n = 0;
    }
    for (int i = tid; i < nt; i += 256) s_dp[i] = S.dp_type[i];
    for (int i = tid; i <= ncut * ncut && i < kPipeGk; i += 256) s_gk[i] = S.gk[i];
    for (int i = tid; i < kHist; i += 256) s_hist_idx[i] = -1;

    const int kw = (nk + G - 1) / G;
    const int kb = g * kw;
    const int ke = min(nk, kb + kw);
    int kn[kPipeMaxKPT];
    double A[kPipeMaxKPT];
    double2 q[kPipeMaxKPT], dq[kPipeMaxKPT], dqh[D + 1][kPipeMaxKPT]; // pending moves, oldest first (D + 1 before a decision);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kPipeMaxKPT; ++s) {
        const int k = kb + s * 256 + tid;
        const bool on = k < ke;
        kn[s] = on ? S.kn[k] : 0;
        q[s] = on ? S.q[k] : make_double2(0.0, 0.0);
        const int nx = unpack_n(kn[s], 0), ny = unpack_n(kn[s], 1), nz = unpack_n(kn[s], 2);
        A[s] = on ? (nx > 0 ? 2.0 : 1.0) * s_gk[nx * nx + ny * ny + nz * nz] : 0.0;
        dq[s] = make_double2(0.0, 0.0);
#pragma unroll
        for (int d = 0; d <= D; ++d) dqh[d][s] = make_double2(0.0, 0.0);
    }
    double4* __restrict__ pos = S.pos;
    const uint32_t* __restrict__ meta = S.meta;
    const unsigned long long move0 = S.moves;
    const double beta = S.beta, pref = S.rec_prefactor;
    double energy = S.energy, sum_du = S.sum_du;
    unsigned long long accepted = S.accepted;
    const unsigned long long seed = S.seed;
    const uint32_t replica = static_cast<uint32_t>(S.replica_id);

    // Decide the oldest pending move (all workgroups), apply it, drop it from the pipeline.
    auto decide_oldest = [&]() -> bool {
        const Pending pe = s_pend[0];
        const unsigned long long tag = pe.mid + 1ull;
        PipeSlot* base = slots + static_cast<size_t>(pe.mid % static_cast<unsigned long long>(R)) * G;
        if (tid < 64) {
            bool ok = true;
            if (tid < G) {
                long long spins = 0;
                while (__hip_atomic_load(&base[tid].seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != tag) {
                    if (++spins > spin_limit) {
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                // the lane that acquired the slot reads its values
                if (ok)
                    for (int v = 0; v < kPipeVals; ++v) s_vals[tid][v] = base[tid].v[v];
            }
            if (!ok) {
                atomicExch(err, 1);
                atomicMax(&S.error, 2);
                s_abort = 1;
            }
        }
        __syncthreads();
        if (s_abort) return false;
        if (tid == 0) {
            // Ewald: sum over workgroups (fixed order) of a + sum_d acc_{mid-d} c_d
            double sde = 0.0;
            for (int w = 0; w < G; ++w) {
                double x = s_vals[w][0];
                for (int d = 1; d <= pe.depth; ++d)
                    if (s_acc_ring[(pe.mid - d) & 63]) x += s_vals[w][d];
                sde += x;
            }
            double pn = s_vals[0][1 + D], po = s_vals[0][2 + D];
            for (int d = 1; d <= pe.depth; ++d) {
                const int b = 3 + D + 4 * (d - 1);
                const bool a = s_acc_ring[(pe.mid - d) & 63] != 0;
                pn += a ? s_vals[0][b + 1] : s_vals[0][b + 0];
                po += a ? s_vals[0][b + 3] : s_vals[0][b + 2];
            }
            const double drec = pref * sde;
            const double e_tot_new = pn + drec;
            double du;
            if (isnan(po) && !isnan(e_tot_new)) du = -__builtin_huge_val();
            else if (isnan(e_tot_new)) du = __builtin_huge_val();
            else if (e_tot_new > 0.0 && isinf(e_tot_new)) du = __builtin_huge_val();
            else {
                du = (pn - po) + drec;
                if (isnan(du)) du = 0.0;
            }
            const double bdu = beta * du;
            bool acc;
            if (isinf(bdu) && bdu < 0.0) acc = true;
            else if (-bdu > 709.782712893384) acc = true;
            else acc = pe.accept_u <= exp(-bdu);
            s_accept = acc ? 1 : 0;
            s_acc_ring[pe.mid & 63] = acc ? 1 : 0;
            if (acc) {
                const int h = s_hist_n % kHist;
                s_hist_idx[h] = pe.idx;
                s_hist_pos[h][0] = pe.nx;
                s_hist_pos[h][1] = pe.ny;
                s_hist_pos[h][2] = pe.nz;
                s_hist_n = s_hist_n + 1;
            }
            if (g == 0) {
                const long long t = static_cast<long long>(pe.mid - move0);
                if (trace && t < trace_cap) {
                    TraceEntry& T = trace[t];
                    T.index = pe.idx;
                    T.accepted = acc ? 1 : 0;
                    T.du = du;
                    T.x = pe.nx;
                    T.y = pe.ny;
                    T.z = pe.nz;
                }
                if (acc) {
                    energy += du;
                    sum_du += du;
                    ++accepted;
                    pos[pe.idx] = make_double4(pe.nx, pe.ny, pe.nz, pe.q);
                    if (kCells) {
                        const int cnew = (pcell(pe.nz, half, S.inv_cell, S.nc2) * S.nc1 +
                                          pcell(pe.ny, half, S.inv_cell, S.nc1)) * S.nc0 +
                                         pcell(pe.nx, half, S.inv_cell, S.nc0);
                        const int cold = S.cell_of[pe.idx];
                        if (cnew != cold) {
                            const int nn = S.cell_count[cnew];
                            if (nn >= S.cap) {
                                // full cell: the lists are left as they are.  The moves in flight
                                // were published by every workgroup, so the launch runs to its
                                // end and the replica is marked inconsistent (refuses later runs).
                                atomicMax(&S.error, 2);
                            } else {
                                const int nold = S.cell_count[cold];
                                for (int kk = 0; kk < nold; ++kk)
                                    if (S.cell_members[cold * S.cap + kk] == pe.idx) {
                                        S.cell_members[cold * S.cap + kk] = S.cell_members[cold * S.cap + nold - 1];
                                        break;
                                    }
                                S.cell_count[cold] = nold - 1;
                                S.cell_members[cnew * S.cap + nn] = pe.idx;
                                S.cell_count[cnew] = nn + 1;
                                S.cell_of[pe.idx] = cnew;
                            }
                        }
                    }
                }
            }
            // drop the decided move from the pending list
            const int p = s_p;
            for (int j = 0; j + 1 < p; ++j) s_pend[j] = s_pend[j + 1];
            s_p = p - 1;
        }
        __syncthreads();
        const bool acc = s_accept != 0;
#pragma unroll
        for (int s = 0; s < kPipeMaxKPT; ++s) {
            if (acc) {
                q[s].x += dqh[0][s].x;
                q[s].y += dqh[0][s].y;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) dqh[d][s] = dqh[d + 1][s];
        }
        return true;
    };

    bool alive = true;
    for (long long t = 0; t < M && alive; ++t) {
        const unsigned long long mid = move0 + static_cast<unsigned long long>(t);
        const Proposal P = make_proposal(seed, replica, mid, n);
        const int i = P.index;
        // a pending move of the same particle: decide everything first
        bool clash = false;
        for (int j = 0; j < s_p; ++j) clash |= (s_pend[j].idx == i);
        if (clash) {
            while (s_p > 0 && alive) alive = decide_oldest();
            if (!alive) break;
        }
        const int p = s_p;
        // old position: newest accepted decision of this particle still in the history, else memory
        double4 po = pos[i];
        {
            const int hn = s_hist_n;
            for (int h = hn - 1; h >= 0 && h >= hn - kHist; --h) {
                const int slot = h % kHist;
                if (s_hist_idx[slot] == i) {
                    po.x = s_hist_pos[slot][0];
                    po.y = s_hist_pos[slot][1];
                    po.z = s_hist_pos[slot][2];
                    break;
                }
            }
        }
        const int ti = static_cast<int>(meta[i] & 0xffu);
        const double scale = s_dp[ti] * P.length_u;
        const double nx_ = pwrap(po.x + P.dir[0] * scale, L, half, inv);
        const double ny_ = pwrap(po.y + P.dir[1] * scale, L, half, inv);
        const double nz_ = pwrap(po.z + P.dir[2] * scale, L, half, inv);
        for (int e = tid; e < 6 * width; e += 256) {
            const int nn = e % width - ncut;
            const int axis = (e / width) % 3;
            const int st = e / (3 * width);
            const double x = st == 0 ? (axis == 0 ? nx_ : (axis == 1 ? ny_ : nz_))
                                     : (axis == 0 ? po.x : (axis == 1 ? po.y : po.z));
            double sn, cs;
            sincos((6.283185307179586 * (static_cast<double>(nn) / L)) * x, &sn, &cs);
            s_phase[e] = make_double2(cs, sn);
        }
        // pair sums (workgroup 0): committed configuration without i and the pending particles
        double acc_a = 0.0, acc_pn = 0.0, acc_po = 0.0;
        double acc_c[D];
#pragma unroll
        for (int jj = 0; jj < D; ++jj) acc_c[jj] = 0.0;
        if (g == 0) {
            auto pending_particle = [&](int j) {
                bool hit = false;
                for (int k = 0; k < p; ++k) hit |= (s_pend[k].idx == j);
                return hit;
            };
            if (kCells) {
                const int combo = tid >> 2, slot0 = tid & 3;
                if (combo < 54) {
                    const int st = combo / 27, c = combo % 27;
                    const double px = st == 0 ? nx_ : po.x, py = st == 0 ? ny_ : po.y, pz = st == 0 ? nz_ : po.z;
                    const int cx = (pcell(px, half, S.inv_cell, S.nc0) + (c % 3) - 1 + S.nc0) % S.nc0;
                    const int cy = (pcell(py, half, S.inv_cell, S.nc1) + ((c / 3) % 3) - 1 + S.nc1) % S.nc1;
                    const int cz = (pcell(pz, half, S.inv_cell, S.nc2) + (c / 9) - 1 + S.nc2) % S.nc2;
                    const int cell = (cz * S.nc1 + cy) * S.nc0 + cx;
                    const int cnt = S.cell_count[cell];
                    double sum = 0.0;
                    for (int k = slot0; k < cnt; k += 4) {
                        const int j = S.cell_members[cell * S.cap + k];
                        if (j == i || pending_particle(j)) continue;
                        const double4 pj = pos[j];
                        sum += pair_energy(s_pairs[ti * nt + static_cast<int>(meta[j] & 0xffu)], s_coul, po.w, pj.w,
                                           dist2(box, px, py, pz, pj.x, pj.y, pj.z));
                    }
                    if (st == 0) acc_pn = sum;
                    else acc_po = sum;
                }
            } else {
                for (int j = tid; j < n; j += 256) {
                    if (j == i || pending_particle(j)) continue;
                    const double4 pj = pos[j];
                    const DevPair& pp = s_pairs[ti * nt + static_cast<int>(meta[j] & 0xffu)];
                    acc_pn += pair_energy(pp, s_coul, po.w, pj.w, dist2(box, nx_, ny_, nz_, pj.x, pj.y, pj.z));
                    acc_po += pair_energy(pp, s_coul, po.w, pj.w, dist2(box, po.x, po.y, po.z, pj.x, pj.y, pj.z));
                }
            }
            // the pair with each pending particle (position jj, oldest first) with this
            // particle new/old and the pending one old/new:
            // s_alt[4 jj + {0,1,2,3}] = {new-vs-old, new-vs-new, old-vs-old, old-vs-new}
            if (tid < 4 * D) {
                double u = 0.0;
                if (tid < 4 * p) {
                    const int jj = tid >> 2, which = tid & 3;
                    const Pending& pe = s_pend[jj];
                    const bool me_new = which < 2, other_new = (which & 1) != 0;
                    const double ax = me_new ? nx_ : po.x, ay = me_new ? ny_ : po.y, az = me_new ? nz_ : po.z;
                    const double bx = other_new ? pe.nx : pe.ox, by = other_new ? pe.ny : pe.oy,
                                 bz = other_new ? pe.nz : pe.oz;
                    u = pair_energy(s_pairs[ti * nt + pe.type], s_coul, po.w, pe.q, dist2(box, ax, ay, az, bx, by, bz));
                }
                s_alt[tid] = u;
            }
        }
        __syncthreads(); // phase tables
#pragma unroll
        for (int s = 0; s < kPipeMaxKPT; ++s) {
            const int k = kb + s * 256 + tid;
            if (k < ke) {
                const int nx = unpack_n(kn[s], 0), ny = unpack_n(kn[s], 1), nz = unpack_n(kn[s], 2);
                const double2 xn = s_phase[nx + ncut], yn = s_phase[width + ny + ncut], zn = s_phase[2 * width + nz + ncut];
                const double2 xo = s_phase[3 * width + nx + ncut], yo = s_phase[4 * width + ny + ncut],
                              zo = s_phase[5 * width + nz + ncut];
                const double an = xn.x * yn.x - xn.y * yn.y, bn = xn.x * yn.y + xn.y * yn.x;
                const double ao = xo.x * yo.x - xo.y * yo.y, bo = xo.x * yo.y + xo.y * yo.x;
                const double dre = po.w * ((an * zn.x - bn * zn.y) - (ao * zo.x - bo * zo.y));
                const double dim = po.w * ((an * zn.y + bn * zn.x) - (ao * zo.y + bo * zo.x));
                dq[s] = make_double2(dre, dim);
                acc_a += A[s] * (2.0 * (q[s].x * dre + q[s].y * dim) + (dre * dre + dim * dim));
#pragma unroll
                for (int jj = 0; jj < D; ++jj) acc_c[jj] += A[s] * 2.0 * (dqh[jj][s].x * dre + dqh[jj][s].y * dim);
            } else {
                dq[s] = make_double2(0.0, 0.0);
            }
        }
        // block reduction: [a, pn, po, c(position 0..D-1)]
        {
            const double ra = wave_sum(acc_a), rpn = wave_sum(acc_pn), rpo = wave_sum(acc_po);
            double rc[D];
#pragma unroll
            for (int jj = 0; jj < D; ++jj) rc[jj] = wave_sum(acc_c[jj]);
            if ((tid & 63) == 0) {
                s_red[0][tid >> 6] = ra;
                s_red[1][tid >> 6] = rpn;
                s_red[2][tid >> 6] = rpo;
#pragma unroll
                for (int jj = 0; jj < D; ++jj) s_red[3 + jj][tid >> 6] = rc[jj];
            }
        }
        __syncthreads();
        PipeSlot* my = slots + static_cast<size_t>(mid % static_cast<unsigned long long>(R)) * G + g;
        if (tid == 0) {
            auto red = [&](int v) { return ((s_red[v][0] + s_red[v][1]) + s_red[v][2]) + s_red[v][3]; };
            // slot layout by distance back d = p - position: [a, c_1..c_D, pn0, po0, alt(d=1..D) x 4]
            for (int v = 0; v < kPipeVals; ++v) my->v[v] = 0.0;
            my->v[0] = red(0);
            my->v[1 + D] = red(1);
            my->v[2 + D] = red(2);
            for (int jj = 0; jj < p; ++jj) {
                const int d = p - jj;
                my->v[d] = red(3 + jj);
                if (g == 0)
                    for (int w = 0; w < 4; ++w) my->v[3 + D + 4 * (d - 1) + w] = s_alt[4 * jj + w];
            }
            __hip_atomic_store(&my->seq, mid + 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            Pending& pe = s_pend[p];
            pe.idx = i;
            pe.type = ti;
            pe.depth = p;
            pe.ox = po.x;
            pe.oy = po.y;
            pe.oz = po.z;
            pe.q = po.w;
            pe.nx = nx_;
            pe.ny = ny_;
            pe.nz = nz_;
            pe.accept_u = P.accept_u;
            pe.mid = mid;
            s_p = p + 1;
        }
#pragma unroll
        for (int s = 0; s < kPipeMaxKPT; ++s)
#pragma unroll
            for (int jj = 0; jj <= D; ++jj)
                if (jj == p) dqh[jj][s] = dq[s];
        __syncthreads();
        if (p + 1 > D) alive = decide_oldest(); // keep at most D moves pending
    }
    while (alive && s_p > 0) alive = decide_oldest();
#pragma unroll
    for (int s = 0; s < kPipeMaxKPT; ++s) {
        const int k = kb + s * 256 + tid;
        if (k < ke) S.q[k] = q[s];
    }
    if (g == 0 && tid == 0) {
        S.moves = move0 + static_cast<unsigned long long>(M);
        S.accepted = accepted;
        S.energy = energy;
        S.sum_du = sum_du;
    }
}

template <int D>
const void* pipe_fn(bool cells) {
    return cells ? reinterpret_cast<const void*>(mc_pipe_kernel<D, true>)
                 : reinterpret_cast<const void*>(mc_pipe_kernel<D, false>);
}

} // namespace

void Batch::run_pipe(long long moves, int groups, int depth, bool trace, int trace_cap) {
    if (!finalized_) finalize();
    if (size() != 1) throw std::runtime_error("run_pipe advances a batch of one replica");
    if (general_ == 1) throw std::runtime_error("run_pipe handles translate-only replicas");
    if (h_states_[0].surf != 0.0) throw std::runtime_error("run_pipe: no surface term (epss = 0)");
    if (h_states_[0].sample_every > 0) throw std::runtime_error("run_pipe takes no energy samples");
    const ReplicaState& s = h_states_[0];
    if (groups < 1 || groups > kPipeMaxG) throw std::runtime_error("run_pipe: 1..64 workgroups");
    if (depth < 1 || depth > kPipeMaxD) throw std::runtime_error("run_pipe: depth 1..4");
    const int kw = (s.nk + groups - 1) / groups;
    if ((kw + 255) / 256 > kPipeMaxKPT) throw std::runtime_error("run_pipe: too many k-vectors per workgroup");
    if (2 * s.ncut + 1 > 65 || s.ncut * s.ncut >= kPipeGk) throw std::runtime_error("run_pipe: ncutoff too large");
    const void* fn = nullptr;
    switch (depth) {
    case 1: fn = pipe_fn<1>(s.cells != 0); break;
    case 2: fn = pipe_fn<2>(s.cells != 0); break;
    case 3: fn = pipe_fn<3>(s.cells != 0); break;
    default: fn = pipe_fn<4>(s.cells != 0); break;
    }
    hipDeviceProp_t prop;
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    int per_cu = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0), "occupancy");
    if (per_cu * prop.multiProcessorCount < groups) throw std::runtime_error("run_pipe: grid cannot be co-resident");
    const int R = 4 * kPipeMaxD + 4;
    if (!pipe_slots_) {
        check_hip(hipMalloc(&pipe_slots_, sizeof(PipeSlot) * R * kPipeMaxG), "hipMalloc slots");
        check_hip(hipMemset(pipe_slots_, 0, sizeof(PipeSlot) * R * kPipeMaxG), "memset slots");
        if (!split_err_) check_hip(hipMalloc(&split_err_, sizeof(int) * 2), "hipMalloc err");
    }
    // every launch starts from zeroed slots (an earlier launch's tags would match a
    // resumed, rewound chain) and a clear time-out flag
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(256), 0, stream_, static_cast<unsigned long long*>(pipe_slots_),
                       static_cast<int>(sizeof(PipeSlot) * R * kPipeMaxG / 8));
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream_, reinterpret_cast<unsigned long long*>(split_err_), 1);
    check_hip(hipGetLastError(), "zero_words_kernel");
    if (trace && trace_cap > trace_cap_) {
        (void)hipFree(d_trace_);
        check_hip(hipMalloc(&d_trace_, sizeof(TraceEntry) * static_cast<size_t>(trace_cap)), "hipMalloc trace");
        trace_cap_ = trace_cap;
    }
    ReplicaState* st = d_states_;
    long long M = moves;
    int G = groups, RR = R;
    PipeSlot* slots = static_cast<PipeSlot*>(pipe_slots_);
    int* err = split_err_;
    TraceEntry* tr = trace ? d_trace_ : nullptr;
    int cap = trace ? trace_cap : 0;
    long long spin = 1LL << 24;
    void* args[] = {&st, &M, &G, &slots, &RR, &err, &tr, &cap, &spin};
    time_begin();
    check_hip(hipLaunchCooperativeKernel(fn, dim3(groups), dim3(256), args, 0, stream_), "hipLaunchCooperativeKernel");
    time_end();
}

} // namespace fmc
