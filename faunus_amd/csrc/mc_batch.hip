// Device-resident Metropolis kernels: one workgroup advances one replica through
// many trial moves (the reference's MetropolisMonteCarlo::sweep/performMove loop,
// montecarlo.cpp:139-227, with AtomicTranslateRotate proposals, move.cpp:225-293),
// computing the nonbonded pair dU and the Ewald reciprocal dU in the same pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "device_math.hpp"
#include "mc_batch.hpp"
#include "mc_common.hpp"
#include "philox.hpp"

namespace fmc {
namespace {

// One round of neighbour exchanges in one ladder (ParallelTempering, move.cpp:844-965):
// rungs k, k+1 for k of the round's parity, accepted with min(1, exp((b_a - b_b)(E_a -
// E_b))) on a Philox deviate of (round, ladder, k); an accepted swap exchanges the betas
// and the rung slots.
__device__ void swap_ladder_rungs(ReplicaState* states, int b, int m, int* rung_slot, int l, unsigned long long seed,
                                  unsigned long long round, unsigned long long* counters) {
    for (int k = static_cast<int>(round & 1); k + 1 < m; k += 2) {
        const int sa = rung_slot[b + k], sb = rung_slot[b + k + 1];
        ReplicaState& A = states[sa];
        ReplicaState& B = states[sb];
        const double x = (A.beta - B.beta) * (A.energy - B.energy);
        const Philox4 r = philox4x32_10(static_cast<uint32_t>(round), static_cast<uint32_t>(round >> 32),
                                        static_cast<uint32_t>(l), 0x5a5a0000u + static_cast<uint32_t>(k),
                                        static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
        const double u = (static_cast<double>(r.v[0] >> 5) * 67108864.0 + static_cast<double>(r.v[1] >> 6)) *
                         1.1102230246251565e-16;
        atomicAdd(&counters[0], 1ull);
        if (x >= 0.0 || u < exp(x)) {
            const double tb = A.beta;
            A.beta = B.beta;
            B.beta = tb;
            rung_slot[b + k] = sb;
            rung_slot[b + k + 1] = sa;
            atomicAdd(&counters[1], 1ull);
        }
    }
}

template <int kThreads, int kMaxKPT, bool kCells, bool kF32>
__global__ __launch_bounds__(kThreads) void mc_advance_kernel(ReplicaState* __restrict__ states,
                                                               const long long* __restrict__ moves,
                                                               TraceEntry* __restrict__ trace, int trace_cap,
                                                               const LadderRun* __restrict__ ladders) {
    ReplicaState& S = states[blockIdx.x];
    const long long M = moves[blockIdx.x];
    // free-running ladders: every replica takes part in its ladder's barriers, also one
    // with no moves or an error
    const int n_iv = ladders != nullptr ? ladders->intervals[blockIdx.x] : 1;
    if (ladders == nullptr && (M <= 0 || S.error)) return;
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevPairF s_pairs_f[kF32 ? kMaxTypes * kMaxTypes : 1];
    __shared__ float4 s_stab[kF32 ? kSTable : 1];
    __shared__ DevCoulomb s_coul;
    __shared__ double s_dp[kMaxTypes];
    __shared__ double s_gk[kMaxGk];
    __shared__ double s_red[3][kThreads / 64];
    __shared__ int s_accept;
    __shared__ int s_nhot;                        // hot pairs queued this move (FP32 form)
    __shared__ int2 s_hot[kF32 ? kHotCap : 1];
    __shared__ Proposal s_prop[64]; // the next 64 proposals, drawn one per lane of wave 0
    extern __shared__ int s_dyn[]; // kn[nk] followed by the phase tables (double2)

    const int tid = threadIdx.x;
    const int n = S.n, nk = S.nk, ncut = S.ncut, nt = S.ntypes;
    const int width = 2 * ncut + 1;
    const int kpt = (nk + kThreads - 1) / kThreads;
    const double L = S.L, half = S.half, inv = S.inv;
    int* s_kn = s_dyn;
    double2* s_phase = reinterpret_cast<double2*>(s_dyn + ((nk + 3) & ~3));

    for (int i = tid; i < nt * nt; i += kThreads) {
        s_pairs[i] = S.pairs[i];
        if (kF32) {
            const DevPair& p = S.pairs[i];
            s_pairs_f[i] = to_f32(p);
        }
    }
    lds_copy(s_coul, S.coul, tid, kThreads);
    for (int i = tid; i < nt; i += kThreads) s_dp[i] = S.dp_type[i];
    for (int i = tid; i <= ncut * ncut && i < kMaxGk; i += kThreads) s_gk[i] = S.gk[i];
    for (int k = tid; k < nk; k += kThreads) s_kn[k] = S.kn[k];
    __syncthreads();
    if (kF32) {
        if (S.stab)
            for (int i = tid; i < kSTable; i += kThreads) s_stab[i] = S.stab[i];
    }
    const CoulF cf = coul_f32(s_coul);

    // this thread's chunk of k-vectors and their running structure factors
    const int k0 = tid * kpt;
    double2 q[kMaxKPT];
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) q[s] = (s < kpt && k0 + s < nk) ? S.q[k0 + s] : make_double2(0.0, 0.0);

    double4* __restrict__ pos = S.pos;
    FixPos* __restrict__ pos32 = S.pos32;
    const uint32_t* __restrict__ meta = S.meta;
    const float hf = static_cast<float>(L * kFixInv);
    unsigned long long move0 = S.moves;
    double beta = S.beta; // thread 0's (a ladder swap changes it between intervals)
    const double pref = S.rec_prefactor;
    const double surf = S.surf;
    double dip0 = S.dip[0], dip1 = S.dip[1], dip2 = S.dip[2]; // thread 0's running dipole
    double energy = S.energy, sum_du = S.sum_du;
    unsigned long long accepted = S.accepted;
    const int replica = S.replica_id;
    const unsigned long long seed = S.seed;
    const long long sample_every = S.sample_every;
    unsigned long long e_samples = S.e_samples;
    double e_sum = S.e_sum, e_sum2 = S.e_sum2;
    if (tid == 0) s_nhot = 0;
    __syncthreads();

    for (int iv = 0; iv < n_iv; ++iv) {
    const long long Mi = (M > 0 && !S.error) ? M : 0;
    long long t_done = Mi; // moves decided in this interval
    for (long long t = 0; t < Mi; ++t) {
        if ((t & 63) == 0) {
            // the previous proposals were last read before this move's barriers; the
            // Philox blocks and sin/cos of 64 moves overlap instead of forming one serial
            // chain per move
            if (tid < 64)
                s_prop[tid] = make_proposal(seed, static_cast<uint32_t>(replica),
                                            move0 + static_cast<unsigned long long>(t + tid), n);
            __syncthreads();
        }
        const Proposal P = s_prop[t & 63];
        const int i = P.index;
        const double4 po = pos[i];
        const int ti = static_cast<int>(meta[i] & 0xffu);
        const double dp = s_dp[ti];
        const double scale = dp * P.length_u;
        const double nx_ = wrap(po.x + P.dir[0] * scale, L, half, inv);
        const double ny_ = wrap(po.y + P.dir[1] * scale, L, half, inv);
        const double nz_ = wrap(po.z + P.dir[2] * scale, L, half, inv);

        // Ewald phase tables e^{i k_a(n) r_a}, n in [-ncut, ncut], new (st=0) and old (st=1)
        for (int e = tid; e < 6 * width; e += kThreads) {
            const int nn = e % width - ncut;
            const int axis = (e / width) % 3;
            const int st = e / (3 * width);
            const double x = st == 0 ? (axis == 0 ? nx_ : (axis == 1 ? ny_ : nz_))
                                     : (axis == 0 ? po.x : (axis == 1 ? po.y : po.z));
            const double kk = 6.283185307179586 * (static_cast<double>(nn) / L);
            double sn, cs;
            sincos(kk * x, &sn, &cs);
            s_phase[e] = make_double2(cs, sn);
        }

        // nonbonded pair sums of the moved particle in the new and old position
        double e_new = 0.0, e_old = 0.0;
        FixPos fa{}, fb{};
        if (kF32) {
            fa = fixed_pos(nx_, ny_, nz_, po.w, inv);
            fb = fixed_pos(po.x, po.y, po.z, po.w, inv);
        }
        if (!kCells) {
            if (kF32) {
                pair_row_f32<kThreads>(pos32, meta, 0, n, i, -1, true, s_pairs_f + ti * nt, cf, s_stab, S.terms, hf,
                                       fa, fb, &s_nhot, s_hot, 0, e_new, e_old);
            } else {
                const Box b{{L, L, L}, {half, half, half}};
                pair_row_f64<kThreads>(pos, meta, 0, n, i, -1, true, s_pairs + ti * nt, s_coul, S.terms, b, nx_, ny_, nz_,
                                       po.x, po.y, po.z, po.w, e_new, e_old);
            }
        } else {
            // 27-cell stencils around the new (st=0) and old (st=1) position
            const int lanes_per = kThreads / 54 > 0 ? kThreads / 54 : 1;
            const int combo = tid / lanes_per;
            const int slot0 = tid % lanes_per;
            if (combo < 54) {
                const int st = combo / 27, c = combo % 27;
                const double px = st == 0 ? nx_ : po.x, py = st == 0 ? ny_ : po.y, pz = st == 0 ? nz_ : po.z;
                const int cx = (cell_of_coord(px, half, S.inv_cell, S.nc0) + (c % 3) - 1 + S.nc0) % S.nc0;
                const int cy = (cell_of_coord(py, half, S.inv_cell, S.nc1) + ((c / 3) % 3) - 1 + S.nc1) % S.nc1;
                const int cz = (cell_of_coord(pz, half, S.inv_cell, S.nc2) + (c / 9) - 1 + S.nc2) % S.nc2;
                const int cell = (cz * S.nc1 + cy) * S.nc0 + cx;
                const int cnt = S.cell_count[cell];
                for (int k = slot0; k < cnt; k += lanes_per) {
                    const int j = S.cell_members[cell * S.cap + k];
                    if (j == i) continue;
                    const int tj = static_cast<int>(meta[j] & 0xffu);
                    double u;
                    if (kF32) {
                        const FixPos pj = pos32[j];
                        const FixPos& A = st == 0 ? fa : fb;
                        bool hot;
                        const float uf = pair_energy_ft(s_pairs_f[ti * nt + tj], cf, s_stab, S.terms, A.q, pj.q,
                                                        dist2_fix(A, pj, hf), hot);
                        u = (hot && push_hot(&s_nhot, s_hot, j, st)) ? 0.0 : static_cast<double>(uf);
                    } else {
                        const double4 pj = pos[j];
                        Box b{{L, L, L}, {half, half, half}};
                        u = pair_energy(s_pairs[ti * nt + tj], s_coul, po.w, pj.w, dist2(b, px, py, pz, pj.x, pj.y, pj.z));
                    }
                    if (st == 0) e_new += u;
                    else e_old += u;
                }
            }
        }
        __syncthreads(); // phase tables complete (and the hot-pair queue)
        if (kF32) {
            // hot pairs in FP64 from the FP64 coordinates
            const Box b{{L, L, L}, {half, half, half}};
            const int nh = min(s_nhot, kHotCap);
            for (int h = tid; h < nh; h += kThreads) {
                const int2 e = s_hot[h];
                const double4 pj = pos[e.x];
                const int tj = static_cast<int>(meta[e.x] & 0xffu);
                const bool nw = (e.y & 1) == 0;
                const double u = pair_energy_t(s_pairs[ti * nt + tj], s_coul, po.w, pj.w,
                                               dist2(b, nw ? nx_ : po.x, nw ? ny_ : po.y, nw ? nz_ : po.z, pj.x, pj.y, pj.z),
                                               S.terms);
                if (nw) e_new += u;
                else e_old += u;
            }
        }

        // Ewald: dE = sum_k A_k (2 Re(conj(Q) dQ) + |dQ|^2) over this thread's k-vectors
        double de = 0.0;
        {
            int prev_col = -1;
            double2 exy_n = make_double2(1.0, 0.0), exy_o = make_double2(1.0, 0.0);
            const double qi = po.w;
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                const int k = k0 + s;
                if (s < kpt && k < nk) {
                    const int packed = s_kn[k];
                    const int nx = unpack_n(packed, 0), ny = unpack_n(packed, 1), nz = unpack_n(packed, 2);
                    const int col = packed & 0xffff;
                    if (col != prev_col) {
                        const double2 xn = s_phase[nx + ncut], yn = s_phase[width + ny + ncut];
                        const double2 xo = s_phase[3 * width + nx + ncut], yo = s_phase[4 * width + ny + ncut];
                        exy_n = make_double2(xn.x * yn.x - xn.y * yn.y, xn.x * yn.y + xn.y * yn.x);
                        exy_o = make_double2(xo.x * yo.x - xo.y * yo.y, xo.x * yo.y + xo.y * yo.x);
                        prev_col = col;
                    }
                    const double2 zn = s_phase[2 * width + nz + ncut], zo = s_phase[5 * width + nz + ncut];
                    const double en_re = exy_n.x * zn.x - exy_n.y * zn.y, en_im = exy_n.x * zn.y + exy_n.y * zn.x;
                    const double eo_re = exy_o.x * zo.x - exy_o.y * zo.y, eo_im = exy_o.x * zo.y + exy_o.y * zo.x;
                    const double dre = qi * (en_re - eo_re), dim = qi * (en_im - eo_im);
                    const int n2 = nx * nx + ny * ny + nz * nz;
                    const double A = (nx > 0 ? 2.0 : 1.0) * s_gk[n2];
                    de += A * (2.0 * (q[s].x * dre + q[s].y * dim) + (dre * dre + dim * dim));
                }
            }
        }

        // block reduction (fixed order) and the acceptance test on thread 0
        {
            const double a = wave_sum(e_new), b = wave_sum(e_old), c = wave_sum(de);
            if ((tid & 63) == 0) {
                s_red[0][tid >> 6] = a;
                s_red[1][tid >> 6] = b;
                s_red[2][tid >> 6] = c;
            }
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, b = 0.0, c = 0.0;
            for (int w = 0; w < kThreads / 64; ++w) {
                a += s_red[0][w];
                b += s_red[1][w];
                c += s_red[2][w];
            }
            // the surface term enters with the reciprocal one (Ewald::energy, energy.cpp:581-587)
            double dsurf = 0.0, mx = dip0, my = dip1, mz = dip2;
            if (surf != 0.0) {
                mx += po.w * (nx_ - po.x);
                my += po.w * (ny_ - po.y);
                mz += po.w * (nz_ - po.z);
                dsurf = surf * ((mx * mx + my * my + mz * mz) - (dip0 * dip0 + dip1 * dip1 + dip2 * dip2));
            }
            const double du = energy_change(a, b, pref * c + dsurf);
            const bool acc = metropolis(beta * du, P.accept_u);
            if (acc) {
                dip0 = mx;
                dip1 = my;
                dip2 = mz;
            }
            s_accept = acc ? 1 : 0;
            if (kF32) s_nhot = 0; // every thread has read the queue before the reduction
            if (trace && t < trace_cap) {
                TraceEntry& T = trace[static_cast<size_t>(blockIdx.x) * trace_cap + t];
                T.index = i;
                T.accepted = acc ? 1 : 0;
                T.du = du;
                T.x = nx_;
                T.y = ny_;
                T.z = nz_;
            }
            if (acc) {
                energy += du;
                sum_du += du;
                ++accepted;
                pos[i] = make_double4(nx_, ny_, nz_, po.w);
                if (kF32) pos32[i] = fa;
                if (kCells) {
                    const int cxn = cell_of_coord(nx_, half, S.inv_cell, S.nc0);
                    const int cyn = cell_of_coord(ny_, half, S.inv_cell, S.nc1);
                    const int czn = cell_of_coord(nz_, half, S.inv_cell, S.nc2);
                    const int cnew = (czn * S.nc1 + cyn) * S.nc0 + cxn;
                    const int cold = S.cell_of[i];
                    if (cnew != cold) {
                        const int nn = S.cell_count[cnew];
                        if (nn >= S.cap) {
                            // full cell: the lists are left as they are and the chain stops
                            // after this move
                            S.error = 1;
                        } else {
                            const int nold = S.cell_count[cold];
                            for (int kk = 0; kk < nold; ++kk)
                                if (S.cell_members[cold * S.cap + kk] == i) {
                                    S.cell_members[cold * S.cap + kk] = S.cell_members[cold * S.cap + nold - 1];
                                    break;
                                }
                            S.cell_count[cold] = nold - 1;
                            S.cell_members[cnew * S.cap + nn] = i;
                            S.cell_count[cnew] = nn + 1;
                            S.cell_of[i] = cnew;
                        }
                    }
                }
            }
            if (sample_every > 0 && isfinite(energy) &&
                (move0 + static_cast<unsigned long long>(t) + 1ull) % static_cast<unsigned long long>(sample_every) == 0ull) {
                e_sum += energy;
                e_sum2 += energy * energy;
                ++e_samples;
            }
        }
        __syncthreads();
        if (s_accept) {
            // Q += dQ for this thread's k-vectors (phase tables still hold this move)
            int prev_col = -1;
            double2 exy_n = make_double2(1.0, 0.0), exy_o = make_double2(1.0, 0.0);
            const double qi = po.w;
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                const int k = k0 + s;
                if (s < kpt && k < nk) {
                    const int packed = s_kn[k];
                    const int nx = unpack_n(packed, 0), ny = unpack_n(packed, 1), nz = unpack_n(packed, 2);
                    const int col = packed & 0xffff;
                    if (col != prev_col) {
                        const double2 xn = s_phase[nx + ncut], yn = s_phase[width + ny + ncut];
                        const double2 xo = s_phase[3 * width + nx + ncut], yo = s_phase[4 * width + ny + ncut];
                        exy_n = make_double2(xn.x * yn.x - xn.y * yn.y, xn.x * yn.y + xn.y * yn.x);
                        exy_o = make_double2(xo.x * yo.x - xo.y * yo.y, xo.x * yo.y + xo.y * yo.x);
                        prev_col = col;
                    }
                    const double2 zn = s_phase[2 * width + nz + ncut], zo = s_phase[5 * width + nz + ncut];
                    const double en_re = exy_n.x * zn.x - exy_n.y * zn.y, en_im = exy_n.x * zn.y + exy_n.y * zn.x;
                    const double eo_re = exy_o.x * zo.x - exy_o.y * zo.y, eo_im = exy_o.x * zo.y + exy_o.y * zo.x;
                    q[s].x += qi * (en_re - eo_re);
                    q[s].y += qi * (en_im - eo_im);
                }
            }
        }
        if (S.error) {
            t_done = t + 1;
            break;
        }
        __syncthreads(); // phase tables and the accepted position are consumed before the next move
    }

    if (tid == 0) {
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
        const int l = ladders != nullptr ? ladders->ladder_of[blockIdx.x] : -1;
        if (l >= 0) {
            const LadderRun& lr = *ladders;
            // the ladder's barrier: the energy above is published before the arrival; the
            // last replica to arrive swaps the rungs, the others wait for its release
            __threadfence();
            const unsigned target = static_cast<unsigned>(lr.size[l]) * static_cast<unsigned>(iv + 1);
            if (atomicAdd(&lr.arrive[l], 1u) + 1u == target) {
                __threadfence();
                swap_ladder_rungs(states, lr.begin[l], lr.size[l], lr.rung_slot, l, lr.seed,
                                  lr.round0 + static_cast<unsigned long long>(iv), lr.counters);
                __threadfence();
                __hip_atomic_store(&lr.gen[l], static_cast<unsigned>(iv + 1), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            } else {
                long long spins = 0;
                while (__hip_atomic_load(&lr.gen[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                       static_cast<unsigned>(iv + 1)) {
                    if (++spins > lr.spin_limit) {
                        atomicMax(&S.error, 2); // a rung did not arrive: this replica stops
                        break;
                    }
                    __builtin_amdgcn_s_sleep(20);
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            beta = __hip_atomic_load(&S.beta, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    move0 += static_cast<unsigned long long>(t_done);
    __syncthreads(); // the interval's state and error are written before the next interval reads them
    }

#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s)
        if (s < kpt && k0 + s < nk) S.q[k0 + s] = q[s];
}

// Replica exchange along each ladder (odd/even neighbour pairs by round parity):
// accept with min(1, exp((beta_a - beta_b)(U_a - U_b))), energies at T_ref; swaps the
// beta labels (ParallelTempering semantics, move.cpp:844-965, without moving data).
// With round_counter non-null (one block holding every ladder, as in a captured
// interval), the round is read from that device counter, which the block then
// advances, so that each replay of the interval takes the next round.
__global__ void ladder_swap_kernel(ReplicaState* states, const int* ladder_begin, const int* ladder_size,
                                   int* rung_slot, int nladders, unsigned long long seed, unsigned long long round,
                                   unsigned long long* counters, unsigned long long* round_counter) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (round_counter) {
        round = *round_counter;
        __syncthreads();
        if (threadIdx.x == 0) *round_counter = round + 1ull;
    }
    if (l >= nladders) return;
    swap_ladder_rungs(states, ladder_begin[l], ladder_size[l], rung_slot, l, seed, round, counters);
}

template <int kThreads, int kMaxKPT, bool kCells, bool kF32>
void launch_advance(int grid, size_t lds, hipStream_t st, ReplicaState* s, const long long* m, TraceEntry* tr, int cap,
                    const LadderRun* lr) {
    auto fn = mc_advance_kernel<kThreads, kMaxKPT, kCells, kF32>;
    // dynamic LDS: what the static LDS leaves of the CU's 160 KB, at most 96 KB (set once)
    static size_t limit = 0;
    if (limit == 0) {
        hipFuncAttributes attr;
        check_hip(hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(fn)), "hipFuncGetAttributes");
        limit = std::min<size_t>(96 * 1024, 160 * 1024 - attr.sharedSizeBytes);
        check_hip(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(limit)),
                  "hipFuncSetAttribute");
    }
    if (lds > limit) throw std::runtime_error("k-vector table too large for LDS");
    if (lr == nullptr) {
        hipLaunchKernelGGL(fn, dim3(grid), dim3(kThreads), lds, st, s, m, tr, cap, lr);
        return;
    }
    // ladders wait for each other: every workgroup must be resident at once
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), kThreads, lds),
              "occupancy");
    if (per_cu * prop.multiProcessorCount < grid)
        throw std::runtime_error("run_ladders: the replicas cannot all be resident at once");
    void* args[] = {&s, &m, &tr, &cap, &lr};
    check_hip(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(fn), dim3(grid), dim3(kThreads), args,
                                         static_cast<unsigned>(lds), st),
              "mc_advance_kernel cooperative launch");
}

template <int kThreads, int kMaxKPT>
void dispatch_cells(bool cells, bool f32, int grid, size_t lds, hipStream_t st, ReplicaState* s, const long long* m,
                    TraceEntry* tr, int cap, const LadderRun* lr) {
    if (cells) {
        if (f32) launch_advance<kThreads, kMaxKPT, true, true>(grid, lds, st, s, m, tr, cap, lr);
        else launch_advance<kThreads, kMaxKPT, true, false>(grid, lds, st, s, m, tr, cap, lr);
    } else {
        if (f32) launch_advance<kThreads, kMaxKPT, false, true>(grid, lds, st, s, m, tr, cap, lr);
        else launch_advance<kThreads, kMaxKPT, false, false>(grid, lds, st, s, m, tr, cap, lr);
    }
}

// Pair-distance histogram of one configuration (AtomRDF::_sample, analysis.cpp:1556-1600):
// types a != b: every (a, b) pair; a == b: pairs i < j.  r = |minimum-image distance|,
// bin floor(r / dr) as Equidistant2DTable::to_bin (equidistant_table.h:32-39), bins >= nbins
// dropped.  Block b takes particles i = b, b + grid, ...; its lanes stride over j; counts
// go to an LDS histogram, then once per block to the global one.
__global__ __launch_bounds__(256) void rdf_kernel(const double4* __restrict__ pos, const uint32_t* __restrict__ meta,
                                                  int n, int ta, int tb, double L, double half, double inv_dr,
                                                  int nbins, unsigned long long* __restrict__ counts) {
    extern __shared__ unsigned int s_hist[];
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_hist[b] = 0u;
    __syncthreads();
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int ti = static_cast<int>(meta[i] & 0xffu);
        if (ti != ta) continue; // uniform over the block
        const double4 pi = pos[i];
        for (int j = (ta == tb ? i + 1 : 0) + static_cast<int>(threadIdx.x); j < n; j += blockDim.x) {
            if (static_cast<int>(meta[j] & 0xffu) != tb) continue;
            const double4 pj = pos[j];
            const double dx = min_image(pi.x - pj.x, L, half);
            const double dy = min_image(pi.y - pj.y, L, half);
            const double dz = min_image(pi.z - pj.z, L, half);
            // x^2 + y^2 + z^2 rounded term by term, as the reference's norm
            const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz)));
            const double x = floor(__dmul_rn(r, inv_dr));
            if (x < static_cast<double>(nbins)) atomicAdd(&s_hist[static_cast<int>(x)], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x)
        if (s_hist[b]) atomicAdd(&counts[b], static_cast<unsigned long long>(s_hist[b]));
}

} // namespace

// ==============================================================================
Batch::Batch(int precision_fp32) : fp32_(precision_fp32) {
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&ev0_), "hipEventCreate");
    check_hip(hipEventCreate(&ev1_), "hipEventCreate");
}

Batch::~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (auto p : graph_owned_) (void)hipFree(p);
    for (auto p : owned_) (void)hipFree(p);
    (void)hipFree(d_states_);
    (void)hipFree(d_moves_);
    (void)hipFree(d_trace_);
    (void)hipFree(split_slots_);
    (void)hipFree(split_priv_);
    (void)hipFree(pipe_slots_);
    (void)hipFree(split_err_);
    (void)hipFree(d_map_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}


namespace {
template <class T> T* upload(const std::vector<T>& v, std::vector<void*>& owned, const char* what) {
    T* d = nullptr;
    check_hip(hipMalloc(&d, sizeof(T) * std::max<size_t>(1, v.size())), what);
    owned.push_back(d);
    if (!v.empty()) check_hip(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), what);
    return d;
}
} // namespace

// S(q) and dS/dq of the engine's Coulomb model: the Andrea spline (value and the
// derivative of its interval polynomial) or the exact function.
static void short_range_and_slope(const DevCoulomb& c, double q, double* s, double* ds) {
    if (c.splined) {
        const int n = c.nknots;
        const int lo = static_cast<int>(std::lower_bound(c.knots, c.knots + n, q) - c.knots);
        int pos = std::min(std::max(lo - 1, 0), n - 2);
        const double* k = c.coeff + 6 * pos;
        const double z = q - c.knots[pos];
        *s = ((((k[5] * z + k[4]) * z + k[3]) * z + k[2]) * z + k[1]) * z + k[0];
        *ds = (((5 * k[5] * z + 4 * k[4]) * z + 3 * k[3]) * z + 2 * k[2]) * z + k[1];
    } else if (c.scheme == kCoulEwald) {
        const double a = c.alpha * c.rc;
        *s = std::erfc(a * q);
        *ds = -2.0 * a / std::sqrt(kPi) * std::exp(-a * a * q * q);
    } else {
        const double q5 = q * q * q * q * q;
        *s = 1.0 - 1.75 * q + 5.25 * q5 - 7.0 * q5 * q + 2.5 * q5 * q * q;
        *ds = -1.75 + 26.25 * q * q * q * q - 42.0 * q5 + 17.5 * q5 * q;
    }
}

// Cubic Hermite coefficients (in the local coordinate of each interval) of S(q).
static std::vector<float4> short_range_table(const DevCoulomb& c) {
    std::vector<float4> t(static_cast<size_t>(kSTable));
    const double h = 1.0 / kSTable;
    for (int i = 0; i < kSTable; ++i) {
        double f0, d0, f1, d1;
        short_range_and_slope(c, i * h, &f0, &d0);
        short_range_and_slope(c, (i + 1) * h, &f1, &d1);
        const double c2 = 3.0 * (f1 - f0) - h * (2.0 * d0 + d1);
        const double c3 = 2.0 * (f0 - f1) + h * (d0 + d1);
        t[static_cast<size_t>(i)] = make_float4(static_cast<float>(f0), static_cast<float>(h * d0),
                                                static_cast<float>(c2), static_cast<float>(c3));
    }
    return t;
}

namespace {
// One pair energy per thread in both device forms: the FP64 one (pair_energy) and the
// FP32-pair form of the device-resident kernels (pair_energy_ft with its S(q) table, and
// the FP64 re-evaluation of hot pairs), at the same distance.
__global__ void pair_probe_kernel(const DevPair* __restrict__ pairs, const DevCoulomb* __restrict__ coul,
                                  const float4* __restrict__ stab, int nt, int terms, int ta, int tb, double qa, double qb,
                                  int n, const double* __restrict__ r2, double* __restrict__ u64,
                                  double* __restrict__ u32) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const DevPair& p = pairs[ta * nt + tb];
    u64[i] = pair_energy(p, *coul, qa, qb, r2[i]);
    const DevPairF pf = to_f32(p);
    const CoulF cf = coul_f32(*coul);
    bool hot;
    const float uf = pair_energy_ft(pf, cf, stab, terms, static_cast<float>(qa), static_cast<float>(qb),
                                    static_cast<float>(r2[i]), hot);
    u32[i] = hot ? pair_energy_t(p, *coul, qa, qb, r2[i], terms) : static_cast<double>(uf);
}
} // namespace

void pair_probe(Engine& e, int ta, int tb, double qa, double qb, int n, const double* r2, double* u64, double* u32) {
    const int nt = e.model().ntypes;
    if (ta < 0 || tb < 0 || ta >= nt || tb >= nt) throw std::runtime_error("pair probe: bad atom type");
    if (n < 0 || n > (1 << 24)) throw std::runtime_error("pair probe: 0 to 2^24 distances");
    if (n == 0) return;
    std::vector<void*> owned;
    try {
        const DevCoulomb& dc = e.dev_coulomb();
        const float4* stab = (dc.scheme == kCoulEwald || dc.scheme == kCoulFanourgakis)
                                 ? upload(short_range_table(dc), owned, "upload S(q) table")
                                 : nullptr;
        int terms = 0;
        for (const DevPair& p : e.dev_pairs()) {
            if (p.lj_e4 != 0.0) terms |= kTermLJ;
            if (p.wca_e4 != 0.0) terms |= kTermWCA;
            if (p.hs_s2 > 0.0) terms |= kTermHS;
        }
        const std::vector<double> x(r2, r2 + n);
        const double* d_r2 = upload(x, owned, "upload r2");
        double* d_out = upload(std::vector<double>(2 * static_cast<size_t>(n), 0.0), owned, "probe output");
        hipLaunchKernelGGL(pair_probe_kernel, dim3((n + 255) / 256), dim3(256), 0, e.stream(), e.d_pairs(), e.d_coul(),
                           stab, nt, terms, ta, tb, qa, qb, n, d_r2, d_out, d_out + n);
        check_hip(hipGetLastError(), "pair_probe_kernel");
        check_hip(hipStreamSynchronize(e.stream()), "pair probe");
        check_hip(hipMemcpy(u64, d_out, sizeof(double) * n, hipMemcpyDeviceToHost), "copy probe");
        check_hip(hipMemcpy(u32, d_out + n, sizeof(double) * n, hipMemcpyDeviceToHost), "copy probe");
    } catch (...) {
        for (void* p : owned) (void)hipFree(p);
        throw;
    }
    for (void* p : owned) (void)hipFree(p);
}

// The dipole sum q r over the engine's particles (positions as stored).
static void replica_dipole(Engine& e, double* dip) {
    const int n = e.num_particles();
    std::vector<double> xyz(3 * static_cast<size_t>(n));
    e.get_positions(xyz.data());
    const auto& q = e.host_charges();
    dip[0] = dip[1] = dip[2] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) dip[d] += xyz[3 * static_cast<size_t>(i) + d] * q[static_cast<size_t>(i)];
}

ReplicaState make_replica_state(Engine& e, bool fp32, unsigned long long seed, int stream_id, double beta, int ladder,
                                std::vector<void*>& owned, const std::vector<MoveKindSpec>* kinds) {
    const Model& m = e.model();
    const int n = e.num_particles();
    if (!(m.box[0] == m.box[1] && m.box[1] == m.box[2])) throw std::runtime_error("device MC needs a cubic box");
    ReplicaState s{};
    s.n = n;
    s.nk = e.num_kvectors();
    s.ncut = m.ewald.present ? static_cast<int>(std::ceil(m.ewald.n_cutoff)) : 0;
    if (s.ncut * s.ncut >= kMaxGk) throw std::runtime_error("ncutoff too large for the device MC (max 32)");
    if (m.ewald.present && !m.ewald.spherical_sum) throw std::runtime_error("device MC needs spherical_sum: true");
    if (!kinds)
        for (const auto& mol : m.molecules)
            if (!mol.atomic) throw std::runtime_error("device MC handles atomic groups (single-particle translations)");
    // Hamiltonian::energy stops summing a state's terms at the first one >= maxenergy
    // (energy.cpp:1227-1241); the device kernels form dU as one difference, so a finite
    // maxenergy is left to the per-move path
    if (std::isfinite(m.maximum_allowed_energy))
        throw std::runtime_error("device MC: maxenergy is supported by the per-move path only");
    s.ntypes = m.ntypes;
    s.L = m.box[0];
    s.half = 0.5 * m.box[0];
    s.inv = 1.0 / m.box[0];
    s.rec_prefactor = m.ewald.present ? 2 * kPi * m.ewald.lB / (m.box[0] * m.box[1] * m.box[2]) : 0.0;
    s.pos = e.d_pos();
    s.meta = e.d_meta();
    s.q = e.d_q_acc();
    s.kn = e.d_kn();
    s.pairs = e.d_pairs();
    s.coul = e.d_coul();
    CellList& cl = e.cells();
    s.cells = cl.enabled ? 1 : 0;
    if (cl.enabled) {
        if (!(cl.nc[0] == cl.nc[1] && cl.nc[1] == cl.nc[2])) throw std::runtime_error("cubic cell grid expected");
        s.nc0 = cl.nc[0];
        s.nc1 = cl.nc[1];
        s.nc2 = cl.nc[2];
        s.cap = cl.cap;
        s.inv_cell = cl.inv_cell[0];
        s.cell_count = cl.count;
        s.cell_members = cl.members;
        s.cell_of = cl.cell_of;
    }
    // A_k as a function of |n|^2 for a cubic box (the n_x>0 factor 2 is applied in the kernel)
    std::vector<double> gk(static_cast<size_t>(kMaxGk), 0.0);
    if (m.ewald.present) {
        const double kb = 2 * kPi / m.box[0];
        for (int n2 = 1; n2 <= s.ncut * s.ncut; ++n2) {
            const double k2 = kb * kb * n2;
            gk[static_cast<size_t>(n2)] = std::exp(-k2 / (4 * m.ewald.alpha * m.ewald.alpha)) / k2;
        }
    }
    double* d_gk = nullptr;
    check_hip(hipMalloc(&d_gk, sizeof(double) * kMaxGk), "hipMalloc gk");
    owned.push_back(d_gk);
    check_hip(hipMemcpy(d_gk, gk.data(), sizeof(double) * kMaxGk, hipMemcpyHostToDevice), "copy gk");
    s.gk = d_gk;
    std::vector<double> dpt(static_cast<size_t>(m.ntypes));
    for (int t = 0; t < m.ntypes; ++t) dpt[static_cast<size_t>(t)] = m.atoms[static_cast<size_t>(t)].dp;
    double* d_dp = nullptr;
    check_hip(hipMalloc(&d_dp, sizeof(double) * m.ntypes), "hipMalloc dp");
    owned.push_back(d_dp);
    check_hip(hipMemcpy(d_dp, dpt.data(), sizeof(double) * m.ntypes, hipMemcpyHostToDevice), "copy dp");
    s.dp_type = d_dp;
    FixPos* d_p32 = nullptr;
    if (fp32) {
        std::vector<double> xyz(3 * static_cast<size_t>(n));
        e.get_positions(xyz.data());
        const std::vector<FixPos> p32 = fixed_positions(xyz.data(), e.host_charges(), n, s.inv);
        check_hip(hipMalloc(&d_p32, sizeof(FixPos) * n), "hipMalloc pos32");
        owned.push_back(d_p32);
        check_hip(hipMemcpy(d_p32, p32.data(), sizeof(FixPos) * n, hipMemcpyHostToDevice), "copy pos32");
    }
    s.pos32 = d_p32;
    s.parts = 1;
    s.gslots = nullptr;
    s.sample_every = 0;
    s.e_samples = 0;
    s.e_sum = s.e_sum2 = 0.0;
    s.pair_range = e.cells().enabled ? e.cell_range() : e.model().pair_range;
    // surface term (PolicyIonIon::surfaceEnergy, energy.cpp:466-482): the prefactor and
    // the dipole sum, kept up to date by the kernels
    s.surf = 0.0;
    s.dip[0] = s.dip[1] = s.dip[2] = 0.0;
    if (m.ewald.present && m.ewald.const_inf > 0.5) {
        const double V = m.box[0] * m.box[1] * m.box[2];
        s.surf = m.ewald.const_inf * 2.0 * kPi / ((2.0 * m.ewald.epss + 1.0) * V) * m.ewald.lB;
        replica_dipole(e, s.dip);
    }
    s.cell_radius = e.cell_radius();
    s.threads = 0;
    const DevCoulomb& dc = e.dev_coulomb();
    if (fp32 && (dc.scheme == kCoulEwald || dc.scheme == kCoulFanourgakis)) s.stab = upload(short_range_table(dc), owned, "upload S(q) table");
    s.terms = 0;
    for (const DevPair& p : e.dev_pairs()) {
        if (p.lj_e4 != 0.0) s.terms |= kTermLJ;
        if (p.wca_e4 != 0.0) s.terms |= kTermWCA;
        if (p.hs_s2 > 0.0) s.terms |= kTermHS;
    }
    s.seed = seed;
    s.replica_id = stream_id;
    s.ladder = ladder;
    s.beta = beta;
    if (kinds) {
        if (kinds->empty()) throw std::runtime_error("empty move set");
        const auto& gb = e.group_begin();
        const auto& gs = e.group_size();
        const auto& gm = e.group_molid();
        std::vector<MoveKindDev> kd;
        std::vector<int> kgroups;
        std::vector<int2> pbonds;
        unsigned long long cum = 0;
        int max_moved = 1;
        std::vector<int> kind_of_molid(m.molecules.size(), -1);
        for (const MoveKindSpec& k : *kinds) {
            if (k.kind < 0 || k.kind > 2) throw std::runtime_error("unknown move kind");
            if (k.molid < 0 || k.molid >= static_cast<int>(m.molecules.size())) throw std::runtime_error("move molid out of range");
            // translate+rotate keeps the mass centre by translation only: its molecule
            // type must not also be changed internally by the other kinds
            int& seen = kind_of_molid[static_cast<size_t>(k.molid)];
            if (seen >= 0 && ((seen == 1) != (k.kind == 1)))
                throw std::runtime_error("a molecule type moved by translate+rotate takes no atom or pivot moves");
            seen = k.kind;
            MoveKindDev d{};
            d.kind = k.kind;
            d.group_first = static_cast<int>(kgroups.size());
            for (size_t g = 0; g < gm.size(); ++g)
                if (gm[g] == k.molid) {
                    kgroups.push_back(static_cast<int>(g));
                    if (k.kind != 0) max_moved = std::max(max_moved, gs[g]);
                }
            d.group_count = static_cast<int>(kgroups.size()) - d.group_first;
            if (d.group_count == 0) throw std::runtime_error("move kind without groups");
            const MoleculeCfg& mol = m.molecules[static_cast<size_t>(k.molid)];
            if (k.kind == 1 && mol.atomic) throw std::runtime_error("translate+rotate needs a molecular group");
            if (k.kind == 2) {
                if (mol.atomic) throw std::runtime_error("pivot needs a molecular group");
                d.bond_first = static_cast<int>(pbonds.size());
                for (const BondCfg& b : mol.bonds) pbonds.push_back(make_int2(b.i, b.j));
                d.bond_count = static_cast<int>(pbonds.size()) - d.bond_first;
            }
            cum += k.weight;
            d.cum = cum;
            d.dp = k.dp;
            d.dprot = k.dprot;
            kd.push_back(d);
        }
        if (cum == 0 || cum > 0xffffffffull) throw std::runtime_error("move weights must sum to 1..2^32-1");
        if (max_moved > kGeneralMaxMoved) throw std::runtime_error("molecule too large for the device MC");
        std::vector<double> dpa(static_cast<size_t>(m.ntypes));
        for (int t = 0; t < m.ntypes; ++t)
            dpa[static_cast<size_t>(t)] = m.atoms[static_cast<size_t>(t)].has_dp ? m.atoms[static_cast<size_t>(t)].dp : -1.0;
        s.nkinds = static_cast<int>(kd.size());
        s.max_moved = max_moved;
        s.total_weight = cum;
        s.kinds = upload(kd, owned, "upload move kinds");
        s.kind_groups = upload(kgroups, owned, "upload kind groups");
        s.pivot_bonds = upload(pbonds, owned, "upload pivot bonds");
        s.gbegin = upload(gb, owned, "upload group begin");
        s.gsize = upload(gs, owned, "upload group size");
        s.gflags = e.d_group_flags();
        s.gcm = e.d_mass_centers();
        s.dp_atom = upload(dpa, owned, "upload atom dp");
        const bool bonded = m.bonded_term >= 0 && e.num_bonds() > 0;
        s.bond_off = bonded ? e.d_bond_offsets() : nullptr;
        s.bonds = bonded ? e.d_bonds() : nullptr;
        s.excl_off = e.d_exclusion_offsets();
        s.excl = e.d_exclusions();
        s.nmol = static_cast<int>(m.molecules.size());
        s.g2g2 = e.d_g2g_cutoffs();
        std::vector<double> mw(static_cast<size_t>(m.ntypes));
        for (int t = 0; t < m.ntypes; ++t) mw[static_cast<size_t>(t)] = m.atoms[static_cast<size_t>(t)].mw;
        s.mw = upload(mw, owned, "upload atom masses");
    }
    std::vector<double> terms(static_cast<size_t>(e.num_terms()));
    e.total_energy(terms.data());
    double u0 = 0;
    for (double v : terms) u0 += v;
    s.energy = u0;
    return s;
}

int advance_threads(const ReplicaState& s) {
    if (s.threads) return s.threads;
    return (!s.cells && s.nk <= 8 * 512) ? 512 : 256;
}

void launch_advance(const ReplicaState* h_states, int count, ReplicaState* d_states, const long long* d_moves, bool f32,
                    TraceEntry* trace, int trace_cap, hipStream_t stream, const LadderRun* ladders) {
    const LadderRun* lr = ladders;
    // one launch per run of replicas with the same workgroup size (a function of each
    // replica alone, so that a replica's arithmetic does not depend on its batch)
    int first = 0;
    while (first < count) {
        const int threads = advance_threads(h_states[first]);
        int last = first;
        int maxkpt = 0, maxnk = 0, maxw = 0;
        while (last < count && advance_threads(h_states[last]) == threads) {
            const ReplicaState& s = h_states[last];
            if (s.cells != h_states[first].cells) throw std::runtime_error("replicas of one launch must agree on the cell list");
            maxkpt = std::max(maxkpt, (s.nk + threads - 1) / threads);
            maxnk = std::max(maxnk, s.nk);
            maxw = std::max(maxw, 2 * s.ncut + 1);
            ++last;
        }
        if (maxkpt > 68) throw std::runtime_error("too many k-vectors per replica for the device MC (max 68 per thread)");
        const size_t lds =
            sizeof(int) * static_cast<size_t>((maxnk + 3) & ~3) + sizeof(double2) * 6 * static_cast<size_t>(maxw);
        if (lds > 96 * 1024) throw std::runtime_error("k-vector table too large for LDS");
        const bool cells = h_states[first].cells != 0;
        const int grid = last - first;
        ReplicaState* ds = d_states + first;
        const long long* dm = d_moves + first;
        TraceEntry* tr = trace ? trace + static_cast<size_t>(first) * trace_cap : nullptr;
        if (lr != nullptr && (first != 0 || last != count))
            throw std::runtime_error("run_ladders: every replica must take the same workgroup size");
        if (threads == 512) dispatch_cells<512, 8>(cells, f32, grid, lds, stream, ds, dm, tr, trace_cap, lr);
        else if (maxkpt <= 8) dispatch_cells<256, 8>(cells, f32, grid, lds, stream, ds, dm, tr, trace_cap, lr);
        else if (maxkpt <= 32) dispatch_cells<256, 32>(cells, f32, grid, lds, stream, ds, dm, tr, trace_cap, lr);
        else dispatch_cells<256, 68>(cells, f32, grid, lds, stream, ds, dm, tr, trace_cap, lr);
        check_hip(hipGetLastError(), "mc_advance_kernel launch");
        first = last;
    }
}

int Batch::add_replica(const std::string& config_json, int n, const double* xyz, const double* charge, const int* ids,
                       int ngroups, const int* gbegin, const int* gsize, const int* gmolid, double beta,
                       unsigned long long seed, int ladder, const std::vector<MoveKindSpec>* kinds) {
    if (finalized_) throw std::runtime_error("batch already finalized");
    const int general = kinds ? 1 : 0;
    if (general_ >= 0 && general_ != general)
        throw std::runtime_error("a batch holds either translate-only replicas or replicas with move sets");
    auto e = std::make_unique<Engine>(config_json);
    e->set_particles(n, xyz, charge, ids);
    e->set_groups(ngroups, gbegin, gsize, gmolid, nullptr);
    e->set_cell_divisions(cell_divisions_);
    e->set_cell_capacity(cell_capacity_);
    e->init();
    if (!general && e->cell_list_enabled() && e->cell_radius() != 1)
        throw std::runtime_error("the finer cell grid is walked by the general move-set kernel only");
    ReplicaState s = make_replica_state(*e, fp32_ != 0, seed, static_cast<int>(engines_.size()), beta, ladder, owned_,
                                        kinds);
    general_ = general;
    h_states_.push_back(s);
    engines_.push_back(std::move(e));
    return s.replica_id;
}

void Batch::set_cell_capacity(int cap) {
    if (cap < 0) throw std::runtime_error("cell capacity must be >= 0");
    cell_capacity_ = cap;
}

void Batch::set_cell_divisions(int d) {
    if (d < 1 || d > 2) throw std::runtime_error("cell divisions: 1 or 2");
    if (!engines_.empty()) throw std::runtime_error("set the cell divisions before adding replicas");
    cell_divisions_ = d;
}

int Batch::threads_for(int r) const {
    if (threads_override_) return threads_override_;
    const ReplicaState& s = h_states_[static_cast<size_t>(r)];
    (void)s;
    return 256;
}

void Batch::finalize() {
    if (engines_.empty()) throw std::runtime_error("empty batch");
    const int threads = threads_for(0);
    for (int r = 0; r < size(); ++r) {
        if (threads_for(r) != threads) throw std::runtime_error("replicas of one batch must share the workgroup size");
        ReplicaState& s = h_states_[static_cast<size_t>(r)];
        s.kpt = (s.nk + threads - 1) / threads;
        if (s.kpt > 68) throw std::runtime_error("too many k-vectors per replica for the device MC (max 68 per thread)");
        if (s.cells != h_states_[0].cells) throw std::runtime_error("replicas must agree on the cell list");
    }
    check_hip(hipMalloc(&d_states_, sizeof(ReplicaState) * size()), "hipMalloc states");
    check_hip(hipMemcpy(d_states_, h_states_.data(), sizeof(ReplicaState) * size(), hipMemcpyHostToDevice), "copy states");
    check_hip(hipMalloc(&d_moves_, sizeof(long long) * size()), "hipMalloc moves");
    finalized_ = true;
}

void Batch::run(const std::vector<long long>& moves, bool trace, int trace_cap) {
    if (!finalized_) finalize();
    if (static_cast<int>(moves.size()) != size()) throw std::runtime_error("run: one move count per replica");
    check_hip(hipMemcpyAsync(d_moves_, moves.data(), sizeof(long long) * size(), hipMemcpyHostToDevice, stream_), "copy moves");
    if (trace) {
        if (trace_cap > trace_cap_) {
            (void)hipFree(d_trace_);
            check_hip(hipMalloc(&d_trace_, sizeof(TraceEntry) * static_cast<size_t>(trace_cap) * size()), "hipMalloc trace");
            trace_cap_ = trace_cap;
        }
    }
    if (threads_for(0) != 256) throw std::runtime_error("unsupported workgroup size");
    TraceEntry* tr = trace ? d_trace_ : nullptr;
    const int cap = trace ? trace_cap : 0;
    time_begin();
    if (general_ == 1) {
        // one workgroup per replica, also for replicas that earlier ran split
        std::vector<ReplicaState> one(h_states_);
        for (auto& s : one) s.parts = 1;
        launch_general(one.data(), size(), d_states_, d_moves_, fp32_ != 0, tr, cap, stream_);
    }
    else launch_advance(h_states_.data(), size(), d_states_, d_moves_, fp32_ != 0, tr, cap, stream_);
    time_end();
}

void Batch::sample_energy(long long every) {
    if (every < 0) throw std::runtime_error("sample_energy: interval >= 0");
    if (!finalized_) finalize();
    check_hip(hipStreamSynchronize(stream_), "batch sync");
    std::vector<ReplicaState> cur(static_cast<size_t>(size()));
    check_hip(hipMemcpy(cur.data(), d_states_, sizeof(ReplicaState) * size(), hipMemcpyDeviceToHost), "copy states");
    for (int r = 0; r < size(); ++r) {
        for (ReplicaState* s : {&cur[static_cast<size_t>(r)], &h_states_[static_cast<size_t>(r)]}) {
            s->sample_every = every;
            s->e_samples = 0;
            s->e_sum = s->e_sum2 = 0.0;
        }
    }
    check_hip(hipMemcpy(d_states_, cur.data(), sizeof(ReplicaState) * size(), hipMemcpyHostToDevice), "copy states");
}

void Batch::rdf(int r, int type_a, int type_b, double dr, int nbins, unsigned long long* counts) {
    if (r < 0 || r >= size()) throw std::runtime_error("replica index out of range");
    if (!(dr > 0.0)) throw std::runtime_error("rdf: dr must be positive");
    if (nbins < 1 || nbins > 8192) throw std::runtime_error("rdf: 1 to 8192 bins");
    const ReplicaState& h = h_states_[static_cast<size_t>(r)];
    if (type_a < 0 || type_b < 0 || type_a >= h.ntypes || type_b >= h.ntypes) throw std::runtime_error("rdf: bad atom type");
    if (!finalized_) finalize();
    unsigned long long* d_counts = nullptr;
    check_hip(hipMalloc(&d_counts, sizeof(unsigned long long) * static_cast<size_t>(nbins)), "hipMalloc rdf");
    std::vector<unsigned long long> part(static_cast<size_t>(nbins));
    try {
        check_hip(hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * static_cast<size_t>(nbins), stream_),
                  "clear rdf");
        const int grid = std::max(1, std::min(h.n, 2048));
        hipLaunchKernelGGL(rdf_kernel, dim3(grid), dim3(256), sizeof(unsigned int) * static_cast<size_t>(nbins), stream_,
                           h.pos, h.meta, h.n, type_a, type_b, h.L, h.half, 1.0 / dr, nbins, d_counts);
        check_hip(hipGetLastError(), "rdf_kernel launch");
        check_hip(hipMemcpyAsync(part.data(), d_counts, sizeof(unsigned long long) * static_cast<size_t>(nbins),
                                 hipMemcpyDeviceToHost, stream_),
                  "copy rdf");
        check_hip(hipStreamSynchronize(stream_), "rdf sync");
    } catch (...) {
        (void)hipFree(d_counts);
        throw;
    }
    check_hip(hipFree(d_counts), "hipFree rdf");
    for (int b = 0; b < nbins; ++b) counts[b] += part[static_cast<size_t>(b)];
}

void Batch::time_begin() {
    if (timing_) check_hip(hipEventRecord(ev0_, stream_), "event");
}

void Batch::time_end() {
    if (!timing_) return;
    check_hip(hipEventRecord(ev1_, stream_), "event");
    ev_pending_ = true;
}

// Waits for the batch's stream; the duration of the last timed launch is read once.
void Batch::sync() {
    check_hip(hipStreamSynchronize(stream_), "batch sync");
    if (ev_pending_) {
        float ms = 0;
        ev_pending_ = false;
        check_hip(hipEventElapsedTime(&ms, ev0_, ev1_), "event elapsed");
        last_ms_ = ms;
    }
}

void Batch::swap_ladders(unsigned long long seed, unsigned long long round) {
    // ladders: replicas with equal ladder id, ordered by decreasing beta at creation
    if (!finalized_) finalize();
    std::vector<int> lb, ls, slots;
    std::vector<int> ids;
    for (auto& s : h_states_) ids.push_back(s.ladder);
    std::vector<int> uniq = ids;
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    sync();
    std::vector<ReplicaState> cur(static_cast<size_t>(size()));
    check_hip(hipMemcpy(cur.data(), d_states_, sizeof(ReplicaState) * size(), hipMemcpyDeviceToHost), "copy states");
    for (int lid : uniq) {
        if (lid < 0) continue;
        std::vector<int> members;
        for (int r = 0; r < size(); ++r)
            if (cur[static_cast<size_t>(r)].ladder == lid) members.push_back(r);
        std::sort(members.begin(), members.end(), [&](int a, int b) { return cur[static_cast<size_t>(a)].beta > cur[static_cast<size_t>(b)].beta; });
        lb.push_back(static_cast<int>(slots.size()));
        ls.push_back(static_cast<int>(members.size()));
        slots.insert(slots.end(), members.begin(), members.end());
    }
    if (lb.empty()) return;
    int *d_lb = nullptr, *d_ls = nullptr, *d_slots = nullptr;
    unsigned long long* d_cnt = nullptr;
    check_hip(hipMalloc(&d_lb, sizeof(int) * lb.size()), "hipMalloc");
    check_hip(hipMalloc(&d_ls, sizeof(int) * ls.size()), "hipMalloc");
    check_hip(hipMalloc(&d_slots, sizeof(int) * slots.size()), "hipMalloc");
    check_hip(hipMalloc(&d_cnt, sizeof(unsigned long long) * 2), "hipMalloc");
    check_hip(hipMemcpy(d_lb, lb.data(), sizeof(int) * lb.size(), hipMemcpyHostToDevice), "copy");
    check_hip(hipMemcpy(d_ls, ls.data(), sizeof(int) * ls.size(), hipMemcpyHostToDevice), "copy");
    check_hip(hipMemcpy(d_slots, slots.data(), sizeof(int) * slots.size(), hipMemcpyHostToDevice), "copy");
    check_hip(hipMemset(d_cnt, 0, sizeof(unsigned long long) * 2), "memset");
    const int nl = static_cast<int>(lb.size());
    hipLaunchKernelGGL(ladder_swap_kernel, dim3((nl + 63) / 64), dim3(64), 0, stream_, d_states_, d_lb, d_ls, d_slots, nl,
                       seed, round, d_cnt, static_cast<unsigned long long*>(nullptr));
    check_hip(hipGetLastError(), "ladder_swap_kernel");
    unsigned long long cnt[2] = {0, 0};
    check_hip(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, stream_), "copy");
    check_hip(hipStreamSynchronize(stream_), "sync");
    swaps_att_ += static_cast<long long>(cnt[0]);
    swaps_acc_ += static_cast<long long>(cnt[1]);
    (void)hipFree(d_lb);
    (void)hipFree(d_ls);
    (void)hipFree(d_slots);
    (void)hipFree(d_cnt);
}

void Batch::get_replica_state(int r, double* xyz, double* mass_center, unsigned long long* moves,
                              unsigned long long* accepted) {
    const ReplicaState s = state(r);
    positions(r, xyz);
    Engine& e = *engines_[static_cast<size_t>(r)];
    if (mass_center) {
        const size_t ng = e.group_begin().size();
        check_hip(hipMemcpy(mass_center, e.d_mass_centers(), sizeof(double) * 3 * ng, hipMemcpyDeviceToHost),
                  "copy mass centres");
    }
    if (moves) *moves = s.moves;
    if (accepted) *accepted = s.accepted;
}

void Batch::set_replica_state(int r, const double* xyz, const double* mass_center, unsigned long long moves,
                              unsigned long long accepted) {
    ReplicaState cur = state(r); // finalizes and waits for the stream
    Engine& e = *engines_[static_cast<size_t>(r)];
    e.reset_positions(xyz, mass_center);
    const CellList& cl = e.cells();
    ReplicaState& h = h_states_[static_cast<size_t>(r)];
    for (ReplicaState* s : {&cur, &h}) {
        if (s->cells != (cl.enabled ? 1 : 0)) throw std::runtime_error("set_replica_state: cell list changed kind");
        if (cl.enabled) {
            s->nc0 = cl.nc[0];
            s->nc1 = cl.nc[1];
            s->nc2 = cl.nc[2];
            s->inv_cell = cl.inv_cell[0];
            s->cell_radius = cl.radius;
            s->cap = cl.cap;
            s->cell_count = cl.count;
            s->cell_members = cl.members;
            s->cell_of = cl.cell_of;
        }
    }
    if (cur.surf != 0.0) {
        replica_dipole(e, cur.dip);
        for (int d = 0; d < 3; ++d) h.dip[d] = cur.dip[d];
    }
    if (cur.pos32) {
        const int n = e.num_particles();
        std::vector<double> now(3 * static_cast<size_t>(n));
        e.get_positions(now.data());
        const std::vector<FixPos> p32 = fixed_positions(now.data(), e.host_charges(), n, cur.inv);
        check_hip(hipMemcpy(cur.pos32, p32.data(), sizeof(FixPos) * n, hipMemcpyHostToDevice), "copy pos32");
    }
    std::vector<double> terms(static_cast<size_t>(e.num_terms()));
    e.total_energy(terms.data());
    double u = 0.0;
    for (double v : terms) u += v;
    cur.energy = u;
    cur.moves = moves;
    cur.accepted = accepted;
    cur.sum_du = 0.0;
    cur.error = 0;
    check_hip(hipMemcpy(d_states_ + r, &cur, sizeof(ReplicaState), hipMemcpyHostToDevice), "copy state");
}

// One replica-exchange interval as a HIP graph: the persistent launch that advances
// replica r by moves[r] moves and one odd/even round of neighbour swaps on every ladder
// (ParallelTempering, move.cpp:844-965), the round taken from a device counter so that
// replays alternate the pairings.  Ladders, their rung order and the counters live on
// the device from here on.
void Batch::capture_interval(const std::vector<long long>& moves, unsigned long long seed, unsigned long long round0) {
    if (!finalized_) finalize();
    if (static_cast<int>(moves.size()) != size()) throw std::runtime_error("capture_interval: one move count per replica");
    for (const auto& s : h_states_)
        if (s.parts > 1) throw std::runtime_error("capture_interval: replicas shared by several workgroups are not captured");
    sync();
    // ladders: replicas with equal ladder id >= 0, rungs ordered by decreasing beta
    std::vector<ReplicaState> cur(static_cast<size_t>(size()));
    check_hip(hipMemcpy(cur.data(), d_states_, sizeof(ReplicaState) * size(), hipMemcpyDeviceToHost), "copy states");
    std::vector<int> ids;
    for (const auto& s : cur) ids.push_back(s.ladder);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    std::vector<int> lb, ls, slots;
    for (int lid : ids) {
        if (lid < 0) continue;
        std::vector<int> members;
        for (int r = 0; r < size(); ++r)
            if (cur[static_cast<size_t>(r)].ladder == lid) members.push_back(r);
        std::sort(members.begin(), members.end(),
                  [&](int a, int b) { return cur[static_cast<size_t>(a)].beta > cur[static_cast<size_t>(b)].beta; });
        lb.push_back(static_cast<int>(slots.size()));
        ls.push_back(static_cast<int>(members.size()));
        slots.insert(slots.end(), members.begin(), members.end());
    }
    const int nl = static_cast<int>(lb.size());
    if (nl > 1024) throw std::runtime_error("capture_interval: at most 1024 ladders");
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    for (void* p : graph_owned_) (void)hipFree(p);
    graph_owned_.clear();
    auto dev_copy = [&](const std::vector<int>& v) {
        int* d = nullptr;
        check_hip(hipMalloc(&d, sizeof(int) * std::max<size_t>(1, v.size())), "hipMalloc ladders");
        graph_owned_.push_back(d);
        if (!v.empty()) check_hip(hipMemcpy(d, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice), "copy ladders");
        return d;
    };
    int* d_lb = dev_copy(lb);
    int* d_ls = dev_copy(ls);
    int* d_rung = dev_copy(slots);
    unsigned long long* d_cnt = nullptr;
    check_hip(hipMalloc(&d_cnt, sizeof(unsigned long long) * 3), "hipMalloc swap counters");
    graph_owned_.push_back(d_cnt);
    const unsigned long long init[3] = {0ull, 0ull, round0};
    check_hip(hipMemcpy(d_cnt, init, sizeof(init), hipMemcpyHostToDevice), "copy swap counters");
    graph_counters_ = d_cnt;
    check_hip(hipMemcpy(d_moves_, moves.data(), sizeof(long long) * size(), hipMemcpyHostToDevice), "copy moves");
    // one launch outside the capture sets the kernels' attributes (not a stream operation)
    const std::vector<long long> zero(static_cast<size_t>(size()), 0);
    check_hip(hipMemcpy(d_moves_, zero.data(), sizeof(long long) * size(), hipMemcpyHostToDevice), "copy moves");
    if (general_ == 1) launch_general(h_states_.data(), size(), d_states_, d_moves_, fp32_ != 0, nullptr, 0, stream_);
    else launch_advance(h_states_.data(), size(), d_states_, d_moves_, fp32_ != 0, nullptr, 0, stream_);
    check_hip(hipStreamSynchronize(stream_), "attribute run");
    check_hip(hipMemcpy(d_moves_, moves.data(), sizeof(long long) * size(), hipMemcpyHostToDevice), "copy moves");
    hipGraph_t g = nullptr;
    check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        if (general_ == 1) launch_general(h_states_.data(), size(), d_states_, d_moves_, fp32_ != 0, nullptr, 0, stream_);
        else launch_advance(h_states_.data(), size(), d_states_, d_moves_, fp32_ != 0, nullptr, 0, stream_);
        if (nl > 0) {
            hipLaunchKernelGGL(ladder_swap_kernel, dim3(1), dim3(((nl + 63) / 64) * 64), 0, stream_, d_states_, d_lb, d_ls,
                               d_rung, nl, seed, 0ull, d_cnt, d_cnt + 2);
            check_hip(hipGetLastError(), "ladder_swap_kernel");
        }
    } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    check_hip(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check_hip(ie, "hipGraphInstantiate");
}

// Free-running ladders: intervals[r] swap intervals of moves[r] moves per replica in one
// cooperative launch of the translate-only kernel.  After each interval the replicas of a
// ladder meet at a barrier in HBM and the last to arrive runs that ladder's round of swaps
// (round0 + interval, the same rule and deviates as the captured interval), so a ladder
// waits only for its own rungs: ladders of different sizes and interval counts share the
// device without a device-wide barrier per interval.  The rungs of a ladder share its
// interval count; replicas outside ladders just run their intervals.
void Batch::run_ladders(const std::vector<long long>& moves, const std::vector<int>& intervals, unsigned long long seed,
                        unsigned long long round0) {
    if (!finalized_) finalize();
    const int n = size();
    if (static_cast<int>(moves.size()) != n || static_cast<int>(intervals.size()) != n)
        throw std::runtime_error("run_ladders: one move count and one interval count per replica");
    if (general_ == 1) throw std::runtime_error("run_ladders: translate-only replicas");
    for (int r = 0; r < n; ++r) {
        if (intervals[static_cast<size_t>(r)] < 0) throw std::runtime_error("run_ladders: interval counts >= 0");
        if (h_states_[static_cast<size_t>(r)].parts > 1)
            throw std::runtime_error("run_ladders: replicas shared by several workgroups");
    }
    sync();
    std::vector<ReplicaState> cur(static_cast<size_t>(n));
    check_hip(hipMemcpy(cur.data(), d_states_, sizeof(ReplicaState) * n, hipMemcpyDeviceToHost), "copy states");
    // ladders as in capture_interval: ids >= 0 in increasing order, rungs by decreasing beta
    std::vector<int> ids;
    for (const auto& s : cur) ids.push_back(s.ladder);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    std::vector<int> lb, ls, slots, ladder_of(static_cast<size_t>(n), -1);
    for (int lid : ids) {
        if (lid < 0) continue;
        std::vector<int> members;
        for (int r = 0; r < n; ++r)
            if (cur[static_cast<size_t>(r)].ladder == lid) members.push_back(r);
        std::sort(members.begin(), members.end(),
                  [&](int a, int b) { return cur[static_cast<size_t>(a)].beta > cur[static_cast<size_t>(b)].beta; });
        for (int r : members) {
            if (intervals[static_cast<size_t>(r)] != intervals[static_cast<size_t>(members[0])])
                throw std::runtime_error("run_ladders: the rungs of a ladder must share the interval count");
            ladder_of[static_cast<size_t>(r)] = static_cast<int>(lb.size());
        }
        lb.push_back(static_cast<int>(slots.size()));
        ls.push_back(static_cast<int>(members.size()));
        slots.insert(slots.end(), members.begin(), members.end());
    }
    const size_t nl = lb.size();
    // one allocation: counters (2 x u64), then intervals, ladder_of, begin, size, rung_slot,
    // arrive, gen (ints); arrive and gen start at zero
    std::vector<int> words;
    words.insert(words.end(), intervals.begin(), intervals.end());
    words.insert(words.end(), ladder_of.begin(), ladder_of.end());
    words.insert(words.end(), lb.begin(), lb.end());
    words.insert(words.end(), ls.begin(), ls.end());
    words.insert(words.end(), slots.begin(), slots.end());
    words.insert(words.end(), 2 * nl, 0);
    // the descriptor first, then the counters (2 x u64), then the words
    const size_t head = (sizeof(LadderRun) + 15) & ~static_cast<size_t>(15);
    const size_t bytes = head + 16 + sizeof(int) * words.size();
    std::vector<unsigned char> host(bytes, 0);
    unsigned char* d = nullptr;
    check_hip(hipMalloc(&d, bytes), "hipMalloc ladders");
    try {
        int* w = reinterpret_cast<int*>(d + head + 16);
        LadderRun lr{};
        lr.intervals = w;
        lr.ladder_of = w + n;
        lr.begin = w + 2 * n;
        lr.size = w + 2 * n + nl;
        lr.rung_slot = w + 2 * n + 2 * nl;
        lr.arrive = reinterpret_cast<unsigned*>(w + 2 * n + 2 * nl + slots.size());
        lr.gen = lr.arrive + nl;
        lr.counters = reinterpret_cast<unsigned long long*>(d + head);
        lr.seed = seed;
        lr.round0 = round0;
        lr.spin_limit = 1LL << 26;
        std::memcpy(host.data(), &lr, sizeof(lr));
        std::memcpy(host.data() + head + 16, words.data(), sizeof(int) * words.size());
        check_hip(hipMemcpy(d_moves_, moves.data(), sizeof(long long) * n, hipMemcpyHostToDevice), "copy moves");
        check_hip(hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice), "copy ladders");
        time_begin();
        launch_advance(h_states_.data(), n, d_states_, d_moves_, fp32_ != 0, nullptr, 0, stream_,
                       reinterpret_cast<const LadderRun*>(d));
        time_end();
        unsigned long long cnt[2] = {0, 0};
        check_hip(hipMemcpyAsync(cnt, d + head, sizeof(cnt), hipMemcpyDeviceToHost, stream_), "copy swap counters");
        sync();
        swaps_att_ += static_cast<long long>(cnt[0]);
        swaps_acc_ += static_cast<long long>(cnt[1]);
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        (void)hipFree(d);
        throw;
    }
    check_hip(hipFree(d), "hipFree ladders");
}

void Batch::replay(int count) {
    if (!graph_) throw std::runtime_error("replay: no interval captured");
    if (count < 0) throw std::runtime_error("replay: count >= 0");
    time_begin();
    for (int i = 0; i < count; ++i) check_hip(hipGraphLaunch(graph_, stream_), "hipGraphLaunch");
    time_end();
}

// Tile autotuner: times each candidate launch shape of this batch on a short run and
// keeps the fastest for the whole batch (one shape per batch keeps every replica in one
// persistent launch).  Translate-only replicas: workgroups of 256 or 512 threads (K/256
// or K/512 k-vectors per thread; 512 holds at most 8 per thread, so K <= 4096).
// Move-set replicas with a cell list: cells of at least the pair range (27-cell stencils)
// or half of it (125 smaller cells).  Every candidate runs `trial` moves per replica from
// checkpoints that are restored afterwards, so the chains continue from their states
// before tuning (Q(k) and the running energy rebuilt from the positions).
std::vector<TuneCandidate> Batch::autotune(long long trial) {
    if (trial < 1) throw std::runtime_error("autotune: trial moves >= 1");
    if (!finalized_) finalize();
    for (const auto& s : h_states_)
        if (s.parts > 1) throw std::runtime_error("autotune: replicas shared by several workgroups");
    sync();
    struct Checkpoint {
        std::vector<double> xyz, cm;
        unsigned long long moves = 0, accepted = 0;
    };
    std::vector<Checkpoint> ck(static_cast<size_t>(size()));
    for (int r = 0; r < size(); ++r) {
        Checkpoint& c = ck[static_cast<size_t>(r)];
        c.xyz.resize(3 * static_cast<size_t>(engines_[static_cast<size_t>(r)]->num_particles()));
        c.cm.resize(3 * engines_[static_cast<size_t>(r)]->group_begin().size());
        get_replica_state(r, c.xyz.data(), c.cm.data(), &c.moves, &c.accepted);
    }
    std::vector<int> cands;
    if (general_ == 1) {
        bool cells = false;
        for (const auto& s : h_states_) cells = cells || s.cells;
        if (!cells) return {};
        cands = {1, 2};
    } else {
        cands = {256};
        bool fits512 = true;
        for (const auto& s : h_states_) fits512 = fits512 && (s.nk + 511) / 512 <= 8;
        if (fits512) cands.push_back(512);
    }
    auto apply = [&](int c) {
        for (int r = 0; r < size(); ++r) {
            if (general_ == 1) engines_[static_cast<size_t>(r)]->regrid(c);
            else h_states_[static_cast<size_t>(r)].threads = c;
        }
    };
    auto restore = [&]() {
        for (int r = 0; r < size(); ++r) {
            const Checkpoint& c = ck[static_cast<size_t>(r)];
            set_replica_state(r, c.xyz.data(), c.cm.data(), c.moves, c.accepted);
        }
    };
    const bool timing = timing_;
    timing_ = true;
    std::vector<TuneCandidate> out;
    try {
        for (int c : cands) {
            apply(c);
            if (general_ == 1) restore(); // the new grid
            run(std::vector<long long>(static_cast<size_t>(size()), trial), false, 0);
            sync();
            out.push_back(TuneCandidate{c, last_ms_ * 1e3 / (static_cast<double>(trial) * size())});
            restore();
        }
    } catch (...) {
        timing_ = timing;
        throw;
    }
    timing_ = timing;
    const auto best = std::min_element(out.begin(), out.end(),
                                       [](const TuneCandidate& a, const TuneCandidate& b) { return a.us_per_move < b.us_per_move; });
    apply(best->value);
    restore();
    tuned_ = best->value;
    return out;
}

void Batch::swap_counts(long long* attempted, long long* accepted) {
    sync();
    long long a = swaps_att_, b = swaps_acc_;
    if (graph_counters_) {
        unsigned long long c[2] = {0, 0};
        check_hip(hipMemcpy(c, graph_counters_, sizeof(c), hipMemcpyDeviceToHost), "copy swap counters");
        a += static_cast<long long>(c[0]);
        b += static_cast<long long>(c[1]);
    }
    *attempted = a;
    *accepted = b;
}

ReplicaState Batch::state(int r) {
    if (r < 0 || r >= size()) throw std::runtime_error("replica index out of range");
    if (!finalized_) finalize();
    sync();
    ReplicaState s;
    check_hip(hipMemcpy(&s, d_states_ + r, sizeof(ReplicaState), hipMemcpyDeviceToHost), "copy state");
    return s;
}

void Batch::positions(int r, double* xyz) {
    if (r < 0 || r >= size()) throw std::runtime_error("replica index out of range");
    sync();
    Engine& e = *engines_[static_cast<size_t>(r)];
    e.refresh_host_positions();
    e.get_positions(xyz);
}

std::vector<TraceEntry> Batch::trace(int r, int count) {
    if (r < 0 || r >= size() || count < 0) throw std::runtime_error("trace: replica index or count out of range");
    sync();
    std::vector<TraceEntry> t(static_cast<size_t>(count));
    if (!d_trace_ || count > trace_cap_) throw std::runtime_error("trace not recorded");
    check_hip(hipMemcpy(t.data(), d_trace_ + static_cast<size_t>(r) * trace_cap_, sizeof(TraceEntry) * count,
                        hipMemcpyDeviceToHost), "copy trace");
    return t;
}

} // namespace fmc
