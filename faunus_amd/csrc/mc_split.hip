// One large replica advanced by many workgroups at once.  Workgroup g owns a
// slice of the k-vectors (its Q(k) stays in registers for the whole launch);
// workgroup 0 also evaluates the pair dU through the cell list and commits
// accepted moves.  After each move every workgroup publishes its partial sums, waits
// for all others (bounded spin), sums the partials with the same butterfly and so
// takes the same decision as every other workgroup.  While wave 0 of workgroup 0
// waits, its other waves load the next move's particle and stencil neighbours; the
// next move uses them unless the accepted move changed one of the loaded cells.
// The launch is cooperative, so all workgroups are resident.
//
// The exchange needs no fences.  Every published double travels as two words, its
// high and its low 32 bits each beside the move's 32-bit tag, written and polled as
// relaxed agent-scope atomics: a reader takes a value when both words carry the tag it
// waits for, so no write has to be ordered before another and no cache has to be
// written back or invalidated per move.  Only workgroup 0 reads or writes the shared
// positions and the cell list after the launch starts.  The other workgroups keep
// private copies of the positions (copied at the launch's start, each accepted move
// applied by every workgroup), so workgroup 0 publishes a single value, the pair energy
// change; without the copies (the no-allocation entry point) it also publishes the next
// move's particle position as committed so far.
//
// A replica whose cell list overflows stops after the move that overflowed: that move
// is committed (position, Q(k), energy), the particle stays listed in its old cell,
// S.error = 1, and workgroup 0 publishes an abort tag in its slot of the next move so
// that the other workgroups stop there too.  A wait that exceeds its spin bound sets
// S.error = 2 (the workgroups may then disagree on the last move).  Either way the
// replica refuses later launches until its state is set again, and S.moves counts the
// moves that were decided.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device_math.hpp"
#include "mc_batch.hpp"
#include "mc_common.hpp"
#include "philox.hpp"

namespace fmc {
namespace {

constexpr int kSplitMaxKPT = 8; // k-vectors per thread at 256 threads
constexpr int kSplitMaxG = 64;
constexpr int kGkMax = 1025;
// published values: workgroup 0 the pair energy change and, without private position
// copies, the next move's particle x, y, z and charge; the others their de
constexpr int kSplitValues = 5;
// tag bit set in a published word: the publishing workgroup (0) stopped before this move
constexpr uint32_t kAbortBit = 0x80000000u;

// One workgroup's exchange slot for one move parity: value v in words 2v (high half)
// and 2v + 1 (low half), one 128-byte line.
struct alignas(128) SplitSlot {
    unsigned long long w[16];
};

// Tag of move `mid`: bit 30 set, so that zeroed slots never carry a valid tag; moves two
// apart (the previous use of a slot) have different tags.
__device__ __forceinline__ uint32_t split_tag(unsigned long long mid) {
    return 0x40000000u | static_cast<uint32_t>((mid + 1ull) & 0x3fffffffull);
}

__device__ __forceinline__ void split_put(SplitSlot* sl, int v, double x, uint32_t tag) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    __hip_atomic_store(&sl->w[2 * v], (b & 0xffffffff00000000ull) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&sl->w[2 * v + 1], (b << 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int cell_coord_s(double x, double half, double inv_cell, int nc) {
    int c = static_cast<int>(floor((x + half) * inv_cell));
    c %= nc;
    if (c < 0) c += nc;
    return c;
}

__device__ __forceinline__ double wrap_s(double x, double L, double half, double inv) {
    if (fabs(x) > half) {
        const double s = x * inv;
        x -= L * static_cast<double>(static_cast<int>(s > 0.0 ? s + 0.5 : s - 0.5));
    }
    return x;
}

template <int kThreads, int kMaxKPT, bool kCells>
__global__ __launch_bounds__(kThreads) void mc_split_kernel(ReplicaState* __restrict__ states, long long M, int G,
                                                       int stride, SplitSlot* __restrict__ slots,
                                                       int* __restrict__ err, TraceEntry* __restrict__ trace,
                                                       int trace_cap, long long spin_limit, double4* __restrict__ priv) {
    if (blockIdx.x % stride != 0) return;
    const int g = blockIdx.x / stride;
    if (g >= G) return;
    ReplicaState& S = states[0];
    if (S.error) return; // a stopped replica (read before any workgroup of this launch can set it)
    __shared__ DevPair s_pairs[kMaxTypes * kMaxTypes];
    __shared__ DevCoulomb s_coul;
    __shared__ double s_dp[kMaxTypes];
    __shared__ double s_gk[kGkMax];
    __shared__ double s_red[3][kThreads / 64];
    __shared__ int s_accept, s_abort, s_stop;
    // proposals of the next 64 moves and the first of the following 64, one lane each
    __shared__ Proposal s_prop[65];
    __shared__ double s_thr[65]; // -log(u) of each proposal's acceptance deviate
    __shared__ double2 s_phase[6 * 65];
    __shared__ double s_npo[4]; // workgroups g > 0: the next move's particle (x, y, z, q)
    __shared__ int s_nti;       // and its type
    // Workgroup 0 at 512 threads with a cell list: waves 1-6 (7 lanes per stencil cell)
    // load the next move's particle and its stencil neighbours while wave 0 waits for the
    // other workgroups' partials.  The loads are kept (s_pf_ok) unless the accepted move
    // was the next particle or changed one of the 54 loaded cells (s_mv_cell).
    constexpr bool kPre = kCells && kThreads >= 512;
    constexpr int kPfLanes = 7;
    // The last wave publishes the workgroup's partials, so that wave 0's polls do not
    // wait for those stores to complete (the vector memory counter covers both).
    constexpr int kPub = kThreads - 64;
    __shared__ double s_pf_pos[4];
    __shared__ int s_pf_ti, s_pf_ok;
    __shared__ int s_pf_cell_of[2]; // the next particle's cell_of entry, by move parity
    __shared__ int s_cnew[2];       // the cell of a move's new position, by move parity
    __shared__ int s_mv_cell[2];

    const int tid = threadIdx.x;
    const int n = S.n, nk = S.nk, ncut = S.ncut, nt = S.ntypes;
    const int width = 2 * ncut + 1;
    const double L = S.L, half = S.half, inv = S.inv;
    for (int i = tid; i < nt * nt; i += kThreads) s_pairs[i] = S.pairs[i];
    lds_copy(s_coul, S.coul, tid, kThreads);
    if (tid == 0) {
        s_abort = 0;
        s_stop = 0;
        s_pf_ok = 0;
    }
    for (int i = tid; i < nt; i += kThreads) s_dp[i] = S.dp_type[i];
    for (int i = tid; i <= ncut * ncut && i < kGkMax; i += kThreads) s_gk[i] = S.gk[i];

    // this workgroup's k-slice, interleaved over the threads.  With G > 1 workgroup 0,
    // whose pair walk and commit form the chain every move waits for, takes none: the
    // k-vectors go to workgroups 1..G-1.
    const int kparts = G > 1 ? G - 1 : 1;
    const int kw = (nk + kparts - 1) / kparts;
    const int kb = G > 1 ? (g == 0 ? 0 : (g - 1) * kw) : 0;
    const int ke = G > 1 && g == 0 ? 0 : min(nk, kb + kw);
    int kn[kMaxKPT];
    double2 q[kMaxKPT], dq[kMaxKPT];
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) {
        const int k = kb + s * kThreads + tid;
        const bool on = k < ke;
        kn[s] = on ? S.kn[k] : 0;
        q[s] = on ? S.q[k] : make_double2(0.0, 0.0);
        dq[s] = make_double2(0.0, 0.0);
    }
    double4* __restrict__ pos = S.pos;
    const uint32_t* __restrict__ meta = S.meta;
    // workgroups g > 0: this workgroup's private copy of the positions (null: none)
    double4* __restrict__ mine = (priv != nullptr && G > 1 && g > 0)
                                     ? priv + static_cast<size_t>(g - 1) * static_cast<size_t>(S.n)
                                     : nullptr;
    if (mine != nullptr)
        for (int j = tid; j < S.n; j += kThreads) mine[j] = pos[j];
    // the cell list, read once
    const double inv_cell = S.inv_cell;
    const int nc0 = S.nc0, nc1 = S.nc1, nc2 = S.nc2, ccap = S.cap;
    int* __restrict__ cell_count = S.cell_count;
    int* __restrict__ cell_members = S.cell_members;
    int* __restrict__ cell_of = S.cell_of;
    const unsigned long long move0 = S.moves;
    const double beta = S.beta, pref = S.rec_prefactor;
    const double surf = S.surf;
    double dip0 = S.dip[0], dip1 = S.dip[1], dip2 = S.dip[2]; // thread 0's running dipole (every workgroup)
    double energy = S.energy, sum_du = S.sum_du;
    unsigned long long accepted = S.accepted;
    const unsigned long long seed = S.seed;
    const uint32_t replica = static_cast<uint32_t>(S.replica_id);
    const long long sample_every = S.sample_every;
    unsigned long long e_samples = S.e_samples;
    double e_sum = S.e_sum, e_sum2 = S.e_sum2;
    // the next move's particle as committed so far: workgroup 0's publisher (published
    // without private copies) or thread 0 of the others (from the private copy); its type
    double4 nxt = make_double4(0.0, 0.0, 0.0, 0.0);
    int nxt_ti = 0;
    // this lane's stencil cell of the next move (waves 1-6 of workgroup 0) and its pair
    // sum there, both from the look-ahead
    int f_cell = -1;
    double f_sum = 0.0;
    long long t_done = M; // moves decided by this launch
    __syncthreads();

    for (long long t = 0; t < M; ++t) {
        const unsigned long long mid = move0 + static_cast<unsigned long long>(t);
        if ((t & 63) == 0) {
            // the previous proposals were last read before this move's barriers
            if (tid < 65) {
                const Proposal Q = make_proposal(seed, replica, mid + static_cast<unsigned long long>(tid), n);
                s_prop[tid] = Q;
                s_thr[tid] = -log(Q.accept_u);
            }
            __syncthreads();
        }
        const Proposal P = s_prop[t & 63];
        const int i = P.index;
        const int i_next = s_prop[(t & 63) + 1].index;
        if (t + 1 < M) {
            // workgroup 0's commits so far are visible to its own threads, and each
            // workgroup's updates of its private copy to itself
            if (g == 0 && tid == kPub && mine == nullptr && priv == nullptr) nxt = pos[i_next];
            if (g != 0 && tid == 0) {
                nxt_ti = static_cast<int>(meta[i_next] & 0xffu);
                if (mine != nullptr) nxt = mine[i_next];
            }
        }
        const bool pre = kPre && g == 0 && s_pf_ok;
        double4 po;
        int ti;
        if (pre) {
            po = make_double4(s_pf_pos[0], s_pf_pos[1], s_pf_pos[2], s_pf_pos[3]);
            ti = s_pf_ti;
        } else if (g != 0 && t > 0) {
            po = make_double4(s_npo[0], s_npo[1], s_npo[2], s_npo[3]);
            ti = s_nti;
        } else {
            po = pos[i]; // workgroup 0, or the launch's first move
            ti = static_cast<int>(meta[i] & 0xffu);
        }
        const double scale = s_dp[ti] * P.length_u;
        const double nx_ = wrap_s(po.x + P.dir[0] * scale, L, half, inv);
        const double ny_ = wrap_s(po.y + P.dir[1] * scale, L, half, inv);
        const double nz_ = wrap_s(po.z + P.dir[2] * scale, L, half, inv);
        for (int e = tid; e < (ke > kb ? 6 * width : 0); e += kThreads) {
            const int nn = e % width - ncut;
            const int axis = (e / width) % 3;
            const int st = e / (3 * width);
            const double x = st == 0 ? (axis == 0 ? nx_ : (axis == 1 ? ny_ : nz_))
                                     : (axis == 0 ? po.x : (axis == 1 ? po.y : po.z));
            double sn, cs;
            sincos((6.283185307179586 * (static_cast<double>(nn) / L)) * x, &sn, &cs);
            s_phase[e] = make_double2(cs, sn);
        }
        double e_new = 0.0, e_old = 0.0;
        if (g == 0) {
            const Box b{{L, L, L}, {half, half, half}};
            if (kPre) {
                // 54 (state, cell) combos x 7 lanes on waves 1-6; a lane whose sum was
                // computed during the previous move takes it from a register
                const int combo = (tid - 64) / kPfLanes, slot0 = (tid - 64) % kPfLanes;
                if (tid >= 64 && combo < 54) {
                    double acc_u = 0.0;
                    if (pre) {
                        acc_u = f_sum;
                    } else {
                        const int st = combo / 27, c = combo % 27;
                        const double px = st == 0 ? nx_ : po.x, py = st == 0 ? ny_ : po.y, pz = st == 0 ? nz_ : po.z;
                        const int cx = (cell_coord_s(px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                        const int cy = (cell_coord_s(py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                        const int cz = (cell_coord_s(pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                        const int cell = (cz * nc1 + cy) * nc0 + cx;
                        if (combo == 13 && slot0 == 0) s_cnew[t & 1] = cell; // the new position's own cell
                        const int cnt = cell_count[cell];
                        for (int k = slot0; k < cnt; k += kPfLanes) {
                            const int j = cell_members[cell * ccap + k];
                            if (j == i) continue;
                            const double4 pj = pos[j];
                            const int tj = static_cast<int>(meta[j] & 0xffu);
                            acc_u += pair_energy(s_pairs[ti * nt + tj], s_coul, po.w, pj.w,
                                                 dist2(b, px, py, pz, pj.x, pj.y, pj.z));
                        }
                    }
                    if (combo < 27) e_new = acc_u;
                    else e_old = acc_u;
                }
            } else if (kCells) {
                // 54 (state, cell) combos x (4 or 8) lanes
                constexpr int kLanes = kThreads >= 512 ? 8 : 4;
                const int combo = tid / kLanes, slot0 = tid % kLanes;
                if (combo < 54) {
                    const int st = combo / 27, c = combo % 27;
                    const double px = st == 0 ? nx_ : po.x, py = st == 0 ? ny_ : po.y, pz = st == 0 ? nz_ : po.z;
                    const int cx = (cell_coord_s(px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                    const int cy = (cell_coord_s(py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                    const int cz = (cell_coord_s(pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                    const int cell = (cz * nc1 + cy) * nc0 + cx;
                    const int cnt = cell_count[cell];
                    for (int k = slot0; k < cnt; k += kLanes) {
                        const int j = cell_members[cell * ccap + k];
                        if (j == i) continue;
                        const double4 pj = pos[j];
                        const double u = pair_energy(s_pairs[ti * nt + static_cast<int>(meta[j] & 0xffu)], s_coul, po.w,
                                                     pj.w, dist2(b, px, py, pz, pj.x, pj.y, pj.z));
                        if (st == 0) e_new += u;
                        else e_old += u;
                    }
                }
            } else {
                for (int j = tid; j < n; j += kThreads) {
                    if (j == i) continue;
                    const double4 pj = pos[j];
                    const DevPair& p = s_pairs[ti * nt + static_cast<int>(meta[j] & 0xffu)];
                    e_new += pair_energy(p, s_coul, po.w, pj.w, dist2(b, nx_, ny_, nz_, pj.x, pj.y, pj.z));
                    e_old += pair_energy(p, s_coul, po.w, pj.w, dist2(b, po.x, po.y, po.z, pj.x, pj.y, pj.z));
                }
            }
        }
        __syncthreads(); // phase tables
        double de = 0.0;
#pragma unroll
        for (int s = 0; s < kMaxKPT; ++s) {
            const int k = kb + s * kThreads + tid;
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
                const double A = (nx > 0 ? 2.0 : 1.0) * s_gk[nx * nx + ny * ny + nz * nz];
                de += A * (2.0 * (q[s].x * dre + q[s].y * dim) + (dre * dre + dim * dim));
            }
        }
        {
            const double a = wave_sum(de), b = wave_sum(e_new), c = wave_sum(e_old);
            if ((tid & 63) == 0) {
                s_red[0][tid >> 6] = a;
                s_red[1][tid >> 6] = b;
                s_red[2][tid >> 6] = c;
            }
        }
        __syncthreads();
        const int parity = static_cast<int>(mid & 1ull);
        const uint32_t tag = split_tag(mid);
        // this workgroup's partials, summed in the same order by the publisher and by thread 0
        double pa = 0.0, pb = 0.0, pc = 0.0;
        if (tid == 0 || tid == kPub) {
            for (int w = 0; w < kThreads / 64; ++w) {
                pa += s_red[0][w];
                pb += s_red[1][w];
                pc += s_red[2][w];
            }
        }
        // workgroup 0's pair energy change as one value: -inf for a NaN old sum (and a new
        // one that is not NaN), +inf for a NaN or +inf new sum, else new - old.  The
        // decision adds the reciprocal change to a finite one; for a finite reciprocal
        // change this is getEnergyChange's handling of the two sums (montecarlo.cpp:193-209),
        // and a NaN one gives +inf as there
        const double pd = (isnan(pc) && !isnan(pb)) ? -__builtin_huge_val()
                          : (isnan(pb) || (pb > 0.0 && isinf(pb))) ? __builtin_huge_val()
                                                                    : pb - pc;
        const bool particle_msg = priv == nullptr && G > 1; // workgroup 0 also sends the next particle
        if (tid == kPub && G > 1) {
            SplitSlot* sl = slots + parity * G + g;
            if (g == 0) {
                if (particle_msg) {
                    // the next particle's load (issued at the move's start) is waited for
                    // here, before the stores: a wait after them would wait for the stores too
                    asm volatile("" ::"v"(nxt.x), "v"(nxt.y), "v"(nxt.z), "v"(nxt.w));
                }
                split_put(sl, 0, pd, tag);
                if (particle_msg) {
                    split_put(sl, 1, nxt.x, tag);
                    split_put(sl, 2, nxt.y, tag);
                    split_put(sl, 3, nxt.z, tag);
                    split_put(sl, 4, nxt.w, tag);
                }
            } else {
                split_put(sl, 0, pa, tag);
            }
        }
        // gather: lane j < G takes workgroup j's value (workgroup 0's pair change, the
        // others' de); lanes 58-61 take workgroup 0's next particle when it is sent.  A
        // workgroup takes its own value from registers, bit for bit what the others read.
        // The des are added with the same butterfly in every workgroup.
        double sde = 0.0, spd = 0.0;                      // totals, valid in thread 0
        double4 npub = make_double4(0.0, 0.0, 0.0, 0.0); // published next particle, thread 0
        if (tid < 64) {
            int ok = 1, stopped = 0;
            double a = 0.0, pdv = 0.0, extra = 0.0;
            const double own_de = readlane_f64(pa, 0), own_pd = readlane_f64(pd, 0);
            auto poll = [&](int src, int v) -> double {
                SplitSlot* sl = slots + parity * G + src;
                long long spins = 0;
                unsigned long long hi, lo;
                for (;;) {
                    hi = __hip_atomic_load(&sl->w[2 * v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    lo = __hip_atomic_load(&sl->w[2 * v + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t th = static_cast<uint32_t>(hi), tl = static_cast<uint32_t>(lo);
                    if (th == tag && tl == tag) break;
                    if (th == (tag | kAbortBit) || tl == (tag | kAbortBit)) {
                        stopped = 1;
                        return 0.0;
                    }
                    if (++spins > spin_limit) {
                        ok = 0;
                        return 0.0;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                return __longlong_as_double(static_cast<long long>((hi & 0xffffffff00000000ull) | (lo >> 32)));
            };
            if (tid < G) {
                if (tid == g) {
                    if (g == 0) {
                        pdv = own_pd;
                        a = G == 1 ? own_de : 0.0; // with G > 1 workgroup 0 holds no k-vectors
                    } else {
                        a = own_de;
                    }
                } else {
                    const double x = poll(tid, 0);
                    if (tid == 0) pdv = x;
                    else a = x;
                }
            }
            const int e = tid - 58; // 0..3 on lanes 58-61
            if (particle_msg && g != 0 && e >= 0 && e < 4) extra = poll(0, 1 + e);
            a = wave_sum(a);
            const double pdt = readlane_f64(pdv, 0);
            const double ex = readlane_f64(extra, 58), ey = readlane_f64(extra, 59);
            const double ez = readlane_f64(extra, 60), eq = readlane_f64(extra, 61);
            const int all_ok = __all(ok);
            const int any_stop = __any(stopped);
            if (tid == 0) {
                sde = a;
                spd = pdt;
                npub = make_double4(ex, ey, ez, eq);
                if (any_stop) {
                    s_stop = 2; // workgroup 0 stopped after the previous move: this one is not decided
                } else if (!all_ok) {
                    atomicExch(err, 1);
                    atomicMax(&S.error, 2);
                    s_abort = 1;
                }
            }
        }
        if (tid == 0 && !s_abort && !s_stop) {
            // the surface term enters with the reciprocal one (Ewald::energy, energy.cpp:581-587)
            double dsurf = 0.0, mx = dip0, my = dip1, mz = dip2;
            if (surf != 0.0) {
                mx += po.w * (nx_ - po.x);
                my += po.w * (ny_ - po.y);
                mz += po.w * (nz_ - po.z);
                dsurf = surf * ((mx * mx + my * my + mz * mz) - (dip0 * dip0 + dip1 * dip1 + dip2 * dip2));
            }
            const double drec = pref * sde + dsurf;
            double du = isnan(drec) ? __builtin_huge_val() : (isinf(spd) ? spd : spd + drec);
            if (isnan(du)) du = 0.0;
            const double bdu = beta * du;
            bool acc;
            if (isinf(bdu) && bdu < 0.0) acc = true;
            else if (-bdu > 709.782712893384) acc = true;
            else {
                // u <= exp(-bdu) decided against the table's -log(u) away from the
                // boundary (the margin exceeds the rounding of exp and log many times),
                // with exp itself only within the margin: the same decision, without exp
                // on the chain
                const double thr = s_thr[t & 63], margin = 1e-12 * (1.0 + fabs(thr));
                if (bdu < thr - margin) acc = true;
                else if (bdu > thr + margin) acc = false;
                else acc = P.accept_u <= exp(-bdu);
            }
            s_accept = acc ? 1 : 0;
            if (acc) {
                dip0 = mx;
                dip1 = my;
                dip2 = mz;
            }
            if (g != 0) {
                // the next move's particle: moved by this move, or as committed so far (the
                // private copy, or workgroup 0's message)
                const bool same = acc && i_next == i;
                const double4 nb = mine != nullptr ? nxt : npub;
                s_npo[0] = same ? nx_ : nb.x;
                s_npo[1] = same ? ny_ : nb.y;
                s_npo[2] = same ? nz_ : nb.z;
                s_npo[3] = nb.w;
                s_nti = nxt_ti;
                if (acc && mine != nullptr) mine[i] = make_double4(nx_, ny_, nz_, po.w);
            }
            if (g == 0) {
                if (kPre) {
                    const long long tn = t + 1;
                    s_pf_ok = (tn < M && !(acc && i_next == i)) ? 1 : 0;
                    s_mv_cell[0] = -1;
                    s_mv_cell[1] = -1;
                }
                if (trace && t < trace_cap) {
                    TraceEntry& T = trace[t];
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
                    if (kCells) {
                        // workgroup 0 with the look-ahead: the stencil lane of the new
                        // position's own cell found it already
                        const int cnew = kPre ? s_cnew[t & 1]
                                              : (cell_coord_s(nz_, half, inv_cell, nc2) * nc1 +
                                                 cell_coord_s(ny_, half, inv_cell, nc1)) * nc0 +
                                                    cell_coord_s(nx_, half, inv_cell, nc0);
                        // cell_of changes only when its particle moves, so the loaded entry is
                        // current whenever the other loads of this move were kept
                        const int cold = pre ? s_pf_cell_of[t & 1] : cell_of[i];
                        if (kPre) {
                            s_mv_cell[0] = cold;
                            s_mv_cell[1] = cnew;
                        }
                        if (cnew != cold) {
                            const int nn = cell_count[cnew];
                            if (nn >= ccap) {
                                // full cell: the lists are left as they are and the chain stops
                                atomicMax(&S.error, 1);
                                s_stop = 1;
                            } else {
                                const int nold = cell_count[cold];
                                for (int kk = 0; kk < nold; ++kk)
                                    if (cell_members[cold * ccap + kk] == i) {
                                        cell_members[cold * ccap + kk] = cell_members[cold * ccap + nold - 1];
                                        break;
                                    }
                                cell_count[cold] = nold - 1;
                                cell_members[cnew * ccap + nn] = i;
                                cell_count[cnew] = nn + 1;
                                cell_of[i] = cnew;
                            }
                        }
                    }
                }
                if (sample_every > 0 && isfinite(energy) &&
                    (mid + 1ull) % static_cast<unsigned long long>(sample_every) == 0ull) {
                    e_sum += energy;
                    e_sum2 += energy * energy;
                    ++e_samples;
                }
            }
        }
        if (kPre && g == 0 && tid >= 64) {
            // while wave 0 waits: load the next move's particle and stencil and sum its
            // pair terms there, as of the commits so far; the sums are kept unless this
            // move's commit changes one of the loaded cells or moves that particle
            const long long tn = t + 1;
            const int combo = (tid - 64) / kPfLanes, slot0 = (tid - 64) % kPfLanes;
            f_cell = -1;
            if (tn < M && combo < 54) {
                const Box b{{L, L, L}, {half, half, half}};
                const Proposal& Q = s_prop[(t & 63) + 1];
                const int in = Q.index;
                const double4 qo = pos[in];
                const int tq = static_cast<int>(meta[in] & 0xffu);
                const double sc = s_dp[tq] * Q.length_u;
                const int st = combo / 27, c = combo % 27;
                double f_px, f_py, f_pz;
                if (st == 0) {
                    f_px = wrap_s(qo.x + Q.dir[0] * sc, L, half, inv);
                    f_py = wrap_s(qo.y + Q.dir[1] * sc, L, half, inv);
                    f_pz = wrap_s(qo.z + Q.dir[2] * sc, L, half, inv);
                } else {
                    f_px = qo.x;
                    f_py = qo.y;
                    f_pz = qo.z;
                }
                if (tid == 64) {
                    s_pf_pos[0] = qo.x;
                    s_pf_pos[1] = qo.y;
                    s_pf_pos[2] = qo.z;
                    s_pf_pos[3] = qo.w;
                    s_pf_ti = tq;
                    s_pf_cell_of[tn & 1] = cell_of[in];
                }
                const int cx = (cell_coord_s(f_px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                const int cy = (cell_coord_s(f_py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                const int cz = (cell_coord_s(f_pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                f_cell = (cz * nc1 + cy) * nc0 + cx;
                if (combo == 13 && slot0 == 0) s_cnew[tn & 1] = f_cell;
                const int f_cnt = cell_count[f_cell];
                // the first two members are loaded at once, further ones in the loop
                const int j0 = slot0 < f_cnt ? cell_members[f_cell * ccap + slot0] : -1;
                const int j1 = slot0 + kPfLanes < f_cnt ? cell_members[f_cell * ccap + slot0 + kPfLanes] : -1;
                double4 p0 = make_double4(0.0, 0.0, 0.0, 0.0), p1 = p0;
                int t0 = 0, t1 = 0;
                if (j0 >= 0) {
                    p0 = pos[j0];
                    t0 = static_cast<int>(meta[j0] & 0xffu);
                }
                if (j1 >= 0) {
                    p1 = pos[j1];
                    t1 = static_cast<int>(meta[j1] & 0xffu);
                }
                double fs = 0.0;
                for (int k = slot0, u = 0; k < f_cnt; k += kPfLanes, ++u) {
                    int j, tj;
                    double4 pj;
                    if (u == 0) {
                        j = j0;
                        pj = p0;
                        tj = t0;
                    } else if (u == 1) {
                        j = j1;
                        pj = p1;
                        tj = t1;
                    } else {
                        j = cell_members[f_cell * ccap + k];
                        if (j == in) continue;
                        pj = pos[j];
                        tj = static_cast<int>(meta[j] & 0xffu);
                    }
                    if (j == in) continue;
                    fs += pair_energy(s_pairs[tq * nt + tj], s_coul, qo.w, pj.w, dist2(b, f_px, f_py, f_pz, pj.x, pj.y, pj.z));
                }
                f_sum = fs;
            }
        }
        __syncthreads();
        if (s_abort) {
            t_done = t;
            break;
        }
        if (s_stop == 2) {
            t_done = t;
            break;
        }
        // the accepted move changed a loaded cell: load again at the next move
        if (kPre && g == 0 && tid >= 64 && f_cell >= 0 && (f_cell == s_mv_cell[0] || f_cell == s_mv_cell[1]))
            s_pf_ok = 0;
        if (s_accept) {
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                q[s].x += dq[s].x;
                q[s].y += dq[s].y;
            }
        }
        if (s_stop == 1) {
            // workgroup 0 committed this move and stopped on a full cell: its words of the
            // next move carry the abort bit, so the other workgroups stop there as well
            if (g == 0 && tid == 0 && t + 1 < M) {
                SplitSlot* sl = slots + static_cast<int>((mid + 1ull) & 1ull) * G;
                for (int v = 0; v < kSplitValues; ++v) split_put(sl, v, 0.0, split_tag(mid + 1ull) | kAbortBit);
                asm volatile("" ::: "memory"); // the abort words are stored before the loop ends
            }
            t_done = t + 1;
            break;
        }
        __syncthreads(); // phase tables are rewritten by the next move
    }
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) {
        const int k = kb + s * kThreads + tid;
        if (k < ke) S.q[k] = q[s];
    }
    if (g == 0 && tid == 0) {
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

// The split kernel for a replica over `groups` workgroups, every `stride`-th block of the
// grid: 512 threads (two waves per SIMD, 7 lanes per stencil cell) when the slice fits 4
// k-vectors per thread, else 256 threads with up to 8; the grid must be co-resident.
struct SplitLaunch {
    const void* fn;
    int threads, grid;
};

SplitLaunch split_launch(const ReplicaState& s, int groups, int stride, const char* who) {
    const std::string w(who);
    if (groups < 1 || groups > kSplitMaxG) throw std::runtime_error(w + ": 1..64 workgroups");
    if (stride < 1) throw std::runtime_error(w + ": stride >= 1");
    const int kparts = groups > 1 ? groups - 1 : 1; // workgroup 0 takes no k-vectors when G > 1
    const int kw = (s.nk + kparts - 1) / kparts;
    const int threads = kw <= 4 * 512 ? 512 : 256;
    if ((kw + threads - 1) / threads > (threads == 512 ? 4 : kSplitMaxKPT))
        throw std::runtime_error(w + ": too many k-vectors per workgroup");
    if (2 * s.ncut + 1 > 65 || s.ncut * s.ncut >= kGkMax) throw std::runtime_error(w + ": ncutoff too large");
    int dev = 0, ncu = 0, per_cu = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const void* fn =
        threads == 512
            ? (s.cells ? reinterpret_cast<const void*>(mc_split_kernel<512, 4, true>)
                       : reinterpret_cast<const void*>(mc_split_kernel<512, 4, false>))
            : (s.cells ? reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, true>)
                       : reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, false>));
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0), "occupancy");
    const int grid = groups * stride;
    if (per_cu * ncu < grid) throw std::runtime_error(w + ": the grid cannot be co-resident");
    return SplitLaunch{fn, threads, grid};
}

// Workspace of the split kernel for callers that cannot allocate
// (fmc_batch_run_split_static): the exchange slots of up to kStaticSlotsG workgroups for
// both move parities and the time-out flag, in a file-scope device array of which each
// GPU holds its own copy for the life of the process.  Host code uses the address that
// hipGetSymbolAddress returns for the current device, never the array's name.
constexpr int kStaticSlotsG = 64;
constexpr size_t kStaticSlotBytes = sizeof(SplitSlot) * 2 * kStaticSlotsG;
constexpr size_t kScratchBytes = kStaticSlotBytes + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per device.
void* scratch_address() {
    static std::mutex mu;
    static void* ptr[kMaxDevices] = {};
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("run_split_static: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        ptr[dev] = p;
    }
    return ptr[dev];
}

} // namespace

// Advance replica 0 of the batch by `moves` moves with G workgroups (every
// `stride`-th block of the grid takes part, so that consecutive participants land on
// the same XCD when stride is the XCD count).
void Batch::run_split(long long moves, int groups, int stride, bool trace, int trace_cap) {
    if (!finalized_) finalize();
    if (size() != 1) throw std::runtime_error("run_split advances a batch of one replica");
    if (general_ == 1) {
        run_split_general(moves, groups, stride, trace, trace_cap);
        return;
    }
    const SplitLaunch sl = split_launch(h_states_[0], groups, stride, "run_split");
    const void* fn = sl.fn;
    const int threads = sl.threads, grid = sl.grid;
    if (!split_slots_) {
        check_hip(hipMalloc(&split_slots_, sizeof(SplitSlot) * 2 * kSplitMaxG), "hipMalloc slots");
        check_hip(hipMalloc(&split_err_, sizeof(int) * 2), "hipMalloc err");
    }
    // slots left by an earlier launch carry tags a resumed (rewound) chain would take for
    // its own: every launch starts from zeroed slots and a clear time-out flag
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(256), 0, stream_, static_cast<unsigned long long*>(split_slots_),
                       static_cast<int>(sizeof(SplitSlot) * 2 * kSplitMaxG / 8));
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, stream_, reinterpret_cast<unsigned long long*>(split_err_),
                       1);
    check_hip(hipGetLastError(), "zero_words_kernel");
    if (trace && trace_cap > trace_cap_) {
        (void)hipFree(d_trace_);
        check_hip(hipMalloc(&d_trace_, sizeof(TraceEntry) * static_cast<size_t>(trace_cap)), "hipMalloc trace");
        trace_cap_ = trace_cap;
    }
    ReplicaState* st = d_states_;
    long long M = moves;
    int G = groups, str = stride;
    SplitSlot* slots = static_cast<SplitSlot*>(split_slots_);
    int* err = split_err_;
    TraceEntry* tr = trace ? d_trace_ : nullptr;
    int cap = trace ? trace_cap : 0;
    long long spin = 1LL << 24;
    // private position copies of workgroups 1..G-1
    const size_t priv_bytes = sizeof(double4) * static_cast<size_t>(std::max(0, G - 1)) *
                              static_cast<size_t>(h_states_[0].n);
    if (priv_bytes > split_priv_bytes_) {
        check_hip(hipStreamSynchronize(stream_), "batch sync");
        (void)hipFree(split_priv_);
        split_priv_ = nullptr;
        split_priv_bytes_ = 0;
        check_hip(hipMalloc(&split_priv_, priv_bytes), "hipMalloc private positions");
        split_priv_bytes_ = priv_bytes;
    }
    double4* priv = static_cast<double4*>(split_priv_);
    void* args[] = {&st, &M, &G, &str, &slots, &err, &tr, &cap, &spin, &priv};
    time_begin();
    check_hip(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, 0, stream_), "hipLaunchCooperativeKernel");
    time_end();
}

// run_split without allocating: the exchange slots and the time-out flag live in the
// file-scope workspace of the current GPU, cleared by a fill kernel on the batch's stream
// before the launch; no trace.  Calls that overlap on one GPU share the workspace and race.
void Batch::run_split_static(long long moves, int groups, int stride) {
    if (!finalized_)
        throw std::runtime_error("run_split_static: the batch has no device state yet (run it, or read a replica's "
                                 "info, once first)");
    if (size() != 1) throw std::runtime_error("run_split_static advances a batch of one replica");
    if (general_ == 1) throw std::runtime_error("run_split_static: translate-only replicas");
    if (groups > kStaticSlotsG)
        throw std::runtime_error("run_split_static: at most " + std::to_string(kStaticSlotsG) +
                                 " workgroups (the static workspace holds " + std::to_string(kStaticSlotsG) +
                                 " exchange slots per move parity)");
    if (moves < 0) throw std::runtime_error("run_split_static: moves >= 0");
    const SplitLaunch sl = split_launch(h_states_[0], groups, stride, "run_split_static");
    void* ws = scratch_address();
    static_assert(kScratchBytes % 8 == 0, "whole words");
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(256), 0, stream_, static_cast<unsigned long long*>(ws),
                       static_cast<int>(kScratchBytes / 8));
    check_hip(hipGetLastError(), "zero_words_kernel");
    ReplicaState* st = d_states_;
    long long M = moves;
    int G = groups, str = stride;
    SplitSlot* slots = static_cast<SplitSlot*>(ws);
    int* err = reinterpret_cast<int*>(static_cast<unsigned char*>(ws) + kStaticSlotBytes);
    TraceEntry* tr = nullptr;
    int cap = 0;
    long long spin = 1LL << 24;
    double4* priv = nullptr; // no private position copies: workgroup 0 sends the next particle
    void* args[] = {&st, &M, &G, &str, &slots, &err, &tr, &cap, &spin, &priv};
    time_begin();
    check_hip(hipLaunchCooperativeKernel(sl.fn, dim3(sl.grid), dim3(sl.threads), args, 0, stream_),
              "hipLaunchCooperativeKernel");
    time_end();
}

// The replica's general-move kernel over `groups` workgroups (cooperative launch).
void Batch::run_split_general(long long moves, int groups, int stride, bool trace, int trace_cap) {
    ReplicaState& h = h_states_[0];
    if (h.parts != groups || (groups > 1 && h.gslots == nullptr)) {
        check_hip(hipStreamSynchronize(stream_), "batch sync");
        set_replica_parts(h, groups, owned_);
        // the device copy holds the chain's dynamic state: update only the split fields
        const size_t off = offsetof(ReplicaState, parts);
        const size_t len = offsetof(ReplicaState, gslots) + sizeof(void*) - off;
        check_hip(hipMemcpy(reinterpret_cast<char*>(d_states_) + off, reinterpret_cast<const char*>(&h) + off, len,
                            hipMemcpyHostToDevice),
                  "copy split fields");
    }
    if (stride < 1) throw std::runtime_error("run_split: stride >= 1");
    const int grid = general_blockmap(h_states_.data(), 1, h_map_, groups > 1 ? stride : 1);
    if (grid > map_cap_) {
        check_hip(hipStreamSynchronize(stream_), "batch sync");
        (void)hipFree(d_map_);
        d_map_ = nullptr;
        check_hip(hipMalloc(&d_map_, sizeof(int2) * static_cast<size_t>(grid)), "hipMalloc block map");
        map_cap_ = grid;
    }
    if (trace && trace_cap > trace_cap_) {
        check_hip(hipStreamSynchronize(stream_), "batch sync");
        (void)hipFree(d_trace_);
        check_hip(hipMalloc(&d_trace_, sizeof(TraceEntry) * static_cast<size_t>(trace_cap)), "hipMalloc trace");
        trace_cap_ = trace_cap;
    }
    check_hip(hipStreamSynchronize(stream_), "batch sync"); // the host copies below are read by the stream
    h_split_moves_ = moves;
    check_hip(hipMemcpyAsync(d_map_, h_map_.data(), sizeof(int2) * static_cast<size_t>(grid), hipMemcpyHostToDevice,
                             stream_),
              "copy block map");
    check_hip(hipMemcpyAsync(d_moves_, &h_split_moves_, sizeof(long long), hipMemcpyHostToDevice, stream_), "copy moves");
    time_begin();
    launch_general(h_states_.data(), 1, d_states_, d_moves_, fp32_ != 0, trace ? d_trace_ : nullptr,
                   trace ? trace_cap : 0, stream_, groups > 1 ? d_map_ : nullptr, groups > 1 ? grid : 0);
    time_end();
}

// The replica's error (1: a cell overflowed, 2: a wait timed out or the state is
// inconsistent), or the sync time-out flag of the last split or pipelined launch.
int Batch::split_error() {
    const int s = state(0).error;
    if (general_ == 1 || !split_err_) return s;
    int e = 0;
    check_hip(hipStreamSynchronize(stream_), "sync");
    check_hip(hipMemcpy(&e, split_err_, sizeof(int), hipMemcpyDeviceToHost), "copy err");
    return std::max(s, e);
}

} // namespace fmc
