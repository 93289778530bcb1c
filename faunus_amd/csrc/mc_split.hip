// One large replica advanced by many workgroups at once.  Workgroups 1..G-1 own slices
// of the k-vectors (Q(k) stays in registers for the whole launch); workgroup 0
// evaluates the pair dU through the cell list and commits accepted moves (with G = 1
// the one workgroup does both).  Every
// workgroup gathers the shares of each move (bounded spin), sums them with the same
// wave sum and so takes the same decision as every other workgroup.  Each workgroup
// has its share of the next move ready before that decision and publishes it right
// after: a k-slice holder sums the next move's de against Q and against Q + dQ of the
// pending move while its wave 0 gathers, and sends the one the decision selects;
// workgroup 0 walks the next move's stencils while its wave 0 gathers and sends the
// result unless the accepted move changed a loaded cell.
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
// change; without the copies (the no-allocation entry point) it also publishes, with each
// move's pair change, the particle of the move after next as committed before that move.
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

constexpr int kSplitMaxKPT = 12; // k-vectors per thread at 256 threads
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

// LDS of the split kernel, shared by its three workgroup roles
template <int kThreads>
struct SplitLds {
    DevPair pairs[kMaxTypes * kMaxTypes];
    DevCoulomb coul;
    double dp[kMaxTypes];
    double gk[kGkMax];
    // per-wave partials: 0/1 the next move's de if the current move is rejected/accepted,
    // 2/3 a pair walk's new/old sums, 4/5 the look-ahead's new/old sums
    double red[6][kThreads / 64];
    int cf[kThreads / 64]; // look-ahead waves: a loaded cell is one the current move changes
    int accept, abort_, stop, ovf;
    // proposals of the next 64 moves and the three after them, one lane each
    Proposal prop[67];
    double thr[67]; // -log(u) of each proposal's acceptance deviate
    double2 phase[6 * 65];
    double kb[4]; // k-slice holders: the particle of the move after next, as of the current move
    int kti;
    double msg[4];       // workgroups g > 0 without private copies: workgroup 0's particle message
    double pf_pos[4][4]; // the look-ahead's particle, by move mod 4
    int pf_ti[4], pf_ok;
    int cold[4], cnew[4]; // a move's old and new cell, by move mod 4
    // the commit's reads, made while wave 0 waits: the new cell's count, the old cell's
    // count, the moved particle's slot there and the old cell's last member
    int cm[4];
    // the particle of the move after next as committed before the current move (position,
    // charge; type and cell), loaded by wave 7 while wave 0 waits
    double la_pos[4];
    int la_ti, la_co;
    // and that move's 54 stencil cells (combo order), each cell's count and first two members
    int la_cell[54], la_cnt[54], la_m0[54], la_m1[54];
};

// Roles: 0 workgroup 0 of several (pair walk, commit), 1 a k-slice holder (g > 0),
// 2 a single workgroup that does both.  Each role is its own instantiation, so that
// registers live in one role do not count against the others.
template <int kThreads, int kMaxKPT, bool kCells, int kRole>
__device__ __attribute__((always_inline)) inline void split_body(SplitLds<kThreads>& sh, ReplicaState* __restrict__ states,
                                                             long long M, int G, int g, SplitSlot* __restrict__ slots,
                                                             int* __restrict__ err, TraceEntry* __restrict__ trace,
                                                             int trace_cap, long long spin_limit,
                                                             double4* __restrict__ priv) {
    ReplicaState& S = states[0];
    constexpr int kWaves = kThreads / 64;
    constexpr int kKT = kThreads - 64; // threads of the k-slice: waves 1.. (wave 0 gathers and decides)
    // Workgroup 0 at 512 threads with a cell list: waves 1-6 (7 lanes per stencil cell)
    // walk the next move's stencils while wave 0 waits for the partials of the current one.
    constexpr bool kPre = kCells && kThreads >= 512;
    constexpr int kPfLanes = 7;
    // The last wave publishes, so that wave 0's polls do not wait for those stores to
    // complete (the vector memory counter covers loads and stores).
    constexpr int kPub = kThreads - 64;
    auto& s_pairs = sh.pairs;
    auto& s_coul = sh.coul;
    auto& s_dp = sh.dp;
    auto& s_gk = sh.gk;
    auto& s_red = sh.red;
    auto& s_cf = sh.cf;
    int& s_accept = sh.accept;
    int& s_abort = sh.abort_;
    int& s_stop = sh.stop;
    int& s_ovf = sh.ovf;
    auto& s_prop = sh.prop;
    auto& s_thr = sh.thr;
    auto& s_phase = sh.phase;
    auto& s_kb = sh.kb;
    int& s_kti = sh.kti;
    auto& s_msg = sh.msg;
    auto& s_pf_pos = sh.pf_pos;
    auto& s_pf_ti = sh.pf_ti;
    int& s_pf_ok = sh.pf_ok;
    auto& s_cold = sh.cold;
    auto& s_cnew = sh.cnew;
    auto& s_cm = sh.cm;
    auto& s_la_pos = sh.la_pos;
    int& s_la_ti = sh.la_ti;
    int& s_la_co = sh.la_co;
    auto& s_la_cell = sh.la_cell;
    auto& s_la_cnt = sh.la_cnt;
    auto& s_la_m0 = sh.la_m0;
    auto& s_la_m1 = sh.la_m1;

    const int tid = threadIdx.x;
    const int wid = tid >> 6;
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

    // Roles.  Workgroup 0 walks the pair stencils and commits accepted moves.  The
    // k-vectors are sliced over workgroups 1..G-1 (over workgroup 0 when G = 1), each
    // slice interleaved over the threads of waves 1.. of its workgroup.
    constexpr bool pairs = kRole != 1;
    constexpr bool hasK = kRole != 0; // also a holder whose slice is empty: it publishes zero
    const int kparts = G > 1 ? G - 1 : 1;
    const int kw = (nk + kparts - 1) / kparts;
    const int kb = kRole == 1 ? (g - 1) * kw : 0;
    const int ke = kRole == 0 ? 0 : min(nk, kb + kw);
    int kn[kMaxKPT];
    double2 q[kMaxKPT], dq[kMaxKPT], dq1[kMaxKPT];
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) {
        const int k = kb + s * kKT + tid - 64;
        const bool on = tid >= 64 && k < ke;
        kn[s] = on ? S.kn[k] : 0;
        q[s] = on ? S.q[k] : make_double2(0.0, 0.0);
        dq[s] = make_double2(0.0, 0.0);
        dq1[s] = make_double2(0.0, 0.0);
    }
    double4* __restrict__ pos = S.pos;
    const uint32_t* __restrict__ meta = S.meta;
    // workgroups g > 0: this workgroup's private copy of the positions (null: none)
    double4* __restrict__ mine = (priv != nullptr && G > 1 && g > 0)
                                     ? priv + static_cast<size_t>(g - 1) * static_cast<size_t>(S.n)
                                     : nullptr;
    if (mine != nullptr)
        for (int j = tid; j < S.n; j += kThreads) mine[j] = pos[j];
    // workgroup 0 also publishes particles when the others keep no copies
    const bool particle_msg = priv == nullptr && G > 1;
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
    long long t_done = M; // moves decided by this launch
    // thread 0's values of the current move (this workgroup's share of the decision):
    // workgroup 0 the pair energy change, the others their de
    double own_pd = 0.0, own_de = 0.0;
    // k-slice holders: the current move's particle (old position, charge; new position)
    // and the next move's, as the phase tables of the next move were built
    double4 cur_po = make_double4(0.0, 0.0, 0.0, 0.0), nb_po = cur_po;
    double cur_nx = 0.0, cur_ny = 0.0, cur_nz = 0.0, nb_nx = 0.0, nb_ny = 0.0, nb_nz = 0.0;
    int cur_ti = 0, nb_ti = 0;
    // the publisher's early loads: the particle of the move after next (k-slice holders
    // with private copies, or workgroup 0 when G = 1 or when it sends particles) and of
    // the move after that (workgroup 0 when it sends particles), as committed so far
    double4 ahead2 = make_double4(0.0, 0.0, 0.0, 0.0), ahead3 = ahead2;
    int ahead2_ti = 0;

    auto build_phases = [&](double px, double py, double pz, double ox, double oy, double oz)
                            __attribute__((always_inline)) {
        for (int e = tid; e < 6 * width; e += kThreads) {
            const int nn = e % width - ncut;
            const int axis = (e / width) % 3;
            const int st = e / (3 * width);
            const double x = st == 0 ? (axis == 0 ? px : (axis == 1 ? py : pz)) : (axis == 0 ? ox : (axis == 1 ? oy : oz));
            double sn, cs;
            sincos((6.283185307179586 * (static_cast<double>(nn) / L)) * x, &sn, &cs);
            s_phase[e] = make_double2(cs, sn);
        }
    };
    // dQ of the move whose phase tables are in LDS (charge w), into dq1, and the thread's
    // share of its de against Q (x0) and against Q + dq (x1): the next move's de if the
    // current move is rejected, and if it is accepted
    auto kslice = [&](double w, double& x0, double& x1) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < kMaxKPT; ++s) {
            const int k = kb + s * kKT + tid - 64;
            if (k < ke) {
                const int nx = unpack_n(kn[s], 0), ny = unpack_n(kn[s], 1), nz = unpack_n(kn[s], 2);
                const double2 xn = s_phase[nx + ncut], yn = s_phase[width + ny + ncut], zn = s_phase[2 * width + nz + ncut];
                const double2 xo = s_phase[3 * width + nx + ncut], yo = s_phase[4 * width + ny + ncut],
                              zo = s_phase[5 * width + nz + ncut];
                const double an = xn.x * yn.x - xn.y * yn.y, bn = xn.x * yn.y + xn.y * yn.x;
                const double ao = xo.x * yo.x - xo.y * yo.y, bo = xo.x * yo.y + xo.y * yo.x;
                const double dre = w * ((an * zn.x - bn * zn.y) - (ao * zo.x - bo * zo.y));
                const double dim = w * ((an * zn.y + bn * zn.x) - (ao * zo.y + bo * zo.x));
                dq1[s] = make_double2(dre, dim);
                const double A = (nx > 0 ? 2.0 : 1.0) * s_gk[nx * nx + ny * ny + nz * nz];
                const double sq = dre * dre + dim * dim;
                x0 += A * (2.0 * (q[s].x * dre + q[s].y * dim) + sq);
                const double ax = q[s].x + dq[s].x, ay = q[s].y + dq[s].y;
                x1 += A * (2.0 * (ax * dre + ay * dim) + sq);
            } else {
                dq1[s] = make_double2(0.0, 0.0);
            }
        }
    };
    // the sum of a row of per-wave partials over waves w0..kWaves-1, in wave order
    auto row_sum = [&](int r, int w0, int w1) __attribute__((always_inline)) {
        double a = 0.0;
        for (int w = w0; w < w1; ++w) a += s_red[r][w];
        return a;
    };
    // workgroup 0's pair energy change as one value: -inf for a NaN old sum (and a new one
    // that is not NaN), +inf for a NaN or +inf new sum, else new - old.  The decision adds
    // the reciprocal change to a finite one; for a finite reciprocal change this is
    // getEnergyChange's handling of the two sums (montecarlo.cpp:193-209), and a NaN one
    // gives +inf as there
    auto pair_change = [](double pb, double pc) __attribute__((always_inline)) {
        return (isnan(pc) && !isnan(pb)) ? -__builtin_huge_val()
               : (isnan(pb) || (pb > 0.0 && isinf(pb))) ? __builtin_huge_val()
                                                         : pb - pc;
    };
    auto refill = [&](unsigned long long m0) __attribute__((always_inline)) {
        if (tid < 67) {
            const Proposal Q = make_proposal(seed, replica, m0 + static_cast<unsigned long long>(tid), n);
            s_prop[tid] = Q;
            s_thr[tid] = -log(Q.accept_u);
        }
    };

    // Workgroup 0's look-ahead (waves 1-6).  Its first part, for the move after next, runs
    // while the publisher commits the current move: the particle and the stencil cell of
    // the lane, and that cell's count and first two member slots.  A value the concurrent
    // commit changes is taken from the move itself (the moved particle, its new cell) or
    // read again before use (the two cells the commit changed).
    int la_tq = 0, la_in = -1, la_cell = -1, la_cnt = 0, la_r0 = -1, la_r1 = -1;
    double la_px = 0.0, la_py = 0.0, la_pz = 0.0, la_w = 0.0;
    auto la_prep = [&](long long tn, int ps, int ci, bool cacc, double cx, double cy, double cz, int ccell,
                       bool staged) __attribute__((always_inline)) {
        const int combo = (tid - 64) % 54, slot0 = (tid - 64) / 54;
        la_cell = -1;
        if (tn < M && slot0 < kPfLanes) {
            const Proposal& Q = s_prop[ps];
            const int in = Q.index;
            double4 qo;
            int tq, co;
            if (staged) {
                qo = make_double4(s_la_pos[0], s_la_pos[1], s_la_pos[2], s_la_pos[3]);
                tq = s_la_ti;
                co = s_la_co;
            } else {
                qo = pos[in];
                tq = static_cast<int>(meta[in] & 0xffu);
                co = tid == 64 ? cell_of[in] : 0;
            }
            if (cacc && in == ci) {
                qo.x = cx;
                qo.y = cy;
                qo.z = cz;
                co = ccell;
            }
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
                s_pf_pos[tn & 3][0] = qo.x;
                s_pf_pos[tn & 3][1] = qo.y;
                s_pf_pos[tn & 3][2] = qo.z;
                s_pf_pos[tn & 3][3] = qo.w;
                s_pf_ti[tn & 3] = tq;
                s_cold[tn & 3] = co;
            }
            // the cells, counts and first two members staged by wave 7, unless this move
            // moved the particle
            const bool fresh = !staged || (cacc && in == ci);
            int f_cell;
            if (fresh) {
                const int cx = (cell_coord_s(f_px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                const int cy = (cell_coord_s(f_py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                const int cz = (cell_coord_s(f_pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                f_cell = (cz * nc1 + cy) * nc0 + cx;
                la_cnt = cell_count[f_cell];
                la_r0 = slot0 < ccap ? cell_members[f_cell * ccap + slot0] : -1;
            } else {
                f_cell = s_la_cell[combo];
                la_cnt = s_la_cnt[combo];
                la_r0 = slot0 == 0 ? s_la_m0[combo]
                                   : (slot0 == 1 ? s_la_m1[combo]
                                                 : (slot0 < ccap ? cell_members[f_cell * ccap + slot0] : -1));
            }
            if (combo == 13 && slot0 == 0) s_cnew[tn & 3] = f_cell; // the new position's own cell
            la_r1 = slot0 + kPfLanes < ccap ? cell_members[f_cell * ccap + slot0 + kPfLanes] : -1;
            la_in = in;
            la_tq = tq;
            la_w = qo.w;
            la_px = f_px;
            la_py = f_py;
            la_pz = f_pz;
            la_cell = f_cell;
        }
    };
    bool prev_acc = false; // workgroup 0: the previous move was accepted (its commit overlapped the loads)

    refill(move0);
    __syncthreads();
    if (hasK && M > 0) {
        // the launch's first move, against the state the launch starts from: no workgroup
        // commits before every k-slice holder has published this move's de
        const Proposal& P0 = s_prop[0];
        const int i0 = P0.index;
        cur_po = pos[i0];
        cur_ti = static_cast<int>(meta[i0] & 0xffu);
        const double sc = s_dp[cur_ti] * P0.length_u;
        cur_nx = wrap_s(cur_po.x + P0.dir[0] * sc, L, half, inv);
        cur_ny = wrap_s(cur_po.y + P0.dir[1] * sc, L, half, inv);
        cur_nz = wrap_s(cur_po.z + P0.dir[2] * sc, L, half, inv);
        if (tid == kPub && M > 1) {
            const int i1 = s_prop[1].index;
            ahead2 = pos[i1];
            ahead2_ti = static_cast<int>(meta[i1] & 0xffu);
        }
        build_phases(cur_nx, cur_ny, cur_nz, cur_po.x, cur_po.y, cur_po.z);
        __syncthreads();
        if (tid >= 64) {
            double x0 = 0.0, x1 = 0.0;
            kslice(cur_po.w, x0, x1);
            const double a = wave_sum(x0);
            if ((tid & 63) == 0) s_red[0][wid] = a;
        }
#pragma unroll
        for (int s = 0; s < kMaxKPT; ++s) dq[s] = dq1[s];
        __syncthreads();
        if (tid == 0 || tid == kPub) {
            const double de0 = row_sum(0, 1, kWaves);
            if (tid == 0) own_de = de0;
            if (tid == kPub && G > 1) split_put(slots + static_cast<int>(move0 & 1ull) * G + g, 0, de0, split_tag(move0));
            if (tid == kPub && M > 1) {
                s_kb[0] = ahead2.x;
                s_kb[1] = ahead2.y;
                s_kb[2] = ahead2.z;
                s_kb[3] = ahead2.w;
                s_kti = ahead2_ti;
            }
        }
        __syncthreads();
    }

    for (long long t = 0; t < M; ++t) {
        const unsigned long long mid = move0 + static_cast<unsigned long long>(t);
        if (t > 0 && (t & 63) == 0) {
            // the previous proposals were last read before the previous move's last barrier
            refill(mid);
            __syncthreads();
        }
        const int slot = static_cast<int>(t & 63);
        const Proposal P = s_prop[slot];
        const int i = P.index;
        const int i1 = s_prop[slot + 1].index, i2 = s_prop[slot + 2].index, i3 = s_prop[slot + 3].index;
        const int parity = static_cast<int>(mid & 1ull);
        const uint32_t tag = split_tag(mid);
        if (tid == kPub) {
            // committed so far: this workgroup's commits and private-copy updates are visible
            // to its own threads
            if (hasK && t + 2 < M && (mine != nullptr || pairs)) {
                ahead2 = mine != nullptr ? mine[i2] : pos[i2];
                ahead2_ti = static_cast<int>(meta[i2] & 0xffu);
            }
            if (pairs && particle_msg) {
                if (!hasK) ahead2 = pos[i2];
                ahead3 = pos[i3];
            }
            if (hasK && mine == nullptr && !pairs) ahead2_ti = static_cast<int>(meta[i2] & 0xffu);
        }
        // the current move's particle: workgroup 0 from the look-ahead or the positions,
        // the k-slice holders from their registers
        const bool pre = kPre && pairs && s_pf_ok;
        double4 po;
        int ti;
        double nx_, ny_, nz_;
        if (pairs) {
            if (pre) {
                po = make_double4(s_pf_pos[t & 3][0], s_pf_pos[t & 3][1], s_pf_pos[t & 3][2], s_pf_pos[t & 3][3]);
                ti = s_pf_ti[t & 3];
            } else {
                po = pos[i];
                ti = static_cast<int>(meta[i] & 0xffu);
            }
            const double scale = s_dp[ti] * P.length_u;
            nx_ = wrap_s(po.x + P.dir[0] * scale, L, half, inv);
            ny_ = wrap_s(po.y + P.dir[1] * scale, L, half, inv);
            nz_ = wrap_s(po.z + P.dir[2] * scale, L, half, inv);
        } else {
            po = cur_po;
            ti = cur_ti;
            nx_ = cur_nx;
            ny_ = cur_ny;
            nz_ = cur_nz;
        }

        // (1) workgroup 0 without a valid look-ahead: walk this move's stencils now and publish
        if (pairs && !pre) {
            double e_new = 0.0, e_old = 0.0;
            const Box b{{L, L, L}, {half, half, half}};
            if (kPre) {
                const int combo = (tid - 64) % 54, slot0 = (tid - 64) / 54;
                if (tid == 64) s_cold[t & 3] = cell_of[i];
                if (tid >= 64 && slot0 < kPfLanes) {
                    const int st = combo / 27, c = combo % 27;
                    const double px = st == 0 ? nx_ : po.x, py = st == 0 ? ny_ : po.y, pz = st == 0 ? nz_ : po.z;
                    const int cx = (cell_coord_s(px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                    const int cy = (cell_coord_s(py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                    const int cz = (cell_coord_s(pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                    const int cell = (cz * nc1 + cy) * nc0 + cx;
                    if (combo == 13 && slot0 == 0) s_cnew[t & 3] = cell; // the new position's own cell
                    const int cnt = cell_count[cell];
                    double acc_u = 0.0;
                    for (int k = slot0; k < cnt; k += kPfLanes) {
                        const int j = cell_members[cell * ccap + k];
                        if (j == i) continue;
                        const double4 pj = pos[j];
                        const int tj = static_cast<int>(meta[j] & 0xffu);
                        acc_u += pair_energy(s_pairs[ti * nt + tj], s_coul, po.w, pj.w, dist2(b, px, py, pz, pj.x, pj.y, pj.z));
                    }
                    if (st == 0) e_new = acc_u;
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
            {
                const double a = wave_sum(e_new), c = wave_sum(e_old);
                if ((tid & 63) == 0) {
                    s_red[2][wid] = a;
                    s_red[3][wid] = c;
                }
            }
            __syncthreads();
            if (tid == 0 || tid == kPub) {
                const double pd = pair_change(row_sum(2, 0, kWaves), row_sum(3, 0, kWaves));
                if (tid == 0) own_pd = pd;
                if (tid == kPub && G > 1) {
                    SplitSlot* sl = slots + parity * G;
                    if (particle_msg) {
                        // the particle's load is waited for here, before the stores: a wait
                        // after them would wait for the stores too
                        asm volatile("" ::"v"(ahead2.x), "v"(ahead2.y), "v"(ahead2.z), "v"(ahead2.w));
                    }
                    split_put(sl, 0, pd, tag);
                    if (particle_msg) {
                        // the particle of the move after next, as committed before this move
                        split_put(sl, 1, ahead2.x, tag);
                        split_put(sl, 2, ahead2.y, tag);
                        split_put(sl, 3, ahead2.z, tag);
                        split_put(sl, 4, ahead2.w, tag);
                    }
                }
            }
        }

        // (2) k-slice holders: the next move's phase tables.  Its particle is taken as
        // committed before this move; if this move is accepted and moves the same particle,
        // the tables are built again after the decision.
        if (hasK) {
            if (t + 1 < M) {
                nb_po = make_double4(s_kb[0], s_kb[1], s_kb[2], s_kb[3]);
                nb_ti = s_kti;
                const Proposal& P1 = s_prop[slot + 1];
                const double sc1 = s_dp[nb_ti] * P1.length_u;
                nb_nx = wrap_s(nb_po.x + P1.dir[0] * sc1, L, half, inv);
                nb_ny = wrap_s(nb_po.y + P1.dir[1] * sc1, L, half, inv);
                nb_nz = wrap_s(nb_po.z + P1.dir[2] * sc1, L, half, inv);
                build_phases(nb_nx, nb_ny, nb_nz, nb_po.x, nb_po.y, nb_po.z);
            }
            __syncthreads(); // phase tables
        }

        // (3) while wave 0 gathers this move's partials and decides: the next move's de for
        // both outcomes (k-slice holders), and the next move's stencil sums as committed
        // before this move, from the loads started during the previous commit (workgroup
        // 0's waves 1-6)
        if (kPre && pairs && tid >= 64 && tid < 448) {
            if (t == 0) la_prep(1, 1, -1, false, 0.0, 0.0, 0.0, 0, false); // the launch's first look-ahead
            const int combo = (tid - 64) % 54, slot0 = (tid - 64) / 54;
            double fs = 0.0;
            int cf = 0;
            if (t + 1 < M && la_cell >= 0) {
                const Box b{{L, L, L}, {half, half, half}};
                const int in = la_in, tq = la_tq, f_cell = la_cell;
                if (t > 0 && prev_acc && (f_cell == s_cold[(t - 1) & 3] || f_cell == s_cnew[(t - 1) & 3])) {
                    // the previous move's commit changed this cell while it was read
                    la_cnt = cell_count[f_cell];
                    la_r0 = slot0 < ccap ? cell_members[f_cell * ccap + slot0] : -1;
                    la_r1 = slot0 + kPfLanes < ccap ? cell_members[f_cell * ccap + slot0 + kPfLanes] : -1;
                }
                cf = (f_cell == s_cold[t & 3] || f_cell == s_cnew[t & 3]) ? 1 : 0;
                const int f_cnt = la_cnt;
                // the first member's position is loaded at once; a second member slot of
                // the lane (cells of more than 7) in the loop
                const int j0 = slot0 < f_cnt ? la_r0 : -1;
                double4 p0 = make_double4(0.0, 0.0, 0.0, 0.0);
                int t0 = 0;
                if (j0 >= 0) {
                    p0 = pos[j0];
                    t0 = static_cast<int>(meta[j0] & 0xffu);
                }
                for (int k = slot0, u = 0; k < f_cnt; k += kPfLanes, ++u) {
                    int j, tj;
                    double4 pj;
                    if (u == 0) {
                        j = j0;
                        pj = p0;
                        tj = t0;
                    } else {
                        j = u == 1 ? la_r1 : cell_members[f_cell * ccap + k];
                        if (j == in) continue;
                        pj = pos[j];
                        tj = static_cast<int>(meta[j] & 0xffu);
                    }
                    if (j == in) continue;
                    fs += pair_energy(s_pairs[tq * nt + tj], s_coul, la_w, pj.w, dist2(b, la_px, la_py, la_pz, pj.x, pj.y, pj.z));
                }
            }
            const double a = wave_sum(combo < 27 ? fs : 0.0), c = wave_sum(combo < 27 ? 0.0 : fs);
            const int any_cf = __any(cf);
            if ((tid & 63) == 0) {
                s_red[4][wid] = a;
                s_red[5][wid] = c;
                s_cf[wid] = any_cf;
            }
        }
        if (kPre && pairs && wid == kWaves - 1) {
            // the commit's reads (wave 7): the cell counts, the moved particle's slot in its
            // old cell (one member per lane) and that cell's last member; and whether the
            // current move, if accepted, overflows its new cell.  Lane 1 also stages the
            // particle of the move after next for the look-ahead that starts at the commit.
            const int cn = s_cnew[t & 3], co = s_cold[t & 3];
            const int lane = tid & 63;
            if (t + 2 < M) {
                // lane c < 54 also stages stencil combo c of that move: its cell, the cell's
                // count and first two members (the look-ahead's combo and state order)
                const Proposal& Q2 = s_prop[slot + 2];
                const int in2 = Q2.index;
                const double4 p2 = pos[in2];
                const int t2 = static_cast<int>(meta[in2] & 0xffu);
                if (lane == 1) {
                    s_la_pos[0] = p2.x;
                    s_la_pos[1] = p2.y;
                    s_la_pos[2] = p2.z;
                    s_la_pos[3] = p2.w;
                    s_la_ti = t2;
                    s_la_co = cell_of[in2];
                }
                if (lane < 54) {
                    const double sc2 = s_dp[t2] * Q2.length_u;
                    const int st = lane / 27, c = lane % 27;
                    double px = p2.x, py = p2.y, pz = p2.z;
                    if (st == 0) {
                        px = wrap_s(p2.x + Q2.dir[0] * sc2, L, half, inv);
                        py = wrap_s(p2.y + Q2.dir[1] * sc2, L, half, inv);
                        pz = wrap_s(p2.z + Q2.dir[2] * sc2, L, half, inv);
                    }
                    const int cx = (cell_coord_s(px, half, inv_cell, nc0) + (c % 3) - 1 + nc0) % nc0;
                    const int cy = (cell_coord_s(py, half, inv_cell, nc1) + ((c / 3) % 3) - 1 + nc1) % nc1;
                    const int cz = (cell_coord_s(pz, half, inv_cell, nc2) + (c / 9) - 1 + nc2) % nc2;
                    const int cell = (cz * nc1 + cy) * nc0 + cx;
                    s_la_cell[lane] = cell;
                    s_la_cnt[lane] = cell_count[cell];
                    s_la_m0[lane] = ccap > 0 ? cell_members[cell * ccap] : -1;
                    s_la_m1[lane] = ccap > 1 ? cell_members[cell * ccap + 1] : -1;
                }
            }
            int nn = 0, nold = 0, kk = -1;
            if (cn != co) {
                nn = cell_count[cn];
                nold = cell_count[co];
                for (int b0 = 0; b0 < nold; b0 += 64) {
                    const int l = b0 + lane;
                    const int mj = l < nold ? cell_members[co * ccap + l] : -1;
                    const unsigned long long hit = __ballot(mj == i);
                    if (hit != 0ull && kk < 0) kk = b0 + static_cast<int>(__builtin_ctzll(hit));
                    if (l == nold - 1) s_cm[3] = mj;
                }
            }
            if (lane == 0) {
                s_ovf = (cn != co && nn >= ccap) ? 1 : 0;
                s_cm[0] = nn;
                s_cm[1] = nold;
                s_cm[2] = kk;
            }
        }
        if (hasK && tid >= 64) {
            double x0 = 0.0, x1 = 0.0;
            if (t + 1 < M) kslice(nb_po.w, x0, x1);
            const double a = wave_sum(x0), c = wave_sum(x1);
            if ((tid & 63) == 0) {
                s_red[0][wid] = a;
                s_red[1][wid] = c;
            }
        }
        // gather: lane j < G takes workgroup j's value (workgroup 0's pair change, the
        // others' de); lanes 58-61 take workgroup 0's particle when it is sent.  A
        // workgroup takes its own value from registers, bit for bit what the others read.
        // The des are added with the same wave sum in every workgroup.
        double du = 0.0;
        if (tid < 64) {
            int ok = 1, stopped = 0;
            double a = 0.0, pdv = 0.0, extra = 0.0;
            const double own_d = readlane_f64(own_de, 0), own_p = readlane_f64(own_pd, 0);
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
            if (G == 1) {
                if (tid == 0) {
                    pdv = own_p;
                    a = own_d;
                }
            } else if (tid < G) {
                if (tid == g) {
                    if (g == 0) pdv = own_p;
                    else a = own_d;
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
                if (any_stop) {
                    s_stop = 2; // workgroup 0 stopped after the previous move: this one is not decided
                } else if (!all_ok) {
                    atomicExch(err, 1);
                    atomicMax(&S.error, 2);
                    s_abort = 1;
                } else {
                    if (particle_msg && g != 0) {
                        s_msg[0] = ex;
                        s_msg[1] = ey;
                        s_msg[2] = ez;
                        s_msg[3] = eq;
                    }
                    // the surface term enters with the reciprocal one (Ewald::energy, energy.cpp:581-587)
                    double dsurf = 0.0, mx = dip0, my = dip1, mz = dip2;
                    if (surf != 0.0) {
                        mx += po.w * (nx_ - po.x);
                        my += po.w * (ny_ - po.y);
                        mz += po.w * (nz_ - po.z);
                        dsurf = surf * ((mx * mx + my * my + mz * mz) - (dip0 * dip0 + dip1 * dip1 + dip2 * dip2));
                    }
                    const double drec = pref * a + dsurf;
                    du = isnan(drec) ? __builtin_huge_val() : (isinf(pdt) ? pdt : pdt + drec);
                    if (isnan(du)) du = 0.0;
                    const double bdu = beta * du;
                    bool acc;
                    if (isinf(bdu) && bdu < 0.0) acc = true;
                    else if (-bdu > 709.782712893384) acc = true;
                    else {
                        // u <= exp(-bdu) decided against the table's -log(u) away from the
                        // boundary (the margin exceeds the rounding of exp and log many times),
                        // with exp itself only within the margin: the same decision, without
                        // exp on the chain
                        const double thr = s_thr[slot], margin = 1e-12 * (1.0 + fabs(thr));
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
                }
            }
        }
        __syncthreads(); // the decision, the partials of the next move
        if (s_abort || s_stop == 2) {
            t_done = t;
            break;
        }
        const bool acc = s_accept != 0;

        // (4) publish the next move's values at once, then apply this move
        const bool refresh = hasK && acc && t + 1 < M && i1 == i; // the next move's particle just moved
        if (hasK) {
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) {
                if (acc) {
                    q[s].x += dq[s].x;
                    q[s].y += dq[s].y;
                }
                dq[s] = dq1[s];
            }
        }
        if (tid == 0 || tid == kPub) {
            const unsigned long long mn = mid + 1ull;
            SplitSlot* sl = slots + static_cast<int>(mn & 1ull) * G + g;
            const uint32_t tag1 = split_tag(mn);
            if (tid == kPub && hasK && t + 2 < M) {
                // the particle of the move after next as of this move: from the private copy
                // or the positions (loaded at this move's start), or from workgroup 0's
                // message, with this move applied.  First, so that the wait for the early
                // load does not also wait for the stores below.
                double4 a2 = ahead2;
                if (mine == nullptr && !pairs) a2 = make_double4(s_msg[0], s_msg[1], s_msg[2], s_msg[3]);
                const bool moved = acc && i2 == i;
                s_kb[0] = moved ? nx_ : a2.x;
                s_kb[1] = moved ? ny_ : a2.y;
                s_kb[2] = moved ? nz_ : a2.z;
                s_kb[3] = a2.w;
                s_kti = ahead2_ti;
            }
            if (kPre && pairs && t + 1 < M) {
                int cf = 0;
                for (int w = 1; w < 7; ++w) cf |= s_cf[w];
                const bool stop = acc && s_ovf;
                const bool valid = !(acc && (i1 == i || s_ovf || cf));
                const double pd1 = pair_change(row_sum(4, 1, 7), row_sum(5, 1, 7));
                if (tid == 0 && valid) own_pd = pd1;
                if (tid == kPub) {
                    s_pf_ok = valid ? 1 : 0;
                    if (G > 1 && stop) {
                        // the cell overflows: workgroup 0 stops after this move, and its words
                        // of the next move carry the abort bit so that the others stop there
                        for (int v = 0; v < kSplitValues; ++v) split_put(sl, v, 0.0, tag1 | kAbortBit);
                    } else if (G > 1 && valid) {
                        if (particle_msg) asm volatile("" ::"v"(ahead3.x), "v"(ahead3.y), "v"(ahead3.z), "v"(ahead3.w));
                        split_put(sl, 0, pd1, tag1);
                        if (particle_msg) {
                            // the particle of the move after next as of this move
                            const bool moved = acc && i3 == i;
                            split_put(sl, 1, moved ? nx_ : ahead3.x, tag1);
                            split_put(sl, 2, moved ? ny_ : ahead3.y, tag1);
                            split_put(sl, 3, moved ? nz_ : ahead3.z, tag1);
                            split_put(sl, 4, ahead3.w, tag1);
                        }
                    }
                }
            }
            if (hasK && t + 1 < M && !refresh) {
                const double de1 = acc ? row_sum(1, 1, kWaves) : row_sum(0, 1, kWaves);
                if (tid == 0) own_de = de1;
                if (tid == kPub && G > 1) split_put(sl, 0, de1, tag1);
            }
            if (tid == kPub && hasK && acc && mine != nullptr) mine[i] = make_double4(nx_, ny_, nz_, po.w);
            if (kPre && pairs && tid == kPub && acc) {
                // the commit, after the publish: stores only, from the reads made during the wait
                pos[i] = make_double4(nx_, ny_, nz_, po.w);
                const int cn = s_cnew[t & 3], co = s_cold[t & 3];
                if (cn != co && !s_ovf) {
                    const int nn = s_cm[0], nold = s_cm[1], kk = s_cm[2];
                    if (kk >= 0) cell_members[co * ccap + kk] = s_cm[3];
                    cell_count[co] = nold - 1;
                    cell_members[cn * ccap + nn] = i;
                    cell_count[cn] = nn + 1;
                    cell_of[i] = cn;
                }
            }
        }
        if (pairs && tid == 0) {
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
                if (kPre) {
                    // the publisher commits; a full new cell stops the chain (the lists are
                    // left as they are)
                    if (s_ovf) {
                        atomicMax(&S.error, 1);
                        s_stop = 1;
                    }
                } else {
                    pos[i] = make_double4(nx_, ny_, nz_, po.w);
                    if (kCells) {
                        const int cnew = (cell_coord_s(nz_, half, inv_cell, nc2) * nc1 +
                                          cell_coord_s(ny_, half, inv_cell, nc1)) * nc0 +
                                         cell_coord_s(nx_, half, inv_cell, nc0);
                        const int cold = cell_of[i];
                        if (cnew != cold) {
                            const int nn = cell_count[cnew];
                            if (nn >= ccap) {
                                // full cell: the lists are left as they are, the chain stops,
                                // and the next move's words carry the abort bit
                                atomicMax(&S.error, 1);
                                s_stop = 1;
                                if (G > 1 && t + 1 < M) {
                                    SplitSlot* sl = slots + static_cast<int>((mid + 1ull) & 1ull) * G;
                                    for (int v = 0; v < kSplitValues; ++v)
                                        split_put(sl, v, 0.0, split_tag(mid + 1ull) | kAbortBit);
                                }
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
            }
            if (sample_every > 0 && isfinite(energy) &&
                (mid + 1ull) % static_cast<unsigned long long>(sample_every) == 0ull) {
                e_sum += energy;
                e_sum2 += energy * energy;
                ++e_samples;
            }
        }
        if (refresh) {
            // the next move's particle is this move's: its phase tables and de again, against
            // Q with this move applied (the tables were last read before the barrier above)
            nb_po = make_double4(nx_, ny_, nz_, po.w);
            nb_ti = ti;
            const Proposal& P1 = s_prop[slot + 1];
            const double sc1 = s_dp[nb_ti] * P1.length_u;
            nb_nx = wrap_s(nb_po.x + P1.dir[0] * sc1, L, half, inv);
            nb_ny = wrap_s(nb_po.y + P1.dir[1] * sc1, L, half, inv);
            nb_nz = wrap_s(nb_po.z + P1.dir[2] * sc1, L, half, inv);
            build_phases(nb_nx, nb_ny, nb_nz, nb_po.x, nb_po.y, nb_po.z);
            __syncthreads();
            if (tid >= 64) {
                double x0 = 0.0, x1 = 0.0;
                kslice(nb_po.w, x0, x1);
                const double a = wave_sum(x0);
                if ((tid & 63) == 0) s_red[0][wid] = a;
            }
#pragma unroll
            for (int s = 0; s < kMaxKPT; ++s) dq[s] = dq1[s];
            __syncthreads();
            if (tid == 0 || tid == kPub) {
                const double de1 = row_sum(0, 1, kWaves);
                if (tid == 0) own_de = de1;
                if (tid == kPub && G > 1)
                    split_put(slots + static_cast<int>((mid + 1ull) & 1ull) * G + g, 0, de1, split_tag(mid + 1ull));
            }
        }
        if (kPre && pairs) {
            // the look-ahead of the move after next starts during the commit
            if (tid >= 64 && tid < 448) la_prep(t + 2, slot + 2, i, acc, nx_, ny_, nz_, s_cnew[t & 3], true);
            prev_acc = acc;
        }
        if (hasK) {
            cur_po = nb_po;
            cur_ti = nb_ti;
            cur_nx = nb_nx;
            cur_ny = nb_ny;
            cur_nz = nb_nz;
        }
        __syncthreads(); // commits, the next move's look-ahead flag and particle
        if (s_stop == 1) {
            t_done = t + 1;
            break;
        }
    }
#pragma unroll
    for (int s = 0; s < kMaxKPT; ++s) {
        const int k = kb + s * kKT + tid - 64;
        if (tid >= 64 && k < ke) S.q[k] = q[s];
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

template <int kThreads, int kMaxKPT, bool kCells, bool kSingle>
__global__ __launch_bounds__(kThreads) void mc_split_kernel(ReplicaState* __restrict__ states, long long M, int G,
                                                       int stride, SplitSlot* __restrict__ slots,
                                                       int* __restrict__ err, TraceEntry* __restrict__ trace,
                                                       int trace_cap, long long spin_limit, double4* __restrict__ priv) {
    if (blockIdx.x % stride != 0) return;
    const int g = blockIdx.x / stride;
    if (g >= G) return;
    if (states[0].error) return; // a stopped replica (read before any workgroup of this launch can set it)
    __shared__ SplitLds<kThreads> sh;
    // one workgroup (kSingle, G = 1) is its own instantiation: it holds both roles' registers
    if (kSingle) split_body<kThreads, kMaxKPT, kCells, 2>(sh, states, M, G, g, slots, err, trace, trace_cap, spin_limit, priv);
    else if (g == 0)
        split_body<kThreads, kMaxKPT, kCells, 0>(sh, states, M, G, g, slots, err, trace, trace_cap, spin_limit, priv);
    else split_body<kThreads, kMaxKPT, kCells, 1>(sh, states, M, G, g, slots, err, trace, trace_cap, spin_limit, priv);
}

// The split kernel for a replica over `groups` workgroups, every `stride`-th block of the
// grid: 512 threads (two waves per SIMD, 7 lanes per stencil cell) when the slice fits 4
// k-vectors per thread of waves 1..7, else 256 threads with up to 12 per thread of waves
// 1..3; the grid must be co-resident.
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
    // the k-slice is spread over waves 1.. of a workgroup (wave 0 gathers and decides)
    const int threads = kw <= 4 * (512 - 64) ? 512 : 256;
    if ((kw + threads - 65) / (threads - 64) > (threads == 512 ? 4 : kSplitMaxKPT))
        throw std::runtime_error(w + ": too many k-vectors per workgroup");
    if (2 * s.ncut + 1 > 65 || s.ncut * s.ncut >= kGkMax) throw std::runtime_error(w + ": ncutoff too large");
    int dev = 0, ncu = 0, per_cu = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const bool one = groups == 1;
    const void* fn =
        threads == 512
            ? (s.cells ? (one ? reinterpret_cast<const void*>(mc_split_kernel<512, 4, true, true>)
                              : reinterpret_cast<const void*>(mc_split_kernel<512, 4, true, false>))
                       : (one ? reinterpret_cast<const void*>(mc_split_kernel<512, 4, false, true>)
                              : reinterpret_cast<const void*>(mc_split_kernel<512, 4, false, false>)))
            : (s.cells ? (one ? reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, true, true>)
                              : reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, true, false>))
                       : (one ? reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, false, true>)
                              : reinterpret_cast<const void*>(mc_split_kernel<256, kSplitMaxKPT, false, false>)));
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
