// Block triangular solve with many right-hand sides: replaces
// BlockPlanner::solve / lowerSolver / upperSolver / updateVector
// (BlockPlanner.cpp:669-862), which recurse over the L and U quadtrees with
// one right-hand side, serially.
//
// Here each block row (with 1-15 right-hand sides: each pair of block rows) is a task
// of a persistent kernel: forward,
//     y_I = Linv_II (b_I - sum_{J<I} L_IJ y_J)
// backward,
//     x_I = Uinv_II (y_I - sum_{J>I} U_IJ x_J)
// with the inverses of the diagonal tiles (and, for row pairs, the off-diagonal block
// of the pair's 2x2 diagonal block's inverse) made once per factor.  Tasks are taken in
// dependency order (ascending / descending) and a task waits only for the rows it
// reads, so independent rows overlap and the chain only serialises where the factor's
// block structure does.  With 1-15 right-hand sides a task waits on the stored values
// themselves (a pending marker until stored); the MFMA form for more right-hand sides
// waits on per-row ready flags.  Sums run in FP64 for every tile precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "device.h"
#include "tile_ops.h"

namespace sogl_dev {

constexpr long long SOLVE_SPIN_LIMIT = 20LL * 100 * 1000 * 1000;
constexpr int TRSM_MIN_RHS = 16;   // from this many right-hand sides the block-row sums run on MFMA
constexpr int TRSM_MAX_RHS = 256;  // columns per MFMA solve launch (4 panels of 64)

struct SolveArgs {
    const void* pool;
    const int32_t* rowptr;
    const int32_t* col;
    const int32_t* slot;
    const int32_t* dinv;  // per block row: slot of the inverted diagonal tile
    const double* in;
    double* out;
    int64_t ld;           // column stride of in/out (padded n)
    int32_t nrhs;
    int32_t nbr;
    int32_t forward;
    int32_t epoch;
    int* ready;
    int* ctrl;            // [0] row counter, [1] abort
    const uint32_t* mask; // per slot: the factor's 16x16 block masks (a zero tile is skipped)
    const int32_t* pairinv;  // per row pair: slot of the 2x2 diagonal block's inverse's
                             // off-diagonal block (the one-to-15-RHS form), -1: zero
};

// The abort word and the wall clock are looked at every SOLVE_SPIN_CHECK polls only, so
// that each poll costs one round trip to the device coherence point.
constexpr int SOLVE_SPIN_CHECK = 64;

__device__ inline void spin_row(const SolveArgs& a, int J) {
    int* flag = a.ready + J;
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.epoch) return;
    const long long t0 = wall_clock64();
    for (int spins = 1;; spins++) {
        __builtin_amdgcn_s_sleep(1);
        if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= a.epoch) return;
        if (spins % SOLVE_SPIN_CHECK) continue;
        if (__hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        if (wall_clock64() - t0 > SOLVE_SPIN_LIMIT) {
            __hip_atomic_store(a.ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
}

// Rows of the output that a one-to-15-RHS pass has not stored yet hold this signalling
// NaN, written by k_fill_pending before the pass.  FP arithmetic never produces a
// signalling NaN (a NaN result is quiet), so a stored value is never mistaken for it.
// A row task sorts the rows it reads into those already published when it starts (an
// oldest-first prefix: one check of their ready flags, one acquire, then ordinary
// cached loads) and the rest, the newest, usually the one on the chain: for those it
// polls the values themselves with agent-scope atomic loads until none of the 16 it
// needs is the marker.  The task stores its row with agent-scope atomic stores (they
// reach the coherence point without waiting for the release), then releases and
// raises its ready flag for the tasks that start later.
constexpr unsigned long long SOLVE_PENDING = 0x7FF4C0DE5EED0001ull;

__global__ void k_fill_pending(double* out, int64_t ld, int64_t rows, int nrhs) {
    const int64_t n = rows * nrhs;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / rows, r = i - c * rows;
        reinterpret_cast<unsigned long long*>(out)[c * ld + r] = SOLVE_PENDING;
    }
}

// The 64 values of one column of row J, which another task stores, into this wave's
// LDS slot ls once all are stored: lane l polls value l (one coalesced load per spin
// for the wave), the wave's ballot says when none is pending.  NC = 2 polls the same
// column of a second row (col + step) in the same spin, into ls[64 ..]: the two rows a
// row-pair task stores together cost one wait.  On the wait bound (or an abort
// elsewhere) it raises the abort flag and leaves zeros.  The caller reads ls after it
// (LDS accesses of one wave complete in order).
template <int NC = 1>
__device__ inline void load_row_stored(const double* col, double* ls, int lane, const SolveArgs& a,
                                       int64_t step = 0) {
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(col) + lane;
    long long t0 = 0;
    unsigned long long u[NC];
    for (int spin = 0;; spin++) {
        bool pend = false;
#pragma unroll
        for (int i = 0; i < NC; i++) {
            u[i] = __hip_atomic_load(q + i * step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pend |= u[i] == SOLVE_PENDING;
        }
        if (!__ballot(pend)) break;
        if (spin == 0) {
            t0 = wall_clock64();
        } else if (spin % SOLVE_SPIN_CHECK == 0 &&
                   (__hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    wall_clock64() - t0 > SOLVE_SPIN_LIMIT)) {
            __hip_atomic_store(a.ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int i = 0; i < NC; i++) u[i] = 0;  // the bit pattern of 0.0
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
#pragma unroll
    for (int i = 0; i < NC; i++) ls[i * TS + lane] = __longlong_as_double((long long)u[i]);
    __builtin_amdgcn_wave_barrier();
}

// Two block rows per task (a row pair p: rows 2p, 2p+1), pairs taken in dependency
// order, so each hand-off along the chain serves two rows.  With A the pair's row solved
// first (forward 2p, backward 2p+1) and B the other, and r_A, r_B the right-hand sides
// less the products of every earlier row (B's entry in column A left out):
//     x_A = D_A^-1 r_A,   x_B = D_B^-1 r_B + C_p r_A,
// where C_p is the off-diagonal block of the inverse of the pair's 2x2 diagonal block
// (forward X_p = -L_BB^-1 L_BA L_AA^-1, backward Y_p = -U_BB^-1 U_BA U_AA^-1; pairinv,
// -1 when L_BA / U_BA is absent and the block is zero), made by two GEMM tasks of the
// factor.  Waves 0-3 (group 0) sum row A, waves 4-7 (group 1) row B, each over its own
// row's entries in dependency order (oldest first), independently up to one barrier
// before the diagonal products.  The entries' columns, slots and masks are read 64 at a
// time, one per lane, and taken from their lane (readlane), so no load chain runs
// between two waits; the next nonzero tile is loaded while the current one's column is
// polled; two consecutive entries stored by the same pair task (the usual last two) are
// polled in one spin.  So once the newest dependency lands only its tile products and
// the diagonal products remain on the chain.  A tile whose masks are all zero is
// skipped with no wait, and every thread skips the 16x16 blocks the factor's masks mark
// zero.  CW right-hand sides per pass over the pair.  An odd last row is a pair of its
// own (B absent).
constexpr int NTP = 2 * NT;  // threads of a row-pair task

// one thread's 16 elements of a tile row.  Callers branch around it rather than
// select per element: a select makes the compiler wait for each load in turn.
template <typename T>
__device__ inline void seg16(const T* p, T* v) {
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = p[k];
}

// Sum over the 4 lanes of a quad (lanes 4i .. 4i+3) in every lane, by two DPP quad
// permutes (lane ^ 1, then lane ^ 2); every lane of the wave must be active.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double quad_sum(double v) {
    v += dpp_f64<0xB1>(v);
    return v + dpp_f64<0x4E>(v);
}

template <typename T, int CW>
__global__ __launch_bounds__(NTP) void k_trsv_persistent(SolveArgs a) {
    __shared__ double vec[2][TS * CW];
    __shared__ double rowls[NTP / 64][2 * TS];  // per wave: one column of one or two polled rows
    __shared__ int s_row;
    const int tid = threadIdx.x, g = tid / NT, tg = tid % NT;
    const int r = tg >> 2, kp = tg & 3, lane = tid & 63;
    const int bsh = 4 * (r >> 4) + kp;  // this thread's 16x16 block in a tile's mask
    double* ls = rowls[tid >> 6];
    const T* pool = reinterpret_cast<const T*>(a.pool);
    const int ntask = (a.nbr + 1) / 2;
    const int rowoff = r * TS + kp * 16;
    for (;;) {
        if (tid == 0) {
            const int abort = __hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_row = abort ? ntask : atomicAdd(a.ctrl, 1);
        }
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_row);  // wave-uniform loop control
        if (q >= ntask) break;
        const int p = a.forward ? q : ntask - 1 - q;
        const int I0 = 2 * p, I1 = 2 * p + 1 < a.nbr ? 2 * p + 1 : -1;
        const int IA = a.forward || I1 < 0 ? I0 : I1;
        const int IB = I1 < 0 ? -1 : (a.forward ? I1 : I0);
        const int I = g == 0 ? IA : IB;  // this group's row, -1: none
        const int cpl = g == 1 && IB >= 0 ? a.pairinv[p] : -1;
        // the diagonal inverse (and for row B the coupling block), loaded while the rows
        // wait; a group without a row or coupling block loads A's inverse and ignores it
        T dg[16], cb[16];
        const bool hcb = cpl >= 0;
        seg16(pool + (size_t)a.dinv[I >= 0 ? I : IA] * TE + rowoff, dg);
        seg16(pool + (size_t)(hcb ? cpl : a.dinv[IA]) * TE + rowoff, cb);
        const int e0 = I >= 0 ? a.rowptr[I] : 0, e1 = I >= 0 ? a.rowptr[I + 1] : 0;
        // entries in dependency order: forward e0, e0+1, ..; backward e1-1, e1-2, ..; row
        // B's last one (column IA: forward the largest column, backward the smallest) is
        // the coupling block's
        int n = e1 - e0;
        if (g == 1 && n > 0 && a.col[a.forward ? e1 - 1 : e0] == IA) n--;
        const int64_t rowI = (int64_t)(I >= 0 ? I : 0) * TS;
        for (int c0 = 0; c0 < a.nrhs; c0 += CW) {
            const int cw = min(CW, a.nrhs - c0);
            double acc[CW];
#pragma unroll
            for (int c = 0; c < CW; c++)
                acc[c] = (I >= 0 && kp == 0 && c < cw) ? a.in[rowI + r + (int64_t)(c0 + c) * a.ld] : 0.0;
            for (int base = 0; base < n; base += 64) {  // group-uniform
                const int m = min(64, n - base);
                int Jv = 0, sv = 0;
                uint32_t mv = 0;
                if (lane < m) {
                    const int e = a.forward ? e0 + base + lane : e1 - 1 - base - lane;
                    Jv = a.col[e];
                    sv = a.slot[e];
                    mv = a.mask[sv];
                }
                uint64_t rem = __ballot(mv != 0u);  // entries with a nonzero tile, in order
                if (!rem) continue;
                // one step: entry J (tile t0), or J and J1 stored by one pair task (t1),
                // their columns polled in one spin, then the tiles' products
                auto consume = [&](int J, int J1, bool both, bool b0, bool b1, const T* u0, const T* u1) {
                    const double* xj = a.out + (int64_t)J * TS;
#pragma unroll
                    for (int c = 0; c < CW; c++) {
                        if (c < cw) {
                            if (both)  // the whole wave
                                load_row_stored<2>(xj + (int64_t)(c0 + c) * a.ld, ls, lane, a, (int64_t)(J1 - J) * TS);
                            else
                                load_row_stored<1>(xj + (int64_t)(c0 + c) * a.ld, ls, lane, a);
                            // the column into registers first (one batch of LDS reads),
                            // then one branch per tile
                            double xc[16], s = 0;
#pragma unroll
                            for (int k = 0; k < 16; k++) xc[k] = ls[kp * 16 + k];
                            if (b0) {
#pragma unroll
                                for (int k = 0; k < 16; k++) s += (double)u0[k] * xc[k];
                            }
                            if (both) {
#pragma unroll
                                for (int k = 0; k < 16; k++) xc[k] = ls[TS + kp * 16 + k];
                                if (b1) {
#pragma unroll
                                    for (int k = 0; k < 16; k++) s += (double)u1[k] * xc[k];
                                }
                            }
                            __builtin_amdgcn_wave_barrier();  // ls is rewritten for the next column
                            acc[c] -= s;
                        }
                    }
                };
                auto on = [&](int j) { return (__builtin_amdgcn_readlane(mv, j) >> bsh) & 1u; };
                auto tile = [&](int j) { return pool + (size_t)__builtin_amdgcn_readlane(sv, j) * TE + rowoff; };
                if constexpr (sizeof(T) == 4) {
                    // FP32 tiles: whole steps, the next step's tiles loaded while the
                    // current step's columns are polled (a step that takes two entries
                    // leaves no tile of its own to load during its wait)
                    auto next_step = [&](uint64_t& r, int& ja, int& jb) {
                        if (!r) return false;
                        ja = __builtin_ctzll(r);
                        r &= r - 1;
                        jb = -1;
                        if (r) {
                            const int c = __builtin_ctzll(r);
                            if ((__builtin_amdgcn_readlane(Jv, c) >> 1) == (__builtin_amdgcn_readlane(Jv, ja) >> 1)) {
                                jb = c;
                                r &= r - 1;
                            }
                        }
                        return true;
                    };
                    T t0[16], t1[16], n0[16], n1[16];  // loaded only where the block is nonzero
                    int ja, jb;
                    next_step(rem, ja, jb);
                    if (on(ja)) seg16(tile(ja), t0);
                    if (jb >= 0 && on(jb)) seg16(tile(jb), t1);
                    for (;;) {
                        int na = 0, nb = -1;
                        const bool more = next_step(rem, na, nb);
                        if (more && on(na)) seg16(tile(na), n0);
                        if (more && nb >= 0 && on(nb)) seg16(tile(nb), n1);
                        const bool both = jb >= 0;
                        const int J = __builtin_amdgcn_readlane(Jv, ja);
                        const int J1 = __builtin_amdgcn_readlane(Jv, both ? jb : ja);
                        consume(J, J1, both, on(ja), both && on(jb), t0, t1);
                        if (!more) break;
                        ja = na;
                        jb = nb;
#pragma unroll
                        for (int k = 0; k < 16; k++) {
                            t0[k] = n0[k];
                            t1[k] = n1[k];
                        }
                    }
                } else {
                    // FP64 tiles (no registers for a second step): the next entry's tile
                    // is loaded while the current one's column is polled
                    int j = __builtin_ctzll(rem);
                    rem &= rem - 1;
                    T t0[16], t1[16];
                    if (on(j)) seg16(tile(j), t0);
                    for (;;) {
                        const int J = __builtin_amdgcn_readlane(Jv, j);
                        const bool has1 = rem != 0;
                        const int j1 = has1 ? __builtin_ctzll(rem) : 0;
                        const int J1 = __builtin_amdgcn_readlane(Jv, j1);
                        const bool on1 = has1 && on(j1);
                        if (on1) seg16(tile(j1), t1);  // the next tile, loaded while J's column is polled
                        const bool both = has1 && (J1 >> 1) == (J >> 1);  // stored by one pair task
                        consume(J, J1, both, on(j), both && on1, t0, t1);
                        if (!has1) break;
                        rem &= rem - 1;
                        if (both) {  // both consumed: the one after is loaded now
                            if (!rem) break;
                            j = __builtin_ctzll(rem);
                            rem &= rem - 1;
                            if (on(j)) seg16(tile(j), t0);
                        } else {
                            j = j1;
#pragma unroll
                            for (int k = 0; k < 16; k++) t0[k] = t1[k];
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < CW; c++) acc[c] = quad_sum(acc[c]);
            if (kp == 0) {
#pragma unroll
                for (int c = 0; c < CW; c++) vec[g][r * CW + c] = acc[c];
            }
            __syncthreads();
            double o[CW];
#pragma unroll
            for (int c = 0; c < CW; c++) {
                double v[16], s = 0;
#pragma unroll
                for (int k = 0; k < 16; k++) v[k] = vec[g][(kp * 16 + k) * CW + c];
#pragma unroll
                for (int k = 0; k < 16; k++) s += (double)dg[k] * v[k];
                if (g == 1 && hcb) {
#pragma unroll
                    for (int k = 0; k < 16; k++) v[k] = vec[0][(kp * 16 + k) * CW + c];
#pragma unroll
                    for (int k = 0; k < 16; k++) s += (double)cb[k] * v[k];
                }
                o[c] = s;
            }
#pragma unroll
            for (int c = 0; c < CW; c++) o[c] = quad_sum(o[c]);
            if (kp == 0 && I >= 0) {
                for (int c = 0; c < cw; c++)
                    __hip_atomic_store(reinterpret_cast<unsigned long long*>(a.out) + rowI + r + (int64_t)(c0 + c) * a.ld,
                                       (unsigned long long)__double_as_longlong(o[c]), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();  // vec is reused by the next chunk
        }
    }
}

// Many right-hand sides: the same block-row schedule with the sums on MFMA.  Row I's
// right-hand sides go in panels of 64 columns; for each panel the workgroup keeps
// R = B_I - sum_J L_IJ X_J as a 64x64 accumulator (wave w: quadrant (w/2, w%2) as 2x2
// v_mfma_f64_16x16x4 blocks), each product staged through LDS like the factor's
// GEMM (tiles of any storage precision widened to FP64), then X_I = Linv_II R.
// Columns past nrhs read as zero and are not stored.
template <typename T>
__device__ inline void tile_to_lds_f64(const T* __restrict__ g, double* __restrict__ s, int ld, int tid,
                                       uint32_t m = MASK_ALL) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    const bool on = (m >> (4 * (r >> 4) + (tid & 3))) & 1u;  // zero blocks are written, not read
#pragma unroll
    for (int q = 0; q < 16; q++) s[r * ld + c0 + q] = on ? (double)g[r * TS + c0 + q] : 0.0;
}

// rows row0 .. row0+63 of the column-major vector block v (column stride ld),
// columns col0 .. col0+63 (zero past ncol), into LDS row-major
__device__ inline void panel_to_lds(const double* __restrict__ v, int64_t ld, int64_t row0, int col0, int ncol,
                                    double* __restrict__ s, int lds_ld, int tid) {
    const int r = tid & 63, cq = tid >> 6;  // lanes walk rows: coalesced within a column
#pragma unroll 4
    for (int k = 0; k < 16; k++) {
        const int cc = cq * 16 + k;
        const int col = col0 + cc;
        s[r * lds_ld + cc] = col < ncol ? v[row0 + r + (int64_t)col * ld] : 0.0;
    }
}

template <typename T, int NP>
__global__ __launch_bounds__(NT, 2) void k_trsm_persistent(SolveArgs a) {
    __shared__ __attribute__((aligned(16))) double la[TS * LDA];
    __shared__ __attribute__((aligned(16))) double lb[TS * LDB];
    __shared__ int s_row, s_abort;
    typedef MfmaOp<double> M;
    typedef M::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, li = lane & 15, lk = lane >> 4;
    const T* pool = reinterpret_cast<const T*>(a.pool);
    auto mma_lds = [&](acc_t (&acc)[2][2], double sg) {
#pragma unroll 4
        for (int kk = 0; kk < TS / 4; kk++) {
            const int k = 4 * kk + lk;
            double af[2], bf[2];
#pragma unroll
            for (int m = 0; m < 2; m++) af[m] = sg * la[(wr + m * 16 + li) * LDA + k];
#pragma unroll
            for (int n = 0; n < 2; n++) bf[n] = lb[k * LDB + wc + n * 16 + li];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++) acc[m][n] = M::mma(af[m], bf[n], acc[m][n]);
        }
    };
    for (;;) {
        if (tid == 0) {
            s_abort = __hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_row = s_abort ? a.nbr : atomicAdd(a.ctrl, 1);
        }
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_row);  // wave-uniform loop control
        if (q >= a.nbr) break;
        const int I = a.forward ? q : a.nbr - 1 - q;
        const int e0 = a.rowptr[I], e1 = a.rowptr[I + 1], cnt = e1 - e0;
        const int64_t rowI = (int64_t)I * TS;
        // R = B_I for every panel of 64 columns (NP panels cover nrhs)
        acc_t acc[NP][2][2];
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = wr + m * 16 + M::row(lk, r), col = p * TS + wc + n * 16 + li;
                        acc[p][m][n][r] = col < a.nrhs ? a.in[rowI + row + (int64_t)col * a.ld] : 0.0;
                    }
        // dependencies oldest first; each L tile is staged once for all panels
        for (int i = 0; i < cnt; i++) {
            const int e = a.forward ? e0 + i : e1 - 1 - i;
            const int J = a.col[e];
            const uint32_t tm = a.mask[a.slot[e]];
            if (tm == 0) continue;  // a zero tile: no product, no wait for row J
            if (tid == 0) {
                spin_row(a, J);
                acquire_agent();
            }
            __syncthreads();  // also: the previous products are done with la / lb
            tile_to_lds_f64(pool + (size_t)a.slot[e] * TE, la, LDA, tid, tm);
#pragma unroll
            for (int p = 0; p < NP; p++) {
                if (p) __syncthreads();  // the previous panel's product is done with lb
                panel_to_lds(a.out, a.ld, (int64_t)J * TS, p * TS, a.nrhs, lb, LDB, tid);
                __syncthreads();
                mma_lds(acc[p], -1.0);
            }
        }
        // X_I = Dinv_II R, panel by panel: R goes through LDS as the right operand
        __syncthreads();
        tile_to_lds_f64(pool + (size_t)a.dinv[I] * TE, la, LDA, tid);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if (p) __syncthreads();
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        lb[(wr + m * 16 + M::row(lk, r)) * LDB + wc + n * 16 + li] = acc[p][m][n][r];
            __syncthreads();
            acc_t o[2][2];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++) o[m][n] = acc_t{0, 0, 0, 0};
            mma_lds(o, 1.0);
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = wr + m * 16 + M::row(lk, r), col = p * TS + wc + n * 16 + li;
                        if (col < a.nrhs) a.out[rowI + row + (int64_t)col * a.ld] = o[m][n][r];
                    }
        }
        __syncthreads();  // every wave's stores are complete (workgroup release)
        if (tid == 0) {   // one device-scope release for the workgroup, then the flag
            release_agent();
            __hip_atomic_store(a.ready + I, a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();  // separates the publish branch from the next fetch branch
    }
}

__global__ void k_solve_ctrl_reset(int* ctrl) {
    if (threadIdx.x < 4) ctrl[threadIdx.x] = 0;
}

__global__ void k_zero_i32(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// out = diag(L) y over the padded rows, every column: the forward result as the
// reference's lowerSolver leaves it in its right-hand side, b_i = l_ii y_i (for a
// unit-diagonal L that is y itself).
template <typename T>
__global__ void k_scale_by_ldiag(const T* __restrict__ pool, const int32_t* __restrict__ l_diag,
                                 const double* __restrict__ y, double* __restrict__ out, int64_t ld, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % ld;
        const int64_t i = r / TS, q = r % TS;
        out[e] = (double)pool[(size_t)l_diag[i] * TE + q * TS + q] * y[e];
    }
}

namespace {

#define HIP_TRY(ctx, expr)                                       \
    do {                                                         \
        int _rc = check_hip((ctx), (expr), #expr);               \
        if (_rc != SOGL_OK) return _rc;                          \
    } while (0)

// Buffers of one solve beyond its input and output: the forward result y (padded n x
// nrhs), the per-block-row ready flags of the MFMA form (they hold the epoch of the pass
// that stored the row, so a fresh set must start at zero) and the control words.
struct SolveWork {
    double* y;
    int* ready;
    int* ctrl;
    int32_t* epoch;  // last pass's epoch of these flags; each pass takes the next one
};

template <typename T>
int launch_pass(Context* c, Factor* f, const SolveWork& w, bool forward, const double* in, double* out, int nrhs) {
    SolveArgs a{};
    a.pool = f->pool.p;
    a.rowptr = (const int32_t*)(forward ? f->l_rowptr.p : f->u_rowptr.p);
    a.col = (const int32_t*)(forward ? f->l_col.p : f->u_col.p);
    a.slot = (const int32_t*)(forward ? f->l_slot.p : f->u_slot.p);
    a.dinv = (const int32_t*)(forward ? f->linv_d.p : f->uinv_d.p);
    a.in = in;
    a.out = out;
    a.ld = (int64_t)f->block_rows * TS;
    a.nrhs = nrhs;
    a.nbr = f->block_rows;
    a.forward = forward ? 1 : 0;
    a.epoch = ++*w.epoch;
    a.ready = w.ready;
    a.ctrl = w.ctrl;
    a.mask = (const uint32_t*)f->mask_d.p;
    a.pairinv = (const int32_t*)(forward ? f->pairx_d.p : f->pairy_d.p);
    hipLaunchKernelGGL(k_solve_ctrl_reset, dim3(1), dim3(64), 0, c->stream, a.ctrl);
    int per_cu = 0;
    const bool mfma = nrhs >= TRSM_MIN_RHS;
    if (!mfma) {  // the rows of out are pending until their task stores them
        const int64_t n = a.ld * nrhs;
        const unsigned fg = (unsigned)std::min<int64_t>((n + 255) / 256, 2048);
        hipLaunchKernelGGL(k_fill_pending, dim3(fg), dim3(256), 0, c->stream, out, a.ld, a.ld, nrhs);
    }
    const int np = (nrhs + TS - 1) / TS;  // panels of 64 columns (at most TRSM_MAX_RHS / 64)
    const void* kern = !mfma ? (nrhs == 1 ? (const void*)k_trsv_persistent<T, 1> : (const void*)k_trsv_persistent<T, 8>)
                       : np == 1 ? (const void*)k_trsm_persistent<T, 1>
                       : np == 2 ? (const void*)k_trsm_persistent<T, 2>
                       : np == 3 ? (const void*)k_trsm_persistent<T, 3> : (const void*)k_trsm_persistent<T, 4>;
    HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, mfma ? NT : NTP, 0));
    if (per_cu < 1) per_cu = 1;
    const int grid = std::max(1, std::min(mfma ? a.nbr : (a.nbr + 1) / 2, c->num_cu * per_cu));
    if (!mfma) {
        if (nrhs == 1) hipLaunchKernelGGL((k_trsv_persistent<T, 1>), dim3(grid), dim3(NTP), 0, c->stream, a);
        else hipLaunchKernelGGL((k_trsv_persistent<T, 8>), dim3(grid), dim3(NTP), 0, c->stream, a);
    } else if (np == 1) {
        hipLaunchKernelGGL((k_trsm_persistent<T, 1>), dim3(grid), dim3(NT), 0, c->stream, a);
    } else if (np == 2) {
        hipLaunchKernelGGL((k_trsm_persistent<T, 2>), dim3(grid), dim3(NT), 0, c->stream, a);
    } else if (np == 3) {
        hipLaunchKernelGGL((k_trsm_persistent<T, 3>), dim3(grid), dim3(NT), 0, c->stream, a);
    } else {
        hipLaunchKernelGGL((k_trsm_persistent<T, 4>), dim3(grid), dim3(NT), 0, c->stream, a);
    }
    HIP_TRY(c, hipGetLastError());
    return SOGL_OK;
}

}  // namespace

// Both passes over nrhs columns on the context stream, timed; d_x may not alias d_b.
int solve_core(Context* c, Factor* f, const SolveWork& w, const double* d_b, double* d_x, int nrhs) {
    const int R = f->block_rows;
    int rc;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    const int64_t ld = (int64_t)R * TS;
    for (int c0 = 0; c0 < nrhs; c0 += TRSM_MAX_RHS) {  // column groups of one MFMA solve launch
        const int m = std::min(TRSM_MAX_RHS, nrhs - c0);
        const int64_t off = (int64_t)c0 * ld;
        if (f->elem == 8) {
            if ((rc = launch_pass<double>(c, f, w, true, d_b + off, w.y + off, m))) return rc;
            if ((rc = launch_pass<double>(c, f, w, false, w.y + off, d_x + off, m))) return rc;
        } else {
            if ((rc = launch_pass<float>(c, f, w, true, d_b + off, w.y + off, m))) return rc;
            if ((rc = launch_pass<float>(c, f, w, false, w.y + off, d_x + off, m))) return rc;
        }
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    f->solve_ms = ms;
    int ctrl[4];
    HIP_TRY(c, hipMemcpy(ctrl, w.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    if (ctrl[1]) return set_error(c, SOGL_ERR_TIMEOUT, "solve: a dependency wait exceeded its bound");
    return SOGL_OK;
}

// The factor's own work buffers (allocated on first use; the flags are zeroed once and
// then ordered by the factor's epoch counter).
int solve_device(Context* c, Factor* f, const double* d_b, double* d_x, double* d_y, int nrhs) {
    const int R = f->block_rows;
    int rc;
    if ((rc = dev_alloc(c, f->solve_ready, (size_t)R * 4, "solve flags"))) return rc;
    if ((rc = dev_alloc(c, f->solve_ctrl, 16, "solve control"))) return rc;
    if (f->solve_epoch == 0) {
        hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, c->stream, (int*)f->solve_ready.p, R);
        HIP_TRY(c, hipGetLastError());
    }
    const SolveWork w{d_y, (int*)f->solve_ready.p, (int*)f->solve_ctrl.p, &f->solve_epoch};
    return solve_core(c, f, w, d_b, d_x, nrhs);
}

int ensure_vectors(Context* c, Factor* f, int nrhs) {
    const size_t need = (size_t)f->block_rows * TS * (size_t)nrhs;
    if (f->vec_cap >= need) return SOGL_OK;
    int rc;
    if ((rc = dev_alloc(c, f->vec_b, need * 8, "rhs")) || (rc = dev_alloc(c, f->vec_x, need * 8, "x")) ||
        (rc = dev_alloc(c, f->vec_y, need * 8, "y")))
        return rc;
    f->vec_cap = need;
    return SOGL_OK;
}

}  // namespace sogl_dev

using namespace sogl_dev;

extern "C" int sogl_solve(sogl_ctx* x, int32_t factor_id, const double* bp, double* xp, double* y,
                          int32_t nrhs) {
    if (!x || !bp || !xp || nrhs <= 0) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "bad argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_vectors(c, f, nrhs))) return rc;
    const size_t bytes = (size_t)f->block_rows * TS * nrhs * 8;
    HIP_TRY(c, hipMemcpyAsync(f->vec_b.p, bp, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = solve_device(c, f, (const double*)f->vec_b.p, (double*)f->vec_x.p, (double*)f->vec_y.p, nrhs)))
        return rc;
    HIP_TRY(c, hipMemcpyAsync(xp, f->vec_x.p, bytes, hipMemcpyDeviceToHost, c->stream));
    if (y && f->llt) {
        // L L^T: the reference's right-hand side ends as diag(L) y (the input b in
        // vec_b is no longer needed)
        const int64_t ld = (int64_t)f->block_rows * TS, n = ld * nrhs;
        const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
        if (f->elem == 8)
            hipLaunchKernelGGL(k_scale_by_ldiag<double>, dim3(grid), dim3(256), 0, c->stream, (const double*)f->pool.p,
                               (const int32_t*)f->l_diag_d.p, (const double*)f->vec_y.p, (double*)f->vec_b.p, ld, n);
        else
            hipLaunchKernelGGL(k_scale_by_ldiag<float>, dim3(grid), dim3(256), 0, c->stream, (const float*)f->pool.p,
                               (const int32_t*)f->l_diag_d.p, (const double*)f->vec_y.p, (double*)f->vec_b.p, ld, n);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(y, f->vec_b.p, bytes, hipMemcpyDeviceToHost, c->stream));
    } else if (y) {
        HIP_TRY(c, hipMemcpyAsync(y, f->vec_y.p, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return SOGL_OK;
}

extern "C" int sogl_solve_device(sogl_ctx* x, int32_t factor_id, const double* d_bp, double* d_xp,
                                 int32_t nrhs) {
    if (!x || !d_bp || !d_xp || nrhs <= 0) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "bad argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc;
    if ((rc = ensure_vectors(c, f, nrhs))) return rc;
    return solve_device(c, f, d_bp, d_xp, (double*)f->vec_y.p, nrhs);
}

// ------------------------------------------------------------------ solve without allocation
// The solve's work buffers (y, flags, control) in one fixed array per GPU, for callers
// that cannot allocate device memory.  In host code `scratch` names a host-side stand-in
// for the array, never its device address: the address each GPU gives its own copy is
// looked up once per device, when the first context on it is created, and only that
// address is handed to kernels and copies.
namespace sogl_dev {
namespace {

static __device__ __align__(256) unsigned char scratch[SOGL_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES];       // device address of `scratch` on each GPU
hipError_t g_scratch_rc[SCRATCH_MAX_DEVICES];   // result of the lookup (valid when looked up)
bool g_scratch_done[SCRATCH_MAX_DEVICES];

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Layout inside the array: control words, the flags of `rows` block rows, then y.
struct ScratchLayout {
    size_t ctrl, ready, y, end;
};
ScratchLayout scratch_layout(int32_t rows, int32_t nrhs) {
    ScratchLayout l;
    l.ctrl = 0;
    l.ready = 256;
    l.y = l.ready + align256((size_t)rows * 4);
    l.end = l.y + (size_t)rows * TS * (size_t)nrhs * 8;
    return l;
}

}  // namespace

// Looks up this GPU's address of `scratch`, once per device (the device is current).
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) return;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_rc[device] != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
    }
    g_scratch_ptr[device] = p;
    g_scratch_done[device] = true;
}

}  // namespace sogl_dev

extern "C" int sogl_solve_device_noalloc(sogl_ctx* x, int32_t factor_id, const double* d_bp, double* d_xp,
                                         int32_t nrhs) {
    if (!x || !d_bp || !d_xp || nrhs <= 0) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "bad argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    const int dev = c->device;
    void* base = nullptr;
    hipError_t lookup = hipErrorNotInitialized;
    {
        std::lock_guard<std::mutex> sg(g_scratch_mu);
        if (dev >= 0 && dev < SCRATCH_MAX_DEVICES && g_scratch_done[dev]) {
            base = g_scratch_ptr[dev];
            lookup = g_scratch_rc[dev];
        }
    }
    if (!base)
        return set_error(c, SOGL_ERR_DEVICE, std::string("solve without allocation: no scratch address for this "
                                                         "device: ") + hipGetErrorString(lookup));
    const ScratchLayout l = scratch_layout(f->block_rows, nrhs);
    if (l.end > (size_t)SOGL_NOALLOC_SCRATCH_BYTES)
        return set_error(c, SOGL_ERR_ARG, "solve without allocation: " + std::to_string(f->block_rows) +
                                              " block rows x " + std::to_string(nrhs) + " right-hand sides need " +
                                              std::to_string(l.end) + " bytes of scratch, more than its " +
                                              std::to_string((size_t)SOGL_NOALLOC_SCRATCH_BYTES));
    HIP_TRY(c, hipSetDevice(dev));
    char* s = (char*)base;
    // the flags start at zero for every call (earlier calls, of any factor, left their
    // epochs in them); a fill kernel, so the call stays capturable
    hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, c->stream, (int*)(s + l.ready), f->block_rows);
    HIP_TRY(c, hipGetLastError());
    int32_t epoch = 0;
    const SolveWork w{(double*)(s + l.y), (int*)(s + l.ready), (int*)(s + l.ctrl), &epoch};
    return solve_core(c, f, w, d_bp, d_xp, nrhs);
}
