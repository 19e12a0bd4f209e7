// Device factor: HBM tile pool, task graph and the three executors that
// replace BlockPlanner::calculate() (BlockPlanner.cpp:376-651).
//
// The reference walks stages and, inside a stage, runs 47 OpenMP "streams"
// (ops bucketed by result % 47) with up to 6 same-result muls fused.  Here the
// unit of work is a *task*: all ops that write one result tile (an lu writes
// L and U).  Writers of a tile always share one stage in the reference's plan,
// and every tile is written once (ids are never reused), so the tasks form a
// DAG over tiles: a task depends only on the tasks producing its sources.
//
//   SOGL_EXEC_PERSISTENT  one kernel; workgroups take task indices from a device
//                         counter in topological order and spin (bounded) on the
//                         ready flags of their source tiles.  Deadlock-free: a
//                         taken task only waits for tasks taken earlier, which
//                         are held by running workgroups.
//   SOGL_EXEC_LEVELS      one launch per dependency level (kernel boundaries
//                         order the levels).
//   SOGL_EXEC_PER_OP      one launch per unit op, in plan order (the unfused
//                         comparison point; C += A*B reads C).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "device.h"
#include "tile_ops.h"

namespace sogl {
void set_last_error(const std::string& s);
bool plan_factor_desc(const sogl_plan* p, int32_t precision, sogl_factor_desc* d);
bool plan_input_cells(const sogl_plan* p, const int64_t** begin, const uint16_t** pos, const double** val);
}

namespace sogl_dev {

// ------------------------------------------------------------------ helpers
int set_error(Context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    sogl::set_last_error(msg);
    return code;
}

int check_hip(Context* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return SOGL_OK;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return set_error(c, SOGL_ERR_OOM, m);
    return set_error(c, SOGL_ERR_DEVICE, m);
}

int dev_alloc(Context* c, DevBuf& b, size_t bytes, const char* what) {
    if (b.p && b.bytes >= bytes) return SOGL_OK;
    dev_free(b);
    if (bytes == 0) return SOGL_OK;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return check_hip(c, e, what);
    }
    b.bytes = bytes;
    return SOGL_OK;
}

void dev_free(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

Factor* find_factor(Context* c, int32_t id) {
    auto it = c->factors.find(id);
    return it == c->factors.end() ? nullptr : it->second;
}

#define HIP_TRY(ctx, expr)                                       \
    do {                                                         \
        int _rc = check_hip((ctx), (expr), #expr);               \
        if (_rc != SOGL_OK) return _rc;                          \
    } while (0)

// ------------------------------------------------------------------ control block
// ctrl[0] task counter, ctrl[1] abort flag, ctrl[2] low-precision tiles refactored in
// FP64, ctrl[3] max growth (float bits, non-negative so int order = float order),
// ctrl[4] tasks run by a workgroup that claimed them (dynamic chaining), ctrl[5] tasks a
// workgroup took from the queue and dropped because another had claimed them, ctrl[6]
// tasks run (each task exactly once per run)
constexpr int CTRL_WORDS = 8;

__global__ void k_ctrl_reset(int* ctrl) {
    if (threadIdx.x < CTRL_WORDS) ctrl[threadIdx.x] = 0;
}

// The reference's check of every upper inverse it makes (inv_check_diag,
// MatrixStdDouble.cpp:2871-2937, after each upperInv, BlockPlanner.cpp:573-577): the
// diagonal of U U^-1 must lie within 1e-3 of 1 (its min and max start at 1), and the
// entries (i, i/2), i >= 2, within 1e-3 of 0 (min and max start at 0); a NaN sum sets
// the max to 10; the middle entries are looked at only when the diagonal passes.  Each
// sum runs over k = 0..63 in FP64, in the reference's order.  One workgroup per block
// row's U_ii and U_ii^-1, after the executor; out[0] counts the tiles out of tolerance
// and out[1 + 4 j] .. holds (block row, 0 diagonal / 1 middle, min, max) of the first
// UINV_KEEP of them (in no particular order).
constexpr int UINV_KEEP = 16;
template <typename T>
__global__ __launch_bounds__(NT) void k_check_uinv(const T* __restrict__ pool, const int32_t* __restrict__ u_diag,
                                                   const int32_t* __restrict__ uinv, double* out) {
    __shared__ double s_v[2][TS];
    __shared__ int s_nan[2][TS];
    const int blk = blockIdx.x, tid = threadIdx.x;
    const T* U = pool + (size_t)u_diag[blk] * TE;
    const T* Z = pool + (size_t)uinv[blk] * TE;
    if (tid < 2 * TS) {
        const int which = tid / TS, i = tid % TS, j = which == 0 ? i : i / 2;
        double sum = 0.0;
        if (which == 0 || i >= 2)
            for (int k = 0; k < TS; k++) sum += (double)U[i * TS + k] * (double)Z[k * TS + j];
        s_v[which][i] = sum;
        s_nan[which][i] = sum != sum;
    }
    __syncthreads();
    if (tid != 0) return;
    const double tor = 1e-3;
    for (int which = 0; which < 2; which++) {
        const double c = which == 0 ? 1.0 : 0.0;
        double mn = c, mx = c;
        for (int i = which == 0 ? 0 : 2; i < TS; i++) {
            if (s_nan[which][i]) {
                mx = 10;
                continue;
            }
            mn = fmin(mn, s_v[which][i]);
            mx = fmax(mx, s_v[which][i]);
        }
        if (mx > c + tor || mx < c - tor || mn > c + tor || mn < c - tor) {
            const int q = atomicAdd(reinterpret_cast<int*>(out), 1);
            if (q < UINV_KEEP) {
                out[1 + 4 * q] = blk;
                out[2 + 4 * q] = which;
                out[3 + 4 * q] = mn;
                out[4 + 4 * q] = mx;
            }
            return;  // as the reference: a failed diagonal ends the check
        }
    }
}

template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

constexpr long long SPIN_LIMIT_TICKS = 20LL * 100 * 1000 * 1000;  // ~20 s of the 100 MHz wall clock

struct Args {
    void* pool;
    const DevTask* tasks;
    const Pair* pairs;
    const int32_t* deps;
    unsigned long long* ready;  // per slot: epoch << 32 | block mask of the tile made in that run
    int* ctrl;
    int32_t ntasks;
    int32_t epoch;
    int32_t task_offset;  // levels executor: first task of the launch
    int32_t chunk;        // persistent executor: tasks taken per fetch from the queue
    const int32_t* queue; // persistent executor: the task indices handed out, in order
    int32_t nqueue;
    double growth_limit;
    long long* trace;     // optional: per task {taken, work done, published, last wait ended}
    uint32_t* mask;       // per slot: 16x16 block mask of the tile (see tile_ops.h)
    const ChainChk* chk;  // persistent executor: chain candidates' check entries
    int32_t chain;        // claims per finished task at most (0: no dynamic chaining)
    const int32_t* war;   // value ids whose readers a reused result slot waits for (DevTask)
};

// Waits (thread 0 polls, relaxed) until the epoch in ready[s] reaches this run's and
// returns the flag (its low word is the tile's block mask); on abort or timeout it
// raises ctrl[1] and returns 0.  Relaxed polling keeps the spin loop free of cache
// invalidations, and the tiles are read with coherent loads afterwards (no acquire).
// The abort word and the wall clock are looked at every SPIN_CHECK polls only: each poll
// is a round trip to the device coherence point (~0.3 us measured between two
// workgroups), and a second one per poll would double the delay before a raised flag is
// seen.
constexpr unsigned SPIN_CHECK = 64;
__device__ inline unsigned long long spin_ready(const Args& a, int s) {
    unsigned long long* flag = a.ready + s;
    const unsigned long long want = (unsigned long long)(unsigned)a.epoch << 32;
    unsigned long long v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v >= want) return v;
    const long long t0 = wall_clock64();
    for (unsigned spins = 1;; spins++) {
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v >= want) return v;
        if (spins % SPIN_CHECK) continue;
        if (__hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return 0;
        if (wall_clock64() - t0 > SPIN_LIMIT_TICKS) {
            __hip_atomic_store(a.ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 0;
        }
    }
}

// No fence on either side of a hand-off.  Release: tile stores are write-through (gst),
// each wave waits for its own before the publish barrier, and the flag stores follow
// that barrier.  Acquire: within one run every pool slot is written once, and no CU
// reads a slot before its flag is up (the loads are issued after the flag was seen),
// so no cache can hold a copy of a tile made in this run that predates it; the only
// stale lines a plain load could hit are left from earlier runs, and each workgroup
// invalidates its CU's L1 and its XCD's L2 once, when it starts, before any tile load.
// (A per-task invalidate measured ~0.45 us median, 2.3 us at p90, on every hand-off;
// agent-coherent loads instead cost more than they saved: they are at most 8 bytes
// wide and bypass the L2.)

// The masks come with the flags.  A window of pairs: wave 0 loads the flags of up to
// PAIR_WINDOW pairs at once (lane 2j / 2j+1: pair s+j's A / B), polls until pair s is
// ready, counts the leading pairs whose flags are both raised and leaves their masks in
// an LDS ring of PAIR_RING pairs (pair j at 2 (j mod PAIR_RING)) and the count in one of
// two words by the window's parity.  Zero products
// are skipped without a barrier, so a wave may still be reading window k's masks and
// count while wave 0 writes window k + 1's; it cannot be further behind (window k + 1
// starts after the barrier that ends the wait for window k), and with windows of at
// most half the ring the two never share an entry.  Single waits (the add tile, the
// bf16 loop): thread 0 leaves the masks in one of two slots by the parity of s (a wave
// may still be reading the previous pair's masks when thread 0 writes the next one;
// every other pair has a barrier in between).
constexpr int PAIR_WINDOW = 32;
constexpr int PAIR_RING = 2 * PAIR_WINDOW;
constexpr int PM_LEN = 2 * PAIR_RING;          // two window counts
constexpr int PM_SINGLE = PM_LEN + 2;          // two single-wait slots of two words
constexpr int PM_ADD = PM_SINGLE + 4;          // the added tile: seen ready, its mask
constexpr int PM_WAR = PM_ADD + 2;             // the result slot's readers all seen done
constexpr int PM_WORDS = PM_WAR + 1;
// A task claimed by the workgroup that made one of its operands (dynamic chaining) has
// every operand published already: `pre` holds the operands' masks (pair lanes 0-3, the
// added tile at 4), taken from the flags the claimer read, so nothing is polled.  A
// queued task that is some task's chain candidate (`ctag` != 0) looks for its claim bit
// in its operands' flags and, when another workgroup has claimed it, sets *s_lost and
// gives up the task (extend returns -1; no operand is read after that point and
// nothing has been stored).
struct PairWait {
    const Args* a;
    long long* t_wait;  // tracing: wall clock when the latest wait ended (thread 0)
    uint32_t* s_pm;     // LDS, PM_WORDS words
    const uint32_t* pre = nullptr;
    int32_t ctag = 0;
    int* s_lost = nullptr;
    int32_t war_b = 0, war_n = 0;  // the task's write-after-read flags (Args::war)
    __device__ inline int extend(const Pair* pairs, int s, int npairs, int window, int fadd = -1) const {
        uint32_t* ring = s_pm;
        int* s_len = reinterpret_cast<int*>(s_pm + PM_LEN + (window & 1));
        if (pre) {  // a claimed task (at most CHAIN_MAX_PAIRS pairs: one window)
            if (threadIdx.x < 64) {
                const int lane = threadIdx.x;
                const int cnt = min(npairs - s, PAIR_WINDOW);
                if (lane < 2 * cnt) ring[2 * ((s + (lane >> 1)) % PAIR_RING) + (lane & 1)] = pre[2 * s + lane];
                if (lane == 0) {
                    *s_len = cnt;
                    if (window == 0) s_pm[PM_WAR] = 0;  // looked at before the stores (war_done)
                    if (a->trace) *t_wait = wall_clock64();
                }
            }
            __syncthreads();
            return s + __builtin_amdgcn_readfirstlane(*s_len);
        }
        if (threadIdx.x < 64) {  // wave 0
            const int lane = threadIdx.x;
            const int cnt = min(npairs - s, PAIR_WINDOW);
            const int j = s + (lane >> 1);
            int slot = -1;
            const bool pl = lane < 2 * cnt;  // a pair's lane
            // lane 63 polls the added tile's flag in the first window when no pair has it,
            // and the lanes between the pairs' and 63 the write-after-read flags
            const bool al = fadd >= 0 && window == 0 && lane == 63 && !pl;
            const int wj = lane - 2 * cnt;
            const bool wl = window == 0 && !pl && lane < 63 && wj < war_n;
            if (pl) {
                const Pair p = pairs[j];
                slot = (lane & 1) ? p.fb : p.fa;  // the flags are per value
            } else if (al) {
                slot = fadd;
            } else if (wl) {
                slot = a->war[war_b + wj];
            }
            const unsigned long long want = (unsigned long long)(unsigned)a->epoch << 32;
            auto poll = [&](bool on, unsigned long long& v) {  // leading pairs with both flags raised
                if (on) v = __hip_atomic_load(a->ready + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long late = __ballot(on && pl && v < want);
                return (late ? __builtin_ctzll(late) : 2 * cnt) >> 1;
            };
            unsigned long long v = 0;
            int len = poll(slot >= 0, v);
            if (len == 0) {
                // spin until pair s is ready, then read the window.  A small window (at
                // most SMALL_WINDOW pairs, the usual case on the critical path) is polled
                // whole each time, so the leading ready pairs are known without a second
                // round trip; a large one polls pair s alone (two flags, not the window's)
                // and then reads the window once.
                constexpr int SMALL_WINDOW = 4;
                const bool small = cnt <= SMALL_WINDOW;
                const long long t0 = wall_clock64();
                for (unsigned spins = 1;; spins++) {
                    __builtin_amdgcn_s_sleep(1);
                    const int got = poll(small ? slot >= 0 : lane < 2, v);
                    if (got > 0) {
                        len = small ? got : poll(slot >= 0, v);
                        break;
                    }
                    if (spins % SPIN_CHECK) continue;
                    // as spin_ready: on abort or timeout the window is taken as ready with
                    // zero masks (the run's results are discarded)
                    const int ab = __hip_atomic_load(a->ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__builtin_amdgcn_readfirstlane(ab) || wall_clock64() - t0 > SPIN_LIMIT_TICKS) {  // wave-uniform
                        if (lane == 0) __hip_atomic_store(a->ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        v = 0;
                        len = cnt;
                        break;
                    }
                }
            }
            if (lane < 2 * len) ring[2 * (j % PAIR_RING) + (lane & 1)] = (uint32_t)v & MASK_ALL;
            if (fadd >= 0 && window == 0 && lane == 63) {  // not polled (a full window): not seen
                s_pm[PM_ADD] = al && v >= want;
                s_pm[PM_ADD + 1] = (uint32_t)v & MASK_ALL;
            }
            if (window == 0 && war_n > 0) {  // wave-uniform
                const bool late = __ballot(wl && v < want) != 0;
                if (lane == 0) s_pm[PM_WAR] = !late && war_n <= 63 - 2 * cnt;
            }
            if (ctag) {  // claimed by the workgroup that made one of these operands?
                const int op = 2 * s + lane;  // operand position in the task
                const int tg = op < 4 && lane < 2 * cnt ? (ctag >> (8 * op)) & 0xFF : 0;
                const bool lost = tg != 0 && v >= want && ((v >> (15 + tg)) & 1ull);
                const bool any_lost = __ballot(lost) != 0;  // every lane of the wave votes
                if (lane == 0) *s_lost = any_lost;
            }
            if (lane == 0) {
                *s_len = len;
                if (a->trace) *t_wait = wall_clock64();
            }
        }
        __syncthreads();
        if (ctag && __builtin_amdgcn_readfirstlane(*s_lost)) return -1;
        return s + __builtin_amdgcn_readfirstlane(*s_len);
    }
    __device__ inline void pair_masks(int j, const Pair&, uint32_t& ma, uint32_t& mb) const {
        ma = s_pm[2 * (j % PAIR_RING)];
        mb = s_pm[2 * (j % PAIR_RING) + 1];
    }
    // Before the first result store of a task whose result slot held another value in
    // this run: every task that read that value has published (its loads are done).
    // Usually the first window saw all these flags up already (polled: extend ran for
    // window 0 and left its verdict in PM_WAR); otherwise wave 0 polls them here, 64 at a
    // time, and the barrier holds the other waves' stores back.
    __device__ inline void war_done(bool polled) const {
        if (war_n == 0) return;                 // wave-uniform (task field)
        if (polled && s_pm[PM_WAR]) return;     // uniform: written before window 0's barrier
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const unsigned long long want = (unsigned long long)(unsigned)a->epoch << 32;
            for (int base = 0; base < war_n; base += 64) {
                const int fl = base + lane < war_n ? a->war[war_b + base + lane] : -1;
                const long long t0 = wall_clock64();
                for (unsigned spins = 1;; spins++) {
                    const bool up = fl < 0 || __hip_atomic_load(a->ready + fl, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT) >= want;
                    if (__ballot(!up) == 0) break;
                    __builtin_amdgcn_s_sleep(1);
                    if (spins % SPIN_CHECK) continue;
                    const int ab = __hip_atomic_load(a->ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__builtin_amdgcn_readfirstlane(ab) || wall_clock64() - t0 > SPIN_LIMIT_TICKS) {
                        if (lane == 0) __hip_atomic_store(a->ctrl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
        }
        __syncthreads();
    }
    // after the first window: was the added tile's flag up (written by that window, read
    // before the next task's first window can write it again)
    __device__ inline bool add_ready(int, uint32_t& m) const {
        if (pre) {
            m = pre[4];
            return true;
        }
        m = s_pm[PM_ADD + 1];
        return s_pm[PM_ADD] != 0;
    }
    __device__ inline void operator()(int s, int fa, int fb, int, int, uint32_t& ma, uint32_t& mb) const {
        uint32_t* pm = s_pm + PM_SINGLE + 2 * (s & 1);
        if (threadIdx.x == 0) {
            if (pre) {  // only the added tile is waited for this way in a claimed task
                pm[0] = pm[1] = pre[4];
            } else {
                pm[0] = (uint32_t)spin_ready(*a, fa) & MASK_ALL;
                pm[1] = fb == fa ? pm[0] : (uint32_t)spin_ready(*a, fb) & MASK_ALL;
            }
            if (a->trace) *t_wait = wall_clock64();
        }
        __syncthreads();
        ma = pm[0];
        mb = pm[1];
    }
};

// LLT: the kernel of an L L^T factor (llt, transpose and A B^T tasks, no lu); the LU
// kernels are built without those, which keeps their register use as it was.
// COH: the factor reuses pool slots, so tiles are read with coherent loads -- except the
// source of an lu / llt run right after the GEMM that made it (chained): that sum is its
// slot's first value in the run (no cache holds an earlier one), and this CU's own
// caches have its stores, which coherent loads could pass (the chain does not wait for
// them).
// xcopy / zcopy (FP32 LU kernels of the persistent executor): LDS for the copies of the
// inverses an lu task with absorbed panels keeps.  Returns an lu task's growth (1
// otherwise).
template <typename T, bool BF16, bool ACC, bool LLT, bool MASKS, bool ONE = false, bool COH = false, typename Wait>
__device__ __attribute__((always_inline)) double run_task(const DevTask& t, const Args& a, char* lds, uint32_t* s_mask,
                                                          const Wait& wait, bool chained = false, T* xcopy = nullptr,
                                                          T* zcopy = nullptr, bool handoff = false) {
    T* pool = reinterpret_cast<T*>(a.pool);
    double g = 1.0;
    // handoff (persistent executor): this GEMM's lu / llt runs next on this workgroup and
    // takes the sum from LDS; chained: this lu / llt is that task (the bf16 kernels load it)
    constexpr bool LDS_HANDOFF = !BF16;
    switch (t.kind) {
        case TK_GEMM:
            if constexpr (BF16 && !ACC) {
                gemm_task_bf16<LLT, MASKS, Wait, COH>(reinterpret_cast<float*>(a.pool), a.pairs + t.a, t.b, t.out0,
                                                      lds, s_mask, wait, t.add, t.add_neg != 0, t.fadd);
            } else {
                gemm_task<T, ACC, LLT, MASKS, Wait, ONE, COH>(pool, a.pairs + t.a, t.b, t.out0, lds, s_mask, wait,
                                                              t.add, t.add_neg != 0, t.fadd, LDS_HANDOFF && handoff);
            }
            break;
        case TK_SUB:
            sub_task<T, MASKS, COH>(pool, t.b, t.a, t.out0, s_mask);
            break;
        case TK_LU:
            if constexpr (!LLT) {
                if constexpr (sizeof(T) == 4) {
                    g = lu_task<T>(pool, t.a, t.out0, t.out1, t.out2, t.out3, lds, a.growth_limit, a.ctrl + 2,
                                   COH && !chained, xcopy, zcopy, LDS_HANDOFF && chained);
                    if (threadIdx.x == 0) {
                        float gf = (float)g;
                        atomicMax(a.ctrl + 3, __float_as_int(gf));
                    }
                } else {
                    g = lu_task<T>(pool, t.a, t.out0, t.out1, t.out2, t.out3, lds, a.growth_limit, a.ctrl + 2,
                                   COH && !chained, nullptr, nullptr, LDS_HANDOFF && chained);
                }
            }
            break;
        case TK_LINV:
            tri_inv_task<T, true, COH>(pool, t.a, t.out0, lds);
            break;
        case TK_UINV:
            tri_inv_task<T, false, COH>(pool, t.a, t.out0, lds);
            break;
        case TK_LLT:
            if constexpr (LLT) llt_task<T>(pool, t.a, t.out0, t.out2, lds, COH && !chained, LDS_HANDOFF && chained);
            break;
        case TK_TRANS:
            if constexpr (LLT) trans_task<T, MASKS, COH>(pool, t.a, t.out0, lds, s_mask);
            break;
        default:
            break;
    }
    return g;
}

// A panel GEMM an lu task runs inside itself (DevTask::pan_l / pan_u, FP32 LU factors):
// the other operand (made earlier: its flag is waited for as a pair's) into LDS where
// the GEMM task puts it, the lu's inverse (lp: U^-1 as the second operand, else L^-1 as
// the first) from the LDS copy made beside its stores (from the stored tile after an
// FP64 refactorisation: copy is then reloaded), then the GEMM task's own product and
// epilogue, so the tile and its masks are the bits the GEMM task makes.  The panel's
// block masks are left in out_mask[w].  The lu's stores may still be in flight.
template <typename T, bool MASKS, bool COH>
__device__ __attribute__((always_inline)) void absorbed_panel(const Args& a, const DevTask& lu, const DevTask& pd, bool lp,
                                                              char* lds, T* copy, bool refactored, uint32_t* out_mask,
                                                              uint32_t* s_word) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    T* pool = reinterpret_cast<T*>(a.pool);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, li = lane & 15, lk = lane >> 4;
    const Pair pr = a.pairs[pd.a];
    T* icopy = copy;
    const int oslot = lp ? pair_a(pr) : pair_b(pr);
    if (tid == 0) *s_word = (uint32_t)spin_ready(a, lp ? pr.fa : pr.fb) & MASK_ALL;
    if (refactored) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the inverse is read back below
    __syncthreads();
    const uint32_t om = *s_word;
    T* la = reinterpret_cast<T*>(lds);
    T* lb = la + TS * LDA;
    tile_to_lds_masked<COH>(pool + (size_t)oslot * TE, lp ? la : lb, lp ? LDA : LDB, tid, om);
    if (refactored) tile_to_lds<true>(pool + (size_t)(lp ? lu.out3 : lu.out2) * TE, icopy, lp ? LDB : LDA, tid);
    __syncthreads();
    acc_t acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) acc[m][n] = acc_t{0, 0, 0, 0};
    const uint32_t ma = lp ? om : MASK_ALL, mb = lp ? MASK_ALL : om;
    if (!MASKS || (ma != 0 && mb != 0)) {  // a zero product is skipped, as in the GEMM task
        uint32_t kuse = 0xFu;
        if constexpr (MASKS) kuse = k_blocks_used(ma, mb, wr >> 4, wc >> 4);
        gemm_lds<T, false>(acc, lp ? la : icopy, lp ? icopy : lb, pr.b < 0 ? T(-1) : T(1), wr, wc, li, lk, kuse);
    }
    T* C = pool + (size_t)pd.out0 * TE;
    bool nz[2][2] = {{false, false}, {false, false}};
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const T v = acc[m][n][r];
                gst(&C[(wr + m * 16 + MfmaOp<T>::row(lk, r)) * TS + wc + n * 16 + li], v);
                if constexpr (MASKS) nz[m][n] |= fabs(v) > T(MASK_TINY);
            }
    if constexpr (MASKS) {
        const uint32_t qm = quad_mask(nz, wr >> 4, wc >> 4);
        if (lane == 0) out_mask[w] = qm;
    }
}

// Block mask of a finished task's outputs (thread 0, after the task's last barrier):
// GEMM, sub and transpose tasks leave their waves' bits in s_mask; the tiles of lu /
// llt and the inverses, and everything when masks are off, count as full.
template <bool MASKS>
__device__ inline uint32_t task_mask(const DevTask& t, const uint32_t* s_mask) {
    if (MASKS && (t.kind == TK_GEMM || t.kind == TK_SUB || t.kind == TK_TRANS))
        return s_mask[0] | s_mask[1] | s_mask[2] | s_mask[3];
    return MASK_ALL;
}
__device__ inline void write_masks(const DevTask& t, const Args& a, uint32_t m) {
    for (int32_t o : {t.out0, t.out1, t.out2, t.out3})
        if (o >= 0) gst(&a.mask[o], m);
}

// Persistent executor: every workgroup loops taking the next task index.  GEMM
// tasks wait for windows of pairs (pairs are ordered by when their operands become
// ready) and multiply each ready pair while loading the next; other tasks wait for
// all their sources first.  After a task, the workgroup may go on with tasks it claims
// (dynamic chaining, see build_chains): it reads the claim candidates' other operands'
// flags when the task starts, and at its end claims every candidate whose operands were
// all published by then, in candidate order, as far as the ring of CHAIN_RING claimed
// tasks has room; it runs them one after the other, each with the operand masks noted
// at the claim (a claimed task never waits).
// ONE: built for one resident workgroup per CU (twice the registers; FP64 GEMMs stage
// their next pair too), launched when the executor asks for one per CU.
constexpr int CHAIN_RING = 8;          // claimed tasks a workgroup holds at once
constexpr int CHAIN_MAX_PAIRS = 2;     // pairs of a chain candidate
constexpr int CHAIN_MAX_CANDS = 16;    // claim bits 16..31 of a ready flag
constexpr int CHAIN_MAX_CHECKS = 64;   // check entries per task: one lane of wave 0 each
template <typename T, bool BF16, bool LLT, bool MASKS, bool ONE, bool COH>
__global__ __launch_bounds__(NT, ONE ? 1 : 2) void k_factor_persistent(Args a) {
    // FP32 LU kernels built for one workgroup per CU: lu tasks run their panel GEMMs
    // (DevTask::pan_l / pan_u) inside themselves, keeping copies of the inverses in LDS
    // beyond what any task kind uses.  The kernels built for two (no registers to spare)
    // run them as after-links instead: right after the lu, in the same workgroup.
    constexpr bool ABSORB = sizeof(T) == 4 && !BF16 && !LLT && ONE;
    constexpr int XCOPY_BYTES = ABSORB ? TS * LDA * (int)sizeof(T) : 0;
    constexpr int ZCOPY_BYTES = ABSORB ? TS * LDB * (int)sizeof(T) : 0;
    __shared__ __attribute__((aligned(16))) char lds[lds_bytes<T>() + XCOPY_BYTES + ZCOPY_BYTES];
    T* const xcopy = ABSORB ? reinterpret_cast<T*>(lds + lds_bytes<T>()) : nullptr;
    T* const zcopy = ABSORB ? reinterpret_cast<T*>(lds + lds_bytes<T>() + XCOPY_BYTES) : nullptr;
    __shared__ uint32_t s_pmask[2][NT / 64];
    __shared__ uint32_t s_word;
    __shared__ int s_pend;  // an absorbed panel still to run as an after-link (kernels without the copies)
    __shared__ int s_task;
    __shared__ int s_abort;
    __shared__ int s_lost;
    __shared__ uint32_t s_mask[NT / 64];
    __shared__ uint32_t s_pm[PM_WORDS];
    __shared__ int s_ring_task[CHAIN_RING];
    __shared__ uint32_t s_ring_mask[CHAIN_RING][5];
    __shared__ int s_ring_head, s_ring_n;
    __shared__ uint32_t s_pre[5];
    __shared__ int s_next_ti, s_next_mode;
    const int tid = threadIdx.x;
    const unsigned long long want = (unsigned long long)(unsigned)a.epoch << 32;
    bool aborted = false;
    int n_ran = 0, n_claimed = 0, n_dropped = 0;  // thread 0: this workgroup's counts
    acquire_agent();  // every wave: lines left from earlier runs (see above), before its first load
    for (;;) {
        if (tid == 0) {
            s_abort = __hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_task = s_abort ? a.nqueue : atomicAdd(a.ctrl, a.chunk);
            s_ring_head = s_ring_n = 0;
            s_pend = -1;
        }
        __syncthreads();
        // wave-uniform loop control: the exit branch must not depend on the lane
        const int base = __builtin_amdgcn_readfirstlane(s_task);
        if (base >= a.nqueue || aborted) break;
        // a chunk is run in queue order; every dependency points to a task handed out
        // earlier (or to one chained to it), so with all workgroups resident no two
        // chunks can wait on each other; a claimed task waits for nothing
        const int end = min(base + a.chunk, a.nqueue);
      for (int qi = base; qi < end && !aborted; qi++) {
        // the task and its chain: a task with `next` hands its sum to the lu / llt that
        // alone reads it, which this workgroup runs at once, without a wait (the sum is
        // its only source, and the barrier below makes it visible to the workgroup);
        // then the tasks it claimed
        int ti = a.queue[qi];
        int mode = 0;  // 0: from the queue, 1: the static chain's next task, 2: claimed,
                       // 3: the after-link of the previous task (waits as a queued task)
        for (;;) {
        const DevTask t = a.tasks[ti];
        const int nxt = __builtin_amdgcn_readfirstlane(t.next);
        long long t_taken = 0, t_ready = 0, t_wait = 0;
        if (a.trace && tid == 0) t_taken = wall_clock64();
        if (t.kind != TK_GEMM && (mode == 0 || mode == 3)) {
            if (tid == 0) {
                for (int d = 0; d < t.dep_count; d++) spin_ready(a, a.deps[t.dep_begin + d]);
            }
            __syncthreads();
        }
        if (a.trace && tid == 0) t_wait = wall_clock64();
        if (mode != 1) {  // a chained task waits for nothing: the next queued task looks
            if (tid == 0) {
                s_abort = __hip_atomic_load(a.ctrl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_lost = 0;
            }
            __syncthreads();
            if (__builtin_amdgcn_readfirstlane(s_abort)) {
                aborted = true;
                break;
            }
        } else if (tid == 0) {
            s_lost = 0;  // read after the task's own barriers
        }
        // the claim candidates' operand flags, read now (wave 0, one lane per entry) and
        // looked at when the task is done
        const int nchk = a.chain ? t.chk_count : 0;
        unsigned long long cv = 0;
        if (tid < nchk) {
            const int32_t cs = a.chk[t.chk_begin + tid].slot;
            if (cs >= 0) cv = __hip_atomic_load(a.ready + cs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // an lu with absorbed panels (wave-uniform task fields): the panels' descriptors
        const bool absorbs = ABSORB && t.kind == TK_LU && (t.pan_l >= 0 || t.pan_u >= 0);
        const bool has_l = absorbs && t.pan_l >= 0, has_u = absorbs && t.pan_u >= 0;
        const DevTask pdl = has_l ? a.tasks[t.pan_l] : t;
        const DevTask pdu = has_u ? a.tasks[t.pan_u] : t;
        // the task this workgroup runs next (the static chain): touch its descriptor now,
        // so that its fetch at the top of the loop hits the cache (the loads are consumed
        // after this task; a GEMM's first wait for its operands covers their latency)
        int pf0 = 0, pf1 = 0, pf2 = 0;
        if (nxt >= 0) {
            const int32_t* pd = reinterpret_cast<const int32_t*>(a.tasks + nxt);
            pf0 = pd[0];
            pf1 = pd[sizeof(DevTask) / 8];
            pf2 = pd[sizeof(DevTask) / 4 - 1];
        }
        const double growth = run_task<T, BF16, false, LLT, MASKS, ONE, COH>(
            t, a, lds, s_mask,
            PairWait{&a, &t_wait, s_pm, mode == 2 ? s_pre : nullptr, mode == 0 ? t.ctag : 0, &s_lost, t.war_begin,
                     t.war_count},
            mode == 1, has_u ? xcopy : nullptr, has_l ? zcopy : nullptr, nxt >= 0);
        if (nxt >= 0) asm volatile("" ::"s"(pf0), "s"(pf1), "s"(pf2));
        if constexpr (ABSORB) {
            if (absorbs) {
                __syncthreads();  // the lu's LDS (the inverses' copies included) is complete
                const bool refactored = !(growth <= a.growth_limit);
                if (has_l) absorbed_panel<T, MASKS, COH>(a, t, pdl, true, lds, zcopy, refactored, s_pmask[0], &s_word);
                if (has_l && has_u) __syncthreads();  // the first panel's operands leave LDS
                if (has_u) absorbed_panel<T, MASKS, COH>(a, t, pdu, false, lds, xcopy, refactored, s_pmask[1], &s_word);
            }
        }
        if (__builtin_amdgcn_readfirstlane(s_lost)) {  // claimed by another workgroup: dropped
            n_dropped++;
            break;
        }
        // publish: every wave waits until its write-through tile stores have reached the
        // device coherence point, the barrier orders those waits before thread 0, and
        // thread 0 raises the flags, which carry the output masks to the consumers and the
        // claims of this workgroup.  A sum that goes to the chained task is read by nobody
        // else: no wait, no flag (the chained task runs on this CU and reads it after the
        // barrier).
        if (nxt < 0) wave_stores_done();
        __syncthreads();
        if (a.trace && tid == 0) t_ready = wall_clock64();  // work done, publish starts
        uint32_t claims = 0;
        if (nchk > 0 && tid < 64) {  // wave 0: which candidates can be claimed
            const uint32_t m_own = task_mask<MASKS>(t, s_mask);
            const bool on = tid < nchk;
            const ChainChk ce = on ? a.chk[t.chk_begin + tid] : ChainChk{-1, -1, -1, -1};
            const bool rdy = on && (ce.slot < 0 || cv >= want);
            uint32_t present = 0, blocked = 0;
            for (int c = 0; c < CHAIN_MAX_CANDS; c++) {
                if (__ballot(on && ce.cand == c)) present |= 1u << c;
                if (__ballot(on && ce.cand == c && !rdy)) blocked |= 1u << c;
            }
            claims = present & ~blocked;
            const int n0 = __builtin_amdgcn_readfirstlane(s_ring_n);
            const int h0 = __builtin_amdgcn_readfirstlane(s_ring_head);
            const int room = min(CHAIN_RING - n0, a.chain);
            while (__popc(claims) > room) claims &= ~(1u << (31 - __clz(claims)));  // keep the first
            if (on && ((claims >> ce.cand) & 1u)) {
                const int r = (h0 + n0 + __popc(claims & ((1u << ce.cand) - 1u))) % CHAIN_RING;
                s_ring_mask[r][ce.pos] = ce.slot < 0 ? m_own : ((uint32_t)cv & MASK_ALL);
                if (ce.pos == 0) s_ring_task[r] = ce.task;
            }
            if (tid == 0) s_ring_n = n0 + __popc(claims);
        }
        if (tid == 0) {
            const uint32_t m = task_mask<MASKS>(t, s_mask);
            if (nxt < 0) {
                const unsigned long long fv = want | ((unsigned long long)claims << 16) | m;
                __hip_atomic_store(a.ready + t.f0, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t.f1 >= 0)
                    __hip_atomic_store(a.ready + t.f1, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t.f2 >= 0)
                    __hip_atomic_store(a.ready + t.f2, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t.f3 >= 0)
                    __hip_atomic_store(a.ready + t.f3, fv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            for (int q = 0; q < 2; q++) {  // the absorbed panels, published with the lu
                if (!(q == 0 ? has_l : has_u)) continue;
                const DevTask& pd = q == 0 ? pdl : pdu;
                const uint32_t* pm = s_pmask[q];
                const uint32_t mp = MASKS ? (pm[0] | pm[1] | pm[2] | pm[3]) : MASK_ALL;
                __hip_atomic_store(a.ready + pd.f0, want | mp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                write_masks(pd, a, mp);
                if (a.trace) {
                    long long* tr = a.trace + (size_t)(q == 0 ? t.pan_l : t.pan_u) * 4;
                    gst(&tr[0], t_taken);
                    gst(&tr[1], t_ready);
                    gst(&tr[2], (long long)wall_clock64());
                    gst(&tr[3], t_wait | (3ll << 60));
                }
                n_ran++;
            }
            // consumers in this run read the masks from the flags; the mask array serves
            // the solve and the other executors, after the kernel
            write_masks(t, a, m);
            if (a.trace) {
                long long* tr = a.trace + (size_t)ti * 4;
                gst(&tr[0], t_taken);
                gst(&tr[1], t_ready);
                gst(&tr[2], (long long)wall_clock64());
                gst(&tr[3], t_wait | ((long long)mode << 60));  // bits 60-61: how the task started
            }
            n_claimed += __popc(claims);
            n_ran++;
            // the next task of this chain, decided here and read by every wave after the
            // barrier below (nothing writes these words again before the next such
            // barrier): the static chain's next task, else the after-link, else the
            // ring's head with its masks
            // (kernels without the inverses' copies run an lu's panels as after-links)
            const int pl = !absorbs && t.kind == TK_LU ? t.pan_l : -1;
            const int pu = !absorbs && t.kind == TK_LU ? t.pan_u : -1;
            if (nxt >= 0) {
                s_next_ti = nxt;
                s_next_mode = 1;
            } else if (pl >= 0 || pu >= 0) {
                s_next_ti = pl >= 0 ? pl : pu;
                s_pend = pl >= 0 ? pu : -1;
                s_next_mode = 3;
            } else if (t.after >= 0) {
                s_next_ti = t.after;
                s_next_mode = 3;
            } else if (s_pend >= 0) {
                s_next_ti = s_pend;
                s_pend = -1;
                s_next_mode = 3;
            } else if (s_ring_n > 0) {
                const int h = s_ring_head;
                s_next_ti = s_ring_task[h];
                for (int q = 0; q < 5; q++) s_pre[q] = s_ring_mask[h][q];
                s_ring_head = (h + 1) % CHAIN_RING;
                s_ring_n--;
                s_next_mode = 2;
            } else {
                s_next_mode = 0;
            }
        }
        // keeps the publishing branch and the next fetch branch apart, so the
        // back-edge cannot join them into a lane-dependent inner loop
        __syncthreads();
        mode = __builtin_amdgcn_readfirstlane(s_next_mode);
        if (mode == 0) break;
        ti = __builtin_amdgcn_readfirstlane(s_next_ti);
        }
      }
    }
    if (tid == 0) {  // one update per workgroup, off the task path
        if (n_claimed) atomicAdd(a.ctrl + 4, n_claimed);
        if (n_dropped) atomicAdd(a.ctrl + 5, n_dropped);
        atomicAdd(a.ctrl + 6, n_ran);
    }
}

// Levels / per-op executors: one workgroup per task of the launch.
template <typename T, bool BF16, bool ACC, bool LLT, bool MASKS, bool COH>
__global__ __launch_bounds__(NT) void k_factor_batch(Args a) {
    __shared__ __attribute__((aligned(16))) char lds[lds_bytes<T>()];
    __shared__ uint32_t s_mask[NT / 64];
    const DevTask t = a.tasks[a.task_offset + blockIdx.x];
    run_task<T, BF16, ACC, LLT, MASKS, false, COH>(t, a, lds, s_mask, NoWait{a.mask});
    __syncthreads();
    if (threadIdx.x == 0) write_masks(t, a, task_mask<MASKS>(t, s_mask));
}

__global__ void k_f64_to_f32(const double* in, float* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)in[i];
}

// Input tiles given as cells: tile t holds cells [begin[t], begin[t+1]), each an
// in-tile position (row*64 + col, < 4096) and a value; positions are distinct within
// a tile, so every store has one writer. The tiles are zeroed beforehand.
template <typename T>
__global__ void k_scatter_cells(T* pool, const int64_t* begin, const uint16_t* pos, const double* val,
                                int64_t n_tiles) {
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        T* tile = pool + (size_t)t * TE;
        const int64_t q1 = begin[t + 1];
        for (int64_t q = begin[t] + threadIdx.x; q < q1; q += blockDim.x) tile[pos[q]] = (T)val[q];
    }
}

// Block masks of tiles already in the pool (the inputs), one workgroup per tile; their
// ready flags say "always ready" (epoch field 0x7fffffff) and carry the mask.
template <typename T>
__global__ __launch_bounds__(NT) void k_tile_masks(const T* __restrict__ pool, uint32_t* mask,
                                                   unsigned long long* ready, int64_t n_tiles) {
    __shared__ uint32_t s[NT / 64];
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t b = tile_mask(pool + (size_t)t * TE, threadIdx.x);
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t m = s[0] | s[1] | s[2] | s[3];
            mask[t] = m;
            ready[t] = (0x7fffffffull << 32) | m;
        }
        __syncthreads();
    }
}

__global__ void k_f32_to_f64(const float* in, double* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (double)in[i];
}

__global__ void k_gather_tiles(const char* pool, const int32_t* slots, char* out, int n,
                               size_t tile_bytes) {
    const int t = blockIdx.x;
    if (t >= n) return;
    const uint4* src = reinterpret_cast<const uint4*>(pool + (size_t)slots[t] * tile_bytes);
    uint4* dst = reinterpret_cast<uint4*>(out + (size_t)t * tile_bytes);
    for (size_t q = threadIdx.x; q < tile_bytes / 16; q += blockDim.x) dst[q] = src[q];
}

// ------------------------------------------------------------------ task graph
namespace {

// Input tiles as cells (the plan's storage): see k_scatter_cells.
struct InputCells {
    const int64_t* begin = nullptr;  // n_inputs + 1
    const uint16_t* pos = nullptr;
    const double* val = nullptr;
};

// Host arrays of a task graph that the factor keeps only on the device.
struct HostGraph {
    std::vector<Pair> pairs, op_pairs;
    std::vector<int32_t> deps, lp, lc, ls, up, uc, us;
    std::vector<int32_t> kept;  // value ids that keep their slot to the end (slot reuse)
    std::vector<int32_t> pair_x, pair_y;  // the solve's two-row inverse blocks per row pair (-1: zero)
};

// Dynamic chaining of the persistent executor.  A workgroup that has just made a tile
// may run a task that reads it at once, without handing it to another workgroup
// (flag store, poll, acquire, operand loads from memory): it claims the task when all
// the task's other operands are already published, marking the claim in the ready
// flag it then publishes (bit 16 + the task's index in its candidate list), and the
// workgroup that took the task from the queue drops it when it sees that bit among its
// operands' flags.  Candidates are GEMM tasks of at most CHAIN_MAX_PAIRS pairs on or
// near the critical path (a longer GEMM's own workgroup has done most of its pairs by
// the time the last operand lands), each listed with every task that makes one of its
// pair operands.  The critical path comes from a cost model in microseconds (task
// work plus a hand-off per edge, measured on the C2 / C5 factors): slack(t) = length
// of the longest path - the longest path through t.
constexpr double CHAIN_SLACK_US = 1.0;

// The cost model of the critical path (microseconds, measured on the C2 / C5 factors):
// task work, and a hand-off per edge that no chain link removes.
double task_cost_us(const DevTask& k) {
    switch (k.kind) {
        case TK_LU: case TK_LLT: return 10.0;
        case TK_LINV: case TK_UINV: return 5.0;
        case TK_GEMM: return 0.8 + 0.5 * k.b + (k.add >= 0 ? 0.2 : 0.0);
        default: return 1.0;
    }
}
constexpr double HANDOFF_US = 2.5;  // publish + hand-off + operand loads from memory

void build_chains(Factor* f, HostGraph& hg) {
    const size_t T = f->tasks.size();
    f->chk.clear();
    f->n_chain_cands = 0;
    for (DevTask& k : f->tasks) {
        k.ctag = 0;
        k.chk_begin = k.chk_count = 0;
    }
    if (T == 0 || f->precision == SOGL_BF16) return;  // the bf16 GEMM waits pair by pair
    std::vector<int32_t> slot_task((size_t)f->n_slots, -1);
    for (size_t t = 0; t < T; t++)
        for (int32_t o : {f->tasks[t].out0, f->tasks[t].out1, f->tasks[t].out2, f->tasks[t].out3})
            if (o >= 0) slot_task[(size_t)o] = (int32_t)t;
    auto cost = [&](const DevTask& k) { return task_cost_us(k); };
    auto edge = [&](size_t p, size_t t) { return f->tasks[p].next == (int32_t)t ? 0.0 : HANDOFF_US; };
    std::vector<double> tl(T, 0.0), bl(T, 0.0);
    std::vector<std::vector<int32_t>> succ(T);
    for (size_t t = 0; t < T; t++) {  // tasks are in dependency-level order
        const DevTask& k = f->tasks[t];
        double st = 0;
        for (int32_t q = k.dep_begin; q < k.dep_begin + k.dep_count; q++) {
            const int32_t p = slot_task[(size_t)hg.deps[(size_t)q]];
            if (p < 0) continue;
            st = std::max(st, tl[(size_t)p] + cost(f->tasks[(size_t)p]) + edge((size_t)p, t));
            succ[(size_t)p].push_back((int32_t)t);
        }
        tl[t] = st;
    }
    double cp = 0;
    for (size_t t = T; t-- > 0;) {
        double b = 0;
        for (int32_t u : succ[t]) b = std::max(b, edge(t, (size_t)u) + bl[(size_t)u]);
        bl[t] = cost(f->tasks[t]) + b;
        cp = std::max(cp, tl[t] + bl[t]);
    }
    // operand slots of a GEMM task by position: pair lanes 0..2n-1, then the added tile
    auto positions = [&](const DevTask& k, int32_t* sl) {
        int n = 0;
        for (int32_t j = 0; j < k.b; j++) {
            const Pair& pr = hg.pairs[(size_t)(k.a + j)];
            sl[n++] = pair_a(pr);
            sl[n++] = pair_b(pr);
        }
        return n;
    };
    std::vector<std::vector<int32_t>> cand(T);
    std::vector<int32_t> nchk(T, 0);
    std::vector<int32_t> order;
    // after-links keep their chains in their heads' workgroups: neither end is claimed
    std::vector<char> linked(T, 0);
    for (size_t t = 0; t < T; t++)
        for (int32_t u : {f->tasks[t].after, f->tasks[t].pan_l, f->tasks[t].pan_u})
            if (u >= 0) linked[(size_t)u] = linked[t] = 1;
    for (size_t u = 0; u < T; u++) {
        const DevTask& k = f->tasks[u];
        if (k.kind == TK_GEMM && k.b >= 1 && k.b <= CHAIN_MAX_PAIRS && !linked[u] &&
            cp - (tl[u] + bl[u]) <= CHAIN_SLACK_US)
            order.push_back((int32_t)u);
    }
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return cp - (tl[(size_t)x] + bl[(size_t)x]) < cp - (tl[(size_t)y] + bl[(size_t)y]);
    });
    for (int32_t u : order) {  // most critical first: they get the low candidate indices
        const DevTask& k = f->tasks[(size_t)u];
        int32_t sl[2 * CHAIN_MAX_PAIRS];
        const int n = positions(k, sl);
        const int nent = n + (k.add >= 0 ? 1 : 0);
        bool any = false;
        for (int l = 0; l < n; l++) {
            const int32_t p = slot_task[(size_t)sl[l]];
            if (p < 0) continue;
            auto& cl = cand[(size_t)p];
            if (std::find(cl.begin(), cl.end(), u) != cl.end()) continue;
            if ((int)cl.size() >= CHAIN_MAX_CANDS || nchk[(size_t)p] + nent > CHAIN_MAX_CHECKS) continue;
            cl.push_back(u);
            nchk[(size_t)p] += nent;
            any = true;
        }
        f->n_chain_cands += any ? 1 : 0;
    }
    for (size_t p = 0; p < T; p++) {
        DevTask& kp = f->tasks[p];
        kp.chk_begin = (int32_t)f->chk.size();
        const int32_t outs[4] = {kp.out0, kp.out1, kp.out2, kp.out3};
        for (size_t i = 0; i < cand[p].size(); i++) {
            const int32_t u = cand[p][i];
            DevTask& ku = f->tasks[(size_t)u];
            int32_t sl[2 * CHAIN_MAX_PAIRS + 1];
            int n = positions(ku, sl);
            for (int l = 0; l < n; l++)
                if (slot_task[(size_t)sl[l]] == (int32_t)p) ku.ctag |= (int32_t)((i + 1) << (8 * l));
            if (ku.add >= 0) sl[n++] = ku.add;
            for (int l = 0; l < n; l++) {
                const bool own = std::find(outs, outs + 4, sl[l]) != outs + 4;
                f->chk.push_back(ChainChk{own ? -1 : sl[l], u, (int32_t)i, l < n - (ku.add >= 0 ? 1 : 0) ? l : 4});
            }
        }
        kp.chk_count = (int32_t)f->chk.size() - kp.chk_begin;
    }
}

// After-links of the persistent executor (off by default: measured slower).  The
// workgroup that ran task t runs t.after right after publishing t's outputs: a
// near-critical GEMM of at most AFTER_MAX_PAIRS pairs whose last operand (by the cost
// model) is t's, so the critical path skips the hand-off to another workgroup and reads
// t's tiles from this XCD's L2.  The linked task waits for its operands as a queued task
// does (its descriptor load and its producer's flag poll then come after the publish,
// where a queued consumer has both done already), and is not queued.
// Deadlock freedom: tasks are taken in index (dependency-level) order, each chain
// belongs to its head (the queued task that starts it), and a link t -> u is made only
// when every other producer of u's operands is an earlier member of t's chain or
// belongs to a chain whose head comes before t's head in the queue.  With every
// workgroup resident, the chain with the first unfinished head then waits only for
// finished work.  AFTER_MAX_LEN bounds a chain, and with it how far its head lies
// behind its last task (slot reuse takes over a slot only from values whose readers
// are REUSE_LEVEL_GAP levels back; the slot check verifies their heads as well).
constexpr int AFTER_MAX_PAIRS = 1;
constexpr int AFTER_MAX_LEN = 24;

// Absorbed panels (absorb_on: FP32 LU factors).  The lu task of a diagonal tile runs
// the one-pair panel GEMMs that multiply one of its inverses by a tile made earlier
// (L_IK = A_IK U_KK^-1 and U_KJ = L_KK^-1 A_KJ, BlockPlanner.cpp:865-939) inside itself,
// from the copies of the inverses it keeps in LDS, and publishes them with its own
// tiles: the hop from the lu to its panels (publish, hand-off, the inverse's load)
// leaves the critical path, for both panels the next Schur update waits for.  At most
// one of each kind, near-critical (by the cost model) with the lu as their last
// operand, under the same head order as after-links.
void build_after_links(Factor* f, const HostGraph& hg, std::vector<char>& linked, bool after_on,
                       bool absorb_on, const std::vector<int64_t>& rank) {
    const size_t T = f->tasks.size();
    f->n_after = 0;
    f->n_absorbed = 0;
    f->chain_head.assign(T, -1);
    std::vector<int32_t> slot_task((size_t)f->n_slots, -1);
    for (size_t t = 0; t < T; t++)
        for (int32_t o : {f->tasks[t].out0, f->tasks[t].out1, f->tasks[t].out2, f->tasks[t].out3})
            if (o >= 0) slot_task[(size_t)o] = (int32_t)t;
    auto edge = [&](size_t p, size_t t) { return f->tasks[p].next == (int32_t)t ? 0.0 : HANDOFF_US; };
    std::vector<double> fin(T, 0.0), bl(T, 0.0);  // earliest finish, longest path to the end
    std::vector<int32_t> crit(T, -1);              // the predecessor that finishes last
    std::vector<std::vector<int32_t>> succ(T);
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        double st = 0;
        for (int32_t q = k.dep_begin; q < k.dep_begin + k.dep_count; q++) {
            const int32_t p = slot_task[(size_t)hg.deps[(size_t)q]];
            if (p < 0) continue;
            const double e = fin[(size_t)p] + edge((size_t)p, t);
            if (crit[t] < 0 || e > st) {
                st = e;
                crit[t] = p;
            }
            if (succ[(size_t)p].empty() || succ[(size_t)p].back() != (int32_t)t) succ[(size_t)p].push_back((int32_t)t);
        }
        fin[t] = st + task_cost_us(k);
    }
    double cp = 0;
    for (size_t t = T; t-- > 0;) {
        double b = 0;
        for (int32_t u : succ[t]) b = std::max(b, edge(t, (size_t)u) + bl[(size_t)u]);
        bl[t] = task_cost_us(f->tasks[t]) + b;
        cp = std::max(cp, fin[t] - task_cost_us(f->tasks[t]) + bl[t]);
    }
    // slack of task t: the longest path minus the longest path through t
    auto slack = [&](size_t t) { return cp - (fin[t] - task_cost_us(f->tasks[t]) + bl[t]); };
    std::vector<int32_t> head(T, -1), len(T, 0), pred(T, -1);
    std::vector<char> absorbed(T, 0);
    // every producer of u's operands other than t is an earlier task of t's chain or
    // belongs to a chain whose head comes first
    auto heads_before = [&](size_t t, const DevTask& ku) {
        for (int32_t q = ku.dep_begin; q < ku.dep_begin + ku.dep_count; q++) {
            const int32_t p = slot_task[(size_t)hg.deps[(size_t)q]];
            if (p < 0 || p == (int32_t)t) continue;
            if ((size_t)p > t) return false;  // not placed yet
            bool member = false;
            for (int32_t x = pred[t]; x >= 0 && !member; x = pred[(size_t)x]) member = x == p;
            if (!member && rank[(size_t)head[(size_t)p]] >= rank[(size_t)head[t]]) return false;
        }
        return true;
    };
    auto link = [&](size_t t, int32_t u) {
        f->tasks[t].after = u;
        linked[(size_t)u] = 1;
        head[(size_t)u] = head[t];
        len[(size_t)u] = len[t] + 1;
        pred[(size_t)u] = (int32_t)t;
    };
    for (size_t t = 0; t < T; t++) {
        DevTask& k = f->tasks[t];
        if (head[t] < 0) head[t] = (int32_t)t;  // queued (or the static chain's head, below)
        if (k.next >= 0) {  // the lu of this GEMM's sum: same workgroup, same chain
            head[(size_t)k.next] = head[t];
            len[(size_t)k.next] = len[t] + 1;
            pred[(size_t)k.next] = (int32_t)t;
            continue;
        }
        if (absorbed[t]) continue;  // runs inside its lu: no link of its own
        if (len[t] + 1 >= AFTER_MAX_LEN) continue;
        if (absorb_on && k.kind == TK_LU) {
            auto own = [&](int32_t v) { return v == k.out0 || v == k.out1 || v == k.out2 || v == k.out3; };
            int32_t best[2] = {-1, -1};  // A U^-1, L^-1 A
            for (int32_t u : succ[t]) {
                const DevTask& ku = f->tasks[(size_t)u];
                if (ku.kind != TK_GEMM || ku.b != 1 || ku.add >= 0 || ku.next >= 0 || linked[(size_t)u] ||
                    crit[(size_t)u] != (int32_t)t || slack((size_t)u) > CHAIN_SLACK_US || !heads_before(t, ku))
                    continue;
                const Pair& pr = hg.pairs[(size_t)ku.a];
                int which = -1;
                if (pr.a >= 0 && k.out3 >= 0 && pair_b(pr) == k.out3 && !own(pair_a(pr))) which = 0;
                else if (pr.a >= 0 && k.out2 >= 0 && pair_a(pr) == k.out2 && !own(pair_b(pr))) which = 1;
                if (which >= 0 && (best[which] < 0 || bl[(size_t)u] > bl[(size_t)best[which]])) best[which] = u;
            }
            for (int which = 0; which < 2; which++) {
                const int32_t u = best[which];
                if (u < 0) continue;
                (which == 0 ? k.pan_l : k.pan_u) = u;
                linked[(size_t)u] = 1;
                head[(size_t)u] = head[t];
                len[(size_t)u] = len[t] + 1;
                pred[(size_t)u] = (int32_t)t;
                absorbed[(size_t)u] = 1;
                f->n_absorbed++;
            }
            if (best[0] >= 0 || best[1] >= 0) continue;
        }
        if (!after_on) continue;
        int32_t best = -1;
        for (int32_t u : succ[t]) {
            const DevTask& ku = f->tasks[(size_t)u];
            if (ku.kind != TK_GEMM || ku.b > AFTER_MAX_PAIRS || linked[(size_t)u] || crit[(size_t)u] != (int32_t)t ||
                slack((size_t)u) > CHAIN_SLACK_US)
                continue;
            if (heads_before(t, ku) && (best < 0 || bl[(size_t)u] > bl[(size_t)best])) best = u;
        }
        if (best < 0) continue;
        link(t, best);
        f->n_after++;
    }
    f->chain_head = head;
}

// The task graph of a factor description, on the host only (no device calls).
int build_host_graph(Context* c, const sogl_factor_desc* d, Factor* f, HostGraph& hg) {
    const int32_t S = d->n_ids;
    std::vector<int32_t> slot(S > 0 ? (size_t)S : 1, -1);
    int32_t next = 0;
    for (int64_t k = 0; k < d->n_inputs; k++) {
        int32_t id = d->input_ids[k];
        if (id <= 0 || id >= S) return set_error(c, SOGL_ERR_ARG, "input id out of range");
        if (slot[(size_t)id] >= 0) return set_error(c, SOGL_ERR_ARG, "duplicate input id");
        slot[(size_t)id] = next++;
        f->input_slots.push_back(slot[(size_t)id]);
    }
    // readers of each id (op operands; factor tiles count as read by the solve): a
    // temporary with exactly one reader can be folded into that reader
    std::vector<int32_t> readers((size_t)S, 0);
    for (int64_t k = 0; k < d->n_ops; k++) {
        const sogl_op& o = d->ops[k];
        for (int32_t r : {o.result, o.result2})
            if (r >= S) return set_error(c, SOGL_ERR_ARG, "result id out of range");
        if (o.src >= S || o.src2 >= S || o.src < 0 || o.src2 < 0)
            return set_error(c, SOGL_ERR_ARG, "source id out of range");
        if (o.src > 0) readers[(size_t)o.src]++;
        if (o.src2 > 0) readers[(size_t)o.src2]++;
    }
    const int32_t R = d->block_rows;
    for (int which = 0; which < 2; which++) {
        const int64_t cnt = which == 0 ? d->n_l : d->n_u;
        const int32_t* bi = which == 0 ? d->l_bi : d->u_bi;
        const int32_t* bj = which == 0 ? d->l_bj : d->u_bj;
        const int32_t* id = which == 0 ? d->l_ids : d->u_ids;
        for (int64_t k = 0; k < cnt; k++) {
            if (bi[k] < 0 || bi[k] >= R || bj[k] < 0 || bj[k] >= R || id[k] <= 0 || id[k] >= S)
                return set_error(c, SOGL_ERR_ARG, "factor tile out of range");
            readers[(size_t)id[k]]++;
        }
    }
    // an op list with llt ops is an L L^T factorisation: its U is L^T, made on the
    // device by transposing the L tiles after they are final
    bool llt = false, has_lu = false, has_mult = false;
    for (int64_t k = 0; k < d->n_ops; k++) {
        llt |= d->ops[k].op == SOGL_OP_LLT;
        has_lu |= d->ops[k].op == SOGL_OP_LU;
        has_mult |= d->ops[k].op == SOGL_OP_MULT;
    }
    if (llt && has_lu) return set_error(c, SOGL_ERR_ARG, "an op list has both lu and llt ops");
    if (has_mult && !llt) return set_error(c, SOGL_ERR_ARG, "mult ops belong to an L L^T op list (llt ops)");
    if (llt && d->n_u != 0) return set_error(c, SOGL_ERR_ARG, "an L L^T factor takes no U tiles");
    f->llt = llt;
    f->zero_slot = next++;

    // producer task of each slot (-1: input or never written); sealed: the slot has
    // been read, so no later op may write it.  Result slots are made on first write.
    std::vector<int32_t> prod((size_t)next, -1);
    std::vector<char> sealed((size_t)next, 0);
    auto new_slot = [&]() -> int32_t {
        prod.push_back(-1);
        sealed.push_back(0);
        return next++;
    };
    std::vector<int32_t> lu_op_task;  // per task: index of its op in op_tasks (lu tasks)
    std::vector<DevTask> tasks;
    std::vector<std::vector<Pair>> tpairs;
    std::vector<int32_t> tstage;
    std::vector<sogl_op> opl(d->ops, d->ops + d->n_ops);
    std::vector<Pair>& op_pairs = hg.op_pairs;
    f->op_tasks.clear();
    // the per-op executor's tables (one entry per unit op): only for factors that may
    // run without slot reuse
    const bool keep_ops = f->keep_ops;
    auto src_slot = [&](int32_t id) -> int32_t { return id > 0 ? slot[(size_t)id] : -1; };
    for (int64_t k = 0; k < d->n_ops; k++) {
        const sogl_op& o = opl[(size_t)k];
        if (o.result <= 0) continue;
        // sub and inverse results may fold into another tile's slot: theirs is made
        // only when they do not (ensure_rs below)
        const bool may_fold = o.op == SOGL_OP_SUB || o.op == SOGL_OP_LOWER_INV || o.op == SOGL_OP_UPPER_INV;
        if (!may_fold && slot[(size_t)o.result] < 0) slot[(size_t)o.result] = new_slot();
        if (o.op == SOGL_OP_LU && o.result2 > 0 && slot[(size_t)o.result2] < 0)
            slot[(size_t)o.result2] = new_slot();
        int32_t rs = slot[(size_t)o.result];
        int kind;
        switch (o.op) {
            case SOGL_OP_MUL: case SOGL_OP_MULNEG: case SOGL_OP_MULT: kind = TK_GEMM; break;
            case SOGL_OP_SUB: kind = TK_SUB; break;
            case SOGL_OP_LU: kind = TK_LU; break;
            case SOGL_OP_LLT: kind = TK_LLT; break;
            case SOGL_OP_LOWER_INV: kind = TK_LINV; break;
            case SOGL_OP_UPPER_INV: kind = TK_UINV; break;
            default:
                return set_error(c, SOGL_ERR_UNSUPPORTED,
                                 "unit op " + std::to_string(o.op) + " is outside the factor path");
        }
        DevTask ot{};
        ot.kind = kind;
        ot.out0 = rs;
        ot.out1 = (kind == TK_LU && o.result2 > 0) ? slot[(size_t)o.result2] : -1;
        ot.level = o.stage;
        int32_t s1 = src_slot(o.src), s2 = src_slot(o.src2);
        if ((o.src > 0 && s1 < 0) || (o.src2 > 0 && s2 < 0))
            return set_error(c, SOGL_ERR_ARG, "op " + std::to_string(k) + " reads a tile that is neither an input "
                                              "nor written by an earlier op");
        Pair pr{0, 0};
        if (kind == TK_GEMM) {
            if (s1 < 0) s1 = f->zero_slot;
            if (s2 < 0) s2 = f->zero_slot;
            // mult: C += A B^T, the transposed second operand flagged on a
            pr = Pair{o.op == SOGL_OP_MULT ? -s1 - 1 : s1, o.op == SOGL_OP_MULNEG ? -s2 - 1 : s2};
            ot.a = (int32_t)op_pairs.size();
            ot.b = 1;
            if (keep_ops) op_pairs.push_back(pr);
        } else if (kind == TK_SUB) {
            ot.a = s2;
            ot.b = s1;
        } else {
            ot.a = s1 < 0 ? f->zero_slot : s1;
            ot.b = 0;
        }
        ot.out2 = ot.out3 = -1;
        if (keep_ops) f->op_tasks.push_back(ot);
        auto seal_sources = [&]() {
            if (s1 >= 0) sealed[(size_t)s1] = 1;
            if (s2 >= 0) sealed[(size_t)s2] = 1;
        };
        // the result id reads slot `to` from now on; the per-op executor runs this op
        // in place on that slot
        auto alias = [&](int32_t to) {
            slot[(size_t)o.result] = to;
            if (keep_ops) f->op_tasks.back().out0 = to;
            seal_sources();
            f->n_folded++;
        };
        auto ensure_rs = [&]() {
            if (rs >= 0) return;
            rs = slot[(size_t)o.result] = new_slot();
            ot.out0 = rs;
            if (keep_ops) f->op_tasks.back().out0 = rs;
        };
        if (kind == TK_SUB && rs < 0) {
            // copy: the result is its source (exact)
            if (o.src == 0 && o.src2 > 0) {
                alias(s2);
                continue;
            }
            // a single-use GEMM sum M folded into this op: src2 - M, -M (src2 == 0) or
            // M - src.  The GEMM task takes the sign and the added tile; exact up to the
            // order of the additions.
            auto foldable = [&](int32_t id) -> int32_t {
                if (id <= 0 || readers[(size_t)id] != 1) return -1;
                const int32_t pt = prod[(size_t)slot[(size_t)id]];
                return pt >= 0 && tasks[(size_t)pt].kind == TK_GEMM ? pt : -1;
            };
            int32_t pt = foldable(o.src);
            if (pt >= 0 && (o.src2 == 0 || (tasks[(size_t)pt].add < 0 && s2 != s1))) {
                DevTask& g = tasks[(size_t)pt];
                for (Pair& q : tpairs[(size_t)pt]) q.b = -q.b - 1;  // mul <-> mulneg
                if (o.src2 == 0) {
                    g.add_neg ^= 1;
                } else {
                    g.add = s2;
                    g.add_neg = 0;
                }
                alias(s1);
                continue;
            }
            pt = foldable(o.src2);
            if (pt >= 0 && o.src > 0 && tasks[(size_t)pt].add < 0 && s1 != s2) {
                tasks[(size_t)pt].add = s1;
                tasks[(size_t)pt].add_neg = 1;
                alias(s2);
                continue;
            }
        }
        // an inverse of an lu output is made by that lu's task (the lu holds the tile
        // already); the per-op executor still runs it as its own op.  The recursive
        // planner inverts some diagonal factor tiles more than once (C2: 3,679
        // lowerInv ops over 2,048 lu ops); a repeat has the same value, so its result
        // id aliases the first inverse's slot and no task is made for it.
        if (kind == TK_LINV || kind == TK_UINV) {
            const int32_t pt = s1 >= 0 ? prod[(size_t)s1] : -1;
            const bool lower = kind == TK_LINV;
            const int32_t pk = pt >= 0 ? tasks[(size_t)pt].kind : -1;
            if ((pk == TK_LU || (pk == TK_LLT && lower)) && rs < 0 &&
                (lower ? tasks[(size_t)pt].out0 : tasks[(size_t)pt].out1) == s1) {
                int32_t inv = lower ? tasks[(size_t)pt].out2 : tasks[(size_t)pt].out3;
                if (inv < 0) {
                    ensure_rs();
                    (lower ? tasks[(size_t)pt].out2 : tasks[(size_t)pt].out3) = rs;
                    prod[(size_t)rs] = pt;
                } else {
                    slot[(size_t)o.result] = inv;
                    if (keep_ops) f->op_tasks.back().out0 = inv;
                    f->n_inv_aliased++;
                }
                seal_sources();
                continue;
            }
        }
        ensure_rs();
        if (sealed[(size_t)rs])
            return set_error(c, SOGL_ERR_ARG, "op list writes a tile after it was read");
        int32_t t = prod[(size_t)rs];
        if (t < 0) {
            t = (int32_t)tasks.size();
            DevTask nt = ot;
            if (kind == TK_GEMM) { nt.a = 0; nt.b = 0; }
            tasks.push_back(nt);
            tpairs.emplace_back();
            tstage.push_back(o.stage);
            lu_op_task.push_back(keep_ops ? (int32_t)f->op_tasks.size() - 1 : -1);
            prod[(size_t)rs] = t;
            if (ot.out1 >= 0) prod[(size_t)ot.out1] = t;
        } else {
            if (tasks[(size_t)t].kind != TK_GEMM || kind != TK_GEMM)
                return set_error(c, SOGL_ERR_ARG, "a tile is written by more than one non-accumulating op");
            tstage[(size_t)t] = std::max(tstage[(size_t)t], o.stage);
        }
        if (kind == TK_GEMM) tpairs[(size_t)t].push_back(pr);
        seal_sources();
    }
    // diagonal factor tiles per block row (after the folds: ids may share slots)
    f->l_diag.assign((size_t)R, -1);
    f->u_diag.assign((size_t)R, -1);
    for (int which = 0; which < 2; which++) {
        const int64_t cnt = which == 0 ? d->n_l : d->n_u;
        const int32_t* bi = which == 0 ? d->l_bi : d->u_bi;
        const int32_t* id = which == 0 ? d->l_ids : d->u_ids;
        for (int64_t k = 0; k < cnt; k++)
            if (bi[k] == (which == 0 ? d->l_bj : d->u_bj)[k])
                (which == 0 ? f->l_diag : f->u_diag)[(size_t)bi[k]] = slot[(size_t)id[k]];
    }
    for (int32_t i = 0; i < R; i++)
        if (f->l_diag[(size_t)i] < 0 || (!llt && f->u_diag[(size_t)i] < 0))
            return set_error(c, SOGL_ERR_ARG, "factor lacks a diagonal tile");
    // inverses of the diagonal factor tiles for the triangular solve: made by the lu
    // task of each diagonal tile, into the plan's inverse tile when the plan has one
    f->linv_slot.assign((size_t)R, -1);
    f->uinv_slot.assign((size_t)R, -1);
    for (int32_t i = 0; i < R; i++) {
        for (int which = 0; which < (llt ? 1 : 2); which++) {
            const int32_t ds = (which == 0 ? f->l_diag : f->u_diag)[(size_t)i];
            const int32_t pt = ds >= 0 && ds < (int32_t)prod.size() ? prod[(size_t)ds] : -1;
            const int32_t pk = pt >= 0 ? tasks[(size_t)pt].kind : -1;
            const bool from_lu = (pk == TK_LU || (pk == TK_LLT && which == 0)) &&
                                 (which == 0 ? tasks[(size_t)pt].out0 : tasks[(size_t)pt].out1) == ds;
            int32_t inv;
            if (from_lu) {
                inv = which == 0 ? tasks[(size_t)pt].out2 : tasks[(size_t)pt].out3;
                if (inv < 0) {  // the plan never inverts this tile: a solve-only slot
                    inv = new_slot();
                    prod[(size_t)inv] = pt;
                    (which == 0 ? tasks[(size_t)pt].out2 : tasks[(size_t)pt].out3) = inv;
                    if (keep_ops) {
                        DevTask& ot = f->op_tasks[(size_t)lu_op_task[(size_t)pt]];
                        (which == 0 ? ot.out2 : ot.out3) = inv;
                    }
                }
            } else {
                // a diagonal tile that no lu produced (hand-built descriptions): its own
                // inverse task, general diagonal
                inv = new_slot();
                DevTask nt{};
                nt.kind = which == 0 ? TK_LINV : TK_UINV;
                nt.out0 = inv;
                nt.out1 = nt.out2 = nt.out3 = -1;
                nt.a = ds;
                tasks.push_back(nt);
                tpairs.emplace_back();
                tstage.push_back(0);
                lu_op_task.push_back(-1);
                prod[(size_t)inv] = (int32_t)tasks.size() - 1;
                if (keep_ops) f->op_tasks.push_back(nt);
            }
            (which == 0 ? f->linv_slot : f->uinv_slot)[(size_t)i] = inv;
        }
    }
    // L L^T: every L tile and every diagonal inverse transposed into a slot of its own
    // (U_JI = L_IJ^T, U_II^-1 = (L_II^-1)^T), by transpose tasks after their sources
    std::map<int32_t, int32_t> trans_of;
    auto transposed = [&](int32_t src) -> int32_t {
        auto it = trans_of.find(src);
        if (it != trans_of.end()) return it->second;
        const int32_t ts = new_slot();
        DevTask nt{};
        nt.kind = TK_TRANS;
        nt.out0 = ts;
        nt.out1 = nt.out2 = nt.out3 = -1;
        nt.a = src;
        tasks.push_back(nt);
        tpairs.emplace_back();
        tstage.push_back(0);
        lu_op_task.push_back(-1);
        prod[(size_t)ts] = (int32_t)tasks.size() - 1;
        if (keep_ops) f->op_tasks.push_back(nt);
        trans_of[src] = ts;
        return ts;
    };
    std::vector<int32_t> l_trans;
    if (llt) {
        for (int64_t k = 0; k < d->n_l; k++) {
            if (d->l_ids[k] <= 0 || d->l_ids[k] >= S) return set_error(c, SOGL_ERR_ARG, "factor tile out of range");
            const int32_t s0 = slot[(size_t)d->l_ids[k]];
            if (s0 < 0) return set_error(c, SOGL_ERR_ARG, "factor tile never produced");
            l_trans.push_back(transposed(s0));
        }
        for (int32_t i = 0; i < R; i++) {
            f->u_diag[(size_t)i] = transposed(f->l_diag[(size_t)i]);
            f->uinv_slot[(size_t)i] = transposed(f->linv_slot[(size_t)i]);
        }
    }
    // The solve's two-row steps (block rows I = 2p, I + 1 together): the off-diagonal
    // block of the inverse of [L_II 0; L_I+1,I L_I+1,I+1],
    //     X_p = -L_I+1,I+1^-1 (L_I+1,I L_II^-1),
    // and of [U_II U_I,I+1; 0 U_I+1,I+1],
    //     Y_p = -U_II^-1 (U_I,I+1 U_I+1,I+1^-1),
    // each as two GEMM tasks of the factor over final factor tiles (nothing in the factor
    // reads them; the solve does).  Without the sub-diagonal tile the block is zero.
    {
        std::map<std::pair<int32_t, int32_t>, int32_t> lt, ut;  // (block row, block col) -> slot
        // (ids out of range or never produced are reported by the solve structure below)
        auto ok_id = [&](int32_t id) { return id > 0 && id < S && slot[(size_t)id] >= 0; };
        for (int64_t k = 0; k < d->n_l; k++) {
            if (!ok_id(d->l_ids[k])) continue;
            lt[{d->l_bi[k], d->l_bj[k]}] = slot[(size_t)d->l_ids[k]];
            if (llt) ut[{d->l_bj[k], d->l_bi[k]}] = l_trans[(size_t)k];  // U_JI = L_IJ^T
        }
        for (int64_t k = 0; k < d->n_u && !llt; k++)
            if (ok_id(d->u_ids[k])) ut[{d->u_bi[k], d->u_bj[k]}] = slot[(size_t)d->u_ids[k]];
        f->n_pair_tasks = 0;
        auto gemm = [&](int32_t x, int32_t y, bool neg) -> int32_t {  // a new slot = (neg ? -1 : 1) x y
            const int32_t out = new_slot();
            f->n_pair_tasks++;
            const Pair pr{x, neg ? -y - 1 : y};
            DevTask nt{};
            nt.kind = TK_GEMM;
            nt.out0 = out;
            nt.out1 = nt.out2 = nt.out3 = -1;
            tasks.push_back(nt);
            tpairs.push_back({pr});
            tstage.push_back(0);
            lu_op_task.push_back(-1);
            prod[(size_t)out] = (int32_t)tasks.size() - 1;
            if (keep_ops) {
                nt.a = (int32_t)op_pairs.size();
                nt.b = 1;
                op_pairs.push_back(pr);
                f->op_tasks.push_back(nt);
            }
            return out;
        };
        hg.pair_x.assign((size_t)(R / 2), -1);
        hg.pair_y.assign((size_t)(R / 2), -1);
        for (int32_t q = 0; q < R / 2; q++) {
            const int32_t I = 2 * q;
            auto lsub = lt.find({I + 1, I});
            if (lsub != lt.end())
                hg.pair_x[(size_t)q] = gemm(f->linv_slot[(size_t)I + 1], gemm(lsub->second, f->linv_slot[(size_t)I], false), true);
            auto usub = ut.find({I, I + 1});
            if (usub != ut.end())
                hg.pair_y[(size_t)q] = gemm(f->uinv_slot[(size_t)I], gemm(usub->second, f->uinv_slot[(size_t)I + 1], false), true);
        }
    }
    f->n_slots = next;

    // dependencies and levels
    const size_t T = tasks.size();
    std::vector<std::vector<int32_t>> dslots(T);
    std::vector<int32_t> level(T, 0);
    auto add_dep = [&](size_t t, int32_t s) {
        if (s < 0) return;
        int32_t p = prod[(size_t)s];
        if (p < 0) return;
        dslots[t].push_back(s);
    };
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = tasks[t];
        if (k.kind == TK_GEMM) {
            for (const Pair& p : tpairs[t]) {
                add_dep(t, pair_a(p));
                add_dep(t, pair_b(p));
            }
            add_dep(t, k.add);
        } else if (k.kind == TK_SUB) {
            add_dep(t, k.a);
            add_dep(t, k.b);
        } else {
            add_dep(t, k.a);
        }
        std::sort(dslots[t].begin(), dslots[t].end());
        dslots[t].erase(std::unique(dslots[t].begin(), dslots[t].end()), dslots[t].end());
    }
    // tasks were created in plan order, which is stage order, so producers precede consumers
    for (size_t t = 0; t < T; t++) {
        int32_t lv = 0;
        for (int32_t s : dslots[t]) {
            int32_t p = prod[(size_t)s];
            if ((size_t)p >= t)
                return set_error(c, SOGL_ERR_ARG, "op list is not in dependency order");
            lv = std::max(lv, level[(size_t)p] + 1);
        }
        level[t] = lv;
    }
    // pairs of a GEMM task run in the order their operands become ready, so the task
    // waits only for its last pair once the earlier ones are done
    auto ready_level = [&](int32_t s) -> int32_t { return prod[(size_t)s] < 0 ? -1 : level[(size_t)prod[(size_t)s]]; };
    auto pair_level = [&](const Pair& x) -> int32_t {
        return std::max(ready_level(pair_a(x)), ready_level(pair_b(x)));
    };
    for (size_t t = 0; t < T; t++) {
        if (tasks[t].kind != TK_GEMM) continue;
        std::stable_sort(tpairs[t].begin(), tpairs[t].end(),
                         [&](const Pair& x, const Pair& y) { return pair_level(x) < pair_level(y); });
    }
    const size_t T2 = tasks.size();
    std::vector<int32_t> order(T2);
    for (size_t t = 0; t < T2; t++) order[t] = (int32_t)t;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return level[(size_t)x] < level[(size_t)y];
    });
    f->tasks.clear();
    f->tasks.reserve(T);
    std::vector<Pair>& pairs = hg.pairs;
    std::vector<int32_t>& deps = hg.deps;
    int32_t maxlv = T2 ? level[(size_t)order.back()] : -1;
    f->level_begin.assign((size_t)maxlv + 2, 0);
    for (int32_t t : order) {
        DevTask k = tasks[(size_t)t];
        k.level = level[(size_t)t];
        if (k.kind == TK_GEMM) {
            k.a = (int32_t)pairs.size();
            k.b = (int32_t)tpairs[(size_t)t].size();
            pairs.insert(pairs.end(), tpairs[(size_t)t].begin(), tpairs[(size_t)t].end());
        }
        k.dep_begin = (int32_t)deps.size();
        k.dep_count = (int32_t)dslots[(size_t)t].size();
        deps.insert(deps.end(), dslots[(size_t)t].begin(), dslots[(size_t)t].end());
        f->level_begin[(size_t)k.level + 1]++;
        f->tasks.push_back(k);
    }
    for (size_t v = 1; v < f->level_begin.size(); v++) f->level_begin[v] += f->level_begin[v - 1];
    f->n_pairs = (int64_t)pairs.size();
    f->n_deps = (int64_t)deps.size();

    // chains for the persistent executor: an lu / llt whose source is a GEMM sum that
    // nothing else reads runs in the workgroup that made the sum, right after it (no
    // publish, hand-off or wait in between); it is then not handed out by the queue
    {
        std::vector<int32_t> nread((size_t)next, 0);
        auto rd = [&](int32_t sl) { if (sl >= 0 && sl < next) nread[(size_t)sl]++; };
        for (size_t t = 0; t < tasks.size(); t++) {
            const DevTask& k = tasks[t];
            if (k.kind == TK_GEMM) {
                for (const Pair& p : tpairs[t]) { rd(pair_a(p)); rd(pair_b(p)); }
                rd(k.add);
            } else if (k.kind == TK_SUB) {
                rd(k.a);
                rd(k.b);
            } else {
                rd(k.a);
            }
        }
        for (int64_t k = 0; k < d->n_l; k++) rd(slot[(size_t)d->l_ids[k]]);
        for (int64_t k = 0; k < d->n_u; k++) rd(slot[(size_t)d->u_ids[k]]);
        for (const auto* v : {&f->l_diag, &f->u_diag, &f->linv_slot, &f->uinv_slot})
            for (int32_t sl : *v) rd(sl);
        std::vector<int32_t> final_of(tasks.size(), -1);
        for (size_t i = 0; i < order.size(); i++) final_of[(size_t)order[i]] = (int32_t)i;
        std::vector<char> chained(f->tasks.size(), 0);
        f->n_chained = 0;
        for (size_t u = 0; u < f->tasks.size(); u++) {
            const DevTask& k = f->tasks[u];
            if (k.kind != TK_LU && k.kind != TK_LLT) continue;
            const int32_t sl = k.a;
            const int32_t po = sl >= 0 && sl < (int32_t)prod.size() ? prod[(size_t)sl] : -1;
            if (po < 0 || tasks[(size_t)po].kind != TK_GEMM || nread[(size_t)sl] != 1) continue;
            DevTask& g = f->tasks[(size_t)final_of[(size_t)po]];
            if (g.next >= 0 || g.out0 != sl) continue;
            g.next = (int32_t)u;
            chained[u] = 1;
            f->n_chained++;
        }
        for (DevTask& k : f->tasks) k.after = k.pan_l = k.pan_u = -1;
        // queue order: the task order (by dependency level).  Handing the GEMMs whose lu
        // follows them out last in their level lets more panels pass the head order
        // below (C5: 23k instead of 8k) but delays the diagonal chain (C5 FP32 factor
        // 375 -> 382 ms with the panels, FP64 485 -> 496 ms without), so the queue keeps
        // the task order; the checks compare queue positions either way.
        std::vector<int64_t> rank(f->tasks.size());
        for (size_t u = 0; u < f->tasks.size(); u++) rank[u] = (int64_t)u;
        const bool absorb = f->absorb_panels && f->precision == SOGL_FP32 && !llt;
        if (f->after_links || absorb) build_after_links(f, hg, chained, f->after_links, absorb, rank);
        else f->chain_head.clear();
        f->queue.clear();
        for (size_t u = 0; u < f->tasks.size(); u++)
            if (!chained[u]) f->queue.push_back((int32_t)u);
        std::sort(f->queue.begin(), f->queue.end(), [&](int32_t x, int32_t y) { return rank[(size_t)x] < rank[(size_t)y]; });
    }
    build_chains(f, hg);

    // triangular-solve structure
    std::vector<std::vector<std::pair<int32_t, int32_t>>> lrow((size_t)R), urow((size_t)R);
    f->l_slots.clear();
    f->u_slots.clear();
    if (llt)  // U = L^T: block (j, i) of U is the transpose of L's block (i, j)
        for (int64_t k = 0; k < d->n_l; k++)
            if (d->l_bi[k] != d->l_bj[k]) {
                if (d->l_bj[k] > d->l_bi[k]) return set_error(c, SOGL_ERR_ARG, "factor tile on the wrong side of the diagonal");
                urow[(size_t)d->l_bj[k]].push_back({d->l_bi[k], l_trans[(size_t)k]});
            }
    for (int which = 0; which < 2; which++) {
        const int64_t cnt = which == 0 ? d->n_l : d->n_u;
        const int32_t* bi = which == 0 ? d->l_bi : d->u_bi;
        const int32_t* bj = which == 0 ? d->l_bj : d->u_bj;
        const int32_t* id = which == 0 ? d->l_ids : d->u_ids;
        for (int64_t k = 0; k < cnt; k++) {
            if (bi[k] < 0 || bi[k] >= R || bj[k] < 0 || bj[k] >= R || id[k] <= 0 || id[k] >= S)
                return set_error(c, SOGL_ERR_ARG, "factor tile out of range");
            int32_t s = slot[(size_t)id[k]];
            if (s < 0) return set_error(c, SOGL_ERR_ARG, "factor tile never produced");
            (which == 0 ? f->l_slots : f->u_slots).push_back(s);
            if (bi[k] == bj[k]) {
                continue;
            } else if ((which == 0) == (bj[k] < bi[k])) {
                (which == 0 ? lrow : urow)[(size_t)bi[k]].push_back({bj[k], s});
            } else {
                return set_error(c, SOGL_ERR_ARG, "factor tile on the wrong side of the diagonal");
            }
        }
    }
    std::vector<int32_t>&lp = hg.lp, &lc = hg.lc, &ls = hg.ls, &up = hg.up, &uc = hg.uc, &us = hg.us;
    lp.assign(1, 0);
    up.assign(1, 0);
    for (int32_t i = 0; i < R; i++) {
        auto& L = lrow[(size_t)i];
        std::sort(L.begin(), L.end());
        for (auto& e : L) { lc.push_back(e.first); ls.push_back(e.second); }
        lp.push_back((int32_t)lc.size());
        auto& U = urow[(size_t)i];
        std::sort(U.begin(), U.end());
        for (auto& e : U) { uc.push_back(e.first); us.push_back(e.second); }
        up.push_back((int32_t)uc.size());
    }
    f->block_rows = R;
    return SOGL_OK;
}

// Checks every index of a host task graph against the pool and every dependency
// against the task order, so that a wrong graph is an error here and never a bad
// address or an endless wait on the device.
int validate_graph(Context* c, const sogl_factor_desc* d, const Factor* f, const HostGraph& hg) {
    const int64_t n = f->n_slots;
    const size_t T = f->tasks.size();
    auto bad = [&](const std::string& m) { return set_error(c, SOGL_ERR_ARG, "task graph check: " + m); };
    auto ok_slot = [&](int64_t s) { return s >= 0 && s < n; };
    auto ok_opt = [&](int64_t s) { return s == -1 || ok_slot(s); };
    if (f->zero_slot < 0 || f->zero_slot >= n) return bad("zero slot");
    std::vector<int32_t> writer((size_t)n, -1);
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        if (!ok_slot(k.out0) || !ok_opt(k.out1) || !ok_opt(k.out2) || !ok_opt(k.out3))
            return bad("output slot of task " + std::to_string(t));
        for (int32_t o : {k.out0, k.out1, k.out2, k.out3}) {
            if (o < 0) continue;
            if (writer[(size_t)o] >= 0) return bad("slot " + std::to_string(o) + " written by two tasks");
            writer[(size_t)o] = (int32_t)t;
        }
    }
    auto is_fixed = [&](int32_t s) { return s == f->zero_slot || (s >= 0 && s < (int64_t)d->n_inputs); };
    if (f->level_begin.empty() || f->level_begin.front() != 0 || (size_t)f->level_begin.back() != T)
        return bad("level ranges");
    for (size_t v = 0; v + 1 < f->level_begin.size(); v++)
        for (int32_t t = f->level_begin[v]; t < f->level_begin[v + 1]; t++)
            if (t < 0 || (size_t)t >= T || f->tasks[(size_t)t].level != (int32_t)v) return bad("level order");
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        if (k.dep_begin < 0 || k.dep_count < 0 || (size_t)k.dep_begin + (size_t)k.dep_count > hg.deps.size())
            return bad("dependency range of task " + std::to_string(t));
        std::vector<int32_t> reads;
        switch (k.kind) {
            case TK_GEMM:
                if (k.a < 0 || k.b < 1 || (size_t)k.a + (size_t)k.b > hg.pairs.size())
                    return bad("pair range of task " + std::to_string(t));
                for (int32_t q = k.a; q < k.a + k.b; q++) {
                    const Pair& p = hg.pairs[(size_t)q];
                    reads.push_back(pair_a(p));
                    reads.push_back(pair_b(p));
                }
                if (!ok_opt(k.add)) return bad("added tile of task " + std::to_string(t));
                if (k.add >= 0) reads.push_back(k.add);
                break;
            case TK_SUB:
                if (!ok_opt(k.a) || !ok_opt(k.b)) return bad("operand of task " + std::to_string(t));
                if (k.a >= 0) reads.push_back(k.a);
                if (k.b >= 0) reads.push_back(k.b);
                break;
            case TK_LU: case TK_LINV: case TK_UINV: case TK_LLT: case TK_TRANS:
                reads.push_back(k.a);
                break;
            default:
                return bad("kind of task " + std::to_string(t));
        }
        for (int32_t s2 : reads) {
            if (!ok_slot(s2)) return bad("operand slot of task " + std::to_string(t));
            const int32_t w = writer[(size_t)s2];
            if (w < 0) {
                if (!is_fixed(s2)) return bad("task " + std::to_string(t) + " reads a tile no task writes");
                continue;
            }
            if (f->tasks[(size_t)w].level >= k.level)
                return bad("task " + std::to_string(t) + " is not after the producer of its operand");
            if (k.kind != TK_GEMM) {  // non-GEMM tasks wait for their dependency list only
                bool found = false;
                for (int32_t q = k.dep_begin; q < k.dep_begin + k.dep_count; q++) found |= hg.deps[(size_t)q] == s2;
                if (!found) return bad("task " + std::to_string(t) + " does not wait for an operand");
            }
        }
        for (int32_t q = k.dep_begin; q < k.dep_begin + k.dep_count; q++)
            if (!ok_slot(hg.deps[(size_t)q])) return bad("dependency slot of task " + std::to_string(t));
    }
    // chains and the queue: a static link goes from a GEMM to the lu / llt of its sum, an
    // after-link from a task that publishes to a GEMM reading one of its outputs, both
    // at a later level; every task is handed out by the queue or reached by one link,
    // once.  Every task a linked GEMM waits for is an earlier task of its own chain or
    // belongs to a chain whose head is queued before its own (see build_after_links).
    {
        std::vector<char> seen(T, 0);
        std::vector<int32_t> head(T, -1), pred(T, -1);
        std::vector<int64_t> qpos(T, -1);
        for (size_t i = 0; i < f->queue.size(); i++) {
            const int32_t q = f->queue[i];
            if (q < 0 || (size_t)q >= T || seen[(size_t)q]) return bad("queue entry " + std::to_string(q));
            seen[(size_t)q] = 1;
            qpos[(size_t)q] = (int64_t)i;
            if (i > 0 && f->tasks[(size_t)q].level < f->tasks[(size_t)f->queue[i - 1]].level)
                return bad("queue order at entry " + std::to_string(i));
        }
        for (size_t t = 0; t < T; t++) {
            const DevTask& k = f->tasks[t];
            if ((k.next >= 0) + (k.after >= 0) + (k.pan_l >= 0 || k.pan_u >= 0) > 1)
                return bad("two kinds of link from task " + std::to_string(t));
            if (k.next >= 0) {
                if ((size_t)k.next >= T || k.kind != TK_GEMM) return bad("chain of task " + std::to_string(t));
                const DevTask& u = f->tasks[(size_t)k.next];
                if ((u.kind != TK_LU && u.kind != TK_LLT) || u.a != k.out0 || u.level <= k.level ||
                    seen[(size_t)k.next])
                    return bad("chain of task " + std::to_string(t));
                seen[(size_t)k.next] = 1;
                pred[(size_t)k.next] = (int32_t)t;
            }
            if (k.after >= 0) {
                if ((size_t)k.after >= T) return bad("after-link of task " + std::to_string(t));
                const DevTask& u = f->tasks[(size_t)k.after];
                if (u.kind != TK_GEMM || u.level <= k.level || seen[(size_t)k.after])
                    return bad("after-link of task " + std::to_string(t));
                seen[(size_t)k.after] = 1;
                pred[(size_t)k.after] = (int32_t)t;
            }
            for (int which = 0; which < 2; which++) {  // an lu running one-pair panel GEMMs on its inverses
                const int32_t p = which == 0 ? k.pan_l : k.pan_u;
                if (p < 0) continue;
                if (k.kind != TK_LU || f->precision != SOGL_FP32 || (size_t)p >= T || seen[(size_t)p])
                    return bad("absorbed panel of task " + std::to_string(t));
                const DevTask& u = f->tasks[(size_t)p];
                if (u.kind != TK_GEMM || u.level <= k.level || u.b != 1 || u.add >= 0 || u.next >= 0 || u.after >= 0 ||
                    u.pan_l >= 0 || u.pan_u >= 0 || u.a < 0 || (size_t)u.a >= hg.pairs.size())
                    return bad("absorbed panel of task " + std::to_string(t));
                const Pair& pr = hg.pairs[(size_t)u.a];
                if (pr.a < 0 || (which == 0 ? pair_b(pr) != k.out3 : pair_a(pr) != k.out2))
                    return bad("absorbed panel of task " + std::to_string(t));
                seen[(size_t)p] = 1;
                pred[(size_t)p] = (int32_t)t;
            }
        }
        for (size_t t = 0; t < T; t++)
            if (!seen[t]) return bad("task " + std::to_string(t) + " is neither queued nor chained");
        // heads: links go to later levels, so a task's predecessor has a lower index
        for (size_t t = 0; t < T; t++) {
            if (pred[t] >= 0 && (size_t)pred[t] >= t) return bad("link order at task " + std::to_string(t));
            head[t] = pred[t] < 0 ? (int32_t)t : head[(size_t)pred[t]];
        }
        std::vector<int32_t> producer((size_t)n, -1);
        for (size_t t = 0; t < T; t++)
            for (int32_t o : {f->tasks[t].out0, f->tasks[t].out1, f->tasks[t].out2, f->tasks[t].out3})
                if (o >= 0) producer[(size_t)o] = (int32_t)t;
        for (size_t t = 0; t < T; t++) {
            const DevTask& k = f->tasks[t];
            if (pred[t] < 0 || k.kind != TK_GEMM || f->tasks[(size_t)pred[t]].next == (int32_t)t) continue;
            for (int32_t q = k.dep_begin; q < k.dep_begin + k.dep_count; q++) {
                const int32_t p = producer[(size_t)hg.deps[(size_t)q]];
                if (p < 0) continue;
                bool member = false;
                for (int32_t x = pred[t]; x >= 0 && !member; x = pred[(size_t)x]) member = x == p;
                if (!member && qpos[(size_t)head[(size_t)p]] >= qpos[(size_t)head[t]])
                    return bad("after-linked task " + std::to_string(t) + " waits for a later chain");
            }
        }
    }
    // chain candidates: each check entry names a queued GEMM of at most CHAIN_MAX_PAIRS
    // pairs, an operand position of it and that operand's slot (-1: an output of the
    // checking task), and the candidate's claim tag for the checking task's operands
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        if (k.chk_count < 0 || k.chk_count > CHAIN_MAX_CHECKS || k.chk_begin < 0 ||
            (size_t)k.chk_begin + (size_t)k.chk_count > f->chk.size())
            return bad("chain checks of task " + std::to_string(t));
        for (int32_t q = k.chk_begin; q < k.chk_begin + k.chk_count; q++) {
            const ChainChk& e = f->chk[(size_t)q];
            if (e.task < 0 || (size_t)e.task >= T || e.cand < 0 || e.cand >= CHAIN_MAX_CANDS || e.pos < 0 || e.pos > 4)
                return bad("chain check " + std::to_string(q));
            const DevTask& u = f->tasks[(size_t)e.task];
            if (u.kind != TK_GEMM || u.b < 1 || u.b > CHAIN_MAX_PAIRS || (e.pos < 4 && e.pos >= 2 * u.b) ||
                (e.pos == 4 && u.add < 0))
                return bad("chain candidate " + std::to_string(e.task));
            const Pair& pr = hg.pairs[(size_t)(u.a + (e.pos < 4 ? e.pos / 2 : 0))];
            const int32_t operand = e.pos == 4 ? u.add : ((e.pos & 1) ? pair_b(pr) : pair_a(pr));
            const bool own = operand == k.out0 || operand == k.out1 || operand == k.out2 || operand == k.out3;
            if (e.slot != (own ? -1 : operand)) return bad("chain check slot " + std::to_string(q));
            if (own && e.pos < 4 && ((u.ctag >> (8 * e.pos)) & 0xFF) != e.cand + 1)
                return bad("claim tag of task " + std::to_string(e.task));
        }
    }
    for (const DevTask& k : f->op_tasks) {
        if (!ok_slot(k.out0) || !ok_opt(k.out1) || !ok_opt(k.out2) || !ok_opt(k.out3)) return bad("per-op output");
        if (k.kind == TK_GEMM) {
            if (k.a < 0 || k.b < 1 || (size_t)k.a + (size_t)k.b > hg.op_pairs.size()) return bad("per-op pairs");
            for (int32_t q = k.a; q < k.a + k.b; q++) {
                const Pair& p = hg.op_pairs[(size_t)q];
                if (!ok_slot(pair_a(p)) || !ok_slot(pair_b(p))) return bad("per-op pair slot");
            }
        } else if (!ok_opt(k.a) || !ok_opt(k.b)) {
            return bad("per-op operand");
        }
    }
    for (const auto* v : {&f->l_diag, &f->u_diag, &f->linv_slot, &f->uinv_slot, &f->l_slots, &f->u_slots,
                          &f->input_slots, &hg.ls, &hg.us})
        for (int32_t s2 : *v)
            if (!ok_slot(s2)) return bad("solve slot");
    for (const auto* v : {&f->linv_slot, &f->uinv_slot})
        for (int32_t s2 : *v)
            if (writer[(size_t)s2] < 0) return bad("a diagonal inverse no task writes");
    const int32_t R = f->block_rows;
    if ((int32_t)hg.lp.size() != R + 1 || (int32_t)hg.up.size() != R + 1 || hg.lp.back() != (int32_t)hg.lc.size() ||
        hg.up.back() != (int32_t)hg.uc.size())
        return bad("solve rows");
    for (int32_t cc : hg.lc) if (cc < 0 || cc >= R) return bad("solve column");
    for (int32_t cc : hg.uc) if (cc < 0 || cc >= R) return bad("solve column");
    return SOGL_OK;
}

// Pool slots by liveness (the reference frees a tile after the last stage that reads it,
// BlockPlanner.cpp:597-634, and recycles it through free lists, memutil.cpp:74-113).
// Tasks are visited in their level order; a result slot is, in this order of choice:
//  * the slot of an operand that the task alone reads and nothing needs later (a
//    GEMM's added tile or operand, a sub's operand, an FP64 lu / llt source; an FP32 or
//    bf16 lu may re-read its source for the FP64 refactorisation).  The task has read
//    that operand in full before it stores any result element (GEMM: every operand load
//    precedes the barrier before the last product, and the added tile is read by the
//    lane that stores the same element; sub: the same thread reads then writes each
//    element; lu / llt: the source is copied to LDS first), so no wait is needed;
//  * for a GEMM, a slot freed REUSE_LEVEL_GAP or more levels earlier: the last value in
//    it has no reader left at a later level.  The levels executors run those readers in
//    earlier launches; the persistent executor makes the task wait for their flags
//    before its first result store (Args::war, polled with the task's first window of
//    pairs).  The gap keeps those waits from holding workgroups up: the persistent
//    executor runs far ahead of the slowest chain on wide graphs (C2 FP32, 1 workgroup
//    per CU: 82 ms with a gap of 2 levels, 64 with 1024, 63 without reuse), and costs
//    little memory (C5 FP64 before the two-row blocks: 13.34 GB at 2, 13.39 at 1024);
//  * a new slot.
// The factor tiles, the diagonal inverses, the solve's two-row blocks and the zero tile
// keep their slots; so do the
// inputs, unless they can be scattered again before each run.  The per-op executor
// accumulates each unit op into the result's slot and runs only without reuse.
// Values keep their own ready flags: the numbers used so far become value ids, and every
// data and mask reference becomes a slot.  The executors of a factor with reuse read
// tiles with coherent loads (COH in tile_ops.h): lines of a slot's earlier values may sit
// in some CU's L1 or some XCD's L2.
constexpr int REUSE_LEVEL_GAP = 1024;

// value ids a task reads (deduplicated), from its slot fields before they become slots
std::vector<int32_t> task_operands(const DevTask& k, const HostGraph& hg) {
    std::vector<int32_t> u;
    if (k.kind == TK_GEMM) {
        if (k.add >= 0) u.push_back(k.add);
        for (int32_t q = k.a; q < k.a + k.b; q++) {
            u.push_back(pair_a(hg.pairs[(size_t)q]));
            u.push_back(pair_b(hg.pairs[(size_t)q]));
        }
    } else if (k.kind == TK_SUB) {
        if (k.a >= 0) u.push_back(k.a);
        if (k.b >= 0) u.push_back(k.b);
    } else if (k.a >= 0) {
        u.push_back(k.a);
    }
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    return u;
}

// the value id whose flag a task's publish raises last (a GEMM chained to its lu raises
// none; the lu does, after it)
int32_t done_flag(const Factor* f, int32_t t) {
    const DevTask& k = f->tasks[(size_t)t];
    return k.next >= 0 ? f->tasks[(size_t)k.next].f0 : k.f0;
}

void assign_slots(Factor* f, HostGraph& hg, bool reuse, bool donate_inputs) {
    const int64_t V = f->n_slots;
    const size_t T = f->tasks.size();
    // a panel GEMM absorbed by its lu stores without write-after-read waits: no slot
    // freed by another value's readers (a sole-reader operand's slot is fine: the
    // operand is in LDS before the stores)
    std::vector<char> absorbed_gemm(T, 0);
    for (size_t t = 0; t < T; t++)
        for (int32_t p : {f->tasks[t].pan_l, f->tasks[t].pan_u})
            if (p >= 0) absorbed_gemm[(size_t)p] = 1;
    for (auto* pv : {&hg.pairs, &hg.op_pairs})
        for (Pair& p : *pv) {
            p.fa = pair_a(p);
            p.fb = pair_b(p);
        }
    auto set_flags = [](DevTask& k) {
        k.f0 = k.out0;
        k.f1 = k.out1;
        k.f2 = k.out2;
        k.f3 = k.out3;
        k.fadd = k.kind == TK_GEMM ? k.add : -1;
        k.fsa = k.kind == TK_GEMM ? -1 : k.a;
        k.fsb = k.kind == TK_SUB ? k.b : -1;
        k.war_begin = k.war_count = 0;
    };
    for (DevTask& k : f->tasks) set_flags(k);
    for (DevTask& k : f->op_tasks) set_flags(k);
    f->war.clear();
    f->reuse = reuse;
    f->rescatter = reuse && donate_inputs && f->n_inputs > 0;
    hg.kept.clear();
    if (!reuse) {
        f->n_locs = V;
        return;
    }
    // readers of each value (tasks, in task order) and the values that keep their slot
    std::vector<int64_t> rb((size_t)V + 1, 0);
    std::vector<std::vector<int32_t>> ops(T);
    for (size_t t = 0; t < T; t++) {
        ops[t] = task_operands(f->tasks[t], hg);
        for (int32_t v : ops[t]) rb[(size_t)v + 1]++;
    }
    for (int64_t v = 0; v < V; v++) rb[(size_t)v + 1] += rb[(size_t)v];
    std::vector<int32_t> rl((size_t)rb[(size_t)V]);
    {
        std::vector<int64_t> fill(rb.begin(), rb.end() - 1);
        for (size_t t = 0; t < T; t++)
            for (int32_t v : ops[t]) rl[(size_t)fill[(size_t)v]++] = (int32_t)t;
    }
    auto nread = [&](int32_t v) { return rb[(size_t)v + 1] - rb[(size_t)v]; };
    std::vector<char> keep((size_t)V, 0);
    keep[(size_t)f->zero_slot] = 1;
    for (const auto* pv : {&f->l_slots, &f->u_slots, &f->l_diag, &f->u_diag, &f->linv_slot, &f->uinv_slot, &hg.ls, &hg.us,
                           &hg.pair_x, &hg.pair_y})
        for (int32_t v : *pv)
            if (v >= 0) keep[(size_t)v] = 1;
    if (!f->rescatter)
        for (int32_t v : f->input_slots) keep[(size_t)v] = 1;
    for (int64_t v = 0; v < V; v++)
        if (keep[(size_t)v]) hg.kept.push_back((int32_t)v);
    // inputs and the zero tile where they are, each the first value of its slot
    std::vector<int32_t> loc((size_t)V, -1);
    std::vector<char> first((size_t)V, 0), given((size_t)V, 0);
    std::vector<int64_t> seen((size_t)V, 0);
    int32_t next = 0;
    for (int32_t v : f->input_slots) {
        loc[(size_t)v] = v;
        first[(size_t)v] = 1;
        next = std::max(next, v + 1);
    }
    loc[(size_t)f->zero_slot] = f->zero_slot;
    first[(size_t)f->zero_slot] = 1;
    next = std::max(next, f->zero_slot + 1);
    struct Freed {
        int32_t slot, level, value;
    };
    std::deque<Freed> freed;  // in the order their last readers were visited (by level)
    const bool src_ok = f->precision == SOGL_FP64;
    std::vector<int32_t> war;
    for (size_t t = 0; t < T; t++) {
        DevTask& k = f->tasks[t];
        // a GEMM chained to its lu / llt leaves its sum to that task without waiting for
        // its stores: the sum must be its slot's first value, which the lu reads with
        // plain loads through this CU's own caches (coherent loads could pass the stores,
        // and no cache holds lines of an earlier value of a first value's slot)
        const bool fresh = k.next >= 0;
        const bool donates = !fresh && (k.kind == TK_GEMM || k.kind == TK_SUB ||
                                        ((k.kind == TK_LU || k.kind == TK_LLT) && src_ok));
        std::vector<int32_t> cand;
        if (donates) {
            if (k.kind == TK_GEMM && k.add >= 0) cand.push_back(k.add);  // update chains stay in one slot
            for (int32_t v : ops[t]) cand.push_back(v);
        }
        war.clear();
        for (int32_t o : {k.out0, k.out1, k.out2, k.out3}) {
            if (o < 0 || loc[(size_t)o] >= 0) continue;
            int32_t from = -1;
            for (int32_t v : cand)
                if (!keep[(size_t)v] && nread(v) == 1 && !given[(size_t)v] && loc[(size_t)v] >= 0) {
                    from = v;
                    break;
                }
            if (from >= 0) {
                given[(size_t)from] = 1;
                loc[(size_t)o] = loc[(size_t)from];
            } else if (!fresh && !absorbed_gemm[t] && k.kind == TK_GEMM && !freed.empty() &&
                       freed.front().level <= k.level - REUSE_LEVEL_GAP) {
                const Freed fr = freed.front();
                freed.pop_front();
                loc[(size_t)o] = fr.slot;
                for (int64_t q = rb[(size_t)fr.value]; q < rb[(size_t)fr.value + 1]; q++)
                    war.push_back(done_flag(f, rl[(size_t)q]));
            } else {
                loc[(size_t)o] = next++;
                first[(size_t)o] = 1;
            }
        }
        std::sort(war.begin(), war.end());
        war.erase(std::unique(war.begin(), war.end()), war.end());
        k.war_begin = (int32_t)f->war.size();
        k.war_count = (int32_t)war.size();
        f->war.insert(f->war.end(), war.begin(), war.end());
        // operands whose last reader this is free their slot (unless kept or handed over)
        for (int32_t v : ops[t])
            if (++seen[(size_t)v] == nread(v) && !keep[(size_t)v] && !given[(size_t)v])
                freed.push_back({loc[(size_t)v], k.level, v});
    }
    for (int64_t v = 0; v < V; v++)
        if (loc[(size_t)v] < 0) loc[(size_t)v] = next++, first[(size_t)v] = 1;  // made and read by no task
    f->n_locs = next;
    // every data and mask reference to its slot
    auto L = [&](int32_t v) { return v < 0 ? v : loc[(size_t)v]; };
    for (size_t t = 0; t < T; t++) {
        DevTask& k = f->tasks[t];
        k.out0 = L(k.out0);
        k.out1 = L(k.out1);
        k.out2 = L(k.out2);
        k.out3 = L(k.out3);
        k.add = L(k.add);
        if (k.kind != TK_GEMM) {
            k.a = L(k.a);
            k.b = L(k.b);
        }
    }
    for (Pair& p : hg.pairs) {
        p.a = p.a < 0 ? -L(p.fa) - 1 : L(p.fa);
        p.b = p.b < 0 ? -L(p.fb) - 1 : L(p.fb);
    }
    for (auto* pv : {&f->l_slots, &f->u_slots, &f->l_diag, &f->u_diag, &f->linv_slot, &f->uinv_slot, &hg.ls, &hg.us,
                     &hg.pair_x, &hg.pair_y})
        for (int32_t& v : *pv) v = L(v);
}

// The slot assignment's invariants, checked on the host: consecutive values of one slot
// either pass from a value to the task that alone reads it (and read it whole), or the
// later value's task is a GEMM at a later level than every reader of the earlier value
// and waits for all their flags; the kept values end their slot's sequence; a GEMM sum
// that goes to a chained lu / llt is its slot's first value (the lu reads it with plain
// loads, see run_task).
int validate_slots(Context* c, const Factor* f, const HostGraph& hg) {
    auto bad = [&](const std::string& m) { return set_error(c, SOGL_ERR_ARG, "slot reuse check: " + m); };
    if (!f->reuse) return SOGL_OK;
    const int64_t V = f->n_slots;
    const size_t T = f->tasks.size();
    std::vector<int64_t> qpos(T, 0);  // queue position of each queued task (chain heads)
    for (size_t i = 0; i < f->queue.size(); i++) qpos[(size_t)f->queue[i]] = (int64_t)i;
    std::vector<int32_t> loc((size_t)V, -1), prod((size_t)V, -1);
    auto put = [&](int32_t v, int32_t l) {
        if (v < 0) return true;
        if (v >= V || l < 0 || l >= f->n_locs) return false;
        if (loc[(size_t)v] >= 0 && loc[(size_t)v] != l) return false;
        loc[(size_t)v] = l;
        return true;
    };
    // each task's reads (value ids, deduplicated), as a CSR over tasks
    std::vector<int64_t> rb(T + 1, 0);
    std::vector<int32_t> rv;
    rv.reserve(hg.pairs.size() * 2 + T);
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        if (!put(k.f0, k.out0) || !put(k.f1, k.out1) || !put(k.f2, k.out2) || !put(k.f3, k.out3))
            return bad("outputs of task " + std::to_string(t));
        for (int32_t v : {k.f0, k.f1, k.f2, k.f3})
            if (v >= 0) prod[(size_t)v] = (int32_t)t;
        const size_t b0 = rv.size();
        if (k.kind == TK_GEMM) {
            if (!put(k.fadd, k.add)) return bad("added tile of task " + std::to_string(t));
            if (k.fadd >= 0) rv.push_back(k.fadd);
            for (int32_t q = k.a; q < k.a + k.b; q++) {
                const Pair& p = hg.pairs[(size_t)q];
                if (!put(p.fa, pair_a(p)) || !put(p.fb, pair_b(p)))
                    return bad("pair of task " + std::to_string(t));
                rv.push_back(p.fa);
                rv.push_back(p.fb);
            }
        } else {
            if (!put(k.fsa, k.a) || (k.kind == TK_SUB && !put(k.fsb, k.b)))
                return bad("operands of task " + std::to_string(t));
            if (k.fsa >= 0) rv.push_back(k.fsa);
            if (k.fsb >= 0) rv.push_back(k.fsb);
        }
        std::sort(rv.begin() + (int64_t)b0, rv.end());
        rv.erase(std::unique(rv.begin() + (int64_t)b0, rv.end()), rv.end());
        rb[t + 1] = (int64_t)rv.size();
        if (k.war_begin < 0 || k.war_count < 0 || (size_t)k.war_begin + (size_t)k.war_count > f->war.size())
            return bad("write-after-read range of task " + std::to_string(t));
        for (int32_t q = k.war_begin; q < k.war_begin + k.war_count; q++)
            if (f->war[(size_t)q] < 0 || f->war[(size_t)q] >= V) return bad("write-after-read flag of task " + std::to_string(t));
    }
    // readers of each value (CSR over values, tasks ascending)
    std::vector<int64_t> qb((size_t)V + 1, 0);
    for (int32_t v : rv) qb[(size_t)v + 1]++;
    for (int64_t v = 0; v < V; v++) qb[(size_t)v + 1] += qb[(size_t)v];
    std::vector<int32_t> ql(rv.size());
    {
        std::vector<int64_t> fill(qb.begin(), qb.end() - 1);
        for (size_t t = 0; t < T; t++)
            for (int64_t q = rb[t]; q < rb[t + 1]; q++) ql[(size_t)fill[(size_t)rv[(size_t)q]]++] = (int32_t)t;
    }
    // occupants of each slot in visiting order (inputs and the zero tile first), CSR
    const int64_t NL = f->n_locs;
    std::vector<int64_t> ob((size_t)NL + 1, 0);
    for (int64_t v = 0; v < V; v++)
        if (loc[(size_t)v] >= 0) ob[(size_t)loc[(size_t)v] + 1]++;
    for (int64_t l = 0; l < NL; l++) ob[(size_t)l + 1] += ob[(size_t)l];
    std::vector<std::pair<int32_t, int32_t>> occ((size_t)ob[(size_t)NL]);
    {
        std::vector<int64_t> fill(ob.begin(), ob.end() - 1);
        for (int64_t v = 0; v < V; v++)
            if (loc[(size_t)v] >= 0) occ[(size_t)fill[(size_t)loc[(size_t)v]]++] = {prod[(size_t)v], (int32_t)v};
    }
    std::vector<char> first((size_t)V, 1);
    std::vector<int32_t> wf;
    for (int64_t l = 0; l < NL; l++) {
        auto* o = occ.data() + ob[(size_t)l];
        const int64_t no = ob[(size_t)l + 1] - ob[(size_t)l];
        std::sort(o, o + no);
        for (int64_t i = 0; i + 1 < no; i++) {
            const int32_t v = o[i].second, w = o[i + 1].second, pw = o[i + 1].first;
            first[(size_t)w] = 0;
            if (v == f->zero_slot) return bad("the zero tile's slot is taken over");
            if (o[i].first < 0 && !f->rescatter) return bad("an input slot is taken over");
            if (pw < 0) return bad("slot " + std::to_string(l) + " holds two values that no task makes");
            const DevTask& kw = f->tasks[(size_t)pw];
            const int64_t r0 = qb[(size_t)v], r1 = qb[(size_t)v + 1];
            if (r1 - r0 == 1 && ql[(size_t)r0] == pw) {  // handed over by its only reader
                if (kw.kind == TK_GEMM || kw.kind == TK_SUB ||
                    ((kw.kind == TK_LU || kw.kind == TK_LLT) && f->precision == SOGL_FP64))
                    continue;
                return bad("task " + std::to_string(pw) + " may read its source again after storing over it");
            }
            if (kw.kind != TK_GEMM)
                return bad("slot " + std::to_string(l) + " passes to a non-GEMM task while another reads it");
            wf.assign(f->war.begin() + kw.war_begin, f->war.begin() + kw.war_begin + kw.war_count);
            std::sort(wf.begin(), wf.end());
            for (int64_t q = r0; q < r1; q++) {
                const int32_t r = ql[(size_t)q];
                if (f->tasks[(size_t)r].level >= kw.level)
                    return bad("slot " + std::to_string(l) + " is reused at the level of a reader");
                if (!std::binary_search(wf.begin(), wf.end(), done_flag(f, r)))
                    return bad("task " + std::to_string(pw) + " does not wait for a reader of its slot");
                // a linked task waits only for chains with earlier heads in the queue
                if (!f->chain_head.empty() && f->chain_head[(size_t)pw] != pw &&
                    qpos[(size_t)f->chain_head[(size_t)r]] >= qpos[(size_t)f->chain_head[(size_t)pw]])
                    return bad("linked task " + std::to_string(pw) + " waits for a reader in a later chain");
            }
        }
    }
    for (int32_t v : hg.kept) {
        const int32_t l = loc[(size_t)v];
        if (l >= 0 && occ[(size_t)ob[(size_t)l + 1] - 1].second != v)
            return bad("kept value " + std::to_string(v) + " is overwritten");
    }
    for (size_t t = 0; t < T; t++) {
        const DevTask& k = f->tasks[t];
        if (k.next >= 0 && !first[(size_t)k.f0])
            return bad("the sum of GEMM task " + std::to_string(t) + " goes to its chained task through a reused slot");
    }
    return SOGL_OK;
}

// Device arrays of a checked graph: pool, tasks, pairs, dependencies, flags, solve
// structure, and the input tiles copied in.
int upload_graph(Context* c, const sogl_factor_desc* d, const InputCells* cells, Factor* f,
                 const HostGraph& hg) {
    const std::vector<Pair>& pairs = hg.pairs;
    const std::vector<Pair>& op_pairs = hg.op_pairs;
    const std::vector<int32_t>& deps = hg.deps;
    const size_t T = f->tasks.size();
    // device arrays
    int rc;
    const size_t tile_bytes = (size_t)TE * f->elem;
    if ((rc = dev_alloc(c, f->pool, (size_t)f->n_locs * tile_bytes, "tile pool")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->tasks_d, std::max<size_t>(T, 1) * sizeof(DevTask), "tasks")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->pairs_d, std::max<size_t>(pairs.size(), 1) * sizeof(Pair), "pairs")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->deps_d, std::max<size_t>(deps.size(), 1) * 4, "deps")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->ready_d, (size_t)f->n_slots * 8, "ready flags")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->mask_d, (size_t)f->n_locs * 4, "block masks")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->ctrl_d, CTRL_WORDS * 4, "control")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->uinv_chk_d, (1 + 4 * UINV_KEEP) * 8, "upper inverse check")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->queue_d, std::max<size_t>(f->queue.size(), 1) * 4, "task queue")) != SOGL_OK) return rc;
    if ((rc = dev_alloc(c, f->chk_d, std::max<size_t>(f->chk.size(), 1) * sizeof(ChainChk), "chain checks")) != SOGL_OK)
        return rc;
    if ((rc = dev_alloc(c, f->war_d, std::max<size_t>(f->war.size(), 1) * 4, "write-after-read flags")) != SOGL_OK)
        return rc;
    if (!f->war.empty()) HIP_TRY(c, hipMemcpy(f->war_d.p, f->war.data(), f->war.size() * 4, hipMemcpyHostToDevice));
    auto up_i32 = [&](DevBuf& b, const std::vector<int32_t>& v, const char* what) -> int {
        int r2 = dev_alloc(c, b, std::max<size_t>(v.size(), 1) * 4, what);
        if (r2 != SOGL_OK) return r2;
        if (!v.empty()) HIP_TRY(c, hipMemcpy(b.p, v.data(), v.size() * 4, hipMemcpyHostToDevice));
        return SOGL_OK;
    };
    hipStream_t st = c->stream;
    if (T) HIP_TRY(c, hipMemcpy(f->tasks_d.p, f->tasks.data(), T * sizeof(DevTask), hipMemcpyHostToDevice));
    if (!pairs.empty()) HIP_TRY(c, hipMemcpy(f->pairs_d.p, pairs.data(), pairs.size() * sizeof(Pair), hipMemcpyHostToDevice));
    if (!deps.empty()) HIP_TRY(c, hipMemcpy(f->deps_d.p, deps.data(), deps.size() * 4, hipMemcpyHostToDevice));
    // the per-op executor's tables (one entry per unit op: 680 MB for a 3D 64^3 plan) go to
    // the device on its first run only; a factor with slot reuse cannot use them
    if (f->reuse) {
        std::vector<DevTask>().swap(f->op_tasks);
    } else {
        f->op_pairs = op_pairs;
    }
    if (!f->queue.empty()) HIP_TRY(c, hipMemcpy(f->queue_d.p, f->queue.data(), f->queue.size() * 4, hipMemcpyHostToDevice));
    if (!f->chk.empty())
        HIP_TRY(c, hipMemcpy(f->chk_d.p, f->chk.data(), f->chk.size() * sizeof(ChainChk), hipMemcpyHostToDevice));
    if ((rc = up_i32(f->l_rowptr, hg.lp, "L rows")) || (rc = up_i32(f->l_col, hg.lc, "L cols")) ||
        (rc = up_i32(f->l_slot, hg.ls, "L slots")) || (rc = up_i32(f->u_rowptr, hg.up, "U rows")) ||
        (rc = up_i32(f->u_col, hg.uc, "U cols")) || (rc = up_i32(f->u_slot, hg.us, "U slots")) ||
        (rc = up_i32(f->l_diag_d, f->l_diag, "L diag")) || (rc = up_i32(f->u_diag_d, f->u_diag, "U diag")) ||
        (rc = up_i32(f->linv_d, f->linv_slot, "L inv")) || (rc = up_i32(f->uinv_d, f->uinv_slot, "U inv")) ||
        (rc = up_i32(f->pairx_d, hg.pair_x, "row pair L blocks")) || (rc = up_i32(f->pairy_d, hg.pair_y, "row pair U blocks")) ||
        (rc = up_i32(f->inputs_d, f->input_slots, "input slots")))
        return rc;
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(256), dim3(256), 0, st, (unsigned long long*)f->ready_d.p,
                       (size_t)f->n_slots, 0ull);
    // input tiles (slots 0..n_inputs-1) and the zero tile never change: always ready
    // (the inputs' flags are set with their masks after the upload, below)
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, st,
                       (unsigned long long*)f->ready_d.p + f->zero_slot, (size_t)1, 0x7fffffffull << 32);
    // the zero tile
    if (f->elem == 8)
        hipLaunchKernelGGL(k_fill<double>, dim3(16), dim3(256), 0, st,
                           (double*)f->pool.p + (size_t)f->zero_slot * TE, (size_t)TE, 0.0);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(16), dim3(256), 0, st,
                           (float*)f->pool.p + (size_t)f->zero_slot * TE, (size_t)TE, 0.0f);
    HIP_TRY(c, hipGetLastError());
    // input tiles: contiguous slots 0..n_inputs-1
    const size_t nin = (size_t)d->n_inputs * TE;
    if (nin && cells) {
        // zero the input slots, then scatter the cells (FP64 values rounded on the device)
        const int64_t ntl = d->n_inputs;
        const int64_t ncell = cells->begin[ntl];
        for (int64_t t = 0; t < ntl; t++)
            if (cells->begin[t + 1] < cells->begin[t]) return set_error(c, SOGL_ERR_ARG, "bad input cells");
        if (f->elem == 8)
            hipLaunchKernelGGL(k_fill<double>, dim3(4096), dim3(256), 0, st, (double*)f->pool.p, nin, 0.0);
        else
            hipLaunchKernelGGL(k_fill<float>, dim3(4096), dim3(256), 0, st, (float*)f->pool.p, nin, 0.0f);
        HIP_TRY(c, hipGetLastError());
        DevBuf cb, cp, cv;
        auto done = [&](int code) { dev_free(cb); dev_free(cp); dev_free(cv); return code; };
        if ((rc = dev_alloc(c, cb, (size_t)(ntl + 1) * 8, "cell index")) ||
            (rc = dev_alloc(c, cp, std::max<int64_t>(ncell, 1) * 2, "cell positions")) ||
            (rc = dev_alloc(c, cv, std::max<int64_t>(ncell, 1) * 8, "cell values")))
            return done(rc);
        hipError_t e = hipMemcpy(cb.p, cells->begin, (size_t)(ntl + 1) * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess && ncell) e = hipMemcpy(cp.p, cells->pos, (size_t)ncell * 2, hipMemcpyHostToDevice);
        if (e == hipSuccess && ncell) e = hipMemcpy(cv.p, cells->val, (size_t)ncell * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) return done(check_hip(c, e, "input cell upload"));
        const unsigned grid = (unsigned)std::min<int64_t>(ntl, 65536);
        if (f->elem == 8)
            hipLaunchKernelGGL(k_scatter_cells<double>, dim3(grid), dim3(256), 0, st, (double*)f->pool.p,
                               (const int64_t*)cb.p, (const uint16_t*)cp.p, (const double*)cv.p, ntl);
        else
            hipLaunchKernelGGL(k_scatter_cells<float>, dim3(grid), dim3(256), 0, st, (float*)f->pool.p,
                               (const int64_t*)cb.p, (const uint16_t*)cp.p, (const double*)cv.p, ntl);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return done(check_hip(c, e, "input cell scatter"));
        if (f->rescatter) {  // kept for the scatter before every run
            std::swap(f->cell_begin, cb);
            std::swap(f->cell_pos, cp);
            std::swap(f->cell_val, cv);
        }
        done(SOGL_OK);
    } else if (nin) {
        if (f->elem == 8) {
            HIP_TRY(c, hipMemcpy(f->pool.p, d->input_tiles, nin * 8, hipMemcpyHostToDevice));
        } else {
            // stage FP64 in chunks and round to FP32 on the device
            const size_t chunk = (size_t)1 << 24;
            DevBuf tmp;
            if ((rc = dev_alloc(c, tmp, std::min(chunk, nin) * 8, "staging")) != SOGL_OK) return rc;
            for (size_t o = 0; o < nin; o += chunk) {
                size_t m = std::min(chunk, nin - o);
                hipError_t e = hipMemcpy(tmp.p, d->input_tiles + o, m * 8, hipMemcpyHostToDevice);
                if (e != hipSuccess) { dev_free(tmp); return check_hip(c, e, "input upload"); }
                hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, st, (const double*)tmp.p,
                                   (float*)f->pool.p + o, m);
                e = hipStreamSynchronize(st);
                if (e != hipSuccess) { dev_free(tmp); return check_hip(c, e, "input convert"); }
            }
            dev_free(tmp);
        }
    }
    // block masks: the inputs' from their values, every other slot (the zero tile
    // included) empty until its task writes it
    hipLaunchKernelGGL(k_fill<int>, dim3(256), dim3(256), 0, st, (int*)f->mask_d.p, (size_t)f->n_locs, 0);
    if (d->n_inputs > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>(d->n_inputs, 65536);
        if (f->elem == 8)
            hipLaunchKernelGGL(k_tile_masks<double>, dim3(grid), dim3(NT), 0, st, (const double*)f->pool.p,
                               (uint32_t*)f->mask_d.p, (unsigned long long*)f->ready_d.p, (int64_t)d->n_inputs);
        else
            hipLaunchKernelGGL(k_tile_masks<float>, dim3(grid), dim3(NT), 0, st, (const float*)f->pool.p,
                               (uint32_t*)f->mask_d.p, (unsigned long long*)f->ready_d.p, (int64_t)d->n_inputs);
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(st));
    return SOGL_OK;
}

int build_graph(Context* c, const sogl_factor_desc* d, const InputCells* cells, Factor* f) {
    HostGraph hg;
    f->keep_ops = c->pool_reuse != SOGL_REUSE_ON;  // a factor with reuse cannot run per op
    f->after_links = c->after_links;
    f->absorb_panels = c->absorb_panels;
    int rc = build_host_graph(c, d, f, hg);
    if (rc == SOGL_OK) rc = validate_graph(c, d, f, hg);
    if (rc == SOGL_OK) {
        f->n_inputs = d->n_inputs;
        // auto: reuse when the values would fill more than a quarter of the device's HBM
        // (reuse costs 2-7% of factor time in the coherent loads; C5 FP64, 35 GB of values,
        // runs without it on a 288 GB device)
        const double value_bytes = (double)f->n_slots * TE * f->elem;
        const bool reuse = c->pool_reuse == SOGL_REUSE_ON ||
                           (c->pool_reuse == SOGL_REUSE_AUTO && value_bytes > 0.25 * (double)c->total_mem);
        assign_slots(f, hg, reuse, cells != nullptr);
        rc = validate_slots(c, f, hg);
    }
    if (rc == SOGL_OK) rc = upload_graph(c, d, cells, f, hg);
    return rc;
}

void free_factor(Factor* f) {
    if (f->graph_exec) (void)hipGraphExecDestroy(f->graph_exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    for (DevBuf* b : {&f->pool, &f->mask_d, &f->queue_d, &f->war_d, &f->uinv_chk_d, &f->pairx_d, &f->pairy_d, &f->cell_begin, &f->cell_pos, &f->cell_val, &f->chk_d, &f->tasks_d, &f->pairs_d, &f->deps_d, &f->ready_d, &f->ctrl_d,
                      &f->op_tasks_d, &f->op_pairs_d, &f->inputs_d, &f->l_rowptr, &f->l_col,
                      &f->l_slot, &f->u_rowptr, &f->u_col, &f->u_slot, &f->l_diag_d, &f->u_diag_d,
                      &f->linv_d, &f->uinv_d, &f->solve_ready, &f->solve_ctrl, &f->vec_b, &f->vec_x,
                      &f->vec_y, &f->csr_ptr, &f->csr_col, &f->csr_val, &f->vec_r, &f->vec_d, &f->red, &f->perm_d, &f->vec_io, &f->trace_d})
        dev_free(*b);
    delete f;
}

template <typename T, bool BF16, bool LLT, bool MASKS, bool COH>
int launch_factor(Context* c, Factor* f, int mode, int chunk, int wg_per_cu) {
    hipStream_t st = c->stream;
    Args a{};
    a.chunk = std::max(1, chunk);
    a.pool = f->pool.p;
    a.ready = (unsigned long long*)f->ready_d.p;
    a.ctrl = (int*)f->ctrl_d.p;
    a.mask = (uint32_t*)f->mask_d.p;
    a.epoch = f->epoch;
    a.growth_limit = f->growth_limit;
    a.deps = (const int32_t*)f->deps_d.p;
    a.trace = (mode == SOGL_EXEC_PERSISTENT && f->trace_on) ? (long long*)f->trace_d.p : nullptr;
    a.chk = (const ChainChk*)f->chk_d.p;
    a.war = (const int32_t*)f->war_d.p;
    a.chain = (mode == SOGL_EXEC_PERSISTENT && !f->chk.empty()) ? f->chain_claims : 0;
    auto launch_levels = [&]() {
        a.tasks = (const DevTask*)f->tasks_d.p;
        a.pairs = (const Pair*)f->pairs_d.p;
        a.ntasks = (int32_t)f->tasks.size();
        for (size_t v = 0; v + 1 < f->level_begin.size(); v++) {
            int cnt = f->level_begin[v + 1] - f->level_begin[v];
            if (cnt <= 0) continue;
            a.task_offset = f->level_begin[v];
            hipLaunchKernelGGL((k_factor_batch<T, BF16, false, LLT, MASKS, COH>), dim3(cnt), dim3(NT), 0, st, a);
        }
    };
    if (mode == SOGL_EXEC_GRAPH && !f->graph_exec) {
        // the level schedule, captured once per factor: the control reset is a kernel
        // (no memset nodes) and every launch's arguments are fixed for the factor
        HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, a.ctrl);
        launch_levels();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (e != hipSuccess) return check_hip(c, e, "level schedule capture");
        f->graph = g;
        HIP_TRY(c, hipGraphInstantiate(&f->graph_exec, g, nullptr, nullptr, 0));
    }
    if (mode == SOGL_EXEC_PER_OP && f->reuse)
        return set_error(c, SOGL_ERR_UNSUPPORTED,
                         "the per-op executor needs a factor loaded with pool slot reuse off "
                         "(sogl_ctx_set_pool_reuse)");
    if (f->rescatter) {
        // input slots were taken over by the last run: scatter the inputs again, with
        // their masks and always-ready flags (the executors' own time excludes this)
        const int64_t ntl = f->n_inputs;
        const unsigned grid = (unsigned)std::min<int64_t>(ntl, 65536);
        hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, st, (T*)f->pool.p, (size_t)ntl * TE, T(0));
        hipLaunchKernelGGL(k_scatter_cells<T>, dim3(grid), dim3(256), 0, st, (T*)f->pool.p,
                           (const int64_t*)f->cell_begin.p, (const uint16_t*)f->cell_pos.p,
                           (const double*)f->cell_val.p, ntl);
        hipLaunchKernelGGL(k_tile_masks<T>, dim3(grid), dim3(NT), 0, st, (const T*)f->pool.p,
                           (uint32_t*)f->mask_d.p, (unsigned long long*)f->ready_d.p, ntl);
    }
    if (mode != SOGL_EXEC_GRAPH) hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, a.ctrl);
    HIP_TRY(c, hipEventRecord(c->ex0, st));  // executor alone, for kernel-level rates
    if (mode == SOGL_EXEC_PERSISTENT) {
        a.tasks = (const DevTask*)f->tasks_d.p;
        a.pairs = (const Pair*)f->pairs_d.p;
        a.ntasks = (int32_t)f->tasks.size();
        a.queue = (const int32_t*)f->queue_d.p;
        a.nqueue = (int32_t)f->queue.size();
        // one workgroup per CU: the kernel built for that occupancy
        const bool one = wg_per_cu == 1;
        auto kern = one ? k_factor_persistent<T, BF16, LLT, MASKS, true, COH>
                        : k_factor_persistent<T, BF16, LLT, MASKS, false, COH>;
        int per_cu = 0;
        HIP_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, 0));
        if (per_cu < 1) per_cu = 1;
        if (wg_per_cu > 0) per_cu = std::min(per_cu, wg_per_cu);  // never more than fit: all must be resident
        int grid = std::max(1, std::min<int>(a.nqueue, c->num_cu * per_cu));
        if (a.nqueue > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), 0, st, a);
    } else if (mode == SOGL_EXEC_LEVELS) {
        launch_levels();
    } else if (mode == SOGL_EXEC_GRAPH) {
        HIP_TRY(c, hipGraphLaunch(f->graph_exec, st));
    } else if (mode == SOGL_EXEC_PER_OP) {
        if (BF16) return set_error(c, SOGL_ERR_UNSUPPORTED, "per-op executor has no bf16 variant");
        // results start at zero; every op then runs alone, C += A*B reading C
        const size_t first_result = f->input_slots.size();  // the zero slot, then results
        const size_t nres = (size_t)f->n_slots - first_result;
        hipLaunchKernelGGL(k_fill<T>, dim3(4096), dim3(256), 0, st,
                           (T*)f->pool.p + first_result * TE, nres * TE, T(0));
        hipLaunchKernelGGL(k_fill<int>, dim3(256), dim3(256), 0, st, (int*)f->mask_d.p + first_result, nres, 0);
        if (!f->op_tasks_d.p) {
            int rc;
            if ((rc = dev_alloc(c, f->op_tasks_d, std::max<size_t>(f->op_tasks.size(), 1) * sizeof(DevTask), "op tasks")) ||
                (rc = dev_alloc(c, f->op_pairs_d, std::max<size_t>(f->op_pairs.size(), 1) * sizeof(Pair), "op pairs")))
                return rc;
            if (!f->op_tasks.empty())
                HIP_TRY(c, hipMemcpy(f->op_tasks_d.p, f->op_tasks.data(), f->op_tasks.size() * sizeof(DevTask),
                                     hipMemcpyHostToDevice));
            if (!f->op_pairs.empty())
                HIP_TRY(c, hipMemcpy(f->op_pairs_d.p, f->op_pairs.data(), f->op_pairs.size() * sizeof(Pair),
                                     hipMemcpyHostToDevice));
        }
        a.tasks = (const DevTask*)f->op_tasks_d.p;
        a.pairs = (const Pair*)f->op_pairs_d.p;
        a.ntasks = (int32_t)f->op_tasks.size();
        for (int32_t k = 0; k < a.ntasks; k++) {
            a.task_offset = k;
            if constexpr (!COH)
                hipLaunchKernelGGL((k_factor_batch<T, false, true, LLT, MASKS, false>), dim3(1), dim3(NT), 0, st, a);
        }
    } else {
        return set_error(c, SOGL_ERR_ARG, "unknown executor mode");
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(c->ex1, st));
    HIP_TRY(c, hipGetLastError());
    return SOGL_OK;
}

}  // namespace

}  // namespace sogl_dev

using namespace sogl_dev;

// ------------------------------------------------------------------ C ABI
extern "C" int sogl_ctx_create(int32_t device, sogl_ctx** out) {
    if (!out) return set_error(nullptr, SOGL_ERR_ARG, "null output");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return set_error(nullptr, SOGL_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= n) return set_error(nullptr, SOGL_ERR_ARG, "device index out of range");
    sogl_ctx* x = new sogl_ctx();
    Context* c = &x->c;
    c->device = device;
    int rc;
    if ((rc = check_hip(c, hipSetDevice(device), "hipSetDevice")) ||
        (rc = check_hip(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream")) ||
        (rc = check_hip(c, hipEventCreate(&c->ev0), "event")) ||
        (rc = check_hip(c, hipEventCreate(&c->ev1), "event")) ||
        (rc = check_hip(c, hipEventCreate(&c->ex0), "event")) ||
        (rc = check_hip(c, hipEventCreate(&c->ex1), "event")) ||
        (rc = check_hip(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device), "attr")) ||
        (rc = check_hip(c, hipMemGetInfo(&c->free_mem, &c->total_mem), "memory info"))) {
        delete x;
        return rc;
    }
    noalloc_scratch_init(device);
    *out = x;
    return SOGL_OK;
}

extern "C" void sogl_ctx_destroy(sogl_ctx* x) {
    if (!x) return;
    Context* c = &x->c;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->factors) free_factor(kv.second);
    c->factors.clear();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ex0) (void)hipEventDestroy(c->ex0);
    if (c->ex1) (void)hipEventDestroy(c->ex1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete x;
}

extern "C" const char* sogl_ctx_error(const sogl_ctx* x) { return x ? x->c.err.c_str() : ""; }

extern "C" int sogl_ctx_mem_info(sogl_ctx* x, size_t* free_bytes, size_t* total_bytes) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    HIP_TRY(c, hipSetDevice(c->device));
    size_t f = 0, t = 0;
    HIP_TRY(c, hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return SOGL_OK;
}

extern "C" int sogl_ctx_sync(sogl_ctx* x) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    HIP_TRY(&x->c, hipStreamSynchronize(x->c.stream));
    return SOGL_OK;
}

namespace {
// Loads a factor whose input tiles are dense (desc) or cells (plan storage).
int load_factor(sogl_ctx* x, const sogl_factor_desc* d, const InputCells* cells, int32_t* factor_id) {
    if (!x || !d || !factor_id) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    if (d->n_ids <= 0 || d->n_ops < 0 || d->n_inputs < 0 || d->block_rows <= 0 ||
        (d->n_ops && !d->ops) || (d->n_inputs && (!d->input_ids || (!d->input_tiles && !cells))))
        return set_error(c, SOGL_ERR_ARG, "bad factor description");
    if (d->precision != SOGL_FP64 && d->precision != SOGL_FP32 && d->precision != SOGL_BF16)
        return set_error(c, SOGL_ERR_ARG, "unknown precision");
    HIP_TRY(c, hipSetDevice(c->device));
    Factor* f = new Factor();
    f->precision = d->precision;
    f->elem = d->precision == SOGL_FP64 ? 8 : 4;
    int rc = build_graph(c, d, cells, f);
    if (rc != SOGL_OK) {
        free_factor(f);
        return rc;
    }
    f->id = c->next_id++;
    c->factors[f->id] = f;
    *factor_id = f->id;
    return SOGL_OK;
}
}  // namespace

extern "C" int sogl_factor_load(sogl_ctx* x, const sogl_factor_desc* d, int32_t* factor_id) {
    return load_factor(x, d, nullptr, factor_id);
}

namespace {
// cells: the inputs come as the plan's cells (they can be scattered again before each
// run, so their slots may be taken over)
int graph_check(const sogl_factor_desc* d, bool cells, int64_t* info) {
    if (!d || d->n_ids <= 0 || d->n_ops < 0 || d->n_inputs < 0 || d->block_rows <= 0 ||
        (d->n_ops && !d->ops) || (d->n_inputs && !d->input_ids))
        return set_error(nullptr, SOGL_ERR_ARG, "bad factor description");
    Factor f;
    f.precision = d->precision;
    f.elem = d->precision == SOGL_FP64 ? 8 : 4;
    f.n_inputs = d->n_inputs;
    f.after_links = true;  // checked with the after-links and absorbed panels (head order included)
    f.absorb_panels = true;
    HostGraph hg;
    int rc = build_host_graph(nullptr, d, &f, hg);
    if (rc == SOGL_OK) rc = validate_graph(nullptr, d, &f, hg);
    const int64_t values = f.n_slots;
    if (rc == SOGL_OK) {
        assign_slots(&f, hg, true, cells);
        rc = validate_slots(nullptr, &f, hg);
    }
    if (rc == SOGL_OK && info) {
        info[0] = (int64_t)f.tasks.size();
        info[1] = values;
        info[2] = (int64_t)f.level_begin.size() - 1;
        info[3] = (int64_t)hg.pairs.size();
        info[4] = (int64_t)hg.deps.size();
        info[5] = f.n_folded;
        info[6] = f.n_inv_aliased;
        info[7] = f.n_chained;
        info[8] = f.n_locs;
        info[9] = f.n_pair_tasks;
        info[10] = f.n_after;
        info[11] = f.n_absorbed;
    }
    return rc;
}
}  // namespace

extern "C" int sogl_graph_check(const sogl_factor_desc* d, int64_t* info) { return graph_check(d, false, info); }

extern "C" int sogl_graph_check_plan(const sogl_plan* plan, int32_t precision, int64_t* info) {
    sogl_factor_desc d;
    if (!sogl::plan_factor_desc(plan, precision, &d)) return set_error(nullptr, SOGL_ERR_ARG, "null plan");
    return graph_check(&d, true, info);
}

extern "C" int sogl_ctx_set_task_links(sogl_ctx* x, int32_t links) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    if (links & ~(SOGL_LINK_AFTER | SOGL_LINK_PANEL)) return set_error(&x->c, SOGL_ERR_ARG, "unknown task links");
    std::lock_guard<std::mutex> g(x->c.mu);
    x->c.after_links = (links & SOGL_LINK_AFTER) != 0;
    x->c.absorb_panels = (links & SOGL_LINK_PANEL) != 0;
    return SOGL_OK;
}

extern "C" int sogl_ctx_set_pool_reuse(sogl_ctx* x, int32_t mode) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    if (mode != SOGL_REUSE_OFF && mode != SOGL_REUSE_ON && mode != SOGL_REUSE_AUTO)
        return set_error(&x->c, SOGL_ERR_ARG, "unknown pool reuse mode");
    std::lock_guard<std::mutex> g(x->c.mu);
    x->c.pool_reuse = mode;
    return SOGL_OK;
}

extern "C" int sogl_factor_load_plan(sogl_ctx* x, const sogl_plan* plan, int32_t precision,
                                     int32_t* factor_id) {
    sogl_factor_desc d;
    InputCells cells;
    if (!sogl::plan_factor_desc(plan, precision, &d) ||
        !sogl::plan_input_cells(plan, &cells.begin, &cells.pos, &cells.val))
        return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "null plan");
    return load_factor(x, &d, &cells, factor_id);
}

namespace sogl_dev {
namespace {
// One factor run on the context stream with an explicit executor configuration
// (caller holds the context lock).
int run_factor(Context* c, Factor* f, int mode, int chunk, int wg_per_cu) {
    HIP_TRY(c, hipSetDevice(c->device));
    f->epoch++;
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    int rc;
    const bool masks = f->mask_mode == SOGL_MASKS_ON || (f->mask_mode == SOGL_MASKS_AUTO && f->masks_on);
    if (f->graph_exec && f->graph_masks != masks) {  // the captured schedule is of the other variant
        (void)hipGraphExecDestroy(f->graph_exec);
        (void)hipGraphDestroy(f->graph);
        f->graph_exec = nullptr;
        f->graph = nullptr;
    }
    f->graph_masks = masks;
    auto go2 = [&](auto llt_c, auto masks_c, auto coh_c) {
        constexpr bool L = decltype(llt_c)::value, M = decltype(masks_c)::value, H = decltype(coh_c)::value;
        if (f->precision == SOGL_FP64) return launch_factor<double, false, L, M, H>(c, f, mode, chunk, wg_per_cu);
        if (f->precision == SOGL_FP32) return launch_factor<float, false, L, M, H>(c, f, mode, chunk, wg_per_cu);
        return launch_factor<float, true, L, M, H>(c, f, mode, chunk, wg_per_cu);
    };
    auto go = [&](auto llt_c, auto masks_c) {  // factors with slot reuse: the coherent-load kernels
        return f->reuse ? go2(llt_c, masks_c, std::true_type{}) : go2(llt_c, masks_c, std::false_type{});
    };
    using TT = std::true_type;
    using FF = std::false_type;
    if (f->llt) rc = masks ? go(TT{}, TT{}) : go(TT{}, FF{});
    else rc = masks ? go(FF{}, TT{}) : go(FF{}, FF{});
    if (rc != SOGL_OK) return rc;
    if (!f->llt && f->block_rows > 0) {  // the upper inverses' check (L L^T factors make none)
        const int R = f->block_rows;
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(256), 0, c->stream, (double*)f->uinv_chk_d.p,
                           (size_t)(1 + 4 * UINV_KEEP), 0.0);
        if (f->elem == 8)
            hipLaunchKernelGGL(k_check_uinv<double>, dim3(R), dim3(NT), 0, c->stream, (const double*)f->pool.p,
                               (const int32_t*)f->u_diag_d.p, (const int32_t*)f->uinv_d.p, (double*)f->uinv_chk_d.p);
        else
            hipLaunchKernelGGL(k_check_uinv<float>, dim3(R), dim3(NT), 0, c->stream, (const float*)f->pool.p,
                               (const int32_t*)f->u_diag_d.p, (const int32_t*)f->uinv_d.p, (double*)f->uinv_chk_d.p);
        HIP_TRY(c, hipGetLastError());
    }
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    HIP_TRY(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    f->factor_ms = ms;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ex0, c->ex1));
    f->exec_ms = ms;
    int ctrl[CTRL_WORDS];
    HIP_TRY(c, hipMemcpy(ctrl, f->ctrl_d.p, sizeof(ctrl), hipMemcpyDeviceToHost));
    f->uinv_check.assign((size_t)(1 + 4 * UINV_KEEP), 0.0);
    if (!f->llt && f->block_rows > 0)
        HIP_TRY(c, hipMemcpy(f->uinv_check.data(), f->uinv_chk_d.p, f->uinv_check.size() * 8, hipMemcpyDeviceToHost));
    f->growth_tiles = ctrl[2];
    f->claimed = mode == SOGL_EXEC_PERSISTENT ? ctrl[4] : 0;
    f->dropped = mode == SOGL_EXEC_PERSISTENT ? ctrl[5] : 0;
    f->ran = mode == SOGL_EXEC_PERSISTENT ? ctrl[6] : (int32_t)f->tasks.size();
    float gm;
    std::memcpy(&gm, &ctrl[3], 4);
    f->max_growth = gm;
    if (ctrl[1]) return set_error(c, SOGL_ERR_TIMEOUT, "factor: a dependency wait exceeded its bound");
    if (masks) {
        // the share of the factor's 16x16 blocks that count as zero
        std::vector<uint32_t> m((size_t)f->n_locs);
        HIP_TRY(c, hipMemcpy(m.data(), f->mask_d.p, m.size() * 4, hipMemcpyDeviceToHost));
        int64_t zero = 0, total = 0;
        for (const auto* v : {&f->l_slots, &f->u_slots})
            for (int32_t sl : *v) {
                zero += 16 - __builtin_popcount(m[(size_t)sl] & MASK_ALL);
                total += 16;
            }
        f->zero_blocks = total ? (double)zero / (double)total : 0.0;
        // automatic: after the first masked run keep the masks where they pay.  FP64
        // factors keep them (measured: as fast or faster at every zero share, C3 with 1%
        // zero blocks 85.6 ms masked against 91.2 dense); FP32 and bf16 factors keep
        // them with a quarter or more of the blocks zero (C2, 4%: 69.5 ms masked
        // against 64.7 dense)
        if (f->mask_mode == SOGL_MASKS_AUTO && !f->masks_decided) {
            f->masks_on = f->precision == SOGL_FP64 || f->zero_blocks >= 0.25;
            f->masks_decided = true;
        }
    }
    return SOGL_OK;
}

// Executor configurations the autotuner times, in this order.
struct TuneCand {
    int mode, chunk, wg_per_cu;
};
constexpr TuneCand kTune[SOGL_TUNE_CANDIDATES] = {
    {SOGL_EXEC_PERSISTENT, 1, 0}, {SOGL_EXEC_PERSISTENT, 2, 0}, {SOGL_EXEC_PERSISTENT, 4, 0},
    {SOGL_EXEC_PERSISTENT, 1, 1}, {SOGL_EXEC_LEVELS, 1, 0},     {SOGL_EXEC_GRAPH, 1, 0},
};

int autotune(Context* c, Factor* f, int trials, double* times_ms) {
    trials = std::max(1, trials);
    int best = -1;
    double best_ms = 0;
    for (int k = 0; k < SOGL_TUNE_CANDIDATES; k++) {
        const TuneCand& q = kTune[k];
        if (q.mode == SOGL_EXEC_GRAPH && !f->graph_exec) {  // capture outside the timed trials
            int rc = run_factor(c, f, q.mode, q.chunk, q.wg_per_cu);
            if (rc != SOGL_OK) return rc;
        }
        double t = 1e30;
        for (int r = 0; r < trials; r++) {
            int rc = run_factor(c, f, q.mode, q.chunk, q.wg_per_cu);
            if (rc != SOGL_OK) return rc;
            t = std::min(t, f->exec_ms);
        }
        if (times_ms) times_ms[k] = t;
        if (best < 0 || t < best_ms) {
            best = k;
            best_ms = t;
        }
    }
    f->tuned = best;
    return SOGL_OK;
}
}  // namespace
}  // namespace sogl_dev

extern "C" int sogl_factor_run_config(sogl_ctx* x, int32_t factor_id, int32_t exec_mode, int32_t chunk,
                                      int32_t wg_per_cu) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (chunk < 1 || chunk > 64 || wg_per_cu < 0) return set_error(c, SOGL_ERR_ARG, "bad executor configuration");
    if (exec_mode == SOGL_EXEC_AUTO) {
        if (f->tuned < 0) {
            int rc = autotune(c, f, 1, nullptr);
            if (rc != SOGL_OK) return rc;
        }
        const TuneCand& q = kTune[f->tuned];
        return run_factor(c, f, q.mode, q.chunk, q.wg_per_cu);
    }
    if (exec_mode < SOGL_EXEC_PERSISTENT || exec_mode > SOGL_EXEC_GRAPH)
        return set_error(c, SOGL_ERR_ARG, "unknown executor mode");
    return run_factor(c, f, exec_mode, chunk, wg_per_cu);
}

extern "C" int sogl_factor_run(sogl_ctx* x, int32_t factor_id, int32_t exec_mode) {
    return sogl_factor_run_config(x, factor_id, exec_mode, 1, 0);
}

extern "C" int sogl_factor_autotune(sogl_ctx* x, int32_t factor_id, int32_t trials, double* times_ms,
                                    int32_t* best) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    int rc = autotune(c, f, trials, times_ms);
    if (rc != SOGL_OK) return rc;
    if (best) *best = f->tuned;
    return SOGL_OK;
}

extern "C" int sogl_factor_release(sogl_ctx* x, int32_t factor_id) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_factor(f);
    c->factors.erase(factor_id);
    return SOGL_OK;
}

extern "C" int sogl_factor_get_stats(sogl_ctx* x, int32_t factor_id, sogl_factor_stats* st) {
    if (!x || !st) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    st->factor_ms = f->factor_ms;
    st->exec_ms = f->exec_ms;
    st->solve_ms = f->solve_ms;
    st->pool_bytes = (int64_t)f->pool.bytes;
    st->n_tasks = (int64_t)f->tasks.size();
    st->n_levels = (int32_t)f->level_begin.size() - 1;
    st->growth_tiles = f->growth_tiles;
    st->max_growth = f->max_growth;
    st->zero_blocks = f->zero_blocks;
    st->block_masks = (f->mask_mode == SOGL_MASKS_ON || (f->mask_mode == SOGL_MASKS_AUTO && f->masks_on)) ? 1 : 0;
    st->claimed_tasks = f->claimed;
    st->dropped_tasks = f->dropped;
    st->tasks_run = f->ran;
    st->pool_tiles = f->n_locs;
    st->value_tiles = f->n_slots;
    int32_t nbad = 0;
    if (!f->uinv_check.empty()) std::memcpy(&nbad, f->uinv_check.data(), 4);
    st->uinv_out_of_tol = nbad;
    return SOGL_OK;
}

extern "C" int sogl_factor_get_uinv_check(sogl_ctx* x, int32_t factor_id, int32_t* n_failed, double* detail) {
    if (!x || !n_failed) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    int32_t nbad = 0;
    if (!f->uinv_check.empty()) std::memcpy(&nbad, f->uinv_check.data(), 4);
    *n_failed = nbad;
    if (detail)
        for (int q = 0; q < std::min<int32_t>(nbad, UINV_KEEP); q++)
            for (int e = 0; e < 4; e++) detail[4 * q + e] = f->uinv_check[(size_t)(1 + 4 * q + e)];
    return SOGL_OK;
}

extern "C" int sogl_factor_set_masks(sogl_ctx* x, int32_t factor_id, int32_t mode) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (mode != SOGL_MASKS_OFF && mode != SOGL_MASKS_ON && mode != SOGL_MASKS_AUTO)
        return set_error(c, SOGL_ERR_ARG, "unknown block-mask mode");
    f->mask_mode = mode;
    if (mode == SOGL_MASKS_AUTO) {  // decide again from the next (masked) run
        f->masks_on = true;
        f->masks_decided = false;
    }
    return SOGL_OK;
}

extern "C" int sogl_factor_set_chaining(sogl_ctx* x, int32_t factor_id, int32_t enable) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (enable < 0 || enable > CHAIN_RING) return set_error(c, SOGL_ERR_ARG, "claims per task: 0 .. 8");
    f->chain_claims = enable;
    return SOGL_OK;
}

extern "C" int sogl_factor_get_tiles(sogl_ctx* x, int32_t factor_id, int32_t which, double* tiles) {
    if (!x || !tiles) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    HIP_TRY(c, hipSetDevice(c->device));
    const std::vector<int32_t>& sl = which == 0 ? f->l_slots : f->u_slots;
    if (sl.empty()) return SOGL_OK;
    const size_t tb = (size_t)TE * f->elem;
    DevBuf ids, out, out64;
    int rc;
    if ((rc = dev_alloc(c, ids, sl.size() * 4, "ids")) || (rc = dev_alloc(c, out, sl.size() * tb, "gather")))
        return rc;
    HIP_TRY(c, hipMemcpy(ids.p, sl.data(), sl.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_gather_tiles, dim3((unsigned)sl.size()), dim3(256), 0, c->stream,
                       (const char*)f->pool.p, (const int32_t*)ids.p, (char*)out.p, (int)sl.size(), tb);
    if (f->elem == 4) {
        if ((rc = dev_alloc(c, out64, sl.size() * TE * 8, "gather64"))) return rc;
        hipLaunchKernelGGL(k_f32_to_f64, dim3(1024), dim3(256), 0, c->stream, (const float*)out.p,
                           (double*)out64.p, sl.size() * TE);
        HIP_TRY(c, hipMemcpyAsync(tiles, out64.p, sl.size() * TE * 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        HIP_TRY(c, hipMemcpyAsync(tiles, out.p, sl.size() * tb, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    dev_free(ids);
    dev_free(out);
    dev_free(out64);
    return SOGL_OK;
}

extern "C" int sogl_factor_set_trace(sogl_ctx* x, int32_t factor_id, int32_t enable) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    HIP_TRY(c, hipSetDevice(c->device));
    if (enable) {
        int rc = dev_alloc(c, f->trace_d, std::max<size_t>(f->tasks.size(), 1) * 4 * 8, "trace");
        if (rc) return rc;
    }
    f->trace_on = enable != 0;
    return SOGL_OK;
}

extern "C" int sogl_factor_get_graph(sogl_ctx* x, int32_t factor_id, int64_t* n_deps, int64_t* n_slots,
                                     int32_t* dep_begin, int32_t* deps, int32_t* slot_task) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    int64_t nd = 0;
    for (const DevTask& k : f->tasks) nd += k.dep_count;
    if (n_deps) *n_deps = nd;
    if (n_slots) *n_slots = f->n_slots;
    if (dep_begin || deps) {
        if (!dep_begin || !deps) return set_error(c, SOGL_ERR_ARG, "dep_begin and deps go together");
        std::vector<int32_t> hd((size_t)std::max<int64_t>(nd, 1));
        if (nd) {
            HIP_TRY(c, hipSetDevice(c->device));
            HIP_TRY(c, hipMemcpy(hd.data(), f->deps_d.p, (size_t)nd * 4, hipMemcpyDeviceToHost));
        }
        for (size_t t = 0; t < f->tasks.size(); t++) dep_begin[t] = f->tasks[t].dep_begin;
        dep_begin[f->tasks.size()] = (int32_t)nd;
        std::copy(hd.begin(), hd.begin() + nd, deps);
    }
    if (slot_task) {
        std::fill(slot_task, slot_task + f->n_slots, -1);
        for (size_t t = 0; t < f->tasks.size(); t++) {
            const DevTask& k = f->tasks[t];
            for (int32_t o : {k.f0, k.f1, k.f2, k.f3})
                if (o >= 0 && o < f->n_slots) slot_task[o] = (int32_t)t;
        }
    }
    return SOGL_OK;
}

extern "C" int sogl_factor_get_trace(sogl_ctx* x, int32_t factor_id, int64_t* tasks4, int64_t* times4) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (tasks4)
        for (size_t t = 0; t < f->tasks.size(); t++) {
            const DevTask& k = f->tasks[t];
            tasks4[4 * t + 0] = k.kind;
            tasks4[4 * t + 1] = k.kind == TK_GEMM ? k.b : 0;
            tasks4[4 * t + 2] = k.level;
            tasks4[4 * t + 3] = k.dep_count;
        }
    if (times4) {
        if (!f->trace_d.p) return set_error(c, SOGL_ERR_ARG, "tracing was not enabled");
        HIP_TRY(c, hipSetDevice(c->device));
        HIP_TRY(c, hipMemcpy(times4, f->trace_d.p, f->tasks.size() * 4 * 8, hipMemcpyDeviceToHost));
    }
    return SOGL_OK;
}
