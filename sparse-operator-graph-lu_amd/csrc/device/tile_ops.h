// Device-side 64x64 tile operations of the factor, executed by one
// 256-thread workgroup (4 wave64s) each.  Tiles are dense, row-major
// 64x64 in the HBM tile pool; element (i,j) sits at [i*64 + j].
//
//   gemm_task : C = sum_s sign_s * A_s * B_s           (mul / mulneg chains)
//   sub_task  : R = A - B, R = A, R = -B               (sub / copy / neg)
//   lu_task   : A = L U, pivot-free, |u_kk| < 1e-9 clamped to +-1e-9
//   linv_task : Y = L^-1 (general diagonal)
//   uinv_task : Y = U^-1
//   llt_task  : A = L L^T (lower triangle read), l_jj = sqrt(max(p, 1e-20)), and L^-1
//   trans_task: R = A^T (the U = L^T tiles of an L L^T factor, for the solve)
//
// A GEMM pair may take its second operand transposed (C += A B^T, the mult op of the
// L L^T path, MatrixStdDouble.cpp mat_mult 704-716, 771-783).
//
// The arithmetic follows the reference's tile kernels (MatrixStdDouble.cpp:
// C+=AB 2043-2507, C-=AB 237-701, ludcmpSimple 2711-2784, inv_lower
// 2787-2802, inv_upper 2829-2866, mat_sub/copy/neg 2948-3121) in a GPU form:
// dense tiles instead of 8x8 sub-block masks (mask-skipped blocks are zero,
// so skipping them is exact), MFMA products for the GEMMs, right-looking
// 64-step eliminations in LDS for the LU and the triangular inverses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sogl_dev {

constexpr int NT = 256;      // threads per workgroup
constexpr int TS = 64;       // tile edge
constexpr int TE = TS * TS;  // tile elements
constexpr int LDA = TS + 1;  // LDS row stride of row-read operands (odd: column reads hit distinct banks)
constexpr int LDB = TS + 2;  // LDS row stride of column-read operands (keeps 16-byte row alignment)

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// LDS elements of the blocked LU (defined with it below)
constexpr int lu_lds_elems();

// LDS bytes one workgroup needs for any task kind, per storage type: the GEMM's A and
// B tiles, the blocked LU in the compute type and, for FP32 tiles, the compact FP64
// refactorisation (tile + 2 x 64 doubles)
template <typename T>
constexpr int lds_bytes() {
    return cmax(cmax((TS * LDA + TS * LDB) * (int)sizeof(T), lu_lds_elems() * (int)sizeof(T)),
                sizeof(T) == 8 ? 0 : (TS * LDA + 2 * TS) * 8);
}

// ------------------------------------------------------------------ publishing
// Device-scope release: write the L2 back, then wait until the write-back is done, so
// a flag stored next is seen only after every earlier store of the workgroup (the
// caller's barrier orders the other waves' stores before it).  Written as asm: the
// compiler's wait-count pass does not count buffer_wbl2 as an outstanding vector
// memory operation and drops the wait that should follow it when no store of the wave
// is pending, and the flag can then overtake the write-back (consumers on other XCDs
// read stale tiles).  vmcnt(0) also waits for every earlier vector memory operation of
// the wave.
__device__ __forceinline__ void release_agent() { asm volatile("buffer_wbl2 sc1\n\ts_waitcnt vmcnt(0)" ::: "memory"); }
// Device-scope acquire made by one wave for its workgroup, after it has seen the flags
// (the caller's barrier follows).  The compiler's fence (buffer_inv sc1, no wait): the
// invalidate is ordered before every later load of the CU, the other waves' loads
// after the barrier included.  Waiting for it to finish as well (measured) added about
// 1 us to every hand-off and 3% to the C5 factor.
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Every tile store of the factor tasks is a relaxed agent-scope atomic store, which
// gfx942 / gfx950 code emit as `global_store ... sc1`: written through the L2 to the
// device coherence point, never left dirty in it.  Once a wave's s_waitcnt vmcnt(0)
// has seen them complete they are visible to any CU that invalidates its caches
// (the acquire above), so publishing a tile needs no L2 write-back (buffer_wbl2, which
// writes back every dirty line of the XCD's L2 and was most of a task's publish).
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The wave's own stores (and loads) have completed: with gst stores, the wave's part
// of a workgroup release at agent scope (a barrier after it orders the flag store).
// The comment marks the wait in the assembly, where a test checks that the publish
// barrier follows it with no store in between.
__device__ __forceinline__ void wave_stores_done() { asm volatile("s_waitcnt vmcnt(0) ; publish: stores done" ::: "memory"); }

// Tile reads in a factor that reuses pool slots (COH): a CU's L1 or an XCD's L2 may
// still hold lines of a slot's earlier value from the same run, so every tile read goes
// to the device coherence point -- sc1 buffer loads, the form an agent-scope relaxed
// atomic load takes, but 16 bytes per lane.  Without reuse every slot is written once
// per run and read only after its flag, and each kernel starts with its caches
// invalidated, so plain loads read the right bytes (and hit in L2 when they can).  The
// form is fixed per kernel: a choice per operand inside the pair loop cost registers
// (measured: spills in the kernels built for two workgroups per CU).
constexpr int CPOL_SC1 = 16;  // the sc1 bit of a buffer instruction's cache policy
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void* tile) {
    // raw buffer over one tile (at most 64 x 64 doubles); dword 3: 32-bit data format
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(tile), 0, TE * 8, 0x00020000);
}
template <bool COH, typename T>
__device__ __forceinline__ T ld1c(const T* tile, int i) {  // element i
    if constexpr (!COH) {
        return tile[i];
    } else if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(tile_rsrc(tile), i * 8, 0, CPOL_SC1));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(tile_rsrc(tile), i * 4, 0, CPOL_SC1));
    }
}
// 16-byte vector at element i (16-byte aligned)
template <bool COH, typename V, typename T>
__device__ __forceinline__ V ld16c(const T* tile, int i) {
    static_assert(sizeof(V) == 16, "16-byte vectors");
    if constexpr (!COH)
        return *reinterpret_cast<const V*>(tile + i);
    else
        return __builtin_bit_cast(
            V, __builtin_amdgcn_raw_buffer_load_b128(tile_rsrc(tile), i * (int)sizeof(T), 0, CPOL_SC1));
}
// 16 consecutive elements from element i (a multiple of 4), converted to C
template <bool COH, typename T, typename C>
__device__ __forceinline__ void load16c(const T* tile, int i, C (&v)[16]) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 x = ld16c<COH, float4>(tile, i + 4 * q);
            v[4 * q] = C(x.x);
            v[4 * q + 1] = C(x.y);
            v[4 * q + 2] = C(x.z);
            v[4 * q + 3] = C(x.w);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double2 x = ld16c<COH, double2>(tile, i + 2 * q);
            v[2 * q] = C(x.x);
            v[2 * q + 1] = C(x.y);
        }
    }
}
template <typename T, typename C>
__device__ __forceinline__ void load16(const T* tile, int i, bool coh, C (&v)[16]) {
    if (coh)
        load16c<true>(tile, i, v);
    else
        load16c<false>(tile, i, v);
}

// ------------------------------------------------------------------ loads
// thread t moves row t/4, columns (t%4)*16 .. +15 of a tile between global and LDS
// (that segment lies in 16x16 block (t/64, t%4))
template <bool COH = false, typename T>
__device__ inline void tile_to_lds(const T* __restrict__ g, T* __restrict__ s, int ld, int tid) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    T v[16];
    load16c<COH>(g, r * TS + c0, v);
    T* dst = s + r * ld + c0;
#pragma unroll
    for (int q = 0; q < 16; q++) dst[q] = v[q];
}

// One thread's segment of a tile (the same row segment as tile_to_lds) held in
// registers: the next pair's operands are loaded into it while the current pair
// multiplies from LDS, then written to LDS.
template <typename T>
struct Seg;
template <>
struct Seg<float> {
    float4 v[4];
};
template <>
struct Seg<double> {
    double2 v[8];
};
// on = false: the segment's 16x16 block is zero by the tile's mask and is not loaded
template <bool COH = false, typename T>
__device__ __forceinline__ void seg_load(const T* __restrict__ g, Seg<T>& s, int tid, bool on) {
    if (!on) return;
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    typedef typename std::remove_reference<decltype(s.v[0])>::type V;
    constexpr int NV = (int)(sizeof(s.v) / sizeof(V));
#pragma unroll
    for (int q = 0; q < NV; q++) s.v[q] = ld16c<COH, V>(g, r * TS + c0 + q * (16 / NV));
}
// on = false: write zeros (the block is zero by the tile's mask)
template <typename T>
__device__ __forceinline__ void seg_store(const Seg<T>& s, T* __restrict__ lds, int ld, int tid, bool on) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    T* dst = lds + r * ld + c0;
    if (on) {
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int q = 0; q < 8; q++) {
                dst[2 * q] = s.v[q].x;
                dst[2 * q + 1] = s.v[q].y;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                dst[4 * q] = s.v[q].x;
                dst[4 * q + 1] = s.v[q].y;
                dst[4 * q + 2] = s.v[q].z;
                dst[4 * q + 3] = s.v[q].w;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) dst[q] = T(0);
    }
}
// whether thread tid's segment lies in a block whose bit is set in mask m
__device__ __forceinline__ bool seg_on(uint32_t m, int tid) { return (m >> (4 * (tid >> 6) + (tid & 3))) & 1u; }

// The same with the blocks whose mask bit is clear written as zeros, not loaded.
template <bool COH = false, typename T>
__device__ inline void tile_to_lds_masked(const T* __restrict__ g, T* __restrict__ s, int ld, int tid, uint32_t m) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    if ((m >> (4 * (r >> 4) + (tid & 3))) & 1u) {
        tile_to_lds<COH>(g, s, ld, tid);
    } else {
        T* dst = s + r * ld + c0;
#pragma unroll
        for (int q = 0; q < 16; q++) dst[q] = T(0);
    }
}

// ------------------------------------------------------------------ 16x16 block masks
// Bit 4*I + J of a tile's mask is set when its 16x16 block (I, J) holds an entry with
// |v| > 1e-18.  A block whose bit is clear counts as zero in every product that reads
// the tile: it is neither loaded nor multiplied, and a tile with no bit set takes no
// part at all.  This is the reference's 8x8 sub-block masking (MatrixStdDouble.cpp:
// 120-154, 268-289; GEMM stores set a block's bit where some |v| > 1e-18, 622-623),
// at the 16x16 granularity of the MFMA blocks.
constexpr double MASK_TINY = 1e-18;
constexpr uint32_t MASK_ALL = 0xFFFFu;

__host__ __device__ inline uint32_t mask_transpose(uint32_t m) {
    uint32_t t = 0;
    for (int I = 0; I < 4; I++)
        for (int J = 0; J < 4; J++)
            if ((m >> (4 * I + J)) & 1u) t |= 1u << (4 * J + I);
    return t;
}

// the four blocks of block row I: bits 4I .. 4I+3
__host__ __device__ inline uint32_t mask_row(uint32_t m, int I) { return (m >> (4 * I)) & 0xFu; }

// ------------------------------------------------------------------ GEMM
template <typename T>
struct MfmaOp;
// Output row held by register r of lane group g = lane/16 (column = lane%16).
// Measured on gfx950: the f32 16x16x4 accumulator keeps 4 consecutive rows per
// lane group (row 4g + r); the f64 one interleaves them (row 4r + g).
template <>
struct MfmaOp<double> {
    typedef v4d acc_t;
    static __device__ inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int g, int r) { return 4 * r + g; }
};
template <>
struct MfmaOp<float> {
    typedef v4f acc_t;
    static __device__ inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int g, int r) { return 4 * g + r; }
};

// One (A,B) pair of a GEMM task.  a, b: the operands' pool slots (tile data and block
// masks); b < 0 encodes a subtracted product (mulneg): slot = -b - 1; a < 0 encodes a
// transposed second operand (A B^T): slot = -a - 1.  fa, fb: the operands' value ids,
// which index the ready flags.  A slot holds several values over a factor run when the
// pool reuses slots (a value's slot taken over by the task that alone reads it); each
// value keeps a flag of its own.
struct Pair {
    int32_t a, b;
    int32_t fa, fb;
};
__host__ __device__ inline int pair_a(const Pair& p) { return p.a < 0 ? -p.a - 1 : p.a; }
__host__ __device__ inline int pair_b(const Pair& p) { return p.b < 0 ? -p.b - 1 : p.b; }

// acc += sg * A B (or A B^T) from LDS: A rows with stride LDA; B row-major with stride
// LDB, or, transposed, stored row-major with stride LDA and read by columns.
// kuse: bit K set when the wave's products need the 16-wide k block K (k ascending,
// as in the dense loop, so nothing changes where every block is used).
template <typename T, bool TB, typename acc_t>
__device__ __forceinline__ void gemm_lds(acc_t (&acc)[2][2], const T* la, const T* lb, T sg, int wr, int wc,
                                         int li, int lk, uint32_t kuse) {
    for (int K = 0; K < 4; K++) {
        if (!((kuse >> K) & 1u)) continue;  // wave-uniform
#pragma unroll
        for (int k4 = 0; k4 < 4; k4++) {
            const int k = 16 * K + 4 * k4 + lk;
            T af[2], bf[2];
#pragma unroll
            for (int m = 0; m < 2; m++) af[m] = sg * la[(wr + m * 16 + li) * LDA + k];
#pragma unroll
            for (int n = 0; n < 2; n++)
                bf[n] = TB ? lb[(wc + n * 16 + li) * LDA + k] : lb[k * LDB + wc + n * 16 + li];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++) acc[m][n] = MfmaOp<T>::mma(af[m], bf[n], acc[m][n]);
        }
    }
}

// k blocks a wave with row blocks I0, I0+1 and column blocks J0, J0+1 needs, from the
// masks of A and of B in (k, j) orientation
__device__ __forceinline__ uint32_t k_blocks_used(uint32_t ma, uint32_t mb, int I0, int J0) {
    const uint32_t acols = mask_row(ma, I0) | mask_row(ma, I0 + 1);             // bit K: A(I, K) nonzero
    const uint32_t bt = mask_transpose(mb);                                      // bit 4J + K: B(K, J)
    const uint32_t brows = mask_row(bt, J0) | mask_row(bt, J0 + 1);             // bit K: B(K, J) nonzero
    return acols & brows;
}

// Mask of a wave's 2 x 2 output blocks (block (I0+m, J0+n) counts when some lane saw
// |v| > 1e-18 in it: nz[m][n]), in the tile's bit positions; every lane returns it.
__device__ __forceinline__ uint32_t quad_mask(const bool (&nz)[2][2], int I0, int J0) {
    uint32_t bits = 0;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
            if (__ballot(nz[m][n]) != 0) bits |= 1u << (4 * (I0 + m) + J0 + n);
    return bits;
}

// C = sum_s sign_s A_s B_s over npairs pairs, full 64x64 output per workgroup.
// Wave w computes the 32x32 quadrant (w/2, w%2) as 2x2 blocks of 16x16 MFMA
// outputs; v_mfma_*_16x16x4 lane layout: A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16g with (g, r) given by MfmaOp<T>::row.
// A GEMM task may start before all of its pairs' operands exist: `wait(s, fa, fb, a, b,
// ...)` is called by every thread before the tiles in slots a, b (value ids fa, fb) are
// read and returns once both are ready (it includes the barrier that precedes the
// operand loads).
// The wait also hands every thread the block masks of both tiles (`ma`, `mb`); `s` is
// the pair's index in the task (the add tile comes last).  NoWait: all operands exist
// already (level / per-op launches), the masks come from the mask array.
// The pair loop asks in windows: `extend(pairs, s, npairs, window, fadd)` (every thread,
// a barrier inside; window counts the calls) returns e > s such that pairs s .. e-1 are
// all ready, and `pair_masks(j, p, ...)` then gives pair j's masks for j < e.  The first
// window also looks at the flag of the added tile (value id fadd): after it,
// `add_ready(add, m)` says whether that tile (slot add) could be read from then on
// (every thread gets the same answer).
struct NoWait {
    const uint32_t* mask;
    __device__ inline void operator()(int, int, int, int sa, int sb, uint32_t& ma, uint32_t& mb) const {
        __syncthreads();
        ma = mask[sa];
        mb = mask[sb];
    }
    __device__ inline int extend(const Pair*, int, int npairs, int, int = -1) const {
        __syncthreads();
        return npairs;
    }
    __device__ inline void pair_masks(int, const Pair& p, uint32_t& ma, uint32_t& mb) const {
        ma = mask[pair_a(p)];
        mb = mask[pair_b(p)];
    }
    __device__ inline bool add_ready(int add, uint32_t& m) const {
        m = mask[add];
        return true;
    }
    // a result slot that held another value: its readers ran at earlier levels, i.e. in
    // earlier launches
    __device__ inline void war_done(bool) const {}
};

// A task made by folding a later sub/neg into the GEMM adds (or subtracts) one more
// tile, `add`, to the sum, after the last product (the same order in every executor).
// Its flag is polled with the first window of pairs; a tile that is ready by then is
// never waited for again (STAGE64 loads it at once, into registers held across the pair
// loop; the other kernels, which lack the registers, read it after the loop); one that
// is not is waited for after the pairs.
template <bool COH, typename T, typename acc_t>
__device__ __forceinline__ void add_tile(acc_t (&acc)[2][2], const T* __restrict__ X, bool neg, int wr, int wc,
                                         int li, int lk) {
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const T x = ld1c<COH>(X, (wr + m * 16 + MfmaOp<T>::row(lk, r)) * TS + wc + n * 16 + li);
                acc[m][n][r] += neg ? -x : x;
            }
}

// TBOK: pairs may carry a transposed second operand (only L L^T factors have them; the
// LU kernels are built without that branch).
// MASKS: skip zero products and zero k blocks by the operands' masks, and leave the
// mask bits of wave w's quadrant of the result in out_mask[w]; without, every pair is
// multiplied in full and nothing is left (the result counts as full).
// STAGE64: stage FP64 operands too, and hold an added tile that is ready early in
// registers (the kernel built for one workgroup per CU has the registers for it).
template <typename T, bool ACC = false, bool TBOK = false, bool MASKS = true, typename Wait = NoWait,
          bool STAGE64 = false, bool COH = false>
__device__ __attribute__((always_inline)) void gemm_task(T* __restrict__ pool, const Pair* __restrict__ pairs, int npairs,
                          int out, char* lds_raw, uint32_t* out_mask, const Wait& wait, int add = -1,
                          bool add_neg = false, int fadd = -1, bool to_lds = false) {
    typedef typename MfmaOp<T>::acc_t acc_t;
    T* la = reinterpret_cast<T*>(lds_raw);
    T* lb = la + TS * LDA;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    acc_t acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) {
            acc[m][n] = acc_t{0, 0, 0, 0};
            if constexpr (ACC) {  // C += ... : start from the stored tile
                const T* C0 = pool + (size_t)out * TE;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    acc[m][n][r] = C0[(wr + m * 16 + MfmaOp<T>::row(lk, r)) * TS + wc + n * 16 + li];
            }
        }
    // A product is zero when either operand's mask is (B taken in (k, j) orientation).
    auto zero_product = [&](const Pair& q, uint32_t qa, uint32_t qb) {
        if constexpr (!MASKS) return false;
        const uint32_t qbk = (TBOK && q.a < 0) ? mask_transpose(qb) : qb;
        return qa == 0 || qbk == 0;
    };
    // Pairs [0, ready) are known to be ready.  The operands of `staged` (the next
    // nonzero pair inside that window) are loaded into ga / gb while the current pair
    // multiplies; a pair outside the window is waited for (and the window extended)
    // when the loop reaches it.  FP64 tiles are staged only in the kernel built for one
    // workgroup per CU: 64 more live registers across the MFMA loop make the FP64
    // kernels built for two spill.
    constexpr bool STAGE = sizeof(T) == 4 || STAGE64;
    Seg<T> ga, gb;
    int ready = 0, staged = -1, window = 0;
    uint32_t sma = 0, smb = 0;
    // the added tile: 0 none or not seen yet, 1 ready (read after the loop), 2 in xr
    int add_state = 0;
    T xr[2][2][4];
    const T* X = pool + (size_t)(add >= 0 ? add : 0) * TE;
    for (int s = 0; s < npairs; s++) {
        const Pair p = pairs[s];
        const bool neg = p.b < 0, tb = TBOK && p.a < 0;
        uint32_t ma, mbs;  // the tiles' masks as stored
        bool synced = false;
        if (s != staged) {
            if (s == ready) {
                ready = wait.extend(pairs, s, npairs, window++, fadd);  // its barrier also frees LDS
                if (ready < 0) return;  // another workgroup took the task over (wave-uniform)
                synced = true;
                uint32_t mx;
                if (s == 0 && add >= 0 && wait.add_ready(add, mx)) {
                    add_state = 1;
                    if constexpr (STAGE64) {  // 16 (FP32) or 32 (FP64) more live registers
#pragma unroll
                        for (int m = 0; m < 2; m++)
#pragma unroll
                            for (int n = 0; n < 2; n++)
#pragma unroll
                                for (int r = 0; r < 4; r++)
                                    xr[m][n][r] =
                                        ld1c<COH>(X, (wr + m * 16 + MfmaOp<T>::row(lk, r)) * TS + wc + n * 16 + li);
                        add_state = 2;
                    }
                }
            }
            wait.pair_masks(s, p, ma, mbs);
            if (zero_product(p, ma, mbs)) continue;  // every thread skips
            seg_load<COH>(pool + (size_t)pair_a(p) * TE, ga, tid, !MASKS || seg_on(ma, tid));
            seg_load<COH>(pool + (size_t)pair_b(p) * TE, gb, tid, !MASKS || seg_on(mbs, tid));
        } else {
            ma = sma;
            mbs = smb;
        }
        if (!synced) __syncthreads();  // every wave is done reading the previous pair from LDS
        seg_store(ga, la, LDA, tid, !MASKS || seg_on(ma, tid));
        seg_store(gb, lb, tb ? LDA : LDB, tid, !MASKS || seg_on(mbs, tid));
        staged = -1;
        for (int n = s + 1; STAGE && n < ready; n++) {
            const Pair q = pairs[n];
            uint32_t qa, qb;
            wait.pair_masks(n, q, qa, qb);
            if (zero_product(q, qa, qb)) continue;
            seg_load<COH>(pool + (size_t)pair_a(q) * TE, ga, tid, !MASKS || seg_on(qa, tid));
            seg_load<COH>(pool + (size_t)pair_b(q) * TE, gb, tid, !MASKS || seg_on(qb, tid));
            staged = n;
            sma = qa;
            smb = qb;
            break;
        }
        __syncthreads();
        uint32_t kuse = 0xFu;
        if constexpr (MASKS) kuse = k_blocks_used(ma, tb ? mask_transpose(mbs) : mbs, wr >> 4, wc >> 4);
        const T sg = neg ? T(-1) : T(1);
        if (tb) gemm_lds<T, TBOK>(acc, la, lb, sg, wr, wc, li, lk, kuse);
        else gemm_lds<T, false>(acc, la, lb, sg, wr, wc, li, lk, kuse);
    }
    if (add_state == 2) {
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int r = 0; r < 4; r++) acc[m][n][r] += add_neg ? -xr[m][n][r] : xr[m][n][r];
    } else if (add >= 0) {
        if (add_state == 0) {
            uint32_t m0, m1;
            wait(npairs, fadd, fadd, add, add, m0, m1);
        }
        add_tile<COH, T>(acc, X, add_neg, wr, wc, li, lk);
    }
    wait.war_done(true);
    T* C = pool + (size_t)out * TE;
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
    if (to_lds) {  // the lu this workgroup runs next takes the sum from LDS (row stride LDA)
        __syncthreads();  // every wave is done reading the staged operands
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int r = 0; r < 4; r++) la[(wr + m * 16 + MfmaOp<T>::row(lk, r)) * LDA + wc + n * 16 + li] = acc[m][n][r];
    }
}

// bf16 experiment: FP32 tiles, operands rounded to bf16, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16: A[i][8q..8q+7] in lane i + 16q, same D layout).
template <bool TBOK = false, bool MASKS = true, typename Wait = NoWait, bool COH = false>
__device__ __attribute__((always_inline)) void gemm_task_bf16(float* __restrict__ pool, const Pair* __restrict__ pairs,
                                      int npairs, int out, char* lds_raw, uint32_t* out_mask, const Wait& wait,
                                      int add = -1, bool add_neg = false, int fadd = -1) {
    float* la = reinterpret_cast<float*>(lds_raw);
    float* lb = la + TS * LDA;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    v4f acc[2][2];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++) acc[m][n] = v4f{0, 0, 0, 0};
    for (int s = 0; s < npairs; s++) {
        const Pair p = pairs[s];
        const bool neg = p.b < 0, tb = p.a < 0;
        const int sa = pair_a(p), sb = pair_b(p);
        uint32_t ma, mbs;
        wait(s, p.fa, p.fb, sa, sb, ma, mbs);
        if constexpr (MASKS) {
            if (ma == 0 || mbs == 0) continue;  // a zero product: every thread skips
            tile_to_lds_masked<COH>(pool + (size_t)sa * TE, la, LDA, tid, ma);
            tile_to_lds_masked<COH>(pool + (size_t)sb * TE, lb, TBOK && tb ? LDA : LDB, tid, mbs);
        } else {
            tile_to_lds<COH>(pool + (size_t)sa * TE, la, LDA, tid);
            tile_to_lds<COH>(pool + (size_t)sb * TE, lb, TBOK && tb ? LDA : LDB, tid);
        }
        __syncthreads();
        const float sg = neg ? -1.f : 1.f;
        // B[k][j] at k * LDB + j, or (transposed operand) at j * LDA + k
        const int bsk = TBOK && tb ? 1 : LDB, bsj = TBOK && tb ? LDA : 1;
#pragma unroll
        for (int kk = 0; kk < TS / 32; kk++) {
            v8bf af[2], bf[2];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int e = 0; e < 8; e++)
                    af[m][e] = (__bf16)(sg * la[(wr + m * 16 + li) * LDA + kk * 32 + lk * 8 + e]);
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int e = 0; e < 8; e++)
                    bf[n][e] = (__bf16)lb[(kk * 32 + lk * 8 + e) * bsk + (wc + n * 16 + li) * bsj];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bf[n], acc[m][n], 0, 0, 0);
        }
    }
    if (add >= 0) {  // the bf16 kernel's accumulator rows follow the f32 layout
        uint32_t m0, m1;
        wait(npairs, fadd, fadd, add, add, m0, m1);
        add_tile<COH, float>(acc, pool + (size_t)add * TE, add_neg, wr, wc, li, lk);
    }
    wait.war_done(false);
    float* C = pool + (size_t)out * TE;
    bool nz[2][2] = {{false, false}, {false, false}};
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float v = acc[m][n][r];
                gst(&C[(wr + m * 16 + 4 * lk + r) * TS + wc + n * 16 + li], v);
                if constexpr (MASKS) nz[m][n] |= fabsf(v) > float(MASK_TINY);
            }
    if constexpr (MASKS) {
        const uint32_t qm = quad_mask(nz, wr >> 4, wc >> 4);
        if (lane == 0) out_mask[w] = qm;
    }
}

// Mask of a stored tile, by one workgroup: wave w covers block row w; lane l reads
// row 16w + l/4 (and + 16 more rows each round), columns 16(l%4) .. +15.
template <typename T>
__device__ inline uint32_t tile_mask(const T* __restrict__ g, int tid) {
    const int lane = tid & 63, w = tid >> 6;
    bool any = false;
    // the thread's segments: rows 16w + lane/4 (lanes cover rows 16w .. 16w+15), cols 16 * (lane % 4)
    const T* p = g + (16 * w + (lane >> 2)) * TS + 16 * (lane & 3);
#pragma unroll
    for (int q = 0; q < 16; q++) any |= fabs((double)p[q]) > MASK_TINY;
    const unsigned long long b = __ballot(any);
    uint32_t bits = 0;
#pragma unroll
    for (int J = 0; J < 4; J++)
        if (b & (0x1111111111111111ull << J)) bits |= 1u << (4 * w + J);
    return bits;
}

// ------------------------------------------------------------------ sub
// sub(src, src2 -> r): r = tile(src2) - tile(src); a missing operand (slot < 0)
// reads as zero, which gives copy and negate (BlockPlanner.cpp:553-572).  All
// loads are issued before the stores (the operands may not alias the result,
// but the compiler cannot know that).
// Wave w leaves the mask bits of the rows it wrote in out_mask[w] (element q*256 + tid
// is row 4q + w, column tid % 64).
template <typename T, bool MASKS = true, bool COH = false>
__device__ __attribute__((always_inline)) void sub_task(T* __restrict__ pool, int src, int src2, int out,
                                                        uint32_t* out_mask) {
    const int tid = threadIdx.x, lane = tid & 63;
    constexpr int Q = TE / NT;  // 16 elements per thread
    T a[Q], b[Q];
    const T* A = src2 >= 0 ? pool + (size_t)src2 * TE : nullptr;
    const T* B = src >= 0 ? pool + (size_t)src * TE : nullptr;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        a[q] = A ? ld1c<COH>(A, q * NT + tid) : T(0);
        b[q] = B ? ld1c<COH>(B, q * NT + tid) : T(0);
    }
    T* R = pool + (size_t)out * TE;
    uint32_t bits = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const T v = a[q] - b[q];
        gst(&R[q * NT + tid], v);
        if constexpr (MASKS) {
            const unsigned long long bb = __ballot(fabs(v) > T(MASK_TINY));
#pragma unroll
            for (int J = 0; J < 4; J++)
                if (bb & (0xFFFFull << (16 * J))) bits |= 1u << (4 * (q >> 2) + J);
        }
    }
    if (MASKS && lane == 0) out_mask[tid >> 6] = bits;
}

// ------------------------------------------------------------------ register-resident tile algebra
// Layout of the LU and inverse routines: wave w owns columns 16w .. 16w+15 and lane
// i owns row i, so a thread keeps the 16 entries (i, 16w + c) in registers.  Row k
// of the wave's columns is read with readlane (lane k), multipliers of the LU cross
// waves through LDS once per step.
__device__ inline double readlane(double x, int l) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffull), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline float readlane(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

template <typename T, typename C>
__device__ inline void store_rowseg(T* g, const C* v, int lane, int w) {
    T* p = g + lane * TS + 16 * w;
#pragma unroll
    for (int c = 0; c < 16; c++) gst(&p[c], T(v[c]));
}

// Row k of the wave's 16 columns, from lane k to every lane of the wave through
// LDS (bro: 16 values per wave).  LDS accesses of one wave complete in order, so
// lane k's stores are seen by the loads that follow; the wave barrier only keeps
// the compiler from reordering them.
template <typename C>
__device__ inline void bcast_row(const C* v, C* bro, int lane, int k, C* out) {
    if (lane == k) {
#pragma unroll
        for (int c = 0; c < 16; c++) bro[c] = v[c];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int c = 0; c < 16; c++) out[c] = bro[c];
    __builtin_amdgcn_wave_barrier();
}

// y = inverse of the lower triangle in ls (LDS, row stride LDA), rows of y in
// lanes, the wave's 16 columns in registers; right-looking over rows, no barrier
// (inv_lower, MatrixStdDouble.cpp:2787-2802: y_ii = 1/l_ii,
// y_ij = -(sum_k l_ik y_kj) / l_ii).  UNIT skips the divisions by l_ii = 1.
template <typename C, bool UNIT>
__device__ inline void inv_lower_regs(const C* ls, C* y, C* bro, int lane, int w) {
#pragma unroll
    for (int c = 0; c < 16; c++) y[c] = (16 * w + c == lane) ? C(1) : C(0);
#pragma unroll 1
    for (int k = 0; k < TS; k++) {
        if (!UNIT) {  // row k is final: scale by 1 / l_kk
            const C q = C(1) / ls[k * LDA + k];
#pragma unroll
            for (int c = 0; c < 16; c++) y[c] = lane == k ? y[c] * q : y[c];
        }
        C yk[16];
        bcast_row(y, bro, lane, k, yk);
        const C f = ls[lane * LDA + k];
        const bool below = lane > k;
#pragma unroll
        for (int c = 0; c < 16; c++)  // branch-free: columns > k of y_k are zero
            y[c] = (below && 16 * w + c <= k) ? y[c] - f * yk[c] : y[c];
    }
}

// y = inverse of the upper triangle in us (LDS): rows from the bottom, row k scaled
// by 1/u_kk once final (inv_upper, MatrixStdDouble.cpp:2829-2866).
template <typename C>
__device__ inline void inv_upper_regs(const C* us, C* y, C* bro, int lane, int w) {
#pragma unroll
    for (int c = 0; c < 16; c++) y[c] = (16 * w + c == lane) ? C(1) : C(0);
#pragma unroll 1
    for (int k = TS - 1; k >= 0; k--) {
        const C q = C(1) / us[k * LDA + k];
#pragma unroll
        for (int c = 0; c < 16; c++) y[c] = lane == k ? y[c] * q : y[c];
        C yk[16];
        bcast_row(y, bro, lane, k, yk);
        const C f = us[lane * LDA + k];
        const bool above = lane < k;
#pragma unroll
        for (int c = 0; c < 16; c++)  // branch-free: columns < k of y_k are zero
            y[c] = (above && 16 * w + c >= k) ? y[c] - f * yk[c] : y[c];
    }
}

// max |x| over the 64 lanes of a wave, in every lane: a DPP butterfly inside each
// 16-lane row (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror), then the
// four row maxima by readlane.  A DPP source lane is always active here (full EXEC).
__device__ __forceinline__ float wave_absmax(float x) {
    x = fabsf(x);
    x = fmaxf(x, __builtin_amdgcn_update_dpp(0.f, x, 0xB1, 0xF, 0xF, false));
    x = fmaxf(x, __builtin_amdgcn_update_dpp(0.f, x, 0x4E, 0xF, 0xF, false));
    x = fmaxf(x, __builtin_amdgcn_update_dpp(0.f, x, 0x141, 0xF, 0xF, false));
    x = fmaxf(x, __builtin_amdgcn_update_dpp(0.f, x, 0x140, 0xF, 0xF, false));
    return fmaxf(fmaxf(readlane(x, 0), readlane(x, 16)), fmaxf(readlane(x, 32), readlane(x, 48)));
}

// ------------------------------------------------------------------ blocked tile LU
// The lu task in 16x16 blocks: for K = 0..3 one wave factors the diagonal block
// (A_KK = L_KK U_KK, plus L_KK^-1 and U_KK^-1), the panel becomes U_KJ = L_KK^-1 A_KJ
// and L_IK = A_IK U_KK^-1, and the trailing blocks take A_IJ -= L_IK U_KJ, the last two
// on MFMA.  L^-1 and U^-1 of the whole tile are then assembled from the blocks by
// recursive doubling (two MFMA rounds).  This is the reference's own recursive 2x2
// block LU with explicit inverses (BlockPlanner.cpp:865-939, 1019-1092) applied one
// level further down, inside the tile; pivots are the same Schur-complement diagonal
// entries as in the unblocked ludcmpSimple (MatrixStdDouble.cpp:2711-2784), clamped
// the same way.  The dependent chain is 4 x 16 short in-wave steps plus 10 barriers,
// instead of 64 barrier-separated steps and two 64-step inverse sweeps.
constexpr int BS = 17;        // LDS row stride of a 16x16 block
constexpr int BE = 16 * BS;   // LDS elements of a 16x16 block

// LDS elements (in the compute type) of lu_inv_blocked: the tile, 8 diagonal-block
// inverses, 12 blocks of partial inverse products, 8 reduction slots, the 2 x 64
// column scales of the Cholesky form, and one slot for the look-ahead's panel counter
constexpr int lu_lds_elems() { return TS * LDA + 8 * BE + 12 * BE + 8 + 2 * TS + 1; }

// Lane k of each 16-lane row to every lane of that row (DPP row_newbcast).  A DPP
// read of a lane that EXEC disables returns 0, so callers evaluate it where every
// lane of the wave is active, never inside one arm of a conditional expression.
template <int K, typename C>
__device__ __forceinline__ C rbcast(C x) {
    return __builtin_amdgcn_update_dpp(C(0), x, 0x150 + K, 0xF, 0xF, true);
}

// x += (x of lane K of this 16-lane row) * s, one v_fmac with the row broadcast as
// its DPP source operand.  The compiler folds a row broadcast into v_add / v_mul but
// not into the tied-destination v_fmac, which would leave a separate v_mov_dpp per
// update.  A DPP source register needs two wait states after a VALU instruction
// writes it, which the compiler does not track through asm statements, so the
// callers keep the rule themselves: dpp_fence before the first update (the values'
// VALU definitions come before it, then s_nop 1), fmac_rbcast_nop for the column the
// next step's pivot broadcast reads, and dpp_step_end after each step (the next
// step's updates read what this step's wrote).  The asm statements are volatile, so
// they keep their order; the emitted assembly is checked for the rule by a test.
// Every lane of the wave must be active, as for rbcast.
template <int K>
__device__ __forceinline__ void fmac_rbcast(float& x, float s) {
    asm volatile("v_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf"
                 : "+v"(x) : "v"(s), "n"(K));
}
template <int K>
__device__ __forceinline__ void fmac_rbcast(double& x, double s) {
    asm volatile("v_fmac_f64_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf"
                 : "+v"(x) : "v"(s), "n"(K));
}
template <int K>
__device__ __forceinline__ void fmac_rbcast_nop(float& x, float s) {
    asm volatile("v_fmac_f32_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf\n\ts_nop 1"
                 : "+v"(x) : "v"(s), "n"(K));
}
template <int K>
__device__ __forceinline__ void fmac_rbcast_nop(double& x, double s) {
    asm volatile("v_fmac_f64_dpp %0, %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf\n\ts_nop 1"
                 : "+v"(x) : "v"(s), "n"(K));
}
__device__ __forceinline__ void dpp_step_end() { asm volatile("s_nop 1"); }
template <typename C>
__device__ __forceinline__ void dpp_fence(C (&v)[16]) {
    asm volatile("s_nop 1"
                 : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]),
                   "+v"(v[7]), "+v"(v[8]), "+v"(v[9]), "+v"(v[10]), "+v"(v[11]), "+v"(v[12]),
                   "+v"(v[13]), "+v"(v[14]), "+v"(v[15]));
}

// 1/p for a clamped pivot (1e-9 <= |p|, finite): v_rcp plus two Newton steps for
// FP64 (each step squares the relative error of the hardware estimate), v_rcp_f32
// (1 ulp) for FP32.  Shorter than the IEEE division, which the pivot chain waits on.
__device__ __forceinline__ double recip(double p) {
    double r = __builtin_amdgcn_rcp(p);
    double e = fma(-p, r, 1.0);
    r = fma(r, e, r);
    e = fma(-p, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float recip(float p) { return __builtin_amdgcn_rcpf(p); }

template <int K, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (K < N) {
        f(std::integral_constant<int, K>());
        static_for<K + 1, N>(f);
    }
}

// acc += sg * A (16 x 4*KS) B (4*KS x 16), A and B row-major in LDS
template <typename C, int KS>
__device__ __forceinline__ void mm16(typename MfmaOp<C>::acc_t& acc, const C* pa, int lda, const C* pb,
                                     int ldb, int li, int lk, C sg = C(1)) {
#pragma unroll
    for (int s = 0; s < KS; s++)
        acc = MfmaOp<C>::mma(sg * pa[li * lda + 4 * s + lk], pb[(4 * s + lk) * ldb + li], acc);
}

template <typename C>
__device__ __forceinline__ void st16(const typename MfmaOp<C>::acc_t& acc, C* pd, int ld, int li, int lk,
                                     C sg = C(1)) {
#pragma unroll
    for (int r = 0; r < 4; r++) pd[MfmaOp<C>::row(lk, r) * ld + li] = sg * acc[r];
}

template <typename T, typename C>
__device__ __forceinline__ void st16_global(const typename MfmaOp<C>::acc_t& acc, T* tile, int I, int J,
                                            int li, int lk, C sg = C(1)) {
#pragma unroll
    for (int r = 0; r < 4; r++)
        gst(&tile[(16 * I + MfmaOp<C>::row(lk, r)) * TS + 16 * J + li], T(sg * acc[r]));
}

// 16x16 block of LDS (stride ld) to global block (I, J) of a tile; 4 elements per lane
template <typename T, typename C>
__device__ __forceinline__ void blk_to_global(const C* pb, int ld, T* tile, int I, int J, int lane) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int r = 4 * q + (lane >> 4), c = lane & 15;
        gst(&tile[(16 * I + r) * TS + 16 * J + c], T(pb[r * ld + c]));
    }
}

// the same two stores with row i of the tile scaled by rs[i] (rs == nullptr: none)
template <typename T, typename C>
__device__ __forceinline__ void st16_global_rs(const typename MfmaOp<C>::acc_t& acc, T* tile, int I, int J,
                                               int li, int lk, C sg, const C* rs) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = 16 * I + MfmaOp<C>::row(lk, r);
        gst(&tile[row * TS + 16 * J + li], T(sg * acc[r] * rs[row]));
    }
}
template <typename T, typename C>
__device__ __forceinline__ void blk_to_global_rs(const C* pb, int ld, T* tile, int I, int J, int lane, const C* rs) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int r = 4 * q + (lane >> 4), c = lane & 15;
        gst(&tile[(16 * I + r) * TS + 16 * J + c], T(pb[r * ld + c] * rs[16 * I + r]));
    }
}

template <typename T>
__device__ __forceinline__ void zero_blk_global(T* tile, int I, int J, int lane) {
#pragma unroll
    for (int q = 0; q < 4; q++) gst(&tile[(16 * I + 4 * q + (lane >> 4)) * TS + 16 * J + (lane & 15)], T(0));
}

// One wave factors the 16x16 diagonal block at (r0, r0) of the LDS tile a in place
// (L strict lower with unit diagonal, U on and above the diagonal, |u_kk| < 1e-9
// clamped to +-1e-9 as in MatrixStdDouble.cpp:2751-2761) and writes L^-1 to xd and
// U^-1 to zd.  Lanes 0-15 and 32-47 hold the block's rows, lanes 16-31 and 48-63 its
// columns, lanes 0-31 the even entries and lanes 32-63 the odd ones; a row_newbcast of
// lane k hands every lane of a 16-lane group row k (or column k) of the block.  In
// both layouts step k updates
// x[c] -= (own_k / p) * bc[c] for lanes past k, with own_k the lane's own entry of
// column (row) k: the row layout keeps own_k / p as the multiplier l_ik, the column
// layout keeps own_k = u_kj.  The same factor f = own_k / p is the step-k multiplier
// of the forward substitution for L^-1 (rows, lanes 0-15: l_ik) and for (U^T)^-1
// (rows, lanes 16-31: u_ki / u_kk, the 1/u_ii scaling applied at the end), so both
// inverses advance in the same 16 steps as the factorisation.
// The pivot rule: |p| < 1e-9 -> +-1e-9 (ludcmpSimple); CHOL: p < 1e-20 -> 1e-20, the
// rule lltdcmpSimple applies before its square root (MatrixStdDouble.cpp:2612).
template <bool CHOL, typename C>
__device__ __forceinline__ C clamp_pivot(C p) {
    if constexpr (CHOL) {
        return p < C(1e-20) ? C(1e-20) : p;
    } else {
        const C TINY2 = C(1e-9);
        return (p < TINY2 && p > -TINY2) ? ((p < 0) ? -TINY2 : TINY2) : p;
    }
}

// The 16 elimination steps on the registers of diag16 (FP32), one scheduled asm
// statement written by gen_diag16.py: the pivot clamp, the multipliers, both
// inverses' updates, x[k] = l_ik for li > k, and pd = the clamped pivot of step li
// (in the half of the wave that holds entry li).  x[j], y[j]: entry 2j + (lane >> 5);
// lg0 / lg1: li in half 0 / 1 and 0 in the other, le0 / le1: li or 99.
template <bool CHOL>
__device__ __forceinline__ void diag16_elim_f32(float (&x)[8], float (&y)[8], float& pd, int lg0, int lg1, int le0,
                                                int le1);
#include "diag16_asm.h"

template <typename C, bool CHOL = false>
__device__ inline void diag16(C* a, int r0, C* xd, C* zd, int lane) {
    static_assert(sizeof(C) == 4, "FP32 only; FP64 tiles use diag16_split");
    const int i = lane & 15, h = lane >> 5;
    const bool colw = (lane >> 4) & 1;
    const int rs = colw ? 1 : LDA, cs = colw ? LDA : 1;
    float x[8], y[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int c = 2 * j + h;
        x[j] = a[r0 * LDA + r0 + i * rs + c * cs];
        y[j] = c == i ? 1.f : 0.f;
    }
    float pd = 1.f;
    diag16_elim_f32<CHOL>(x, y, pd, h == 0 ? i : 0, h == 1 ? i : 0, h == 0 ? i : 99, h == 1 ? i : 99);
    // pd of row / column i was set in half i & 1: lanes 32-63 of the swap's first result
    // take lanes 0-31 of its second and the other way round, so the first holds half 0's
    // pd in both halves and the second half 1's
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(pd), __float_as_uint(pd), false, false);
    pd = __uint_as_float((i & 1) ? sw[1] : sw[0]);
    if (!colw) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            a[(r0 + i) * LDA + r0 + 2 * j + h] = x[j];
            xd[i * BS + 2 * j + h] = y[j];
        }
        if ((i & 1) == h) a[(r0 + i) * LDA + r0 + i] = pd;  // the clamped pivot over u_ii (same lane, after)
    } else {
        const float dq = recip(pd);  // 1 / u_ii as taken at step i
#pragma unroll
        for (int j = 0; j < 8; j++) zd[(2 * j + h) * BS + i] = y[j] * dq;
    }
}

// The same outputs in two sweeps (LU, then both inverses), for FP64: holding the
// FP64 rows of the LU and of the inverses at once would exceed the registers the
// executor has at two workgroups per CU.  Lanes i, i+16, i+32, i+48 hold row i;
// after the LU, lanes 0-15 invert L by rows and lanes 16-31 invert U^T by rows with
// the same instructions (forward substitution, 1/u_ii scaling deferred to the end).
template <typename C, bool CHOL = false>
__device__ inline void diag16_split(C* a, int r0, C* xd, C* zd, int lane) {
    const int i = lane & 15, g = lane >> 4;
    C v[16];
#pragma unroll
    for (int c = 0; c < 16; c++) v[c] = a[(r0 + i) * LDA + r0 + c];
    dpp_fence(v);
    static_for<0, 16>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const C p = clamp_pivot<CHOL>(rbcast<k>(v[k]));
        const C nl = i > k ? v[k] * -recip(p) : C(-0.0);  // -l_ik; lanes <= k unchanged
        v[k] = i > k ? -nl : (i == k ? p : v[k]);
        if constexpr (k < 15) fmac_rbcast_nop<k>(v[k + 1], nl);  // the next pivot
        static_for<k + 2, 16>([&](auto cc) { fmac_rbcast<k>(v[decltype(cc)::value], nl); });
        dpp_step_end();
    });
    if (g == 0) {
#pragma unroll
        for (int c = 0; c < 16; c++) a[(r0 + i) * LDA + r0 + c] = v[c];
    }
    __builtin_amdgcn_wave_barrier();
    C m[16];
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const C ut = c <= i ? a[(r0 + c) * LDA + r0 + i] : C(0);
        m[c] = g == 0 ? (c < i ? v[c] : (c == i ? C(1) : C(0))) : ut;
    }
    const C di = C(1) / (g == 0 ? C(1) : a[(r0 + i) * LDA + r0 + i]);
    C y[16];
#pragma unroll
    for (int c = 0; c < 16; c++) y[c] = c == i ? C(1) : C(0);
    dpp_fence(y);
    static_for<0, 16>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const C qk = rbcast<k>(di);  // outside the conditional: every lane must run it
        const C nf = i > k ? m[k] * -qk : C(-0.0);
        static_for<0, k + 1>([&](auto cc) { fmac_rbcast<k>(y[decltype(cc)::value], nf); });
        dpp_step_end();
    });
    if (g == 0) {
#pragma unroll
        for (int c = 0; c < 16; c++) xd[i * BS + c] = y[c] * di;
    } else if (g == 1) {
#pragma unroll
        for (int c = 0; c < 16; c++) zd[c * BS + i] = y[c] * di;
    }
}

// Factor blocks of the LDS tile a (row stride LDA) to global block (I, J) of L or U,
// 4 elements per lane (rows 4q + lane/16, column lane%16).  L (I >= J): the strict
// lower part of a, 1 on the diagonal, 0 above; CHOL: L_u diag(d) with d = dsc.
// U (I <= J): a on and above the diagonal (the clamped pivots on it), 0 below; um
// collects max |u| for the growth check.
template <typename T, typename C, bool CHOL>
__device__ __forceinline__ void l_blk_to_global(const C* a, T* Lg, int I, int J, int lane, const C* dsc) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int r = 4 * q + (lane >> 4), c = lane & 15;
        const int gr = 16 * I + r, gc = 16 * J + c;
        const C x = a[gr * LDA + gc];
        C v = (I > J || r > c) ? x : (r == c ? C(1) : C(0));
        if constexpr (CHOL) v = (I == J && r == c) ? dsc[gc] : v * dsc[gc];
        gst(&Lg[gr * TS + gc], T(v));
    }
}
template <typename T, typename C>
__device__ __forceinline__ void u_blk_to_global(const C* a, T* Ug, int I, int J, int lane, C& um) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int r = 4 * q + (lane >> 4), c = lane & 15;
        const int gr = 16 * I + r, gc = 16 * J + c;
        const C x = a[gr * LDA + gc];
        const C v = (I < J || c >= r) ? x : C(0);
        gst(&Ug[gr * TS + gc], T(v));
        um = fabs(v) > um ? fabs(v) : um;
    }
}

// LU of one tile in compute type C plus, when their slots are given (>= 0), the
// inverses of its unit-lower L and upper U: the reference's lowerInv / upperInv
// of the same diagonal tile, fused into the producing task.
// CHOL: the Cholesky factor of a symmetric tile from the same elimination.  Only the
// lower triangle is read (mirrored on load), pivots follow the Cholesky clamp, and
// with d_j = sqrt(u_jj) the outputs are L = L_u diag(d) and L^-1 = diag(1/d) L_u^-1
// (A = L_u U with U = diag(u) L_u^T, so L L^T = A); outu and outui are unused.
//
// Schedule: step K factors diagonal block K on wave 0 (diag16), then all waves form
// block row / column K (panel) and update the trailing blocks.  While wave 0 factors
// block K, waves 1-3 finish what block row / column K-1 allows: the stores of its L,
// U and inverse blocks, and the block products of the inverses
//   X_IJ = -X_II N_IJ,  N_IJ = sum_{J <= K < I} L_IK X_KJ      (X = L^-1)
//   Z_IJ = -M_IJ Z_JJ,  M_IJ = sum_{I <= K < J} Z_IK U_KJ      (Z = U^-1)
// accumulated term by term as the factors of column / row K appear (N and M live in
// LDS; X_IJ and Z_IJ replace them once complete).  After block 3, X_3J and Z_I3 take
// one product each, spread over the four waves.
template <typename T, typename C, bool CHOL = false>
__device__ __attribute__((always_inline)) void lu_inv_blocked(T* __restrict__ pool, int src, int outl, int outu,
                                                              int outli, int outui, char* lds_raw,
                                                              float* amax_out, float* umax_out, bool coh = false,
                                                              C* xcopy = nullptr, C* zcopy = nullptr, bool in_lds = false) {
    typedef typename MfmaOp<C>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int li = lane & 15, lk = lane >> 4;
    // The stores and products of waves 1-3 use bl, bli, blk, re-defined opaquely where
    // that work starts in each step: their block addresses then stay inside the step
    // instead of being hoisted out of the step loop and held in registers throughout
    // (wave 0's own loop-invariant values stay hoisted).
    int bl = lane, bli = li, blk = lk;
    C* a = reinterpret_cast<C*>(lds_raw);  // the tile, row stride LDA
    C* xd = a + TS * LDA;                  // X_II = L_II^-1, I = 0..3
    C* zd = xd + 4 * BE;                   // Z_II = U_II^-1
    C* nm = zd + 4 * BE;                   // 12 blocks: N_IJ / X_IJ and M_IJ / Z_IJ
    C* red = nm + 12 * BE;                 // 8 reduction slots
    C* dsc = red + 8;                      // CHOL: d_j = sqrt(u_jj), then 1 / d_j
    int* pcnt = reinterpret_cast<int*>(dsc + 2 * TS);  // panel blocks written (look-ahead)
    enum { N10, N20, N30, N21, N31, N32, M01, M02, M03, M12, M13, M23 };
    T* Lg = pool + (size_t)outl * TE;
    T* Ug = (!CHOL && outu >= 0) ? pool + (size_t)outu * TE : nullptr;
    T* Xg = outli >= 0 ? pool + (size_t)outli * TE : nullptr;
    T* Zg = (!CHOL && outui >= 0) ? pool + (size_t)outui * TE : nullptr;
    auto ablk = [&](int I, int J) { return a + 16 * I * LDA + 16 * J; };
    auto slot = [&](int s) { return nm + s * BE; };
    auto ldacc = [&](const C* p) {
        acc_t acc;
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] = p[MfmaOp<C>::row(blk, r) * BS + bli];
        return acc;
    };
    // N_s (+)= L_IK X and M_s (+)= Z U_KJ, X and Z blocks of row stride BS
    auto n_add = [&](int s, bool first, int I, int K, const C* X) {
        acc_t acc = first ? acc_t{0, 0, 0, 0} : ldacc(slot(s));
        mm16<C, 4>(acc, ablk(I, K), LDA, X, BS, bli, blk);
        st16(acc, slot(s), BS, bli, blk);
    };
    auto m_add = [&](int s, bool first, const C* Z, int K, int J) {
        acc_t acc = first ? acc_t{0, 0, 0, 0} : ldacc(slot(s));
        mm16<C, 4>(acc, Z, BS, ablk(K, J), LDA, bli, blk);
        st16(acc, slot(s), BS, bli, blk);
    };
    // X_IJ = -X_II N_IJ and Z_IJ = -M_IJ Z_JJ to HBM, and in place over N / M (keep)
    // xcopy / zcopy (not CHOL; nullptr: none): copies of X (row stride LDA) and Z (row
    // stride LDB) in LDS beside the stores, for panel products the caller runs next
    const bool cpx = !CHOL && xcopy, cpz = !CHOL && zcopy;
    auto x_fin = [&](int s, int I, int J, bool keep) {
        acc_t acc = acc_t{0, 0, 0, 0};
        mm16<C, 4>(acc, xd + I * BE, BS, slot(s), BS, bli, blk);
        if (keep) st16(acc, slot(s), BS, bli, blk, C(-1));
        if constexpr (CHOL) st16_global_rs(acc, Xg, I, J, bli, blk, C(-1), (const C*)(dsc + TS));
        else st16_global(acc, Xg, I, J, bli, blk, C(-1));
        if (cpx) st16(acc, xcopy + 16 * I * LDA + 16 * J, LDA, bli, blk, C(-1));
    };
    auto z_fin = [&](int s, int I, int J, bool keep) {
        acc_t acc = acc_t{0, 0, 0, 0};
        mm16<C, 4>(acc, slot(s), BS, zd + J * BE, BS, bli, blk);
        if (keep) st16(acc, slot(s), BS, bli, blk, C(-1));
        st16_global(acc, Zg, I, J, bli, blk, C(-1));
        if (cpz) st16(acc, zcopy + 16 * I * LDB + 16 * J, LDB, bli, blk, C(-1));
    };
    // 16x16 block (I, J) of a copy: from an LDS block of stride BS, or zeros
    auto cp_blk = [&](C* cp, const C* pb, int I, int J, int ld) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = 4 * q + (bl >> 4), c = bl & 15;
            cp[(16 * I + r) * ld + 16 * J + c] = pb ? pb[r * BS + c] : C(0);
        }
    };
    // block column K of L, block row K of U, the diagonal blocks of X and Z, and the zero
    // halves of the LDS copies of X and Z (the outputs' zero blocks, L and L^-1 above the
    // diagonal blocks and U and U^-1 below, are stored by waves 1-3 while wave 0 factors
    // diagonal block 0)
    auto l_col = [&](int K) {
        for (int I = K; I < 4; I++) l_blk_to_global<T, C, CHOL>(a, Lg, I, K, bl, dsc);
    };
    C um = 0;
    auto u_row = [&](int K) {
        for (int J = K; J < 4; J++) u_blk_to_global<T, C>(a, Ug, K, J, bl, um);
    };
    auto inv_diag = [&](int K) {
        if (Xg) {
            if constexpr (CHOL) blk_to_global_rs(xd + K * BE, BS, Xg, K, K, bl, (const C*)(dsc + TS));
            else blk_to_global(xd + K * BE, BS, Xg, K, K, bl);
            if (cpx) {
                cp_blk(xcopy, xd + K * BE, K, K, LDA);
                for (int J = K + 1; J < 4; J++) cp_blk(xcopy, nullptr, K, J, LDA);
            }
        }
        if (Zg) {
            blk_to_global(zd + K * BE, BS, Zg, K, K, bl);
            if (cpz) {
                cp_blk(zcopy, zd + K * BE, K, K, LDB);
                for (int I = K + 1; I < 4; I++) cp_blk(zcopy, nullptr, I, K, LDB);
            }
        }
    };
    {
        const T* A = pool + (size_t)src * TE;
        const int r = tid >> 2, c0 = (tid & 3) * 16;
        if (tid == 0) *pcnt = 0;
        C m = 0;
        if (in_lds) {  // the GEMM that made the source left it in a (CHOL: mirrored below)
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const C x = a[r * LDA + c0 + q];
                m = fabs(x) > m ? fabs(x) : m;
            }
        } else {
            C av[16];
            load16(A, r * TS + c0, coh, av);
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = c0 + q;
                const C x = av[q];
                a[r * LDA + c] = x;
                m = fabs(x) > m ? fabs(x) : m;
            }
        }
        if constexpr (CHOL) {  // only the lower triangle is read: mirror it
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = c0 + q;
                if (c > r) a[r * LDA + c] = a[c * LDA + r];
            }
        }
        if (amax_out) {  // the wave's max |a| to red[w]; max over the waves at the end
            const float wm = wave_absmax((float)m);
            if (lane == 0) red[w] = C(wm);
        }
        __syncthreads();
    }
    // diagonal block K on wave 0 (and, CHOL, the scales of its columns)
    auto diag_step = [&](int K) {
        const int r0 = 16 * K;
        if constexpr (sizeof(C) == 8) diag16_split<C, CHOL>(a, r0, xd + K * BE, zd + K * BE, lane);
        else diag16<C, CHOL>(a, r0, xd + K * BE, zd + K * BE, lane);
        if constexpr (CHOL) {
            if (lane < 16) {  // this wave stored the pivot of row r0 + lane (FP32: lane or lane + 32)
                const C d = sqrt(a[(r0 + lane) * LDA + r0 + lane]);
                dsc[r0 + lane] = d;
                dsc[TS + r0 + lane] = C(1) / d;
            }
        }
    };
    // panel blocks, in place: U_KJ = L_KK^-1 A_KJ and L_IK = A_IK U_KK^-1; trailing update
    // A_IJ -= L_IK U_KJ
    auto u_panel = [&](int K, int J, int ci, int ck) {
        acc_t acc = acc_t{0, 0, 0, 0};
        C* d = ablk(K, J);
        mm16<C, 4>(acc, xd + K * BE, BS, d, LDA, ci, ck);
        st16(acc, d, LDA, ci, ck);
    };
    auto l_panel = [&](int I, int K, int ci, int ck) {
        acc_t acc = acc_t{0, 0, 0, 0};
        C* d = ablk(I, K);
        mm16<C, 4>(acc, d, LDA, zd + K * BE, BS, ci, ck);
        st16(acc, d, LDA, ci, ck);
    };
    auto trail = [&](int I, int J, int K, int ci, int ck) {
        C* d = ablk(I, J);
        acc_t acc;
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] = d[MfmaOp<C>::row(ck, r) * LDA + ci];
        mm16<C, 4>(acc, ablk(I, K), LDA, ablk(K, J), LDA, ci, ck, C(-1));
        st16(acc, d, LDA, ci, ck);
    };
    // Panel blocks of step K written by all four waves: each wave adds 1 to the LDS
    // counter after its blocks; waves 1-3 wait for 4(K+1) before the trailing update.
    // Wave 0 never waits on it, so the wait always ends; the spin is bounded all the same.
    auto panel_done = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_fetch_add(pcnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto panel_wait = [&](int K) {
        if (lane == 0) {
            for (int spin = 0; spin < (1 << 22); spin++) {
                if (__hip_atomic_load(pcnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= 4 * (K + 1)) break;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    if (w == 0) {
        diag_step(0);
    } else {
        // the 24 zero blocks (6 per output: block (K, J), J > K, of L and L^-1 and its
        // transpose of U and U^-1), 8 per wave; the tile is in LDS, so an output slot that
        // is the source's as well is not read again
        for (int e = w - 1; e < 24; e += 3) {
            T* const o = e < 6 ? Lg : (e < 12 ? Xg : (e < 18 ? Ug : Zg));
            const int q = e % 6, K = q < 3 ? 0 : (q < 5 ? 1 : 2), J = q < 3 ? q + 1 : (q < 5 ? q - 1 : 3);
            if (o) zero_blk_global(o, e < 12 ? K : J, e < 12 ? J : K, lane);
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int K = 0; K < 3; K++) {
        const int nb = 3 - K;  // blocks right of / below diagonal block K
        if (w == 0) {
            // look-ahead: the two panel blocks and the update of diagonal block K+1 first,
            // then its elimination, while waves 1-3 finish step K
            u_panel(K, K + 1, li, lk);
            l_panel(K + 1, K, li, lk);
            panel_done();
            __builtin_amdgcn_wave_barrier();
            trail(K + 1, K + 1, K, li, lk);
            __builtin_amdgcn_wave_barrier();
            diag_step(K + 1);
            // FP32: the elimination ends before waves 1 and 2 finish their share of step
            // 0, so wave 0 forms N10 = L10 X00 from its own two blocks (and wave 3 M01)
            if (sizeof(C) == 4 && K == 0 && Xg) n_add(N10, true, 1, 0, xd);
        } else {
            asm volatile("" : "+v"(bl));
            bli = bl & 15;
            blk = bl >> 4;
            for (int p = w - 1; p < 2 * (nb - 1); p += 3) {  // the other panel blocks
                if (p < nb - 1) u_panel(K, K + 2 + p, bli, blk);
                else l_panel(K + 2 + p - (nb - 1), K, bli, blk);
            }
            panel_done();
            panel_wait(K);
            for (int q = w; q < nb * nb; q += 3) {  // the other trailing blocks (q = 0 is wave 0's)
                trail(K + 1 + q / nb, K + 1 + q % nb, K, bli, blk);
            }
            // block row / column K: stores and the inverse products it allows
            if (K == 0) {
                if (w == 1) {
                    l_col(0);
                    if (Xg) {
                        if (sizeof(C) != 4) n_add(N10, true, 1, 0, xd);
                        n_add(N20, true, 2, 0, xd);
                    }
                } else if (w == 2) {
                    if (Ug) u_row(0);
                    if (Zg) {
                        if (sizeof(C) != 4) m_add(M01, true, zd, 0, 1);
                        m_add(M02, true, zd, 0, 2);
                    }
                } else {  // the diagonal inverses' stores are the shortest share
                    inv_diag(0);
                    if (sizeof(C) == 4 && Zg) m_add(M01, true, zd, 0, 1);
                    if (Xg) n_add(N30, true, 3, 0, xd);
                    if (Zg) m_add(M03, true, zd, 0, 3);
                }
            } else if (K == 1) {
                if (w == 1) {
                    l_col(1);
                    if (Xg) {
                        x_fin(N10, 1, 0, true);  // X10
                        __builtin_amdgcn_wave_barrier();
                        n_add(N20, false, 2, 1, slot(N10));
                        n_add(N30, false, 3, 1, slot(N10));
                    }
                } else if (w == 2) {
                    if (Ug) u_row(1);
                    if (Zg) {
                        z_fin(M01, 0, 1, true);  // Z01
                        __builtin_amdgcn_wave_barrier();
                        m_add(M02, false, slot(M01), 1, 2);
                        m_add(M03, false, slot(M01), 1, 3);
                    }
                } else {
                    inv_diag(1);
                    if (Xg) {
                        n_add(N21, true, 2, 1, xd + BE);
                        n_add(N31, true, 3, 1, xd + BE);
                    }
                    if (Zg) {
                        m_add(M12, true, zd + BE, 1, 2);
                        m_add(M13, true, zd + BE, 1, 3);
                    }
                }
            } else {
                if (w == 1) {
                    l_col(2);
                    if (Xg) {
                        x_fin(N20, 2, 0, true);  // X20
                        x_fin(N21, 2, 1, true);  // X21
                        __builtin_amdgcn_wave_barrier();
                        n_add(N30, false, 3, 2, slot(N20));
                        n_add(N31, false, 3, 2, slot(N21));
                    }
                } else if (w == 2) {
                    if (Ug) u_row(2);
                    if (Zg) {
                        z_fin(M02, 0, 2, true);  // Z02
                        z_fin(M12, 1, 2, true);  // Z12
                        __builtin_amdgcn_wave_barrier();
                        m_add(M03, false, slot(M02), 2, 3);
                        m_add(M13, false, slot(M12), 2, 3);
                    }
                } else {
                    inv_diag(2);
                    if (Xg) n_add(N32, true, 3, 2, xd + 2 * BE);
                    if (Zg) m_add(M23, true, zd + 2 * BE, 2, 3);
                }
            }
        }
        __syncthreads();
    }
    // block row / column 3: waves 0-2 finish X_3w and Z_w3, wave 3 stores the rest
    if (w == 3) {
        l_col(3);
        if (Ug) u_row(3);
        inv_diag(3);
    } else {
        if (Xg) x_fin(w == 0 ? N30 : (w == 1 ? N31 : N32), 3, w, false);
        if (Zg) z_fin(w == 0 ? M03 : (w == 1 ? M13 : M23), w, 3, false);
    }
    if (amax_out || umax_out) {  // the growth check's max |a| and max |u| over the four waves
        const float wu = wave_absmax((float)um);
        if (lane == 0) red[4 + w] = C(wu);
        __syncthreads();
        if (amax_out) *amax_out = (float)fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        if (umax_out) *umax_out = (float)fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    }
}

// FP64 refactorisation of a low-precision tile whose pivot-free growth is too large.
// It runs rarely, so it uses a compact LDS-resident form (two barriers per step)
// rather than a second copy of the register-resident code; L and U stay in the
// one LDS tile, which the inverse routines read directly (strict lower part as L
// with unit diagonal, upper part as U).
template <typename T>
__device__ inline void lu_inv_fp64_lds(T* __restrict__ pool, int src, int outl, int outu, int outli,
                                       int outui, char* lds_raw, bool coh) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    double* a = reinterpret_cast<double*>(lds_raw);
    double* lcol = a + TS * LDA;
    const T* A = pool + (size_t)src * TE;
    const int i = tid >> 2, j0 = (tid & 3) * 16;
    {
        double av[16];
        load16(A, i * TS + j0, coh, av);
#pragma unroll
        for (int q = 0; q < 16; q++) a[i * LDA + j0 + q] = av[q];
    }
    __syncthreads();
    const double TINY2 = 1e-9;
#pragma unroll 1
    for (int k = 0; k < TS; k++) {
        double p = a[k * LDA + k];
        if (p < TINY2 && p > -TINY2) p = (p < 0) ? -TINY2 : TINY2;
        if (tid < TS && tid > k) lcol[tid] = a[tid * LDA + k] * (1.0 / p);
        __syncthreads();
        if (i > k) {
            const double l = lcol[i];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int j = j0 + q;
                if (j > k) a[i * LDA + j] -= l * a[k * LDA + j];
                else if (j == k) a[i * LDA + j] = l;
            }
        } else if (i == k && j0 <= k && k < j0 + 16) {
            a[k * LDA + k] = p;
        }
        __syncthreads();
    }
    T* L = pool + (size_t)outl * TE;
    T* U = pool + (size_t)outu * TE;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int j = j0 + q;
        const double v = a[i * LDA + j];
        gst(&L[i * TS + j], i > j ? T(v) : (i == j ? T(1) : T(0)));
        gst(&U[i * TS + j], j >= i ? T(v) : T(0));
    }
    double y[16];
    double* bro = lcol + TS + 16 * w;
    __syncthreads();
    if (outli >= 0) {
        inv_lower_regs<double, true>(a, y, bro, lane, w);
        store_rowseg(pool + (size_t)outli * TE, y, lane, w);
    }
    if (outui >= 0) {
        inv_upper_regs<double>(a, y, bro, lane, w);
        store_rowseg(pool + (size_t)outui * TE, y, lane, w);
    }
}

// The lu task: FP64 tiles factor in FP64; FP32 tiles factor in FP32 and a tile whose
// pivot-free growth max|U| / max|A| exceeds growth_limit is refactored from its input
// in FP64 (counted in n_fp64).  Returns the growth seen.
// xcopy / zcopy (FP32 tiles): copies of the inverses in LDS (see lu_inv_blocked); after
// an FP64 refactorisation they are stale (the growth returned exceeds growth_limit).
// in_lds: the source is already in LDS (row stride LDA), left there by the GEMM that made
// it (its global copy is still read by an FP64 refactorisation).
template <typename T>
__device__ __attribute__((always_inline)) double lu_task(T* __restrict__ pool, int src, int outl, int outu, int outli, int outui,
                          char* lds_raw, double growth_limit, int* n_fp64, bool coh = false,
                          float* xcopy = nullptr, float* zcopy = nullptr, bool in_lds = false) {
    if constexpr (sizeof(T) == 8) {
        lu_inv_blocked<T, double>(pool, src, outl, outu, outli, outui, lds_raw, nullptr, nullptr, coh, nullptr, nullptr,
                                  in_lds);
        return 1.0;
    } else {
        float amax = 0, umax = 0;
        lu_inv_blocked<T, float>(pool, src, outl, outu, outli, outui, lds_raw, &amax, &umax, coh, xcopy, zcopy, in_lds);
        const double growth = amax > 0 ? (double)umax / (double)amax : 1.0;
        if (!(growth <= growth_limit)) {
            __syncthreads();
            lu_inv_fp64_lds<T>(pool, src, outl, outu, outli, outui, lds_raw, coh);
            if (threadIdx.x == 0) atomicAdd(n_fp64, 1);
        }
        return growth;
    }
}

// Standalone inverse of a lower (general diagonal) or upper tile.
template <typename T, bool LOWER, bool COH = false>
__device__ __attribute__((always_inline)) void tri_inv_task(T* __restrict__ pool, int src, int out, char* lds_raw) {
    typedef typename std::conditional<sizeof(T) == 8, double, float>::type C;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    C* ts = reinterpret_cast<C*>(lds_raw);
    C* bro = ts + TS * LDA + 16 * w;
    C v[16];
    load16c<COH>(pool + (size_t)src * TE, lane * TS + 16 * w, v);
#pragma unroll
    for (int c = 0; c < 16; c++) ts[lane * LDA + 16 * w + c] = v[c];
    __syncthreads();
    if (LOWER) inv_lower_regs<C, false>(ts, v, bro, lane, w);
    else inv_upper_regs<C>(ts, v, bro, lane, w);
    store_rowseg(pool + (size_t)out * TE, v, lane, w);
}

// The llt task: Cholesky of one tile (its lower triangle) and, when outinv >= 0, L^-1,
// from the blocked elimination of the lu task in its Cholesky form (lu_inv_blocked,
// CHOL): the pivots are lltdcmpSimple's p_j (MatrixStdDouble.cpp:2629-2668), clamped
// to 1e-20 before the square root as it does (2612).  FP32 tiles factor in FP32.
template <typename T>
__device__ __attribute__((always_inline)) void llt_task(T* __restrict__ pool, int src, int out, int outinv,
                                                        char* lds_raw, bool coh = false, bool in_lds = false) {
    typedef typename std::conditional<sizeof(T) == 8, double, float>::type C;
    lu_inv_blocked<T, C, true>(pool, src, out, -1, outinv, -1, lds_raw, nullptr, nullptr, coh, nullptr, nullptr, in_lds);
}

// R = A^T through LDS (odd row stride: the column reads hit distinct banks); wave w
// leaves the mask bits of block row w of R in out_mask[w].
template <typename T, bool MASKS = true, bool COH = false>
__device__ __attribute__((always_inline)) void trans_task(T* __restrict__ pool, int src, int out, char* lds_raw,
                                                          uint32_t* out_mask) {
    T* s = reinterpret_cast<T*>(lds_raw);
    const int tid = threadIdx.x, lane = tid & 63;
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    tile_to_lds<COH>(pool + (size_t)src * TE, s, LDA, tid);
    __syncthreads();
    T* R = pool + (size_t)out * TE;
    bool any = false;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const T v = s[(c0 + q) * LDA + r];
        gst(&R[r * TS + c0 + q], v);
        any |= fabs(v) > T(MASK_TINY);
    }
    if constexpr (!MASKS) return;
    const unsigned long long b = __ballot(any);  // lane l: block (w, l % 4)
    uint32_t bits = 0;
#pragma unroll
    for (int J = 0; J < 4; J++)
        if (b & (0x1111111111111111ull << J)) bits |= 1u << (4 * (tid >> 6) + J);
    if (lane == 0) out_mask[tid >> 6] = bits;
}

}  // namespace sogl_dev
