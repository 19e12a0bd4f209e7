// Device-side 64x64 tile operations of the factor, executed by one
// 256-thread workgroup (4 wave64s) each.  Tiles are dense, row-major
// 64x64 in the HBM tile pool; element (i,j) sits at [i*64 + j].
//
//   gemm_task : C = sum_s sign_s * A_s * B_s           (mul / mulneg chains)
//   sub_task  : R = A - B, R = A, R = -B               (sub / copy / neg)
//   lu_task   : A = L U, pivot-free, |u_kk| < 1e-9 clamped to +-1e-9
//   linv_task : Y = L^-1 (general diagonal)
//   uinv_task : Y = U^-1
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

// LDS bytes one workgroup needs for any task kind, per storage type
template <typename T>
constexpr int lds_bytes() {
    // gemm: A (64 x LDA) + B (64 x LDB) in the MFMA operand type; the lu / inverses use
    // 2 x 64 x LDA doubles at most (the FP64 fallback of a low-precision factor)
    return (TS * LDA + TS * LDB) * 8 + 2 * TS * 8 + 8 * 8 + 64 * 8;
}

// ------------------------------------------------------------------ loads
// thread t moves row t/4, columns (t%4)*16 .. +15 of a tile between global and LDS
template <typename T>
__device__ inline void tile_to_lds(const T* __restrict__ g, T* __restrict__ s, int ld, int tid) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
    const T* src = g + r * TS + c0;
    T* dst = s + r * ld + c0;
    if constexpr (sizeof(T) == 8) {
        const double2* v = reinterpret_cast<const double2*>(src);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            double2 x = v[q];
            dst[2 * q] = x.x;
            dst[2 * q + 1] = x.y;
        }
    } else {
        const float4* v = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float4 x = v[q];
            dst[4 * q] = x.x;
            dst[4 * q + 1] = x.y;
            dst[4 * q + 2] = x.z;
            dst[4 * q + 3] = x.w;
        }
    }
}

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

// One (A,B) pair of a GEMM task.  b < 0 encodes a subtracted product (mulneg):
// slot = -b - 1.
struct Pair {
    int32_t a, b;
};

// C = sum_s sign_s A_s B_s over npairs pairs, full 64x64 output per workgroup.
// Wave w computes the 32x32 quadrant (w/2, w%2) as 2x2 blocks of 16x16 MFMA
// outputs; v_mfma_*_16x16x4 lane layout: A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16g with (g, r) given by MfmaOp<T>::row.
// A GEMM task may start before all of its pairs' operands exist: `wait(a, b)` is
// called by every thread before pair (a, b) is read and returns once both tiles
// are ready (it includes the barrier that precedes the operand loads).
struct NoWait {
    __device__ inline void operator()(int, int) const { __syncthreads(); }
};

template <typename T, bool ACC = false, typename Wait = NoWait>
__device__ __attribute__((always_inline)) void gemm_task(T* __restrict__ pool, const Pair* __restrict__ pairs, int npairs,
                          int out, char* lds_raw, const Wait& wait = Wait()) {
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
    for (int s = 0; s < npairs; s++) {
        const Pair p = pairs[s];
        const bool neg = p.b < 0;
        const int sb = neg ? -p.b - 1 : p.b;
        wait(p.a, sb);
        tile_to_lds(pool + (size_t)p.a * TE, la, LDA, tid);
        tile_to_lds(pool + (size_t)sb * TE, lb, LDB, tid);
        __syncthreads();
        const T sg = neg ? T(-1) : T(1);
#pragma unroll 4
        for (int kk = 0; kk < TS / 4; kk++) {
            const int k = 4 * kk + lk;
            T af[2], bf[2];
#pragma unroll
            for (int m = 0; m < 2; m++) af[m] = sg * la[(wr + m * 16 + li) * LDA + k];
#pragma unroll
            for (int n = 0; n < 2; n++) bf[n] = lb[k * LDB + wc + n * 16 + li];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++) acc[m][n] = MfmaOp<T>::mma(af[m], bf[n], acc[m][n]);
        }
    }
    T* C = pool + (size_t)out * TE;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                C[(wr + m * 16 + MfmaOp<T>::row(lk, r)) * TS + wc + n * 16 + li] = acc[m][n][r];
}

// bf16 experiment: FP32 tiles, operands rounded to bf16, FP32 accumulation
// (v_mfma_f32_16x16x32_bf16: A[i][8q..8q+7] in lane i + 16q, same D layout).
template <typename Wait = NoWait>
__device__ __attribute__((always_inline)) void gemm_task_bf16(float* __restrict__ pool, const Pair* __restrict__ pairs,
                                      int npairs, int out, char* lds_raw, const Wait& wait = Wait()) {
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
        const bool neg = p.b < 0;
        const int sb = neg ? -p.b - 1 : p.b;
        wait(p.a, sb);
        tile_to_lds(pool + (size_t)p.a * TE, la, LDA, tid);
        tile_to_lds(pool + (size_t)sb * TE, lb, LDB, tid);
        __syncthreads();
        const float sg = neg ? -1.f : 1.f;
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
                    bf[n][e] = (__bf16)lb[(kk * 32 + lk * 8 + e) * LDB + wc + n * 16 + li];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int n = 0; n < 2; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bf[n], acc[m][n], 0, 0, 0);
        }
    }
    float* C = pool + (size_t)out * TE;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int n = 0; n < 2; n++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                C[(wr + m * 16 + 4 * lk + r) * TS + wc + n * 16 + li] = acc[m][n][r];
}

// ------------------------------------------------------------------ sub
// sub(src, src2 -> r): r = tile(src2) - tile(src); a missing operand (slot < 0)
// reads as zero, which gives copy and negate (BlockPlanner.cpp:553-572).  All
// loads are issued before the stores (the operands may not alias the result,
// but the compiler cannot know that).
template <typename T>
__device__ __attribute__((always_inline)) void sub_task(T* __restrict__ pool, int src, int src2, int out) {
    const int tid = threadIdx.x;
    constexpr int Q = TE / NT;  // 16 elements per thread
    T a[Q], b[Q];
    const T* A = src2 >= 0 ? pool + (size_t)src2 * TE : nullptr;
    const T* B = src >= 0 ? pool + (size_t)src * TE : nullptr;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        a[q] = A ? A[q * NT + tid] : T(0);
        b[q] = B ? B[q * NT + tid] : T(0);
    }
    T* R = pool + (size_t)out * TE;
#pragma unroll
    for (int q = 0; q < Q; q++) R[q * NT + tid] = a[q] - b[q];
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
__device__ inline void load_rowseg(const T* g, C* v, int lane, int w) {
    const T* p = g + lane * TS + 16 * w;
#pragma unroll
    for (int c = 0; c < 16; c++) v[c] = C(p[c]);
}

template <typename T, typename C>
__device__ inline void store_rowseg(T* g, const C* v, int lane, int w) {
    T* p = g + lane * TS + 16 * w;
#pragma unroll
    for (int c = 0; c < 16; c++) p[c] = T(v[c]);
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

// Pivot-free right-looking LU of the tile held in v (ludcmpSimple,
// MatrixStdDouble.cpp:2711-2784): afterwards v holds U on and above the diagonal
// and the multipliers of L below it.  |u_kk| < 1e-9 becomes -1e-9 if u_kk < 0,
// else +1e-9 (MatrixStdDouble.cpp:2751-2761).  One barrier per step (lcol, a
// double buffer of 2 x 64 multipliers in LDS); the pivot row reaches the wave's
// lanes through its bro buffer.  The 16 steps of a column panel are unrolled so
// the pivot column's register index is a constant; the 4 panels loop.
template <typename C>
__device__ inline void lu_regs(C* v, C* lcol, C* bro, int lane, int w) {
    const C TINY2 = C(1e-9);
#pragma unroll 1
    for (int kb = 0; kb < 4; kb++) {
#pragma unroll
        for (int kc = 0; kc < 16; kc++) {
            const int k = kb * 16 + kc;
            C* lc = lcol + (kc & 1) * TS;
            C u[16];
            bcast_row(v, bro, lane, k, u);
            if (w == kb) {  // the wave owning column k makes the multipliers
                C p = u[kc];
                if (p < TINY2 && p > -TINY2) p = (p < 0) ? -TINY2 : TINY2;
                const C l = v[kc] * (C(1) / p);
                v[kc] = lane > k ? l : (lane == k ? p : v[kc]);
                lc[lane] = l;  // read only by the lanes below the pivot
            }
            __syncthreads();
            const C l = lc[lane];
            const bool below = lane > k;
#pragma unroll
            for (int c = 0; c < 16; c++)  // branch-free: columns <= k keep their value
                v[c] = (below && 16 * w + c > k) ? v[c] - l * u[c] : v[c];
        }
    }
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

// block-wide max of |x|; red has >= 4 slots
template <typename C>
__device__ inline C block_absmax(C x, C* red, int tid) {
    x = x < 0 ? -x : x;
    for (int o = 32; o > 0; o >>= 1) {
        C y = __shfl_xor(x, o);
        x = y > x ? y : x;
    }
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = x;
    __syncthreads();
    C m = red[0];
    for (int q = 1; q < NT / 64; q++) m = red[q] > m ? red[q] : m;
    return m;
}

// LU of one tile in compute type C plus, when their slots are given (>= 0), the
// inverses of its unit-lower L and upper U: the reference's lowerInv / upperInv
// of the same diagonal tile, fused into the producing task.
template <typename T, typename C>
__device__ __attribute__((always_inline)) void lu_inv_core(T* __restrict__ pool, int src, int outl, int outu, int outli, int outui,
                            char* lds_raw, float* amax_out, float* umax_out) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    C* ls = reinterpret_cast<C*>(lds_raw);       // L (strict lower) then U, row stride LDA
    C* us = ls + TS * LDA;
    C* lcol = us + TS * LDA;                      // 2 x 64 multipliers
    C* red = lcol + 2 * TS;                       // 8 reduction slots
    C* bro = red + 8 + 16 * w;                    // 16 per wave: row broadcasts
    C v[16];
    load_rowseg(pool + (size_t)src * TE, v, lane, w);
    if (amax_out) {
        C m = 0;
#pragma unroll
        for (int c = 0; c < 16; c++) m = fabs(v[c]) > m ? fabs(v[c]) : m;
        *amax_out = (float)block_absmax(m, red, tid);
    }
    lu_regs(v, lcol, bro, lane, w);
    // L (unit lower) and U out of the same registers, to HBM and, for the inverses, to LDS
    const bool inv = outli >= 0 || outui >= 0;
    T* Lg = pool + (size_t)outl * TE + lane * TS + 16 * w;
    T* Ug = pool + (size_t)outu * TE + lane * TS + 16 * w;
    C um = 0;
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const int j = 16 * w + c;
        const C l = lane > j ? v[c] : (lane == j ? C(1) : C(0));
        const C u = j >= lane ? v[c] : C(0);
        Lg[c] = T(l);
        Ug[c] = T(u);
        um = fabs(u) > um ? fabs(u) : um;
        if (inv) {
            ls[lane * LDA + j] = l;
            us[lane * LDA + j] = u;
        }
    }
    if (umax_out) *umax_out = (float)block_absmax(um, red + 4, tid);
    if (!inv) return;
    __syncthreads();
    C y[16];
    if (outli >= 0) {
        inv_lower_regs<C, true>(ls, y, bro, lane, w);
        store_rowseg(pool + (size_t)outli * TE, y, lane, w);
    }
    if (outui >= 0) {
        inv_upper_regs<C>(us, y, bro, lane, w);
        store_rowseg(pool + (size_t)outui * TE, y, lane, w);
    }
}

// FP64 refactorisation of a low-precision tile whose pivot-free growth is too large.
// It runs rarely, so it uses a compact LDS-resident form (two barriers per step)
// rather than a second copy of the register-resident code; L and U stay in the
// one LDS tile, which the inverse routines read directly (strict lower part as L
// with unit diagonal, upper part as U).
template <typename T>
__device__ inline void lu_inv_fp64_lds(T* __restrict__ pool, int src, int outl, int outu, int outli,
                                       int outui, char* lds_raw) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    double* a = reinterpret_cast<double*>(lds_raw);
    double* lcol = a + TS * LDA;
    const T* A = pool + (size_t)src * TE;
    const int i = tid >> 2, j0 = (tid & 3) * 16;
#pragma unroll
    for (int q = 0; q < 16; q++) a[i * LDA + j0 + q] = (double)A[i * TS + j0 + q];
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
        L[i * TS + j] = i > j ? T(v) : (i == j ? T(1) : T(0));
        U[i * TS + j] = j >= i ? T(v) : T(0);
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
template <typename T>
__device__ __attribute__((always_inline)) double lu_task(T* __restrict__ pool, int src, int outl, int outu, int outli, int outui,
                          char* lds_raw, double growth_limit, int* n_fp64) {
    if constexpr (sizeof(T) == 8) {
        lu_inv_core<T, double>(pool, src, outl, outu, outli, outui, lds_raw, nullptr, nullptr);
        return 1.0;
    } else {
        float amax = 0, umax = 0;
        lu_inv_core<T, float>(pool, src, outl, outu, outli, outui, lds_raw, &amax, &umax);
        const double growth = amax > 0 ? (double)umax / (double)amax : 1.0;
        if (!(growth <= growth_limit)) {
            __syncthreads();
            lu_inv_fp64_lds<T>(pool, src, outl, outu, outli, outui, lds_raw);
            if (threadIdx.x == 0) atomicAdd(n_fp64, 1);
        }
        return growth;
    }
}

// Standalone inverse of a lower (general diagonal) or upper tile.
template <typename T, bool LOWER>
__device__ __attribute__((always_inline)) void tri_inv_task(T* __restrict__ pool, int src, int out, char* lds_raw) {
    typedef typename std::conditional<sizeof(T) == 8, double, float>::type C;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    C* ts = reinterpret_cast<C*>(lds_raw);
    C* bro = ts + TS * LDA + 16 * w;
    C v[16];
    load_rowseg(pool + (size_t)src * TE, v, lane, w);
#pragma unroll
    for (int c = 0; c < 16; c++) ts[lane * LDA + 16 * w + c] = v[c];
    __syncthreads();
    if (LOWER) inv_lower_regs<C, false>(ts, v, bro, lane, w);
    else inv_upper_regs<C>(ts, v, bro, lane, w);
    store_rowseg(pool + (size_t)out * TE, v, lane, w);
}

}  // namespace sogl_dev
