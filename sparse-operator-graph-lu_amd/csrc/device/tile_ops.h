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
    return (TS * LDA + TS * LDB) * 8 + 2 * TS * 8 + 64;
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
__device__ void gemm_task(T* __restrict__ pool, const Pair* __restrict__ pairs, int npairs,
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
__device__ inline void gemm_task_bf16(float* __restrict__ pool, const Pair* __restrict__ pairs,
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
// reads as zero, which gives copy and negate (BlockPlanner.cpp:553-572).
template <typename T>
__device__ void sub_task(T* __restrict__ pool, int src, int src2, int out) {
    const int tid = threadIdx.x;
    const T* A = src2 >= 0 ? pool + (size_t)src2 * TE : nullptr;
    const T* B = src >= 0 ? pool + (size_t)src * TE : nullptr;
    T* R = pool + (size_t)out * TE;
#pragma unroll
    for (int q = 0; q < TE / NT; q++) {
        const int e = q * NT + tid;
        T a = A ? A[e] : T(0);
        T b = B ? B[e] : T(0);
        R[e] = a - b;
    }
}

// ------------------------------------------------------------------ LU
// Pivot-free tile LU (ludcmpSimple, MatrixStdDouble.cpp:2711-2784): L unit lower,
// U upper; a pivot with |u_kk| < 1e-9 becomes -1e-9 if negative, else +1e-9
// (MatrixStdDouble.cpp:2751-2761).  Right-looking in LDS, compute type C.
// Returns (through lds) max|A| and max|U| for the growth check.
template <typename C>
__device__ void lu_core(C* a, C* lcol, int tid) {
    const C TINY2 = C(1e-9);
    const int i = tid >> 2, j0 = (tid & 3) * 16;
    for (int k = 0; k < TS; k++) {
        C p = a[k * LDA + k];
        if (p < TINY2 && p > -TINY2) p = (p < 0) ? -TINY2 : TINY2;
        const C rp = C(1) / p;
        if (tid < TS && tid > k) lcol[tid] = a[tid * LDA + k] * rp;
        __syncthreads();
        if (i > k) {
            const C l = lcol[i];
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
}

template <typename T, typename C>
__device__ void lu_load(const T* A, C* a, int tid) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
#pragma unroll
    for (int q = 0; q < 16; q++) a[r * LDA + c0 + q] = C(A[r * TS + c0 + q]);
}

template <typename T, typename C>
__device__ void lu_store(const C* a, T* L, T* U, int tid) {
    const int r = tid >> 2, c0 = (tid & 3) * 16;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int c = c0 + q;
        const C v = a[r * LDA + c];
        L[r * TS + c] = r > c ? T(v) : (r == c ? T(1) : T(0));
        U[r * TS + c] = c >= r ? T(v) : T(0);
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

// growth = max|U| / max|A|; a low-precision tile above `growth_limit` is refactored
// in FP64 from its (low-precision) input.  Returns the growth seen.
template <typename T>
__device__ double lu_task(T* __restrict__ pool, int src, int outl, int outu, char* lds_raw,
                          double growth_limit, int* n_fp64) {
    const int tid = threadIdx.x;
    const T* A = pool + (size_t)src * TE;
    T* L = pool + (size_t)outl * TE;
    T* U = pool + (size_t)outu * TE;
    if constexpr (sizeof(T) == 8) {
        double* a = reinterpret_cast<double*>(lds_raw);
        double* lcol = a + TS * LDA;
        lu_load(A, a, tid);
        __syncthreads();
        lu_core(a, lcol, tid);
        lu_store(a, L, U, tid);
        return 1.0;
    } else {
        float* a = reinterpret_cast<float*>(lds_raw);
        float* lcol = a + TS * LDA;
        float* red = lcol + TS;
        const int r = tid >> 2, c0 = (tid & 3) * 16;
        float mx = 0;
        lu_load(A, a, tid);
#pragma unroll
        for (int q = 0; q < 16; q++) {
            float v = a[r * LDA + c0 + q];
            v = v < 0 ? -v : v;
            mx = v > mx ? v : mx;
        }
        const float amax = block_absmax(mx, red, tid);
        lu_core(a, lcol, tid);
        mx = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (c0 + q >= r) {
                float v = a[r * LDA + c0 + q];
                v = v < 0 ? -v : v;
                mx = v > mx ? v : mx;
            }
        }
        const float umax = block_absmax(mx, red + 8, tid);
        const double growth = amax > 0 ? (double)umax / (double)amax : 1.0;
        if (!(growth <= growth_limit)) {
            __syncthreads();
            double* ad = reinterpret_cast<double*>(lds_raw);
            double* lc = ad + TS * LDA;
            lu_load(A, ad, tid);
            __syncthreads();
            lu_core(ad, lc, tid);
            lu_store(ad, L, U, tid);
            if (tid == 0) atomicAdd(n_fp64, 1);
        } else {
            lu_store(a, L, U, tid);
        }
        return growth;
    }
}

// ------------------------------------------------------------------ inverses
// Y = L^-1 for lower L with general diagonal (inv_lower, MatrixStdDouble.cpp:2787-2802:
// y_ii = 1/l_ii, y_ij = -(sum_k l_ik y_kj) / l_ii): right-looking over rows of Y.
template <typename T, bool LOWER>
__device__ void tri_inv_task(T* __restrict__ pool, int src, int out, char* lds_raw) {
    typedef typename std::conditional<sizeof(T) == 8, double, float>::type C;
    C* l = reinterpret_cast<C*>(lds_raw);
    C* y = l + TS * LDA;
    const int tid = threadIdx.x;
    const T* Lg = pool + (size_t)src * TE;
    const int i = tid >> 2, j0 = (tid & 3) * 16;
    lu_load(Lg, l, tid);
#pragma unroll
    for (int q = 0; q < 16; q++) y[i * LDA + j0 + q] = (i == j0 + q) ? C(1) : C(0);
    __syncthreads();
    for (int s = 0; s < TS; s++) {
        const int k = LOWER ? s : TS - 1 - s;
        // row k of Y is final once scaled: y_k = r_k / t_kk
        if (tid < TS) {
            const int j = tid;
            if (LOWER ? (j <= k) : (j >= k)) y[k * LDA + j] = y[k * LDA + j] * (C(1) / l[k * LDA + k]);
        }
        __syncthreads();
        // eliminate column k from the remaining rows: r_i -= t_ik y_k
        if (LOWER ? (i > k) : (i < k)) {
            const C f = l[i * LDA + k];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int j = j0 + q;
                if (LOWER ? (j <= k) : (j >= k)) y[i * LDA + j] -= f * y[k * LDA + j];
            }
        }
        __syncthreads();
    }
    T* Y = pool + (size_t)out * TE;
#pragma unroll
    for (int q = 0; q < 16; q++) Y[i * TS + j0 + q] = T(y[i * LDA + j0 + q]);
}

}  // namespace sogl_dev
