// Mixed precision: an FP32 (or bf16-GEMM) factor is corrected by FP64
// iterative refinement against the original matrix,
//     x_0 = solve(b);  repeat  r = b - A x (FP64 CSR SpMV);  x += solve(r)
// until the normwise backward error is <= tol in both of its usual forms,
//     ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf)  and
//     ||r||_2 / (||A||_F ||x||_2 + ||b||_2).
// (The second alone is about sqrt(n) looser at large n, since ||A||_F grows with
// sqrt(n) where ||A||_inf, a row sum, does not.)  FP64 lives only in the
// residual, the correction accumulation and the solve's vector sums; the
// factor tiles stay in their low precision.  The reference has no
// counterpart (it factors in FP64 only); its residual check is the max-abs
// mtx::checkResult (mtx.cpp:183-215).
//
// Also the drop-in sogl_solveLU, which mirrors SOGLU::solveLU (solver.h:24).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "device.h"
#include "tile_ops.h"

namespace sogl_dev {
int solve_device(Context* c, Factor* f, const double* d_b, double* d_x, double* d_y, int nrhs);
int ensure_vectors(Context* c, Factor* f, int nrhs);
}  // namespace sogl_dev

using namespace sogl_dev;

namespace {

#define HIP_TRY(ctx, expr)                                       \
    do {                                                         \
        int _rc = check_hip((ctx), (expr), #expr);               \
        if (_rc != SOGL_OK) return _rc;                          \
    } while (0)

constexpr int RED_BLOCKS = 256;

// r = b - A x on the permuted padded system, one row per thread, every rhs column
__global__ void k_residual(const int64_t* ptr, const int32_t* col, const double* val,
                           const double* b, const double* x, double* r, int64_t N, int nrhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int c = 0; c < nrhs; c++) {
        const double* xc = x + (int64_t)c * N;
        double s = 0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) s = fma(val[e], xc[col[e]], s);
        r[i + (int64_t)c * N] = b[i + (int64_t)c * N] - s;
    }
}

// bp[i] = b[new_order[i]] for i < n, 1 on the padding rows (solver.cpp:155-161 + GPSOrder.cpp:433-447)
__global__ void k_permute_in(const double* b, const int32_t* new_order, double* bp, int32_t n, int64_t N,
                             int nrhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int c = 0; c < nrhs; c++) bp[i + c * N] = i < n ? b[new_order[i] + (int64_t)c * n] : 1.0;
}

// x[new_order[i]] = xp[i] (GOrder::reOrderResult, GPSOrder.cpp:41-53)
__global__ void k_permute_out(const double* xp, const int32_t* new_order, double* x, int32_t n, int64_t N,
                              int nrhs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < nrhs; c++) x[new_order[i] + (int64_t)c * n] = xp[i + c * N];
}

__global__ void k_axpy(double* x, const double* d, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// Per column and block, over the first n rows: the sums of squares of r and x and
// their largest magnitudes, at out[(c*4 + {0,1,2,3}) * RED_BLOCKS + block].  A NaN
// reaches the maxima through the sums (fmax drops NaN; the host checks the sums).
constexpr int RED_VALUES = 4;
__global__ void k_norms(const double* r, const double* x, int64_t n, int64_t N, double* out) {
    __shared__ double sr[256], sx[256], mr[256], mx[256];
    const int c = blockIdx.y;
    double ar = 0, ax = 0, br = 0, bx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = r[i + c * N], b = x[i + c * N];
        ar += a * a;
        ax += b * b;
        br = fmax(br, fabs(a));
        bx = fmax(bx, fabs(b));
    }
    sr[threadIdx.x] = ar;
    sx[threadIdx.x] = ax;
    mr[threadIdx.x] = br;
    mx[threadIdx.x] = bx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            sx[threadIdx.x] += sx[threadIdx.x + o];
            mr[threadIdx.x] = fmax(mr[threadIdx.x], mr[threadIdx.x + o]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[(c * RED_VALUES + 0) * RED_BLOCKS + blockIdx.x] = sr[0];
        out[(c * RED_VALUES + 1) * RED_BLOCKS + blockIdx.x] = sx[0];
        out[(c * RED_VALUES + 2) * RED_BLOCKS + blockIdx.x] = mr[0];
        out[(c * RED_VALUES + 3) * RED_BLOCKS + blockIdx.x] = mx[0];
    }
}

// Column c's reduced norms from the per-block partials: {||r||_2^2, ||x||_2^2,
// ||r||_inf, ||x||_inf}.
void reduce_norms(const std::vector<double>& part, int c, double out[RED_VALUES]) {
    for (int v = 0; v < RED_VALUES; v++) {
        double a = 0;
        for (int q = 0; q < RED_BLOCKS; q++) {
            const double p = part[(size_t)((c * RED_VALUES + v) * RED_BLOCKS + q)];
            a = v < 2 ? a + p : std::max(a, p);
        }
        out[v] = a;
    }
}

}  // namespace

namespace sogl {
struct PlanView {
    int32_t n;
    int64_t N;
    const int32_t* new_order;
    int64_t nnz;
    const int32_t* row;
    const int32_t* col;
    const double* val;
    int64_t nnz_orig;  // entries of the original matrix (the rest is padding)
    int64_t n_inputs;  // input tiles: block coordinates and cells (position row*64+col, value)
    const int32_t* in_bi;
    const int32_t* in_bj;
    const int64_t* cell_begin;
    const uint16_t* cell_pos;
    const double* cell_val;
};
bool plan_view(const sogl_plan* p, PlanView* v);
void set_last_error(const std::string& s);
}  // namespace sogl

extern "C" int sogl_refine_attach(sogl_ctx* x, int32_t factor_id, const sogl_plan* plan) {
    if (!x || !plan) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    sogl::PlanView v;
    if (!sogl::plan_view(plan, &v)) return set_error(c, SOGL_ERR_ARG, "bad plan");
    if (v.N != (int64_t)f->block_rows * TS) return set_error(c, SOGL_ERR_ARG, "plan does not match factor");
    HIP_TRY(c, hipSetDevice(c->device));
    // CSR of the permuted padded matrix, taken from the input tiles' cells so that the
    // residual uses the matrix the factor holds (where the COO repeats a position, the
    // tiles keep one value; the rows >= n hold the padding diagonal only)
    const int64_t nc = v.n_inputs ? v.cell_begin[v.n_inputs] : 0;
    std::vector<int64_t> ptr((size_t)v.N + 1, 0);
    for (int64_t t = 0; t < v.n_inputs; t++)
        for (int64_t q = v.cell_begin[t]; q < v.cell_begin[t + 1]; q++)
            ptr[(size_t)((int64_t)v.in_bi[t] * TS + (v.cell_pos[q] >> 6)) + 1]++;
    for (int64_t i = 0; i < v.N; i++) ptr[(size_t)i + 1] += ptr[(size_t)i];
    std::vector<int32_t> col((size_t)nc);
    std::vector<double> val((size_t)nc);
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t t = 0; t < v.n_inputs; t++)
        for (int64_t q = v.cell_begin[t]; q < v.cell_begin[t + 1]; q++) {
            const int64_t r = (int64_t)v.in_bi[t] * TS + (v.cell_pos[q] >> 6);
            const int64_t p = pos[(size_t)r]++;
            col[(size_t)p] = v.in_bj[t] * TS + (v.cell_pos[q] & 63);
            val[(size_t)p] = v.cell_val[q];
        }
    // each row by column, and the Frobenius and infinity norms over the rows of the
    // original matrix
    double fro = 0, inf = 0;
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t i = 0; i < v.N; i++) {
        const int64_t b = ptr[(size_t)i], e = ptr[(size_t)i + 1];
        row.clear();
        for (int64_t p = b; p < e; p++) row.emplace_back(col[(size_t)p], val[(size_t)p]);
        std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        double rsum = 0;
        for (int64_t p = b; p < e; p++) {
            col[(size_t)p] = row[(size_t)(p - b)].first;
            val[(size_t)p] = row[(size_t)(p - b)].second;
            if (i < v.n) {
                fro += val[(size_t)p] * val[(size_t)p];
                rsum += std::fabs(val[(size_t)p]);
            }
        }
        inf = std::max(inf, rsum);
    }
    int rc;
    if ((rc = dev_alloc(c, f->csr_ptr, ptr.size() * 8, "csr ptr")) ||
        (rc = dev_alloc(c, f->csr_col, std::max<size_t>(col.size(), 1) * 4, "csr col")) ||
        (rc = dev_alloc(c, f->csr_val, std::max<size_t>(val.size(), 1) * 8, "csr val")))
        return rc;
    HIP_TRY(c, hipMemcpy(f->csr_ptr.p, ptr.data(), ptr.size() * 8, hipMemcpyHostToDevice));
    if (!col.empty()) {
        HIP_TRY(c, hipMemcpy(f->csr_col.p, col.data(), col.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(f->csr_val.p, val.data(), val.size() * 8, hipMemcpyHostToDevice));
    }
    f->n = v.n;
    f->nnz = nc;
    f->new_order.assign(v.new_order, v.new_order + v.n);
    if ((rc = dev_alloc(c, f->perm_d, (size_t)v.n * 4, "permutation"))) return rc;
    HIP_TRY(c, hipMemcpy(f->perm_d.p, v.new_order, (size_t)v.n * 4, hipMemcpyHostToDevice));
    f->a_fro = std::sqrt(fro);
    f->a_inf = inf;
    f->has_refine = true;
    return SOGL_OK;
}

static int refined_core(Context* c, Factor* f, const double* d_b, double* d_x, int32_t nrhs, double tol,
                        int32_t max_iter, int32_t* iters, double* resid) {
    const int32_t n = f->n;
    const int64_t N = (int64_t)f->block_rows * TS;
    int rc;
    if ((rc = ensure_vectors(c, f, nrhs))) return rc;
    if ((rc = dev_alloc(c, f->vec_r, (size_t)N * nrhs * 8, "residual")) ||
        (rc = dev_alloc(c, f->vec_d, (size_t)N * nrhs * 8, "correction")) ||
        (rc = dev_alloc(c, f->red, (size_t)nrhs * RED_VALUES * RED_BLOCKS * 8, "norms")))
        return rc;
    double* db = (double*)f->vec_b.p;
    double* dx = (double*)f->vec_x.p;
    double* dr = (double*)f->vec_r.p;
    double* dd = (double*)f->vec_d.p;
    const unsigned gN = (unsigned)((N + 255) / 256);
    const unsigned gA = (unsigned)((N * nrhs + 255) / 256);
    hipLaunchKernelGGL(k_permute_in, dim3(gN), dim3(256), 0, c->stream, d_b, (const int32_t*)f->perm_d.p, db, n, N,
                       nrhs);
    HIP_TRY(c, hipGetLastError());
    std::vector<double> part((size_t)nrhs * RED_VALUES * RED_BLOCKS);
    // ||b||_2 and ||b||_inf over the real rows, once
    hipLaunchKernelGGL(k_norms, dim3(RED_BLOCKS, nrhs), dim3(256), 0, c->stream, db, db, (int64_t)n, N,
                       (double*)f->red.p);
    HIP_TRY(c, hipMemcpyAsync(part.data(), f->red.p, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<double> bnorm((size_t)nrhs, 0.0), bmax((size_t)nrhs, 0.0);
    for (int32_t r = 0; r < nrhs; r++) {
        double v[RED_VALUES];
        reduce_norms(part, r, v);
        bnorm[(size_t)r] = std::sqrt(v[0]);
        bmax[(size_t)r] = v[2];
    }
    if ((rc = solve_device(c, f, db, dx, (double*)f->vec_y.p, nrhs))) return rc;
    double solve_ms = f->solve_ms;
    int it = 0;
    double worst = 0;
    for (;;) {
        hipLaunchKernelGGL(k_residual, dim3(gN), dim3(256), 0, c->stream, (const int64_t*)f->csr_ptr.p,
                           (const int32_t*)f->csr_col.p, (const double*)f->csr_val.p, db, dx, dr, N, nrhs);
        hipLaunchKernelGGL(k_norms, dim3(RED_BLOCKS, nrhs), dim3(256), 0, c->stream, dr, dx, (int64_t)n, N,
                           (double*)f->red.p);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(part.data(), f->red.p, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        worst = 0;
        for (int32_t r = 0; r < nrhs; r++) {
            double v[RED_VALUES];
            reduce_norms(part, r, v);
            const double den2 = f->a_fro * std::sqrt(v[1]) + bnorm[(size_t)r];
            const double rel2 = den2 > 0 ? std::sqrt(v[0]) / den2 : std::sqrt(v[0]);
            const double deni = f->a_inf * v[3] + bmax[(size_t)r];
            const double reli = deni > 0 ? v[2] / deni : v[2];
            const double rel = std::isnan(rel2) ? rel2 : std::max(rel2, reli);
            if (!(rel <= worst)) worst = rel;  // NaN propagates
        }
        if (worst <= tol || it >= max_iter || std::isnan(worst)) break;
        if ((rc = solve_device(c, f, dr, dd, (double*)f->vec_y.p, nrhs))) return rc;
        solve_ms += f->solve_ms;
        hipLaunchKernelGGL(k_axpy, dim3(gA), dim3(256), 0, c->stream, dx, (const double*)dd, N * nrhs);
        HIP_TRY(c, hipGetLastError());
        it++;
    }
    hipLaunchKernelGGL(k_permute_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (const double*)dx,
                       (const int32_t*)f->perm_d.p, d_x, n, N, nrhs);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    f->solve_ms = solve_ms;
    if (iters) *iters = it;
    if (resid) *resid = worst;
    if (std::isnan(worst)) return set_error(c, SOGL_ERR_NAN, "refinement: NaN in the residual");
    if (!(worst <= tol)) return set_error(c, SOGL_ERR_NOT_CONVERGED, "refinement did not reach the tolerance");
    return SOGL_OK;
}

extern "C" int sogl_solve_refined_device(sogl_ctx* x, int32_t factor_id, const double* d_b, double* d_x,
                                         int32_t nrhs, double tol, int32_t max_iter, int32_t* iters,
                                         double* resid) {
    if (!x || !d_b || !d_x || nrhs <= 0) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "bad argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (!f->has_refine) return set_error(c, SOGL_ERR_ARG, "sogl_refine_attach was not called");
    HIP_TRY(c, hipSetDevice(c->device));
    return refined_core(c, f, d_b, d_x, nrhs, tol, max_iter, iters, resid);
}

extern "C" int sogl_solve_refined(sogl_ctx* x, int32_t factor_id, const double* b, double* xo,
                                  int32_t nrhs, double tol, int32_t max_iter, int32_t* iters,
                                  double* resid) {
    if (!x || !b || !xo || nrhs <= 0) return set_error(x ? &x->c : nullptr, SOGL_ERR_ARG, "bad argument");
    Context* c = &x->c;
    std::lock_guard<std::mutex> g(c->mu);
    Factor* f = find_factor(c, factor_id);
    if (!f) return set_error(c, SOGL_ERR_ARG, "unknown factor id");
    if (!f->has_refine) return set_error(c, SOGL_ERR_ARG, "sogl_refine_attach was not called");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)f->n * nrhs * 8;
    int rc;
    if ((rc = dev_alloc(c, f->vec_io, 2 * bytes, "refine io"))) return rc;
    double* db = (double*)f->vec_io.p;
    double* dx = db + (size_t)f->n * nrhs;
    HIP_TRY(c, hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, c->stream));
    int rc2 = refined_core(c, f, db, dx, nrhs, tol, max_iter, iters, resid);
    if (rc2 != SOGL_OK && rc2 != SOGL_ERR_NOT_CONVERGED && rc2 != SOGL_ERR_NAN) return rc2;
    HIP_TRY(c, hipMemcpyAsync(xo, dx, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return rc2;
}

// ------------------------------------------------------------------ drop-in solveLU
extern "C" int sogl_factor_load_plan(sogl_ctx* x, const sogl_plan* plan, int32_t precision,
                                     int32_t* factor_id);

extern "C" int sogl_solveLU(int32_t dim, int64_t valcount, int32_t symmetric, const int32_t* index_i,
                            const int32_t* index_j, const double* vals, const double* b, double* xo) {
    static std::mutex mu;
    static sogl_ctx* ctx = nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (!xo || !b) return set_error(nullptr, SOGL_ERR_ARG, "null b or x");
    sogl_plan* plan = nullptr;
    int rc = sogl_plan_create(dim, valcount, symmetric, index_i, index_j, vals, b, 0, &plan);
    if (rc) return rc;
    if (!ctx && (rc = sogl_ctx_create(0, &ctx))) {
        sogl_plan_destroy(plan);
        return rc;
    }
    int32_t fid = 0;
    if ((rc = sogl_factor_load_plan(ctx, plan, SOGL_FP64, &fid))) {
        sogl_plan_destroy(plan);
        return rc;
    }
    sogl_plan_info info;
    sogl_plan_get_info(plan, &info);
    const int64_t N = (int64_t)info.block_rows * 64;
    std::vector<double> bp((size_t)N), xp((size_t)N);
    rc = sogl_factor_run(ctx, fid, SOGL_EXEC_PERSISTENT);
    if (!rc) rc = sogl_plan_permute_rhs(plan, b, 1, dim, bp.data());
    if (!rc) rc = sogl_solve(ctx, fid, bp.data(), xp.data(), nullptr, 1);
    if (!rc) rc = sogl_plan_unpermute(plan, xp.data(), 1, xo, dim);
    sogl_factor_release(ctx, fid);
    sogl_plan_destroy(plan);
    if (!rc)
        for (int32_t i = 0; i < dim; i++)
            if (std::isnan(xo[i])) return set_error(nullptr, SOGL_ERR_NAN, "NaN in the solution");
    return rc;
}

// ------------------------------------------------------------------ device memory helpers
extern "C" int sogl_device_alloc(sogl_ctx* x, size_t bytes, void** ptr) {
    if (!x || !ptr) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMalloc(ptr, bytes));
    return SOGL_OK;
}

extern "C" int sogl_device_free(sogl_ctx* x, void* ptr) {
    if (!x) return set_error(nullptr, SOGL_ERR_ARG, "null context");
    Context* c = &x->c;
    HIP_TRY(c, hipSetDevice(c->device));
    if (ptr) HIP_TRY(c, hipFree(ptr));
    return SOGL_OK;
}

extern "C" int sogl_device_copy(sogl_ctx* x, void* dst, const void* src, size_t bytes, int32_t to_device) {
    if (!x || (bytes && (!dst || !src))) return set_error(nullptr, SOGL_ERR_ARG, "null argument");
    Context* c = &x->c;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return SOGL_OK;
}
