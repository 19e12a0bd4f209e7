// ./solve file.mtx -- drop-in for the reference CLI (main.cpp:34-66).
//
// Reads file.mtx and file_b.mtx, plans on the host, factors and solves on
// GPU 0, prints the reference's log lines, the max-abs residual and a sample
// of x, and writes file_x.mtx exactly as the reference does: the reference
// writes data::b after the solve, i.e. the permuted forward-substitution
// vector y at 6 significant digits (main.cpp:61).  The true x is written at
// full precision to file_xsol.mtx, which the reference has no counterpart for.
//
// Options after the file name (none in the reference):
//   --fp32 | --bf16        low-precision factor with FP64 iterative refinement
//   --tol T                refinement tolerance (default 1e-12)
//   --exec persistent|levels|perop|auto   (auto: tune the executor configuration
//                          on the factor first; 'kernel time' is the tuned factor)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/sogl.h"
#include "host/planner.h"

static void sample(const double* a, int dim) {  // mtx::sampleArray (mtx.cpp:119-128)
    if (dim > 3) std::cout << a[0] << " " << a[1] << " " << a[2] << " ";
    std::cout << "...";
    if (dim > 3) std::cout << " " << a[dim - 3] << " " << a[dim - 2] << " " << a[dim - 1] << " ";
    std::cout << '\n';
}

static int die(const char* what, int rc, const char* msg) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, msg ? msg : "");
    return 1;
}

int main(int argc, char** argv) {
    if (argc <= 1 || std::string(argv[1]).find(".mtx") == std::string::npos) {
        std::cout << "usage: ./solve filename.mtx" << std::endl;
        return 0;
    }
    int precision = SOGL_FP64, mode = SOGL_EXEC_PERSISTENT;
    double tol = 1e-12;
    for (int i = 2; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--fp32") precision = SOGL_FP32;
        else if (a == "--bf16") precision = SOGL_BF16;
        else if (a == "--tol" && i + 1 < argc) tol = std::atof(argv[++i]);
        else if (a == "--exec" && i + 1 < argc) {
            std::string m = argv[++i];
            mode = m == "levels" ? SOGL_EXEC_LEVELS
                   : m == "perop" ? SOGL_EXEC_PER_OP
                   : m == "auto"  ? SOGL_EXEC_AUTO
                                   : SOGL_EXEC_PERSISTENT;
        }
    }
    const std::string fname = argv[1];
    const std::string base = fname.substr(0, fname.find(".mtx"));
    sogl::Coo a;
    std::string log;
    if (sogl::read_mtx(fname, a, &log) == 0) {
        std::cout << log;
        return 0;
    }
    std::cout << log;
    log.clear();
    std::vector<double> b;
    sogl::read_array(base + "_b.mtx", a.n, b, &log);
    std::cout << log;
    auto t0 = std::chrono::steady_clock::now();
    sogl_plan* plan = nullptr;
    int rc = sogl_plan_create(a.n, (int64_t)a.val.size(), a.symmetric ? 1 : 0, a.row.data(), a.col.data(),
                              a.val.data(), b.data(), 1, &plan);
    if (rc) return die("plan", rc, sogl_last_error());
    sogl_plan_info info;
    sogl_plan_get_info(plan, &info);
    const int64_t N = (int64_t)info.block_rows * 64;

    sogl_ctx* ctx = nullptr;
    if ((rc = sogl_ctx_create(0, &ctx))) return die("device", rc, sogl_last_error());
    int32_t fid = 0;
    if ((rc = sogl_factor_load_plan(ctx, plan, precision, &fid))) return die("load", rc, sogl_ctx_error(ctx));
    if (mode == SOGL_EXEC_AUTO && (rc = sogl_factor_autotune(ctx, fid, 1, nullptr, nullptr)))
        return die("autotune", rc, sogl_ctx_error(ctx));
    auto tc = std::chrono::steady_clock::now();
    if ((rc = sogl_factor_run(ctx, fid, mode))) return die("factor", rc, sogl_ctx_error(ctx));
    double t_calc = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
    {  // the reference prints these while it factors (BlockPlanner.cpp:573-577, MatrixStdDouble.cpp:2908-2934)
        int32_t nbad = 0;
        double det[64];
        if ((rc = sogl_factor_get_uinv_check(ctx, fid, &nbad, det))) return die("check", rc, sogl_ctx_error(ctx));
        for (int q = 0; q < std::min<int32_t>(nbad, 16); q++) {
            const char* what = det[4 * q + 1] == 0 ? "diag" : "middle";
            std::cout << std::string(what) + " min:" + std::to_string(det[4 * q + 2]) + "  " + what +
                             " max:" + std::to_string(det[4 * q + 3])
                      << std::endl;
            std::cout << " upper out of tolerance " << std::endl;
        }
    }
    std::cout << "kernel time: " << t_calc << '\n';

    std::vector<double> bp((size_t)N), xp((size_t)N), y((size_t)N), x((size_t)a.n);
    sogl_plan_permute_rhs(plan, b.data(), 1, a.n, bp.data());
    auto ts = std::chrono::steady_clock::now();
    if ((rc = sogl_solve(ctx, fid, bp.data(), xp.data(), y.data(), 1))) return die("solve", rc, sogl_ctx_error(ctx));
    sogl_plan_unpermute(plan, xp.data(), 1, x.data(), a.n);
    if (precision != SOGL_FP64) {
        int32_t it = 0;
        double res = 0;
        if ((rc = sogl_refine_attach(ctx, fid, plan))) return die("refine", rc, sogl_ctx_error(ctx));
        rc = sogl_solve_refined(ctx, fid, b.data(), x.data(), 1, tol, 50, &it, &res);
        std::cout << "refinement iterations: " << it << " residual: " << res << '\n';
        if (rc && rc != SOGL_ERR_NOT_CONVERGED) return die("refine", rc, sogl_ctx_error(ctx));
    }
    double t_solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
    std::cout << "solve triangled :" << t_solve << '\n';
    int nan = 0;
    for (int32_t i = 0; i < a.n; i++) nan += std::isnan(x[(size_t)i]) ? 1 : 0;
    if (nan > 0) std::cout << "found NaN:  " + std::to_string(nan) << std::endl;
    double t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "totaltime: " << t_total << " reordertime: " << info.t_reorder
              << " calculationtime: " << (t_total - info.t_reorder) << std::endl;

    std::cout << "max rhs error:" << sogl::check_result(a, b.data(), x.data()) << std::endl;
    sogl::write_array(base + "_x.mtx", y.data(), a.n);  // the reference's file content (y)
    sogl::write_array_exact(base + "_xsol.mtx", x.data(), a.n);
    sample(x.data(), a.n);
    sogl_factor_release(ctx, fid);
    sogl_ctx_destroy(ctx);
    sogl_plan_destroy(plan);
    return 0;
}
