// Golden-fixture harness (test infrastructure, runs only where the reference
// sources are present).  It links the reference's own objects -- built from
// /root/reference by oracle/Makefile into oracle/_ref -- and drives the
// reference's public static API through the same sequence as SOGLU::solveLU
// and decompose_solveLU (solver.cpp:50-184), dumping the intermediate state
// that those functions clear before returning:
//
//   perm.i32        newOrder[new]=old then reverseOrder[old]=new (2*n)
//   bperm.f64       permuted padded b before the factor (N)
//   coarse_ops.i32  coarse op list after its blockPlan, 8 ints per op
//   ops.i32         fine op list after the final blockPlan, 8 ints per op
//   tiles.i32       storageCount, then laststage[S], then stage[S]
//   inputs.i32      per input tile: id, bi, bj
//   inputs.f64      input tiles, dense 64x64 row-major (mask-filtered)
//   ltiles.i32 / utiles.i32  factor tree leaves: bi, bj, id
//   x.f64           data::x (permuted solution, n) -- unless --plan-only
//   y.f64           data::b after solve (forward substitution result, N)
//   xorig.f64       x in the original order (GOrder::reOrderResult)
//   meta.txt        sizes and timings
//
// Op records are written as: stage group seq src src2 op result result2.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "operation.h"
#include "matrix.h"
#include "data.h"
#include "memutil.h"
#include "BlockPlanner.h"
#include "GPSOrder.h"
#include "config.h"
#include "mtx.h"
#include "solver.h"

using namespace SOGLU;

// Built with SOGL_GPU_SITES (oracle/Makefile target `integration`): the two call
// sites of INTEGRATION.md section 2, extracted from it verbatim, replace
// BlockPlanner::calculate() and BlockPlanner::solve() below.
#ifdef SOGL_GPU_SITES
#include "integration_sites.inc"
#endif

static std::string g_out;

template <typename T>
static void dump(const char* name, const T* p, size_t n) {
    std::string path = g_out + "/" + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    if (n) std::fwrite(p, sizeof(T), n, f);
    std::fclose(f);
}

static void dump_ops(const char* name) {
    std::vector<int32_t> v;
    v.reserve(data::graph.size() * 8);
    for (operation* o : data::graph) {
        int32_t r[8] = {o->stage, o->groupNum, o->sequenceNum, o->src, o->src2, (int32_t)o->op,
                        o->result, o->result2};
        v.insert(v.end(), r, r + 8);
    }
    dump(name, v.data(), v.size());
}

static void leaves(matrix* t, int r0, int c0, int size, std::vector<int32_t>& out) {
    if (t == NULL) return;
    if (size == 1) {
        if (t->blockindex > 0) {
            out.push_back(r0);
            out.push_back(c0);
            out.push_back((int32_t)t->blockindex);
        }
        return;
    }
    int h = size / 2;
    for (int i = 0; i < 4; i++) leaves(t->submatrix[i], r0 + (i >> 1) * h, c0 + (i & 1) * h, h, out);
}

// dense copy of a reference tile, honouring the row-detail masks (Appendix of
// the tile format: bit p of the row's uint16 covers columns 8p..8p+7)
static void dense_tile(const double* t, double* out) {
    const uint16_t* detail = (const uint16_t*)(t + DETAILOFFSET);
    for (int i = 0; i < BLOCK64; i++) {
        uint16_t m = detail[DETAILSKIPSHORT * (i / 32) + i % 32];
        for (int j = 0; j < BLOCK64; j++)
            out[i * 64 + j] = (m & (1u << (j / 8))) ? t[i * BLOCKCOL + j] : 0.0;
    }
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_dump file.mtx outdir [--plan-only]\n");
        return 1;
    }
    std::string fname = argv[1];
    g_out = argv[2];
    bool plan_only = argc > 3 && std::string(argv[3]) == "--plan-only";

    iniData();
    std::string base = fname.substr(0, fname.find(".mtx"));
    if (mtx::readMTX(fname) == 0) return 1;
    mtx::readArray(base + "_b.mtx", mtx::mSize);

    auto t0 = std::chrono::steady_clock::now();
    // --- solveLU (solver.cpp:121-161)
    const int dim = mtx::mSize, valcount = mtx::valcount;
    data::mSize = dim;
    data::mSize_out = dim;
    data::valcount = valcount;
    data::symmetric = mtx::symmetric;
    // symmetric input: each off-diagonal entry followed by its mirror (solver.cpp:130-149)
    const int cap = data::symmetric ? valcount * 2 : valcount;
    data::indexi = (int*)memutil::getSmallMem(1, sizeof(int) * cap);
    data::indexj = (int*)memutil::getSmallMem(1, sizeof(int) * cap);
    data::vals = (double*)memutil::getSmallMem(1, sizeof(double) * cap);
    int ii = 0;
    for (int i = 0; i < valcount; i++) {
        data::indexi[ii] = mtx::indexi[i];
        data::indexj[ii] = mtx::indexj[i];
        data::vals[ii] = mtx::vals[i];
        ii++;
        if (data::symmetric && mtx::indexi[i] != mtx::indexj[i]) {
            data::indexi[ii] = mtx::indexj[i];
            data::indexj[ii] = mtx::indexi[i];
            data::vals[ii] = mtx::vals[i];
            ii++;
        }
    }
    data::valcount = ii;
    config::set(dim);
    data::blockRows = config::blockRows;
    data::blockSize = config::blockSize;
    const int N = data::blockRows * data::blockSize;
    data::b = (double*)memutil::getSmallMem(1, N * sizeof(double));
    for (int i = 0; i < dim; i++) data::b[i] = mtx::b[i];
    for (int i = dim; i < N; i++) data::b[i] = 1;
    GOrder::Reorder();
    GOrder::sortInBlock();
    double t_reorder = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    dump("perm.i32", GOrder::newOrder, dim);
    {
        std::vector<int32_t> rv(GOrder::reverseOrder, GOrder::reverseOrder + dim);
        std::string p = g_out + "/perm.i32";
        FILE* f = std::fopen(p.c_str(), "ab");
        std::fwrite(rv.data(), 4, rv.size(), f);
        std::fclose(f);
    }
    dump("bperm.f64", data::b, N);
    GOrder::clearmost();

    // --- decompose_solveLU planning (solver.cpp:50-100)
    auto tp = std::chrono::steady_clock::now();
    BlockPlanner::appendBlockStorage(NULL);
    data::PlanL2 = true;
    data::blockSize = config::blockSizeL2;
    data::blockRows = config::blockRowsL2;
    BlockPlanner::checkBlock();
    BlockPlanner::iniBlockStorage();
    int n = data::blockRows;
    int levels = __builtin_popcount(data::blockRows - 1);
    matrix* bl = memutil::newmatrix(0, n, levels);
    matrix* bu = memutil::newmatrix(0, n, levels);
    if (data::symmetric) {
        BlockPlanner::blockMatrixLLT(data::blocks, bl, n, NULL);
        BlockPlanner::blockPlan(bl, NULL);
    } else {
        BlockPlanner::blockMatrixLU(data::blocks, bl, bu, n, NULL, NULL);
        BlockPlanner::blockPlan(bl, bu);
    }
    dump_ops("coarse_ops.i32");
    data::PlanL2 = false;
    data::blockSize = BLOCK64;
    int levelall = __builtin_popcount(data::blockRows - 1);
    matrix* bl2 = memutil::newmatrix(0, data::blockRows, levelall);
    matrix* bu2 = memutil::newmatrix(0, data::blockRows, levelall);
    BlockPlanner::copyOperatorL2(data::blocks, bl, data::symmetric ? NULL : bu, bl2, bu2, n);
    memutil::freeAllSmallMem(0);
    BlockPlanner::blockPlan(bl2, bu2);
    double t_plan = std::chrono::duration<double>(std::chrono::steady_clock::now() - tp).count();
    dump_ops("ops.i32");
    {
        std::vector<int32_t> v;
        const int S = data::storageCount;
        v.push_back(S);
        for (int i = 0; i < S; i++) v.push_back(data::laststage[i]);
        for (int i = 0; i < S; i++) v.push_back(data::stage[i]);
        dump("tiles.i32", v.data(), v.size());
    }
    {
        // input tiles: ids 1..T are the tiles built from the COO (BlockPlanner.cpp:1491-1541)
        std::vector<int32_t> meta;
        std::vector<double> vals;
        std::vector<double> one(64 * 64);
        const int R = data::blockRows;
        std::vector<int32_t> pos;
        leaves(data::blocks, 0, 0, R, pos);
        // order by id
        std::vector<int32_t> byid(pos.size());
        size_t cnt = pos.size() / 3;
        std::vector<int> idx(cnt);
        for (size_t k = 0; k < cnt; k++) idx[k] = (int)k;
        std::sort(idx.begin(), idx.end(), [&](int a, int b) { return pos[3 * a + 2] < pos[3 * b + 2]; });
        for (int k : idx) {
            meta.push_back(pos[3 * k + 2]);
            meta.push_back(pos[3 * k]);
            meta.push_back(pos[3 * k + 1]);
            dense_tile(data::blockstorage[pos[3 * k + 2]], one.data());
            vals.insert(vals.end(), one.begin(), one.end());
        }
        dump("inputs.i32", meta.data(), meta.size());
        dump("inputs.f64", vals.data(), vals.size());
        std::vector<int32_t> lv, uv;
        leaves(bl2, 0, 0, R, lv);
        leaves(bu2, 0, 0, R, uv);
        dump("ltiles.i32", lv.data(), lv.size());
        dump("utiles.i32", uv.data(), uv.size());
    }
    double t_calc = 0, t_solve = 0;
    if (!plan_only) {
        auto tc = std::chrono::steady_clock::now();
#ifdef SOGL_GPU_SITES
        if (gpu_calculate(bl2, bu2) != 0) {
            std::fprintf(stderr, "gpu_calculate: %s\n", sogl_last_error());
            return 3;
        }
#else
        BlockPlanner::calculate();
#endif
        t_calc = std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
        data::clearOperations();
        auto ts = std::chrono::steady_clock::now();
#ifdef SOGL_GPU_SITES
        std::vector<double> xs((size_t)N);
        if (gpu_solve(data::b, xs.data()) != 0) {
            std::fprintf(stderr, "gpu_solve: %s\n", sogl_last_error());
            return 3;
        }
        data::x = xs.data();
#else
        BlockPlanner::solve(bl2, bu2, data::b, data::blockRows * data::blockSize);
#endif
        t_solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
        dump("x.f64", data::x, dim);
        dump("y.f64", data::b, N);
        double* nr = GOrder::reOrderResult(data::x);
        dump("xorig.f64", nr, dim);
        mtx::checkResult(nr);
    }
    std::ofstream m(g_out + "/meta.txt");
    m << "n " << dim << "\nnnz " << valcount << "\nblock_rows " << config::blockRows
      << "\nblock_rows_l2 " << config::blockRowsL2 << "\nblock_size_l2 " << config::blockSizeL2
      << "\nstorage " << data::storageCount << "\nt_reorder " << t_reorder << "\nt_plan " << t_plan
      << "\nt_factor " << t_calc << "\nt_solve " << t_solve << "\n";
    return 0;
}
