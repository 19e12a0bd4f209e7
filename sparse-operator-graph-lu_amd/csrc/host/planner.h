// Host side of the sparse block LU path: Matrix Market reader, size config,
// Gibbs-Poole-Stockmeyer ordering and the two-level block planner that turns
// a sparse matrix into a staged list of 64x64 unit operations.
//
// Everything here is integer/byte work that must reproduce the reference
// bit for bit (permutation, tile ids, op list, stages), so each function
// cites the reference file:line whose behaviour it restates.  Unlike the
// reference, all state lives in a per-solve object (no static singletons),
// so several plans can be built in one process.
#pragma once

#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "../../../include/sogl.h"

namespace sogl {

constexpr int kTile = 64;  // BLOCK64 (const.h:19)

// ---------------------------------------------------------------- reader
// Coordinate matrix as produced by mtx::readMTX (mtx.cpp:44-117): 0-based,
// explicit zeros and out-of-range entries dropped.
struct Coo {
    int32_t n = 0;
    bool symmetric = false;
    std::vector<int32_t> row, col;
    std::vector<double> val;
};

// mtx::readMTX (mtx.cpp:44-117).  Returns the number of entries kept
// (0 on an unreadable file, like the reference).
int64_t read_mtx(const std::string& path, Coo& out, std::string* log);
// mtx::readArray (mtx.cpp:145-182): missing values become 1.0.
int64_t read_array(const std::string& path, int32_t dim, std::vector<double>& b,
                   std::string* log);
// mtx::writeArray (mtx.cpp:130-143): 6 significant digits, like std::ostream.
bool write_array(const std::string& path, const double* a, int32_t dim);
// Full-precision (%.17g) array writer; has no reference counterpart.
bool write_array_exact(const std::string& path, const double* a, int32_t dim);
// mtx::checkResult (mtx.cpp:183-215): max_i |b_i - (A x)_i|.
double check_result(const Coo& a, const double* b, const double* x);

// ---------------------------------------------------------------- config
// config::set + roundup (config.cpp:30-49).
struct SizeConfig {
    int32_t mSize = 0;
    int32_t blockRows = 0;    // fine block rows (power of two, > n/64)
    int32_t blockSize = kTile;
    int32_t blockRowsL2 = 0;  // coarse block rows
    int32_t blockSizeL2 = 0;  // coarse block edge in unknowns
    static SizeConfig set(int32_t dim);
};

// ---------------------------------------------------------------- quadtree
// struct matrix (matrix.h:22-44): Z-order quadtree of tile ids; children
// are b11, b12, b21, b22.  Nodes live in a deque so pointers stay valid.
struct Node {
    int32_t blockrows2 = 0;
    int32_t level = 0;
    int64_t id = 0;
    Node* sub[4] = {nullptr, nullptr, nullptr, nullptr};
};

// ---------------------------------------------------------------- plan
struct Plan {
    SizeConfig cfg;
    // ordering: newOrder[new] = old, reverseOrder[old] = new (GPSOrder.cpp:448-449)
    std::vector<int32_t> newOrder, reverseOrder;
    int32_t gpsLevels = 0, gpsBandwidth = 0, gpsStart = 0;
    bool structureSymmetric = false;
    bool symmetric = false;  // L L^T plan: factor tiles are L only (U = L^T)

    std::vector<sogl_op> coarseOps;  // coarse list after its blockPlan
    std::vector<sogl_op> ops;        // fine 64x64 unit operations, staged and sorted
    int32_t storageCount = 0;        // number of tile ids (id 0 = null tile)
    std::vector<int32_t> laststage;  // per id: last stage that reads it
    std::vector<int32_t> tileStage;  // per id: stage of its last write (data::stage)
    int32_t maxStage = 0;

    // input tiles (fine ids in allocation order) as their nonzero cells: the cells of
    // input k are [inputCellBegin[k], inputCellBegin[k+1]), each a position
    // row*64 + col inside the tile and a value, one cell per position (where the COO
    // repeats an entry, the value scattered last in the reference's order).
    std::vector<int32_t> inputIds, inputBi, inputBj;
    std::vector<int64_t> inputCellBegin;
    std::vector<uint16_t> cellPos;
    std::vector<double> cellVal;
    std::vector<uint32_t> cellSrc;  // per cell: the entry of the padded permuted COO it holds
    // dense 64x64 row-major tiles of all inputs (n_inputs * 4096 doubles)
    void input_tiles_dense(double* out) const;

    // factor tiles reachable from the L and U quadtrees
    std::vector<int32_t> lBi, lBj, lId, uBi, uBj, uId;

    // permuted COO of the padded problem (indices after GPS + sortInBlock)
    std::vector<int32_t> prow, pcol;
    std::vector<double> pval;

    int64_t nnzOrig = 0;        // COO entries given (the rest of prow/pcol/pval is padding)
    bool blockSorted = false;   // sortInBlock changed the GPS order (it depends on values)

    std::vector<double> bPerm;  // permuted, padded right-hand side (may be empty)
    double tReorder = 0, tPlan = 0;
    std::string log;            // the reference's stdout lines, in order
};

struct PlanOptions {
    bool verbose = false;        // also echo log lines to stdout as they happen
    bool keepCoarse = true;
    int threads = 0;             // OpenMP threads for the GPS candidate search (0 = default)
};

// solveLU's host part (solver.cpp:121-161) + GOrder::Reorder/sortInBlock +
// decompose_solveLU's planning (solver.cpp:50-100).  b may be null.
// Throws std::runtime_error on inputs the reference cannot handle.
void build_plan(const Coo& a, const double* b, const PlanOptions& opt, Plan& plan);

// The plan of a matrix with the same sparsity pattern as `like` (the same n and the
// same row/column sequence) from `like`'s ordering and op list: only the values, the
// right-hand side and the input cells are new.  The result equals build_plan's.  The
// ordering depends on values only through sortInBlock, so this returns false (and
// leaves `out` untouched) where the pattern differs or sortInBlock would reorder for
// either matrix; build_plan is the way then.  Has no reference counterpart.
bool rebind_plan(const Plan& like, const Coo& a, const double* b, Plan& out);

// 64-bit key of a sparsity pattern (n and the row/column sequence) for plan caches.
uint64_t pattern_hash(int32_t n, int64_t nnz, const int32_t* row, const int32_t* col);

// GOrder::reOrderResult (GPSOrder.cpp:41-53): x_orig[i] = x_perm[reverseOrder[i]].
void unpermute(const Plan& plan, const double* xperm, double* xorig);

}  // namespace sogl
