// extern "C" entry points of the host half of sogl.h: reader and planner.
#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "planner.h"

struct sogl_mtx {
    sogl::Coo coo;
    std::vector<double> rhs;
    std::string log;
};

struct sogl_plan {
    sogl::Plan plan;
    std::string err;
};

namespace {
thread_local std::string g_host_err;
int fail(int code, const std::string& msg) {
    g_host_err = msg;
    return code;
}
}  // namespace

namespace sogl {
void set_last_error(const std::string& s) { g_host_err = s; }

// read-only views of a plan for the device half of the library
struct PlanView {
    int32_t n;
    int64_t N;
    const int32_t* new_order;
    int64_t nnz;
    const int32_t* row;
    const int32_t* col;
    const double* val;
    int64_t nnz_orig;
    // the matrix as the factor holds it: input tiles as cells (a COO position that
    // repeats keeps one value, as in the tiles)
    int64_t n_inputs;
    const int32_t* in_bi;
    const int32_t* in_bj;
    const int64_t* cell_begin;
    const uint16_t* cell_pos;
    const double* cell_val;
};
bool plan_view(const sogl_plan* p, PlanView* v) {
    if (!p || !v) return false;
    const Plan& P = p->plan;
    v->n = P.cfg.mSize;
    v->N = (int64_t)P.cfg.blockRows * P.cfg.blockSize;
    v->new_order = P.newOrder.data();
    v->nnz = (int64_t)P.pval.size();
    v->row = P.prow.data();
    v->col = P.pcol.data();
    v->val = P.pval.data();
    v->nnz_orig = v->nnz - (v->N - v->n);
    v->n_inputs = (int64_t)P.inputIds.size();
    v->in_bi = P.inputBi.data();
    v->in_bj = P.inputBj.data();
    v->cell_begin = P.inputCellBegin.data();
    v->cell_pos = P.cellPos.data();
    v->cell_val = P.cellVal.data();
    return true;
}
bool plan_input_cells(const sogl_plan* p, const int64_t** begin, const uint16_t** pos, const double** val) {
    if (!p) return false;
    *begin = p->plan.inputCellBegin.data();
    *pos = p->plan.cellPos.data();
    *val = p->plan.cellVal.data();
    return true;
}
bool plan_factor_desc(const sogl_plan* p, int32_t precision, sogl_factor_desc* d) {
    if (!p || !d) return false;
    const Plan& P = p->plan;
    d->n_ids = P.storageCount;
    d->block_rows = P.cfg.blockRows;
    d->ops = P.ops.data();
    d->n_ops = (int64_t)P.ops.size();
    d->input_ids = P.inputIds.data();
    d->input_tiles = nullptr;  // the plan keeps cells: see plan_input_cells
    d->n_inputs = (int64_t)P.inputIds.size();
    d->l_bi = P.lBi.data();
    d->l_bj = P.lBj.data();
    d->l_ids = P.lId.data();
    d->n_l = (int64_t)P.lId.size();
    d->u_bi = P.uBi.data();
    d->u_bj = P.uBj.data();
    d->u_ids = P.uId.data();
    d->n_u = (int64_t)P.uId.size();
    d->precision = precision;
    return true;
}
}  // namespace sogl

extern "C" const char* sogl_last_error(void) { return g_host_err.c_str(); }

extern "C" int sogl_mtx_read(const char* path, sogl_mtx** out) {
    if (!path || !out) return fail(SOGL_ERR_ARG, "null argument");
    *out = nullptr;
    sogl_mtx* m = new (std::nothrow) sogl_mtx();
    if (!m) return fail(SOGL_ERR_OOM, "out of host memory");
    try {
        if (sogl::read_mtx(path, m->coo, &m->log) == 0) {
            std::string msg = m->log.empty() ? "no entries read" : m->log;
            delete m;
            return fail(SOGL_ERR_IO, std::string(path) + ": " + msg);
        }
    } catch (const std::exception& e) {
        delete m;
        return fail(SOGL_ERR_IO, e.what());
    }
    *out = m;
    return SOGL_OK;
}

extern "C" int sogl_mtx_read_rhs(sogl_mtx* m, const char* path) {
    if (!m || !path) return fail(SOGL_ERR_ARG, "null argument");
    sogl::read_array(path, m->coo.n, m->rhs, &m->log);
    return SOGL_OK;
}

extern "C" int sogl_mtx_info(const sogl_mtx* m, int32_t* n, int64_t* nnz, int32_t* symmetric) {
    if (!m) return fail(SOGL_ERR_ARG, "null matrix");
    if (n) *n = m->coo.n;
    if (nnz) *nnz = (int64_t)m->coo.val.size();
    if (symmetric) *symmetric = m->coo.symmetric ? 1 : 0;
    return SOGL_OK;
}

extern "C" int sogl_mtx_arrays(const sogl_mtx* m, const int32_t** row, const int32_t** col,
                               const double** val, const double** rhs) {
    if (!m) return fail(SOGL_ERR_ARG, "null matrix");
    if (row) *row = m->coo.row.data();
    if (col) *col = m->coo.col.data();
    if (val) *val = m->coo.val.data();
    if (rhs) *rhs = m->rhs.empty() ? nullptr : m->rhs.data();
    return SOGL_OK;
}

extern "C" void sogl_mtx_free(sogl_mtx* m) { delete m; }

extern "C" int sogl_plan_create(int32_t dim, int64_t valcount, int32_t symmetric,
                                const int32_t* index_i, const int32_t* index_j,
                                const double* vals, const double* b, int32_t verbose,
                                sogl_plan** out) {
    if (!out) return fail(SOGL_ERR_ARG, "null output");
    *out = nullptr;
    if (dim <= 0 || valcount < 0 || (valcount > 0 && (!index_i || !index_j || !vals)))
        return fail(SOGL_ERR_ARG, "bad matrix arguments");
    sogl_plan* p = new (std::nothrow) sogl_plan();
    if (!p) return fail(SOGL_ERR_OOM, "out of host memory");
    try {
        sogl::Coo a;
        a.n = dim;
        a.symmetric = symmetric != 0;
        a.row.assign(index_i, index_i + valcount);
        a.col.assign(index_j, index_j + valcount);
        a.val.assign(vals, vals + valcount);
        sogl::PlanOptions opt;
        opt.verbose = verbose != 0;
        sogl::build_plan(a, b, opt, p->plan);
    } catch (const std::bad_alloc&) {
        delete p;
        return fail(SOGL_ERR_OOM, "out of host memory while planning");
    } catch (const std::exception& e) {
        delete p;
        std::string w = e.what();
        return fail(SOGL_ERR_ARG, w);
    }
    *out = p;
    return SOGL_OK;
}

extern "C" int sogl_plan_create_like(const sogl_plan* like, int32_t dim, int64_t valcount,
                                     const int32_t* index_i, const int32_t* index_j, const double* vals,
                                     const double* b, sogl_plan** out, int32_t* reused) {
    if (reused) *reused = 0;
    if (!out) return fail(SOGL_ERR_ARG, "null output");
    *out = nullptr;
    if (dim <= 0 || valcount < 0 || (valcount > 0 && (!index_i || !index_j || !vals)))
        return fail(SOGL_ERR_ARG, "bad matrix arguments");
    if (like) {
        sogl_plan* p = new (std::nothrow) sogl_plan();
        if (!p) return fail(SOGL_ERR_OOM, "out of host memory");
        try {
            sogl::Coo a;
            a.n = dim;
            a.row.assign(index_i, index_i + valcount);
            a.col.assign(index_j, index_j + valcount);
            a.val.assign(vals, vals + valcount);
            if (sogl::rebind_plan(like->plan, a, b, p->plan)) {
                if (reused) *reused = 1;
                *out = p;
                return SOGL_OK;
            }
        } catch (const std::bad_alloc&) {
            delete p;
            return fail(SOGL_ERR_OOM, "out of host memory while planning");
        }
        delete p;
    }
    return sogl_plan_create(dim, valcount, 0, index_i, index_j, vals, b, 0, out);
}

extern "C" int sogl_pattern_hash(int32_t dim, int64_t valcount, const int32_t* index_i,
                                 const int32_t* index_j, uint64_t* hash) {
    if (!hash || valcount < 0 || (valcount > 0 && (!index_i || !index_j))) return fail(SOGL_ERR_ARG, "bad arguments");
    *hash = sogl::pattern_hash(dim, valcount, index_i, index_j);
    return SOGL_OK;
}

extern "C" void sogl_plan_destroy(sogl_plan* p) { delete p; }

extern "C" const char* sogl_plan_log(const sogl_plan* p) {
    return p ? p->plan.log.c_str() : g_host_err.c_str();
}

extern "C" int sogl_plan_get_info(const sogl_plan* p, sogl_plan_info* info) {
    if (!p || !info) return fail(SOGL_ERR_ARG, "null argument");
    const sogl::Plan& P = p->plan;
    std::memset(info, 0, sizeof(*info));
    info->m_size = P.cfg.mSize;
    info->block_rows = P.cfg.blockRows;
    info->block_rows_l2 = P.cfg.blockRowsL2;
    info->block_size_l2 = P.cfg.blockSizeL2;
    info->storage_count = P.storageCount;
    info->n_stages = P.maxStage;
    info->gps_levels = P.gpsLevels;
    info->gps_bandwidth = P.gpsBandwidth;
    info->n_ops = (int64_t)P.ops.size();
    info->n_coarse_ops = (int64_t)P.coarseOps.size();
    info->n_inputs = (int64_t)P.inputIds.size();
    info->n_l_tiles = (int64_t)P.lId.size();
    info->n_u_tiles = (int64_t)P.uId.size();
    for (const sogl_op& o : P.ops) {
        switch (o.op) {
            case SOGL_OP_MUL: info->n_mul++; break;
            case SOGL_OP_MULNEG: info->n_mulneg++; break;
            case SOGL_OP_SUB: info->n_sub++; break;
            case SOGL_OP_LU: info->n_lu++; break;
            case SOGL_OP_LOWER_INV: info->n_lower_inv++; break;
            case SOGL_OP_UPPER_INV: info->n_upper_inv++; break;
            case SOGL_OP_LLT: info->n_llt++; break;
            case SOGL_OP_MULT: info->n_mult++; break;
            default: break;
        }
    }
    info->symmetric = P.symmetric ? 1 : 0;
    info->t_reorder = P.tReorder;
    info->t_plan = P.tPlan;
    return SOGL_OK;
}

extern "C" int sogl_plan_get_perm(const sogl_plan* p, int32_t* new_order, int32_t* reverse_order) {
    if (!p) return fail(SOGL_ERR_ARG, "null plan");
    const size_t n = p->plan.newOrder.size();
    if (new_order) std::memcpy(new_order, p->plan.newOrder.data(), n * sizeof(int32_t));
    if (reverse_order) std::memcpy(reverse_order, p->plan.reverseOrder.data(), n * sizeof(int32_t));
    return SOGL_OK;
}

extern "C" int sogl_plan_get_ops(const sogl_plan* p, int32_t level, sogl_op* out) {
    if (!p || !out) return fail(SOGL_ERR_ARG, "null argument");
    const std::vector<sogl_op>& v = level == 0 ? p->plan.ops : p->plan.coarseOps;
    std::memcpy(out, v.data(), v.size() * sizeof(sogl_op));
    return SOGL_OK;
}

extern "C" int sogl_plan_get_tile_stages(const sogl_plan* p, int32_t* laststage, int32_t* stage) {
    if (!p) return fail(SOGL_ERR_ARG, "null plan");
    const size_t n = (size_t)p->plan.storageCount;
    if (laststage) std::memcpy(laststage, p->plan.laststage.data(), n * sizeof(int32_t));
    if (stage) std::memcpy(stage, p->plan.tileStage.data(), n * sizeof(int32_t));
    return SOGL_OK;
}

extern "C" int sogl_plan_get_inputs(const sogl_plan* p, int32_t* ids, int32_t* bi, int32_t* bj,
                                    double* tiles) {
    if (!p) return fail(SOGL_ERR_ARG, "null plan");
    const sogl::Plan& P = p->plan;
    const size_t n = P.inputIds.size();
    if (ids) std::memcpy(ids, P.inputIds.data(), n * sizeof(int32_t));
    if (bi) std::memcpy(bi, P.inputBi.data(), n * sizeof(int32_t));
    if (bj) std::memcpy(bj, P.inputBj.data(), n * sizeof(int32_t));
    if (tiles) P.input_tiles_dense(tiles);
    return SOGL_OK;
}

extern "C" int sogl_plan_get_factor_tiles(const sogl_plan* p, int32_t which, int32_t* bi,
                                          int32_t* bj, int32_t* ids) {
    if (!p) return fail(SOGL_ERR_ARG, "null plan");
    const sogl::Plan& P = p->plan;
    const std::vector<int32_t>& vi = which == 0 ? P.lBi : P.uBi;
    const std::vector<int32_t>& vj = which == 0 ? P.lBj : P.uBj;
    const std::vector<int32_t>& vd = which == 0 ? P.lId : P.uId;
    if (bi) std::memcpy(bi, vi.data(), vi.size() * sizeof(int32_t));
    if (bj) std::memcpy(bj, vj.data(), vj.size() * sizeof(int32_t));
    if (ids) std::memcpy(ids, vd.data(), vd.size() * sizeof(int32_t));
    return SOGL_OK;
}

// solveLU pads b with 1.0 (solver.cpp:155-161) and permutes it with the ordering
// (GPSOrder.cpp:433-447, 126-132): bp[new] = b[newOrder[new]]
extern "C" int sogl_plan_permute_rhs(const sogl_plan* p, const double* b, int32_t nrhs,
                                     int64_t ldb, double* bp) {
    if (!p || !b || !bp || nrhs <= 0) return fail(SOGL_ERR_ARG, "bad argument");
    const sogl::Plan& P = p->plan;
    const int32_t n = P.cfg.mSize;
    const int64_t N = (int64_t)P.cfg.blockRows * P.cfg.blockSize;
    if (ldb < n) return fail(SOGL_ERR_ARG, "ldb below n");
    for (int32_t r = 0; r < nrhs; r++) {
        const double* bc = b + (int64_t)r * ldb;
        double* pc = bp + (int64_t)r * N;
        for (int32_t i = 0; i < n; i++) pc[i] = bc[P.newOrder[(size_t)i]];
        for (int64_t i = n; i < N; i++) pc[i] = 1.0;
    }
    return SOGL_OK;
}

extern "C" int sogl_plan_unpermute(const sogl_plan* p, const double* xp, int32_t nrhs, double* x,
                                   int64_t ldx) {
    if (!p || !xp || !x || nrhs <= 0) return fail(SOGL_ERR_ARG, "bad argument");
    const sogl::Plan& P = p->plan;
    const int32_t n = P.cfg.mSize;
    const int64_t N = (int64_t)P.cfg.blockRows * P.cfg.blockSize;
    if (ldx < n) return fail(SOGL_ERR_ARG, "ldx below n");
    for (int32_t r = 0; r < nrhs; r++) sogl::unpermute(P, xp + (int64_t)r * N, x + (int64_t)r * ldx);
    return SOGL_OK;
}

extern "C" int64_t sogl_plan_perm_nnz(const sogl_plan* p) {
    return p ? (int64_t)p->plan.pval.size() : -1;
}

extern "C" int sogl_plan_get_perm_coo(const sogl_plan* p, int32_t* row, int32_t* col, double* val) {
    if (!p) return fail(SOGL_ERR_ARG, "null plan");
    const sogl::Plan& P = p->plan;
    const size_t n = P.pval.size();
    if (row) std::memcpy(row, P.prow.data(), n * sizeof(int32_t));
    if (col) std::memcpy(col, P.pcol.data(), n * sizeof(int32_t));
    if (val) std::memcpy(val, P.pval.data(), n * sizeof(double));
    return SOGL_OK;
}
