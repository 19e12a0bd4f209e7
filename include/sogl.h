/*
 * sogl.h -- C ABI of the MI355X-native sparse block LU factorise-and-solve path.
 *
 * Plain C types only: pointers, sizes and status codes.  Two halves:
 *
 *   1. Host plan (CPU, no GPU needed): Matrix Market reader, GPS ordering and
 *      the two-level BlockPlanner that emits the staged list of 64x64 unit ops.
 *      These replace the reference's host pipeline and must match it bit for
 *      bit (permutation, tile ids, op list, stages).
 *   2. Device execution (HIP on one gfx950 GPU): the tile pool, the factor
 *      executor and the triangular solve, plus mixed precision with FP64
 *      iterative refinement.  These replace BlockPlanner::calculate() and
 *      BlockPlanner::solve(), the two call sites in decompose_solveLU.
 *
 * Every function returns SOGL_OK (0) or a positive error code; the message of
 * the last error is available from sogl_plan_log / sogl_ctx_error /
 * sogl_last_error.
 */
#ifndef SOGL_H
#define SOGL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define SOGL_OK 0
#define SOGL_ERR_ARG 1           /* bad argument or input the reference cannot handle */
#define SOGL_ERR_IO 2            /* file could not be read or written */
#define SOGL_ERR_OOM 3           /* host/device allocation failed or admission refused */
#define SOGL_ERR_DEVICE 4        /* HIP runtime error */
#define SOGL_ERR_NAN 5           /* NaN found in the solution */
#define SOGL_ERR_UNSUPPORTED 6   /* feature outside this path (e.g. a unit op kind it lacks) */
#define SOGL_ERR_NOT_CONVERGED 7 /* iterative refinement missed its tolerance */
#define SOGL_ERR_TIMEOUT 8       /* a device dependency wait hit its bound */
#define SOGL_ERR_GROWTH 9        /* pivot-free growth above threshold in a low-precision factor */

/* unit-operation codes: the same numbering as enum blockOp (operation.h:20) */
enum {
    SOGL_OP_INV = 0, SOGL_OP_LU = 1, SOGL_OP_LOWER_INV = 2, SOGL_OP_UPPER_INV = 3,
    SOGL_OP_SUB = 4, SOGL_OP_ADD = 5, SOGL_OP_NEG = 6, SOGL_OP_COPY = 7,
    SOGL_OP_MUL = 8, SOGL_OP_MULNEG = 9, SOGL_OP_LLT = 10, SOGL_OP_MULT = 11,
    SOGL_OP_NOOP = 12
};

/* factor precisions */
#define SOGL_FP64 0
#define SOGL_FP32 1
#define SOGL_BF16 2  /* experimental: bf16 MFMA operands, fp32 accumulation and storage */

/* One 64x64 unit operation; mirrors struct operation (operation.h:37-52).
 * sub: result = tile(src2) - tile(src); src == 0 means copy, src2 == 0 means negate
 * (BlockPlanner.cpp:553-572).  mul: result += src*src2; mulneg: result -= src*src2.
 * lu: src -> (result = L, result2 = U). */
typedef struct sogl_op {
    int32_t src, src2, op, result, result2, stage, seq, group;
} sogl_op;

/* ============================================================ host: reader */
typedef struct sogl_mtx sogl_mtx;

/* mtx::readMTX (mtx.cpp:44-117): coordinate file -> 0-based COO with
 * explicit zeros and out-of-range entries dropped. */
int sogl_mtx_read(const char* path, sogl_mtx** out);
/* mtx::readArray (mtx.cpp:145-182) into the matrix object; missing -> 1.0 */
int sogl_mtx_read_rhs(sogl_mtx* m, const char* path);
int sogl_mtx_info(const sogl_mtx* m, int32_t* n, int64_t* nnz, int32_t* symmetric);
int sogl_mtx_arrays(const sogl_mtx* m, const int32_t** row, const int32_t** col,
                    const double** val, const double** rhs);
void sogl_mtx_free(sogl_mtx* m);

/* ============================================================ host: plan */
typedef struct sogl_plan sogl_plan;

typedef struct sogl_plan_info {
    int32_t m_size;         /* n */
    int32_t block_rows;     /* fine block rows (padded n = 64*block_rows) */
    int32_t block_rows_l2;  /* coarse block rows */
    int32_t block_size_l2;  /* coarse block edge, unknowns */
    int32_t storage_count;  /* number of tile ids; id 0 is the null tile */
    int32_t n_stages;       /* last stage number */
    int32_t gps_levels, gps_bandwidth;
    int64_t n_ops, n_coarse_ops, n_inputs, n_l_tiles, n_u_tiles;
    int64_t n_mul, n_mulneg, n_sub, n_lu, n_lower_inv, n_upper_inv;
    double t_reorder, t_plan; /* seconds, the reference's timer boundaries */
    int64_t n_llt, n_mult;    /* L L^T plans (symmetric input): llt and mult ops */
    int32_t symmetric;        /* 1: L L^T plan */
    int32_t reserved;
} sogl_plan_info;

/* solveLU's host part (solver.cpp:121-161, GPSOrder.cpp, solver.cpp:50-100):
 * copy COO, size config, GPS + sortInBlock, coarse plan, L2 expansion, fine
 * plan.  b may be NULL.  verbose != 0 echoes the reference's log lines.
 * symmetric != 0: the COO holds one triangle of a symmetric matrix; each off-diagonal
 * entry is mirrored (solver.cpp:136-149) and the plan is the L L^T one
 * (blockMatrixLLT, BlockPlanner.cpp:941-989): llt / mult ops and no U tiles. */
int sogl_plan_create(int32_t dim, int64_t valcount, int32_t symmetric,
                     const int32_t* index_i, const int32_t* index_j,
                     const double* vals, const double* b, int32_t verbose,
                     sogl_plan** out);
/* Plan cache support (no reference counterpart: the reference plans every solve from
 * scratch, GPSOrder.cpp:494-635 and BlockPlanner.cpp:103-374, 1332-1460).
 * sogl_plan_create_like gives the same plan as sogl_plan_create (general matrix,
 * quiet) for a matrix with the same sparsity pattern as `like` -- the same dim and
 * the same (index_i, index_j) sequence -- by reusing `like`'s ordering, op list and
 * tile maps; only values, b and input tiles are new.  Where the pattern differs, or
 * the value-dependent sortInBlock step (GPSOrder.cpp:55-150) reorders for either
 * matrix, it plans from scratch.  *reused (may be NULL) is 1 when the plan was reused.
 * `like` may be NULL.  sogl_pattern_hash keys a cache of plans by pattern. */
int sogl_plan_create_like(const sogl_plan* like, int32_t dim, int64_t valcount,
                          const int32_t* index_i, const int32_t* index_j, const double* vals,
                          const double* b, sogl_plan** out, int32_t* reused);
int sogl_pattern_hash(int32_t dim, int64_t valcount, const int32_t* index_i,
                      const int32_t* index_j, uint64_t* hash);
void sogl_plan_destroy(sogl_plan* p);
const char* sogl_plan_log(const sogl_plan* p);  /* reference stdout lines, or last error */
int sogl_plan_get_info(const sogl_plan* p, sogl_plan_info* info);
/* newOrder[new] = old and reverseOrder[old] = new, length m_size */
int sogl_plan_get_perm(const sogl_plan* p, int32_t* new_order, int32_t* reverse_order);
/* level 0: fine ops (n_ops), level 1: coarse ops (n_coarse_ops) */
int sogl_plan_get_ops(const sogl_plan* p, int32_t level, sogl_op* out);
/* per tile id (storage_count): last reading stage, and stage of last write */
int sogl_plan_get_tile_stages(const sogl_plan* p, int32_t* laststage, int32_t* stage);
/* input tiles in id order; tiles (n_inputs*64*64 doubles, row-major) may be NULL */
int sogl_plan_get_inputs(const sogl_plan* p, int32_t* ids, int32_t* bi, int32_t* bj,
                         double* tiles);
/* which = 0: L factor tiles, 1: U factor tiles (block row, block col, id) */
int sogl_plan_get_factor_tiles(const sogl_plan* p, int32_t which, int32_t* bi,
                               int32_t* bj, int32_t* ids);
/* b (n x nrhs, column stride ldb) -> permuted padded bp (64*block_rows x nrhs) */
int sogl_plan_permute_rhs(const sogl_plan* p, const double* b, int32_t nrhs, int64_t ldb,
                          double* bp);
/* permuted padded x (64*block_rows x nrhs) -> original order x (n x nrhs, stride ldx);
 * GOrder::reOrderResult (GPSOrder.cpp:41-53) */
int sogl_plan_unpermute(const sogl_plan* p, const double* xp, int32_t nrhs, double* x,
                        int64_t ldx);
/* permuted padded COO of the problem (n_perm_nnz entries incl. padding diagonal) */
int64_t sogl_plan_perm_nnz(const sogl_plan* p);
int sogl_plan_get_perm_coo(const sogl_plan* p, int32_t* row, int32_t* col, double* val);

/* ============================================================ device */
typedef struct sogl_ctx sogl_ctx;

int sogl_ctx_create(int32_t device, sogl_ctx** out);
void sogl_ctx_destroy(sogl_ctx* ctx);
const char* sogl_ctx_error(const sogl_ctx* ctx);
/* free / total HBM bytes of the context's device */
int sogl_ctx_mem_info(sogl_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* synchronise the context's stream */
int sogl_ctx_sync(sogl_ctx* ctx);

/* Execution modes of sogl_factor_run */
#define SOGL_EXEC_PERSISTENT 0 /* one persistent kernel pulling ready tasks from a device queue */
#define SOGL_EXEC_LEVELS 1     /* one fused launch per dependency level */
#define SOGL_EXEC_PER_OP 2     /* one launch per unit op (the unfused comparison point) */
#define SOGL_EXEC_GRAPH 3      /* the level schedule captured once into a hipGraph, then replayed */
#define SOGL_EXEC_AUTO 4       /* the configuration sogl_factor_autotune picked (tunes on first use) */

/* 16x16 block masks (the reference's 8x8 sub-block masks, MatrixStdDouble.cpp:120-154,
 * at MFMA block size): a block with no |v| > 1e-18 counts as zero, and products and
 * substitutions skip it.  ON always, OFF never (dense tiles), AUTO (default): the first
 * factor run is masked, later runs keep the masks when at least a quarter of the factor
 * tiles' blocks came out zero.  sogl_factor_stats reports the share and the choice. */
#define SOGL_MASKS_OFF 0
#define SOGL_MASKS_ON 1
#define SOGL_MASKS_AUTO 2

typedef struct sogl_factor_desc {
    int32_t n_ids;          /* storage_count */
    int32_t block_rows;     /* fine block rows */
    const sogl_op* ops;     /* staged, sorted op list */
    int64_t n_ops;
    const int32_t* input_ids;
    const double* input_tiles; /* n_inputs * 4096 doubles, row-major 64x64 */
    int64_t n_inputs;
    const int32_t* l_bi; const int32_t* l_bj; const int32_t* l_ids; int64_t n_l;
    const int32_t* u_bi; const int32_t* u_bj; const int32_t* u_ids; int64_t n_u;
    int32_t precision;      /* SOGL_FP64 / SOGL_FP32 / SOGL_BF16 */
} sogl_factor_desc;

/* Build the device task graph and tile pool for one factorisation and copy the
 * input tiles to HBM.  The factor stays resident under *factor_id until released. */
int sogl_factor_load(sogl_ctx* ctx, const sogl_factor_desc* desc, int32_t* factor_id);
/* Pool slot reuse for the factors this context loads from now on: the reference's
 * liveness-based tile recycling (BlockPlanner.cpp:597-634, memutil.cpp:74-113).  A
 * result takes over the HBM tile of a value that only its own task reads (a GEMM's added
 * tile or operand, a sub's operand, an FP64 lu / llt source), or, for a GEMM, a tile
 * whose last value's readers all lie 1024 or more dependency levels earlier, after
 * waiting for those readers' flags.  The factor tiles and diagonal inverses keep theirs for the
 * solve.  A factor loaded from a plan keeps its input cells on the device and scatters
 * them into the input tiles again before every run (inside factor_ms, outside exec_ms),
 * so their tiles are reused as well.  Reuse reads tiles with device-coherent loads
 * (2-7% slower factors); the factor's bits are the same either way.  ON / OFF, or AUTO
 * (the default): reuse when the factor's values would take more than a quarter of the
 * device's HBM.  The per-op executor accumulates unit ops in the result's tile and runs
 * only without reuse. */
#define SOGL_REUSE_OFF 0
#define SOGL_REUSE_ON 1
#define SOGL_REUSE_AUTO 2
int sogl_ctx_set_pool_reuse(sogl_ctx* ctx, int32_t mode);
/* Task links of the persistent executor, for the factors loaded from now on (a mask of
 * SOGL_LINK_*; default SOGL_LINK_PANEL).  No reference counterpart (the reference runs
 * a stage's ops on a thread pool); results are the same bits with any mask.
 *   SOGL_LINK_PANEL: the lu task of an FP32 LU factor runs its most critical panel
 *     GEMM (A U^-1 or L^-1 A, one pair) itself, from the inverse it keeps in LDS, and
 *     publishes both tiles at once.
 *   SOGL_LINK_AFTER: the workgroup that made a tile on the critical path goes on, right
 *     after publishing it, to a one-pair GEMM that reads it, instead of handing it to
 *     another workgroup (measured slower: off by default).
 * The links are chosen on the host by a cost model of the critical path, only where
 * they keep the queue deadlock-free (checked on every load). */
#define SOGL_LINK_AFTER 1
#define SOGL_LINK_PANEL 2
int sogl_ctx_set_task_links(sogl_ctx* ctx, int32_t links);
/* Builds the executor's task graph of a description on the host only (no device is
 * touched) and checks it: every slot index inside the pool, one writer per tile,
 * every operand produced at an earlier dependency level and waited for, and the pool
 * slot reuse (see sogl_ctx_set_pool_reuse) handing a slot over only from a value to the
 * task that alone reads it.  sogl_factor_load runs the same checks before any upload.
 * info (may be NULL) receives 12 counts: tasks, value tiles (the pool without slot
 * reuse), dependency levels, GEMM pairs, dependencies, folded copy/negate/subtract ops,
 * aliased repeated inverses, lu / llt tasks chained to the GEMM that makes their source
 * (same workgroup, no hand-off), pool tiles with slot reuse (the _plan form's inputs
 * can be scattered again, so their slots are reused too), and the GEMM tasks (one pair
 * each) that make the solve's two-row blocks: per pair of block rows 2p, 2p+1, two for
 * the L block when L_2p+1,2p is present and two for the U block when U_2p,2p+1 is, and
 * the after-links and the absorbed panels (see sogl_ctx_set_task_links), which the check
 * builds whatever the context's setting. */
int sogl_graph_check(const sogl_factor_desc* desc, int64_t* info);
int sogl_graph_check_plan(const sogl_plan* plan, int32_t precision, int64_t* info);
/* Convenience: sogl_factor_load from a host plan. */
int sogl_factor_load_plan(sogl_ctx* ctx, const sogl_plan* plan, int32_t precision,
                          int32_t* factor_id);
/* Factor on the device: replaces BlockPlanner::calculate() (BlockPlanner.cpp:376-651),
 * called from decompose_solveLU (solver.cpp:106).  Re-running refactors from the
 * resident input tiles. */
int sogl_factor_run(sogl_ctx* ctx, int32_t factor_id, int32_t exec_mode);
/* Same with the persistent executor's queue settings: chunk = tasks taken per fetch
 * (1..64), wg_per_cu = resident workgroups per CU (0: as many as fit). */
int sogl_factor_run_config(sogl_ctx* ctx, int32_t factor_id, int32_t exec_mode, int32_t chunk,
                           int32_t wg_per_cu);
/* Autotuner: times each executor configuration `trials` times (best of) on this
 * factor, keeps the fastest for SOGL_EXEC_AUTO and returns its index.  Candidates, in
 * order: persistent chunk 1 / 2 / 4 at full occupancy, persistent chunk 1 at one
 * workgroup per CU, levels, captured level graph.  times_ms (may be NULL) receives
 * SOGL_TUNE_CANDIDATES executor times.  The factor is left computed. */
#define SOGL_TUNE_CANDIDATES 6
int sogl_factor_autotune(sogl_ctx* ctx, int32_t factor_id, int32_t trials, double* times_ms,
                         int32_t* best);
/* Block-mask policy of a factor (SOGL_MASKS_*); takes effect at the next factor run.
 * No reference counterpart beyond the masks themselves. */
int sogl_factor_set_masks(sogl_ctx* ctx, int32_t factor_id, int32_t mode);
/* Triangular solve: replaces BlockPlanner::solve() (BlockPlanner.cpp:834-862), called
 * from decompose_solveLU (solver.cpp:115).  bp, xp: host arrays in the permuted padded
 * order, 64*block_rows x nrhs column-major.  y (may be NULL) receives what the
 * reference leaves in data::b: the forward substitution result y of an LU factor, and
 * diag(L) y for an L L^T factor (lowerSolver stores b_i - sum before dividing by l_ii,
 * BlockPlanner.cpp:741-767).  An L L^T factor (op lists with llt ops, symmetric input)
 * solves L y = b, then L^T x = y (upperSolverT, BlockPlanner.cpp:769-797). */
int sogl_solve(sogl_ctx* ctx, int32_t factor_id, const double* bp, double* xp, double* y,
               int32_t nrhs);
/* Same, on device-resident arrays (hipMalloc'd by the caller) on the ctx stream. */
int sogl_solve_device(sogl_ctx* ctx, int32_t factor_id, const double* d_bp, double* d_xp,
                      int32_t nrhs);
int sogl_factor_release(sogl_ctx* ctx, int32_t factor_id);

/* No counterpart in the reference.  sogl_solve_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same kernels, but the solve's intermediates (the forward
 * result y, per-block-row flags and control words) live in a fixed per-GPU scratch
 * array of SOGL_NOALLOC_SCRATCH_BYTES instead of buffers the factor allocates, and the
 * call allocates nothing.  It needs 256 + 4*block_rows (rounded up to a multiple of
 * 256) + 512*block_rows*nrhs bytes of the array; a larger solve returns SOGL_ERR_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (several
 * host threads or contexts) race.  The factor must be loaded and factored (loading
 * allocates). */
#define SOGL_NOALLOC_SCRATCH_BYTES (64u << 20)
int sogl_solve_device_noalloc(sogl_ctx* ctx, int32_t factor_id, const double* d_bp, double* d_xp,
                              int32_t nrhs);

typedef struct sogl_factor_stats {
    double factor_ms;       /* last sogl_factor_run, device time (executor + diagonal inverses) */
    double exec_ms;         /* last sogl_factor_run, executor kernel(s) alone */
    double solve_ms;        /* last sogl_solve*, device time */
    int64_t pool_bytes;     /* HBM held by the tile pool */
    int64_t n_tasks;        /* device tasks (one per result tile) */
    int32_t n_levels;       /* dependency levels of the task graph */
    int32_t growth_tiles;   /* low precision only: diagonal tiles factored in FP64 */
    double max_growth;      /* largest pivot-free growth factor seen */
    double zero_blocks;     /* share of the factor tiles' 16x16 blocks that are zero (last masked run) */
    int32_t block_masks;    /* 1 when the next run uses block masks */
    int32_t claimed_tasks;  /* last persistent run: tasks run by the workgroup that made an operand */
    int32_t dropped_tasks;  /* last persistent run: queued tasks given up because another claimed them */
    int32_t tasks_run;      /* last run: tasks executed (n_tasks when each ran exactly once) */
    int64_t pool_tiles;     /* 64x64 tiles in the pool (pool_bytes / tile bytes) */
    int64_t value_tiles;    /* tiles the factor's values would need without slot reuse */
    int32_t uinv_out_of_tol; /* last run: diagonal U tiles whose inverse fails the check below */
} sogl_factor_stats;
int sogl_factor_get_stats(sogl_ctx* ctx, int32_t factor_id, sogl_factor_stats* st);
/* The reference's check of every upper inverse (inv_check_diag, MatrixStdDouble.cpp:
 * 2871-2937, run after each upperInv at BlockPlanner.cpp:573-577, which prints
 * "diag min:..  diag max:.." or "middle min:..  middle max:.." and " upper out of
 * tolerance " and goes on): after each LU factor run the diagonal of U_ii U_ii^-1 must
 * lie within 1e-3 of 1 and its entries (i, i/2) within 1e-3 of 0 (a NaN counts as 10).
 * n_failed receives the number of block rows that fail; detail (may be NULL; up to 16 x
 * 4 doubles) receives block row, 0 (diagonal) / 1 (middle), min and max of each of the
 * first 16 recorded.  L L^T factors make no upper inverses and report 0. */
int sogl_factor_get_uinv_check(sogl_ctx* ctx, int32_t factor_id, int32_t* n_failed, double* detail);
/* Dynamic chaining of the persistent executor: a workgroup that has made a tile runs a
 * near-critical GEMM task that reads it right away, when that task's other operands are
 * already published, instead of handing it over through memory.  claims = the most such
 * tasks one finished task hands to its own workgroup (0: off, the default; at most 8).  No
 * reference counterpart (BlockPlanner::calculate runs stage by stage); the results are
 * bit-identical either way. */
int sogl_factor_set_chaining(sogl_ctx* ctx, int32_t factor_id, int32_t claims);
/* Copy factor tiles back (FP64 view).  which 0 = L, 1 = U, in the order given at load. */
int sogl_factor_get_tiles(sogl_ctx* ctx, int32_t factor_id, int32_t which, double* tiles);

/* Diagnostics: per-task device timestamps of the persistent executor.
 * tasks4[4t..]: kind (0 gemm, 1 sub, 2 lu, 3 lower inverse, 4 upper inverse), pair
 * count, dependency level, dependency count; times4[4t..]: wall-clock ticks (100 MHz)
 * when the task was taken, when its work was done (before publishing), when its
 * outputs were published, and when its last dependency wait ended (GEMM tasks wait
 * pair by pair); bits 60-61 of that last entry say how the task was started (0 taken
 * from the queue, 1 run after the GEMM that made its only source, 2 claimed by the
 * workgroup that made one of its operands).  Arrays hold n_tasks * 4 entries. */
int sogl_factor_set_trace(sogl_ctx* ctx, int32_t factor_id, int32_t enable);
int sogl_factor_get_trace(sogl_ctx* ctx, int32_t factor_id, int64_t* tasks4, int64_t* times4);
/* Diagnostics: the executor's task graph.  n_deps / n_slots receive the sizes; with
 * dep_begin (n_tasks + 1 entries) and deps (n_deps entries) given, task t waits for the
 * pool slots deps[dep_begin[t] .. dep_begin[t+1]); slot_task (n_slots entries) receives
 * the task that writes each slot (-1 for inputs).  Pass NULL arrays to query sizes. */
int sogl_factor_get_graph(sogl_ctx* ctx, int32_t factor_id, int64_t* n_deps, int64_t* n_slots,
                          int32_t* dep_begin, int32_t* deps, int32_t* slot_task);

/* ------------------------------------------------ mixed precision + refinement */
/* Attach the original (unpermuted) matrix for the FP64 residual r = b - A x. */
int sogl_refine_attach(sogl_ctx* ctx, int32_t factor_id, const sogl_plan* plan);
/* Solve with FP64 iterative refinement in the original order: x = A^-1 b until the
 * normwise backward error is <= tol in both forms, ||b-Ax||_inf / (||A||_inf ||x||_inf +
 * ||b||_inf) and ||b-Ax||_2 / (||A||_F ||x||_2 + ||b||_2), at most max_iter corrections;
 * *resid receives the larger of the two (the worst column).  b, x: n x nrhs column-major
 * host arrays. */
int sogl_solve_refined(sogl_ctx* ctx, int32_t factor_id, const double* b, double* x,
                       int32_t nrhs, double tol, int32_t max_iter, int32_t* iters,
                       double* resid);

/* Same on device-resident arrays (original order, n x nrhs column-major, allocated
 * by the caller, e.g. with sogl_device_alloc): the permutation, padding, solves,
 * FP64 residuals and corrections all stay in HBM.  max_iter = 0 gives a plain
 * solve plus its residual. */
int sogl_solve_refined_device(sogl_ctx* ctx, int32_t factor_id, const double* d_b, double* d_x,
                              int32_t nrhs, double tol, int32_t max_iter, int32_t* iters,
                              double* resid);

/* ------------------------------------------------ device memory helpers */
int sogl_device_alloc(sogl_ctx* ctx, size_t bytes, void** ptr);
int sogl_device_free(sogl_ctx* ctx, void* ptr);
/* to_device != 0: host -> device, else device -> host; synchronous */
int sogl_device_copy(sogl_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t to_device);

/* ------------------------------------------------ drop-in solveLU */
/* Mirrors SOGLU::solveLU (solver.h:24, solver.cpp:121-184): plan on the host,
 * factor and solve on device 0 in FP64, x (length dim) in the original order;
 * symmetric != 0 takes the L L^T path. */
int sogl_solveLU(int32_t dim, int64_t valcount, int32_t symmetric, const int32_t* index_i,
                 const int32_t* index_j, const double* vals, const double* b, double* x);
const char* sogl_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SOGL_H */
