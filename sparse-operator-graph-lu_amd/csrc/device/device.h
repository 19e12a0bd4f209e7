// Internal structures shared by the device translation units (not part of
// the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/sogl.h"

namespace sogl_dev {

struct Pair;  // tile_ops.h

enum TaskKind : int32_t { TK_GEMM = 0, TK_SUB = 1, TK_LU = 2, TK_LINV = 3, TK_UINV = 4, TK_LLT = 5, TK_TRANS = 6 };

// One device task = every unit op that writes one result tile (an lu writes two).
// Writers of one tile always share a stage in the reference's plan (the "fix"
// pass, BlockPlanner.cpp:269-276), so a tile is produced by exactly one task.
// Slots address tile data and block masks in the pool; value ids (f0..f3, fadd, the
// dependency list, Pair::fa / fb) index the ready flags.  Without slot reuse the two
// are the same numbers.
struct DevTask {
    int32_t kind;
    int32_t out0, out1;   // result slots (out1: U of an lu, else -1)
    int32_t a, b;         // GEMM: pair begin / pair count; SUB: minuend / subtrahend slot
                          // (-1 = zero); LU, LINV, UINV, LLT, TRANS: source slot
    int32_t dep_begin, dep_count;  // value ids this task waits for (persistent executor)
    int32_t level;        // dependency level (0-based)
    int32_t out2, out3;   // lu: slots of L^-1 and U^-1 made by the same task, llt: of L^-1
                          // (-1: none)
    int32_t add = -1;     // GEMM: slot of a tile added to the sum (-1: none)
    int32_t add_neg = 0;  // GEMM: the added tile is subtracted instead
    int32_t next = -1;    // persistent executor: the task the same workgroup runs right after
                          // this one (the lu / llt that alone reads this GEMM's sum), or -1
    // Dynamic chaining (persistent executor).  ctag: byte l (pair operand lane l = 2j or
    // 2j+1 of pair j) holds 1 + this task's index in the candidate list of the task that
    // makes that operand, 0 if none; chk_begin / chk_count: this task's own candidates,
    // as check entries (ChainChk), one per candidate operand.
    int32_t ctag = 0;
    int32_t chk_begin = 0, chk_count = 0;
    int32_t f0 = -1, f1 = -1, f2 = -1, f3 = -1;  // value ids of out0..out3 (-1: none)
    int32_t fadd = -1;                            // value id of the added tile
    int32_t fsa = -1, fsb = -1;                   // value ids of a / b (other than GEMM)
    // slot reuse: before storing its result, a task whose result slot held another value
    // waits for the tasks that read that value; their flags are war[war_begin ..
    // war_begin + war_count) (value ids)
    int32_t war_begin = 0, war_count = 0;
    // persistent executor: the task the same workgroup runs after this one has published
    // (a GEMM on the critical path that reads an output of this one; it waits for its
    // other operands like a queued task and is not handed out by the queue), or -1
    int32_t after = -1;
    // lu (FP32 LU factors, persistent executor): the panel GEMMs this task runs inside
    // itself, before publishing, from the inverses it keeps in LDS: pan_l = A U^-1 (its
    // U^-1 the second operand), pan_u = L^-1 A (its L^-1 the first); -1: none
    int32_t pan_l = -1, pan_u = -1;
    int32_t pad[3] = {0, 0, 0};
};

// One check entry of a task's chain candidates: candidate `task` (index `cand` in the
// list) reads `slot` at operand position `pos` (0-3: pair lanes, 4: the added tile);
// slot -1: an output of the checking task itself.
struct ChainChk {
    int32_t slot, task, cand, pos;
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Factor {
    int32_t id = 0;
    int32_t precision = SOGL_FP64;
    int32_t block_rows = 0;
    int64_t n_slots = 0;        // values: inputs + results + zero + diagonal inverses (one flag each)
    int64_t n_locs = 0;         // tiles in the pool: n_slots, or fewer when slots are reused
    bool reuse = false;         // a result may take over the slot of a value only its task reads
    bool rescatter = false;     // input slots are taken over too: the inputs are scattered
                                // again from the cells kept on the device before every run
    int64_t n_inputs = 0;
    DevBuf cell_begin, cell_pos, cell_val;  // rescatter: the input cells
    std::vector<int32_t> war;   // write-after-read flags of reused result slots (DevTask)
    DevBuf war_d;
    DevBuf uinv_chk_d;                // the upper inverses' check of the last run (k_check_uinv)
    std::vector<double> uinv_check;   // its host copy
    int32_t zero_slot = -1;
    size_t elem = 8;            // bytes per tile element in the pool
    // host copies
    std::vector<DevTask> tasks;            // topological order (by level, then stage)
    std::vector<int32_t> level_begin;      // tasks of level v: [level_begin[v], level_begin[v+1])
    std::vector<DevTask> op_tasks;         // per-op executor: one entry per unit op, list order
    std::vector<Pair> op_pairs;            // and their pairs (uploaded on the first per-op run)
    bool keep_ops = true;                  // build those tables (not when reuse is forced on)
    std::vector<int32_t> input_slots;
    int64_t n_pairs = 0, n_deps = 0;
    int64_t n_inv_aliased = 0;  // repeated inverses of one factor tile that share its slot
    int64_t n_folded = 0;       // copies, negations and subtractions folded into other tasks
    int64_t n_chained = 0;      // lu / llt tasks run by the workgroup that made their source
    int64_t n_pair_tasks = 0;   // GEMM tasks making the solve's two-row blocks X_p / Y_p
    // device arrays
    DevBuf pool, tasks_d, pairs_d, deps_d, ready_d, ctrl_d, op_tasks_d, op_pairs_d, inputs_d, queue_d, chk_d;
    std::vector<ChainChk> chk;             // chain candidates' check entries (see DevTask)
    int64_t n_chain_cands = 0;             // tasks that are some task's chain candidate
    bool after_links = false;              // build after-links (set from the context at load)
    bool absorb_panels = false;            // let lu tasks absorb a panel GEMM (set from the context)
    int64_t n_absorbed = 0;                // panel GEMMs absorbed by their lu task
    int64_t n_after = 0;                   // after-links (tasks run by their producer's workgroup)
    std::vector<int32_t> chain_head;       // per task: the queued task whose workgroup runs it
    int32_t chain_claims = 0;              // persistent executor: claims per finished task (0: off)
    std::vector<int32_t> queue;            // persistent executor: task indices handed out, in order
    DevBuf mask_d;                         // per slot: 16x16 block mask (uint32)
    int32_t epoch = 0;
    // triangular solve structure (block CSR by block row, off-diagonal only)
    std::vector<int32_t> l_diag, u_diag, linv_slot, uinv_slot;  // per block row
    DevBuf l_rowptr, l_col, l_slot, u_rowptr, u_col, u_slot, l_diag_d, u_diag_d, linv_d, uinv_d;
    DevBuf pairx_d, pairy_d;  // per row pair: slot of X_p / Y_p (the two-row solve steps), -1: zero
    DevBuf solve_ready, solve_ctrl, vec_b, vec_x, vec_y;
    size_t vec_cap = 0;  // doubles per vector buffer
    int32_t solve_epoch = 0;
    // factor tile lists in load order
    std::vector<int32_t> l_slots, u_slots;
    // refinement
    bool has_refine = false;
    int32_t n = 0;                // unknowns of the original problem
    int64_t nnz = 0;
    std::vector<int32_t> new_order;  // new -> old
    double a_fro = 0, a_inf = 0;  // ||A||_F and ||A||_inf of the original matrix
    DevBuf csr_ptr, csr_col, csr_val, vec_r, vec_d, red, perm_d, vec_io;
    // stats
    double factor_ms = 0, solve_ms = 0, exec_ms = 0;
    int32_t growth_tiles = 0;
    int32_t claimed = 0;  // last persistent run: tasks run by the workgroup that claimed them
    int32_t dropped = 0;  // last persistent run: queued tasks dropped because claimed
    int32_t ran = 0;      // last run: tasks run (the task count when each ran once)
    double max_growth = 0;
    double growth_limit = 1e4;
    bool trace_on = false;
    DevBuf trace_d;
    // captured level schedule (SOGL_EXEC_GRAPH) and the autotuner's choice
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int32_t tuned = -1;  // index into the tuning candidates, -1: not tuned
    bool llt = false;    // L L^T factor: the U tiles of the solve are transposes of L tiles
    // 16x16 block masks: SOGL_MASKS_OFF / ON / AUTO (masked first run, then kept when a
    // quarter or more of the factor's blocks are zero)
    int32_t mask_mode = SOGL_MASKS_AUTO;
    bool masks_on = true, masks_decided = false, graph_masks = false;
    double zero_blocks = 0;  // share of zero blocks in the last masked run's factor tiles
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ex0 = nullptr, ex1 = nullptr;
    int num_cu = 0;
    int32_t pool_reuse = SOGL_REUSE_AUTO;  // for the factors loaded from now on (sogl_ctx_set_pool_reuse)
    bool after_links = false;              // for the factors loaded from now on (sogl_ctx_set_task_links)
    bool absorb_panels = true;             // for the factors loaded from now on (sogl_ctx_set_task_links)
    size_t free_mem = 0, total_mem = 0;    // device HBM when the context was made
    std::string err;
    std::map<int32_t, Factor*> factors;
    int32_t next_id = 1;
    std::mutex mu;
};

// helpers implemented in factor.hip
int set_error(Context* c, int code, const std::string& msg);
int check_hip(Context* c, hipError_t e, const char* what);
int dev_alloc(Context* c, DevBuf& b, size_t bytes, const char* what);
void dev_free(DevBuf& b);
Factor* find_factor(Context* c, int32_t id);
// implemented in solve.hip: looks up the device address of the no-allocation solve's
// scratch array on `device` (current), once per device
void noalloc_scratch_init(int device);

}  // namespace sogl_dev

struct sogl_ctx {
    sogl_dev::Context c;
};
