// Two-level block planner: recursive 2x2 block LU with explicit triangular
// inverses, emitted as 64x64 unit operations, then liveness pruning, ASAP
// staging, sorting and band splitting (BlockPlanner.cpp:73-374, 865-1577,
// solver.cpp:50-100, config.cpp:30-49).
//
// The op list, the tile ids, the sequence numbers and the stages must equal
// the reference's, so every counter side effect of the reference is kept:
// operation::seq advances once per emitted op (memutil.cpp:297-299), twice
// per op kept by a compaction (operation.cpp:117 + BlockPlanner.cpp:216) and
// once per coarse op re-created by the L2 expansion (BlockPlanner.cpp:1362).
// The counter starts at 0 for every plan, which is the reference's state in
// its one-solve-per-process use.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "planner.h"

namespace sogl {

void gps_reorder(int n, std::vector<int32_t>& prow, std::vector<int32_t>& pcol,
                 std::vector<int32_t>& newOrder, std::vector<int32_t>& reverseOrder,
                 int& outLevels, int& outWidth, int& outStart, bool& symmetricPattern,
                 int threads, std::string& log);
bool sort_in_block(int n, int blockSize, std::vector<int32_t>& prow, std::vector<int32_t>& pcol,
                   const std::vector<double>& pval, std::vector<double>* bperm,
                   std::vector<int32_t>& newOrder, std::vector<int32_t>& reverseOrder);
bool block_needs_sort(const double* diag, int i, int n, int blockSize);

// config::set + roundup (config.cpp:30-49)
SizeConfig SizeConfig::set(int32_t dim) {
    SizeConfig c;
    c.mSize = dim;
    c.blockSize = kTile;
    int32_t m = dim / c.blockSize, r = 1;
    while (m > 0) { m /= 2; r *= 2; }  // strictly greater power of two
    c.blockRows = r;
    c.blockRowsL2 = 2 << (__builtin_popcount((unsigned)(c.blockRows - 1)) / 2);
    c.blockSizeL2 = c.blockSize * c.blockRows / c.blockRowsL2;
    return c;
}

namespace {

inline int ilog2(int v) { return __builtin_popcount((unsigned)(v - 1)); }

struct Planner {
    std::deque<Node> arena;
    int32_t seq = 0;        // operation::seq
    int32_t storage = 1;    // data::blockstorage.size(); id 0 is the null tile
    bool planL2 = false;    // data::PlanL2
    std::vector<sogl_op> graph;
    std::vector<int32_t> stage, laststage;
    int32_t maxStage = 0;
    std::string* log = nullptr;
    int threads = 1;        // OpenMP threads for the per-stage op sort

    Node* node(int32_t blockrows2, int32_t level) {  // memutil::newmatrix
        arena.emplace_back();
        Node* p = &arena.back();
        p->blockrows2 = blockrows2;
        p->level = level;
        return p;
    }
    int32_t append_storage() { return storage++; }  // appendBlockStorage(NULL)
    void emit(int32_t src, int32_t src2, int32_t op, int32_t r, int32_t r2, int32_t group) {
        graph.push_back(sogl_op{src, src2, op, r, r2, 0, seq, group});  // memutil::newoperation
        seq++;
    }

    // matrix::set (matrix.cpp:62-86)
    static void tree_set(Node* root, Planner& P, int64_t key, int64_t val) {
        Node* cur = root;
        uint32_t mask = (1u << root->level) - 1;
        uint32_t row = (uint32_t)(key >> root->level);
        uint32_t col = (uint32_t)(key & mask);
        int step = root->level;
        while (step > 0) {
            step--;
            uint32_t idx = ((row >> step) << 1) + (col >> step);
            if (cur->sub[idx] == nullptr) cur->sub[idx] = P.node(cur->blockrows2 / 2, cur->level - 1);
            cur = cur->sub[idx];
            mask >>= 1;
            row &= mask;
            col &= mask;
        }
        cur->id = val;
    }
    // matrix::operator[] (matrix.cpp:34-60); 0 when absent
    static int64_t tree_get(Node* root, int64_t key) {
        Node* cur = root;
        uint32_t mask = (1u << root->level) - 1;
        uint32_t row = (uint32_t)(key >> root->level);
        uint32_t col = (uint32_t)(key & mask);
        int step = root->level;
        while (step > 0) {
            step--;
            uint32_t idx = ((row >> step) << 1) + (col >> step);
            cur = cur->sub[idx];
            mask >>= 1;
            row &= mask;
            col &= mask;
            if (cur == nullptr) return 0;
        }
        return cur->id;
    }

    static void need(const Node* p, const char* what) {
        if (p == nullptr)
            throw std::runtime_error(std::string("structurally singular block (missing ") + what +
                                     "): the pivot-free block LU needs every diagonal block");
    }

    // ---------------------------------------------------------- emitters
    // blockMatrixMul / MulNeg (BlockPlanner.cpp:1230-1283): c (+/-)= a*b over present children
    void mul(Node* a, Node* b, Node* c, int32_t group, int32_t op) {
        if (a == nullptr || b == nullptr) return;
        if (a->level == 0) {
            if (a->id > 0 && b->id > 0) {
                if (c->id <= 0) c->id = append_storage();
                emit((int32_t)a->id, (int32_t)b->id, op, (int32_t)c->id, 0, group);
            }
            return;
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    if (a->sub[i * 2 + k] != nullptr && b->sub[k * 2 + j] != nullptr) {
                        if (c->sub[i * 2 + j] == nullptr)
                            c->sub[i * 2 + j] = node(c->blockrows2 / 2, c->level - 1);
                        mul(a->sub[i * 2 + k], b->sub[k * 2 + j], c->sub[i * 2 + j], group, op);
                    }
                }
    }
    // blockMatrixCopy (BlockPlanner.cpp:1158-1179): sub with src = 0
    void copy(Node* a, Node* c, int32_t group) {
        if (a->level == 0) {
            if (a->id > 0) {
                if (c->id <= 0) c->id = append_storage();
                emit(0, (int32_t)a->id, SOGL_OP_SUB, (int32_t)c->id, 0, group);
            }
            return;
        }
        for (int i = 0; i < 4; i++)
            if (a->sub[i] != nullptr) {
                if (c->sub[i] == nullptr) c->sub[i] = node(c->blockrows2 / 2, c->level - 1);
                copy(a->sub[i], c->sub[i], group);
            }
    }
    // blockMatrixNeg (BlockPlanner.cpp:1180-1201): sub with src2 = 0
    void neg(Node* b, Node* c, int32_t group) {
        if (b->level == 0) {
            if (b->id > 0) {
                if (c->id <= 0) c->id = append_storage();
                emit((int32_t)b->id, 0, SOGL_OP_SUB, (int32_t)c->id, 0, group);
            }
            return;
        }
        for (int i = 0; i < 4; i++)
            if (b->sub[i] != nullptr) {
                if (c->sub[i] == nullptr) c->sub[i] = node(c->blockrows2 / 2, c->level - 1);
                neg(b->sub[i], c->sub[i], group);
            }
    }
    // blockMatrixSub (BlockPlanner.cpp:1126-1156): c = a - b, emitted as sub(src=b, src2=a)
    void sub(Node* a, Node* b, Node* c, int32_t group) {
        if (a->level == 0) {
            if (a->id > 0 || b->id > 0) {
                if (c->id <= 0) c->id = append_storage();
                emit((int32_t)b->id, (int32_t)a->id, SOGL_OP_SUB, (int32_t)c->id, 0, group);
            }
            return;
        }
        for (int i = 0; i < 4; i++) {
            if (a->sub[i] != nullptr || b->sub[i] != nullptr)
                if (c->sub[i] == nullptr) c->sub[i] = node(c->blockrows2 / 2, c->level - 1);
            if (a->sub[i] != nullptr && b->sub[i] != nullptr) sub(a->sub[i], b->sub[i], c->sub[i], group);
            else if (a->sub[i] != nullptr) copy(a->sub[i], c->sub[i], group);
            else if (b->sub[i] != nullptr) neg(b->sub[i], c->sub[i], group);
        }
    }
    // blockMatrixInvLower (BlockPlanner.cpp:1019-1054): y = l^-1 for block lower l;
    // a3, when given, already holds the inverse of l's top-left quadrant
    void inv_lower(Node* l, Node* y, Node* a3, int32_t group) {
        if (l->level == 0) {
            if (l->id > 0) {
                y->id = append_storage();
                emit((int32_t)l->id, 0, SOGL_OP_LOWER_INV, (int32_t)y->id, 0, group);
            }
            return;
        }
        Node* a = l->sub[0];
        Node* c = l->sub[2];
        Node* d = l->sub[3];
        need(a, "L11");
        need(d, "L22");
        Node* d1 = node(a->blockrows2, a->level);
        inv_lower(d, d1, nullptr, group);
        y->sub[3] = d1;
        Node* d1c = node(a->blockrows2, a->level);
        mul(d1, c, d1c, group, SOGL_OP_MUL);
        Node* a1 = a3;
        if (a1 == nullptr) {
            a1 = node(a->blockrows2, a->level);
            inv_lower(a, a1, nullptr, group);
        }
        y->sub[0] = a1;
        Node* c21 = node(a->blockrows2, a->level);
        mul(d1c, a1, c21, group, SOGL_OP_MULNEG);
        y->sub[2] = c21;
    }
    // blockMatrixInvUpper (BlockPlanner.cpp:1056-1092)
    void inv_upper(Node* u, Node* y, Node* a3, int32_t group) {
        if (u->level == 0) {
            if (u->id > 0) {
                y->id = append_storage();
                emit((int32_t)u->id, 0, SOGL_OP_UPPER_INV, (int32_t)y->id, 0, group);
            }
            return;
        }
        Node* a = u->sub[0];
        Node* b = u->sub[1];
        Node* d = u->sub[3];
        need(a, "U11");
        need(d, "U22");
        Node* a2 = a3;
        if (a2 == nullptr) {
            a2 = node(a->blockrows2, a->level);
            inv_upper(a, a2, nullptr, group);
        }
        y->sub[0] = a2;
        Node* a1b = node(a->blockrows2, a->level);
        mul(a2, b, a1b, group, SOGL_OP_MUL);
        Node* d1 = node(a->blockrows2, a->level);
        inv_upper(d, d1, nullptr, group);
        y->sub[3] = d1;
        Node* a1bd1 = node(a->blockrows2, a->level);
        mul(a1b, d1, a1bd1, group, SOGL_OP_MULNEG);
        y->sub[1] = a1bd1;
    }
    // blockMatrixLU (BlockPlanner.cpp:865-939): a = l*u; l3/u3, when given, receive the
    // inverses of the top-left quadrant's factors for the caller to reuse
    void lu(Node* a, Node* l, Node* u, Node* l3, Node* u3, int32_t group) {
        if (a->level == 0) {
            if (a->id > 0) {
                if (l->id <= 0) l->id = append_storage();
                if (u->id <= 0) u->id = append_storage();
                emit((int32_t)a->id, 0, SOGL_OP_LU, (int32_t)l->id, (int32_t)u->id, group);
            }
            return;
        }
        const int32_t br = a->blockrows2, lv = a->level;
        Node* ha = a->sub[0];
        need(ha, "A11");
        Node* l1 = node(br / 2, lv - 1);
        Node* u1 = node(br / 2, lv - 1);
        Node* l2 = l3;
        Node* u2 = u3;
        if (l2 == nullptr) {
            l2 = node(br / 2, lv - 1);
            u2 = node(br / 2, lv - 1);
        }
        Node* hl1 = nullptr;
        Node* hu1 = nullptr;
        if (lv > 1) {
            hl1 = node(br / 4, lv - 2);
            hu1 = node(br / 4, lv - 2);
        }
        lu(ha, l1, u1, hl1, hu1, group);
        inv_lower(l1, l2, hl1, group);
        inv_upper(u1, u2, hu1, group);
        l->sub[0] = l1;
        u->sub[0] = u1;
        Node* aua1b = node(br / 2, lv - 1);
        Node* b = a->sub[1];
        if (b != nullptr) {
            mul(l2, b, aua1b, group, SOGL_OP_MUL);  // U12 = L11^-1 A12
            u->sub[1] = aua1b;
        }
        Node* d = a->sub[3];
        Node* c = a->sub[2];
        need(d, "A22");
        Node* dsub;
        if (c != nullptr) {
            Node* ca1al = node(br / 2, lv - 1);
            mul(c, u2, ca1al, group, SOGL_OP_MUL);  // L21 = A21 U11^-1
            l->sub[2] = ca1al;
            Node* ca1b = node(br / 2, lv - 1);
            mul(ca1al, aua1b, ca1b, group, SOGL_OP_MUL);  // L21 U12
            dsub = node(br / 2, lv - 1);
            sub(d, ca1b, dsub, group);  // Schur complement A22 - L21 U12
        } else {
            dsub = d;
        }
        l1 = node(br / 2, lv - 1);
        u1 = node(br / 2, lv - 1);
        Node* dl1 = nullptr;
        Node* du1 = nullptr;
        if (lv > 1) {
            dl1 = node(br / 4, lv - 2);
            du1 = node(br / 4, lv - 2);
        }
        lu(dsub, l1, u1, dl1, du1, group);
        l->sub[3] = l1;
        u->sub[3] = u1;
    }

    // blockMatrixMulT (BlockPlanner.cpp:1203-1229): c += a * b^T over present children
    void mul_t(Node* a, Node* b, Node* c, int32_t group) {
        if (a == nullptr || b == nullptr) return;
        if (a->level == 0) {
            if (a->id > 0 && b->id > 0) {
                if (c->id <= 0) c->id = append_storage();
                emit((int32_t)a->id, (int32_t)b->id, SOGL_OP_MULT, (int32_t)c->id, 0, group);
            }
            return;
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    if (a->sub[i * 2 + k] != nullptr && b->sub[j * 2 + k] != nullptr) {
                        if (c->sub[i * 2 + j] == nullptr)
                            c->sub[i * 2 + j] = node(c->blockrows2 / 2, c->level - 1);
                        mul_t(a->sub[i * 2 + k], b->sub[j * 2 + k], c->sub[i * 2 + j], group);
                    }
                }
    }
    // blockMatrixLLT (BlockPlanner.cpp:941-989): a = l l^T; l3, when given, receives
    // the inverse of the top-left quadrant's factor for the caller to reuse
    void llt(Node* a, Node* l, Node* l3, int32_t group) {
        if (a->level == 0) {
            if (a->id > 0) {
                if (l->id <= 0) l->id = append_storage();
                emit((int32_t)a->id, 0, SOGL_OP_LLT, (int32_t)l->id, 0, group);
            }
            return;
        }
        const int32_t br = a->blockrows2, lv = a->level;
        Node* ha = a->sub[0];
        need(ha, "A11");
        Node* l1 = node(br / 2, lv - 1);
        Node* l2 = l3 != nullptr ? l3 : node(br / 2, lv - 1);
        Node* hl1 = lv > 1 ? node(br / 4, lv - 2) : nullptr;
        llt(ha, l1, hl1, group);
        inv_lower(l1, l2, hl1, group);  // L11^-1
        l->sub[0] = l1;
        Node* d = a->sub[3];
        Node* c = a->sub[2];
        need(d, "A22");
        Node* dsub;
        if (c != nullptr) {
            Node* ca1 = node(br / 2, lv - 1);
            Node* ca2 = node(br / 2, lv - 1);
            mul_t(c, l2, ca1, group);  // L21 = A21 L11^-T
            l->sub[2] = ca1;
            mul_t(ca1, ca1, ca2, group);  // L21 L21^T
            dsub = node(br / 2, lv - 1);
            sub(d, ca2, dsub, group);  // Schur complement A22 - L21 L21^T
        } else {
            dsub = d;
        }
        l1 = node(br / 2, lv - 1);
        Node* dl1 = lv > 1 ? node(br / 4, lv - 2) : nullptr;
        llt(dsub, l1, dl1, group);
        l->sub[3] = l1;
    }

    // ---------------------------------------------------------- blockPlan
    void add_input(Node* a, std::vector<int32_t>& lastuse) {  // BlockPlanner.cpp:73-87
        if (a->level == 0 && a->id > 0) {
            stage[(size_t)a->id] = 1;
            laststage[(size_t)a->id] = 1;
            lastuse[(size_t)a->id] = 0;
        } else if (a->level > 0) {
            for (int i = 0; i < 4; i++)
                if (a->sub[i] != nullptr) add_input(a->sub[i], lastuse);
        }
    }
    void add_output(Node* a, std::vector<int32_t>& lastuse, int32_t marker) {  // :88-101
        if (a->level == 0 && a->id > 0) {
            laststage[(size_t)a->id] = marker;
            lastuse[(size_t)a->id] = marker;
        } else if (a->level > 0) {
            for (int i = 0; i < 4; i++)
                if (a->sub[i] != nullptr) add_output(a->sub[i], lastuse, marker);
        }
    }

    // BlockPlanner::blockPlan (BlockPlanner.cpp:103-374)
    void block_plan(Node* blocks, Node* saved, Node* savedu) {
        const int32_t S = storage;
        std::vector<int32_t> lastuse((size_t)S, 0);
        stage.assign((size_t)S, 0);
        laststage.assign((size_t)S, 0);
        stage[0] = 1;
        add_input(blocks, lastuse);
        if (saved) add_output(saved, lastuse, S);
        if (savedu) add_output(savedu, lastuse, S);

        // (ii) backward liveness: an op is live if any of its results is used later
        for (int64_t i = (int64_t)graph.size() - 1; i >= 0; i--) {
            const sogl_op& o = graph[(size_t)i];
            int32_t stg = 0;
            if (o.result > 0) {
                stg = o.result;
                if (o.result2 > 0 && o.result2 > stg) stg = o.result2;
            }
            if (o.result > 0 && o.result2 > 0)
                if (lastuse[(size_t)o.result] == 0 && lastuse[(size_t)o.result2] == 0) continue;
            if (o.result > 0 && o.result2 == 0)
                if (lastuse[(size_t)o.result] == 0) continue;
            if (stg > 0) {
                if (o.src > 0 && lastuse[(size_t)o.src] < stg) lastuse[(size_t)o.src] = stg;
                if (o.src2 > 0) {
                    if (o.src2 >= S) continue;
                    if (lastuse[(size_t)o.src2] < stg) lastuse[(size_t)o.src2] = stg;
                }
            }
        }
        // (iii) compaction: always on the coarse plan, on the fine plan only when more
        // than 1/25 of the ops are dead (BlockPlanner.cpp:180-221)
        int64_t waste = 0;
        for (const sogl_op& o : graph) {
            if (lastuse[(size_t)o.result] > 0) continue;
            if (o.result2 > 0 && lastuse[(size_t)o.result2] > 0) continue;
            ++waste;
        }
        if (planL2 || waste > (int64_t)(graph.size() / 25)) {
            std::vector<sogl_op> kept;
            kept.reserve(graph.size() - (size_t)waste);
            for (const sogl_op& o : graph) {
                bool live = lastuse[(size_t)o.result] > 0 ||
                            (o.result2 > 0 && lastuse[(size_t)o.result2] > 0);
                if (!live) continue;
                kept.push_back(o);
                seq += 2;  // operation::sett pre-increment + memutil::newoperation
            }
            graph.swap(kept);
        }
        if (log) *log += "reduced ops to: " + std::to_string(graph.size()) + "\n";

        // (iv) ASAP stage, at least 2 (BlockPlanner.cpp:225-267)
        int32_t maxstg = 1;
        for (sogl_op& o : graph) {
            int32_t stg = 0;
            if (o.src > 0) {
                stg = stage[(size_t)o.src] + 1;
                if (stg < 2) stg = 2;
                if (o.src2 > 0 && stage[(size_t)o.src2] + 1 > stg) stg = stage[(size_t)o.src2] + 1;
            } else {
                if (o.src2 > 0 && stage[(size_t)o.src2] + 1 > stg) stg = stage[(size_t)o.src2] + 1;
                if (stg < 2) stg = 2;
            }
            if (stg > 1) {
                o.stage = stg;
                if (o.result > 0 && stage[(size_t)o.result] < stg) stage[(size_t)o.result] = stg;
                if (o.result2 > 0 && stage[(size_t)o.result2] < stg) stage[(size_t)o.result2] = stg;
                if (stg > maxstg) maxstg = stg;
            }
        }
        // (v) single-result ops move to their result's final stage (:269-276)
        for (int64_t i = (int64_t)graph.size() - 1; i >= 0; i--) {
            sogl_op& o = graph[(size_t)i];
            if (o.result2 > 0) continue;
            if (o.result > 0 && stage[(size_t)o.result] > o.stage) o.stage = stage[(size_t)o.result];
        }
        // (vi) order by (stage, group, result, src, seq).  The reference buckets by
        // stage%700, sorts each bucket and merges by stage (:278-333); seq is unique,
        // so this is a total order and any sort gives the same sequence.  Here: a
        // counting pass by stage, then each stage's ops sorted on their own, the stages
        // spread over threads (3D 64^3: 7.1M ops, 1.3 s in one global sort).
        {
            std::vector<int64_t> at((size_t)maxstg + 2, 0);
            for (const sogl_op& o : graph) at[(size_t)o.stage + 1]++;
            for (size_t v = 1; v < at.size(); v++) at[v] += at[v - 1];
            std::vector<sogl_op> by((size_t)graph.size());
            {
                std::vector<int64_t> fill(at.begin(), at.end() - 1);
                for (const sogl_op& o : graph) by[(size_t)fill[(size_t)o.stage]++] = o;
            }
            const int64_t nst = (int64_t)at.size() - 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
            for (int64_t v = 0; v < nst; v++)
                std::sort(by.begin() + at[(size_t)v], by.begin() + at[(size_t)v + 1],
                          [](const sogl_op& x, const sogl_op& y) {
                              if (x.group != y.group) return x.group < y.group;
                              if (x.result != y.result) return x.result < y.result;
                              if (x.src != y.src) return x.src < y.src;
                              return x.seq < y.seq;
                          });
            graph.swap(by);
        }
        // (vii) band split: after every 8001 ops of one stage, later ops move one stage on
        int32_t jump = 0, band = 0, stgc = 0;
        for (sogl_op& o : graph) {
            if (o.stage > stgc) { stgc = o.stage; band = 0; }
            if (band > 8000) { jump++; band = 0; }
            o.stage += jump;
            band++;
        }
        // (viii) laststage = last reading stage (:356-371)
        for (const sogl_op& o : graph) {
            if (o.src > 0 && stage[(size_t)o.src] >= 0 && laststage[(size_t)o.src] < o.stage)
                laststage[(size_t)o.src] = o.stage;
            if (o.src2 > 0 && stage[(size_t)o.src2] >= 0 && laststage[(size_t)o.src2] < o.stage)
                laststage[(size_t)o.src2] = o.stage;
        }
        maxStage = graph.empty() ? 0 : graph.back().stage;
    }
};

// matrixZoomSet (BlockPlanner.cpp:1305-1315): coarse input leaf -> fine subtree
void zoom_set(Node* a, Node* detail, std::vector<Node*>& l2) {
    if (a->level == 0 && a->id > 0) {
        l2[(size_t)a->id] = detail;
    } else if (a->level > 0) {
        for (int i = 0; i < 4; i++)
            if (a->sub[i] != nullptr) {
                if (detail == nullptr || detail->sub[i] == nullptr)
                    throw std::runtime_error("coarse/fine tile trees disagree");
                zoom_set(a->sub[i], detail->sub[i], l2);
            }
    }
}
// matrixZoomUpdate (BlockPlanner.cpp:1316-1330): coarse factor tree -> fine factor tree
void zoom_update(Planner& P, Node* l, Node* bl2, std::vector<Node*>& l2) {
    if (l->level == 1) {
        for (int i = 0; i < 4; i++)
            if (l->sub[i] != nullptr && l->sub[i]->id > 0) bl2->sub[i] = l2[(size_t)l->sub[i]->id];
    } else if (l->level > 1) {
        for (int i = 0; i < 4; i++)
            if (l->sub[i] != nullptr) {
                bl2->sub[i] = P.node(l->blockrows2 / 2, l->level - 1);
                zoom_update(P, l->sub[i], bl2->sub[i], l2);
            }
    }
}

// collect the leaves of a factor tree with their block coordinates
void collect(const Node* t, int32_t r0, int32_t c0, int32_t size, std::vector<int32_t>& bi,
             std::vector<int32_t>& bj, std::vector<int32_t>& id) {
    if (t == nullptr) return;
    if (size == 1) {
        if (t->level != 0) throw std::runtime_error("factor tree depth mismatch");
        if (t->id > 0) { bi.push_back(r0); bj.push_back(c0); id.push_back((int32_t)t->id); }
        return;
    }
    int32_t h = size / 2;
    for (int i = 0; i < 4; i++) collect(t->sub[i], r0 + (i >> 1) * h, c0 + (i & 1) * h, h, bi, bj, id);
}

std::string blocks_line(int32_t blockRows, int32_t blockSize, int32_t mSize, size_t nops,
                        int32_t storage) {
    return "blocks: " + std::to_string(blockRows) + " blockSize: " + std::to_string(blockSize) +
           " inputSize: " + std::to_string(mSize) + " extend: " +
           std::to_string((int64_t)blockRows * blockSize) + " op count: " + std::to_string(nops) +
           " storage: " + std::to_string(storage) + "\n";
}

std::string fmt_seconds(const char* label, double s) {
    char buf[128];
    // std::ostream default formatting: 6 significant digits, %g style
    std::snprintf(buf, sizeof(buf), "%s%g\n", label, s);
    return buf;
}

}  // namespace

void build_plan(const Coo& a, const double* b, const PlanOptions& opt, Plan& plan) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    plan = Plan();
    const int32_t n = a.n;
    if (n < kTile) throw std::runtime_error("matrix dimension below 64 is outside the planner's sizing");
    for (size_t k = 0; k < a.val.size(); k++)
        if (a.row[k] < 0 || a.row[k] >= n || a.col[k] < 0 || a.col[k] >= n)
            throw std::runtime_error("COO index out of range");
    // symmetric input: every off-diagonal entry is followed by its mirror
    // (solver.cpp:136-149), and the factorisation is L L^T
    const bool sym = a.symmetric;
    Coo mirrored;
    if (sym) {
        mirrored.n = n;
        mirrored.symmetric = true;
        const size_t m = a.val.size();
        mirrored.row.reserve(2 * m);
        mirrored.col.reserve(2 * m);
        mirrored.val.reserve(2 * m);
        for (size_t k = 0; k < m; k++) {
            mirrored.row.push_back(a.row[k]);
            mirrored.col.push_back(a.col[k]);
            mirrored.val.push_back(a.val[k]);
            if (a.row[k] != a.col[k]) {
                mirrored.row.push_back(a.col[k]);
                mirrored.col.push_back(a.row[k]);
                mirrored.val.push_back(a.val[k]);
            }
        }
    }
    const Coo& in = sym ? mirrored : a;
    const size_t nnz = in.val.size();
    plan.symmetric = sym;
    std::string& log = plan.log;
    auto flush = [&](size_t from) {
        if (opt.verbose) { std::fputs(log.c_str() + from, stdout); std::fflush(stdout); }
    };

    // solveLU (solver.cpp:121-161): copy, size, pad b with 1.0
    plan.cfg = SizeConfig::set(n);
    const SizeConfig& cfg = plan.cfg;
    const int64_t N = (int64_t)cfg.blockRows * cfg.blockSize;
    plan.prow = in.row;
    plan.pcol = in.col;
    plan.pval = in.val;
    if (b) {
        plan.bPerm.assign((size_t)N, 1.0);
        for (int32_t i = 0; i < n; i++) plan.bPerm[(size_t)i] = b[i];
    }
    size_t mark = log.size();
    gps_reorder(n, plan.prow, plan.pcol, plan.newOrder, plan.reverseOrder, plan.gpsLevels,
                plan.gpsBandwidth, plan.gpsStart, plan.structureSymmetric, opt.threads, log);
    if (b) {  // updatewithlevel (GPSOrder.cpp:433-447)
        std::vector<double> t(plan.bPerm);
        for (int32_t i = 0; i < n; i++) plan.bPerm[(size_t)i] = t[(size_t)plan.newOrder[(size_t)i]];
    }
    plan.nnzOrig = (int64_t)nnz;
    plan.blockSorted = sort_in_block(n, cfg.blockSize, plan.prow, plan.pcol, plan.pval,
                                     b ? &plan.bPerm : nullptr, plan.newOrder, plan.reverseOrder);
    plan.tReorder = std::chrono::duration<double>(clk::now() - t0).count();
    log += fmt_seconds("re Order time: ", plan.tReorder);
    flush(mark);

    // decompose_solveLU planning (solver.cpp:50-100)
    auto tp = clk::now();
    mark = log.size();
    Planner P;
    P.log = &log;
    P.threads = opt.threads > 0 ? opt.threads : std::max(1, omp_get_max_threads());
    // coarse level: ids in COO order of first touch, no tile memory (BlockPlanner.cpp:1480-1490)
    P.planL2 = true;
    int32_t blockRows = cfg.blockRowsL2, blockSize = cfg.blockSizeL2;
    Node* cblocks = P.node(blockRows, ilog2(blockRows));
    {
        const int e = ilog2(blockSize);
        for (size_t k = 0; k < nnz; k++) {
            int64_t bi = plan.prow[k] >> e, bj = plan.pcol[k] >> e;
            int64_t key = bi * blockRows + bj;
            if (Planner::tree_get(cblocks, key) == 0) Planner::tree_set(cblocks, P, key, P.storage++);
        }
        for (int64_t i = n; i < N; i++) {  // padding rows: identity diagonal
            int64_t bi = i / blockSize;
            int64_t key = bi * blockRows + bi;
            if (Planner::tree_get(cblocks, key) == 0) Planner::tree_set(cblocks, P, key, P.storage++);
        }
    }
    const int levels = ilog2(blockRows);
    Node* bl = P.node(blockRows, levels);
    Node* bu = P.node(blockRows, levels);
    if (sym) {
        log += "symmetric\n";  // solver.cpp:52
        P.llt(cblocks, bl, nullptr, 0);
    } else {
        P.lu(cblocks, bl, bu, nullptr, nullptr, 0);
    }
    log += blocks_line(blockRows, blockSize, n, P.graph.size(), P.storage);
    P.block_plan(cblocks, bl, sym ? nullptr : bu);
    if (opt.keepCoarse) plan.coarseOps = P.graph;

    // L2 expansion (BlockPlanner::copyOperatorL2, BlockPlanner.cpp:1332-1460)
    P.planL2 = false;
    const int32_t coarseStorage = P.storage;
    blockRows = cfg.blockRows;
    blockSize = cfg.blockSize;
    const int32_t scaleL2 = cfg.blockSizeL2 / cfg.blockSize;
    const int32_t levelL2 = ilog2(scaleL2);
    Node* LL2 = P.node(blockRows, ilog2(blockRows));
    Node* UL2 = P.node(blockRows, ilog2(blockRows));
    std::vector<Node*> l2((size_t)coarseStorage, nullptr);
    P.storage = 1;  // checkBlock
    Node* fblocks = P.node(blockRows, ilog2(blockRows));
    {
        // fine tiles: cells sorted block-major with std::sort (cellCompare, operation.cpp:86-92);
        // ids follow first touch in that order, values scatter into dense tiles
        struct Cell { int row; int col; double val; uint32_t src; };
        if (nnz + (size_t)(N - n) >= ((size_t)1 << 32)) throw std::runtime_error("too many matrix entries");
        std::vector<Cell> cells(nnz);
        // The reference orders the cells by tile only (BlockPlanner.cpp:1498,
        // std::sort), so the order inside a tile is the sort's own and matters only where
        // the COO repeats a position: the value stored last wins. Without repeats any
        // order by tile stores the same tiles, and two stable counting passes (column
        // tile, then row tile) do it in linear time; with repeats the same std::sort runs.
        {
            const size_t nb = (size_t)blockRows + 1;
            std::vector<uint32_t> a(nnz), b(nnz);
            std::vector<int64_t> cnt(nb + 1);
            auto pass = [&](const int32_t* key, const std::vector<uint32_t>* in, std::vector<uint32_t>& out) {
                std::fill(cnt.begin(), cnt.end(), 0);
                for (size_t k = 0; k < nnz; k++) cnt[(size_t)(key[in ? (*in)[k] : k] >> 6) + 1]++;
                for (size_t t = 0; t < nb; t++) cnt[t + 1] += cnt[t];
                for (size_t k = 0; k < nnz; k++) {
                    const uint32_t e = in ? (*in)[k] : (uint32_t)k;
                    out[(size_t)cnt[(size_t)(key[e] >> 6)]++] = e;
                }
            };
            pass(plan.pcol.data(), nullptr, a);
            pass(plan.prow.data(), &a, b);
            bool repeats = false;
            std::vector<int64_t> seen(kTile * kTile, -1);
            int64_t group = -1, prevKey = -1;
            for (size_t k = 0; k < nnz && !repeats; k++) {
                const int32_t r = plan.prow[b[k]], c = plan.pcol[b[k]];
                const int64_t key = (int64_t)(r >> 6) * blockRows + (c >> 6);
                if (key != prevKey) { prevKey = key; group++; }
                int64_t& m = seen[(size_t)((r & 63) * kTile + (c & 63))];
                if (m == group) repeats = true;
                m = group;
            }
            if (!repeats) {
                for (size_t k = 0; k < nnz; k++) cells[k] = Cell{plan.prow[b[k]], plan.pcol[b[k]], plan.pval[b[k]], b[k]};
            } else {
                for (size_t k = 0; k < nnz; k++) cells[k] = Cell{plan.prow[k], plan.pcol[k], plan.pval[k], (uint32_t)k};
                std::sort(cells.begin(), cells.end(), [](const Cell& x, const Cell& y) {
                    if ((x.row >> 6) == (y.row >> 6)) return (x.col >> 6) < (y.col >> 6);
                    return (x.row >> 6) < (y.row >> 6);
                });
            }
        }
        std::vector<int64_t> tileOf;  // id -> input index
        auto tile_for = [&](int64_t bi, int64_t bj) -> int32_t {
            int64_t key = bi * blockRows + bj;
            int64_t id = Planner::tree_get(fblocks, key);
            if (id == 0) {
                id = P.storage++;
                Planner::tree_set(fblocks, P, key, id);
                plan.inputIds.push_back((int32_t)id);
                plan.inputBi.push_back((int32_t)bi);
                plan.inputBj.push_back((int32_t)bj);
                if ((int64_t)tileOf.size() <= id) tileOf.resize((size_t)id + 1, -1);
                tileOf[(size_t)id] = (int64_t)plan.inputIds.size() - 1;
            }
            return (int32_t)tileOf[(size_t)id];
        };
        // cells in scatter order: the sorted COO, then the padding diagonal
        const size_t npad = (size_t)(N - n);
        std::vector<int32_t> ck(nnz + npad);
        std::vector<uint16_t> cp(nnz + npad);
        std::vector<double> cv(nnz + npad);
        std::vector<uint32_t> cs(nnz + npad);
        for (size_t k = 0; k < nnz; k++) {
            const Cell& c = cells[k];
            ck[k] = tile_for(c.row >> 6, c.col >> 6);
            cp[k] = (uint16_t)((c.row & 63) * kTile + (c.col & 63));
            cv[k] = c.val;
            cs[k] = c.src;
        }
        for (int64_t i = n; i < N; i++) {
            int64_t bi = i / blockSize;
            int64_t ri = i - bi * blockSize;
            const size_t k = nnz + (size_t)(i - n);
            ck[k] = tile_for(bi, bi);
            cp[k] = (uint16_t)(ri * kTile + ri);
            cv[k] = 1.0;
            cs[k] = (uint32_t)k;  // the padding entries follow the COO in prow/pcol/pval
        }
        // group by input tile keeping scatter order (counting sort), then keep only the
        // last cell of each position, which is the value a dense scatter leaves
        const size_t ntiles = plan.inputIds.size();
        std::vector<int64_t> start(ntiles + 1, 0);
        for (int32_t k : ck) start[(size_t)k + 1]++;
        for (size_t t = 0; t < ntiles; t++) start[t + 1] += start[t];
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        std::vector<int64_t> order(ck.size());
        for (size_t k = 0; k < ck.size(); k++) order[(size_t)fill[(size_t)ck[k]]++] = (int64_t)k;
        plan.inputCellBegin.assign(ntiles + 1, 0);
        plan.cellPos.reserve(ck.size());
        plan.cellVal.reserve(ck.size());
        plan.cellSrc.reserve(ck.size());
        std::vector<int32_t> lastAt(kTile * kTile, -1);
        for (size_t t = 0; t < ntiles; t++) {
            const int64_t b0 = start[t], b1 = start[t + 1];
            for (int64_t q = b0; q < b1; q++) lastAt[cp[(size_t)order[(size_t)q]]] = (int32_t)q;
            for (int64_t q = b0; q < b1; q++) {
                const size_t k = (size_t)order[(size_t)q];
                if (lastAt[cp[k]] != (int32_t)q) continue;  // overwritten later
                plan.cellPos.push_back(cp[k]);
                plan.cellVal.push_back(cv[k]);
                plan.cellSrc.push_back(cs[k]);
            }
            for (int64_t q = b0; q < b1; q++) lastAt[cp[(size_t)order[(size_t)q]]] = -1;
            plan.inputCellBegin[t + 1] = (int64_t)plan.cellPos.size();
        }
    }
    zoom_set(cblocks, fblocks, l2);
    // re-create every coarse op (one seq step each), then expand in staged order
    std::vector<sogl_op> coarse;
    coarse.swap(P.graph);
    P.seq += (int32_t)coarse.size();
    for (const sogl_op& o : coarse) {
        if (o.result > 0 && l2[(size_t)o.result] == nullptr) l2[(size_t)o.result] = P.node(scaleL2, levelL2);
        if (o.result2 > 0 && l2[(size_t)o.result2] == nullptr) l2[(size_t)o.result2] = P.node(scaleL2, levelL2);
        Node* s1 = o.src > 0 ? l2[(size_t)o.src] : nullptr;
        Node* s2 = o.src2 > 0 ? l2[(size_t)o.src2] : nullptr;
        if ((o.src > 0 && s1 == nullptr) || (o.src2 > 0 && s2 == nullptr))
            throw std::runtime_error("coarse op reads a tile that was never produced");
        Node* r1 = o.result > 0 ? l2[(size_t)o.result] : nullptr;
        Node* r2 = o.result2 > 0 ? l2[(size_t)o.result2] : nullptr;
        const int32_t g = o.seq;
        switch (o.op) {
            case SOGL_OP_LOWER_INV: P.inv_lower(s1, r1, nullptr, g); break;
            case SOGL_OP_UPPER_INV: P.inv_upper(s1, r1, nullptr, g); break;
            case SOGL_OP_LU: P.lu(s1, r1, r2, nullptr, nullptr, g); break;
            case SOGL_OP_LLT: P.llt(s1, r1, nullptr, g); break;
            case SOGL_OP_MULT: P.mul_t(s1, s2, r1, g); break;
            case SOGL_OP_MUL: P.mul(s1, s2, r1, g, SOGL_OP_MUL); break;
            case SOGL_OP_MULNEG: P.mul(s1, s2, r1, g, SOGL_OP_MULNEG); break;
            case SOGL_OP_SUB:
                if (o.src > 0) {
                    if (o.src2 > 0) P.sub(s2, s1, r1, g);
                    else P.neg(s1, r1, g);
                } else if (o.src2 > 0) {
                    P.copy(s2, r1, g);
                }
                break;
            default:
                throw std::runtime_error("unexpected coarse op " + std::to_string(o.op));
        }
    }
    if (!sym) zoom_update(P, bu, UL2, l2);
    zoom_update(P, bl, LL2, l2);
    log += blocks_line(blockRows, blockSize, n, P.graph.size(), P.storage);
    P.block_plan(fblocks, LL2, UL2);
    plan.ops.swap(P.graph);
    plan.storageCount = P.storage;
    plan.laststage.swap(P.laststage);
    plan.tileStage.swap(P.stage);
    plan.maxStage = P.maxStage;
    collect(LL2, 0, 0, blockRows, plan.lBi, plan.lBj, plan.lId);
    collect(UL2, 0, 0, blockRows, plan.uBi, plan.uBj, plan.uId);
    // padded permuted COO: the padding diagonal completes the problem
    for (int64_t i = n; i < N; i++) {
        plan.prow.push_back((int32_t)i);
        plan.pcol.push_back((int32_t)i);
        plan.pval.push_back(1.0);
    }
    plan.tPlan = std::chrono::duration<double>(clk::now() - tp).count();
    log += fmt_seconds("plan time: ", plan.tPlan);
    flush(mark);
}

bool rebind_plan(const Plan& like, const Coo& a, const double* b, Plan& out) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    const int32_t n = a.n;
    const size_t nnz = a.val.size();
    // an L L^T plan factors one triangle: never the plan of a general matrix, whatever
    // its pattern
    if (a.symmetric || like.symmetric || n != like.cfg.mSize || like.blockSorted ||
        (int64_t)nnz != like.nnzOrig || a.row.size() != nnz || a.col.size() != nnz)
        return false;
    // the same pattern: every entry lands where the earlier plan put it
    const int32_t* rev = like.reverseOrder.data();
    for (size_t k = 0; k < nnz; k++) {
        const int32_t r = a.row[k], c = a.col[k];
        if (r < 0 || r >= n || c < 0 || c >= n) return false;
        if (rev[r] != like.prow[k] || rev[c] != like.pcol[k]) return false;
    }
    // sortInBlock stays the identity for the new diagonal (GPSOrder.cpp:55-150)
    std::vector<double> diag((size_t)n, 0.0);
    for (size_t k = 0; k < nnz; k++)
        if (like.prow[k] == like.pcol[k]) diag[(size_t)like.prow[k]] = a.val[k];
    for (int32_t i = 0; i < n; i += like.cfg.blockSize)
        if (block_needs_sort(diag.data(), i, n, like.cfg.blockSize)) return false;
    out = like;
    for (size_t k = 0; k < nnz; k++) out.pval[k] = a.val[k];
    const int64_t N = (int64_t)like.cfg.blockRows * like.cfg.blockSize;
    out.bPerm.clear();
    if (b) {
        out.bPerm.assign((size_t)N, 1.0);
        for (int32_t i = 0; i < n; i++) out.bPerm[(size_t)i] = b[like.newOrder[(size_t)i]];
    }
    for (size_t q = 0; q < out.cellSrc.size(); q++) out.cellVal[q] = out.pval[out.cellSrc[q]];
    out.tReorder = 0;
    out.tPlan = std::chrono::duration<double>(clk::now() - t0).count();
    out.log = "plan reused: same sparsity pattern as an earlier plan\n" + fmt_seconds("plan time: ", out.tPlan);
    return true;
}

uint64_t pattern_hash(int32_t n, int64_t nnz, const int32_t* row, const int32_t* col) {
    // splitmix64 finaliser over (n, nnz, then each (row, col) pair), chained
    auto mix = [](uint64_t z) {
        z += 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    uint64_t h = mix((uint64_t)(uint32_t)n ^ ((uint64_t)nnz << 32));
    for (int64_t k = 0; k < nnz; k++)
        h = mix(h ^ (((uint64_t)(uint32_t)row[k] << 32) | (uint32_t)col[k]));
    return h;
}

void Plan::input_tiles_dense(double* out) const {
    const size_t n = inputIds.size();
    std::memset(out, 0, n * kTile * kTile * sizeof(double));
    for (size_t t = 0; t < n; t++) {
        double* tile = out + t * kTile * kTile;
        for (int64_t q = inputCellBegin[t]; q < inputCellBegin[t + 1]; q++) tile[cellPos[(size_t)q]] = cellVal[(size_t)q];
    }
}

void unpermute(const Plan& plan, const double* xperm, double* xorig) {
    const int32_t n = plan.cfg.mSize;
    for (int32_t i = 0; i < n; i++) xorig[i] = xperm[plan.reverseOrder[(size_t)i]];
}

}  // namespace sogl
