// Gibbs-Poole-Stockmeyer bandwidth-reducing ordering (GPSOrder.cpp) and the
// per-64-block diagonal swap network (GOrder::sortInBlock).
//
// The permutation has to be identical to the reference's, including how ties
// are broken.  The reference numbers nodes by std::sort with a value-only
// comparator (operation.cpp:31-39), which is not stable: tie order is a
// property of libstdc++'s introsort.  This file therefore builds the same
// candidate sequences in the same order and sorts them with std::sort and a
// value-only comparator too (compiled with g++/libstdc++, never libc++ or
// stable_sort).  introsort's result is a function of the comparison outcomes
// only, so the payload layout of the sorted records does not matter.
//
// Differences in form, not in result: adjacency is a CSR array rather than
// vector<oneRow*>, BFS is iterative rather than recursive (nextLevel recurses
// once per level), and the candidate search keeps only depth/width/last level
// of each trial structure instead of every full structure.
#include <algorithm>
#include <cstdio>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

#include "planner.h"

namespace sogl {

namespace {

struct Position {  // class position (operation.h:22-31): value is the sort key
    int value;
    int index;
};
inline bool position_less(const Position& x, const Position& y) { return x.value < y.value; }

struct Graph {
    int n = 0;
    std::vector<int64_t> off;  // CSR offsets
    std::vector<int> adj;      // sorted, de-duplicated, no self loops
    int deg(int i) const { return (int)(off[i + 1] - off[i]); }
};

// GOrder::setupEdge (GPSOrder.cpp:279-338): symmetrised off-diagonal pattern.
// Returns true when the pattern was already symmetric ("structure symetric").
bool setup_edges(const int32_t* ri, const int32_t* ci, int64_t nnz, int n, Graph& g) {
    g.n = n;
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    int64_t shadow = 0;
    for (int64_t k = 0; k < nnz; k++)
        if (ri[k] != ci[k]) { cnt[ri[k] + 1]++; cnt[ci[k] + 1]++; shadow++; }
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    std::vector<int> raw((size_t)cnt[n]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t k = 0; k < nnz; k++)
        if (ri[k] != ci[k]) raw[(size_t)pos[ri[k]]++] = ci[k];
    for (int64_t k = 0; k < nnz; k++)
        if (ri[k] != ci[k]) raw[(size_t)pos[ci[k]]++] = ri[k];
    g.off.assign((size_t)n + 1, 0);
    g.adj.clear();
    g.adj.reserve(raw.size());
    for (int i = 0; i < n; i++) {
        auto b = raw.begin() + cnt[i], e = raw.begin() + cnt[i + 1];
        std::sort(b, e);
        for (auto it = b; it != e; ++it)
            if (it == b || *it != g.adj.back() || (int64_t)g.adj.size() == g.off[i])
                g.adj.push_back(*it);
        g.off[i + 1] = (int64_t)g.adj.size();
    }
    return shadow == (int64_t)g.adj.size();
}

// Rooted level structure: GOrder::getLevelList + nextLevel + addMissing
// (GPSOrder.cpp:185-277).  Levels are contiguous runs of `order`.
struct Levels {
    std::vector<int> order;
    std::vector<int64_t> start;  // level k = order[start[k] .. start[k+1])
    int count() const { return (int)start.size() - 1; }
    int64_t size(int k) const { return start[k + 1] - start[k]; }
    int width() const {
        int64_t w = 0;
        for (int k = 0; k < count(); k++) w = std::max(w, size(k));
        return (int)w;
    }
};

void level_list(const Graph& g, int root, Levels& L) {
    const int n = g.n;
    std::vector<char> colored((size_t)n, 0);
    L.order.clear();
    L.order.reserve((size_t)n);
    L.start.assign(1, 0);
    auto expand = [&]() {
        for (;;) {
            int64_t b = L.start.back(), e = (int64_t)L.order.size();
            for (int64_t p = b; p < e; p++) {
                int row = L.order[(size_t)p];
                for (int64_t q = g.off[row]; q < g.off[row + 1]; q++) {
                    int col = g.adj[(size_t)q];
                    if (colored[col]) continue;
                    colored[col] = 1;
                    L.order.push_back(col);
                }
            }
            if ((int64_t)L.order.size() > e) L.start.push_back(e);
            else break;
        }
    };
    L.order.push_back(root);
    colored[root] = 1;
    expand();
    // addMissing: each still-uncoloured node, lowest id first, roots a new run of levels
    if ((int)L.order.size() < n) {
        for (int i = 0; i < n; i++) {
            if (colored[i]) continue;
            L.start.push_back((int64_t)L.order.size());
            L.order.push_back(i);
            colored[i] = 1;
            expand();
        }
    }
    L.start.push_back((int64_t)L.order.size());
}

struct Trial {  // what the candidate search needs of one level structure
    int levels = 0, width = 0;
    std::vector<int> last;
};

void run_trials(const Graph& g, const std::vector<int>& roots, size_t from,
                std::vector<Trial>& out, int threads) {
    out.resize(roots.size());
    const long cnt = (long)roots.size() - (long)from;
    if (cnt <= 0) return;
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
    for (long t = 0; t < cnt; t++) {
        Levels L;
        level_list(g, roots[from + (size_t)t], L);
        Trial& tr = out[from + (size_t)t];
        tr.levels = L.count();
        tr.width = L.width();
        int k = L.count() - 1;
        tr.last.assign(L.order.begin() + L.start[k], L.order.begin() + L.start[k + 1]);
    }
}

}  // namespace

// GOrder::Reorder + updatewithlevel + reorderOneMoreLevel (GPSOrder.cpp:340-635).
// Permutes prow/pcol in place and returns newOrder/reverseOrder.
void gps_reorder(int n, std::vector<int32_t>& prow, std::vector<int32_t>& pcol,
                 std::vector<int32_t>& newOrder, std::vector<int32_t>& reverseOrder,
                 int& outLevels, int& outWidth, int& outStart, bool& symmetricPattern,
                 int threads, std::string& log) {
    const int LIST100 = 100, LIST500 = 500;
    Graph g;
    symmetricPattern = setup_edges(prow.data(), pcol.data(), (int64_t)prow.size(), n, g);
    if (symmetricPattern) log += "structure symetric\n";

    // getLeastConnected (GPSOrder.cpp:471-484): first node of minimum nonzero degree
    int startnode = 0;
    {
        int count = n;
        for (int i = 0; i < n; i++) {
            if (g.deg(i) < count) {
                if (g.deg(i) == 0) continue;
                startnode = i;
                count = g.deg(i);
            }
        }
    }
    Trial cur;
    {
        std::vector<int> r(1, startnode);
        std::vector<Trial> t;
        run_trials(g, r, 0, t, threads);
        cur = t[0];
    }
    std::set<int> Gset(cur.last.begin(), cur.last.end());
    for (int v = 0; v < n; v++)
        if (g.deg(v) == g.deg(startnode)) Gset.insert(v);

    // pseudo-peripheral search (GPSOrder.cpp:514-570)
    std::vector<int> ppnode;
    std::vector<Trial> pplevel;
    for (;;) {
        int lastsize = (int)cur.last.size();
        int lastskip = 0;
        if (lastsize > LIST100) {
            lastskip = (lastsize + LIST100 - 1) / LIST100;
            lastskip = (lastskip / 2 + 1) * 2 - 1;
        }
        ppnode.clear();
        int maxsize = cur.levels;
        for (int t = 0; t < lastsize; t++) {
            if (lastskip > 0 && t % lastskip != 0) continue;
            ppnode.push_back(cur.last[(size_t)t]);
        }
        run_trials(g, ppnode, 0, pplevel, threads);
        bool updated = false;
        for (size_t t = 0; t < ppnode.size(); t++) {
            const Trial& lvl = pplevel[t];
            if (lvl.levels < maxsize) continue;
            Gset.insert(lvl.last.begin(), lvl.last.end());
            if (lvl.levels > cur.levels) {
                cur = lvl;
                startnode = ppnode[t];
                updated = true;
                break;
            }
        }
        if (updated) continue;
        break;
    }
    if (ppnode.empty()) log += " can not find ppnode \n";

    // candidate set from Gset (GPSOrder.cpp:572-612)
    const int ppnodecount = (int)ppnode.size();
    const int vdegree = g.deg(startnode);
    Gset.insert(n - 1);
    Gset.insert(0);
    int nodeindex = ppnodecount;
    int nodeskip = 0;
    int cap = LIST500 - ppnodecount;
    if (cap > 0) {
        nodeskip = ((int)Gset.size() + cap - 1) / cap;
        nodeskip = (nodeskip / 2 + 1) * 2 - 1;
    }
    for (int ss : Gset) {
        if (g.deg(ss) > vdegree && ss != n - 1 && ss != 0) continue;
        bool found = false;
        for (size_t q = 0; q < ppnode.size(); q++)
            if (ppnode[q] == ss) { found = true; break; }
        nodeindex++;
        if (nodeskip > 1 && nodeindex % nodeskip != 0) continue;
        if (found) continue;
        if ((int)ppnode.size() >= LIST500) break;
        ppnode.push_back(ss);
    }
    run_trials(g, ppnode, (size_t)ppnodecount, pplevel, threads);

    // best = max(levels - width), first wins ties (GPSOrder.cpp:614-624)
    int minmax = -n, best = -1;
    for (size_t i = 0; i < ppnode.size(); i++) {
        int eff = pplevel[i].levels - pplevel[i].width;
        if (minmax < eff) { minmax = eff; best = (int)i; }
    }
    if (best < 0) throw std::runtime_error("GPS: no candidate level structure");
    Levels L;
    level_list(g, ppnode[(size_t)best], L);
    outLevels = L.count();
    outWidth = L.width();
    outStart = ppnode[(size_t)best];
    {
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "GGPS reorder: levels: %d bandwidth: %d last level count: %lld total accounted: %lld start from %d\n",
                      L.count(), L.width(), (long long)L.size(L.count() - 1),
                      (long long)L.order.size(), L.order[0]);
        log += buf;
    }

    // updatewithlevel / reorderOneMoreLevel (GPSOrder.cpp:340-459).  The reference's
    // per-node tbd sets only short-cut the search for still-unnumbered neighbours;
    // the numbering below takes the unnumbered neighbours of each previous-level node
    // in ascending id order, which is exactly the tbd iteration order.
    std::vector<int> nodesum((size_t)n);
    for (int v = 0; v < n; v++) {
        int s = 0;
        for (int64_t q = g.off[v]; q < g.off[v + 1]; q++) s += g.deg(g.adj[(size_t)q]);
        nodesum[(size_t)v] = s;
    }
    std::vector<int32_t> newOrd((size_t)n, -1), revOrd((size_t)n, -1);
    std::vector<Position> rank;
    rank.reserve(256);
    for (int64_t p = L.start[0]; p < L.start[1]; p++) {
        int v = L.order[(size_t)p];
        rank.push_back(Position{nodesum[(size_t)v], v});
    }
    std::sort(rank.begin(), rank.end(), position_less);
    int cursor = 0;
    for (const Position& r : rank) { newOrd[(size_t)r.index] = cursor; revOrd[(size_t)cursor] = r.index; cursor++; }
    std::vector<int> prev;
    for (int k = 1; k < L.count(); k++) {
        const int start = cursor;
        const int64_t shouldAdd = L.size(k);
        // previous level, by ascending new number (positionReverseCompare + reverse walk)
        prev.assign(L.order.begin() + L.start[k - 1], L.order.begin() + L.start[k]);
        std::sort(prev.begin(), prev.end(),
                  [&](int a, int b) { return newOrd[(size_t)a] < newOrd[(size_t)b]; });
        for (int w : prev) {
            rank.clear();
            for (int64_t q = g.off[w]; q < g.off[w + 1]; q++) {
                int v = g.adj[(size_t)q];
                if (newOrd[(size_t)v] < 0) rank.push_back(Position{nodesum[(size_t)v], v});
            }
            if (rank.empty()) continue;
            std::sort(rank.begin(), rank.end(), position_less);
            for (const Position& r : rank) {
                newOrd[(size_t)r.index] = cursor;
                revOrd[(size_t)cursor] = r.index;
                cursor++;
            }
        }
        if (cursor - start != shouldAdd) {
            rank.clear();
            for (int64_t p = L.start[k]; p < L.start[k + 1]; p++) {
                int v = L.order[(size_t)p];
                if (newOrd[(size_t)v] < 0) rank.push_back(Position{nodesum[(size_t)v], v});
            }
            std::sort(rank.begin(), rank.end(), position_less);
            for (const Position& r : rank) {
                newOrd[(size_t)r.index] = cursor;
                revOrd[(size_t)cursor] = r.index;
                cursor++;
            }
        }
    }
    if (cursor != n) throw std::runtime_error("GPS: numbering incomplete");
    for (size_t k = 0; k < prow.size(); k++) {
        prow[k] = newOrd[(size_t)prow[k]];
        pcol[k] = newOrd[(size_t)pcol[k]];
    }
    // GPSOrder.cpp:448-449: the names are swapped on purpose
    newOrder = revOrd;       // new -> old
    reverseOrder = newOrd;   // old -> new
}

// GOrder::sortInBlock (GPSOrder.cpp:55-150).  For every full 64-block whose first 32
// diagonal entries include one with |d| <= 1e-8, a compare-swap network moves larger
// signed diagonals forward; the block-local permutation composes into the ordering.
// The block test of GOrder::sortInBlock (GPSOrder.cpp:79-90): a block strictly inside
// [0, n) is reordered when one of its first blockSize/2 diagonal entries has |d| <= 1e-8.
bool block_needs_sort(const double* diag, int i, int n, int blockSize) {
    const double TINY = 1e-8;
    if (i + blockSize >= n) return false;
    for (int j = 0; j < blockSize / 2; j++)
        if (!(diag[i + j] > TINY || diag[i + j] < -TINY)) return true;
    return false;
}

bool sort_in_block(int n, int blockSize, std::vector<int32_t>& prow, std::vector<int32_t>& pcol,
                   const std::vector<double>& pval, std::vector<double>* bperm,
                   std::vector<int32_t>& newOrder, std::vector<int32_t>& reverseOrder) {
    std::vector<double> tmpv((size_t)n, 0.0);
    std::vector<int> tmpi((size_t)n, -1);
    for (size_t k = 0; k < prow.size(); k++)
        if (prow[k] == pcol[k]) { tmpv[(size_t)prow[k]] = pval[k]; tmpi[(size_t)prow[k]] = prow[k]; }
    for (int i = 0; i < n; i++)
        if (tmpi[(size_t)i] == -1) tmpi[(size_t)i] = i;
    bool any = false;
    for (int i = 0; i < n; i += blockSize) {
        if (!block_needs_sort(tmpv.data(), i, n, blockSize)) continue;
        int step = blockSize / 2;
        any = true;
        while (step > 0) {
            for (int j = 0; j < step; j++) {
                int k = i + j, m = i + j + step;
                if (tmpv[(size_t)m] > tmpv[(size_t)k]) {
                    std::swap(tmpv[(size_t)m], tmpv[(size_t)k]);
                    std::swap(tmpi[(size_t)m], tmpi[(size_t)k]);
                }
            }
            step /= 2;
        }
    }
    if (!any) return false;  // identity: nothing below changes
    std::vector<int> tmpn((size_t)n, -1);
    for (int i = 0; i < n; i++) tmpn[(size_t)tmpi[(size_t)i]] = i;
    for (size_t k = 0; k < prow.size(); k++) {
        prow[k] = tmpn[(size_t)prow[k]];
        pcol[k] = tmpn[(size_t)pcol[k]];
    }
    if (bperm && !bperm->empty()) {
        std::vector<double> bn((size_t)n);
        for (int i = 0; i < n; i++) bn[(size_t)i] = (*bperm)[(size_t)tmpi[(size_t)i]];
        for (int i = 0; i < n; i++) (*bperm)[(size_t)i] = bn[(size_t)i];
    }
    std::vector<int32_t> norder((size_t)n);
    for (int i = 0; i < n; i++) norder[(size_t)i] = newOrder[(size_t)tmpi[(size_t)i]];
    for (int i = 0; i < n; i++) {
        newOrder[(size_t)i] = norder[(size_t)i];
        reverseOrder[(size_t)norder[(size_t)i]] = i;
    }
    return true;
}

}  // namespace sogl
