// schedule.cpp -- host-side sweep scheduling: the reference's per-direction
// element order, and the level-set (wavefront) schedule the device sweep runs.
//
// Element order.  The reference (solver-python/pybte/ordering.py:61-140)
// repeatedly scans the elements in ascending index, emitting every element
// whose inflow faces are all cleared and clearing the reciprocal flag at
// once.  An element is emitted in the first pass q in which every upwind
// neighbour j was emitted earlier: in a previous pass, or in the same pass
// at a smaller index.  Hence
//     pass(i) = max(1, max_j (pass(j) + [j > i]))
// and the emitted order is the elements sorted by (pass, index).  That is
// a longest-path computation on the upwind DAG, done here in O(N + E) per
// direction by Kahn's algorithm instead of O(N * passes).  A direction whose
// graph has a cycle falls back to the literal scan with the reference's
// cycle-breaking rule (fewest remaining inflow faces, lowest index).
//
// Level sets.  Within one direction the sweep reads, for each element, the
// distribution of its upwind neighbours and (through periodic faces, or
// through inflow faces lagged by a broken cycle) possibly of elements that
// the reference order visits later, whose *old* value is then read.  Every
// such read becomes an edge that agrees with the reference order:
// new-value reads j -> i and old-value reads i -> j.  Levels of that DAG,
// separated by a barrier, reproduce the reference's values exactly, because
// every element sees precisely the neighbour values it saw in the
// reference order.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "pbte_hip.h"

namespace {

// upwind interior neighbours of element i in direction (cx, cy); a face is
// inflow when c.n < 0 (ordering.py:75-82)
inline bool is_inflow(double cx, double cy, const double* nx, const double* ny, int i, int il)
{
    return cx * nx[i * 3 + il] + cy * ny[i * 3 + il] < 0.0;
}

// literal scan, used only for a direction whose upwind graph has a cycle
int scan_order(int n, const int32_t* nb, const double* nx, const double* ny, double cx, double cy,
               int allow_break, int32_t* order, int32_t* broken_out)
{
    std::vector<int8_t> flag((size_t)n * 3, 0);
    std::vector<uint8_t> queued((size_t)n, 1);
    for (int i = 0; i < n; ++i)
        for (int il = 0; il < 3; ++il)
            if (nb[i * 3 + il] >= 0) {
                double fl = cx * nx[i * 3 + il] + cy * ny[i * 3 + il];
                flag[i * 3 + il] = fl > 0.0 ? 1 : (fl < 0.0 ? -1 : 0);
            }
    auto clear_from = [&](int i) {
        for (int il = 0; il < 3; ++il) {
            int t = nb[i * 3 + il];
            if (t < 0) continue;
            for (int k = 0; k < 3; ++k)
                if (nb[t * 3 + k] == i) { flag[t * 3 + k] = 0; break; }
        }
    };
    int num = 0, broken = 0;
    while (num < n) {
        bool progressed = false;
        for (int i = 0; i < n; ++i) {
            if (!queued[i]) continue;
            if (flag[i * 3] == -1 || flag[i * 3 + 1] == -1 || flag[i * 3 + 2] == -1) continue;
            order[num++] = i;
            queued[i] = 0;
            progressed = true;
            clear_from(i);
        }
        if (progressed) continue;
        if (!allow_break) return -1;
        int best = -1, best_in = 4;
        for (int i = 0; i < n; ++i) {
            if (!queued[i]) continue;
            int cnt = (flag[i * 3] == -1) + (flag[i * 3 + 1] == -1) + (flag[i * 3 + 2] == -1);
            if (cnt < best_in) { best_in = cnt; best = i; }
        }
        if (best < 0) return -1;
        order[num++] = best;
        queued[best] = 0;
        ++broken;
        clear_from(best);
    }
    *broken_out = broken;
    return 0;
}

// closed-form pass numbers; returns false when the graph has a cycle
bool pass_order(int n, const int32_t* nb, const double* nx, const double* ny, double cx, double cy,
                int32_t* order)
{
    std::vector<int32_t> indeg((size_t)n, 0), pass((size_t)n, 1), stack;
    stack.reserve(n);
    for (int i = 0; i < n; ++i)
        for (int il = 0; il < 3; ++il)
            if (nb[i * 3 + il] >= 0 && is_inflow(cx, cy, nx, ny, i, il)) ++indeg[i];
    for (int i = 0; i < n; ++i)
        if (indeg[i] == 0) stack.push_back(i);
    int done = 0, maxpass = 1;
    while (!stack.empty()) {
        int j = stack.back();
        stack.pop_back();
        ++done;
        maxpass = std::max(maxpass, pass[j]);
        // downwind neighbours of j: faces of j with c.n > 0 whose neighbour
        // sees the shared face as inflow
        for (int il = 0; il < 3; ++il) {
            int i = nb[j * 3 + il];
            if (i < 0) continue;
            for (int k = 0; k < 3; ++k) {
                if (nb[i * 3 + k] != j || !is_inflow(cx, cy, nx, ny, i, k)) continue;
                int cand = pass[j] + (j > i ? 1 : 0);
                if (cand > pass[i]) pass[i] = cand;
                if (--indeg[i] == 0) stack.push_back(i);
            }
        }
    }
    if (done != n) return false;
    // counting sort by pass, stable in the element index
    std::vector<int32_t> count((size_t)maxpass + 2, 0);
    for (int i = 0; i < n; ++i) ++count[pass[i] + 1];
    for (int p = 1; p <= maxpass + 1; ++p) count[p] += count[p - 1];
    for (int i = 0; i < n; ++i) order[count[pass[i]]++] = i;
    return true;
}

template <class F>
void parallel_for(int n, F&& f)
{
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nt = (int)std::min<unsigned>(hw, 16u);
    if (nt <= 1 || n < 2) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic_int next{0};
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (int i = next++; i < n; i = next++) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" int pbte_build_order(int32_t ndir, int32_t n_tris, const int32_t* neighbour,
                                const double* nx, const double* ny, const double* cx,
                                const double* cy, int32_t allow_break, int32_t* order,
                                int32_t* cycles)
{
    std::atomic_int status{0};
    parallel_for(ndir, [&](int d) {
        int32_t* ord = order + (size_t)d * n_tris;
        cycles[d] = 0;
        if (pass_order(n_tris, neighbour, nx, ny, cx[d], cy[d], ord)) return;
        int32_t broken = 0;
        if (scan_order(n_tris, neighbour, nx, ny, cx[d], cy[d], allow_break, ord, &broken) != 0) {
            for (int k = 0; k < n_tris; ++k) ord[k] = -1;
            status = -1;
            return;
        }
        cycles[d] = broken;
    });
    return status.load();
}

// Level schedule of one direction.  reads(i) are the elements whose
// distribution element i reads during the sweep: interior inflow neighbours
// and, unless the periodic inflow comes from an explicit buffer, the
// partner element behind an inflow periodic face.
int pbte_detail_levels(int n, const int32_t* order, const int32_t* nb, const int32_t* tri_faces,
                       const int32_t* face_bc, const int32_t* bc_type, const int32_t* periodic_tri,
                       const double* nx, const double* ny, double cx, double cy,
                       std::vector<int32_t>& level_of)
{
    // work arrays kept per thread across calls (one call per direction
    // class; fresh ones page-fault on every call)
    static thread_local std::vector<int32_t> pos, lower;
    pos.resize((size_t)n);
    for (int k = 0; k < n; ++k) pos[order[k]] = k;
    level_of.assign((size_t)n, 0);
    // edges point forward in the reference order, so one pass in that order
    // finalises each element's level before it is used.  Old-value reads
    // i -> j are pushed forward as a lower bound on j.
    lower.assign((size_t)n, 0);
    int maxlev = 0;
    for (int k = 0; k < n; ++k) {
        int i = order[k];
        int lev = lower[i];
        for (int il = 0; il < 3; ++il) {
            if (!is_inflow(cx, cy, nx, ny, i, il)) continue;
            int j = nb[i * 3 + il];
            if (j < 0) {
                int fc = tri_faces[i * 3 + il];
                int bc = face_bc[fc];
                if (bc < 0 || bc_type[bc] != 3) continue;
                j = periodic_tri[fc];
                if (j < 0) continue;
            }
            if (pos[j] < k) lev = std::max(lev, level_of[j] + 1);
        }
        level_of[i] = lev;
        maxlev = std::max(maxlev, lev);
        for (int il = 0; il < 3; ++il) {
            if (!is_inflow(cx, cy, nx, ny, i, il)) continue;
            int j = nb[i * 3 + il];
            if (j < 0) {
                int fc = tri_faces[i * 3 + il];
                int bc = face_bc[fc];
                if (bc < 0 || bc_type[bc] != 3) continue;
                j = periodic_tri[fc];
                if (j < 0) continue;
            }
            if (pos[j] > k) lower[j] = std::max(lower[j], lev + 1);
        }
    }
    return maxlev + 1;
}
