// cluster.h -- host code: boxes of the next AMR level from refinement tags
// (the step AmrCore::MakeNewGrids performs with amrex::ClusterList,
// src/Integrator/Integrator.cpp:23-140 sets its parameters: amr.blocking_factor,
// amr.grid_eff, amr.n_error_buf).
//
// 1. Tags are grown by `buffer` cells (n_error_buf) inside the array.
// 2. They are coarsened by q = blocking_factor / 2 parent cells on the global
//    index grid, so every box that comes back is aligned to q parent cells:
//    its extent on the refined level is a multiple of the blocking factor.
// 3. Berger-Rigoutsos clustering (M. Berger, I. Rigoutsos, IEEE Trans. SMC
//    21(5), 1991) of the coarse tags: a box is shrunk to its tags' bounding
//    box and kept when its tagged fraction reaches grid_eff; otherwise it is
//    cut at a hole of the tag signature (the plane counts along an axis), or,
//    without holes, at the strongest inflection of the signature (sign change
//    of its second difference), or, without either, in the middle of its
//    longest axis; both halves are clustered again.  Ties: the longest axis
//    first (then the lowest), the cut nearest the middle (then the lowest).
// 4. Boxes are refined back to parent cells and clipped to the array.
// 5. Boxes closer than one cell (touching or overlapping once one is grown by
//    one cell) are merged into their bounding box until none are: the
//    composite solve couples a patch to its parent only, so two patches of a
//    level must not share nodes or reach each other's ghost layer.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace ael {
namespace cluster {

struct Box {
    int lo[3], hi[3];  // inclusive cell indices
    long volume() const { return (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1); }
};

struct Grid {
    int n[3];
    std::vector<uint8_t> t;  // t[(k n1 + j) n0 + i]
    uint8_t at(int i, int j, int k) const { return t[((size_t)k * n[1] + j) * n[0] + i]; }
};

// Bounding box of the tags inside b; false when there are none.
inline bool shrink(const Grid& g, const Box& b, Box& out, long& count) {
    int lo[3] = {b.hi[0] + 1, b.hi[1] + 1, b.hi[2] + 1}, hi[3] = {b.lo[0] - 1, b.lo[1] - 1, b.lo[2] - 1};
    count = 0;
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (g.at(i, j, k)) {
                    ++count;
                    const int p[3] = {i, j, k};
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], p[a]);
                        hi[a] = std::max(hi[a], p[a]);
                    }
                }
    if (count == 0) return false;
    for (int a = 0; a < 3; ++a) {
        out.lo[a] = lo[a];
        out.hi[a] = hi[a];
    }
    return true;
}

// Tag counts of the planes of b normal to axis a.
inline std::vector<long> signature(const Grid& g, const Box& b, int a) {
    std::vector<long> s((size_t)(b.hi[a] - b.lo[a] + 1), 0);
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                if (g.at(i, j, k)) {
                    const int p[3] = {i, j, k};
                    ++s[(size_t)(p[a] - b.lo[a])];
                }
    return s;
}

// Axis order of the cut search: longest extent first, then the lowest axis.
inline void axis_order(const Box& b, int ord[3]) {
    ord[0] = 0;
    ord[1] = 1;
    ord[2] = 2;
    std::stable_sort(ord, ord + 3, [&](int x, int y) { return b.hi[x] - b.lo[x] > b.hi[y] - b.lo[y]; });
}

// Cut of b: axis and the first cell of the upper half; false when b is a
// single cell along every axis.
inline bool choose_cut(const Grid& g, const Box& b, int& axis, int& cut) {
    int ord[3];
    axis_order(b, ord);
    // twice the distance of a cut before cell c (c = lo + m) from the middle
    auto dist2 = [&](int a, int m) {
        const int len = b.hi[a] - b.lo[a] + 1;
        return std::abs(2 * m - len);
    };
    // 1. holes: an empty plane m (never the first or last: b is a bounding box)
    for (int q = 0; q < 3; ++q) {
        const int a = ord[q];
        const std::vector<long> s = signature(g, b, a);
        int best = -1;
        for (int m = 1; m + 1 < (int)s.size(); ++m)
            if (s[(size_t)m] == 0 && (best < 0 || dist2(a, m) < dist2(a, best))) best = m;
        if (best >= 0) {
            axis = a;
            cut = b.lo[a] + best;
            return true;
        }
    }
    // 2. strongest inflection: lap[m] = s[m-1] - 2 s[m] + s[m+1] changes sign
    // between m and m+1; the cut goes before cell m+1
    long best_strength = -1;
    int best_axis = -1, best_m = -1;
    for (int q = 0; q < 3; ++q) {
        const int a = ord[q];
        const std::vector<long> s = signature(g, b, a);
        const int len = (int)s.size();
        for (int m = 1; m + 2 < len; ++m) {
            const long l0 = s[(size_t)m - 1] - 2 * s[(size_t)m] + s[(size_t)m + 1];
            const long l1 = s[(size_t)m] - 2 * s[(size_t)m + 1] + s[(size_t)m + 2];
            if (!((l0 < 0 && l1 > 0) || (l0 > 0 && l1 < 0))) continue;
            const long st = std::labs(l1 - l0);
            if (st > best_strength || (st == best_strength && dist2(a, m + 1) < dist2(best_axis, best_m))) {
                best_strength = st;
                best_axis = a;
                best_m = m + 1;
            }
        }
    }
    if (best_axis >= 0) {
        axis = best_axis;
        cut = b.lo[best_axis] + best_m;
        return true;
    }
    // 3. the middle of the longest axis
    const int a = ord[0];
    const int len = b.hi[a] - b.lo[a] + 1;
    if (len < 2) return false;
    axis = a;
    cut = b.lo[a] + len / 2;
    return true;
}

inline void berger_rigoutsos(const Grid& g, const Box& in, double eff, std::vector<Box>& out) {
    Box b;
    long count = 0;
    if (!shrink(g, in, b, count)) return;
    if ((double)count >= eff * (double)b.volume()) {
        out.push_back(b);
        return;
    }
    int axis = 0, cut = 0;
    if (!choose_cut(g, b, axis, cut)) {
        out.push_back(b);
        return;
    }
    Box lo = b, hi = b;
    lo.hi[axis] = cut - 1;
    hi.lo[axis] = cut;
    berger_rigoutsos(g, lo, eff, out);
    berger_rigoutsos(g, hi, eff, out);
}

inline bool near(const Box& x, const Box& y) {
    for (int a = 0; a < 3; ++a)
        if (x.hi[a] + 1 < y.lo[a] || y.hi[a] + 1 < x.lo[a]) return false;
    return true;
}

// The whole procedure on the tags of an n[0] x n[1] x n[2] cell array whose
// cell (0, 0, 0) is global cell origin; boxes in global cells.
inline std::vector<Box> make_boxes(const uint8_t* tags, const int n[3], const int origin[3], int buffer,
                                   int blocking_factor, double grid_eff) {
    // 1. buffer
    Grid t{{n[0], n[1], n[2]}, std::vector<uint8_t>((size_t)n[0] * n[1] * n[2], 0)};
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                if (!tags[((size_t)k * n[1] + j) * n[0] + i]) continue;
                for (int c = std::max(0, k - buffer); c <= std::min(n[2] - 1, k + buffer); ++c)
                    for (int b = std::max(0, j - buffer); b <= std::min(n[1] - 1, j + buffer); ++b)
                        for (int a = std::max(0, i - buffer); a <= std::min(n[0] - 1, i + buffer); ++a)
                            t.t[((size_t)c * n[1] + b) * n[0] + a] = 1;
            }
    // 2. coarsen by q on the global grid
    const int q = std::max(1, blocking_factor / 2);
    auto fdiv = [](int v, int d) { return v >= 0 ? v / d : -((-v + d - 1) / d); };
    int clo[3], cn[3];
    for (int a = 0; a < 3; ++a) {
        clo[a] = fdiv(origin[a], q);
        cn[a] = fdiv(origin[a] + n[a] - 1, q) - clo[a] + 1;
    }
    Grid g{{cn[0], cn[1], cn[2]}, std::vector<uint8_t>((size_t)cn[0] * cn[1] * cn[2], 0)};
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                if (t.at(i, j, k)) {
                    const int I = fdiv(origin[0] + i, q) - clo[0], J = fdiv(origin[1] + j, q) - clo[1],
                              K = fdiv(origin[2] + k, q) - clo[2];
                    g.t[((size_t)K * cn[1] + J) * cn[0] + I] = 1;
                }
    // 3. Berger-Rigoutsos
    std::vector<Box> coarse;
    berger_rigoutsos(g, Box{{0, 0, 0}, {cn[0] - 1, cn[1] - 1, cn[2] - 1}}, grid_eff, coarse);
    // 4. back to global cells, clipped to the array
    std::vector<Box> boxes;
    for (const Box& c : coarse) {
        Box b;
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = std::max((c.lo[a] + clo[a]) * q, origin[a]);
            b.hi[a] = std::min((c.hi[a] + clo[a] + 1) * q - 1, origin[a] + n[a] - 1);
        }
        boxes.push_back(b);
    }
    // 5. merge boxes closer than one cell
    bool merged = true;
    while (merged) {
        merged = false;
        for (size_t x = 0; x < boxes.size() && !merged; ++x)
            for (size_t y = x + 1; y < boxes.size() && !merged; ++y)
                if (near(boxes[x], boxes[y])) {
                    for (int a = 0; a < 3; ++a) {
                        boxes[x].lo[a] = std::min(boxes[x].lo[a], boxes[y].lo[a]);
                        boxes[x].hi[a] = std::max(boxes[x].hi[a], boxes[y].hi[a]);
                    }
                    boxes.erase(boxes.begin() + (long)y);
                    merged = true;
                }
    }
    std::sort(boxes.begin(), boxes.end(), [](const Box& x, const Box& y) {
        for (int a = 2; a >= 0; --a)
            if (x.lo[a] != y.lo[a]) return x.lo[a] < y.lo[a];
        return false;
    });
    return boxes;
}

}  // namespace cluster
}  // namespace ael
