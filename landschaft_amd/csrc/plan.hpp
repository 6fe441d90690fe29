// Host-side symbolic plans of the multifrontal Cholesky, laid out for the GPU.
//
// A plan orders the unknowns of a set of grid cells by geometric nested
// dissection (the scheme of the reference's TreeBuilder,
// src/cpp/circuit/system.cpp:25-94: boxes shrunk to their cells, cut by the
// middle row or column, leaves of <= 16 cells), builds every front's update
// rows and its child -> parent row map (system.cpp:277-339), and then adds
// what the device kernels need:
//   * fronts grouped by height (0 = leaves), so one launch or one barrier
//     phase factors every front of a level;
//   * L stored level by level, each front's block as packed columns
//     (column j holds rows j..m-1), so a level's factor is one contiguous
//     stream of HBM;
//   * per (unknown, neighbour direction): the front row of that matrix entry
//     (-1 when the entry belongs to the neighbour's front);
//   * offsets of the update matrices (nu x nu, column-major) and update
//     vectors (nu) of every front.
#pragma once

#include <cstdint>
#include <vector>

#include "host.hpp"

namespace lsc {

using i64 = std::int64_t;
using i32 = std::int32_t;

// Neighbours in increasing column-major index (the reference's kNeighbours8/4,
// src/cpp/circuit/grid.hpp:29-31): this order fixes the diagonal sums.
struct Offset {
    int dy, dx;
    bool diagonal;
};
constexpr Offset kNb8[8] = {{-1, -1, true}, {0, -1, false}, {1, -1, true}, {-1, 0, false},
                            {1, 0, false},  {-1, 1, true},  {0, 1, false}, {1, 1, true}};
constexpr Offset kNb4[4] = {{0, -1, false}, {-1, 0, false}, {1, 0, false}, {0, 1, false}};

struct Plan {
    i64 bx = 0, by = 0;   // box (row-major bx x by)
    bool four = false;
    int nd = 8;
    i32 n = 0;            // unknowns
    i32 nfront = 0;
    // unknowns
    std::vector<i32> cell;      // unknown -> box cell
    std::vector<i32> unk_of;    // box cell -> unknown (-1)
    std::vector<i32> nbr;       // [n*nd] neighbouring unknown (-1: none)
    std::vector<i32> amap;      // [n*nd] front row of the entry (p, nbr) when nbr > p, else -1
    // fronts, in post-order (children before parents)
    std::vector<i32> first, npiv, nupd, upd, cb, ce, height;
    std::vector<i64> lpos, upos, vpos;
    std::vector<i32> kids, rows, relmap;
    // Off-diagonal entries assembled into each front: entries fe_ptr[t] ..
    // fe_ptr[t+1] go to front-local (fe_row, fe_col); fe_p / fe_q are the
    // unknowns of the entry, fe_src its source: p * nd + direction (grids) or
    // the caller's adjacency entry (graphs).
    std::vector<i32> fe_ptr;
    RawVec<i32> fe_row, fe_col, fe_p, fe_q;
    RawVec<i64> fe_src;
    // graphs: unknown -> caller's node, and back
    std::vector<i32> perm, iperm;
    // levels: fronts of height h are order[lev_ptr[h] .. lev_ptr[h+1])
    std::vector<i32> lev_ptr, order;
    i64 lsize = 0, usize = 0, vsize = 0;
    i32 max_front = 0;
    double flops = 0;     // sum np^3/3 + np^2 nu + np nu^2
};

// The plan of the marked cells of a bx x by box; boxes of at most `leaf`
// cells are not cut further (the reference cuts down to 16).
Plan make_plan(i64 bx, i64 by, const std::uint8_t* mark, bool four, int leaf = 16);

// An Omniscape window: W = 2D + 1 square, the cells within D of the centre
// (every cell when disk is false), without the centre (the ground).
Plan window_plan(int D, bool disk, bool four, int leaf = 16);

// The plan of a graph of n nodes with geometry (the reference's
// frontal_analyze_graph, src/cpp/circuit/system.cpp:373-440): node c is the
// box cell cell[c] (row-major bx x by) or, when cell[c] < 0, a merged node
// (a short-circuit region) eliminated last in one front at the root.
// Adjacency in CSR: node c's neighbours adj[ptr[c] .. ptr[c+1]) (symmetric,
// -1 ignored). Cell nodes that are grid neighbours may be the only edges
// between cell nodes. A graph with no cell node at all (a network, or a
// component made only of short-circuit regions) is ordered by a nested
// dissection of the graph itself instead (breadth-first level structures).
Plan graph_plan(i64 bx, i64 by, const std::vector<i64>& cell, const std::vector<i64>& ptr, const i32* adj,
                int leaf = 16);

}  // namespace lsc
