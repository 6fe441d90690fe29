// The batched window engine: many circuits that share one symbolic plan (an
// Omniscape window shape) solved together on one GPU. Host side.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"
#include "plan.hpp"

namespace lsc {

constexpr int kThreads = 256;      // 4 wavefronts per workgroup
constexpr int kWaves = 4;
constexpr int kCurThreads = 1024;  // node currents: 16 wavefronts share a window's voltages in LDS
constexpr int kSmallRows = 64;     // fronts of <= 64 rows (RHS row included): one wavefront each
constexpr int kSmallWaves = 1;     // wavefronts (small fronts) per workgroup: LDS is allotted per wavefront
constexpr int kBigTiles = 10;      // fronts of <= 160 rows: the whole workgroup, tiles in LDS
constexpr int kBigWaves = 4;       // wavefronts per large-front workgroup (the panel needs 16 + 48 kBigWaves >= 160)
static_assert(16 + 48 * kBigWaves >= kBigTiles * 16, "the panel's rows must fit the large-front team");
// A mixed-precision solve that ends more than this many times above its
// tolerance has failed (FP32-preconditioned refinement that stalls at the
// FP64 floor of the residual ends within it): it is solved again in FP64.
constexpr double kFallbackGap = 1000.0;
constexpr int kMaxW = 127;         // window side of the register flood fill (2 x 64-bit words per row) and of
                                   // the voltages in LDS for node currents; wider windows: k_flood_wide, HBM
constexpr int kMaxWideW = 511;     // window side of k_flood_wide (8 words a row: 3 x 511 x 64 B of LDS)

// Everything a factorisation or substitution task reads about its front, one
// record per level-order position (the plan's `order`), so that a task
// starts with one load instead of a chain of dependent ones (level start,
// front id, then its sizes and offsets, then its children's): the front, its
// pivots, assembled entries and offsets, and its first two children's update
// sizes and offsets (nk = -1: more than two children, read from the lists).
struct alignas(16) FrontDesc {
    int t, first, np, nu;
    int fe0, fe1, cb, ce;
    long long lpos, uposA;
    long long vpos, reserved0;
    int nk, ku[2], reserved1;
    int kupd[2], kid[2];
    long long kuposA[2];
    long long kvpos[2];
};
static_assert(sizeof(FrontDesc) == 128, "front records are two 64-byte scalar loads");

// Device copy of a Plan plus the per-level chunking of the solve kernels.
struct DPlan {
    int n, nd, W, D, nfront, H;
    const int *cell, *nbr;
    const int *first, *npiv, *nupd, *upd, *cb, *ce;
    const long long *lpos, *uposA, *vpos;
    const int *kids, *rows, *relmap;
    const int *lev_ptr, *lev_small, *lev_fit, *order;  // per level: fronts, small ones, ones of <= 159 rows
    const int *fe_ptr, *fe_row, *fe_col, *fe_p, *fe_q;  // entries assembled into each front
    const long long* fe_src;
    const FrontDesc* desc;  // [nfront] in level order
    // Window plans: the LDS offset (in the front's tiles) of every stored
    // entry, so that stores and extend-adds need no index arithmetic:
    // lidx[lpos[t] + e] for L's packed entry e of front t, uidx[uposA[t] + e]
    // for its update matrix's, xidx[uposA[c] + e] for where entry e of child
    // c's update matrix lands in its parent, eidx[e] for assembled entry e
    // (null for graph plans).
    const unsigned short *lidx, *uidx, *xidx, *eidx;
    int nent;
    long long lsize, usizeA, vsize;
};

// A run of a level's fronts whose LDS tiles are the same (the plan orders a
// level's fronts by size): `first` is its offset in the level, `count` its
// fronts, `tiles` the 16 x 16 tiles of each.
struct FrontGroup {
    int first, count, tiles;
};

struct DevPlan {
    Plan host;
    DPlan d{};
    // per level: the small (<= 64 rows) and large (65-159 rows) fronts in runs
    // of equal tile counts, each factorised by a launch with that much LDS per
    // front, so that smaller fronts are not given the level's largest share
    std::vector<std::vector<FrontGroup>> lev_sgroups, lev_bgroups;
    // per level: the fronts of at most 159 rows in runs for the back
    // substitution, `tiles` holding the run's largest L block (entries)
    std::vector<std::vector<FrontGroup>> lev_lgroups;
    std::vector<int> lev_fit_host;  // per level: fronts of at most 159 rows (RHS row included), the level's first ones
    std::vector<int> lev_small_host, lev_stile, lev_btile, lev_lmax, lev_mmax;  // per level: small fronts, their
                                                                                // tiles, the big fronts' tiles,
                                                                                // largest L block, largest front
    std::vector<DevBuf<int>> ibufs;
    std::vector<DevBuf<long long>> lbufs;
    DevBuf<FrontDesc> dbuf;
    DevBuf<unsigned short> tbuf[4];
    std::vector<int> huge;  // fronts of more than 159 rows (RHS row included), in level order
    // `tables`: build lidx / uidx / xidx (window plans, whose fronts share
    // every window of a batch; graph plans read the lists instead)
    void upload(Plan p, hipStream_t s, bool tables = false);
};

// Per-window arrays of a batch (window slot k of B).
struct WindowBatchBuffers {
    int B = 0, W2 = 0, n = 0;
    long long lsize = 0;
    int elem = 4;  // bytes of an L entry
    DevBuf<double> g, src, dg, sc, b, x, y, z, rs, cur, rnorm, rprev;
    DevBuf<unsigned char> present;
    DevBuf<int> active, fail, done, iters;
    DevBuf<char> L;
    void ensure(int B_, int W2_, int n_, long long lsize_, int elem_);
};

// Geometry of an Omniscape run on the target lattice (reference
// omniscape.cpp:406-421): target (ix, iy) sits at (br + ix bs, br + iy bs),
// lattice index ix * nly + iy.
struct Lattice {
    long long nx, ny;
    int R, D, W, bs, br, nly;
    long long nlat;
    int rc;  // resistance_is_conductance
};

// The conditions of a run on the device: present maps, the value of every
// target of the lattice, the comparison.
struct CondArgs {
    int n;
    const double* present[2];
    const double* value[2];
    int equal[2];
    double lower[2], upper[2];
};

struct IrStats {
    long long solved = 0, unconverged = 0, over_bar = 0, iter_sum = 0, solve_launches = 0;
    int batches = 0;
    int iter_max = 0;
    int streams = 1;  // batches in flight at once
    int fallback = 0; // circuits (graphs: factorisations) solved again in FP64 after the FP32 factor failed or stalled
    double resid_max = 0, resid_sum = 0;
    std::vector<float> resids;  // every solved circuit's relative residual (for the median)
    double resid_median() const {
        if (resids.empty()) return 0.0;
        std::vector<float> v(resids);
        const std::size_t h = v.size() / 2;
        std::nth_element(v.begin(), v.begin() + h, v.end());
        if (v.size() % 2) return v[h];
        const float hi = v[h];
        return 0.5 * (static_cast<double>(*std::max_element(v.begin(), v.begin() + h)) + hi);
    }
    double ms_fill = 0, ms_flood = 0, ms_factor = 0, ms_solve = 0, ms_residual = 0, ms_currents = 0,
           ms_accum = 0;
};

}  // namespace lsc
