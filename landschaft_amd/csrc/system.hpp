// One sparse SPD system (a grounded raster-graph Laplacian) factorised once
// on the GPU and solved for any number of right-hand sides, with FP64
// iterative refinement against the FP64 matrix. Host side.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "host.hpp"
#include "plan.hpp"
#include "window.hpp"

namespace lsc {

// A factorisation met a non-positive pivot (status 2, the reference's
// message, src/cpp/circuit/grid.cpp:208).
struct NotPositiveDefinite : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Nested-dissection leaf size of the graph plans (cells of a leaf front).
constexpr int kGraphLeaf = 32;

// The matrix in the caller's node order: diagonal, and CSR off-diagonal
// values (the negated edge conductances) over a symmetric adjacency.
struct HostMatrix {
    int n = 0;
    i64 bx = 1, by = 1;
    std::vector<i64> cell;   // node -> box cell (row-major bx x by), -1: merged node
    std::vector<i64> ptr;    // CSR
    RawVec<i32> adj;
    std::vector<double> diag;
    RawVec<double> off;
};

class SystemSolver {
public:
    SystemSolver(int precision, hipStream_t s);
    ~SystemSolver();
    SystemSolver(const SystemSolver&) = delete;
    SystemSolver& operator=(const SystemSolver&) = delete;

    // Symbolic analysis, upload and numeric factorisation of `m`.
    void factor(const HostMatrix& m);
    // Numeric factorisation of new values on the structure (n, ptr, adj) of
    // the last factor(): the symbolic plan is kept.
    void refactor(const HostMatrix& m);
    // X = A^-1 B for k right-hand sides in the caller's node order (device
    // pointers, k x n row-major), refined to ||b - Ax|| / ||b|| <= tol.
    void solve_device(int k, const double* d_b, double* d_x, double tol, int maxit, IrStats& st);
    // The same with host arrays.
    void solve(int k, const double* b, double* x, double tol, int maxit, IrStats& st);

    int n() const;
    // FP64 refactorisations after the FP32 factor failed (a non-positive
    // pivot) or a mixed solve stalled: 0 or 1.
    int fallbacks() const;
    const Plan& plan() const;
    double analyze_ms = 0;  // host symbolic analysis and upload (wall clock)
    double factor_ms = 0;   // numeric factorisation (device)
    double solve_ms = 0;    // solves with refinement (device)

private:
    struct Impl;
    Impl* impl_;
};

}  // namespace lsc
