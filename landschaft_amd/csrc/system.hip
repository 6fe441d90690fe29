// The graph engine: one grounded Laplacian of any size factorised by the
// batched multifrontal kernels (fronts of <= 159 rows in LDS) plus a
// global-memory path for larger fronts, then solved for many right-hand
// sides with FP64 iterative refinement against the FP64 CSR matrix.
//
// Large fronts (the top separators of big grids, the reference's
// dense_front / factorize_large territory, src/cpp/circuit/system.cpp:517-846)
// keep their 16 x 16 tiles in HBM. A panel of 16 columns is solved by as many
// workgroups as it has row blocks of 192 (each factors the 16 x 16 diagonal
// block redundantly, so they need no exchange), then the trailing tiles are
// updated with MFMA, one tile per wavefront.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "huge.hpp"
#include "system.hpp"
#include "window_kernels.hpp"

namespace lsc {

namespace {

void sys_reset(int* p, int n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

// Raises a kernel's dynamic LDS limit once per size (the calls repeat in
// every sweep; none is made again while a sweep is captured into a graph).
void sys_allow(const void* fn, std::size_t bytes) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, std::size_t> set;
    int dev = 0;
    LSC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    std::size_t& have = set[{dev, fn}];
    if (have >= bytes) return;
    LSC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
    have = bytes;
}

unsigned nblocks(long long tasks, int per) { return static_cast<unsigned>((tasks + per - 1) / per); }

// s[j] = diag^-1/2 of unknown j (elimination order).
__global__ void k_g_scale(int n, const int* __restrict__ perm, const double* __restrict__ diag,
                          double* __restrict__ sc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) sc[j] = 1.0 / sqrt(diag[perm[j]]);
}

// The scaled entries of S A S in front order.
template <typename T>
__global__ void k_g_assemble(int nent, const int* __restrict__ fe_p, const int* __restrict__ fe_q,
                             const long long* __restrict__ fe_src, const double* __restrict__ off,
                             const double* __restrict__ sc, T* __restrict__ aval) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nent) aval[e] = static_cast<T>(off[fe_src[e]] * sc[fe_p[e]] * sc[fe_q[e]]);
}

// b (caller order, k x n) -> bp (elimination order) and the scaled RHS s b.
__global__ void k_g_perm_in(int n, int k, const int* __restrict__ perm, const double* __restrict__ b,
                            const double* __restrict__ sc, double* __restrict__ bp, double* __restrict__ rs,
                            double* __restrict__ x) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long long>(n) * k) return;
    const int inst = static_cast<int>(e / n), j = static_cast<int>(e - static_cast<long long>(inst) * n);
    const double v = b[static_cast<long long>(inst) * n + perm[j]];
    bp[e] = v;
    rs[e] = sc[j] * v;
    x[e] = 0.0;
}

__global__ void k_g_perm_out(int n, int k, const int* __restrict__ perm, const double* __restrict__ x,
                             double* __restrict__ out) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long long>(n) * k) return;
    const int inst = static_cast<int>(e / n), j = static_cast<int>(e - static_cast<long long>(inst) * n);
    out[static_cast<long long>(inst) * n + perm[j]] = x[e];
}

// ---- preconditioned conjugate gradients in FP64, the factor as the
// preconditioner M^-1 = S (L L^T)^-1 S. Every vector is in elimination
// order, right-hand side blockIdx.y; dot products are summed per block and
// then over blocks in a fixed order, so results do not depend on scheduling.

// Sum of a and b over the 256 threads of a block into out[0], out[1].
__device__ __forceinline__ void block_sum2(double a, double b, double* __restrict__ out) {
    __shared__ double red[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int lane = threadIdx.x & 63, wave = wave_id();
    if (lane == 0) red[0][wave] = a, red[1][wave] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        out[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// The products run on a copy of the FP64 matrix in elimination order (row j
// of the copy is the caller's row perm[j], its entries in the same order,
// their columns in elimination order), kSpmvRhs right-hand sides per thread:
// a row's entries are read once for all of them.
constexpr int kSpmvRhs = 8;

// The elimination-order copy: row j takes row perm[j] of the caller's CSR.
__global__ __launch_bounds__(256) void k_g_permute(int n, const int* __restrict__ perm, const int* __restrict__ iperm,
                                                   const long long* __restrict__ ptr, const int* __restrict__ adj,
                                                   const double* __restrict__ diag, const double* __restrict__ off,
                                                   const long long* __restrict__ eptr, int* __restrict__ eadj,
                                                   double* __restrict__ ediag, double* __restrict__ eoff) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int c = perm[j];
    ediag[j] = diag[c];
    long long w = eptr[j];
    for (long long e = ptr[c]; e < ptr[c + 1]; ++e, ++w) {
        eadj[w] = iperm[adj[e]];
        eoff[w] = off[e];
    }
}

// Per right-hand side q < kSpmvRhs of the block (inst = r0 + q): the block's
// sums of a[q] and b[q] into part[(inst * gridDim.x + blockIdx.x) * 2 + {0, 1}],
// in the order of block_sum2.
__device__ __forceinline__ void block_sum2_rhs(const double (&a)[kSpmvRhs], const double (&b)[kSpmvRhs], int r0,
                                               int k, const int* __restrict__ done, double* __restrict__ part) {
    __shared__ double red[kSpmvRhs][2][4];
    const int lane = threadIdx.x & 63, wave = wave_id();
#pragma unroll
    for (int q = 0; q < kSpmvRhs; ++q) {
        double x = a[q], y = b[q];
        for (int o = 32; o > 0; o >>= 1) {
            x += __shfl_xor(x, o);
            y += __shfl_xor(y, o);
        }
        if (lane == 0) red[q][0][wave] = x, red[q][1][wave] = y;
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < kSpmvRhs && r0 + q < k && !done[r0 + q]) {
        double* out = part + (static_cast<long long>(r0 + q) * gridDim.x + blockIdx.x) * 2;
        out[0] = (red[q][0][0] + red[q][0][1]) + (red[q][0][2] + red[q][0][3]);
        out[1] = (red[q][1][0] + red[q][1][1]) + (red[q][1][2] + red[q][1][3]);
    }
}

// r = b - A x; rs = s r (the preconditioner's input); partial r.r and b.b.
// Right-hand sides kSpmvRhs y .. kSpmvRhs y + kSpmvRhs - 1.
__global__ __launch_bounds__(256) void k_g_residual(int n, int k, const long long* __restrict__ eptr,
                                                    const int* __restrict__ eadj, const double* __restrict__ ediag,
                                                    const double* __restrict__ eoff, const double* __restrict__ bp,
                                                    const double* __restrict__ x, const double* __restrict__ sc,
                                                    double* __restrict__ rs, const int* __restrict__ done,
                                                    double* __restrict__ part) {
    const int r0 = blockIdx.y * kSpmvRhs;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double rr[kSpmvRhs], bb[kSpmvRhs];
#pragma unroll
    for (int q = 0; q < kSpmvRhs; ++q) rr[q] = 0.0, bb[q] = 0.0;
    if (j < n) {
        const long long e0 = eptr[j], e1 = eptr[j + 1];
        const double d = ediag[j], sj = sc[j];
#pragma unroll
        for (int q = 0; q < kSpmvRhs; ++q) {
            const int inst = r0 + q;
            if (inst >= k || done[inst]) continue;
            const long long o = static_cast<long long>(inst) * n;
            double a = d * x[o + j];
            for (long long e = e0; e < e1; ++e) a += eoff[e] * x[o + eadj[e]];
            const double b = bp[o + j];
            const double r = b - a;
            rs[o + j] = sj * r;
            rr[q] = r * r;
            bb[q] = b * b;
        }
    }
    block_sum2_rhs(rr, bb, r0, k, done, part);
}

// q = A p; partial p.q. Right-hand sides kSpmvRhs y .. kSpmvRhs y + kSpmvRhs - 1.
__global__ __launch_bounds__(256) void k_g_spmv(int n, int k, const long long* __restrict__ eptr,
                                                const int* __restrict__ eadj, const double* __restrict__ ediag,
                                                const double* __restrict__ eoff, const double* __restrict__ p,
                                                double* __restrict__ q, const int* __restrict__ done,
                                                double* __restrict__ part) {
    const int r0 = blockIdx.y * kSpmvRhs;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double pq[kSpmvRhs], zero[kSpmvRhs];
#pragma unroll
    for (int t = 0; t < kSpmvRhs; ++t) pq[t] = 0.0, zero[t] = 0.0;
    if (j < n) {
        const long long e0 = eptr[j], e1 = eptr[j + 1];
        const double d = ediag[j];
#pragma unroll
        for (int t = 0; t < kSpmvRhs; ++t) {
            const int inst = r0 + t;
            if (inst >= k || done[inst]) continue;
            const long long o = static_cast<long long>(inst) * n;
            double a = d * p[o + j];
            for (long long e = e0; e < e1; ++e) a += eoff[e] * p[o + eadj[e]];
            q[o + j] = a;
            pq[t] = p[o + j] * a;
        }
    }
    block_sum2_rhs(pq, zero, r0, k, done, part);
}

// partial r.z = (s r).(L L^T)^-1 (s r), from rs and the sweep's output zz.
__global__ __launch_bounds__(256) void k_g_dot(int n, const double* __restrict__ rs, const double* __restrict__ zz,
                                               const int* __restrict__ done, double* __restrict__ part) {
    const int inst = blockIdx.y;
    if (done[inst]) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const long long o = static_cast<long long>(inst) * n;
    const double d = j < n ? rs[o + j] * zz[o + j] : 0.0;
    block_sum2(d, 0.0, part + (static_cast<long long>(inst) * gridDim.x + blockIdx.x) * 2);
}

// out[inst * 2 + t] = sum over the nb blocks of part, t = 0, 1, in a fixed order.
__global__ __launch_bounds__(256) void k_g_sum(int nb, const double* __restrict__ part, const int* __restrict__ done,
                                               double* __restrict__ out) {
    const int inst = blockIdx.x;
    if (done[inst]) return;
    const double* pp = part + static_cast<long long>(inst) * nb * 2;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) a += pp[2 * i], b += pp[2 * i + 1];
    block_sum2(a, b, out + 2 * inst);
}

// p = s zz + beta p, beta = rz_new / rz_old (0 on the first step).
__global__ __launch_bounds__(256) void k_g_dir(int n, int first, const double* __restrict__ sc,
                                               const double* __restrict__ zz, const double* __restrict__ rz_new,
                                               const double* __restrict__ rz_old, double* __restrict__ p,
                                               const int* __restrict__ done) {
    const int inst = blockIdx.y;
    if (done[inst]) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long long o = static_cast<long long>(inst) * n;
    const double beta = first ? 0.0 : rz_new[2 * inst] / rz_old[2 * inst];
    p[o + j] = sc[j] * zz[o + j] + (first ? 0.0 : beta * p[o + j]);
}

// x += alpha p, alpha = rz / p.q.
__global__ __launch_bounds__(256) void k_g_step(int n, const double* __restrict__ p, const double* __restrict__ rz,
                                                const double* __restrict__ pq, double* __restrict__ x,
                                                const int* __restrict__ done) {
    const int inst = blockIdx.y;
    if (done[inst]) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double den = pq[2 * inst];
    const double alpha = den > 0 ? rz[2 * inst] / den : 0.0;
    const long long o = static_cast<long long>(inst) * n;
    x[o + j] += alpha * p[o + j];
}

// Relative residual of every open right-hand side. The best iterate so far
// is kept (keep[i] = 1: copy x); a right-hand side is done at tol, after
// kStall iterations without a new best (the residual has reached what FP64
// arithmetic can resolve for this matrix; beyond it CG only drifts), or on a
// breakdown.
constexpr int kStall = 4;
__global__ void k_g_check(int k, const double* __restrict__ sums, double* __restrict__ rnorm,
                          int* __restrict__ iters, int* __restrict__ stall, int* __restrict__ keep,
                          int* __restrict__ done, double tol, int first, int* __restrict__ open) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k || done[i]) return;
    keep[i] = 0;
    const double rr = sums[2 * i], bb = sums[2 * i + 1];
    const double rel = bb > 0 ? sqrt(rr / bb) : 0.0;
    if (!(rel == rel)) {
        done[i] = 1;
        return;
    }
    if (first || rel < rnorm[i]) {
        rnorm[i] = rel;
        stall[i] = 0;
        keep[i] = first ? 0 : 1;
    } else {
        stall[i] += 1;
    }
    if (rel <= tol || stall[i] >= kStall) {
        done[i] = 1;
    } else {
        if (!first) iters[i] += 1;
        atomicAdd(open, 1);
    }
}

// xb = x for the right-hand sides whose iterate is a new best.
__global__ __launch_bounds__(256) void k_g_keep(int n, const int* __restrict__ keep, const double* __restrict__ x,
                                                double* __restrict__ xb) {
    const int inst = blockIdx.y;
    if (!keep[inst]) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long long o = static_cast<long long>(inst) * n;
    xb[o + j] = x[o + j];
}

// ---- very large fronts, row-parallel: the vectors of every right-hand side
// live in HBM (W, front f at woff[f] * k, right-hand side r at r * m), and
// each panel of 32 pivots is two launches over every front of the level: one
// workgroup per front solves the diagonal block for all right-hand sides,
// then workgroups over row chunks of 256 update the rows below (forward) or
// take partial dot products (backward, reduced in chunk order by the next
// diagonal launch). Every L block is read once per sweep for all right-hand
// sides, and the results do not depend on the launch's scheduling.

constexpr int kWidePanel = 64;  // pivots per panel: one wavefront's lanes

__global__ __launch_bounds__(256) void k_w_fgather(DPlan P, const int* __restrict__ ht,
                                                   const long long* __restrict__ woff, int n, int k,
                                                   const double* __restrict__ rs, double* __restrict__ W) {
    const int f = blockIdx.y, t = ht[f], r = blockIdx.z;
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    W[woff[f] * k + static_cast<long long>(r) * m + i] = i < np ? rs[static_cast<long long>(r) * n + P.first[t] + i] : 0.0;
}

// One child of each listed front (a child rank): its update vector added into the parent's rows.
__global__ __launch_bounds__(256) void k_w_extend(DPlan P, const int* __restrict__ chf, const int* __restrict__ chc,
                                                  const int* __restrict__ ht, const long long* __restrict__ woff,
                                                  int k, long long vstride, const double* __restrict__ V,
                                                  double* __restrict__ W) {
    const int f = chf[blockIdx.y], c = chc[blockIdx.y], t = ht[f], r = blockIdx.z;
    const int u = P.nupd[c], m = P.npiv[t] + P.nupd[t];
    const int ii = blockIdx.x * 256 + threadIdx.x;
    if (ii >= u) return;
    W[woff[f] * k + static_cast<long long>(r) * m + P.relmap[P.upd[c] + ii]] +=
        V[static_cast<long long>(r) * vstride + P.vpos[c] + ii];
}

// tri[c * kWidePanel + r] = L(j0 + r, j0 + c) of the jb x jb diagonal block
// (0 elsewhere), dinv[c] = 1 / L(j0 + c, j0 + c).
template <typename T>
__device__ __forceinline__ void stage_tri(const T* __restrict__ Lt, int m, int j0, int jb, double* __restrict__ tri,
                                          double* __restrict__ dinv) {
    for (int e = threadIdx.x; e < kWidePanel * kWidePanel; e += blockDim.x) {
        const int c = e / kWidePanel, r = e - c * kWidePanel;
        tri[e] = (r < jb && c < jb && r >= c) ? static_cast<double>(Lt[hcol(j0 + c, m) + (r - c)]) : 0.0;
    }
    for (int c = threadIdx.x; c < kWidePanel; c += blockDim.x)
        dinv[c] = c < jb ? 1.0 / static_cast<double>(Lt[hcol(j0 + c, m)]) : 0.0;
}

// Forward, panel s: the diagonal block solved for right-hand sides 16 y .. 16 y + 15.
template <typename T>
__global__ __launch_bounds__(256) void k_w_fdiag(DPlan P, const int* __restrict__ ht,
                                                 const long long* __restrict__ woff, int k, int s,
                                                 const T* __restrict__ L, double* __restrict__ W) {
    __shared__ double tri[kWidePanel * kWidePanel], dinv[kWidePanel];
    const int f = blockIdx.x, t = ht[f];
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int j0 = kWidePanel * s;
    if (j0 >= np) return;
    const int jb = min(kWidePanel, np - j0);
    stage_tri<T>(L + P.lpos[t], m, j0, jb, tri, dinv);
    __syncthreads();
    double* Wf = W + woff[f] * k;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int rend = min(k, static_cast<int>(blockIdx.y) * 16 + 16);
    for (int r = blockIdx.y * 16 + wave; r < rend; r += 4) {
        double* w = Wf + static_cast<long long>(r) * m + j0;
        double v = lane < jb ? w[lane] : 0.0;
        for (int q = 0; q < jb; ++q) {
            const double yq = bcast(v, q) * dinv[q];
            if (lane > q && lane < jb) v -= tri[q * kWidePanel + lane] * yq;
            if (lane == q) v = yq;
        }
        if (lane < jb) w[lane] = v;
    }
}

// Forward, panel s: rows below the panel, 256 per workgroup, for right-hand sides 64 z .. 64 z + 63.
template <typename T>
__global__ __launch_bounds__(256) void k_w_fupdate(DPlan P, const int* __restrict__ ht,
                                                   const long long* __restrict__ woff, int k, int s,
                                                   const T* __restrict__ L, double* __restrict__ W) {
    __shared__ double ys[64 * kWidePanel];
    const int f = blockIdx.y, t = ht[f];
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int j0 = kWidePanel * s;
    if (j0 >= np) return;
    const int jb = min(kWidePanel, np - j0);
    const int r0 = 64 * blockIdx.z, rn = min(64, k - r0);
    double* Wf = W + woff[f] * k;
    for (int e = threadIdx.x; e < rn * kWidePanel; e += 256) {
        const int r = e / kWidePanel, q = e - r * kWidePanel;
        ys[e] = q < jb ? Wf[static_cast<long long>(r0 + r) * m + j0 + q] : 0.0;
    }
    __syncthreads();
    const int i = j0 + jb + blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const T* Lt = L + P.lpos[t];
    T Lq[kWidePanel];
#pragma unroll
    for (int q = 0; q < kWidePanel; ++q) Lq[q] = q < jb ? Lt[hcol(j0 + q, m) + i - j0 - q] : T(0);
    for (int r = 0; r < rn; ++r) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < kWidePanel; ++q) a[q & 3] += static_cast<double>(Lq[q]) * ys[r * kWidePanel + q];
        Wf[static_cast<long long>(r0 + r) * m + i] -= (a[0] + a[1]) + (a[2] + a[3]);
    }
}

__global__ __launch_bounds__(256) void k_w_fscatter(DPlan P, const int* __restrict__ ht,
                                                    const long long* __restrict__ woff, int n, int k,
                                                    long long vstride, const double* __restrict__ W,
                                                    double* __restrict__ y, double* __restrict__ V,
                                                    const int* __restrict__ done) {
    const int f = blockIdx.y, t = ht[f], r = blockIdx.z;
    if (done[r]) return;
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const double v = W[woff[f] * k + static_cast<long long>(r) * m + i];
    if (i < np) y[static_cast<long long>(r) * n + P.first[t] + i] = v;
    else V[static_cast<long long>(r) * vstride + P.vpos[t] + i - np] = v;
}

__global__ __launch_bounds__(256) void k_w_bgather(DPlan P, const int* __restrict__ ht,
                                                   const long long* __restrict__ woff, int n, int k,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   double* __restrict__ W) {
    const int f = blockIdx.y, t = ht[f], r = blockIdx.z;
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const long long off = static_cast<long long>(r) * n;
    W[woff[f] * k + static_cast<long long>(r) * m + i] =
        i < np ? y[off + P.first[t] + i] : z[off + P.rows[P.upd[t] + i - np]];
}

// Backward, step s (the block of pivots s-th from the last): partial dot
// products of the block's columns with the solved rows below, chunk c of 256
// rows, right-hand sides 64 z .. 64 z + 63; thread (q, g) takes column q and
// right-hand sides kRg g .. kRg g + kRg - 1.
template <typename T>
__global__ __launch_bounds__(256) void k_w_bdots(DPlan P, const int* __restrict__ ht,
                                                 const long long* __restrict__ woff,
                                                 const long long* __restrict__ cpre, int k, int s,
                                                 const T* __restrict__ L, const double* __restrict__ W,
                                                 double* __restrict__ part) {
    __shared__ double xs[64 * 64];
    const int f = blockIdx.y, t = ht[f];
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int b = (np + kWidePanel - 1) / kWidePanel - 1 - s;
    if (b < 0) return;
    const int jlo = kWidePanel * b, jb = min(kWidePanel, np - jlo);
    const int i0 = jlo + jb + 256 * static_cast<int>(blockIdx.x);
    if (i0 >= m) return;
    const int i1 = min(m, i0 + 256);
    const int r0 = 64 * blockIdx.z, rn = min(64, k - r0);
    constexpr int kRg = 64 * kWidePanel / 256;  // right-hand sides of a thread
    const int q = threadIdx.x % kWidePanel, g = threadIdx.x / kWidePanel;
    const T* Lt = L + P.lpos[t];
    const long long colbase = hcol(jlo + q, m) - (jlo + q);  // L(i, jlo + q) = Lt[colbase + i], i > jlo + q
    const double* Wf = W + woff[f] * k;
    double acc[kRg];
#pragma unroll
    for (int j = 0; j < kRg; ++j) acc[j] = 0.0;
    for (int s0 = i0; s0 < i1; s0 += 64) {
        const int cnt = min(64, i1 - s0);
        __syncthreads();
        for (int e = threadIdx.x; e < rn * 64; e += 256) {
            const int rr = e >> 6, ii = e & 63;
            xs[e] = ii < cnt ? Wf[static_cast<long long>(r0 + rr) * m + s0 + ii] : 0.0;
        }
        __syncthreads();
        if (q < jb)
            for (int ii = 0; ii < cnt; ++ii) {
                const double l = static_cast<double>(Lt[colbase + s0 + ii]);
#pragma unroll
                for (int j = 0; j < kRg; ++j) acc[j] += l * xs[(g * kRg + j < rn ? g * kRg + j : 0) * 64 + ii];
            }
    }
    double* pf = part + (cpre[f] + blockIdx.x) * static_cast<long long>(k) * kWidePanel;
#pragma unroll
    for (int j = 0; j < kRg; ++j) {
        const int rr = g * kRg + j;
        if (rr < rn) pf[static_cast<long long>(r0 + rr) * kWidePanel + q] = acc[j];
    }
}

// Backward, step s: the partial dots summed in chunk order, then the block's
// transposed triangular solve, right-hand sides 16 y .. 16 y + 15.
template <typename T>
__global__ __launch_bounds__(256) void k_w_bdiag(DPlan P, const int* __restrict__ ht,
                                                 const long long* __restrict__ woff,
                                                 const long long* __restrict__ cpre, int k, int s,
                                                 const T* __restrict__ L, const double* __restrict__ part,
                                                 double* __restrict__ W) {
    __shared__ double tri[kWidePanel * kWidePanel], dinv[kWidePanel];
    const int f = blockIdx.x, t = ht[f];
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int b = (np + kWidePanel - 1) / kWidePanel - 1 - s;
    if (b < 0) return;
    const int jlo = kWidePanel * b, jb = min(kWidePanel, np - jlo);
    const int nch = (m - jlo - jb + 255) / 256;
    stage_tri<T>(L + P.lpos[t], m, jlo, jb, tri, dinv);
    __syncthreads();
    double* Wf = W + woff[f] * k;
    const double* pf = part + cpre[f] * static_cast<long long>(k) * kWidePanel;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int rend = min(k, static_cast<int>(blockIdx.y) * 16 + 16);
    for (int r = blockIdx.y * 16 + wave; r < rend; r += 4) {
        double d = 0.0;
        if (lane < jb)
            for (int c = 0; c < nch; ++c) d += pf[(static_cast<long long>(c) * k + r) * kWidePanel + lane];
        double* w = Wf + static_cast<long long>(r) * m + jlo;
        double v = lane < jb ? w[lane] - d : 0.0;
        for (int qq = jb - 1; qq >= 0; --qq) {
            const double xq = bcast(v, qq) * dinv[qq];
            if (lane < qq) v -= tri[lane * kWidePanel + qq] * xq;
            if (lane == qq) v = xq;
        }
        if (lane < jb) w[lane] = v;
    }
}

__global__ __launch_bounds__(256) void k_w_bscatter(DPlan P, const int* __restrict__ ht,
                                                    const long long* __restrict__ woff, int n, int k,
                                                    const double* __restrict__ W, double* __restrict__ z,
                                                    const int* __restrict__ done) {
    const int f = blockIdx.y, t = ht[f], r = blockIdx.z;
    if (done[r]) return;
    const int np = P.npiv[t], m = np + P.nupd[t];
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= np) return;
    z[static_cast<long long>(r) * n + P.first[t] + j] = W[woff[f] * k + static_cast<long long>(r) * m + j];
}

// The large fronts of one level, their tile arena and child lists.
}  // namespace

struct SystemSolver::Impl {
    int precision;
    hipStream_t s;
    DevPlan plan;
    std::vector<HugeLevel> lev_huge;  // large fronts of each level
    long long huge_tiles = 0;          // largest arena of a level
    int huge_rows = 0;                 // largest large front (rows, RHS row included)
    // matrix (caller order)
    DevBuf<long long> ptr;
    DevBuf<int> adj, perm, iperm;
    DevBuf<double> diag, off, sc;
    DevBuf<long long> eptr;    // the matrix in elimination order (products)
    DevBuf<int> eadj;
    DevBuf<double> ediag, eoff;
    // factor
    DevBuf<char> L, U, A, F, D;
    DevBuf<double> rs0, y0;
    DevBuf<int> one, zero, fail;
    // solve scratch (grown on demand)
    int kcap = 0;
    DevBuf<double> bp, x, y, z, rs, V, pv, qv, part, sums, rz, pq, rnorm;
    DevBuf<double> xb;
    DevBuf<int> active, done, iters, stall, keep, open;

    template <typename T>
    void factor_t(const HostMatrix& m, double& analyze_ms, double& factor_ms);
    template <typename T>
    void numeric_t(const HostMatrix* m, double& analyze_ms, double& factor_ms);
    // FP64 in place of a failed FP32 factor: the same plan and values, L in FP64 from now on
    void to_fp64(double& analyze_ms, double& factor_ms);
    int fallbacks = 0;  // FP64 refactorisations after a failed or stalled FP32 factor
    // the structure the plan was built for (a refactorisation must match it)
    std::vector<i64> ptr_host;
    RawVec<i32> adj_host;
    template <typename T>
    void sweep_t(int k, int mode);
    template <typename T>
    void huge_forward(const HugeLevel& H, int k);
    template <typename T>
    void huge_backward(const HugeLevel& H, int k);
    DevBuf<double> wbuf, pbuf;  // row-parallel substitution: vectors and partial dots
    // The preconditioner sweep replayed as a graph from the second PCG
    // iteration on; `sweep_gen` counts the factorisations and scratch
    // allocations its captured arguments depend on.
    LaunchGraph g_sweep;
    long long sweep_gen = 0;
    hipStream_t owned = nullptr;  // the solver's stream when the caller gave the null stream
    Impl() = default;
    Impl(const Impl&) = delete;
    Impl& operator=(const Impl&) = delete;
    ~Impl() {
        g_sweep.clear();
        if (owned) {
            (void)hipStreamSynchronize(owned);
            (void)hipStreamDestroy(owned);
        }
    }
    template <typename T>
    void wide_forward(const HugeLevel& H, int k);
    template <typename T>
    void wide_backward(const HugeLevel& H, int k);
    template <typename T>
    void solve_t(int k, const double* d_b, double* d_x, double tol, int maxit, IrStats& st, std::vector<double>& rn,
                 std::vector<int>& its);
};

template <typename T>
void SystemSolver::Impl::factor_t(const HostMatrix& m, double& analyze_ms, double& factor_ms) {
    const auto h0 = std::chrono::steady_clock::now();
    g_sweep.clear();
    ++sweep_gen;
    progress("analysis of %d unknowns", m.n);
    Plan p = graph_plan(m.bx, m.by, m.cell, m.ptr, m.adj.data(), kGraphLeaf);
    progress("plan: %d fronts, L %lld entries, largest front %d", p.nfront, static_cast<long long>(p.lsize),
             p.max_front);
    plan.upload(std::move(p), s);
    const Plan& h = plan.host;
    huge_tiles = build_huge_levels(h, kWidePanel, lev_huge, huge_rows, s);
    std::vector<long long> ptrl(m.ptr.begin(), m.ptr.end());
    ptr = to_device(ptrl, s);
    adj = to_device(m.adj, s);
    ptr_host = m.ptr;
    adj_host = m.adj;
    // the solve scratch is sized by n and the plan's update vectors: a new
    // plan sizes it again at the next solve
    kcap = 0;
    analyze_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    numeric_t<T>(&m, analyze_ms, factor_ms);
}

// The values of `m` (the structure of the last analysis) uploaded, permuted
// and scaled, then the numeric factorisation.
template <typename T>
void SystemSolver::Impl::numeric_t(const HostMatrix* m, double& analyze_ms, double& factor_ms) {
    const auto h0 = std::chrono::steady_clock::now();
    g_sweep.clear();
    ++sweep_gen;
    const Plan& h = plan.host;
    const DPlan& P = plan.d;
    const int H = P.H, n = P.n;
    if (m) {  // new values (m == nullptr: the values already on the device)
        diag = to_device(m->diag, s);
        off = m->off.empty() ? to_device(std::vector<double>{0.0}, s) : to_device(m->off, s);
    }
    perm = to_device(h.perm, s);
    iperm = to_device(h.iperm, s);
    {
        std::vector<long long> ep(static_cast<std::size_t>(n) + 1, 0);
        for (int j = 0; j < n; ++j) ep[j + 1] = ep[j] + (ptr_host[h.perm[j] + 1] - ptr_host[h.perm[j]]);
        eptr = to_device(ep, s);
        eadj.alloc(ep[n] ? ep[n] : 1);
        eoff.alloc(ep[n] ? ep[n] : 1);
        ediag.alloc(n ? n : 1);
        if (n)
            hipLaunchKernelGGL(k_g_permute, dim3(nblocks(n, 256)), dim3(256), 0, s, n, perm.p, iperm.p, ptr.p, adj.p,
                               diag.p, off.p, eptr.p, eadj.p, ediag.p, eoff.p);
    }
    sc.alloc(n ? n : 1);
    if (n) hipLaunchKernelGGL(k_g_scale, dim3(nblocks(n, 256)), dim3(256), 0, s, n, perm.p, diag.p, sc.p);
    L.alloc(static_cast<std::size_t>(P.lsize + 4) * sizeof(T));  // + 4: whole 16-byte loads at the end
    U.alloc(static_cast<std::size_t>(P.usizeA ? P.usizeA : 1) * sizeof(T));
    A.alloc(static_cast<std::size_t>(P.nent ? P.nent : 1) * sizeof(T));
    F.alloc(static_cast<std::size_t>(huge_tiles ? huge_tiles : 1) * 256 * sizeof(T));
    std::size_t dmax = 1;
    for (const HugeLevel& hl : lev_huge) dmax = std::max(dmax, hl.t.size());
    D.alloc(dmax * 256 * sizeof(T));
    rs0.alloc(n ? n : 1);
    y0.alloc(n ? n : 1);
    LSC_HIP(hipMemsetAsync(rs0.p, 0, (n ? n : 1) * sizeof(double), s));
    one = to_device(std::vector<int>{1}, s);
    zero = to_device(std::vector<int>{0}, s);
    fail.alloc(1);
    sys_reset(fail.p, 1, s);
    LSC_HIP(hipStreamSynchronize(s));
    analyze_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    hipEvent_t ev0, ev1;
    LSC_HIP(hipEventCreate(&ev0));
    LSC_HIP(hipEventCreate(&ev1));
    LSC_HIP(hipEventRecord(ev0, s));
    T* Ap = reinterpret_cast<T*>(A.p);
    T* Lp = reinterpret_cast<T*>(L.p);
    T* Up = reinterpret_cast<T*>(U.p);
    T* Fp = reinterpret_cast<T*>(F.p);
    T* Dp = reinterpret_cast<T*>(D.p);
    if (P.nent)
        hipLaunchKernelGGL(k_g_assemble<T>, dim3(nblocks(P.nent, 256)), dim3(256), 0, s, P.nent, P.fe_p, P.fe_q,
                           P.fe_src, off.p, sc.p, Ap);
    int smax = 1;
    for (int v : plan.lev_stile) smax = std::max(smax, v);
    sys_allow(reinterpret_cast<const void*>(&k_factor_small<T>),
              static_cast<std::size_t>(kSmallWaves) * smax * 256 * sizeof(T));
    sys_allow(reinterpret_cast<const void*>(&k_factor_big<T>),
              static_cast<std::size_t>(kBigTiles * (kBigTiles + 1) / 2) * 256 * sizeof(T));
    for (int lv = 0; lv < H; ++lv) {
        const int ns = plan.lev_small_host[lv];
        const int nbig = plan.lev_fit_host[lv] - ns;  // fronts of 65-159 rows
        if (ns > 0)
            hipLaunchKernelGGL(k_factor_small<T>, dim3(nblocks(ns, kSmallWaves)), dim3(kSmallWaves * 64),
                               static_cast<std::size_t>(kSmallWaves) * plan.lev_stile[lv] * 256 * sizeof(T), s, P,
                               plan.host.lev_ptr[lv], ns, 1, plan.lev_stile[lv], Ap, rs0.p, one.p, Lp, y0.p, Up, fail.p);
        if (nbig > 0)
            hipLaunchKernelGGL(k_factor_big<T>, dim3(nbig), dim3(kBigWaves * 64),
                               static_cast<std::size_t>(plan.lev_btile[lv]) * 256 * sizeof(T), s, P,
                               plan.host.lev_ptr[lv] + ns, nbig, 1, Ap, rs0.p, one.p, Lp, y0.p, Up, fail.p);
        HugeLevel& hl = lev_huge[lv];
        if (hl.t.empty()) continue;
        const unsigned nh = static_cast<unsigned>(hl.t.size());
        HugeInst one;  // a single factorisation: no instance strides, one failure flag
        one.nh = static_cast<int>(nh);
        one.fail = fail.p;
        hipLaunchKernelGGL(k_h_zero<T>, dim3(nblocks(hl.tiles * 256, 256)), dim3(256), 0, s, Fp, hl.tiles * 256);
        hipLaunchKernelGGL(k_h_init<T>, dim3(64, nh), dim3(256), 0, s, P, one, hl.dt.p, hl.doff.p, Fp, Ap,
                           static_cast<const double*>(nullptr));
        for (std::size_t r = 0; r < hl.chf.size(); ++r)
            hipLaunchKernelGGL(k_h_extend<T>, dim3(static_cast<unsigned>(hl.cbk[r].back()), 1), dim3(256), 0, s, P, one,
                               static_cast<int>(hl.chf[r].size()), hl.dchf[r].p, hl.dchc[r].p, hl.dcb[r].p, hl.dt.p,
                               hl.doff.p, Fp, Up);
        for (int Kt = 0; Kt < hl.maxKt; ++Kt) {
            hipLaunchKernelGGL(k_h_panel<T>, dim3(hl.rb[Kt], nh), dim3(256), 0, s, P, one, Kt, hl.dt.p, hl.doff.p, Fp,
                               Dp);
            hipLaunchKernelGGL(k_h_diag<T>, dim3(nh), dim3(256), 0, s, P, one, Kt, hl.dt.p, hl.doff.p, Fp, Dp);
            if (hl.tb[Kt] > 0)
                hipLaunchKernelGGL(k_h_trailing<T>, dim3(hl.tb[Kt], nh), dim3(256), 0, s, P, one, Kt, hl.dt.p,
                                   hl.doff.p, Fp);
        }
        hipLaunchKernelGGL(k_h_store<T>, dim3(static_cast<unsigned>(hl.store_blocks), nh), dim3(256), 0, s, P, one,
                           hl.dt.p, hl.doff.p, Fp, Lp, Up, static_cast<double*>(nullptr));
    }
    LSC_HIP(hipEventRecord(ev1, s));
    progress("factor launched");
    int h_fail = 0;
    LSC_HIP(hipMemcpyAsync(&h_fail, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipStreamSynchronize(s));
    float fms = 0;
    LSC_HIP(hipEventElapsedTime(&fms, ev0, ev1));
    factor_ms += fms;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (h_fail) throw NotPositiveDefinite("a circuit component is not positive definite");
    // the update matrices are no longer needed
    U.release();
    A.release();
    F.release();
    D.release();
}

// Right-hand sides per workgroup of a level's large-front solves: as many as
// their vectors fit in 120 KB of LDS, at most 8.
inline int huge_rb(int mmax, int k) {
    const int fit = static_cast<int>((120 * 1024 / 8 - 272) / std::max(mmax, 1));
    int rb = 1;
    while (rb * 2 <= std::min(fit, 8) && rb < k) rb *= 2;
    return rb;
}

template <typename T, int RB>
void huge_forward_rb(const DPlan& P, const HugeLevel& H, int k, const T* L, const double* rs, double* y, double* V,
                     const int* done, hipStream_t s) {
    const std::size_t lds = (256 + static_cast<std::size_t>(RB) * H.mmax) * sizeof(double);
    sys_allow(reinterpret_cast<const void*>(&k_h_forward<T, RB>), lds);
    hipLaunchKernelGGL((k_h_forward<T, RB>), dim3(static_cast<unsigned>(H.t.size()), (k + RB - 1) / RB), dim3(256),
                       lds, s, P, H.dt.p, P.n, k, P.vsize, 0ll, L, rs, y, V, done, static_cast<const int*>(nullptr));
}

template <typename T, int RB>
void huge_backward_rb(const DPlan& P, const HugeLevel& H, int k, const T* L, const double* y, double* z,
                      const int* done, hipStream_t s) {
    const std::size_t lds = (256 + 16 * RB + static_cast<std::size_t>(RB) * H.mmax) * sizeof(double);
    sys_allow(reinterpret_cast<const void*>(&k_h_backward<T, RB>), lds);
    hipLaunchKernelGGL((k_h_backward<T, RB>), dim3(static_cast<unsigned>(H.t.size()), (k + RB - 1) / RB), dim3(256),
                       lds, s, P, H.dt.p, P.n, k, 0ll, L, y, z, done, static_cast<const int*>(nullptr));
}

// Rows from which a level's large fronts take the row-parallel substitution
// (LSC_WIDE_ROWS, default 1024): a workgroup per front then no longer keeps
// every right-hand side's vector in LDS, nor reads L once per group of them.
inline int wide_rows() {
    const char* e = std::getenv("LSC_WIDE_ROWS");
    return e ? std::atoi(e) : 1024;
}

inline unsigned cdiv(long long a, long long b) { return static_cast<unsigned>((a + b - 1) / b); }

template <typename T>
void SystemSolver::Impl::wide_forward(const HugeLevel& H, int k) {
    const DPlan& P = plan.d;
    const T* Lp = reinterpret_cast<const T*>(L.p);
    const unsigned F = static_cast<unsigned>(H.t.size());
    const std::size_t need = static_cast<std::size_t>(k) * H.msum;
    if (wbuf.n < need) wbuf.alloc(need), ++sweep_gen;
    hipLaunchKernelGGL(k_w_fgather, dim3(cdiv(H.mmax, 256), F, k), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, P.n, k,
                       rs.p, wbuf.p);
    for (std::size_t r = 0; r < H.chf.size(); ++r)
        hipLaunchKernelGGL(k_w_extend, dim3(cdiv(H.chmax[r], 256), static_cast<unsigned>(H.chf[r].size()), k),
                           dim3(256), 0, s, P, H.dchf[r].p, H.dchc[r].p, H.dt.p, H.dwoff.p, k,
                           static_cast<long long>(P.vsize), V.p, wbuf.p);
    for (int st = 0; st < H.wkt; ++st) {
        hipLaunchKernelGGL(k_w_fdiag<T>, dim3(F, cdiv(k, 16)), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, k, st, Lp,
                           wbuf.p);
        const long long below = H.mmax - static_cast<long long>(kWidePanel) * st;
        if (below > 0)
            hipLaunchKernelGGL(k_w_fupdate<T>, dim3(cdiv(below, 256), F, cdiv(k, 64)), dim3(256), 0, s, P, H.dt.p,
                               H.dwoff.p, k, st, Lp, wbuf.p);
    }
    hipLaunchKernelGGL(k_w_fscatter, dim3(cdiv(H.mmax, 256), F, k), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, P.n, k,
                       static_cast<long long>(P.vsize), wbuf.p, y.p, V.p, done.p);
}

template <typename T>
void SystemSolver::Impl::wide_backward(const HugeLevel& H, int k) {
    const DPlan& P = plan.d;
    const T* Lp = reinterpret_cast<const T*>(L.p);
    const unsigned F = static_cast<unsigned>(H.t.size());
    const std::size_t need = static_cast<std::size_t>(k) * H.msum;
    if (wbuf.n < need) wbuf.alloc(need), ++sweep_gen;
    const std::size_t pneed = static_cast<std::size_t>(k) * kWidePanel * H.csum;
    if (pbuf.n < pneed) pbuf.alloc(pneed), ++sweep_gen;
    hipLaunchKernelGGL(k_w_bgather, dim3(cdiv(H.mmax, 256), F, k), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, P.n, k,
                       y.p, z.p, wbuf.p);
    for (int st = 0; st < H.wkt; ++st) {
        hipLaunchKernelGGL(k_w_bdots<T>, dim3(cdiv(H.mmax, 256), F, cdiv(k, 64)), dim3(256), 0, s, P, H.dt.p,
                           H.dwoff.p, H.dcpre.p, k, st, Lp, wbuf.p, pbuf.p);
        hipLaunchKernelGGL(k_w_bdiag<T>, dim3(F, cdiv(k, 16)), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, H.dcpre.p, k, st,
                           Lp, pbuf.p, wbuf.p);
    }
    hipLaunchKernelGGL(k_w_bscatter, dim3(cdiv(H.mmax, 256), F, k), dim3(256), 0, s, P, H.dt.p, H.dwoff.p, P.n, k,
                       wbuf.p, z.p, done.p);
}

template <typename T>
void SystemSolver::Impl::huge_forward(const HugeLevel& H, int k) {
    if (H.t.empty()) return;
    if (H.mmax >= wide_rows()) return wide_forward<T>(H, k);
    const T* Lp = reinterpret_cast<const T*>(L.p);
    switch (huge_rb(H.mmax, k)) {
        case 8: huge_forward_rb<T, 8>(plan.d, H, k, Lp, rs.p, y.p, V.p, done.p, s); break;
        case 4: huge_forward_rb<T, 4>(plan.d, H, k, Lp, rs.p, y.p, V.p, done.p, s); break;
        case 2: huge_forward_rb<T, 2>(plan.d, H, k, Lp, rs.p, y.p, V.p, done.p, s); break;
        default: huge_forward_rb<T, 1>(plan.d, H, k, Lp, rs.p, y.p, V.p, done.p, s); break;
    }
}

template <typename T>
void SystemSolver::Impl::huge_backward(const HugeLevel& H, int k) {
    if (H.t.empty()) return;
    if (H.mmax >= wide_rows()) return wide_backward<T>(H, k);
    const T* Lp = reinterpret_cast<const T*>(L.p);
    switch (huge_rb(H.mmax, k)) {
        case 8: huge_backward_rb<T, 8>(plan.d, H, k, Lp, y.p, z.p, done.p, s); break;
        case 4: huge_backward_rb<T, 4>(plan.d, H, k, Lp, y.p, z.p, done.p, s); break;
        case 2: huge_backward_rb<T, 2>(plan.d, H, k, Lp, y.p, z.p, done.p, s); break;
        default: huge_backward_rb<T, 1>(plan.d, H, k, Lp, y.p, z.p, done.p, s); break;
    }
}

template <typename T>
void SystemSolver::Impl::sweep_t(int k, int mode) {
    const DPlan& P = plan.d;
    const T* Lp = reinterpret_cast<const T*>(L.p);
    auto lds = [&](int lv) {
        return plan.lev_mmax[lv] <= 64 ? rhs_lds<T, 1>(plan.lev_lmax[lv]) : rows_lds<T>(plan.lev_lmax[lv]);
    };
    auto rgrid = [&](int lv) {
        return dim3(static_cast<unsigned>(plan.lev_fit_host[lv]), static_cast<unsigned>((k + kRhs - 1) / kRhs));
    };
    if (mode == 1)
        for (int lv = 0; lv < P.H; ++lv) {
            auto kern = plan.lev_mmax[lv] <= 64 ? k_forward_rhs<T, 1> : k_forward_rows<T, 3>;
            if (plan.lev_fit_host[lv] > 0)
                hipLaunchKernelGGL(kern, rgrid(lv), dim3(64), lds(lv), s, P, lv, k, done.p, Lp, rs.p, y.p, V.p);
            huge_forward<T>(lev_huge[lv], k);
        }
    for (int lv = P.H - 1; lv >= 0; --lv) {
        huge_backward<T>(lev_huge[lv], k);
        auto kern = plan.lev_mmax[lv] <= 64 ? k_backward_rhs<T, 1> : k_backward_rows<T, 3>;
        if (plan.lev_fit_host[lv] > 0)
            hipLaunchKernelGGL(kern, rgrid(lv), dim3(64), lds(lv), s, P, lv, k, done.p, Lp, y.p, z.p);
    }
}

// X = A^-1 B by preconditioned conjugate gradients in FP64: the first
// iterate is the direct solve with the factor, and every iteration applies
// the factor once more, until ||b - A x|| / ||b|| <= tol (FP64, from the CSR
// matrix), `maxit` further iterations, or kStall iterations without a better
// residual; the best iterate is returned. Unlike plain refinement, this
// converges when the FP32 factor of a badly conditioned Laplacian is only a
// rough inverse.
template <typename T>
void SystemSolver::Impl::solve_t(int k, const double* d_b, double* d_x, double tol, int maxit, IrStats& st,
                                 std::vector<double>& rn, std::vector<int>& its) {
    const DPlan& P = plan.d;
    const int n = P.n;
    if (k <= 0 || n == 0) return;
    const unsigned nb = nblocks(n, 256);
    if (k > kcap) {
        const std::size_t kn = static_cast<std::size_t>(k) * n;
        bp.alloc(kn), x.alloc(kn), y.alloc(kn), z.alloc(kn), rs.alloc(kn), pv.alloc(kn), qv.alloc(kn);
        V.alloc(static_cast<std::size_t>(k) * (P.vsize ? P.vsize : 1));
        part.alloc(static_cast<std::size_t>(k) * nb * 2);
        sums.alloc(2 * k), rz.alloc(4 * k), pq.alloc(2 * k), rnorm.alloc(k);
        xb.alloc(kn);
        active.alloc(k), done.alloc(k), iters.alloc(k), stall.alloc(k), keep.alloc(k);
        if (!open.p) open.alloc(1);
        kcap = k;
        ++sweep_gen;
        std::vector<int> ones(k, 1);
        active.upload(ones.data(), k, s);
    }
    int lmax = 1;
    for (int v : plan.lev_lmax) lmax = std::max(lmax, v);
    sys_allow(reinterpret_cast<const void*>(&k_forward_rhs<T, 1>), rhs_lds<T, 1>(lmax));
    sys_allow(reinterpret_cast<const void*>(&k_forward_rows<T, 3>), rows_lds<T>(lmax));
    sys_allow(reinterpret_cast<const void*>(&k_backward_rhs<T, 1>), rhs_lds<T, 1>(lmax));
    sys_allow(reinterpret_cast<const void*>(&k_backward_rows<T, 3>), rows_lds<T>(lmax));
    const long long kn = static_cast<long long>(k) * n;
    const dim3 g2(nb, k);
    const dim3 gk(nb, cdiv(k, kSpmvRhs));  // the products: kSpmvRhs right-hand sides per thread
    sys_reset(done.p, k, s);
    sys_reset(iters.p, k, s);
    hipLaunchKernelGGL(k_g_perm_in, dim3(nblocks(kn, 256)), dim3(256), 0, s, n, k, perm.p, d_b, sc.p, bp.p, rs.p, x.p);
    hipLaunchKernelGGL(k_g_perm_in, dim3(nblocks(kn, 256)), dim3(256), 0, s, n, k, perm.p, d_b, sc.p, bp.p, rs.p, xb.p);
    auto residual_check = [&](int first) {
        sys_reset(open.p, 1, s);
        hipLaunchKernelGGL(k_g_residual, gk, dim3(256), 0, s, n, k, eptr.p, eadj.p, ediag.p, eoff.p, bp.p, x.p, sc.p,
                           rs.p, done.p, part.p);
        hipLaunchKernelGGL(k_g_sum, dim3(k), dim3(256), 0, s, static_cast<int>(nb), part.p, done.p, sums.p);
        hipLaunchKernelGGL(k_g_check, dim3(nblocks(k, 256)), dim3(256), 0, s, k, sums.p, rnorm.p, iters.p, stall.p,
                           keep.p, done.p, tol, first, open.p);
        if (!first) hipLaunchKernelGGL(k_g_keep, g2, dim3(256), 0, s, n, keep.p, x.p, xb.p);
        int h_open = 0;
        LSC_HIP(hipMemcpyAsync(&h_open, open.p, sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipStreamSynchronize(s));
        return h_open;
    };
    const bool trace = std::getenv("LSC_PCG_TRACE") != nullptr;
    auto show = [&](int it, const double* rzp) {
        if (!trace) return;
        double hs[2], hr[2], hp[2], hn;
        LSC_HIP(hipMemcpy(hs, sums.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
        LSC_HIP(hipMemcpy(hr, rzp, 2 * sizeof(double), hipMemcpyDeviceToHost));
        LSC_HIP(hipMemcpy(hp, pq.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
        LSC_HIP(hipMemcpy(&hn, rnorm.p, sizeof(double), hipMemcpyDeviceToHost));
        std::fprintf(stderr, "pcg n=%d k=%d it=%d rel=%.3e rr=%.3e bb=%.3e rz=%.3e pq=%.3e\n", n, k, it, hn, hs[0],
                     hs[1], hr[0], hp[0]);
    };
    int h_open = residual_check(1);  // r = b: zero right-hand sides are done at once
    for (int it = 0; h_open > 0 && it <= maxit; ++it) {
        double* rz_new = rz.p + 2 * k * (it & 1);
        double* rz_old = rz.p + 2 * k * ((it + 1) & 1);
        // z (scaled space) = (L L^T)^-1 s r; the first sweep runs eagerly and
        // grows the scratch buffers, later ones replay its graph
        if (it == 0)
            sweep_t<T>(k, 1);
        else
            g_sweep.run(s, (sweep_gen << 32) | k, graphs_enabled(), [&] { sweep_t<T>(k, 1); });
        st.solve_launches += 1;
        hipLaunchKernelGGL(k_g_dot, g2, dim3(256), 0, s, n, rs.p, z.p, done.p, part.p);
        hipLaunchKernelGGL(k_g_sum, dim3(k), dim3(256), 0, s, static_cast<int>(nb), part.p, done.p, rz_new);
        hipLaunchKernelGGL(k_g_dir, g2, dim3(256), 0, s, n, it == 0 ? 1 : 0, sc.p, z.p, rz_new, rz_old, pv.p, done.p);
        hipLaunchKernelGGL(k_g_spmv, gk, dim3(256), 0, s, n, k, eptr.p, eadj.p, ediag.p, eoff.p, pv.p, qv.p, done.p,
                           part.p);
        hipLaunchKernelGGL(k_g_sum, dim3(k), dim3(256), 0, s, static_cast<int>(nb), part.p, done.p, pq.p);
        hipLaunchKernelGGL(k_g_step, g2, dim3(256), 0, s, n, pv.p, rz_new, pq.p, x.p, done.p);
        h_open = residual_check(0);
        show(it, rz_new);
        progress("solve of %d right-hand sides: iteration %d, %d open", k, it + 1, h_open);
    }
    hipLaunchKernelGGL(k_g_perm_out, dim3(nblocks(kn, 256)), dim3(256), 0, s, n, k, perm.p, xb.p, d_x);
    its.assign(k, 0);
    rn.assign(k, 0.0);
    LSC_HIP(hipMemcpyAsync(its.data(), iters.p, k * sizeof(int), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipMemcpyAsync(rn.data(), rnorm.p, k * sizeof(double), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipStreamSynchronize(s));
}

void SystemSolver::Impl::to_fp64(double& analyze_ms, double& factor_ms) {
    progress("FP64 refactorisation of %d unknowns after the FP32 factor failed or stalled", plan.d.n);
    precision = 1;
    ++fallbacks;
    L.release();
    numeric_t<double>(nullptr, analyze_ms, factor_ms);
}

SystemSolver::SystemSolver(int precision, hipStream_t s) : impl_(new Impl) {
    impl_->precision = precision;
    impl_->s = s;
    // The null stream cannot be captured into graphs: work instead on a
    // blocking stream of our own, which the null stream orders both ways.
    if (!s) {
        LSC_HIP(hipStreamCreateWithFlags(&impl_->owned, hipStreamDefault));
        impl_->s = impl_->owned;
    }
}
SystemSolver::~SystemSolver() { delete impl_; }

void SystemSolver::factor(const HostMatrix& m) {
    if (m.n <= 0) {
        impl_->plan = DevPlan{};
        return;
    }
    if (impl_->precision == 1) {
        impl_->factor_t<double>(m, analyze_ms, factor_ms);
        return;
    }
    try {
        impl_->factor_t<float>(m, analyze_ms, factor_ms);
    } catch (const NotPositiveDefinite&) {
        impl_->to_fp64(analyze_ms, factor_ms);  // throws again if the matrix is not positive definite
    }
}

void SystemSolver::refactor(const HostMatrix& m) {
    if (m.n != impl_->plan.host.n || m.ptr != impl_->ptr_host || m.adj != impl_->adj_host)
        throw std::invalid_argument("a refactorisation needs the structure of the analysed matrix");
    if (m.n <= 0) return;
    if (impl_->precision == 1) {
        impl_->numeric_t<double>(&m, analyze_ms, factor_ms);
        return;
    }
    try {
        impl_->numeric_t<float>(&m, analyze_ms, factor_ms);
    } catch (const NotPositiveDefinite&) {
        impl_->to_fp64(analyze_ms, factor_ms);
    }
}

int SystemSolver::fallbacks() const { return impl_->fallbacks; }

void SystemSolver::solve_device(int k, const double* d_b, double* d_x, double tol, int maxit, IrStats& st) {
    if (impl_->plan.host.n == 0) return;
    hipEvent_t a, b;
    LSC_HIP(hipEventCreate(&a));
    LSC_HIP(hipEventCreate(&b));
    LSC_HIP(hipEventRecord(a, impl_->s));
    std::vector<double> rn;
    std::vector<int> its;
    if (impl_->precision == 1) {
        impl_->solve_t<double>(k, d_b, d_x, tol, maxit, st, rn, its);
    } else {
        impl_->solve_t<float>(k, d_b, d_x, tol, maxit, st, rn, its);
        // A right-hand side whose FP32-preconditioned iteration ended far above tol (more than
        // kFallbackGap times) has failed rather than met the FP64 floor of the residual: the solver
        // refactorises in FP64 (for this and every later solve) and solves those again, keeping the
        // better of the two solutions.
        std::vector<int> bad;
        for (int i = 0; i < k; ++i)
            if (!(rn[i] <= kFallbackGap * tol)) bad.push_back(i);
        if (!bad.empty()) {
            const int kb = static_cast<int>(bad.size());
            const std::size_t n = static_cast<std::size_t>(impl_->plan.d.n);
            DevBuf<double> b2(kb * n), x2(kb * n);
            // the right-hand sides from the solve's own copy in elimination order (d_b may be d_x)
            for (int j = 0; j < kb; ++j)
                hipLaunchKernelGGL(k_g_perm_out, dim3(nblocks(static_cast<long long>(n), 256)), dim3(256), 0, impl_->s,
                                   static_cast<int>(n), 1, impl_->perm.p, impl_->bp.p + bad[j] * n, b2.p + j * n);
            impl_->to_fp64(analyze_ms, factor_ms);
            std::vector<double> rn2;
            std::vector<int> its2;
            impl_->solve_t<double>(kb, b2.p, x2.p, tol, maxit, st, rn2, its2);
            for (int j = 0; j < kb; ++j) {
                const int i = bad[j];
                its[i] += its2[j];
                if (rn2[j] < rn[i] || !(rn[i] == rn[i])) {
                    LSC_HIP(hipMemcpyAsync(d_x + i * n, x2.p + j * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                                           impl_->s));
                    rn[i] = rn2[j];
                }
            }
            LSC_HIP(hipStreamSynchronize(impl_->s));
        }
    }
    for (int i = 0; i < k; ++i) {
        ++st.solved;
        st.iter_sum += its[i];
        st.iter_max = std::max(st.iter_max, its[i]);
        st.resid_max = std::max(st.resid_max, rn[i]);
        st.resid_sum += rn[i];
        st.resids.push_back(static_cast<float>(rn[i]));
        if (!(rn[i] <= tol)) ++st.unconverged;
        if (!(rn[i] <= 1e-10)) ++st.over_bar;
    }
    LSC_HIP(hipEventRecord(b, impl_->s));
    LSC_HIP(hipEventSynchronize(b));
    float ms = 0;
    LSC_HIP(hipEventElapsedTime(&ms, a, b));
    solve_ms += ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
}

void SystemSolver::solve(int k, const double* b, double* x, double tol, int maxit, IrStats& st) {
    const int nn = impl_->plan.host.n;
    if (nn == 0 || k <= 0) return;
    const std::size_t kn = static_cast<std::size_t>(k) * nn;
    DevBuf<double> db(kn), dx(kn);
    db.upload(b, kn, impl_->s);
    solve_device(k, db.p, dx.p, tol, maxit, st);
    LSC_HIP(hipMemcpyAsync(x, dx.p, kn * sizeof(double), hipMemcpyDeviceToHost, impl_->s));
    LSC_HIP(hipStreamSynchronize(impl_->s));
}

int SystemSolver::n() const { return impl_->plan.host.n; }
const Plan& SystemSolver::plan() const { return impl_->plan.host; }

}  // namespace lsc
