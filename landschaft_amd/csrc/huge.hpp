// Fronts of more than 159 rows ("large"): tiles in HBM, every large front of
// one tree level in the same launches, for one factorisation (the graph
// engine) or for a batch of instances of one plan (the window engine's
// windows of large radius). Shared by both engines.
//
// A launch covers `nh` large fronts of a level times a run of instances:
// slot s of the launch (a grid row) is front i = s % nh of instance
// inst = inst0 + s / nh; its tiles start at F + ((s / nh) * ftiles +
// hoff[i]) * 256, its diagonal-block scratch at D + s * 256. An instance's L,
// update matrices, assembled entries and vectors start at inst times their
// strides (all 0 for one factorisation).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "window_kernels.hpp"

namespace lsc {

namespace {

struct HugeInst {
    int nh = 1;              // large fronts of the level
    int inst0 = 0;           // first instance of the launch
    long long ftiles = 0;    // arena tiles per instance
    long long lstride = 0;   // L entries per instance
    long long ustride = 0;   // update-matrix entries per instance
    long long astride = 0;   // assembled entries per instance
    long long nstride = 0;   // vector entries (unknowns) per instance
    const int* active = nullptr;  // instances that run (nullptr: all)
    int* fail = nullptr;          // non-positive pivot flags
    int fail_stride = 0;          // per instance (0: one flag)
};

struct HugeSlot {
    int i, il, inst;  // front of the level, instance in the launch, instance
    bool run;
};

__device__ __forceinline__ HugeSlot huge_slot(const HugeInst& h, int s) {
    HugeSlot r;
    r.i = s % h.nh;
    r.il = s / h.nh;
    r.inst = h.inst0 + r.il;
    r.run = !h.active || h.active[r.inst];
    return r;
}

template <typename T>
__global__ void k_h_zero(T* __restrict__ F, long long count) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < count) F[e] = T(0);
}

// The unit diagonal of the pivots and the front's assembled entries; with
// `rs` (the window engine), the right-hand side row m of the pivot columns,
// so that the factorisation carries the forward substitution.
template <typename T>
__global__ void k_h_init(DPlan P, HugeInst H, const int* __restrict__ ht, const long long* __restrict__ hoff,
                         T* __restrict__ F, const T* __restrict__ aval, const double* __restrict__ rs) {
    const HugeSlot q = huge_slot(H, blockIdx.y);
    if (!q.run) return;
    const int t = ht[q.i];
    T* Ft = F + (q.il * H.ftiles + hoff[q.i]) * 256;
    const T* av = aval + q.inst * H.astride;
    const int np = P.npiv[t], m = np + P.nupd[t], first = P.first[t];
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (int j = tid; j < np; j += nt) {
        Ft[tix(j, j)] = T(1);
        if (rs) Ft[tix(m, j)] = static_cast<T>(rs[q.inst * H.nstride + first + j]);
    }
    for (int e = P.fe_ptr[t] + tid; e < P.fe_ptr[t + 1]; e += nt) Ft[tix(P.fe_row[e], P.fe_col[e])] = av[e];
}

// Extend-add of the r-th child of every large front that has one (children
// of one front go in separate launches, so no entry is written twice at
// once). Workgroup (b, il) belongs to entry j of the child list with
// cb[j] <= b < cb[j + 1], in instance il of the launch.
template <typename T>
__global__ void k_h_extend(DPlan P, HugeInst H, int nch, const int* __restrict__ chf, const int* __restrict__ chc,
                           const long long* __restrict__ cb, const int* __restrict__ ht,
                           const long long* __restrict__ hoff, T* __restrict__ F, const T* __restrict__ U) {
    const long long b = blockIdx.x;
    const int il = blockIdx.y, inst = H.inst0 + il;
    if (H.active && !H.active[inst]) return;
    int lo = 0, hi = nch - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cb[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int f = chf[lo], c = chc[lo], t = ht[f];
    const int m = P.npiv[t] + P.nupd[t];
    const int u = P.nupd[c], u1 = u + 1;
    const long long e = (b - cb[lo]) * blockDim.x + threadIdx.x;
    if (e >= static_cast<long long>(u1) * u) return;
    const int jj = static_cast<int>(e / u1), ii = static_cast<int>(e - static_cast<long long>(jj) * u1);
    if (ii < jj) return;
    const int* rel = P.relmap + P.upd[c];
    T* Ft = F + (il * H.ftiles + hoff[f]) * 256;
    Ft[tix(ii < u ? rel[ii] : m, rel[jj])] += U[inst * H.ustride + P.uposA[c] + colstart(jj, u1) + (ii - jj)];
}

// Panel step Kt of every large front still pivoting: workgroup (bx, s)
// solves rows 192 bx.. of the panel below the diagonal block; every one
// factors the (unfactored) diagonal block for itself, and workgroup 0 leaves
// the factored block in D for k_h_diag, after the launch.
template <typename T>
__global__ __launch_bounds__(256) void k_h_panel(DPlan P, HugeInst H, int Kt, const int* __restrict__ ht,
                                                 const long long* __restrict__ hoff, T* __restrict__ F,
                                                 T* __restrict__ D) {
    const HugeSlot q = huge_slot(H, blockIdx.y);
    if (!q.run) return;
    const int t = ht[q.i];
    const int np = P.npiv[t], ma = np + P.nupd[t] + 1;
    const int k0 = 16 * Kt;
    if (k0 >= np) return;
    const int below = ma - k0 - 16;
    if (blockIdx.x > 0 && static_cast<int>(192 * blockIdx.x) >= below) return;
    const int lane = threadIdx.x & 63, wave = wave_id();
    __shared__ __align__(16) T s_bc[4][64];
    const bool ok = panel<T>(F + (q.il * H.ftiles + hoff[q.i]) * 256, ma, k0, min(k0 + 16, np), wave, lane,
                             192 * blockIdx.x, blockIdx.x == 0 ? D + static_cast<long long>(blockIdx.y) * 256 : nullptr,
                             s_bc[wave], [] __device__() {});
    if (!ok && lane == 0) atomicOr(H.fail + static_cast<long long>(q.inst) * H.fail_stride, 1);
}

template <typename T>
__global__ void k_h_diag(DPlan P, HugeInst H, int Kt, const int* __restrict__ ht, const long long* __restrict__ hoff,
                         T* __restrict__ F, const T* __restrict__ D) {
    const HugeSlot q = huge_slot(H, blockIdx.x);
    if (!q.run) return;
    const int t = ht[q.i];
    if (16 * Kt >= P.npiv[t]) return;
    F[(q.il * H.ftiles + hoff[q.i]) * 256 + tile_index(Kt, Kt) * 256 + threadIdx.x] =
        D[static_cast<long long>(blockIdx.x) * 256 + threadIdx.x];
}

// Trailing update after panel step Kt: one 16 x 16 tile per wavefront.
template <typename T>
__global__ __launch_bounds__(256) void k_h_trailing(DPlan P, HugeInst H, int Kt, const int* __restrict__ ht,
                                                    const long long* __restrict__ hoff, T* __restrict__ F) {
    const HugeSlot q = huge_slot(H, blockIdx.y);
    if (!q.run) return;
    const int t = ht[q.i];
    const int np = P.npiv[t], ma = np + P.nupd[t] + 1;
    const int k0 = 16 * Kt;
    if (k0 >= np) return;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int MT = (ma + 15) >> 4, Tn = MT - Kt - 1;
    const long long ntiles = static_cast<long long>(Tn) * (Tn + 1) / 2;
    const long long tt = static_cast<long long>(blockIdx.x) * 4 + wave;
    if (tt >= ntiles) return;
    int I, J;
    tri_tile(static_cast<int>(tt), I, J);
    tile_update<T>(F + (q.il * H.ftiles + hoff[q.i]) * 256, Kt + 1 + I, Kt + 1 + J, Kt, k0, min(k0 + 16, np), lane);
}

// L (packed by columns) and the update matrix (its lower triangle packed by
// columns, the right-hand side row included) out of the tiles; with `y` (the
// window engine) the right-hand side row's pivot entries, the forward
// substitution's values.
template <typename T>
__global__ void k_h_store(DPlan P, HugeInst H, const int* __restrict__ ht, const long long* __restrict__ hoff,
                          const T* __restrict__ F, T* __restrict__ L, T* __restrict__ U, double* __restrict__ y) {
    const HugeSlot q = huge_slot(H, blockIdx.y);
    if (!q.run) return;
    const int t = ht[q.i];
    const T* Ft = F + (q.il * H.ftiles + hoff[q.i]) * 256;
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, u1 = nu + 1;
    const long long tid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long nt = static_cast<long long>(gridDim.x) * blockDim.x;
    T* Lt = L + q.inst * H.lstride + P.lpos[t];
    for (long long e = tid; e < static_cast<long long>(np) * m; e += nt) {
        const int j = static_cast<int>(e / m), i = static_cast<int>(e - static_cast<long long>(j) * m);
        if (i < j) continue;
        Lt[static_cast<long long>(j) * m - ((static_cast<long long>(j) * (j - 1)) >> 1) + (i - j)] = Ft[tix(i, j)];
    }
    T* Ut = U + q.inst * H.ustride + P.uposA[t];
    for (long long e = tid; e < static_cast<long long>(u1) * u1; e += nt) {
        const int jj = static_cast<int>(e / u1), ii = static_cast<int>(e - static_cast<long long>(jj) * u1);
        if (ii < jj) continue;
        Ut[colstart(jj, u1) + (ii - jj)] = Ft[tix(np + ii, np + jj)];
    }
    if (y)
        for (long long j = tid; j < np; j += nt) y[q.inst * H.nstride + P.first[t] + j] = static_cast<double>(Ft[tix(m, j)]);
}

// Start of column j of a front's L block (column j holds rows j..m-1).
__device__ __forceinline__ long long hcol(int j, int m) {
    return static_cast<long long>(j) * m - ((static_cast<long long>(j) * (j - 1)) >> 1);
}

// tri[c * 16 + r] = L(j0 + r, j0 + c) of the jb x jb diagonal block (0 above).
template <typename T>
__device__ __forceinline__ void stage_tri(const T* __restrict__ Lt, int m, int j0, int jb, double* __restrict__ tri) {
    const int r = threadIdx.x & 15, c = threadIdx.x >> 4;
    tri[threadIdx.x] = (r < jb && c < jb && r >= c) ? static_cast<double>(Lt[hcol(j0 + c, m) + (r - c)]) : 0.0;
}

// Forward substitution of the large fronts of one level for RB right-hand
// sides at once: workgroup (f, g) holds the vectors of right-hand sides
// g RB .. in LDS side by side, so every L column block is read once for all
// of them. Per block of 16 pivots the diagonal block is staged in LDS, the
// wavefronts solve it for one right-hand side each, and the workgroup
// updates the rows below (coalesced along the columns of L). With RB = 1,
// right-hand side g may have an L of its own at g * lstride (the window
// engine's instances) and skip when it is not `active`.
template <typename T, int RB>
__global__ __launch_bounds__(256) void k_h_forward(DPlan P, const int* __restrict__ ht, int n, int k,
                                                   long long vstride, long long lstride, const T* __restrict__ L,
                                                   const double* __restrict__ rs, double* __restrict__ y,
                                                   double* __restrict__ V, const int* __restrict__ done,
                                                   const int* __restrict__ active) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int t = ht[blockIdx.x];
    const int g0 = blockIdx.y * RB;
    if (RB == 1 && (g0 >= k || done[g0] || (active && !active[g0]))) return;
    double* tri = reinterpret_cast<double*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    double* fv = tri + 256;  // fv[r * m + i]
    for (int r = 0; r < RB; ++r) {
        const int inst = g0 + r < k ? g0 + r : k - 1;
        const long long off = static_cast<long long>(inst) * n;
        for (int i = threadIdx.x; i < m; i += 256) fv[r * m + i] = i < np ? rs[off + first + i] : 0.0;
    }
    __syncthreads();
    for (int ci = P.cb[t]; ci < P.ce[t]; ++ci) {
        const int c = P.kids[ci];
        const int u = P.nupd[c];
        const int* rel = P.relmap + P.upd[c];
        for (int r = 0; r < RB; ++r) {
            const int inst = g0 + r < k ? g0 + r : k - 1;
            const double* Vi = V + static_cast<long long>(inst) * vstride + P.vpos[c];
            for (int ii = threadIdx.x; ii < u; ii += 256) fv[r * m + rel[ii]] += Vi[ii];
        }
        __syncthreads();
    }
    const T* Lt = L + static_cast<long long>(g0) * lstride + P.lpos[t];
    for (int j0 = 0; j0 < np; j0 += 16) {
        const int jb = min(16, np - j0);
        stage_tri<T>(Lt, m, j0, jb, tri);
        __syncthreads();
        for (int r = wave; r < RB; r += 4) {
            double f = lane < jb ? fv[r * m + j0 + lane] : 0.0;
            for (int q = 0; q < jb; ++q) {
                const double yq = bcast(f, q) / tri[q * 17];
                if (lane > q && lane < jb) f -= tri[q * 16 + lane] * yq;
                if (lane == q) f = yq;
            }
            if (lane < jb) fv[r * m + j0 + lane] = f;
        }
        __syncthreads();
        for (int i = j0 + jb + threadIdx.x; i < m; i += 256) {
            double Lq[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                Lq[q] = q < jb ? static_cast<double>(Lt[hcol(j0 + q, m) + i - j0 - q]) : 0.0;
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                double a = 0.0;
#pragma unroll
                for (int q = 0; q < 16; ++q) a += Lq[q] * fv[r * m + j0 + (q < jb ? q : 0)];
                fv[r * m + i] -= a;
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < RB; ++r) {
        const int inst = g0 + r;
        if (inst >= k || done[inst]) continue;
        const long long off = static_cast<long long>(inst) * n;
        double* Vi = V + static_cast<long long>(inst) * vstride;
        for (int i = threadIdx.x; i < m; i += 256) {
            if (i < np) y[off + first + i] = fv[r * m + i];
            else Vi[P.vpos[t] + i - np] = fv[r * m + i];
        }
    }
}

// Back substitution of the large fronts of one level for RB right-hand sides
// at once, left-looking: per block of 16 pivots (from the last) each
// column's dot products with the solved part below are taken by one
// wavefront along the column (coalesced, one L read for all right-hand
// sides), then the wavefronts solve the staged diagonal block. lstride and
// active as in k_h_forward.
template <typename T, int RB>
__global__ __launch_bounds__(256) void k_h_backward(DPlan P, const int* __restrict__ ht, int n, int k,
                                                    long long lstride, const T* __restrict__ L,
                                                    const double* __restrict__ y, double* __restrict__ z,
                                                    const int* __restrict__ done, const int* __restrict__ active) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int t = ht[blockIdx.x];
    const int g0 = blockIdx.y * RB;
    if (RB == 1 && (g0 >= k || done[g0] || (active && !active[g0]))) return;
    double* tri = reinterpret_cast<double*>(smem_raw);
    double* dots = tri + 256;       // dots[r * 16 + cc]
    double* fv = dots + 16 * RB;    // fv[r * m + i]
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int* rows = P.rows + P.upd[t];
    for (int r = 0; r < RB; ++r) {
        const int inst = g0 + r < k ? g0 + r : k - 1;
        const long long off = static_cast<long long>(inst) * n;
        for (int i = threadIdx.x; i < m; i += 256)
            fv[r * m + i] = i < np ? y[off + first + i] : z[off + rows[i - np]];
    }
    __syncthreads();
    const T* Lt = L + static_cast<long long>(g0) * lstride + P.lpos[t];
    for (int j1 = np - 1; j1 >= 0; j1 -= 16) {
        const int jlo = max(0, j1 - 15), jb = j1 - jlo + 1;
        for (int cc = wave; cc < jb; cc += 4) {
            const int i = jlo + cc;
            const long long base = hcol(i, m) - i;  // L(r, i) = Lt[base + r]
            double a[RB];
#pragma unroll
            for (int r = 0; r < RB; ++r) a[r] = 0.0;
            for (int row = j1 + 1 + lane; row < m; row += 64) {
                const double l = static_cast<double>(Lt[base + row]);
#pragma unroll
                for (int r = 0; r < RB; ++r) a[r] += l * fv[r * m + row];
            }
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                double v = a[r];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                if (lane == 0) dots[r * 16 + cc] = v;
            }
        }
        stage_tri<T>(Lt, m, jlo, jb, tri);
        __syncthreads();
        for (int r = wave; r < RB; r += 4) {
            double v = lane < jb ? fv[r * m + jlo + lane] - dots[r * 16 + lane] : 0.0;
            for (int q = jb - 1; q >= 0; --q) {
                const double xq = bcast(v, q) / tri[q * 17];
                if (lane < q) v -= tri[lane * 16 + q] * xq;
                if (lane == q) v = xq;
            }
            if (lane < jb) fv[r * m + jlo + lane] = v;
        }
        __syncthreads();
    }
    for (int r = 0; r < RB; ++r) {
        const int inst = g0 + r;
        if (inst >= k || done[inst]) continue;
        const long long off = static_cast<long long>(inst) * n;
        for (int j = threadIdx.x; j < np; j += 256) z[off + first + j] = fv[r * m + j];
    }
}

}  // namespace

// The large fronts of every level of a plan, with their tile arenas, launch
// grids and child lists (host side). `wide_from`: fronts of this many rows
// or more get the graph engine's row-parallel substitution layout.
struct HugeLevel {
    std::vector<int> t;                 // fronts
    std::vector<long long> off;         // tile offsets in the arena
    long long tiles = 0;                // arena size of the level
    int maxKt = 0;                      // panel steps of the level's largest front
    int mmax = 0;                       // rows of the level's largest front (RHS row excluded)
    std::vector<int> rb, tb;            // per panel step: largest row-block and tile-block grids
    std::vector<std::vector<int>> chf, chc;  // per child rank: (front index, child)
    std::vector<std::vector<long long>> cbk; // per child rank: workgroup prefix
    std::vector<int> chmax;                  // per child rank: largest child update size
    long long store_blocks = 1;
    // row-parallel substitution: row prefix (woff) and 256-row chunk prefix (cpre) per front
    std::vector<long long> woff, cpre;
    long long msum = 0, csum = 0;
    int wkt = 0;                             // panels of the level's largest front (kWidePanel pivots)
    DevBuf<long long> dwoff, dcpre;
    DevBuf<int> dt;
    DevBuf<long long> doff;
    std::vector<DevBuf<int>> dchf, dchc;
    std::vector<DevBuf<long long>> dcb;
};

// Fills `lev` (one entry per level) for the plan's fronts of more than 159
// rows (RHS row included) and uploads their tables; returns the largest
// level arena (tiles) and sets `rows` to the largest front's rows.
inline long long build_huge_levels(const Plan& h, int wide_panel, std::vector<HugeLevel>& lev, int& rows,
                                   hipStream_t s) {
    const int H = static_cast<int>(h.lev_ptr.size()) - 1;
    lev.clear();
    lev.resize(H);
    long long max_tiles = 0;
    rows = 0;
    for (int lv = 0; lv < H; ++lv) {
        HugeLevel& hl = lev[lv];
        for (int k = h.lev_ptr[lv]; k < h.lev_ptr[lv + 1]; ++k) {
            const int t = h.order[k];
            const int ma = h.npiv[t] + h.nupd[t] + 1;
            if (ma <= kBigTiles * 16) continue;
            const long long MT = (ma + 15) / 16;
            hl.t.push_back(t);
            hl.off.push_back(hl.tiles);
            hl.tiles += MT * (MT + 1) / 2;
            rows = std::max(rows, ma);
        }
        if (hl.t.empty()) continue;
        max_tiles = std::max(max_tiles, hl.tiles);
        for (int t : hl.t) {
            hl.maxKt = std::max(hl.maxKt, (h.npiv[t] + 15) / 16);
            hl.mmax = std::max(hl.mmax, h.npiv[t] + h.nupd[t]);
        }
        hl.rb.assign(hl.maxKt, 1);
        hl.tb.assign(hl.maxKt, 0);
        long long work = 1;
        for (std::size_t f = 0; f < hl.t.size(); ++f) {
            const int t = hl.t[f];
            const int np = h.npiv[t], ma = np + h.nupd[t] + 1, MT = (ma + 15) / 16;
            for (int Kt = 0; Kt * 16 < np; ++Kt) {
                hl.rb[Kt] = std::max(hl.rb[Kt], std::max(1, (ma - 16 * Kt - 16 + 191) / 192));
                const long long Tn = MT - Kt - 1;
                hl.tb[Kt] = std::max<long long>(hl.tb[Kt], (Tn * (Tn + 1) / 2 + 3) / 4);
            }
            work = std::max<long long>(work, std::max<long long>(static_cast<long long>(np) * (ma - 1),
                                                                 static_cast<long long>(ma - np) * (ma - np)));
            for (int ci = h.cb[t], r = 0; ci < h.ce[t]; ++ci) {
                const int c = h.kids[ci];
                const long long u = h.nupd[c];
                if (u == 0) continue;
                if (static_cast<int>(hl.chf.size()) <= r) hl.chf.emplace_back(), hl.chc.emplace_back(),
                                                          hl.cbk.emplace_back(1, 0), hl.chmax.push_back(0);
                hl.chf[r].push_back(static_cast<int>(f));
                hl.chc[r].push_back(c);
                hl.cbk[r].push_back(hl.cbk[r].back() + ((u + 1) * u + 255) / 256);
                hl.chmax[r] = std::max(hl.chmax[r], static_cast<int>(u));
                ++r;
            }
            const long long mf = h.npiv[t] + h.nupd[t];
            hl.woff.push_back(hl.msum);
            hl.cpre.push_back(hl.csum);
            hl.msum += mf;
            hl.csum += (mf + 255) / 256;
            hl.wkt = std::max(hl.wkt, (h.npiv[t] + wide_panel - 1) / wide_panel);
        }
        hl.dwoff = to_device(hl.woff, s);
        hl.dcpre = to_device(hl.cpre, s);
        hl.store_blocks = std::min<long long>(1024, (work + 255) / 256);
        hl.dt = to_device(hl.t, s);
        hl.doff = to_device(hl.off, s);
        for (std::size_t r = 0; r < hl.chf.size(); ++r) {
            hl.dchf.push_back(to_device(hl.chf[r], s));
            hl.dchc.push_back(to_device(hl.chc[r], s));
            hl.dcb.push_back(to_device(hl.cbk[r], s));
        }
    }
    return max_tiles;
}

}  // namespace lsc
