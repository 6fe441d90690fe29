// Device kernels of the batched window engine (gfx950).
//
// One window = one circuit of the reference's solve_window
// (src/cpp/circuit/omniscape.cpp:72-122): the cells of a W x W box within
// radius D of the centre, grounded at the centre, NoData cells out of the
// circuit (identity rows, system.cpp:466-469), only the ground's connected
// component solved. Every window of a run shares one symbolic plan, so a
// batch of windows is a uniform batch of sparse Cholesky factorisations.
//
// Pipeline per batch (one kernel each, so rocprof times each phase):
//   k_fill      conductance and scaled sources of every window from the maps
//   k_flood     the ground's component (row bit masks), Laplacian diagonal,
//               Jacobi scale s = diag^-1/2, right-hand side
//   k_factor    multifrontal LL^T of S A S (T = float: MFMA f32, or double),
//               forward substitution fused (RHS row of every front)
//   k_solve     level-by-level triangular solves, L staged through LDS
//   k_residual  r = b - A x in FP64 from the conductances (never from L)
//   k_currents  node currents (the reference's node_currents, grid.cpp:19-71),
//               times the block-artifact correction
//   k_accumulate  cumulative map, every cell summing its windows in a fixed
//               order (deterministic: no atomics)
#pragma once

#include <hip/hip_runtime.h>

#include "front.hpp"
#include "window.hpp"

namespace lsc {

namespace {

__constant__ int c_dy8[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int c_dx8[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__constant__ int c_dg8[8] = {1, 0, 1, 0, 0, 1, 0, 1};
__constant__ int c_dy4[4] = {0, -1, 1, 0};
__constant__ int c_dx4[4] = {-1, 0, 0, 1};

__device__ __forceinline__ int ndy(int nd, int d) { return nd == 8 ? c_dy8[d] : c_dy4[d]; }
__device__ __forceinline__ int ndx(int nd, int d) { return nd == 8 ? c_dx8[d] : c_dx4[d]; }
__device__ __forceinline__ bool ndiag(int nd, int d) { return nd == 8 && c_dg8[d]; }

// Circuitscape's edge weight (reference grid.hpp:38-40).
__device__ __forceinline__ double edge_weight(double a, double b, bool diag) {
    return diag ? (a + b) / 2.8284271247461903 : (a + b) / 2;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = wave_id();
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < kWaves; ++w) t += red[w];
    return t;
}

template <int NW = kWaves>
__device__ __forceinline__ double block_max(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wave = wave_id();
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = red[0];
    for (int w = 1; w < NW; ++w) t = fmax(t, red[w]);
    return t;
}

}  // namespace

__device__ __forceinline__ void target_of(const Lattice& L, long long k, long long& tx, long long& ty) {
    tx = L.br + (k / L.nly) * L.bs;
    ty = L.br + (k % L.nly) * L.bs;
}

// Conductance within D (NoData -9999) and sources within R of every window,
// the target block zeroed at D from the clipped window's corner and the
// block's strength shared out (reference omniscape.cpp:406-421, 495-525).
static __global__ __launch_bounds__(kThreads) void k_fill(const double* __restrict__ res, const double* __restrict__ source,
                                                   Lattice Lt, long long k0, int B, CondArgs cond,
                                                   double* __restrict__ g, double* __restrict__ src,
                                                   int* __restrict__ active,
                                                   unsigned long long* __restrict__ n_targets) {
    __shared__ double red[kWaves];
    __shared__ double s_amps;
    const int k = blockIdx.x;
    if (k >= B) return;
    const long long kk = k0 + k;
    const int W = Lt.W, D = Lt.D, W2 = W * W;
    double* gw = g + static_cast<long long>(k) * W2;
    double* sw = src + static_cast<long long>(k) * W2;
    if (kk >= Lt.nlat) {
        if (threadIdx.x == 0) active[k] = 0;
        return;
    }
    long long tx, ty;
    target_of(Lt, kk, tx, ty);
    if (threadIdx.x == 0) {
        double amps = 0.0;  // column-major block sum, NaN skipped (omniscape.cpp:410-415)
        for (long long xx = tx - Lt.br; xx <= tx + Lt.br; ++xx)
            for (long long yy = ty - Lt.br; yy <= ty + Lt.br; ++yy) {
                const double v = source[yy * Lt.nx + xx];
                if (!isnan(v)) amps += v;
            }
        s_amps = amps;
    }
    __syncthreads();
    const double amps = s_amps;
    const long long oy = ty - D, ox = tx - D;
    const long long y0c = oy > 0 ? oy : 0, x0c = ox > 0 ? ox : 0;
    const long long zy1 = (D + Lt.br < Lt.ny - 1) ? D + Lt.br : Lt.ny - 1;
    const long long zx1 = (D + Lt.br < Lt.nx - 1) ? D + Lt.br : Lt.nx - 1;
    double part = 0.0;
    for (int c = threadIdx.x; c < W2; c += kThreads) {
        const int y = c / W, x = c - (c / W) * W;
        const long long Y = oy + y, X = ox + x;
        const bool on = Y >= 0 && Y < Lt.ny && X >= 0 && X < Lt.nx;
        const long long d2 = static_cast<long long>(y - D) * (y - D) + static_cast<long long>(x - D) * (x - D);
        double gv = -9999.0, sv = 0.0;
        if (on) {
            const double r = res[Y * Lt.nx + X];
            if (d2 <= static_cast<long long>(D) * D && !isnan(r)) gv = Lt.rc ? r : 1 / r;
            if (d2 <= static_cast<long long>(Lt.R) * Lt.R) {
                sv = source[Y * Lt.nx + X];
                if (isnan(sv)) sv = 0.0;
            }
            const long long yc = Y - y0c, xc = X - x0c;  // the clipped window's frame
            if (yc >= D - Lt.br && yc <= zy1 && xc >= D - Lt.br && xc <= zx1) sv = 0.0;
        }
        gw[c] = gv;
        sw[c] = sv;
        if (sv > 0) part += sv;
    }
    const double sum = block_sum(part, red);
    const bool target = amps > 0;
    if (target) {
        for (int c = threadIdx.x; c < W2; c += kThreads) {
            const double v = sw[c];
            if (v > 0) sw[c] = (v * amps) / sum;
        }
        // conditional connectivity (reference omniscape.cpp:217-232, 526-530):
        // sources whose present condition does not match the block's value
        for (int ci = 0; ci < cond.n; ++ci) {
            const double value = cond.value[ci][kk];
            for (int c = threadIdx.x; c < W2; c += kThreads) {
                const int y = c / W, x = c - (c / W) * W;
                const long long Y = oy + y, X = ox + x;
                if (Y < 0 || Y >= Lt.ny || X < 0 || X >= Lt.nx) continue;
                const double v = cond.present[ci][Y * Lt.nx + X];
                if (isnan(v)) continue;
                const bool keep = cond.equal[ci] ? v == value : (v - value <= cond.upper[ci] && v - value >= cond.lower[ci]);
                if (!keep) sw[c] = 0.0;
            }
        }
    }
    if (threadIdx.x == 0) {
        active[k] = target ? 1 : 0;
        if (target) atomicAdd(n_targets, 1ull);
    }
}

// The summed edge weights from cell c (row y, column xx) to its present
// neighbours, in the neighbour order; every neighbour's conductance is
// requested before any is used (outside the window the cell's own, unused).
template <int ND, typename Bit>
__device__ __forceinline__ double present_edge_sum(int W, int c, int y, int xx, const double* gw, const Bit& bit) {
    bool on[ND];
    double gn[ND];
#pragma unroll
    for (int e = 0; e < ND; ++e) {
        const int yy = y + ndy(ND, e), x2 = xx + ndx(ND, e);
        const bool in = yy >= 0 && yy < W && x2 >= 0 && x2 < W;
        on[e] = in && bit(in ? yy : y, in ? x2 : xx);
        gn[e] = gw[in ? yy * W + x2 : c];
    }
    const double gc = gw[c];
    double sum = 0.0;
#pragma unroll
    for (int e = 0; e < ND; ++e)
        if (on[e]) sum += edge_weight(gc, gn[e], ndiag(ND, e));
    return sum;
}

// After the flood: the present mask of every cell from the row masks (`words`
// 64-bit words a row), the window dropped when its component holds no source
// (omniscape.cpp:101), then the Laplacian diagonal of every unknown (edges to
// present cells in column-major neighbour order + the 10 eps shift,
// reference system.cpp:471-485), its Jacobi scale and the right-hand side.
__device__ void flood_finish(const DPlan& P, int k, const unsigned long long* pm, int words, const double* gw,
                             const double* sw, int* active, unsigned char* present, double* dg, double* sc,
                             double* bvec, double* rs, double* x, double shift, double* red) {
    const int W = P.W, W2 = W * W, D = P.D, n = P.n, nd = P.nd;
    const int ground = D * W + D;
    auto bit = [&](int y, int xx) { return ((pm[y * words + (xx >> 6)] >> (xx & 63)) & 1ull) != 0; };
    unsigned char* pw = present + static_cast<long long>(k) * W2;
    double part = 0.0;
    for (int c = threadIdx.x; c < W2; c += kThreads) {
        const int y = c / W, xx = c - (c / W) * W;
        const bool p = bit(y, xx);
        pw[c] = p;
        if (p && c != ground) part += sw[c];
    }
    const double total = block_sum(part, red);
    if (total == 0) {
        if (threadIdx.x == 0) active[k] = 0;
        return;
    }
    const long long off = static_cast<long long>(k) * n;
    for (int p = threadIdx.x; p < n; p += kThreads) {
        const int c = P.cell[p];
        const int y = c / W, xx = c - (c / W) * W;
        double d = 1.0, s = 1.0, b = 0.0;
        if (bit(y, xx)) {
            const double sum = nd == 8 ? present_edge_sum<8>(W, c, y, xx, gw, bit) : present_edge_sum<4>(W, c, y, xx, gw, bit);
            d = sum + shift;
            s = 1.0 / sqrt(d);
            b = sw[c];
        }
        dg[off + p] = d;
        sc[off + p] = s;
        bvec[off + p] = b;
        rs[off + p] = s * b;
        x[off + p] = 0.0;
    }
}

// The ground's component (cells with g > 0 joined to the centre), by row bit
// masks: runs are filled along rows with one 128-bit add, then grown between
// rows, until nothing changes. Then the Laplacian diagonal of every unknown
// (edges to present cells in column-major neighbour order + the 10 eps
// shift, reference system.cpp:471-485), its Jacobi scale and the RHS.
static __global__ __launch_bounds__(kThreads) void k_flood(DPlan P, int B, const double* __restrict__ g,
                                                    const double* __restrict__ src, int* __restrict__ active,
                                                    unsigned char* __restrict__ present, double* __restrict__ dg,
                                                    double* __restrict__ sc, double* __restrict__ bvec,
                                                    double* __restrict__ rs, double* __restrict__ x, double shift) {
    __shared__ unsigned long long om[kMaxW + 1][2], pm[kMaxW + 1][2];
    __shared__ double red[kWaves];
    __shared__ int s_ok;
    const int k = blockIdx.x;
    if (k >= B || !active[k]) return;
    const int W = P.W, W2 = W * W, D = P.D, nd = P.nd;
    const double* gw = g + static_cast<long long>(k) * W2;
    const double* sw = src + static_cast<long long>(k) * W2;
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int ground = D * W + D;
    if (threadIdx.x == 0) s_ok = gw[ground] > 0;
    __syncthreads();
    if (!s_ok) {
        if (threadIdx.x == 0) active[k] = 0;
        return;
    }
    // open masks
    for (int y = wave; y < W; y += kWaves) {
        for (int h = 0; h < 2; ++h) {
            const int xx = h * 64 + lane;
            const bool open = xx < W && gw[y * W + xx] > 0;
            const unsigned long long m = __ballot(open);
            if (lane == 0) om[y][h] = m;
        }
    }
    __syncthreads();
    if (wave == 0) {
        unsigned long long O[2][2], Pm[2][2];
        for (int r = 0; r < 2; ++r) {
            const int y = lane + 64 * r;
            for (int h = 0; h < 2; ++h) {
                O[r][h] = y < W ? om[y][h] : 0ull;
                Pm[r][h] = 0ull;
            }
            if (y == D) Pm[r][D >> 6] = 1ull << (D & 63);
        }
        for (int it = 0; it < W2 + 2; ++it) {
            // along rows, both directions
            for (int r = 0; r < 2; ++r) {
                const unsigned long long o0 = O[r][0], o1 = O[r][1];
                unsigned long long s0 = Pm[r][0], s1 = Pm[r][1];
                unsigned long long a0 = o0 + s0;
                unsigned long long a1 = o1 + s1 + (a0 < o0 ? 1ull : 0ull);
                const unsigned long long f0 = ((a0 ^ o0) & o0) | s0, f1 = ((a1 ^ o1) & o1) | s1;
                // reversed rows for the other direction
                const unsigned long long ro0 = __builtin_bitreverse64(o1), ro1 = __builtin_bitreverse64(o0);
                const unsigned long long rs0 = __builtin_bitreverse64(s1), rs1 = __builtin_bitreverse64(s0);
                const unsigned long long b0 = ro0 + rs0;
                const unsigned long long b1 = ro1 + rs1 + (b0 < ro0 ? 1ull : 0ull);
                const unsigned long long e0 = ((b0 ^ ro0) & ro0) | rs0, e1 = ((b1 ^ ro1) & ro1) | rs1;
                Pm[r][0] = f0 | __builtin_bitreverse64(e1);
                Pm[r][1] = f1 | __builtin_bitreverse64(e0);
            }
            for (int r = 0; r < 2; ++r) {
                const int y = lane + 64 * r;
                if (y < W) pm[y][0] = Pm[r][0], pm[y][1] = Pm[r][1];
            }
            wave_sync();
            bool changed = false;
            for (int r = 0; r < 2; ++r) {
                const int y = lane + 64 * r;
                if (y >= W) continue;
                unsigned long long n0 = 0, n1 = 0;
                if (y > 0) n0 |= pm[y - 1][0], n1 |= pm[y - 1][1];
                if (y + 1 < W) n0 |= pm[y + 1][0], n1 |= pm[y + 1][1];
                if (nd == 8) {
                    const unsigned long long l0 = n0 << 1, l1 = (n1 << 1) | (n0 >> 63);
                    const unsigned long long r0 = (n0 >> 1) | (n1 << 63), r1 = n1 >> 1;
                    n0 |= l0 | r0;
                    n1 |= l1 | r1;
                }
                const unsigned long long p0 = Pm[r][0] | (n0 & O[r][0]), p1 = Pm[r][1] | (n1 & O[r][1]);
                changed = changed || p0 != Pm[r][0] || p1 != Pm[r][1];
                Pm[r][0] = p0, Pm[r][1] = p1;
            }
            wave_sync();
            if (!__any(changed)) break;
        }
        for (int r = 0; r < 2; ++r) {
            const int y = lane + 64 * r;
            if (y < W) pm[y][0] = Pm[r][0], pm[y][1] = Pm[r][1];
        }
    }
    __syncthreads();
    flood_finish(P, k, &pm[0][0], 2, gw, sw, active, present, dg, sc, bvec, rs, x, shift, red);
}

// Wider windows (more than 127 cells a side, up to 64 * kWideWords): the
// same component by row bit masks of `words` 64-bit words in LDS (open,
// present and the next present mask, W rows each), rows spread over the
// workgroup's threads: each pass fills along every row (a carry add over the
// row's words, both directions), then grows every row by its neighbours',
// until no row changes.
static __global__ __launch_bounds__(kThreads) void k_flood_wide(DPlan P, int B, int words,
                                                         const double* __restrict__ g,
                                                         const double* __restrict__ src, int* __restrict__ active,
                                                         unsigned char* __restrict__ present, double* __restrict__ dg,
                                                         double* __restrict__ sc, double* __restrict__ bvec,
                                                         double* __restrict__ rs, double* __restrict__ x, double shift) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[kWaves];
    const int k = blockIdx.x;
    if (k >= B || !active[k]) return;
    const int W = P.W, D = P.D, nd = P.nd;
    unsigned long long* om = reinterpret_cast<unsigned long long*>(smem_raw);
    unsigned long long* pm = om + static_cast<long long>(W) * words;
    unsigned long long* pn = pm + static_cast<long long>(W) * words;
    const double* gw = g + static_cast<long long>(k) * W * W;
    const double* sw = src + static_cast<long long>(k) * W * W;
    if (!(gw[D * W + D] > 0)) {
        if (threadIdx.x == 0) active[k] = 0;
        return;
    }
    for (int e = threadIdx.x; e < W * words; e += kThreads) {
        const int y = e / words, w = e - y * words;
        unsigned long long m = 0;
        for (int b = 0; b < 64; ++b) {
            const int xx = w * 64 + b;
            if (xx < W && gw[y * W + xx] > 0) m |= 1ull << b;
        }
        om[e] = m;
        pm[e] = (y == D && w == (D >> 6)) ? 1ull << (D & 63) : 0ull;
    }
    __syncthreads();
    for (;;) {
        // along the rows: seeds spread over the open runs they touch, rightwards by a carry add, leftwards
        // by the same add on the row reversed
        for (int y = threadIdx.x; y < W; y += kThreads) {
            unsigned long long* p = pm + static_cast<long long>(y) * words;
            const unsigned long long* o = om + static_cast<long long>(y) * words;
            unsigned long long carry = 0;
            for (int w = 0; w < words; ++w) {
                const unsigned long long a = o[w] + p[w];
                const unsigned long long a2 = a + carry;
                carry = (a < o[w] || a2 < a) ? 1ull : 0ull;
                p[w] |= (a2 ^ o[w]) & o[w];
            }
            carry = 0;
            for (int r = 0; r < words; ++r) {  // word r of the reversed row is word words-1-r, bits reversed
                const int w = words - 1 - r;
                const unsigned long long ro = __builtin_bitreverse64(o[w]), rp = __builtin_bitreverse64(p[w]);
                const unsigned long long a = ro + rp;
                const unsigned long long a2 = a + carry;
                carry = (a < ro || a2 < a) ? 1ull : 0ull;
                p[w] |= __builtin_bitreverse64((a2 ^ ro) & ro);
            }
        }
        __syncthreads();
        // across the rows: a cell joins when a neighbour in the row above or below is in
        int grew = 0;
        for (int e = threadIdx.x; e < W * words; e += kThreads) {
            const int y = e / words, w = e - y * words;
            unsigned long long n0 = 0, lo = 0, hi = 0;
            for (int dy = -1; dy <= 1; dy += 2) {
                const int yy = y + dy;
                if (yy < 0 || yy >= W) continue;
                const unsigned long long* q = pm + static_cast<long long>(yy) * words;
                n0 |= q[w];
                if (nd == 8) {
                    lo |= (q[w] << 1) | (w > 0 ? q[w - 1] >> 63 : 0ull);
                    hi |= (q[w] >> 1) | (w + 1 < words ? q[w + 1] << 63 : 0ull);
                }
            }
            const unsigned long long next = pm[e] | ((n0 | lo | hi) & om[e]);
            grew |= next != pm[e];
            pn[e] = next;
        }
        grew = __syncthreads_or(grew);
        for (int e = threadIdx.x; e < W * words; e += kThreads) pm[e] = pn[e];
        __syncthreads();
        if (!grew) break;
    }
    flood_finish(P, k, pm, words, gw, sw, active, present, dg, sc, bvec, rs, x, shift, red);
}

// The off-diagonal entries of S A S (S = diag^-1/2) of every window, in the
// plan's front order: the "Laplacian assembly" of the reference's
// CellSystem::assemble (system.cpp:458-487) with the Jacobi scaling folded in.
// Entries touching a cell outside the circuit are 0 (identity rows).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_assemble(DPlan P, int B, const double* __restrict__ g,
                                                       const unsigned char* __restrict__ present,
                                                       const double* __restrict__ sc, const int* __restrict__ active,
                                                       T* __restrict__ aval) {
    const int k = blockIdx.y;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (k >= B || e >= P.nent || !active[k]) return;
    const int W = P.W, W2 = W * W, nd = P.nd;
    const int p = P.fe_p[e], d = static_cast<int>(P.fe_src[e] - static_cast<long long>(p) * nd);
    const int c = P.cell[p];
    const int c2 = c + ndy(nd, d) * W + ndx(nd, d);  // an entry's neighbour lies inside the window
    const unsigned char* pw = present + static_cast<long long>(k) * W2;
    double v = 0.0;
    if (pw[c] && pw[c2]) {
        const double* gw = g + static_cast<long long>(k) * W2;
        const double* sw = sc + static_cast<long long>(k) * P.n;
        v = -edge_weight(gw[c], gw[c2], ndiag(nd, d)) * sw[p] * sw[P.fe_q[e]];
    }
    aval[static_cast<long long>(k) * P.nent + e] = static_cast<T>(v);
}

namespace {

__device__ __forceinline__ long long colstart(int j, int m) {
    return static_cast<long long>(j) * m - ((static_cast<long long>(j) * (j - 1)) >> 1);
}

// Column of entry p of a lower triangle of order m packed by columns: the
// largest j with colstart(j, m) <= p.
__device__ __forceinline__ int packed_col(long long p, int m) {
    const double b = 2.0 * m + 1.0;
    int j = static_cast<int>((b - sqrt(b * b - 8.0 * static_cast<double>(p))) * 0.5);
    if (j < 0) j = 0;
    while (j > 0 && colstart(j, m) > p) --j;
    while (j + 1 < m && colstart(j + 1, m) <= p) ++j;
    return j;
}

// Element r of a 16-byte chunk of T values.
template <typename T>
__device__ __forceinline__ T chunk_elem(const uint4& v, int r);
template <>
__device__ __forceinline__ float chunk_elem<float>(const uint4& v, int r) {
    return __builtin_bit_cast(float, r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w);
}
template <>
__device__ __forceinline__ double chunk_elem<double>(const uint4& v, int r) {
    const unsigned long long lo = r == 0 ? v.x : v.z, hi = r == 0 ? v.y : v.w;
    return __builtin_bit_cast(double, lo | (hi << 32));
}

// One front: assembled (unit diagonal, the scaled entries, RHS row, the
// children's update matrices), factorised, written out (L packed by columns,
// y, the update matrix). Team of nw wavefronts (tid / nt over the team).
// Columns 0..nc-1 of the lower triangle of the tiled front's trailing block
// F[o.., o..] of `rows` rows, packed column after column (column j holds
// rows j..rows-1) from `out`: entry e of the packed array is taken by thread
// e mod nt, four at a time so that the LDS reads overlap.
template <typename T>
__device__ __forceinline__ void store_packed(const T* F, int o, int rows, int nc, T* out, int tid, int nt) {
    const int sz = nc * rows - ((nc * (nc - 1)) >> 1);
    int j = 0, r = tid;  // entry e = column j, row j + r
    while (j < nc && r >= rows - j) r -= rows - j, ++j;
    for (int e0 = tid; e0 < sz; e0 += 4 * nt) {
        T v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool in = j < nc;
            v[q] = F[in ? tix(o + j + r, o + j) : 0];
            r += nt;
            while (j < nc && r >= rows - j) r -= rows - j, ++j;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = e0 + q * nt;
            if (e < sz) out[e] = v[q];
        }
    }
}

// A front record, which is the same for the whole wavefront: read with
// vector loads (the plan is not provably unwritten, so no scalar loads) and
// moved to scalar registers, so that everything derived from it stays
// wave-uniform (scalar branches, no vector registers held).
__device__ __forceinline__ FrontDesc load_desc(const FrontDesc* p) {
    struct Raw {
        int v[sizeof(FrontDesc) / 4];
    } r;
    const int4* q = reinterpret_cast<const int4*>(p);
    int4 w[sizeof(FrontDesc) / 16];
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(FrontDesc) / 16); ++i) w[i] = q[i];
#pragma unroll
    for (int i = 0; i < static_cast<int>(sizeof(FrontDesc) / 16); ++i) {
        r.v[4 * i] = __builtin_amdgcn_readfirstlane(w[i].x);
        r.v[4 * i + 1] = __builtin_amdgcn_readfirstlane(w[i].y);
        r.v[4 * i + 2] = __builtin_amdgcn_readfirstlane(w[i].z);
        r.v[4 * i + 3] = __builtin_amdgcn_readfirstlane(w[i].w);
    }
    return __builtin_bit_cast(FrontDesc, r);
}

// The child update matrix Uc (u rows, packed by columns, 16-byte aligned)
// added into the front at the rows rel (LDS relbuf), streamed in 16-byte
// chunks, eight in flight per lane, every lane loading (clamped) so the
// chunks stay in registers.
template <typename T, typename Sync>
__device__ __forceinline__ void extend_add(T* F, int m, int u, const int* rel, const T* Uc, int* relbuf, int tid,
                                           int nt, Sync sync) {
    const int u1 = u + 1;
    for (int i = tid; i < u; i += nt) relbuf[i] = rel[i];
    if (tid == 0) relbuf[u] = m;
    sync();
    constexpr int CH = 16 / static_cast<int>(sizeof(T));
    const int tot = static_cast<int>(colstart(u, u1));  // columns 0..u-1 of the packed triangle
    const int nch = (tot + CH - 1) / CH;
    const uint4* Uv = reinterpret_cast<const uint4*>(Uc);
    for (int base = tid; base < nch; base += 8 * nt) {
        uint4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = base + q * nt;
            v[q] = Uv[c < nch ? c : nch - 1];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = base + q * nt;
            if (c < nch) {
                const int e0 = c * CH;
                int jj = packed_col(e0, u1);
                int ii = jj + static_cast<int>(e0 - colstart(jj, u1));
#pragma unroll
                for (int r = 0; r < CH; ++r) {
                    if (e0 + r < tot) F[tix(relbuf[ii], relbuf[jj])] += chunk_elem<T>(v[q], r);
                    if (++ii == u1) ++jj, ii = jj;
                }
            }
        }
    }
    sync();
}

// out[e] = F[idx[e]] for the sz entries and the padding up to a multiple of
// 4 (`out` and `idx` are 4-entry aligned; a padding index reads some entry
// of the front, a finite value that no reader of L or of the update matrix
// uses): a thread takes four consecutive entries (one 8-byte index load, one
// vector store), four such chunks in flight, the team walking the chunks in
// order (coalesced). A thread past the last chunk stores the last chunk
// again, the same values to the same place.
template <typename T>
__device__ __forceinline__ void store_by_table(const T* F, const unsigned short* idx, int sz, T* out, int tid,
                                               int nt) {
    typedef T V4 __attribute__((ext_vector_type(4)));
    typedef unsigned short I4 __attribute__((ext_vector_type(4)));
    const int nch = (sz + 3) >> 2;
    const I4* ix = reinterpret_cast<const I4*>(idx);
    V4* o = reinterpret_cast<V4*>(out);
    for (int c0 = tid; c0 < nch; c0 += 4 * nt) {
        int cc[4];
        I4 id[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = c0 + q * nt;
            cc[q] = c < nch ? c : nch - 1;
            id[q] = ix[cc[q]];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            V4 v;
            v.x = F[id[q].x];
            v.y = F[id[q].y];
            v.z = F[id[q].z];
            v.w = F[id[q].w];
            o[cc[q]] = v;
        }
    }
}

// F[idx[e]] += Uc[e] for the tot entries of a child's update matrix, padded
// to whole 4-entry chunks: the entries land on distinct front entries (the
// right-hand-side corner and the padding all on the front's unused corner),
// so the team adds without conflicts and without per-entry tests; a thread
// takes four consecutive entries (one vector load of values, one of
// offsets), eight such chunks in flight.
template <typename T, typename Sync>
__device__ __forceinline__ void extend_by_table(T* F, const unsigned short* idx, const T* Uc, int tot, int tid,
                                                int nt, Sync sync) {
    typedef T V4 __attribute__((ext_vector_type(4)));
    typedef unsigned short I4 __attribute__((ext_vector_type(4)));
    const int nch = (tot + 3) >> 2;
    const I4* ix = reinterpret_cast<const I4*>(idx);
    const V4* uv = reinterpret_cast<const V4*>(Uc);
    for (int c0 = tid; c0 < nch; c0 += 8 * nt) {
        I4 id[8];
        V4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = c0 + q * nt, cc = c < nch ? c : nch - 1;
            id[q] = ix[cc];
            v[q] = uv[cc];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (c0 + q * nt < nch) {
                F[id[q].x] += v[q].x;
                F[id[q].y] += v[q].y;
                F[id[q].z] += v[q].z;
                F[id[q].w] += v[q].w;
            }
        }
    }
    sync();
}

template <typename T, typename Sync>
__device__ void process_front(const DPlan& P, const FrontDesc& D, T* F, int* relbuf, T* bc, int tid, int nt,
                              int wave, int nw, int lane, int* flag, Sync sync, const T* aval, const double* rsw,
                              T* Lw, double* yw, T* U) {
    const int np = D.np, nu = D.nu, m = np + nu, ma = m + 1, first = D.first;
    const int MT = (ma + 15) >> 4;
    const int nz = ((MT * (MT + 1)) >> 1) * 256;
    // the first entries and the pivots' RHS requested before the tiles are
    // cleared (clamped indices: every lane loads, the values stay in registers)
    const int e0 = D.fe0, e1 = D.fe1;
    const int elast = e1 > e0 ? e1 - 1 : 0;  // entry arrays hold at least one element
    const bool tab = P.eidx != nullptr;
    int eo[4];  // the entries' LDS offsets
    T ev[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = e0 + tid + q * nt;
        const int eq = e < e1 ? e : elast;
        eo[q] = tab ? P.eidx[eq] : tix(P.fe_row[eq], P.fe_col[eq]);
        ev[q] = aval[eq];
    }
    const double rhs0 = np > 0 ? rsw[first + (tid < np ? tid : np - 1)] : 0.0;
    {  // the tiles cleared 16 bytes at a time (nz is a multiple of 256), in
       // rounds counted by a wave-uniform bound
        uint4* Fz = reinterpret_cast<uint4*>(F);
        const int nz16 = nz * static_cast<int>(sizeof(T)) / 16;
        const int rounds = nz16 / nt;
        for (int r = 0; r < rounds; ++r) Fz[tid + r * nt] = uint4{0u, 0u, 0u, 0u};
        if (tid + rounds * nt < nz16) Fz[tid + rounds * nt] = uint4{0u, 0u, 0u, 0u};
    }
    sync();
    for (int j = tid; j < np; j += nt) {
        F[tix(j, j)] = T(1);
        F[tix(m, j)] = static_cast<T>(j == tid ? rhs0 : rsw[first + j]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (e0 + tid + q * nt < e1) F[eo[q]] = ev[q];
    for (int e = e0 + tid + 4 * nt; e < e1; e += nt) F[tab ? P.eidx[e] : tix(P.fe_row[e], P.fe_col[e])] = aval[e];
    sync();
    // the children's update matrices. A record's (at most two) children are
    // told apart by arithmetic on the child number: a select between the two
    // fields would be turned into an indexed read of the record, which puts
    // the record in scratch memory.
    const int nk = D.nk >= 0 ? D.nk : D.ce - D.cb;
    for (int ci = 0; ci < nk; ++ci) {
        int u, upd;
        long long up;
        if (D.nk >= 0) {
            u = D.ku[0] + ci * (D.ku[1] - D.ku[0]);
            upd = D.kupd[0] + ci * (D.kupd[1] - D.kupd[0]);
            up = D.kuposA[0] + ci * (D.kuposA[1] - D.kuposA[0]);
        } else {
            const int c = P.kids[D.cb + ci];
            u = P.nupd[c], upd = P.upd[c], up = P.uposA[c];
        }
        if (P.xidx) {
            if (u > 0) extend_by_table<T>(F, P.xidx + up, U + up, (u + 1) * (u + 2) / 2, tid, nt, sync);
        } else {
            extend_add<T>(F, m, u, P.relmap + upd, U + up, relbuf, tid, nt, sync);
        }
    }
    factor_front<T>(F, ma, np, wave, nw, lane, flag, bc, sync);
    // L block and update matrix, both packed by columns: the team walks the
    // packed entries in order (coalesced stores), four LDS reads in flight
    // per thread
    if (P.lidx) {
        store_by_table<T>(F, P.lidx + D.lpos, static_cast<int>(colstart(np, m)), Lw + D.lpos, tid, nt);
        store_by_table<T>(F, P.uidx + D.uposA, (nu + 1) * (nu + 2) / 2, U + D.uposA, tid, nt);
    } else {
        store_packed<T>(F, 0, m, np, Lw + D.lpos, tid, nt);
        store_packed<T>(F, np, nu + 1, nu + 1, U + D.uposA, tid, nt);
    }
    for (int j = tid; j < np; j += nt) yw[first + j] = static_cast<double>(F[tix(m, j)]);
    sync();
}

}  // namespace

// Level-synchronous batched factorisation: one launch per tree level and
// front class, every (instance, front) pair of the level an independent task,
// so thousands of fronts are in flight and their memory latencies overlap.
// Small fronts (<= 64 rows with the RHS row): one wavefront each, its own LDS
// region of `stile` tiles. Large fronts: one workgroup each, up to 55 tiles.
template <typename T>
__global__ __launch_bounds__(kSmallWaves * 64) void k_factor_small(DPlan P, int base, int ns, int B, int stile,
                                                           const T* __restrict__ aval, const double* __restrict__ rs,
                                                           const int* __restrict__ active, T* __restrict__ L,
                                                           double* __restrict__ y, T* __restrict__ Uall,
                                                           int* __restrict__ fail) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_flag[kSmallWaves];
    __shared__ int s_rel[kSmallWaves][kSmallRows + 1];
    __shared__ __align__(16) T s_bc[kSmallWaves][64];
    const int lane = threadIdx.x & 63, wave = wave_id();
    const long long task = static_cast<long long>(blockIdx.x) * kSmallWaves + wave;
    const int k = static_cast<int>(task / ns);
    if (k >= B || !active[k]) return;
    const FrontDesc D = load_desc(P.desc + base + static_cast<int>(task - static_cast<long long>(k) * ns));
    T* F = reinterpret_cast<T*>(smem_raw) + wave * stile * 256;
    if (lane == 0) s_flag[wave] = 0;
    wave_sync();
    auto wsync = [] { wave_sync(); };
    process_front<T>(P, D, F, s_rel[wave], s_bc[wave], lane, 64, 0, 1, lane, &s_flag[wave], wsync,
                     aval + static_cast<long long>(k) * P.nent,
                     rs + static_cast<long long>(k) * P.n, L + static_cast<long long>(k) * P.lsize,
                     y + static_cast<long long>(k) * P.n, Uall + static_cast<long long>(k) * P.usizeA);
    if (lane == 0 && s_flag[wave]) atomicOr(&fail[k], 1);
}

// `base`: the level-order position of the level's first front of 65-159
// rows, `nb` their number. FP32 fronts are held to 80 registers (six
// wavefronts per SIMD, no spills), so that fronts whose tiles take at most
// 26 KB of LDS fit six workgroups per CU.
template <typename T>
__global__ __launch_bounds__(kBigWaves * 64, sizeof(T) == 4 ? 6 : 1) void k_factor_big(DPlan P, int base, int nb, int B,
                                                         const T* __restrict__ aval, const double* __restrict__ rs,
                                                         const int* __restrict__ active, T* __restrict__ L,
                                                         double* __restrict__ y, T* __restrict__ Uall,
                                                         int* __restrict__ fail) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_flag;
    __shared__ int s_rel[kBigTiles * 16 + 1];
    __shared__ __align__(16) T s_bc[kBigWaves][64];
    const int lane = threadIdx.x & 63, wave = wave_id();
    const int k = blockIdx.x / nb;
    if (k >= B || !active[k]) return;
    const FrontDesc D = load_desc(P.desc + base + (blockIdx.x - k * nb));
    if (threadIdx.x == 0) s_flag = 0;
    __syncthreads();
    auto bsync = [] { __syncthreads(); };
    process_front<T>(P, D, reinterpret_cast<T*>(smem_raw), s_rel, s_bc[wave], threadIdx.x, kBigWaves * 64, wave,
                     kBigWaves, lane, &s_flag, bsync,
                     aval + static_cast<long long>(k) * P.nent, rs + static_cast<long long>(k) * P.n,
                     L + static_cast<long long>(k) * P.lsize, y + static_cast<long long>(k) * P.n,
                     Uall + static_cast<long long>(k) * P.usizeA);
    if (threadIdx.x == 0 && s_flag) atomicOr(&fail[k], 1);
}

namespace {

// 1 / pivot for the substitutions: for a single-precision pivot, the
// hardware reciprocal refined by one FP64 Newton step (relative error near
// 1e-14, two FMAs instead of a double division); the exact quotient for an
// FP64 pivot.
__device__ __forceinline__ double inv_pivot(float d) {
    const double dd = static_cast<double>(d);
    const double r = static_cast<double>(__builtin_amdgcn_rcpf(d));
    return fma(r, fma(-dd, r, 1.0), r);
}
__device__ __forceinline__ double inv_pivot(double d) { return 1.0 / d; }

__device__ __forceinline__ double pick(const double (&v)[3], int slot) {
    return slot == 0 ? v[0] : (slot == 1 ? v[1] : v[2]);
}

// Columns j0..j0+KB-1 of front t's L block straight from HBM into
// registers, rows on lanes (slot r of lane l holds row l + 64 r; 0 above the
// diagonal and past the front): every lane loads from a valid address
// (clamped), so the values stay in registers.
template <typename T, int R, int KB>
__device__ __forceinline__ void load_cols(const T* Lt, int np, int m, int j0, int lane, T (&Lr)[R][KB]) {
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const int j = j0 + q;
        const bool okj = j < np;
        const int jj = okj ? j : 0;
        const T* col = Lt + colstart(jj, m) - jj;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = lane + 64 * r;
            const int i = row < jj ? jj : (row < m ? row : m - 1);
            const T v = col[i];
            Lr[r][q] = okj && row >= j && row < m ? v : T(0);
        }
    }
}

// v[slot] as selects (an indexed read would put v in private memory).
template <typename T, int R>
__device__ __forceinline__ T pick_slot(const T (&v)[R], int slot) {
    static_assert(R == 1 || R == 3, "one or three row slots");
    if constexpr (R == 1) return v[0];
    else return slot == 0 ? v[0] : (slot == 1 ? v[1] : v[2]);
}

// Forward substitution of a front of at most 64 R rows (one wavefront, rows
// on lanes; the reference's system.cpp:974-989): its vector assembled from
// the pivots' RHS and the children's update vectors, then y = L11^-1 f and
// the update vector f_upd - L21 y. L is read from HBM into registers KB
// columns at a time, the next KB requested before the current ones are
// eliminated; only the assembly uses LDS (m doubles).
template <typename T, int R, int KB>
__device__ void forward_front_direct(const DPlan& P, int t, const T* Lw, double* fv, int lane, const double* rsw,
                                     double* yw, double* V) {
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int cb = P.cb[t], ce = P.ce[t];
    const T* Lt = Lw + P.lpos[t];
    T Lr[R][KB];
    if (np > 0) load_cols<T, R, KB>(Lt, np, m, 0, lane, Lr);
    double pre[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        const double v = rsw[i < np ? first + i : 0];
        pre[r] = i < np ? v : 0.0;
    }
    double cv[2][R];
    int cr[2][R];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int ci = cb + k;
        const int c = ci < ce ? P.kids[ci] : 0;  // (front 0 when absent: no lane reads for it)
        const int u = ci < ce ? P.nupd[c] : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {  // loads from clamped indices: no per-lane branch
            const int ii = lane + 64 * r;
            const bool ok = ii < u;
            const int rv = P.relmap[ok ? P.upd[c] + ii : 0];
            const double vv = V[ok ? P.vpos[c] + ii : 0];
            cr[k][r] = ok ? rv : -1;
            cv[k][r] = ok ? vv : 0.0;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < m) fv[i] = pre[r];
    }
    wave_sync();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (cr[k][r] >= 0) fv[cr[k][r]] += cv[k][r];
        wave_sync();
    }
    for (int ci = cb + 2; ci < ce; ++ci) {
        const int c = P.kids[ci];
        const int u = P.nupd[c];
        for (int ii = lane; ii < u; ii += 64) fv[P.relmap[P.upd[c] + ii]] += V[P.vpos[c] + ii];
        wave_sync();
    }
    double f[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        f[r] = i < m ? fv[i] : 0.0;
    }
    for (int j0 = 0; j0 < np; j0 += KB) {
        T Ln[R][KB];
        const bool more = j0 + KB < np;
        if (more) load_cols<T, R, KB>(Lt, np, m, j0 + KB, lane, Ln);
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const int j = j0 + q;
            if (j < np) {
                const int slot = j >> 6, jl = j & 63;
                T dcol[R];
#pragma unroll
                for (int r = 0; r < R; ++r) dcol[r] = Lr[r][q];
                const double yj = bcast(pick_slot<double, R>(f, slot), jl) *
                                  inv_pivot(bcast(pick_slot<T, R>(dcol, slot), jl));
#pragma unroll
                for (int r = 0; r < R; ++r) f[r] -= static_cast<double>(Lr[r][q]) * yj;
                if (lane == jl) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (slot == r) f[r] = yj;
                }
            }
        }
        if (more) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int q = 0; q < KB; ++q) Lr[r][q] = Ln[r][q];
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < np) yw[first + i] = f[r];
        else if (i < m) V[P.vpos[t] + i - np] = f[r];
    }
}

// Back substitution of front t (one wavefront, L block Lc in LDS): z_piv =
// L11^-T (y - L21^T z_upd) (reference system.cpp:992-1016). The L block,
// the solved update rows and the pivots' y are requested together.
template <typename T, int R>
__device__ void backward_front(const DPlan& P, int t, const T* Lw, T* Lc, double* xu, int lane, const double* yw,
                               double* zw) {
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int* rows = P.rows + P.upd[t];
    const uint4* src = reinterpret_cast<const uint4*>(Lw + P.lpos[t]);
    const long long sz = static_cast<long long>(np) * m - ((static_cast<long long>(np) * (np - 1)) >> 1);
    const int nv = static_cast<int>((sz * static_cast<long long>(sizeof(T)) + 15) / 16);
    int ridx[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ii = lane + 64 * r;
        ridx[r] = ii < nu ? rows[ii] : -1;
    }
    // every lane loads (clamped index): a per-lane conditional load would
    // put lv in private memory. A front without pivots has no L block and
    // reads the start of the (non-empty) L array instead.
    constexpr int kPre = R == 1 ? 4 : 8;  // 16-byte requests per lane before the first use
    uint4 lv[kPre];
    const int last = nv > 0 ? nv - 1 : 0;
    const uint4* srcl = nv > 0 ? src : reinterpret_cast<const uint4*>(Lw);
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int e = lane + q * 64;
        lv[q] = srcl[e < nv ? e : last];
    }
    double yv[R], zu[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = lane + 64 * r;
        yv[r] = j < np ? yw[first + j] : 0.0;
        zu[r] = ridx[r] >= 0 ? zw[ridx[r]] : 0.0;
    }
    uint4* dst = reinterpret_cast<uint4*>(Lc);
#pragma unroll
    for (int q = 0; q < kPre; ++q) {
        const int e = lane + q * 64;
        if (e < nv) dst[e] = lv[q];
    }
#pragma unroll 8
    for (int e = lane + kPre * 64; e < nv; e += 64) dst[e] = src[e];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ii = lane + 64 * r;
        if (ii < nu) xu[ii] = zu[r];
    }
    wave_sync();
    double tv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = lane + 64 * r;
        double acc = 0.0;
        if (j < np) {
            const T* col = Lc + colstart(j, m) + (np - j);
            for (int ii = 0; ii < nu; ++ii) acc += static_cast<double>(col[ii]) * xu[ii];
            acc = yv[r] - acc;
        }
        tv[r] = acc;
    }
    // 8 pivots at a time from the last: L[j][i] of the lane's columns i and
    // the reciprocal pivots in registers, then a register-only chain.
    constexpr int kB = 8;  // pivots per register block
    for (int j1 = np - 1; j1 >= 0; j1 -= kB) {
        T Lr[R][kB];
        double dinv[kB];
#pragma unroll
        for (int q = 0; q < kB; ++q) {
            const int j = j1 - q;
            const bool okj = j >= 0;
            dinv[q] = okj ? inv_pivot(Lc[colstart(j, m)]) : 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                Lr[r][q] = (okj && i < j) ? Lc[colstart(i, m) + (j - i)] : T(0);
            }
        }
#pragma unroll
        for (int q = 0; q < kB; ++q) {
            const int j = j1 - q;
            if (j >= 0) {
                const int slot = j >> 6;
                const double xj = bcast(pick(tv, slot), j & 63) * dinv[q];
#pragma unroll
                for (int r = 0; r < R; ++r) tv[r] -= static_cast<double>(Lr[r][q]) * xj;
                if (lane == (j & 63)) {
                    if (slot == 0) tv[0] = xj;
                    else if (slot == 1) tv[1] = xj;
                    else tv[2] = xj;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < np) zw[first + i] = tv[r];
    }
}

}  // namespace

// Batched triangular solves, one launch per level: every (instance, front)
// pair of the level is the task of a one-wavefront workgroup. The forward
// sweep reads the front's L columns into registers (LDS holds the front's
// vector, 64 R doubles); the backward sweep stages the L block (contiguous
// in HBM) into LDS, then substitutes.
template <typename T, int R>
__global__ __launch_bounds__(64) void k_forward_level(DPlan P, int h, int B, long long lstride,
                                                      const int* __restrict__ active,
                                                      const int* __restrict__ done, const T* __restrict__ L,
                                                      const double* __restrict__ rs, double* __restrict__ y,
                                                      double* __restrict__ Vall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* fv = reinterpret_cast<double*>(smem_raw);
    const int lane = threadIdx.x;
    const int nf = P.lev_fit[h];  // the level's fronts of at most 159 rows (larger ones: huge.hpp)
    const int k = blockIdx.x / nf;
    if (k >= B || !active[k] || done[k]) return;
    const int t = P.order[P.lev_ptr[h] + (blockIdx.x - k * nf)];
    const long long off = static_cast<long long>(k) * P.n;
    forward_front_direct<T, R, R == 1 ? 16 : 4>(P, t, L + static_cast<long long>(k) * lstride, fv, lane, rs + off,
                                                y + off, Vall + static_cast<long long>(k) * P.vsize);
}

// `first`, `nf`: the run of the level's fronts (of at most 159 rows) this
// launch substitutes, its LDS sized to the run's largest L block.
template <typename T, int R>
__global__ __launch_bounds__(64) void k_backward_level(DPlan P, int h, int first, int nf, int B, long long lstride,
                                                       const int* __restrict__ active,
                                                       const int* __restrict__ done, const T* __restrict__ L,
                                                       const double* __restrict__ y, double* __restrict__ z) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* xu = reinterpret_cast<double*>(smem_raw);
    T* Ls = reinterpret_cast<T*>(smem_raw + (R == 1 ? 64 : 192) * sizeof(double));
    const int lane = threadIdx.x;
    const int k = blockIdx.x / nf;
    if (k >= B || !active[k] || done[k]) return;
    const int t = P.order[P.lev_ptr[h] + first + (blockIdx.x - k * nf)];
    const long long off = static_cast<long long>(k) * P.n;
    backward_front<T, R>(P, t, L + static_cast<long long>(k) * lstride, Ls, xu, lane, y + off, z + off);
}

// Solves of one system for many right-hand sides (graphs: one L shared by
// every instance): workgroup (f, g) is one wavefront for front f and the
// kRhs right-hand sides g kRhs.., the front's L block staged in LDS once.
// Lane 4 r + q holds rows 16 R q .. 16 R q + 16 R - 1 of right-hand side r's
// front vector (R = 1: fronts of at most 64 rows; R = 3: at most 192), so a
// pivot is eliminated for all of them at once: the forward substitution
// broadcasts y_j from its row's lane to the three others of its quad, the
// backward sums a pivot's dot product over the quad; each lane updates its
// own rows. The dependency chain of a front is one pivot step long per
// pivot, not one per pivot and right-hand side.
constexpr int kRhs = 16;

// v of lane `src` (LDS crossbar; no LDS memory is touched).
__device__ __forceinline__ double lane_read(double v, int src) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b >> 32));
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// v summed over the four lanes of its quad (DPP quad permutations), the
// same sum in every lane of the quad.
template <int Ctrl>
__device__ __forceinline__ double quad_perm(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b & 0xffffffffll), Ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), Ctrl, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double quad_sum(double v) {
    v += quad_perm<0xB1>(v);  // lanes (1, 0, 3, 2)
    return v + quad_perm<0x4E>(v);  // lanes (2, 3, 0, 1)
}

// Levels of small fronts solved four lanes per window: lane 4 w + q of
// workgroup (g, f) substitutes front f of the level for window 16 g + w and
// holds rows q, q + 4, q + 8, .. of the front's vector in registers (rows
// statically indexed, loops unrolled to kLaneNp pivots and M rows). The four
// lanes of a window read four neighbouring entries of a column, so one load
// instruction touches 16 windows' segments of L rather than 64 windows' single
// entries, and the loads of a column are in flight together. A pivot's value
// is broadcast over its quad (forward) or its dot product summed over the quad
// (backward) with DPP; each lane updates only its own rows. A segment that
// straddles the diagonal reads rows above it (inside the same L block), and
// rows past the front's m read the entries that follow the column in the same
// window's L (the host checks that they lie inside it); the forward discards
// them, the backward masks their terms.
constexpr int kLaneNp = 32;

// every lane of the quad takes lane r's value (r folds to a constant in the
// unrolled loops that call this)
__device__ __forceinline__ double quad_bcast(double v, int r) {
    switch (r) {
        case 0: return quad_perm<0x00>(v);
        case 1: return quad_perm<0x55>(v);
        case 2: return quad_perm<0xAA>(v);
        default: return quad_perm<0xFF>(v);
    }
}

// Forward substitution of a front without children (reference
// system.cpp:974-989): y = L11^-1 rs, update vector -L21 y.
template <typename T, int M>
__global__ __launch_bounds__(64) void k_forward_quads(DPlan P, int h, int B, long long lstride,
                                                      const int* __restrict__ active, const int* __restrict__ done,
                                                      const T* __restrict__ L, const double* __restrict__ rs,
                                                      double* __restrict__ y, double* __restrict__ Vall) {
    constexpr int MQ = M / 4;
    static_assert(M % 4 == 0, "four lanes a window");
    const int q = threadIdx.x & 3;
    const int k = blockIdx.x * 16 + (threadIdx.x >> 2);
    const int t = P.order[P.lev_ptr[h] + blockIdx.y];
    // the four lanes of a window leave together: DPP reads stay in live quads
    if (k >= B || !active[k] || done[k]) return;
    const int np = P.npiv[t], m = np + P.nupd[t], first = P.first[t];
    const T* Lt = L + static_cast<long long>(k) * lstride + P.lpos[t] + q;
    const double* rw = rs + static_cast<long long>(k) * P.n + first + q;
    double f[MQ];
#pragma unroll
    for (int s = 0; s < MQ; ++s) f[s] = 4 * s + q < np ? rw[4 * s] : 0.0;
#pragma unroll
    for (int j = 0; j < kLaneNp; ++j) {
        if (j < np) {
            const T* col = Lt + colstart(j, m) - j;  // col[4 s] = L[4 s + q][j]
            T lv[MQ];
#pragma unroll
            for (int s = j >> 2; s < MQ; ++s) lv[s] = col[4 * s];
            const double yj = quad_bcast(f[j >> 2] * inv_pivot(lv[j >> 2]), j & 3);
#pragma unroll
            for (int s = j >> 2; s < MQ; ++s) {
                const int i = 4 * s + q;
                if (i > j) f[s] -= static_cast<double>(lv[s]) * yj;
                else if (i == j) f[s] = yj;
            }
        }
    }
    double* yw = y + static_cast<long long>(k) * P.n + first + q;
    double* Vt = Vall + static_cast<long long>(k) * P.vsize + P.vpos[t] - np + q;
#pragma unroll
    for (int s = 0; s < MQ; ++s) {
        const int i = 4 * s + q;
        if (i < np) yw[4 * s] = f[s];
        else if (i < m) Vt[4 * s] = f[s];
    }
}

// Back substitution of a front (reference system.cpp:992-1016): z_piv =
// L11^-T (y - L21^T z_upd), column j of L dotted with the solved rows below
// it, from the last pivot up.
template <typename T, int M>
__global__ __launch_bounds__(64) void k_backward_quads(DPlan P, int h, int B, long long lstride,
                                                       const int* __restrict__ active, const int* __restrict__ done,
                                                       const T* __restrict__ L, const double* __restrict__ y,
                                                       double* __restrict__ z) {
    constexpr int MQ = M / 4;
    static_assert(M % 4 == 0, "four lanes a window");
    const int q = threadIdx.x & 3;
    const int k = blockIdx.x * 16 + (threadIdx.x >> 2);
    const int t = P.order[P.lev_ptr[h] + blockIdx.y];
    if (k >= B || !active[k] || done[k]) return;
    const int np = P.npiv[t], m = np + P.nupd[t], first = P.first[t];
    const int* rows = P.rows + P.upd[t];
    const T* Lt = L + static_cast<long long>(k) * lstride + P.lpos[t] + q;
    const long long off = static_cast<long long>(k) * P.n;
    double x[MQ];
#pragma unroll
    for (int s = 0; s < MQ; ++s) {
        const int i = 4 * s + q;
        if (i < np) x[s] = y[off + first + i];
        else if (i < m) x[s] = z[off + rows[i - np]];
        else x[s] = 0.0;
    }
#pragma unroll
    for (int j = kLaneNp - 1; j >= 0; --j) {
        if (j < np) {
            const T* col = Lt + colstart(j, m) - j;
            T lv[MQ];
#pragma unroll
            for (int s = j >> 2; s < MQ; ++s) lv[s] = col[4 * s];
            double acc[2] = {0.0, 0.0};
#pragma unroll
            for (int s = j >> 2; s < MQ; ++s) {
                const int i = 4 * s + q;
                const double l = (i > j && i < m) ? static_cast<double>(lv[s]) : 0.0;
                acc[s & 1] = fma(l, x[s], acc[s & 1]);
            }
            const double dot = quad_sum(acc[0] + acc[1]);
            if (q == (j & 3)) x[j >> 2] = (x[j >> 2] - dot) * inv_pivot(lv[j >> 2]);
        }
    }
    double* zw = z + off + first + q;
#pragma unroll
    for (int s = 0; s < MQ; ++s)
        if (4 * s + q < np) zw[4 * s] = x[s];
}

// The L block of front t into LDS (16-byte chunks; blocks start on 16 bytes).
template <typename T>
__device__ __forceinline__ void stage_L(const T* L, long long lpos, int np, int m, int lane, T* Lc) {
    const long long sz = static_cast<long long>(np) * m - ((static_cast<long long>(np) * (np - 1)) >> 1);
    const int nv = static_cast<int>((sz * static_cast<long long>(sizeof(T)) + 15) / 16);
    const uint4* src = reinterpret_cast<const uint4*>(L + lpos);
    uint4* dst = reinterpret_cast<uint4*>(Lc);
#pragma unroll 8
    for (int e = lane; e < nv; e += 64) dst[e] = src[e];
}

// LDS of the multi-right-hand-side kernels: kRhs front vectors of 64 R rows
// (doubles), then the L block with 64 R entries of slack (the row loops read
// up to that far past a column; those terms are masked).
template <typename T, int R>
__host__ __device__ constexpr std::size_t rhs_lds(int lmax) {
    return static_cast<std::size_t>(kRhs) * 64 * R * sizeof(double) +
           static_cast<std::size_t>(lmax + 64 * R + 4) * sizeof(T);
}

template <typename T, int R>
__global__ __launch_bounds__(64) void k_forward_rhs(DPlan P, int h, int k, const int* __restrict__ done,
                                                    const T* __restrict__ L, const double* __restrict__ rs,
                                                    double* __restrict__ y, double* __restrict__ Vall) {
    constexpr int MR = 64 * R;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* fs = reinterpret_cast<double*>(smem_raw);  // [kRhs][MR]
    T* Lc = reinterpret_cast<T*>(smem_raw + kRhs * MR * sizeof(double));
    const int lane = threadIdx.x, rq = lane >> 2, q = lane & 3;
    const int t = P.order[P.lev_ptr[h] + blockIdx.x];
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int g0 = blockIdx.y * kRhs;
    stage_L<T>(L, P.lpos[t], np, m, lane, Lc);
    // the front vectors: the pivots' right-hand sides, then the children's
    // update vectors added at their rows (in child order)
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = g0 + rr;
        const bool on = inst < k && !done[inst];
        const double* rsw = rs + static_cast<long long>(on ? inst : 0) * P.n + first;
        for (int i = lane; i < m; i += 64) fs[rr * MR + i] = on && i < np ? rsw[i] : 0.0;
    }
    for (int ci = P.cb[t]; ci < P.ce[t]; ++ci) {
        const int c = P.kids[ci];
        const int u = P.nupd[c];
        const int* rel = P.relmap + P.upd[c];
        for (int rr = 0; rr < kRhs; ++rr) {
            const int inst = g0 + rr;
            if (inst >= k || done[inst]) continue;
            const double* Vc = Vall + static_cast<long long>(inst) * P.vsize + P.vpos[c];
            for (int ii = lane; ii < u; ii += 64) fs[rr * MR + rel[ii]] += Vc[ii];
        }
    }
    wave_sync();
    // rows 64 sb + 16 q + ii of right-hand side rq in f[sb][ii]
    double f[R][16];
#pragma unroll
    for (int sb = 0; sb < R; ++sb)
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int gi = 64 * sb + 16 * q + ii;
            f[sb][ii] = gi < m ? fs[rq * MR + gi] : 0.0;
        }
    const int quad = lane & ~3;
    for (int j0 = 0; j0 < np; j0 += 16) {
        const int tb = j0 >> 6, qb = (j0 >> 4) & 3;  // the pivots' slot block and quad
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int j = j0 + s;
            if (j < np) {
                const T* col = Lc + colstart(j, m) - j;  // col[i] = L[i][j] for i >= j
                double fj = f[0][s];
#pragma unroll
                for (int sb = 1; sb < R; ++sb) fj = tb == sb ? f[sb][s] : fj;
                const double yj = lane_read(fj * inv_pivot(col[j]), quad | qb);
#pragma unroll
                for (int sb = 0; sb < R; ++sb) {
                    if (tb == sb && q == qb) f[sb][s] = yj;
#pragma unroll
                    for (int ii = 0; ii < 16; ++ii) {
                        const int gi = 64 * sb + 16 * q + ii;
                        const double lv = static_cast<double>(col[gi]);
                        f[sb][ii] = gi > j && gi < m ? fma(-lv, yj, f[sb][ii]) : f[sb][ii];
                    }
                }
            }
        }
    }
    // out through LDS: y of the pivots, the update vector of the rest
    wave_sync();
#pragma unroll
    for (int sb = 0; sb < R; ++sb)
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int gi = 64 * sb + 16 * q + ii;
            if (gi < m) fs[rq * MR + gi] = f[sb][ii];
        }
    wave_sync();
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = g0 + rr;
        if (inst >= k || done[inst]) continue;
        double* yw = y + static_cast<long long>(inst) * P.n + first;
        double* Vt = Vall + static_cast<long long>(inst) * P.vsize + P.vpos[t] - np;
        for (int i = lane; i < m; i += 64) {
            const double v = fs[rr * MR + i];
            if (i < np) yw[i] = v;
            else Vt[i] = v;
        }
    }
}

template <typename T, int R>
__global__ __launch_bounds__(64) void k_backward_rhs(DPlan P, int h, int k, const int* __restrict__ done,
                                                     const T* __restrict__ L, const double* __restrict__ y,
                                                     double* __restrict__ z) {
    constexpr int MR = 64 * R;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* xs = reinterpret_cast<double*>(smem_raw);  // [kRhs][MR]
    T* Lc = reinterpret_cast<T*>(smem_raw + kRhs * MR * sizeof(double));
    const int lane = threadIdx.x, rq = lane >> 2, q = lane & 3;
    const int t = P.order[P.lev_ptr[h] + blockIdx.x];
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int* rows = P.rows + P.upd[t];
    const int g0 = blockIdx.y * kRhs;
    stage_L<T>(L, P.lpos[t], np, m, lane, Lc);
    // the pivots' y and the solved update rows' z
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = g0 + rr;
        const bool on = inst < k && !done[inst];
        const long long off = static_cast<long long>(on ? inst : 0) * P.n;
        for (int i = lane; i < m; i += 64)
            xs[rr * MR + i] = !on ? 0.0 : i < np ? y[off + first + i] : z[off + rows[i - np]];
    }
    wave_sync();
    double x[R][16];
#pragma unroll
    for (int sb = 0; sb < R; ++sb)
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int gi = 64 * sb + 16 * q + ii;
            x[sb][ii] = gi < m ? xs[rq * MR + gi] : 0.0;
        }
    for (int j0 = ((np - 1) >> 4) << 4; j0 >= 0; j0 -= 16) {
        const int tb = j0 >> 6, qb = (j0 >> 4) & 3;
#pragma unroll
        for (int s = 15; s >= 0; --s) {
            const int j = j0 + s;
            if (j < np) {
                const T* col = Lc + colstart(j, m) - j;
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int sb = 0; sb < R; ++sb)
#pragma unroll
                    for (int ii = 0; ii < 16; ++ii) {
                        const int gi = 64 * sb + 16 * q + ii;
                        const double lv = gi > j && gi < m ? static_cast<double>(col[gi]) : 0.0;
                        acc[ii & 3] = fma(lv, x[sb][ii], acc[ii & 3]);
                    }
                const double dot = quad_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
                const double d = inv_pivot(col[j]);
#pragma unroll
                for (int sb = 0; sb < R; ++sb)
                    if (tb == sb && q == qb) x[sb][s] = (x[sb][s] - dot) * d;
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int sb = 0; sb < R; ++sb)
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int gi = 64 * sb + 16 * q + ii;
            if (gi < np) xs[rq * MR + gi] = x[sb][ii];
        }
    wave_sync();
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = g0 + rr;
        if (inst >= k || done[inst]) continue;
        double* zw = z + static_cast<long long>(inst) * P.n + first;
        for (int i = lane; i < np; i += 64) zw[i] = xs[rr * MR + i];
    }
}

// Fronts of 65 to 192 rows (R = 3): one right-hand side at a time, rows on
// lanes (three a lane), the pivot broadcast by a lane read; their 48 rows a
// lane would hold too many registers in the layout above.
template <typename T, int R>
__device__ __forceinline__ void forward_compute(const T* Lc, const double* fv, int np, int m, int lane, double* yw,
                                                double* Vt) {
    double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        f[r] = i < m ? fv[i] : 0.0;
    }
    for (int j0 = 0; j0 < np; j0 += 16) {
        T Lr[R][16];
        double dinv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = j0 + q;
            const bool okj = j < np;
            const long long cs = okj ? colstart(j, m) : 0;
            dinv[q] = okj ? inv_pivot(Lc[cs]) : 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                Lr[r][q] = (okj && i > j && i < m) ? Lc[cs + i - j] : T(0);
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = j0 + q;
            if (j < np) {
                const int slot = j >> 6;
                const double yj = bcast(pick(f, slot), j & 63) * dinv[q];
#pragma unroll
                for (int r = 0; r < R; ++r) f[r] -= static_cast<double>(Lr[r][q]) * yj;
                if (lane == (j & 63)) {
                    if (slot == 0) f[0] = yj;
                    else if (slot == 1) f[1] = yj;
                    else f[2] = yj;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < np) yw[i] = f[r];
        else if (i < m) Vt[i - np] = f[r];
    }
}

template <typename T, int R>
__device__ __forceinline__ void backward_compute(const T* Lc, const double* xu, const double* ywf, int np, int nu,
                                                 int m, int lane, double* zw) {
    double tv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = lane + 64 * r;
        double acc = 0.0;
        if (j < np) {
            const T* col = Lc + colstart(j, m) + (np - j);
            for (int ii = 0; ii < nu; ++ii) acc += static_cast<double>(col[ii]) * xu[ii];
            acc = ywf[j] - acc;
        }
        tv[r] = acc;
    }
    for (int j1 = np - 1; j1 >= 0; j1 -= 16) {
        T Lr[R][16];
        double dinv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = j1 - q;
            const bool okj = j >= 0;
            dinv[q] = okj ? inv_pivot(Lc[colstart(j, m)]) : 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                Lr[r][q] = (okj && i < j) ? Lc[colstart(i, m) + (j - i)] : T(0);
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = j1 - q;
            if (j >= 0) {
                const int slot = j >> 6;
                const double xj = bcast(pick(tv, slot), j & 63) * dinv[q];
#pragma unroll
                for (int r = 0; r < R; ++r) tv[r] -= static_cast<double>(Lr[r][q]) * xj;
                if (lane == (j & 63)) {
                    if (slot == 0) tv[0] = xj;
                    else if (slot == 1) tv[1] = xj;
                    else tv[2] = xj;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < np) zw[i] = tv[r];
    }
}

// LDS of the rows-on-lanes kernels: one front vector (192 doubles), then
// the L block.
template <typename T>
__host__ __device__ constexpr std::size_t rows_lds(int lmax) {
    return 192 * sizeof(double) + static_cast<std::size_t>(lmax + 4) * sizeof(T);
}

template <typename T, int R>
__global__ __launch_bounds__(64) void k_forward_rows(DPlan P, int h, int k, const int* __restrict__ done,
                                                    const T* __restrict__ L, const double* __restrict__ rs,
                                                    double* __restrict__ y, double* __restrict__ Vall) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* fv = reinterpret_cast<double*>(smem_raw);
    T* Lc = reinterpret_cast<T*>(smem_raw + 192 * sizeof(double));
    const int lane = threadIdx.x;
    const int t = P.order[P.lev_ptr[h] + blockIdx.x];
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int cb = P.cb[t], ce = P.ce[t];
    stage_L<T>(L, P.lpos[t], np, m, lane, Lc);
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = blockIdx.y * kRhs + rr;
        if (inst >= k) break;
        if (done[inst]) continue;
        const double* rsw = rs + static_cast<long long>(inst) * P.n;
        double* V = Vall + static_cast<long long>(inst) * P.vsize;
        for (int i = lane; i < m; i += 64) fv[i] = i < np ? rsw[first + i] : 0.0;
        wave_sync();
        for (int ci = cb; ci < ce; ++ci) {
            const int c = P.kids[ci];
            const int u = P.nupd[c];
            const int* rel = P.relmap + P.upd[c];
            const double* Vc = V + P.vpos[c];
            for (int ii = lane; ii < u; ii += 64) fv[rel[ii]] += Vc[ii];
            wave_sync();
        }
        forward_compute<T, R>(Lc, fv, np, m, lane, y + static_cast<long long>(inst) * P.n + first, V + P.vpos[t]);
        wave_sync();
    }
}

template <typename T, int R>
__global__ __launch_bounds__(64) void k_backward_rows(DPlan P, int h, int k, const int* __restrict__ done,
                                                     const T* __restrict__ L, const double* __restrict__ y,
                                                     double* __restrict__ z) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* xu = reinterpret_cast<double*>(smem_raw);
    T* Lc = reinterpret_cast<T*>(smem_raw + 192 * sizeof(double));
    const int lane = threadIdx.x;
    const int t = P.order[P.lev_ptr[h] + blockIdx.x];
    const int np = P.npiv[t], nu = P.nupd[t], m = np + nu, first = P.first[t];
    const int* rows = P.rows + P.upd[t];
    stage_L<T>(L, P.lpos[t], np, m, lane, Lc);
    for (int rr = 0; rr < kRhs; ++rr) {
        const int inst = blockIdx.y * kRhs + rr;
        if (inst >= k) break;
        if (done[inst]) continue;
        const long long off = static_cast<long long>(inst) * P.n;
        for (int ii = lane; ii < nu; ii += 64) xu[ii] = z[off + rows[ii]];
        wave_sync();
        backward_compute<T, R>(Lc, xu, y + off + first, np, nu, m, lane, z + off + first);
        wave_sync();
    }
}

// The neighbour offsets of ndy / ndx as arithmetic on a compile-time count,
// so that unrolled loops fold them into immediates.
template <int ND>
__device__ __forceinline__ int fixed_dy(int d) {
    if (ND == 8) return d == 1 || d == 6 ? 0 : (d == 0 || d == 3 || d == 5 ? -1 : 1);
    return d == 1 ? -1 : (d == 2 ? 1 : 0);
}
template <int ND>
__device__ __forceinline__ int fixed_dx(int d) {
    if (ND == 8) return d < 3 ? -1 : (d < 5 ? 0 : 1);
    return d == 0 ? -1 : (d == 3 ? 1 : 0);
}

// Rows of r = b - A x of one window, a row per thread. The neighbour loop is
// unrolled and branch-free (clamped indices, masked terms) so that a row's
// loads are requested in two rounds (its neighbours' indices, then their
// values) instead of one round trip per neighbour; the terms are added in
// the neighbour order of the reference.
template <int ND>
__device__ __forceinline__ void residual_rows(const DPlan& P, int W, int n, long long off, const double* gw,
                                              const unsigned char* pw, const double* dg, const double* sc,
                                              const double* bvec, const double* x, double* rs, double& rr,
                                              double& bb) {
    for (int p = threadIdx.x; p < n; p += kThreads) {
        const int c = P.cell[p];
        int q[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) q[d] = P.nbr[p * ND + d];
        const double xp = x[off + p], dgp = dg[off + p], bp = bvec[off + p], scp = sc[off + p];
        const bool pr = pw[c];
        const double gc = gw[c];
        double gn[ND], xn[ND];
        bool on[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const bool in = q[d] >= 0;  // the neighbour is a cell of the window: c plus its offset
            const int c2 = in ? c + fixed_dy<ND>(d) * W + fixed_dx<ND>(d) : c;
            const bool pc = pw[c2];  // loaded for every neighbour (c2 is clamped)
            on[d] = in & pc;
            gn[d] = gw[c2];
            xn[d] = x[off + (in ? q[d] : p)];
        }
        double acc = dgp * xp;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const bool diag = ND == 8 && fixed_dx<ND>(d) != 0 && fixed_dy<ND>(d) != 0;
            if (on[d]) acc -= edge_weight(gc, gn[d], diag) * xn[d];
        }
        const double r = pr ? bp - acc : 0.0;
        if (pr) bb += bp * bp;
        rr += r * r;
        rs[off + p] = scp * r;
    }
}

// The substitutions' solution z (of the scaled system) taken into x first:
// x = s z after the first solve, x += s z after a refinement step (the
// window's step count tells which); then r = b - A x in FP64 from the
// conductances (the reference's matrix, shifted, system.cpp:458-487);
// relative residual ||r|| / ||b||; the scaled residual s r for the next
// refinement step. A window is done when it reaches `tol` or stops
// improving. Every substitution sweep is followed by this kernel, so x is
// the solution once the last residual is taken.
static __global__ __launch_bounds__(kThreads) void k_residual(DPlan P, int B, const double* __restrict__ g,
                                                       const unsigned char* __restrict__ present,
                                                       const double* __restrict__ dg, const double* __restrict__ sc,
                                                       const double* __restrict__ bvec, const double* __restrict__ z,
                                                       double* __restrict__ x, double* __restrict__ rs,
                                                       const int* __restrict__ active, int* __restrict__ done,
                                                       double* __restrict__ rnorm, double* __restrict__ rprev,
                                                       int* __restrict__ iters, double tol, int* __restrict__ open) {
    __shared__ double red[kWaves];
    const int k = blockIdx.x;
    if (k >= B || !active[k] || done[k]) return;
    const int W = P.W, W2 = W * W, nd = P.nd, n = P.n;
    const double* gw = g + static_cast<long long>(k) * W2;
    const unsigned char* pw = present + static_cast<long long>(k) * W2;
    const long long off = static_cast<long long>(k) * n;
    {
        const bool first = iters[k] == 0;
        for (int p = threadIdx.x; p < n; p += kThreads) {
            const double d = sc[off + p] * z[off + p];
            x[off + p] = first ? d : x[off + p] + d;
        }
        __syncthreads();
    }
    double rr = 0.0, bb = 0.0;
    if (nd == 8) residual_rows<8>(P, W, n, off, gw, pw, dg, sc, bvec, x, rs, rr, bb);
    else residual_rows<4>(P, W, n, off, gw, pw, dg, sc, bvec, x, rs, rr, bb);
    rr = block_sum(rr, red);
    bb = block_sum(bb, red);
    if (threadIdx.x == 0) {
        const double rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        const double prev = rprev[k];
        rnorm[k] = rel;
        rprev[k] = rel;
        const bool stalled = iters[k] > 0 && rel > 0.5 * prev;
        if (rel <= tol || stalled || !(rel == rel)) {
            done[k] = 1;
        } else {
            iters[k] += 1;
            atomicAdd(open, 1);
        }
    }
}

// The neighbours of cell q of a W x W window for the current kernels: the
// neighbour's cell (q itself when outside the window) and whether the
// branch exists (inside, present). Loads are unconditional so that a cell's
// neighbours are requested together.
template <int ND>
__device__ __forceinline__ void cell_neighbours(int W, int q, int y, int xx, const unsigned char* pw, int (&r)[ND],
                                                bool (&on)[ND]) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const int yy = y + fixed_dy<ND>(d), x2 = xx + fixed_dx<ND>(d);
        const bool in = yy >= 0 && yy < W && x2 >= 0 && x2 < W;
        r[d] = in ? yy * W + x2 : q;
        const bool pr = pw[r[d]];
        on[d] = in & pr;
    }
}

// Largest branch current in each direction over the window (the reference's
// maxout / maxin of node_currents, each branch once: its lower neighbours).
template <int ND>
__device__ __forceinline__ void branch_extremes(int W, const double* gw, const unsigned char* pw, const double* v,
                                                double& mo, double& mi) {
    constexpr int lower = ND == 8 ? 4 : 2;
    for (int q = threadIdx.x; q < W * W; q += kCurThreads) {
        if (!pw[q]) continue;
        const int y = q / W, xx = q - y * W;
        int r[ND];
        bool on[ND];
        cell_neighbours<ND>(W, q, y, xx, pw, r, on);
        const double gq = gw[q], vq = v[q];
        double gr[ND];
#pragma unroll
        for (int d = lower; d < ND; ++d) gr[d] = gw[r[d]];
#pragma unroll
        for (int d = lower; d < ND; ++d) {
            if (!on[d]) continue;
            const bool diag = ND == 8 && fixed_dx<ND>(d) != 0 && fixed_dy<ND>(d) != 0;
            const double b = edge_weight(gq, gr[d], diag) * (v[r[d]] - vq);
            mo = fmax(mo, b);
            mi = fmax(mi, -b);
        }
    }
}

// A node's current: the larger of its inflow and outflow over branches not
// negligible against maxin / maxout (reference grid.cpp:45-70).
template <int ND>
__device__ __forceinline__ double node_current(int W, int q, int y, int xx, const double* gw,
                                               const unsigned char* pw, const double* v, double maxin,
                                               double maxout) {
    constexpr int lower = ND == 8 ? 4 : 2;
    int r[ND];
    bool on[ND];
    cell_neighbours<ND>(W, q, y, xx, pw, r, on);
    const double gq = gw[q], vq = v[q];
    double gr[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) gr[d] = gw[r[d]];
    double in = 0.0, outf = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        if (!on[d]) continue;
        const bool dgl = ND == 8 && fixed_dx<ND>(d) != 0 && fixed_dy<ND>(d) != 0;
        const double vr = v[r[d]];
        const double b = d < lower ? -(edge_weight(gr[d], gq, dgl) * (vq - vr)) : edge_weight(gq, gr[d], dgl) * (vr - vq);
        if (b > 0) {
            if (!(fabs(b / maxin) < 1e-8)) in += b;
        } else if (b < 0) {
            if (!(fabs(b / maxout) < 1e-8)) outf -= b;
        }
    }
    return in > outf ? in : outf;
}

// Node currents of every window (reference grid.cpp:19-71), times the
// correction factors of the clipped window (omniscape.cpp:553-571).
static __global__ __launch_bounds__(kCurThreads) void k_currents(DPlan P, Lattice Lt, long long k0, int B,
                                                       const double* __restrict__ g,
                                                       const unsigned char* __restrict__ present,
                                                       const double* __restrict__ x, const int* __restrict__ active,
                                                       const double* __restrict__ corr, double* __restrict__ cur,
                                                       double* __restrict__ vgrid) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the window's voltages by cell: in LDS, or (windows wider than kMaxW) in vgrid, W^2 per window
    double* v = vgrid ? vgrid + static_cast<long long>(blockIdx.x) * P.W * P.W : reinterpret_cast<double*>(smem_raw);
    __shared__ double red[kCurThreads / 64];
    const int k = blockIdx.x;
    if (k >= B || !active[k]) return;
    const int W = P.W, W2 = W * W, nd = P.nd, D = P.D;
    const double* gw = g + static_cast<long long>(k) * W2;
    const unsigned char* pw = present + static_cast<long long>(k) * W2;
    const long long off = static_cast<long long>(k) * P.n;
    double* cw = cur + static_cast<long long>(k) * W2;
    for (int c = threadIdx.x; c < W2; c += kCurThreads) v[c] = 0.0;
    __syncthreads();
    for (int p = threadIdx.x; p < P.n; p += kCurThreads) {
        const int c = P.cell[p];
        if (pw[c]) v[c] = x[off + p];
    }
    __syncthreads();
    double mo = -INFINITY, mi = -INFINITY;
    if (nd == 8) branch_extremes<8>(W, gw, pw, v, mo, mi);
    else branch_extremes<4>(W, gw, pw, v, mo, mi);
    const double maxout = block_max<kCurThreads / 64>(mo, red);
    const double maxin = block_max<kCurThreads / 64>(mi, red);
    // correction window offsets (clipped windows keep Omniscape's slice of it)
    long long tx = 0, ty = 0;
    long long cx0 = 0, cy0 = 0, ox = 0, oy = 0, x0c = 0, y0c = 0;
    if (corr) {
        target_of(Lt, k0 + k, tx, ty);
        oy = ty - D, ox = tx - D;
        y0c = oy > 0 ? oy : 0, x0c = ox > 0 ? ox : 0;
        const long long y1c = ty + D < Lt.ny - 1 ? ty + D : Lt.ny - 1;
        const long long x1c = tx + D < Lt.nx - 1 ? tx + D : Lt.nx - 1;
        const long long h = y1c - y0c + 1, w = x1c - x0c + 1;
        cx0 = (tx < D && tx + D <= Lt.nx - 1) ? W - w : 0;
        cy0 = (ty < D && ty + D <= Lt.ny - 1) ? W - h : 0;
    }
    for (int q = threadIdx.x; q < W2; q += kCurThreads) {
        double out = 0.0;
        const bool pq = pw[q];
        if (pq) {
            const int y = q / W, xx = q - (q / W) * W;
            out = nd == 8 ? node_current<8>(W, q, y, xx, gw, pw, v, maxin, maxout)
                          : node_current<4>(W, q, y, xx, gw, pw, v, maxin, maxout);
            if (corr) {
                const long long yc = oy + y - y0c, xc = ox + xx - x0c;
                out *= corr[(cy0 + yc) * W + (cx0 + xc)];
            }
        }
        cw[q] = out;
    }
}

// Cumulative map: every map cell adds the currents of the batch's windows
// that cover it, in lattice order (reference omniscape.cpp:581-605 sums in
// target order; any fixed order keeps runs reproducible).
static __global__ __launch_bounds__(kThreads) void k_accumulate(Lattice Lt, long long k0, int B, long long X0, long long ncol,
                                                         const int* __restrict__ active,
                                                         const double* __restrict__ cur, double* __restrict__ map) {
    const long long e = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (e >= ncol * Lt.ny) return;
    const long long Y = e / ncol, X = X0 + (e - (e / ncol) * ncol);
    const int D = Lt.D, W = Lt.W, bs = Lt.bs, br = Lt.br;
    const long long k1 = k0 + B;
    // lattice columns / rows whose window covers (Y, X)
    long long ixa = X - D - br, ixb = X + D - br;
    ixa = ixa <= 0 ? 0 : (ixa + bs - 1) / bs;
    ixb = ixb < 0 ? -1 : ixb / bs;
    long long iya = Y - D - br, iyb = Y + D - br;
    iya = iya <= 0 ? 0 : (iya + bs - 1) / bs;
    iyb = iyb < 0 ? -1 : iyb / bs;
    if (iyb > Lt.nly - 1) iyb = Lt.nly - 1;
    const long long kfirst = k0 / Lt.nly, klast = (k1 - 1) / Lt.nly;
    if (ixa < kfirst) ixa = kfirst;
    if (ixb > klast) ixb = klast;
    double sum = 0.0;
    for (long long ix = ixa; ix <= ixb; ++ix)
        for (long long iy = iya; iy <= iyb; ++iy) {
            const long long kk = ix * Lt.nly + iy;
            if (kk < k0 || kk >= k1) continue;
            const int k = static_cast<int>(kk - k0);
            if (!active[k]) continue;
            const long long tx = br + ix * bs, ty = br + iy * bs;
            const long long wy = Y - ty + D, wx = X - tx + D;
            sum += cur[static_cast<long long>(k) * W * W + wy * W + wx];
        }
    if (sum != 0.0) map[Y * Lt.nx + X] += sum;
}

static __global__ void k_reset(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace lsc
