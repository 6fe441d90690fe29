// Omniscape on the window engine: host driver and C-ABI entry points.
//
// The reference runs Omniscape's targets in parallel CPU chunks, one sparse
// factorisation per window in double precision
// (src/cpp/circuit/omniscape.cpp:394-608). Here every window of a run, also
// the ones clipped at the map edges, is the same W x W disk with the cells
// off the map out of the circuit, so all windows share one symbolic plan and
// a batch of thousands of windows is factorised level by level, every front
// of a level in one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/landschaft_circuit.h"
#include "common.hpp"
#include "huge.hpp"
#include "window.hpp"
#include "window_kernels.hpp"

namespace lsc {

// Zeroes n doubles (a kernel rather than a memset, so a caller's capture of
// the stream records kernels only).
__global__ void k_zero_f64(long long n, double* a) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) a[q] = 0.0;
}

__global__ void k_zero_u64(int n, unsigned long long* a) {
    if (static_cast<int>(threadIdx.x) < n) a[threadIdx.x] = 0ull;
}

// Flow potential circuits: conductance 1 on every on-map cell of the window
// (reference omniscape.cpp:127-166 uses a window of ones).
__global__ void k_unit_conductance(Lattice Lt, long long k0, int B, double* g) {
    const int k = blockIdx.x;
    if (k >= B || k0 + k >= Lt.nlat) return;
    const long long kk = k0 + k;
    const long long tx = Lt.br + (kk / Lt.nly) * Lt.bs, ty = Lt.br + (kk % Lt.nly) * Lt.bs;
    const int W = Lt.W, D = Lt.D;
    double* gw = g + static_cast<long long>(k) * W * W;
    for (int c = threadIdx.x; c < W * W; c += blockDim.x) {
        const long long Y = ty - D + c / W, X = tx - D + c % W;
        gw[c] = (Y >= 0 && Y < Lt.ny && X >= 0 && X < Lt.nx) ? 1.0 : -9999.0;
    }
}

namespace {

template <typename T>
std::size_t big_lds() {
    return static_cast<std::size_t>(kBigTiles * (kBigTiles + 1) / 2) * 256 * sizeof(T);
}
template <typename T>
std::size_t small_lds(int stile) {
    return static_cast<std::size_t>(kSmallWaves) * stile * 256 * sizeof(T);
}
// LDS of k_flood_wide: open, present and next-present row masks.
std::size_t wide_flood_lds(int W) { return 3 * static_cast<std::size_t>(W) * ((W + 63) / 64) * 8; }

// LDS of k_currents: the window's voltages, in HBM (no LDS) beyond kMaxW.
std::size_t currents_lds(int W) { return W <= kMaxW ? static_cast<std::size_t>(W) * W * sizeof(double) : 0; }

template <typename T>
std::size_t solve_lds(int lmax) {
    return 192 * sizeof(double) + static_cast<std::size_t>(lmax + 4) * sizeof(T);  // + the 16-byte tail
}

// Cells of the nested-dissection leaves of the window plans (LSC_ND_LEAF,
// default 32: half the fronts of 16-cell leaves for 25 % more flops).
// Batches in flight at once (LSC_BATCH_STREAMS, 1-4; default 3: on the
// 4096^2 radius-50 map 2, 3 and 4 give 161k, 169-170k and 163-167k
// circuits/s with the current kernels).
long long batch_streams() {
    if (active_arena()) return 1;  // the static scratch serves one batch on the caller's stream
    const char* e = std::getenv("LSC_BATCH_STREAMS");
    const int v = e ? std::atoi(e) : 3;
    return v >= 1 && v <= 4 ? v : 3;
}

int nd_leaf() {
    const char* e = std::getenv("LSC_ND_LEAF");
    const int v = e ? std::atoi(e) : 32;
    return v >= 4 ? v : 32;
}

void reset(int* p, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

// Dynamic LDS beyond the default limit must be allowed per kernel.
void allow_lds(const void* fn, std::size_t bytes) {
    LSC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

}  // namespace

namespace {

// Host copy of the device tile layout (front.hpp tix): element (i, j), i >= j.
int host_tix(int i, int j) { return ((((i >> 4) * ((i >> 4) + 1)) >> 1) + (j >> 4)) * 256 + ((j & 15) << 4) + (i & 15); }

}  // namespace

void DevPlan::upload(Plan p, hipStream_t s, bool tables) {
    host = std::move(p);
    const Plan& h = host;
    const int H = static_cast<int>(h.lev_ptr.size()) - 1;
    lev_small_host.assign(H, 0);
    lev_fit_host.assign(H, 0);
    lev_stile.assign(H, 1);
    lev_btile.assign(H, 1);
    lev_lmax.assign(H, 1);
    lev_mmax.assign(H, 1);
    lev_sgroups.assign(H, {});
    lev_bgroups.assign(H, {});
    lev_lgroups.assign(H, {});
    auto group = [](std::vector<FrontGroup>& g, int off, int tiles) {
        if (!g.empty() && g.back().tiles == tiles && g.back().first + g.back().count == off) ++g.back().count;
        else g.push_back(FrontGroup{off, 1, tiles});
    };
    huge.clear();
    for (int lv = 0; lv < H; ++lv)
        for (int k = h.lev_ptr[lv]; k < h.lev_ptr[lv + 1]; ++k) {
            const int t = h.order[k];
            const long long np = h.npiv[t], m = np + h.nupd[t];
            if (m + 1 <= kSmallRows) {
                ++lev_small_host[lv];
                const int MT = static_cast<int>((m + 1 + 15) / 16);
                lev_stile[lv] = std::max(lev_stile[lv], MT * (MT + 1) / 2);
                group(lev_sgroups[lv], k - h.lev_ptr[lv], MT * (MT + 1) / 2);
            }
            if (m + 1 > kBigTiles * 16) {
                huge.push_back(t);
                continue;
            }
            // the plan orders a level's fronts by size: the large ones come last
            if (k - h.lev_ptr[lv] != lev_fit_host[lv]) throw std::logic_error("a level's large fronts are not last");
            ++lev_fit_host[lv];
            if (m + 1 > kSmallRows) {
                const int MT = static_cast<int>((m + 1 + 15) / 16);
                lev_btile[lv] = std::max(lev_btile[lv], MT * (MT + 1) / 2);
                group(lev_bgroups[lv], k - h.lev_ptr[lv], MT * (MT + 1) / 2);
            }
            const int lsz = static_cast<int>(np * m - np * (np - 1) / 2);
            lev_lmax[lv] = std::max(lev_lmax[lv], lsz);
            lev_mmax[lv] = std::max(lev_mmax[lv], static_cast<int>(m));
            // back-substitution runs: a new run where the L block needs a
            // larger power-of-two share of LDS (from 4 KB) than the run so far
            auto cls = [](int e) {
                int c = 1024;
                while (c < e) c <<= 1;
                return c;
            };
            std::vector<FrontGroup>& lg = lev_lgroups[lv];
            const int off = k - h.lev_ptr[lv];
            if (!lg.empty() && cls(lsz) <= cls(lg.back().tiles)) {
                ++lg.back().count;
                lg.back().tiles = std::max(lg.back().tiles, lsz);
            } else {
                lg.push_back(FrontGroup{off, 1, lsz});
            }
        }
    std::vector<long long> uposA(h.nfront, 0);
    long long ua = 0;
    for (int k = 0; k < h.nfront; ++k) {
        const int t = h.order[k];
        // the lower triangle of the (nu+1)^2 update matrix packed by columns,
        // 4-entry aligned: the extend-add reads 16-byte chunks
        uposA[t] = ua;
        const long long u1 = h.nupd[t] + 1;
        ua += (u1 * (u1 + 1) / 2 + 3) & ~3ll;
    }
    ibufs.clear();
    lbufs.clear();
    ibufs.reserve(40);
    lbufs.reserve(8);
    auto I = [&](const auto& v) {
        ibufs.push_back(to_device(v, s));
        return static_cast<const int*>(ibufs.back().p);
    };
    auto Lg = [&](const std::vector<long long>& v) {
        lbufs.push_back(to_device(v, s));
        return static_cast<const long long*>(lbufs.back().p);
    };
    std::vector<long long> lpos(h.lpos.begin(), h.lpos.end()), vpos(h.vpos.begin(), h.vpos.end());
    d.n = h.n, d.nd = h.nd, d.W = static_cast<int>(h.bx), d.D = static_cast<int>(h.bx / 2), d.nfront = h.nfront;
    d.H = H;
    d.cell = I(h.cell), d.nbr = I(h.nbr);
    d.first = I(h.first), d.npiv = I(h.npiv), d.nupd = I(h.nupd), d.upd = I(h.upd), d.cb = I(h.cb), d.ce = I(h.ce);
    d.lpos = Lg(lpos), d.uposA = Lg(uposA), d.vpos = Lg(vpos);
    d.kids = I(h.kids), d.rows = I(h.rows), d.relmap = I(h.relmap);
    d.lev_ptr = I(h.lev_ptr), d.lev_small = I(lev_small_host), d.lev_fit = I(lev_fit_host), d.order = I(h.order);
    d.fe_ptr = I(h.fe_ptr), d.fe_row = I(h.fe_row), d.fe_col = I(h.fe_col), d.fe_p = I(h.fe_p), d.fe_q = I(h.fe_q);
    static_assert(sizeof(long long) == sizeof(h.fe_src[0]), "64-bit entry sources");
    lbufs.emplace_back(h.fe_src.empty() ? 1 : h.fe_src.size());
    if (!h.fe_src.empty()) lbufs.back().upload(reinterpret_cast<const long long*>(h.fe_src.data()), h.fe_src.size(), s);
    d.fe_src = lbufs.back().p;
    d.nent = static_cast<int>(h.fe_row.size());
    {
        std::vector<FrontDesc> desc(h.nfront > 0 ? h.nfront : 1, FrontDesc{});
        for (int k = 0; k < h.nfront; ++k) {
            const int t = h.order[k];
            FrontDesc& f = desc[k];
            f.t = t, f.first = h.first[t], f.np = h.npiv[t], f.nu = h.nupd[t];
            f.fe0 = h.fe_ptr[t], f.fe1 = h.fe_ptr[t + 1], f.cb = h.cb[t], f.ce = h.ce[t];
            f.lpos = h.lpos[t], f.uposA = uposA[t], f.vpos = h.vpos[t];
            const int nk = h.ce[t] - h.cb[t];
            f.nk = nk <= 2 ? nk : -1;
            for (int c = 0; c < 2; ++c) {
                const int kid = c < nk ? h.kids[h.cb[t] + c] : -1;
                f.kid[c] = kid;
                f.ku[c] = kid >= 0 ? h.nupd[kid] : 0;
                f.kupd[c] = kid >= 0 ? h.upd[kid] : 0;
                f.kuposA[c] = kid >= 0 ? uposA[kid] : 0;
                f.kvpos[c] = kid >= 0 ? h.vpos[kid] : 0;
            }
        }
        dbuf = to_device(desc, s);
        d.desc = dbuf.p;
    }
    d.lidx = d.uidx = d.xidx = d.eidx = nullptr;
    if (tables) {
        std::vector<unsigned short> li(static_cast<std::size_t>(h.lsize) + 4, 0), ui(static_cast<std::size_t>(ua) + 4, 0),
            xi(static_cast<std::size_t>(ua) + 4, 0);
        for (int t = 0; t < h.nfront; ++t) {
            const int np = h.npiv[t], nu = h.nupd[t], m = np + nu;
            if (m + 1 > kBigTiles * 16) continue;  // tiled in HBM (huge.hpp): no tables
            long long e = h.lpos[t];
            for (int j = 0; j < np; ++j)
                for (int i = j; i < m; ++i) li[e++] = static_cast<unsigned short>(host_tix(i, j));
            e = uposA[t];
            for (int j = 0; j <= nu; ++j)
                for (int i = j; i <= nu; ++i) ui[e++] = static_cast<unsigned short>(host_tix(np + i, np + j));
            // where each entry of this front's update matrix lands in its parent
            // (columns 0..nu-1; the last row, the right-hand side, goes to the
            // parent's right-hand side row)
        }
        for (int t = 0; t < h.nfront; ++t) {
            const int m = h.npiv[t] + h.nupd[t];
            if (m + 1 > kBigTiles * 16) continue;
            for (int ci = h.cb[t]; ci < h.ce[t]; ++ci) {
                const int c = h.kids[ci], u = h.nupd[c];
                auto rel = [&](int i) { return i < u ? h.relmap[h.upd[c] + i] : m; };
                long long e = uposA[c];
                for (int j = 0; j < u; ++j)
                    for (int i = j; i <= u; ++i) xi[e++] = static_cast<unsigned short>(host_tix(rel(i), rel(j)));
                // the child's right-hand-side corner and its padding to four
                // entries go to the parent's corner (m, m), which no result
                // reads: the extend-add then needs no per-entry test
                const long long u1 = u + 1, end = uposA[c] + ((u1 * (u1 + 1) / 2 + 3) & ~3ll);
                while (e < end) xi[e++] = static_cast<unsigned short>(host_tix(m, m));
            }
        }
        std::vector<unsigned short> ei(h.fe_row.size() + 1, 0);
        for (std::size_t e = 0; e < h.fe_row.size(); ++e)
            ei[e] = static_cast<unsigned short>(host_tix(h.fe_row[e], h.fe_col[e]));
        tbuf[0] = to_device(li, s), tbuf[1] = to_device(ui, s), tbuf[2] = to_device(xi, s), tbuf[3] = to_device(ei, s);
        d.lidx = tbuf[0].p, d.uidx = tbuf[1].p, d.xidx = tbuf[2].p, d.eidx = tbuf[3].p;
    }
    d.lsize = h.lsize, d.usizeA = ua, d.vsize = h.vsize;
}

void WindowBatchBuffers::ensure(int B_, int W2_, int n_, long long lsize_, int elem_) {
    if (B_ <= B && W2_ == W2 && n_ == n && lsize_ == lsize && elem_ == elem) return;
    B = B_, W2 = W2_, n = n_, lsize = lsize_, elem = elem_;
    const std::size_t bw = static_cast<std::size_t>(B) * W2, bn = static_cast<std::size_t>(B) * n;
    g.alloc(bw), src.alloc(bw), cur.alloc(bw), present.alloc(bw);
    dg.alloc(bn), sc.alloc(bn), b.alloc(bn), x.alloc(bn), y.alloc(bn), z.alloc(bn), rs.alloc(bn);
    rnorm.alloc(B), rprev.alloc(B);
    active.alloc(B), fail.alloc(B), done.alloc(B), iters.alloc(B);
    L.alloc((static_cast<std::size_t>(B) * lsize + 4) * elem);  // + 4: whole 16-byte loads at the end
}

// One batch of windows whose g, src and active flags are filled: flood,
// factor, solve, refine. Returns per-window refinement statistics in `st`.
template <typename T>
struct WindowSolver {
    hipStream_t s;
    DevPlan plan;
    WindowBatchBuffers wb;
    DevBuf<char> U, A;  // update matrices, assembled entries (per instance)
    DevBuf<double> V;
    DevBuf<int> open;
    int Bcap = 0;
    // the batch's fixed launch sequences, replayed as graphs (keyed by the
    // batch size; set_plan drops them with the buffers they name)
    LaunchGraph g_flood, g_factor, g_sweep0, g_residual, g_sweep1;
    double residual_tol = -1.0;
    // per level: rows M of the one-window-per-lane substitutions (0: the
    // one-front-per-wavefront kernels), forward and backward
    std::vector<int> quad_fwd, quad_bwd;
    // mixed precision: the FP64 solver of the windows whose FP32 factor
    // failed or whose refinement stalled (made on first use)
    std::unique_ptr<WindowSolver<double>> fb;
    int fb_cap = 0;
    // fronts of more than 159 rows (windows of large radius): per level, in
    // HBM tiles, for huge_chunk windows at a time
    std::vector<HugeLevel> lev_huge;
    long long huge_tiles = 0;  // largest level arena per window (tiles)
    int huge_rows = 0, huge_nh = 0, huge_chunk = 0;
    DevBuf<char> HF, HD;       // tile arena and diagonal-block scratch of a chunk

    explicit WindowSolver(hipStream_t st) : s(st) {}

    void set_plan(Plan p, int B) {
        for (LaunchGraph* g : {&g_flood, &g_factor, &g_sweep0, &g_residual, &g_sweep1}) g->clear();
        const std::size_t used0 = active_arena() ? active_arena()->used : 0;
        plan.upload(std::move(p), s, true);
        huge_tiles = build_huge_levels(plan.host, 32, lev_huge, huge_rows, s);
        huge_nh = 0;
        for (const HugeLevel& hl : lev_huge) huge_nh = std::max(huge_nh, static_cast<int>(hl.t.size()));
        if (huge_tiles > 0) {
            // windows per chunk: launch rows (fronts x windows) within 65535, the arena within 2 GiB
            // (one window in the static scratch)
            const long long per = huge_tiles * 256 * static_cast<long long>(sizeof(T));
            const long long cap = active_arena() ? 1 : std::max<long long>(1, (2ll << 30) / per);
            huge_chunk = static_cast<int>(std::max<long long>(1, std::min<long long>({cap, 65535 / huge_nh, B})));
            HF.alloc(static_cast<std::size_t>(huge_chunk) * per);
            HD.alloc(static_cast<std::size_t>(huge_chunk) * huge_nh * 256 * sizeof(T));
            const std::size_t hl = (256 + 16 + static_cast<std::size_t>(huge_rows)) * sizeof(double);
            allow_lds(reinterpret_cast<const void*>(&k_h_forward<T, 1>), hl);
            allow_lds(reinterpret_cast<const void*>(&k_h_backward<T, 1>), hl);
        }
        if (active_arena()) plan_bytes = static_cast<long long>(active_arena()->used - used0);
        int smax = 1, lmax = 1;
        for (int v : plan.lev_stile) smax = std::max(smax, v);
        for (int v : plan.lev_lmax) lmax = std::max(lmax, v);
        allow_lds(reinterpret_cast<const void*>(&k_factor_small<T>), small_lds<T>(smax));
        allow_lds(reinterpret_cast<const void*>(&k_factor_big<T>), big_lds<T>());
        allow_lds(reinterpret_cast<const void*>(&k_forward_level<T, 1>), solve_lds<T>(lmax));
        allow_lds(reinterpret_cast<const void*>(&k_backward_level<T, 1>), solve_lds<T>(lmax));
        allow_lds(reinterpret_cast<const void*>(&k_forward_level<T, 3>), solve_lds<T>(lmax));
        allow_lds(reinterpret_cast<const void*>(&k_backward_level<T, 3>), solve_lds<T>(lmax));
        allow_lds(reinterpret_cast<const void*>(&k_currents), static_cast<std::size_t>(kMaxW) * kMaxW * sizeof(double));
        if (plan.d.W > kMaxW) allow_lds(reinterpret_cast<const void*>(&k_flood_wide), wide_flood_lds(plan.d.W));
        quad_levels();
        const int W = plan.d.W;
        if (active_arena()) B = arena_batch(B);
        Bcap = B;
        wb.ensure(B, W * W, plan.d.n, plan.d.lsize, sizeof(T));
        U.alloc(static_cast<std::size_t>(B) * plan.d.usizeA * sizeof(T));
        V.alloc(static_cast<std::size_t>(B) * (plan.d.vsize ? plan.d.vsize : 1));
        A.alloc(static_cast<std::size_t>(B) * (plan.d.nent ? plan.d.nent : 1) * sizeof(T));
        if (!open.p) open.alloc(4);
    }

    // A level is solved four lanes per window when its fronts have at most
    // kLaneNp pivots and 64 rows (the forward also needs fronts without
    // children), and every column's reads up to M rows stay inside the
    // window's L. LSC_LANE_LEVELS=0 keeps every level on the per-front kernels.
    void quad_levels() {
        const Plan& h = plan.host;
        const int H = static_cast<int>(h.lev_ptr.size()) - 1;
        quad_fwd.assign(H, 0), quad_bwd.assign(H, 0);
        const char* e = std::getenv("LSC_LANE_LEVELS");
        if (e && e[0] == '0') return;
        for (int lv = 0; lv < H; ++lv) {
            int mmax = 1;
            bool fits = true, leaves = true;
            for (int q = h.lev_ptr[lv]; q < h.lev_ptr[lv + 1]; ++q) {
                const int t = h.order[q];
                const int np = h.npiv[t], m = np + h.nupd[t];
                fits = fits && np <= kLaneNp && m <= 64;
                leaves = leaves && h.cb[t] == h.ce[t];
                mmax = std::max(mmax, m);
            }
            if (!fits) continue;
            const int M = mmax <= 32 ? 32 : (mmax <= 48 ? 48 : (mmax <= 56 ? 56 : 64));
            for (int q = h.lev_ptr[lv]; q < h.lev_ptr[lv + 1]; ++q) {
                const int t = h.order[q];
                const long long np = h.npiv[t], m = np + h.nupd[t];
                const long long j = np - 1;  // the last column reads furthest
                if (np > 0 && h.lpos[t] + j * m - ((j * (j - 1)) >> 1) - j + M > h.lsize) fits = false;
            }
            if (!fits) continue;
            quad_bwd[lv] = M;
            if (leaves) quad_fwd[lv] = M;
        }
    }

    void time_phase(double& acc, hipEvent_t a, hipEvent_t b, bool timing) {
        if (!timing) return;
        float ms = 0;
        LSC_HIP(hipEventSynchronize(b));
        LSC_HIP(hipEventElapsedTime(&ms, a, b));
        acc += ms;
    }

    static unsigned blocks(long long tasks, int per) { return static_cast<unsigned>((tasks + per - 1) / per); }

    void factor(int nb) {
        const DPlan& P = plan.d;
        T* L = reinterpret_cast<T*>(wb.L.p);
        T* Up = reinterpret_cast<T*>(U.p);
        T* Ap = reinterpret_cast<T*>(A.p);
        if (P.nent > 0)
            hipLaunchKernelGGL(k_assemble<T>, dim3(blocks(P.nent, kThreads), nb), dim3(kThreads), 0, s, P, nb, wb.g.p,
                               wb.present.p, wb.sc.p, wb.active.p, Ap);
        for (int h = 0; h < P.H; ++h) {
            for (const FrontGroup& g : plan.lev_sgroups[h])
                hipLaunchKernelGGL(k_factor_small<T>, dim3(blocks(static_cast<long long>(nb) * g.count, kSmallWaves)),
                                   dim3(kSmallWaves * 64), small_lds<T>(g.tiles), s, P,
                                   plan.host.lev_ptr[h] + g.first, g.count, nb, g.tiles,
                                   Ap, wb.rs.p, wb.active.p, L, wb.y.p, Up, wb.fail.p);
            for (const FrontGroup& g : plan.lev_bgroups[h])
                hipLaunchKernelGGL(k_factor_big<T>, dim3(static_cast<unsigned>(static_cast<long long>(nb) * g.count)),
                                   dim3(kBigWaves * 64), static_cast<std::size_t>(g.tiles) * 256 * sizeof(T), s,
                                   P, plan.host.lev_ptr[h] + g.first, g.count, nb, Ap, wb.rs.p, wb.active.p, L, wb.y.p,
                                   Up, wb.fail.p);
            if (!lev_huge[h].t.empty()) huge_factor(h, nb, Ap, L, Up);
        }
    }

    // The level's fronts of more than 159 rows for every window, huge_chunk
    // windows at a time: tiles in HBM, panel steps with MFMA trailing
    // updates, the right-hand side row carried (the fused forward
    // substitution), L, y and the update matrices written as the LDS fronts
    // write them.
    void huge_factor(int h, int nb, T* Ap, T* L, T* Up) {
        const DPlan& P = plan.d;
        HugeLevel& hl = lev_huge[h];
        const int nh = static_cast<int>(hl.t.size());
        T* F = reinterpret_cast<T*>(HF.p);
        T* D = reinterpret_cast<T*>(HD.p);
        for (int i0 = 0; i0 < nb; i0 += huge_chunk) {
            const int ni = std::min(huge_chunk, nb - i0);
            HugeInst hi;
            hi.nh = nh, hi.inst0 = i0, hi.ftiles = hl.tiles;
            hi.lstride = P.lsize, hi.ustride = P.usizeA, hi.astride = P.nent, hi.nstride = P.n;
            hi.active = wb.active.p, hi.fail = wb.fail.p, hi.fail_stride = 1;
            const unsigned slots = static_cast<unsigned>(nh * ni);
            const long long zero = hl.tiles * ni * 256;
            hipLaunchKernelGGL(k_h_zero<T>, dim3(blocks(zero, 256)), dim3(256), 0, s, F, zero);
            hipLaunchKernelGGL(k_h_init<T>, dim3(4, slots), dim3(256), 0, s, P, hi, hl.dt.p, hl.doff.p, F,
                               static_cast<const T*>(Ap), static_cast<const double*>(wb.rs.p));
            for (std::size_t r = 0; r < hl.chf.size(); ++r)
                hipLaunchKernelGGL(k_h_extend<T>, dim3(static_cast<unsigned>(hl.cbk[r].back()), ni), dim3(256), 0, s, P,
                                   hi, static_cast<int>(hl.chf[r].size()), hl.dchf[r].p, hl.dchc[r].p, hl.dcb[r].p,
                                   hl.dt.p, hl.doff.p, F, static_cast<const T*>(Up));
            for (int Kt = 0; Kt < hl.maxKt; ++Kt) {
                hipLaunchKernelGGL(k_h_panel<T>, dim3(hl.rb[Kt], slots), dim3(256), 0, s, P, hi, Kt, hl.dt.p, hl.doff.p,
                                   F, D);
                hipLaunchKernelGGL(k_h_diag<T>, dim3(slots), dim3(256), 0, s, P, hi, Kt, hl.dt.p, hl.doff.p, F,
                                   static_cast<const T*>(D));
                if (hl.tb[Kt] > 0)
                    hipLaunchKernelGGL(k_h_trailing<T>, dim3(hl.tb[Kt], slots), dim3(256), 0, s, P, hi, Kt, hl.dt.p,
                                       hl.doff.p, F);
            }
            hipLaunchKernelGGL(k_h_store<T>, dim3(static_cast<unsigned>(std::min<long long>(hl.store_blocks, 16)), slots),
                               dim3(256), 0, s, P, hi, hl.dt.p, hl.doff.p, static_cast<const T*>(F), L, Up, wb.y.p);
        }
    }

    void triangular(int nb, int mode) {
        const DPlan& P = plan.d;
        const T* L = reinterpret_cast<const T*>(wb.L.p);
        auto grid = [&](int h) {
            return dim3(static_cast<unsigned>(static_cast<long long>(nb) * plan.lev_fit_host[h]));
        };
        auto quadg = [&](int h) {
            return dim3((static_cast<unsigned>(nb) + 15) / 16,
                        static_cast<unsigned>(plan.host.lev_ptr[h + 1] - plan.host.lev_ptr[h]));
        };
        if (mode == 1)
            for (int h = 0; h < P.H; ++h) {
                if (const int M = quad_fwd[h]) {
                    auto kern = M == 32   ? k_forward_quads<T, 32>
                                : M == 48 ? k_forward_quads<T, 48>
                                : M == 56 ? k_forward_quads<T, 56>
                                          : k_forward_quads<T, 64>;
                    hipLaunchKernelGGL(kern, quadg(h), dim3(64), 0, s, P, h, nb, P.lsize, wb.active.p, wb.done.p,
                                       L, wb.rs.p, wb.y.p, V.p);
                    continue;
                }
                const bool small = plan.lev_mmax[h] <= 64;  // rows on lanes, L in registers: the front's vector in LDS
                auto kern = small ? k_forward_level<T, 1> : k_forward_level<T, 3>;
                if (plan.lev_fit_host[h] > 0)
                    hipLaunchKernelGGL(kern, grid(h), dim3(64), (small ? 64 : 192) * sizeof(double),
                                       s, P, h, nb, P.lsize,
                                       wb.active.p, wb.done.p, L, wb.rs.p, wb.y.p, V.p);
                const HugeLevel& hl = lev_huge[h];
                if (!hl.t.empty())
                    hipLaunchKernelGGL((k_h_forward<T, 1>), dim3(static_cast<unsigned>(hl.t.size()), nb), dim3(256),
                                       (256 + static_cast<std::size_t>(hl.mmax)) * sizeof(double), s, P, hl.dt.p, P.n,
                                       nb, P.vsize, P.lsize, L, static_cast<const double*>(wb.rs.p), wb.y.p, V.p,
                                       static_cast<const int*>(wb.done.p), static_cast<const int*>(wb.active.p));
            }
        for (int h = P.H - 1; h >= 0; --h) {
            if (const int M = quad_bwd[h]) {
                auto kern = M == 32   ? k_backward_quads<T, 32>
                            : M == 48 ? k_backward_quads<T, 48>
                            : M == 56 ? k_backward_quads<T, 56>
                                      : k_backward_quads<T, 64>;
                hipLaunchKernelGGL(kern, quadg(h), dim3(64), 0, s, P, h, nb, P.lsize, wb.active.p, wb.done.p, L,
                                   wb.y.p, wb.z.p);
                continue;
            }
            const HugeLevel& hl = lev_huge[h];
            if (!hl.t.empty())
                hipLaunchKernelGGL((k_h_backward<T, 1>), dim3(static_cast<unsigned>(hl.t.size()), nb), dim3(256),
                                   (256 + 16 + static_cast<std::size_t>(hl.mmax)) * sizeof(double), s, P, hl.dt.p, P.n,
                                   nb, P.lsize, L, static_cast<const double*>(wb.y.p), wb.z.p,
                                   static_cast<const int*>(wb.done.p), static_cast<const int*>(wb.active.p));
            const bool small = plan.lev_mmax[h] <= 64;  // update rows in 64 doubles of LDS
            auto kern = small ? k_backward_level<T, 1> : k_backward_level<T, 3>;
            for (const FrontGroup& g : plan.lev_lgroups[h]) {
                const std::size_t lds = solve_lds<T>(g.tiles) - (small ? 128 * sizeof(double) : 0);
                hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(static_cast<long long>(nb) * g.count)), dim3(64),
                                   lds, s, P, h, g.first, g.count, nb, P.lsize, wb.active.p, wb.done.p, L, wb.y.p,
                                   wb.z.p);
            }
        }
        // x takes z in the residual that follows every sweep (k_residual)
    }

    // Device bytes of one window's batch buffers with L entries of `elem` bytes.
    long long window_bytes(int elem) const {
        const DPlan& P = plan.d;
        const long long W2 = static_cast<long long>(P.W) * P.W;
        return W2 * 25 + static_cast<long long>(P.n) * 56 + 48 +
               (P.lsize + P.usizeA + std::max(P.nent, 1)) * elem + std::max<long long>(P.vsize, 1) * 8;
    }
    static constexpr long long kBufferSlack = 24 * 256 + 64;  // alignment of the batch buffers
    static constexpr int kArenaFallback = 8;                  // FP64 fallback windows at once in the scratch

    // The largest batch (at most `want`) whose buffers fit what is left of
    // the active arena once the plan is in it; with FP32 fronts, room is left
    // for the FP64 fallback (its plan and kArenaFallback windows).
    int arena_batch(int want) const {
        const Arena* a = active_arena();
        long long left = static_cast<long long>(a->size) - static_cast<long long>(arena_round(a->used));
        if (std::is_same<T, float>::value)  // the fallback's plan, its FP64 large-front arena, its windows
            left -= 3 * plan_bytes + kBufferSlack + kArenaFallback * window_bytes(8) + 4 * 8;
        const long long fit = (left - kBufferSlack - 4 * static_cast<long long>(sizeof(T))) / window_bytes(sizeof(T));
        if (fit < 1)
            throw std::invalid_argument("one window of radius " + std::to_string(plan.d.D) +
                                        " needs more device memory than the static scratch of " +
                                        std::to_string(a->size >> 20) + " MiB holds");
        return static_cast<int>(std::min<long long>(want, fit));
    }
    long long plan_bytes = 0;  // arena bytes the plan took (the arena's)

    void fallback_fp64(const std::vector<int>& bad, double tol, int maxit, hipEvent_t* ev, std::vector<double>& rn,
                       std::vector<int>& its, std::vector<int>& fl, IrStats& st);

    void solve(int nb, double tol, int maxit, IrStats& st, hipEvent_t* ev, bool timing) {
        const DPlan& P = plan.d;
        const bool gr = graphs_enabled();
        if (tol != residual_tol) g_residual.clear(), residual_tol = tol;  // the residual graph holds tol
        if (timing) LSC_HIP(hipEventRecord(ev[0], s));
        g_flood.run(s, nb, gr, [&] {
            reset(wb.done.p, nb, s);
            reset(wb.iters.p, nb, s);
            reset(wb.fail.p, nb, s);
            const double shift = 10 * std::numeric_limits<double>::epsilon();
            if (P.W <= kMaxW)
                hipLaunchKernelGGL(k_flood, dim3(nb), dim3(kThreads), 0, s, P, nb, wb.g.p, wb.src.p, wb.active.p,
                                   wb.present.p, wb.dg.p, wb.sc.p, wb.b.p, wb.rs.p, wb.x.p, shift);
            else
                hipLaunchKernelGGL(k_flood_wide, dim3(nb), dim3(kThreads), wide_flood_lds(P.W), s, P, nb,
                                   (P.W + 63) / 64, wb.g.p, wb.src.p, wb.active.p, wb.present.p, wb.dg.p, wb.sc.p,
                                   wb.b.p, wb.rs.p, wb.x.p, shift);
        });
        if (timing) LSC_HIP(hipEventRecord(ev[1], s));
        g_factor.run(s, nb, gr, [&] { factor(nb); });
        if (timing) LSC_HIP(hipEventRecord(ev[2], s));
        g_sweep0.run(s, nb, gr, [&] { triangular(nb, 0); });
        if (timing) LSC_HIP(hipEventRecord(ev[3], s));
        st.batches += 1;
        st.solve_launches += 1;
        time_phase(st.ms_flood, ev[0], ev[1], timing);
        time_phase(st.ms_factor, ev[1], ev[2], timing);
        time_phase(st.ms_solve, ev[2], ev[3], timing);
        for (int it = 0;; ++it) {
            if (timing) LSC_HIP(hipEventRecord(ev[0], s));
            g_residual.run(s, nb, gr, [&] {
                reset(open.p, 1, s);
                hipLaunchKernelGGL(k_residual, dim3(nb), dim3(kThreads), 0, s, P, nb, wb.g.p, wb.present.p,
                                   wb.dg.p, wb.sc.p, wb.b.p, wb.z.p, wb.x.p, wb.rs.p, wb.active.p, wb.done.p, wb.rnorm.p,
                                   wb.rprev.p, wb.iters.p, tol, open.p);
            });
            if (timing) LSC_HIP(hipEventRecord(ev[1], s));
            int h_open = 0;
            LSC_HIP(hipMemcpyAsync(&h_open, open.p, sizeof(int), hipMemcpyDeviceToHost, s));
            LSC_HIP(hipStreamSynchronize(s));
            time_phase(st.ms_residual, ev[0], ev[1], timing);
            if (h_open == 0 || it >= maxit) break;
            st.solve_launches += 1;
            if (timing) LSC_HIP(hipEventRecord(ev[0], s));
            g_sweep1.run(s, nb, gr, [&] { triangular(nb, 1); });
            if (timing) LSC_HIP(hipEventRecord(ev[1], s));
            time_phase(st.ms_solve, ev[0], ev[1], timing);
        }
        // statistics of the batch
        std::vector<int> act(nb), fl(nb), its(nb);
        std::vector<double> rn(nb);
        LSC_HIP(hipMemcpyAsync(act.data(), wb.active.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(fl.data(), wb.fail.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(its.data(), wb.iters.p, nb * sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(rn.data(), wb.rnorm.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipStreamSynchronize(s));
        if constexpr (std::is_same<T, float>::value) {
            // A window whose FP32 factor met a non-positive pivot, or whose refinement ended more than
            // kFallbackGap times above tol (failed, rather than stopped at the FP64 floor of its
            // residual), is solved again in FP64; its solution replaces the mixed one.
            std::vector<int> bad;
            for (int k = 0; k < nb; ++k)
                if (act[k] && (fl[k] || !(rn[k] <= kFallbackGap * tol))) bad.push_back(k);
            if (!bad.empty()) fallback_fp64(bad, tol, maxit, ev, rn, its, fl, st);
        }
        for (int k = 0; k < nb; ++k) {
            if (!act[k]) continue;
            if (fl[k]) throw std::runtime_error("an Omniscape window is not positive definite");
            ++st.solved;
            st.iter_sum += its[k];
            st.iter_max = std::max(st.iter_max, its[k]);
            st.resid_max = std::max(st.resid_max, rn[k]);
            st.resid_sum += rn[k];
            st.resids.push_back(static_cast<float>(rn[k]));
            if (!(rn[k] <= tol)) ++st.unconverged;
            if (!(rn[k] <= 1e-10)) ++st.over_bar;
        }
    }
};

template <typename T>
void WindowSolver<T>::fallback_fp64(const std::vector<int>& bad, double tol, int maxit, hipEvent_t* ev,
                                    std::vector<double>& rn, std::vector<int>& its, std::vector<int>& fl,
                                    IrStats& st) {
    const int W2 = plan.d.W * plan.d.W, n = plan.d.n;
    const int nbad = static_cast<int>(bad.size());
    // in chunks of at most the batch size (kArenaFallback windows in the static scratch)
    const int chunk = active_arena() ? kArenaFallback : Bcap;
    for (int c0 = 0; c0 < nbad; c0 += chunk) {
        const int m = std::min(chunk, nbad - c0);
        if (!fb || fb_cap < m) {
            fb_cap = active_arena() ? chunk : std::max(m, std::min(Bcap, 256));
            fb.reset(new WindowSolver<double>(s));
            fb->set_plan(Plan(plan.host), fb_cap);
        }
        auto& w = fb->wb;
        for (int j = 0; j < m; ++j) {
            const std::size_t k = static_cast<std::size_t>(bad[c0 + j]);
            LSC_HIP(hipMemcpyAsync(w.g.p + static_cast<std::size_t>(j) * W2, wb.g.p + k * W2, W2 * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
            LSC_HIP(hipMemcpyAsync(w.src.p + static_cast<std::size_t>(j) * W2, wb.src.p + k * W2,
                                   W2 * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        const std::vector<int> ones(m, 1);
        LSC_HIP(hipMemcpyAsync(w.active.p, ones.data(), m * sizeof(int), hipMemcpyHostToDevice, s));
        IrStats fst;  // the windows are counted once, by the caller
        fb->solve(m, tol, maxit, fst, ev, false);
        std::vector<int> fact(m), fits(m);
        std::vector<double> frn(m);
        LSC_HIP(hipMemcpyAsync(fact.data(), w.active.p, m * sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(fits.data(), w.iters.p, m * sizeof(int), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(frn.data(), w.rnorm.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
        for (int j = 0; j < m; ++j) {
            const std::size_t k = static_cast<std::size_t>(bad[c0 + j]);
            LSC_HIP(hipMemcpyAsync(wb.x.p + k * n, w.x.p + static_cast<std::size_t>(j) * n, n * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
        }
        LSC_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < m; ++j) {
            const int k = bad[c0 + j];
            // the same window floods to the same component: it stays active
            if (!fact[j]) throw std::logic_error("a window's FP64 solve dropped it");
            rn[k] = frn[j], its[k] += fits[j], fl[k] = 0;
        }
        st.fallback += m;
        st.solve_launches += fst.solve_launches;
    }
}

struct OmniArgs {
    int radius, buffer, block, four, rc, correct, fp, precision;
    double tol;
    int maxit;
    long long batch;
    int timing;
};

namespace {

Lattice make_lattice(long long nx, long long ny, const OmniArgs& a) {
    Lattice L{};
    L.nx = nx, L.ny = ny;
    L.R = a.radius, L.D = a.radius + a.buffer, L.W = 2 * L.D + 1;
    L.bs = a.block, L.br = (a.block - 1) / 2;
    const long long nlx = nx >= a.block ? (nx - a.block) / a.block + 1 : 0;
    const long long nly = ny >= a.block ? (ny - a.block) / a.block + 1 : 0;
    L.nly = static_cast<int>(nly);
    L.nlat = nlx * nly;
    L.rc = a.rc;
    return L;
}

// The correction circuits of omniscape_correction (reference
// omniscape.cpp:326-392): conductance 1 within D, a block-source and a
// null-source circuit; their currents give the factors.
template <typename T>
std::vector<double> correction(const OmniArgs& a, long long nx, long long ny, hipStream_t s, IrStats& st) {
    const long long R = a.radius, D = a.radius + a.buffer, W = 2 * D + 1, bs = a.block, br = (bs - 1) / 2;
    const long long n = W * W;
    auto within = [&](long long c, long long rad) {
        const long long dy = c / W - D, dx = c % W - D;
        return dy * dy + dx * dx <= rad * rad;
    };
    // the target block around the centre, as the reference clips it: at the
    // window and at the map's last row and column
    auto in_block = [&](long long c) {
        const long long y = c / W, x = c % W;
        return y >= D - br && y <= std::min(D + br, ny - 1) && x >= D - br && x <= std::min(D + br, nx - 1);
    };
    long long n_src = 0, n_block = 0;  // sources within R; those outside the block
    for (long long c = 0; c < n; ++c)
        if (within(c, R)) ++n_src, n_block += !in_block(c);
    // every source 1 / (n_src - 1) except at the ground (null circuit); the
    // block's amps bs^2 shared by the sources outside it (block circuit)
    const double null_amps = 1 / static_cast<double>(n_src - 1);
    const double block_amps = static_cast<double>(bs * bs) / static_cast<double>(n_block ? n_block : 1);
    std::vector<double> g(n), block_src(n, 0.0), null_src(n, 0.0);
    for (long long c = 0; c < n; ++c) {
        g[c] = within(c, D) ? 1.0 : -9999.0;
        if (!within(c, R)) continue;
        if (c != D * W + D) null_src[c] = null_amps;
        if (!in_block(c)) block_src[c] = block_amps;
    }

    WindowSolver<T> ws(s);
    ws.set_plan(window_plan(static_cast<int>(D), true, a.four != 0), 2);
    ws.wb.g.upload(g.data(), n, s);
    LSC_HIP(hipMemcpyAsync(ws.wb.g.p + n, g.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    LSC_HIP(hipMemcpyAsync(ws.wb.src.p, block_src.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    LSC_HIP(hipMemcpyAsync(ws.wb.src.p + n, null_src.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    const int one[2] = {1, 1};
    LSC_HIP(hipMemcpyAsync(ws.wb.active.p, one, sizeof(one), hipMemcpyHostToDevice, s));
    hipEvent_t ev[4] = {};
    IrStats cst;
    ws.solve(2, std::min(a.tol, 1e-14), std::max(a.maxit, 30), cst, ev, false);
    Lattice Lt{};
    hipLaunchKernelGGL(k_currents, dim3(2), dim3(kCurThreads), currents_lds(static_cast<int>(W)), s,
                       ws.plan.d, Lt, 0ll, 2, ws.wb.g.p, ws.wb.present.p, ws.wb.x.p, ws.wb.active.p,
                       static_cast<const double*>(nullptr), ws.wb.cur.p, W > kMaxW ? ws.wb.src.p : nullptr);
    std::vector<double> cur(2 * n);
    std::vector<int> act(2);
    LSC_HIP(hipMemcpyAsync(cur.data(), ws.wb.cur.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipMemcpyAsync(act.data(), ws.wb.active.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipStreamSynchronize(s));
    std::vector<double> bc(cur.begin(), cur.begin() + n), nc(cur.begin() + n, cur.end());
    if (!act[0]) std::fill(bc.begin(), bc.end(), 0.0);
    if (!act[1]) std::fill(nc.begin(), nc.end(), 0.0);
    (void)st;
    // factor = (null currents summed over the bs x bs block of targets
    // centred on the cell) / (block circuit's current), 1 where the block
    // current is within 1e-6 of its maximum from 0; the block sum adds its
    // terms by block row, then column, as the reference's shifted sums do
    const double floor = 1e-6 * *std::max_element(bc.begin(), bc.end());
    std::vector<double> corr(n);
    for (long long c = 0; c < n; ++c) {
        const double b = bc[c];
        if (b <= floor) {
            corr[c] = 1.0;
            continue;
        }
        const long long y = c / W, x = c % W;
        double t = 0.0;
        for (long long i = 0; i < bs; ++i) {
            const long long yy = y + br - i;
            if (yy < 0 || yy >= W) continue;
            for (long long j = 0; j < bs; ++j) {
                const long long xx = x + br - j;
                if (xx >= 0 && xx < W) t += nc[yy * W + xx];
            }
        }
        corr[c] = (t == 0.0 ? 1.0 : t) / b;
    }
    return corr;
}

// One pass over the target lattice: `kind` 0 the landscape circuits, 1 the
// flow potential's (conductance 1 on the whole clipped window).
template <typename T>
void lattice_pass(const double* d_res, const double* d_src, const Lattice& Lt, const OmniArgs& a, const CondArgs& cond,
                  int kind, const double* d_corr, double* d_map, unsigned long long* d_ntargets, hipStream_t s,
                  IrStats& st) {
    if (Lt.nlat == 0) return;
    Plan plan = window_plan(Lt.D, kind == 0, a.four != 0, nd_leaf());
    // per-window device memory: maps, vectors, L, update matrices and vectors
    long long ua = 0;
    for (int t = 0; t < plan.nfront; ++t) {
        const long long u1 = plan.nupd[t] + 1;
        ua += (u1 * (u1 + 1) / 2 + 3) & ~3ll;
    }
    const double per_window = static_cast<double>(Lt.W) * Lt.W * 33.0 + plan.n * 7.0 * 8.0 +
                              static_cast<double>(plan.lsize + ua + plan.fe_row.size()) * sizeof(T) +
                              plan.vsize * 8.0;
    std::size_t free_b = 0, total_b = 0;
    if (!active_arena()) LSC_HIP(hipMemGetInfo(&free_b, &total_b));
    // in the static scratch the solver sizes the batch to what is left of it
    const long long fit = active_arena() ? std::numeric_limits<int>::max()
                                         : std::max<long long>(1, static_cast<long long>(0.5 * static_cast<double>(free_b) / per_window));
    const long long want = std::min<long long>(a.batch > 0 ? a.batch : 16384, Lt.nlat);
    // Several batches in flight on their own streams (three by default,
    // LSC_BATCH_STREAMS), each driven by a host thread: the window kernels
    // are latency-bound (one front per wavefront, the factor limited by LDS),
    // so other batches fill issue slots one batch leaves idle. The
    // cumulative map is added to in batch order (events), so the result is
    // the same as with one stream, bit for bit.
    const int nsol = static_cast<int>(std::max<long long>(1, std::min<long long>({batch_streams(), fit,
                                                                                  (Lt.nlat + want - 1) / want})));
    // the batches in flight share the one-batch budget: allocating (and
    // freeing) HBM costs about 13 ms per GB, more than extra batches gain
    long long B = std::max<long long>(1, std::min<long long>(want, fit) / nsol);
    // The first batch runs on the caller's stream, the others on streams of
    // their own that start after the work already queued there; the host
    // waits for all of them below. (A stream of our own for the first batch
    // too made five streams on the box's four hardware queues: 109k instead
    // of 117k circuits/s on the 4096^2 map.) The caller's stream is often
    // the null stream, which cannot be captured: its batches launch eagerly.
    std::vector<std::unique_ptr<WindowSolver<T>>> ws;
    std::vector<hipStream_t> own(nsol - 1, nullptr);
    for (auto& x : own) LSC_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    hipEvent_t queued = nullptr;
    std::vector<hipEvent_t> acc_ev(nsol, nullptr);
    if (nsol > 1) {  // the other streams start after the caller's queued work; accumulation in batch order
        LSC_HIP(hipEventCreateWithFlags(&queued, hipEventDisableTiming));
        LSC_HIP(hipEventRecord(queued, s));
        for (auto& x : own) LSC_HIP(hipStreamWaitEvent(x, queued, 0));
    }
    for (int j = 0; j < nsol; ++j) {
        ws.emplace_back(new WindowSolver<T>(j == 0 ? s : own[j - 1]));
        ws.back()->set_plan(j == 0 ? plan : Plan(plan), static_cast<int>(B));
        if (nsol > 1) LSC_HIP(hipEventCreateWithFlags(&acc_ev[j], hipEventDisableTiming));
    }
    B = ws[0]->Bcap;  // the static scratch may hold fewer windows
    const long long nbatch = (Lt.nlat + B - 1) / B;
    std::vector<IrStats> part(nsol);
    std::vector<std::exception_ptr> err(nsol);
    std::mutex mu;
    std::condition_variable cv;
    long long acc_done = -1;  // last batch whose accumulation is enqueued
    bool failed = false;
    const bool timing = a.timing != 0;
    int dev = 0;
    LSC_HIP(hipGetDevice(&dev));
    auto run = [&](int j) {
        WindowSolver<T>& w = *ws[j];
        const hipStream_t sj = w.s;
        hipEvent_t ev[4] = {};
        try {
            LSC_HIP(hipSetDevice(dev));  // the helper thread works on the caller's device
            if (timing)
                for (auto& e : ev) LSC_HIP(hipEventCreate(&e));
            for (long long bi = j; bi < nbatch; bi += nsol) {
                const long long k0 = bi * B;
                const int nb = static_cast<int>(std::min<long long>(B, Lt.nlat - k0));
                if (timing) LSC_HIP(hipEventRecord(ev[0], sj));
                hipLaunchKernelGGL(k_fill, dim3(nb), dim3(kThreads), 0, sj, d_res, d_src, Lt, k0, nb, cond, w.wb.g.p,
                                   w.wb.src.p, w.wb.active.p, d_ntargets);
                if (kind == 1)  // conductance 1 on every on-map cell of the box
                    hipLaunchKernelGGL(k_unit_conductance, dim3(nb), dim3(kThreads), 0, sj, Lt, k0, nb, w.wb.g.p);
                if (timing) {
                    LSC_HIP(hipEventRecord(ev[1], sj));
                    w.time_phase(part[j].ms_fill, ev[0], ev[1], true);
                }
                w.solve(nb, a.tol, a.maxit, part[j], ev, timing);
                if (timing) LSC_HIP(hipEventRecord(ev[0], sj));
                // wider windows keep their voltages in the batch's source maps (no longer needed)
                hipLaunchKernelGGL(k_currents, dim3(nb), dim3(kCurThreads), currents_lds(Lt.W), sj, w.plan.d, Lt, k0,
                                   nb, w.wb.g.p, w.wb.present.p, w.wb.x.p, w.wb.active.p, d_corr, w.wb.cur.p,
                                   Lt.W > kMaxW ? w.wb.src.p : nullptr);
                if (timing) LSC_HIP(hipEventRecord(ev[1], sj));
                const long long ix0 = k0 / Lt.nly, ix1 = (k0 + nb - 1) / Lt.nly;
                const long long X0 = std::max<long long>(0, Lt.br + ix0 * Lt.bs - Lt.D);
                const long long X1 = std::min<long long>(Lt.nx - 1, Lt.br + ix1 * Lt.bs + Lt.D);
                const long long ncol = X1 - X0 + 1;
                const long long cells = ncol * Lt.ny;
                if (nsol > 1 && bi > 0) {  // after the previous batch's accumulation
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return acc_done >= bi - 1 || failed; });
                    if (failed) break;
                    lk.unlock();
                    std::lock_guard<std::mutex> cl(capture_mutex());
                    LSC_HIP(hipStreamWaitEvent(sj, acc_ev[(j + nsol - 1) % nsol], 0));
                }
                hipLaunchKernelGGL(k_accumulate, dim3(static_cast<unsigned>((cells + kThreads - 1) / kThreads)),
                                   dim3(kThreads), 0, sj, Lt, k0, nb, X0, ncol, w.wb.active.p, w.wb.cur.p, d_map);
                if (nsol > 1) LSC_HIP(hipEventRecord(acc_ev[j], sj));
                {
                    std::lock_guard<std::mutex> lk(mu);
                    acc_done = bi;
                }
                cv.notify_all();
                if (timing) {
                    LSC_HIP(hipEventRecord(ev[2], sj));
                    w.time_phase(part[j].ms_currents, ev[0], ev[1], true);
                    w.time_phase(part[j].ms_accum, ev[1], ev[2], true);
                }
            }
            LSC_HIP(hipStreamSynchronize(sj));
        } catch (...) {
            err[j] = std::current_exception();
            {
                std::lock_guard<std::mutex> lk(mu);
                failed = true;
            }
            cv.notify_all();
            (void)hipStreamSynchronize(sj);
        }
        if (timing)
            for (auto& e : ev)
                if (e) (void)hipEventDestroy(e);
    };
    std::vector<std::thread> helpers;
    for (int j = 1; j < nsol; ++j) helpers.emplace_back(run, j);
    run(0);
    for (auto& h : helpers) h.join();
    for (auto& e : acc_ev)
        if (e) (void)hipEventDestroy(e);
    ws.clear();
    if (queued) (void)hipEventDestroy(queued);
    for (auto& x : own) (void)hipStreamDestroy(x);
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    st.streams = std::max(st.streams, nsol);
    for (const IrStats& q : part) {
        st.solved += q.solved, st.unconverged += q.unconverged, st.over_bar += q.over_bar;
        st.iter_sum += q.iter_sum, st.solve_launches += q.solve_launches, st.batches += q.batches;
        st.iter_max = std::max(st.iter_max, q.iter_max);
        st.resid_max = std::max(st.resid_max, q.resid_max);
        st.resid_sum += q.resid_sum;
        st.resids.insert(st.resids.end(), q.resids.begin(), q.resids.end());
        st.ms_fill += q.ms_fill, st.ms_flood += q.ms_flood, st.ms_factor += q.ms_factor;
        st.ms_solve += q.ms_solve, st.ms_residual += q.ms_residual, st.ms_currents += q.ms_currents;
        st.ms_accum += q.ms_accum;
        st.fallback += q.fallback;
    }
}

}  // namespace

// The condition value of every lattice point that is a target (a positive
// block source sum, NaN skipped, column-major): the median ("within", the two
// middle values halved separately for an even count) or the mode ("equal",
// values counted by bit pattern, the first to pass the running top count
// wins) of the block's future values in column-major order, NaN left out
// (reference omniscape.cpp:168-215); NaN for the other lattice points.
std::vector<double> condition_values(const double* source, const double* future, const Lattice& L, bool equal) {
    std::vector<double> out(L.nlat, std::numeric_limits<double>::quiet_NaN());
    std::vector<double> v;
    std::vector<std::pair<std::uint64_t, int>> count;
    for (long long kk = 0; kk < L.nlat; ++kk) {
        const long long x = L.br + (kk / L.nly) * L.bs, y = L.br + (kk % L.nly) * L.bs;
        double amps = 0.0;
        for (long long xx = x - L.br; xx <= x + L.br; ++xx)
            for (long long yy = y - L.br; yy <= y + L.br; ++yy) {
                const double q = source[yy * L.nx + xx];
                if (!std::isnan(q)) amps += q;
            }
        if (!(amps > 0)) continue;
        v.clear();
        for (long long xx = x - L.br; xx <= std::min(x + L.br, L.nx - 1); ++xx)
            for (long long yy = y - L.br; yy <= std::min(y + L.br, L.ny - 1); ++yy) {
                const double q = future[yy * L.nx + xx];
                if (!std::isnan(q)) v.push_back(q);
            }
        if (v.empty()) throw std::runtime_error("a target block has no value of a condition (all NoData)");
        if (!equal) {
            std::sort(v.begin(), v.end());
            const std::size_t n = v.size();
            out[kk] = n % 2 ? v[n / 2] : v[n / 2 - 1] / 2 + v[n / 2] / 2;
            continue;
        }
        count.clear();
        double best = v[0];
        int top = 0;
        for (double q : v) {
            std::uint64_t bits;
            std::memcpy(&bits, &q, sizeof bits);
            auto it = std::find_if(count.begin(), count.end(), [&](const auto& e) { return e.first == bits; });
            if (it == count.end()) it = count.insert(count.end(), {bits, 0});
            if (++it->second > top) top = it->second, best = q;
        }
        out[kk] = best;
    }
    return out;
}

// The device form of a run's conditions; `on_device`: the maps (and source)
// are device pointers, else host pointers.
struct Conditions {
    std::vector<DevBuf<double>> bufs;
    CondArgs args{};
};

void prepare_conditions(const lsc_omniscape_condition* c, bool on_device, const double* source, const Lattice& L,
                        hipStream_t s, Conditions& out) {
    out.args = CondArgs{};
    if (!c) return;
    const std::size_t n = static_cast<std::size_t>(L.nx) * L.ny;
    auto host_of = [&](const double* p) {
        std::vector<double> h(n);
        if (n) LSC_HIP(hipMemcpyAsync(h.data(), p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipStreamSynchronize(s));
        return h;
    };
    std::vector<double> src_h;
    if (on_device) src_h = host_of(source);
    const double* hs = on_device ? src_h.data() : source;
    out.bufs.reserve(4);
    for (int k = 0; k < 2; ++k) {
        if (!c[k].present) continue;
        const double* fut = c[k].future ? c[k].future : c[k].present;
        std::vector<double> fut_h;
        if (on_device) fut_h = host_of(fut);
        const std::vector<double> val = condition_values(hs, on_device ? fut_h.data() : fut, L, c[k].equal != 0);
        out.bufs.push_back(to_device(val, s));
        const double* vp = out.bufs.back().p;
        const double* pp = c[k].present;
        if (!on_device) {
            DevBuf<double> d(n ? n : 1);
            if (n) d.upload(c[k].present, n, s);
            out.bufs.push_back(std::move(d));
            pp = out.bufs.back().p;
        }
        const int i = out.args.n++;
        out.args.present[i] = pp;
        out.args.value[i] = vp;
        out.args.equal[i] = c[k].equal != 0;
        out.args.lower[i] = c[k].lower;
        out.args.upper[i] = c[k].upper;
    }
}

template <typename T>
void omniscape_device(const double* d_res, const double* d_src, long long nx, long long ny, const OmniArgs& a,
                      const CondArgs& cond, double* d_cur, double* d_fp, long long* n_targets, IrStats& st,
                      hipStream_t s) {
    if (a.radius < 0 || a.buffer < 0 || a.block < 1 || a.block % 2 == 0)
        throw std::invalid_argument("radius and buffer must be >= 0 and block_size odd");
    if (2 * (a.radius + a.buffer) + 1 > kMaxWideW)
        throw std::invalid_argument("window side 2 (radius + buffer) + 1 exceeds " + std::to_string(kMaxWideW));
    const Lattice Lt = make_lattice(nx, ny, a);
    const long long ncell = nx * ny;
    const unsigned zb = static_cast<unsigned>((ncell + 255) / 256);
    if (ncell > 0) hipLaunchKernelGGL(k_zero_f64, dim3(zb), dim3(256), 0, s, ncell, d_cur);
    if (a.fp && ncell > 0) hipLaunchKernelGGL(k_zero_f64, dim3(zb), dim3(256), 0, s, ncell, d_fp);
    DevBuf<unsigned long long> cnt(2);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, 2, cnt.p);
    DevBuf<double> corr;
    const bool correct = a.correct && a.block > 1;
    if (correct && Lt.nlat > 0) {
        IrStats cst;
        std::vector<double> c;
        {
            ArenaMark m;  // the correction's circuits give their scratch back
            c = correction<T>(a, nx, ny, s, cst);
        }
        corr = to_device(c, s);
    }
    {
        ArenaMark m;
        lattice_pass<T>(d_res, d_src, Lt, a, cond, 0, corr.p, d_cur, cnt.p, s, st);
    }
    if (a.fp) {
        IrStats fst;
        ArenaMark m;
        lattice_pass<T>(d_res, d_src, Lt, a, cond, 1, corr.p, d_fp, cnt.p + 1, s, fst);
    }
    unsigned long long h = 0;
    LSC_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    LSC_HIP(hipStreamSynchronize(s));
    *n_targets = static_cast<long long>(h);
}

// ---- the static scratch of the allocation-free entry point

// Device memory of lsc_circuit_omniscape_device_static: a file-scope array
// each GPU holds from the library's load until the process ends. Host code
// never uses its name as an address (it is a host-side stand-in): the
// address on each device comes from hipGetSymbolAddress.
constexpr std::size_t kScratchBytes = std::size_t{128} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};
std::mutex g_device_mu[kMaxDevices];  // one run at a time per device: they share its scratch
}  // namespace

// The scratch's address on the current device, looked up once per device
// (with that device current).
char* scratch_on_current_device(int& dev) {
    LSC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        LSC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("the static scratch has no device address");
        g_scratch[dev] = static_cast<char*>(p);
    }
    return g_scratch[dev];
}

std::mutex& device_run_mutex(int dev) { return g_device_mu[dev]; }

// Makes `a` the calling thread's arena for the scope.
struct ArenaScope {
    explicit ArenaScope(Arena* a) { active_arena() = a; }
    ~ArenaScope() { active_arena() = nullptr; }
    ArenaScope(const ArenaScope&) = delete;
    ArenaScope& operator=(const ArenaScope&) = delete;
};

}  // namespace lsc

namespace {

lsc::OmniArgs omni_args(int radius, int buffer, int block_size, int four, int rc, int correct, int fp,
                        const lsc_solve_options* opt) {
    lsc::OmniArgs a{};
    a.radius = radius, a.buffer = buffer, a.block = block_size, a.four = four, a.rc = rc, a.correct = correct,
    a.fp = fp;
    a.precision = opt ? opt->precision : LSC_PRECISION_MIXED;
    if (a.precision == LSC_PRECISION_AUTO) a.precision = LSC_PRECISION_MIXED;  // windows: mixed
    a.tol = opt && opt->tol > 0 ? opt->tol : 1e-12;
    a.maxit = opt && opt->max_refine > 0 ? opt->max_refine : 30;
    a.batch = opt && opt->batch > 0 ? opt->batch : 16384;
    a.timing = opt ? opt->timing : 0;
    return a;
}

void fill_stats(const lsc::IrStats& st, long long n_targets, lsc_solve_stats* out) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->n_targets = n_targets;
    out->n_solved = st.solved;
    out->n_unconverged = st.unconverged;
    out->n_over_1e10 = st.over_bar;
    out->refine_max = st.iter_max;
    out->n_batches = st.batches;
    out->n_solve_launches = st.solve_launches;
    out->n_factored = st.solved;
    out->refine_mean = st.solved ? static_cast<double>(st.iter_sum) / st.solved : 0.0;
    out->resid_max = st.resid_max;
    out->resid_mean = st.solved ? st.resid_sum / st.solved : 0.0;
    out->resid_median = st.resid_median();
    out->ms_fill = st.ms_fill, out->ms_flood = st.ms_flood, out->ms_factor = st.ms_factor;
    out->ms_solve = st.ms_solve, out->ms_residual = st.ms_residual, out->ms_currents = st.ms_currents;
    out->ms_accumulate = st.ms_accum;
    out->n_streams = st.streams;
    out->n_fallback = st.fallback;
}

// The device entry with the condition maps on the host (`cond_on_device`
// false) or on the device.
int omniscape_run(const double* resistance, const double* source, const double* host_source, int64_t nx, int64_t ny,
                  int radius, int buffer, int block_size, int four_neighbors, int resistance_is_conductance,
                  int correct_artifacts, int flow_potential, const lsc_omniscape_condition* conditions,
                  bool cond_on_device, const lsc_solve_options* options, double* current, double* flow_potential_out,
                  int64_t* n_targets, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        const lsc::OmniArgs a = omni_args(radius, buffer, block_size, four_neighbors, resistance_is_conductance,
                                          correct_artifacts, flow_potential, options);
        hipStream_t s = options ? reinterpret_cast<hipStream_t>(options->stream) : nullptr;
        lsc::Conditions cond;
        if (conditions) {
            if (a.block < 1 || a.block % 2 == 0) throw std::invalid_argument("block_size must be odd");
            lsc::prepare_conditions(conditions, cond_on_device, cond_on_device ? source : host_source,
                                    lsc::make_lattice(nx, ny, a), s, cond);
        }
        lsc::IrStats st;
        long long nt = 0;
        if (a.precision == LSC_PRECISION_FP64)
            lsc::omniscape_device<double>(resistance, source, nx, ny, a, cond.args, current, flow_potential_out, &nt,
                                          st, s);
        else
            lsc::omniscape_device<float>(resistance, source, nx, ny, a, cond.args, current, flow_potential_out, &nt,
                                         st, s);
        if (n_targets) *n_targets = nt;
        fill_stats(st, nt, stats);
    });
}

}  // namespace

extern "C" {

int lsc_circuit_omniscape_device(const double* resistance, const double* source, int64_t nx, int64_t ny, int radius,
                                 int buffer, int block_size, int four_neighbors, int resistance_is_conductance,
                                 int correct_artifacts, int flow_potential, const lsc_omniscape_condition* conditions,
                                 const lsc_solve_options* options, double* current, double* flow_potential_out,
                                 int64_t* n_targets, lsc_solve_stats* stats) {
    return omniscape_run(resistance, source, nullptr, nx, ny, radius, buffer, block_size, four_neighbors,
                         resistance_is_conductance, correct_artifacts, flow_potential, conditions, true, options,
                         current, flow_potential_out, n_targets, stats);
}

int lsc_circuit_omniscape_device_static(const double* resistance, const double* source, int64_t nx, int64_t ny,
                                        int radius, int buffer, int block_size, int four_neighbors,
                                        int resistance_is_conductance, int correct_artifacts, int flow_potential,
                                        const lsc_omniscape_condition* conditions, const lsc_solve_options* options,
                                        double* current, double* flow_potential_out, int64_t* n_targets,
                                        lsc_solve_stats* stats) {
    int inner = 0;
    const int rc = lsc::guarded([&] {
        int dev = 0;
        char* base = lsc::scratch_on_current_device(dev);
        std::lock_guard<std::mutex> lk(lsc::device_run_mutex(dev));
        lsc::Arena arena{base, lsc::kScratchBytes, 0};
        lsc::ArenaScope scope(&arena);
        inner = omniscape_run(resistance, source, nullptr, nx, ny, radius, buffer, block_size, four_neighbors,
                              resistance_is_conductance, correct_artifacts, flow_potential, conditions, true, options,
                              current, flow_potential_out, n_targets, stats);
    });
    return inner != 0 ? inner : rc;
}

int lsc_circuit_omniscape(const double* resistance, const double* source, int64_t nx, int64_t ny, int radius,
                          int buffer, int block_size, int four_neighbors, int resistance_is_conductance,
                          int correct_artifacts, int flow_potential, const lsc_omniscape_condition* conditions,
                          const lsc_solve_options* options, double* current, double* flow_potential_out,
                          int64_t* n_targets, lsc_solve_stats* stats) {
    int inner = 0;  // the device run's status: an argument error stays status 1
    const int rc = lsc::guarded([&] {
        if (nx < 0 || ny < 0) throw std::invalid_argument("negative raster size");
        const std::size_t n = static_cast<std::size_t>(nx) * ny;
        hipStream_t s = options ? reinterpret_cast<hipStream_t>(options->stream) : nullptr;
        lsc::DevBuf<double> r(n ? n : 1), src(n ? n : 1), cur(n ? n : 1), fp(flow_potential && n ? n : 1);
        if (n) {
            r.upload(resistance, n, s);
            src.upload(source, n, s);
        }
        inner = omniscape_run(r.p, src.p, source, nx, ny, radius, buffer, block_size, four_neighbors,
                              resistance_is_conductance, correct_artifacts, flow_potential, conditions, false,
                              options, cur.p, fp.p, n_targets, stats);
        if (inner != 0) return;  // its message is set
        if (n) {
            LSC_HIP(hipMemcpyAsync(current, cur.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
            if (flow_potential)
                LSC_HIP(hipMemcpyAsync(flow_potential_out, fp.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        LSC_HIP(hipStreamSynchronize(s));
    });
    return inner != 0 ? inner : rc;
}

int lsc_window_plan_stats(int radius, int buffer, int four_neighbors, int disk, double* out) {
    return lsc::guarded([&] {
        if (radius < 0 || buffer < 0) throw std::invalid_argument("radius and buffer must be >= 0");
        const lsc::Plan p = lsc::window_plan(radius + buffer, disk != 0, four_neighbors != 0);
        double unpacked = 0;
        for (int t = 0; t < p.nfront; ++t) unpacked += static_cast<double>(p.npiv[t]) * (p.npiv[t] + p.nupd[t]);
        out[0] = p.n, out[1] = p.nfront, out[2] = static_cast<double>(p.lsize), out[3] = p.flops;
        out[4] = p.max_front, out[5] = static_cast<double>(p.lev_ptr.size() - 1), out[6] = unpacked;
    });
}

int lsc_circuit_omniscape_correction(int radius, int buffer, int block_size, int four_neighbors, int64_t nx,
                                     int64_t ny, const lsc_solve_options* options, double* out) {
    return lsc::guarded([&] {
        lsc::OmniArgs a = omni_args(radius, buffer, block_size, four_neighbors, 0, 1, 0, options);
        if (a.radius < 0 || a.buffer < 0 || a.block < 1 || a.block % 2 == 0)
            throw std::invalid_argument("radius and buffer must be >= 0 and block_size odd");
        const long long W = 2 * (static_cast<long long>(radius) + buffer) + 1;
        if (a.block == 1) {
            std::fill(out, out + W * W, 1.0);
            return;
        }
        hipStream_t s = options ? reinterpret_cast<hipStream_t>(options->stream) : nullptr;
        lsc::IrStats st;
        std::vector<double> c = a.precision == LSC_PRECISION_FP64 ? lsc::correction<double>(a, nx, ny, s, st)
                                                                  : lsc::correction<float>(a, nx, ny, s, st);
        std::copy(c.begin(), c.end(), out);
    });
}

}  // extern "C"
