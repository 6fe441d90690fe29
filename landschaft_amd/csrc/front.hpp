// Dense frontal kernels for gfx950: a front is a symmetric m x m matrix of
// which only the lower triangle is kept, stored as 16 x 16 tiles, tile (I, J)
// (I >= J) at tile index I (I + 1) / 2 + J, each tile column-major. A
// partial Cholesky of the first np columns (the reference's LLT / TRSM / SYRK
// of F11, F21, F22: src/cpp/circuit/system.cpp:899-908) runs tile column by
// tile column:
//   * the panel (the 16 columns of one tile column, rows k0..m-1) is held in
//     registers, a row per lane, and factored step by step, each step's
//     multipliers of the diagonal block broadcast to the wavefront through
//     a small LDS buffer of its own - no barrier inside the panel;
//   * the trailing tiles are updated with v_mfma_f32_16x16x4f32 (or the f64
//     MFMA), C(I,J) -= P_I P_J^T, spread over the waves of the team.
// Fronts carry one extra row, the right-hand side: eliminating it with the
// matrix performs the forward substitution (y = L11^-1 b, the update rows'
// b -= L21 y) as part of the factorisation, as the reference fuses its
// forward solve into factorize (system.cpp:909-921).
#pragma once

#include <hip/hip_runtime.h>

namespace lsc {

__device__ __forceinline__ int tile_index(int I, int J) { return ((I * (I + 1)) >> 1) + J; }
// Offset of element (i, j), i >= j, in a tiled lower triangle.
__device__ __forceinline__ int tix(int i, int j) {
    return tile_index(i >> 4, j >> 4) * 256 + ((j & 15) << 4) + (i & 15);
}

// The wavefront's index in its workgroup as a wave-uniform (scalar) value,
// so that what depends on it stays in scalar registers and scalar branches.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float bcast(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ __forceinline__ double bcast(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template <typename T>
struct Mfma;
// Both 16x16x4 MFMAs take A[i][k] from lane i + 16 k and B[k][j] from lane
// j + 16 k; result element r of a lane is D[row(lane, r)][lane % 16]: rows
// 4 (lane / 16) + r for f32, interleaved lane / 16 + 4 r for f64 (checked on
// the device by lsc_probe_mfma).
template <>
struct Mfma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return ((lane >> 4) << 2) + r; }
    // the four rows of a lane are consecutive in a column-major tile: one 16-byte access
    static __device__ __forceinline__ acc load(const float* col, int lane) {
        return *reinterpret_cast<const acc*>(col + ((lane >> 4) << 2));
    }
    static __device__ __forceinline__ void store(float* col, int lane, acc v) {
        *reinterpret_cast<acc*>(col + ((lane >> 4) << 2)) = v;
    }
    // pivot inv = 1 / sqrt(a) and d = a inv: the hardware reciprocal square
    // root (1 ulp, biased) refined by one Newton step (an unbiased error of
    // about 1 ulp; the unrefined one makes the factor a measurably worse
    // preconditioner of the FP64 refinement, and the correctly rounded square
    // root and division cost four times the instructions)
    static __device__ __forceinline__ void root(float a, float& d, float& inv) {
        float r = __builtin_amdgcn_rsqf(a);
        r = fmaf(r, fmaf(-0.5f * a * r, r, 0.5f), r);
        inv = r;
        d = a * r;
    }
};
template <>
struct Mfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + (r << 2); }
    static __device__ __forceinline__ acc load(const double* col, int lane) {
        acc v;
        v[0] = col[row(lane, 0)], v[1] = col[row(lane, 1)], v[2] = col[row(lane, 2)], v[3] = col[row(lane, 3)];
        return v;
    }
    static __device__ __forceinline__ void store(double* col, int lane, acc v) {
        col[row(lane, 0)] = v[0], col[row(lane, 1)] = v[1], col[row(lane, 2)] = v[2], col[row(lane, 3)] = v[3];
    }
    static __device__ __forceinline__ void root(double a, double& d, double& inv) {
        d = sqrt(a);
        inv = 1.0 / d;
    }
};

// The 16 elimination steps of a panel held as P[16] (a row per lane, row i).
// A step's column of the diagonal block (the multipliers of rows k0..k0+15,
// lanes 0..15) reaches every lane through the wavefront's 64-entry LDS
// buffer `bc`: one store of every lane's value, then 16-byte reads of the
// entries the step needs, the same address on every lane (a broadcast). The
// update then costs one FMA per later column; a lane read (v_readlane) per
// column would cost a second vector instruction. A wavefront's LDS
// operations complete in order, so the next step's store cannot overtake
// this step's reads.
template <typename T, bool kFull>
__device__ __forceinline__ bool panel_steps(T (&P)[16], int i, int lane, int k0, int kend, T* bc) {
    constexpr int V = 16 / static_cast<int>(sizeof(T));  // entries per 16-byte read
    typedef T vec __attribute__((ext_vector_type(V)));
    bool ok = true;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        if (!kFull && k0 + kk >= kend) break;  // the tile column's remaining columns are not pivots
        const T dkk = bcast(P[kk], kk);
        ok = ok && (dkk > T(0));
        T d, inv;
        Mfma<T>::root(dkk, d, inv);
        const T l = P[kk] * inv;
        P[kk] = i == k0 + kk ? d : l;
        if (kk < 15) {
            bc[lane] = l;
            T lj[16];
#pragma unroll
            for (int g = (kk + 1) / V; g < 16 / V; ++g) {
                const vec v = *reinterpret_cast<const vec*>(bc + g * V);
#pragma unroll
                for (int e = 0; e < V; ++e) lj[g * V + e] = v[e];
            }
            // every column, without a test: past the front's last row (column
            // k0 + c >= m) the lanes' rows are zero, so l and lj[c] are zero
            // there, and those columns are never stored
#pragma unroll
            for (int c = kk + 1; c < 16; ++c) P[c] -= l * lj[c];
        }
    }
    return ok;
}

// The panel of tile column Kt (columns k0 = 16 Kt .. cend-1 with cend =
// min(k0 + 16, m)): columns k0..kend-1 factorised (pivots), the others of
// the tile column updated, in registers. Every wavefront of the team holds
// the 16 rows of the diagonal block on lanes 0..15 and factors that block
// redundantly, so it can solve its own 48 rows below (lanes 16..63 of wave w:
// rows k0 + 16 + rowbase + 48 w + lane - 16) without any exchange between
// waves (or workgroups: `rowbase` offsets a workgroup's rows). Requires
// m - k0 - 16 <= rowbase + 48 nw. Returns false when a pivot is not positive.
// `sync` orders the team's loads of the diagonal block before wave 0 writes
// it back. With `diag_out`, the factored diagonal block goes there instead
// (16 x 16, column-major, zero above the diagonal) and F keeps the
// unfactored one, for teams spread over several workgroups that cannot be
// ordered inside one launch. `bc`: the wavefront's 64-entry LDS buffer
// (16-byte aligned) for the steps' broadcasts.
// The updates run on every lane without a branch: entries above the
// diagonal of the diagonal block then hold values that are written back but
// never read (the broadcasts read rows at or below their column; L, the
// update matrix and the next panels read the lower triangle only).
template <typename T, typename Sync>
__device__ bool panel(T* F, int m, int k0, int kend, int wave, int lane, int rowbase, T* diag_out, T* bc,
                      Sync sync) {
    const int i = lane < 16 ? k0 + lane : k0 + 16 + rowbase + 48 * wave + (lane - 16);
    const bool mine = i < m && (lane >= 16 || (wave == 0 && rowbase == 0 && diag_out == nullptr));
    T P[16];
    if (i < m) {
        // columns past the front's last column (k0 + c >= m) hold zeros
#pragma unroll
        for (int c = 0; c < 16; ++c) P[c] = F[tix(i, k0 + c)];
    } else {
#pragma unroll
        for (int c = 0; c < 16; ++c) P[c] = T(0);
    }
    sync();
    // a full panel (16 pivots, so 16 columns) without any column test
    const bool ok = kend == k0 + 16 ? panel_steps<T, true>(P, i, lane, k0, kend, bc)
                                    : panel_steps<T, false>(P, i, lane, k0, kend, bc);
    // every column of the lane's row, without a test: above the diagonal
    // block's diagonal the values are never read, and past the front's last
    // column they are the zeros that were loaded
    if (mine) {
#pragma unroll
        for (int c = 0; c < 16; ++c) F[tix(i, k0 + c)] = P[c];
    }
    if (diag_out != nullptr && wave == 0 && lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) diag_out[(c << 4) + lane] = i >= k0 + c ? P[c] : T(0);
    }
    return ok;
}

// C(I, J) -= P_I P_J^T over the panel columns k0..kend-1 of tile column Kt,
// in two steps so that the operands of the next tile can be requested before
// this one's products: `load` issues every LDS (or global) read of the tile,
// `apply` runs the MFMAs and writes C back. Columns of the tile column past
// kend are multiplied by 0. MFMA operands as described at Mfma.
template <typename T>
struct TileJob {
    typedef typename Mfma<T>::acc Acc;
    T* C;
    Acc acc;
    T a[4], b[4];

    __device__ __forceinline__ void load(T* F, int I, int J, int Kt, int lane) {
        C = F + tile_index(I, J) * 256 + ((lane & 15) << 4);
        const T* PI = F + tile_index(I, Kt) * 256 + (lane & 15);
        const T* PJ = F + tile_index(J, Kt) * 256 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = 4 * q + (lane >> 4);
            a[q] = PI[col << 4];
            b[q] = PJ[col << 4];
        }
        acc = Mfma<T>::load(C, lane);
    }
    __device__ __forceinline__ void apply(int nk, int lane) {
        if (nk >= 16) {
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = Mfma<T>::run(-a[q], b[q], acc);
        } else {
            // the panel's columns past kend (not pivots) are not multipliers:
            // their A operand is 0 (B holds finite values of the front)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (4 * q < nk) acc = Mfma<T>::run(4 * q + (lane >> 4) < nk ? -a[q] : T(0), b[q], acc);
        }
        Mfma<T>::store(C, lane, acc);
    }
};

template <typename T>
__device__ __forceinline__ void tile_update(T* F, int I, int J, int Kt, int k0, int kend, int lane) {
    TileJob<T> job;
    job.load(F, I, J, Kt, lane);
    job.apply(kend - k0, lane);
}

// Lower-triangle tile (I, J), I >= J, of index t in a T x T triangle.
__device__ __forceinline__ void tri_tile(int t, int& I, int& J) {
    int i = static_cast<int>((sqrtf(8.0f * t + 1.0f) - 1.0f) * 0.5f);
    while (((i + 1) * (i + 2)) / 2 <= t) ++i;
    while ((i * (i + 1)) / 2 > t) --i;
    I = i;
    J = t - (i * (i + 1)) / 2;
}

// Moves (I, J) to the next lower-triangle tile `step` tiles on in row order.
__device__ __forceinline__ void next_tile(int& I, int& J, int step, int Tn) {
    J += step;
    while (J > I && I < Tn) {
        J -= I + 1;
        ++I;
    }
}

// Partial Cholesky of the first np columns of a tiled front of m rows by a
// team of `nw` wavefronts (this one is `wave`); `sync` orders the team. The
// result flag is returned to every wave through `flag` (team-visible memory);
// `bc` is this wave's 64-entry LDS buffer for the panel's broadcasts.
// A wave's trailing tiles are pipelined: the next tile's operands are read
// while the current tile's MFMAs run.
template <typename T, typename Sync>
__device__ bool factor_front(T* F, int m, int np, int wave, int nw, int lane, int* flag, T* bc, Sync sync) {
    const int MT = (m + 15) >> 4;
    for (int Kt = 0; Kt * 16 < np; ++Kt) {
        const int k0 = Kt * 16, kend = min(k0 + 16, np);
        const bool ok = panel<T>(F, m, k0, kend, wave, lane, 0, static_cast<T*>(nullptr), bc, sync);
        if (!ok && wave == 0 && lane == 0) *flag = 1;
        sync();
        const int Tn = MT - Kt - 1;
        int I = 0, J = 0;
        next_tile(I, J, wave, Tn);
        if (I < Tn) {
            TileJob<T> cur, nxt;
            cur.load(F, Kt + 1 + I, Kt + 1 + J, Kt, lane);
            for (;;) {
                next_tile(I, J, nw, Tn);
                const bool more = I < Tn;
                if (more) nxt.load(F, Kt + 1 + I, Kt + 1 + J, Kt, lane);
                cur.apply(kend - k0, lane);
                if (!more) break;
                cur = nxt;
            }
        }
        sync();
    }
    return *flag == 0;
}

}  // namespace lsc
