// Dense frontal kernels for gfx950: a front is a symmetric m x m matrix of
// which only the lower triangle is kept, stored as 16 x 16 tiles, tile (I, J)
// (I >= J) at tile index I (I + 1) / 2 + J, each tile column-major. A
// partial Cholesky of the first np columns (the reference's LLT / TRSM / SYRK
// of F11, F21, F22: src/cpp/circuit/system.cpp:899-908) runs tile column by
// tile column:
//   * the panel (the 16 columns of one tile column, rows k0..m-1) is held in
//     registers, a row per lane, and factored with lane broadcasts
//     (v_readlane) - no LDS traffic and no barrier inside the panel;
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
};
template <>
struct Mfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + (r << 2); }
};

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
// (16 x 16, column-major) and F keeps the unfactored one, for teams spread
// over several workgroups that cannot be ordered inside one launch.
template <typename T, typename Sync>
__device__ bool panel(T* F, int m, int k0, int kend, int wave, int lane, int rowbase, T* diag_out, Sync sync) {
    const int cend = min(k0 + 16, m);
    const int i = lane < 16 ? k0 + lane : k0 + 16 + rowbase + 48 * wave + (lane - 16);
    const bool mine = i < m && (lane >= 16 || (wave == 0 && rowbase == 0 && diag_out == nullptr));
    T P[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) P[c] = (i < m && k0 + c < cend && i >= k0 + c) ? F[tix(i, k0 + c)] : T(0);
    sync();
    bool ok = true;
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        if (k0 + kk < kend) {
            const T dkk = bcast(P[kk], kk);
            ok = ok && (dkk > T(0));
            const T d = sqrt(dkk), inv = T(1) / d;
            if (i > k0 + kk) P[kk] *= inv;
            else if (i == k0 + kk) P[kk] = d;
#pragma unroll
            for (int c = kk + 1; c < 16; ++c) {
                if (k0 + c < cend) {
                    const T ljk = bcast(P[kk], c);
                    if (i >= k0 + c) P[c] -= P[kk] * ljk;
                }
            }
        }
    }
    if (mine) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (k0 + c < cend && i >= k0 + c) F[tix(i, k0 + c)] = P[c];
    }
    if (diag_out != nullptr && wave == 0 && lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) diag_out[(c << 4) + lane] = P[c];
    }
    return ok;
}

// C(I, J) -= P_I P_J^T over the panel columns k0..kend-1 of tile column Kt.
// MFMA operands as described at Mfma.
template <typename T>
__device__ __forceinline__ void tile_update(T* F, int I, int J, int Kt, int k0, int kend, int lane) {
    typedef typename Mfma<T>::acc Acc;
    T* C = F + tile_index(I, J) * 256 + ((lane & 15) << 4);
    const int r0 = Mfma<T>::row(lane, 0), r1 = Mfma<T>::row(lane, 1), r2 = Mfma<T>::row(lane, 2),
              r3 = Mfma<T>::row(lane, 3);
    Acc acc;
    acc[0] = C[r0], acc[1] = C[r1], acc[2] = C[r2], acc[3] = C[r3];
    const T* PI = F + tile_index(I, Kt) * 256;
    const T* PJ = F + tile_index(J, Kt) * 256;
#pragma unroll
    for (int kc = 0; kc < 16; kc += 4) {
        if (k0 + kc < kend) {
            const int col = kc + (lane >> 4);
            const bool valid = k0 + col < kend;
            const T a = valid ? -PI[(col << 4) + (lane & 15)] : T(0);
            const T b = valid ? PJ[(col << 4) + (lane & 15)] : T(0);
            acc = Mfma<T>::run(a, b, acc);
        }
    }
    C[r0] = acc[0], C[r1] = acc[1], C[r2] = acc[2], C[r3] = acc[3];
}

// Lower-triangle tile (I, J), I >= J, of index t in a T x T triangle.
__device__ __forceinline__ void tri_tile(int t, int& I, int& J) {
    int i = static_cast<int>((sqrtf(8.0f * t + 1.0f) - 1.0f) * 0.5f);
    while (((i + 1) * (i + 2)) / 2 <= t) ++i;
    while ((i * (i + 1)) / 2 > t) --i;
    I = i;
    J = t - (i * (i + 1)) / 2;
}

// Partial Cholesky of the first np columns of a tiled front of m rows by a
// team of `nw` wavefronts (this one is `wave`); `sync` orders the team. The
// result flag is returned to every wave through `flag` (team-visible memory).
template <typename T, typename Sync>
__device__ bool factor_front(T* F, int m, int np, int wave, int nw, int lane, int* flag, Sync sync) {
    const int MT = (m + 15) >> 4;
    for (int Kt = 0; Kt * 16 < np; ++Kt) {
        const int k0 = Kt * 16, kend = min(k0 + 16, np);
        const bool ok = panel<T>(F, m, k0, kend, wave, lane, 0, static_cast<T*>(nullptr), sync);
        if (!ok && wave == 0 && lane == 0) *flag = 1;
        sync();
        const int Tn = MT - Kt - 1;
        const int ntiles = (Tn * (Tn + 1)) >> 1;
        for (int t = wave; t < ntiles; t += nw) {
            int I, J;
            tri_tile(t, I, J);
            tile_update<T>(F, Kt + 1 + I, Kt + 1 + J, Kt, k0, kend, lane);
        }
        sync();
    }
    return *flag == 0;
}

}  // namespace lsc
