// precomp.h -- projection with a cached pointing matrix (ProjEng_Precomp,
// Projection.cxx:2012-2253).  No trigonometry: these kernels stream the
// int32 pixel indices, float32 spin factors and float32 signal, gather or
// scatter-add the float64 map, and are bound by HBM bandwidth.
#pragma once

#include "kernels.h"

namespace so3g {

// Map addressing for cached pixel tuples of k = 1 (1-D map), 2 (y, x) or
// 3 (tile, y, x) indices.
struct PcMap {
    int64_t ny, nx;            // non-tiled: plane = ny * nx (k=1: nx = 1)
    const int32_t *tile_dims;  // tiled
};

template <int K, typename T>
__device__ __forceinline__ T *pc_elem(T *map, T *const *tiles, const PcMap &pm,
                                      const int32_t *pi, int64_t *plane) {
    if constexpr (K == 1) {
        *plane = pm.ny * pm.nx;
        return map + pi[0];
    } else if constexpr (K == 2) {
        *plane = pm.ny * pm.nx;
        return map + ((int64_t)pi[0] * pm.nx + pi[1]);
    } else {
        T *b = tiles[pi[0]];
        if (b == nullptr) return nullptr;
        const int th = pm.tile_dims[2 * pi[0]], tw = pm.tile_dims[2 * pi[0] + 1];
        *plane = (int64_t)th * tw;
        return b + ((int64_t)pi[1] * tw + pi[2]);
    }
}

struct PcIn {
    int64_t n_t, n_det;
    int32_t dpb;
    const Chunk *chunks;
};

template <class F>
__device__ __forceinline__ void for_idx(const PcIn &in, F &&f) {
    if (in.chunks != nullptr) {
        const Chunk c = in.chunks[blockIdx.x];
        const int t = c.t0 + (int)threadIdx.x;
        if (t < c.t1) f((int64_t)c.det, (int64_t)t);
        return;
    }
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= in.n_t) return;
    const int64_t d0 = (int64_t)blockIdx.y * in.dpb;
    const int64_t d1 = min(d0 + in.dpb, in.n_det);
    for (int64_t d = d0; d < d1; d++) f(d, t);
}

// Precomp from_map, Projection.cxx:2066-2082: s accumulates in float32
// through float64 sums, then signal += s.
template <int K, int NC>
__global__ __launch_bounds__(kBlock) void k_pc_from_map(PcIn in, PcMap pm, const double *map,
                                                        const double *const *tiles,
                                                        const int32_t *pix, const float *spin,
                                                        float *sig, int accumulate,
                                                        int *flags) {
    for_idx(in, [&](int64_t d, int64_t t) {
        const int64_t i = d * in.n_t + t;
        int32_t pi[K];
#pragma unroll
        for (int j = 0; j < K; j++) pi[j] = pix[i * K + j];
        float *sp = sig + i;
        if (pi[0] < 0) {
            if (!accumulate) *sp = 0.f;
            return;
        }
        int64_t plane;
        const double *e = pc_elem<K>(map, tiles, pm, pi, &plane);
        if (e == nullptr) {
            raise_flag(flags, FLAG_TILE);
            if (!accumulate) *sp = 0.f;
            return;
        }
        float s = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) s = s + spin[i * NC + c] * e[c * plane];
        *sp = accumulate ? *sp + s : s;
    });
}

// Precomp to_map, Projection.cxx:2100-2121: (det_wt * sig) * pf in float32.
template <int K, int NC>
__global__ __launch_bounds__(kBlock) void k_pc_to_map(PcIn in, PcMap pm, double *map,
                                                      double *const *tiles, const int32_t *pix,
                                                      const float *spin, const float *sig,
                                                      const float *det_weights, int *flags) {
    for_idx(in, [&](int64_t d, int64_t t) {
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) return;
        }
        const int64_t i = d * in.n_t + t;
        int32_t pi[K];
#pragma unroll
        for (int j = 0; j < K; j++) pi[j] = pix[i * K + j];
        if (pi[0] < 0) return;
        int64_t plane;
        double *e = pc_elem<K>(map, tiles, pm, pi, &plane);
        if (e == nullptr) { raise_flag(flags, FLAG_TILE); return; }
        const float s = sig[i];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const float v = det_wt * s * spin[i * NC + c];
            unsafeAtomicAdd(e + c * plane, (double)v);
        }
    });
}

// Precomp to_weight_map, Projection.cxx:2135-2154: pf_a * pf_b * det_wt.
template <int K, int NC>
__global__ __launch_bounds__(kBlock) void k_pc_to_weights(PcIn in, PcMap pm, double *map,
                                                          double *const *tiles,
                                                          const int32_t *pix,
                                                          const float *spin,
                                                          const float *det_weights,
                                                          int *flags) {
    for_idx(in, [&](int64_t d, int64_t t) {
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) return;
        }
        const int64_t i = d * in.n_t + t;
        int32_t pi[K];
#pragma unroll
        for (int j = 0; j < K; j++) pi[j] = pix[i * K + j];
        if (pi[0] < 0) return;
        int64_t plane;
        double *e = pc_elem<K>(map, tiles, pm, pi, &plane);
        if (e == nullptr) { raise_flag(flags, FLAG_TILE); return; }
        float pf[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) pf[c] = spin[i * NC + c];
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
            for (int b = a; b < NC; b++)
                unsafeAtomicAdd(e + (a * NC + b) * plane, (double)(pf[a] * pf[b] * det_wt));
    });
}

}  // namespace so3g
