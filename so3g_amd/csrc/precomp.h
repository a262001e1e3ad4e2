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
    const int32_t *tile_dims;  // tiled: (rows, cols) per tile
    int32_t n_tile;            // tiled: tiles in tile_dims
};

// Status of a cached index tuple whose slot 0 is >= 0.  The reference
// indexes the map without checking (Projection.cxx:2066-2154); here an
// index outside the map is reported (PC_RANGE, raised as FLAG_RANGE and
// turned into a RuntimeError after the launch) and the sample is skipped.
enum { PC_OK = 0, PC_NOTILE = 1, PC_RANGE = 2 };

template <int K, typename T>
__device__ __forceinline__ int pc_elem(T *map, T *const *tiles, const PcMap &pm,
                                       const int32_t *pi, T **e, int64_t *plane) {
    if constexpr (K == 1) {
        *plane = pm.ny * pm.nx;
        if (pi[0] >= *plane) return PC_RANGE;
        *e = map + pi[0];
    } else if constexpr (K == 2) {
        *plane = pm.ny * pm.nx;
        if (pi[0] >= pm.ny || pi[1] < 0 || pi[1] >= pm.nx) return PC_RANGE;
        *e = map + ((int64_t)pi[0] * pm.nx + pi[1]);
    } else {
        if (pi[0] >= pm.n_tile) return PC_RANGE;
        T *b = tiles[pi[0]];
        if (b == nullptr) return PC_NOTILE;
        const int th = pm.tile_dims[2 * pi[0]], tw = pm.tile_dims[2 * pi[0] + 1];
        if (pi[1] < 0 || pi[1] >= th || pi[2] < 0 || pi[2] >= tw) return PC_RANGE;
        *plane = (int64_t)th * tw;
        *e = b + ((int64_t)pi[1] * tw + pi[2]);
    }
    return PC_OK;
}

__device__ __forceinline__ void pc_flag(int *flags, int st) {
    raise_flag(flags, st == PC_NOTILE ? FLAG_TILE : FLAG_RANGE);
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
        const double *e = nullptr;
        const int st = pc_elem<K>(map, tiles, pm, pi, &e, &plane);
        if (st != PC_OK) {
            pc_flag(flags, st);
            if (!accumulate) *sp = 0.f;
            return;
        }
        float s = 0;
#pragma unroll
        for (int c = 0; c < NC; c++) s = s + spin[i * NC + c] * e[c * plane];
        *sp = accumulate ? *sp + s : s;
    });
}

// Pixel-major copy of a (NC, npix) map, padded to 4 components, so that the
// from_map gather of one sample reads one 32-byte span instead of NC lines.
template <int NC>
__global__ __launch_bounds__(kBlock) void k_pack_map(const double *map, int64_t plane,
                                                     double *packed) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < plane;
         p += (int64_t)gridDim.x * kBlock) {
        double v[4] = {0., 0., 0., 0.};
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = map[c * plane + p];
        reinterpret_cast<double2 *>(packed + 4 * p)[0] = double2{v[0], v[1]};
        reinterpret_cast<double2 *>(packed + 4 * p)[1] = double2{v[2], v[3]};
    }
}

// Precomp from_map for k = 2 maps and n_t % 4 == 0, reading the packed map:
// four consecutive samples per lane (16-byte loads of the pixel pairs, spin
// factors and signal), two detectors per iteration with every independent
// load issued before the map gathers.  Same arithmetic as k_pc_from_map.
template <int NC>
__global__ __launch_bounds__(kBlock) void k_pc_from_map_packed(PcIn in, PcMap pm,
                                                               const double *packed,
                                                               const int32_t *pix,
                                                               const float *spin, float *sig,
                                                               int accumulate, int *flags) {
    const int64_t t = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 4;
    if (t >= in.n_t) return;
    const int64_t d0 = (int64_t)blockIdx.y * in.dpb;
    const int64_t d1 = min(d0 + in.dpb, in.n_det);
    for (int64_t d = d0; d < d1; d += 2) {
        const bool two = d + 1 < d1;
        int4 pa[2], pb[2];
        float sp[2][4 * NC];
        float o[2][4];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int64_t i = (d + (two ? u : 0)) * in.n_t + t;
            pa[u] = *reinterpret_cast<const int4 *>(pix + 2 * i);
            pb[u] = *reinterpret_cast<const int4 *>(pix + 2 * i + 4);
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const float4 f = reinterpret_cast<const float4 *>(spin + i * NC)[q];
                sp[u][4 * q] = f.x; sp[u][4 * q + 1] = f.y;
                sp[u][4 * q + 2] = f.z; sp[u][4 * q + 3] = f.w;
            }
            const float4 a = accumulate ? *reinterpret_cast<const float4 *>(sig + i)
                                        : float4{0.f, 0.f, 0.f, 0.f};
            o[u][0] = a.x; o[u][1] = a.y; o[u][2] = a.z; o[u][3] = a.w;
        }
        double mv[2][4][4];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int py[4] = {pa[u].x, pa[u].z, pb[u].x, pb[u].z};
            const int px[4] = {pa[u].y, pa[u].w, pb[u].y, pb[u].w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const bool ok = py[r] >= 0 && py[r] < pm.ny && px[r] >= 0 && px[r] < pm.nx;
                if (py[r] >= 0 && !ok) raise_flag(flags, FLAG_RANGE);
                const int64_t p = ok ? ((int64_t)py[r] * pm.nx + px[r]) : 0;
                const double2 lo = reinterpret_cast<const double2 *>(packed + 4 * p)[0];
                mv[u][r][0] = lo.x; mv[u][r][1] = lo.y;
                if (NC > 2) {
                    const double2 hi = reinterpret_cast<const double2 *>(packed + 4 * p)[1];
                    mv[u][r][2] = hi.x; mv[u][r][3] = hi.y;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if (u == 1 && !two) break;
            const int py[4] = {pa[u].x, pa[u].z, pb[u].x, pb[u].z};
            const int px[4] = {pa[u].y, pa[u].w, pb[u].y, pb[u].w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (py[r] < 0 || py[r] >= pm.ny || px[r] < 0 || px[r] >= pm.nx) continue;
                float sr = 0;
#pragma unroll
                for (int c = 0; c < NC; c++) sr = sr + sp[u][r * NC + c] * mv[u][r][c];
                o[u][r] = accumulate ? o[u][r] + sr : sr;
            }
            const int64_t i = (d + u) * in.n_t + t;
            *reinterpret_cast<float4 *>(sig + i) = float4{o[u][0], o[u][1], o[u][2], o[u][3]};
        }
    }
}

// Precomp to_map, Projection.cxx:2100-2121: (det_wt * sig) * pf in float32,
// added to the float64 map.  AGG sums runs of equal pixels within a wave
// before the atomics (wave_runs); XCD sends each XCD's atomics to its own
// private copy of the map (copies + xcc * copy_stride), summed afterwards
// by k_reduce_copies.  All lanes of a block stay alive to the end (the
// wave-level scan needs them); lanes without a sample add nothing.
template <int K, int NC, bool AGG, bool XCD>
__global__ __launch_bounds__(kBlock) void k_pc_to_map(PcIn in, PcMap pm, double *map,
                                                      double *const *tiles, const int32_t *pix,
                                                      const float *spin, const float *sig,
                                                      const float *det_weights, double *copies,
                                                      int64_t copy_stride, int *flags) {
    double *mbase = map;
    if constexpr (XCD) mbase = copies + (int64_t)xcc_id() * copy_stride;
    int64_t d0, d1, t;
    int t_end;
    if (in.chunks != nullptr) {
        const Chunk c = in.chunks[blockIdx.x];
        d0 = c.det;
        d1 = d0 + 1;
        t = c.t0 + (int64_t)threadIdx.x;
        t_end = c.t1;
    } else {
        t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        t_end = (int)in.n_t;
        d0 = (int64_t)blockIdx.y * in.dpb;
        d1 = min(d0 + in.dpb, in.n_det);
    }
    const bool valid = t < t_end;
    for (int64_t d = d0; d < d1; d++) {
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) continue;
        }
        int64_t key = -1;
        double v[NC];
        int64_t plane = 0;
        double *e = nullptr;
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = 0.;
        if (valid) {
            const int64_t i = d * in.n_t + t;
            int32_t pi[K];
#pragma unroll
            for (int j = 0; j < K; j++) pi[j] = pix[i * K + j];
            if (pi[0] >= 0) {
                const int st = pc_elem<K>(mbase, tiles, pm, pi, &e, &plane);
                if (st != PC_OK) {
                    pc_flag(flags, st);
                    e = nullptr;
                } else {
                    key = (int64_t)(uintptr_t)e;
                    const float s = sig[i];
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        v[c] = (double)(det_wt * s * spin[i * NC + c]);
                }
            }
        }
        bool emit = key >= 0;
        if constexpr (AGG) emit = wave_runs<NC>(key, v) && key >= 0;
        if (emit) {
#pragma unroll
            for (int c = 0; c < NC; c++) unsafeAtomicAdd(e + c * plane, v[c]);
        }
    }
}

// Element and plane of a Precomp table key (k_pc_scatter_lds).
template <int K>
__device__ __forceinline__ double *pc_key_elem(double *map, double *const *tiles,
                                               const PcMap &pm, int32_t key, int64_t *plane) {
    if constexpr (K == 3) {
        const int32_t t = key >> 16;
        *plane = (int64_t)pm.tile_dims[2 * t] * pm.tile_dims[2 * t + 1];
        return tiles[t] + (key & 0xFFFF);
    } else {
        *plane = pm.ny * pm.nx;
        return map + key;
    }
}

// Precomp to_map (W = false) / to_weight_map (W = true) through an LDS
// table per block, as the on-the-fly scatter's small tables.  Keys: the
// element offset in the plane (K = 1, 2), or tile << 16 | offset in the
// tile (K = 3; an element of a tile past 65536 pixels or of a tile numbered
// from 32768 up is added to the map directly).  A block takes `dets`
// detectors (in `order`: detectors
// grouped by the pixel they see at mid-observation, so that a block's
// detectors sweep overlapping pixels) x 256 samples; runs of equal pixels
// within a wave are summed first, each run's total is added to the table
// (key = element offset, home slot = key mod H; a run that finds no slot
// adds to the map directly) and the occupied slots are added to the map at
// the end, one FP64 atomic per value.  The float32 products are those of
// k_pc_to_map / k_pc_to_weights; the float64 sums are the same values added
// in another order.
template <int K, int NC, bool W, int H>
__global__ __launch_bounds__(kBlock) void k_pc_scatter_lds(PcIn in, PcMap pm, double *map,
                                                           double *const *tiles,
                                                           const int32_t *pix,
                                                           const float *spin, const float *sig,
                                                           const float *det_weights,
                                                           const int32_t *order, int32_t dets,
                                                           int *flags) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    __shared__ int32_t hk[H];
    __shared__ double hv[NV * H];
    lds_table_clear<NV, H>(hk, hv, kBlock);
    __syncthreads();
    const int64_t ng = (in.n_det + dets - 1) / dets;
    const int64_t tc = (int64_t)blockIdx.x / ng;
    const int64_t g0 = ((int64_t)blockIdx.x - tc * ng) * dets;
    const int nd = (int)min((int64_t)dets, in.n_det - g0);
    const int64_t t = tc * kBlock + threadIdx.x;
    const bool valid = t < in.n_t;
    for (int j = 0; j < nd; j++) {
        const int64_t d = order[g0 + j];
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) continue;
        }
        int32_t key = -1;
        double v[NV];
#pragma unroll
        for (int c = 0; c < NV; c++) v[c] = 0.;
        if (valid) {
            const int64_t i = d * in.n_t + t;
            int32_t pi[K];
#pragma unroll
            for (int q = 0; q < K; q++) pi[q] = pix[i * K + q];
            if (pi[0] >= 0) {
                double *e = nullptr;
                int64_t pl = 0;
                const int st = pc_elem<K>(map, tiles, pm, pi, &e, &pl);
                if (st != PC_OK) {
                    pc_flag(flags, st);
                } else {
                    bool direct = false;
                    if constexpr (K == 3) {
                        const int64_t off = e - tiles[pi[0]];
                        direct = off >= 65536 || pi[0] >= 32768;
                        key = direct ? -1 : (int32_t)((pi[0] << 16) | (int32_t)off);
                    } else {
                        key = (int32_t)(e - map);
                    }
                    if constexpr (W) {
                        float pf[NC];
#pragma unroll
                        for (int c = 0; c < NC; c++) pf[c] = spin[i * NC + c];
                        int k = 0;
#pragma unroll
                        for (int a = 0; a < NC; a++)
#pragma unroll
                            for (int b = a; b < NC; b++, k++)
                                v[k] = (double)(pf[a] * pf[b] * det_wt);
                    } else {
                        const float s = sig[i];
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            v[c] = (double)(det_wt * s * spin[i * NC + c]);
                    }
                    if (direct) {
#pragma unroll
                        for (int c = 0; c < NV; c++) {
                            const int lead = W ? hash_lead<NC>(c) : c;
                            unsafeAtomicAdd(e + lead * pl, v[c]);
                            v[c] = 0.;
                        }
                    }
                }
            }
        }
        const bool emit = wave_runs<NV>((int64_t)key, v) && key >= 0;
        if (emit && !hash_add<NV, H, true>(hk, hv, key, v)) {
            int64_t pl;
            double *e = pc_key_elem<K>(map, tiles, pm, key, &pl);
#pragma unroll
            for (int c = 0; c < NV; c++) {
                const int lead = W ? hash_lead<NC>(c) : c;
                unsafeAtomicAdd(e + lead * pl, v[c]);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < H; i += kBlock) {
        const int32_t key = hk[i];
        if (key == kHashEmpty) continue;
        int64_t pl;
        double *e = pc_key_elem<K>(map, tiles, pm, key, &pl);
#pragma unroll
        for (int c = 0; c < NV; c++) {
            const int lead = W ? hash_lead<NC>(c) : c;
            unsafeAtomicAdd(e + lead * pl, hv[c * H + i]);
        }
    }
}

// map[i] += sum over the n_copy XCD-private partial maps, fixed order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reduce_copies(T *map, const T *copies, int64_t n,
                                                          int64_t stride, int n_copy) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        T acc = map[i];
        for (int c = 0; c < n_copy; c++) acc += copies[c * stride + i];
        map[i] = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_zero(T *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = T(0);
}

// Precomp to_weight_map, Projection.cxx:2135-2154: pf_a * pf_b * det_wt,
// upper triangle; runs of equal pixels summed per wave when agg is set.
template <int K, int NC>
__global__ __launch_bounds__(kBlock) void k_pc_to_weights(PcIn in, PcMap pm, double *map,
                                                          double *const *tiles,
                                                          const int32_t *pix,
                                                          const float *spin,
                                                          const float *det_weights, int agg,
                                                          int *flags) {
    constexpr int NV = NC * (NC + 1) / 2;
    int64_t d0, d1, t, t_end;
    if (in.chunks != nullptr) {
        const Chunk c = in.chunks[blockIdx.x];
        d0 = c.det;
        d1 = d0 + 1;
        t = c.t0 + (int64_t)threadIdx.x;
        t_end = c.t1;
    } else {
        t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
        t_end = in.n_t;
        d0 = (int64_t)blockIdx.y * in.dpb;
        d1 = min(d0 + in.dpb, in.n_det);
    }
    const bool valid = t < t_end;
    for (int64_t d = d0; d < d1; d++) {
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) continue;
        }
        double *e[1] = {nullptr};
        int64_t key[1] = {-1};
        int64_t off[1][NV];
        double v[1][NV];
#pragma unroll
        for (int k = 0; k < NV; k++) { v[0][k] = 0.; off[0][k] = 0; }
        if (valid) {
            const int64_t i = d * in.n_t + t;
            int32_t pi[K];
#pragma unroll
            for (int j = 0; j < K; j++) pi[j] = pix[i * K + j];
            if (pi[0] >= 0) {
                int64_t plane;
                double *ej = nullptr;
                const int st = pc_elem<K>(map, tiles, pm, pi, &ej, &plane);
                if (st != PC_OK) {
                    pc_flag(flags, st);
                } else {
                    e[0] = ej;
                    key[0] = (int64_t)(uintptr_t)ej;
                    float pf[NC];
#pragma unroll
                    for (int c = 0; c < NC; c++) pf[c] = spin[i * NC + c];
                    int k = 0;
#pragma unroll
                    for (int a = 0; a < NC; a++)
#pragma unroll
                        for (int b = a; b < NC; b++, k++) {
                            off[0][k] = (a * NC + b) * plane;
                            v[0][k] = (double)(pf[a] * pf[b] * det_wt);
                        }
                }
            }
        }
        scatter_points<1, NV, double, double>(agg != 0, e, key, off, v);
    }
}

}  // namespace so3g
