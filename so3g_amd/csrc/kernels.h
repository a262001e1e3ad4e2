// kernels.h -- on-the-fly projection kernels (pointing fused with the
// pixel/spin computation and the map gather or scatter).
//
// Work mapping.  A block owns 256 consecutive time samples; each thread
// keeps its boresight quaternion in registers and walks a group of
// detectors, so the boresight stream (32 B/sample) is read once per
// detector group instead of once per detector, the detector quaternion
// and response are wave-uniform (scalar loads), and every signal access
// is a coalesced 1 KiB row segment of the det-major float32 array.
// With an explicit segment list (so3g's thread_intervals) a block owns
// one chunk (det, t0, t1) of at most 256 samples instead.
#pragma once

#include <type_traits>

#include "ptg.h"

namespace so3g {

constexpr int kBlock = 256;

struct Chunk { int32_t det, t0, t1; };

struct PtgIn {
    const double *bore;
    const double *dets;
    int64_t n_t, n_det;
    Tabs tabs;
    int32_t dpb;               // detectors per block (dense mode)
    int32_t agg;               // wave-level run aggregation of scatter-adds
    const Chunk *chunks;       // segment mode when non-null
};

// Dense-mode block -> (detector group, 256-sample chunk).  The grid is 1-D
// and the detector group varies fastest, so the blocks that share a chunk of
// the boresight run together, one after another across the XCDs: each XCD's
// L2 fetches the chunk once instead of once per detector group (with the
// chunk varying fastest, 250 groups re-read the whole 23 MB boresight of a
// 720000-sample observation from beyond the L2, 5.8 GB).
__device__ __forceinline__ void dense_block(const PtgIn &in, int64_t &t0, int64_t &d0) {
    const int64_t ng = (in.n_det + in.dpb - 1) / in.dpb;
    const int64_t b = blockIdx.x;
    const int64_t tc = b / ng;
    d0 = (b - tc * ng) * in.dpb;
    t0 = tc * kBlock;
}

// Calls f(det, t, qbore) for the samples this thread owns.
template <class F>
__device__ __forceinline__ void for_samples(const PtgIn &in, F &&f) {
    if (in.chunks != nullptr) {
        const Chunk c = in.chunks[blockIdx.x];
        const int t = c.t0 + (int)threadIdx.x;
        if (t < c.t1) f(c.det, (int64_t)t, load_q(in.bore + 4 * (int64_t)t));
        return;
    }
    int64_t t0, d0;
    dense_block(in, t0, d0);
    const int64_t t = t0 + threadIdx.x;
    if (t >= in.n_t) return;
    const Q4 qb = load_q(in.bore + 4 * t);
    const int64_t d1 = min(d0 + in.dpb, in.n_det);
    for (int64_t d = d0; d < d1; d++) f(d, t, qb);
}

// Hardware id (0..7 on MI355X) of the XCD -- the chiplet with its own L2 --
// that runs the calling wave (HW_REG_XCC_ID, gfx940 and later).
__device__ __forceinline__ int xcc_id() {
    unsigned x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return (int)(x & 7);
}

// Wave-level run aggregation for scatter-adds.  The 64 lanes of a wave hold
// (key, v[NC]) for consecutive samples; a key < 0 marks a lane with nothing
// to add.  Lanes in a run of equal keys sum their values (segmented
// inclusive scan) so that the run's last lane holds the run total; returns
// true for that lane.  A detector crossing a map pixel stays in it for a
// few consecutive samples, so this divides the number of memory atomics by
// the mean run length.  All 64 lanes must call this together.
// One-lane shifts across the wave by DPP (GFX9 whole-wave shifts): the
// previous lane's value (lane 0 gets `old`) and the next lane's (lane 63
// gets `old`).
__device__ __forceinline__ int dpp_prev(int v, int old) {
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_next(int v, int old) {
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ int64_t dpp_prev64(int64_t v) {
    const int lo = dpp_prev((int)(v & 0xffffffffll), 0), hi = dpp_prev((int)(v >> 32), 0);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}
__device__ __forceinline__ int64_t dpp_next64(int64_t v) {
    const int lo = dpp_next((int)(v & 0xffffffffll), 0), hi = dpp_next((int)(v >> 32), 0);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}
template <typename VT>
__device__ __forceinline__ VT dpp_prev_val(VT v) {
    if constexpr (sizeof(VT) == 8) {
        return __longlong_as_double(dpp_prev64(__double_as_longlong(v)));
    } else {
        return __int_as_float(dpp_prev(__float_as_int(v), 0));
    }
}

// Wave-level run aggregation for scatter-adds.  The 64 lanes of a wave hold
// (key, v[NC]) for consecutive samples; a key < 0 marks a lane with nothing
// to add.  Lanes in a run of equal keys sum their values (segmented
// inclusive scan, Hillis-Steele: at step `off` a lane adds the value `off`
// lanes back when that lane is in its run) so that the run's last lane
// holds the run total; returns true for that lane.  A detector crossing a
// map pixel stays in it for a few consecutive samples, so this divides the
// number of memory atomics by the mean run length.  Run starts come from one
// ballot, one-lane shifts from DPP, and the scan stops after the last step
// in which some lane still adds (runs are short: the steps skipped would add
// nothing, so the sums are those of the full scan).  All 64 lanes must call
// this together.
template <int NC, typename VT>
__device__ __forceinline__ bool wave_runs(int64_t key, VT (&v)[NC]) {
    const int lane = (int)__lane_id();
    const int64_t kprev = dpp_prev64(key);
    const uint64_t starts = __ballot(lane == 0 || kprev != key);
    const int head = 63 - __clzll((long long)(starts & (~0ull >> (63 - lane))));
    const int rank = lane - head;
    if (__any(rank >= 1)) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const VT u = dpp_prev_val(v[c]);
            if (rank >= 1) v[c] += u;
        }
#pragma unroll
        for (int off = 2; off < 64; off <<= 1) {
            if (!__any(rank >= off)) break;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const VT u = __shfl_up(v[c], off);
                if (rank >= off) v[c] += u;
            }
        }
    }
    const int64_t knext = dpp_next64(key);
    return lane == 63 || knext != key;
}

// Library-owned device flags.
enum { FLAG_TILE = 0, FLAG_OFFMAP = 1, FLAG_RANGE = 2, FLAG_COUNT = 4 };

__device__ __forceinline__ void raise_flag(int *flags, int which) {
    __hip_atomic_fetch_or(flags + which, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// --------------------------------------------------------------------------
// coords / pixels / pointing_matrix
// --------------------------------------------------------------------------
template <int PROJ>
__global__ __launch_bounds__(kBlock) void k_coords(PtgIn in, double *out) {
    for_samples(in, [&](int64_t d, int64_t t, const Q4 &qb) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ>(qb, qd, in.tabs, co);
        double *o = out + 4 * (d * in.n_t + t);
        *reinterpret_cast<double2 *>(o) = double2{co[0], co[1]};
        *reinterpret_cast<double2 *>(o + 2) = double2{co[2], co[3]};
    });
}

// pixels() and pointing_matrix() in one kernel (spin == nullptr for
// pixels()).  Off-map samples write slot 0 = -1 and raise FLAG_OFFMAP;
// their other slots are filled afterwards by k_fill_stale.
template <int PROJ, int PK, int NC, bool SPIN>
// SPIN: pointing_matrix() (spin factors written) rather than pixels(); a
// template flag, so that pixels() does not compute the spin terms.
__global__ __launch_bounds__(kBlock) void k_pixels(PtgIn in, Geo g, const float *resp,
                                                   int32_t *pix, float *spin, int *flags) {
    constexpr int K = index_count<PK>();
    for_samples(in, [&](int64_t d, int64_t t, const Q4 &qb) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ, SPIN>(qb, qd, in.tabs, co);
        int32_t pi[3] = {0, 0, 0};
        const bool on = get_pixel<PK, SPIN>(g, co, pi);
        int32_t *po = pix + (d * in.n_t + t) * K;
        if (on) {
#pragma unroll
            for (int j = 0; j < K; j++) po[j] = pi[j];
        } else {
            po[0] = -1;
            raise_flag(flags, FLAG_OFFMAP);
        }
        if constexpr (SPIN) {
            float pf[3];
            spin_factors(NC, co, resp[2 * d], resp[2 * d + 1], pf);
            float *so = spin + (d * in.n_t + t) * NC;
#pragma unroll
            for (int j = 0; j < NC; j++) so[j] = pf[j];
        }
    });
}

// Forward-fill of the non-leading index slots of off-map samples: the
// reference keeps one index buffer per detector, initialised {-1, 0, ...},
// and an off-map sample only overwrites slot 0 (Projection.cxx:1372,
// 1418, 848-859), so slots 1.. repeat the last on-map sample's values.
// One block per detector; a block-wide max-scan carries the index of the
// last on-map sample across 1024-sample chunks.  Exits at once when no
// sample was off-map.
constexpr int kFillBlock = 1024;

template <int K>
__global__ __launch_bounds__(kFillBlock) void k_fill_stale(int32_t *pix, int64_t n_t,
                                                           const int *flags) {
    if (flags[FLAG_OFFMAP] == 0) return;
    __shared__ int64_t scan[kFillBlock];
    const int64_t d = blockIdx.x;
    int32_t *row = pix + d * n_t * K;
    int64_t carry = -1;
    for (int64_t s = 0; s < n_t; s += kFillBlock) {
        const int64_t t = s + threadIdx.x;
        const bool in_range = t < n_t;
        const bool on = in_range && row[t * K] >= 0;
        scan[threadIdx.x] = on ? t : -1;
        __syncthreads();
        for (int off = 1; off < kFillBlock; off <<= 1) {
            int64_t v = scan[threadIdx.x];
            if ((int)threadIdx.x >= off) v = max(v, scan[threadIdx.x - off]);
            __syncthreads();
            scan[threadIdx.x] = v;
            __syncthreads();
        }
        const int64_t last = max(carry, scan[threadIdx.x]);
        if (in_range && !on) {
#pragma unroll
            for (int j = 1; j < K; j++) row[t * K + j] = (last >= 0) ? row[last * K + j] : 0;
        }
        carry = max(carry, scan[kFillBlock - 1]);
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// from_map: sig = (float)((double)sig + map * pf * w) per point and
// component (Projection.cxx:1761-1765); deterministic per sample.
// --------------------------------------------------------------------------
template <int PROJ, int PK, int INTERP, int NC>
__device__ __forceinline__ void from_map_sample(const PtgIn &in, const Geo &g,
                                                const MapView<const double> &m,
                                                const float *resp, float *sig, int accumulate,
                                                int *flags, int64_t d, int64_t t,
                                                const Q4 &qb) {
    const Q4 qd = load_q(in.dets + 4 * d);
    double co[4];
    get_coords<PROJ>(qb, qd, in.tabs, co);
    float pf[3];
    spin_factors(NC, co, resp[2 * d], resp[2 * d + 1], pf);
    Pt p[4];
    const int np = get_pixels<PK, INTERP>(g, co, p);
    float *sp = sig + d * in.n_t + t;
    float s = accumulate ? *sp : 0.f;
    for (int ip = 0; ip < np; ip++) {
        const double *e = elem<PK>(g, m, p[ip], 0);
        if (e == nullptr) { raise_flag(flags, FLAG_TILE); continue; }
        const int64_t plane = (PK == PK_FLAT_TILED)
            ? (int64_t)m.tile_dims[2 * p[ip].i0] * m.tile_dims[2 * p[ip].i0 + 1]
            : (PK == PK_HP_TILED ? (int64_t)g.npt : g.npix);
#pragma unroll
        for (int c = 0; c < NC; c++) s = s + e[c * plane] * pf[c] * p[ip].w;
    }
    *sp = s;
}

template <int PROJ, int PK, int INTERP, int NC>
__global__ __launch_bounds__(kBlock) void k_from_map(PtgIn in, Geo g, MapView<const double> m,
                                                     const float *resp, float *sig,
                                                     int accumulate, int *flags) {
    for_samples(in, [&](int64_t d, int64_t t, const Q4 &qb) {
        from_map_sample<PROJ, PK, INTERP, NC>(in, g, m, resp, sig, accumulate, flags, d, t, qb);
    });
}

// --------------------------------------------------------------------------
// to_map / to_weights: the per-sample product is formed in float32 exactly
// as the reference does (Projection.cxx:1807, 1850) and added to the
// float64 map with a hardware FP64 atomic.  PREC selects the precision
// study variants (float32 map; float64 products).
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void atomic_add(T *a, T v) {
    unsafeAtomicAdd(a, v);
}

template <int PREC>
struct PrecT { using map_t = double; };
template <>
struct PrecT<1> { using map_t = float; };

// Calls f(det, t, qbore, valid) on every lane of the block for every
// detector of the block's group (segment mode: the chunk's detector);
// lanes past the end of the samples get valid == false, so that wave-level
// collectives see all 64 lanes.
template <class F>
__device__ __forceinline__ void for_samples_all(const PtgIn &in, F &&f) {
    int64_t d0, d1, t, t_end;
    if (in.chunks != nullptr) {
        const Chunk c = in.chunks[blockIdx.x];
        d0 = c.det;
        d1 = d0 + 1;
        t = c.t0 + (int64_t)threadIdx.x;
        t_end = c.t1;
    } else {
        int64_t t0;
        dense_block(in, t0, d0);
        t = t0 + threadIdx.x;
        t_end = in.n_t;
        d1 = min(d0 + in.dpb, in.n_det);
    }
    const bool valid = t < t_end;
    const Q4 qb = valid ? load_q(in.bore + 4 * t) : Q4{1., 0., 0., 0.};
    for (int64_t d = d0; d < d1; d++) f(d, t, qb, valid);
}

// Scatter-add of NV values per point to e + off[k] for the NP interpolation
// points of each lane's sample (key < 0: no point).  With agg, runs of equal
// keys within the wave are summed first (wave_runs) and only each run's
// last lane issues the atomics.
template <int NP, int NV, typename MT, typename VT>
__device__ __forceinline__ void scatter_points(bool agg, MT *const (&e)[NP],
                                               const int64_t (&key)[NP],
                                               const int64_t (&off)[NP][NV],
                                               VT (&v)[NP][NV]) {
#pragma unroll
    for (int j = 0; j < NP; j++) {
        bool emit = key[j] >= 0;
        if (agg) emit = wave_runs<NV>(key[j], v[j]) && key[j] >= 0;
        if (emit) {
#pragma unroll
            for (int k = 0; k < NV; k++) atomic_add(e[j] + off[j][k], (MT)v[j][k]);
        }
    }
}

template <int PROJ, int PK, int INTERP, int NC, int PREC>
__device__ __forceinline__ void to_map_sample(const PtgIn &in, const Geo &g,
                                              const MapView<typename PrecT<PREC>::map_t> &m,
                                              const float *resp, const float *sig,
                                              const float *det_weights, int *flags, int64_t d,
                                              int64_t t, const Q4 &qb, bool valid) {
    using MT = typename PrecT<PREC>::map_t;
    using VT = typename std::conditional<PREC == 1, float, double>::type;
    constexpr int NP = INTERP ? 4 : 1;
    float det_wt = 1.f;
    if (det_weights != nullptr) {
        det_wt = det_weights[d];
        if (det_wt == 0.f) return;
    }
    MT *e[NP];
    int64_t key[NP], off[NP][NC];
    VT v[NP][NC];
#pragma unroll
    for (int j = 0; j < NP; j++) {
        key[j] = -1;
        e[j] = nullptr;
#pragma unroll
        for (int c = 0; c < NC; c++) { v[j][c] = 0; off[j][c] = 0; }
    }
    if (valid) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ>(qb, qd, in.tabs, co);
        float pf[3];
        spin_factors(NC, co, resp[2 * d], resp[2 * d + 1], pf);
        const float s = sig[d * in.n_t + t];
        Pt p[4];
        const int np = get_pixels<PK, INTERP>(g, co, p);
#pragma unroll
        for (int j = 0; j < NP; j++) {
            if (j >= np) break;
            MT *ej = elem<PK>(g, m, p[j], 0);
            if (ej == nullptr) { raise_flag(flags, FLAG_TILE); continue; }
            const int64_t plane = (PK == PK_FLAT_TILED)
                ? (int64_t)m.tile_dims[2 * p[j].i0] * m.tile_dims[2 * p[j].i0 + 1]
                : (PK == PK_HP_TILED ? (int64_t)g.npt : g.npix);
            e[j] = ej;
            key[j] = (int64_t)(uintptr_t)ej;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                off[j][c] = c * plane;
                if constexpr (PREC == 2)
                    v[j][c] = (double)s * (double)pf[c] * (double)p[j].w * (double)det_wt;
                else
                    v[j][c] = (VT)(s * pf[c] * p[j].w * det_wt);
            }
        }
    }
    scatter_points<NP, NC, MT, VT>(in.agg, e, key, off, v);
}

template <int PROJ, int PK, int INTERP, int NC, int PREC>
__global__ __launch_bounds__(kBlock) void k_to_map(PtgIn in, Geo g,
                                                   MapView<typename PrecT<PREC>::map_t> m,
                                                   const float *resp, const float *sig,
                                                   const float *det_weights, int *flags) {
    for_samples_all(in, [&](int64_t d, int64_t t, const Q4 &qb, bool valid) {
        to_map_sample<PROJ, PK, INTERP, NC, PREC>(in, g, m, resp, sig, det_weights, flags, d,
                                                  t, qb, valid);
    });
}

template <int PROJ, int PK, int INTERP, int NC, int PREC>
__global__ __launch_bounds__(kBlock) void k_to_weights(PtgIn in, Geo g,
                                                       MapView<typename PrecT<PREC>::map_t> m,
                                                       const float *resp,
                                                       const float *det_weights, int *flags) {
    using MT = typename PrecT<PREC>::map_t;
    using VT = typename std::conditional<PREC == 1, float, double>::type;
    constexpr int NP = INTERP ? 4 : 1;
    constexpr int NV = NC * (NC + 1) / 2;
    for_samples_all(in, [&](int64_t d, int64_t t, const Q4 &qb, bool valid) {
        float det_wt = 1.f;
        if (det_weights != nullptr) {
            det_wt = det_weights[d];
            if (det_wt == 0.f) return;
        }
        MT *e[NP];
        int64_t key[NP], off[NP][NV];
        VT v[NP][NV];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            key[j] = -1;
            e[j] = nullptr;
#pragma unroll
            for (int k = 0; k < NV; k++) { v[j][k] = 0; off[j][k] = 0; }
        }
        if (valid) {
            const Q4 qd = load_q(in.dets + 4 * d);
            double co[4];
            get_coords<PROJ>(qb, qd, in.tabs, co);
            float pf[3];
            spin_factors(NC, co, resp[2 * d], resp[2 * d + 1], pf);
            Pt p[4];
            const int np = get_pixels<PK, INTERP>(g, co, p);
#pragma unroll
            for (int j = 0; j < NP; j++) {
                if (j >= np) break;
                MT *ej = elem<PK>(g, m, p[j], 0);
                if (ej == nullptr) { raise_flag(flags, FLAG_TILE); continue; }
                const int64_t plane = (PK == PK_FLAT_TILED)
                    ? (int64_t)m.tile_dims[2 * p[j].i0] * m.tile_dims[2 * p[j].i0 + 1]
                    : (PK == PK_HP_TILED ? (int64_t)g.npt : g.npix);
                e[j] = ej;
                key[j] = (int64_t)(uintptr_t)ej;
                const float w = p[j].w;
                int k = 0;
#pragma unroll
                for (int a = 0; a < NC; a++)
#pragma unroll
                    for (int b = a; b < NC; b++, k++) {
                        off[j][k] = (a * NC + b) * plane;
                        if constexpr (PREC == 2)
                            v[j][k] = (double)pf[a] * (double)pf[b] * (double)w * (double)w *
                                      (double)det_wt;
                        else
                            v[j][k] = (VT)(pf[a] * pf[b] * w * w * det_wt);
                    }
            }
        }
        scatter_points<NP, NV, MT, VT>(in.agg, e, key, off, v);
    });
}

// --------------------------------------------------------------------------
// to_map / to_weights through LDS-privatised partial maps.
//
// A block owns G detectors, taken in `order` (detectors sorted by the map
// row they see at mid-observation, then by column: row-mates, which in a
// constant-elevation scan sweep the same pixels), over T consecutive
// samples.  Each thread keeps its sample's boresight in registers and walks
// the block's detectors; runs of equal pixels within a wave are summed
// first (wave_runs) and each run's total is added to an open-addressing
// hash table of the block's pixels in LDS (int32 pixel keys, NV float64
// sums per slot, linear probing).  A long time span lets the scan's
// repeated sweeps add into the same slots, so each touched pixel costs one
// global FP64 atomic per value per block when the table is flushed instead
// of one per run.  A run that finds no slot within kProbes probes (table
// full) adds straight to the map, so the result does not depend on the
// table size; the sums are float64 sums of the same float32 products as
// k_to_map / k_to_weights in another order.
// --------------------------------------------------------------------------
struct HashWork {
    const int32_t *order;   // detector order, n_det entries
    int32_t dets;           // G: detectors per block
    int32_t span;           // T: samples per block
    int32_t r;              // samples per lane: 1 (runs across the wave) or 4
};

constexpr int32_t kHashEmpty = -1;
constexpr int kProbes = 32;

// int32 key of a map element (map pixel numbering of the pixelizor; tiled
// flat maps number tile t's pixels from t * ty * tx, rows of the nominal
// tile width tx, so that no tile table is read per sample).
template <int PK>
__device__ __forceinline__ int32_t hash_key(const Geo &g, const Pt &p) {
    if constexpr (PK == PK_FLAT) return p.i0 * g.nx + p.i1;
    else if constexpr (PK == PK_HP) return p.i0;
    else if constexpr (PK == PK_HP_TILED) return p.i0 * g.npt + p.i1;
    else return p.i0 * (g.ty * g.tx) + p.i1 * g.tx + p.i2;
}

// Map element of a key's component 0 and the distance between its
// components (inverse of hash_key; one tile lookup for all components).
template <int PK>
__device__ __forceinline__ double *hash_elem(const Geo &g, const MapView<double> &m,
                                             int32_t key, int64_t &plane) {
    if constexpr (PK == PK_FLAT || PK == PK_HP) {
        plane = g.npix;
        return m.map + key;
    } else if constexpr (PK == PK_HP_TILED) {
        const int32_t t = key / g.npt;
        plane = g.npt;
        return m.tiles[t] + (key - t * g.npt);
    } else {
        const int32_t ts = g.ty * g.tx;
        const int32_t t = key / ts;
        const int32_t r = key - t * ts;
        const int32_t row = r / g.tx;
        const int32_t tw = m.tile_dims[2 * t + 1];
        plane = (int64_t)m.tile_dims[2 * t] * tw;
        return m.tiles[t] + ((int64_t)row * tw + (r - row * g.tx));
    }
}

// Adds v to the slot of key; false when no slot was found in kProbes probes.
// MOD: the home slot is key mod H, so consecutive keys (a stretch of map
// pixels) take consecutive slots and the flush adds them to the map in
// coalesced runs; otherwise a multiplicative hash spreads them.
template <int NV, int H, bool MOD>
__device__ __forceinline__ bool hash_add(int32_t *hk, double *hv, int32_t key,
                                         const double (&v)[NV]) {
    constexpr int kBits = __builtin_ctz(H);
    uint32_t h = MOD ? ((uint32_t)key & (H - 1)) : ((uint32_t)key * 2654435761u) >> (32 - kBits);
    for (int probe = 0; probe < kProbes; probe++) {
        // One LDS round trip per probe: the compare-and-swap claims an empty
        // slot and returns the key already there otherwise.
        int32_t cur = kHashEmpty;
        __hip_atomic_compare_exchange_strong(hk + h, &cur, key, __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == kHashEmpty || cur == key) {
#pragma unroll
            for (int c = 0; c < NV; c++)
                __hip_atomic_fetch_add(hv + c * H + h, v[c], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            return true;
        }
        h = (h + 1) & (H - 1);
    }
    return false;
}

// Map slot (a * NC + b) of the k-th upper-triangle product (a <= b) of a
// weights map, in the order the products are formed.
template <int NC>
__device__ __forceinline__ int hash_lead(int k) {
    for (int a = 0; a < NC; a++)
        for (int b = a; b < NC; b++, k--)
            if (k == 0) return a * NC + b;
    return 0;
}

// W = false: to_map (NV = NC values per sample, Projection.cxx:1807);
// W = true: to_weights (the NC (NC + 1) / 2 upper-triangle products,
// Projection.cxx:1850).
template <int NV, int H>
__device__ __forceinline__ void lds_table_clear(int32_t *hk, double *hv, int nt) {
    for (int i = threadIdx.x; i < H; i += nt) {
        hk[i] = kHashEmpty;
#pragma unroll
        for (int c = 0; c < NV; c++) hv[c * H + i] = 0.;
    }
}

// One detector's sample t into the block's table (all lanes of the block
// call it together for the same detector).
template <int PROJ, int PK, int INTERP, int NC, bool W, int H, bool MOD>
__device__ __forceinline__ void lds_scatter_det(const PtgIn &in, const Geo &g,
                                                const MapView<double> &m, const float *resp,
                                                const float *sig, const float *det_weights,
                                                int *flags, int64_t d, int64_t t, const Q4 &qb,
                                                bool valid, int32_t *hk, double *hv) {
    constexpr int NP = INTERP ? 4 : 1;
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    float det_wt = 1.f;
    if (det_weights != nullptr) {
        det_wt = det_weights[d];
        if (det_wt == 0.f) return;
    }
    int32_t key[NP];
    double v[NP][NV];
#pragma unroll
    for (int jp = 0; jp < NP; jp++) {
        key[jp] = -1;
#pragma unroll
        for (int c = 0; c < NV; c++) v[jp][c] = 0.;
    }
    if (valid) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ>(qb, qd, in.tabs, co);
        float pf[3];
        spin_factors(NC, co, resp[2 * d], resp[2 * d + 1], pf);
        float s = 0.f;
        if constexpr (!W) s = sig[d * in.n_t + t];
        Pt p[4];
        const int np = get_pixels<PK, INTERP>(g, co, p);
#pragma unroll
        for (int jp = 0; jp < NP; jp++) {
            if (jp >= np) break;
            if constexpr (PK == PK_FLAT_TILED || PK == PK_HP_TILED) {
                if (m.tiles[p[jp].i0] == nullptr) {
                    raise_flag(flags, FLAG_TILE);
                    continue;
                }
            }
            key[jp] = hash_key<PK>(g, p[jp]);
            const float w = p[jp].w;
            if constexpr (W) {
                int k = 0;
#pragma unroll
                for (int a = 0; a < NC; a++)
#pragma unroll
                    for (int b = a; b < NC; b++, k++)
                        v[jp][k] = (double)(pf[a] * pf[b] * w * w * det_wt);
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) v[jp][c] = (double)(s * pf[c] * w * det_wt);
            }
        }
    }
#pragma unroll
    for (int jp = 0; jp < NP; jp++) {
        const bool emit = wave_runs<NV>((int64_t)key[jp], v[jp]) && key[jp] >= 0;
        if (emit && !hash_add<NV, H, MOD>(hk, hv, key[jp], v[jp])) {
            int64_t plane;
            double *e = hash_elem<PK>(g, m, key[jp], plane);
#pragma unroll
            for (int c = 0; c < NV; c++) {
                const int lead = W ? hash_lead<NC>(c) : c;
                unsafeAtomicAdd(e + lead * plane, v[jp][c]);
            }
        }
    }
}

// One detector's R consecutive samples t0 .. t0 + R - 1 of this lane into
// the block's table (nearest-neighbour pixels): the lane sums its own runs
// of equal keys in registers and adds each run to the table (or to the map
// when no slot is found), without a wave-level scan.
template <int PROJ, int PK, int NC, bool W, int H, bool MOD, int R>
__device__ __forceinline__ void lds_scatter_det_r(const PtgIn &in, const Geo &g,
                                                  const MapView<double> &m, const float *resp,
                                                  const float *sig, const float *det_weights,
                                                  int *flags, int64_t d, int64_t t0,
                                                  int64_t t_end, const Q4 (&qb)[R],
                                                  int32_t *hk, double *hv) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    float det_wt = 1.f;
    if (det_weights != nullptr) {
        det_wt = det_weights[d];
        if (det_wt == 0.f) return;
    }
    const Q4 qd = load_q(in.dets + 4 * d);
    const float r0 = resp[2 * d], r1 = resp[2 * d + 1];
    int32_t ck = -1;
    double cv[NV];
#pragma unroll
    for (int c = 0; c < NV; c++) cv[c] = 0.;
    auto emit = [&](int32_t key, const double (&v)[NV]) {
        if (key < 0) return;
        if (!hash_add<NV, H, MOD>(hk, hv, key, v)) {
            int64_t plane;
            double *e = hash_elem<PK>(g, m, key, plane);
#pragma unroll
            for (int c = 0; c < NV; c++) {
                const int lead = W ? hash_lead<NC>(c) : c;
                unsafeAtomicAdd(e + lead * plane, v[c]);
            }
        }
    };
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (t0 + r >= t_end) break;
        double co[4];
        get_coords<PROJ>(qb[r], qd, in.tabs, co);
        float pf[3];
        spin_factors(NC, co, r0, r1, pf);
        Pt p[4];
        if (get_pixels<PK, 0>(g, co, p) == 0) continue;
        if constexpr (PK == PK_FLAT_TILED || PK == PK_HP_TILED) {
            if (m.tiles[p[0].i0] == nullptr) {
                raise_flag(flags, FLAG_TILE);
                continue;
            }
        }
        double v[NV];
        if constexpr (W) {
            int k = 0;
#pragma unroll
            for (int a = 0; a < NC; a++)
#pragma unroll
                for (int b = a; b < NC; b++, k++)
                    v[k] = (double)(pf[a] * pf[b] * p[0].w * p[0].w * det_wt);
        } else {
            const float s = sig[d * in.n_t + t0 + r];
#pragma unroll
            for (int c = 0; c < NC; c++) v[c] = (double)(s * pf[c] * p[0].w * det_wt);
        }
        const int32_t key = hash_key<PK>(g, p[0]);
        if (key != ck) {
            emit(ck, cv);
            ck = key;
#pragma unroll
            for (int c = 0; c < NV; c++) cv[c] = v[c];
        } else {
#pragma unroll
            for (int c = 0; c < NV; c++) cv[c] += v[c];
        }
    }
    emit(ck, cv);
}

// Adds the table's occupied slots to the map, one FP64 atomic per value.
template <int PK, int NC, bool W, int H>
__device__ __forceinline__ void lds_table_flush(const Geo &g, const MapView<double> &m,
                                                const int32_t *hk, const double *hv, int nt) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    for (int i = threadIdx.x; i < H; i += nt) {
        const int32_t key = hk[i];
        if (key == kHashEmpty) continue;
        int64_t plane;
        double *e = hash_elem<PK>(g, m, key, plane);
#pragma unroll
        for (int c = 0; c < NV; c++) {
            const int lead = W ? hash_lead<NC>(c) : c;
            unsafeAtomicAdd(e + lead * plane, hv[c * H + i]);
        }
    }
}

// R = 1: one sample per lane, runs summed across the wave (wave_runs);
// R = 4 (nearest neighbour only): four consecutive samples per lane, runs
// summed in the lane (lds_scatter_det_r).
template <int PROJ, int PK, int INTERP, int NC, bool W, int NT, int H, bool MOD, int R>
__global__ __launch_bounds__(NT) void k_scatter_lds(PtgIn in, Geo g, MapView<double> m,
                                                    const float *resp, const float *sig,
                                                    const float *det_weights, HashWork hw,
                                                    int *flags) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    __shared__ int32_t hk[H];
    __shared__ double hv[NV * H];
    lds_table_clear<NV, H>(hk, hv, NT);
    __syncthreads();
    // 1-D grid, detector group fastest (as dense_block): the blocks over one
    // span of the boresight run together
    const int64_t ng = (in.n_det + hw.dets - 1) / hw.dets;
    const int64_t ts = (int64_t)blockIdx.x / ng;
    const int64_t g0 = ((int64_t)blockIdx.x - ts * ng) * hw.dets;
    const int nd = (int)min((int64_t)hw.dets, in.n_det - g0);
    const int64_t t_beg = ts * hw.span;
    const int64_t t_end = min(t_beg + (int64_t)hw.span, in.n_t);
    if constexpr (R == 1 || INTERP != 0) {
        for (int64_t tb = t_beg; tb < t_end; tb += NT) {
            const int64_t t = tb + threadIdx.x;
            const bool valid = t < t_end;
            const Q4 qb = valid ? load_q(in.bore + 4 * t) : Q4{1., 0., 0., 0.};
            for (int j = 0; j < nd; j++)
                lds_scatter_det<PROJ, PK, INTERP, NC, W, H, MOD>(in, g, m, resp, sig,
                                                                 det_weights, flags,
                                                                 hw.order[g0 + j], t, qb,
                                                                 valid, hk, hv);
        }
    } else {
        for (int64_t tb = t_beg; tb < t_end; tb += (int64_t)NT * R) {
            const int64_t t0 = tb + (int64_t)threadIdx.x * R;
            Q4 qb[R];
#pragma unroll
            for (int r = 0; r < R; r++)
                qb[r] = t0 + r < t_end ? load_q(in.bore + 4 * (t0 + r)) : Q4{1., 0., 0., 0.};
            for (int j = 0; j < nd; j++)
                lds_scatter_det_r<PROJ, PK, NC, W, H, MOD, R>(in, g, m, resp, sig, det_weights,
                                                              flags, hw.order[g0 + j], t0,
                                                              t_end, qb, hk, hv);
        }
    }
    __syncthreads();
    lds_table_flush<PK, NC, W, H>(g, m, hk, hv, NT);
}

// --------------------------------------------------------------------------
// to_map / to_weights through a map row held in LDS (flat maps).
//
// A constant-elevation scan in horizon coordinates keeps each detector on
// one map row.  A block owns a group of up to G detectors that share a home
// row (their row at mid-observation) over T consecutive samples, and keeps
// the NV float64 sums per pixel of a segment of that row -- xw columns from
// x0, the whole row where it fits in LDS -- in LDS (direct indexing: no
// keys, no probing).  Each thread keeps its samples' boresights in
// registers and walks the group's detectors; each run of equal pixels is
// added to the segment in LDS, or to the map with float64 atomics when it
// lies elsewhere.  The segment's touched pixels are then added to the map,
// one atomic per pixel and value (tiled maps: the segment crosses the
// tiles of one tile row, the samples' tiles are checked as in k_to_map).
// Sums are float64 sums of the same float32 products as k_to_map /
// k_to_weights, in another order.
// --------------------------------------------------------------------------
struct RowWork {
    const int32_t *gdet;    // [n_groups * G] detectors of each group, -1 padded
    const int32_t *grow;    // [n_groups] home row (-1: none)
    const int32_t *gxc;     // [n_groups] home column (segment centre)
    int32_t n_groups, G, span;
    int32_t xw;             // segment width in columns (<= nx)
};

// First column of a segment of xw columns centred on column xc.
__device__ __forceinline__ int row_x0(int xc, int xw, int nx) {
    return max(0, min(xc - xw / 2, nx - xw));
}

// LDS of one row segment: NV x xw float64 sums, then (tiled maps) xw
// active-tile flags.
template <int PK>
constexpr size_t row_lds_bytes(int nv, int xw) {
    return (size_t)nv * xw * sizeof(double) + (PK == PK_FLAT_TILED ? (size_t)xw : 0);
}

// One (group, sample span) item of the row scatter, by the whole block.
// Each lane takes R consecutive samples (no wave-level scan): it sums its
// own runs of equal pixels over them in registers and adds each run to the
// segment in LDS (or to the map off it).  A wave-level scan over one
// sample per lane (wave_runs) measured slower: to_map 9.0 vs 7.2 ms,
// to_weights 8.2 vs 6.9 ms on 7.2e8 samples; R = 8 needs more than 128
// registers and measured 7.5 / 7.7 ms.  LEAN: the lane's R boresight
// quaternions are reloaded for each detector (cache hits) instead of held
// across the detector loop, and the sample loop is not unrolled, which
// lets the kernel fit 64 registers (k_scatter_row_x2).
template <int PROJ, int PK, int NC, bool W, int NT, int R, bool LEAN = false>
__device__ __forceinline__ void row_item(const PtgIn &in, const Geo &g, const MapView<double> &m,
                                         const float *resp, const float *sig,
                                         const float *det_weights, const int32_t *gd, int G,
                                         int row, int x0, int xw, int64_t t_beg,
                                         int64_t t_end, int *flags, double *racc) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    constexpr bool TILED = PK == PK_FLAT_TILED;
    uint8_t *ract = reinterpret_cast<uint8_t *>(racc + NV * xw);
    for (int q = threadIdx.x; q < NV * xw; q += NT) racc[q] = 0.;
    if constexpr (TILED) {
        if (row >= 0) {
            const int t0 = (row / g.ty) * g.n_tx;
            for (int x = threadIdx.x; x < xw; x += NT)
                ract[x] = m.tiles[t0 + (x0 + x) / g.tx] != nullptr;
        }
    }
    __syncthreads();
    // adds one run's sums (map row iy, column ix): to the segment in LDS,
    // or to the map off it; a run in an inactive tile raises the tile flag
    auto emit = [&](int iy, int ix, bool on, const double (&v)[NV]) {
        if (!on) return;
        const unsigned lx = (unsigned)(ix - x0);
        if (iy == row && lx < (unsigned)xw) {
            if constexpr (TILED) {
                if (!ract[lx]) {
                    raise_flag(flags, FLAG_TILE);
                    return;
                }
            }
#pragma unroll
            for (int c = 0; c < NV; c++) atomicAdd(racc + c * xw + lx, v[c]);
        } else {
            double *e;
            int64_t plane;
            if constexpr (TILED) {
                const int sy = iy / g.ty, sx = ix / g.tx;
                const int t = sy * g.n_tx + sx;
                e = m.tiles[t];
                if (e == nullptr) {
                    raise_flag(flags, FLAG_TILE);
                    return;
                }
                const int tw = m.tile_dims[2 * t + 1];
                plane = (int64_t)m.tile_dims[2 * t] * tw;
                e += (int64_t)(iy - sy * g.ty) * tw + (ix - sx * g.tx);
            } else {
                e = m.map + (int64_t)iy * g.nx + ix;
                plane = g.npix;
            }
#pragma unroll
            for (int c = 0; c < NV; c++) {
                const int lead = W ? hash_lead<NC>(c) : c;
                unsafeAtomicAdd(e + lead * plane, v[c]);
            }
        }
    };
    for (int64_t tb = t_beg; tb < t_end; tb += (int64_t)NT * R) {
        const int64_t t0 = tb + (int64_t)threadIdx.x * R;
        Q4 qb[LEAN ? 1 : R];
        if constexpr (!LEAN) {
#pragma unroll
            for (int r = 0; r < R; r++)
                qb[r] = t0 + r < t_end ? load_q(in.bore + 4 * (t0 + r)) : Q4{1., 0., 0., 0.};
        }
        for (int j = 0; j < G; j++) {
            const int32_t d = gd[j];
            if (d < 0) break;
            float det_wt = 1.f;
            if (det_weights != nullptr) {
                det_wt = det_weights[d];
                if (det_wt == 0.f) continue;   // skips the detector (Projection.cxx:1798)
            }
            const Q4 qd = load_q(in.dets + 4 * d);
            const float r0 = resp[2 * d], r1 = resp[2 * d + 1];
            bool con = false;
            int cy = -1, cx = 0;
            double cv[NV];
#pragma unroll
            for (int c = 0; c < NV; c++) cv[c] = 0.;
            // one sample: its pointing, then its run sums
            auto sample = [&](int r, const Q4 &qbr) {
                double co[4];
                get_coords<PROJ>(qbr, qd, in.tabs, co);
                float pf[3];
                spin_factors(NC, co, r0, r1, pf);
                // the pixel of get_pixels<PK, 0> as (row, column) of the
                // whole map; tiled maps truncate before the bounds test
                const double fx = flat_x(g, co), fy = flat_y(g, co);
                int gx, gy;
                if constexpr (TILED) {
                    gx = int(fx);
                    gy = int(fy);
                    if (gx < 0 || gx >= g.nx || gy < 0 || gy >= g.ny) return;
                } else {
                    if (fx < 0 || fx >= g.nx || fy < 0 || fy >= g.ny) return;
                    gx = int(fx);
                    gy = int(fy);
                }
                double v[NV];
                if constexpr (W) {
                    int k = 0;
#pragma unroll
                    for (int a = 0; a < NC; a++)
#pragma unroll
                        for (int b = a; b < NC; b++, k++)
                            v[k] = (double)(pf[a] * pf[b] * det_wt);
                } else {
                    const float s = sig[d * in.n_t + t0 + r];
#pragma unroll
                    for (int c = 0; c < NC; c++) v[c] = (double)(s * pf[c] * det_wt);
                }
                if (!(con && gy == cy && gx == cx)) {
                    emit(cy, cx, con, cv);
                    con = true;
                    cy = gy;
                    cx = gx;
#pragma unroll
                    for (int c = 0; c < NV; c++) cv[c] = v[c];
                } else {
#pragma unroll
                    for (int c = 0; c < NV; c++) cv[c] += v[c];
                }
            };
            if constexpr (LEAN) {
#pragma unroll 1
                for (int r = 0; r < R; r++) {
                    if (t0 + r >= t_end) break;
                    sample(r, load_q(in.bore + 4 * (t0 + r)));
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (t0 + r >= t_end) break;
                    sample(r, qb[r]);
                }
            }
            emit(cy, cx, con, cv);
        }
    }
    __syncthreads();
    if (row < 0) return;
    for (int lx = threadIdx.x; lx < xw; lx += NT) {
        const int x = x0 + lx;
        double *e;
        int64_t plane;
        if constexpr (TILED) {
            if (!ract[lx]) continue;   // nothing was added there
            const int sy = row / g.ty, sx = x / g.tx;
            const int t = sy * g.n_tx + sx;
            const int tw = m.tile_dims[2 * t + 1];
            plane = (int64_t)m.tile_dims[2 * t] * tw;
            e = m.tiles[t] + (int64_t)(row - sy * g.ty) * tw + (x - sx * g.tx);
        } else {
            e = m.map + (int64_t)row * g.nx + x;
            plane = g.npix;
        }
#pragma unroll
        for (int c = 0; c < NV; c++) {
            const double v = racc[c * xw + lx];
            const int lead = W ? hash_lead<NC>(c) : c;
            if (v != 0.) unsafeAtomicAdd(e + lead * plane, v);
        }
    }
}

// One observation: block = (sample span, group), group fastest.
template <int PROJ, int PK, int NC, bool W, int NT, int R, bool LEAN>
__device__ __forceinline__ void scatter_row_block(const PtgIn &in, const Geo &g,
                                                  const MapView<double> &m, const float *resp,
                                                  const float *sig, const float *det_weights,
                                                  const RowWork &rw, int *flags,
                                                  double *racc) {
    const int64_t ts = (int64_t)blockIdx.x / rw.n_groups;
    const int gi = (int)((int64_t)blockIdx.x - ts * rw.n_groups);
    const int64_t t_beg = ts * rw.span;
    const int64_t t_end = min(t_beg + (int64_t)rw.span, in.n_t);
    row_item<PROJ, PK, NC, W, NT, R, LEAN>(in, g, m, resp, sig, det_weights,
                                           rw.gdet + (int64_t)gi * rw.G, rw.G, rw.grow[gi],
                                           row_x0(rw.gxc[gi], rw.xw, g.nx), rw.xw, t_beg,
                                           t_end, flags, racc);
}

template <int PROJ, int PK, int NC, bool W, int NT, int R>
__global__ __launch_bounds__(NT) void k_scatter_row(PtgIn in, Geo g, MapView<double> m,
                                                    const float *resp, const float *sig,
                                                    const float *det_weights, RowWork rw,
                                                    int *flags) {
    extern __shared__ double racc[];
    scatter_row_block<PROJ, PK, NC, W, NT, R, false>(in, g, m, resp, sig, det_weights, rw,
                                                     flags, racc);
}

// The same with at most 64 registers per lane (8 waves per SIMD), so two
// 1024-thread blocks share a compute unit where their row segments fit
// its LDS together: CAR 0.5' TQU to_map on 1000 x 720000 samples 7.04 ->
// 6.4 ms (held at one block per compute unit by 94 registers before).
template <int PROJ, int PK, int NC, bool W, int NT, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_scatter_row_x2(PtgIn in, Geo g, MapView<double> m, const float *resp, const float *sig,
                 const float *det_weights, RowWork rw, int *flags) {
    extern __shared__ double racc[];
    scatter_row_block<PROJ, PK, NC, W, NT, R, true>(in, g, m, resp, sig, det_weights, rw,
                                                    flags, racc);
}

// --------------------------------------------------------------------------
// tile_hits and per-sample domains (thread-assignment helpers).
// --------------------------------------------------------------------------
// A wave's samples mostly fall in one or two tiles: the active lanes that
// share a tile add their count with one atomic (a per-sample atomic on a few
// counters serialises, ~3.5 s for 7.2e8 samples).
template <int PROJ, int PK, int INTERP>
__global__ __launch_bounds__(kBlock) void k_tile_hits(PtgIn in, Geo g, int32_t *hits) {
    const int lane = (int)__lane_id();
    for_samples(in, [&](int64_t d, int64_t t, const Q4 &qb) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ>(qb, qd, in.tabs, co);
        Pt p[4];
        const int np = get_pixels<PK, INTERP>(g, co, p);
#pragma unroll
        for (int ip = 0; ip < (INTERP ? 4 : 1); ip++) {
            const int key = ip < np ? (int)p[ip].i0 : -1;
            bool mine = key >= 0;
            uint64_t todo = __ballot(mine);
            while (todo) {
                const int leader = __ffsll((unsigned long long)todo) - 1;
                const int lk = __shfl(key, leader);
                const uint64_t same = __ballot(mine && key == lk);
                if (lane == leader) atomicAdd(hits + lk, (int)__popcll(same));
                if (key == lk) mine = false;
                todo &= ~same;
            }
        }
    });
}

__device__ __forceinline__ int wrap_mul(int a, int b) {
    return (int)((uint32_t)a * (uint32_t)b);
}

template <int PROJ, int PK, int INTERP>
__global__ __launch_bounds__(kBlock) void k_domains(PtgIn in, Geo g, int mode, int n_domain,
                                                    MapView<const double> tm,
                                                    const int32_t *tile_dom, int32_t *out,
                                                    int *flags) {
    for_samples(in, [&](int64_t d, int64_t t, const Q4 &qb) {
        const Q4 qd = load_q(in.dets + 4 * d);
        double co[4];
        get_coords<PROJ>(qb, qd, in.tabs, co);
        Pt p[4];
        const int np = get_pixels<PK, INTERP>(g, co, p);
        int dom = -1;
        for (int ip = 0; ip < np; ip++) {
            int pd;
            if (mode == 1) {
                const double *e = elem<PK>(g, tm, p[ip], 0);
                if (e == nullptr) { raise_flag(flags, FLAG_TILE); dom = -2; break; }
                pd = (int)*e;
            } else if (mode == 2) {
                pd = tile_dom[p[ip].i0];
            } else {
                // stripe(), in the reference's 32-bit int arithmetic.
                if constexpr (PK == PK_HP) {
                    pd = wrap_mul(p[ip].i0, n_domain) / (int)g.npix;
                } else if constexpr (PK == PK_HP_TILED) {
                    const int ntiles = (int)(g.npix / g.npt);
                    const int ipix = (int)((uint32_t)wrap_mul(p[ip].i0, g.npt) + (uint32_t)p[ip].i1);
                    pd = wrap_mul(ipix, n_domain) / wrap_mul(g.npt, ntiles);
                } else if constexpr (PK == PK_FLAT) {
                    pd = wrap_mul(p[ip].i1, n_domain) / g.nx;
                } else {
                    pd = p[ip].i0 % n_domain;
                }
            }
            if (ip > 0 && pd != dom) { dom = n_domain; break; }
            dom = pd;
        }
        out[d * in.n_t + t] = dom;
    });
}

}  // namespace so3g
