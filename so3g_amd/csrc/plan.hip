// plan.hip -- pixel-sorted projection plans for repeated cached-pointing
// projections with a fixed pointing matrix (the CG hot loop).
//
// ProjEng_Precomp's to_map scatters every sample's float32 contributions into
// the float64 map with atomics.  Over a whole observation each map pixel is
// hit by hundreds of detector passes, but passes that share a pixel are far
// apart in time, so no time-local block sees much of that reuse.  A plan
// sorts the on-map samples by (time block, pixel) once (rocPRIM radix sort,
// stable, so each detector pass stays contiguous) and to_map then streams
// the samples in that order, gathers the signal through the stored sample
// index, and sums each run of equal pixels with a wave-level segmented scan:
// one atomic per pixel per wave instead of one per sample.  Time blocks keep
// the signal slice a block gathers from (t_block_mb) inside the Infinity
// Cache.
//
// Compact storage (bytes per on-map sample in the sorted stream):
//   idx    4     sample index d * n_t + t
//   spin   4 ns  spin factors in sorted order; the T factor is not stored
//                when it is constant over each detector's samples (so3g's
//                pointing matrices: T = the detector's T response), ns = nc - 1
//   heads  1/8   bit per sample: a new pixel run starts here
//   runs         pixel of each run (one per distinct (time block, pixel))
//                and the run count before each 512-sample group
// so that a sample's pixel is run_pix[run index], the run index being a
// popcount over the head bits.  from_map reads a time-ordered copy: compact
// pixel index (4 B, -1 off the map) and the ns stored spin factors.
//
// Tiled layout ("plan_layout" 1, the default from 2^25 samples up): the
// samples are sorted by
// (time block, map tile) instead -- 32 x 64 pixels of a flat map, or 2048
// consecutive pixel numbers of a HEALPix map -- and keep their sample order
// inside a tile, so a detector's pass over a tile is a stretch of
// consecutive samples (~100 for the 1.5 degree scan) and to_map's signal
// loads are nearly contiguous.  Work items of up to 256 K positions sum into
// the tile held in LDS (two copies, by lane parity) and add it to the map
// once.  Per position: the slot
// in the tile (2 B) and the stored factors; per 64 positions: a stretch-start
// mask and a start count; per start (stretches and chunk starts): its sample.
// from_map keeps the time-ordered gather (or, with "plan_time_copy" 0 and a
// smaller plan, gathers from the tile in LDS).
//
// Pixels are numbered compactly over the pixels the observation visits (a
// bitmap of the map and a prefix count of its words): sort keys, the
// time-ordered copy and from_map's packed map then scale with the visited
// pixels, not the map -- a 1.5 degree scan visits 0.3 M of a HEALPix nside
// 4096 map's 201 M pixels.  Maps are (ny, nx) with (iy, ix) pixel pairs
// (k = 2) or (npix, 1) with pixel numbers (k = 1).
//
// Arithmetic per sample is the Precomp one (Projection.cxx:2100-2121 for
// to_map, (det_wt * sig) * pf in float32 summed in float64; :2060-2085 for
// from_map, sig += pf * map per component into a float32 partial sum).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/so3g_hip.h"
#include "kernels.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

struct so3g_pc_plan {
    int64_t n_det, n_t, ny, nx;
    int32_t nc, ns, tconst;
    int64_t n_on;          // on-map samples
    int64_t n_runs;        // pixel runs in sorted order
    int64_t t_block;       // samples per time block
    uint32_t *idx;         // [n_on] sample index, sorted by (time block, pixel)
    float *sspin;          // [n_on * ns] stored spin factors, sorted order
    float *tval;           // [n_det] T factor per detector (tconst)
    uint8_t *heads;        // [n_groups * 64] run-start bits
    uint32_t *wbase;       // [n_groups] runs before each 512-sample group
    uint32_t *run_pix;     // [n_runs] flat pixel iy * nx + ix of each run
    // Time-ordered compact pixels in groups of kTGroup samples: tpb[g] is
    // the group's base (its smallest on-map pixel) and
    // tpd[i] = pixel - base (kTOff off the map), or, for a group whose
    // pixels span more than kTOff - 1 values, -1 - k with the group's pixels
    // (-1 off the map) in tpw[k * kTGroup ...] and its number in wgrp[k].
    uint16_t *tpd;         // [n]
    int32_t *tpb;          // [n_groups]
    int32_t *tpw;          // [n_wide * kTGroup]
    int32_t *wgrp;         // [n_wide]
    int64_t n_wide;
    float *tspin;          // [n * ns] time-ordered stored spin factors
    double *packed;        // [n_vis * 4] pixel-major copy of the visited pixels
    int32_t *comp2pix;     // [n_vis] flat pixel of each compact index (ascending)
    int64_t n_vis;         // visited pixels
    int64_t bytes;
    // Tiled layout (layout 1): idx, heads, wbase and run_pix are unused,
    // and the time-ordered copy (tpd ... tspin, packed) is optional; sspin holds the factors in tiled sorted order.
    int32_t layout;        // 0: pixel-sorted runs, 1: tiles
    uint16_t *tslot;       // [n_on] slot within the tile
    uint64_t *tmask;       // [n_chunks] stretch starts per 64 positions
    uint32_t *thead;       // [n_chunks] stretch starts before each chunk
    uint32_t *head_i;      // [n_heads] sample of each stretch start
    int32_t *tile_pix;     // [n_tiles * kTS] flat pixel of each slot, -1 off the map
    void *items;           // [n_items] work items (first position, count, tile)
    uint32_t *off_idx;     // [n_off] samples off the map
    int64_t n_items, n_tiles, n_heads, n_off;
    // Tiled maps (k = 3 pixel arrays; tiled layout only): pixels are
    // numbered by concatenating the map tiles.
    int32_t n_tile_map;    // map tiles (0: a non-tiled map)
    int64_t n_sub;         // plan-tile-sized sub-blocks over all map tiles
    int32_t *tdims_d;      // [n_tile_map * 2] tile shapes
    int64_t *tbase_d;      // [n_tile_map + 1] first pixel of each tile
    int32_t *sbase_d;      // [n_tile_map + 1] first sub-block of each tile
    int32_t *tile_map;     // [n_tiles] map tile of each plan tile
    int32_t *comp_tile;    // [n_vis] map tile of each visited pixel (comp2pix: offset in it)
    double **tiles_d;      // [n_tile_map] tile pointers of the current call
    std::vector<uint8_t> tile_used;   // map tiles the plan's samples touch
    // Windowed layout (layout 2): samples in time order (tspin holds their
    // stored factors), work items of one detector's consecutive samples.
    uint16_t *wslot;       // [n] offset in the item's window, kTOff off the map
    void *witems;          // [n_witems] WItem
    int64_t n_witems;
    void *wfitems;         // [n_wfitems] from_map's pieces of the items (WItem)
    int64_t n_wfitems;
    int32_t win_max;       // largest window
    int64_t n_wslots;      // windows summed: partial slots
    double *wpart;         // [n_wslots * max(nc, nc (nc + 1) / 2)] per-item sums
    uint32_t *wcsr;        // [n_vis + 1] first entry of each visited pixel
    uint32_t *wse;         // [n_wslots] entry of each item slot (partials in entry order)
    uint32_t *wbitmap;     // [npix / 32] visited pixels (overwriting to_map)
    uint32_t *wwscan;      // [npix / 32] visited pixels before each word
    int32_t wvec;          // n_t % 8 == 0: vector loads per 8 samples
};

namespace so3g {

#define PLAN_TRY(expr)                                                               \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess)                                                        \
            return set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                               hipGetErrorString(_e));               \
    } while (0)

constexpr int kPlanBlock = 256;

// Tunables (so3g_set_option "plan_div", "plan_from_v", "plan_from_layout"):
// detector index by float reciprocal (1) or integer division (0); samples per
// lane of the planned from_map (4 or 8; the consecutive layout needs a sample
// count that allows it) and whether a lane's samples are 64 apart (strided)
// or consecutive.
int g_plan_div = 1;
int g_plan_from_v = 4;
int g_plan_from_layout = 1;
// Planned to_map: per-detector tables staged in LDS ("plan_to_lds") and
// 2048-sample chunks per block ("plan_to_cpb").
int g_plan_to_lds = 1;
int g_plan_to_cpb = 2;
// Plan layout chosen at creation ("plan_layout"): 0 pixel-sorted runs, 1
// map tiles in LDS, 2 time-ordered windows, -1 (default) tiles from 2^25
// samples up (below that a plan has too few tile work items to fill the
// GPU: 56.8 vs 48.6 us for the to_map + from_map pair of 4e6 samples with
// pixel runs), and below it windows where every window fits, else pixel
// runs; the tiled layout's positions per work item; the windowed layout's
// samples per work item ("plan_win_len").
int g_plan_layout = -1;
constexpr int64_t kPlanTiledMin = (int64_t)1 << 25;
int64_t g_plan_tile_item = 262144;
int64_t g_plan_win_len = 65536;
// Windowed layout: threads per work item ("plan_win_block", 256 or 1024) and
// how to_map adds the window sums ("plan_win_flush": 1 float64 atomics into
// the map, 0 partial sums and a second pass per pixel -- no atomics and a
// fixed summation order, but the second pass is latency-bound: 42.8 vs
// 54.8 us for the cold from_map + overwriting to_map pair of 4e6 samples).
int g_plan_win_block = 1024;
int g_plan_win_flush = 1;
// Tiled layout: also keep the time-ordered copy for from_map ("plan_time_copy",
// 1: 2.18 ms gather vs 2.54 ms over the tiles; 0: a smaller plan).
int g_plan_time_copy = 1;
constexpr int64_t kPlanLdsBytes = 32768;
// The planned gather's streams (time-ordered pixels and spin factors, read
// once per pass) and its signal stores are non-temporal, so that they do not
// displace the gathered map from the caches (2.60 vs 2.70 ms; the same hint
// on the scatter's streams costs 4.03 vs 3.18 ms, so it keeps plain loads).
typedef float f32x2 __attribute__((ext_vector_type(2)));
constexpr int kPlanR = 8;                 // sorted samples per lane
constexpr int kGroup = 64 * kPlanR;       // samples per wave = per head group
// Time-ordered pixel copy: samples per base, and the offset of an off-map
// sample.  The planned gathers take 64 x {1, 4, 8} samples per wave from an
// aligned start, so a wave never spans two groups.
constexpr int kTGroup = 512;
constexpr uint16_t kTOff = 0xFFFF;
constexpr int32_t kTWideMark = INT32_MIN;

// Pixel array geometry: k = 1 pixel numbers on an (ny = npix, 1) map, k = 2
// (iy, ix) pairs on an (ny, nx) map, k = 3 (tile, iy, ix) triples on a tiled
// map whose tiles are numbered by concatenation (ny = total pixels, nx = 1;
// tdims: per tile (rows, cols); tbase: first pixel of each tile; sbase: first
// plan-tile-sized (32 x 64) sub-block of each tile, n_tile + 1 entries each).
struct PixGeom {
    int32_t k;
    int32_t n_tile;
    const int32_t *tdims;
    const int64_t *tbase;
    const int32_t *sbase;
};

// Tile, row and column of a k = 3 sample, false off the map (tile or
// position outside the tiles' shapes).
__device__ __forceinline__ bool tile_pos(const int32_t *pix, const PixGeom &g, int64_t i,
                                         int32_t &t, int32_t &iy, int32_t &ix, int32_t &tx) {
    t = pix[3 * i]; iy = pix[3 * i + 1]; ix = pix[3 * i + 2];
    if (t < 0 || t >= g.n_tile) return false;
    tx = g.tdims[2 * t + 1];
    return iy >= 0 && iy < g.tdims[2 * t] && ix >= 0 && ix < tx;
}

// Flat pixel of sample i, -1 off the map.
__device__ __forceinline__ int64_t flat_pixel(const int32_t *pix, const PixGeom &g, int64_t i,
                                              int64_t ny, int64_t nx) {
    if (g.k == 1) {
        const int32_t p = pix[i];
        return (p >= 0 && p < ny) ? (int64_t)p : -1;
    }
    if (g.k == 3) {
        int32_t t, iy, ix, tx;
        return tile_pos(pix, g, i, t, iy, ix, tx) ? g.tbase[t] + (int64_t)iy * tx + ix : -1;
    }
    const int2 p = *reinterpret_cast<const int2 *>(pix + 2 * i);
    return (p.x >= 0 && p.x < ny && p.y >= 0 && p.y < nx) ? (int64_t)p.x * nx + p.y : -1;
}

// Compact index of a visited pixel: visited pixels before it.
__device__ __forceinline__ uint32_t compact_of(int64_t p, const uint32_t *bitmap,
                                               const uint32_t *wscan) {
    const uint32_t w = bitmap[p >> 5];
    return wscan[p >> 5] + __popc(w & ((1u << (p & 31)) - 1u));
}

// Visited-pixel bitmap.  A wave's 64 consecutive samples fall in a few
// bitmap words: the lanes that share a word OR their bits across the wave
// and one of them adds them with a single atomic (one atomic per sample on
// the few thousand words of a map took 12.8 ms for 7.2e8 samples).  The
// loop bound is uniform per wave, so every lane takes part in the shuffles.
// A sample whose slot 0 is >= 0 but whose index lies outside the map sets
// *range_bad (plan creation then fails, as the Precomp calls do).
__global__ __launch_bounds__(kPlanBlock) void k_plan_mark(const int32_t *pix, PixGeom pg, int64_t n,
                                                          int64_t ny, int64_t nx,
                                                          uint32_t *bitmap, int *range_bad) {
    const int lane = (int)__lane_id();
    for (int64_t i0 = (int64_t)blockIdx.x * kPlanBlock + (threadIdx.x & ~63); i0 < n;
         i0 += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t i = i0 + lane;
        const int64_t p = i < n ? flat_pixel(pix, pg, i, ny, nx) : -1;
        if (__any(i < n && p < 0 && pix[(int64_t)pg.k * i] >= 0) && lane == 0) *range_bad = 1;
        const int32_t w = p >= 0 ? (int32_t)(p >> 5) : -1;
        const uint32_t bit = p >= 0 ? 1u << (p & 31) : 0u;
        uint64_t todo = __ballot(w >= 0);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int32_t lw = __shfl(w, leader);
            uint32_t v = w == lw ? bit : 0u;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v |= (uint32_t)__shfl_xor((int)v, off);
            if (lane == leader && (bitmap[lw] & v) != v) atomicOr(bitmap + lw, v);
            todo &= ~__ballot(w == lw);
        }
    }
}

__global__ __launch_bounds__(kPlanBlock) void k_plan_wcount(const uint32_t *bitmap,
                                                            int64_t n_words, uint32_t *cnt) {
    for (int64_t w = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * kPlanBlock)
        cnt[w] = __popc(bitmap[w]);
}

// comp2pix: the visited pixels in ascending order.
__global__ __launch_bounds__(kPlanBlock) void k_plan_compact(const uint32_t *bitmap,
                                                             const uint32_t *wscan,
                                                             int64_t n_words, int32_t *comp2pix) {
    for (int64_t w = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * kPlanBlock) {
        uint32_t bits = bitmap[w];
        uint32_t at = wscan[w];
        while (bits) {
            const int b = __ffs(bits) - 1;
            comp2pix[at++] = (int32_t)(w * 32 + b);
            bits &= bits - 1;
        }
    }
}

// Sort keys: (time block, compact pixel) as tb * n_vis + compact, or the
// largest key for an off-map sample (sorted last).
__global__ __launch_bounds__(kPlanBlock) void k_plan_keys(const int32_t *pix, PixGeom pg, int64_t n,
                                                          int64_t ny, int64_t nx, int64_t n_t,
                                                          int64_t t_block, int32_t n_vis,
                                                          int32_t off_key,
                                                          const uint32_t *bitmap,
                                                          const uint32_t *wscan, int32_t *keys,
                                                          uint32_t *vals) {
    const int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t p = flat_pixel(pix, pg, i, ny, nx);
    const int64_t tb = (i % n_t) / t_block;
    keys[i] = p >= 0 ? (int32_t)(tb * n_vis + compact_of(p, bitmap, wscan)) : off_key;
    vals[i] = (uint32_t)i;
}

// Number of sorted keys below off_key (the on-map samples): binary search.
__global__ void k_plan_count(const int32_t *keys, int64_t n, int32_t off_key,
                             unsigned long long *n_on) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < off_key) lo = mid + 1;
        else hi = mid;
    }
    *n_on = (unsigned long long)lo;
}

// bad[d] = 1 if the T factor (component 0) of detector d is not the same
// float (bit for bit) over all its samples.
__global__ __launch_bounds__(kPlanBlock) void k_plan_tcheck(const float *spin, int nc,
                                                            int64_t n_t, int64_t n,
                                                            int *bad) {
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t d = i / n_t;
        if (__float_as_uint(spin[i * nc]) != __float_as_uint(spin[d * n_t * nc])) bad[d] = 1;
    }
}

__global__ void k_plan_tval(const float *spin, int nc, int64_t n_t, int64_t n_det,
                            float *tval) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d < n_det) tval[d] = spin[d * n_t * nc];
}

// Stored spin factors (components c0.. of nc) in sorted order.
__global__ __launch_bounds__(kPlanBlock) void k_plan_sspin(const float *spin, int nc, int c0,
                                                           const uint32_t *idx, int64_t n_on,
                                                           float *out) {
    const int ns = nc - c0;
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n_on;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t s = idx[i];
        for (int c = 0; c < ns; c++) out[i * ns + c] = spin[s * nc + c0 + c];
    }
}

// Time-ordered copy: compact pixel (or -1) and the stored spin factors.
__global__ __launch_bounds__(kPlanBlock) void k_plan_tcopy(const int32_t *pix, PixGeom pg,
                                                           const float *spin, int nc, int c0,
                                                           int64_t n, int64_t ny, int64_t nx,
                                                           const uint32_t *bitmap,
                                                           const uint32_t *wscan, int32_t *tpix,
                                                           float *tspin) {
    const int ns = nc - c0;
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t p = flat_pixel(pix, pg, i, ny, nx);
        tpix[i] = p >= 0 ? (int32_t)compact_of(p, bitmap, wscan) : -1;
        for (int c = 0; c < ns; c++) tspin[i * ns + c] = spin[i * nc + c0 + c];
    }
}

// Base of each kTGroup-sample group of the time-ordered compact pixels
// (smallest on-map pixel, 0 for a group off the map), or kTWideMark when
// they span more than kTOff - 1 values; counts the wide groups.  One wave per
// group.
__global__ __launch_bounds__(kPlanBlock) void k_plan_tbase(const int32_t *tpix, int64_t n,
                                                           int64_t n_groups, int32_t *tpb,
                                                           uint32_t *n_wide) {
    const int lane = threadIdx.x & 63;
    constexpr int WPB = kPlanBlock / 64;
    for (int64_t g = (int64_t)blockIdx.x * WPB + threadIdx.x / 64; g < n_groups;
         g += (int64_t)gridDim.x * WPB) {
        int32_t lo = INT32_MAX, hi = -1;
        for (int r = 0; r < kTGroup / 64; r++) {
            const int64_t i = g * kTGroup + r * 64 + lane;
            const int32_t v = i < n ? tpix[i] : -1;
            if (v >= 0) { lo = min(lo, v); hi = max(hi, v); }
        }
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        if (lane == 0) {
            const bool wide = hi >= 0 && hi - lo > (int32_t)kTOff - 1;
            tpb[g] = hi < 0 ? 0 : (wide ? kTWideMark : lo);
            if (wide) atomicAdd(n_wide, 1u);
        }
    }
}

// Offsets of the narrow groups; each wide group takes the next index k
// (in no set order), tpb[g] = -1 - k, wgrp[k] = g and its pixels go to
// tpw[k * kTGroup ...] (-1 past n).  One wave per group.
__global__ __launch_bounds__(kPlanBlock) void k_plan_tpack(const int32_t *tpix, int64_t n,
                                                           int64_t n_groups, int32_t *tpb,
                                                           uint16_t *tpd, int32_t *tpw,
                                                           int32_t *wgrp, uint32_t *next) {
    const int lane = threadIdx.x & 63;
    constexpr int WPB = kPlanBlock / 64;
    for (int64_t g = (int64_t)blockIdx.x * WPB + threadIdx.x / 64; g < n_groups;
         g += (int64_t)gridDim.x * WPB) {
        const int32_t b = tpb[g];
        if (b == kTWideMark) {
            uint32_t k = 0;
            if (lane == 0) {
                k = atomicAdd(next, 1u);
                tpb[g] = -1 - (int32_t)k;
                wgrp[k] = (int32_t)g;
            }
            k = __shfl(k, 0);
            for (int r = 0; r < kTGroup / 64; r++) {
                const int64_t i = g * kTGroup + r * 64 + lane;
                tpw[(int64_t)k * kTGroup + r * 64 + lane] = i < n ? tpix[i] : -1;
            }
        } else {
            for (int r = 0; r < kTGroup / 64; r++) {
                const int64_t i = g * kTGroup + r * 64 + lane;
                if (i < n) {
                    const int32_t v = tpix[i];
                    tpd[i] = v < 0 ? kTOff : (uint16_t)(v - b);
                }
            }
        }
    }
}

// Head bits of the sorted keys, one byte per 8 samples (bit r: sample
// 8 j + r starts a run), and the head count of each 512-sample group.  One
// thread per byte; a wave covers one group.  Grid covers n_groups * 64
// threads exactly.
__global__ __launch_bounds__(kPlanBlock) void k_plan_heads(const int32_t *key, int64_t n_on,
                                                           int64_t n_bytes, uint8_t *heads,
                                                           uint32_t *gcount) {
    const int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (j >= n_bytes) return;   // whole waves: n_bytes is a multiple of 64
    const int64_t s0 = j * 8;
    uint32_t bits = 0;
    if (s0 < n_on) {
        int32_t prev = s0 > 0 ? key[s0 - 1] : -1;
        for (int r = 0; r < 8 && s0 + r < n_on; r++) {
            const int32_t k = key[s0 + r];
            if (s0 + r == 0 || k != prev) bits |= 1u << r;
            prev = k;
        }
    }
    heads[j] = (uint8_t)bits;
    uint32_t c = __popc(bits);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63) == 0) gcount[j / 64] = c;
}

// x / d for x < 2^32, d > 0, via a float reciprocal and one correction
// step: while x / d < 2^22 the float estimate (relative error below 2e-7)
// is off by less than 1, so one step fixes it.  Plans hold fewer than 2^22
// detectors.
struct FastDiv {
    uint32_t d;
    float inv;
};

template <int DIV>
__device__ __forceinline__ uint32_t fast_div(uint32_t x, const FastDiv &f) {
    if constexpr (DIV == 0) {
        return x / f.d;
    } else {
        uint32_t q = (uint32_t)((float)x * f.inv);
        const int32_t r = (int32_t)(x - q * f.d);   // -d < r < 2 d < 2^31 (d < 2^30)
        if (r < 0) q--;
        else if (r >= (int32_t)f.d) q++;
        return q;
    }
}

// Exclusive prefix of v over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    const int lane = (int)__lane_id();
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    return x - v;
}

// run_pix[run] = pixel of the run's first sample.
__global__ __launch_bounds__(kPlanBlock) void k_plan_runs(const int32_t *key, int64_t n_bytes,
                                                          int32_t n_vis, const int32_t *comp2pix,
                                                          const uint8_t *heads,
                                                          const uint32_t *wbase,
                                                          uint32_t *run_pix) {
    const int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (j >= n_bytes) return;   // whole waves: n_bytes is a multiple of 64
    const uint32_t bits = heads[j];
    const uint32_t e = wave_excl_scan(__popc(bits));
    const uint32_t base = wbase[j / 64] + e;
    uint32_t k = 0;
    for (int r = 0; r < 8; r++) {
        if (bits & (1u << r)) {
            run_pix[base + k] = (uint32_t)comp2pix[key[j * 8 + r] % n_vis];
            k++;
        }
    }
}

// Plane of each output value in the output map, in units of the map plane
// (to_map: component c; to_weights: entry (a, b), b >= a, at a * nc + b).
struct WPlanes {
    int32_t o[6];
};

// Run sums and atomics of one lane's nr (<= kPlanR) consecutive sorted
// samples, whose run-start bits are hb and whose first run index is
// lane_base + (bit 0 of hb ? 0 : -1); xf(r, x) gives sample r's NC
// contributions.  The lane sums its samples run by run: its first run may
// continue from earlier lanes and goes through the wave-level run
// aggregation; runs that start inside the lane and end inside it, and its
// last run when it holds more than one, are added to the map directly.
template <int NC, typename XF>
__device__ __forceinline__ void plan_lane_scatter(uint32_t hb, uint32_t lane_base, int nr,
                                                  const uint32_t *__restrict__ run_pix,
                                                  uint32_t plane, const WPlanes &pl,
                                                  double *map, XF &&xf) {
    int64_t hk = -1, ck = -1;
    uint32_t cpix = 0;
    double hv[NC], cv[NC];
    bool closed = false;   // has the lane's first run ended?
#pragma unroll
    for (int c = 0; c < NC; c++) { hv[c] = 0.; cv[c] = 0.; }
#pragma unroll
    for (int r = 0; r < kPlanR; r++) {
        if (r >= nr) break;
        const int64_t rid = (int64_t)lane_base + __popc(hb & ((2u << r) - 1u)) - 1;
        double x[NC];
        xf(r, x);
        if (rid != ck) {
            if (ck >= 0) {
                if (!closed) {
                    hk = ck;
#pragma unroll
                    for (int c = 0; c < NC; c++) hv[c] = cv[c];
                    closed = true;
                } else {
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        unsafeAtomicAdd(map + (int64_t)pl.o[c] * plane + cpix, cv[c]);
                }
            }
            ck = rid;
            cpix = run_pix[rid];
#pragma unroll
            for (int c = 0; c < NC; c++) cv[c] = x[c];
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) cv[c] += x[c];
        }
    }
    uint32_t hpix = cpix;
    if (!closed) {
        hk = ck;
#pragma unroll
        for (int c = 0; c < NC; c++) hv[c] = cv[c];
    } else if (ck >= 0) {
#pragma unroll
        for (int c = 0; c < NC; c++)
            unsafeAtomicAdd(map + (int64_t)pl.o[c] * plane + cpix, cv[c]);
        hpix = hk >= 0 ? run_pix[hk] : 0u;
    }
    if (wave_runs<NC>(hk, hv) && hk >= 0) {
#pragma unroll
        for (int c = 0; c < NC; c++)
            unsafeAtomicAdd(map + (int64_t)pl.o[c] * plane + hpix, hv[c]);
    }
}

// Planned to_map over the compact sorted stream, kPlanR consecutive samples
// per lane.  Each block takes cpb consecutive 2048-sample chunks (adjacent
// blocks stay on adjacent pixels and signal rows); with LDT the per-detector
// tables (T factor, weights) are staged in LDS once per block so their
// lookups, which differ lane to lane, stay off the vector memory path.
template <int NC, bool TC, typename ST, int DIV, bool LDT>
__global__ __launch_bounds__(kPlanBlock, 6) void k_plan_to_map_c(
    const uint32_t *__restrict__ idx, const float *__restrict__ sspin,
    const float *__restrict__ tval, const uint8_t *__restrict__ heads,
    const uint32_t *__restrict__ wbase, const uint32_t *__restrict__ run_pix, int64_t n_on,
    FastDiv n_t, uint32_t plane, const ST *__restrict__ sig,
    const float *__restrict__ det_weights, int32_t n_det, int32_t cpb, double *map) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int R = kPlanR;
    extern __shared__ float s_tab[];   // LDT: [n_det] T factors (TC), [n_det] weights
    const int w_off = TC ? n_det : 0;
    if constexpr (LDT) {
        for (int d = threadIdx.x; d < n_det; d += kPlanBlock) {
            if (TC) s_tab[d] = tval[d];
            if (det_weights != nullptr) s_tab[w_off + d] = det_weights[d];
        }
        __syncthreads();
    }
    const int64_t c0 = (int64_t)blockIdx.x * cpb;
    for (int64_t blk = c0; blk < c0 + cpb && blk * (kPlanBlock * R) < n_on; blk++) {
        const int64_t base = (blk * kPlanBlock + threadIdx.x) * R;
        const int nr = (int)max((int64_t)0, min((int64_t)R, n_on - base));
        uint32_t ii[R];
        float sp[R * (NS > 0 ? NS : 1)];
        if (nr == R) {
#pragma unroll
            for (int q = 0; q < R / 4; q++) {
                const uint4 b = reinterpret_cast<const uint4 *>(idx + base)[q];
                ii[4 * q] = b.x; ii[4 * q + 1] = b.y; ii[4 * q + 2] = b.z; ii[4 * q + 3] = b.w;
            }
            if constexpr (NS > 0) {
#pragma unroll
                for (int q = 0; q < R * NS / 4; q++) {
                    const float4 f = reinterpret_cast<const float4 *>(sspin + base * NS)[q];
                    sp[4 * q] = f.x; sp[4 * q + 1] = f.y; sp[4 * q + 2] = f.z; sp[4 * q + 3] = f.w;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                ii[r] = r < nr ? idx[base + r] : 0u;
#pragma unroll
                for (int c = 0; c < NS; c++)
                    sp[r * NS + c] = r < nr ? sspin[(base + r) * NS + c] : 0.f;
            }
        }
        // Run index of each sample: the runs before the wave's group, the heads
        // of the earlier lanes of the wave, and the heads up to it in this lane.
        const uint32_t hb = nr > 0 ? heads[base / 8] : 0u;
        const uint32_t wb = nr > 0 ? wbase[base / kGroup] : 0u;
        const uint32_t lane_base = wb + wave_excl_scan(__popc(hb));
        const WPlanes pl{{0, 1, 2, 0, 0, 0}};
        plan_lane_scatter<NC>(hb, lane_base, nr, run_pix, plane, pl, map, [&](int r, double *x) {
            const uint32_t si = ii[r];
            uint32_t det = 0;
            if (TC || det_weights != nullptr) det = fast_div<DIV>(si, n_t);
            float pf[NC];
            if constexpr (TC) {
                if constexpr (LDT) pf[0] = s_tab[det];
                else pf[0] = tval[det];
#pragma unroll
                for (int c = 1; c < NC; c++) pf[c] = sp[r * NS + c - 1];
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) pf[c] = sp[r * NS + c];
            }
            float w = 1.f;
            if (det_weights != nullptr) {
                if constexpr (LDT) w = s_tab[w_off + det];
                else w = det_weights[det];
            }
            if (w != 0.f) {  // a zero weight skips the detector (Projection.cxx:2104)
                const ST sg = sig[si];
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    if constexpr (std::is_same<ST, double>::value)
                        x[c] = (double)w * sg * (double)pf[c];
                    else
                        x[c] = (double)(w * sg * pf[c]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) x[c] = 0.;
            }
        });
    }
}

// Planned to_weights over the compact sorted stream (pixel-run layout):
// per sample the NC (NC + 1) / 2 upper-triangle products (pf_a * pf_b) * w
// in float32 (Projection.cxx:2135-2154), run sums and atomics as in
// k_plan_to_map_c, into the (NC, NC, ...) map at planes a * NC + b.
template <int NC, bool TC, int DIV>
__global__ __launch_bounds__(kPlanBlock, 6) void k_plan_to_weights_c(
    const uint32_t *__restrict__ idx, const float *__restrict__ sspin,
    const float *__restrict__ tval, const uint8_t *__restrict__ heads,
    const uint32_t *__restrict__ wbase, const uint32_t *__restrict__ run_pix, int64_t n_on,
    FastDiv n_t, uint32_t plane, const float *__restrict__ det_weights, int32_t cpb,
    double *map) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int NW = NC * (NC + 1) / 2;
    constexpr int R = kPlanR;
    WPlanes pl{{0, 0, 0, 0, 0, 0}};
    {
        int k = 0;
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
            for (int b = a; b < NC; b++) pl.o[k++] = a * NC + b;
    }
    const int64_t c0 = (int64_t)blockIdx.x * cpb;
    for (int64_t blk = c0; blk < c0 + cpb && blk * (kPlanBlock * R) < n_on; blk++) {
        const int64_t base = (blk * kPlanBlock + threadIdx.x) * R;
        const int nr = (int)max((int64_t)0, min((int64_t)R, n_on - base));
        uint32_t ii[R];
        float sp[R * (NS > 0 ? NS : 1)];
#pragma unroll
        for (int r = 0; r < R; r++) {
            ii[r] = r < nr ? idx[base + r] : 0u;
#pragma unroll
            for (int c = 0; c < NS; c++) sp[r * NS + c] = r < nr ? sspin[(base + r) * NS + c] : 0.f;
        }
        const uint32_t hb = nr > 0 ? heads[base / 8] : 0u;
        const uint32_t wb = nr > 0 ? wbase[base / kGroup] : 0u;
        const uint32_t lane_base = wb + wave_excl_scan(__popc(hb));
        plan_lane_scatter<NW>(hb, lane_base, nr, run_pix, plane, pl, map, [&](int r, double *x) {
            const uint32_t si = ii[r];
            uint32_t det = 0;
            if (TC || det_weights != nullptr) det = fast_div<DIV>(si, n_t);
            float pf[NC];
            if constexpr (TC) {
                pf[0] = tval[det];
#pragma unroll
                for (int c = 1; c < NC; c++) pf[c] = sp[r * NS + c - 1];
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) pf[c] = sp[r * NS + c];
            }
            const float w = det_weights != nullptr ? det_weights[det] : 1.f;
            int k = 0;
#pragma unroll
            for (int a = 0; a < NC; a++)
#pragma unroll
                for (int b = a; b < NC; b++, k++)
                    x[k] = w != 0.f ? (double)(pf[a] * pf[b] * w) : 0.;   // zero weight: skipped
        });
    }
}

// Pixel-major copy of the visited pixels of an (NC, plane) map, 4 doubles
// per pixel, in compact order (comp2pix[i] = flat pixel of compact index i).
template <int NC>
__global__ __launch_bounds__(kPlanBlock) void k_plan_pack(const double *map, int64_t plane,
                                                          const int32_t *comp2pix, int64_t n_vis,
                                                          double *packed) {
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n_vis;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t p = comp2pix[i];
        double v[4] = {0., 0., 0., 0.};
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = map[c * plane + p];
        reinterpret_cast<double2 *>(packed + 4 * i)[0] = double2{v[0], v[1]};
        reinterpret_cast<double2 *>(packed + 4 * i)[1] = double2{v[2], v[3]};
    }
}

// Tiled maps: split each visited pixel's concatenated number into its map
// tile (comp_tile) and its offset in the tile (written over comp2pix).
__global__ __launch_bounds__(kPlanBlock) void k_plan_comp_tiles(int32_t *comp2pix, int64_t n_vis,
                                                                PixGeom pg, int32_t *comp_tile) {
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n_vis;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t p = comp2pix[i];
        int lo = 0, hi = pg.n_tile;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (pg.tbase[mid] <= p) lo = mid;
            else hi = mid;
        }
        comp_tile[i] = lo;
        comp2pix[i] = (int32_t)(p - pg.tbase[lo]);
    }
}

// Pixel-major copy of the visited pixels of a tiled map (tiles[t]: the
// (NC, rows, cols) tile t).
template <int NC>
__global__ __launch_bounds__(kPlanBlock) void k_plan_pack_tiles(const double *const *tiles,
                                                                const int32_t *tdims,
                                                                const int32_t *comp_tile,
                                                                const int32_t *comp2pix,
                                                                int64_t n_vis, double *packed) {
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n_vis;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int32_t t = comp_tile[i];
        const int64_t pl = (int64_t)tdims[2 * t] * tdims[2 * t + 1];
        const double *base = tiles[t] + comp2pix[i];
        double v[4] = {0., 0., 0., 0.};
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = base[c * pl];
        reinterpret_cast<double2 *>(packed + 4 * i)[0] = double2{v[0], v[1]};
        reinterpret_cast<double2 *>(packed + 4 * i)[1] = double2{v[2], v[3]};
    }
}

// Compact pixel of a time-ordered offset (-1 off the map).
__device__ __forceinline__ int32_t tpix_of(uint32_t d, int32_t base) {
    return d == kTOff ? -1 : base + (int32_t)d;
}

// Planned from_map over the time-ordered copy: V consecutive samples per
// lane (V = 8 with vector loads when n % 8 == 0, else V = 1), narrow groups
// only (the wide ones are left to k_plan_from_map_s<..., true>).  sr = sum_c
// pf_c * map_c accumulated into a float (each addition rounded to float),
// then sig = accumulate ? sig + sr : sr, as ProjEng_Precomp's from_map does;
// off-map samples keep sig (or 0).
template <int NC, bool TC, int V, int DIV>
__global__ __launch_bounds__(kPlanBlock) void k_plan_from_map(
    const uint16_t *__restrict__ tpd, const int32_t *__restrict__ tpb,
    const float *__restrict__ tspin,
    const float *__restrict__ tval, const double *__restrict__ packed, int64_t n,
    FastDiv n_t, float *__restrict__ sig, int accumulate) {
    constexpr int NS = TC ? NC - 1 : NC;
    for (int64_t i0 = ((int64_t)blockIdx.x * kPlanBlock + threadIdx.x) * V; i0 < n;
         i0 += (int64_t)gridDim.x * kPlanBlock * V) {
        // a wave's 64 V samples lie in one group
        const int32_t b = tpb[__builtin_amdgcn_readfirstlane((int32_t)(i0 / kTGroup))];
        if (b < 0) continue;
        int32_t px[V];
        float sp[V * (NS > 0 ? NS : 1)];
        float o[V];
        if constexpr (V == 8) {
            const uint4 a = *reinterpret_cast<const uint4 *>(tpd + i0);
            const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                px[2 * q] = tpix_of(w[q] & 0xFFFFu, b);
                px[2 * q + 1] = tpix_of(w[q] >> 16, b);
            }
        } else if constexpr (V == 4) {
            const uint2 a = *reinterpret_cast<const uint2 *>(tpd + i0);
            px[0] = tpix_of(a.x & 0xFFFFu, b); px[1] = tpix_of(a.x >> 16, b);
            px[2] = tpix_of(a.y & 0xFFFFu, b); px[3] = tpix_of(a.y >> 16, b);
        }
        if constexpr (V > 1) {
            if constexpr (NS > 0) {
#pragma unroll
                for (int q = 0; q < V * NS / 4; q++) {
                    const float4 f = reinterpret_cast<const float4 *>(tspin + i0 * NS)[q];
                    sp[4 * q] = f.x; sp[4 * q + 1] = f.y; sp[4 * q + 2] = f.z; sp[4 * q + 3] = f.w;
                }
            }
#pragma unroll
            for (int q = 0; q < V / 4; q++) {
                const float4 s = accumulate ? reinterpret_cast<const float4 *>(sig + i0)[q]
                                            : float4{0.f, 0.f, 0.f, 0.f};
                o[4 * q] = s.x; o[4 * q + 1] = s.y; o[4 * q + 2] = s.z; o[4 * q + 3] = s.w;
            }
        } else {
            px[0] = tpix_of(tpd[i0], b);
#pragma unroll
            for (int c = 0; c < NS; c++) sp[c] = tspin[i0 * NS + c];
            o[0] = accumulate ? sig[i0] : 0.f;
        }
        // Detector of the first sample; a later sample of the lane belongs to
        // the next detector once it passes the row end (V <= n_t is not
        // assumed: the step is recomputed per sample from t0 + r).
        uint32_t d0 = 0, t0 = 0;
        if constexpr (TC) {
            d0 = fast_div<DIV>((uint32_t)i0, n_t);
            t0 = (uint32_t)i0 - d0 * n_t.d;
        }
        double mv[V][4];
#pragma unroll
        for (int r = 0; r < V; r++) {
            const int64_t p = px[r] >= 0 ? px[r] : 0;
            const double2 lo = reinterpret_cast<const double2 *>(packed + 4 * p)[0];
            mv[r][0] = lo.x; mv[r][1] = lo.y;
            if (NC > 2) {
                const double2 hi = reinterpret_cast<const double2 *>(packed + 4 * p)[1];
                mv[r][2] = hi.x; mv[r][3] = hi.y;
            }
        }
#pragma unroll
        for (int r = 0; r < V; r++) {
            if (px[r] < 0) continue;
            float pf[NC];
            if constexpr (TC) {
                const uint32_t k = t0 + r;
                pf[0] = tval[d0 + (k < n_t.d ? 0u : k / n_t.d)];
#pragma unroll
                for (int c = 1; c < NC; c++) pf[c] = sp[r * NS + c - 1];
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) pf[c] = sp[r * NS + c];
            }
            float sr = 0;
#pragma unroll
            for (int c = 0; c < NC; c++) sr = sr + pf[c] * mv[r][c];
            o[r] = accumulate ? o[r] + sr : sr;
        }
        if constexpr (V > 1) {
#pragma unroll
            for (int q = 0; q < V / 4; q++)
                reinterpret_cast<float4 *>(sig + i0)[q] =
                    float4{o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]};
        } else {
            sig[i0] = o[0];
        }
    }
}

// Lane-strided planned from_map: a wave takes 64·V consecutive samples and
// lane l the samples l, l + 64, ..., so one gather instruction covers 64
// consecutive samples of a detector (a few map cache lines) instead of 64·V.
// WIDE = false takes the narrow groups (u16 offsets from the group's base;
// a wide group's waves leave after issuing their loads), WIDE = true the
// n_wide groups listed in wgrp (int32 pixels).  Arithmetic and off-map handling are those of
// k_plan_from_map.
template <int NC, bool TC, int V, int DIV, bool WIDE>
__global__ __launch_bounds__(kPlanBlock) void k_plan_from_map_s(
    const uint16_t *__restrict__ tpd, const int32_t *__restrict__ tpb,
    const int32_t *__restrict__ tpw, const int32_t *__restrict__ wgrp, int64_t n_wide,
    const float *__restrict__ tspin,
    const float *__restrict__ tval, const double *__restrict__ packed, int64_t n,
    FastDiv n_t, float *__restrict__ sig, int accumulate) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int WPB = kPlanBlock / 64;
    constexpr int WPG = kTGroup / (64 * V);   // waves per group
    const int lane = threadIdx.x & 63;
    const int64_t n_waves = WIDE ? n_wide * WPG : (n + 64 * V - 1) / (64 * V);
    for (int64_t q = (int64_t)blockIdx.x * WPB + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
         q < n_waves; q += (int64_t)gridDim.x * WPB) {
        int64_t w = q, woff = 0;   // woff: tpw index minus sample index
        int32_t b = 0;
        if constexpr (WIDE) {
            const int64_t k = q / WPG;
            const int64_t g = wgrp[k];
            w = g * WPG + q % WPG;
            woff = (k - g) * kTGroup;
        } else {
            b = tpb[w / WPG];
        }
        const int64_t i0 = w * 64 * V + lane;
        int32_t px[V];
        uint32_t dv[V];
        float sp[V * (NS > 0 ? NS : 1)];
        float o[V];
#pragma unroll
        for (int r = 0; r < V; r++) {
            const int64_t i = i0 + 64 * r;
            const bool ok = i < n;
            if constexpr (WIDE)
                px[r] = ok ? tpw[i + woff] : -1;
            else
                dv[r] = ok ? (uint32_t)__builtin_nontemporal_load(tpd + i) : kTOff;
            if constexpr (NS == 2) {
                const f32x2 f = ok ? __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(tspin) + i)
                                   : f32x2{0.f, 0.f};
                sp[2 * r] = f.x; sp[2 * r + 1] = f.y;
            } else {
#pragma unroll
                for (int c = 0; c < NS; c++) sp[r * NS + c] = ok ? tspin[i * NS + c] : 0.f;
            }
            o[r] = (accumulate && ok) ? sig[i] : 0.f;
        }
        // the group's base is tested once the streams' loads are in flight
        // (a wide group's wave leaves here, its loads unused)
        if constexpr (!WIDE) {
            if (b < 0) continue;
#pragma unroll
            for (int r = 0; r < V; r++) px[r] = tpix_of(dv[r], b);
        }
        double mv[V][4];
#pragma unroll
        for (int r = 0; r < V; r++) {
            const int64_t p = px[r] >= 0 ? px[r] : 0;
            const double2 lo = reinterpret_cast<const double2 *>(packed + 4 * p)[0];
            mv[r][0] = lo.x; mv[r][1] = lo.y;
            if (NC > 2) {
                const double2 hi = reinterpret_cast<const double2 *>(packed + 4 * p)[1];
                mv[r][2] = hi.x; mv[r][3] = hi.y;
            }
        }
#pragma unroll
        for (int r = 0; r < V; r++) {
            const int64_t i = i0 + 64 * r;
            if (i >= n) continue;
            if (px[r] >= 0) {
                float pf[NC];
                if constexpr (TC) {
                    pf[0] = tval[fast_div<DIV>((uint32_t)i, n_t)];
#pragma unroll
                    for (int c = 1; c < NC; c++) pf[c] = sp[r * NS + c - 1];
                } else {
#pragma unroll
                    for (int c = 0; c < NC; c++) pf[c] = sp[r * NS + c];
                }
                float sr = 0;
#pragma unroll
                for (int c = 0; c < NC; c++) sr = sr + pf[c] * mv[r][c];
                o[r] = accumulate ? o[r] + sr : sr;
            }
            __builtin_nontemporal_store(o[r], sig + i);
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled layout.  Samples are sorted by (time block, map tile) -- 32 x 64
// pixels of a flat map, or 2048 consecutive pixel numbers of a HEALPix map
// (two aligned 32 x 32 sub-squares in NEST order) -- and stay in sample order inside a tile, so a wave's 64 positions fall in
// one or two stretches of consecutive samples of one detector and its signal
// loads are nearly contiguous.  Work items (tile, positions [b, e)) sum into
// or gather from the tile's values held in LDS.
//
// Per position: the slot within the tile (u16) and the stored factors.  The
// sample index of position j is encoded by stretches: bit j % 64 of
// tmask[j / 64] marks a start -- a position that does not continue the
// previous one's stretch (different tile or time block, or not the next
// sample), and every chunk's first position --, thead[j / 64] counts the
// starts before the chunk and head_i lists each start's sample.  With a
// start at every chunk's first position a lane always finds one at or before
// it, so the lookup has no branch and a wave's loads for several chunks
// overlap.
// ---------------------------------------------------------------------------
#ifndef SO3G_TILE_BITS_Y
#define SO3G_TILE_BITS_Y 5
#endif
#ifndef SO3G_TILE_BITS_X
#define SO3G_TILE_BITS_X 6
#endif
constexpr int kTileBY = SO3G_TILE_BITS_Y;       // 32 x 64 tiles of a flat map
constexpr int kTileBX = SO3G_TILE_BITS_X;
constexpr int kTS = 1 << (kTileBY + kTileBX);  // slots per tile
// Plan tile shape, read through so3g_get_option ("plan_tile_rows", "_cols").
extern const int g_plan_tile_rows = 1 << kTileBY;
extern const int g_plan_tile_cols = 1 << kTileBX;
constexpr int kTBlock = 1024;                  // threads per work item

struct TItem {
    int64_t b;      // first sorted position
    int32_t n;      // positions
    int32_t tile;   // compact tile id
};

// Sample index of this lane's position in a 64-position chunk (bit 0 of
// mask is always set).
__device__ __forceinline__ uint32_t tile_sample(uint64_t mask, uint32_t hbase,
                                                const uint32_t *__restrict__ head_i, int lane) {
    const uint64_t h = mask & (~0ull >> (63 - lane));
    const int p = 63 - __clzll((long long)h);
    return head_i[hbase + __popcll(h) - 1] + (uint32_t)(lane - p);
}

// Map tile of a sample (-1 off the map) before compaction and its slot.  k = 2:
// 32 x 64 pixel tiles, numbered row-major over a gtx-wide tile grid; k = 1:
// blocks of kTS consecutive pixel numbers (in HEALPix NEST order an aligned
// pair of adjacent 32 x 32 sub-squares of a base face).
__device__ __forceinline__ int64_t flat_tile(const int32_t *pix, const PixGeom &pg, int64_t i,
                                             int64_t ny, int64_t nx, int64_t gtx,
                                             int32_t &slot) {
    if (pg.k == 3) {
        // plan-tile-sized sub-blocks of each map tile, numbered from sbase[t]
        int32_t t, iy, ix, tx;
        slot = 0;
        if (!tile_pos(pix, pg, i, t, iy, ix, tx)) return -1;
        slot = ((iy & ((1 << kTileBY) - 1)) << kTileBX) | (ix & ((1 << kTileBX) - 1));
        const int32_t nbx = (tx + (1 << kTileBX) - 1) >> kTileBX;
        return (int64_t)pg.sbase[t] + (iy >> kTileBY) * nbx + (ix >> kTileBX);
    }
    if (pg.k == 1) {
        const int32_t q = pix[i];
        slot = q & (kTS - 1);
        return (q >= 0 && q < ny) ? (int64_t)(q >> (kTileBY + kTileBX)) : -1;
    }
    const int2 q = *reinterpret_cast<const int2 *>(pix + 2 * i);
    slot = ((q.x & ((1 << kTileBY) - 1)) << kTileBX) | (q.y & ((1 << kTileBX) - 1));
    if (q.x < 0 || q.x >= ny || q.y < 0 || q.y >= nx) return -1;
    return (int64_t)(q.x >> kTileBY) * gtx + (q.y >> kTileBX);
}

// Tiles touched by the samples (byte flags; every writer stores 1).
__global__ __launch_bounds__(kPlanBlock) void k_tp_tmark(const int32_t *pix, PixGeom pg, int64_t n,
                                                         int64_t ny, int64_t nx, int64_t gtx,
                                                         uint8_t *used) {
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kPlanBlock) {
        int32_t sl;
        const int64_t tf = flat_tile(pix, pg, i, ny, nx, gtx, sl);
        if (tf >= 0) used[tf] = 1;
    }
}

// Compact tile (tile_of maps the flat tile number to the visited tiles'
// numbering, -1 off the map) and slot of sample i.
__device__ __forceinline__ void tile_slot(const int32_t *pix, const PixGeom &pg, int64_t i,
                                          int64_t ny, int64_t nx, int64_t gtx,
                                          const int32_t *tile_of, int32_t &tile,
                                          int32_t &slot) {
    const int64_t tf = flat_tile(pix, pg, i, ny, nx, gtx, slot);
    tile = tf >= 0 ? tile_of[tf] : -1;
}

__global__ __launch_bounds__(kPlanBlock) void k_tp_keys(const int32_t *pix, PixGeom pg, int64_t n,
                                                        int64_t ny, int64_t nx, int64_t gtx,
                                                        const int32_t *tile_of, int64_t n_t,
                                                        int64_t t_block, int32_t n_tiles,
                                                        int32_t off_key, int32_t *keys,
                                                        uint32_t *vals) {
    const int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (i >= n) return;
    int32_t tile, slot;
    tile_slot(pix, pg, i, ny, nx, gtx, tile_of, tile, slot);
    const int64_t tb = (i % n_t) / t_block;
    keys[i] = tile >= 0 ? (int32_t)(tb * n_tiles + tile) : off_key;
    vals[i] = (uint32_t)i;
}

// Slot and stored factors of each sorted position.
__global__ __launch_bounds__(kPlanBlock) void k_tp_fill(const int32_t *pix, PixGeom pg,
                                                        const float *spin, int nc, int c0,
                                                        const uint32_t *idx, int64_t n_on,
                                                        int64_t ny, int64_t nx, int64_t gtx,
                                                        const int32_t *tile_of, uint16_t *slot,
                                                        float *fac) {
    const int ns = nc - c0;
    for (int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; j < n_on;
         j += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t i = idx[j];
        int32_t tile, sl;
        tile_slot(pix, pg, i, ny, nx, gtx, tile_of, tile, sl);
        slot[j] = (uint16_t)sl;
        for (int c = 0; c < ns; c++) fac[j * ns + c] = spin[i * nc + c0 + c];
    }
}

// Per 64-position chunk: stretch-start mask and start count.  One wave per
// chunk; the grid covers n_chunks * 64 threads exactly.
__global__ __launch_bounds__(kPlanBlock) void k_tp_chunks(const int32_t *key,
                                                          const uint32_t *idx, int64_t n_on,
                                                          int64_t n_chunks, uint64_t *mask,
                                                          uint32_t *cnt) {
    const int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    const int64_t c = j >> 6;
    if (c >= n_chunks) return;   // whole waves
    const int lane = (int)(j & 63);
    bool head = false;
    if (j < n_on)
        head = lane == 0 || key[j] != key[j - 1] || idx[j] != idx[j - 1] + 1u;
    const uint64_t m = __ballot(head);
    if (lane == 0) {
        mask[c] = m;
        cnt[c] = (uint32_t)__popcll(m);
    }
}

__global__ __launch_bounds__(kPlanBlock) void k_tp_heads(const uint64_t *mask,
                                                         const uint32_t *hbase,
                                                         const uint32_t *idx, int64_t n_on,
                                                         uint32_t *head_i) {
    for (int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; j < n_on;
         j += (int64_t)gridDim.x * kPlanBlock) {
        const uint64_t m = mask[j >> 6];
        const int l = (int)(j & 63);
        if ((m >> l) & 1ull)
            head_i[hbase[j >> 6] + __popcll(m & ((1ull << l) - 1ull))] = idx[j];
    }
}

// seg[v] = first sorted position with key >= v, v in [0, n_keys].
__global__ __launch_bounds__(kPlanBlock) void k_tp_segs(const int32_t *key, int64_t n_on,
                                                        int32_t n_keys, int64_t *seg) {
    const int64_t v = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (v > n_keys) return;
    int64_t lo = 0, hi = n_on;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (key[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    seg[v] = lo;
}

// Flat pixel of each (compact tile, slot), -1 for slots off the map.
// k = 3: the slot's offset in its map tile (row * cols + col), and per plan
// tile the map tile (tile_map) -- the last t with sbase[t] <= sub-block, so
// that tiles without sub-blocks are skipped.
__global__ __launch_bounds__(kPlanBlock) void k_tp_tilepix(PixGeom pg, int64_t ny, int64_t nx,
                                                           int64_t gtx,
                                                           const int32_t *tile_flat,
                                                           int64_t n_slots,
                                                           int32_t *tile_pix,
                                                           int32_t *tile_map) {
    for (int64_t q = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; q < n_slots;
         q += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t tf = tile_flat[q >> (kTileBY + kTileBX)];
        const int sl = (int)(q & (kTS - 1));
        if (tf < 0) {
            tile_pix[q] = -1;
        } else if (pg.k == 3) {
            int lo = 0, hi = pg.n_tile;
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (pg.sbase[mid] <= tf) lo = mid;
                else hi = mid;
            }
            const int32_t ty = pg.tdims[2 * lo], tx = pg.tdims[2 * lo + 1];
            const int32_t nbx = (tx + (1 << kTileBX) - 1) >> kTileBX;
            const int32_t local = (int32_t)(tf - pg.sbase[lo]);
            const int32_t iy = (local / nbx) * (1 << kTileBY) + (sl >> kTileBX);
            const int32_t ix = (local % nbx) * (1 << kTileBX) + (sl & ((1 << kTileBX) - 1));
            tile_pix[q] = (iy < ty && ix < tx) ? iy * tx + ix : -1;
            if (sl == 0) tile_map[q >> (kTileBY + kTileBX)] = lo;
        } else if (pg.k == 1) {
            const int64_t px = tf * kTS + sl;
            tile_pix[q] = px < ny ? (int32_t)px : -1;
        } else {
            const int64_t iy = (tf / gtx) * (1 << kTileBY) + (sl >> kTileBX);
            const int64_t ix = (tf % gtx) * (1 << kTileBX) + (sl & ((1 << kTileBX) - 1));
            tile_pix[q] = (iy < ny && ix < nx) ? (int32_t)(iy * nx + ix) : -1;
        }
    }
}

// Tiled to_map: one work item per block; contributions summed in the
// item's tile in LDS (float64 LDS atomics), then added to the map, one
// atomic per touched slot and component.  Each wave takes kTU consecutive
// 64-position chunks at a time and issues each stage's loads for all of them
// before using any (chunk tables, then stretch starts, slots and factors,
// then the signal), so the dependent loads overlap.  LDT: per-detector
// tables in LDS.
#ifndef SO3G_TILE_U
#define SO3G_TILE_U 8
#endif
constexpr int kTU = SO3G_TILE_U;
// Copies of the tile sums in LDS: lane parity picks the copy, so that two
// neighbouring lanes (often the same pixel) never add to one address in one
// instruction; the flush adds the copies in order.
#ifndef SO3G_TILE_COPIES
#define SO3G_TILE_COPIES 2
#endif
constexpr int kTCopies = SO3G_TILE_COPIES;
// Doubles between the copies: a copy holds NC x 2048 doubles, a multiple of
// the LDS bank period (128 bytes), so without a gap a slot's two copies sit in
// the same bank and neighbouring lanes adding to one pixel (consecutive
// samples, alternate copies) conflict.  64 bytes put them 16 banks apart
// (HEALPix nside 4096 scatter 2.55 -> 2.12 ms, CAR 2.05 -> 1.99 ms; LDS
// bank-conflict cycles 4.8e8 -> 2.1e8; gaps of 0 / 4 / 8 / 12 / 16 / 24
// doubles: 2.55 / 2.74 / 2.12 / 2.74 / 2.52 / 2.12 ms on HEALPix).
constexpr int kTPad = 8;

template <int NC, bool TC, typename ST, bool LDT>
__global__ __launch_bounds__(kTBlock) void k_tp_to_map(
    const TItem *__restrict__ items, const uint16_t *__restrict__ slot,
    const float *__restrict__ fac, const uint64_t *__restrict__ tmask,
    const uint32_t *__restrict__ thead,
    const uint32_t *__restrict__ head_i, const int32_t *__restrict__ tile_pix,
    const float *__restrict__ tval, FastDiv n_t, uint32_t plane, const ST *__restrict__ sig,
    const float *__restrict__ det_weights, int32_t n_det, double *map,
    double *const *__restrict__ tiles, const int32_t *__restrict__ tile_map,
    const int32_t *__restrict__ tdims) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int CS = NC * kTS + kTPad;   // doubles per copy and gap
    __shared__ double acc[kTCopies * CS];
    extern __shared__ float s_tab[];   // LDT: [n_det] T factors (TC), [n_det] weights
    const int w_off = TC ? n_det : 0;
    const TItem it = items[blockIdx.x];
    for (int q = threadIdx.x; q < kTCopies * CS; q += kTBlock) acc[q] = 0.;
    if constexpr (LDT) {
        for (int d = threadIdx.x; d < n_det; d += kTBlock) {
            if (TC) s_tab[d] = tval[d];
            if (det_weights != nullptr) s_tab[w_off + d] = det_weights[d];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t e = it.b + it.n;
    const int64_t c_end = (e + 63) >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t cb = (it.b >> 6) + (int64_t)wave * kTU; cb < c_end;
         cb += (int64_t)(kTBlock / 64) * kTU) {
        uint64_t m[kTU];
        uint32_t hb[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            const int64_t c = min(cb + u, c_end - 1);
            m[u] = tmask[c]; hb[u] = thead[c];
        }
        uint32_t si[kTU];
        int sl[kTU];
        float sp[kTU][NS > 0 ? NS : 1];
        bool act[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            // every lane loads (inactive lanes at the item's first position)
            const int64_t j0 = (cb + u) * 64 + lane;
            act[u] = j0 >= it.b && j0 < e;
            const int64_t j = act[u] ? j0 : it.b;
            const uint32_t ts = tile_sample(m[u], hb[u], head_i, lane);
            si[u] = act[u] ? ts : 0u;
            sl[u] = slot[j];
            if constexpr (NS == 2) {
                const f32x2 v = reinterpret_cast<const f32x2 *>(fac)[j];
                sp[u][0] = v.x; sp[u][1] = v.y;
            } else {
#pragma unroll
                for (int q = 0; q < NS; q++) sp[u][q] = fac[j * NS + q];
            }
        }
        ST sg[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) sg[u] = sig[si[u]];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            if (!act[u]) continue;
            float pf[NC];
            uint32_t det = 0;
            if (TC || det_weights != nullptr) det = fast_div<1>(si[u], n_t);
            if constexpr (TC) {
                if constexpr (LDT) pf[0] = s_tab[det];
                else pf[0] = tval[det];
#pragma unroll
                for (int q = 1; q < NC; q++) pf[q] = sp[u][q - 1];
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) pf[q] = sp[u][q];
            }
            float w = 1.f;
            if (det_weights != nullptr) {
                if constexpr (LDT) w = s_tab[w_off + det];
                else w = det_weights[det];
            }
            if (w == 0.f) continue;   // a zero weight skips the detector (Projection.cxx:2104)
#pragma unroll
            for (int q = 0; q < NC; q++) {
                double x;
                if constexpr (std::is_same<ST, double>::value)
                    x = (double)w * sg[u] * (double)pf[q];
                else
                    x = (double)(w * sg[u] * pf[q]);
                atomicAdd(acc + (lane & (kTCopies - 1)) * CS + q * kTS + sl[u], x);
            }
        }
    }
    __syncthreads();
    const int32_t *tp = tile_pix + (int64_t)it.tile * kTS;
    double *mbase = map;     // non-tiled: the map; tiled: the item's map tile
    int64_t pl = plane;
    if (tiles != nullptr) {
        const int32_t mt = tile_map[it.tile];
        mbase = tiles[mt];
        pl = (int64_t)tdims[2 * mt] * tdims[2 * mt + 1];
    }
    for (int q = threadIdx.x; q < kTS; q += kTBlock) {
        const int32_t px = tp[q];
        if (px < 0) continue;
#pragma unroll
        for (int cc = 0; cc < NC; cc++) {
            double v = acc[cc * kTS + q];
#pragma unroll
            for (int cp = 1; cp < kTCopies; cp++) v += acc[cp * CS + cc * kTS + q];
            if (v != 0.) unsafeAtomicAdd(mbase + cc * pl + px, v);
        }
    }
}

// Tiled to_weights (Precomp to_weight_map, Projection.cxx:2135-2154): per
// position (pf_a * pf_b) * det_wt in float32 for b >= a, summed per work item
// in the tile held in LDS (one copy: NC (NC + 1) / 2 planes), then added to
// the (NC, NC, ...) weights map's upper triangle, one atomic per touched slot
// and entry.  Loads as k_tp_to_map (no signal).
template <int NC, bool TC, bool LDT>
__global__ __launch_bounds__(kTBlock) void k_tp_to_weights(
    const TItem *__restrict__ items, const uint16_t *__restrict__ slot,
    const float *__restrict__ fac, const uint64_t *__restrict__ tmask,
    const uint32_t *__restrict__ thead, const uint32_t *__restrict__ head_i,
    const int32_t *__restrict__ tile_pix, const float *__restrict__ tval, FastDiv n_t,
    uint32_t plane, const float *__restrict__ det_weights, int32_t n_det, double *map,
    double *const *__restrict__ tiles, const int32_t *__restrict__ tile_map,
    const int32_t *__restrict__ tdims) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int NW = NC * (NC + 1) / 2;
    __shared__ double acc[NW * kTS];
    extern __shared__ float s_tab[];   // LDT: [n_det] T factors (TC), [n_det] weights
    const int w_off = TC ? n_det : 0;
    const TItem it = items[blockIdx.x];
    for (int q = threadIdx.x; q < NW * kTS; q += kTBlock) acc[q] = 0.;
    if constexpr (LDT) {
        for (int d = threadIdx.x; d < n_det; d += kTBlock) {
            if (TC) s_tab[d] = tval[d];
            if (det_weights != nullptr) s_tab[w_off + d] = det_weights[d];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t e = it.b + it.n;
    const int64_t c_end = (e + 63) >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t cb = (it.b >> 6) + (int64_t)wave * kTU; cb < c_end;
         cb += (int64_t)(kTBlock / 64) * kTU) {
        uint64_t m[kTU];
        uint32_t hb[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            const int64_t c = min(cb + u, c_end - 1);
            m[u] = tmask[c]; hb[u] = thead[c];
        }
        uint32_t si[kTU];
        int sl[kTU];
        float sp[kTU][NS > 0 ? NS : 1];
        bool act[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            const int64_t j0 = (cb + u) * 64 + lane;
            act[u] = j0 >= it.b && j0 < e;
            const int64_t j = act[u] ? j0 : it.b;
            const uint32_t ts = tile_sample(m[u], hb[u], head_i, lane);
            si[u] = act[u] ? ts : 0u;
            sl[u] = slot[j];
            if constexpr (NS == 2) {
                const f32x2 v = reinterpret_cast<const f32x2 *>(fac)[j];
                sp[u][0] = v.x; sp[u][1] = v.y;
            } else {
#pragma unroll
                for (int q = 0; q < NS; q++) sp[u][q] = fac[j * NS + q];
            }
        }
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            if (!act[u]) continue;
            float pf[NC];
            uint32_t det = 0;
            if (TC || det_weights != nullptr) det = fast_div<1>(si[u], n_t);
            if constexpr (TC) {
                if constexpr (LDT) pf[0] = s_tab[det];
                else pf[0] = tval[det];
#pragma unroll
                for (int q = 1; q < NC; q++) pf[q] = sp[u][q - 1];
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) pf[q] = sp[u][q];
            }
            float w = 1.f;
            if (det_weights != nullptr) {
                if constexpr (LDT) w = s_tab[w_off + det];
                else w = det_weights[det];
            }
            if (w == 0.f) continue;   // a zero weight skips the detector (Projection.cxx:2139)
            int k = 0;
#pragma unroll
            for (int a = 0; a < NC; a++)
#pragma unroll
                for (int b = a; b < NC; b++, k++)
                    atomicAdd(acc + k * kTS + sl[u], (double)(pf[a] * pf[b] * w));
        }
    }
    __syncthreads();
    const int32_t *tp = tile_pix + (int64_t)it.tile * kTS;
    double *mbase = map;
    int64_t pl = plane;
    if (tiles != nullptr) {
        const int32_t mt = tile_map[it.tile];
        mbase = tiles[mt];
        pl = (int64_t)tdims[2 * mt] * tdims[2 * mt + 1];
    }
    for (int q = threadIdx.x; q < kTS; q += kTBlock) {
        const int32_t px = tp[q];
        if (px < 0) continue;
        int k = 0;
#pragma unroll
        for (int a = 0; a < NC; a++)
#pragma unroll
            for (int b = a; b < NC; b++, k++) {
                const double v = acc[k * kTS + q];
                if (v != 0.) unsafeAtomicAdd(mbase + (a * NC + b) * pl + px, v);
            }
    }
}

// Tiled from_map: the item's tile values are loaded into LDS, then each
// position's signal is written (or accumulated into), kTU chunks per wave
// at a time as in k_tp_to_map; arithmetic as k_plan_from_map.
template <int NC, bool TC>
__global__ __launch_bounds__(kTBlock) void k_tp_from_map(
    const TItem *__restrict__ items, const uint16_t *__restrict__ slot,
    const float *__restrict__ fac, const uint64_t *__restrict__ tmask,
    const uint32_t *__restrict__ thead,
    const uint32_t *__restrict__ head_i, const int32_t *__restrict__ tile_pix,
    const float *__restrict__ tval, FastDiv n_t, uint32_t plane, const double *__restrict__ map,
    float *__restrict__ sig, int accumulate, const double *const *__restrict__ tiles,
    const int32_t *__restrict__ tile_map, const int32_t *__restrict__ tdims) {
    constexpr int NS = TC ? NC - 1 : NC;
    __shared__ double mv[NC * kTS];
    const TItem it = items[blockIdx.x];
    const int32_t *tp = tile_pix + (int64_t)it.tile * kTS;
    const double *mbase = map;
    int64_t pl = plane;
    if (tiles != nullptr) {
        const int32_t mt = tile_map[it.tile];
        mbase = tiles[mt];
        pl = (int64_t)tdims[2 * mt] * tdims[2 * mt + 1];
    }
    for (int q = threadIdx.x; q < kTS; q += kTBlock) {
        const int32_t px = tp[q];
#pragma unroll
        for (int cc = 0; cc < NC; cc++)
            mv[cc * kTS + q] = px >= 0 ? mbase[cc * pl + px] : 0.;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t e = it.b + it.n;
    const int64_t c_end = (e + 63) >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int64_t cb = (it.b >> 6) + (int64_t)wave * kTU; cb < c_end;
         cb += (int64_t)(kTBlock / 64) * kTU) {
        uint64_t m[kTU];
        uint32_t hb[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            const int64_t c = min(cb + u, c_end - 1);
            m[u] = tmask[c]; hb[u] = thead[c];
        }
        uint32_t si[kTU];
        int sl[kTU];
        float sp[kTU][NS > 0 ? NS : 1];
        bool act[kTU];
        float o[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            // every lane loads (inactive lanes at the item's first position)
            const int64_t j0 = (cb + u) * 64 + lane;
            act[u] = j0 >= it.b && j0 < e;
            const int64_t j = act[u] ? j0 : it.b;
            const uint32_t ts = tile_sample(m[u], hb[u], head_i, lane);
            si[u] = act[u] ? ts : 0u;
            sl[u] = slot[j];
            if constexpr (NS == 2) {
                const f32x2 v = reinterpret_cast<const f32x2 *>(fac)[j];
                sp[u][0] = v.x; sp[u][1] = v.y;
            } else {
#pragma unroll
                for (int q = 0; q < NS; q++) sp[u][q] = fac[j * NS + q];
            }
        }
#pragma unroll
        for (int u = 0; u < kTU; u++) o[u] = accumulate ? sig[si[u]] : 0.f;
        float tv[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) tv[u] = TC ? tval[fast_div<1>(si[u], n_t)] : 0.f;
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            if (!act[u]) continue;
            float pf[NC];
            if constexpr (TC) {
                pf[0] = tv[u];
#pragma unroll
                for (int q = 1; q < NC; q++) pf[q] = sp[u][q - 1];
            } else {
#pragma unroll
                for (int q = 0; q < NC; q++) pf[q] = sp[u][q];
            }
            float sr = 0;
#pragma unroll
            for (int q = 0; q < NC; q++) sr = sr + pf[q] * mv[q * kTS + sl[u]];
            sig[si[u]] = accumulate ? o[u] + sr : sr;
        }
    }
}

__global__ __launch_bounds__(kPlanBlock) void k_tp_zero(const uint32_t *off, int64_t n_off,
                                                        float *sig) {
    for (int64_t q = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; q < n_off;
         q += (int64_t)gridDim.x * kPlanBlock)
        sig[off[q]] = 0.f;
}

// ---------------------------------------------------------------------------
// Windowed layout.  The samples stay in time order.  Each work item is a
// stretch of one detector's consecutive samples whose compact pixels span at
// most kWinMax values (the item's window); per sample the plan keeps the u16
// offset into the window (kTOff off the map) and the stored factors.  For a
// scan whose detectors sweep one strip of the map (a constant-elevation scan
// in horizon coordinates puts each detector on one map row) a window covers
// the pixels of the sweep, and:
//   from_map  loads the item's window of the map into LDS and gathers from it;
//   to_map    sums the item's samples into its window in LDS and writes the
//             window's sums to a per-item partial buffer (plain stores), and a
//             second kernel adds, per visited pixel, the partial sums of the
//             items whose windows hold it (a fixed order: items in sample
//             order) -- no global atomics, and the second kernel can write
//             every map pixel, so an overwriting to_map needs no zeroing pass.
// Items are formed on the host from per-512-sample window bounds: sub-chunks
// of a detector are merged while the window fits and the item stays under
// its length limit.  The layout is used only when every sub-chunk fits.
// ---------------------------------------------------------------------------
constexpr int kWSub = 512;       // samples per sub-chunk of the window bounds
constexpr int kWinMax = 2048;    // largest window (LDS: 48 KB of TQU sums)
constexpr int kWFrom = 8192;     // samples per from_map piece (8 per thread of 1024)

struct WItem {
    int64_t b;       // first sample (d * n_t + t)
    int32_t n;       // samples
    int32_t det;     // detector
    int32_t lo;      // first compact pixel of the window
    int32_t win;     // window size (0: no sample on the map)
    int64_t pbase;   // first partial slot of the item
};

// Smallest and largest compact pixel of each sub-chunk of kWSub samples of a
// detector (hi = -1 when no sample is on the map); one wave per sub-chunk.
__global__ __launch_bounds__(kPlanBlock) void k_tw_bounds(const int32_t *pix, PixGeom pg,
                                                          int64_t n_t, int64_t n_sub_det,
                                                          int64_t n_sub, int64_t ny, int64_t nx,
                                                          const uint32_t *bitmap,
                                                          const uint32_t *wscan, int32_t *lo,
                                                          int32_t *hi) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * (kPlanBlock / 64) + (threadIdx.x >> 6);
    if (q >= n_sub) return;   // uniform per wave
    const int64_t d = q / n_sub_det, t0 = (q % n_sub_det) * kWSub;
    const int64_t t1 = min(n_t, t0 + (int64_t)kWSub);
    int32_t mn = INT32_MAX, mx = -1;
    for (int64_t t = t0 + lane; t < t1; t += 64) {
        const int64_t p = flat_pixel(pix, pg, d * n_t + t, ny, nx);
        if (p >= 0) {
            const int32_t c = (int32_t)compact_of(p, bitmap, wscan);
            mn = min(mn, c);
            mx = max(mx, c);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, __shfl_xor(mn, o));
        mx = max(mx, __shfl_xor(mx, o));
    }
    if (lane == 0) { lo[q] = mn; hi[q] = mx; }
}

// Per sample: its window offset (kTOff off the map) and its stored factors
// (spin components c0 .. nc-1), in time order.  sub_item: the item of each
// sub-chunk; item_lo: each item's first compact pixel.
__global__ __launch_bounds__(kPlanBlock) void k_tw_fill(const int32_t *pix, PixGeom pg,
                                                        const float *spin, int nc, int c0,
                                                        int64_t n, int64_t n_t,
                                                        int64_t n_sub_det, int64_t ny,
                                                        int64_t nx, const uint32_t *bitmap,
                                                        const uint32_t *wscan,
                                                        const int32_t *sub_item,
                                                        const int32_t *item_lo, uint16_t *wslot,
                                                        float *tspin, unsigned long long *n_on) {
    const int ns = nc - c0;
    uint32_t on = 0;
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t d = i / n_t, t = i - d * n_t;
        const int64_t p = flat_pixel(pix, pg, i, ny, nx);
        uint16_t sl = kTOff;
        if (p >= 0) {
            sl = (uint16_t)((int32_t)compact_of(p, bitmap, wscan) -
                            item_lo[sub_item[d * n_sub_det + t / kWSub]]);
            on++;
        }
        wslot[i] = sl;
        for (int c = 0; c < ns; c++) tspin[i * ns + c] = spin[i * nc + c0 + c];
    }
    // on-map samples: summed over the block, one atomic per block
    __shared__ uint32_t s_on;
    if (threadIdx.x == 0) s_on = 0;
    __syncthreads();
    if (on) atomicAdd(&s_on, on);
    __syncthreads();
    if (threadIdx.x == 0 && s_on) atomicAdd(n_on, (unsigned long long)s_on);
}

// The R = 8 consecutive samples of one lane starting at sample i0 (of an
// item ending at sample e): window slots, stored factors and, for to_map and
// accumulating from_map, the signal.  VEC: i0 % 8 == 0 (n_t % 8 == 0), so
// whole groups load as vectors.
template <int NS, bool VEC, typename ST>
__device__ __forceinline__ void tw_load(int64_t i0, int64_t e, const uint16_t *__restrict__ wslot,
                                        const float *__restrict__ tspin,
                                        const ST *__restrict__ sig, bool with_sig,
                                        uint32_t (&sl)[8], float (&sp)[8][NS > 0 ? NS : 1],
                                        ST (&sg)[8]) {
    if (VEC && i0 + 8 <= e) {
        const uint4 a = *reinterpret_cast<const uint4 *>(wslot + i0);
        const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int q = 0; q < 4; q++) { sl[2 * q] = w[q] & 0xFFFFu; sl[2 * q + 1] = w[q] >> 16; }
        if constexpr (NS > 0) {
#pragma unroll
            for (int q = 0; q < 2 * NS; q++) {
                const float4 f = reinterpret_cast<const float4 *>(tspin + i0 * NS)[q];
                const float v[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
                for (int k = 0; k < 4; k++) sp[(4 * q + k) / NS][(4 * q + k) % NS] = v[k];
            }
        }
        if (with_sig) {
            if constexpr (std::is_same<ST, float>::value) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const float4 s = reinterpret_cast<const float4 *>(sig + i0)[q];
                    sg[4 * q] = s.x; sg[4 * q + 1] = s.y; sg[4 * q + 2] = s.z; sg[4 * q + 3] = s.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double2 s = reinterpret_cast<const double2 *>(sig + i0)[q];
                    sg[2 * q] = s.x; sg[2 * q + 1] = s.y;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++) sg[r] = (ST)0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const int64_t i = i0 + r;
            const bool ok = i < e;
            sl[r] = ok ? (uint32_t)wslot[i] : (uint32_t)kTOff;
#pragma unroll
            for (int c = 0; c < NS; c++) sp[r][c] = ok ? tspin[i * NS + c] : 0.f;
            sg[r] = (ok && with_sig) ? sig[i] : (ST)0;
        }
    }
}

// Windowed to_map, first pass: one block of BS threads per item.  Each lane
// sums its runs of equal slots among its 8 consecutive samples and adds each
// run to the window in LDS (float64 LDS atomics).  The window's sums then go
// to the partial buffer in entry order (wse: slot -> entry; ATOM = false) or
// are added to the map with one float64 atomic per touched slot and
// component (ATOM = true).  Arithmetic per sample as ProjEng_Precomp's
// to_map (Projection.cxx:2100-2121): (w * sig) * pf in float32 (float64 for
// a float64 signal), summed in float64; a zero detector weight skips the
// detector.
// Doubles between the two copies of a window's sums (see kTPad).
constexpr int kWPad = 8;

template <int NC, bool TC, typename ST, bool VEC, int BS, bool ATOM, int CP>
__global__ __launch_bounds__(1024) void k_tw_to_map(const WItem *__restrict__ items,
                                                    const uint16_t *__restrict__ wslot,
                                                    const float *__restrict__ tspin,
                                                    const float *__restrict__ tval,
                                                    const ST *__restrict__ sig,
                                                    const float *__restrict__ det_weights,
                                                    const uint32_t *__restrict__ wse,
                                                    double *__restrict__ part,
                                                    const int32_t *__restrict__ comp2pix,
                                                    int64_t plane, double *map) {
    constexpr int NS = TC ? NC - 1 : NC;
    // [CP][NC][win] with kWPad doubles between the copies: lane parity picks
    // the copy (CP = 2)
    extern __shared__ double wsum[];
    const WItem it = items[blockIdx.x];
    const int win = it.win;
    const int cs = NC * win + kWPad;   // doubles per copy and gap
    double *wacc = wsum + (CP > 1 ? (threadIdx.x & 1) * cs : 0);
    const float w = det_weights != nullptr ? det_weights[it.det] : 1.f;
    const float pt = TC ? tval[it.det] : 0.f;
    const bool work = win > 0 && w != 0.f;   // a zero weight skips the detector (Projection.cxx:2104)
    const int64_t e = it.b + it.n;
    int64_t i0 = it.b + (int64_t)threadIdx.x * 8;
    uint32_t sl[8];
    float sp[8][NS > 0 ? NS : 1];
    ST sg[8];
    // the first 8 samples' loads are issued before the window is zeroed
    if (work && i0 < e) tw_load<NS, VEC, ST>(i0, e, wslot, tspin, sig, true, sl, sp, sg);
    for (int q = threadIdx.x; q < CP * cs; q += BS) wsum[q] = 0.;
    __syncthreads();
    if (work) {
        for (bool first = true; i0 < e; i0 += BS * 8, first = false) {
            if (!first) tw_load<NS, VEC, ST>(i0, e, wslot, tspin, sig, true, sl, sp, sg);
            uint32_t cur = kTOff;
            double cv[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) cv[c] = 0.;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                if (sl[r] == kTOff) continue;
                float pf[NC];
                if constexpr (TC) {
                    pf[0] = pt;
#pragma unroll
                    for (int c = 1; c < NC; c++) pf[c] = sp[r][c - 1];
                } else {
#pragma unroll
                    for (int c = 0; c < NC; c++) pf[c] = sp[r][c];
                }
                double x[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    if constexpr (std::is_same<ST, double>::value)
                        x[c] = (double)w * sg[r] * (double)pf[c];
                    else
                        x[c] = (double)(w * sg[r] * pf[c]);
                }
                if (sl[r] != cur) {
                    if (cur != kTOff) {
#pragma unroll
                        for (int c = 0; c < NC; c++) atomicAdd(wacc + c * win + cur, cv[c]);
                    }
                    cur = sl[r];
#pragma unroll
                    for (int c = 0; c < NC; c++) cv[c] = x[c];
                } else {
#pragma unroll
                    for (int c = 0; c < NC; c++) cv[c] += x[c];
                }
            }
            if (cur != kTOff) {
#pragma unroll
                for (int c = 0; c < NC; c++) atomicAdd(wacc + c * win + cur, cv[c]);
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < win; q += BS) {
        double v[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            v[c] = wsum[c * win + q];
            if (CP > 1) v[c] += wsum[cs + c * win + q];
        }
        if constexpr (ATOM) {
            const int64_t p = comp2pix[it.lo + q];
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (v[c] != 0.) unsafeAtomicAdd(map + c * plane + p, v[c]);
        } else {
            double *pp = part + (int64_t)wse[it.pbase + q] * NC;
#pragma unroll
            for (int c = 0; c < NC; c++) pp[c] = v[c];
        }
    }
}

// Windowed to_weights, first pass: as k_tw_to_map with the NC (NC + 1) / 2
// upper-triangle products (pf_a * pf_b) * w in float32 per sample
// (Projection.cxx:2135-2154); partial sums in entry order.
template <int NC, bool TC, bool VEC, int BS>
__global__ __launch_bounds__(1024) void k_tw_to_weights(
    const WItem *__restrict__ items, const uint16_t *__restrict__ wslot,
    const float *__restrict__ tspin, const float *__restrict__ tval,
    const float *__restrict__ det_weights, const uint32_t *__restrict__ wse,
    double *__restrict__ part) {
    constexpr int NS = TC ? NC - 1 : NC;
    constexpr int NW = NC * (NC + 1) / 2;
    extern __shared__ double wacc[];   // [NW][win]
    const WItem it = items[blockIdx.x];
    const int win = it.win;
    for (int q = threadIdx.x; q < NW * win; q += BS) wacc[q] = 0.;
    __syncthreads();
    const float w = det_weights != nullptr ? det_weights[it.det] : 1.f;
    const float pt = TC ? tval[it.det] : 0.f;
    if (win > 0 && w != 0.f) {   // a zero weight skips the detector (Projection.cxx:2139)
        const int64_t e = it.b + it.n;
        for (int64_t i0 = it.b + (int64_t)threadIdx.x * 8; i0 < e; i0 += BS * 8) {
            uint32_t sl[8];
            float sp[8][NS > 0 ? NS : 1];
            float sg[8];
            tw_load<NS, VEC, float>(i0, e, wslot, tspin, nullptr, false, sl, sp, sg);
            uint32_t cur = kTOff;
            double cv[NW];
#pragma unroll
            for (int k = 0; k < NW; k++) cv[k] = 0.;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                if (sl[r] == kTOff) continue;
                float pf[NC];
                if constexpr (TC) {
                    pf[0] = pt;
#pragma unroll
                    for (int c = 1; c < NC; c++) pf[c] = sp[r][c - 1];
                } else {
#pragma unroll
                    for (int c = 0; c < NC; c++) pf[c] = sp[r][c];
                }
                double x[NW];
                int k = 0;
#pragma unroll
                for (int a = 0; a < NC; a++)
#pragma unroll
                    for (int b = a; b < NC; b++, k++) x[k] = (double)(pf[a] * pf[b] * w);
                if (sl[r] != cur) {
                    if (cur != kTOff) {
#pragma unroll
                        for (int q = 0; q < NW; q++) atomicAdd(wacc + q * win + cur, cv[q]);
                    }
                    cur = sl[r];
#pragma unroll
                    for (int q = 0; q < NW; q++) cv[q] = x[q];
                } else {
#pragma unroll
                    for (int q = 0; q < NW; q++) cv[q] += x[q];
                }
            }
            if (cur != kTOff) {
#pragma unroll
                for (int q = 0; q < NW; q++) atomicAdd(wacc + q * win + cur, cv[q]);
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < win; q += BS) {
        double *pp = part + (int64_t)wse[it.pbase + q] * NW;
#pragma unroll
        for (int k = 0; k < NW; k++) pp[k] = wacc[k * win + q];
    }
}

// Sum of entries [e0, e1) of one component q of the partial buffer (NP
// values per entry), four loads in flight.
template <int NP>
__device__ __forceinline__ double tw_sum(const double *__restrict__ part, uint32_t e0,
                                         uint32_t e1, int q) {
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    uint32_t e = e0;
    for (; e + 4 <= e1; e += 4) {
        s0 += part[(int64_t)e * NP + q];
        s1 += part[(int64_t)(e + 1) * NP + q];
        s2 += part[(int64_t)(e + 2) * NP + q];
        s3 += part[(int64_t)(e + 3) * NP + q];
    }
    for (; e < e1; e++) s0 += part[(int64_t)e * NP + q];
    return (s0 + s1) + (s2 + s3);
}

// Windowed second pass, adding into the map: one thread per visited pixel
// and value sums the pixel's entries (contiguous; items in sample order).
template <int NP>
__global__ __launch_bounds__(kPlanBlock) void k_tw_add(const uint32_t *__restrict__ wcsr,
                                                       const double *__restrict__ part,
                                                       const int32_t *__restrict__ comp2pix,
                                                       int64_t n_vis, WPlanes pl, int64_t plane,
                                                       double *map) {
    for (int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; j < n_vis * NP;
         j += (int64_t)gridDim.x * kPlanBlock) {
        const int64_t c = j / NP;
        const int q = (int)(j - c * NP);
        const double s = tw_sum<NP>(part, wcsr[c], wcsr[c + 1], q);
        map[pl.o[q] * plane + comp2pix[c]] += s;
    }
}

// Windowed second pass, overwriting the map: one thread per map pixel and
// component writes its sum (0 for a pixel no sample visits), so no zeroing
// pass is needed.
template <int NP>
__global__ __launch_bounds__(kPlanBlock) void k_tw_set(const uint32_t *__restrict__ wcsr,
                                                       const double *__restrict__ part,
                                                       const uint32_t *__restrict__ bitmap,
                                                       const uint32_t *__restrict__ wscan,
                                                       int64_t npix, double *map) {
    for (int64_t j = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; j < npix * NP;
         j += (int64_t)gridDim.x * kPlanBlock) {
        const int q = (int)(j / npix);
        const int64_t p = j - (int64_t)q * npix;
        double s = 0.;
        if ((bitmap[p >> 5] >> (p & 31)) & 1u) {
            const uint32_t c = compact_of(p, bitmap, wscan);
            s = tw_sum<NP>(part, wcsr[c], wcsr[c + 1], q);
        }
        map[j] = s;
    }
}

// Windowed from_map: one block of BS threads per item; the item's window of
// the map (NC values per slot, through comp2pix) is loaded into LDS, then
// each lane takes 8 consecutive samples.  Arithmetic and off-map handling as
// k_plan_from_map (ProjEng_Precomp's from_map, Projection.cxx:2060-2085).
template <int NC, bool TC, bool VEC, int BS>
__global__ __launch_bounds__(1024) void k_tw_from_map(const WItem *__restrict__ items,
                                                      const uint16_t *__restrict__ wslot,
                                                      const float *__restrict__ tspin,
                                                      const float *__restrict__ tval,
                                                      const int32_t *__restrict__ comp2pix,
                                                      const double *__restrict__ map,
                                                      int64_t plane, float *__restrict__ sig,
                                                      int accumulate) {
    constexpr int NS = TC ? NC - 1 : NC;
    extern __shared__ double wmv[];   // [NC][win]
    const WItem it = items[blockIdx.x];
    const int win = it.win;
    const float pt = TC ? tval[it.det] : 0.f;
    const int64_t e = it.b + it.n;
    int64_t i0 = it.b + (int64_t)threadIdx.x * 8;
    uint32_t sl[8];
    float sp[8][NS > 0 ? NS : 1];
    float o[8];
    // the first 8 samples' loads are issued before the window's
    if (i0 < e) tw_load<NS, VEC, float>(i0, e, wslot, tspin, sig, accumulate != 0, sl, sp, o);
    for (int q = threadIdx.x; q < win; q += BS) {
        const int64_t p = comp2pix[it.lo + q];
#pragma unroll
        for (int c = 0; c < NC; c++) wmv[c * win + q] = map[c * plane + p];
    }
    __syncthreads();
    for (bool first = true; i0 < e; i0 += BS * 8, first = false) {
        if (!first) tw_load<NS, VEC, float>(i0, e, wslot, tspin, sig, accumulate != 0, sl, sp, o);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            if (sl[r] == kTOff) continue;
            float pf[NC];
            if constexpr (TC) {
                pf[0] = pt;
#pragma unroll
                for (int c = 1; c < NC; c++) pf[c] = sp[r][c - 1];
            } else {
#pragma unroll
                for (int c = 0; c < NC; c++) pf[c] = sp[r][c];
            }
            float sr = 0;
#pragma unroll
            for (int c = 0; c < NC; c++) sr = sr + pf[c] * wmv[c * win + sl[r]];
            o[r] = accumulate ? o[r] + sr : sr;
        }
        if (VEC && i0 + 8 <= e) {
#pragma unroll
            for (int q = 0; q < 2; q++)
                reinterpret_cast<float4 *>(sig + i0)[q] =
                    float4{o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]};
        } else {
#pragma unroll
            for (int r = 0; r < 8; r++)
                if (i0 + r < e) sig[i0 + r] = o[r];
        }
    }
}

// Plain fill of n doubles (the zeroing pass of an overwriting to_map on the
// other layouts).
__global__ __launch_bounds__(kPlanBlock) void k_plan_fill0(double *x, int64_t n) {
    if (n <= 0) return;
    // 16-byte stores from the first 16-byte boundary; the ends one by one
    const int64_t h = (int64_t)(((16 - ((uintptr_t)x & 15)) & 15) / 8);
    const int64_t nv = (n - h) / 2;
    const int64_t tid = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    const int64_t nt = (int64_t)gridDim.x * kPlanBlock;
    for (int64_t i = tid; i < nv; i += nt)
        reinterpret_cast<double2 *>(x + h)[i] = double2{0., 0.};
    if (tid < h) x[tid] = 0.;
    if (tid == 0 && h + 2 * nv < n) x[n - 1] = 0.;
}

}  // namespace so3g

namespace {

unsigned grid_of(int64_t n, int64_t per = kPlanBlock, int64_t cap = 1 << 20) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, cap));
}

void free_plan_arrays(so3g_pc_plan *p) {
    (void)hipFree(p->idx); (void)hipFree(p->sspin); (void)hipFree(p->tval);
    (void)hipFree(p->heads); (void)hipFree(p->wbase); (void)hipFree(p->run_pix);
    (void)hipFree(p->tpd); (void)hipFree(p->tpb); (void)hipFree(p->tpw);
    (void)hipFree(p->wgrp); (void)hipFree(p->tspin); (void)hipFree(p->packed);
    (void)hipFree(p->comp2pix);
    (void)hipFree(p->tslot); (void)hipFree(p->tmask);
    (void)hipFree(p->thead); (void)hipFree(p->head_i); (void)hipFree(p->tile_pix);
    (void)hipFree(p->items); (void)hipFree(p->off_idx);
    (void)hipFree(p->tdims_d); (void)hipFree(p->tbase_d); (void)hipFree(p->sbase_d);
    (void)hipFree(p->tile_map); (void)hipFree(p->comp_tile); (void)hipFree(p->tiles_d);
    (void)hipFree(p->wslot); (void)hipFree(p->witems); (void)hipFree(p->wpart);
    (void)hipFree(p->wfitems);
    (void)hipFree(p->wcsr); (void)hipFree(p->wse); (void)hipFree(p->wbitmap);
    (void)hipFree(p->wwscan);
}

// Time-ordered copy for from_map: spin factors, and the compact pixels as
// u16 offsets from a base per kTGroup samples (int32 for the wide groups).
// Arrays land in p (freed with it on failure); temporaries are freed here.
hipError_t plan_build_tcopy(so3g_pc_plan *p, const int32_t *pixel, PixGeom pg, const float *spin,
                            int nc, int c0, int64_t n, int64_t ny, int64_t nx,
                            const uint32_t *bitmap, const uint32_t *wscan, hipStream_t s) {
    const int64_t n_groups = (n + kTGroup - 1) / kTGroup;
    int32_t *tpix = nullptr;
    uint32_t *cnt = nullptr;
    auto done = [&](hipError_t e) {
        (void)hipFree(tpix); (void)hipFree(cnt);
        return e;
    };
    hipError_t e;
#define TC_STEP(expr) if ((e = (expr)) != hipSuccess) return done(e)
    TC_STEP(hipMalloc(&tpix, std::max<int64_t>(1, n) * 4));
    TC_STEP(hipMalloc(&cnt, 8));
    TC_STEP(hipMalloc(&p->tspin, std::max<int64_t>(1, n * p->ns) * 4));
    TC_STEP(hipMalloc(&p->tpd, std::max<int64_t>(1, n) * 2));
    TC_STEP(hipMalloc(&p->tpb, std::max<int64_t>(1, n_groups) * 4));
    TC_STEP(hipMemsetAsync(cnt, 0, 8, s));
    hipLaunchKernelGGL(k_plan_tcopy, dim3(grid_of(n)), dim3(kPlanBlock), 0, s, pixel, pg, spin,
                       nc, c0, n, ny, nx, bitmap, wscan, tpix, p->tspin);
    const dim3 gg(grid_of(n_groups * 64));
    hipLaunchKernelGGL(k_plan_tbase, gg, dim3(kPlanBlock), 0, s, tpix, n, n_groups, p->tpb, cnt);
    TC_STEP(hipGetLastError());
    uint32_t n_wide = 0;
    TC_STEP(hipMemcpyAsync(&n_wide, cnt, 4, hipMemcpyDeviceToHost, s));
    TC_STEP(hipStreamSynchronize(s));
    p->n_wide = n_wide;
    TC_STEP(hipMalloc(&p->tpw, std::max<int64_t>(1, p->n_wide) * kTGroup * 4));
    TC_STEP(hipMalloc(&p->wgrp, std::max<int64_t>(1, p->n_wide) * 4));
    hipLaunchKernelGGL(k_plan_tpack, gg, dim3(kPlanBlock), 0, s, tpix, n, n_groups, p->tpb,
                       p->tpd, p->tpw, p->wgrp, cnt + 1);
    TC_STEP(hipGetLastError());
    TC_STEP(hipMalloc(&p->packed, std::max<int64_t>(1, p->n_vis) * 4 * sizeof(double)));
    TC_STEP(hipStreamSynchronize(s));
#undef TC_STEP
    return done(hipSuccess);
}

// Device bytes of the time-ordered copy.
int64_t tcopy_bytes(const so3g_pc_plan *p, int64_t n) {
    return n * (2 + 4 * p->ns) + (n + kTGroup - 1) / kTGroup * 4 +
           p->n_wide * (kTGroup * 4 + 4) + p->n_vis * 32;
}

// Windowed to_map: the items' window sums into the partial buffer, then
// per pixel into the map (overwrite: every map pixel is written); or, with
// "plan_win_flush" 1, added to the map by atomics (overwrite: zeroed first).
template <typename ST>
int plan_tw_to_map(const so3g_pc_plan *p, double *map, const ST *signal,
                   const float *det_weights, hipStream_t s, bool overwrite) {
    const int64_t npix = p->ny * p->nx;
    const bool vec = p->wvec && ((uintptr_t)signal % 16) == 0;
    const bool atom = g_plan_win_flush != 0;
    if (atom && overwrite) {
        const int64_t n_map = (int64_t)p->nc * npix;
        hipLaunchKernelGGL(k_plan_fill0, dim3(grid_of(n_map / 2 + 1)), dim3(kPlanBlock), 0, s,
                           map, n_map);
    }
    if (p->n_witems > 0) {
        // two copies of the window sums (by lane parity) where they leave room
        // for two blocks per compute unit
        const size_t lds1 = (size_t)(p->nc * std::max(1, p->win_max) + kWPad) * sizeof(double);
        const bool two = 2 * lds1 <= 65536;
        const size_t lds = two ? 2 * lds1 : lds1;
        const dim3 gi((unsigned)p->n_witems);
        const WItem *it = (const WItem *)p->witems;
        const int bs = g_plan_win_block;
#define TW_TO4(NC, TC, V, BS, A, CP)                                                        \
    hipLaunchKernelGGL((k_tw_to_map<NC, TC, ST, V, BS, A, CP>), gi, dim3(BS), lds, s, it,   \
                       p->wslot, p->tspin, p->tval, signal, det_weights, p->wse, p->wpart,  \
                       p->comp2pix, npix, map)
#define TW_TO3(NC, TC, V, BS, A)                                                            \
    do {                                                                                    \
        if (two) TW_TO4(NC, TC, V, BS, A, 2); else TW_TO4(NC, TC, V, BS, A, 1);             \
    } while (0)
#define TW_TO2(NC, TC, V)                                                                   \
    do {                                                                                    \
        if (bs == 1024) {                                                                   \
            if (atom) TW_TO3(NC, TC, V, 1024, true); else TW_TO3(NC, TC, V, 1024, false);   \
        } else {                                                                            \
            if (atom) TW_TO3(NC, TC, V, 256, true); else TW_TO3(NC, TC, V, 256, false);     \
        }                                                                                   \
    } while (0)
#define TW_TO(NC, TC)                                                                       \
    do {                                                                                    \
        if (vec) TW_TO2(NC, TC, true); else TW_TO2(NC, TC, false);                          \
    } while (0)
        if (p->tconst) {
            if (p->nc == 1) TW_TO(1, true);
            else TW_TO(3, true);
        } else {
            if (p->nc == 1) TW_TO(1, false);
            else if (p->nc == 2) TW_TO(2, false);
            else TW_TO(3, false);
        }
#undef TW_TO
#undef TW_TO2
#undef TW_TO3
#undef TW_TO4
        PLAN_TRY(hipGetLastError());
    }
    if (atom) return SO3G_OK;
    if (overwrite) {
        const dim3 g(grid_of(npix * p->nc));
        switch (p->nc) {
        case 1: hipLaunchKernelGGL(k_tw_set<1>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->wbitmap, p->wwscan, npix, map); break;
        case 2: hipLaunchKernelGGL(k_tw_set<2>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->wbitmap, p->wwscan, npix, map); break;
        default: hipLaunchKernelGGL(k_tw_set<3>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->wbitmap, p->wwscan, npix, map); break;
        }
    } else if (p->n_vis > 0) {
        const dim3 g(grid_of(p->n_vis * p->nc));
        const WPlanes pl{{0, 1, 2, 0, 0, 0}};
        switch (p->nc) {
        case 1: hipLaunchKernelGGL(k_tw_add<1>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
        case 2: hipLaunchKernelGGL(k_tw_add<2>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
        default: hipLaunchKernelGGL(k_tw_add<3>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
        }
    }
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

// Windowed to_weights: upper-triangle products per item window, then added
// per pixel into the (nc, nc, ...) map.
int plan_tw_to_weights(const so3g_pc_plan *p, double *map, const float *det_weights,
                       hipStream_t s) {
    if (p->n_witems == 0 || p->n_vis == 0) return SO3G_OK;
    const int nc = p->nc, nw = nc * (nc + 1) / 2;
    const size_t lds = (size_t)nw * std::max(1, p->win_max) * sizeof(double);
    const dim3 gi((unsigned)p->n_witems);
    const WItem *it = (const WItem *)p->witems;
    const int bs = g_plan_win_block;
#define TW_W2(NC, TC, V)                                                                    \
    do {                                                                                    \
        if (bs == 1024)                                                                     \
            hipLaunchKernelGGL((k_tw_to_weights<NC, TC, V, 1024>), gi, dim3(1024), lds, s,  \
                               it, p->wslot, p->tspin, p->tval, det_weights, p->wse,        \
                               p->wpart);                                                   \
        else                                                                                \
            hipLaunchKernelGGL((k_tw_to_weights<NC, TC, V, 256>), gi, dim3(256), lds, s, it, \
                               p->wslot, p->tspin, p->tval, det_weights, p->wse, p->wpart); \
    } while (0)
#define TW_W(NC, TC)                                                                        \
    do {                                                                                    \
        if (p->wvec) TW_W2(NC, TC, true); else TW_W2(NC, TC, false);                        \
    } while (0)
    if (p->tconst) {
        if (nc == 1) TW_W(1, true);
        else TW_W(3, true);
    } else {
        if (nc == 1) TW_W(1, false);
        else if (nc == 2) TW_W(2, false);
        else TW_W(3, false);
    }
#undef TW_W
#undef TW_W2
    PLAN_TRY(hipGetLastError());
    WPlanes pl{{0, 0, 0, 0, 0, 0}};
    int k = 0;
    for (int a = 0; a < nc; a++)
        for (int b = a; b < nc; b++) pl.o[k++] = a * nc + b;
    const dim3 g(grid_of(p->n_vis * nw));
    const int64_t npix = p->ny * p->nx;
    switch (nw) {
    case 1: hipLaunchKernelGGL(k_tw_add<1>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
    case 3: hipLaunchKernelGGL(k_tw_add<3>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
    default: hipLaunchKernelGGL(k_tw_add<6>, g, dim3(kPlanBlock), 0, s, p->wcsr, p->wpart, p->comp2pix, p->n_vis, pl, npix, map); break;
    }
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

template <typename ST>
int plan_to_map(const so3g_pc_plan *p, double *map, const ST *signal,
                const float *det_weights, void *stream, double *const *tiles = nullptr) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    if (p->layout == 2)
        return plan_tw_to_map<ST>(p, map, signal, det_weights, (hipStream_t)stream, false);
    if (p->n_on == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t plane = (uint32_t)(p->ny * p->nx);
    const FastDiv nt{(uint32_t)p->n_t, 1.f / (float)p->n_t};
    const int tabs = (p->tconst ? 1 : 0) + (det_weights != nullptr ? 1 : 0);
    const bool ldt = g_plan_to_lds && tabs > 0 &&
                     p->n_det * tabs * (int64_t)sizeof(float) <= kPlanLdsBytes;
    const size_t lds = ldt ? (size_t)p->n_det * tabs * sizeof(float) : 0;
    const int32_t cpb = g_plan_to_cpb;
    const dim3 g(grid_of(p->n_on, (int64_t)kPlanBlock * kPlanR * cpb, INT32_MAX));
    const int32_t nd = (int32_t)p->n_det;
    if (p->layout == 1) {
        const dim3 gi((unsigned)p->n_items);
        const TItem *items = (const TItem *)p->items;
#define TP_TO(NC, TC)                                                                       \
    do {                                                                                    \
        if (ldt)                                                                            \
            hipLaunchKernelGGL((k_tp_to_map<NC, TC, ST, true>), gi, dim3(kTBlock), lds, s,  \
                               items, p->tslot, p->sspin, p->tmask, p->thead,   \
                               p->head_i, p->tile_pix, p->tval, nt, plane, signal,          \
                               det_weights, nd, map, tiles, p->tile_map, p->tdims_d);       \
        else                                                                                \
            hipLaunchKernelGGL((k_tp_to_map<NC, TC, ST, false>), gi, dim3(kTBlock), 0, s,   \
                               items, p->tslot, p->sspin, p->tmask, p->thead,   \
                               p->head_i, p->tile_pix, p->tval, nt, plane, signal,          \
                               det_weights, nd, map, tiles, p->tile_map, p->tdims_d);       \
    } while (0)
        if (p->tconst) {
            if (p->nc == 1) TP_TO(1, true);
            else TP_TO(3, true);
        } else {
            if (p->nc == 1) TP_TO(1, false);
            else if (p->nc == 2) TP_TO(2, false);
            else TP_TO(3, false);
        }
#undef TP_TO
        PLAN_TRY(hipGetLastError());
        return SO3G_OK;
    }
#define PLAN_TO(NC, TC)                                                                     \
    do {                                                                                    \
        if (ldt)                                                                            \
            hipLaunchKernelGGL((k_plan_to_map_c<NC, TC, ST, 1, true>), g, dim3(kPlanBlock),  \
                               lds, s, p->idx, p->sspin, p->tval, p->heads, p->wbase,       \
                               p->run_pix, p->n_on, nt, plane, signal, det_weights, nd, cpb, \
                               map);                                                        \
        else if (g_plan_div)                                                                \
            hipLaunchKernelGGL((k_plan_to_map_c<NC, TC, ST, 1, false>), g, dim3(kPlanBlock), \
                               0, s, p->idx, p->sspin, p->tval, p->heads, p->wbase,         \
                               p->run_pix, p->n_on, nt, plane, signal, det_weights, nd, cpb, \
                               map);                                                        \
        else                                                                                \
            hipLaunchKernelGGL((k_plan_to_map_c<NC, TC, ST, 0, false>), g, dim3(kPlanBlock), \
                               0, s, p->idx, p->sspin, p->tval, p->heads, p->wbase,         \
                               p->run_pix, p->n_on, nt, plane, signal, det_weights, nd, cpb, \
                               map);                                                        \
    } while (0)
    if (p->tconst) {
        if (p->nc == 1) PLAN_TO(1, true);
        else PLAN_TO(3, true);
    } else {
        if (p->nc == 1) PLAN_TO(1, false);
        else if (p->nc == 2) PLAN_TO(2, false);
        else PLAN_TO(3, false);
    }
#undef PLAN_TO
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

// Float64 gather for the map-maker: sig[i] (+)= sum_c map[c, pix_i] * spin[i, c]
// in float64; samples off the map (pixel row < 0 or outside the map) add 0.
template <int NC>
__global__ __launch_bounds__(kPlanBlock) void k_pc_from_map_f64(int64_t n, int32_t ny,
                                                                int32_t nx,
                                                                const double *map,
                                                                const int32_t *pix,
                                                                const float *spin,
                                                                double *sig,
                                                                int accumulate) {
    const int64_t plane = (int64_t)ny * nx;
    for (int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kPlanBlock) {
        const int2 q = reinterpret_cast<const int2 *>(pix)[i];
        double v = accumulate ? sig[i] : 0.;
        if (q.x >= 0 && q.x < ny && q.y >= 0 && q.y < nx) {
            const int64_t px = (int64_t)q.x * nx + q.y;
#pragma unroll
            for (int c = 0; c < NC; c++) v += map[c * plane + px] * (double)spin[i * NC + c];
        }
        sig[i] = v;
    }
}

int plan_to_weights(const so3g_pc_plan *p, double *map, const float *det_weights,
                    void *stream, double *const *tiles) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    if (p->layout == 2) return plan_tw_to_weights(p, map, det_weights, (hipStream_t)stream);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t plane = (uint32_t)(p->ny * p->nx);
    const FastDiv nt{(uint32_t)p->n_t, 1.f / (float)p->n_t};
    if (p->layout == 0) {
        if (p->n_on == 0) return SO3G_OK;
        const int32_t cpb = g_plan_to_cpb;
        const dim3 g(grid_of(p->n_on, (int64_t)kPlanBlock * kPlanR * cpb, INT32_MAX));
#define PW(NC, TC)                                                                          \
    hipLaunchKernelGGL((k_plan_to_weights_c<NC, TC, 1>), g, dim3(kPlanBlock), 0, s, p->idx, \
                       p->sspin, p->tval, p->heads, p->wbase, p->run_pix, p->n_on, nt,       \
                       plane, det_weights, cpb, map)
        if (p->tconst) {
            if (p->nc == 1) PW(1, true);
            else PW(3, true);
        } else {
            if (p->nc == 1) PW(1, false);
            else if (p->nc == 2) PW(2, false);
            else PW(3, false);
        }
#undef PW
        PLAN_TRY(hipGetLastError());
        return SO3G_OK;
    }
    if (p->n_on == 0 || p->n_items == 0) return SO3G_OK;
    const int tabs = (p->tconst ? 1 : 0) + (det_weights != nullptr ? 1 : 0);
    const bool ldt = g_plan_to_lds && tabs > 0 &&
                     p->n_det * tabs * (int64_t)sizeof(float) <= kPlanLdsBytes;
    const size_t lds = ldt ? (size_t)p->n_det * tabs * sizeof(float) : 0;
    const int32_t nd = (int32_t)p->n_det;
    const dim3 gi((unsigned)p->n_items);
    const TItem *items = (const TItem *)p->items;
#define TP_W(NC, TC)                                                                        \
    do {                                                                                    \
        if (ldt)                                                                            \
            hipLaunchKernelGGL((k_tp_to_weights<NC, TC, true>), gi, dim3(kTBlock), lds, s,  \
                               items, p->tslot, p->sspin, p->tmask, p->thead, p->head_i,    \
                               p->tile_pix, p->tval, nt, plane, det_weights, nd, map,       \
                               tiles, p->tile_map, p->tdims_d);                             \
        else                                                                                \
            hipLaunchKernelGGL((k_tp_to_weights<NC, TC, false>), gi, dim3(kTBlock), 0, s,   \
                               items, p->tslot, p->sspin, p->tmask, p->thead, p->head_i,    \
                               p->tile_pix, p->tval, nt, plane, det_weights, nd, map,       \
                               tiles, p->tile_map, p->tdims_d);                             \
    } while (0)
    if (p->tconst) {
        if (p->nc == 1) TP_W(1, true);
        else TP_W(3, true);
    } else {
        if (p->nc == 1) TP_W(1, false);
        else if (p->nc == 2) TP_W(2, false);
        else TP_W(3, false);
    }
#undef TP_W
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

}  // namespace

namespace {

// Shared plan construction.  p holds the map description (and, for tiled
// maps, its device tables); it is freed here on failure.  ny, nx: the map's
// flat pixel grid ((npix, 1) for k = 1 and 3).
int plan_create_impl(const int32_t *pixel, PixGeom pg, const float *spin, int32_t nc,
                     int64_t n_det, int64_t n_t, int64_t ny, int64_t nx, int64_t t_block_mb,
                     so3g_pc_plan *p, so3g_pc_plan **out, hipStream_t s) {
    const int k = pg.k;
    const int64_t n = n_det * n_t;
    const int64_t npix = ny * nx;
    p->n_det = n_det; p->n_t = n_t; p->ny = ny; p->nx = nx; p->nc = nc;
    int32_t *keys_in = nullptr, *keys_out = nullptr;
    uint32_t *vals_in = nullptr, *gcount = nullptr, *bitmap = nullptr, *wcnt = nullptr,
             *wscan = nullptr;
    uint8_t *used = nullptr;
    int32_t *tile_of = nullptr, *tile_flat = nullptr;
    int64_t *seg = nullptr;
    unsigned long long *d_on = nullptr;
    int *bad = nullptr, *rbad = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    auto cleanup = [&]() {
        (void)hipFree(keys_in); (void)hipFree(keys_out); (void)hipFree(vals_in);
        (void)hipFree(gcount); (void)hipFree(d_on); (void)hipFree(bad); (void)hipFree(tmp);
        (void)hipFree(rbad); rbad = nullptr;
        (void)hipFree(bitmap); (void)hipFree(wcnt); (void)hipFree(wscan);
        (void)hipFree(used); (void)hipFree(tile_of); (void)hipFree(tile_flat); (void)hipFree(seg);
        keys_in = keys_out = nullptr; vals_in = gcount = bitmap = wcnt = wscan = nullptr;
        d_on = nullptr; bad = nullptr; used = nullptr; tile_of = tile_flat = nullptr;
        seg = nullptr; tmp = nullptr;
    };
    auto oom = [&]() {
        cleanup();
        free_plan_arrays(p);
        delete p;
        (void)hipGetLastError();
        return set_error(SO3G_ERR_HIP, "out of device memory building the plan");
    };
#define PLAN_ALLOC(ptr, bytes) \
    if (hipMalloc(&(ptr), std::max<int64_t>((int64_t)(bytes), 1)) != hipSuccess) return oom()
#define PLAN_STEP(expr)                                                              \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            cleanup();                                                               \
            free_plan_arrays(p);                                                     \
            delete p;                                                                \
            return set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                               hipGetErrorString(_e));               \
        }                                                                            \
    } while (0)

    // 0. visited pixels: bitmap, per-word counts, their prefix, comp2pix
    const int64_t n_words = (npix + 31) / 32;
    PLAN_ALLOC(bitmap, n_words * 4);
    PLAN_ALLOC(wcnt, n_words * 4);
    PLAN_ALLOC(wscan, n_words * 4);
    PLAN_ALLOC(rbad, sizeof(int));
    PLAN_STEP(hipMemsetAsync(bitmap, 0, n_words * 4, s));
    PLAN_STEP(hipMemsetAsync(rbad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_plan_mark, dim3(grid_of(n)), dim3(kPlanBlock), 0, s, pixel, pg, n, ny,
                       nx, bitmap, rbad);
    hipLaunchKernelGGL(k_plan_wcount, dim3(grid_of(n_words)), dim3(kPlanBlock), 0, s, bitmap,
                       n_words, wcnt);
    PLAN_STEP(hipGetLastError());
    PLAN_STEP(rocprim::exclusive_scan(nullptr, tmp_bytes, wcnt, wscan, 0u, (size_t)n_words,
                                      rocprim::plus<uint32_t>(), s));
    PLAN_ALLOC(tmp, tmp_bytes);
    PLAN_STEP(rocprim::exclusive_scan(tmp, tmp_bytes, wcnt, wscan, 0u, (size_t)n_words,
                                      rocprim::plus<uint32_t>(), s));
    {
        uint32_t lb = 0, lc = 0;
        int hbad = 0;
        PLAN_STEP(hipMemcpyAsync(&lb, wscan + n_words - 1, 4, hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipMemcpyAsync(&lc, wcnt + n_words - 1, 4, hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipMemcpyAsync(&hbad, rbad, sizeof(int), hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipStreamSynchronize(s));
        p->n_vis = (int64_t)lb + lc;
        if (hbad) {
            cleanup();
            free_plan_arrays(p);
            delete p;
            return set_error(SO3G_ERR_RUNTIME, "pixel_index holds an index outside the map");
        }
    }
    (void)hipFree(tmp); tmp = nullptr;
    PLAN_ALLOC(p->comp2pix, p->n_vis * 4);
    hipLaunchKernelGGL(k_plan_compact, dim3(grid_of(n_words)), dim3(kPlanBlock), 0, s, bitmap,
                       wscan, n_words, p->comp2pix);
    PLAN_STEP(hipGetLastError());
    if (k == 3) {
        PLAN_ALLOC(p->comp_tile, p->n_vis * 4);
        hipLaunchKernelGGL(k_plan_comp_tiles, dim3(grid_of(p->n_vis)), dim3(kPlanBlock), 0, s,
                           p->comp2pix, p->n_vis, pg, p->comp_tile);
        PLAN_STEP(hipGetLastError());
    }
    const int32_t n_vis = (int32_t)std::max<int64_t>(1, p->n_vis);

    // Is the T factor constant per detector (bit for bit)?  Then it is
    // stored once per detector (every so3g pointing matrix: T = the
    // detector's T response).
    p->tconst = 0;
    if (nc != 2) {
        PLAN_ALLOC(bad, n_det * sizeof(int));
        PLAN_STEP(hipMemsetAsync(bad, 0, n_det * sizeof(int), s));
        hipLaunchKernelGGL(k_plan_tcheck, dim3(grid_of(n)), dim3(kPlanBlock), 0, s, spin, nc, n_t,
                           n, bad);
        PLAN_STEP(hipGetLastError());
        std::vector<int> hbad(n_det);
        PLAN_STEP(hipMemcpyAsync(hbad.data(), bad, n_det * sizeof(int), hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipStreamSynchronize(s));
        p->tconst = std::all_of(hbad.begin(), hbad.end(), [](int b) { return b == 0; });
    }
    const int c0 = p->tconst ? 1 : 0;
    p->ns = nc - c0;
    if (p->tconst) {
        PLAN_ALLOC(p->tval, n_det * sizeof(float));
        hipLaunchKernelGGL(k_plan_tval, dim3(grid_of(n_det)), dim3(kPlanBlock), 0, s, spin, nc, n_t,
                           n_det, p->tval);
        PLAN_STEP(hipGetLastError());
    }

    // ---- windowed layout: time order, per-item windows ----
    if (k != 3 && (g_plan_layout == 2 || (g_plan_layout < 0 && n < kPlanTiledMin))) {
        const int64_t n_sub_det = (n_t + kWSub - 1) / kWSub;
        const int64_t n_sub = n_det * n_sub_det;
        PLAN_ALLOC(keys_in, n_sub * 4);
        PLAN_ALLOC(keys_out, n_sub * 4);
        hipLaunchKernelGGL(k_tw_bounds, dim3(grid_of(n_sub * 64, kPlanBlock, INT32_MAX)),
                           dim3(kPlanBlock), 0, s, pixel, pg, n_t, n_sub_det, n_sub, ny, nx,
                           bitmap, wscan, keys_in, keys_out);
        PLAN_STEP(hipGetLastError());
        std::vector<int32_t> hlo(n_sub), hhi(n_sub);
        PLAN_STEP(hipMemcpyAsync(hlo.data(), keys_in, n_sub * 4, hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipMemcpyAsync(hhi.data(), keys_out, n_sub * 4, hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipStreamSynchronize(s));
        (void)hipFree(keys_in); keys_in = nullptr;
        (void)hipFree(keys_out); keys_out = nullptr;
        bool fits = true;
        for (int64_t q = 0; q < n_sub && fits; q++)
            fits = hhi[q] < 0 || (int64_t)hhi[q] - hlo[q] + 1 <= kWinMax;
        // Items: a detector's sub-chunks merged while the window fits and the
        // item stays within plan_win_len samples.
        std::vector<WItem> hit;
        std::vector<int32_t> sub_item(fits ? n_sub : 0), item_lo;
        int64_t n_slots = 0;
        // a detector's samples in k = ceil(n_t / plan_win_len) items of about
        // equal length (whole sub-chunks), at least 64 items in all: the
        // items' window flushes (one float64 atomic per touched slot and
        // component) cost more than the parallelism of shorter items gains
        // (4e6 samples: from_map + overwriting to_map 39.5 us cold with 100
        // items, 41.2 with 200, 44.2 with 300)
        const int64_t wl = std::max<int64_t>(kWSub, g_plan_win_len);
        const int64_t k_items = std::min<int64_t>(
            n_sub_det, std::max<int64_t>((n_t + wl - 1) / wl, (64 + n_det - 1) / n_det));
        const int64_t len_max = ((n_sub_det + k_items - 1) / k_items) * kWSub;
        for (int64_t d = 0; d < n_det && fits; d++) {
            int64_t lo = INT32_MAX, hi = -1, start = 0, len = 0;
            auto emit = [&]() {
                const int32_t win = hi >= 0 ? (int32_t)(hi - lo + 1) : 0;
                hit.push_back(WItem{d * n_t + start, (int32_t)len, (int32_t)d,
                                    hi >= 0 ? (int32_t)lo : 0, win, n_slots});
                item_lo.push_back(hi >= 0 ? (int32_t)lo : 0);
                n_slots += win;
            };
            for (int64_t j = 0; j < n_sub_det; j++) {
                const int64_t q = d * n_sub_det + j;
                const int64_t lj = std::min<int64_t>(n_t, (j + 1) * kWSub) - j * kWSub;
                int64_t nlo = lo, nhi = hi;
                if (hhi[q] >= 0) { nlo = std::min<int64_t>(lo, hlo[q]); nhi = std::max<int64_t>(hi, hhi[q]); }
                if (len > 0 && ((nhi >= 0 && nhi - nlo + 1 > kWinMax) || len + lj > len_max)) {
                    emit();
                    start = j * kWSub;
                    len = 0;
                    lo = INT32_MAX; hi = -1;
                    nlo = hhi[q] >= 0 ? hlo[q] : lo;
                    nhi = hhi[q];
                }
                lo = nlo; hi = nhi; len += lj;
                sub_item[q] = (int32_t)hit.size();
            }
            emit();
        }
        fits = fits && n_slots < ((int64_t)1 << 31) && (int64_t)hit.size() < INT32_MAX;
        if (!fits && g_plan_layout == 2) {
            cleanup();
            free_plan_arrays(p);
            delete p;
            return set_error(SO3G_ERR_RUNTIME,
                             "windowed plan (plan_layout 2): some detector's 512 consecutive "
                             "samples visit pixels spanning more than 2048 visited pixels");
        }
        if (fits) {
            p->layout = 2;
            p->n_witems = (int64_t)hit.size();
            p->n_wslots = n_slots;
            p->win_max = 0;
            for (const WItem &w : hit) p->win_max = std::max(p->win_max, w.win);
            p->wvec = n_t % 8 == 0;
            // entries of each visited pixel: the items whose windows hold it,
            // in item order; ent: the entry of each item slot
            std::vector<uint32_t> csr(p->n_vis + 1, 0), ent(std::max<int64_t>(1, n_slots));
            {
                std::vector<int64_t> diff(p->n_vis + 1, 0);
                for (const WItem &w : hit)
                    if (w.win > 0) { diff[w.lo] += 1; diff[w.lo + w.win] -= 1; }
                int64_t run = 0, acc = 0;
                for (int64_t c = 0; c < p->n_vis; c++) {
                    run += diff[c];
                    csr[c] = (uint32_t)acc;
                    acc += run;
                }
                csr[p->n_vis] = (uint32_t)acc;
                std::vector<uint32_t> pos(csr.begin(), csr.end() - 1);
                for (const WItem &w : hit)
                    for (int32_t q = 0; q < w.win; q++) ent[w.pbase + q] = pos[w.lo + q]++;
            }
            // from_map takes the items in pieces of at most kWFrom samples (one
            // pass of a 1024-thread block), each with its item's window
            std::vector<WItem> hfit;
            for (const WItem &w : hit)
                for (int32_t o = 0; o < w.n || (o == 0 && w.n == 0); o += kWFrom)
                    hfit.push_back(WItem{w.b + o, std::min<int32_t>(kWFrom, w.n - o), w.det,
                                         w.lo, w.win, w.pbase});
            p->n_wfitems = (int64_t)hfit.size();
            PLAN_ALLOC(p->wfitems, std::max<int64_t>(1, p->n_wfitems) * (int64_t)sizeof(WItem));
            PLAN_STEP(hipMemcpyAsync(p->wfitems, hfit.data(), p->n_wfitems * sizeof(WItem),
                                     hipMemcpyHostToDevice, s));
            const int np = std::max(nc, nc * (nc + 1) / 2);
            PLAN_ALLOC(p->witems, std::max<int64_t>(1, p->n_witems) * (int64_t)sizeof(WItem));
            PLAN_ALLOC(p->wcsr, (p->n_vis + 1) * 4);
            PLAN_ALLOC(p->wse, std::max<int64_t>(1, n_slots) * 4);
            PLAN_ALLOC(p->wpart, std::max<int64_t>(1, n_slots) * np * (int64_t)sizeof(double));
            PLAN_ALLOC(p->wslot, n * 2);
            PLAN_ALLOC(p->tspin, std::max<int64_t>(1, n * p->ns) * 4);
            PLAN_ALLOC(keys_in, n_sub * 4);
            PLAN_ALLOC(keys_out, std::max<int64_t>(1, p->n_witems) * 4);
            PLAN_STEP(hipMemcpyAsync(p->witems, hit.data(), p->n_witems * sizeof(WItem),
                                     hipMemcpyHostToDevice, s));
            PLAN_STEP(hipMemcpyAsync(p->wcsr, csr.data(), (p->n_vis + 1) * 4,
                                     hipMemcpyHostToDevice, s));
            PLAN_STEP(hipMemcpyAsync(p->wse, ent.data(), std::max<int64_t>(1, n_slots) * 4,
                                     hipMemcpyHostToDevice, s));
            PLAN_STEP(hipMemcpyAsync(keys_in, sub_item.data(), n_sub * 4, hipMemcpyHostToDevice,
                                     s));
            PLAN_STEP(hipMemcpyAsync(keys_out, item_lo.data(), p->n_witems * 4,
                                     hipMemcpyHostToDevice, s));
            PLAN_ALLOC(d_on, sizeof(unsigned long long));
            PLAN_STEP(hipMemsetAsync(d_on, 0, sizeof(unsigned long long), s));
            hipLaunchKernelGGL(k_tw_fill, dim3(grid_of(n)), dim3(kPlanBlock), 0, s, pixel, pg,
                               spin, nc, c0, n, n_t, n_sub_det, ny, nx, bitmap, wscan, keys_in,
                               keys_out, p->wslot, p->tspin, d_on);
            PLAN_STEP(hipGetLastError());
            unsigned long long n_on = 0;
            PLAN_STEP(hipMemcpyAsync(&n_on, d_on, sizeof(n_on), hipMemcpyDeviceToHost, s));
            PLAN_STEP(hipStreamSynchronize(s));
            p->n_on = (int64_t)n_on;
            // the visited bitmap stays with the plan (overwriting to_map)
            p->wbitmap = bitmap; bitmap = nullptr;
            p->wwscan = wscan; wscan = nullptr;
            cleanup();
            p->bytes = n * (2 + 4 * p->ns) + (p->n_witems + p->n_wfitems) * (int64_t)sizeof(WItem) +
                       n_slots * (4 + 8 * np) + (p->n_vis + 1) * 4 + p->n_vis * 4 +
                       n_words * 8 + (p->tconst ? n_det * 4 : 0);
            *out = p;
            return SO3G_OK;
        }
    }

    if (k == 3 || g_plan_layout == 1 || (g_plan_layout < 0 && n >= kPlanTiledMin)) {
        // ---- tiled layout ----
        p->layout = 1;
        int64_t gtx = 1, n_tflat;
        if (k == 2) {
            gtx = (nx + (1 << kTileBX) - 1) >> kTileBX;
            n_tflat = ((ny + (1 << kTileBY) - 1) >> kTileBY) * gtx;
        } else if (k == 3) {
            n_tflat = p->n_sub;
        } else {
            n_tflat = (ny + kTS - 1) / kTS;
        }
        std::vector<int32_t> h_tile_flat;
        {
            PLAN_ALLOC(used, n_tflat);
            PLAN_STEP(hipMemsetAsync(used, 0, n_tflat, s));
            hipLaunchKernelGGL(k_tp_tmark, dim3(grid_of(n)), dim3(kPlanBlock), 0, s, pixel, pg, n,
                               ny, nx, gtx, used);
            PLAN_STEP(hipGetLastError());
            std::vector<uint8_t> h_used(n_tflat);
            PLAN_STEP(hipMemcpyAsync(h_used.data(), used, n_tflat, hipMemcpyDeviceToHost, s));
            PLAN_STEP(hipStreamSynchronize(s));
            std::vector<int32_t> h_tile_of(n_tflat, -1);
            for (int64_t t = 0; t < n_tflat; t++)
                if (h_used[t]) {
                    h_tile_of[t] = (int32_t)h_tile_flat.size();
                    h_tile_flat.push_back((int32_t)t);
                }
            PLAN_ALLOC(tile_of, n_tflat * 4);
            PLAN_STEP(hipMemcpyAsync(tile_of, h_tile_of.data(), n_tflat * 4,
                                     hipMemcpyHostToDevice, s));
            p->n_tiles = (int64_t)h_tile_flat.size();
            PLAN_ALLOC(tile_flat, std::max<int64_t>(1, p->n_tiles) * 4);
            PLAN_STEP(hipMemsetAsync(tile_flat, 0xff, 4, s));
            if (p->n_tiles > 0)
                PLAN_STEP(hipMemcpyAsync(tile_flat, h_tile_flat.data(), p->n_tiles * 4,
                                         hipMemcpyHostToDevice, s));
        }
        const int64_t n_tiles = std::max<int64_t>(1, p->n_tiles);
        if (t_block_mb < 0) t_block_mb = 0;   // one time block: stretches are contiguous
        int64_t t_block = std::max<int64_t>(1, (t_block_mb << 20) / (4 * n_det));
        if (t_block_mb <= 0 || t_block > n_t) t_block = n_t;
        int64_t n_tb = (n_t + t_block - 1) / t_block;
        while (n_tb * n_tiles >= (int64_t)1 << 24) {  // keys and the segment table stay small
            t_block *= 2;
            n_tb = (n_t + t_block - 1) / t_block;
        }
        p->t_block = t_block;
        const int32_t off_key = (int32_t)(n_tb * n_tiles);
        int end_bit = 1;
        while ((1LL << end_bit) <= off_key) end_bit++;
        // sort by (time block, tile); stable, so sample order within a key
        PLAN_ALLOC(keys_in, n * 4);
        PLAN_ALLOC(vals_in, n * 4);
        PLAN_ALLOC(keys_out, n * 4);
        PLAN_ALLOC(p->idx, n * 4);
        PLAN_ALLOC(d_on, sizeof(unsigned long long));
        hipLaunchKernelGGL(k_tp_keys, dim3(grid_of(n, kPlanBlock, INT32_MAX)), dim3(kPlanBlock),
                           0, s, pixel, pg, n, ny, nx, gtx, tile_of, n_t, t_block,
                           (int32_t)n_tiles, off_key, keys_in, vals_in);
        PLAN_STEP(hipGetLastError());
        PLAN_STEP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, vals_in,
                                            p->idx, (size_t)n, 0, end_bit, s));
        PLAN_ALLOC(tmp, tmp_bytes);
        PLAN_STEP(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, p->idx,
                                            (size_t)n, 0, end_bit, s));
        hipLaunchKernelGGL(k_plan_count, dim3(1), dim3(1), 0, s, keys_out, n, off_key, d_on);
        PLAN_STEP(hipGetLastError());
        unsigned long long n_on = 0;
        PLAN_STEP(hipMemcpyAsync(&n_on, d_on, sizeof(n_on), hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipStreamSynchronize(s));
        p->n_on = (int64_t)n_on;
        (void)hipFree(keys_in); keys_in = nullptr;
        (void)hipFree(vals_in); vals_in = nullptr;
        (void)hipFree(tmp); tmp = nullptr;
        // slots and stored factors per sorted position
        PLAN_ALLOC(p->tslot, p->n_on * 2);
        PLAN_ALLOC(p->sspin, p->n_on * p->ns * sizeof(float));
        if (p->n_on > 0) {
            hipLaunchKernelGGL(k_tp_fill, dim3(grid_of(p->n_on)), dim3(kPlanBlock), 0, s, pixel,
                               pg, spin, nc, c0, p->idx, p->n_on, ny, nx, gtx, tile_of,
                               p->tslot, p->sspin);
            PLAN_STEP(hipGetLastError());
        }
        // stretches
        const int64_t n_chunks = (p->n_on + 63) / 64;
        PLAN_ALLOC(p->tmask, n_chunks * 8);
        PLAN_ALLOC(p->thead, n_chunks * 4);
        PLAN_ALLOC(gcount, n_chunks * 4);
        p->n_heads = 0;
        if (n_chunks > 0) {
            hipLaunchKernelGGL(k_tp_chunks, dim3(grid_of(n_chunks * 64, kPlanBlock, INT32_MAX)),
                               dim3(kPlanBlock), 0, s, keys_out, p->idx, p->n_on, n_chunks,
                               p->tmask, gcount);
            PLAN_STEP(hipGetLastError());
            PLAN_STEP(rocprim::exclusive_scan(nullptr, tmp_bytes, gcount, p->thead, 0u,
                                              (size_t)n_chunks, rocprim::plus<uint32_t>(), s));
            PLAN_ALLOC(tmp, tmp_bytes);
            PLAN_STEP(rocprim::exclusive_scan(tmp, tmp_bytes, gcount, p->thead, 0u,
                                              (size_t)n_chunks, rocprim::plus<uint32_t>(), s));
            uint32_t lb = 0, lc = 0;
            PLAN_STEP(hipMemcpyAsync(&lb, p->thead + n_chunks - 1, 4, hipMemcpyDeviceToHost, s));
            PLAN_STEP(hipMemcpyAsync(&lc, gcount + n_chunks - 1, 4, hipMemcpyDeviceToHost, s));
            PLAN_STEP(hipStreamSynchronize(s));
            p->n_heads = (int64_t)lb + lc;
            (void)hipFree(tmp); tmp = nullptr;
        }
        PLAN_ALLOC(p->head_i, p->n_heads * 4);
        if (p->n_on > 0) {
            hipLaunchKernelGGL(k_tp_heads, dim3(grid_of(p->n_on)), dim3(kPlanBlock), 0, s,
                               p->tmask, p->thead, p->idx, p->n_on, p->head_i);
            PLAN_STEP(hipGetLastError());
        }
        // work items: each (time block, tile) segment cut into pieces
        PLAN_ALLOC(seg, ((int64_t)off_key + 1) * 8);
        hipLaunchKernelGGL(k_tp_segs, dim3(grid_of((int64_t)off_key + 1, kPlanBlock, INT32_MAX)),
                           dim3(kPlanBlock), 0, s, keys_out, p->n_on, off_key, seg);
        PLAN_STEP(hipGetLastError());
        std::vector<int64_t> h_seg((size_t)off_key + 1);
        PLAN_STEP(hipMemcpyAsync(h_seg.data(), seg, h_seg.size() * 8, hipMemcpyDeviceToHost, s));
        PLAN_STEP(hipStreamSynchronize(s));
        std::vector<TItem> h_items;
        // at most plan_tile_item positions, and about 1024 items or more so
        // that a small observation still fills the GPU
        const int64_t per = std::max<int64_t>(
            64, std::min<int64_t>(g_plan_tile_item, std::max<int64_t>(8192, p->n_on / 1024)));
        for (int64_t v = 0; v < off_key; v++)
            for (int64_t b0 = h_seg[v]; b0 < h_seg[v + 1]; b0 += per)
                h_items.push_back(TItem{b0, (int32_t)std::min<int64_t>(per, h_seg[v + 1] - b0),
                                        (int32_t)(v % n_tiles)});
        p->n_items = (int64_t)h_items.size();
        PLAN_ALLOC(p->items, std::max<int64_t>(1, p->n_items) * (int64_t)sizeof(TItem));
        if (p->n_items > 0)
            PLAN_STEP(hipMemcpyAsync(p->items, h_items.data(), p->n_items * sizeof(TItem),
                                     hipMemcpyHostToDevice, s));
        // flat pixel of every slot
        PLAN_ALLOC(p->tile_pix, n_tiles * kTS * 4);
        if (k == 3) PLAN_ALLOC(p->tile_map, n_tiles * 4);
        hipLaunchKernelGGL(k_tp_tilepix, dim3(grid_of(n_tiles * kTS)), dim3(kPlanBlock), 0, s, pg,
                           ny, nx, gtx, tile_flat, n_tiles * kTS, p->tile_pix, p->tile_map);
        PLAN_STEP(hipGetLastError());
        if (k == 3) {
            // the map tiles the samples touch, checked against each call's tiles
            std::vector<int32_t> h_map(n_tiles, 0);
            if (p->n_tiles > 0)
                PLAN_STEP(hipMemcpyAsync(h_map.data(), p->tile_map, p->n_tiles * 4,
                                         hipMemcpyDeviceToHost, s));
            PLAN_STEP(hipStreamSynchronize(s));
            p->tile_used.assign(p->n_tile_map, 0);
            for (int64_t t = 0; t < p->n_tiles; t++) p->tile_used[h_map[t]] = 1;
        }
        // samples off the map (zeroed by a fresh from_map)
        p->n_off = n - p->n_on;
        PLAN_ALLOC(p->off_idx, p->n_off * 4);
        if (p->n_off > 0)
            PLAN_STEP(hipMemcpyAsync(p->off_idx, p->idx + p->n_on, p->n_off * 4,
                                     hipMemcpyDeviceToDevice, s));
        if (g_plan_time_copy) {
            PLAN_STEP(plan_build_tcopy(p, pixel, pg, spin, nc, c0, n, ny, nx, bitmap, wscan, s));
        }
        PLAN_STEP(hipStreamSynchronize(s));
        (void)hipFree(p->idx); p->idx = nullptr;
        cleanup();
        p->bytes = p->n_on * (2 + 4 * p->ns) + n_chunks * 12 + p->n_heads * 4 +
                   p->n_items * (int64_t)sizeof(TItem) + n_tiles * kTS * 4 + p->n_off * 4 +
                   p->n_vis * 4 + (p->tconst ? n_det * 4 : 0) +
                   (p->tpd != nullptr ? tcopy_bytes(p, n) : 0) +
                   (k == 3 ? (int64_t)p->n_tile_map * 28 + 12 + p->n_tiles * 4 + p->n_vis * 4
                           : 0);
        *out = p;
        return SO3G_OK;
    }
    // Time blocks whose signal slice (n_det x t_block float32) stays resident
    // in the Infinity Cache while the block's samples are gathered.
    if (t_block_mb < 0) t_block_mb = 64;
    int64_t t_block = std::max<int64_t>(1, (t_block_mb << 20) / (4 * n_det));
    if (t_block_mb <= 0 || t_block > n_t) t_block = n_t;
    int64_t n_tb = (n_t + t_block - 1) / t_block;
    while (n_tb * n_vis >= INT32_MAX) {  // keys must fit int32
        t_block *= 2;
        n_tb = (n_t + t_block - 1) / t_block;
    }
    p->t_block = t_block;
    const int32_t off_key = (int32_t)(n_tb * n_vis);
    int end_bit = 1;
    while ((1LL << end_bit) <= off_key) end_bit++;

    // 1. sort the samples by (time block, compact pixel)
    PLAN_ALLOC(keys_in, n * 4);
    PLAN_ALLOC(vals_in, n * 4);
    PLAN_ALLOC(keys_out, n * 4);
    PLAN_ALLOC(p->idx, n * 4);
    PLAN_ALLOC(d_on, sizeof(unsigned long long));
    hipLaunchKernelGGL(k_plan_keys, dim3(grid_of(n, kPlanBlock, INT32_MAX)), dim3(kPlanBlock), 0,
                       s, pixel, pg, n, ny, nx, n_t, t_block, n_vis, off_key, bitmap, wscan,
                       keys_in, vals_in);
    PLAN_STEP(hipGetLastError());
    PLAN_STEP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, vals_in, p->idx,
                                        (size_t)n, 0, end_bit, s));
    PLAN_ALLOC(tmp, tmp_bytes);
    PLAN_STEP(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, p->idx,
                                        (size_t)n, 0, end_bit, s));
    hipLaunchKernelGGL(k_plan_count, dim3(1), dim3(1), 0, s, keys_out, n, off_key, d_on);
    PLAN_STEP(hipGetLastError());
    unsigned long long n_on = 0;
    PLAN_STEP(hipMemcpyAsync(&n_on, d_on, sizeof(n_on), hipMemcpyDeviceToHost, s));
    PLAN_STEP(hipStreamSynchronize(s));
    p->n_on = (int64_t)n_on;
    (void)hipFree(keys_in); keys_in = nullptr;
    (void)hipFree(vals_in); vals_in = nullptr;
    (void)hipFree(tmp); tmp = nullptr;

    // 3. stored spin factors in sorted order
    PLAN_ALLOC(p->sspin, p->n_on * p->ns * sizeof(float));
    if (p->n_on > 0 && p->ns > 0) {
        hipLaunchKernelGGL(k_plan_sspin, dim3(grid_of(p->n_on)), dim3(kPlanBlock), 0, s, spin, nc,
                           c0, p->idx, p->n_on, p->sspin);
        PLAN_STEP(hipGetLastError());
    }

    // 4. head bits, run counts per group, run pixels
    const int64_t n_groups = std::max<int64_t>(1, (p->n_on + kGroup - 1) / kGroup);
    PLAN_ALLOC(p->heads, n_groups * (kGroup / 8));
    PLAN_ALLOC(gcount, n_groups * sizeof(uint32_t));
    PLAN_ALLOC(p->wbase, n_groups * sizeof(uint32_t));
    const int64_t n_bytes = n_groups * (kGroup / 8);
    const dim3 gb(grid_of(n_bytes, kPlanBlock, INT32_MAX));
    hipLaunchKernelGGL(k_plan_heads, gb, dim3(kPlanBlock), 0, s, keys_out, p->n_on, n_bytes,
                       p->heads, gcount);
    PLAN_STEP(hipGetLastError());
    PLAN_STEP(rocprim::exclusive_scan(nullptr, tmp_bytes, gcount, p->wbase, 0u,
                                      (size_t)n_groups, rocprim::plus<uint32_t>(), s));
    PLAN_ALLOC(tmp, tmp_bytes);
    PLAN_STEP(rocprim::exclusive_scan(tmp, tmp_bytes, gcount, p->wbase, 0u, (size_t)n_groups,
                                      rocprim::plus<uint32_t>(), s));
    uint32_t last_base = 0, last_count = 0;
    PLAN_STEP(hipMemcpyAsync(&last_base, p->wbase + n_groups - 1, 4, hipMemcpyDeviceToHost, s));
    PLAN_STEP(hipMemcpyAsync(&last_count, gcount + n_groups - 1, 4, hipMemcpyDeviceToHost, s));
    PLAN_STEP(hipStreamSynchronize(s));
    p->n_runs = (int64_t)last_base + last_count;
    PLAN_ALLOC(p->run_pix, p->n_runs * sizeof(uint32_t));
    hipLaunchKernelGGL(k_plan_runs, gb, dim3(kPlanBlock), 0, s, keys_out, n_bytes, n_vis,
                       p->comp2pix, p->heads, p->wbase, p->run_pix);
    PLAN_STEP(hipGetLastError());

    // 5. time-ordered copy for from_map, and the packed map buffer
    PLAN_STEP(plan_build_tcopy(p, pixel, pg, spin, nc, c0, n, ny, nx, bitmap, wscan, s));
    cleanup();
#undef PLAN_ALLOC
#undef PLAN_STEP
    p->bytes = p->n_on * (4 + 4 * p->ns) + n_groups * (kGroup / 8 + 4) + p->n_runs * 4 +
               tcopy_bytes(p, n) + p->n_vis * 4 + (p->tconst ? n_det * 4 : 0);
    *out = p;
    return SO3G_OK;
}

// Arguments shared by both entry points.
int plan_check_args(int32_t nc, int64_t n_det, int64_t n_t, so3g_pc_plan **out) {
    if (out == nullptr) return set_error(SO3G_ERR_VALUE, "out is NULL");
    *out = nullptr;
    if (nc < 1 || nc > 3) return set_error(SO3G_ERR_RUNTIME, "spin component count must be 1..3");
    if (n_det <= 0 || n_t <= 0 || n_det * n_t > (int64_t)UINT32_MAX)
        return set_error(SO3G_ERR_RUNTIME, "plan needs 0 < n_det * n_t < 2^32");
    if (n_det >= (1 << 22)) return set_error(SO3G_ERR_RUNTIME, "plan needs n_det < 2^22");
    // fast_div's remainder x - q * n_t lies in (-n_t, 2 n_t) and is held in
    // an int32.
    if (n_t >= (1 << 30)) return set_error(SO3G_ERR_RUNTIME, "plan needs n_t < 2^30");
    return SO3G_OK;
}

}  // namespace

extern "C" {

int so3g_pc_plan_create(const int32_t *pixel, int32_t k, const float *spin, int32_t nc,
                        int64_t n_det, int64_t n_t, const int64_t *map_shape,
                        int64_t t_block_mb, so3g_pc_plan **out, void *stream) {
    if (int rc = plan_check_args(nc, n_det, n_t, out)) return rc;
    if (k != 1 && k != 2)
        return set_error(SO3G_ERR_RUNTIME, "plans take k = 1 or 2 pixel indices here "
                                           "(tiled maps: so3g_pc_plan_create_tiled)");
    const int64_t ny = map_shape[0], nx = map_shape[1];
    if (ny <= 0 || nx <= 0 || ny * nx >= INT32_MAX || (k == 1 && nx != 1))
        return set_error(SO3G_ERR_RUNTIME, "map too large for a plan, or not (npix, 1) for k = 1");
    so3g_pc_plan *p = new so3g_pc_plan();
    return plan_create_impl(pixel, PixGeom{k, 0, nullptr, nullptr, nullptr}, spin, nc, n_det, n_t,
                            ny, nx, t_block_mb, p, out, (hipStream_t)stream);
}

int so3g_pc_plan_create_tiled(const int32_t *pixel, const float *spin, int32_t nc,
                              int64_t n_det, int64_t n_t, const int32_t *tile_dims,
                              int32_t n_tile, int64_t t_block_mb, so3g_pc_plan **out,
                              void *stream) {
    if (int rc = plan_check_args(nc, n_det, n_t, out)) return rc;
    if (tile_dims == nullptr || n_tile <= 0)
        return set_error(SO3G_ERR_VALUE, "tiled plan needs tile_dims for n_tile > 0 tiles");
    std::vector<int64_t> tbase(n_tile + 1, 0);
    std::vector<int32_t> sbase(n_tile + 1, 0);
    for (int32_t t = 0; t < n_tile; t++) {
        const int64_t ty = tile_dims[2 * t], tx = tile_dims[2 * t + 1];
        if (ty < 0 || tx < 0) return set_error(SO3G_ERR_VALUE, "negative tile shape");
        tbase[t + 1] = tbase[t] + ty * tx;
        sbase[t + 1] = (int32_t)(sbase[t] + ((ty + (1 << kTileBY) - 1) >> kTileBY) *
                                                ((tx + (1 << kTileBX) - 1) >> kTileBX));
        if (tbase[t + 1] >= INT32_MAX) return set_error(SO3G_ERR_RUNTIME, "map too large for a plan");
    }
    if (tbase[n_tile] == 0) return set_error(SO3G_ERR_VALUE, "tiles hold no pixels");
    hipStream_t s = (hipStream_t)stream;
    so3g_pc_plan *p = new so3g_pc_plan();
    p->n_tile_map = n_tile;
    p->n_sub = sbase[n_tile];
    auto fail = [&](hipError_t e) {
        free_plan_arrays(p);
        delete p;
        (void)hipGetLastError();
        return set_error(SO3G_ERR_HIP, std::string("HIP error setting up a tiled plan: ") +
                                           hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipMalloc(&p->tdims_d, (size_t)n_tile * 8)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->tbase_d, (size_t)(n_tile + 1) * 8)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->sbase_d, (size_t)(n_tile + 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->tiles_d, (size_t)n_tile * 8)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(p->tdims_d, tile_dims, (size_t)n_tile * 8, hipMemcpyHostToDevice,
                            s)) != hipSuccess ||
        (e = hipMemcpyAsync(p->tbase_d, tbase.data(), (size_t)(n_tile + 1) * 8,
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(p->sbase_d, sbase.data(), (size_t)(n_tile + 1) * 4,
                            hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
        return fail(e);
    return plan_create_impl(pixel, PixGeom{3, n_tile, p->tdims_d, p->tbase_d, p->sbase_d}, spin,
                            nc, n_det, n_t, tbase[n_tile], 1, t_block_mb, p, out, s);
}

static int plan_untiled(const so3g_pc_plan *p) {
    if (p != nullptr && p->n_tile_map > 0)
        return set_error(SO3G_ERR_RUNTIME, "plan was built for a tiled map: use the *_tiles calls");
    return SO3G_OK;
}

// Host array of the map's tile pointers into the plan's device table; every
// tile the plan's samples touch must be present.
static int plan_upload_tiles(const so3g_pc_plan *p, const void *const *tiles, hipStream_t s) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    if (p->n_tile_map == 0)
        return set_error(SO3G_ERR_RUNTIME, "plan was built for a non-tiled map");
    if (tiles == nullptr) return set_error(SO3G_ERR_VALUE, "tiles is NULL");
    for (int32_t t = 0; t < p->n_tile_map; t++)
        if (p->tile_used[t] && tiles[t] == nullptr)
            return set_error(SO3G_ERR_RUNTIME, "Tiling problem (index " + std::to_string(t) +
                                                   "): the plan's samples hit this tile but "
                                                   "it is missing.");
    PLAN_TRY(hipMemcpyAsync(p->tiles_d, tiles, (size_t)p->n_tile_map * sizeof(void *),
                            hipMemcpyHostToDevice, s));
    return SO3G_OK;
}

int so3g_pc_plan_to_map(const so3g_pc_plan *p, double *map, const float *signal,
                        const float *det_weights, void *stream) {
    if (int rc = plan_untiled(p)) return rc;
    return plan_to_map<float>(p, map, signal, det_weights, stream);
}

int so3g_pc_plan_to_map_set(const so3g_pc_plan *p, double *map, const float *signal,
                            const float *det_weights, void *stream) {
    if (int rc = plan_untiled(p)) return rc;
    if (p->layout == 2)
        return plan_tw_to_map<float>(p, map, signal, det_weights, (hipStream_t)stream, true);
    const int64_t n_map = (int64_t)p->nc * p->ny * p->nx;
    hipLaunchKernelGGL(k_plan_fill0, dim3(grid_of(n_map / 2 + 1)), dim3(kPlanBlock), 0,
                       (hipStream_t)stream, map, n_map);
    PLAN_TRY(hipGetLastError());
    return plan_to_map<float>(p, map, signal, det_weights, stream);
}

int so3g_pc_plan_to_map_f64(const so3g_pc_plan *p, double *map, const double *signal,
                            const float *det_weights, void *stream) {
    if (int rc = plan_untiled(p)) return rc;
    return plan_to_map<double>(p, map, signal, det_weights, stream);
}

int so3g_pc_plan_to_map_tiles(const so3g_pc_plan *p, double *const *tiles, const float *signal,
                              const float *det_weights, void *stream) {
    if (int rc = plan_upload_tiles(p, (const void *const *)tiles, (hipStream_t)stream)) return rc;
    return plan_to_map<float>(p, nullptr, signal, det_weights, stream, p->tiles_d);
}

static int plan_from_map_impl(const so3g_pc_plan *p, const double *map,
                              const double *const *tiles, float *signal, int accumulate,
                              void *stream) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int64_t plane = p->ny * p->nx;
    const int64_t n = p->n_det * p->n_t;
    if (p->layout == 2) {
        if (p->n_wfitems == 0) return SO3G_OK;
        const size_t lds = (size_t)p->nc * std::max(1, p->win_max) * sizeof(double);
        const dim3 gi((unsigned)p->n_wfitems);
        const WItem *it = (const WItem *)p->wfitems;
        const bool vec = p->wvec && ((uintptr_t)signal % 16) == 0;
#define TW_FROM2(NC, TC, V)                                                                 \
    do {                                                                                    \
        if (g_plan_win_block == 1024)                                                       \
            hipLaunchKernelGGL((k_tw_from_map<NC, TC, V, 1024>), gi, dim3(1024), lds, s, it, \
                               p->wslot, p->tspin, p->tval, p->comp2pix, map, plane,        \
                               signal, accumulate);                                         \
        else                                                                                \
            hipLaunchKernelGGL((k_tw_from_map<NC, TC, V, 256>), gi, dim3(256), lds, s, it,  \
                               p->wslot, p->tspin, p->tval, p->comp2pix, map, plane,        \
                               signal, accumulate);                                         \
    } while (0)
#define TW_FROM(NC, TC)                                                                     \
    do {                                                                                    \
        if (vec) TW_FROM2(NC, TC, true); else TW_FROM2(NC, TC, false);                      \
    } while (0)
        if (p->tconst) {
            if (p->nc == 1) TW_FROM(1, true);
            else TW_FROM(3, true);
        } else {
            if (p->nc == 1) TW_FROM(1, false);
            else if (p->nc == 2) TW_FROM(2, false);
            else TW_FROM(3, false);
        }
#undef TW_FROM2
#undef TW_FROM
        PLAN_TRY(hipGetLastError());
        return SO3G_OK;
    }
    if (p->layout == 1 && p->tpd == nullptr) {
        if (!accumulate && p->n_off > 0) {
            hipLaunchKernelGGL(k_tp_zero, dim3(grid_of(p->n_off)), dim3(kPlanBlock), 0, s,
                               p->off_idx, p->n_off, signal);
            PLAN_TRY(hipGetLastError());
        }
        if (p->n_items == 0) return SO3G_OK;
        const dim3 gi((unsigned)p->n_items);
        const TItem *items = (const TItem *)p->items;
        const FastDiv nt{(uint32_t)p->n_t, 1.f / (float)p->n_t};
        const uint32_t pl = (uint32_t)plane;
#define TP_FROM(NC, TC)                                                                     \
    hipLaunchKernelGGL((k_tp_from_map<NC, TC>), gi, dim3(kTBlock), 0, s, items, p->tslot,    \
                       p->sspin, p->tmask, p->thead, p->head_i, p->tile_pix,     \
                       p->tval, nt, pl, map, signal, accumulate, tiles, p->tile_map, p->tdims_d)
        if (p->tconst) {
            if (p->nc == 1) TP_FROM(1, true);
            else TP_FROM(3, true);
        } else {
            if (p->nc == 1) TP_FROM(1, false);
            else if (p->nc == 2) TP_FROM(2, false);
            else TP_FROM(3, false);
        }
#undef TP_FROM
        PLAN_TRY(hipGetLastError());
        return SO3G_OK;
    }
    if (tiles != nullptr) {
        const dim3 gp(grid_of(p->n_vis));
        switch (p->nc) {
        case 1: hipLaunchKernelGGL(k_plan_pack_tiles<1>, gp, dim3(kPlanBlock), 0, s, tiles, p->tdims_d, p->comp_tile, p->comp2pix, p->n_vis, p->packed); break;
        case 2: hipLaunchKernelGGL(k_plan_pack_tiles<2>, gp, dim3(kPlanBlock), 0, s, tiles, p->tdims_d, p->comp_tile, p->comp2pix, p->n_vis, p->packed); break;
        default: hipLaunchKernelGGL(k_plan_pack_tiles<3>, gp, dim3(kPlanBlock), 0, s, tiles, p->tdims_d, p->comp_tile, p->comp2pix, p->n_vis, p->packed); break;
        }
    } else switch (p->nc) {
    case 1: hipLaunchKernelGGL(k_plan_pack<1>, dim3(grid_of(p->n_vis)), dim3(kPlanBlock), 0, s, map, plane, p->comp2pix, p->n_vis, p->packed); break;
    case 2: hipLaunchKernelGGL(k_plan_pack<2>, dim3(grid_of(p->n_vis)), dim3(kPlanBlock), 0, s, map, plane, p->comp2pix, p->n_vis, p->packed); break;
    default: hipLaunchKernelGGL(k_plan_pack<3>, dim3(grid_of(p->n_vis)), dim3(kPlanBlock), 0, s, map, plane, p->comp2pix, p->n_vis, p->packed); break;
    }
    PLAN_TRY(hipGetLastError());
    const bool strided = g_plan_from_layout != 0;
    const int vv = strided ? g_plan_from_v : (n % 8 == 0 ? g_plan_from_v : (n % 4 == 0 ? 4 : 1));
    const dim3 g(grid_of(n, (int64_t)kPlanBlock * vv, INT32_MAX));
    const dim3 gw(grid_of(p->n_wide * (kTGroup / 256), kPlanBlock / 64));
    const FastDiv nt{(uint32_t)p->n_t, 1.f / (float)p->n_t};
#define PLAN_FROM(NC, TC)                                                                   \
    do {                                                                                    \
        if (strided && vv == 8)                                                             \
            hipLaunchKernelGGL((k_plan_from_map_s<NC, TC, 8, 1, false>), g, dim3(kPlanBlock), \
                               0, s, p->tpd, p->tpb, p->tpw, p->wgrp, p->n_wide, p->tspin,  \
                               p->tval, p->packed, n, nt, signal, accumulate);              \
        else if (strided)                                                                   \
            hipLaunchKernelGGL((k_plan_from_map_s<NC, TC, 4, 1, false>), g, dim3(kPlanBlock), \
                               0, s, p->tpd, p->tpb, p->tpw, p->wgrp, p->n_wide, p->tspin,  \
                               p->tval, p->packed, n, nt, signal, accumulate);              \
        else if (vv == 8)                                                                   \
            hipLaunchKernelGGL((k_plan_from_map<NC, TC, 8, 1>), g, dim3(kPlanBlock), 0, s,   \
                               p->tpd, p->tpb, p->tspin, p->tval, p->packed, n, nt, signal, \
                               accumulate);                                                 \
        else if (vv == 4)                                                                   \
            hipLaunchKernelGGL((k_plan_from_map<NC, TC, 4, 1>), g, dim3(kPlanBlock), 0, s,   \
                               p->tpd, p->tpb, p->tspin, p->tval, p->packed, n, nt, signal, \
                               accumulate);                                                 \
        else                                                                                \
            hipLaunchKernelGGL((k_plan_from_map<NC, TC, 1, 1>), g, dim3(kPlanBlock), 0, s,   \
                               p->tpd, p->tpb, p->tspin, p->tval, p->packed, n, nt, signal, \
                               accumulate);                                                 \
        if (p->n_wide > 0)                                                                  \
            hipLaunchKernelGGL((k_plan_from_map_s<NC, TC, 4, 1, true>), gw, dim3(kPlanBlock), \
                               0, s, p->tpd, p->tpb, p->tpw, p->wgrp, p->n_wide, p->tspin,  \
                               p->tval, p->packed, n, nt, signal, accumulate);              \
    } while (0)
    if (p->tconst) {
        if (p->nc == 1) PLAN_FROM(1, true);
        else PLAN_FROM(3, true);
    } else {
        if (p->nc == 1) PLAN_FROM(1, false);
        else if (p->nc == 2) PLAN_FROM(2, false);
        else PLAN_FROM(3, false);
    }
#undef PLAN_FROM
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

int so3g_precomp_from_map_f64(const double *map, const int64_t *map_shape,
                              const int32_t *pixel, const float *spin, int32_t nc,
                              int64_t n_det, int64_t n_t, double *signal, int accumulate,
                              void *stream) {
    if (nc < 1 || nc > 3) return set_error(SO3G_ERR_RUNTIME, "spin component count must be 1..3");
    if (n_det < 0 || n_t < 0) return set_error(SO3G_ERR_RUNTIME, "negative sample count");
    const int64_t n = n_det * n_t;
    if (n == 0) return SO3G_OK;
    const int64_t ny = map_shape[0], nx = map_shape[1];
    if (ny <= 0 || nx <= 0 || ny > INT32_MAX || nx > INT32_MAX)
        return set_error(SO3G_ERR_RUNTIME, "bad map shape");
    hipStream_t s = (hipStream_t)stream;
    const dim3 g(grid_of(n));
    switch (nc) {
    case 1: hipLaunchKernelGGL(k_pc_from_map_f64<1>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    case 2: hipLaunchKernelGGL(k_pc_from_map_f64<2>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    default: hipLaunchKernelGGL(k_pc_from_map_f64<3>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    }
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

int so3g_pc_plan_to_weights(const so3g_pc_plan *p, double *map, const float *det_weights,
                            void *stream) {
    if (int rc = plan_untiled(p)) return rc;
    return plan_to_weights(p, map, det_weights, stream, nullptr);
}

int so3g_pc_plan_to_weights_tiles(const so3g_pc_plan *p, double *const *tiles,
                                  const float *det_weights, void *stream) {
    if (int rc = plan_upload_tiles(p, (const void *const *)tiles, (hipStream_t)stream)) return rc;
    return plan_to_weights(p, nullptr, det_weights, stream, p->tiles_d);
}

int so3g_pc_plan_from_map(const so3g_pc_plan *p, const double *map, float *signal,
                          int accumulate, void *stream) {
    if (int rc = plan_untiled(p)) return rc;
    return plan_from_map_impl(p, map, nullptr, signal, accumulate, stream);
}

int so3g_pc_plan_from_map_tiles(const so3g_pc_plan *p, const double *const *tiles,
                                float *signal, int accumulate, void *stream) {
    if (int rc = plan_upload_tiles(p, (const void *const *)tiles, (hipStream_t)stream)) return rc;
    return plan_from_map_impl(p, nullptr, p->tiles_d, signal, accumulate, stream);
}

int so3g_pc_plan_info(const so3g_pc_plan *p, int64_t *n_on, int64_t *bytes) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    *n_on = p->n_on;
    *bytes = p->bytes;
    return SO3G_OK;
}

int so3g_pc_plan_stats(const so3g_pc_plan *p, int64_t *stats) {
    if (p == nullptr || stats == nullptr) return set_error(SO3G_ERR_VALUE, "NULL argument");
    stats[0] = p->n_on;
    stats[1] = p->n_runs;
    stats[2] = p->tconst;
    stats[3] = p->ns;
    stats[4] = p->t_block;
    stats[5] = p->bytes;
    stats[6] = p->n_vis;
    stats[7] = p->layout;
    stats[8] = p->n_heads;
    stats[9] = p->layout == 2 ? p->n_witems : p->n_items;
    stats[10] = p->n_wide;
    stats[11] = p->win_max;
    stats[12] = p->n_wslots;
    return SO3G_OK;
}

int so3g_pc_plan_destroy(so3g_pc_plan *p) {
    if (p == nullptr) return SO3G_OK;
    free_plan_arrays(p);
    delete p;
    return SO3G_OK;
}

}  // extern "C"
