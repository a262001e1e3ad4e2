// batch.h -- one launch over many observations projected into one map.
//
// A batch of observations (each its own boresight, focal plane, responses
// and signal) is cut into equal work items of kBlock time samples x dpb
// detectors.  Items are numbered observation by observation (obs.item0 is
// the first item of each, time tiles fastest) and a persistent grid walks
// them with a block stride, so that a single launch -- one node of a
// captured graph -- covers observations of any mix of sizes without a
// host round trip between them.  Each item runs the same per-sample body
// as the single-observation kernels (from_map_sample, to_map_sample,
// lds_scatter_det with a table per item, or row_item over row-segment
// items), so results match one launch per observation.
#pragma once

#include "kernels.h"

namespace so3g {

struct ObsDev {
    const double *bore;        // (n_t, 4) native-frame boresight
    const double *dets;        // (n_det, 4)
    const float *resp;         // (n_det, 2)
    const float *sig_in;       // to_map input (n_det, n_t)
    float *sig_out;            // from_map output (n_det, n_t)
    const float *det_w;        // (n_det) or nullptr
    int64_t n_t, n_det;
    int64_t item0;             // first work item of this observation
    int64_t n_tt;              // time tiles, ceil(n_t / kBlock)
    const int32_t *order;      // detector order of the LDS-table scatter
    // row scatter: groups of G detectors sharing a home row
    const int32_t *rgdet;      // [n_groups * G], -1 padded
    const int32_t *rgrow;      // [n_groups] home row (-1: none)
    const int32_t *rgxc;       // [n_groups] home column
    int64_t n_groups;
    int64_t ritem0;            // first row item of this observation
};

// Calls f(obs, in, d, t, qbore, valid) for every detector of the block's
// current item; all lanes take part (valid == false past the end of the
// observation) so that the wave-level aggregation sees 64 lanes.
template <class F>
__device__ __forceinline__ void for_batch_items(const ObsDev *obs, int n_obs, int64_t n_items,
                                                int dpb, const Tabs &tabs, int agg, F &&f) {
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        int lo = 0, hi = n_obs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (obs[mid].item0 <= it) lo = mid;
            else hi = mid - 1;
        }
        const ObsDev &o = obs[lo];
        const int64_t local = it - o.item0;
        const int64_t dg = local / o.n_tt;
        const int64_t t = (local - dg * o.n_tt) * kBlock + threadIdx.x;
        const bool valid = t < o.n_t;
        const Q4 qb = valid ? load_q(o.bore + 4 * t) : Q4{1., 0., 0., 0.};
        const PtgIn in{o.bore, o.dets, o.n_t, o.n_det, tabs, dpb, agg, nullptr};
        const int64_t d0 = dg * dpb;
        const int64_t d1 = min(d0 + (int64_t)dpb, o.n_det);
        for (int64_t d = d0; d < d1; d++) f(o, in, d, t, qb, valid);
    }
}

template <int PROJ, int PK, int INTERP, int NC>
__global__ __launch_bounds__(kBlock) void k_batch_to_map(const ObsDev *obs, int n_obs,
                                                         int64_t n_items, int dpb, Tabs tabs,
                                                         int agg, Geo g, MapView<double> m,
                                                         int *flags) {
    for_batch_items(obs, n_obs, n_items, dpb, tabs, agg,
                    [&](const ObsDev &o, const PtgIn &in, int64_t d, int64_t t, const Q4 &qb,
                        bool valid) {
                        to_map_sample<PROJ, PK, INTERP, NC, 0>(in, g, m, o.resp, o.sig_in,
                                                               o.det_w, flags, d, t, qb,
                                                               valid);
                    });
}

// to_map of a batch through LDS-privatised partial maps: each work item
// (256 samples x dpb detectors of one observation, the detectors taken in
// the observation's row order, so that an item's detectors sweep the same
// pixels) adds into a table of H slots in LDS (home slot = element number
// mod H, so a stretch of pixels takes a stretch of slots), flushed to the
// map with one FP64 atomic per occupied slot and value before the block's
// next item; a run that finds no slot adds to the map directly.
template <int PROJ, int PK, int INTERP, int NC, int H>
__global__ __launch_bounds__(kBlock) void k_batch_to_map_lds(const ObsDev *obs, int n_obs,
                                                             int64_t n_items, int dpb,
                                                             Tabs tabs, Geo g,
                                                             MapView<double> m, int *flags) {
    __shared__ int32_t hk[H];
    __shared__ double hv[NC * H];
    for (int64_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        lds_table_clear<NC, H>(hk, hv, kBlock);
        __syncthreads();
        int lo = 0, hi = n_obs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (obs[mid].item0 <= it) lo = mid;
            else hi = mid - 1;
        }
        const ObsDev &o = obs[lo];
        const int64_t local = it - o.item0;
        const int64_t dg = local / o.n_tt;
        const int64_t t = (local - dg * o.n_tt) * kBlock + threadIdx.x;
        const bool valid = t < o.n_t;
        const Q4 qb = valid ? load_q(o.bore + 4 * t) : Q4{1., 0., 0., 0.};
        const PtgIn in{o.bore, o.dets, o.n_t, o.n_det, tabs, dpb, 1, nullptr};
        const int64_t d0 = dg * dpb;
        const int64_t d1 = min(d0 + (int64_t)dpb, o.n_det);
        for (int64_t j = d0; j < d1; j++)
            lds_scatter_det<PROJ, PK, INTERP, NC, false, H, true>(
                in, g, m, o.resp, o.sig_in, o.det_w, flags, o.order ? o.order[j] : j, t, qb,
                valid, hk, hv);
        __syncthreads();
        lds_table_flush<PK, NC, false, H>(g, m, hk, hv, kBlock);
        __syncthreads();
    }
}

// to_map of a batch through map-row segments in LDS: one block per item
// (observation, sample span of `span`, group of G row-mate detectors; group
// fastest within an observation), each summing into a segment of xw columns
// of its group's home row (row_item), so the whole batch runs in one launch
// with the single observation's row scatter.
template <int PROJ, int PK, int NC, int NT, int R, bool LEAN>
__device__ __forceinline__ void batch_row_block(const ObsDev *obs, int n_obs, const Tabs &tabs,
                                                const Geo &g, const MapView<double> &m, int G,
                                                int span, int xw, int *flags, double *racc) {
    const int64_t it = blockIdx.x;
    int lo = 0, hi = n_obs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (obs[mid].ritem0 <= it) lo = mid;
        else hi = mid - 1;
    }
    const ObsDev &o = obs[lo];
    const int64_t local = it - o.ritem0;
    const int64_t ts = local / o.n_groups;
    const int64_t gi = local - ts * o.n_groups;
    const int64_t t_beg = ts * span;
    const int64_t t_end = min(t_beg + (int64_t)span, o.n_t);
    const PtgIn in{o.bore, o.dets, o.n_t, o.n_det, tabs, 1, 0, nullptr};
    row_item<PROJ, PK, NC, false, NT, R, LEAN>(in, g, m, o.resp, o.sig_in, o.det_w,
                                               o.rgdet + gi * G, G, o.rgrow[gi],
                                               row_x0(o.rgxc[gi], xw, g.nx), xw, t_beg, t_end,
                                               flags, racc);
}

template <int PROJ, int PK, int NC, int NT, int R>
__global__ __launch_bounds__(NT) void k_batch_to_map_row(const ObsDev *obs, int n_obs,
                                                         Tabs tabs, Geo g, MapView<double> m,
                                                         int G, int span, int xw, int *flags) {
    extern __shared__ double racc[];
    batch_row_block<PROJ, PK, NC, NT, R, false>(obs, n_obs, tabs, g, m, G, span, xw, flags,
                                                racc);
}

// The same with at most 64 registers per lane, two blocks per compute unit
// where their segments fit its LDS together (as k_scatter_row_x2).
template <int PROJ, int PK, int NC, int NT, int R>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_batch_to_map_row_x2(const ObsDev *obs, int n_obs, Tabs tabs, Geo g, MapView<double> m, int G,
                      int span, int xw, int *flags) {
    extern __shared__ double racc[];
    batch_row_block<PROJ, PK, NC, NT, R, true>(obs, n_obs, tabs, g, m, G, span, xw, flags,
                                               racc);
}

template <int PROJ, int PK, int INTERP, int NC>
__global__ __launch_bounds__(kBlock) void k_batch_from_map(const ObsDev *obs, int n_obs,
                                                           int64_t n_items, int dpb, Tabs tabs,
                                                           Geo g, MapView<const double> m,
                                                           int accumulate, int *flags) {
    for_batch_items(obs, n_obs, n_items, dpb, tabs, 0,
                    [&](const ObsDev &o, const PtgIn &in, int64_t d, int64_t t, const Q4 &qb,
                        bool valid) {
                        if (valid)
                            from_map_sample<PROJ, PK, INTERP, NC>(in, g, m, o.resp, o.sig_out,
                                                                  accumulate, flags, d, t, qb);
                    });
}

}  // namespace so3g
