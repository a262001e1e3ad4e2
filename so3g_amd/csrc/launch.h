// launch.h -- template dispatch from a runtime engine description to the
// kernel instantiations of one coordinate system.  Each k_<proj>.hip
// translation unit instantiates dispatch<PROJ> for its projection, so the
// kernel variants of the eight coordinate systems compile in parallel.
#pragma once

#include "batch.h"
#include "kernels.h"
#include "precomp.h"

namespace so3g {

enum { OP_COORDS = 0, OP_PIXELS, OP_PMAT, OP_FROM_MAP, OP_TO_MAP, OP_TO_WEIGHTS,
       OP_TILE_HITS, OP_DOMAINS, OP_BATCH_FROM_MAP, OP_BATCH_TO_MAP };

struct OpArgs {
    int op;
    PtgIn in;
    Geo g;
    int pk, interp, nc, prec;
    int64_t n_chunks;          // segment mode when in.chunks != nullptr
    hipStream_t stream;
    int *flags;
    const float *resp;
    const float *sig_in;
    float *sig_out;
    const float *det_w;
    void *map;
    void *const *tiles;
    const int32_t *tile_dims;
    double *coords_out;
    int32_t *pix_out;
    float *spin_out;
    int accumulate;
    int mode, n_domain;
    const int32_t *tile_dom;
    int32_t *dom_out;
    int32_t *hits;
    // LDS-privatised scatter (to_map / to_weights): hash_cls 0 = global
    // atomics, 1 = large tables, 2 = small tables (launch_scatter_cls),
    // 3 = map rows in LDS (launch_scatter_row).
    int hash_cls;
    HashWork hw;
    RowWork rw;
    // batch ops
    int batch_slots = 0;       // batch to_map: 0 global atomics, 512 / 1024 LDS
                               // table slots, < 0 row segments (width in rw.xw)
    const ObsDev *obs;
    int n_obs;
    int64_t n_items;
    int grid_blocks;
};

// Slots of an LDS hash table of `budget` bytes with NV float64 sums and an
// int32 key per slot (a power of two).
constexpr int hash_slots(int nv, int budget) {
    int h = 1;
    while (2 * h * (8 * nv + 4) <= budget) h *= 2;
    return h;
}

template <int PROJ, int PK, int INTERP, int NC, bool W, int NT, int H, bool MOD>
void launch_scatter_lds(const OpArgs &a) {
    const dim3 grid((unsigned)(((a.in.n_t + a.hw.span - 1) / a.hw.span) *
                               ((a.in.n_det + a.hw.dets - 1) / a.hw.dets)));
    MapView<double> m{(double *)a.map, (double *const *)a.tiles, a.tile_dims};
    if (INTERP == 0 && a.hw.r == 4)
        hipLaunchKernelGGL((k_scatter_lds<PROJ, PK, INTERP, NC, W, NT, H, MOD, 4>), grid,
                           dim3(NT), 0, a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w, a.hw,
                           a.flags);
    else
        hipLaunchKernelGGL((k_scatter_lds<PROJ, PK, INTERP, NC, W, NT, H, MOD, 1>), grid,
                           dim3(NT), 0, a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w, a.hw,
                           a.flags);
}

// hash_cls 1: 1024 threads and a 112 KB table per block (one block per
// compute unit), multiplicative hash; hash_cls 2: 256 threads and 512 slots
// (14 KB for to_map TQU), home slot = key mod 512.
template <int PROJ, int PK, int INTERP, int NC, bool W>
void launch_scatter_cls(const OpArgs &a) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    if (a.hash_cls == 2)
        launch_scatter_lds<PROJ, PK, INTERP, NC, W, 256, 512, true>(a);
    else
        launch_scatter_lds<PROJ, PK, INTERP, NC, W, 1024, hash_slots(NV, 112 * 1024), false>(a);
}

// Rows in LDS: 1024 threads of 4 samples each, a row segment's NV x xw
// float64 sums as dynamic LDS (flat maps, nearest-neighbour pixels).
template <int PROJ, int PK, int NC, bool W>
void launch_scatter_row(const OpArgs &a) {
    constexpr int NV = W ? NC * (NC + 1) / 2 : NC;
    constexpr int NT = 1024;
    const size_t lds = row_lds_bytes<PK>(NV, a.rw.xw);
    const dim3 grid((unsigned)(((a.in.n_t + a.rw.span - 1) / a.rw.span) * a.rw.n_groups));
    MapView<double> m{(double *)a.map, (double *const *)a.tiles, a.tile_dims};
    // two blocks per compute unit (160 KB of LDS) where the segments fit:
    // the 64-register instance; the weights (6 sums per pixel) and CEA
    // (its correctly rounded atan2) would spill under that cap
    if constexpr (!W && PROJ != PROJ_CEA) {
        if (2 * lds <= 160 * 1024) {
            (void)hipFuncSetAttribute((const void *)k_scatter_row_x2<PROJ, PK, NC, W, NT, 4>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipLaunchKernelGGL((k_scatter_row_x2<PROJ, PK, NC, W, NT, 4>), grid, dim3(NT), lds,
                               a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w, a.rw,
                               a.flags);
            return;
        }
    }
    (void)hipFuncSetAttribute((const void *)k_scatter_row<PROJ, PK, NC, W, NT, 4>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_scatter_row<PROJ, PK, NC, W, NT, 4>), grid, dim3(NT), lds, a.stream,
                       a.in, a.g, m, a.resp, a.sig_in, a.det_w, a.rw, a.flags);
}

// Grid of the on-the-fly kernels: one block per chunk (segment mode), or
// one per (detector group, 256-sample chunk), group fastest (dense_block).
inline dim3 grid_for(const OpArgs &a) {
    if (a.in.chunks != nullptr) return dim3((unsigned)a.n_chunks);
    const int64_t bt = (a.in.n_t + kBlock - 1) / kBlock;
    const int64_t bg = (a.in.n_det + a.in.dpb - 1) / a.in.dpb;
    return dim3((unsigned)(bt * bg));
}

template <int PROJ, int PK, int INTERP, int NC>
int launch_map_ops(const OpArgs &a) {
    const dim3 block(kBlock);
    if (a.op == OP_BATCH_TO_MAP || a.op == OP_BATCH_FROM_MAP) {
        if (PK == PK_FLAT_TILED || PK == PK_HP_TILED ? a.tiles == nullptr : a.map == nullptr)
            return -1;   // no map to project from or into
        const dim3 bgrid((unsigned)a.grid_blocks);
        if (a.op == OP_BATCH_TO_MAP) {
            MapView<double> m{(double *)a.map, (double *const *)a.tiles, a.tile_dims};
            if (a.batch_slots < 0) {
                if constexpr ((PK == PK_FLAT || PK == PK_FLAT_TILED) && INTERP == 0) {
                    constexpr int NT = 1024;
                    const size_t lds = row_lds_bytes<PK>(NC, a.rw.xw);
                    if constexpr (PROJ != PROJ_CEA) {
                        if (2 * lds <= 160 * 1024) {
                            (void)hipFuncSetAttribute(
                                (const void *)k_batch_to_map_row_x2<PROJ, PK, NC, NT, 4>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                            hipLaunchKernelGGL((k_batch_to_map_row_x2<PROJ, PK, NC, NT, 4>),
                                               bgrid, dim3(NT), lds, a.stream, a.obs, a.n_obs,
                                               a.in.tabs, a.g, m, a.rw.G, a.rw.span, a.rw.xw,
                                               a.flags);
                            return 0;
                        }
                    }
                    (void)hipFuncSetAttribute(
                        (const void *)k_batch_to_map_row<PROJ, PK, NC, NT, 4>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                    hipLaunchKernelGGL((k_batch_to_map_row<PROJ, PK, NC, NT, 4>), bgrid,
                                       dim3(NT), lds, a.stream, a.obs, a.n_obs, a.in.tabs,
                                       a.g, m, a.rw.G, a.rw.span, a.rw.xw, a.flags);
                    return 0;
                } else {
                    return -1;   // the host offers row segments for such maps only
                }
            }
            if (a.batch_slots == 512)
                hipLaunchKernelGGL((k_batch_to_map_lds<PROJ, PK, INTERP, NC, 512>), bgrid,
                                   block, 0, a.stream, a.obs, a.n_obs, a.n_items, a.in.dpb,
                                   a.in.tabs, a.g, m, a.flags);
            else if (a.batch_slots == 1024)
                hipLaunchKernelGGL((k_batch_to_map_lds<PROJ, PK, INTERP, NC, 1024>), bgrid,
                                   block, 0, a.stream, a.obs, a.n_obs, a.n_items, a.in.dpb,
                                   a.in.tabs, a.g, m, a.flags);
            else
                hipLaunchKernelGGL((k_batch_to_map<PROJ, PK, INTERP, NC>), bgrid, block, 0,
                                   a.stream, a.obs, a.n_obs, a.n_items, a.in.dpb, a.in.tabs,
                                   a.in.agg, a.g, m, a.flags);
        } else {
            MapView<const double> m{(const double *)a.map, (const double *const *)a.tiles,
                                    a.tile_dims};
            hipLaunchKernelGGL((k_batch_from_map<PROJ, PK, INTERP, NC>), bgrid, block, 0,
                               a.stream, a.obs, a.n_obs, a.n_items, a.in.dpb, a.in.tabs, a.g,
                               m, a.accumulate, a.flags);
        }
        return 0;
    }
    if constexpr ((PK == PK_FLAT || PK == PK_FLAT_TILED) && INTERP == 0) {
        if ((a.op == OP_TO_MAP || a.op == OP_TO_WEIGHTS) && a.prec == 0 && a.hash_cls == 3 &&
            a.in.chunks == nullptr) {
            if (a.op == OP_TO_MAP) launch_scatter_row<PROJ, PK, NC, false>(a);
            else launch_scatter_row<PROJ, PK, NC, true>(a);
            return 0;
        }
    }
    if ((a.op == OP_TO_MAP || a.op == OP_TO_WEIGHTS) && a.prec == 0 && a.hash_cls > 0 &&
        a.hash_cls < 3 && a.in.chunks == nullptr) {
        if (a.op == OP_TO_MAP) launch_scatter_cls<PROJ, PK, INTERP, NC, false>(a);
        else launch_scatter_cls<PROJ, PK, INTERP, NC, true>(a);
        return 0;
    }
    const dim3 grid = grid_for(a);
    switch (a.op) {
    case OP_FROM_MAP: {
        MapView<const double> m{(const double *)a.map, (const double *const *)a.tiles,
                                a.tile_dims};
        hipLaunchKernelGGL((k_from_map<PROJ, PK, INTERP, NC>), grid, block, 0, a.stream,
                           a.in, a.g, m, a.resp, a.sig_out, a.accumulate, a.flags);
        return 0;
    }
    case OP_TO_MAP:
        if (a.prec == 1) {
            MapView<float> m{(float *)a.map, (float *const *)a.tiles, a.tile_dims};
            hipLaunchKernelGGL((k_to_map<PROJ, PK, INTERP, NC, 1>), grid, block, 0,
                               a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w, a.flags);
        } else {
            MapView<double> m{(double *)a.map, (double *const *)a.tiles, a.tile_dims};
            if (a.prec == 2)
                hipLaunchKernelGGL((k_to_map<PROJ, PK, INTERP, NC, 2>), grid, block, 0,
                                   a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w,
                                   a.flags);
            else
                hipLaunchKernelGGL((k_to_map<PROJ, PK, INTERP, NC, 0>), grid, block, 0,
                                   a.stream, a.in, a.g, m, a.resp, a.sig_in, a.det_w,
                                   a.flags);
        }
        return 0;
    case OP_TO_WEIGHTS:
        if (a.prec == 1) {
            MapView<float> m{(float *)a.map, (float *const *)a.tiles, a.tile_dims};
            hipLaunchKernelGGL((k_to_weights<PROJ, PK, INTERP, NC, 1>), grid, block, 0,
                               a.stream, a.in, a.g, m, a.resp, a.det_w, a.flags);
        } else {
            MapView<double> m{(double *)a.map, (double *const *)a.tiles, a.tile_dims};
            if (a.prec == 2)
                hipLaunchKernelGGL((k_to_weights<PROJ, PK, INTERP, NC, 2>), grid, block, 0,
                                   a.stream, a.in, a.g, m, a.resp, a.det_w, a.flags);
            else
                hipLaunchKernelGGL((k_to_weights<PROJ, PK, INTERP, NC, 0>), grid, block, 0,
                                   a.stream, a.in, a.g, m, a.resp, a.det_w, a.flags);
        }
        return 0;
    }
    return -1;
}

template <int PROJ, int PK, int INTERP>
int launch_interp(const OpArgs &a) {
    const dim3 grid = grid_for(a);
    const dim3 block(kBlock);
    if (a.op == OP_TILE_HITS) {
        hipLaunchKernelGGL((k_tile_hits<PROJ, PK, INTERP>), grid, block, 0, a.stream, a.in,
                           a.g, a.hits);
        return 0;
    }
    if (a.op == OP_DOMAINS) {
        MapView<const double> m{(const double *)a.map, (const double *const *)a.tiles,
                                a.tile_dims};
        hipLaunchKernelGGL((k_domains<PROJ, PK, INTERP>), grid, block, 0, a.stream, a.in,
                           a.g, a.mode, a.n_domain, m, a.tile_dom, a.dom_out, a.flags);
        return 0;
    }
    switch (a.nc) {
    case 1: return launch_map_ops<PROJ, PK, INTERP, 1>(a);
    case 2: return launch_map_ops<PROJ, PK, INTERP, 2>(a);
    case 3: return launch_map_ops<PROJ, PK, INTERP, 3>(a);
    }
    return -1;
}

template <int PROJ, int PK>
int launch_pk(const OpArgs &a) {
    const dim3 grid = grid_for(a);
    const dim3 block(kBlock);
    if (a.op == OP_PIXELS || a.op == OP_PMAT) {
        float *spin = (a.op == OP_PMAT) ? a.spin_out : nullptr;
        if (spin == nullptr) {
            hipLaunchKernelGGL((k_pixels<PROJ, PK, 1, false>), grid, block, 0, a.stream, a.in,
                               a.g, a.resp, a.pix_out, spin, a.flags);
        } else {
            switch (a.nc) {
            case 1:
                hipLaunchKernelGGL((k_pixels<PROJ, PK, 1, true>), grid, block, 0, a.stream,
                                   a.in, a.g, a.resp, a.pix_out, spin, a.flags);
                break;
            case 2:
                hipLaunchKernelGGL((k_pixels<PROJ, PK, 2, true>), grid, block, 0, a.stream,
                                   a.in, a.g, a.resp, a.pix_out, spin, a.flags);
                break;
            default:
                hipLaunchKernelGGL((k_pixels<PROJ, PK, 3, true>), grid, block, 0, a.stream,
                                   a.in, a.g, a.resp, a.pix_out, spin, a.flags);
                break;
            }
        }
        constexpr int K = index_count<PK>();
        if constexpr (K > 1)
            hipLaunchKernelGGL((k_fill_stale<K>), dim3((unsigned)a.in.n_det), dim3(kFillBlock),
                               0, a.stream, a.pix_out, a.in.n_t, a.flags);
        return 0;
    }
    if constexpr (PK == PK_FLAT || PK == PK_FLAT_TILED) {
        if (a.interp) return launch_interp<PROJ, PK, 1>(a);
    }
    return launch_interp<PROJ, PK, 0>(a);
}

template <int PROJ>
int dispatch(const OpArgs &a) {
    if (a.op == OP_COORDS) {
        hipLaunchKernelGGL((k_coords<PROJ>), grid_for(a), dim3(kBlock), 0, a.stream, a.in,
                           a.coords_out);
        return 0;
    }
    switch (a.pk) {
    case PK_FLAT: return launch_pk<PROJ, PK_FLAT>(a);
    case PK_FLAT_TILED: return launch_pk<PROJ, PK_FLAT_TILED>(a);
    }
    if constexpr (PROJ == PROJ_CAR) {
        switch (a.pk) {
        case PK_HP: return launch_pk<PROJ, PK_HP>(a);
        case PK_HP_TILED: return launch_pk<PROJ, PK_HP_TILED>(a);
        }
    }
    return -1;
}

}  // namespace so3g
