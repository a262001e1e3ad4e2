/* so3g_hip.h -- C ABI of the MI355X projection engine (libso3g_hip.so).
 *
 * This is the drop-in boundary behind so3g_amd.ProjEng_* (the Python
 * mirror of so3g's ProjectionEngine classes).  Each entry point replaces
 * one method of the reference's boost-python engine; the method it
 * replaces is cited on each declaration.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (hipMalloc'd or torch-ROCm
 *     tensors), C-contiguous, unless stated otherwise.  Segment lists are
 *     HOST pointers.
 *   - bore:  (n_t, 4) float64 boresight quaternions (a,b,c,d) = (1,i,j,k)
 *   - dets:  (n_det, 4) float64 detector offset quaternions
 *   - resp:  (n_det, 2) float32 (T, P) responses
 *   - signal:(n_det, n_t) float32
 *   - det_weights: (n_det) float32 or NULL
 *   - maps: float64.  Non-tiled flat (lead..., ny, nx); non-tiled HEALPix
 *     (lead..., npix); lead = (ncomp) for maps, (ncomp, ncomp) for weight
 *     maps.  Tiled maps are given as a device array `tiles` of n_tile
 *     device pointers (NULL = inactive tile) plus a device array
 *     `tile_dims` of n_tile (th, tw) pairs (HEALPix tiles: (1, npix/tile)).
 *   - segs: host array of n_seg (det, t_start, t_end) int32 triples; every
 *     listed segment is processed (the flattened form of so3g's
 *     [bunch][thread][det] RangesInt32 lists, Projection.cxx:1856-1922).
 *     segs == NULL means every sample of every detector.
 *   - stream: a hipStream_t (NULL = default stream).  Calls are ordered
 *     on that stream; calls that can raise an in-kernel condition (inactive
 *     tile hit) synchronise the stream before returning.
 *   - Return value: SO3G_OK or an SO3G_ERR_* code; so3g_last_error() then
 *     holds a message (thread-local).
 */
#ifndef SO3G_HIP_H
#define SO3G_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Coordinate systems: the CoordSys template options of
 * ProjectionEngine<C,P,S> (Projection.cxx:152-159, 238-507). */
enum {
    SO3G_PROJ_FLAT = 0,
    SO3G_PROJ_QUAT = 1,
    SO3G_PROJ_CAR = 2,
    SO3G_PROJ_CEA = 3,
    SO3G_PROJ_ARC = 4,
    SO3G_PROJ_SIN = 5,
    SO3G_PROJ_TAN = 6,
    SO3G_PROJ_ZEA = 7
};

/* Pixelizors: Pixelizor2_Flat (Projection.cxx:802-1166) and
 * Pixelizor_Healpix (:510-800). */
enum { SO3G_PIX_FLAT = 0, SO3G_PIX_HEALPIX = 1 };

/* Status codes; the Python layer maps them onto the exception classes
 * that so3g's exceptions.cxx:10-32 translates to. */
enum {
    SO3G_OK = 0,
    SO3G_ERR_VALUE = 1,   /* ValueError  */
    SO3G_ERR_RUNTIME = 2, /* RuntimeError (shape / tiling problems) */
    SO3G_ERR_TYPE = 3,    /* TypeError */
    SO3G_ERR_HIP = 4      /* HIP runtime failure (RuntimeError) */
};

/* Sample-domain modes for so3g_sample_domains. */
enum {
    SO3G_DOMAIN_STRIPE = 0, /* pixelizor stripe() (pixel_ranges, map=None) */
    SO3G_DOMAIN_MAP = 1,    /* thread map value   (pixel_ranges, map given) */
    SO3G_DOMAIN_TILE = 2    /* tile -> domain table (tile_ranges) */
};

/* Precision modes for the so3g_to_map_prec / so3g_to_weights_prec study. */
enum {
    SO3G_PREC_REF = 0,  /* float32 products, float64 accumulation (so3g) */
    SO3G_PREC_F32 = 1,  /* float32 products, float32 accumulation (map is float32) */
    SO3G_PREC_F64 = 2   /* float64 products, float64 accumulation */
};

/* The template parameters of one ProjectionEngine<C,P,S> plus its
 * pixelizor constructor arguments (Projection.cxx:821-829, 971-994,
 * 534-542, 640-675). */
typedef struct so3g_geom {
    int32_t proj;          /* SO3G_PROJ_* (HEALPix engines use CAR) */
    int32_t pixelizor;     /* SO3G_PIX_* */
    int32_t tiled;         /* 0 / 1 */
    int32_t interp;        /* 0 nearest, 1 bilinear (flat only) */
    int32_t ncomp;         /* 1 (T), 2 (QU), 3 (TQU) */
    int32_t naxis[2];      /* flat: ny, nx */
    double cdelt[2];       /* flat: dy, dx (radians) */
    double crpix[2];       /* flat: iy0, ix0 */
    int32_t tile_shape[2]; /* flat tiled: ty, tx */
    int32_t nside;         /* HEALPix */
    int32_t nest;          /* HEALPix: 1 NEST, 0 RING */
    int32_t nside_tile;    /* HEALPix tiled */
    int32_t allow_nside_max; /* opt-in: accept nside == 8192 */
    /* Derived by so3g_geom_init: */
    int32_t n_tile;        /* tiles (0 if not tiled) */
    int32_t index_count;   /* P::index_count */
    int64_t npix;          /* pixels per map plane (non-tiled) */
} so3g_geom;

/* Library set-up: selects the device, builds the 16384-entry asin/atan2
 * interpolation tables on the host with the C library (as
 * so_linterp.h:35-56 does at static init) and uploads them.  Idempotent. */
int so3g_init(int device);
const char *so3g_last_error(void);
int so3g_abi_version(void);
/* Copies the trig tables (2 x 16384 float64) for verification: the device
 * copies the kernels read once the library is initialised, else tables
 * freshly built on the host. */
int so3g_get_tables(double *asin_vals, double *atan2_vals);

/* Validates the geometry (check_nside, Projection.cxx:607-614, 785-792;
 * nside_tile <= nside, :670-674; RING + tiles, :646-650) and fills the
 * derived fields.  Replaces the engine constructors. */
int so3g_geom_init(so3g_geom *g);

/* ProjectionEngine::coords (Projection.cxx:1316-1349).
 * coords: (n_det, n_t, 4) float64. */
int so3g_coords(const so3g_geom *g, const double *bore, int64_t n_t,
                const double *dets, int64_t n_det, double *coords, void *stream);

/* ProjectionEngine::pixels (Projection.cxx:1351-1385).
 * pixel: (n_det, n_t, index_count) int32.  Off-map samples get -1 in slot
 * 0 and keep the previous sample's other slots (initially 0). */
int so3g_pixels(const so3g_geom *g, const double *bore, int64_t n_t,
                const double *dets, int64_t n_det, int32_t *pixel, void *stream);

/* ProjectionEngine::pointing_matrix (Projection.cxx:1391-1438).
 * pixel as above; spin_proj: (n_det, n_t, ncomp) float32. */
int so3g_pointing_matrix(const so3g_geom *g, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         int32_t *pixel, float *spin_proj, void *stream);

/* ProjectionEngine::from_map (Projection.cxx:1728-1770).  Accumulates into
 * signal; with accumulate == 0 the signal is overwritten instead (the
 * result equals accumulating into zeros). */
int so3g_from_map(const so3g_geom *g, const double *map, const double *const *tiles,
                  const int32_t *tile_dims, const double *bore, int64_t n_t,
                  const double *dets, int64_t n_det, const float *resp,
                  float *signal, int accumulate, void *stream);

/* ProjectionEngine::to_map (Projection.cxx:1924-1966, 1772-1811). */
int so3g_to_map(const so3g_geom *g, double *map, double *const *tiles,
                const int32_t *tile_dims, const double *bore, int64_t n_t,
                const double *dets, int64_t n_det, const float *resp,
                const float *signal, const float *det_weights,
                const int32_t *segs, int64_t n_seg, void *stream);

/* ProjectionEngine::to_weight_map (Projection.cxx:1968-2010, 1813-1854).
 * Only the upper triangle (j >= i) of the (ncomp, ncomp) block is filled. */
int so3g_to_weights(const so3g_geom *g, double *map, double *const *tiles,
                    const int32_t *tile_dims, const double *bore, int64_t n_t,
                    const double *dets, int64_t n_det, const float *resp,
                    const float *det_weights, const int32_t *segs, int64_t n_seg,
                    void *stream);

/* Precision-study variants of to_map / to_weights (no counterpart in the
 * reference, whose arithmetic is SO3G_PREC_REF).  For SO3G_PREC_F32 the
 * map (and tiles) are float32 arrays passed through the same pointers. */
int so3g_to_map_prec(const so3g_geom *g, void *map, void *const *tiles,
                     const int32_t *tile_dims, const double *bore, int64_t n_t,
                     const double *dets, int64_t n_det, const float *resp,
                     const float *signal, const float *det_weights, int prec,
                     void *stream);
int so3g_to_weights_prec(const so3g_geom *g, void *map, void *const *tiles,
                         const int32_t *tile_dims, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         const float *det_weights, int prec, void *stream);

/* ProjectionEngine::tile_hits (Projection.cxx:1543-1601).  hits: n_tile
 * int32, overwritten. */
int so3g_tile_hits(const so3g_geom *g, const double *bore, int64_t n_t,
                   const double *dets, int64_t n_det, int32_t *hits, void *stream);

/* Per-sample domain for pixel_ranges / tile_ranges (Projection.cxx:
 * 1440-1541, 1606-1705): out (n_det, n_t) int32 = the common domain of the
 * sample's points, n_domain if they disagree, -1 if the sample has none.
 * The run-length grouping into ranges is done by the caller. */
int so3g_sample_domains(const so3g_geom *g, const double *bore, int64_t n_t,
                        const double *dets, int64_t n_det, int mode, int n_domain,
                        const double *tmap, const double *const *ttiles,
                        const int32_t *tile_dims, const int32_t *tile_dom,
                        int32_t *out, void *stream);

/* ProjEng_Precomp (Projection.cxx:2012-2253): projection with a cached
 * pointing matrix.  pixel: (n_det, n_t, k) int32 (k = 2 non-tiled flat,
 * 3 tiled flat, 1 for a 1-D (ncomp, npix) map); spin: (n_det, n_t, nc).
 * map_shape (host int64 pair): (ny, nx) for k = 2, (npix, 1) for k = 1,
 * (n_tile, 1) for k = 3 (tile_dims then holds n_tile (rows, cols) pairs).
 * Samples with pixel slot 0 < 0 are off the map.  An index outside the map
 * (the reference does not check) skips its sample and makes the call
 * return SO3G_ERR_RUNTIME; every call synchronises its stream to report it. */
int so3g_precomp_from_map(const double *map, const int64_t *map_shape,
                          const double *const *tiles, const int32_t *tile_dims,
                          const int32_t *pixel, int32_t k, const float *spin,
                          int32_t nc, int64_t n_det, int64_t n_t, float *signal,
                          int accumulate, void *stream);
/* so3g_precomp_from_map for a non-tiled (ny, nx) map (k = 2) that allocates
 * nothing: no counterpart in the reference; for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  The pixel-major packed copy of the map lives in
 * a fixed 32 MiB device array per GPU (kept until the process ends), so the
 * map may hold at most 1048576 pixels (SO3G_ERR_VALUE beyond that).  The
 * library must be initialised (so3g_init) first.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several streams or host
 * threads) race and must be serialised by the caller.  Same results,
 * checks and errors as so3g_precomp_from_map, bit for bit. */
int so3g_precomp_from_map_noalloc(const double *map, const int64_t *map_shape,
                                  const int32_t *pixel, const float *spin, int32_t nc,
                                  int64_t n_det, int64_t n_t, float *signal, int accumulate,
                                  void *stream);
int so3g_precomp_to_map(double *map, const int64_t *map_shape, double *const *tiles,
                        const int32_t *tile_dims, const int32_t *pixel, int32_t k,
                        const float *spin, int32_t nc, int64_t n_det, int64_t n_t,
                        const float *signal, const float *det_weights,
                        const int32_t *segs, int64_t n_seg, void *stream);
int so3g_precomp_to_weights(double *map, const int64_t *map_shape,
                            double *const *tiles, const int32_t *tile_dims,
                            const int32_t *pixel, int32_t k, const float *spin,
                            int32_t nc, int64_t n_det, int64_t n_t,
                            const float *det_weights, const int32_t *segs,
                            int64_t n_seg, void *stream);

/* Projection plan for repeated cached-pointing projections with one
 * pointing matrix (no counterpart in the reference; for CG-style loops).
 * Built once from a Precomp pixel array and spin array: k = 2 with a
 * non-tiled flat map of map_shape = (ny, nx), or k = 1 (HEALPix pixel
 * numbers) with map_shape = (npix, 1); host array.  The arrays are copied,
 * so they may be freed afterwards.  The pixels the samples visit are numbered
 * in ascending order (compact pixels), so the plan's pixel tables scale with
 * the visited pixels, not the map.  Samples are grouped in time blocks whose
 * signal slice is about t_block_mb MiB (0: one block; < 0: the layout's
 * default), then sorted within each block: by map tile (32 x 64 pixels, or
 * 2048 consecutive HEALPix pixel numbers), keeping sample order, in the tiled layout
 * (option "plan_layout" 1, the default from 2^25 samples up; one time
 * block); by pixel in layout 0
 * (64 MiB blocks, sized to stay in the Infinity Cache).  Layout 2 (the
 * default below 2^25 samples when it applies, else layout 0) keeps the
 * samples in time order, in work items of one detector's consecutive
 * samples (up to "plan_win_len") whose compact pixels span at most 2048
 * values: from_map gathers from the item's window of the map held in LDS,
 * to_map sums into it and a second pass adds the items' sums per pixel (no
 * atomics, a fixed summation order); it is refused (layout 2 asked for) or
 * skipped (automatic choice) when some 512 consecutive samples of a detector
 * visit pixels that span more than 2048 compact values.  Stored: per on-map
 * sample the spin factors (the T factor once per detector when it is
 * constant over the detector, as in so3g's pointing matrices) and the tile
 * slot (tiled) or the sample index and run-start bits (layout 0); per
 * sample a time-ordered compact pixel and the same spin factors (unless
 * "plan_time_copy" is 0 in the tiled layout).  so3g_pc_plan_to_map adds the same
 * contributions as so3g_precomp_to_map over every sample (float64 sums in
 * another order); so3g_pc_plan_from_map writes what so3g_precomp_from_map
 * writes, bit for bit. */
typedef struct so3g_pc_plan so3g_pc_plan;
int so3g_pc_plan_from_map(const so3g_pc_plan *plan, const double *map, float *signal,
                          int accumulate, void *stream);
int so3g_pc_plan_create(const int32_t *pixel, int32_t k, const float *spin, int32_t nc,
                        int64_t n_det, int64_t n_t, const int64_t *map_shape,
                        int64_t t_block_mb, so3g_pc_plan **out, void *stream);
int so3g_pc_plan_to_map(const so3g_pc_plan *plan, double *map, const float *signal,
                        const float *det_weights, void *stream);
/* map = P^T signal: so3g_pc_plan_to_map into a map that is first set to 0,
 * as ProjEng_Precomp.to_map does when called with map=None
 * (Projection.cxx:2094-2099 allocates a zero map); a layout-2 plan writes
 * every pixel in its second pass, the other layouts zero the map first. */
int so3g_pc_plan_to_map_set(const so3g_pc_plan *plan, double *map, const float *signal,
                            const float *det_weights, void *stream);
/* Float64-signal variants for map-making solvers (no counterpart in the
 * reference, whose signals are float32): products and sums in float64, so
 * that P^T N^-1 P stays symmetric to float64 rounding when N^-1 has a large
 * dynamic range (1/f noise).  so3g_precomp_from_map_f64 takes a non-tiled
 * (ny, nx) map and a k = 2 pixel array; samples off the map add 0. */
int so3g_pc_plan_to_map_f64(const so3g_pc_plan *plan, double *map, const double *signal,
                            const float *det_weights, void *stream);
int so3g_precomp_from_map_f64(const double *map, const int64_t *map_shape,
                              const int32_t *pixel, const float *spin, int32_t nc,
                              int64_t n_det, int64_t n_t, double *signal, int accumulate,
                              void *stream);
/* Plans for tiled maps (ProjEng_Precomp<Tiled>, Projection.cxx:2012-2253:
 * pixel (n_det, n_t, 3) = (tile, row, column within the tile)).  tile_dims:
 * host (n_tile, 2) int32 tile shapes (rows, cols) of the whole tiling,
 * inactive tiles included; always the tiled layout.  The _tiles calls take a
 * HOST array of n_tile device tile pointers ((nc, rows, cols) float64 each;
 * NULL for an inactive tile) and fail with a RuntimeError when a tile that
 * the plan's samples hit is NULL, as the reference does for an inactive tile;
 * the pointer array may be reused once the call returns.  The non-tiled
 * calls refuse a tiled plan and the other way round. */
int so3g_pc_plan_create_tiled(const int32_t *pixel, const float *spin, int32_t nc,
                              int64_t n_det, int64_t n_t, const int32_t *tile_dims,
                              int32_t n_tile, int64_t t_block_mb, so3g_pc_plan **out,
                              void *stream);
int so3g_pc_plan_to_map_tiles(const so3g_pc_plan *plan, double *const *tiles,
                              const float *signal, const float *det_weights, void *stream);
int so3g_pc_plan_from_map_tiles(const so3g_pc_plan *plan, const double *const *tiles,
                                float *signal, int accumulate, void *stream);
/* Precomp to_weight_map (Projection.cxx:2124-2154) through a tiled-layout
 * plan: map (nc, nc, ...) += sum of (pf_a * pf_b) * det_wt (float32) for
 * b >= a, the upper triangle as the reference fills it; summed per work item
 * in LDS (float64, another order than the reference's sequential sums).
 * Every layout (pixel runs: per-run sums and atomics as to_map). */
int so3g_pc_plan_to_weights(const so3g_pc_plan *plan, double *map, const float *det_weights,
                            void *stream);
int so3g_pc_plan_to_weights_tiles(const so3g_pc_plan *plan, double *const *tiles,
                                  const float *det_weights, void *stream);
int so3g_pc_plan_info(const so3g_pc_plan *plan, int64_t *n_on, int64_t *bytes);
/* stats[13]: n_on, pixel runs (layout 0), T factor constant per detector
 * (0/1), stored spin factors per sample, samples per time block, device
 * bytes, visited pixels, layout, stretch starts (counting each 64-position
 * chunk's first position; tiled layout), work items (tiled and windowed
 * layouts), the 512-sample groups of the time-ordered copy whose compact
 * pixels span more than 65535 values (stored as int32 instead of u16
 * offsets), and for the windowed layout its largest window and the windows'
 * total size (partial sums per pass) */
int so3g_pc_plan_stats(const so3g_pc_plan *plan, int64_t *stats);
int so3g_pc_plan_destroy(so3g_pc_plan *plan);

/* Batched projection of many observations into one map (no counterpart in
 * the reference, which projects one observation per call,
 * python/proj/wcs.py:415-476 looped by the caller).  An observation list is
 * turned into one table of equal work items (256 samples x dpb detectors);
 * so3g_batch_to_map / so3g_batch_from_map then cover the whole batch in a
 * single persistent launch.  mode SO3G_BATCH_EAGER launches directly,
 * SO3G_BATCH_GRAPH captures the launch in a HIP graph (re-captured when the
 * map pointers, shape or op change) and replays it, SO3G_BATCH_PER_OBS
 * launches the same kernel once per observation (for comparison).  Results
 * equal so3g_to_map / so3g_from_map called per observation (to_map up to
 * the order of the float64 additions).  The observation arrays must stay
 * allocated while the batch is used.  to_map can sum each work item in an
 * LDS-privatised partial map (so3g_batch_set_table: 0 global atomics, 512
 * or 1024 slots; the detectors of an item taken in map-row order, computed
 * at creation), or -w: segments of w columns (64 <= w <= the widest whose
 * float64 sums fit 160 KB of LDS) of each group's home map row in LDS, with
 * items of 65536 samples x 2 row-mate detectors of one observation, one
 * block of 1024 lanes each (flat maps with nearest-neighbour pixels; the
 * groups are computed at creation).  so3g_batch_tune times every launch
 * shape (dpb in {1,2,4,8,16} x blocks per CU in {2,4,8,16, one per item},
 * and for to_map x table slots in {0, 512, 1024}, then row segments of
 * 1024, 2048, 4096 and the widest number of columns) and keeps the fastest
 * for the op (0 from_map, 1 to_map); results holds (dpb, blocks per CU,
 * slots, ms) quadruples (row segments: (2, 0, -w, ms)).  Tuning to_map adds
 * into `map`: pass a scratch map. */
typedef struct so3g_batch so3g_batch;
typedef struct {
    const double *bore;        /* (n_t, 4) native-frame boresight */
    int64_t n_t;
    const double *dets;        /* (n_det, 4) */
    int64_t n_det;
    const float *resp;         /* (n_det, 2) */
    const float *signal;       /* (n_det, n_t) to_map input, or NULL */
    float *signal_out;         /* (n_det, n_t) from_map output, or NULL */
    const float *det_weights;  /* (n_det) or NULL */
} so3g_obs;
enum { SO3G_BATCH_EAGER = 0, SO3G_BATCH_GRAPH = 1, SO3G_BATCH_PER_OBS = 2 };
int so3g_batch_create(const so3g_geom *g, const so3g_obs *obs, int64_t n_obs,
                      so3g_batch **out);
int so3g_batch_to_map(so3g_batch *b, double *map, double *const *tiles,
                      const int32_t *tile_dims, int mode, void *stream);
int so3g_batch_from_map(so3g_batch *b, const double *map, const double *const *tiles,
                        const int32_t *tile_dims, int accumulate, int mode, void *stream);
int so3g_batch_set_shape(so3g_batch *b, int op, int dpb, int grid_mult);
int so3g_batch_set_table(so3g_batch *b, int slots);
int so3g_batch_tune(so3g_batch *b, int op, double *map, double *const *tiles,
                    const int32_t *tile_dims, int reps, double *results, int max_results,
                    int *n_results, void *stream);
/* info: n_obs, det*samples, from (dpb, blocks/CU), to (dpb, blocks/CU), CUs,
 * to_map table slots */
int so3g_batch_info(const so3g_batch *b, int64_t *info);
int so3g_batch_destroy(so3g_batch *b);

/* nmat_detvecs_apply (src/array_ops.cxx:39-88): in place on ft (ndet,
 * nmode) float32 (device), the real view of a Fourier transform unless
 * dct_binning; bins (nbin, 2) int32 on the HOST; iD (nbin, ndet) and iV
 * (nbin, ndet, nvec) float32 on the device, nvec <= 64.  For each bin in
 * order, columns [min(k b0, nmode - 1), min(k b1, nmode)) (k = 2, or 1 with
 * dct_binning) become ft * (iD / norm) + (s / norm) iV (iV^T ft), float32. */
int so3g_nmat_detvecs_apply(float *ft, int64_t ndet, int64_t nmode, const int32_t *bins,
                            int64_t nbin, const float *iD, const float *iV, int64_t nvec,
                            float s, float norm, int dct_binning, void *stream);

/* ML cut handling: process_cuts (src/array_ops.cxx:130-183) and
 * translate_cuts (:185-217).  A RangesMatrix is passed as a HOST segment
 * list segs[n_seg][3] = (det, start, end) in its order (detector by
 * detector, ranges ascending); tod (ndet, nsamp) and vals are device arrays
 * of float32 (f64 = 0) or float64 (f64 = 1).  op: 0 insert (vals -> tod),
 * 1 extract (tod -> vals), 2 clear; model: 0 "full" (one value per cut
 * sample), 1 "poly" (Legendre coefficients per range, count from
 * resolution and nmax, array_ops.cxx:233-240).  so3g_cuts_measure gives the
 * value count ("measure").  Poly extract adds into vals for one-coefficient
 * ranges, as the reference does.  translate maps values between sample
 * rates (full: nearest / average; poly: copy, zero-pad, truncate). */
int so3g_cuts_measure(const int32_t *segs, int64_t n_seg, int model, int resolution, int nmax,
                      int64_t *n_vals);
int so3g_process_cuts(const int32_t *segs, int64_t n_seg, int op, int model, int resolution,
                      int nmax, void *tod, int64_t ndet, int64_t nsamp, void *vals,
                      int64_t n_vals, int f64, void *stream);
int so3g_translate_cuts(const int32_t *isegs, const int32_t *osegs, int64_t n_seg, int model,
                        int resolution, int nmax, int64_t insamp, int64_t onsamp,
                        const void *ivals, int64_t n_ivals, void *ovals, int64_t n_ovals,
                        int f64, void *stream);

/* Boresight quaternions on the device (no counterpart in the reference's
 * C++; CelestialSightLine.for_horizon, python/proj/coords.py:190-209, for
 * mode 0 and naive_az_el, :99-125, for mode 1).  az, el, t and roll are
 * device arrays of n float64 (roll may be NULL: roll0 for every sample);
 * site (host, mode 1): {era slope, era offset, era epoch (unix s), site lon
 * (rad), site lat (rad)}; pre (host, 4 doubles, may be NULL) premultiplies
 * every quaternion; out (device) is (n, 4). */
int so3g_boresight(int mode, const double *t, const double *az, const double *el,
                   const double *roll, double roll0, int64_t n, const double *site,
                   const double *pre, double *out, void *stream);

/* G3SuperTimestream array blobs (src/G3SuperTimestream.cxx).
 *
 * so3g_g3s_split replaces the quantise-and-rebranch half of encode_array
 * (G3SuperTimestream.cxx:72-92 rebranch, :218-231): per channel, float data
 * are divided by the quantum and rounded (roundf of the float-converted
 * quotient for float32, round for float64), then each value v is split as
 * v = r + y, r a multiple of 2^24 changing only when v leaves its multiple,
 * y in [-2^23, 2^23).  type: 0 int32, 1 int64, 2 float32, 3 float64
 * (host arrays, (n_det, n)); y int32, r int64 (n_det, n); changes[d] = 1
 * when r is constant over the channel.
 *
 * so3g_flac_encode replaces the libFLAC stream encoder call of encode_array
 * (:236-260: mono, bps bits per sample, compression level 0..8): writes a
 * FLAC stream of the n host samples x into out (capacity bytes; out NULL
 * or capacity < 0 to size only) and its length into n_bytes.  Returns
 * SO3G_ERR_VALUE if the stream would exceed capacity, as libFLAC's write
 * callback refuses (:99-104).
 *
 * so3g_g3s_decode replaces Extract (:664-812) and Decode (:636-662): the
 * host has walked the blob's per-channel blocks into chans (one entry per
 * extracted channel); payload (device, 16-byte aligned) holds the blob;
 * offsets (device, n_offset_rows x n_samp, full-width integers of the
 * array type) the inflated bz2 blocks; dest (device) dest_rows rows of
 * dest_stride elements receives samples [start, stop) of each channel in
 * row dest_row.  Channels with neither FLAC nor raw data add to dest's
 * contents, as Extract does.  Synchronises the stream. */
typedef struct {
    int64_t flac_pos, flac_end;  /* first FLAC frame and end of the FLAC block; -1: none */
    int64_t raw_pos;             /* ALGO_NONE: payload offset of the raw data; -1: none */
    int64_t off_row;             /* row of offsets added to the channel; -1: none */
    int64_t const_bits;          /* CONST datum (full-width two's complement) */
    double quantum;              /* float types: the channel's quantum */
    int64_t dest_row;
    int32_t block;               /* FLAC: STREAMINFO maximum block size */
    int32_t bps;                 /* FLAC: STREAMINFO bits per sample */
    int32_t has_const;
    int32_t pad_;
} so3g_g3s_chan;
int so3g_g3s_split(const void *data, int32_t type, int64_t n_det, int64_t n,
                   const double *quanta, int32_t *y, int64_t *r, int32_t *changes);
int so3g_flac_encode(const int32_t *x, int64_t n, int32_t bps, int32_t level, uint8_t *out,
                     int64_t capacity, int64_t *n_bytes);
int so3g_g3s_decode(const uint8_t *payload, int64_t payload_bytes, const so3g_g3s_chan *chans,
                    int64_t n_chan, int32_t type, int64_t n_samp, int64_t start, int64_t stop,
                    const void *offsets, int64_t n_offset_rows, void *dest, int64_t dest_rows,
                    int64_t dest_stride, void *stream);
/* Diagnostics of the last so3g_g3s_decode call (no counterpart in the
 * reference): out[0] = FLAC frames decoded frame-parallel, out[1] = channels
 * decoded again frame after frame, out[2] = why, OR of 1 variable block
 * size, 2 a frame without sync, 4 a malformed frame, 8 a frame not ending
 * where the next begins (false sync), 16 an unexpected frame number or
 * block size. */
int so3g_g3s_stats(int64_t *out);

/* Layers of the learned map-making preconditioner (no counterpart in the
 * reference).  Activations are bf16 NHWC (n, h, w, c) with c in {16, 32, 48,
 * 64}; conv weights bf16 (9, co, ci) (tap = 3 * (dy + 1) + (dx + 1)), bias
 * float32 (co).  so3g_nn_conv3x3: out = act(bias + conv), zero padding,
 * act = ReLU if relu, float32 accumulation on the matrix cores.
 * so3g_nn_conv3x3_wgrad ADDS dW (float32, (9, co, ci)) and db (float32, co;
 * may be NULL) of sum(dy * conv(in)).  so3g_nn_weight_flip makes the (9, ci,
 * co) weights whose conv is the data gradient.  so3g_nn_relu_mask: dz = dy *
 * (act > 0).  pool/up: (fy, fx) average pooling and nearest upsampling of an
 * (n, h, w, c) input (h, w: the input's for the forward, the forward
 * input's for backward = 1).  from/to_planes convert float32 (n, cp, hy, wx)
 * planes to/from the zero-padded bf16 layout.  so3g_nn_adam: one Adam step
 * (step counts from 1) on float32 weights, optionally refreshing a bf16
 * copy. */
int so3g_nn_conv3x3(const void *in, int n, int h, int w, int ci, const void *wt,
                    const float *bias, int co, int relu, void *out, void *stream);
int so3g_nn_conv3x3_wgrad(const void *in, const void *dy, int n, int h, int w, int ci, int co,
                          float *dw, float *db, void *stream);
int so3g_nn_weight_flip(const void *wt, int ci, int co, void *wt_t, void *stream);
int so3g_nn_relu_mask(const void *dy, const void *act, int64_t n, void *dz, void *stream);
int so3g_nn_add(const void *a, const void *b, int64_t n, void *out, void *stream);
int so3g_nn_pool(const void *in, int n, int h, int w, int c, int fy, int fx, void *out,
                 int backward, void *stream);
int so3g_nn_up(const void *in, int n, int h, int w, int c, int fy, int fx, void *out,
               int backward, void *stream);
int so3g_nn_from_planes(const float *x, int n, int cp, int hy, int wx, int h, int w, int c,
                        void *out, void *stream);
int so3g_nn_to_planes(const void *in, int n, int h, int w, int c, int cp, int hy, int wx,
                      float *x, void *stream);
int so3g_nn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1,
                 float b2, float eps, int step, void *w_bf16, void *stream);
int so3g_nn_to_bf16(const float *x, int64_t n, void *out, void *stream);

/* Implementation tunables (no counterpart in the reference).  They are
 * process-wide values read when a call starts (the plan options when a plan
 * is created): set them before, not during, calls from other threads.
 * so3g_get_option also reads the build constants "plan_tile_rows" and
 * "plan_tile_cols".
 *   "scatter_agg"   1: sum runs of equal pixels within a wave before the
 *                   map atomics (default 1)
 *   "scatter_xcd"   1: XCD-private partial maps for non-tiled scatters,
 *                   reduced after the kernel (default 0)
 *   "xcd_budget_mb" device memory allowed for those copies (default 4096)
 *   "scatter_hash"  on-the-fly to_map / to_weights through LDS-privatised
 *                   partial maps: 0 never, 1 (default) where measured
 *                   faster than global atomics, 2 always with small tables
 *                   (256 threads, 512 slots), 3 always with large ones
 *                   (1024 threads, 112 KB), 4 the map row of a group of
 *                   detectors held in LDS where the map allows it
 *                   (flat, tiled or not, nearest neighbour; the whole
 *                   row where its sums fit 160 KB, else a segment of the
 *                   widest width that fits, centred on the group's home
 *                   column; other maps as 1).  Mode 1 uses the rows where
 *                   the map allows them.  Precomp to_map /
 *                   to_weights use the small tables in modes 1, 2 and 4
 *                   (tiled maps: to_weights only in modes 1 and 4)
 *   "hash_r"        samples per lane of the hash tables in modes 2 and 3
 *                   (1 default: runs summed across the wave; 4: four
 *                   consecutive samples per lane, runs summed in the lane;
 *                   mode 1 takes 4 for HEALPix to_map and to_weights)
 *   "hash_dets", "hash_span"  detectors and samples per block of modes 2
 *                   to 4 (0, the default: 4 x 256, 2 x up to 131072, and for map
 *                   rows 2 x a span giving about 8 blocks per compute unit,
 *                   4096 to 131072 samples);
 *                   also the detectors per block of the Precomp tables (0:
 *                   measured per case, 2 to 8)
 *   "pc_from_dpb"   detectors per block of the Precomp gather (default 2)
 *   "plan_div"      plan kernels' detector index by float reciprocal (1,
 *                   default) or integer division (0)
 *   "plan_from_v"   samples per lane of the planned gather (4 default, 8)
 *   "plan_to_lds"   1 (default): the planned scatter stages the per-detector
 *                   tables (T factor, weights) in LDS when they fit in 32 KB
 *   "plan_to_cpb"   2048-sample chunks per block of the planned scatter
 *                   (default 2, 1..64)
 *   "nmat_dchunk"   detectors per block of nmat_detvecs_apply's passes
 *                   (default 64, 1..65536)
 *   "plan_layout"   layout of plans created afterwards: 1 map tiles summed
 *                   in LDS, 0 pixel runs, 2 time-ordered windows, -1
 *                   (default) tiles from 2^25 samples up, below that windows
 *                   where they fit, else pixel runs
 *   "plan_win_len"  samples per work item of the windowed layout's to_map
 *                   at most (default 65536, 512..2^30): a detector's samples
 *                   go to ceil(n_t / plan_win_len) items of about equal
 *                   length (more where the plan would have fewer than 64
 *                   items); from_map takes them in pieces of 8192
 *   "plan_win_block"  threads per work item of the windowed layout (1024
 *                   default, or 256)
 *   "plan_win_flush"  windowed to_map: 1 (default) float64 atomics from
 *                   each window into the map; 0 window sums to a partial
 *                   buffer and a second pass per pixel (no atomics, a fixed
 *                   summation order, slower)
 *   "plan_tile_item"  sorted positions per work item of the tiled layout
 *                   (default 262144, 64..2^30)
 *   "plan_tile_rows", "plan_tile_cols"  (read only) the tiled plan's tile
 *                   shape, a build constant (32 x 64)
 *   "plan_time_copy"  1 (default): tiled plans also keep the time-ordered
 *                   copy for from_map; 0: from_map gathers from the tiles
 *                   (smaller plan, slower gather)
 *   "plan_from_layout"  1 (default): a lane's gather samples are 64 apart, so
 *                   one gather instruction covers 64 consecutive samples;
 *                   0: consecutive samples per lane (measured slower, 3.51 vs
 *                   3.05 ms for pack + gather) */
int so3g_set_option(const char *name, int64_t value);
int so3g_get_option(const char *name, int64_t *value);

#ifdef __cplusplus
}
#endif

#endif /* SO3G_HIP_H */
