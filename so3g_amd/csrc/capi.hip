// capi.hip -- the extern "C" boundary of libso3g_hip.so (declared in
// include/so3g_hip.h): argument validation with the reference's error
// classes, engine-geometry set-up, trig-table construction, segment
// chunking, kernel dispatch and the device error flags.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/so3g_hip.h"
#include "launch.h"

namespace so3g {
extern int g_plan_div;
extern int g_plan_from_v;
extern int g_plan_from_layout;
extern int g_plan_to_lds;
extern int g_plan_to_cpb;
extern int g_plan_layout;
extern int64_t g_plan_tile_item;
extern int64_t g_plan_win_len;
extern int g_plan_win_block;
extern int g_plan_win_flush;
extern int g_plan_time_copy;
extern int g_nmat_dchunk;
extern const int g_plan_tile_rows, g_plan_tile_cols;
extern template int dispatch<PROJ_FLAT>(const OpArgs &);
extern template int dispatch<PROJ_QUAT>(const OpArgs &);
extern template int dispatch<PROJ_CAR>(const OpArgs &);
extern template int dispatch<PROJ_CEA>(const OpArgs &);
extern template int dispatch<PROJ_ARC>(const OpArgs &);
extern template int dispatch<PROJ_SIN>(const OpArgs &);
extern template int dispatch<PROJ_TAN>(const OpArgs &);
extern template int dispatch<PROJ_ZEA>(const OpArgs &);
}  // namespace so3g

using namespace so3g;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess)                                                  \
            return fail(SO3G_ERR_HIP, std::string("HIP error in ") + #expr +   \
                                          ": " + hipGetErrorString(_e));       \
    } while (0)

struct State {
    int device = -1;
    int n_cu = 256;            // compute units of the device
    std::vector<double> asin_h, atan2_h;
    double *asin_d = nullptr, *atan2_d = nullptr;
    int *flags_d = nullptr;
    int *flags_h = nullptr;     // pinned
    Chunk *chunks_d = nullptr;
    size_t chunk_cap = 0;
    double *copies_d = nullptr;   // XCD-private partial maps
    size_t copies_cap = 0;        // in doubles
    double *packed_d = nullptr;   // pixel-major map copy for gathers
    size_t packed_cap = 0;
    double *scratch_d = nullptr;  // device address of `scratch` (so3g_init)
    // Detector order of the LDS-privatised scatter, cached for the last
    // (boresight, detectors, geometry) it was computed for.  It only steers
    // which detectors share a block, so a stale entry costs speed, never
    // correctness.
    int32_t *order_d = nullptr;
    double *order_co_d = nullptr;
    int64_t order_cap = 0;
    const void *order_bore = nullptr, *order_dets = nullptr;
    int64_t order_nt = -1, order_ndet = -1;
    so3g_geom order_geom{};
    // The same for the Precomp LDS scatter, keyed on the pixel array: the
    // order (n_det entries) then the middle samples' indices (k n_det, k <= 3).
    int32_t *pc_order_d = nullptr;
    int64_t pc_order_cap = 0;
    const int32_t *pc_order_pix = nullptr;
    int64_t pc_order_nt = -1, pc_order_ndet = -1;
    int32_t pc_order_k = 0;
    // Detector groups of the row scatter (home row, up to G detectors
    // each), cached like the order above.
    int32_t *rg_det_d = nullptr, *rg_row_d = nullptr, *rg_xc_d = nullptr;
    int64_t rg_cap = 0;
    const void *rg_bore = nullptr, *rg_dets = nullptr;
    int64_t rg_nt = -1, rg_ndet = -1;
    int32_t rg_G = 0, rg_n = 0;
    so3g_geom rg_geom{};
};

// Tunables (so3g_set_option): wave-level run aggregation of scatter-adds,
// XCD-private partial maps, and the memory budget for the latter.
struct Options {
    int scatter_agg = 1;
    int scatter_xcd = 0;
    int64_t xcd_budget_mb = 4096;
    int pc_from_dpb = 2;       // detectors per block of the packed-map from_map
    // On-the-fly to_map / to_weights through LDS-privatised partial maps:
    // 0 never, 1 where measured faster than global atomics (all but tiled
    // flat to_map, the table per case), 2 always with the small table
    // (256 threads, 512 slots), 3 always with the large one (1024 threads,
    // 112 KB), 4 map rows in LDS where the geometry allows it (non-tiled
    // flat maps, nearest neighbour; else as 1).  Mode 1 takes the rows
    // where the geometry allows them.
    int scatter_hash = 1;
    int hash_dets = 0;         // detectors per block in modes 2-4 (0: 4 / 2 / 2)
    int hash_span = 0;         // samples per block in modes 2-4 (0: 256 / 131072 / 65536)
    int hash_r = 1;            // hash tables: samples per lane (1: wave-level runs, 4)
};
Options g_opt;
constexpr int kNumXcd = 8;

// Scratch of so3g_precomp_from_map_noalloc: the pixel-major packed copy of
// the map (4 float64 per pixel), 32 MiB, held by every device the library
// runs on until the process ends.  Host code uses g_state.scratch_d (its
// device address, resolved at so3g_init), never the array's name.
constexpr int64_t kScratchPixels = 1 << 20;
static __device__ __align__(256) double scratch[4 * kScratchPixels];

State g_state;
std::mutex g_mutex;

// The 16384-entry tables of so_linterp.h:35-39 (asin(i*dx)) and :51-56
// (atan2(i*dx, 1.)), dx = 1./(n-1), evaluated with the host C library.
void build_tables(std::vector<double> &as, std::vector<double> &at) {
    const int n = kTrigN;
    as.resize(n);
    at.resize(n);
    const double dx = 1. / (n - 1);
    for (int i = 0; i < n; i++) as[i] = asin(i * dx);
    for (int i = 0; i < n; i++) at[i] = atan2(i * dx, 1.);
}

__global__ void k_zero_flags(int *flags) {
    if (threadIdx.x < FLAG_COUNT) flags[threadIdx.x] = 0;
}

// The k indices of each detector's middle sample: out[d * k + q].
__global__ void k_pc_mid(const int32_t *pix, int32_t k, int64_t n_det, int64_t n_t,
                         int32_t *out) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n_det) return;
    for (int q = 0; q < k; q++) out[d * k + q] = pix[(d * n_t + n_t / 2) * k + q];
}

int ensure_init() {
    if (g_state.device >= 0) return SO3G_OK;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    return so3g_init(dev);
}

int geom_to_launch(const so3g_geom *g, Geo &geo, int &pk) {
    if (g == nullptr) return fail(SO3G_ERR_VALUE, "geometry is NULL");
    if (g->index_count <= 0)
        return fail(SO3G_ERR_VALUE, "geometry was not initialised (call so3g_geom_init)");
    memset(&geo, 0, sizeof(geo));
    geo.ncomp = g->ncomp;
    geo.interp = g->interp;
    if (g->pixelizor == SO3G_PIX_FLAT) {
        geo.ny = g->naxis[0];
        geo.nx = g->naxis[1];
        geo.cdelt_y = g->cdelt[0];
        geo.cdelt_x = g->cdelt[1];
        geo.icdelt_y = 1. / g->cdelt[0];
        geo.icdelt_x = 1. / g->cdelt[1];
        geo.crpix_y = g->crpix[0];
        geo.crpix_x = g->crpix[1];
        geo.npix = (int64_t)geo.ny * geo.nx;
        if (g->tiled) {
            geo.ty = g->tile_shape[0];
            geo.tx = g->tile_shape[1];
            geo.n_tx = (geo.nx + geo.tx - 1) / geo.tx;
        }
        pk = g->tiled ? PK_FLAT_TILED : PK_FLAT;
    } else {
        geo.nside = g->nside;
        geo.nest = g->tiled ? 1 : g->nest;
        geo.npix = 12LL * g->nside * g->nside;
        if (g->tiled) geo.npt = (int32_t)(geo.npix / (12LL * g->nside_tile * g->nside_tile));
        pk = g->tiled ? PK_HP_TILED : PK_HP;
    }
    return SO3G_OK;
}

int run_dispatch(int proj, const OpArgs &a) {
    int r = -1;
    switch (proj) {
    case PROJ_FLAT: r = dispatch<PROJ_FLAT>(a); break;
    case PROJ_QUAT: r = dispatch<PROJ_QUAT>(a); break;
    case PROJ_CAR: r = dispatch<PROJ_CAR>(a); break;
    case PROJ_CEA: r = dispatch<PROJ_CEA>(a); break;
    case PROJ_ARC: r = dispatch<PROJ_ARC>(a); break;
    case PROJ_SIN: r = dispatch<PROJ_SIN>(a); break;
    case PROJ_TAN: r = dispatch<PROJ_TAN>(a); break;
    case PROJ_ZEA: r = dispatch<PROJ_ZEA>(a); break;
    }
    if (r != 0) return fail(SO3G_ERR_VALUE, "no kernel for this engine configuration");
    HIP_TRY(hipGetLastError());
    return SO3G_OK;
}

// Splits host segments (det, t0, t1) into chunks of at most kBlock samples
// and uploads them.  Returns the chunk count in *n_chunks.
int upload_chunks(const int32_t *segs, int64_t n_seg, int64_t n_det, int64_t n_t,
                  hipStream_t stream, int64_t *n_chunks) {
    std::vector<Chunk> ch;
    for (int64_t i = 0; i < n_seg; i++) {
        const int32_t d = segs[3 * i], t0 = segs[3 * i + 1], t1 = segs[3 * i + 2];
        if (d < 0 || d >= n_det)
            return fail(SO3G_ERR_RUNTIME, "segment detector index out of range");
        if (t0 < 0 || t1 > n_t || t0 > t1)
            return fail(SO3G_ERR_RUNTIME, "segment sample range out of bounds");
        for (int32_t s = t0; s < t1; s += kBlock)
            ch.push_back(Chunk{d, s, std::min<int32_t>(s + kBlock, t1)});
    }
    *n_chunks = (int64_t)ch.size();
    if (ch.empty()) return SO3G_OK;
    if (ch.size() > g_state.chunk_cap) {
        if (g_state.chunks_d) HIP_TRY(hipFree(g_state.chunks_d));
        g_state.chunks_d = nullptr;
        size_t cap = std::max<size_t>(ch.size(), 1 << 16);
        HIP_TRY(hipMalloc(&g_state.chunks_d, cap * sizeof(Chunk)));
        g_state.chunk_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(g_state.chunks_d, ch.data(), ch.size() * sizeof(Chunk),
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return SO3G_OK;
}

int zero_flags(hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, stream, g_state.flags_d);
    HIP_TRY(hipGetLastError());
    return SO3G_OK;
}

// Synchronises and reports an inactive-tile hit the way the reference's
// tiling_exception does (a RuntimeError), and a cached pixel index outside
// the map (which the reference does not check) as a RuntimeError too.
int check_tile_flag(hipStream_t stream) {
    HIP_TRY(hipMemcpyAsync(g_state.flags_h, g_state.flags_d, FLAG_COUNT * sizeof(int),
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (g_state.flags_h[FLAG_TILE])
        return fail(SO3G_ERR_RUNTIME,
                    "Tiling problem: Attempted pointing operation on non-instantiated tile.");
    if (g_state.flags_h[FLAG_RANGE])
        return fail(SO3G_ERR_RUNTIME, "pixel_index holds an index outside the map");
    return SO3G_OK;
}

int base_args(const so3g_geom *g, const double *bore, int64_t n_t, const double *dets,
              int64_t n_det, void *stream, OpArgs &a) {
    int r = ensure_init();
    if (r) return r;
    memset(&a, 0, sizeof(a));
    r = geom_to_launch(g, a.g, a.pk);
    if (r) return r;
    if (n_t < 0 || n_det < 0) return fail(SO3G_ERR_RUNTIME, "negative sample or detector count");
    if (n_t > INT32_MAX) return fail(SO3G_ERR_RUNTIME, "n_t exceeds int32 range");
    a.in.bore = bore;
    a.in.dets = dets;
    a.in.n_t = n_t;
    a.in.n_det = n_det;
    a.in.tabs = Tabs{g_state.asin_d, g_state.atan2_d};
    // Detectors per block: at least 4 (boresight reuse), more when the
    // 1-D grid of (detector group, 256-sample chunk) blocks would exceed
    // 2^31 - 1 blocks.
    a.in.dpb = 4;
    const int64_t n_chunk = (n_t + kBlock - 1) / kBlock;
    while (n_chunk * ((n_det + a.in.dpb - 1) / a.in.dpb) > INT32_MAX) a.in.dpb *= 2;
    a.in.agg = g_opt.scatter_agg;
    a.interp = g->interp;
    a.nc = g->ncomp;
    a.stream = (hipStream_t)stream;
    a.flags = g_state.flags_d;
    return SO3G_OK;
}

}  // namespace

namespace so3g {
int set_error(int code, const std::string &msg) { return fail(code, msg); }

// Entry points for the other translation units of the library (batch.hip).
int op_args(const so3g_geom *g, void *stream, OpArgs &a) {
    return base_args(g, nullptr, 1, nullptr, 1, stream, a);
}
int launch(int proj, const OpArgs &a) { return run_dispatch(proj, a); }
int flags_zero(hipStream_t stream) { return zero_flags(stream); }
int flags_check_tile(hipStream_t stream) { return check_tile_flag(stream); }

// Detectors sorted by the map row of their coordinates co (n_det x 4, lon
// lat ...; rows of the pixel height, or of the HEALPix pixel size), then by
// longitude: row-mates of a constant-elevation scan sweep the same pixels.
void row_order(const so3g_geom *g, const double *co, int64_t n_det, int32_t *idx) {
    double row = 1.;
    if (g->pixelizor == SO3G_PIX_FLAT) row = std::fabs(g->cdelt[0]);
    else row = std::sqrt(4. * M_PI / (12. * (double)g->nside * g->nside));
    if (!(row > 0.)) row = 1.;
    std::vector<int64_t> rows(n_det);
    for (int64_t d = 0; d < n_det; d++) {
        const double y = std::floor(co[4 * d + 1] / row);
        rows[d] = std::isfinite(y) ? (int64_t)std::max(-1e15, std::min(1e15, y)) : 0;
    }
    for (int64_t i = 0; i < n_det; i++) idx[i] = (int32_t)i;
    std::stable_sort(idx, idx + n_det, [&](int32_t x, int32_t y) {
        if (rows[x] != rows[y]) return rows[x] < rows[y];
        return co[4 * x] < co[4 * y];
    });
}

// True when every map element of the geometry has an int32 key.
bool hash_keys_fit(const so3g_geom *g) {
    if (g->pixelizor != SO3G_PIX_FLAT) return true;   // 12 nside^2 <= 8.1e8
    if (!g->tiled) return g->npix < INT32_MAX;
    return (int64_t)g->n_tile * g->tile_shape[0] * g->tile_shape[1] < INT32_MAX;
}
// Detector groups of the row scatter from each detector's coordinates at
// one sample (co, n_det x 4): the home row and column are the pixel formula
// of get_pixel on the host; detectors are sorted by home row, then along
// it, and split into groups of at most G that share a row (gdet: G entries
// per group, -1 padded; grow: the row, -1 for detectors off the map; gxc:
// the mean home column).  The groups only steer which samples add into LDS
// rather than to the map, so any grouping gives the same sums.
void make_row_groups(const so3g_geom *g, const double *co, int64_t n_det, int G,
                     std::vector<int32_t> &gdet, std::vector<int32_t> &grow,
                     std::vector<int32_t> &gxc) {
    std::vector<int32_t> row(n_det), col(n_det), idx(n_det);
    for (int64_t d = 0; d < n_det; d++) {
        const double x = co[4 * d] / g->cdelt[1] + g->crpix[1] - 1 + 0.5;
        const double y = co[4 * d + 1] / g->cdelt[0] + g->crpix[0] - 1 + 0.5;
        const bool on = x >= 0 && x < g->naxis[1] && y >= 0 && y < g->naxis[0];
        row[d] = on ? (int32_t)y : -1;
        col[d] = on ? (int32_t)x : 0;
        idx[d] = (int32_t)d;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t p, int32_t q) {
        if (row[p] != row[q]) return row[p] < row[q];
        return co[4 * p] < co[4 * q];
    });
    gdet.clear();
    grow.clear();
    gxc.clear();
    int64_t csum = 0, cnt = 0;
    auto close_group = [&]() {
        while (gdet.size() % G) gdet.push_back(-1);
        if (cnt) gxc.push_back((int32_t)(csum / cnt));
        csum = cnt = 0;
    };
    for (int64_t i = 0; i < n_det; i++) {
        const int32_t d = idx[i];
        const bool fresh = grow.empty() || grow.back() != row[d] || gdet.size() % G == 0;
        if (fresh) {
            close_group();
            grow.push_back(row[d]);
        }
        gdet.push_back(d);
        csum += col[d];
        cnt++;
    }
    close_group();
}

// Widest row segment whose sums (and tiled maps' active flags) fit in LDS.
int row_width_max(const so3g_geom *g, int nv) {
    return (int)((160 * 1024) / (nv * 8 + (g->tiled ? 1 : 0)));
}

}  // namespace so3g

extern "C" {

int so3g_abi_version(void) { return 1; }

const char *so3g_last_error(void) { return g_err.c_str(); }

int so3g_init(int device) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_state.device >= 0) {
        if (device != g_state.device)
            return fail(SO3G_ERR_RUNTIME, "library already initialised on another device");
        return SO3G_OK;
    }
    HIP_TRY(hipSetDevice(device));
    build_tables(g_state.asin_h, g_state.atan2_h);
    HIP_TRY(hipMalloc(&g_state.asin_d, kTrigN * sizeof(double)));
    HIP_TRY(hipMalloc(&g_state.atan2_d, kTrigN * sizeof(double)));
    HIP_TRY(hipMemcpy(g_state.asin_d, g_state.asin_h.data(), kTrigN * sizeof(double),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g_state.atan2_d, g_state.atan2_h.data(), kTrigN * sizeof(double),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&g_state.flags_d, FLAG_COUNT * sizeof(int)));
    HIP_TRY(hipMemset(g_state.flags_d, 0, FLAG_COUNT * sizeof(int)));
    HIP_TRY(hipHostMalloc(&g_state.flags_h, FLAG_COUNT * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipGetSymbolAddress((void **)&g_state.scratch_d, HIP_SYMBOL(scratch)));
    (void)hipDeviceGetAttribute(&g_state.n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (g_state.n_cu <= 0) g_state.n_cu = 256;
    g_state.device = device;
    return SO3G_OK;
}

int so3g_get_tables(double *asin_vals, double *atan2_vals) {
    if (g_state.device >= 0) {
        // The tables the kernels read: copied back from the device.
        HIP_TRY(hipMemcpy(asin_vals, g_state.asin_d, kTrigN * sizeof(double),
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(atan2_vals, g_state.atan2_d, kTrigN * sizeof(double),
                          hipMemcpyDeviceToHost));
        return SO3G_OK;
    }
    std::vector<double> as, at;
    build_tables(as, at);
    memcpy(asin_vals, as.data(), kTrigN * sizeof(double));
    memcpy(atan2_vals, at.data(), kTrigN * sizeof(double));
    return SO3G_OK;
}

int so3g_geom_init(so3g_geom *g) {
    if (g == nullptr) return fail(SO3G_ERR_VALUE, "geometry is NULL");
    if (g->proj < SO3G_PROJ_FLAT || g->proj > SO3G_PROJ_ZEA)
        return fail(SO3G_ERR_VALUE, "unknown coordinate system");
    if (g->ncomp < 1 || g->ncomp > 3) return fail(SO3G_ERR_VALUE, "ncomp must be 1, 2 or 3");
    if (g->interp < 0 || g->interp > 1) return fail(SO3G_ERR_VALUE, "unknown interpolation");
    g->n_tile = 0;
    if (g->pixelizor == SO3G_PIX_FLAT) {
        g->index_count = g->tiled ? 3 : 2;
        g->npix = (int64_t)g->naxis[0] * g->naxis[1];
        if (g->tiled) {
            if (g->tile_shape[0] <= 0 || g->tile_shape[1] <= 0)
                return fail(SO3G_ERR_VALUE, "tile shape must be positive");
            const int n_ty = (g->naxis[0] + g->tile_shape[0] - 1) / g->tile_shape[0];
            const int n_tx = (g->naxis[1] + g->tile_shape[1] - 1) / g->tile_shape[1];
            g->n_tile = n_ty * n_tx;
        }
        return SO3G_OK;
    }
    if (g->pixelizor != SO3G_PIX_HEALPIX) return fail(SO3G_ERR_VALUE, "unknown pixelizor");
    if (g->proj != SO3G_PROJ_CAR)
        return fail(SO3G_ERR_VALUE, "HEALPix engines use the CAR pointer");
    if (g->interp != 0)
        return fail(SO3G_ERR_VALUE, "Only 'nearest' interpolation is supported for Healpix");
    // check_nside, Projection.cxx:607-614: 0 < nside < 8192, power of 2.
    auto nside_ok = [&](int n) {
        const int lim = g->allow_nside_max ? 8192 : 8191;
        return n > 0 && n <= lim && (n & (n - 1)) == 0;
    };
    if (g->tiled && !g->nest) return fail(SO3G_ERR_VALUE, "RING not supported for tiled maps");
    if (!nside_ok(g->nside))
        return fail(SO3G_ERR_VALUE, "Invalid nside " + std::to_string(g->nside));
    g->npix = 12LL * g->nside * g->nside;
    if (g->tiled) {
        if (!nside_ok(g->nside_tile))
            return fail(SO3G_ERR_VALUE, "Invalid nside " + std::to_string(g->nside_tile));
        if (g->nside_tile > g->nside)
            return fail(SO3G_ERR_VALUE, "Invalid nside_tile " + std::to_string(g->nside_tile) +
                                            " > nside " + std::to_string(g->nside));
        g->n_tile = 12 * g->nside_tile * g->nside_tile;
        g->index_count = 2;
    } else {
        g->index_count = 1;
    }
    return SO3G_OK;
}

int so3g_coords(const so3g_geom *g, const double *bore, int64_t n_t, const double *dets,
                int64_t n_det, double *coords, void *stream) {
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    a.op = OP_COORDS;
    a.coords_out = coords;
    return run_dispatch(g->proj, a);
}

static int pixels_common(const so3g_geom *g, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         int32_t *pixel, float *spin, void *stream) {
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    r = zero_flags(a.stream);
    if (r) return r;
    a.op = spin ? OP_PMAT : OP_PIXELS;
    a.resp = resp;
    a.pix_out = pixel;
    a.spin_out = spin;
    return run_dispatch(g->proj, a);
}

int so3g_pixels(const so3g_geom *g, const double *bore, int64_t n_t, const double *dets,
                int64_t n_det, int32_t *pixel, void *stream) {
    return pixels_common(g, bore, n_t, dets, n_det, nullptr, pixel, nullptr, stream);
}

int so3g_pointing_matrix(const so3g_geom *g, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         int32_t *pixel, float *spin_proj, void *stream) {
    if (spin_proj == nullptr) return fail(SO3G_ERR_VALUE, "spin_proj is NULL");
    return pixels_common(g, bore, n_t, dets, n_det, resp, pixel, spin_proj, stream);
}

int so3g_from_map(const so3g_geom *g, const double *map, const double *const *tiles,
                  const int32_t *tile_dims, const double *bore, int64_t n_t,
                  const double *dets, int64_t n_det, const float *resp, float *signal,
                  int accumulate, void *stream) {
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    if (g->tiled && (tiles == nullptr || tile_dims == nullptr))
        return fail(SO3G_ERR_RUNTIME, "tiled map requires tile pointers");
    if (g->tiled) {
        r = zero_flags(a.stream);
        if (r) return r;
    }
    a.op = OP_FROM_MAP;
    a.map = (void *)map;
    a.tiles = (void *const *)tiles;
    a.tile_dims = tile_dims;
    a.resp = resp;
    a.sig_out = signal;
    a.accumulate = accumulate;
    r = run_dispatch(g->proj, a);
    if (r) return r;
    return g->tiled ? check_tile_flag(a.stream) : SO3G_OK;
}

// Detector order for k_scatter_lds: each detector's coordinates at the
// middle sample (one k_coords launch over n_det samples), sorted by map row
// (the coordinate across the scan, in units of the pixel height) and then
// along the row, so that a block's detectors sweep the same pixels.

static int scatter_order(const so3g_geom *g, const OpArgs &a0, const int32_t **order) {
    const int64_t n_det = a0.in.n_det, n_t = a0.in.n_t;
    if (g_state.order_bore == a0.in.bore && g_state.order_dets == a0.in.dets &&
        g_state.order_nt == n_t && g_state.order_ndet == n_det &&
        memcmp(&g_state.order_geom, g, sizeof(*g)) == 0) {
        *order = g_state.order_d;
        return SO3G_OK;
    }
    if (n_det > g_state.order_cap) {
        if (g_state.order_d) HIP_TRY(hipFree(g_state.order_d));
        if (g_state.order_co_d) HIP_TRY(hipFree(g_state.order_co_d));
        g_state.order_d = nullptr;
        g_state.order_co_d = nullptr;
        g_state.order_cap = 0;
        HIP_TRY(hipMalloc(&g_state.order_d, n_det * sizeof(int32_t)));
        HIP_TRY(hipMalloc(&g_state.order_co_d, n_det * 4 * sizeof(double)));
        g_state.order_cap = n_det;
    }
    OpArgs a = a0;
    a.op = OP_COORDS;
    a.in.bore = a0.in.bore + 4 * (n_t / 2);
    a.in.n_t = 1;
    a.in.chunks = nullptr;
    a.coords_out = g_state.order_co_d;
    int r = run_dispatch(g->proj, a);
    if (r) return r;
    std::vector<double> co(n_det * 4);
    HIP_TRY(hipMemcpyAsync(co.data(), g_state.order_co_d, n_det * 4 * sizeof(double),
                           hipMemcpyDeviceToHost, a.stream));
    HIP_TRY(hipStreamSynchronize(a.stream));
    std::vector<int32_t> idx(n_det);
    so3g::row_order(g, co.data(), n_det, idx.data());
    HIP_TRY(hipMemcpyAsync(g_state.order_d, idx.data(), n_det * sizeof(int32_t),
                           hipMemcpyHostToDevice, a.stream));
    HIP_TRY(hipStreamSynchronize(a.stream));
    g_state.order_bore = a0.in.bore;
    g_state.order_dets = a0.in.dets;
    g_state.order_nt = n_t;
    g_state.order_ndet = n_det;
    g_state.order_geom = *g;
    *order = g_state.order_d;
    return SO3G_OK;
}

// Row groups of one observation (its middle sample; one k_coords launch over
// n_det samples), cached on the device for repeated calls.  A stale cache
// entry costs speed, never correctness.
static int row_groups(const so3g_geom *g, const OpArgs &a0, int G, RowWork &rw) {
    const int64_t n_det = a0.in.n_det, n_t = a0.in.n_t;
    if (!(g_state.rg_bore == a0.in.bore && g_state.rg_dets == a0.in.dets &&
          g_state.rg_nt == n_t && g_state.rg_ndet == n_det && g_state.rg_G == G &&
          memcmp(&g_state.rg_geom, g, sizeof(*g)) == 0)) {
        if (n_det > g_state.order_cap) {
            if (g_state.order_d) HIP_TRY(hipFree(g_state.order_d));
            if (g_state.order_co_d) HIP_TRY(hipFree(g_state.order_co_d));
            g_state.order_d = nullptr;
            g_state.order_co_d = nullptr;
            g_state.order_cap = 0;
            g_state.order_bore = nullptr;
            HIP_TRY(hipMalloc(&g_state.order_d, n_det * sizeof(int32_t)));
            HIP_TRY(hipMalloc(&g_state.order_co_d, n_det * 4 * sizeof(double)));
            g_state.order_cap = n_det;
        }
        OpArgs a = a0;
        a.op = OP_COORDS;
        a.in.bore = a0.in.bore + 4 * (n_t / 2);
        a.in.n_t = 1;
        a.in.chunks = nullptr;
        a.coords_out = g_state.order_co_d;
        int r = run_dispatch(g->proj, a);
        if (r) return r;
        std::vector<double> co(n_det * 4);
        HIP_TRY(hipMemcpyAsync(co.data(), g_state.order_co_d, n_det * 4 * sizeof(double),
                               hipMemcpyDeviceToHost, a.stream));
        HIP_TRY(hipStreamSynchronize(a.stream));
        std::vector<int32_t> gdet, grow, gxc;
        make_row_groups(g, co.data(), n_det, G, gdet, grow, gxc);
        const int64_t ng = (int64_t)grow.size();
        if (ng * G > g_state.rg_cap || ng > g_state.rg_cap) {
            if (g_state.rg_det_d) HIP_TRY(hipFree(g_state.rg_det_d));
            if (g_state.rg_row_d) HIP_TRY(hipFree(g_state.rg_row_d));
            if (g_state.rg_xc_d) HIP_TRY(hipFree(g_state.rg_xc_d));
            g_state.rg_det_d = g_state.rg_row_d = g_state.rg_xc_d = nullptr;
            g_state.rg_cap = 0;
            g_state.rg_bore = nullptr;
            HIP_TRY(hipMalloc(&g_state.rg_det_d, ng * G * sizeof(int32_t)));
            HIP_TRY(hipMalloc(&g_state.rg_row_d, ng * G * sizeof(int32_t)));
            HIP_TRY(hipMalloc(&g_state.rg_xc_d, ng * G * sizeof(int32_t)));
            g_state.rg_cap = ng * G;
        }
        HIP_TRY(hipMemcpyAsync(g_state.rg_det_d, gdet.data(), ng * G * sizeof(int32_t),
                               hipMemcpyHostToDevice, a.stream));
        HIP_TRY(hipMemcpyAsync(g_state.rg_row_d, grow.data(), ng * sizeof(int32_t),
                               hipMemcpyHostToDevice, a.stream));
        HIP_TRY(hipMemcpyAsync(g_state.rg_xc_d, gxc.data(), ng * sizeof(int32_t),
                               hipMemcpyHostToDevice, a.stream));
        HIP_TRY(hipStreamSynchronize(a.stream));
        g_state.rg_bore = a0.in.bore;
        g_state.rg_dets = a0.in.dets;
        g_state.rg_nt = n_t;
        g_state.rg_ndet = n_det;
        g_state.rg_G = G;
        g_state.rg_n = (int32_t)ng;
        g_state.rg_geom = *g;
    }
    rw.gdet = g_state.rg_det_d;
    rw.grow = g_state.rg_row_d;
    rw.gxc = g_state.rg_xc_d;
    rw.n_groups = g_state.rg_n;
    rw.G = G;
    return SO3G_OK;
}

static int to_map_common(const so3g_geom *g, void *map, void *const *tiles,
                         const int32_t *tile_dims, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         const float *signal, const float *det_weights, const int32_t *segs,
                         int64_t n_seg, int prec, int op, void *stream) {
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    if (g->tiled && (tiles == nullptr || tile_dims == nullptr))
        return fail(SO3G_ERR_RUNTIME, "tiled map requires tile pointers");
    if (segs != nullptr) {
        r = upload_chunks(segs, n_seg, n_det, n_t, a.stream, &a.n_chunks);
        if (r) return r;
        if (a.n_chunks == 0) return SO3G_OK;
        a.in.chunks = g_state.chunks_d;
    }
    if (g->tiled) {
        r = zero_flags(a.stream);
        if (r) return r;
    }
    a.op = op;
    a.prec = prec;
    a.map = map;
    a.tiles = tiles;
    a.tile_dims = tile_dims;
    a.resp = resp;
    a.sig_in = signal;
    a.det_w = det_weights;
    a.hash_cls = 0;
    // LDS-privatised partial maps (k_scatter_lds) where they measured faster
    // than global atomics (1000 x 720000 samples, ms; global / large table /
    // small table): CAR 0.5' to_map 9.0 / 11.6 / 8.6; CAR to_weights
    // 17.7 / 15.3 / 11.8 (8 detectors per block); HEALPix nside 4096 to_map
    // 17.4 / 14.9 / 15.4; HEALPix to_weights 33.2 / 20.0 / 17.3; tiled CAR
    // to_map 9.4 / 13.7 / 11.5; tiled CAR to_weights 17.7 / 19.7 / 13.9.
    // Every global atomic is forwarded to the
    // memory side, about 6 lanes per 64 B request for a wave's run ends; a
    // small table (256 threads, 512 slots, home slot = pixel mod 512) sums
    // four row-mates over one 256-sample chunk and flushes consecutive
    // pixels together, and keeps more blocks per compute unit than the
    // large one (1024 threads, 112 KB, a span of up to 131072 samples).
    int cls = 0, per_block = 0, span = 0, lane_r = g_opt.hash_r;
    const int nv_row = op == OP_TO_MAP ? g->ncomp : g->ncomp * (g->ncomp + 1) / 2;
    const bool row_ok = g->pixelizor == SO3G_PIX_FLAT && g->interp == 0 &&
                        prec == SO3G_PREC_REF && segs == nullptr;
    // Map rows in LDS (k_scatter_row) in modes 1 and 4 (1000 x 720000
    // samples, CAR 0.5' TQU: to_map 7.2 ms against 8.7 with the small
    // tables, to_weights 6.9 against 11.7; groups of 2 detectors over 65536
    // samples: 4 / 8 detectors 7.4 / 7.7 ms, spans of 16384 / 32768 7.5 /
    // 7.35 ms for to_map; in 100 x 100 tiles to_map 7.7 against 11.2,
    // to_weights 7.6 against 12.4).
    if (row_ok && (g_opt.scatter_hash == 4 || g_opt.scatter_hash == 1)) {
        a.hash_cls = 3;
        r = row_groups(g, a, g_opt.hash_dets > 0 ? g_opt.hash_dets : 2, a.rw);
        if (r) return r;
        if (g_opt.hash_span > 0) {
            a.rw.span = g_opt.hash_span;
        } else {
            // about 8 blocks per compute unit, spans a multiple of 4096
            // samples (1024 lanes x 4) from 4096 to 131072: 100 x 40000
            // samples 0.22 -> 0.10 ms (spans of 65536 / 4096), 1000 x 720000
            // 7.25 -> 7.16 ms (65536 / 131072)
            const int64_t want = 8LL * g_state.n_cu;
            int64_t sp = (n_t * (int64_t)a.rw.n_groups + want - 1) / std::max<int64_t>(want, 1);
            sp = (sp + 4095) / 4096 * 4096;
            a.rw.span = (int32_t)std::min<int64_t>(std::max<int64_t>(sp, 4096), 131072);
        }
        a.rw.xw = (int32_t)std::min<int64_t>(g->naxis[1], row_width_max(g, nv_row));
        while (((n_t + a.rw.span - 1) / a.rw.span) * (int64_t)a.rw.n_groups > INT32_MAX)
            a.rw.span *= 2;
        r = run_dispatch(g->proj, a);
        if (r) return r;
        return g->tiled ? check_tile_flag(a.stream) : SO3G_OK;
    }
    if (g_opt.scatter_hash == 1 || g_opt.scatter_hash == 4) {
        if (g->pixelizor == SO3G_PIX_HEALPIX) {
            // to_map: 4 samples per lane with in-lane run sums (13.7 ms
            // against 14.4 with the wave-level scan; with the small tables
            // and to_weights 4 samples per lane overflow the 512 slots and
            // lose); spans up to 131072 samples (1000 x 720000 on nside
            // 4096 NEST, 2 detectors per block: 8192 / 16384 / 32768 /
            // 65536 / 131072 samples 14.5 / 11.7 / 10.7 / 10.3 / 10.1 ms;
            // 1 or 4 detectors slower at every span)
            // to_weights: the large tables too, one detector per block (its
            // 6 sums per pixel leave 2048 slots), 4 samples per lane: 13.8
            // ms with the small tables (4 detectors x 256 samples) against
            // 11.9; 2 / 4 detectors per block 13.6 / 15.4 ms
            if (op == OP_TO_MAP) cls = 1, per_block = 2, span = 131072, lane_r = 4;
            else cls = 1, per_block = 1, span = 131072, lane_r = 4;
        } else if (op == OP_TO_WEIGHTS) {
            cls = 2, per_block = g->tiled ? 4 : 8, span = 256;
        } else if (!g->tiled) {
            cls = 2, per_block = 4, span = 256;
        }
    } else if (g_opt.scatter_hash >= 2) {
        cls = g_opt.scatter_hash == 2 ? 2 : 1;
        per_block = g_opt.hash_dets > 0 ? g_opt.hash_dets : (cls == 2 ? 4 : 2);
        span = g_opt.hash_span > 0 ? g_opt.hash_span : (cls == 2 ? 256 : 131072);
    }
    if (cls > 0 && prec == SO3G_PREC_REF && segs == nullptr && hash_keys_fit(g)) {
        r = scatter_order(g, a, &a.hw.order);
        if (r) return r;
        a.hash_cls = cls;
        a.hw.dets = per_block;
        a.hw.span = span;
        a.hw.r = g->interp == 0 ? lane_r : 1;
        if (cls == 1 && g_opt.hash_span == 0) {
            // large tables: as for the rows, about 8 blocks per compute unit
            // (spans of 4096 to 131072 samples, a multiple of 1024 lanes x 4)
            const int64_t want = 8LL * g_state.n_cu;
            const int64_t ng = (n_det + per_block - 1) / per_block;
            int64_t sp = (n_t * ng + want - 1) / want;
            sp = (sp + 4095) / 4096 * 4096;
            a.hw.span = (int32_t)std::min<int64_t>(std::max<int64_t>(sp, 4096), span);
        }
        // the 1-D grid of (group, span) blocks stays below 2^31
        while (((n_t + a.hw.span - 1) / a.hw.span) * ((n_det + a.hw.dets - 1) / a.hw.dets) >
               INT32_MAX)
            a.hw.dets *= 2;
    }
    r = run_dispatch(g->proj, a);
    if (r) return r;
    return g->tiled ? check_tile_flag(a.stream) : SO3G_OK;
}

int so3g_to_map(const so3g_geom *g, double *map, double *const *tiles,
                const int32_t *tile_dims, const double *bore, int64_t n_t, const double *dets,
                int64_t n_det, const float *resp, const float *signal,
                const float *det_weights, const int32_t *segs, int64_t n_seg, void *stream) {
    return to_map_common(g, map, (void *const *)tiles, tile_dims, bore, n_t, dets, n_det,
                         resp, signal, det_weights, segs, n_seg, SO3G_PREC_REF, OP_TO_MAP,
                         stream);
}

int so3g_to_weights(const so3g_geom *g, double *map, double *const *tiles,
                    const int32_t *tile_dims, const double *bore, int64_t n_t,
                    const double *dets, int64_t n_det, const float *resp,
                    const float *det_weights, const int32_t *segs, int64_t n_seg,
                    void *stream) {
    return to_map_common(g, map, (void *const *)tiles, tile_dims, bore, n_t, dets, n_det,
                         resp, nullptr, det_weights, segs, n_seg, SO3G_PREC_REF,
                         OP_TO_WEIGHTS, stream);
}

int so3g_to_map_prec(const so3g_geom *g, void *map, void *const *tiles,
                     const int32_t *tile_dims, const double *bore, int64_t n_t,
                     const double *dets, int64_t n_det, const float *resp,
                     const float *signal, const float *det_weights, int prec, void *stream) {
    if (prec < 0 || prec > 2) return fail(SO3G_ERR_VALUE, "unknown precision mode");
    return to_map_common(g, map, tiles, tile_dims, bore, n_t, dets, n_det, resp, signal,
                         det_weights, nullptr, 0, prec, OP_TO_MAP, stream);
}

int so3g_to_weights_prec(const so3g_geom *g, void *map, void *const *tiles,
                         const int32_t *tile_dims, const double *bore, int64_t n_t,
                         const double *dets, int64_t n_det, const float *resp,
                         const float *det_weights, int prec, void *stream) {
    if (prec < 0 || prec > 2) return fail(SO3G_ERR_VALUE, "unknown precision mode");
    return to_map_common(g, map, tiles, tile_dims, bore, n_t, dets, n_det, resp, nullptr,
                         det_weights, nullptr, 0, prec, OP_TO_WEIGHTS, stream);
}

int so3g_tile_hits(const so3g_geom *g, const double *bore, int64_t n_t, const double *dets,
                   int64_t n_det, int32_t *hits, void *stream) {
    if (g != nullptr && !g->tiled)
        return fail(SO3G_ERR_RUNTIME, "No tiles in this pixelization.");
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    HIP_TRY(hipMemsetAsync(hits, 0, sizeof(int32_t) * g->n_tile, a.stream));
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    a.op = OP_TILE_HITS;
    a.hits = hits;
    return run_dispatch(g->proj, a);
}

int so3g_sample_domains(const so3g_geom *g, const double *bore, int64_t n_t,
                        const double *dets, int64_t n_det, int mode, int n_domain,
                        const double *tmap, const double *const *ttiles,
                        const int32_t *tile_dims, const int32_t *tile_dom, int32_t *out,
                        void *stream) {
    OpArgs a;
    int r = base_args(g, bore, n_t, dets, n_det, stream, a);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    if (mode < 0 || mode > 2) return fail(SO3G_ERR_VALUE, "unknown domain mode");
    if (mode == SO3G_DOMAIN_TILE && !g->tiled)
        return fail(SO3G_ERR_RUNTIME, "No tiles in this pixelization.");
    if (n_domain <= 0) return fail(SO3G_ERR_VALUE, "n_domain must be positive");
    r = zero_flags(a.stream);
    if (r) return r;
    a.op = OP_DOMAINS;
    a.mode = mode;
    a.n_domain = n_domain;
    a.map = (void *)tmap;
    a.tiles = (void *const *)ttiles;
    a.tile_dims = tile_dims;
    a.tile_dom = tile_dom;
    a.dom_out = out;
    r = run_dispatch(g->proj, a);
    if (r) return r;
    return check_tile_flag(a.stream);
}

// ---------------------------------------------------------------------------
// ProjEng_Precomp.
// ---------------------------------------------------------------------------
static int pc_setup(const int64_t *map_shape, const int32_t *tile_dims, int32_t k,
                    int32_t nc, int64_t n_det, int64_t n_t, void *stream, PcIn &in,
                    PcMap &pm) {
    int r = ensure_init();
    if (r) return r;
    if (k < 1 || k > 3) return fail(SO3G_ERR_RUNTIME, "pixel index count must be 1, 2 or 3");
    if (nc < 1 || nc > 3) return fail(SO3G_ERR_RUNTIME, "spin component count must be 1..3");
    if (n_t < 0 || n_det < 0 || n_t > INT32_MAX)
        return fail(SO3G_ERR_RUNTIME, "bad sample or detector count");
    memset(&in, 0, sizeof(in));
    memset(&pm, 0, sizeof(pm));
    in.n_t = n_t;
    in.n_det = n_det;
    in.dpb = (int32_t)std::max<int64_t>(1, (n_det + 65534) / 65535);
    if (k < 3) {
        if (map_shape == nullptr) return fail(SO3G_ERR_RUNTIME, "map shape required");
        pm.ny = map_shape[0];
        pm.nx = map_shape[1];
    } else {
        if (tile_dims == nullptr) return fail(SO3G_ERR_RUNTIME, "tile dims required");
        if (map_shape == nullptr || map_shape[0] < 0 || map_shape[0] > INT32_MAX)
            return fail(SO3G_ERR_RUNTIME, "tiled maps need map_shape = (n_tile, 1)");
        pm.tile_dims = tile_dims;
        pm.n_tile = (int32_t)map_shape[0];
    }
    (void)stream;
    return SO3G_OK;
}

static dim3 pc_grid(const PcIn &in, int64_t n_chunks) {
    if (in.chunks) return dim3((unsigned)n_chunks);
    return dim3((unsigned)((in.n_t + kBlock - 1) / kBlock),
                (unsigned)((in.n_det + in.dpb - 1) / in.dpb));
}

#define PC_SWITCH(K, NC, KERNEL, ...)                                                   \
    do {                                                                               \
        int _ok = 1;                                                                   \
        switch ((K) * 4 + (NC)) {                                                      \
        case 5: hipLaunchKernelGGL((KERNEL<1, 1>), __VA_ARGS__); break;                \
        case 6: hipLaunchKernelGGL((KERNEL<1, 2>), __VA_ARGS__); break;                \
        case 7: hipLaunchKernelGGL((KERNEL<1, 3>), __VA_ARGS__); break;                \
        case 9: hipLaunchKernelGGL((KERNEL<2, 1>), __VA_ARGS__); break;                \
        case 10: hipLaunchKernelGGL((KERNEL<2, 2>), __VA_ARGS__); break;               \
        case 11: hipLaunchKernelGGL((KERNEL<2, 3>), __VA_ARGS__); break;               \
        case 13: hipLaunchKernelGGL((KERNEL<3, 1>), __VA_ARGS__); break;               \
        case 14: hipLaunchKernelGGL((KERNEL<3, 2>), __VA_ARGS__); break;               \
        case 15: hipLaunchKernelGGL((KERNEL<3, 3>), __VA_ARGS__); break;               \
        default: _ok = 0;                                                              \
        }                                                                              \
        if (!_ok) return fail(SO3G_ERR_RUNTIME, "unsupported pixel/spin layout");       \
    } while (0)

// The k = 2, n_t % 4 == 0 gather: map packed pixel-major into pk (4 * plane
// doubles), then the packed gather.
static int pc_from_map_packed(const double *map, double *pk, const PcIn &in0, const PcMap &pm,
                              const int32_t *pixel, const float *spin, int32_t nc,
                              float *signal, int accumulate, hipStream_t s) {
    PcIn in = in0;
    in.dpb = std::max<int32_t>(in.dpb, (int32_t)g_opt.pc_from_dpb);
    const int64_t plane = pm.ny * pm.nx;
    int *fl = g_state.flags_d;
    const dim3 gp(1024), g4((unsigned)((in.n_t / 4 + kBlock - 1) / kBlock),
                            (unsigned)((in.n_det + in.dpb - 1) / in.dpb));
    switch (nc) {
    case 1:
        hipLaunchKernelGGL((k_pack_map<1>), gp, dim3(kBlock), 0, s, map, plane, pk);
        hipLaunchKernelGGL((k_pc_from_map_packed<1>), g4, dim3(kBlock), 0, s, in, pm, pk, pixel, spin, signal, accumulate, fl);
        break;
    case 2:
        hipLaunchKernelGGL((k_pack_map<2>), gp, dim3(kBlock), 0, s, map, plane, pk);
        hipLaunchKernelGGL((k_pc_from_map_packed<2>), g4, dim3(kBlock), 0, s, in, pm, pk, pixel, spin, signal, accumulate, fl);
        break;
    default:
        hipLaunchKernelGGL((k_pack_map<3>), gp, dim3(kBlock), 0, s, map, plane, pk);
        hipLaunchKernelGGL((k_pc_from_map_packed<3>), g4, dim3(kBlock), 0, s, in, pm, pk, pixel, spin, signal, accumulate, fl);
        break;
    }
    HIP_TRY(hipGetLastError());
    return check_tile_flag(s);
}

int so3g_precomp_from_map(const double *map, const int64_t *map_shape,
                          const double *const *tiles, const int32_t *tile_dims,
                          const int32_t *pixel, int32_t k, const float *spin, int32_t nc,
                          int64_t n_det, int64_t n_t, float *signal, int accumulate,
                          void *stream) {
    PcIn in;
    PcMap pm;
    int r = pc_setup(map_shape, tile_dims, k, nc, n_det, n_t, stream, in, pm);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    r = zero_flags(s);
    if (r) return r;
    if (k == 2 && n_t % 4 == 0) {
        const int64_t plane = pm.ny * pm.nx;
        if ((size_t)(plane * 4) > g_state.packed_cap) {
            if (g_state.packed_d) HIP_TRY(hipFree(g_state.packed_d));
            g_state.packed_d = nullptr;
            g_state.packed_cap = 0;
            HIP_TRY(hipMalloc(&g_state.packed_d, plane * 4 * sizeof(double)));
            g_state.packed_cap = plane * 4;
        }
        return pc_from_map_packed(map, g_state.packed_d, in, pm, pixel, spin, nc, signal,
                                  accumulate, s);
    }
    PC_SWITCH(k, nc, k_pc_from_map, pc_grid(in, 0), dim3(kBlock), 0, s, in, pm, map, tiles,
              pixel, spin, signal, accumulate, g_state.flags_d);
    HIP_TRY(hipGetLastError());
    return check_tile_flag(s);
}

int so3g_precomp_from_map_noalloc(const double *map, const int64_t *map_shape,
                                  const int32_t *pixel, const float *spin, int32_t nc,
                                  int64_t n_det, int64_t n_t, float *signal, int accumulate,
                                  void *stream) {
    // Initialisation allocates (tables, flags): it must have happened.
    if (g_state.device < 0 || g_state.scratch_d == nullptr)
        return fail(SO3G_ERR_RUNTIME, "so3g_precomp_from_map_noalloc: call so3g_init first");
    PcIn in;
    PcMap pm;
    int r = pc_setup(map_shape, nullptr, 2, nc, n_det, n_t, stream, in, pm);
    if (r) return r;
    if (pm.ny < 0 || pm.nx < 0 || pm.ny * pm.nx > kScratchPixels)
        return fail(SO3G_ERR_VALUE, "so3g_precomp_from_map_noalloc: the map has " +
                                        std::to_string(pm.ny * pm.nx) + " pixels, more than "
                                        "the " + std::to_string(kScratchPixels) +
                                        " its scratch holds");
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    r = zero_flags(s);
    if (r) return r;
    if (n_t % 4 == 0)
        return pc_from_map_packed(map, g_state.scratch_d, in, pm, pixel, spin, nc, signal,
                                  accumulate, s);
    PC_SWITCH(2, nc, k_pc_from_map, pc_grid(in, 0), dim3(kBlock), 0, s, in, pm, map, nullptr,
              pixel, spin, signal, accumulate, g_state.flags_d);
    HIP_TRY(hipGetLastError());
    return check_tile_flag(s);
}

// Launches the precomp to_map variant selected by the options.
extern "C++" template <int K, int NC>
static int pc_to_map_launch(dim3 grid, hipStream_t s, const PcIn &in, const PcMap &pm,
                            double *map, double *const *tiles, const int32_t *pixel,
                            const float *spin, const float *signal,
                            const float *det_weights) {
    const int64_t n = (int64_t)NC * pm.ny * pm.nx;
    bool xcd = g_opt.scatter_xcd && K < 3 &&
               n * kNumXcd * 8 <= g_opt.xcd_budget_mb * (int64_t)(1 << 20);
    if (xcd && (size_t)(n * kNumXcd) > g_state.copies_cap) {
        if (g_state.copies_d) HIP_TRY(hipFree(g_state.copies_d));
        g_state.copies_d = nullptr;
        g_state.copies_cap = 0;
        if (hipMalloc(&g_state.copies_d, n * kNumXcd * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            xcd = false;
        } else {
            g_state.copies_cap = n * kNumXcd;
        }
    }
    const dim3 block(kBlock);
    double *cp = xcd ? g_state.copies_d : nullptr;
    if (xcd)
        hipLaunchKernelGGL(k_fill_zero<double>, dim3(2048), block, 0, s, cp, n * kNumXcd);
    if (g_opt.scatter_agg) {
        if (xcd)
            hipLaunchKernelGGL((k_pc_to_map<K, NC, true, true>), grid, block, 0, s, in, pm, map,
                               tiles, pixel, spin, signal, det_weights, cp, n, g_state.flags_d);
        else
            hipLaunchKernelGGL((k_pc_to_map<K, NC, true, false>), grid, block, 0, s, in, pm,
                               map, tiles, pixel, spin, signal, det_weights, cp, n,
                               g_state.flags_d);
    } else {
        if (xcd)
            hipLaunchKernelGGL((k_pc_to_map<K, NC, false, true>), grid, block, 0, s, in, pm,
                               map, tiles, pixel, spin, signal, det_weights, cp, n,
                               g_state.flags_d);
        else
            hipLaunchKernelGGL((k_pc_to_map<K, NC, false, false>), grid, block, 0, s, in, pm,
                               map, tiles, pixel, spin, signal, det_weights, cp, n,
                               g_state.flags_d);
    }
    if (xcd)
        hipLaunchKernelGGL(k_reduce_copies<double>, dim3(2048), block, 0, s, map, cp, n, n, kNumXcd);
    HIP_TRY(hipGetLastError());
    return SO3G_OK;
}

// Detector order of the Precomp LDS scatter: detectors sorted by the pixel
// they see at the middle sample (k = 2: row, then column; k = 3: tile, row,
// column; k = 1: pixel number; off-map detectors last), cached for the
// pixel array.  The order
// only decides which detectors share a block: any order gives the same sums,
// so an edit of the array in place costs speed, never correctness.
static int pc_order(const int32_t *pixel, int32_t k, int64_t n_det, int64_t n_t, hipStream_t s,
                    const int32_t **order) {
    if (g_state.pc_order_pix == pixel && g_state.pc_order_ndet == n_det &&
        g_state.pc_order_nt == n_t && g_state.pc_order_k == k) {
        *order = g_state.pc_order_d;
        return SO3G_OK;
    }
    if (n_det > g_state.pc_order_cap) {
        if (g_state.pc_order_d) HIP_TRY(hipFree(g_state.pc_order_d));
        g_state.pc_order_d = nullptr;
        g_state.pc_order_cap = 0;
        HIP_TRY(hipMalloc(&g_state.pc_order_d, 4 * n_det * sizeof(int32_t)));
        g_state.pc_order_cap = n_det;
    }
    g_state.pc_order_pix = nullptr;
    int32_t *mid_d = g_state.pc_order_d + n_det;
    hipLaunchKernelGGL(k_pc_mid, dim3((unsigned)((n_det + 255) / 256)), dim3(256), 0, s, pixel,
                       k, n_det, n_t, mid_d);
    HIP_TRY(hipGetLastError());
    std::vector<int32_t> mid(n_det * k);
    HIP_TRY(hipMemcpyAsync(mid.data(), mid_d, n_det * k * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<int64_t> key(n_det);
    for (int64_t d = 0; d < n_det; d++) {
        const int32_t a = mid[d * k];
        if (a < 0) key[d] = INT64_MAX;
        else if (k == 2) key[d] = ((int64_t)a << 32) | (uint32_t)mid[d * k + 1];
        else if (k == 3)
            key[d] = ((int64_t)a << 40) | ((int64_t)(mid[d * k + 1] & 0xFFFFF) << 20) |
                     (mid[d * k + 2] & 0xFFFFF);
        else key[d] = a;
    }
    std::vector<int32_t> idx(n_det);
    for (int64_t d = 0; d < n_det; d++) idx[d] = (int32_t)d;
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return key[x] < key[y]; });
    HIP_TRY(hipMemcpyAsync(g_state.pc_order_d, idx.data(), n_det * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    g_state.pc_order_pix = pixel;
    g_state.pc_order_ndet = n_det;
    g_state.pc_order_nt = n_t;
    g_state.pc_order_k = k;
    *order = g_state.pc_order_d;
    return SO3G_OK;
}

// Whether a Precomp to_map / to_weights goes through the LDS tables: modes
// 1 (auto) and 2 (small tables), without thread segments (segment lists
// keep the run-aggregated global atomics); non-tiled maps whose elements
// have int32 keys, and tiled maps for to_weights (tiled to_map, 1000 x
// 720000 samples into 100 x 100 tiles: 12.1 ms with global atomics, 12.6 ms
// at best with the tables; to_weights 21.6 -> 16.8 ms).
static bool pc_lds_applies(int32_t k, const PcMap &pm, const int32_t *segs, bool w) {
    if (!(g_opt.scatter_hash == 1 || g_opt.scatter_hash == 2 || g_opt.scatter_hash == 4) ||
        segs != nullptr)
        return false;
    if (k == 3) return w || g_opt.scatter_hash == 2;
    return pm.ny * pm.nx < INT32_MAX;
}

extern "C++" template <int K, int NC, bool W>
static void pc_lds_launch(dim3 grid, hipStream_t s, const PcIn &in, const PcMap &pm,
                          double *map, double *const *tiles, const int32_t *pixel,
                          const float *spin, const float *signal, const float *det_weights,
                          const int32_t *order, int32_t dets) {
    hipLaunchKernelGGL((k_pc_scatter_lds<K, NC, W, 512>), grid, dim3(kBlock), 0, s, in, pm,
                       map, tiles, pixel, spin, signal, det_weights, order, dets,
                       g_state.flags_d);
}

static int pc_lds_scatter(bool w, hipStream_t s, const PcIn &in, const PcMap &pm, int32_t k,
                          int32_t nc, double *map, double *const *tiles, const int32_t *pixel,
                          const float *spin, const float *signal, const float *det_weights) {
    const int32_t *order = nullptr;
    int r = pc_order(pixel, k, in.n_det, in.n_t, s, &order);
    if (r) return r;
    // detectors per block measured best on 1000 x 720000 samples (ms, global
    // atomics -> table): CAR 0.5' to_map 10.6 -> 6.3 (2 detectors; 6.7 with
    // 4), to_weights 20.1 -> 9.9 (4); HEALPix nside 4096 to_map 18.9 -> 10.0
    // (4), to_weights 34.8 -> 17.5 (8); tiled CAR to_weights 21.6 -> 16.8 (8)
    int64_t dets = g_opt.hash_dets > 0 ? g_opt.hash_dets
                                        : (w ? (k == 2 ? 4 : 8) : (k == 1 ? 4 : 2));
    const int64_t n_tc = (in.n_t + kBlock - 1) / kBlock;
    while (n_tc * ((in.n_det + dets - 1) / dets) > INT32_MAX) dets *= 2;
    const dim3 grid((unsigned)(n_tc * ((in.n_det + dets - 1) / dets)));
    const int32_t dd = (int32_t)dets;
#define PC_LDS(K, NC)                                                                   \
    (w ? pc_lds_launch<K, NC, true>(grid, s, in, pm, map, tiles, pixel, spin, signal,   \
                                    det_weights, order, dd)                             \
       : pc_lds_launch<K, NC, false>(grid, s, in, pm, map, tiles, pixel, spin, signal,  \
                                     det_weights, order, dd))
    switch (k * 4 + nc) {
    case 5: PC_LDS(1, 1); break;
    case 6: PC_LDS(1, 2); break;
    case 7: PC_LDS(1, 3); break;
    case 9: PC_LDS(2, 1); break;
    case 10: PC_LDS(2, 2); break;
    case 11: PC_LDS(2, 3); break;
    case 13: PC_LDS(3, 1); break;
    case 14: PC_LDS(3, 2); break;
    case 15: PC_LDS(3, 3); break;
    default: return fail(SO3G_ERR_RUNTIME, "unsupported pixel/spin layout");
    }
#undef PC_LDS
    HIP_TRY(hipGetLastError());
    return check_tile_flag(s);
}

int so3g_set_option(const char *name, int64_t value) {
    std::string n(name ? name : "");
    if (n == "scatter_agg") g_opt.scatter_agg = (int)(value != 0);
    else if (n == "scatter_xcd") g_opt.scatter_xcd = (int)(value != 0);
    else if (n == "pc_from_dpb") g_opt.pc_from_dpb = (int)std::max<int64_t>(1, value);
    else if (n == "scatter_hash") {
        if (value < 0 || value > 4) return fail(SO3G_ERR_VALUE, "scatter_hash must be 0..4");
        g_opt.scatter_hash = (int)value;
    } else if (n == "hash_r") {
        if (value != 1 && value != 4) return fail(SO3G_ERR_VALUE, "hash_r must be 1 or 4");
        g_opt.hash_r = (int)value;
    } else if (n == "hash_dets") {
        if (value < 0 || value > 4096)
            return fail(SO3G_ERR_VALUE, "hash_dets must be 0 (default) or 1..4096");
        g_opt.hash_dets = (int)value;
    } else if (n == "hash_span") {
        if (value != 0 && (value < 64 || value > (1 << 24)))
            return fail(SO3G_ERR_VALUE, "hash_span must be 0 (default) or 64..2^24");
        g_opt.hash_span = (int)value;
    }
    else if (n == "xcd_budget_mb") g_opt.xcd_budget_mb = value;
    else if (n == "plan_div") so3g::g_plan_div = (int)(value != 0);
    else if (n == "plan_from_layout") so3g::g_plan_from_layout = (int)(value != 0);
    else if (n == "plan_to_lds") so3g::g_plan_to_lds = (int)(value != 0);
    else if (n == "plan_layout") {
        if (value < -1 || value > 2)
            return fail(SO3G_ERR_VALUE, "plan_layout must be -1 (auto), 0, 1 or 2");
        so3g::g_plan_layout = (int)value;
    }
    else if (n == "plan_time_copy") so3g::g_plan_time_copy = (int)(value != 0);
    else if (n == "nmat_dchunk") {
        if (value < 1 || value > 65536) return fail(SO3G_ERR_VALUE, "nmat_dchunk must be 1..65536");
        so3g::g_nmat_dchunk = (int)value;
    }
    else if (n == "plan_tile_item") {
        if (value < 64 || value > (1 << 30))
            return fail(SO3G_ERR_VALUE, "plan_tile_item must be 64..2^30");
        so3g::g_plan_tile_item = value;
    }
    else if (n == "plan_win_block") {
        if (value != 256 && value != 1024)
            return fail(SO3G_ERR_VALUE, "plan_win_block must be 256 or 1024");
        so3g::g_plan_win_block = (int)value;
    }
    else if (n == "plan_win_flush") so3g::g_plan_win_flush = (int)(value != 0);
    else if (n == "plan_win_len") {
        if (value < 512 || value > (1 << 30))
            return fail(SO3G_ERR_VALUE, "plan_win_len must be 512..2^30");
        so3g::g_plan_win_len = value;
    }
    else if (n == "plan_to_cpb") {
        if (value < 1 || value > 64) return fail(SO3G_ERR_VALUE, "plan_to_cpb must be 1..64");
        so3g::g_plan_to_cpb = (int)value;
    }
    else if (n == "plan_from_v") {
        if (value != 4 && value != 8) return fail(SO3G_ERR_VALUE, "plan_from_v must be 4 or 8");
        so3g::g_plan_from_v = (int)value;
    } else return fail(SO3G_ERR_VALUE, "unknown option " + n);
    return SO3G_OK;
}

int so3g_get_option(const char *name, int64_t *value) {
    std::string n(name ? name : "");
    if (n == "scatter_agg") *value = g_opt.scatter_agg;
    else if (n == "scatter_xcd") *value = g_opt.scatter_xcd;
    else if (n == "pc_from_dpb") *value = g_opt.pc_from_dpb;
    else if (n == "scatter_hash") *value = g_opt.scatter_hash;
    else if (n == "hash_r") *value = g_opt.hash_r;
    else if (n == "hash_dets") *value = g_opt.hash_dets;
    else if (n == "hash_span") *value = g_opt.hash_span;
    else if (n == "xcd_budget_mb") *value = g_opt.xcd_budget_mb;
    else if (n == "plan_div") *value = so3g::g_plan_div;
    else if (n == "plan_from_layout") *value = so3g::g_plan_from_layout;
    else if (n == "plan_to_lds") *value = so3g::g_plan_to_lds;
    else if (n == "plan_to_cpb") *value = so3g::g_plan_to_cpb;
    else if (n == "plan_layout") *value = so3g::g_plan_layout;
    else if (n == "plan_tile_item") *value = so3g::g_plan_tile_item;
    else if (n == "plan_win_len") *value = so3g::g_plan_win_len;
    else if (n == "plan_win_block") *value = so3g::g_plan_win_block;
    else if (n == "plan_win_flush") *value = so3g::g_plan_win_flush;
    else if (n == "plan_time_copy") *value = so3g::g_plan_time_copy;
    else if (n == "nmat_dchunk") *value = so3g::g_nmat_dchunk;
    else if (n == "plan_tile_rows") *value = so3g::g_plan_tile_rows;
    else if (n == "plan_tile_cols") *value = so3g::g_plan_tile_cols;
    else if (n == "plan_from_v") *value = so3g::g_plan_from_v;
    else return fail(SO3G_ERR_VALUE, "unknown option " + n);
    return SO3G_OK;
}

int so3g_precomp_to_map(double *map, const int64_t *map_shape, double *const *tiles,
                        const int32_t *tile_dims, const int32_t *pixel, int32_t k,
                        const float *spin, int32_t nc, int64_t n_det, int64_t n_t,
                        const float *signal, const float *det_weights, const int32_t *segs,
                        int64_t n_seg, void *stream) {
    PcIn in;
    PcMap pm;
    int r = pc_setup(map_shape, tile_dims, k, nc, n_det, n_t, stream, in, pm);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    int64_t n_chunks = 0;
    if (segs != nullptr) {
        r = upload_chunks(segs, n_seg, n_det, n_t, s, &n_chunks);
        if (r) return r;
        if (n_chunks == 0) return SO3G_OK;
        in.chunks = g_state.chunks_d;
    }
    r = zero_flags(s);
    if (r) return r;
    if (pc_lds_applies(k, pm, segs, false))
        return pc_lds_scatter(false, s, in, pm, k, nc, map, tiles, pixel, spin, signal,
                              det_weights);
    {
        const dim3 grid = pc_grid(in, n_chunks);
        switch (k * 4 + nc) {
        case 5: r = pc_to_map_launch<1, 1>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 6: r = pc_to_map_launch<1, 2>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 7: r = pc_to_map_launch<1, 3>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 9: r = pc_to_map_launch<2, 1>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 10: r = pc_to_map_launch<2, 2>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 11: r = pc_to_map_launch<2, 3>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 13: r = pc_to_map_launch<3, 1>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 14: r = pc_to_map_launch<3, 2>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        case 15: r = pc_to_map_launch<3, 3>(grid, s, in, pm, map, tiles, pixel, spin, signal, det_weights); break;
        default: return fail(SO3G_ERR_RUNTIME, "unsupported pixel/spin layout");
        }
        if (r) return r;
    }
    return check_tile_flag(s);
}

int so3g_precomp_to_weights(double *map, const int64_t *map_shape, double *const *tiles,
                            const int32_t *tile_dims, const int32_t *pixel, int32_t k,
                            const float *spin, int32_t nc, int64_t n_det, int64_t n_t,
                            const float *det_weights, const int32_t *segs, int64_t n_seg,
                            void *stream) {
    PcIn in;
    PcMap pm;
    int r = pc_setup(map_shape, tile_dims, k, nc, n_det, n_t, stream, in, pm);
    if (r) return r;
    if (n_t == 0 || n_det == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    int64_t n_chunks = 0;
    if (segs != nullptr) {
        r = upload_chunks(segs, n_seg, n_det, n_t, s, &n_chunks);
        if (r) return r;
        if (n_chunks == 0) return SO3G_OK;
        in.chunks = g_state.chunks_d;
    }
    r = zero_flags(s);
    if (r) return r;
    if (pc_lds_applies(k, pm, segs, true))
        return pc_lds_scatter(true, s, in, pm, k, nc, map, tiles, pixel, spin, nullptr,
                              det_weights);
    PC_SWITCH(k, nc, k_pc_to_weights, pc_grid(in, n_chunks), dim3(kBlock), 0, s, in, pm,
              map, tiles, pixel, spin, det_weights, g_opt.scatter_agg, g_state.flags_d);
    HIP_TRY(hipGetLastError());
    return check_tile_flag(s);
}

}  // extern "C"
