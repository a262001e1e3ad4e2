// batch.hip -- C ABI of the batched projection: many observations, one
// map, one launch (optionally replayed from a captured HIP graph), with a
// launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/so3g_hip.h"
#include "launch.h"

namespace so3g {
int set_error(int code, const std::string &msg);
int op_args(const so3g_geom *g, void *stream, OpArgs &a);
int launch(int proj, const OpArgs &a);
int flags_zero(hipStream_t stream);
int flags_check_tile(hipStream_t stream);
void row_order(const so3g_geom *g, const double *co, int64_t n_det, int32_t *idx);
bool hash_keys_fit(const so3g_geom *g);
void make_row_groups(const so3g_geom *g, const double *co, int64_t n_det, int G,
                     std::vector<int32_t> &gdet, std::vector<int32_t> &grow,
                     std::vector<int32_t> &gxc);
int row_width_max(const so3g_geom *g, int nv);
}  // namespace so3g

using namespace so3g;

#define BT_TRY(expr)                                                                  \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess)                                                         \
            return set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                               hipGetErrorString(_e));                \
    } while (0)

namespace {

constexpr int kDpbCand[] = {1, 2, 4, 8, 16};
constexpr int kGridMult[] = {2, 4, 8, 16, 0};   // blocks per CU; 0 = one block per item
constexpr int kSlots[] = {0, 512, 1024};         // LDS table of to_map; 0 = global atomics
// row segments of to_map; 0 = widest, -1 = the widest of which two fit the
// LDS of a compute unit together (two blocks per compute unit)
constexpr int kRowWidth[] = {1024, 2048, -1, 4096, 0};
constexpr int kRowG = 2;                         // detectors per row group
constexpr int kRowSpan = 65536;                  // samples per row item

struct GraphKey {
    int op = -1;
    const void *map = nullptr, *tiles = nullptr, *dims = nullptr;
    int dpb = 0, grid = 0, acc = 0, slots = 0;
    bool operator==(const GraphKey &o) const {
        return op == o.op && map == o.map && tiles == o.tiles && dims == o.dims &&
               dpb == o.dpb && grid == o.grid && acc == o.acc && slots == o.slots;
    }
};

}  // namespace

struct so3g_batch {
    so3g_geom g;
    std::vector<so3g_obs> obs;
    std::map<int, ObsDev *> table;     // per dpb: item0 numbered across the batch
    std::map<int, ObsDev *> single;    // per dpb: item0 = 0 (one launch per observation)
    std::map<int, std::vector<int64_t>> items;   // per dpb: items of each observation
    std::map<int, int64_t> n_items;
    int n_cu = 256;
    int dpb[2] = {4, 4};               // chosen shape, [0] from_map, [1] to_map
    int grid[2] = {0, 0};              // blocks (0: one per item)
    int slots = 0;                     // to_map: 0 global atomics, 512 / 1024 LDS
                                       // table slots, -w row segments of w columns
    bool lds_ok = false;               // every map element has an int32 key
    bool row_ok = false;               // flat map, nearest neighbour: row segments
    int32_t *order_d = nullptr;        // per observation: detectors in row order
    int32_t *rg_d = nullptr;           // per observation: row groups, rows, columns
    std::vector<int64_t> row_items;    // row items of each observation
    int64_t n_row_items = 0;
    int64_t samples = 0;
    hipStream_t cap = nullptr;         // capture stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    GraphKey key;
};

namespace {

void drop_graph(so3g_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key = GraphKey{};
}

int blocks_for(const so3g_batch *b, int gm, int64_t items) {
    int64_t n = gm > 0 ? (int64_t)gm * b->n_cu : items;
    n = std::min<int64_t>(std::min<int64_t>(n, items), INT32_MAX);
    return (int)std::max<int64_t>(n, 1);
}

// Queues the batch kernel(s) on stream s: one persistent launch over the
// whole batch, or (per_obs) one launch per observation.
int enqueue(so3g_batch *b, int op, void *map, void *const *tiles, const int32_t *dims,
            int accumulate, int dpb, int gm, int slots, int per_obs, hipStream_t s) {
    OpArgs a;
    int r = op_args(&b->g, s, a);
    if (r) return r;
    a.op = op ? OP_BATCH_TO_MAP : OP_BATCH_FROM_MAP;
    a.batch_slots = op ? slots : 0;
    a.in.dpb = dpb;
    a.map = map;
    a.tiles = tiles;
    a.tile_dims = dims;
    a.accumulate = accumulate;
    if (a.batch_slots < 0) {
        // row segments: one block per (observation, span, group) item
        a.rw.G = kRowG;
        a.rw.span = kRowSpan;
        a.rw.xw = (int32_t)std::min<int64_t>(-slots, b->g.naxis[1]);
        if (!per_obs) {
            if (b->n_row_items == 0) return SO3G_OK;
            a.obs = b->table[dpb];
            a.n_obs = (int)b->obs.size();
            a.grid_blocks = (int)b->n_row_items;
            return launch(b->g.proj, a);
        }
        for (size_t i = 0; i < b->obs.size(); i++) {
            if (b->row_items[i] == 0) continue;
            a.obs = b->single[dpb] + i;
            a.n_obs = 1;
            a.grid_blocks = (int)b->row_items[i];
            r = launch(b->g.proj, a);
            if (r) return r;
        }
        return SO3G_OK;
    }
    if (!per_obs) {
        const int64_t n = b->n_items[dpb];
        if (n == 0) return SO3G_OK;
        a.obs = b->table[dpb];
        a.n_obs = (int)b->obs.size();
        a.n_items = n;
        a.grid_blocks = blocks_for(b, gm, n);
        return launch(b->g.proj, a);
    }
    const std::vector<int64_t> &it = b->items[dpb];
    for (size_t i = 0; i < b->obs.size(); i++) {
        if (it[i] == 0) continue;
        a.obs = b->single[dpb] + i;
        a.n_obs = 1;
        a.n_items = it[i];
        a.grid_blocks = blocks_for(b, gm, it[i]);
        r = launch(b->g.proj, a);
        if (r) return r;
    }
    return SO3G_OK;
}

int check_map(const so3g_batch *b, const void *map, void *const *tiles, const int32_t *dims) {
    if (b->g.tiled ? (tiles == nullptr || dims == nullptr) : map == nullptr)
        return set_error(SO3G_ERR_RUNTIME, b->g.tiled ? "tiled map requires tile pointers"
                                                      : "map is NULL");
    return SO3G_OK;
}

int run_op(so3g_batch *b, int op, void *map, void *const *tiles, const int32_t *dims,
           int accumulate, int mode, void *stream) {
    if (b == nullptr) return set_error(SO3G_ERR_VALUE, "batch is NULL");
    int r = check_map(b, map, tiles, dims);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    const int dpb = b->dpb[op], gm = b->grid[op], slots = op ? b->slots : 0;
    if (mode == SO3G_BATCH_GRAPH) {
        GraphKey k;
        k.op = op;
        k.map = map;
        k.tiles = tiles;
        k.dims = dims;
        k.dpb = dpb;
        k.grid = gm;
        k.acc = accumulate;
        k.slots = slots;
        if (!(b->exec && b->key == k)) {
            drop_graph(b);
            BT_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
            r = b->g.tiled ? flags_zero(b->cap) : SO3G_OK;
            if (!r) r = enqueue(b, op, map, tiles, dims, accumulate, dpb, gm, slots, 0, b->cap);
            hipGraph_t gr = nullptr;
            const hipError_t e = hipStreamEndCapture(b->cap, &gr);
            if (r) {
                if (gr) (void)hipGraphDestroy(gr);
                return r;
            }
            if (e != hipSuccess)
                return set_error(SO3G_ERR_HIP, std::string("graph capture failed: ") +
                                                   hipGetErrorString(e));
            b->graph = gr;
            BT_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->key = k;
        }
        BT_TRY(hipGraphLaunch(b->exec, s));
    } else {
        if (b->g.tiled) {
            r = flags_zero(s);
            if (r) return r;
        }
        r = enqueue(b, op, map, tiles, dims, accumulate, dpb, gm, slots,
                    mode == SO3G_BATCH_PER_OBS, s);
        if (r) return r;
    }
    return b->g.tiled ? flags_check_tile(s) : SO3G_OK;
}

}  // namespace

extern "C" {

int so3g_batch_create(const so3g_geom *g, const so3g_obs *obs, int64_t n_obs,
                      so3g_batch **out) {
    if (out == nullptr) return set_error(SO3G_ERR_VALUE, "out is NULL");
    *out = nullptr;
    if (g == nullptr || g->index_count <= 0)
        return set_error(SO3G_ERR_VALUE, "geometry was not initialised (call so3g_geom_init)");
    if (n_obs < 0 || n_obs > INT32_MAX || (n_obs > 0 && obs == nullptr))
        return set_error(SO3G_ERR_VALUE, "bad observation list");
    for (int64_t i = 0; i < n_obs; i++) {
        const so3g_obs &o = obs[i];
        if (o.n_t < 0 || o.n_det < 0 || o.n_t > INT32_MAX)
            return set_error(SO3G_ERR_RUNTIME, "observation " + std::to_string(i) +
                                                   ": bad sample or detector count");
        if (o.n_t > 0 && o.n_det > 0 && (o.bore == nullptr || o.dets == nullptr ||
                                         o.resp == nullptr))
            return set_error(SO3G_ERR_VALUE, "observation " + std::to_string(i) +
                                                 ": missing pointing or response");
    }
    OpArgs probe;
    int r = op_args(g, nullptr, probe);   // initialises the library, checks the geometry
    if (r) return r;
    so3g_batch *b = new so3g_batch();
    b->g = *g;
    b->obs.assign(obs, obs + n_obs);
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    auto cleanup = [&](int code, const char *msg) {
        so3g_batch_destroy(b);
        return set_error(code, msg);
    };
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
        return cleanup(SO3G_ERR_HIP, "cannot create the capture stream");
    // Detector order of each observation for the LDS tables: the map row
    // (then column) of each detector at the observation's middle sample.
    // Row groups of each observation for the row segments (flat maps,
    // nearest neighbour), from the same coordinates.
    std::vector<int64_t> order_off(n_obs, 0);
    b->lds_ok = hash_keys_fit(g);
    b->row_ok = g->pixelizor == SO3G_PIX_FLAT && g->interp == 0;
    std::vector<int32_t> rg_host;
    std::vector<int64_t> rg_off(n_obs, 0), rg_n(n_obs, 0);
    b->row_items.assign(n_obs, 0);
    if (b->lds_ok || b->row_ok) {
        int64_t total = 0, most = 1;
        for (int64_t i = 0; i < n_obs; i++) {
            order_off[i] = total;
            total += obs[i].n_det;
            most = std::max<int64_t>(most, obs[i].n_det);
        }
        if (hipMalloc(&b->order_d, std::max<int64_t>(total, 1) * sizeof(int32_t)) != hipSuccess)
            return cleanup(SO3G_ERR_HIP, "out of memory");
        double *co_d = nullptr;
        if (hipMalloc(&co_d, most * 4 * sizeof(double)) != hipSuccess)
            return cleanup(SO3G_ERR_HIP, "out of memory");
        std::vector<double> co(most * 4);
        std::vector<int32_t> idx(std::max<int64_t>(total, 1));
        for (int64_t i = 0; i < n_obs && r == SO3G_OK; i++) {
            const so3g_obs &o = obs[i];
            if (o.n_t == 0 || o.n_det == 0) continue;
            OpArgs a;
            r = op_args(g, nullptr, a);
            if (r) break;
            a.op = OP_COORDS;
            a.in.bore = o.bore + 4 * (o.n_t / 2);
            a.in.n_t = 1;
            a.in.dets = o.dets;
            a.in.n_det = o.n_det;
            a.in.dpb = 4;
            a.in.chunks = nullptr;
            a.coords_out = co_d;
            r = launch(g->proj, a);
            if (!r && hipMemcpy(co.data(), co_d, o.n_det * 4 * sizeof(double),
                                hipMemcpyDeviceToHost) != hipSuccess)
                r = set_error(SO3G_ERR_HIP, "cannot read detector coordinates");
            if (!r) row_order(g, co.data(), o.n_det, idx.data() + order_off[i]);
            if (!r && b->row_ok) {
                std::vector<int32_t> gdet, grow, gxc;
                make_row_groups(g, co.data(), o.n_det, kRowG, gdet, grow, gxc);
                rg_off[i] = (int64_t)rg_host.size();
                rg_n[i] = (int64_t)grow.size();
                rg_host.insert(rg_host.end(), gdet.begin(), gdet.end());
                rg_host.insert(rg_host.end(), grow.begin(), grow.end());
                rg_host.insert(rg_host.end(), gxc.begin(), gxc.end());
                b->row_items[i] = rg_n[i] * ((o.n_t + kRowSpan - 1) / kRowSpan);
                b->n_row_items += b->row_items[i];
            }
        }
        (void)hipFree(co_d);
        if (r) {
            so3g_batch_destroy(b);
            return r;
        }
        if (total > 0 && hipMemcpy(b->order_d, idx.data(), total * sizeof(int32_t),
                                   hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(SO3G_ERR_HIP, "cannot upload the detector order");
        if (b->n_row_items > INT32_MAX) b->row_ok = false;   // one block per item
        if (b->row_ok) {
            if (hipMalloc(&b->rg_d, std::max<size_t>(rg_host.size(), 1) * sizeof(int32_t)) !=
                hipSuccess)
                return cleanup(SO3G_ERR_HIP, "out of memory");
            if (!rg_host.empty() &&
                hipMemcpy(b->rg_d, rg_host.data(), rg_host.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(SO3G_ERR_HIP, "cannot upload the row groups");
        }
    }
    for (int dpb : kDpbCand) {
        std::vector<ObsDev> tab(n_obs), one(n_obs);
        std::vector<int64_t> items(n_obs);
        int64_t item0 = 0, ritem0 = 0;
        for (int64_t i = 0; i < n_obs; i++) {
            const so3g_obs &o = obs[i];
            ObsDev d{};
            d.bore = o.bore;
            d.dets = o.dets;
            d.resp = o.resp;
            d.sig_in = o.signal;
            d.sig_out = o.signal_out;
            d.det_w = o.det_weights;
            d.n_t = o.n_t;
            d.n_det = o.n_det;
            d.n_tt = (o.n_t + kBlock - 1) / kBlock;
            d.order = b->lds_ok ? b->order_d + order_off[i] : nullptr;
            if (b->row_ok) {
                const int64_t ng = rg_n[i];
                d.rgdet = b->rg_d + rg_off[i];
                d.rgrow = d.rgdet + ng * kRowG;
                d.rgxc = d.rgrow + ng;
                d.n_groups = ng;
            }
            items[i] = d.n_tt * ((o.n_det + dpb - 1) / dpb);
            d.item0 = 0;
            d.ritem0 = 0;
            one[i] = d;
            d.item0 = item0;
            d.ritem0 = ritem0;
            tab[i] = d;
            item0 += items[i];
            ritem0 += b->row_items[i];
        }
        ObsDev *t_d = nullptr, *o_d = nullptr;
        const size_t bytes = std::max<size_t>(1, n_obs) * sizeof(ObsDev);
        if (hipMalloc(&t_d, bytes) != hipSuccess) return cleanup(SO3G_ERR_HIP, "out of memory");
        b->table[dpb] = t_d;
        if (hipMalloc(&o_d, bytes) != hipSuccess) return cleanup(SO3G_ERR_HIP, "out of memory");
        b->single[dpb] = o_d;
        if (n_obs > 0 &&
            (hipMemcpy(t_d, tab.data(), n_obs * sizeof(ObsDev), hipMemcpyHostToDevice) !=
                 hipSuccess ||
             hipMemcpy(o_d, one.data(), n_obs * sizeof(ObsDev), hipMemcpyHostToDevice) !=
                 hipSuccess))
            return cleanup(SO3G_ERR_HIP, "cannot upload the observation table");
        b->items[dpb] = items;
        b->n_items[dpb] = item0;
    }
    for (int64_t i = 0; i < n_obs; i++) b->samples += obs[i].n_t * obs[i].n_det;
    *out = b;
    return SO3G_OK;
}

int so3g_batch_to_map(so3g_batch *b, double *map, double *const *tiles,
                      const int32_t *tile_dims, int mode, void *stream) {
    return run_op(b, 1, map, (void *const *)tiles, tile_dims, 0, mode, stream);
}

int so3g_batch_from_map(so3g_batch *b, const double *map, const double *const *tiles,
                        const int32_t *tile_dims, int accumulate, int mode, void *stream) {
    return run_op(b, 0, (void *)map, (void *const *)tiles, tile_dims, accumulate, mode,
                  stream);
}

int so3g_batch_set_shape(so3g_batch *b, int op, int dpb, int grid_mult) {
    if (b == nullptr) return set_error(SO3G_ERR_VALUE, "batch is NULL");
    if (op != 0 && op != 1) return set_error(SO3G_ERR_VALUE, "op must be 0 (from) or 1 (to)");
    if (b->table.count(dpb) == 0)
        return set_error(SO3G_ERR_VALUE, "detectors per block must be 1, 2, 4, 8 or 16");
    if (grid_mult < 0) return set_error(SO3G_ERR_VALUE, "grid multiplier must be >= 0");
    b->dpb[op] = dpb;
    b->grid[op] = grid_mult;
    return SO3G_OK;
}

int so3g_batch_set_table(so3g_batch *b, int slots) {
    if (b == nullptr) return set_error(SO3G_ERR_VALUE, "batch is NULL");
    if (slots < 0) {
        if (!b->row_ok)
            return set_error(SO3G_ERR_VALUE,
                             "row segments need a flat map with nearest-neighbour pixels");
        const int wmax = row_width_max(&b->g, b->g.ncomp);
        if (-slots < 64 || -slots > wmax)
            return set_error(SO3G_ERR_VALUE, "row segment width must be 64.." +
                                                 std::to_string(wmax) + " columns");
        b->slots = slots;
        return SO3G_OK;
    }
    if (slots != 0 && slots != 512 && slots != 1024)
        return set_error(SO3G_ERR_VALUE, "table slots must be 0, 512 or 1024, or -w");
    if (slots > 0 && !b->lds_ok)
        return set_error(SO3G_ERR_VALUE, "map elements do not fit int32 table keys");
    b->slots = slots;
    return SO3G_OK;
}

int so3g_batch_tune(so3g_batch *b, int op, double *map, double *const *tiles,
                    const int32_t *tile_dims, int reps, double *results, int max_results,
                    int *n_results, void *stream) {
    if (b == nullptr) return set_error(SO3G_ERR_VALUE, "batch is NULL");
    if (op != 0 && op != 1) return set_error(SO3G_ERR_VALUE, "op must be 0 (from) or 1 (to)");
    int r = check_map(b, map, (void *const *)tiles, tile_dims);
    if (r) return r;
    hipStream_t s = (hipStream_t)stream;
    if (b->g.tiled) {
        r = flags_zero(s);
        if (r) return r;
    }
    hipEvent_t e0, e1;
    BT_TRY(hipEventCreate(&e0));
    BT_TRY(hipEventCreate(&e1));
    double best = -1.;
    int n = 0;
    reps = std::max(reps, 1);
    // candidate to_map scatters: the LDS tables, then row segments of each
    // width (their launch shape is fixed: one block of 1024 lanes per item)
    std::vector<int> cand(std::begin(kSlots), std::end(kSlots));
    if (op == 1 && b->row_ok) {
        const int wmax = row_width_max(&b->g, b->g.ncomp);
        std::vector<int> ws;
        const int w2 = (int)(80 * 1024 / (b->g.ncomp * 8 + (b->g.tiled ? 1 : 0))) / 64 * 64;
        for (int w : kRowWidth) {
            int x = w > 0 ? std::min(w, wmax) : w < 0 ? std::min(w2, wmax) : wmax;
            x = (int)std::min<int64_t>(x, b->g.naxis[1]);
            x = std::max(x, 64);
            if (std::find(ws.begin(), ws.end(), x) == ws.end()) ws.push_back(x);
        }
        for (int x : ws) cand.push_back(-x);
    }
    for (int slots : cand) {
        if (slots > 0 && (op == 0 || !b->lds_ok)) continue;
        for (int dpb : kDpbCand) {
            for (int gm : kGridMult) {
                if (slots < 0 && (dpb != kDpbCand[0] || gm != kGridMult[0])) continue;
                // warm-up, then reps timed launches
                r = enqueue(b, op, map, (void *const *)tiles, tile_dims, 0, dpb, gm, slots, 0,
                            s);
                if (!r) r = (hipEventRecord(e0, s) == hipSuccess) ? 0 : SO3G_ERR_HIP;
                for (int i = 0; !r && i < reps; i++)
                    r = enqueue(b, op, map, (void *const *)tiles, tile_dims, 0, dpb, gm, slots,
                                0, s);
                if (!r) r = (hipEventRecord(e1, s) == hipSuccess) ? 0 : SO3G_ERR_HIP;
                if (!r) r = (hipEventSynchronize(e1) == hipSuccess) ? 0 : SO3G_ERR_HIP;
                float ms = 0.f;
                if (!r) r = (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) ? 0 : SO3G_ERR_HIP;
                if (r) {
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    return r == SO3G_ERR_HIP ? set_error(r, "timing failed during tuning") : r;
                }
                const double t = ms / reps;
                if (results != nullptr && n < max_results) {
                    results[4 * n] = slots < 0 ? kRowG : dpb;
                    results[4 * n + 1] = slots < 0 ? 0 : gm;
                    results[4 * n + 2] = slots;
                    results[4 * n + 3] = t;
                }
                n++;
                if (best < 0 || t < best) {
                    best = t;
                    b->dpb[op] = dpb;
                    b->grid[op] = gm;
                    if (op == 1) b->slots = slots;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (n_results) *n_results = n;
    if (b->g.tiled) return flags_check_tile(s);
    return SO3G_OK;
}

int so3g_batch_info(const so3g_batch *b, int64_t *info) {
    if (b == nullptr || info == nullptr) return set_error(SO3G_ERR_VALUE, "NULL argument");
    info[0] = (int64_t)b->obs.size();
    info[1] = b->samples;
    info[2] = b->dpb[0];
    info[3] = b->grid[0];
    info[4] = b->dpb[1];
    info[5] = b->grid[1];
    info[6] = b->n_cu;
    info[7] = b->slots;
    return SO3G_OK;
}

int so3g_batch_destroy(so3g_batch *b) {
    if (b == nullptr) return SO3G_OK;
    drop_graph(b);
    for (auto &kv : b->table) (void)hipFree(kv.second);
    for (auto &kv : b->single) (void)hipFree(kv.second);
    if (b->order_d) (void)hipFree(b->order_d);
    if (b->rg_d) (void)hipFree(b->rg_d);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return SO3G_OK;
}

}  // extern "C"
