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
    int slots = 0;                     // LDS table of to_map (0: global atomics)
    bool lds_ok = false;               // every map element has an int32 key
    int32_t *order_d = nullptr;        // per observation: detectors in row order
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
    std::vector<int64_t> order_off(n_obs, 0);
    b->lds_ok = hash_keys_fit(g);
    if (b->lds_ok) {
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
        }
        (void)hipFree(co_d);
        if (r) {
            so3g_batch_destroy(b);
            return r;
        }
        if (total > 0 && hipMemcpy(b->order_d, idx.data(), total * sizeof(int32_t),
                                   hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(SO3G_ERR_HIP, "cannot upload the detector order");
    }
    for (int dpb : kDpbCand) {
        std::vector<ObsDev> tab(n_obs), one(n_obs);
        std::vector<int64_t> items(n_obs);
        int64_t item0 = 0;
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
            items[i] = d.n_tt * ((o.n_det + dpb - 1) / dpb);
            d.item0 = 0;
            one[i] = d;
            d.item0 = item0;
            tab[i] = d;
            item0 += items[i];
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
    if (slots != 0 && slots != 512 && slots != 1024)
        return set_error(SO3G_ERR_VALUE, "table slots must be 0, 512 or 1024");
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
    for (int slots : kSlots) {
        if (slots > 0 && (op == 0 || !b->lds_ok)) continue;
        for (int dpb : kDpbCand) {
            for (int gm : kGridMult) {
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
                    results[4 * n] = dpb;
                    results[4 * n + 1] = gm;
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
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return SO3G_OK;
}

}  // extern "C"
