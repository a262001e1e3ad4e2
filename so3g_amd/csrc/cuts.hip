// cuts.hip -- maximum-likelihood cut handling (src/array_ops.cxx:130-217,
// helpers :233-400): the cut part of the pointing matrix.  A RangesMatrix
// arrives as a host segment list (det, start, end) in the reference's walk
// order; the host turns it into a table with each segment's sample offset,
// value offset and polynomial count, and the kernels work from that table:
//
//   insert  (P: vals -> tod)   one thread per cut sample
//   clear                      one thread per cut sample
//   extract (P^T: tod -> vals) full model: one thread per cut sample;
//                              poly model: one thread per segment, summing
//                              its samples in order (the reference's float
//                              order, so results are bit-identical)
//   translate                  one thread per segment
//
// Arithmetic in T with the reference's expression order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/so3g_hip.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

namespace {

constexpr int kNT = 256;

struct CutSeg {
    int32_t det, a, b, np;     // detector, [a, b), polynomial count (poly model)
    int64_t soff, voff;        // first cut sample, first value
};

// array_ops.cxx:233-240
int npoly_host(int gap, int resolution, int nmax) {
    if (nmax < 1) nmax = 1;
    if (gap <= 1) return std::min(1, nmax);
    if (gap <= 3) return std::min(2, nmax);
    if (gap <= 6) return std::min(3, nmax);
    if (gap <= 20) return std::min(4, nmax);
    return std::min(5 + gap / resolution, nmax);
}

// Segment holding cut sample i: largest k with seg[k].soff <= i.
__device__ __forceinline__ int find_seg(const CutSeg *seg, int n, int64_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].soff <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// op 0 insert, 1 extract (full model only), 2 clear; model 0 full, 1 poly.
template <typename T, int OP, int MODEL>
__global__ __launch_bounds__(kNT) void k_cut_samples(const CutSeg *seg, int n_seg, int64_t total,
                                                     T *tod, int64_t nsamp, T *vals) {
    for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kNT) {
        const CutSeg &g = seg[find_seg(seg, n_seg, i)];
        const int s = g.a + (int)(i - g.soff);
        T *t = tod + (int64_t)g.det * nsamp + s;
        if constexpr (OP == 2) {
            *t = 0;
        } else if constexpr (MODEL == 0) {
            if constexpr (OP == 0) *t = vals[g.voff + (i - g.soff)];
            else vals[g.voff + (i - g.soff)] = *t;
        } else {   // poly insert, array_ops.cxx:326-350
            const T *v = vals + g.voff;
            if (g.np <= 1) {
                *t = v[0];
            } else {
                T x = -1 + 2 * (s - g.a) / T(g.b - g.a - 1);
                T acc = v[0];
                if (g.np > 1) acc += x * v[1];
                if (g.np > 2) {
                    T Pa = x, Pb = 1, Pc = 0;
                    for (int p = 2; p < g.np; p++) {
                        Pc = Pb; Pb = Pa; Pa = ((2 * p - 1) * x * Pb - (p - 1) * Pc) / p;
                        acc += Pa * v[p];
                    }
                }
                *t = acc;
            }
        }
    }
}

// poly extract, array_ops.cxx:287-321: one thread per segment, in order.
template <typename T>
__global__ __launch_bounds__(kNT) void k_cut_poly_extract(const CutSeg *seg, int n_seg,
                                                          const T *tod, int64_t nsamp, T *vals) {
    const int k = blockIdx.x * kNT + threadIdx.x;
    if (k >= n_seg) return;
    const CutSeg g = seg[k];
    const T *row = tod + (int64_t)g.det * nsamp;
    T *v = vals + g.voff;
    if (g.np <= 1) {
        for (int s = g.a; s < g.b; s++) v[0] += row[s];
        return;
    }
    for (int p = 0; p < g.np; p++) v[p] = 0;
    for (int s = g.a; s < g.b; s++) {
        T x = -1 + 2 * (s - g.a) / T(g.b - g.a - 1);
        T t = row[s];
        v[0] += t;
        if (g.np > 1) v[1] += t * x;
        if (g.np > 2) {
            T Pa = x, Pb = 1, Pc = 0;
            for (int p = 2; p < g.np; p++) {
                Pc = Pb; Pb = Pa; Pa = ((2 * p - 1) * x * Pb - (p - 1) * Pc) / p;
                v[p] += t * Pa;
            }
        }
    }
}

// translate, array_ops.cxx:243-277 (full) and :364-383 (poly).
struct CutPair {
    int32_t i1, i2, o1, o2;
    int64_t ioff, ooff;
    int32_t inp, onp;
};

template <typename T, int MODEL>
__global__ __launch_bounds__(kNT) void k_cut_translate(const CutPair *pr, int n_seg,
                                                       int64_t insamp, int64_t onsamp,
                                                       const T *ivals, T *ovals) {
    const int k = blockIdx.x * kNT + threadIdx.x;
    if (k >= n_seg) return;
    const CutPair g = pr[k];
    if constexpr (MODEL == 1) {
        int i;
        for (i = 0; i < min(g.inp, g.onp); i++) ovals[g.ooff + i] = ivals[g.ioff + i];
        for (; i < g.onp; i++) ovals[g.ooff + i] = 0;
    } else if (onsamp >= insamp) {
        for (int64_t os = g.o1, oi = g.ooff; os < g.o2; os++, oi++)
            ovals[oi] = ivals[g.ioff + os * insamp / onsamp - g.i1];
    } else {
        // Accumulate, then divide each output sample by its number of
        // inputs (0 for a sample no input reaches, as in the reference).
        // The output index is monotone in the input index, so each output
        // sample's inputs are consecutive and it is finished (divided) as
        // soon as the next output sample starts; the per-sample sums are
        // the reference's, in its order.
        int64_t cur = g.o1;
        int hits = 0;
        for (int64_t is = g.i1, ii = g.ioff; is < g.i2; is++, ii++) {
            const int64_t os = is * onsamp / insamp;
            while (cur < os) {
                ovals[g.ooff + cur - g.o1] /= hits;
                hits = 0;
                cur++;
            }
            ovals[g.ooff + os - g.o1] += ivals[ii];
            hits++;
        }
        for (; cur < g.o2; cur++) {
            ovals[g.ooff + cur - g.o1] /= hits;
            hits = 0;
        }
    }
}

#define CUT_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                               hipGetErrorString(_e));                  \
    } while (0)

unsigned grid_of(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kNT - 1) / kNT, 1 << 20));
}

template <typename T>
int run_process(const std::vector<CutSeg> &tab, int64_t total, int op, int model, T *tod,
                int64_t nsamp, T *vals, hipStream_t st) {
    const int n = (int)tab.size();
    CutSeg *d = nullptr;
    CUT_TRY(hipMallocAsync((void **)&d, n * sizeof(CutSeg), st));
    if (hipMemcpyAsync(d, tab.data(), n * sizeof(CutSeg), hipMemcpyHostToDevice, st) !=
        hipSuccess) {
        (void)hipFreeAsync(d, st);
        return set_error(SO3G_ERR_HIP, "cannot upload the cut table");
    }
    const dim3 g(grid_of(total)), b(kNT);
    if (op == 2) hipLaunchKernelGGL((k_cut_samples<T, 2, 0>), g, b, 0, st, d, n, total, tod, nsamp, vals);
    else if (op == 0 && model == 0) hipLaunchKernelGGL((k_cut_samples<T, 0, 0>), g, b, 0, st, d, n, total, tod, nsamp, vals);
    else if (op == 0) hipLaunchKernelGGL((k_cut_samples<T, 0, 1>), g, b, 0, st, d, n, total, tod, nsamp, vals);
    else if (model == 0) hipLaunchKernelGGL((k_cut_samples<T, 1, 0>), g, b, 0, st, d, n, total, tod, nsamp, vals);
    else hipLaunchKernelGGL((k_cut_poly_extract<T>), dim3(grid_of(n)), b, 0, st, d, n, tod, nsamp, vals);
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_error(SO3G_ERR_HIP, std::string("cut kernel: ") + hipGetErrorString(e));
    return SO3G_OK;
}

}  // namespace

extern "C" {

int so3g_cuts_measure(const int32_t *segs, int64_t n_seg, int model, int resolution, int nmax,
                      int64_t *n_vals) {
    if (n_vals == nullptr || (n_seg > 0 && segs == nullptr))
        return set_error(SO3G_ERR_VALUE, "NULL argument");
    if (model != 0 && model != 1)
        return set_error(SO3G_ERR_VALUE, "process_cuts model can only be 'full' or 'poly'");
    if (model == 1 && resolution <= 0)
        return set_error(SO3G_ERR_VALUE, "poly cuts need resolution > 0");
    int64_t n = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        const int gap = segs[3 * k + 2] - segs[3 * k + 1];
        n += model == 0 ? gap : npoly_host(gap, resolution, nmax);
    }
    *n_vals = n;
    return SO3G_OK;
}

int so3g_process_cuts(const int32_t *segs, int64_t n_seg, int op, int model, int resolution,
                      int nmax, void *tod, int64_t ndet, int64_t nsamp, void *vals,
                      int64_t n_vals, int f64, void *stream) {
    if (op < 0 || op > 2)
        return set_error(SO3G_ERR_VALUE,
                         "process_cuts operation can only be 'measure', 'insert' or 'extract'");
    int64_t need = 0;
    int r = so3g_cuts_measure(segs, n_seg, model, resolution, nmax, &need);
    if (r) return r;
    if (n_seg > INT32_MAX) return set_error(SO3G_ERR_RUNTIME, "too many cut ranges");
    std::vector<CutSeg> tab;
    tab.reserve(n_seg);
    int64_t soff = 0, voff = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        CutSeg g;
        g.det = segs[3 * k];
        g.a = segs[3 * k + 1];
        g.b = segs[3 * k + 2];
        if (g.det < 0 || g.det >= ndet || g.a < 0 || g.b > nsamp || g.a > g.b)
            return set_error(SO3G_ERR_RUNTIME, "cut range outside the tod");
        g.np = model == 1 ? npoly_host(g.b - g.a, resolution, nmax) : 0;
        g.soff = soff;
        g.voff = voff;
        soff += g.b - g.a;
        voff += model == 0 ? g.b - g.a : g.np;
        if (g.b > g.a) tab.push_back(g);
    }
    if (op != 2 && n_vals < need)
        return set_error(SO3G_ERR_RUNTIME, "vals is shorter than the cut model (" +
                                               std::to_string(need) + " values)");
    if (tab.empty()) return SO3G_OK;
    hipStream_t st = (hipStream_t)stream;
    if (f64)
        return run_process<double>(tab, soff, op, model, (double *)tod, nsamp, (double *)vals, st);
    return run_process<float>(tab, soff, op, model, (float *)tod, nsamp, (float *)vals, st);
}

int so3g_translate_cuts(const int32_t *isegs, const int32_t *osegs, int64_t n_seg, int model,
                        int resolution, int nmax, int64_t insamp, int64_t onsamp,
                        const void *ivals, int64_t n_ivals, void *ovals, int64_t n_ovals,
                        int f64, void *stream) {
    int64_t ni = 0, no = 0;
    int r = so3g_cuts_measure(isegs, n_seg, model, resolution, nmax, &ni);
    if (!r) r = so3g_cuts_measure(osegs, n_seg, model, resolution, nmax, &no);
    if (r) return r;
    if (n_ivals < ni || n_ovals < no)
        return set_error(SO3G_ERR_RUNTIME, "vals shorter than the cut model");
    if (insamp <= 0 || onsamp <= 0) return set_error(SO3G_ERR_RUNTIME, "sample counts must be > 0");
    std::vector<CutPair> tab(n_seg);
    int64_t ioff = 0, ooff = 0;
    for (int64_t k = 0; k < n_seg; k++) {
        CutPair &g = tab[k];
        g.i1 = isegs[3 * k + 1]; g.i2 = isegs[3 * k + 2];
        g.o1 = osegs[3 * k + 1]; g.o2 = osegs[3 * k + 2];
        if (isegs[3 * k] != osegs[3 * k])
            return set_error(SO3G_ERR_RUNTIME, "input and output cuts must pair up range by range");
        if (model == 0 && g.i2 > g.i1 && g.o2 > g.o1) {
            // every index the reference's loops form must stay in the pair
            const bool up = onsamp >= insamp;
            const bool ok = up ? ((int64_t)g.o1 * insamp / onsamp >= g.i1 &&
                                  ((int64_t)g.o2 - 1) * insamp / onsamp < g.i2)
                               : ((int64_t)g.i1 * onsamp / insamp >= g.o1 &&
                                  ((int64_t)g.i2 - 1) * onsamp / insamp < g.o2);
            if (!ok)
                return set_error(SO3G_ERR_RUNTIME,
                                 "input and output cut ranges do not correspond");
        }
        g.inp = model == 1 ? npoly_host(g.i2 - g.i1, resolution, nmax) : 0;
        g.onp = model == 1 ? npoly_host(g.o2 - g.o1, resolution, nmax) : 0;
        g.ioff = ioff;
        g.ooff = ooff;
        ioff += model == 1 ? g.inp : g.i2 - g.i1;
        ooff += model == 1 ? g.onp : g.o2 - g.o1;
    }
    if (n_seg == 0) return SO3G_OK;
    hipStream_t st = (hipStream_t)stream;
    CutPair *d = nullptr;
    CUT_TRY(hipMallocAsync((void **)&d, n_seg * sizeof(CutPair), st));
    if (hipMemcpyAsync(d, tab.data(), n_seg * sizeof(CutPair), hipMemcpyHostToDevice, st) !=
        hipSuccess) {
        (void)hipFreeAsync(d, st);
        return set_error(SO3G_ERR_HIP, "cannot upload the cut table");
    }
    const dim3 g(grid_of(n_seg)), b(kNT);
    const int n = (int)n_seg;
    if (f64) {
        if (model == 1) hipLaunchKernelGGL((k_cut_translate<double, 1>), g, b, 0, st, d, n, insamp, onsamp, (const double *)ivals, (double *)ovals);
        else hipLaunchKernelGGL((k_cut_translate<double, 0>), g, b, 0, st, d, n, insamp, onsamp, (const double *)ivals, (double *)ovals);
    } else {
        if (model == 1) hipLaunchKernelGGL((k_cut_translate<float, 1>), g, b, 0, st, d, n, insamp, onsamp, (const float *)ivals, (float *)ovals);
        else hipLaunchKernelGGL((k_cut_translate<float, 0>), g, b, 0, st, d, n, insamp, onsamp, (const float *)ivals, (float *)ovals);
    }
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(d, st);
    if (e != hipSuccess) return set_error(SO3G_ERR_HIP, std::string("cut kernel: ") + hipGetErrorString(e));
    return SO3G_OK;
}

}  // extern "C"
