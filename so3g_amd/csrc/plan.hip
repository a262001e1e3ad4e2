// plan.hip -- pixel-sorted projection plans for repeated to_map calls with a
// fixed pointing matrix (the CG hot loop).
//
// ProjEng_Precomp's to_map scatters every sample's float32 contributions into
// the float64 map with atomics.  Over a whole observation each map pixel is
// hit by hundreds of detector passes, but passes that share a pixel are far
// apart in time, so no time-local block sees much of that reuse.  A plan
// sorts the on-map samples by pixel once (rocPRIM radix sort, stable, so each
// detector pass stays contiguous) and stores the spin factors in that order.
// to_map then streams (pixel, sample index, spin) in sorted order, gathers
// the signal through the sample index, and sums each pixel's contributions
// with a wave-level segmented scan: one atomic per pixel per wave instead of
// one per sample.
//
// Arithmetic per sample is the Precomp one, (det_wt * sig) * pf in float32
// (Projection.cxx:2117), summed in float64.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <string>
#include <type_traits>

#include "../../include/so3g_hip.h"
#include "kernels.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

struct so3g_pc_plan {
    int64_t n_det, n_t, ny, nx;
    int32_t nc;
    int64_t n_on;          // on-map samples
    int64_t t_block;       // samples per time block
    int32_t *key;          // [n_on] tb * npix + iy * nx + ix, ascending
    uint32_t *idx;         // [n_on] sample index d * n_t + t
    float *spin;           // [n_on * nc] spin factors in sorted order
    int64_t bytes;
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

// Sort keys: (time block, row-major pixel) as tb * npix + pixel, or the
// largest key for an off-map sample (sorted last).
__global__ __launch_bounds__(kPlanBlock) void k_plan_keys(const int32_t *pix, int64_t n,
                                                          int64_t nx, int32_t npix, int64_t n_t,
                                                          int64_t t_block, int32_t off_key,
                                                          int32_t *keys, uint32_t *vals) {
    const int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (i >= n) return;
    const int2 p = *reinterpret_cast<const int2 *>(pix + 2 * i);
    const int64_t tb = (i % n_t) / t_block;
    keys[i] = p.x >= 0 ? (int32_t)(tb * npix + p.x * nx + p.y) : off_key;
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

template <int NC>
__global__ __launch_bounds__(kPlanBlock) void k_plan_spin(const float *spin, const uint32_t *idx,
                                                          int64_t n_on, float *out) {
    const int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    if (i >= n_on) return;
    const int64_t s = idx[i];
#pragma unroll
    for (int c = 0; c < NC; c++) out[i * NC + c] = spin[s * NC + c];
}

// Planned to_map.  Every lane of the grid handles one sorted sample (lanes
// past the end add nothing but still take part in the wave scan).
template <int NC>
__global__ __launch_bounds__(kPlanBlock) void k_plan_to_map(const int32_t *key,
                                                            const uint32_t *idx,
                                                            const float *spin, int64_t n_on,
                                                            uint32_t n_t, uint32_t plane,
                                                            const float *sig,
                                                            const float *det_weights,
                                                            double *map) {
    const int64_t i = (int64_t)blockIdx.x * kPlanBlock + threadIdx.x;
    int64_t k = -1;
    double v[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) v[c] = 0.;
    if (i < n_on) {
        k = key[i];
        const uint32_t s = idx[i];
        float w = 1.f;
        if (det_weights != nullptr) w = det_weights[s / n_t];
        if (w != 0.f) {  // a zero weight skips the detector (Projection.cxx:2104)
            const float sg = sig[s];
#pragma unroll
            for (int c = 0; c < NC; c++) v[c] = (double)(w * sg * spin[i * NC + c]);
        }
    }
    if (wave_runs<NC>(k, v) && k >= 0) {
        const uint32_t px = (uint32_t)k % plane;
#pragma unroll
        for (int c = 0; c < NC; c++) unsafeAtomicAdd(map + (int64_t)c * plane + px, v[c]);
    }
}

// Planned to_map, R sorted samples per lane.  Each lane sums its samples
// segment by segment (keys are sorted, so a segment is a run of equal keys):
// its first segment may continue from earlier lanes and goes through the
// wave-level run aggregation; segments that start and end inside the lane,
// and the lane's last segment when the lane holds more than one, are added
// to the map directly.
template <int NC, int R, typename ST = float>
__global__ __launch_bounds__(kPlanBlock) void k_plan_to_map_r(const int32_t *key,
                                                              const uint32_t *idx,
                                                              const float *spin, int64_t n_on,
                                                              uint32_t n_t, uint32_t plane,
                                                              const ST *sig,
                                                              const float *det_weights,
                                                              double *map) {
    static_assert(R % 4 == 0 && (R * NC) % 4 == 0, "vector loads");
    const int64_t base = ((int64_t)blockIdx.x * kPlanBlock + threadIdx.x) * R;
    int32_t kk[R];
    uint32_t ii[R];
    float sp[R * NC];
    const int nr = (int)max((int64_t)0, min((int64_t)R, n_on - base));
    if (nr == R) {
#pragma unroll
        for (int q = 0; q < R / 4; q++) {
            const int4 a = reinterpret_cast<const int4 *>(key + base)[q];
            kk[4 * q] = a.x; kk[4 * q + 1] = a.y; kk[4 * q + 2] = a.z; kk[4 * q + 3] = a.w;
            const uint4 b = reinterpret_cast<const uint4 *>(idx + base)[q];
            ii[4 * q] = b.x; ii[4 * q + 1] = b.y; ii[4 * q + 2] = b.z; ii[4 * q + 3] = b.w;
        }
#pragma unroll
        for (int q = 0; q < R * NC / 4; q++) {
            const float4 f = reinterpret_cast<const float4 *>(spin + base * NC)[q];
            sp[4 * q] = f.x; sp[4 * q + 1] = f.y; sp[4 * q + 2] = f.z; sp[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) {
            kk[r] = r < nr ? key[base + r] : -1;
            ii[r] = r < nr ? idx[base + r] : 0u;
#pragma unroll
            for (int c = 0; c < NC; c++) sp[r * NC + c] = r < nr ? spin[(base + r) * NC + c] : 0.f;
        }
    }
    int64_t hk = -1, ck = -1;
    double hv[NC], cv[NC];
    bool closed = false;   // has the lane's first segment ended?
#pragma unroll
    for (int c = 0; c < NC; c++) { hv[c] = 0.; cv[c] = 0.; }
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (r >= nr) break;
        const uint32_t si = ii[r];
        float w = 1.f;
        if (det_weights != nullptr) w = det_weights[si / n_t];
        double x[NC];
        if (w != 0.f) {  // a zero weight skips the detector (Projection.cxx:2104)
            const ST sg = sig[si];
#pragma unroll
            for (int c = 0; c < NC; c++) {
                if constexpr (std::is_same<ST, double>::value)
                    x[c] = (double)w * sg * (double)sp[r * NC + c];
                else
                    x[c] = (double)(w * sg * sp[r * NC + c]);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) x[c] = 0.;
        }
        const int64_t k = kk[r];
        if (k != ck) {
            if (ck >= 0) {
                if (!closed) {
                    hk = ck;
#pragma unroll
                    for (int c = 0; c < NC; c++) hv[c] = cv[c];
                    closed = true;
                } else {
                    const uint32_t px = (uint32_t)ck % plane;
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        unsafeAtomicAdd(map + (int64_t)c * plane + px, cv[c]);
                }
            }
            ck = k;
#pragma unroll
            for (int c = 0; c < NC; c++) cv[c] = x[c];
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) cv[c] += x[c];
        }
    }
    if (!closed) {
        hk = ck;
#pragma unroll
        for (int c = 0; c < NC; c++) hv[c] = cv[c];
    } else if (ck >= 0) {
        const uint32_t px = (uint32_t)ck % plane;
#pragma unroll
        for (int c = 0; c < NC; c++) unsafeAtomicAdd(map + (int64_t)c * plane + px, cv[c]);
    }
    if (wave_runs<NC>(hk, hv) && hk >= 0) {
        const uint32_t px = (uint32_t)hk % plane;
#pragma unroll
        for (int c = 0; c < NC; c++) unsafeAtomicAdd(map + (int64_t)c * plane + px, hv[c]);
    }
}

constexpr int kPlanR = 8;

}  // namespace so3g

extern "C" {

int so3g_pc_plan_create(const int32_t *pixel, int32_t k, const float *spin, int32_t nc,
                        int64_t n_det, int64_t n_t, const int64_t *map_shape,
                        int64_t t_block_mb, so3g_pc_plan **out, void *stream) {
    *out = nullptr;
    if (k != 2) return set_error(SO3G_ERR_RUNTIME, "plans support non-tiled flat maps (k = 2)");
    if (nc < 1 || nc > 3) return set_error(SO3G_ERR_RUNTIME, "spin component count must be 1..3");
    const int64_t n = n_det * n_t;
    if (n <= 0 || n > (int64_t)UINT32_MAX)
        return set_error(SO3G_ERR_RUNTIME, "plan needs 0 < n_det * n_t < 2^32");
    const int64_t ny = map_shape[0], nx = map_shape[1];
    if (ny * nx >= INT32_MAX) return set_error(SO3G_ERR_RUNTIME, "map too large for a plan");
    hipStream_t s = (hipStream_t)stream;
    const int32_t npix = (int32_t)(ny * nx);
    // Time blocks whose signal slice (n_det x t_block float32) stays resident
    // in the Infinity Cache while the block's samples are gathered.
    int64_t t_block = std::max<int64_t>(1, (t_block_mb << 20) / (4 * n_det));
    if (t_block_mb <= 0 || t_block > n_t) t_block = n_t;
    int64_t n_tb = (n_t + t_block - 1) / t_block;
    while (n_tb * npix >= INT32_MAX) {  // keys must fit int32
        t_block *= 2;
        n_tb = (n_t + t_block - 1) / t_block;
    }
    const int32_t off_key = (int32_t)(n_tb * npix);
    int end_bit = 1;
    while ((1LL << end_bit) <= off_key) end_bit++;

    int32_t *keys_in = nullptr, *keys_out = nullptr;
    uint32_t *vals_in = nullptr, *vals_out = nullptr;
    unsigned long long *d_on = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    auto cleanup = [&]() {
        (void)hipFree(keys_in); (void)hipFree(vals_in); (void)hipFree(d_on); (void)hipFree(tmp);
    };
    if (hipMalloc(&keys_in, n * 4) != hipSuccess || hipMalloc(&vals_in, n * 4) != hipSuccess ||
        hipMalloc(&keys_out, n * 4) != hipSuccess || hipMalloc(&vals_out, n * 4) != hipSuccess ||
        hipMalloc(&d_on, sizeof(unsigned long long)) != hipSuccess) {
        cleanup();
        (void)hipFree(keys_out); (void)hipFree(vals_out);
        (void)hipGetLastError();
        return set_error(SO3G_ERR_HIP, "out of device memory building the plan");
    }
    hipLaunchKernelGGL(k_plan_keys, dim3((unsigned)((n + kPlanBlock - 1) / kPlanBlock)),
                       dim3(kPlanBlock), 0, s, pixel, n, nx, npix, n_t, t_block, off_key,
                       keys_in, vals_in);
    PLAN_TRY(hipGetLastError());
    PLAN_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys_in, keys_out, vals_in, vals_out,
                                       (size_t)n, 0, end_bit, s));
    PLAN_TRY(hipMalloc(&tmp, tmp_bytes));
    PLAN_TRY(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out,
                                       (size_t)n, 0, end_bit, s));
    hipLaunchKernelGGL(k_plan_count, dim3(1), dim3(1), 0, s, keys_out, n, off_key, d_on);
    PLAN_TRY(hipGetLastError());
    unsigned long long n_on = 0;
    PLAN_TRY(hipMemcpyAsync(&n_on, d_on, sizeof(n_on), hipMemcpyDeviceToHost, s));
    PLAN_TRY(hipStreamSynchronize(s));
    cleanup();

    so3g_pc_plan *p = new so3g_pc_plan();
    p->n_det = n_det; p->n_t = n_t; p->ny = ny; p->nx = nx; p->nc = nc;
    p->n_on = (int64_t)n_on;
    p->t_block = t_block;
    p->key = keys_out;
    p->idx = vals_out;
    p->spin = nullptr;
    if (hipMalloc(&p->spin, std::max<int64_t>(p->n_on, 1) * nc * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        so3g_pc_plan_destroy(p);
        return set_error(SO3G_ERR_HIP, "out of device memory building the plan");
    }
    const dim3 g((unsigned)((p->n_on + kPlanBlock - 1) / kPlanBlock));
    if (p->n_on > 0) {
        if (nc == 1) hipLaunchKernelGGL(k_plan_spin<1>, g, dim3(kPlanBlock), 0, s, spin, p->idx, p->n_on, p->spin);
        else if (nc == 2) hipLaunchKernelGGL(k_plan_spin<2>, g, dim3(kPlanBlock), 0, s, spin, p->idx, p->n_on, p->spin);
        else hipLaunchKernelGGL(k_plan_spin<3>, g, dim3(kPlanBlock), 0, s, spin, p->idx, p->n_on, p->spin);
        PLAN_TRY(hipGetLastError());
    }
    PLAN_TRY(hipStreamSynchronize(s));
    p->bytes = n * 8 + p->n_on * nc * 4;
    *out = p;
    return SO3G_OK;
}

}  // extern "C"

namespace {
template <typename ST>
int plan_to_map(const so3g_pc_plan *p, double *map, const ST *signal,
                const float *det_weights, void *stream) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    if (p->n_on == 0) return SO3G_OK;
    hipStream_t s = (hipStream_t)stream;
    const int64_t per_block = (int64_t)kPlanBlock * kPlanR;
    const dim3 g((unsigned)((p->n_on + per_block - 1) / per_block));
    const uint32_t plane = (uint32_t)(p->ny * p->nx);
    const uint32_t nt = (uint32_t)p->n_t;
    switch (p->nc) {
    case 1: hipLaunchKernelGGL((k_plan_to_map_r<1, kPlanR, ST>), g, dim3(kPlanBlock), 0, s, p->key, p->idx, p->spin, p->n_on, nt, plane, signal, det_weights, map); break;
    case 2: hipLaunchKernelGGL((k_plan_to_map_r<2, kPlanR, ST>), g, dim3(kPlanBlock), 0, s, p->key, p->idx, p->spin, p->n_on, nt, plane, signal, det_weights, map); break;
    default: hipLaunchKernelGGL((k_plan_to_map_r<3, kPlanR, ST>), g, dim3(kPlanBlock), 0, s, p->key, p->idx, p->spin, p->n_on, nt, plane, signal, det_weights, map); break;
    }
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
}  // namespace

extern "C" {

int so3g_pc_plan_to_map(const so3g_pc_plan *p, double *map, const float *signal,
                        const float *det_weights, void *stream) {
    return plan_to_map<float>(p, map, signal, det_weights, stream);
}

int so3g_pc_plan_to_map_f64(const so3g_pc_plan *p, double *map, const double *signal,
                            const float *det_weights, void *stream) {
    return plan_to_map<double>(p, map, signal, det_weights, stream);
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
    const int64_t nb = std::min<int64_t>((n + kPlanBlock - 1) / kPlanBlock, 1 << 20);
    const dim3 g((unsigned)nb);
    switch (nc) {
    case 1: hipLaunchKernelGGL(k_pc_from_map_f64<1>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    case 2: hipLaunchKernelGGL(k_pc_from_map_f64<2>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    default: hipLaunchKernelGGL(k_pc_from_map_f64<3>, g, dim3(kPlanBlock), 0, s, n, (int32_t)ny, (int32_t)nx, map, pixel, spin, signal, accumulate); break;
    }
    PLAN_TRY(hipGetLastError());
    return SO3G_OK;
}

int so3g_pc_plan_info(const so3g_pc_plan *p, int64_t *n_on, int64_t *bytes) {
    if (p == nullptr) return set_error(SO3G_ERR_VALUE, "plan is NULL");
    *n_on = p->n_on;
    *bytes = p->bytes;
    return SO3G_OK;
}

int so3g_pc_plan_destroy(so3g_pc_plan *p) {
    if (p == nullptr) return SO3G_OK;
    (void)hipFree(p->key);
    (void)hipFree(p->idx);
    (void)hipFree(p->spin);
    delete p;
    return SO3G_OK;
}

}  // extern "C"
