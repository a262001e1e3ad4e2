// nmat.hip -- the detector-vector noise model's inverse-covariance
// application in Fourier space (nmat_detvecs_apply, src/array_ops.cxx:39-88),
// the other operator of every map-making CG iteration.
//
// For each frequency bin b with (real-view) column range [c1, c2):
//     Q  = iV_b^T ft[:, c1:c2]                       (nvec x nm)
//     ft[:, c1:c2] = ft * (iD_b / norm) + (s / norm) * iV_b Q
// in float32, bins in order (a column in two bins is transformed twice, as
// in the reference when bin ends are clamped).
//
// ft is read twice and written once (3 x 4 B per element); Q, nvec rows of
// nmode floats, and its per-chunk partials live in a library buffer that is
// reused across calls (calls on one device share it: do not overlap calls on
// different streams).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/so3g_hip.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

namespace {

constexpr int kNT = 256;        // threads per block
constexpr int kMaxVec = 64;

// Two streaming passes, each thread owning C consecutive columns and
// walking all detector rows, so that a block reads 256 C contiguous floats
// of every row (whole DRAM pages; a narrow column strip over 1000 rows 2.9 MB
// apart would open a page per 64 B):
//   pass 1   Q[v][col] = sum_d iV[bin(col)][d][v] * ft[d][col]
//   pass 2   ft[d][col] = ft * (iD[bin][d] / norm) + (s / norm) sum_v iV[bin][d][v] Q[v][col]
// cbin[col] is the column's bin (-1: untouched).  NV >= nvec accumulators
// per column stay in registers (NV x C <= 64).
// A thread's C = 4 columns of one row in one 16-byte access where the
// address allows it, else two 8-byte ones (a row of an even number of modes
// starts 8-byte aligned; col0 is a multiple of 4, so the alignment is the
// row's and the same for every lane of the wave), else four 4-byte ones.
// Four 4-byte loads per row made the texture addresser the limit: 10x the
// L1 accesses of a plain read of the array.
template <int NV, int C>
__device__ __forceinline__ void load_cols(const float *row, int64_t col0, int nc,
                                          float (&x)[C]) {
    const float *p = row + col0;
    if constexpr (C == 4) {
        if (nc == 4) {
            if (((uintptr_t)p & 15) == 0) {
                const float4 f = *reinterpret_cast<const float4 *>(p);
                x[0] = f.x; x[1] = f.y; x[2] = f.z; x[3] = f.w;
                return;
            }
            if (((uintptr_t)p & 7) == 0) {
                const float2 f = *reinterpret_cast<const float2 *>(p);
                const float2 g = *reinterpret_cast<const float2 *>(p + 2);
                x[0] = f.x; x[1] = f.y; x[2] = g.x; x[3] = g.y;
                return;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < C; j++) x[j] = j < nc ? p[j] : 0.f;
}

// The fast loops' 4 columns of one row with the access width W chosen for
// the whole launch: 4 (one 16-byte access: every row 16-byte aligned), 2
// (two 8-byte ones: an even number of modes) or 1.
template <int W>
__device__ __forceinline__ void load4(const float *p, float (&x)[4]) {
    if constexpr (W == 4) {
        const float4 f = *reinterpret_cast<const float4 *>(p);
        x[0] = f.x; x[1] = f.y; x[2] = f.z; x[3] = f.w;
    } else if constexpr (W == 2) {
        const float2 f = *reinterpret_cast<const float2 *>(p);
        const float2 g = *reinterpret_cast<const float2 *>(p + 2);
        x[0] = f.x; x[1] = f.y; x[2] = g.x; x[3] = g.y;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = p[j];
    }
}
template <int W>
__device__ __forceinline__ void store4(float *p, const float (&o)[4]) {
    if constexpr (W == 4) {
        *reinterpret_cast<float4 *>(p) = float4{o[0], o[1], o[2], o[3]};
    } else if constexpr (W == 2) {
        *reinterpret_cast<float2 *>(p) = float2{o[0], o[1]};
        *reinterpret_cast<float2 *>(p + 2) = float2{o[2], o[3]};
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) p[j] = o[j];
    }
}

// iV with each detector's row padded with zeros to NV values:
// iVp[(b * ndet + d) * NV + v].
__global__ __launch_bounds__(kNT) void k_nmat_pad(const float *__restrict__ iV, int64_t n,
                                                  int nvec, int nv, float *__restrict__ iVp) {
    for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n * nv;
         i += (int64_t)gridDim.x * kNT) {
        const int64_t r = i / nv;
        const int v = (int)(i - r * nv);
        iVp[i] = v < nvec ? iV[r * nvec + v] : 0.f;
    }
}

// Stores a thread's C columns of one row the same way; o holds the row's
// own values in columns outside every bin, so the wide stores rewrite them
// unchanged.
template <int C>
__device__ __forceinline__ void store_cols(float *row, int64_t col0, int nc, const int (&b)[C],
                                           const float (&o)[C]) {
    float *p = row + col0;
    if constexpr (C == 4) {
        if (nc == 4) {
            if (((uintptr_t)p & 15) == 0) {
                *reinterpret_cast<float4 *>(p) = float4{o[0], o[1], o[2], o[3]};
                return;
            }
            if (((uintptr_t)p & 7) == 0) {
                *reinterpret_cast<float2 *>(p) = float2{o[0], o[1]};
                *reinterpret_cast<float2 *>(p + 2) = float2{o[2], o[3]};
                return;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < C; j++)
        if (j < nc && b[j] >= 0) p[j] = o[j];
}

}  // namespace

namespace so3g {
// Detectors per block of both passes ("nmat_dchunk"): pass 1 writes one
// partial Q per chunk of detectors, which pass 1's reduction then sums.
int g_nmat_dchunk = 64;
}  // namespace so3g

namespace {

template <int NV, int C, int W>
__global__ __launch_bounds__(kNT) void k_nmat_q(const float *__restrict__ ft, int ndet,
                                                int64_t nmode, const int32_t *__restrict__ cbin,
                                                const float *__restrict__ iV,
                                                const float *__restrict__ iVp, int nvec,
                                                float *__restrict__ Qp, int dchunk) {
    const int d0 = blockIdx.y * dchunk, d1 = min(ndet, d0 + dchunk);
    float *Q = Qp + (int64_t)blockIdx.y * nvec * nmode;   // this chunk's partial
    const int64_t col0 = ((int64_t)blockIdx.x * kNT + threadIdx.x) * C;
    if (col0 >= nmode) return;
    const int nc = (int)min((int64_t)C, nmode - col0);
    int b[C];
    bool same = true;
#pragma unroll
    for (int j = 0; j < C; j++) {
        b[j] = j < nc ? cbin[col0 + j] : -1;
        same = same && b[j] == b[0];
    }
    if (same && b[0] < 0) return;
    // When the whole wave's columns share one bin (the common case), the
    // iV row address is wave-uniform and its loads go through the scalar unit.
    const int bw = __builtin_amdgcn_readfirstlane(b[0]);
    const bool uni = __all(same && b[0] == bw);
    float acc[NV][C];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int j = 0; j < C; j++) acc[v][j] = 0.f;
    if constexpr (C == 4) {
        // Common case: one bin for the wave, 4 whole columns per lane.  No
        // branch in the row loop (padded iV rows, the access width fixed for
        // the launch), so the next rows' loads issue ahead of this row's
        // multiply-adds; the sums are those of the general loop (the padding
        // adds zeros to accumulators that are not stored).
        if (uni && nc == 4) {
            const float *vr = iVp + ((int64_t)bw * ndet + d0) * NV;
            const float *xp = ft + (int64_t)d0 * nmode + col0;
#pragma unroll 8
            for (int d = d0; d < d1; d++, vr += NV, xp += nmode) {
                float x[4];
                load4<W>(xp, x);
#pragma unroll
                for (int v = 0; v < NV; v++) {
                    const float r = vr[v];
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[v][j] += r * x[j];
                }
            }
#pragma unroll
            for (int v = 0; v < NV; v++)
                if (v < nvec) {
                    float o[4] = {acc[v][0], acc[v][1], acc[v][2], acc[v][3]};
                    store4<W>(Q + (int64_t)v * nmode + col0, o);
                }
            return;
        }
    }
#pragma unroll 4
    for (int d = d0; d < d1; d++) {
        float x[C];
        load_cols<NV, C>(ft + (int64_t)d * nmode, col0, nc, x);
        if (uni) {
            const float *row = iV + ((int64_t)bw * ndet + d) * nvec;
#pragma unroll
            for (int v = 0; v < NV; v++) {
                if (v < nvec) {
                    const float r = row[v];
#pragma unroll
                    for (int j = 0; j < C; j++) acc[v][j] += r * x[j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < C; j++) {
                if (b[j] < 0) continue;
                const float *row = iV + ((int64_t)b[j] * ndet + d) * nvec;
#pragma unroll
                for (int v = 0; v < NV; v++)
                    if (v < nvec) acc[v][j] += row[v] * x[j];
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
        if (v < nvec) {
#pragma unroll
            for (int j = 0; j < C; j++)
                if (j < nc) Q[(int64_t)v * nmode + col0 + j] = acc[v][j];
        }
    }
}

// Q = sum over detector chunks of the partials, in chunk order.
__global__ __launch_bounds__(kNT) void k_nmat_reduce(const float *__restrict__ Qp, int nchunk,
                                                     int64_t n, const int32_t *__restrict__ cbin,
                                                     int64_t nmode, float *__restrict__ Q) {
    for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kNT) {
        if (cbin[i % nmode] < 0) continue;
        float acc = 0.f;
        for (int k = 0; k < nchunk; k++) acc += Qp[(int64_t)k * n + i];
        Q[i] = acc;
    }
}

template <int NV, int C>
__device__ __forceinline__ void update_row(const float (&x)[C], float (&o)[C], const int (&b)[C],
                                           bool uni, int bw, int d, int ndet,
                                           const float *__restrict__ iD,
                                           const float *__restrict__ iV, int nvec, float norm,
                                           float a, const float (&q)[NV][C]) {
    if (uni) {
        const float *vr = iV + ((int64_t)bw * ndet + d) * nvec;
        const float dn = iD[(int64_t)bw * ndet + d] / norm;
        float acc[C];
#pragma unroll
        for (int j = 0; j < C; j++) acc[j] = 0.f;
#pragma unroll
        for (int v = 0; v < NV; v++) {
            if (v < nvec) {
                const float r = vr[v];
#pragma unroll
                for (int j = 0; j < C; j++) acc[j] += r * q[v][j];
            }
        }
#pragma unroll
        for (int j = 0; j < C; j++) o[j] = x[j] * dn + a * acc[j];
    } else {
#pragma unroll
        for (int j = 0; j < C; j++) {
            if (b[j] < 0) { o[j] = x[j]; continue; }
            const float *vr = iV + ((int64_t)b[j] * ndet + d) * nvec;
            float acc = 0.f;
#pragma unroll
            for (int v = 0; v < NV; v++)
                if (v < nvec) acc += vr[v] * q[v][j];
            o[j] = x[j] * (iD[(int64_t)b[j] * ndet + d] / norm) + a * acc;
        }
    }
}

// Pass 2 walks its detector rows kRU at a time and loads the next group's
// rows before storing the current group's: a load's wait (vmcnt) also
// covers the wave's older stores, so loads issued after the stores would
// wait for them.
constexpr int kRU = 4;

template <int NV, int C, int W>
__global__ __launch_bounds__(kNT) void k_nmat_update(float *__restrict__ ft, int ndet,
                                                     int64_t nmode,
                                                     const int32_t *__restrict__ cbin,
                                                     const float *__restrict__ iD,
                                                     const float *__restrict__ iV,
                                                     const float *__restrict__ iVp, int nvec,
                                                     float s, float norm,
                                                     const float *__restrict__ Q, int dchunk) {
    const int64_t col0 = ((int64_t)blockIdx.x * kNT + threadIdx.x) * C;
    if (col0 >= nmode) return;
    const int nc = (int)min((int64_t)C, nmode - col0);
    int b[C];
    bool any = false, same = true;
#pragma unroll
    for (int j = 0; j < C; j++) {
        b[j] = j < nc ? cbin[col0 + j] : -1;
        any = any || b[j] >= 0;
        same = same && b[j] == b[0];
    }
    if (!any) return;
    const int bw = __builtin_amdgcn_readfirstlane(b[0]);
    const bool uni = __all(same && b[0] == bw && bw >= 0);
    float q[NV][C];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int j = 0; j < C; j++)
            q[v][j] = (v < nvec && j < nc) ? Q[(int64_t)v * nmode + col0 + j] : 0.f;
    const float a = s / norm;
    const int d0 = blockIdx.y * dchunk, d1 = min(ndet, d0 + dchunk);
    if constexpr (C == 4) {
        // Common case (one bin for the wave, 4 whole columns per lane): the
        // general loop's arithmetic (update_row's uniform branch; the padded
        // iV values add zero products to the sum) without its branches.
        if (uni && nc == 4) {
            const float *vr = iVp + ((int64_t)bw * ndet + d0) * NV;
            const float *dr = iD + (int64_t)bw * ndet;
            float xa[kRU][4], xb[kRU][4];
#pragma unroll
            for (int r = 0; r < kRU; r++)
                if (d0 + r < d1) load4<W>(ft + (int64_t)(d0 + r) * nmode + col0, xa[r]);
            for (int d = d0; d < d1; d += kRU) {
                // next group's loads first
#pragma unroll
                for (int r = 0; r < kRU; r++)
                    if (d + kRU + r < d1)
                        load4<W>(ft + (int64_t)(d + kRU + r) * nmode + col0, xb[r]);
#pragma unroll
                for (int r = 0; r < kRU; r++) {
                    if (d + r >= d1) break;
                    const float *rv = vr + (int64_t)(d + r - d0) * NV;
                    const float dn = dr[d + r] / norm;
                    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        const float rr = rv[v];
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[j] += rr * q[v][j];
                    }
                    float o[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) o[j] = xa[r][j] * dn + a * acc[j];
                    store4<W>(ft + (int64_t)(d + r) * nmode + col0, o);
                }
#pragma unroll
                for (int r = 0; r < kRU; r++)
#pragma unroll
                    for (int j = 0; j < 4; j++) xa[r][j] = xb[r][j];
            }
            return;
        }
    }
    float x[kRU][C], xn[kRU][C];
#pragma unroll
    for (int r = 0; r < kRU; r++)
        if (d0 + r < d1) load_cols<NV, C>(ft + (int64_t)(d0 + r) * nmode, col0, nc, x[r]);
    for (int d = d0; d < d1; d += kRU) {
        // next group's loads first
#pragma unroll
        for (int r = 0; r < kRU; r++)
            if (d + kRU + r < d1)
                load_cols<NV, C>(ft + (int64_t)(d + kRU + r) * nmode, col0, nc, xn[r]);
#pragma unroll
        for (int r = 0; r < kRU; r++) {
            if (d + r >= d1) break;
            float *row = ft + (int64_t)(d + r) * nmode;
            float o[C];
            update_row<NV, C>(x[r], o, b, uni, bw, d + r, ndet, iD, iV, nvec, norm, a, q);
            store_cols<C>(row, col0, nc, b, o);
        }
#pragma unroll
        for (int r = 0; r < kRU; r++)
#pragma unroll
            for (int j = 0; j < C; j++) x[r][j] = xn[r][j];
    }
}

struct NmCache {
    std::vector<int32_t> key;
    int64_t nmode = -1;
    std::vector<int32_t *> dev;    // one column -> bin map per group
};
// Q and its per-chunk partials, grown as needed and kept (a CG loop calls
// this every iteration with the same shapes).
float *g_nm_q = nullptr;
size_t g_nm_q_cap = 0;
float *g_nm_ivp = nullptr;     // iV padded to NV values per detector
size_t g_nm_ivp_cap = 0;
NmCache g_nm_cache;
std::mutex g_nm_mutex;

#define NM_TRY(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                               hipGetErrorString(_e));                 \
    } while (0)

template <int NV, int C, int W>
void launch_pair(float *ft, int ndet, int64_t nmode, const int32_t *cbin, const float *iD,
                 const float *iV, const float *iVp, int nvec, float s, float norm, float *Qp,
                 float *Q, int dchunk, hipStream_t st) {
    const int nchunk = (ndet + dchunk - 1) / dchunk;
    const dim3 g((unsigned)((nmode + (int64_t)kNT * C - 1) / ((int64_t)kNT * C)),
                 (unsigned)nchunk);
    hipLaunchKernelGGL((k_nmat_q<NV, C, W>), g, dim3(kNT), 0, st, ft, ndet, nmode, cbin, iV,
                       iVp, nvec, Qp, dchunk);
    const int64_t n = (int64_t)nvec * nmode;
    if (n > 0)
        hipLaunchKernelGGL(k_nmat_reduce, dim3((unsigned)std::min<int64_t>((n + kNT - 1) / kNT, 65536)),
                           dim3(kNT), 0, st, Qp, nchunk, n, cbin, nmode, Q);
    hipLaunchKernelGGL((k_nmat_update<NV, C, W>), g, dim3(kNT), 0, st, ft, ndet, nmode, cbin,
                       iD, iV, iVp, nvec, s, norm, Q, dchunk);
}

// Access width of the fast loops: 16 bytes when every row starts 16-byte
// aligned, 8 when every row starts 8-byte aligned, else 4.
template <int NV, int C>
void launch_pair_w(float *ft, int ndet, int64_t nmode, const int32_t *cbin, const float *iD,
                   const float *iV, const float *iVp, int nvec, float s, float norm,
                   float *Qp, float *Q, int dchunk, hipStream_t st) {
    const uintptr_t base = (uintptr_t)ft;
    if (nmode % 4 == 0 && base % 16 == 0)
        launch_pair<NV, C, 4>(ft, ndet, nmode, cbin, iD, iV, iVp, nvec, s, norm, Qp, Q, dchunk,
                              st);
    else if (nmode % 2 == 0 && base % 8 == 0)
        launch_pair<NV, C, 2>(ft, ndet, nmode, cbin, iD, iV, iVp, nvec, s, norm, Qp, Q, dchunk,
                              st);
    else
        launch_pair<NV, C, 1>(ft, ndet, nmode, cbin, iD, iV, iVp, nvec, s, norm, Qp, Q, dchunk,
                              st);
}

}  // namespace

extern "C" {

int so3g_nmat_detvecs_apply(float *ft, int64_t ndet, int64_t nmode, const int32_t *bins,
                            int64_t nbin, const float *iD, const float *iV, int64_t nvec,
                            float s, float norm, int dct_binning, void *stream) {
    if (ndet < 0 || nmode < 0 || nbin < 0 || nvec < 0)
        return set_error(SO3G_ERR_RUNTIME, "negative dimension");
    if (ndet == 0 || nmode == 0 || nbin == 0) return SO3G_OK;
    if (nvec > kMaxVec) return set_error(SO3G_ERR_RUNTIME, "nvec above 64 is not supported");
    if (ndet > INT32_MAX || nbin > INT32_MAX)
        return set_error(SO3G_ERR_RUNTIME, "too many detectors or bins");
    if (bins == nullptr) return set_error(SO3G_ERR_VALUE, "bins is NULL");
    const int64_t scale = dct_binning ? 1 : 2;
    hipStream_t st = (hipStream_t)stream;
    // Column ranges as the reference clamps them (array_ops.cxx:65-66).  Bins
    // run in order; consecutive bins whose ranges do not overlap form one
    // group (one pass pair), a column's bin recorded in cbin.
    std::vector<std::vector<int32_t>> groups;
    std::vector<std::pair<int64_t, int64_t>> used;
    for (int64_t b = 0; b < nbin; b++) {
        const int64_t b1 = std::min(scale * (int64_t)bins[2 * b], nmode - 1);
        const int64_t b2 = std::min(scale * (int64_t)bins[2 * b + 1], nmode);
        if (b2 <= b1) continue;
        bool clash = groups.empty();
        for (auto &r : used)
            if (b1 < r.second && r.first < b2) clash = true;
        if (clash) {
            groups.emplace_back(nmode, -1);
            used.clear();
        }
        used.emplace_back(b1, b2);
        std::fill(groups.back().begin() + b1, groups.back().begin() + b2, (int32_t)b);
    }
    if (groups.empty()) return SO3G_OK;
    // The bin maps are uploaded once per (bins, nmode, binning): a CG loop
    // calls this with the same bins every iteration.
    std::lock_guard<std::mutex> lock(g_nm_mutex);
    std::vector<int32_t> key(bins, bins + 2 * nbin);
    key.push_back((int32_t)scale);
    if (!(g_nm_cache.key == key && g_nm_cache.nmode == nmode)) {
        for (int32_t *p : g_nm_cache.dev) (void)hipFree(p);
        g_nm_cache = NmCache{};
        for (auto &g : groups) {
            int32_t *d = nullptr;
            NM_TRY(hipMalloc((void **)&d, nmode * sizeof(int32_t)));
            g_nm_cache.dev.push_back(d);
            NM_TRY(hipMemcpy(d, g.data(), nmode * sizeof(int32_t), hipMemcpyHostToDevice));
        }
        g_nm_cache.key = key;
        g_nm_cache.nmode = nmode;
    }
    const int dchunk = std::max(1, so3g::g_nmat_dchunk);
    const int64_t nchunk = (ndet + dchunk - 1) / dchunk;
    const size_t need = (size_t)std::max<int64_t>(1, (nchunk + 1) * nvec * nmode);
    if (need > g_nm_q_cap) {
        // earlier calls on other streams may still read the old buffer
        NM_TRY(hipDeviceSynchronize());
        if (g_nm_q) (void)hipFree(g_nm_q);
        g_nm_q = nullptr;
        g_nm_q_cap = 0;
        NM_TRY(hipMalloc((void **)&g_nm_q, need * sizeof(float)));
        g_nm_q_cap = need;
    }
    float *Q = g_nm_q, *Qp = g_nm_q + nvec * nmode;
    int rc = SO3G_OK;
    const int nv = nvec <= 4 ? 4 : nvec <= 8 ? 8 : nvec <= 12 ? 12 : nvec <= 16 ? 16
                                                                  : nvec <= 32 ? 32 : 64;
    const size_t nivp = (size_t)std::max<int64_t>(1, nbin * ndet * nv);
    if (nivp > g_nm_ivp_cap) {
        NM_TRY(hipDeviceSynchronize());
        if (g_nm_ivp) (void)hipFree(g_nm_ivp);
        g_nm_ivp = nullptr;
        g_nm_ivp_cap = 0;
        NM_TRY(hipMalloc((void **)&g_nm_ivp, nivp * sizeof(float)));
        g_nm_ivp_cap = nivp;
    }
    const float *iVp = g_nm_ivp;
    {
        const int64_t n = nbin * ndet;
        hipLaunchKernelGGL(k_nmat_pad,
                           dim3((unsigned)std::min<int64_t>((n * nv + kNT - 1) / kNT, 65536)),
                           dim3(kNT), 0, st, iV, n, (int)nvec, nv, g_nm_ivp);
    }
    for (size_t gi = 0; gi < groups.size(); gi++) {
        const int32_t *cbin = g_nm_cache.dev[gi];
        const int nd = (int)ndet, nvi = (int)nvec;
        switch (nv) {
        case 4: launch_pair_w<4, 4>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        case 8: launch_pair_w<8, 4>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        case 12: launch_pair_w<12, 4>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        case 16: launch_pair_w<16, 4>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        case 32: launch_pair_w<32, 2>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        default: launch_pair<64, 1, 1>(ft, nd, nmode, cbin, iD, iV, iVp, nvi, s, norm, Qp, Q, dchunk, st); break;
        }
        if (hipGetLastError() != hipSuccess) {
            rc = set_error(SO3G_ERR_HIP, "nmat kernel launch failed");
            break;
        }
    }
    return rc;
}

}  // extern "C"
