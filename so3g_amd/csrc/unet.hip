// unet.hip -- the layers of the learned CG preconditioner: 3x3 convolutions
// on bf16 NHWC activations as implicit GEMMs on the CDNA matrix cores
// (v_mfma_f32_16x16x16_bf16, float32 accumulation), their weight and data
// gradients, average pooling, nearest upsampling, and an Adam step.
//
// MFMA 16x16x16 bf16 operand layout (one wave, lane l = 0..63):
//   A (16 x 16, row i, col k): lane l holds A[l % 16][4 * (l / 16) + v], v = 0..3
//   B (16 x 16, row k, col j): lane l holds B[4 * (l / 16) + v][l % 16]
//   D (16 x 16 float):         lane l holds D[4 * (l / 16) + v][l % 16]
// Convolution: rows i = 16 consecutive output pixels of one image row,
// k = 16 input channels of one tap, j = 16 output channels.  Channel counts
// are multiples of 16 (the caller pads the map's 3 components).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/so3g_hip.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}

// float -> bf16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int kNB = 256;   // threads per block (4 waves)

// out[n, y, x, co] = act(bias[co] + sum_{tap, ci} in[n, y + dy, x + dx, ci] * wt[tap][co][ci])
// (zero padding), act = ReLU if relu.  One wave computes 16 pixels of one
// row for all CO channels; a block covers 64 pixels.
template <int CO>
__global__ __launch_bounds__(kNB) void k_conv3x3(const uint16_t *__restrict__ in, int H, int W,
                                                 int CI, const uint16_t *__restrict__ wt,
                                                 const float *__restrict__ bias, int relu,
                                                 uint16_t *__restrict__ out, int tiles_x) {
    constexpr int NT = CO / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tx = blockIdx.x % tiles_x;
    const int y = (blockIdx.x / tiles_x) % H;
    const int n = blockIdx.x / (tiles_x * H);
    const int x0 = tx * 64 + wave * 16;
    if (x0 >= W) return;
    float4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
    const int li = lane & 15, lk = 4 * (lane >> 4);
    for (int tap = 0; tap < 9; tap++) {
        const int yy = y + tap / 3 - 1;
        if (yy < 0 || yy >= H) continue;
        const int xi = x0 + li + tap % 3 - 1;
        const bool ok = xi >= 0 && xi < W;
        const uint16_t *row = in + ((int64_t)(n * H + yy) * W + (ok ? xi : 0)) * CI;
        const uint16_t *wtap = wt + (int64_t)tap * CO * CI;
        for (int cb = 0; cb < CI; cb += 16) {
            short4_t a = {0, 0, 0, 0};
            if (ok) a = *reinterpret_cast<const short4_t *>(row + cb + lk);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const short4_t b =
                    *reinterpret_cast<const short4_t *>(wtap + (t * 16 + li) * CI + cb + lk);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int co = t * 16 + li;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int x = x0 + lk + v;
            if (x >= W) continue;
            float r = acc[t][v] + bv;
            if (relu) r = fmaxf(r, 0.f);
            out[((int64_t)(n * H + y) * W + x) * CO + co] = f2bf(r);
        }
    }
}

// dW[tap][co][ci] += sum_p dy[p, co] * in[p + off(tap), ci] and
// db[co] += sum_p dy[p, co] over a chunk of kWgChunk pixels (flattened n, y,
// x) per block; blockIdx.y = tap.  Each wave owns (co tile, ci tile) pairs.
constexpr int kWgChunk = 1024;

__global__ __launch_bounds__(kNB) void k_conv3x3_wgrad(const uint16_t *__restrict__ in,
                                                       const uint16_t *__restrict__ dy,
                                                       int N, int H, int W, int CI, int CO,
                                                       float *__restrict__ dw,
                                                       float *__restrict__ db) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.y;
    const int ddy = tap / 3 - 1, ddx = tap % 3 - 1;
    const int64_t npx = (int64_t)N * H * W;
    const int64_t p0 = (int64_t)blockIdx.x * kWgChunk;
    const int64_t p1 = min(p0 + kWgChunk, npx);
    const int li = lane & 15, lk = 4 * (lane >> 4);
    const int nco = CO / 16, nci = CI / 16;
    for (int pair = wave; pair < nco * nci; pair += 4) {
        const int tco = pair / nci, tci = pair % nci;
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        for (int64_t p = p0; p < p1; p += 16) {
            short4_t a, b;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t q = p + lk + v;   // this lane's k (pixel) index
                uint16_t av = 0, bv = 0;
                if (q < p1) {
                    av = dy[q * CO + tco * 16 + li];
                    const int x = (int)(q % W);
                    const int y = (int)((q / W) % H);
                    const int yy = y + ddy, xx = x + ddx;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                        bv = in[(q + (int64_t)ddy * W + ddx) * CI + tci * 16 + li];
                }
                a[v] = (short)av;
                b[v] = (short)bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        // D[i = co][j = ci]: lane holds rows 4*(l/16)+v, column l%16.
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int co = tco * 16 + lk + v, ci = tci * 16 + li;
            atomicAdd(dw + ((int64_t)tap * CO + co) * CI + ci, acc[v]);
        }
    }
    if (tap == 4 && db != nullptr && (int)threadIdx.x < CO) {
        float s = 0.f;
        for (int64_t p = p0; p < p1; p++) s += bf2f(dy[p * CO + threadIdx.x]);
        atomicAdd(db + threadIdx.x, s);
    }
}

// Weights for the data gradient: wb[tap][ci][co] = wf[8 - tap][co][ci].
__global__ void k_weight_flip(const uint16_t *wf, int CI, int CO, uint16_t *wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 9 * CI * CO) return;
    const int ci = i % CI, co = (i / CI) % CO, tap = i / (CI * CO);
    wb[((int64_t)(8 - tap) * CI + ci) * CO + co] = wf[i];
}

// dz = dy * (act > 0) (ReLU backward, act = the layer's output).
__global__ void k_relu_mask(const uint16_t *dy, const uint16_t *act, int64_t n, uint16_t *dz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dz[i] = bf2f(act[i]) > 0.f ? dy[i] : (uint16_t)0;
}

__global__ void k_add(const uint16_t *a, const uint16_t *b, int64_t n, uint16_t *o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        o[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// fy x fx average pooling of (N, H, W, C) -> (N, H/fy, W/fx, C) and its
// gradient; nearest (fy, fx) upsampling of (N, H, W, C) -> (N, H fy, W fx, C)
// and its gradient (the sum over each block).
__global__ void k_pool(const uint16_t *in, int N, int H, int W, int C, int fy, int fx,
                       uint16_t *out) {
    const int Ho = H / fy, Wo = W / fx;
    const float inv = 1.f / (float)(fy * fx);
    const int64_t n_out = (int64_t)N * Ho * Wo * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho);
        const int64_t n = p / ((int64_t)Wo * Ho);
        float s = 0.f;
        for (int a = 0; a < fy; a++)
            for (int b = 0; b < fx; b++)
                s += bf2f(in[((n * H + (int64_t)y * fy + a) * W + (int64_t)x * fx + b) * C + c]);
        out[i] = f2bf(inv * s);
    }
}

__global__ void k_pool_bwd(const uint16_t *dy, int N, int H, int W, int C, int fy, int fx,
                           uint16_t *dx) {
    const int Ho = H / fy, Wo = W / fx;
    const float inv = 1.f / (float)(fy * fx);
    const int64_t n_in = (int64_t)N * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_in;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const int64_t n = p / ((int64_t)W * H);
        dx[i] = f2bf(inv * bf2f(dy[((n * Ho + y / fy) * Wo + x / fx) * C + c]));
    }
}

__global__ void k_up(const uint16_t *in, int N, int H, int W, int C, int fy, int fx,
                     uint16_t *out) {
    const int Hu = H * fy, Wu = W * fx;
    const int64_t n_out = (int64_t)N * Hu * Wu * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % Wu), y = (int)((p / Wu) % Hu);
        const int64_t n = p / ((int64_t)Wu * Hu);
        out[i] = in[((n * H + y / fy) * W + x / fx) * C + c];
    }
}

__global__ void k_up_bwd(const uint16_t *dy, int N, int H, int W, int C, int fy, int fx,
                         uint16_t *dx) {
    const int Hu = H * fy, Wu = W * fx;
    const int64_t n_in = (int64_t)N * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_in;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int x = (int)(p % W), y = (int)((p / W) % H);
        const int64_t n = p / ((int64_t)W * H);
        float s = 0.f;
        for (int a = 0; a < fy; a++)
            for (int b = 0; b < fx; b++)
                s += bf2f(dy[((n * Hu + (int64_t)y * fy + a) * Wu + (int64_t)x * fx + b) * C + c]);
        dx[i] = f2bf(s);
    }
}

// float32 planes (N, CP, HY, WX) <-> bf16 NHWC (N, H, W, C), C >= CP, H >= HY,
// W >= WX; the extra channels and the pad region are zero.
__global__ void k_from_planes(const float *x, int N, int CP, int HY, int WX, int H, int W, int C,
                              uint16_t *out) {
    const int64_t n_out = (int64_t)N * H * W * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const int64_t p = i / C;
        const int xx = (int)(p % W), yy = (int)((p / W) % H);
        const int64_t n = p / ((int64_t)W * H);
        float v = 0.f;
        if (c < CP && yy < HY && xx < WX) v = x[((n * CP + c) * HY + yy) * WX + xx];
        out[i] = f2bf(v);
    }
}

__global__ void k_to_planes(const uint16_t *in, int N, int H, int W, int C, int CP, int HY,
                            int WX, float *x) {
    const int64_t n_out = (int64_t)N * CP * HY * WX;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_out;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int xx = (int)(i % WX), yy = (int)((i / WX) % HY);
        const int c = (int)((i / ((int64_t)WX * HY)) % CP);
        const int64_t n = i / ((int64_t)WX * HY * CP);
        x[i] = bf2f(in[((n * H + yy) * W + xx) * C + c]);
    }
}

// Adam (bias-corrected) on float32 master weights; writes the bf16 copy.
__global__ void k_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, uint16_t *wb) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ void k_to_bf16(const float *x, int64_t n, uint16_t *o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        o[i] = f2bf(x[i]);
}

dim3 ew_grid(int64_t n) {
    int64_t b = (n + kNB - 1) / kNB;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

#define NN_CHECK()                                                                      \
    do {                                                                                \
        hipError_t _e = hipGetLastError();                                              \
        if (_e != hipSuccess)                                                           \
            return set_error(SO3G_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(_e)); \
    } while (0)

bool ch_ok(int c) { return c > 0 && c <= 64 && c % 16 == 0; }

}  // namespace

extern "C" {

int so3g_nn_conv3x3(const void *in, int n, int h, int w, int ci, const void *wt,
                    const float *bias, int co, int relu, void *out, void *stream) {
    if (!ch_ok(ci) || !ch_ok(co)) return set_error(SO3G_ERR_VALUE, "channels must be 16, 32, 48 or 64");
    if (n <= 0 || h <= 0 || w <= 0) return set_error(SO3G_ERR_VALUE, "bad image shape");
    const int tiles_x = (w + 63) / 64;
    const int64_t nb = (int64_t)n * h * tiles_x;
    if (nb > INT32_MAX) return set_error(SO3G_ERR_VALUE, "image too large");
    hipStream_t s = (hipStream_t)stream;
    const uint16_t *i = (const uint16_t *)in, *wp = (const uint16_t *)wt;
    uint16_t *o = (uint16_t *)out;
    const dim3 g((unsigned)nb);
    switch (co) {
    case 16: hipLaunchKernelGGL(k_conv3x3<16>, g, dim3(kNB), 0, s, i, h, w, ci, wp, bias, relu, o, tiles_x); break;
    case 32: hipLaunchKernelGGL(k_conv3x3<32>, g, dim3(kNB), 0, s, i, h, w, ci, wp, bias, relu, o, tiles_x); break;
    case 48: hipLaunchKernelGGL(k_conv3x3<48>, g, dim3(kNB), 0, s, i, h, w, ci, wp, bias, relu, o, tiles_x); break;
    default: hipLaunchKernelGGL(k_conv3x3<64>, g, dim3(kNB), 0, s, i, h, w, ci, wp, bias, relu, o, tiles_x); break;
    }
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_conv3x3_wgrad(const void *in, const void *dy, int n, int h, int w, int ci, int co,
                          float *dw, float *db, void *stream) {
    if (!ch_ok(ci) || !ch_ok(co)) return set_error(SO3G_ERR_VALUE, "channels must be 16, 32, 48 or 64");
    if (n <= 0 || h <= 0 || w <= 0) return set_error(SO3G_ERR_VALUE, "bad image shape");
    const int64_t npx = (int64_t)n * h * w;
    const dim3 g((unsigned)((npx + kWgChunk - 1) / kWgChunk), 9);
    hipLaunchKernelGGL(k_conv3x3_wgrad, g, dim3(kNB), 0, (hipStream_t)stream,
                       (const uint16_t *)in, (const uint16_t *)dy, n, h, w, ci, co, dw, db);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_weight_flip(const void *wt, int ci, int co, void *wt_t, void *stream) {
    const int n = 9 * ci * co;
    hipLaunchKernelGGL(k_weight_flip, dim3((n + kNB - 1) / kNB), dim3(kNB), 0,
                       (hipStream_t)stream, (const uint16_t *)wt, ci, co, (uint16_t *)wt_t);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_relu_mask(const void *dy, const void *act, int64_t n, void *dz, void *stream) {
    hipLaunchKernelGGL(k_relu_mask, ew_grid(n), dim3(kNB), 0, (hipStream_t)stream,
                       (const uint16_t *)dy, (const uint16_t *)act, n, (uint16_t *)dz);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_add(const void *a, const void *b, int64_t n, void *out, void *stream) {
    hipLaunchKernelGGL(k_add, ew_grid(n), dim3(kNB), 0, (hipStream_t)stream,
                       (const uint16_t *)a, (const uint16_t *)b, n, (uint16_t *)out);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_pool(const void *in, int n, int h, int w, int c, int fy, int fx, void *out,
                 int backward, void *stream) {
    if (fy < 1 || fx < 1 || fy > 8 || fx > 8)
        return set_error(SO3G_ERR_VALUE, "pooling factors must be 1..8");
    if (h % fy || w % fx) return set_error(SO3G_ERR_VALUE, "image not divisible by the pooling factors");
    const int64_t cnt = (int64_t)n * h * w * c / (backward ? 1 : fy * fx);
    if (backward)
        hipLaunchKernelGGL(k_pool_bwd, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream,
                           (const uint16_t *)in, n, h, w, c, fy, fx, (uint16_t *)out);
    else
        hipLaunchKernelGGL(k_pool, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream,
                           (const uint16_t *)in, n, h, w, c, fy, fx, (uint16_t *)out);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_up(const void *in, int n, int h, int w, int c, int fy, int fx, void *out,
               int backward, void *stream) {
    if (fy < 1 || fx < 1 || fy > 8 || fx > 8)
        return set_error(SO3G_ERR_VALUE, "upsampling factors must be 1..8");
    const int64_t cnt = (int64_t)n * h * w * c * (backward ? 1 : fy * fx);
    if (backward)
        hipLaunchKernelGGL(k_up_bwd, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream,
                           (const uint16_t *)in, n, h, w, c, fy, fx, (uint16_t *)out);
    else
        hipLaunchKernelGGL(k_up, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream,
                           (const uint16_t *)in, n, h, w, c, fy, fx, (uint16_t *)out);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_from_planes(const float *x, int n, int cp, int hy, int wx, int h, int w, int c,
                        void *out, void *stream) {
    if (cp > c || hy > h || wx > w) return set_error(SO3G_ERR_VALUE, "planes exceed the image");
    const int64_t cnt = (int64_t)n * h * w * c;
    hipLaunchKernelGGL(k_from_planes, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream, x, n,
                       cp, hy, wx, h, w, c, (uint16_t *)out);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_to_planes(const void *in, int n, int h, int w, int c, int cp, int hy, int wx,
                      float *x, void *stream) {
    if (cp > c || hy > h || wx > w) return set_error(SO3G_ERR_VALUE, "planes exceed the image");
    const int64_t cnt = (int64_t)n * cp * hy * wx;
    hipLaunchKernelGGL(k_to_planes, ew_grid(cnt), dim3(kNB), 0, (hipStream_t)stream,
                       (const uint16_t *)in, n, h, w, c, cp, hy, wx, x);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1,
                 float b2, float eps, int step, void *w_bf16, void *stream) {
    if (step < 1) return set_error(SO3G_ERR_VALUE, "Adam step counts from 1");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, ew_grid(n), dim3(kNB), 0, (hipStream_t)stream, w, g, m, v, n, lr,
                       b1, b2, eps, c1, c2, (uint16_t *)w_bf16);
    NN_CHECK();
    return SO3G_OK;
}

int so3g_nn_to_bf16(const float *x, int64_t n, void *out, void *stream) {
    hipLaunchKernelGGL(k_to_bf16, ew_grid(n), dim3(kNB), 0, (hipStream_t)stream, x, n,
                       (uint16_t *)out);
    NN_CHECK();
    return SO3G_OK;
}

}  // extern "C"
