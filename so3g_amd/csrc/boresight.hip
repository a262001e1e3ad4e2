// boresight.hip -- per-sample boresight quaternions on the GPU
// (CelestialSightLine.for_horizon, python/proj/coords.py:190-209, and
// naive_az_el, :99-125), optionally premultiplied by a fixed rotation (the
// Projectionist's celestial -> native quaternion), so that the engines'
// q_native input is produced where it is used.
//
//   horizon:   Q = euler(2, -az) euler(1, pi/2 - el) euler(2, roll)
//   naive:     Q = euler(2, lst) euler(1, pi/2 - lat) euler(2, pi) [horizon]
//              lst = (era0 J + era1) + lon,  J = (t - epoch) / 86400
//   output:    pre Q
// Products left to right in boost's term order (ptg.h qmul); euler(axis, a)
// = (cos(a/2), sin(a/2) on the axis).  cos/sin are the device library's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/so3g_hip.h"
#include "ptg.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}
using namespace so3g;

namespace {

__device__ __forceinline__ Q4 euler(int axis, double a) {
    const double c = cos(a / 2), s = sin(a / 2);
    Q4 q{c, 0., 0., 0.};
    if (axis == 0) q.b = s;
    else if (axis == 1) q.c = s;
    else q.d = s;
    return q;
}

struct BoreArgs {
    const double *t, *az, *el, *roll;
    double roll0;
    int64_t n;
    int mode;                 // 0 horizon, 1 naive az/el
    double era0, era1, epoch, lon, lat;
    Q4 pre;
    int has_pre;
};

__global__ __launch_bounds__(256) void k_boresight(BoreArgs a, double *out) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * 256) {
        const double roll = a.roll ? a.roll[i] : a.roll0;
        Q4 q;
        if (a.mode == 1) {
            const double J = (a.t[i] - a.epoch) / 86400;
            const double lst = (a.era0 * J + a.era1) + a.lon;
            q = euler(2, lst);
            q = qmul(q, euler(1, M_PI / 2 - a.lat));
            q = qmul(q, euler(2, M_PI));
            q = qmul(q, euler(2, -a.az[i]));
        } else {
            q = euler(2, -a.az[i]);
        }
        q = qmul(q, euler(1, M_PI / 2 - a.el[i]));
        q = qmul(q, euler(2, roll));
        if (a.has_pre) q = qmul(a.pre, q);
        double2 *o = reinterpret_cast<double2 *>(out + 4 * i);
        o[0] = double2{q.a, q.b};
        o[1] = double2{q.c, q.d};
    }
}

}  // namespace

extern "C" {

int so3g_boresight(int mode, const double *t, const double *az, const double *el,
                   const double *roll, double roll0, int64_t n, const double *site,
                   const double *pre, double *out, void *stream) {
    if (n < 0) return set_error(SO3G_ERR_RUNTIME, "negative sample count");
    if (n == 0) return SO3G_OK;
    if (az == nullptr || el == nullptr || out == nullptr)
        return set_error(SO3G_ERR_VALUE, "az, el and out are required");
    if (mode != 0 && mode != 1) return set_error(SO3G_ERR_VALUE, "mode must be 0 or 1");
    if (mode == 1 && (t == nullptr || site == nullptr))
        return set_error(SO3G_ERR_VALUE, "naive az/el needs t and the site");
    BoreArgs a{};
    a.t = t; a.az = az; a.el = el; a.roll = roll; a.roll0 = roll0; a.n = n; a.mode = mode;
    if (mode == 1) {
        a.era0 = site[0]; a.era1 = site[1]; a.epoch = site[2]; a.lon = site[3]; a.lat = site[4];
    }
    a.has_pre = pre != nullptr;
    if (pre) a.pre = Q4{pre[0], pre[1], pre[2], pre[3]};
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_boresight, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, a, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SO3G_ERR_HIP, std::string("boresight: ") + hipGetErrorString(e));
    return SO3G_OK;
}

}  // extern "C"
