// san_driver.cpp -- exercises the CPU oracle (so3g_oracle.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer: every projection, flat
// non-tiled / tiled maps with nearest and bilinear interpolation, HEALPix
// NEST / RING / tiled, clipped and unclipped footprints, the Precomp
// operations and the thread-segment path.  Test infrastructure only (built
// by `make -C oracle sanitize`, never linked into the product); exits 0
// when no sanitizer fires.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
typedef struct {
    int32_t proj, pixelizor, tiled, interp, ncomp, ny, nx;
    double cdelt_y, cdelt_x, crpix_y, crpix_x;
    int32_t tile_y, tile_x;
    int32_t nside, nest, nside_tile;
} oracle_geom;
void oracle_coords(const oracle_geom *, const double *, int64_t, const double *, int64_t,
                   double *);
void oracle_pixels(const oracle_geom *, const double *, int64_t, const double *, int64_t,
                   int32_t *);
void oracle_pointing_matrix(const oracle_geom *, const double *, int64_t, const double *,
                            int64_t, const float *, int32_t *, float *);
int oracle_from_map(const oracle_geom *, const double *, double *const *, const double *,
                    int64_t, const double *, int64_t, const float *, float *);
int oracle_to_map(const oracle_geom *, double *, double *const *, const double *, int64_t,
                  const double *, int64_t, const float *, const float *, const float *,
                  const int32_t *, int64_t);
int oracle_to_weights(const oracle_geom *, double *, double *const *, const double *,
                      int64_t, const double *, int64_t, const float *, const float *,
                      const int32_t *, int64_t);
void oracle_tile_hits(const oracle_geom *, const double *, int64_t, const double *, int64_t,
                      int32_t *);
int oracle_precomp_from_map(const oracle_geom *, const double *, double *const *,
                            const int32_t *, int, const float *, int, int64_t, int64_t,
                            float *);
int oracle_precomp_to_map(const oracle_geom *, double *, double *const *, const int32_t *, int,
                          const float *, int, int64_t, int64_t, const float *, const float *,
                          const int32_t *, int64_t);
int oracle_precomp_to_weights(const oracle_geom *, double *, double *const *, const int32_t *,
                              int, const float *, int, int64_t, int64_t, const float *,
                              const int32_t *, int64_t);
}

namespace {

struct Q { double a, b, c, d; };

Q mul(const Q &l, const Q &r) {
    return {l.a * r.a - l.b * r.b - l.c * r.c - l.d * r.d,
            l.a * r.b + l.b * r.a + l.c * r.d - l.d * r.c,
            l.a * r.c - l.b * r.d + l.c * r.a + l.d * r.b,
            l.a * r.d + l.b * r.c - l.c * r.b + l.d * r.a};
}
Q rz(double x) { return {std::cos(x / 2), 0, 0, std::sin(x / 2)}; }
Q ry(double x) { return {std::cos(x / 2), 0, std::sin(x / 2), 0}; }

int failures = 0;

void check(int rc, const char *what) {
    if (rc != 0) {
        std::printf("%s returned %d\n", what, rc);
        failures++;
    }
}

// One engine configuration over the synthetic scan.
void run(const oracle_geom &g, const std::vector<double> &bore, const std::vector<double> &dets,
         int64_t n_t, int64_t n_det) {
    const int nc = g.ncomp;
    const int k = g.pixelizor ? (g.tiled ? 2 : 1) : (g.tiled ? 3 : 2);
    std::vector<float> resp(2 * n_det, 1.f), sig(n_det * n_t), dw(n_det, 1.f);
    for (int64_t i = 0; i < n_det * n_t; i++) sig[i] = (float)std::sin(0.1 * i);
    std::vector<double> co(4 * n_det * n_t);
    oracle_coords(&g, bore.data(), n_t, dets.data(), n_det, co.data());
    std::vector<int32_t> pix(k * n_det * n_t);
    std::vector<float> spin(nc * n_det * n_t);
    oracle_pointing_matrix(&g, bore.data(), n_t, dets.data(), n_det, resp.data(), pix.data(),
                           spin.data());
    oracle_pixels(&g, bore.data(), n_t, dets.data(), n_det, pix.data());
    // maps: one plane per component (and per weight slot), tiles all active
    int64_t plane = 0, n_tile = 0, tpix = 0;
    if (g.pixelizor) {
        plane = 12LL * g.nside * g.nside;
        if (g.tiled) {
            n_tile = 12LL * g.nside_tile * g.nside_tile;
            tpix = plane / n_tile;
        }
    } else {
        plane = (int64_t)g.ny * g.nx;
        if (g.tiled) n_tile = (int64_t)((g.ny + g.tile_y - 1) / g.tile_y) *
                              ((g.nx + g.tile_x - 1) / g.tile_x);
    }
    std::vector<std::vector<double>> tiles, wtiles;
    std::vector<double *> tp, wp;
    std::vector<double> map, wmap;
    if (g.tiled) {
        const int ntx = g.pixelizor ? 0 : (g.nx + g.tile_x - 1) / g.tile_x;
        for (int64_t t = 0; t < n_tile; t++) {
            int64_t sz = tpix;
            if (!g.pixelizor) {
                const int iy = (int)(t / ntx), ix = (int)(t % ntx);
                sz = (int64_t)std::min(g.tile_y, g.ny - iy * g.tile_y) *
                     std::min(g.tile_x, g.nx - ix * g.tile_x);
            }
            tiles.emplace_back(nc * sz, 0.5);
            wtiles.emplace_back(nc * nc * sz, 0.);
        }
        for (auto &t : tiles) tp.push_back(t.data());
        for (auto &t : wtiles) wp.push_back(t.data());
    } else {
        map.assign(nc * plane, 0.5);
        wmap.assign(nc * nc * plane, 0.);
    }
    double *m = g.tiled ? nullptr : map.data();
    double *w = g.tiled ? nullptr : wmap.data();
    double *const *mt = g.tiled ? tp.data() : nullptr;
    double *const *wt = g.tiled ? wp.data() : nullptr;
    check(oracle_from_map(&g, m, mt, bore.data(), n_t, dets.data(), n_det, resp.data(),
                          sig.data()), "from_map");
    check(oracle_to_map(&g, m, mt, bore.data(), n_t, dets.data(), n_det, resp.data(),
                        sig.data(), dw.data(), nullptr, 0), "to_map");
    const int32_t seg[6] = {0, 0, (int32_t)(n_t / 2), (int32_t)(n_det - 1), 3, (int32_t)n_t};
    check(oracle_to_map(&g, m, mt, bore.data(), n_t, dets.data(), n_det, resp.data(),
                        sig.data(), nullptr, seg, 2), "to_map segments");
    check(oracle_to_weights(&g, w, wt, bore.data(), n_t, dets.data(), n_det, resp.data(),
                            dw.data(), nullptr, 0), "to_weights");
    if (g.tiled && !g.pixelizor) {
        std::vector<int32_t> hits(n_tile);
        oracle_tile_hits(&g, bore.data(), n_t, dets.data(), n_det, hits.data());
    }
    if (!g.interp) {
        check(oracle_precomp_from_map(&g, m, mt, pix.data(), k, spin.data(), nc, n_det, n_t,
                                      sig.data()), "precomp_from_map");
        check(oracle_precomp_to_map(&g, m, mt, pix.data(), k, spin.data(), nc, n_det, n_t,
                                    sig.data(), dw.data(), nullptr, 0), "precomp_to_map");
        check(oracle_precomp_to_weights(&g, w, wt, pix.data(), k, spin.data(), nc, n_det, n_t,
                                        dw.data(), seg, 2), "precomp_to_weights");
    }
}

}  // namespace

int main() {
    // constant-elevation scan: el 50 deg, az +-10 deg, 3 detectors
    const int64_t n_t = 3000, n_det = 3;
    const double deg = M_PI / 180;
    std::vector<double> bore(4 * n_t), dets(4 * n_det);
    for (int64_t t = 0; t < n_t; t++) {
        const double ph = std::fmod(t * 0.01, 4.0);
        const double az = (ph < 2 ? ph - 1 : 3 - ph) * 10 * deg;
        const Q q = mul(rz(-az), ry(M_PI / 2 - 50 * deg));
        bore[4 * t] = q.a; bore[4 * t + 1] = q.b; bore[4 * t + 2] = q.c; bore[4 * t + 3] = q.d;
    }
    for (int64_t d = 0; d < n_det; d++) {
        const Q q = mul(rz(0.7 * d), ry((0.2 + 0.1 * d) * deg));
        dets[4 * d] = q.a; dets[4 * d + 1] = q.b; dets[4 * d + 2] = q.c; dets[4 * d + 3] = q.d;
    }
    for (int proj = 0; proj < 8; proj++)
        for (int tiled = 0; tiled < 2; tiled++)
            for (int interp = 0; interp < 2; interp++)
                for (int clip = 0; clip < 2; clip++) {
                    oracle_geom g{};
                    g.proj = proj;
                    g.tiled = tiled;
                    g.interp = interp;
                    g.ncomp = 3;
                    g.ny = 40;
                    g.nx = 60;
                    // centre the map on this projection's mean coordinates
                    std::vector<double> co(4 * n_det * n_t);
                    oracle_coords(&g, bore.data(), n_t, dets.data(), n_det, co.data());
                    double mx = 0, my = 0;
                    for (int64_t i = 0; i < n_det * n_t; i++) {
                        mx += co[4 * i];
                        my += co[4 * i + 1];
                    }
                    mx /= (double)(n_det * n_t);
                    my /= (double)(n_det * n_t);
                    if (!std::isfinite(mx) || !std::isfinite(my)) mx = my = 0;
                    g.cdelt_y = (clip ? 0.2 : 1.) * std::max(1e-9, std::fabs(my)) * 0.05;
                    g.cdelt_x = -(clip ? 0.2 : 1.) * std::max(1e-9, std::fabs(mx) + 0.2) * 0.05;
                    g.crpix_y = 20.5 - my / g.cdelt_y;
                    g.crpix_x = 30.5 - mx / g.cdelt_x;
                    g.tile_y = 7;
                    g.tile_x = 11;
                    std::vector<int32_t> pix(2 * n_det * n_t);
                    oracle_geom gp = g;
                    gp.tiled = 0;
                    oracle_pixels(&gp, bore.data(), n_t, dets.data(), n_det, pix.data());
                    int64_t on = 0;
                    for (int64_t i = 0; i < n_det * n_t; i++) on += pix[2 * i] >= 0;
                    std::printf("proj %d tiled %d interp %d clip %d: %.3f on the map\n", proj,
                                tiled, interp, clip, (double)on / (double)(n_det * n_t));
                    run(g, bore, dets, n_t, n_det);
                }
    for (int nest = 0; nest < 2; nest++)
        for (int tiled = 0; tiled < 2; tiled++) {
            if (tiled && !nest) continue;
            oracle_geom g{};
            g.proj = 2;
            g.pixelizor = 1;
            g.ncomp = 3;
            g.nside = 64;
            g.nest = nest;
            g.tiled = tiled;
            g.nside_tile = 4;
            run(g, bore, dets, n_t, n_det);
        }
    std::printf("sanitized oracle run: %d failures\n", failures);
    return failures ? 1 : 0;
}
