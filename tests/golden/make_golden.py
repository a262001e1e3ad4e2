"""Write the golden vectors under tests/golden/ from the reference's own
trig-table and HEALPix code (include/so_linterp.h and src/healpix_bare.c,
compiled from /root/reference by `make -C oracle ref` into
oracle/_ref/libso3g_ref.so).  Inputs are seeded, so re-running reproduces
the files bit for bit on the same C library.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import hashlib
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(ROOT, 'oracle', '_ref', 'libso3g_ref.so')


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def trig_inputs(rng):
    n = 1 << 14
    dx = 1. / (n - 1)
    # asin: grid nodes, their neighbours, random points, signs, the clamp.
    nodes = np.arange(0, n, 37) * dx
    xa = np.concatenate([
        nodes, np.nextafter(nodes, 2), np.nextafter(nodes, -2),
        rng.uniform(-1, 1, 3000), [0., -0., 1., -1., 1 + 1e-15, -1 - 1e-15, 1e-300,
                                   -1e-300, 0.5, -0.5, 1 - dx / 2, dx / 3],
    ])
    # atan2: all quadrants, y == 0, |x| < |y| and |x| > |y|, x == y.
    ang = rng.uniform(-np.pi, np.pi, 3000)
    r = rng.uniform(1e-3, 3, 3000)
    ya = np.concatenate([r * np.sin(ang), [0., 0., -0., 1., -1., 1., -1., 0.5, 2., 1e-12]])
    xb = np.concatenate([r * np.cos(ang), [1., -1., 1., 1., 1., -1., -1., 0.5, 1e-12, 2.]])
    return xa, ya, xb


def healpix_inputs(rng, nside):
    k = 1000
    z = rng.uniform(-1, 1, k)
    theta = np.arccos(z)
    phi = rng.uniform(-2 * np.pi, 2 * np.pi, k)
    # Edge-adjacent cases: the polar/equatorial boundary z = +-2/3, the
    # poles, phi wrapping, and exact multiples of pi/2.
    t23 = np.arccos(2. / 3.)
    edges_t = np.array([0., 1e-12, np.pi, np.pi - 1e-12, t23, np.nextafter(t23, 0),
                        np.nextafter(t23, 4), np.pi - t23, np.pi / 2, np.pi / 2, np.pi / 2,
                        np.pi / 4, 3 * np.pi / 4, 1.0, 2.0])
    edges_p = np.array([0., 0., 0., 0., 0.3, 0.3, 0.3, 1.0, 0., np.pi / 2, -np.pi, 2 * np.pi,
                        -1e-14, np.pi, 7.0])
    return np.concatenate([theta, edges_t]), np.concatenate([phi, edges_p])


def main():
    if not os.path.exists(REF):
        subprocess.run(['make', '-s', '-C', os.path.join(ROOT, 'oracle'), 'ref'], check=True)
    lib = ctypes.CDLL(REF)
    rng = np.random.default_rng(20240601)

    # Tables.
    n = 1 << 14
    a = np.zeros(n)
    b = np.zeros(n)
    lib.ref_tables(p(a), p(b))
    sha_a = hashlib.sha256(a.tobytes()).hexdigest()
    sha_b = hashlib.sha256(b.tobytes()).hexdigest()
    idx = np.arange(0, n, 64)
    np.savez_compressed(os.path.join(OUT, 'trig_tables.npz'), idx=idx, asin=a[idx],
                        atan2=b[idx], sha_asin=np.array(sha_a), sha_atan2=np.array(sha_b))

    # Lookups.
    xa, ya, xb = trig_inputs(rng)
    oa = np.zeros_like(xa)
    lib.ref_asin_lookup(p(xa), p(oa), ctypes.c_int64(len(xa)))
    ob = np.zeros_like(xb)
    lib.ref_atan2_lookup(p(ya), p(xb), p(ob), ctypes.c_int64(len(xb)))
    np.savez_compressed(os.path.join(OUT, 'trig_lookup.npz'), asin_x=xa, asin_out=oa,
                        atan2_y=ya, atan2_x=xb, atan2_out=ob)

    # HEALPix ang2nest / ang2ring.
    rec = {}
    for nside in [1, 2, 4, 16, 128, 1024, 4096, 8192]:
        th, ph = healpix_inputs(rng, nside)
        for nest in (0, 1):
            out = np.zeros(len(th), np.int64)
            lib.ref_ang2pix(ctypes.c_int64(nside), ctypes.c_int(nest), p(th), p(ph), p(out),
                            ctypes.c_int64(len(th)))
            rec[f'pix_{nside}_{nest}'] = out
        rec[f'theta_{nside}'] = th
        rec[f'phi_{nside}'] = ph
    # Pixel-centre angles for small nside (brute-force round trips).
    for nside in [1, 2, 4, 8]:
        npix = 12 * nside * nside
        for nest in (0, 1):
            pix = np.arange(npix, dtype=np.int64)
            th = np.zeros(npix)
            ph = np.zeros(npix)
            lib.ref_pix2ang(ctypes.c_int64(nside), ctypes.c_int(nest), p(pix), p(th), p(ph),
                            ctypes.c_int64(npix))
            rec[f'centre_theta_{nside}_{nest}'] = th
            rec[f'centre_phi_{nside}_{nest}'] = ph
        ring = np.zeros(npix, np.int64)
        allpix = np.arange(npix, dtype=np.int64)
        lib.ref_nest2ring(ctypes.c_int64(nside), p(allpix), p(ring), ctypes.c_int64(npix))
        rec[f'nest2ring_{nside}'] = ring
    np.savez_compressed(os.path.join(OUT, 'healpix.npz'), **rec)
    print('wrote', sorted(os.listdir(OUT)))


if __name__ == '__main__':
    sys.exit(main())
