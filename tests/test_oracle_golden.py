"""Pin the CPU oracle against the reference's own code: golden vectors from
include/so_linterp.h and src/healpix_bare.c compiled in this container
(tests/golden/make_golden.py), the compiled reference itself where it is
available (oracle/_ref), and what mathematics fixes (table accuracy bound,
brute-force HEALPix pixel-centre round trips, RING/NEST consistency)."""
import ctypes
import hashlib
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
    return np.load(os.path.join(GOLD, name))


def test_tables_match_reference_hash():
    a, b = oracle.tables()
    g = _load('trig_tables.npz')
    assert hashlib.sha256(a.tobytes()).hexdigest() == str(g['sha_asin'])
    assert hashlib.sha256(b.tobytes()).hexdigest() == str(g['sha_atan2'])
    assert np.array_equal(a[g['idx']], g['asin'])
    assert np.array_equal(b[g['idx']], g['atan2'])


def test_lookups_match_golden():
    g = _load('trig_lookup.npz')
    assert np.array_equal(oracle.asin_lookup(g['asin_x']).view(np.int64),
                          g['asin_out'].view(np.int64))
    assert np.array_equal(oracle.atan2_lookup(g['atan2_y'], g['atan2_x']).view(np.int64),
                          g['atan2_out'].view(np.int64))


def test_lookup_accuracy_bound():
    """Interpolated asin/atan2 within 0.2 arcsec of the math library
    (test/test_proj_astro.py:155-163, test_trig over its grid)."""
    x = np.arange(-1., 1., .01)
    worst = np.abs(oracle.asin_lookup(x) - np.arcsin(x)).max()
    xs, ys = np.meshgrid(np.arange(-3, 3.1, 0.5), np.arange(-1, 1, .01))
    worst = max(worst, np.abs(oracle.atan2_lookup(ys.ravel(), xs.ravel()) -
                              np.arctan2(ys.ravel(), xs.ravel())).max())
    assert worst * 180 / np.pi * 3600 < 0.2


def test_healpix_matches_golden():
    h = _load('healpix.npz')
    for ns in [1, 2, 4, 16, 128, 1024, 4096, 8192]:
        for nest in (0, 1):
            out = oracle.ang2pix(ns, nest, h[f'theta_{ns}'], h[f'phi_{ns}'])
            assert np.array_equal(out, h[f'pix_{ns}_{nest}']), (ns, nest)


def test_healpix_pixel_centres_round_trip():
    """Brute force: every pixel centre of small nside maps to itself."""
    h = _load('healpix.npz')
    for ns in [1, 2, 4, 8]:
        npix = 12 * ns * ns
        for nest in (0, 1):
            th = h[f'centre_theta_{ns}_{nest}']
            ph = h[f'centre_phi_{ns}_{nest}']
            assert np.array_equal(oracle.ang2pix(ns, nest, th, ph), np.arange(npix))


def test_healpix_ring_nest_consistency():
    """RING index of a point == nest2ring(NEST index) (demo_proj_hp.py:58)."""
    h = _load('healpix.npz')
    rng = np.random.default_rng(5)
    for ns in [1, 2, 4, 8]:
        th = np.arccos(rng.uniform(-1, 1, 4000))
        ph = rng.uniform(-np.pi, np.pi, 4000)
        nest = oracle.ang2pix(ns, 1, th, ph)
        ring = oracle.ang2pix(ns, 0, th, ph)
        assert np.array_equal(h[f'nest2ring_{ns}'][nest], ring)


def test_healpix_equal_area():
    """Uniform points on the sphere fill HEALPix pixels uniformly."""
    rng = np.random.default_rng(6)
    n = 400000
    th = np.arccos(rng.uniform(-1, 1, n))
    ph = rng.uniform(-np.pi, np.pi, n)
    counts = np.bincount(oracle.ang2pix(4, 1, th, ph), minlength=192)
    assert abs(counts.mean() - n / 192) < 1e-9
    assert counts.std() < 5 * np.sqrt(n / 192)


@pytest.mark.skipif(oracle.ref_lib() is None, reason='reference not compiled here')
def test_against_compiled_reference_random():
    """Oracle == the reference's compiled code on fresh random inputs."""
    ref = oracle.ref_lib()
    rng = np.random.default_rng(99)
    x = rng.uniform(-1.2, 1.2, 20000)
    out = np.zeros_like(x)
    ref.ref_asin_lookup(ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(out.ctypes.data),
                        ctypes.c_int64(len(x)))
    assert np.array_equal(out.view(np.int64), oracle.asin_lookup(x).view(np.int64))
    y = rng.normal(size=20000)
    xx = rng.normal(size=20000)
    ref.ref_atan2_lookup(ctypes.c_void_p(y.ctypes.data), ctypes.c_void_p(xx.ctypes.data),
                         ctypes.c_void_p(out.ctypes.data), ctypes.c_int64(len(x)))
    assert np.array_equal(out.view(np.int64), oracle.atan2_lookup(y, xx).view(np.int64))
    th = np.arccos(rng.uniform(-1, 1, 20000))
    ph = rng.uniform(-np.pi, np.pi, 20000)
    for ns in (64, 2048, 8192):
        for nest in (0, 1):
            o = np.zeros(len(th), np.int64)
            ref.ref_ang2pix(ctypes.c_int64(ns), ctypes.c_int(nest),
                            ctypes.c_void_p(th.ctypes.data), ctypes.c_void_p(ph.ctypes.data),
                            ctypes.c_void_p(o.ctypes.data), ctypes.c_int64(len(th)))
            assert np.array_equal(o, oracle.ang2pix(ns, nest, th, ph))
