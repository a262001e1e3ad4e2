"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/so3g_hip.h declares, engine classes and geometry
validation mirror the reference's names and errors, and the Python
helpers (quat, Ranges, geometry, thread-interval flattening) behave."""
import os
import re

import numpy as np
import pytest

import so3g_amd
from so3g_amd import _native, engine
from so3g_amd.proj import quat, Ranges, RangesMatrix
from so3g_amd.proj.wcs import Projectionist, ProjectionistHealpix, _Tiling, wrap_ivals

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    lib = _native.load()
    hdr = open(os.path.join(ROOT, 'include', 'so3g_hip.h')).read()
    names = set(re.findall(r'\b(so3g_[a-z0-9_]+)\s*\(', hdr))
    assert len(names) >= 18
    for n in sorted(names):
        assert hasattr(lib, n), n
        # and a ctypes signature, so 64-bit parameters are not passed as ints
        assert n in _native.exported_symbols(), n
    assert lib.so3g_abi_version() == 1


def test_host_tables_match_golden():
    import hashlib
    a, b = _native.get_tables()
    g = np.load(os.path.join(ROOT, 'tests', 'golden', 'trig_tables.npz'))
    assert hashlib.sha256(a.tobytes()).hexdigest() == str(g['sha_asin'])
    assert hashlib.sha256(b.tobytes()).hexdigest() == str(g['sha_atan2'])


def test_engine_names():
    # 8 projections x 3 spins x 2 tilings x 2 interpolations + 6 HEALPix + 2 Precomp,
    # the names registered at Projection.cxx:2337-2357.
    names = [n for n in dir(so3g_amd) if n.startswith('ProjEng_')]
    assert len(names) == 104
    for n in ['ProjEng_CAR_TQU_NonTiled', 'ProjEng_ZEA_QU_Tiled_Bilinear',
              'ProjEng_HP_T_Tiled', 'ProjEng_Precomp_NonTiled', 'ProjEng_Flat_T_NonTiled']:
        assert n in names
    e = so3g_amd.ProjEng_CAR_QU_Tiled((10, 20, 1., 1., 1., 1., 4, 6))
    assert (e.index_count, e.comp_count) == (3, 2)
    assert so3g_amd.ProjEng_HP_T_NonTiled((64, 1)).index_count == 1


def test_geometry_errors():
    with pytest.raises(ValueError, match='Invalid nside 100'):
        so3g_amd.ProjEng_HP_T_NonTiled((100, 1))
    with pytest.raises(ValueError):
        so3g_amd.ProjEng_HP_T_NonTiled((8192, 1))
    assert so3g_amd.ProjEng_HP_T_NonTiled((8192, 1), allow_nside_max=True).index_count == 1
    with pytest.raises(ValueError, match='RING not supported'):
        so3g_amd.ProjEng_HP_T_Tiled((64, 0, 4, None))
    with pytest.raises(ValueError, match='nside_tile'):
        so3g_amd.ProjEng_HP_T_Tiled((64, 1, 128, None))


def test_zeros_shapes():
    e = so3g_amd.ProjEng_CAR_T_Tiled((10, 21, 1., 1., 1., 1., 4, 6, [0, 3, 11]))
    z = e.zeros((2,))
    assert len(z) == 12
    assert z[0].shape == (2, 4, 6) and z[3].shape == (2, 4, 3) and z[11].shape == (2, 2, 3)
    assert z[1] is None
    with pytest.raises(RuntimeError):
        so3g_amd.ProjEng_CAR_T_Tiled((10, 21, 1., 1., 1., 1., 4, 6)).zeros((1,))
    assert so3g_amd.ProjEng_TAN_TQU_NonTiled((5, 7, 1., 1., 1., 1.)).zeros(3).shape == (3, 5, 7)
    hp = so3g_amd.ProjEng_HP_QU_Tiled((16, 1, 2, [1, 5]))
    z = hp.zeros((2,))
    assert len(z) == 48 and z[1].shape == (2, 64) and z[0] is None


def test_quat_algebra():
    rng = np.random.default_rng(0)
    a = quat.Quat(rng.normal(size=(50, 4)))
    b = quat.Quat(rng.normal(size=(50, 4)))
    c = quat.Quat(rng.normal(size=(50, 4)))
    one = quat.quat(1, 0, 0, 0)
    assert np.allclose(np.asarray(a * one), np.asarray(a))
    assert np.allclose(np.asarray((a * b) * c), np.asarray(a * (b * c)))
    # |ab| = |a||b|
    n = lambda q: np.sqrt((np.asarray(q) ** 2).sum(-1))
    assert np.allclose(n(a * b), n(a) * n(b))
    # i*j = k, j*i = -k
    i, j = quat.quat(0, 1, 0, 0), quat.quat(0, 0, 1, 0)
    assert np.array_equal(np.asarray(i * j), [0, 0, 0, 1])
    assert np.array_equal(np.asarray(j * i), [0, 0, 0, -1])
    # Scalars multiply elementwise, as on spt3g quats.
    assert np.all(np.isnan(np.asarray(a[1] * np.nan)))
    x, y, z, w = np.transpose(a)
    assert isinstance(x, np.ndarray) and not isinstance(x, quat.Quat)


def test_rotation_round_trips():
    """decompose_* inverts rotation_* (test/test_proj_quat.py:23-69)."""
    rng = np.random.default_rng(1)
    lon = rng.uniform(-np.pi, np.pi, 100)
    lat = rng.uniform(-1.4, 1.4, 100)
    psi = rng.uniform(-np.pi, np.pi, 100)
    q = quat.rotation_lonlat(lon, lat, psi)
    l2, b2, p2 = quat.decompose_lonlat(q)
    assert np.allclose(l2, lon) and np.allclose(b2, lat) and np.allclose(p2, psi)
    xi, eta, gamma = rng.uniform(-.1, .1, (3, 100))
    x2, e2, g2 = quat.decompose_xieta(quat.rotation_xieta(xi, eta, gamma))
    assert np.allclose(x2, xi) and np.allclose(e2, eta)
    assert np.allclose(np.angle(np.exp(1j * (g2 - gamma))), 0)


def test_ranges_basic():
    r = Ranges(100)
    r.add_interval(10, 20).add_interval(15, 30).add_interval(50, 60)
    assert r.ranges().tolist() == [[10, 30], [50, 60]]
    assert (~r).ranges().tolist() == [[0, 10], [30, 50], [60, 100]]
    assert Ranges.from_mask(r.mask()).ranges().tolist() == r.ranges().tolist()
    rm = RangesMatrix.zeros((3, 4, 100))
    assert rm.shape == (3, 4, 100) and rm.mask().sum() == 0
    r2 = Ranges(100).append_interval_no_check(0, 5).append_interval_no_check(5, 9)
    assert r2.ranges().tolist() == [[0, 9]]


def test_flatten_intervals():
    n_det, n_t = 3, 50
    rows = [Ranges(n_t).add_interval(d, d + 10) for d in range(n_det)]
    segs = engine._flatten_intervals(RangesMatrix(rows), n_det, n_t)
    assert segs.tolist() == [[0, 0, 10], [1, 1, 11], [2, 2, 12]]
    nested = [RangesMatrix([RangesMatrix(rows), RangesMatrix(rows)])]
    assert len(engine._flatten_intervals(nested, n_det, n_t)) == 6
    with pytest.raises(RuntimeError, match='n_det'):
        engine._flatten_intervals(RangesMatrix(rows[:2]), n_det, n_t)
    with pytest.raises(RuntimeError, match='n_time'):
        engine._flatten_intervals(RangesMatrix(rows), n_det, n_t + 1)


def test_rle_ranges():
    dom = np.array([[-1, 0, 0, 1, 1, 3, 0, -1],
                    [2, 2, 2, 2, 2, 2, 2, 2]], np.int32)
    out = engine._rle_ranges(dom, 3, 8)
    assert out[0][0][0].ranges().tolist() == [[1, 3], [6, 7]]
    assert out[0][1][0].ranges().tolist() == [[3, 5]]
    assert out[1][0][0].ranges().tolist() == [[5, 6]]
    assert out[0][2][1].ranges().tolist() == [[0, 8]]
    w = wrap_ivals(out)
    assert w[0].shape == (3, 2, 8) and w[1].shape == (1, 2, 8)


def test_projectionist_args_and_errors():
    from so3g_amd.proj.wcslite import make_wcs
    wcs = make_wcs('CAR', (0., 0.), (-1 / 60., 1 / 60.), (100.5, -2990.))
    p = Projectionist.for_geom((60, 200), wcs)
    args = p._get_pixelizor_args()
    assert args[:2] == (60, 200)
    assert np.isclose(args[2], np.pi / 180 / 60) and np.isclose(args[3], -np.pi / 180 / 60)
    assert p.get_ProjEng('TQU', get=False) == 'ProjEng_CAR_TQU_NonTiled'
    pt = Projectionist.for_tiled((60, 200), wcs, (20, 30))
    assert pt.tiling.tile_count == 3 * 7 and len(pt._get_pixelizor_args()) == 9
    with pytest.raises(ValueError):
        p.to_map(np.zeros((1, 5), np.float32), None)
    with pytest.raises(ValueError, match='interpol'):
        p.get_ProjEng('T', interpol='cubic')
    with pytest.raises(ValueError, match='no projector'):
        p.get_ProjEng('IQUV')
    h = ProjectionistHealpix.for_healpix(64, ordering='RING')
    assert h._get_pixelizor_args() == (64, 0, None, None)
    with pytest.raises(NotImplementedError):
        ProjectionistHealpix.for_healpix(64, 4, ordering='RING')
    t = _Tiling((60, 200), (20, 30))
    assert t.tile_dims(6) == (20, 20) and t.tile_offset(8) == (20, 30)


def test_compute_path_refuses_without_gpu():
    """No CPU fallback: without a GPU the engines raise."""
    import torch
    if torch.cuda.is_available():
        pytest.skip('GPU present')
    e = so3g_amd.ProjEng_CAR_T_NonTiled((5, 5, 1., 1., 1., 1.))
    with pytest.raises(RuntimeError, match='no CPU fallback'):
        e.pixels(np.zeros((3, 4)), np.zeros((2, 4)))


def test_plan_options():
    """Plan tunables documented in include/so3g_hip.h: defaults, round trip,
    and rejection of values outside their ranges."""
    defaults = {'plan_layout': -1, 'plan_tile_item': 262144, 'plan_time_copy': 1,
                'plan_to_lds': 1, 'plan_to_cpb': 2, 'plan_from_layout': 1, 'plan_from_v': 4,
                'nmat_dchunk': 64, 'plan_win_len': 65536, 'plan_win_block': 1024,
                'plan_win_flush': 1, 'scatter_hash': 1, 'hash_r': 1}
    for name, v in defaults.items():
        assert _native.get_option(name) == v, name
    for name, good, bad in (('plan_layout', 2, 3), ('plan_layout', 1, -2),
                            ('plan_win_len', 512, 511), ('plan_win_block', 256, 512),
                            ('plan_tile_item', 4096, 63), ('plan_to_cpb', 5, 0),
                            ('plan_from_v', 8, 5), ('nmat_dchunk', 128, 0),
                            ('hash_r', 4, 2), ('scatter_hash', 0, 5)):
        old = _native.get_option(name)
        try:
            _native.set_option(name, good)
            assert _native.get_option(name) == good
            with pytest.raises(ValueError):
                _native.set_option(name, bad)
            assert _native.get_option(name) == good
        finally:
            _native.set_option(name, old)
    with pytest.raises(ValueError):
        _native.set_option('plan_no_such_option', 1)
    # the tiled plan's tile shape is a build constant, readable only
    assert (_native.get_option('plan_tile_rows'), _native.get_option('plan_tile_cols')) == (32, 64)
    with pytest.raises(ValueError):
        _native.set_option('plan_tile_rows', 64)


def test_plan_argument_errors():
    """Plan entry points reject bad arguments on the host, before any device
    work, with the reference's exception classes."""
    import ctypes
    lib = _native.load()
    h = ctypes.c_void_p()
    shape = np.array([10, 10], np.int64)
    dims = np.array([[4, 4]], np.int32)
    cases = [
        # k must be 1 or 2 here (tiled maps have their own entry point)
        (RuntimeError, lambda: lib.so3g_pc_plan_create(None, 3, None, 3, 1, 1, shape.ctypes.data,
                                                       -1, ctypes.byref(h), None)),
        (RuntimeError, lambda: lib.so3g_pc_plan_create(None, 2, None, 4, 1, 1, shape.ctypes.data,
                                                       -1, ctypes.byref(h), None)),
        (RuntimeError, lambda: lib.so3g_pc_plan_create(None, 2, None, 3, 1 << 22, 1,
                                                       shape.ctypes.data, -1, ctypes.byref(h),
                                                       None)),
        # the detector index's float-reciprocal division needs n_t < 2^30,
        # and sample indices must fit 32 bits
        (RuntimeError, lambda: lib.so3g_pc_plan_create(None, 2, None, 3, 1, 1 << 30,
                                                       shape.ctypes.data, -1, ctypes.byref(h),
                                                       None)),
        (RuntimeError, lambda: lib.so3g_pc_plan_create(None, 2, None, 3, 3, (1 << 30) - 1,
                                                       shape.ctypes.data, -1, ctypes.byref(h),
                                                       None)),
        (ValueError, lambda: lib.so3g_pc_plan_create_tiled(None, None, 3, 1, 1, None, 0, -1,
                                                           ctypes.byref(h), None)),
        (ValueError, lambda: lib.so3g_pc_plan_create_tiled(None, None, 3, 1, 1,
                                                           np.array([[-1, 4]], np.int32).ctypes.data,
                                                           1, -1, ctypes.byref(h), None)),
        (RuntimeError, lambda: lib.so3g_pc_plan_create_tiled(None, None, 0, 1, 1, dims.ctypes.data,
                                                             1, -1, ctypes.byref(h), None)),
        (ValueError, lambda: lib.so3g_pc_plan_to_map_tiles(None, None, None, None, None)),
        (ValueError, lambda: lib.so3g_pc_plan_from_map_tiles(None, None, None, 0, None)),
        (ValueError, lambda: lib.so3g_pc_plan_from_map(None, None, None, 0, None)),
    ]
    for exc, call in cases:
        with pytest.raises(exc):
            _native.check(call())
