"""ProjEng_Precomp contract checks beyond the oracle comparisons of
test_gpu_parity: cached pixel indices outside the map, and the calls
staying usable afterwards."""
import numpy as np
import pytest

import oracle
import so3g_amd
from helpers import Case, flat_map

pytestmark = pytest.mark.gpu


def _pmat(c):
    return [np.array(x) for x in c.engine.pointing_matrix(c.bore, c.dets, c.resps)]


@pytest.mark.parametrize('n_samp', [3000, 2999])
def test_precomp_index_out_of_range_flat(gpu, n_samp):
    """An index outside a non-tiled (ny, nx) map -- which the reference
    does not check (Projection.cxx:2066-2154 index the map directly) --
    raises RuntimeError from every Precomp method and from plan creation.
    n_samp % 4 == 0 takes the packed-map gather, odd n_samp the generic one."""
    c = Case('CAR', comps='TQU', n_samp=n_samp)
    pix, spin = _pmat(c)
    ny, nx = c.pixargs[0], c.pixargs[1]
    pe = so3g_amd.ProjEng_Precomp_NonTiled()
    m = c.map_like([3])
    for det, t, slot, val in ((0, 5, 0, ny), (1, 7, 1, nx), (2, 9, 1, -1), (3, n_samp - 1, 0,
                                                                           1 << 30)):
        p = pix.copy()
        p[det, t, slot] = val
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.from_map(m, p, spin)
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.to_map(c.zeros([3]), p, spin, c.signal)
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.to_weight_map(c.zeros([3, 3]), p, spin)
        with pytest.raises(RuntimeError, match='outside the map'):
            so3g_amd.PrecompPlan(p, spin, (ny, nx))
    # the flags are reset per call: valid indices work again
    got = np.array(pe.from_map(m, pix, spin))
    assert np.array_equal(got.view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin).view(np.int32))
    g = pe.to_map(c.zeros([3]), pix, spin, c.signal)
    r = oracle.precomp_to_map(c.og, c.zeros([3]), pix, spin, c.signal)
    assert np.abs(g - r).max() < 1e-10


def test_precomp_index_out_of_range_healpix(gpu):
    c = Case('HP', comps='TQU', nside=128)
    pix, spin = _pmat(c)
    npix = 12 * 128 * 128
    pe = so3g_amd.ProjEng_Precomp_NonTiled()
    m = np.random.default_rng(1).standard_normal((3, npix))
    p = pix.copy()
    p[1, 11, 0] = npix
    with pytest.raises(RuntimeError, match='outside the map'):
        pe.from_map(m, p, spin)
    with pytest.raises(RuntimeError, match='outside the map'):
        pe.to_map(np.zeros((3, npix)), p, spin, c.signal)
    with pytest.raises(RuntimeError, match='outside the map'):
        so3g_amd.PrecompPlan(p, spin, (npix,))
    got = np.array(pe.from_map(m, pix, spin))
    assert np.array_equal(got.view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin).view(np.int32))


def test_precomp_index_out_of_range_tiled(gpu):
    """Tiled maps: a tile number past the tiling, and a row or column past
    the tile's own shape."""
    c = Case('CAR', comps='TQU', tiled=True)
    pix, spin = _pmat(c)
    pe = so3g_amd.ProjEng_Precomp_Tiled()
    m = c.map_like([3])
    n_tile = len(m)
    th, tw = c.pixargs[6], c.pixargs[7]
    for det, t, slot, val in ((0, 3, 0, n_tile), (1, 4, 1, th), (2, 5, 2, tw), (2, 6, 2, -3)):
        p = pix.copy()
        p[det, t, slot] = val
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.from_map(m, p, spin)
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.to_map(c.zeros([3]), p, spin, c.signal)
        with pytest.raises(RuntimeError, match='outside the map'):
            pe.to_weight_map(c.zeros([3, 3]), p, spin)
    got = np.array(pe.from_map(m, pix, spin))
    assert np.array_equal(got.view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin).view(np.int32))
    g = pe.to_map(c.zeros([3]), pix, spin, c.signal)
    r = oracle.precomp_to_map(c.og, c.zeros([3]), pix, spin, c.signal)
    assert np.abs(flat_map(g) - flat_map(r)).max() < 1e-10


def _torch_pmat(c, gpu):
    import torch
    bore, dets, resp = [torch.from_numpy(np.ascontiguousarray(a)).to(gpu)
                        for a in (c.bore, c.dets, c.resps)]
    return c.engine.pointing_matrix(bore, dets, resp)


@pytest.mark.parametrize('proj,tiled', [('CAR', False), ('CAR', True), ('HP', False)])
def test_precomp_cached_plan(gpu, proj, tiled):
    """ProjEng_Precomp with torch pointing matrices routes through a cached
    PrecompPlan: from_map bit-exact with the oracle's Precomp (fresh and
    accumulating), to_map and to_weight_map within float64 reordering, the
    plan reused while the tensors are unchanged and rebuilt after an
    in-place edit; thread_intervals and plan_cache = False take the direct
    kernels with the same results."""
    import torch
    c = Case(proj, comps='TQU', tiled=tiled, n_det=8, n_samp=4000)
    pix_t, spin_t = _torch_pmat(c, gpu)
    pix, spin = pix_t.cpu().numpy(), spin_t.cpu().numpy()
    pe = (so3g_amd.ProjEng_Precomp_Tiled if tiled else so3g_amd.ProjEng_Precomp_NonTiled)()
    if proj == 'HP':
        m = np.random.default_rng(1).standard_normal((3, 12 * 128 * 128))
        zeros = lambda *lead: np.zeros(lead + (12 * 128 * 128,))
    else:
        m = c.map_like([3])
        zeros = lambda *lead: c.zeros(list(lead))
    ref_sig = oracle.precomp_from_map(c.og, m, pix, spin)
    got = pe.from_map(m, pix_t, spin_t)
    assert isinstance(got, torch.Tensor)
    assert np.array_equal(got.cpu().numpy().view(np.int32), ref_sig.view(np.int32))
    plan = pe._plan
    assert plan is not None
    base = c.signal.copy()
    acc = pe.from_map(m, pix_t, spin_t, base)
    assert pe._plan is plan
    assert np.array_equal(np.asarray(acc).view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin, c.signal).view(np.int32))
    g = pe.to_map(zeros(3), pix_t, spin_t, c.signal, c.det_weights)
    r = oracle.precomp_to_map(c.og, zeros(3), pix, spin, c.signal, c.det_weights)
    assert np.abs(flat_map(g) - flat_map(r)).max() < 1e-10
    g = pe.to_weight_map(zeros(3, 3), pix_t, spin_t, c.det_weights)
    r = oracle.precomp_to_weights(c.og, zeros(3, 3), pix, spin, c.det_weights)
    assert np.abs(flat_map(g) - flat_map(r)).max() < 1e-10
    assert pe._plan is plan
    # an in-place edit of the pointing matrix rebuilds the plan
    pix_t[0, :100, 0] = -1
    pix[0, :100, 0] = -1
    got = pe.from_map(m, pix_t, spin_t)
    assert pe._plan is not plan
    assert np.array_equal(got.cpu().numpy().view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin).view(np.int32))
    # thread_intervals: direct kernels, same sums
    from so3g_amd.proj import RangesMatrix
    full = RangesMatrix.zeros((c.obs.n_det, c.obs.n_samp))
    for r_ in full:
        r_.add_interval(0, c.obs.n_samp)
    g = pe.to_map(zeros(3), pix_t, spin_t, c.signal, None, [full])
    r = oracle.precomp_to_map(c.og, zeros(3), pix, spin, c.signal)
    assert np.abs(flat_map(g) - flat_map(r)).max() < 1e-10
    pe.plan_cache = False
    pe.clear_plan()
    got = pe.from_map(m, pix_t, spin_t)
    assert pe._plan is None
    assert np.array_equal(got.cpu().numpy().view(np.int32),
                          oracle.precomp_from_map(c.og, m, pix, spin).view(np.int32))


def test_precomp_cached_plan_inactive_tile(gpu):
    """A tile list with an inactive (None) tile takes the direct kernels,
    which raise for samples that hit it, as the reference does."""
    c = Case('CAR', comps='TQU', tiled=True, n_det=8, n_samp=4000)
    pix_t, spin_t = _torch_pmat(c, gpu)
    pe = so3g_amd.ProjEng_Precomp_Tiled()
    m = c.map_like([3])
    hit = int(pix_t[0, 0, 0])
    m[hit] = None
    with pytest.raises(RuntimeError, match='Tiling problem'):
        pe.from_map(m, pix_t, spin_t)
    assert pe._plan is None
