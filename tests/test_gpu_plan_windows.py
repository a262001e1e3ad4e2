"""GPU parity of the windowed projection plan (plan layout 2): samples kept
in time order, work items of one detector's consecutive samples whose
visited pixels span at most 2048 compact values, from_map gathered from the
item's window in LDS, to_map / to_weights summed per window and added per
pixel in a second pass (or written over the whole map: to_map overwrite).

Bar, against the oracle's ProjEng_Precomp (Projection.cxx:2012-2253):
from_map bit-exact, fresh and accumulating; to_map and to_weights within
1e-9 absolute (float64 sums of the same float32 products in another order).
"""
import numpy as np
import pytest

import oracle
from helpers import Case
from test_gpu_parity import assert_bit_equal, maps_close

pytestmark = pytest.mark.gpu

WIN_MAX = 2048
SUB = 512


def window_items(comp, win_len):
    """Work items of the windowed layout from the compact pixel of every
    sample (n_det, n_t; -1 off the map): 512-sample sub-chunks of a detector
    merged while the merged window spans at most WIN_MAX compact values and
    the item holds at most ceil(n_sub / k) sub-chunks, k = max(ceil(n_t /
    win_len), ceil(64 / n_det)) capped at n_sub, the sub-chunks per
    detector.  Returns [(lo, win)], or None
    when a sub-chunk alone spans more than WIN_MAX."""
    n_det, n_t = comp.shape
    n_sub = -(-n_t // SUB)
    # about equal items of whole sub-chunks, at least 64 items in all
    k = min(n_sub, max(-(-n_t // max(SUB, win_len)), -(-64 // n_det)))
    win_len = -(-n_sub // k) * SUB
    items = []
    for d in range(n_det):
        lo, hi, length = None, None, 0
        for t0 in range(0, n_t, SUB):
            c = comp[d, t0:t0 + SUB]
            c = c[c >= 0]
            slo, shi = (int(c.min()), int(c.max())) if c.size else (None, None)
            if slo is not None and shi - slo + 1 > WIN_MAX:
                return None
            lj = min(n_t, t0 + SUB) - t0
            nlo = lo if slo is None else (slo if lo is None else min(lo, slo))
            nhi = hi if shi is None else (shi if hi is None else max(hi, shi))
            if length and ((nhi is not None and nhi - nlo + 1 > WIN_MAX) or
                           length + lj > win_len):
                items.append((lo, hi))
                length, nlo, nhi = 0, slo, shi
            lo, hi, length = nlo, nhi, length + lj
        items.append((lo, hi))
    return [(0, 0) if lo is None else (lo, hi - lo + 1) for lo, hi in items]


def compact(flat, on):
    c = np.full(flat.shape, -1, np.int64)
    c[on] = np.searchsorted(np.unique(flat[on]), flat[on])
    return c


class options:
    """Library options set for the duration of a with-block."""

    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        from so3g_amd import _native
        self.old = {k: _native.get_option(k) for k in self.kw}
        for k, v in self.kw.items():
            _native.set_option(k, v)

    def __exit__(self, *a):
        from so3g_amd import _native
        for k, v in self.old.items():
            _native.set_option(k, v)


def set_win_len(n):
    from so3g_amd import _native
    old = _native.get_option('plan_win_len')
    _native.set_option('plan_win_len', n)
    return old


@pytest.mark.parametrize('win_len', [512, 8192, 20000])
@pytest.mark.parametrize('n_det,n_samp,comps,tconst', [
    (30, 7000, 'TQU', True),       # n_t % 8 == 0: vector loads
    (21, 3001, 'TQU', True),       # odd sample count: per-sample loads
    (8, 1001, 'TQU', True),        # items shorter than a block's 2048 samples
    (40, 2000, 'TQU', False),      # T factor varies: all factors stored
    (40, 2000, 'T', True),         # T only: nothing stored per sample
    (40, 2000, 'QU', False)])
def test_windowed_plan(gpu, n_det, n_samp, comps, tconst, win_len):
    """Flat non-tiled map, part of the scan off the map: item table, to_map
    (adding and overwriting), to_weights, from_map (fresh and accumulating),
    the float64-signal to_map, with and without detector weights (one zero)."""
    import torch
    import so3g_amd
    from so3g_amd import _native as N
    c = Case('CAR', comps=comps, n_det=n_det, n_samp=n_samp, ny=100, nx=300, clip=True)
    pix, spin = [np.array(x) for x in c.engine.pointing_matrix(c.bore, c.dets, c.resps)]
    nc = spin.shape[2]
    if not tconst and comps != 'QU':
        spin[..., 0] *= np.linspace(0.5, 1.5, n_samp, dtype=np.float32)
    old = set_win_len(win_len)
    try:
        plan = so3g_amd.PrecompPlan(pix, spin, (100, 300), layout=2)
    finally:
        set_win_len(old)
    info = plan.info()
    on = pix[..., 0] >= 0
    flat = pix[..., 0].astype(np.int64) * 300 + pix[..., 1]
    items = window_items(compact(flat, on), win_len)
    assert items is not None
    assert info['layout'] == 2
    assert info['n_on'] == int(on.sum())
    assert info['n_vis'] == len(np.unique(flat[on]))
    assert info['items'] == len(items)
    assert info['window_max'] == max(w for _, w in items) <= WIN_MAX
    assert info['window_slots'] == sum(w for _, w in items)
    assert info['t_const'] == (tconst and comps != 'QU')
    dw = c.det_weights.copy()
    dw[3] = 0.
    m = c.map_like([nc])
    s0 = np.array(c.signal)
    rf = oracle.precomp_from_map(c.og, m, pix, spin)
    rf_acc = oracle.precomp_from_map(c.og, m, pix, spin, s0.copy())
    # 1024 (default) or 256 threads per item; window sums by a second pass
    # (default) or by atomics
    for block, flush in ((1024, 0), (256, 0), (1024, 1)):
        with options(plan_win_block=block, plan_win_flush=flush):
            tag = f'{comps} block={block} flush={flush}'
            for w in (None, dw):
                r = oracle.precomp_to_map(c.og, np.zeros((nc, 100, 300)), pix, spin, c.signal,
                                          w)
                m0 = np.random.default_rng(5).standard_normal((nc, 100, 300))
                g = plan.to_map(m0.copy(), c.signal, w)
                maps_close(g - m0, r, f'windowed to_map {tag}', atol=1e-9)
                g = plan.to_map(np.full((nc, 100, 300), 7.), c.signal, w, overwrite=True)
                maps_close(g, r, f'windowed to_map overwrite {tag}', atol=1e-9)
                gw = plan.to_weights(np.zeros((nc, nc, 100, 300)), w)
                rw = oracle.precomp_to_weights(c.og, np.zeros((nc, nc, 100, 300)), pix, spin, w)
                maps_close(gw, rw, f'windowed to_weights {tag}', atol=1e-9)
            assert_bit_equal(np.array(plan.from_map(m)), rf, f'windowed from_map {tag}')
            assert_bit_equal(np.array(plan.from_map(m, s0.copy())), rf_acc,
                             f'windowed from_map accumulate {tag}')
    # float64 signal
    dev = torch.device('cuda')
    mf = torch.zeros((nc, 100, 300), dtype=torch.float64, device=dev)
    sig64 = torch.from_numpy(c.signal.astype(np.float64)).to(dev)
    N.check(N.load().so3g_pc_plan_to_map_f64(plan._h, mf.data_ptr(), sig64.data_ptr(), None,
                                             None))
    r = oracle.precomp_to_map(c.og, np.zeros((nc, 100, 300)), pix, spin, c.signal)
    maps_close(mf.cpu().numpy(), r, 'windowed to_map f64', atol=1e-5)


def test_windowed_plan_unaligned_signal(gpu):
    """A signal whose first element is not 16-byte aligned (a torch view one
    float into its storage) takes the per-sample loads: same results."""
    import torch
    import so3g_amd
    c = Case('CAR', comps='TQU', n_det=12, n_samp=4000, ny=100, nx=300)
    pix, spin = c.engine.pointing_matrix(c.bore, c.dets, c.resps)
    pix, spin = np.array(pix), np.array(spin)
    plan = so3g_amd.PrecompPlan(pix, spin, (100, 300), layout=2)
    dev = torch.device('cuda')
    store = torch.zeros(12 * 4000 + 1, dtype=torch.float32, device=dev)
    sig = store[1:].view(12, 4000)
    assert sig.data_ptr() % 16 != 0
    sig.copy_(torch.from_numpy(c.signal))
    r = oracle.precomp_to_map(c.og, np.zeros((3, 100, 300)), pix, spin, c.signal)
    g = torch.zeros((3, 100, 300), dtype=torch.float64, device=dev)
    plan.to_map(g, sig)
    maps_close(g.cpu().numpy(), r, 'to_map unaligned', atol=1e-9)
    m = c.map_like([3])
    plan.from_map(torch.from_numpy(m).to(dev), sig, fresh=True)
    assert_bit_equal(sig.cpu().numpy(), oracle.precomp_from_map(c.og, m, pix, spin),
                     'from_map unaligned')


@pytest.mark.parametrize('case', ['all_off', 'tiny', 'one_pixel', 'half_off'])
def test_windowed_plan_edges(gpu, case):
    """Degenerate pointing on the windowed layout: every sample off the map
    (empty windows; to_map adds nothing, overwrite writes zeros, from_map
    fresh is zero), 5 samples, a detector parked on one pixel, detectors off
    the map for whole sub-chunks."""
    import so3g_amd
    rng = np.random.default_rng(11)
    ny, nx = 40, 90
    n_det, n_t = {'tiny': (1, 5)}.get(case, (5, 1600))
    # each detector sweeps one row back and forth
    t = np.arange(n_t)
    col = (np.abs((t % 400) - 200) * (nx - 1) // 200).astype(np.int32)
    pix = np.zeros((n_det, n_t, 2), np.int32)
    for d in range(n_det):
        pix[d, :, 0] = (7 * d) % ny
        pix[d, :, 1] = col
    if case == 'all_off':
        pix[..., 0] = -1
    elif case == 'one_pixel':
        pix[2, :, 0], pix[2, :, 1] = 33, 77
    elif case == 'half_off':
        pix[1, 512:1100] = -1
        pix[3, :600] = -1
    spin = np.empty((n_det, n_t, 3), np.float32)
    spin[..., 0] = rng.uniform(0.5, 1.5, (n_det, 1))
    spin[..., 1:] = rng.uniform(-1, 1, (n_det, n_t, 2))
    sig = rng.standard_normal((n_det, n_t)).astype(np.float32)
    og = oracle.geom('CAR', 3, ny=ny, nx=nx, cdelt=(1., 1.), crpix=(1., 1.), tiled=False,
                     tile_shape=(1, 1))
    plan = so3g_amd.PrecompPlan(pix, spin, (ny, nx))
    info = plan.info()
    assert info['layout'] == 2
    assert info['n_on'] == int((pix[..., 0] >= 0).sum())
    m0 = rng.standard_normal((3, ny, nx))
    r = oracle.precomp_to_map(og, np.zeros((3, ny, nx)), pix, spin, sig)
    maps_close(plan.to_map(m0.copy(), sig) - m0, r, f'to_map {case}', atol=1e-9)
    maps_close(plan.to_map(m0.copy(), sig, overwrite=True), r, f'to_map overwrite {case}',
               atol=1e-9)
    m = rng.standard_normal((3, ny, nx))
    assert_bit_equal(np.array(plan.from_map(m)), oracle.precomp_from_map(og, m, pix, spin),
                     f'from_map {case}')
    s0 = rng.standard_normal((n_det, n_t)).astype(np.float32)
    assert_bit_equal(np.array(plan.from_map(m, s0.copy())),
                     oracle.precomp_from_map(og, m, pix, spin, s0.copy()),
                     f'from_map accumulate {case}')
    if case == 'all_off':
        assert info['window_slots'] == 0
        assert not np.any(plan.to_map(m0.copy(), sig, overwrite=True))


def test_windowed_plan_refused_for_scattered_pixels(gpu):
    """Pixels hopping over the whole map: a 512-sample sub-chunk spans more
    than 2048 compact values, so layout 2 is refused (RuntimeError) and the
    automatic choice falls back to pixel runs (layout 0); both plans' results
    match the oracle."""
    import so3g_amd
    rng = np.random.default_rng(3)
    ny, nx, n_det, n_t = 80, 100, 4, 3000
    pix = np.stack([rng.integers(0, ny, (n_det, n_t)), rng.integers(0, nx, (n_det, n_t))],
                   axis=-1).astype(np.int32)
    spin = rng.uniform(-1, 1, (n_det, n_t, 3)).astype(np.float32)
    sig = rng.standard_normal((n_det, n_t)).astype(np.float32)
    with pytest.raises(RuntimeError):
        so3g_amd.PrecompPlan(pix, spin, (ny, nx), layout=2)
    plan = so3g_amd.PrecompPlan(pix, spin, (ny, nx))
    assert plan.info()['layout'] == 0
    og = oracle.geom('CAR', 3, ny=ny, nx=nx, cdelt=(1., 1.), crpix=(1., 1.), tiled=False,
                     tile_shape=(1, 1))
    r = oracle.precomp_to_map(og, np.zeros((3, ny, nx)), pix, spin, sig)
    maps_close(plan.to_map(np.full((3, ny, nx), 2.), sig, overwrite=True), r,
               'runs to_map overwrite', atol=1e-9)


@pytest.mark.parametrize('nest,nside', [(True, 512), (False, 4096)])
def test_windowed_plan_healpix(gpu, nest, nside):
    """HEALPix (k = 1 pixel numbers): where the windows fit, the layout-2 plan
    matches the oracle's Precomp (to_map, overwrite, to_weights, from_map bit
    for bit); where they do not, it is refused."""
    import so3g_amd
    c = Case('HP', comps='TQU', nest=nest, nside=nside, n_det=16, n_samp=3000)
    pix, spin = [np.array(x) for x in c.engine.pointing_matrix(c.bore, c.dets, c.resps)]
    pix[2, 100:400, 0] = -1
    npix = 12 * nside * nside
    on = pix[..., 0] >= 0
    from so3g_amd import _native
    items = window_items(compact(pix[..., 0].astype(np.int64), on),
                         _native.get_option('plan_win_len'))
    if items is None:
        with pytest.raises(RuntimeError):
            so3g_amd.PrecompPlan(pix, spin, (npix,), layout=2)
        return
    plan = so3g_amd.PrecompPlan(pix, spin, (npix,), layout=2)
    assert plan.info()['items'] == len(items)
    dw = c.det_weights.copy()
    dw[1] = 0.
    r = oracle.precomp_to_map(c.og, np.zeros((3, npix)), pix, spin, c.signal, dw)
    maps_close(plan.to_map(np.zeros((3, npix)), c.signal, dw), r, 'HP to_map', atol=1e-9)
    maps_close(plan.to_map(np.ones((3, npix)), c.signal, dw, overwrite=True), r,
               'HP to_map overwrite', atol=1e-9)
    rw = oracle.precomp_to_weights(c.og, np.zeros((3, 3, npix)), pix, spin, dw)
    maps_close(plan.to_weights(np.zeros((3, 3, npix)), dw), rw, 'HP to_weights', atol=1e-9)
    m = c.map_like([3])
    assert_bit_equal(np.array(plan.from_map(m)), oracle.precomp_from_map(c.og, m, pix, spin),
                     'HP from_map')


def test_overwrite_other_layouts(gpu):
    """to_map overwrite on the pixel-run (0) and tiled (1) layouts: the map is
    zeroed, then the plan's sums added."""
    import so3g_amd
    c = Case('CAR', comps='TQU', n_det=10, n_samp=2000, ny=100, nx=300, clip=True)
    pix, spin = [np.array(x) for x in c.engine.pointing_matrix(c.bore, c.dets, c.resps)]
    r = oracle.precomp_to_map(c.og, np.zeros((3, 100, 300)), pix, spin, c.signal)
    for layout in (0, 1):
        plan = so3g_amd.PrecompPlan(pix, spin, (100, 300), layout=layout)
        g = plan.to_map(np.full((3, 100, 300), -3.), c.signal, overwrite=True)
        maps_close(g, r, f'overwrite layout {layout}', atol=1e-9)
