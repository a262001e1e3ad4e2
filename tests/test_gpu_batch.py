"""Batched projection (many observations, one launch) against per-observation
calls and the oracle."""
import numpy as np
import pytest

import oracle
import so3g_amd
from so3g_amd import workloads
from so3g_amd.proj.wcs import Projectionist
from so3g_amd.proj.wcslite import car_geometry

torch = pytest.importorskip('torch')

# (n_det, n_samp, az0 deg, el deg): uneven sizes, one empty observation.
OBS = [(5, 3000, 0., 50.), (7, 1500, 4., 49.), (0, 100, 0., 50.), (3, 700, -3., 51.),
       (9, 257, 2., 50.5)]


def make_batch(tiled):
    obs = []
    for i, (n_det, n_t, az0, el) in enumerate(OBS):
        o = workloads.Observation(max(n_det, 1), n_t, 1.0, 2.0, el_deg=el, az0_deg=az0,
                                  throw_deg=3., seed=10 + i)
        obs.append((o, n_det))
    shape, wcs = car_geometry((-17., 17.), (48., 53.), 2. / 60., margin=2)
    if tiled:
        p = Projectionist.for_tiled(shape, wcs, (40, 60))
    else:
        p = Projectionist.for_geom(shape, wcs)
    rng = np.random.default_rng(3)
    items = []
    for (o, n_det) in obs:
        q = np.ascontiguousarray(np.asarray(p._cache_q_fp_to_native(o.asm.Q)), np.float64)
        dets = o.dets[:n_det]
        resp = np.stack([rng.uniform(0.5, 1.5, n_det), rng.uniform(0.2, 1., n_det)],
                        1).astype(np.float32)
        sig = workloads.signal(n_det, o.n_samp, seed=int(rng.integers(1000)))
        dw = rng.uniform(0.5, 2., n_det).astype(np.float32)
        if n_det > 1:
            dw[1] = 0.
        items.append(dict(bore=q, dets=dets, resp=resp, signal=sig, det_weights=dw))
    return p, items


def pixargs_and_engine(p, items):
    if p.tiling is None:
        return p._get_pixelizor_args(), p.get_ProjEng('TQU')
    # active tiles = union of the tiles every observation touches
    eng0 = p.get_ProjEng('TQU')
    hits = 0
    for it in items:
        if len(it['dets']):
            hits = hits + np.asarray(eng0.tile_hits(it['bore'], it['dets']))
    active = [int(i) for i in np.nonzero(hits)[0]]
    p.active_tiles = active
    return p._get_pixelizor_args(), p.get_ProjEng('TQU')


def dev_map(eng, seed=5):
    rng = np.random.default_rng(seed)
    z = eng.zeros((3,))
    dev = torch.device('cuda')
    if isinstance(z, list):
        return [None if t is None else torch.from_numpy(rng.standard_normal(t.shape)).to(dev)
                for t in z]
    return torch.from_numpy(rng.standard_normal(z.shape)).to(dev)


def zero_like(m):
    if isinstance(m, list):
        return [None if t is None else torch.zeros_like(t) for t in m]
    return torch.zeros_like(m)


def host(m):
    if isinstance(m, list):
        return [None if t is None else t.cpu().numpy() for t in m]
    return m.cpu().numpy()


def flat(m):
    if isinstance(m, list):
        return np.concatenate([t.ravel() for t in m if t is not None])
    return np.asarray(m).ravel()


@pytest.mark.gpu
@pytest.mark.parametrize('tiled', [False, True], ids=['flat', 'tiled'])
def test_batch_matches_per_observation(gpu, tiled):
    p, items = make_batch(tiled)
    pixargs, eng = pixargs_and_engine(p, items)
    og = oracle.geom_from_args('CAR', 'TQU', pixargs, 0)
    sky = dev_map(eng)
    batch = so3g_amd.ObservationBatch(eng, items)
    assert batch.info()['n_obs'] == len(OBS)
    dev = torch.device('cuda')

    # from_map: bit-exact against the per-observation engine call and the oracle.
    sky_h = host(sky)
    for mode in ['eager', 'graph', 'per_obs', 'graph']:
        for s in batch.signals:
            s.fill_(7.)
        sigs = batch.from_map(sky, mode=mode)
        for it, s in zip(items, sigs):
            if len(it['dets']) == 0:
                continue
            ref = oracle.from_map(og, sky_h, it['bore'], it['dets'], it['resp'])
            np.testing.assert_array_equal(s.cpu().numpy(), ref, err_msg=mode)

    # to_map: the same sums as per-observation engine calls.
    for s, it in zip(batch.signals, items):
        s.copy_(torch.from_numpy(it['signal']))
    ref = zero_like(sky)
    for it in items:
        if len(it['dets']) == 0:
            continue
        ref = eng.to_map(ref, torch.from_numpy(it['bore']).to(dev),
                         torch.from_numpy(it['dets']).to(dev),
                         torch.from_numpy(it['resp']).to(dev),
                         torch.from_numpy(it['signal']).to(dev),
                         torch.from_numpy(it['det_weights']).to(dev))
    ref_h = flat(host(ref))
    # and the oracle, observation by observation
    om = eng.zeros((3,))
    for it in items:
        if len(it['dets']):
            om = oracle.to_map(og, om, it['bore'], it['dets'], it['resp'], it['signal'],
                               it['det_weights'])
    np.testing.assert_allclose(ref_h, flat(om), rtol=0, atol=1e-9)
    # global atomics, then each work item summed in an LDS table of 512 or
    # 1024 slots first, then row segments of 64 columns (most samples off the
    # segment), 700 and 6000 (wider than the map: the whole row)
    for slots in (0, 512, 1024, -64, -700, -6000):
        batch.set_table(slots)
        for mode in ['eager', 'per_obs', 'graph', 'graph', 'graph']:
            m = zero_like(sky)
            batch.to_map(m, mode=mode)
            torch.cuda.synchronize()
            _ = torch.ones(1000, device=dev).sum().item()   # eager work between replays
            np.testing.assert_allclose(flat(host(m)), ref_h, rtol=0, atol=1e-9,
                                       err_msg=f'{mode} {slots}')
    batch.set_table(0)
    # Graph replays into the same map accumulate.
    m = zero_like(sky)
    batch.to_map(m, mode='graph')
    batch.to_map(m, mode='graph')
    np.testing.assert_allclose(flat(host(m)), 2 * ref_h, rtol=0, atol=2e-9)
    batch.close()


@pytest.mark.gpu
def test_batch_tune(gpu):
    p, items = make_batch(False)
    pixargs, eng = pixargs_and_engine(p, items)
    sky = dev_map(eng)
    batch = so3g_amd.ObservationBatch(eng, items)
    res = batch.tune('from_map', sky)
    assert len(res) == 25 and all(r[3] > 0 and r[2] == 0 for r in res)
    best = min(res, key=lambda r: r[3])
    assert batch.info()['from_map_shape'] == best[:2]
    scratch = zero_like(sky)
    res = batch.tune('to_map', scratch)
    # 25 launch shapes x (global atomics, 512- and 1024-slot LDS tables),
    # then row segments of 1024 columns and of the whole row
    nx = pixargs[1]
    widths = sorted({min(1024, nx), nx})
    assert len(res) == 75 + len(widths)
    assert sorted({r[2] for r in res}) == sorted([-w for w in widths] + [0, 512, 1024])
    assert all(r[:2] == (2, 0) for r in res if r[2] < 0)
    best = min(res, key=lambda r: r[3])
    assert batch.info()['to_map_shape'] == best[:2]
    assert batch.info()['to_map_slots'] == best[2]
    # every tuned shape gives the same result
    ref = None
    for dpb, gm, slots in [(1, 2, 0), (16, 0, 512), (4, 8, 1024), (2, 0, 1024), (1, 2, -100)]:
        batch.set_shape('to_map', dpb, gm)
        batch.set_table(slots)
        m = zero_like(sky)
        batch.to_map(m, mode='eager')
        h = flat(host(m))
        if ref is None:
            ref = h
        np.testing.assert_allclose(h, ref, rtol=0, atol=1e-9)
    with pytest.raises(ValueError):
        batch.set_shape('to_map', 3, 2)
    with pytest.raises(ValueError):
        batch.set_table(256)
    with pytest.raises(ValueError):
        batch.set_table(-63)
    with pytest.raises(ValueError):
        batch.set_table(-7000)   # 3 x 7000 float64 sums exceed 160 KB


@pytest.mark.gpu
def test_batch_inactive_tile_raises(gpu):
    p, items = make_batch(True)
    pixargs, eng = pixargs_and_engine(p, items)
    # drop one active tile
    act = list(p.active_tiles)
    p.active_tiles = act[1:]
    eng2 = p.get_ProjEng('TQU')
    sky = dev_map(eng2)
    batch = so3g_amd.ObservationBatch(eng2, items)
    with pytest.raises(RuntimeError):
        batch.from_map(sky, mode='eager')
    with pytest.raises(RuntimeError):
        batch.to_map(zero_like(sky), mode='graph')
    batch.set_table(-300)
    for mode in ('eager', 'per_obs', 'graph'):
        with pytest.raises(RuntimeError):
            batch.to_map(zero_like(sky), mode=mode)
