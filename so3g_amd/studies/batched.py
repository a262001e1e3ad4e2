"""Batched observations: 64 uneven observations projected into one tiled
CAR 0.5 arcmin TQU map, one persistent launch (and its HIP-graph replay)
against one launch per observation.

Draw (numpy default_rng(seed)): n_det = round(exp U[ln 100, ln 2000]),
n_samp = round(exp U[ln 5e4, ln 5e5]), az0 ~ U[-30, 30] deg,
el ~ U[40, 60] deg; constant-elevation scans (throw 10 deg, 1 deg/s,
200 Hz), detectors in a 1.5 deg disk, float32 white-noise signal.  Map:
CAR 0.5 arcmin over the union footprint, 100 x 100 tiles, active tiles =
union of the observations' tiles.

Timed (to_map and from_map each, the whole batch per pass):
  api       the engine's to_map / from_map called once per observation
            (from_map accumulating into the signal, as the engine does; the
            batch modes are timed with accumulate=True too)
  per_obs   the batch kernel launched once per observation, no host sync
  eager     one persistent launch over all observations
  graph     the same launch captured in a HIP graph and replayed
after the autotuner has picked the launch shape of each op (detectors per
work item, blocks per compute unit and, for to_map, global atomics or an
LDS table of 512 or 1024 slots per work item).

Run: python -m so3g_amd.studies.batched [--n-obs 64] [--seed 4] [--reps 3]
"""
import argparse
import json
import sys
import time

import numpy as np


def draw(n_obs, seed):
    rng = np.random.default_rng(seed)
    n_det = np.round(np.exp(rng.uniform(np.log(100), np.log(2000), n_obs))).astype(int)
    n_samp = np.round(np.exp(rng.uniform(np.log(5e4), np.log(5e5), n_obs))).astype(int)
    az0 = rng.uniform(-30, 30, n_obs)
    el = rng.uniform(40, 60, n_obs)
    return n_det, n_samp, az0, el


def build(n_obs=64, seed=4, res_arcmin=0.5, tile=100, disk_deg=1.5, throw=10.):
    import torch
    from .. import workloads
    from ..proj.wcs import Projectionist
    from ..proj.wcslite import car_geometry
    n_det, n_samp, az0, el = draw(n_obs, seed)
    r = disk_deg / 2
    lon_pad = r / np.cos(np.radians(el.max() + r)) * 1.02
    lon = (-(az0.max() + throw) - lon_pad, -(az0.min() - throw) + lon_pad)
    lat = (el.min() - r * 1.01, el.max() + r * 1.01)
    shape, wcs = car_geometry(lon, lat, res_arcmin / 60., margin=4)
    p = Projectionist.for_tiled(shape, wcs, (tile, tile))
    eng_all = p.get_ProjEng('TQU')
    dev = torch.device('cuda')
    gen = torch.Generator(device=dev).manual_seed(seed)
    items = []
    hits = np.zeros(p.tiling.tile_count, np.int64)
    from ..boresight import horizon
    qc2n = getattr(p, 'q_celestial_to_native', None)
    for i in range(n_obs):
        t, az, e = workloads.scan(int(n_samp[i]), el_deg=el[i], az0_deg=az0[i],
                                  throw_deg=throw)
        fp = workloads.focal_plane(int(n_det[i]), disk_deg, seed=seed * 1000 + i)
        # native-frame boresight made on the GPU (for_horizon, premultiplied)
        q = horizon(az, e, q_pre=None if qc2n is None else np.asarray(qc2n))
        dets = torch.from_numpy(np.ascontiguousarray(np.asarray(fp.quats), np.float64)).to(dev)
        resp = torch.from_numpy(np.ascontiguousarray(fp.resps, np.float32)).to(dev)
        hits += np.asarray(eng_all.tile_hits(q, dets), np.int64)
        sig = torch.randn((int(n_det[i]), int(n_samp[i])), dtype=torch.float32, device=dev,
                          generator=gen)
        items.append(dict(bore=q, dets=dets, resp=resp, signal=sig))
    p.active_tiles = [int(j) for j in np.nonzero(hits)[0]]
    eng = p.get_ProjEng('TQU')
    return p, eng, items, (n_det, n_samp, az0, el)


def _tiles(eng, dev, fill=None, seed=5):
    import torch
    z = eng.zeros((3,))
    out = []
    g = torch.Generator(device=dev).manual_seed(seed)
    for t in z:
        if t is None:
            out.append(None)
        elif fill == 'randn':
            out.append(torch.randn(t.shape, dtype=torch.float64, device=dev, generator=g))
        else:
            out.append(torch.zeros(t.shape, dtype=torch.float64, device=dev))
    return out


def run(n_obs=64, seed=4, reps=3):
    import torch
    from ..batch import ObservationBatch
    dev = torch.device('cuda')
    t0 = time.time()
    p, eng, items, (n_det, n_samp, az0, el) = build(n_obs, seed)
    torch.cuda.synchronize()
    t_build = time.time() - t0
    N_s = int(np.sum(n_det.astype(np.int64) * n_samp))
    sky = _tiles(eng, dev, 'randn')
    acc = _tiles(eng, dev)
    batch = ObservationBatch(eng, items)
    n_active = len(p.active_tiles)
    out = {'workload': f'{n_obs} observations (seed {seed}), {N_s:.4g} det*samples, '
                       f'tiled CAR 0.5 arcmin TQU {tuple(p.naxis[::-1])}, 100x100 tiles, '
                       f'{n_active} of {p.tiling.tile_count} active',
           'det_samples': N_s, 'n_det_range': [int(n_det.min()), int(n_det.max())],
           'n_samp_range': [int(n_samp.min()), int(n_samp.max())],
           'setup_s': t_build}
    print(f'built in {t_build:.1f} s: {out["workload"]}', file=sys.stderr, flush=True)

    t0 = time.time()
    tune_from = batch.tune('from_map', sky)
    tune_to = batch.tune('to_map', acc)
    out['tune_s'] = time.time() - t0
    out['tune'] = {'from_map': tune_from, 'to_map': tune_to}
    # best to_map launch per LDS table size (0: global atomics)
    out['to_map_best_by_table'] = {
        str(k): min(r[3] for r in tune_to if r[2] == k) for k in sorted({r[2] for r in tune_to})}
    out['shape'] = batch.info()
    print(f'tuned in {out["tune_s"]:.1f} s: {out["shape"]}', file=sys.stderr, flush=True)

    def api_from():
        for it, s in zip(items, batch.signals):
            eng.from_map(sky, it['bore'], it['dets'], it['resp'], s)

    def api_to():
        for it, s in zip(items, batch.signals):
            eng.to_map(acc, it['bore'], it['dets'], it['resp'], s)

    def timed(f):
        f()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            f()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    res = {}
    for op in ['from_map', 'to_map']:
        res[op] = {}
        for mode in ['api', 'per_obs', 'eager', 'graph']:
            if mode == 'api':
                f = api_from if op == 'from_map' else api_to
            elif op == 'from_map':
                # accumulate, as the engine's from_map does into a given signal
                f = (lambda m=mode: batch.from_map(sky, mode=m, accumulate=True))
            else:
                f = (lambda m=mode: batch.to_map(acc, mode=m))
            s = timed(f)
            res[op][mode] = {'s': s, 'det_samples_per_s': N_s / s}
            print(f'{op} {mode}: {s * 1e3:.1f} ms  {N_s / s:.3g} det*samples/s',
                  file=sys.stderr, flush=True)
    out['timing'] = res

    # Agreement: batch graph vs per-observation API.
    for s in batch.signals:
        s.zero_()
    batch.from_map(sky, mode='graph')
    ref = [s.clone() for s in batch.signals[:4]]
    for s in batch.signals:
        s.zero_()
    api_from()
    out['from_map_bit_exact'] = all(bool(torch.equal(a, b))
                                    for a, b in zip(ref, batch.signals[:4]))
    m1, m2 = _tiles(eng, dev), _tiles(eng, dev)
    batch.to_map(m1, mode='graph')
    for it, s in zip(items, batch.signals):
        eng.to_map(m2, it['bore'], it['dets'], it['resp'], s)
    num = max(float((a - b).abs().max()) for a, b in zip(m1, m2) if a is not None)
    den = max(float(b.abs().max()) for b in m2 if b is not None)
    out['to_map_max_rel_diff'] = num / den
    batch.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--n-obs', type=int, default=64)
    ap.add_argument('--seed', type=int, default=4)
    ap.add_argument('--reps', type=int, default=3)
    a = ap.parse_args(argv)
    print(json.dumps(run(a.n_obs, a.seed, a.reps)), flush=True)


if __name__ == '__main__':
    main()
