"""Precision study of the scatter arithmetic.

to_map and to_weights of one observation are run on the GPU three ways:

  ref   float32 products added to a float64 map -- so3g's arithmetic
        (Projection.cxx:1807, 1850)
  f32   float32 products accumulated in a float32 map
  f64   float64 products accumulated in a float64 map (the comparator)

and for every output component (T, Q, U; the six upper-triangle weight
entries) the max and RMS errors of ref and f32 relative to f64 are reported:
  max_rel = max|x - x64| / max|x64|,   rms_rel = rms(x - x64) / rms(x64)
over the hit pixels.

Geometries: ZEA and TAN at 0.25 arcmin centred on the scan, and HEALPix RING
nside 8192 (full-sky maps, admitted with allow_nside_max=True).  The
HEALPix 3x3 weight maps (58 GB per float64 copy) are made one arithmetic at
a time and reduced to the hit pixels before the next is made.

With keep=True the float64 comparator maps (hit pixels only for HEALPix)
are returned too, under '_f64', for checks against an independent CPU
accumulation.

Run: python -m so3g_amd.studies.precision [--n-det N] [--n-samp N] [--out FILE]
"""
import argparse
import json
import sys

import numpy as np

COMPS = ['T', 'Q', 'U']
W_ENTRIES = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _errors(x, x64, hit):
    d = (x.astype(np.float64) - x64)[hit]
    ref = x64[hit]
    scale = np.abs(ref).max()
    return {'max_rel': float(np.abs(d).max() / scale) if scale > 0 else 0.,
            'rms_rel': float(np.sqrt((d ** 2).mean()) / np.sqrt((ref ** 2).mean()))
            if ref.size else 0.}


def zenithal_projectionist(proj, res_arcmin, el_deg=50., throw_deg=10., disk_deg=1.0):
    """Square zenithal map centred on the scan centre (lon 0, lat el_deg)
    covering the az throw seen at that elevation plus the focal plane."""
    from ..proj.wcs import Projectionist
    from ..proj.wcslite import zenithal_geometry
    half = throw_deg * np.cos(np.radians(el_deg)) + disk_deg
    shape, wcs = zenithal_geometry((0., el_deg), half, res_arcmin / 60., proj=proj)
    return Projectionist.for_geom(shape, wcs)


def run(n_det=100, n_samp=40000, res_arcmin=0.25, nside=8192, seed=1, keep=False):
    import so3g_amd
    from .. import workloads
    from ..proj.wcs import ProjectionistHealpix
    obs = workloads.make('small', n_det=n_det, n_samp=n_samp)
    sig = workloads.signal(n_det, n_samp, seed=seed)
    # Non-unit responses, so that T products round too.
    rng = np.random.default_rng(seed + 7)
    resp = np.stack([rng.uniform(0.5, 1.5, n_det), rng.uniform(0.2, 1.0, n_det)],
                    1).astype(np.float32)
    out = {'workload': f'{n_det} det x {n_samp} samples (el 50 deg, az +-10 deg, 200 Hz), '
                       'responses T U[0.5,1.5], P U[0.2,1]',
           'cases': {}}
    for proj in ['ZEA', 'TAN']:
        p = zenithal_projectionist(proj, res_arcmin)
        eng = p.get_ProjEng('TQU')
        q = np.asarray(p._cache_q_fp_to_native(obs.asm.Q))
        dets = obs.dets
        maps = {pr: eng.to_map_prec(None, q, dets, resp, sig, None, pr) for pr in (0, 1, 2)}
        wts = {pr: eng.to_weight_map_prec(None, q, dets, resp, None, pr) for pr in (0, 1, 2)}
        hit = wts[2][0, 0] != 0
        case = {'shape': list(maps[2].shape), 'hit_pixels': int(hit.sum())}
        for name, pr in (('ref', 0), ('f32', 1)):
            case[name] = {
                'map': {c: _errors(maps[pr][i], maps[2][i], hit) for i, c in enumerate(COMPS)},
                'weights': {f'W{a}{b}': _errors(wts[pr][a, b], wts[2][a, b], hit)
                            for a, b in W_ENTRIES},
            }
        out['cases'][f'{proj} {res_arcmin} arcmin'] = case
        if keep:
            out.setdefault('_f64', {})[proj] = {'map': maps[2], 'weights': wts[2],
                                                'projectionist': p, 'resp': resp}
    # HEALPix RING nside 8192.
    hp = ProjectionistHealpix.for_healpix(nside, ordering='RING', allow_nside_max=True)
    eng = hp.get_ProjEng('TQU')
    q = np.asarray(hp._cache_q_fp_to_native(obs.asm.Q))
    import torch
    dev = torch.device('cuda')
    npix = 12 * nside * nside
    res = {}
    for pr in (0, 1, 2):
        dt = torch.float32 if pr == 1 else torch.float64
        m = torch.zeros((3, npix), dtype=dt, device=dev)
        eng.to_map_prec(m, torch.from_numpy(q).to(dev), torch.from_numpy(obs.dets).to(dev),
                        torch.from_numpy(resp).to(dev), torch.from_numpy(sig).to(dev),
                        None, pr)
        res[pr] = m
    hitt = (res[2][0] != 0) | (res[2][1] != 0)
    idx = torch.nonzero(hitt).squeeze(1)
    m64 = res[2][:, idx].cpu().numpy()
    case = {'npix': npix, 'hit_pixels': int(idx.numel())}
    for name, pr in (('ref', 0), ('f32', 1)):
        mm = res[pr][:, idx].cpu().numpy()
        allhit = np.ones(idx.numel(), bool)
        case[name] = {'map': {c: _errors(mm[i], m64[i], allhit) for i, c in enumerate(COMPS)}}
    del res
    wh = {}
    for pr in (2, 0, 1):
        dt = torch.float32 if pr == 1 else torch.float64
        w = torch.zeros((3, 3, npix), dtype=dt, device=dev)
        eng.to_weight_map_prec(w, torch.from_numpy(q).to(dev),
                               torch.from_numpy(obs.dets).to(dev),
                               torch.from_numpy(resp).to(dev), None, pr)
        wh[pr] = w[:, :, idx].cpu().numpy()
        del w
        torch.cuda.empty_cache()
    allhit = np.ones(idx.numel(), bool)
    for name, pr in (('ref', 0), ('f32', 1)):
        case[name]['weights'] = {f'W{a}{b}': _errors(wh[pr][a, b], wh[2][a, b], allhit)
                                 for a, b in W_ENTRIES}
    out['cases'][f'HEALPix RING nside {nside}'] = case
    if keep:
        out.setdefault('_f64', {})['HP'] = {'map': m64, 'weights': wh[2],
                                            'pixels': idx.cpu().numpy(), 'projectionist': hp,
                                            'resp': resp}
    return out


def to_markdown(r):
    lines = [f"Workload: {r['workload']}.", '',
             '| case | output | ref max rel | ref rms rel | f32 max rel | f32 rms rel |',
             '|---|---|---|---|---|---|']
    for cname, c in r['cases'].items():
        for kind in ('map', 'weights'):
            if kind not in c['ref']:
                continue
            for k in c['ref'][kind]:
                a, b = c['ref'][kind][k], c['f32'][kind][k]
                lines.append(f"| {cname} | {k} | {a['max_rel']:.2e} | {a['rms_rel']:.2e} | "
                             f"{b['max_rel']:.2e} | {b['rms_rel']:.2e} |")
    return '\n'.join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--n-det', type=int, default=100)
    ap.add_argument('--n-samp', type=int, default=40000)
    ap.add_argument('--nside', type=int, default=8192)
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)
    r = run(a.n_det, a.n_samp, nside=a.nside)
    print(json.dumps(r))
    md = to_markdown(r)
    print(md, file=sys.stderr)
    if a.out:
        open(a.out, 'w').write(md + '\n')


if __name__ == '__main__':
    main()
