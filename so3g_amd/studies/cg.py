"""Map-making study: CG to ||r||/||b|| = 1e-6 on the small observation
(100 detectors x 40000 samples, CAR 1 arcmin TQU) with 1/f noise:
binned-preconditioned CG, the two-level (binned + along-scan coarse space)
preconditioner, the two-level preconditioner with a learned deflation space
(Ritz vectors of its smallest eigenvalues, learned from the Lanczos
processes of solves on three independent realisations), and flexible CG
with the U-Net
correction (trained on residual/error pairs of the deflated solves) on top
of that, falling back to it when it stalls.

Noise: PSD_d(f) = sigma_d^2 (1 + (f / 1 Hz)^-3), sigma_d ~ lognormal(0, 0.3).
Sky: flat-sky Gaussian random field on the map grid, C_l ~ l^-2.5 for T and
0.1 x that for E and B, rotated to Q/U.

Each solve reports its true residual ||b - A x|| / ||b|| and its distance
from a reference solution (the deflated PCG run to 1e-10) in the A-norm.

Run: python -m so3g_amd.studies.cg [--no-learned] [--no-unet] [--steps N] [--n-rhs 4 --per-rhs 500]
"""
import argparse
import json
import sys
import time

import numpy as np


def grf_tqu(shape, res_rad, seed, slope=-2.5, pol_frac=0.1):
    """(3, ny, nx) float64 T, Q, U with C_l ~ l^slope (T) and pol_frac x C_l
    for E and B (flat sky, l = 2 pi |k| / res)."""
    ny, nx = shape
    rng = np.random.default_rng(seed)
    ly = np.fft.fftfreq(ny, res_rad) * 2 * np.pi
    lx = np.fft.fftfreq(nx, res_rad) * 2 * np.pi
    LY, LX = np.meshgrid(ly, lx, indexing='ij')
    ell = np.hypot(LY, LX)
    with np.errstate(divide='ignore'):
        amp = np.where(ell > 0, ell ** (slope / 2), 0.)
    def field(scale):
        w = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        return w * amp * np.sqrt(scale)
    T = field(1.)
    E = field(pol_frac)
    B = field(pol_frac)
    phi = np.arctan2(LY, LX)
    Q = E * np.cos(2 * phi) - B * np.sin(2 * phi)
    U = E * np.sin(2 * phi) + B * np.cos(2 * phi)
    out = np.stack([np.fft.ifft2(x).real for x in (T, Q, U)])
    return out / out[0].std()


def setup(n_det=100, n_samp=40000, seed=5):
    import torch
    import so3g_amd
    from .. import workloads
    from ..mapmaker import MapMaker, NoiseModel, PointingOperator
    dev = torch.device('cuda')
    obs = workloads.make('small', n_det=n_det, n_samp=n_samp)
    p = obs.projectionist()
    eng = p.get_ProjEng('TQU')
    q = torch.from_numpy(np.ascontiguousarray(np.asarray(p._cache_q_fp_to_native(obs.asm.Q)),
                                              np.float64)).to(dev)
    dets = torch.from_numpy(obs.dets).to(dev)
    resp = torch.from_numpy(obs.resps).to(dev)
    pix, spin = eng.pointing_matrix(q, dets, resp)
    ny, nx = obs.shape
    pointing = PointingOperator(pix, spin, (ny, nx))
    rng = np.random.default_rng(seed)
    sigma = np.exp(rng.normal(0., 0.3, n_det))
    noise = NoiseModel(sigma, n_samp, device=dev)
    mm = MapMaker(pointing, noise)
    ctx = dict(mm=mm, shape=(ny, nx), obs=obs)
    ctx['sky'], ctx['d'] = make_data(ctx, seed)
    return ctx


def make_data(ctx, seed):
    """(sky, d = P sky + noise) for one realisation."""
    import torch
    mm = ctx['mm']
    P, noise = mm.P, mm.noise
    dev = P.pix.device
    res = ctx['obs'].res_arcmin / 60 * np.pi / 180
    sky = torch.from_numpy(grf_tqu(ctx['shape'], res, seed + 1)).to(dev)
    gen = torch.Generator(device=dev).manual_seed(seed + 2)
    d = P.P(sky, torch.empty((P.n_det, P.n_t), dtype=torch.float32, device=dev))
    d += noise.simulate(gen)
    return sky, d


def solve_binned(ctx, tol=1e-6, maxiter=3000):
    from ..mapmaker import pcg
    mm = ctx['mm']
    t0 = time.perf_counter()
    M = mm.binned()
    b = mm.rhs(ctx['d'])
    t_setup = time.perf_counter() - t0
    x, info = pcg(mm.apply_A, b, M, tol=tol, maxiter=maxiter)
    info['setup_s'] = t_setup
    return x, info, M


def map_error(x, sky, mask):
    d = (x - sky)[:, mask]
    return [float(v) for v in d.pow(2).mean(dim=1).sqrt()]


def true_residual(mm, x, b):
    """||b - A x|| / ||b||, recomputed from x (CG's own residual is updated
    recursively and drifts from it)."""
    r = b - mm.apply_A(x)
    return float((r.norm() / b.norm()).item())


def a_norm_error(mm, x, x_ref):
    """||x - x_ref||_A / ||x_ref||_A: the error CG minimises, weighted by A,
    so the near-null modes that the data barely constrain count little."""
    e = x - x_ref
    num = float((e * mm.apply_A(e)).sum().item())
    den = float((x_ref * mm.apply_A(x_ref)).sum().item())
    return float(np.sqrt(max(num, 0.) / den))


def summary(x, info, ctx, M):
    mm = ctx['mm']
    return {'iters': info['iters'], 'seconds': info['seconds'],
            'final_residual': info['residuals'][-1],
            'true_residual': true_residual(mm, x, mm.rhs(ctx['d'])),
            'residual_every_100': [info['residuals'][i]
                                   for i in range(0, len(info['residuals']), 100)],
            'fallback_at': info.get('fallback_at'),
            'map_rms_error_TQU': map_error(x, ctx['sky'], M.mask),
            'ms_per_iter': info['seconds'] / max(info['iters'], 1) * 1e3}


def run(tol=1e-6, maxiter=40000, learned=True, n_rhs=4, per_rhs=500, steps=3000,
        batch=8, lr=3e-4, factor=(1, 4), h=16, harvest_iters=1200, k_defl=32,
        n_train_solves=3, ref_tol=1e-10, unet=True):
    import torch
    from ..mapmaker import (DeflatedPreconditioner, LearnedPreconditioner,
                            collect_residual_error_pairs, harvest_ritz, pcg, train_correction)
    from ..unet import UNet
    ctx = setup()
    mm = ctx['mm']
    mm.apply_A(mm.zeros())     # warm-up of every kernel on the path
    torch.cuda.synchronize()
    x, info, M = solve_binned(ctx, tol, maxiter)
    out = {'workload': '100 det x 40000 samples, CAR 1 arcmin TQU %dx%d, 1/f noise '
                       '(fknee 1 Hz, alpha -3), float64 time streams in A' % ctx['shape'],
           'binned': summary(x, info, ctx, M)}
    out['binned']['setup_s'] = info['setup_s']
    print('binned', json.dumps(out['binned']), file=sys.stderr, flush=True)
    # two-level: binned + along-scan coarse space (knots every h pixels)
    from ..mapmaker import TwoLevelPreconditioner
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T = TwoLevelPreconditioner(mm.apply_A, M, mm.P.map_shape, h=h)
    torch.cuda.synchronize()
    t_tl = time.perf_counter() - t0
    b = mm.rhs(ctx['d'])
    x2, info2 = pcg(mm.apply_A, b, T, tol=tol, maxiter=maxiter)
    out['two_level'] = summary(x2, info2, ctx, M)
    out['two_level'].update({'setup_s': t_tl, 'knot_spacing': h,
                             'coarse_dofs': int(3 * T.nk * ctx['shape'][0])})
    out['speedup_two_level'] = (out['binned']['seconds'] + out['binned']['setup_s']) / \
        (info2['seconds'] + t_tl)
    print('two_level', json.dumps(out['two_level']), file=sys.stderr, flush=True)
    if not learned:
        return out
    # learned deflation space: Ritz vectors of the two-level-preconditioned
    # operator's smallest eigenvalues, from the Lanczos process of a solve on
    # an independent realisation (the training solve)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Us, thetas = [], []
    for j in range(n_train_solves):
        b_train = mm.rhs(make_data(ctx, 100 + j)[1])
        th, Uj = harvest_ritz(mm.apply_A, T, b_train, harvest_iters, k_defl)
        Us.append(Uj)
        thetas.append(th)
    D = DeflatedPreconditioner(mm.apply_A, T, torch.cat(Us), mode='additive')
    theta = np.concatenate(thetas)
    del Us
    torch.cuda.synchronize()
    t_defl = time.perf_counter() - t0
    x4, info4 = pcg(mm.apply_A, b, D, tol=tol, maxiter=maxiter)
    out['deflated'] = summary(x4, info4, ctx, M)
    out['deflated'].update({'learn_s': t_defl, 'harvest_iters': harvest_iters,
                            'training_solves': n_train_solves, 'ritz_vectors': D.k,
                            'ritz_values_max': float(theta.max())})
    out['speedup_deflated_vs_binned'] = out['binned']['seconds'] / out['deflated']['seconds']
    out['speedup_deflated_vs_two_level'] = out['two_level']['seconds'] / \
        out['deflated']['seconds']
    print('deflated', json.dumps(out['deflated']), file=sys.stderr, flush=True)
    # reference solution: the deflated PCG run to ref_tol.  Solutions that
    # meet the same tol differ in the near-null modes of A (Q/U on rows seen
    # at one angle, per-row offsets), whose components of A^-1 b are mostly
    # noise; the closer a solve gets to A^-1 b, the more of that noise its
    # map carries, so the sky error is compared with the reference's own.
    xr, infor = pcg(mm.apply_A, b, D, tol=ref_tol, maxiter=maxiter)
    out['reference'] = {'tol': ref_tol, 'iters': infor['iters'],
                        'final_residual': infor['residuals'][-1],
                        'true_residual': true_residual(mm, xr, b),
                        'map_rms_error_TQU': map_error(xr, ctx['sky'], M.mask)}
    for k, xk in (('binned', x), ('two_level', x2), ('deflated', x4)):
        out[k]['a_norm_error_vs_reference'] = a_norm_error(mm, xk, xr)
        out[k]['map_rms_diff_vs_reference_TQU'] = map_error(xk, xr, M.mask)
    print('reference', json.dumps(out['reference']), file=sys.stderr, flush=True)
    if not unet:
        return out
    # U-Net correction on top of the deflated preconditioner, trained on
    # residual/error pairs of its own CG solves on independent realisations
    t0 = time.perf_counter()
    rhs = [mm.rhs(make_data(ctx, 200 + j)[1]) for j in range(n_rhs)]
    R, E = collect_residual_error_pairs(mm, D, rhs, per_rhs)
    torch.cuda.synchronize()
    t_pairs = time.perf_counter() - t0
    net = UNet(seed=0, factor=factor)
    t0 = time.perf_counter()
    hist = train_correction(net, D, R, E, steps=steps, batch=batch, lr=lr,
                            log=lambda s, l: print(f'step {s}: loss {l:.4f}', file=sys.stderr,
                                                   flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    n_pairs = int(R.shape[0])
    del R, E
    L = LearnedPreconditioner(D, net, mm.P.map_shape)
    L(mm.zeros() + 1.)       # warm-up
    x3, info3 = pcg(mm.apply_A, b, L, tol=tol, maxiter=maxiter, flexible=True, fallback=D)
    out['learned'] = summary(x3, info3, ctx, M)
    out['learned']['a_norm_error_vs_reference'] = a_norm_error(mm, x3, xr)
    out['learned']['map_rms_diff_vs_reference_TQU'] = map_error(x3, xr, M.mask)
    out['learned'].update({'pairs': n_pairs, 'pairs_s': t_pairs,
                           'pool_factor': list(factor),
                           'train_steps': steps, 'train_s': t_train,
                           'loss_first_last': [float(np.mean(hist[:20])),
                                               float(np.mean(hist[-20:]))]})
    out['speedup_learned_vs_binned'] = out['binned']['seconds'] / out['learned']['seconds']
    out['speedup_learned_vs_two_level'] = out['two_level']['seconds'] / out['learned']['seconds']
    out['speedup_learned_vs_deflated'] = out['deflated']['seconds'] / out['learned']['seconds']
    print('learned', json.dumps(out['learned']), file=sys.stderr, flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--maxiter', type=int, default=40000)
    ap.add_argument('--no-learned', action='store_true')
    ap.add_argument('--no-unet', action='store_true')
    ap.add_argument('--steps', type=int, default=3000)
    ap.add_argument('--per-rhs', type=int, default=500)
    ap.add_argument('--n-rhs', type=int, default=4)
    ap.add_argument('--lr', type=float, default=3e-4)
    ap.add_argument('--factor', type=int, nargs=2, default=(1, 4))
    a = ap.parse_args(argv)
    print(json.dumps(run(a.tol, a.maxiter, not a.no_learned, a.n_rhs, a.per_rhs, a.steps,
                         lr=a.lr, factor=tuple(a.factor), unet=not a.no_unet)), flush=True)


if __name__ == '__main__':
    main()
