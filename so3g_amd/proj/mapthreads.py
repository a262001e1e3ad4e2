"""Thread assignments: disjoint RangesMatrix objects of shape
(n_threads, n_dets, n_samps) for the threads= argument
(python/proj/mapthreads.py:1-203).

On the GPU every listed sample is processed with atomic map updates, so a
thread assignment no longer drives parallelism; it is still honoured as a
work list (samples outside it are skipped), and these routines still
produce assignments with the reference's coverage properties.
"""
import os

import numpy as np

from .ranges import RangesMatrix


def get_num_threads(n_threads=None):
    """n_threads if given, else the so3g_amd.useful_info() thread count."""
    if n_threads is None:
        from .. import useful_info
        return useful_info()['omp_num_threads']
    return n_threads


def get_threads_domdir(sight, fplane, shape, wcs, tile_shape=None, active_tiles=None,
                       n_threads=None, fplane_rep=None, plot_prefix=None):
    """Assign samples to threads in bands parallel to the dominant scan
    direction, balanced by hit count (the 'domdir' method).

    The scan direction comes from the Q/U of a map made with detectors whose
    polarisation is aligned with elevation; pixels are then labelled by
    their coordinate across that direction, the hit-weighted CDF of the
    label is cut into n_threads equal parts, and the resulting per-pixel
    thread map is turned into sample ranges with pixel_ranges."""
    from . import wcs as _wcs
    from . import coords as _coords
    from . import quat as _quat
    n_threads = get_num_threads(n_threads)
    if fplane_rep is None:
        fplane_rep = fplane
    if tile_shape is None:
        tile_shape = shape
        active_tiles = [0]
    asm_full = _coords.Assembly.attach(sight, fplane)
    pmat = _wcs.Projectionist.for_tiled(shape, wcs, tile_shape=tile_shape,
                                        active_tiles=active_tiles)
    if active_tiles is None:
        pmat.active_tiles = pmat.get_active_tiles(asm_full, assign=True)['active_tiles']
    tiling = pmat.tiling

    xi, eta, gamma = _quat.decompose_xieta(fplane_rep.quats)
    fplane_xl = _coords.FocalPlane.from_xieta(xi, eta, gamma * 0 + 90 * _quat.DEG)
    asm_rep = _coords.Assembly.attach(sight, fplane_xl)
    sig = np.ones((fplane_xl.ndet, len(asm_rep.Q)), dtype='float32')
    scan_maps = pmat.to_map(sig, asm_rep, comps='TQU')

    if len(pmat.active_tiles) > 0:
        T, Q, U = np.sum([_m.reshape((3, -1)).sum(axis=-1)
                          for _m in scan_maps if _m is not None], axis=0)
    else:
        T, Q, U = 0., 0., 0.
    if T == 0:
        n_dets, n_samps = len(asm_full.dets), len(asm_full.Q)
        return [RangesMatrix.zeros(shape=(n_threads, n_dets, n_samps))]

    phi = np.arctan2(U, Q) / 2
    idx_maps = pmat.zeros((1,))
    lims = None
    for t in pmat.active_tiles:
        y0, x0 = tiling.tile_offset(t)
        ny, nx = tiling.tile_dims(t)
        y, x = y0 + np.arange(ny), x0 + np.arange(nx)
        idx_maps[t] = y[:, None] * np.sin(phi) - x[None, :] * np.cos(phi)
        lo, hi = idx_maps[t].min(), idx_maps[t].max()
        lims = (lo, hi) if lims is None else (min(lo, lims[0]), max(hi, lims[1]))

    n_bins = 200
    bins = np.linspace(lims[0], lims[1], n_bins + 1)
    H = np.zeros(n_bins)
    for t in pmat.active_tiles:
        H += np.histogram(idx_maps[t].ravel(), bins=bins, weights=scan_maps[t][0].ravel())[0]
    del scan_maps

    # Strictly increasing CDF of hits along the label.
    H = np.hstack((0, H.cumsum()))
    bias = .00001
    H = H / H.max() * (1 - bias) + np.linspace(0, bias, len(H))
    xs = np.linspace(0, 1, n_threads + 1)
    ys = np.interp(xs, H, bins)
    superbins = np.hstack((bins[0], ys[1:-1], bins[-1]))

    for t in pmat.active_tiles:
        temp = np.zeros(idx_maps[t].shape)
        for i in range(n_threads):
            s = (superbins[i] <= idx_maps[t]) * (idx_maps[t] < superbins[i + 1])
            temp[s] = i
        idx_maps[t][:] = temp
        idx_maps[t].shape = (1,) + tuple(idx_maps[t].shape)

    return pmat.assign_threads_from_map(asm_full, idx_maps, n_threads=n_threads)
