"""so3g's array operations on the map-making path, on the GPU.

nmat_detvecs_apply: the detector-vector noise model's N^-1 in Fourier space
(src/array_ops.cxx:39-88, bound as so3g.nmat_detvecs_apply at :1267), in
place on ft, with the reference's argument checks (dtype -> ValueError,
shape -> RuntimeError, layout -> TypeError).

process_cuts / translate_cuts: the cut degrees of freedom of ML map-making
(src/array_ops.cxx:130-217, bound at :1268-1269).
"""
import numpy as np

from . import _native as N
from .engine import _Call, _is_torch, _stream


def _c_contig(x, name):
    ok = x.is_contiguous() if _is_torch(x) else np.asarray(x).flags['C_CONTIGUOUS']
    if not ok:
        raise TypeError(f'{name} must be C-contiguous along last axis')


def nmat_detvecs_apply(ft, bins, iD, iV, s, norm, dct_binning=False):
    """ft[ndet, nmode] float32 (the real view of an rfft unless dct_binning),
    bins[nbin, 2] int32, iD[nbin, ndet], iV[nbin, ndet, nvec] float32: for
    each bin, ft[:, b] = ft * iD / norm + (s / norm) iV (iV^T ft), in place."""
    for x, name in ((ft, 'ft'), (bins, 'bins'), (iD, 'iD'), (iV, 'iV')):
        if x is not None and hasattr(x, 'shape'):
            _c_contig(x, name)
    call = _Call()
    ftd = call.array(ft, 'float32', 'ft', spec=[-1, -1], out=True)
    ndet, nmode = ftd.shape
    b = bins.cpu().numpy() if _is_torch(bins) else np.asarray(bins)
    if b.dtype != np.int32:
        raise ValueError("Expected buffer 'bins' to contain items of type int32.")
    if b.ndim != 2 or b.shape[1] != 2:
        raise RuntimeError("Buffer 'bins' has incompatible shape: Expected (*, 2).")
    b = np.ascontiguousarray(b)
    nbin = b.shape[0]
    iDd = call.array(iD, 'float32', 'iD', spec=[nbin, ndet])
    iVd = call.array(iV, 'float32', 'iV', spec=[nbin, ndet, -1])
    nvec = iVd.shape[2]
    N.check(N.load().so3g_nmat_detvecs_apply(ftd.data_ptr(), ndet, nmode, b.ctypes.data, nbin,
                                             iDd.data_ptr(), iVd.data_ptr(), nvec, float(s),
                                             float(norm), int(bool(dct_binning)), _stream()))
    call.finish()


_CUT_OPS = {'insert': 0, 'extract': 1, 'clear': 2}
_CUT_MODELS = {'full': 0, 'poly': 1}


def _cut_segments(range_matrix):
    """(n, 3) int32 (det, start, end) of a RangesMatrix / list of Ranges, in
    the order the reference walks it; also returns the sample count."""
    segs = []
    count = None
    for d, r in enumerate(range_matrix):
        count = r.count if count is None else count
        rr = np.asarray(r.ranges(), np.int32).reshape(-1, 2)
        if len(rr):
            seg = np.empty((len(rr), 3), np.int32)
            seg[:, 0] = d
            seg[:, 1:] = rr
            segs.append(seg)
    out = np.concatenate(segs) if segs else np.zeros((0, 3), np.int32)
    return np.ascontiguousarray(out, np.int32), (count or 0)


def _cut_params(model, params):
    if model not in _CUT_MODELS:
        raise ValueError("process_cuts model can only be 'full' or 'poly'")
    if model == 'poly':
        return int(params['resolution']), int(params['nmax'])
    return 0, 0


def _vals_dtype(x, name):
    dt = x.dtype if not _is_torch(x) else {torch_float32(): np.float32,
                                            torch_float64(): np.float64}.get(x.dtype)
    if dt not in (np.float32, np.float64):
        raise TypeError('process_cuts only supports float32 and float64')
    return 'float64' if dt == np.float64 else 'float32'


def torch_float32():
    import torch
    return torch.float32


def torch_float64():
    import torch
    return torch.float64


def process_cuts(range_matrix, operation, model, params, tod, vals):
    """array_ops.cxx:130-183.  operation 'measure' returns the number of cut
    values; 'insert' writes vals into the cut samples of tod (P), 'extract'
    gathers them into vals (P^T), 'clear' zeroes them.  tod (ndet, nsamp)
    and vals (n,) float32 or float64, numpy (updated in place) or torch."""
    resolution, nmax = _cut_params(model, params)
    segs, _ = _cut_segments(range_matrix)
    if operation == 'measure':
        import ctypes
        n = ctypes.c_int64()
        N.check(N.load().so3g_cuts_measure(segs.ctypes.data, len(segs), _CUT_MODELS[model],
                                           resolution, nmax, ctypes.byref(n)))
        return int(n.value)
    if operation not in _CUT_OPS:
        raise ValueError("process_cuts operation can only be 'measure', 'insert' or 'extract'")
    dt = _vals_dtype(tod, 'tod')
    call = _Call()
    t = call.array(tod, dt, 'tod', spec=[-1, -1], out=operation != 'extract')
    v = call.array(vals, dt, 'vals', spec=[-1], out=operation == 'extract')
    N.check(N.load().so3g_process_cuts(segs.ctypes.data, len(segs), _CUT_OPS[operation],
                                       _CUT_MODELS[model], resolution, nmax, t.data_ptr(),
                                       t.shape[0], t.shape[1], v.data_ptr(), v.shape[0],
                                       int(dt == 'float64'), _stream()))
    call.finish()
    return 0


def translate_cuts(irange_matrix, orange_matrix, model, params, ivals, ovals):
    """array_ops.cxx:185-217: cut values of irange_matrix (sample rate
    insamp = its count) mapped onto orange_matrix's ranges, into ovals."""
    resolution, nmax = _cut_params(model, params)
    isegs, insamp = _cut_segments(irange_matrix)
    osegs, onsamp = _cut_segments(orange_matrix)
    if len(isegs) != len(osegs):
        raise RuntimeError('input and output cuts must pair up range by range')
    dt = _vals_dtype(ivals, 'ivals')
    call = _Call()
    iv = call.array(ivals, dt, 'ivals', spec=[-1])
    ov = call.array(ovals, dt, 'ovals', spec=[-1], out=True)
    N.check(N.load().so3g_translate_cuts(isegs.ctypes.data, osegs.ctypes.data, len(isegs),
                                         _CUT_MODELS[model], resolution, nmax, insamp, onsamp,
                                         iv.data_ptr(), iv.shape[0], ov.data_ptr(), ov.shape[0],
                                         int(dt == 'float64'), _stream()))
    call.finish()
