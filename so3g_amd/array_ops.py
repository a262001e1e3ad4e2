"""so3g's array operations on the map-making path, on the GPU.

nmat_detvecs_apply: the detector-vector noise model's N^-1 in Fourier space
(src/array_ops.cxx:39-88, bound as so3g.nmat_detvecs_apply at :1267), in
place on ft, with the reference's argument checks (dtype -> ValueError,
shape -> RuntimeError, layout -> TypeError).
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
