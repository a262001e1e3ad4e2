"""CPU restatement of nmat_detvecs_apply (src/array_ops.cxx:39-88) in
numpy float64.  Test infrastructure only.

For each bin in order, columns [min(k b0, nmode - 1), min(k b1, nmode))
(k = 2 for the real view of an rfft, 1 with dct_binning) are replaced by
ft * iD / norm + (s / norm) iV (iV^T ft); an empty range is skipped.
"""
import numpy as np


def nmat_detvecs_apply(ft, bins, iD, iV, s, norm, dct_binning=False):
    out = np.array(ft, np.float64)
    nmode = out.shape[1]
    k = 1 if dct_binning else 2
    iD = np.asarray(iD, np.float64)
    iV = np.asarray(iV, np.float64)
    for bi, (a, b) in enumerate(np.asarray(bins)):
        b1 = min(k * int(a), nmode - 1)
        b2 = min(k * int(b), nmode)
        if b2 <= b1:
            continue
        f = out[:, b1:b2]
        q = iV[bi].T @ f
        out[:, b1:b2] = f * (iD[bi][:, None] / norm) + (s / norm) * (iV[bi] @ q)
    return out
