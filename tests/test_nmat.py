"""nmat_detvecs_apply (src/array_ops.cxx:39-88): the numpy restatement
pinned by closed forms, then the HIP kernel against it."""
import numpy as np
import pytest

from oracle import nmat_ref as R

torch = pytest.importorskip('torch')


def blas_case():
    """test/test_blas.py:15-28: ones, iD = 2, iV = 0, two half bins."""
    ndet, nfreq, nbin, nvec = 50, 100000, 2, 3
    ft = np.ones((ndet, nfreq), np.float32)
    bins = np.array([[0, nfreq // 2], [nfreq // 2, nfreq]], np.int32)
    iD = np.ones((nbin, ndet), np.float32) * 2
    iV = np.zeros((nbin, ndet, nvec), np.float32)
    return ft, bins, iD, iV, 1., 1.2


def test_oracle_blas_case_closed_form():
    ft, bins, iD, iV, s, norm = blas_case()
    out = R.nmat_detvecs_apply(ft, bins, iD, iV, s, norm)
    # bin 0 covers every column (2 x 50000 = nmode); bin 1 starts at the
    # clamped nmode - 1 and transforms the last column a second time
    np.testing.assert_allclose(out[:, :-1], 2 / 1.2, rtol=1e-15)
    np.testing.assert_allclose(out[:, -1], (2 / 1.2) ** 2, rtol=1e-15)
    assert out[0, 0] != 1.


def test_oracle_dense_matrix_form():
    """Each bin applies N_b^-1 = diag(iD_b) / norm + (s / norm) iV_b iV_b^T to
    every column: built here as a dense matrix."""
    rng = np.random.default_rng(0)
    ndet, nmode, nvec = 7, 40, 3
    ft = rng.standard_normal((ndet, nmode))
    bins = np.array([[0, 5], [5, 12], [12, 20]], np.int32)     # x2: [0,10) [10,24) [24,40)
    iD = rng.uniform(0.5, 2, (3, ndet))
    iV = rng.standard_normal((3, ndet, nvec))
    s, norm = 0.7, 1.3
    out = R.nmat_detvecs_apply(ft, bins, iD, iV, s, norm)
    for bi, (c1, c2) in enumerate([(0, 10), (10, 24), (24, 40)]):
        Nb = np.diag(iD[bi]) / norm + (s / norm) * iV[bi] @ iV[bi].T
        np.testing.assert_allclose(out[:, c1:c2], Nb @ ft[:, c1:c2], rtol=1e-12)
    # symmetric operator: <x, N y> = <N x, y>
    y = rng.standard_normal((ndet, nmode))
    lhs = np.sum(ft * R.nmat_detvecs_apply(y, bins, iD, iV, s, norm))
    rhs = np.sum(out * y)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # dct binning: bin edges in columns, not pairs
    out1 = R.nmat_detvecs_apply(ft, bins, iD, iV, s, norm, dct_binning=True)
    Nb = np.diag(iD[1]) / norm + (s / norm) * iV[1] @ iV[1].T
    np.testing.assert_allclose(out1[:, 5:12], Nb @ ft[:, 5:12], rtol=1e-12)
    np.testing.assert_array_equal(out1[:, 20:], ft[:, 20:])


def test_argument_errors_cpu():
    import so3g_amd
    ft, bins, iD, iV, s, norm = blas_case()
    with pytest.raises(TypeError):
        so3g_amd.nmat_detvecs_apply(np.asfortranarray(ft[:4, :6]), bins, iD[:, :4],
                                    iV[:, :4], s, norm)


@pytest.mark.gpu
@pytest.mark.parametrize('ndet,nfreq,nvec,nbin', [(50, 1000, 3, 4), (1000, 5000, 20, 7),
                                                  (3000, 800, 5, 3), (40, 333, 0, 2),
                                                  (100, 2000, 64, 5), (67, 2501, 10, 6)])
def test_nmat_gpu_against_oracle(gpu, ndet, nfreq, nvec, nbin):
    import so3g_amd
    rng = np.random.default_rng(ndet + nfreq)
    nmode = 2 * nfreq
    ft = rng.standard_normal((ndet, nmode)).astype(np.float32)
    edges = np.unique(np.concatenate([[0, nfreq], rng.integers(1, nfreq, nbin - 1)]))
    bins = np.stack([edges[:-1], edges[1:]], 1).astype(np.int32)
    bins = np.concatenate([bins, [[nfreq - 3, nfreq + 10]]]).astype(np.int32)  # clamped, overlaps
    nb = len(bins)
    iD = rng.uniform(0.5, 2, (nb, ndet)).astype(np.float32)
    iV = (rng.standard_normal((nb, ndet, nvec)) / np.sqrt(ndet)).astype(np.float32)
    s, norm = 0.9, 1.1
    ref = R.nmat_detvecs_apply(ft, bins, iD, iV, s, norm)
    g = ft.copy()
    so3g_amd.nmat_detvecs_apply(g, bins, iD, iV, s, norm)
    err = np.abs(g - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
    # torch tensors in place, dct binning
    gt = torch.from_numpy(ft).cuda()
    so3g_amd.nmat_detvecs_apply(gt, bins, torch.from_numpy(iD).cuda(),
                                torch.from_numpy(iV).cuda(), s, norm, dct_binning=True)
    ref1 = R.nmat_detvecs_apply(ft, bins, iD, iV, s, norm, dct_binning=True)
    assert np.abs(gt.cpu().numpy() - ref1).max() / np.abs(ref1).max() < 2e-6
    # rows whose columns start 4 bytes past an 8-byte boundary (a view one
    # float into a larger buffer): 4-byte accesses
    buf = torch.zeros(ndet * nmode + 1, dtype=torch.float32, device='cuda')
    gv = buf[1:].view(ndet, nmode)
    gv.copy_(torch.from_numpy(ft))
    so3g_amd.nmat_detvecs_apply(gv, bins, torch.from_numpy(iD).cuda(),
                                torch.from_numpy(iV).cuda(), s, norm)
    assert np.abs(gv.cpu().numpy() - ref).max() / np.abs(ref).max() < 2e-6
    assert float(buf[0]) == 0.


@pytest.mark.gpu
def test_nmat_blas_smoke(gpu):
    """test/test_blas.py:test_smoke on the GPU."""
    import so3g_amd
    ft, bins, iD, iV, s, norm = blas_case()
    so3g_amd.nmat_detvecs_apply(ft, bins, iD, iV, s, norm)
    assert ft[0, 0] != 1.
    np.testing.assert_allclose(ft[:, :-1], np.float32(2) / np.float32(1.2), rtol=1e-6)
    np.testing.assert_allclose(ft[:, -1], (np.float32(2) / np.float32(1.2)) ** 2, rtol=1e-6)
