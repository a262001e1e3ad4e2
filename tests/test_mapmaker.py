"""Map-maker pieces: noise model and sky generator (CPU), and the CG solve
against a dense CPU restatement of A = P^T N^-1 P (GPU)."""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip('torch')


def test_noise_model_psd_cpu():
    from so3g_amd.mapmaker import NoiseModel
    n_det, n_t, fs = 3, 4096, 200.
    sigma = np.array([0.5, 1., 2.])
    nm = NoiseModel(sigma, n_t, fsample=fs, device=torch.device('cpu'))
    g = torch.Generator().manual_seed(0)
    acc = 0
    n_real = 200
    for _ in range(n_real):
        x = nm.simulate(g).double()
        acc = acc + torch.fft.rfft(x, dim=1).abs() ** 2 / n_t
    per = (acc / n_real).numpy()
    f = np.fft.rfftfreq(n_t, 1 / fs)
    psd = sigma[:, None] ** 2 * (1 + (f[None, 1:] / 1.) ** -3)
    ratio = per[:, 1:] / psd
    # averaged periodogram / PSD: mean 1 within the scatter of 200 realisations
    assert abs(ratio.mean() - 1) < 0.02
    assert np.all(per[:, 0] < 1e-9 * per[:, 1:].mean())
    # N^-1 is the inverse of the simulated covariance: whitened noise has unit
    # variance per Fourier mode (after the 1/PSD^(1/2) halves)
    x = nm.simulate(g).double()
    y = torch.fft.irfft(torch.fft.rfft(x, dim=1) * nm.inv_psd.sqrt(), n=n_t, dim=1)
    assert abs(float(y[:, :].std()) - 1) < 0.05


def test_grf_spectrum_cpu():
    from so3g_amd.studies.cg import grf_tqu
    ny, nx, res = 256, 256, np.radians(1 / 60)
    m = grf_tqu((ny, nx), res, seed=3)
    ly = np.fft.fftfreq(ny, res) * 2 * np.pi
    lx = np.fft.fftfreq(nx, res) * 2 * np.pi
    ell = np.hypot(*np.meshgrid(ly, lx, indexing='ij'))
    P = np.abs(np.fft.fft2(m[0])) ** 2
    sel = (ell > 2000) & (ell < 40000)
    slope = np.polyfit(np.log(ell[sel]), np.log(P[sel]), 1)[0]
    assert abs(slope + 2.5) < 0.1
    # E and B at 0.1 x the T power: Q^2 + U^2 = E^2 + B^2 per mode
    pq = (np.abs(np.fft.fft2(m[1])) ** 2 + np.abs(np.fft.fft2(m[2])) ** 2)[sel].sum()
    assert abs(pq / P[sel].sum() / 0.2 - 1) < 0.1


def dense_problem(n_det=16, n_t=1500):
    """Tiny observation (a 1 deg throw crossed twice by 16 detectors, 10
    arcmin pixels) with its dense P (rows det-major)."""
    from so3g_amd import workloads
    obs = workloads.Observation(n_det, n_t, 1.0, 10.0, throw_deg=1., seed=1)
    p = obs.projectionist()
    q = np.asarray(p._cache_q_fp_to_native(obs.asm.Q), np.float64)
    args = p._get_pixelizor_args()
    og = oracle.geom_from_args('CAR', 'TQU', args, 0)
    pix, spin = oracle.pointing_matrix(og, q, obs.dets, obs.resps)
    ny, nx = obs.shape
    npx = ny * nx
    P = np.zeros((n_det * n_t, 3 * npx))
    for d in range(n_det):
        for t in range(n_t):
            iy, ix = pix[d, t]
            if iy < 0:
                continue
            for c in range(3):
                P[d * n_t + t, c * npx + iy * nx + ix] = spin[d, t, c]
    return obs, q, P, (ny, nx)


def dense_A(P, nm, n_det, n_t):
    """P^T N^-1 P with each detector's N^-1 block built as a dense circulant
    (the Fourier filter applied to the identity), and N^-1 applied to a
    vector the same way."""
    inv = nm.inv_psd.cpu().numpy()
    eye = np.eye(n_t)
    A = 0
    blocks = []
    for d in range(n_det):
        blk = np.fft.irfft(np.fft.rfft(eye, axis=0) * inv[d][:, None], n=n_t, axis=0)
        Pd = P[d * n_t:(d + 1) * n_t]
        A = A + Pd.T @ blk @ Pd
        blocks.append(blk)
    def ninv(v):
        return np.concatenate([blocks[d] @ v[d * n_t:(d + 1) * n_t] for d in range(n_det)])
    return A, ninv


@pytest.mark.gpu
def test_cg_against_dense(gpu):
    import so3g_amd
    from so3g_amd.mapmaker import MapMaker, NoiseModel, PointingOperator, pcg
    n_det, n_t = 16, 1500
    obs, q, P, (ny, nx) = dense_problem(n_det, n_t)
    dev = torch.device('cuda')
    eng = obs.projectionist().get_ProjEng('TQU')
    pix, spin = eng.pointing_matrix(torch.from_numpy(q).to(dev),
                                    torch.from_numpy(obs.dets).to(dev),
                                    torch.from_numpy(obs.resps).to(dev))
    sigma = np.exp(np.random.default_rng(4).normal(0, 0.3, n_det))
    nm = NoiseModel(sigma, n_t, device=dev)
    A, ninv = dense_A(P, nm, n_det, n_t)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(3 * ny * nx)
    ref = A @ x
    for use_plan, precise, tol in [(True, False, 1e-5), (False, False, 1e-5),
                                   (True, True, 1e-12)]:
        mm = MapMaker(PointingOperator(pix, spin, (ny, nx), use_plan=use_plan), nm,
                      precise=precise)
        Ax = mm.apply_A(torch.from_numpy(x.reshape(3, ny, nx)).to(dev)).cpu().numpy().ravel()
        assert np.abs(Ax - ref).max() < tol * np.abs(ref).max(), (use_plan, precise)
    # binned preconditioner = inverse of the per-pixel white-noise blocks
    M = mm.binned()
    W = P.T @ np.diag(np.repeat(1 / sigma ** 2, n_t)) @ P
    good = M.mask.cpu().numpy().ravel()
    npx = ny * nx
    assert good.sum() > 20
    for ipx in np.nonzero(good)[0][:20]:
        idx = [c * npx + ipx for c in range(3)]
        blk = W[np.ix_(idx, idx)]
        inv = M.inv.reshape(3, 3, -1)[:, :, ipx].cpu().numpy()
        # the GPU weights are so3g's float32 products (and float32 1/sigma^2)
        np.testing.assert_allclose(inv @ blk, np.eye(3), atol=2e-6)
    # CG to 1e-6 solves the dense system
    sky = rng.standard_normal((3, ny, nx))
    d = (P @ sky.ravel()).reshape(n_det, n_t).astype(np.float32)
    b = mm.rhs(torch.from_numpy(d).to(dev))
    bd = P.T @ ninv(d.astype(np.float64).ravel())
    assert np.abs(b.cpu().numpy().ravel() - bd).max() < 1e-12 * np.abs(bd).max()
    xs, info = pcg(mm.apply_A, b, M, tol=1e-8, maxiter=2000)
    assert info['residuals'][-1] < 1e-8
    xv = xs.cpu().numpy().ravel()
    assert np.linalg.norm(A @ xv - bd) / np.linalg.norm(bd) < 1e-7
    # flexible CG with the same (linear) preconditioner also converges
    xs2, info2 = pcg(mm.apply_A, b, M, tol=1e-6, maxiter=500, flexible=True)
    assert info2['residuals'][-1] < 1e-6


@pytest.mark.gpu
def test_two_level_preconditioner(gpu):
    """M_b r + Z (Z^T A Z)^-1 Z^T r against a dense build (A from the dense
    problem; rows do not couple for a constant-elevation scan, so probing
    all rows at once is exact), and PCG with it."""
    from so3g_amd.mapmaker import (MapMaker, NoiseModel, PointingOperator,
                                   TwoLevelPreconditioner, pcg)
    n_det, n_t = 16, 1500
    obs, q, P, (ny, nx) = dense_problem(n_det, n_t)
    dev = torch.device('cuda')
    eng = obs.projectionist().get_ProjEng('TQU')
    pix, spin = eng.pointing_matrix(torch.from_numpy(q).to(dev),
                                    torch.from_numpy(obs.dets).to(dev),
                                    torch.from_numpy(obs.resps).to(dev))
    sigma = np.exp(np.random.default_rng(4).normal(0, 0.3, n_det))
    nm = NoiseModel(sigma, n_t, device=dev)
    A, ninv = dense_A(P, nm, n_det, n_t)
    mm = MapMaker(PointingOperator(pix, spin, (ny, nx)), nm)
    M = mm.binned()
    T = TwoLevelPreconditioner(mm.apply_A, M, (3, ny, nx), h=4, ridge=1e-8)
    H = T.H.cpu().numpy()
    nk = H.shape[1]
    npx = ny * nx
    # Z: hat k on row j, component c -> map index c * npx + j * nx + x
    Z = np.zeros((3 * npx, 3 * ny * nk))
    for c in range(3):
        for j in range(ny):
            Z[c * npx + j * nx:c * npx + (j + 1) * nx, (j * 3 + c) * nk:(j * 3 + c + 1) * nk] = H
    Ac = Z.T @ A @ Z
    rng = np.random.default_rng(1)
    r = rng.standard_normal((3, ny, nx))
    rt = torch.from_numpy(r).to(dev)
    z = T(rt).cpu().numpy().ravel()
    zb = M(rt).cpu().numpy().ravel()
    # the same regularisation: ridge x the diagonal, 1 where it is empty
    d = np.diag(Ac)
    Ac_reg = Ac + np.diag(np.where(d > 0, 1e-8 * d, 1.))
    coarse = Z @ np.linalg.solve(Ac_reg, Z.T @ r.ravel())
    err = np.abs((z - zb) - coarse).max() / np.abs(coarse).max()
    assert err < 1e-5, err
    b = mm.rhs(torch.from_numpy(
        (P @ rng.standard_normal(3 * npx)).reshape(n_det, n_t).astype(np.float32)).to(dev))
    _, info_b = pcg(mm.apply_A, b, M, tol=1e-8, maxiter=3000)
    _, info_t = pcg(mm.apply_A, b, T, tol=1e-8, maxiter=3000)
    assert info_t['residuals'][-1] < 1e-8
    assert info_t['iters'] < info_b['iters']
