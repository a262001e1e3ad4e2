"""Maximum-likelihood map-making on the projection kernels.

Solves  P^T N^-1 P m = P^T N^-1 d  by preconditioned conjugate gradients,
where P is the cached pointing matrix (ProjEng_Precomp pixel_index /
spin_proj, applied with the HIP from_map / pixel-sorted to_map kernels) and
N the per-detector stationary noise covariance, applied in Fourier space.
The reference has no map-maker; this is the caller the projection
operators exist for (sotodlib's ML mapmaker uses so3g's ProjEng_Precomp the
same way: Projection.cxx:2012-2253 for P and P^T).

Preconditioners: the binned (per-pixel nc x nc block-diagonal) inverse of
P^T diag(1/sigma^2) P, and a learned correction on top of it (see unet.py),
used with flexible CG.
"""
import ctypes
import time

import numpy as np
import torch

from . import _native as N
from .engine import _stream


class NoiseModel:
    """Stationary noise with PSD_d(f) = sigma_d^2 (1 + (f / fknee)^alpha)
    per detector (white level sigma_d^2 per sample).  N^-1 is applied as
    irfft(rfft(x) / PSD) in float64; the f = 0 bin is projected out."""

    def __init__(self, sigma, n_t, fsample=200., fknee=1., alpha=-3., device=None):
        self.n_t = int(n_t)
        self.sigma = np.asarray(sigma, np.float64)
        f = np.fft.rfftfreq(self.n_t, 1. / fsample)
        with np.errstate(divide='ignore'):
            shape = 1. + (f / fknee) ** alpha
        shape[0] = np.inf
        psd = self.sigma[:, None] ** 2 * shape[None, :]
        dev = device or torch.device('cuda')
        self.inv_psd = torch.from_numpy(1. / psd).to(dev)
        amp = np.sqrt(psd)
        amp[:, 0] = 0.
        self.amp = torch.from_numpy(amp).to(dev)

    def apply_inverse(self, sig, out=None):
        """N^-1 sig, returned in sig's dtype (or written to out)."""
        x = torch.fft.rfft(sig.to(torch.float64), dim=1)
        x *= self.inv_psd
        y = torch.fft.irfft(x, n=self.n_t, dim=1)
        if out is None:
            return y.to(sig.dtype)
        out.copy_(y)
        return out

    def simulate(self, generator):
        w = torch.randn(self.inv_psd.shape[0], self.n_t, dtype=torch.float64,
                        device=self.inv_psd.device, generator=generator)
        x = torch.fft.rfft(w, dim=1) * self.amp
        return torch.fft.irfft(x, n=self.n_t, dim=1).to(torch.float32)


class PointingOperator:
    """P (map -> signal) and P^T (signal -> map) for a non-tiled flat map of
    shape (nc, ny, nx), through the cached pointing matrix; P^T runs on the
    pixel-sorted plan."""

    def __init__(self, pixel_index, spin_proj, map_shape, use_plan=True):
        from .engine import PrecompPlan
        self.pix = pixel_index.contiguous()
        self.spin = spin_proj.contiguous()
        self.n_det, self.n_t, _ = self.pix.shape
        self.nc = int(self.spin.shape[2])
        self.map_shape = (self.nc, int(map_shape[0]), int(map_shape[1]))
        self._shape = np.array(map_shape[:2], np.int64)
        self.plan = PrecompPlan(self.pix, self.spin, map_shape[:2]) if use_plan else None
        self.lib = N.load()

    def P(self, m, out):
        """out = P m (fresh)."""
        N.check(self.lib.so3g_precomp_from_map(m.data_ptr(), self._shape.ctypes.data, None,
                                               None, self.pix.data_ptr(), 2,
                                               self.spin.data_ptr(), self.nc, self.n_det,
                                               self.n_t, out.data_ptr(), 0, _stream()))
        return out

    def PT(self, sig, out, det_weights=None):
        """out = P^T sig (out is overwritten)."""
        out.zero_()
        dw = None if det_weights is None else det_weights.data_ptr()
        if self.plan is not None:
            N.check(self.lib.so3g_pc_plan_to_map(self.plan._h, out.data_ptr(), sig.data_ptr(),
                                                 dw, _stream()))
        else:
            N.check(self.lib.so3g_precomp_to_map(out.data_ptr(), self._shape.ctypes.data,
                                                 None, None, self.pix.data_ptr(), 2,
                                                 self.spin.data_ptr(), self.nc, self.n_det,
                                                 self.n_t, sig.data_ptr(), dw, None, 0,
                                                 _stream()))
        return out

    def P64(self, m, out):
        """out = P m in float64 (out: float64 (n_det, n_t))."""
        N.check(self.lib.so3g_precomp_from_map_f64(m.data_ptr(), self._shape.ctypes.data,
                                                   self.pix.data_ptr(), self.spin.data_ptr(),
                                                   self.nc, self.n_det, self.n_t,
                                                   out.data_ptr(), 0, _stream()))
        return out

    def PT64(self, sig, out, det_weights=None):
        """out = P^T sig with float64 products (sig: float64)."""
        out.zero_()
        dw = None if det_weights is None else det_weights.data_ptr()
        N.check(self.lib.so3g_pc_plan_to_map_f64(self.plan._h, out.data_ptr(), sig.data_ptr(),
                                                 dw, _stream()))
        return out

    def weights(self, det_weights=None):
        """P^T diag(det_weights) P as (nc, nc, ny, nx), upper triangle
        filled (so3g's to_weights) and mirrored here."""
        nc = self.nc
        w = torch.zeros((nc, nc) + self.map_shape[1:], dtype=torch.float64,
                        device=self.pix.device)
        dw = None if det_weights is None else det_weights.data_ptr()
        N.check(self.lib.so3g_precomp_to_weights(w.data_ptr(), self._shape.ctypes.data, None,
                                                 None, self.pix.data_ptr(), 2,
                                                 self.spin.data_ptr(), nc, self.n_det,
                                                 self.n_t, dw, None, 0, _stream()))
        for a in range(nc):
            for b in range(a):
                w[a, b] = w[b, a]
        return w


class BinnedPreconditioner:
    """M^-1 = per-pixel inverse of the nc x nc block of P^T diag(w) P.
    Pixels whose block is ill-conditioned (eigenvalue ratio below rcond: too
    few hits or angles) get its diagonal inverse instead, so that M stays
    positive definite on every hit pixel; unhit pixels get 0.  `mask` marks
    the well-conditioned pixels."""

    def __init__(self, weights, rcond=1e-3):
        nc = weights.shape[0]
        w = weights.reshape(nc, nc, -1).permute(2, 0, 1)          # (npix, nc, nc)
        ev = torch.linalg.eigvalsh(w)
        ok = (ev[:, 0] > rcond * ev[:, -1]) & (ev[:, -1] > 0)
        eye = torch.eye(nc, dtype=w.dtype, device=w.device)
        inv = torch.linalg.inv(torch.where(ok[:, None, None], w, eye))
        diag = torch.diagonal(w, dim1=1, dim2=2)
        dinv = torch.where(diag > 0, 1. / torch.where(diag > 0, diag, 1.), 0.)
        inv = torch.where(ok[:, None, None], inv, torch.diag_embed(dinv))
        self.inv = inv.permute(1, 2, 0).reshape(weights.shape).contiguous()
        self.mask = ok.reshape(weights.shape[2:])
        self.nc = nc

    def __call__(self, r):
        nc = self.nc
        out = torch.zeros_like(r)
        for a in range(nc):
            for b in range(nc):
                out[a] += self.inv[a, b] * r[b]
        return out


class TwoLevelPreconditioner:
    """Additive two-level preconditioner: M^-1 r = M_b r + Z A_c^-1 Z^T r.

    Z spans, on every map row and component, hat functions along the row
    (the scan direction of a constant-elevation scan) with knots every h
    pixels: the large-scale along-scan structure that 1/f weighting leaves
    poorly conditioned and a per-pixel preconditioner cannot reach.  The
    coarse operator A_c = Z^T A Z is probed: one application of A per
    (component, knot, row colour), on every row of the colour at once (rows
    k = colour mod colors); it is exact when rows within colors - 1 of each
    other do not couple (one row per detector for a constant-elevation scan
    in a CAR map).  A_c is block-diagonal by row, each block regularised by
    ridge x its diagonal and inverted on the host in float64."""

    def __init__(self, apply_A, binned, map_shape, h=16, colors=1, ridge=1e-8, device=None):
        nc, ny, nx = map_shape
        dev = device or torch.device('cuda')
        knots = np.arange(0, nx - 1 + h, h, dtype=np.float64)
        knots[-1] = min(knots[-1], nx - 1)
        knots = np.unique(knots)
        nk = len(knots)
        x = np.arange(nx, dtype=np.float64)
        H = np.zeros((nx, nk))
        for k in range(nk):
            lo = knots[k - 1] if k > 0 else knots[k]
            hi = knots[k + 1] if k < nk - 1 else knots[k]
            left = (x >= lo) & (x <= knots[k])
            right = (x >= knots[k]) & (x <= hi)
            if knots[k] > lo:
                H[left, k] = (x[left] - lo) / (knots[k] - lo)
            if hi > knots[k]:
                H[right, k] = (hi - x[right]) / (hi - knots[k])
            H[int(knots[k]), k] = 1.
        self.H = torch.from_numpy(H).to(dev)                       # (nx, nk)
        m = nc * nk
        Ac = np.zeros((ny, m, m))
        e = torch.zeros(map_shape, dtype=torch.float64, device=dev)
        for col in range(colors):
            rows = torch.arange(col, ny, colors, device=dev)
            rows_h = rows.cpu().numpy()
            for c in range(nc):
                for k in range(nk):
                    e.zero_()
                    e[c, rows, :] = self.H[:, k]
                    y = apply_A(e)                                   # (nc, ny, nx)
                    proj = torch.einsum('cyx,xk->yck', y[:, rows, :], self.H)  # (nr, nc, nk)
                    Ac[rows_h, :, c * nk + k] = proj.reshape(len(rows_h), m).cpu().numpy()
        Ac = 0.5 * (Ac + Ac.transpose(0, 2, 1))
        d = np.diagonal(Ac, axis1=1, axis2=2)
        reg = np.where(d > 0, ridge * d, 1.)
        inv = np.linalg.inv(Ac + np.einsum('ij,jk->ijk', reg, np.eye(m)))
        self.inv = torch.from_numpy(inv).to(dev)                   # (ny, m, m)
        self.M = binned
        self.shape = map_shape
        self.nk = nk

    def __call__(self, r):
        nc, ny, nx = self.shape
        zc = torch.einsum('cyx,xk->yck', r, self.H).reshape(ny, nc * self.nk, 1)
        vc = torch.bmm(self.inv, zc).reshape(ny, nc, self.nk)
        coarse = torch.einsum('yck,xk->cyx', vc, self.H)
        return self.M(r) + coarse


def _dot(a, b):
    return float(torch.sum(a * b))


def harvest_ritz(apply_A, M, b, iters, k, tol=1e-14):
    """Ritz vectors for the k smallest eigenvalues of the M-preconditioned
    operator M A, learned from the Lanczos process of an M-preconditioned CG
    run on the right-hand side b (Paige 1971; the CG <-> Lanczos relation of
    Golub & Van Loan, Matrix Computations, sec. 10.2).  With z_j = M r_j, the
    Lanczos vectors are q_j = (-1)^j z_j / sqrt(r_j . z_j) and the Lanczos
    tridiagonal has T_jj = 1/alpha_j + beta_(j-1)/alpha_(j-1), T_j,j+1 =
    sqrt(beta_j)/alpha_j.  Returns (theta (k,), U (k, *b.shape)) with U = Q Y,
    Y the eigenvectors of T for its k smallest eigenvalues theta.  The
    eigenvectors M A has for isolated small eigenvalues -- near-null modes of
    A that M does not resolve -- are properties of the operator, the same for
    every right-hand side, so a run on one realisation learns them for all."""
    x = torch.zeros_like(b)
    r = b.clone()
    z = M(r)
    rz = _dot(r, z)
    bn = np.sqrt(_dot(b, b))
    Q, alphas, betas = [], [], []
    p = z.clone()
    for j in range(iters):
        Q.append(((-1) ** j / np.sqrt(rz)) * z)
        Ap = apply_A(p)
        alpha = rz / _dot(p, Ap)
        alphas.append(alpha)
        x.add_(p, alpha=alpha)
        r.sub_(Ap, alpha=alpha)
        if np.sqrt(_dot(r, r)) / bn < tol:
            break
        z = M(r)
        rz_new = _dot(r, z)
        betas.append(rz_new / rz)
        rz = rz_new
        p.mul_(betas[-1]).add_(z)
    m = len(alphas)
    T = np.zeros((m, m))
    for j in range(m):
        T[j, j] = 1 / alphas[j] + (betas[j - 1] / alphas[j - 1] if j > 0 else 0.)
        if j + 1 < m:
            T[j, j + 1] = T[j + 1, j] = np.sqrt(betas[j]) / alphas[j]
    theta, Y = np.linalg.eigh(T)
    k = min(k, m)
    Yk = torch.from_numpy(np.ascontiguousarray(Y[:, :k])).to(b.device)
    Qm = torch.stack(Q).reshape(m, -1)
    U = (Yk.T @ Qm).reshape((k,) + tuple(b.shape))
    return theta[:k], U


class DeflatedPreconditioner:
    """A base SPD preconditioner M plus a deflation space W (vectors learned
    from earlier solves, harvest_ritz):

      balanced (default, "A-DEF2", Tang, Nabben, Vuik & Erlangga 2009):
        z = P^T M P r + W E^-1 W^T r,  P = I - A W E^-1 W^T,  E = W^T A W;
      additive:  z = M r + W E^-1 W^T r.

    Both are symmetric positive definite maps of r (plain PCG applies): the
    deflated directions get an exact coarse solve, the rest M.  W is first
    orthonormalised (directions whose Gram eigenvalue is below rtol of the
    largest are dropped: Lanczos repeats converged Ritz vectors) and A W is
    kept, so an application costs one M, a few dot products of k map
    vectors, and no extra A."""

    def __init__(self, apply_A, base, W, mode='balanced', rtol=1e-10):
        self.M = base
        self.mode = mode
        k = W.shape[0]
        self.shape = tuple(W.shape[1:])
        Wf = W.reshape(k, -1)
        G = Wf @ Wf.T
        ev, V = torch.linalg.eigh(0.5 * (G + G.T))
        keep = ev > rtol * ev.max()
        B = V[:, keep] / ev[keep].sqrt()
        self.W = (B.T @ Wf).contiguous()                        # (k', n), orthonormal
        self.AW = torch.stack([apply_A(w.reshape(self.shape)).reshape(-1) for w in self.W])
        E = self.W @ self.AW.T
        E = 0.5 * (E + E.T)
        ee, EV = torch.linalg.eigh(E)
        ee = ee.clamp_min(ee.max() * 1e-14)
        self.Einv = (EV / ee) @ EV.T
        # W E^-1 W^T = V V^T with V^T = E^-1/2 W: the additive correction is
        # two matrix-vector products
        self.Vt = ((EV / ee.sqrt()) @ EV.T @ self.W).contiguous()
        self.k = int(self.W.shape[0])

    def __call__(self, r):
        rf = r.reshape(-1)
        if self.mode == 'additive':
            z = self.M(r)
            z.view(-1).addmv_(self.Vt.T, self.Vt @ rf)
            return z
        c = self.Einv @ (self.W @ rf)                            # E^-1 W^T r
        z1 = self.M((rf - self.AW.T @ c).reshape(self.shape)).reshape(-1)
        z = z1 - self.W.T @ (self.Einv @ (self.AW @ z1)) + self.W.T @ c
        return z.reshape(self.shape)


def pcg(apply_A, b, precond, tol=1e-6, maxiter=2000, x0=None, flexible=False,
        callback=None, fallback=None, stall=(200, 0.9)):
    """(Flexible) preconditioned CG.  Stops when ||r|| / ||b|| < tol.
    flexible: Polak-Ribiere beta (Notay's FCG(1)), for preconditioners that
    are not fixed linear maps.  fallback: a second preconditioner; when the
    residual has not fallen below stall[1] x its value stall[0] iterations
    earlier, CG restarts from the current x with the fallback (plain PCG).
    callback(it, x, r, p, Ap) after each iteration.  Returns (x, info) with
    info = dict(iters, residuals, seconds, fallback_at)."""
    t0 = time.perf_counter()
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = b - apply_A(x) if x0 is not None else b.clone()
    bnorm = np.sqrt(_dot(b, b))
    res = [np.sqrt(_dot(r, r)) / bnorm]
    z = precond(r)
    p = z.clone()
    rz = _dot(r, z)
    it = 0
    fallback_at = None
    while res[-1] >= tol and it < maxiter:
        Ap = apply_A(p)
        pAp = _dot(p, Ap)
        if not pAp > 0:
            break       # breakdown: p is in the null space of A (or M is not SPD)
        alpha = rz / pAp
        x.add_(p, alpha=alpha)
        r_old = r.clone() if flexible else None
        r.sub_(Ap, alpha=alpha)
        it += 1
        res.append(np.sqrt(_dot(r, r)) / bnorm)
        if callback is not None:
            callback(it, x, r, p, Ap)
        if res[-1] < tol:
            break
        if (fallback is not None and fallback_at is None and it > stall[0] and
                res[-1] > stall[1] * res[-1 - stall[0]]):
            # stalled: restart as plain PCG with the fallback preconditioner
            fallback_at = it
            precond, flexible = fallback, False
            z = precond(r)
            p = z.clone()
            rz = _dot(r, z)
            continue
        z = precond(r)
        if flexible:
            beta = _dot(z, r - r_old) / rz
            rz = _dot(r, z)
        else:
            rz_new = _dot(r, z)
            beta = rz_new / rz
            rz = rz_new
        p.mul_(beta).add_(z)
    torch.cuda.synchronize()
    return x, {'iters': it, 'residuals': res, 'seconds': time.perf_counter() - t0,
               'fallback_at': fallback_at}


class MapMaker:
    """A = P^T N^-1 P on one observation; rhs(d) = P^T N^-1 d.

    precise (default): the time-ordered data inside A are float64 (the
    float64 gather and plan scatter), so A is symmetric to float64 rounding.
    With float32 time streams (so3g's own arithmetic) the rounding of P m,
    of order 1e-7 of the signal and white, passes N^-1 at full weight while
    the 1/f-dominated signal modes are suppressed by up to 1e-7, and CG
    stalls near ||r||/||b|| ~ 1e-4."""

    def __init__(self, pointing, noise, precise=True):
        self.P = pointing
        self.noise = noise
        self.precise = precise
        if precise and pointing.plan is None:
            raise ValueError('precise map-making needs the pointing plan')
        self._sig = torch.empty((pointing.n_det, pointing.n_t),
                                dtype=torch.float64 if precise else torch.float32,
                                device=pointing.pix.device)

    def zeros(self):
        return torch.zeros(self.P.map_shape, dtype=torch.float64, device=self.P.pix.device)

    def apply_A(self, m):
        if self.precise:
            self.P.P64(m, self._sig)
            return self.P.PT64(self.noise.apply_inverse(self._sig), self.zeros())
        self.P.P(m, self._sig)
        return self.P.PT(self.noise.apply_inverse(self._sig), self.zeros())

    def rhs(self, d):
        if self.precise:
            return self.P.PT64(self.noise.apply_inverse(d.to(torch.float64)), self.zeros())
        return self.P.PT(self.noise.apply_inverse(d), self.zeros())

    def binned(self, rcond=1e-3):
        w = torch.from_numpy(1. / self.noise.sigma ** 2).to(torch.float32).to(
            self.P.pix.device)
        return BinnedPreconditioner(self.P.weights(w), rcond=rcond)


SMOOTH = (15, 63, 255)


def features(mr):
    """Network input for binned-preconditioned residuals mr (n, nc, ny, nx):
    mr / s and its running means along the scan direction (map rows) over
    SMOOTH pixels, each scaled to unit rms, so that the large-scale part --
    the part the binned preconditioner leaves unresolved -- is not lost to
    bf16 rounding next to the pixel-scale part.  Returns (features
    (n, nc * (1 + len(SMOOTH)), ny, nx) float32, s (n,))."""
    mr = mr.to(torch.float32)
    s = mr.pow(2).mean(dim=(1, 2, 3)).sqrt().clamp_min(1e-30)
    out = [mr / s[:, None, None, None]]
    n, nc, ny, nx = mr.shape
    for w in SMOOTH:
        sm = torch.nn.functional.avg_pool2d(mr, (1, w), stride=1, padding=(0, w // 2),
                                            count_include_pad=False)
        ss = sm.pow(2).mean(dim=(1, 2, 3)).sqrt().clamp_min(1e-30)
        out.append(sm / ss[:, None, None, None])
    return torch.cat(out, dim=1), s


class LearnedPreconditioner:
    """z = M r + s * U(features(M r)), s = rms(M r): a base preconditioner M
    (binned or two-level) plus a U-Net correction trained on (A e, e) pairs.
    Not a fixed linear map: use with flexible CG."""

    def __init__(self, base, net, map_shape):
        self.M = base
        self.net = net
        self.shape = map_shape      # (nc, ny, nx)

    def __call__(self, r):
        mr = self.M(r)
        x, s = features(mr[None])
        nc, ny, nx = self.shape
        y, _ = self.net.forward(self.net.to_input(x))
        c = self.net.from_output(y, nc, ny, nx)[0].to(torch.float64)
        return mr + float(s[0]) * c


def collect_pairs(mm, M, rhs_list, per_rhs, tol=1e-12):
    """(A p, p) pairs from the search directions of binned PCG runs on each
    right-hand side: r = A e with e = p, the directions CG still has to
    resolve.  Returns two float32 tensors (n, nc, ny, nx)."""
    R, E = [], []
    def cb(it, x, r, p, Ap):
        R.append(Ap.to(torch.float32))
        E.append(p.to(torch.float32))
    for b in rhs_list:
        pcg(mm.apply_A, b, M, tol=tol, maxiter=per_rhs, callback=cb)
    return torch.stack(R), torch.stack(E)


def collect_residual_error_pairs(mm, M, rhs_list, per_rhs, tol_exact=1e-8, rmin=1e-5):
    """(r_j, e_j) pairs along M-preconditioned CG runs on each right-hand
    side: r_j = b - A x_j and e_j = x* - x_j, with x* the solve to tol_exact
    (a tighter solve amplifies rounding along A's near-null space, where
    the errors would reach 1e20 without meaning).  Pairs are taken while
    ||r_j|| > rmin ||b|| and scaled by 1/||r_j|| (e = A^-1 r is linear), so
    float32 holds them.  Returns two float32 tensors (n, nc, ny, nx)."""
    R, E = [], []
    for b in rhs_list:
        xs, _ = pcg(mm.apply_A, b, M, tol=tol_exact, maxiter=20000)
        bn = float(b.norm())

        def cb(it, x, r, p, Ap):
            nr = float(r.norm())
            if nr > rmin * bn:
                R.append((r / nr).to(torch.float32))
                E.append(((xs - x) / nr).to(torch.float32))
        pcg(mm.apply_A, b, M, tol=rmin, maxiter=per_rhs, callback=cb)
    return torch.stack(R), torch.stack(E)


def train_correction(net, M, R, E, steps=1500, batch=8, lr=1e-3, seed=0, log=None):
    """Fit the U-Net so that M r + s U(M r / s) ~ e on the pairs (R, E):
    relative squared error of the corrected estimate per pair, Adam.
    Returns the loss history."""
    g = torch.Generator(device=R.device).manual_seed(seed)
    n = R.shape[0]
    nc, ny, nx = R.shape[1:]
    # binned estimate and scale per pair
    MR = torch.stack([M(R[i].to(torch.float64)).to(torch.float32) for i in range(n)])
    X, S = features(MR)
    T = (E - MR) / S[:, None, None, None]          # target correction
    TN = (E / S[:, None, None, None]).pow(2).sum(dim=(1, 2, 3))
    hist = []
    for step in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g, device=R.device)
        x = net.to_input(X[idx])
        y, tape = net.forward(x)
        out = net.from_output(y, nc, ny, nx)
        d = out - T[idx]
        per = d.pow(2).sum(dim=(1, 2, 3)) / TN[idx]
        hist.append(float(per.mean()))
        gy = net.to_input(2 * d / (TN[idx][:, None, None, None] * batch))
        dw, db = net.backward(tape, gy)
        net.adam_step(dw, db, lr=lr)
        if log is not None and step % 100 == 0:
            log(step, hist[-1])
    return hist
