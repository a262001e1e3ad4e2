"""ML cut handling (process_cuts / translate_cuts, src/array_ops.cxx:130-400):
the C++ restatement pinned by closed forms, then the HIP kernels against it
bit for bit."""
import numpy as np
import pytest

import oracle
from so3g_amd.proj.ranges import Ranges

torch = pytest.importorskip('torch')


def random_cuts(ndet, nsamp, seed, max_len=60):
    rng = np.random.default_rng(seed)
    rm = []
    for d in range(ndet):
        r = Ranges(nsamp)
        for _ in range(rng.integers(0, 6)):
            a = int(rng.integers(0, nsamp - 1))
            r.add_interval(a, min(nsamp, a + int(rng.integers(1, max_len))))
        rm.append(r)
    return rm


def test_npoly_table():
    # array_ops.cxx:233-240 with resolution 10, nmax 8
    gaps = [1, 2, 3, 4, 6, 7, 20, 21, 45, 200]
    want = [1, 2, 2, 3, 3, 4, 4, 7, 8, 8]
    for g, w in zip(gaps, want):
        segs = np.array([[0, 0, g]], np.int32)
        assert oracle.cuts_measure(segs, 'poly', 10, 8) == w, g
    assert oracle.cuts_measure(np.array([[0, 0, 5]], np.int32), 'poly', 10, 0) == 1
    assert oracle.cuts_measure(np.array([[0, 0, 5], [1, 3, 9]], np.int32), 'full') == 11


def test_full_model_is_a_gather():
    rm = random_cuts(4, 300, 1)
    segs = oracle.cut_segments(rm)
    tod = np.random.default_rng(2).standard_normal((4, 300))
    n = oracle.cuts_measure(segs, 'full')
    vals = np.zeros(n)
    oracle.process_cuts(segs, 'extract', 'full', tod, vals)
    mask = np.stack([r.mask() for r in rm])
    np.testing.assert_array_equal(vals, tod[mask])
    tod2 = np.zeros_like(tod)
    oracle.process_cuts(segs, 'insert', 'full', tod2, vals)
    np.testing.assert_array_equal(tod2[mask], tod[mask])
    assert np.all(tod2[~mask] == 0)
    oracle.process_cuts(segs, 'clear', 'full', tod, vals)
    assert np.all(tod[mask] == 0)


def test_poly_model_is_legendre_projection():
    """extract = sum_s t_s P_p(x_s), insert = sum_p v_p P_p(x_s), x in
    [-1, 1] across the range (numpy's Legendre basis), and they are
    adjoint."""
    from numpy.polynomial import legendre
    rm = random_cuts(3, 400, 3, max_len=120)
    segs = oracle.cut_segments(rm)
    res, nmax = 10, 9
    n = oracle.cuts_measure(segs, 'poly', res, nmax)
    rng = np.random.default_rng(4)
    tod = rng.standard_normal((3, 400))
    vals = np.zeros(n)
    oracle.process_cuts(segs, 'extract', 'poly', tod, vals, res, nmax)
    off = 0
    for d, a, b in segs:
        g = b - a
        np_ = oracle.cuts_measure(np.array([[0, a, b]], np.int32), 'poly', res, nmax)
        if np_ <= 1:
            np.testing.assert_allclose(vals[off], tod[d, a:b].sum(), rtol=1e-12)
        else:
            x = -1 + 2 * np.arange(g) / (g - 1)
            V = legendre.legvander(x, np_ - 1)
            np.testing.assert_allclose(vals[off:off + np_], V.T @ tod[d, a:b], rtol=1e-10,
                                       atol=1e-12)
        off += np_
    v = rng.standard_normal(n)
    t2 = np.zeros_like(tod)
    oracle.process_cuts(segs, 'insert', 'poly', t2, v, res, nmax)
    assert abs(np.sum(t2 * tod) - np.sum(v * vals)) < 1e-10 * abs(np.sum(v * vals))


def test_translate_closed_forms():
    # full: downsample 2x averages pairs, upsample 2x repeats
    i = np.array([[0, 10, 20]], np.int32)
    o = np.array([[0, 5, 10]], np.int32)
    iv = np.arange(10, dtype=np.float64)
    ov = np.zeros(5)
    oracle.translate_cuts(i, o, 'full', 100, 50, iv, ov)
    np.testing.assert_array_equal(ov, [0.5, 2.5, 4.5, 6.5, 8.5])
    back = np.zeros(10)
    oracle.translate_cuts(o, i, 'full', 50, 100, ov, back)
    np.testing.assert_array_equal(back, np.repeat(ov, 2))
    # poly: copy, truncate, zero-pad
    i = np.array([[0, 0, 40]], np.int32)      # 5 + 40 / 10 = 9 -> nmax 6
    o = np.array([[0, 0, 5]], np.int32)       # 3
    iv = np.arange(1., 7.)
    ov = np.full(3, 9.)
    oracle.translate_cuts(i, o, 'poly', 100, 100, iv, ov, 10, 6)
    np.testing.assert_array_equal(ov, [1., 2., 3.])
    ov2 = np.full(6, 9.)
    oracle.translate_cuts(o, i, 'poly', 100, 100, ov, ov2, 10, 6)
    np.testing.assert_array_equal(ov2, [1., 2., 3., 0., 0., 0.])


def test_argument_errors_cpu():
    import so3g_amd
    rm = random_cuts(2, 50, 0)
    with pytest.raises(ValueError):
        so3g_amd.process_cuts(rm, 'measure', 'spline', {}, None, None)
    assert so3g_amd.process_cuts(rm, 'measure', 'full', {}, None, None) == \
        oracle.cuts_measure(oracle.cut_segments(rm), 'full')


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('model', ['full', 'poly'])
def test_process_cuts_gpu(gpu, dtype, model):
    import so3g_amd
    ndet, nsamp = 20, 3000
    rm = random_cuts(ndet, nsamp, 7, max_len=300)
    segs = oracle.cut_segments(rm)
    params = {'resolution': 25, 'nmax': 7}
    res, nmax = (25, 7) if model == 'poly' else (0, 0)
    n = so3g_amd.process_cuts(rm, 'measure', model, params, None, None)
    assert n == oracle.cuts_measure(segs, model, res, nmax)
    rng = np.random.default_rng(8)
    tod = rng.standard_normal((ndet, nsamp)).astype(dtype)
    # extract (poly: one-coefficient ranges add into what vals holds)
    v0 = rng.standard_normal(n).astype(dtype)
    g, r = v0.copy(), v0.copy()
    so3g_amd.process_cuts(rm, 'extract', model, params, tod, g)
    oracle.process_cuts(segs, 'extract', model, tod, r, res, nmax)
    np.testing.assert_array_equal(g, r)
    # insert
    tg, tr = tod.copy(), tod.copy()
    so3g_amd.process_cuts(rm, 'insert', model, params, tg, v0)
    oracle.process_cuts(segs, 'insert', model, tr, v0.copy(), res, nmax)
    np.testing.assert_array_equal(tg, tr)
    # clear, torch in place
    tt = torch.from_numpy(tod.copy()).cuda()
    so3g_amd.process_cuts(rm, 'clear', model, params, tt, torch.from_numpy(v0).cuda())
    tc = tod.copy()
    oracle.process_cuts(segs, 'clear', model, tc, v0.copy(), res, nmax)
    np.testing.assert_array_equal(tt.cpu().numpy(), tc)
    with pytest.raises(ValueError):
        so3g_amd.process_cuts(rm, 'scatter', model, params, tod, g)
    with pytest.raises(TypeError):
        so3g_amd.process_cuts(rm, 'insert', model, params, tod.astype(np.int32), g)


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_translate_cuts_gpu(gpu, dtype):
    import so3g_amd
    ndet, insamp, onsamp = 6, 2000, 500
    rng = np.random.default_rng(9)
    ri, ro = [], []
    for d in range(ndet):
        a, b = Ranges(insamp), Ranges(onsamp)
        for k in range(3):
            lo = int(rng.integers(0, 150)) * 4 + 600 * k
            hi = lo + 4 * int(rng.integers(1, 30))
            a.add_interval(lo, hi)
            b.add_interval(lo // 4, hi // 4)
        ri.append(a)
        ro.append(b)
    for model, params, rs in [('full', {}, (0, 0)), ('poly', {'resolution': 5, 'nmax': 6}, (5, 6))]:
        isegs, osegs = oracle.cut_segments(ri), oracle.cut_segments(ro)
        ni = oracle.cuts_measure(isegs, model, *rs)
        no = oracle.cuts_measure(osegs, model, *rs)
        iv = rng.standard_normal(ni).astype(dtype)
        for (src, dst, sv, dv, n_in, n_out, ns) in [
                (ri, ro, isegs, osegs, ni, no, (insamp, onsamp)),
                (ro, ri, osegs, isegs, no, ni, (onsamp, insamp))]:
            x = rng.standard_normal(n_in).astype(dtype)
            g = np.zeros(n_out, dtype)
            r = np.zeros(n_out, dtype)
            so3g_amd.translate_cuts(src, dst, model, params, x, g)
            oracle.translate_cuts(sv, dv, model, ns[0], ns[1], x, r, *rs)
            np.testing.assert_array_equal(g, r)
