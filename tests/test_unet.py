"""U-Net layers of the learned preconditioner: HIP kernels against the
numpy restatement (oracle/unet_ref.py), and the restatement's own algebra."""
import numpy as np
import pytest

from oracle import unet_ref as R

torch = pytest.importorskip('torch')


@pytest.fixture
def exact():
    R.set_rounding(False)
    yield
    R.set_rounding(True)


def test_bf16_emulation_matches_torch():
    x = np.random.default_rng(0).standard_normal(10000) * 10.0 ** np.arange(-3, 3).repeat(
        10000 // 6 + 1)[:10000]
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(R.bf16(x), ref)


def test_conv_adjoint_and_weight_gradient(exact):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 7, 3))
    w = rng.standard_normal((9, 4, 3))
    gz = rng.standard_normal((2, 5, 7, 4))
    y = R.conv(x, w, None, False)
    dw, db, gx = R.conv_grads(x, gz, w)
    # <conv(x), gz> = <x, gx>  (the data gradient is the adjoint)
    assert abs(np.sum(y * gz) - np.sum(x * gx)) < 1e-10 * abs(np.sum(y * gz))
    # d<conv(x), gz>/dw by finite differences (exact: linear in w)
    for (t, o, c) in [(0, 1, 2), (4, 0, 0), (8, 3, 1)]:
        e = np.zeros_like(w)
        e[t, o, c] = 1.
        fd = np.sum(R.conv(x, w + e, None, False) * gz) - np.sum(y * gz)
        assert abs(fd - dw[t, o, c]) < 1e-9 * max(1, abs(fd))
    # bias gradient, pool / up adjoints
    assert np.allclose(db, gz.sum(axis=(0, 1, 2)))
    a = rng.standard_normal((2, 6, 8, 3))
    b = rng.standard_normal((2, 3, 4, 3))
    assert abs(np.sum(R.pool(a) * b) - np.sum(a * R.pool_bwd(b))) < 1e-10
    assert abs(np.sum(R.up(b) * a) - np.sum(b * R.up_bwd(a))) < 1e-10
    # a tap of the 3x3 kernel sees the neighbour it names: (dy, dx) = (-1, 0) is the row above
    w1 = np.zeros((9, 1, 1))
    w1[1, 0, 0] = 1.
    img = np.arange(20.).reshape(1, 4, 5, 1)
    out = R.conv(img, w1, None, False)
    assert out[0, 2, 3, 0] == img[0, 1, 3, 0] and out[0, 0, 3, 0] == 0.


# ---------------------------------------------------------------------------
def dev():
    return torch.device('cuda')


def bf(x):
    return torch.from_numpy(np.asarray(x, np.float32)).to(dev()).to(torch.bfloat16).contiguous()


def np64(t):
    return t.to(torch.float32).cpu().numpy().astype(np.float64)


def lib():
    from so3g_amd import _native as N
    return N


def ulp_close(g, r):
    return np.all(np.abs(g - r) <= 2. ** -7 * np.abs(r) + 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize('ci,co,relu', [(16, 16, 1), (32, 48, 0), (16, 64, 1), (64, 32, 1)])
def test_conv_forward_and_grads(gpu, ci, co, relu):
    N = lib()
    from so3g_amd.engine import _stream
    rng = np.random.default_rng(ci + co)
    n, h, w = 2, 6, 70
    x = R.bf16(rng.standard_normal((n, h, w, ci)))
    wt = R.bf16(rng.standard_normal((9, co, ci)) / np.sqrt(9 * ci))
    b = rng.standard_normal(co).astype(np.float32)
    xd, wd = bf(x), bf(wt)
    bd = torch.from_numpy(b).to(dev())
    out = torch.empty((n, h, w, co), dtype=torch.bfloat16, device=dev())
    N.check(N.load().so3g_nn_conv3x3(xd.data_ptr(), n, h, w, ci, wd.data_ptr(), bd.data_ptr(),
                                     co, relu, out.data_ptr(), _stream()))
    ref = R.conv(x, wt, b.astype(np.float64), relu)
    assert ulp_close(np64(out), ref)
    # weight and bias gradients of sum(gz * conv(x))
    gz = R.bf16(rng.standard_normal((n, h, w, co)))
    gzd = bf(gz)
    dw = torch.zeros((9, co, ci), dtype=torch.float32, device=dev())
    db = torch.zeros(co, dtype=torch.float32, device=dev())
    N.check(N.load().so3g_nn_conv3x3_wgrad(xd.data_ptr(), gzd.data_ptr(), n, h, w, ci, co,
                                           dw.data_ptr(), db.data_ptr(), _stream()))
    rdw, rdb, rgx = R.conv_grads(x, gz, wt)
    assert np.abs(dw.cpu().numpy() - rdw).max() < 1e-5 * np.abs(rdw).max()
    assert np.abs(db.cpu().numpy() - rdb).max() < 1e-5 * np.abs(rdb).max()
    # data gradient: conv with the flipped weights
    wf = torch.empty((9, ci, co), dtype=torch.bfloat16, device=dev())
    N.check(N.load().so3g_nn_weight_flip(wd.data_ptr(), ci, co, wf.data_ptr(), _stream()))
    gx = torch.empty((n, h, w, ci), dtype=torch.bfloat16, device=dev())
    N.check(N.load().so3g_nn_conv3x3(gzd.data_ptr(), n, h, w, co, wf.data_ptr(), None, ci, 0,
                                     gx.data_ptr(), _stream()))
    assert ulp_close(np64(gx), rgx)


@pytest.mark.gpu
def test_pool_up_relu_add(gpu):
    from so3g_amd.unet import UNet
    net = UNet(device=dev())
    rng = np.random.default_rng(3)
    a = R.bf16(rng.standard_normal((2, 8, 12, 16)))
    b = R.bf16(rng.standard_normal((2, 8, 12, 16)))
    ad, bd = bf(a), bf(b)
    np.testing.assert_array_equal(np64(net._pool(ad)), R.pool(a))
    np.testing.assert_array_equal(np64(net._pool(net._pool(ad), backward=True)),
                                  R.pool_bwd(R.pool(a)))
    np.testing.assert_array_equal(np64(net._up(ad)), R.up(a))
    np.testing.assert_array_equal(np64(net._up(net._up(ad), backward=True)), R.up_bwd(R.up(a)))
    np.testing.assert_array_equal(np64(net._add(ad, bd)), R.bf16(a + b))
    net.factor = (1, 4)
    f = (1, 4)
    np.testing.assert_array_equal(np64(net._pool(ad)), R.pool(a, f))
    np.testing.assert_array_equal(np64(net._pool(net._pool(ad), backward=True)),
                                  R.pool_bwd(R.pool(a, f), f))
    np.testing.assert_array_equal(np64(net._up(ad)), R.up(a, f))
    np.testing.assert_array_equal(np64(net._up(net._up(ad), backward=True)),
                                  R.up_bwd(R.up(a, f), f))


def rel_err(g, r):
    return np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize('factor,h,w', [((2, 2), 16, 40), ((1, 4), 6, 128)])
def test_unet_gradients_against_cpu(gpu, factor, h, w):
    """Full forward/backward of the 4-level network against the numpy
    restatement with the same bf16 rounding points."""
    from so3g_amd.unet import UNet, layer_specs, program
    net = UNet(device=dev(), seed=2, factor=factor)
    rng = np.random.default_rng(5)
    n = 2
    planes = rng.standard_normal((n, 3, h - 2, w - 3)).astype(np.float32)
    x = net.to_input(torch.from_numpy(planes).to(dev()))
    y, tape = net.forward(x)
    weights = {k: (net.w[k].cpu().numpy().astype(np.float64), net.b[k].cpu().numpy())
               for k in net.specs}
    xr = np.zeros((n,) + net.padded(h - 2, w - 3) + (16,))
    xr[:, :h - 2, :w - 3, :3] = planes.transpose(0, 2, 3, 1)
    specs = {s[0]: s[1:] for s in layer_specs()}
    t = R.forward(program(), specs, weights, xr, factor)
    assert rel_err(np64(y), t['y']) < 2e-2
    gy = R.bf16(rng.standard_normal(t['y'].shape))
    dw, db = net.backward(tape, bf(gy))
    rdw, rdb = R.backward(program(), specs, weights, t, gy, factor)
    for k in specs:
        e = rel_err(dw[k].cpu().numpy(), rdw[k])
        assert e < 3e-2, (k, e)
        e = rel_err(db[k].cpu().numpy(), rdb[k])
        assert e < 3e-2, (k, e)


@pytest.mark.gpu
def test_adam_step(gpu):
    from so3g_amd.engine import _stream
    N = lib()
    rng = np.random.default_rng(0)
    w = rng.standard_normal(1000).astype(np.float32)
    gs = [rng.standard_normal(1000).astype(np.float32) for _ in range(3)]
    wd = torch.from_numpy(w.copy()).to(dev())
    m = torch.zeros_like(wd)
    v = torch.zeros_like(wd)
    wb = torch.empty(1000, dtype=torch.bfloat16, device=dev())
    mr, vr, wr = np.zeros(1000), np.zeros(1000), w.astype(np.float64)
    for k, g in enumerate(gs, 1):
        gd = torch.from_numpy(g).to(dev())
        N.check(N.load().so3g_nn_adam(wd.data_ptr(), gd.data_ptr(), m.data_ptr(), v.data_ptr(),
                                      1000, 1e-2, 0.9, 0.999, 1e-8, k, wb.data_ptr(),
                                      _stream()))
        mr = 0.9 * mr + 0.1 * g
        vr = 0.999 * vr + 0.001 * g.astype(np.float64) ** 2
        wr = wr - 1e-2 * (mr / (1 - 0.9 ** k)) / (np.sqrt(vr / (1 - 0.999 ** k)) + 1e-8)
    np.testing.assert_allclose(wd.cpu().numpy(), wr, rtol=0, atol=1e-5)
    np.testing.assert_array_equal(np64(wb), R.bf16(wd.cpu().numpy()))


@pytest.mark.gpu
def test_training_reduces_loss(gpu):
    """A few Adam steps on a smoothing target lower the relative error."""
    from so3g_amd.unet import UNet
    net = UNet(device=dev(), seed=1)
    g = torch.Generator(device=dev()).manual_seed(0)
    def batch():
        x = torch.randn((4, 3, 24, 64), device=dev(), generator=g)
        t = 0.5 * (x + torch.roll(x, 1, dims=3))
        return x, t
    losses = []
    for it in range(60):
        x, t = batch()
        y, tape = net.forward(net.to_input(x))
        out = net.from_output(y, 3, 24, 64)
        d = out - t
        losses.append(float((d ** 2).sum() / (t ** 2).sum()))
        gy = net.to_input(2 * d / (t ** 2).sum())
        dw, db = net.backward(tape, gy)
        net.adam_step(dw, db, lr=3e-3)
    assert np.mean(losses[-10:]) < 0.5 * np.mean(losses[:5])
