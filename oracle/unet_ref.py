"""CPU restatement of the U-Net layers (so3g_amd/unet.py, csrc/unet.hip) in
numpy, for gradient checks.  Test infrastructure only.

bf16 storage is emulated: every tensor the GPU stores in bf16 (activations,
gradients, weights) is rounded to bf16 (nearest even) at the same point;
sums are float64 (the GPU accumulates in float32).  The data gradient of a
convolution is computed here as the direct adjoint (scatter of each output
gradient through the taps), not through flipped weights as on the GPU.
"""
import numpy as np


_ROUND = [True]


def set_rounding(enabled):
    """Turn the bf16 emulation off for exact float64 checks of the algebra."""
    _ROUND[0] = bool(enabled)


def bf16(x):
    """Round float values to the nearest bf16 (ties to even), as float64."""
    if not _ROUND[0]:
        return np.asarray(x, np.float64)
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def _shift(x, dy, dx):
    """s[n, y, x] = x[n, y + dy, x + dx] (zero outside)."""
    n, h, w, c = x.shape
    s = np.zeros_like(x)
    ys, ye = max(0, -dy), min(h, h - dy)
    xs, xe = max(0, -dx), min(w, w - dx)
    s[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return s


def conv(x, w, b, relu):
    """x (n,h,w,ci), w (9, co, ci), b (co) -> bf16(act(b + sum))."""
    out = 0.
    for t, (dy, dx) in enumerate(TAPS):
        out = out + np.einsum('nhwc,oc->nhwo', _shift(x, dy, dx), w[t])
    if b is not None:
        out = out + b
    if relu:
        out = np.maximum(out, 0.)
    return bf16(out)


def conv_grads(x, gz, w):
    """(dW (9,co,ci), db (co), dx (n,h,w,ci) rounded to bf16)."""
    dw = np.stack([np.einsum('nhwo,nhwc->oc', gz, _shift(x, dy, dx)) for dy, dx in TAPS])
    db = gz.sum(axis=(0, 1, 2))
    gx = 0.
    for t, (dy, dx) in enumerate(TAPS):
        gx = gx + _shift(np.einsum('nhwo,oc->nhwc', gz, w[t]), -dy, -dx)
    return dw, db, bf16(gx)


def pool(x, f=(2, 2)):
    n, h, w, c = x.shape
    fy, fx = f
    return bf16(x.reshape(n, h // fy, fy, w // fx, fx, c).sum(axis=(2, 4)) / (fy * fx))


def pool_bwd(g, f=(2, 2)):
    fy, fx = f
    return bf16(np.repeat(np.repeat(g, fy, axis=1), fx, axis=2) / (fy * fx))


def up(x, f=(2, 2)):
    return np.repeat(np.repeat(x, f[0], axis=1), f[1], axis=2)


def up_bwd(g, f=(2, 2)):
    n, h, w, c = g.shape
    fy, fx = f
    return bf16(g.reshape(n, h // fy, fy, w // fx, fx, c).sum(axis=(2, 4)))


def forward(prog, specs, weights, x, f=(2, 2)):
    """weights: name -> (w (9,co,ci) float, b (co)).  Returns the tape."""
    t = {'x': bf16(x)}
    for op, o, ins, layer in prog:
        if op == 'conv':
            w, b = weights[layer]
            t[o] = conv(t[ins[0]], bf16(w), np.asarray(b, np.float64), specs[layer][2])
        elif op == 'pool':
            t[o] = pool(t[ins[0]], f)
        elif op == 'up':
            t[o] = up(t[ins[0]], f)
        else:
            t[o] = bf16(t[ins[0]] + t[ins[1]])
    return t


def backward(prog, specs, weights, tape, gy, f=(2, 2)):
    dw, db = {}, {}
    g = {'y': bf16(gy)}
    def acc(name, v):
        g[name] = v if name not in g else bf16(g[name] + v)
    for op, o, ins, layer in reversed(prog):
        go = g.pop(o)
        if op == 'conv':
            relu = specs[layer][2]
            gz = go * (tape[o] > 0) if relu else go
            w, _ = weights[layer]
            dw[layer], db[layer], gx = conv_grads(tape[ins[0]], gz, bf16(w))
            if ins[0] != 'x':
                acc(ins[0], gx)
        elif op == 'pool':
            acc(ins[0], pool_bwd(go, f))
        elif op == 'up':
            acc(ins[0], up_bwd(go, f))
        else:
            acc(ins[0], go)
            acc(ins[1], go)
    return dw, db
