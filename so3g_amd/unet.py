"""Learned correction for the map-making preconditioner: a 4-level U-Net
(16/32/48/64 channels) on bf16 NHWC activations, run with the HIP layer
kernels of csrc/unet.hip (3x3 convolutions on the matrix cores, float32
accumulation), trained on the GPU with Adam.

Architecture (every conv 3x3, ReLU unless noted; '+' adds a skip):
    e0 = conv(conv(x, 16), 16)                  H x W
    e1 = conv(conv(pool(e0), 32), 32)           H/2
    e2 = conv(conv(pool(e1), 48), 48)           H/4
    b  = conv(conv(pool(e2), 64), 64)           H/8
    d2 = conv(conv(up(b), 48) + e2, 48)
    d1 = conv(conv(up(d2), 32) + e1, 32)
    d0 = conv(conv(up(d1), 16) + e0, 16)
    out = conv(d0, 16)  (no ReLU; the first 3 channels are T, Q, U)
pool / up change the resolution by `factor` = (rows, columns) per level,
(2, 2) by default; (1, 4) pools only along map rows, the scan direction.
The input planes are zero-padded to 16 channels and to H, W divisible by
factor^3.

No counterpart in the reference (which has no map-maker).
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from .engine import _stream

CH = (16, 32, 48, 64)
NIN = 16


def layer_specs(ch=CH):
    """(name, ci, co, relu) in forward order."""
    c0, c1, c2, c3 = ch
    return [('e0a', NIN, c0, 1), ('e0b', c0, c0, 1),
            ('e1a', c0, c1, 1), ('e1b', c1, c1, 1),
            ('e2a', c1, c2, 1), ('e2b', c2, c2, 1),
            ('e3a', c2, c3, 1), ('e3b', c3, c3, 1),
            ('d2a', c3, c2, 1), ('d2b', c2, c2, 1),
            ('d1a', c2, c1, 1), ('d1b', c1, c1, 1),
            ('d0a', c1, c0, 1), ('d0b', c0, c0, 1),
            ('out', c0, NIN, 0)]


def program():
    """The forward program as (op, out, inputs, layer): op in conv, pool,
    up, add.  Tensor 'x' is the input."""
    P = []
    def conv(o, i, l):
        P.append(('conv', o, (i,), l))
    conv('a0', 'x', 'e0a'); conv('e0', 'a0', 'e0b')
    P.append(('pool', 'p1', ('e0',), None))
    conv('a1', 'p1', 'e1a'); conv('e1', 'a1', 'e1b')
    P.append(('pool', 'p2', ('e1',), None))
    conv('a2', 'p2', 'e2a'); conv('e2', 'a2', 'e2b')
    P.append(('pool', 'p3', ('e2',), None))
    conv('a3', 'p3', 'e3a'); conv('b', 'a3', 'e3b')
    P.append(('up', 'u2', ('b',), None))
    conv('c2', 'u2', 'd2a'); P.append(('add', 's2', ('c2', 'e2'), None)); conv('d2', 's2', 'd2b')
    P.append(('up', 'u1', ('d2',), None))
    conv('c1', 'u1', 'd1a'); P.append(('add', 's1', ('c1', 'e1'), None)); conv('d1', 's1', 'd1b')
    P.append(('up', 'u0', ('d1',), None))
    conv('c0', 'u0', 'd0a'); P.append(('add', 's0', ('c0', 'e0'), None)); conv('d0', 's0', 'd0b')
    conv('y', 'd0', 'out')
    return P


class UNet:
    def __init__(self, ch=CH, seed=0, device=None, init_scale=1.0, factor=(2, 2)):
        self.dev = device or torch.device('cuda')
        self.factor = tuple(int(f) for f in factor)   # pooling (rows, columns) per level
        self.specs = {s[0]: s[1:] for s in layer_specs(ch)}
        self.prog = program()
        self.lib = N.load()
        g = torch.Generator().manual_seed(seed)
        self.w, self.b = {}, {}
        for name, (ci, co, relu) in self.specs.items():
            fan_in = 9 * ci
            std = np.sqrt(2. / fan_in) * (init_scale if name != 'out' else 0.1)
            w = torch.randn((9, co, ci), generator=g, dtype=torch.float32) * std
            self.w[name] = w.to(self.dev)
            self.b[name] = torch.zeros(co, dtype=torch.float32, device=self.dev)
        self._m = {k: (torch.zeros_like(v), torch.zeros_like(v)) for k, v in self._params()}
        self.step_count = 0
        self._sync_bf16()

    def _params(self):
        for k in self.specs:
            yield 'w:' + k, self.w[k]
            yield 'b:' + k, self.b[k]

    def _sync_bf16(self):
        self.wbf = {k: self.w[k].to(torch.bfloat16).contiguous() for k in self.specs}
        self.wflip = {}
        for k, (ci, co, _) in self.specs.items():
            t = torch.empty((9, ci, co), dtype=torch.bfloat16, device=self.dev)
            N.check(self.lib.so3g_nn_weight_flip(self.wbf[k].data_ptr(), ci, co, t.data_ptr(),
                                                 _stream()))
            self.wflip[k] = t

    # -- layers ------------------------------------------------------------
    def _conv(self, x, wt, bias, ci, co, relu):
        n, h, w, _ = x.shape
        y = torch.empty((n, h, w, co), dtype=torch.bfloat16, device=self.dev)
        N.check(self.lib.so3g_nn_conv3x3(x.data_ptr(), n, h, w, ci, wt.data_ptr(),
                                         None if bias is None else bias.data_ptr(), co,
                                         int(relu), y.data_ptr(), _stream()))
        return y

    def _pool(self, x, backward=False):
        n, h, w, c = x.shape
        fy, fx = self.factor
        shape = (n, fy * h, fx * w, c) if backward else (n, h // fy, w // fx, c)
        y = torch.empty(shape, dtype=torch.bfloat16, device=self.dev)
        hh, ww = (fy * h, fx * w) if backward else (h, w)
        N.check(self.lib.so3g_nn_pool(x.data_ptr(), n, hh, ww, c, fy, fx, y.data_ptr(),
                                      int(backward), _stream()))
        return y

    def _up(self, x, backward=False):
        n, h, w, c = x.shape
        fy, fx = self.factor
        shape = (n, h // fy, w // fx, c) if backward else (n, fy * h, fx * w, c)
        y = torch.empty(shape, dtype=torch.bfloat16, device=self.dev)
        hh, ww = (h // fy, w // fx) if backward else (h, w)
        N.check(self.lib.so3g_nn_up(x.data_ptr(), n, hh, ww, c, fy, fx, y.data_ptr(),
                                    int(backward), _stream()))
        return y

    def _add(self, a, b):
        y = torch.empty_like(a)
        N.check(self.lib.so3g_nn_add(a.data_ptr(), b.data_ptr(), a.numel(), y.data_ptr(),
                                     _stream()))
        return y

    # -- forward / backward --------------------------------------------------
    def forward(self, x):
        """x: bf16 (n, H, W, 16) with H, W divisible by factor^3.  Returns (y, tape)."""
        t = {'x': x}
        for op, o, ins, layer in self.prog:
            if op == 'conv':
                ci, co, relu = self.specs[layer]
                t[o] = self._conv(t[ins[0]], self.wbf[layer], self.b[layer], ci, co, relu)
            elif op == 'pool':
                t[o] = self._pool(t[ins[0]])
            elif op == 'up':
                t[o] = self._up(t[ins[0]])
            else:
                t[o] = self._add(t[ins[0]], t[ins[1]])
        return t['y'], t

    def backward(self, tape, gy):
        """Gradients (float32) of sum(gy * y) with respect to every weight
        and bias, as dicts keyed by layer name."""
        dw = {k: torch.zeros_like(self.w[k]) for k in self.specs}
        db = {k: torch.zeros_like(self.b[k]) for k in self.specs}
        g = {'y': gy}
        def acc(name, v):
            g[name] = v if name not in g else self._add(g[name], v)
        for op, o, ins, layer in reversed(self.prog):
            go = g.pop(o)
            if op == 'conv':
                ci, co, relu = self.specs[layer]
                x = tape[ins[0]]
                if relu:
                    gz = torch.empty_like(go)
                    N.check(self.lib.so3g_nn_relu_mask(go.data_ptr(), tape[o].data_ptr(),
                                                       go.numel(), gz.data_ptr(), _stream()))
                else:
                    gz = go
                n, h, w, _ = x.shape
                N.check(self.lib.so3g_nn_conv3x3_wgrad(x.data_ptr(), gz.data_ptr(), n, h, w, ci,
                                                       co, dw[layer].data_ptr(),
                                                       db[layer].data_ptr(), _stream()))
                if ins[0] != 'x':
                    acc(ins[0], self._conv(gz, self.wflip[layer], None, co, ci, 0))
            elif op == 'pool':
                acc(ins[0], self._pool(go, backward=True))
            elif op == 'up':
                acc(ins[0], self._up(go, backward=True))
            else:
                acc(ins[0], go)
                acc(ins[1], go)
        return dw, db

    def adam_step(self, dw, db, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for k in self.specs:
            for key, p, gr in (('w:' + k, self.w[k], dw[k]), ('b:' + k, self.b[k], db[k])):
                m, v = self._m[key]
                out = self.wbf[k] if key[0] == 'w' else None
                N.check(self.lib.so3g_nn_adam(p.data_ptr(), gr.data_ptr(), m.data_ptr(),
                                              v.data_ptr(), p.numel(), lr, b1, b2, eps,
                                              self.step_count,
                                              None if out is None else out.data_ptr(),
                                              _stream()))
        for k, (ci, co, _) in self.specs.items():
            N.check(self.lib.so3g_nn_weight_flip(self.wbf[k].data_ptr(), ci, co,
                                                 self.wflip[k].data_ptr(), _stream()))

    # -- planes <-> activations ---------------------------------------------
    def padded(self, h, w):
        my, mx = self.factor[0] ** 3, self.factor[1] ** 3
        return (h + my - 1) // my * my, (w + mx - 1) // mx * mx

    def to_input(self, planes):
        """float32 (n, cp <= 16, hy, wx) -> bf16 (n, H, W, 16), zero-padded."""
        n, cp, hy, wx = planes.shape
        H, W = self.padded(hy, wx)
        x = torch.empty((n, H, W, NIN), dtype=torch.bfloat16, device=self.dev)
        p = planes.to(torch.float32).contiguous()
        N.check(self.lib.so3g_nn_from_planes(p.data_ptr(), n, cp, hy, wx, H, W, NIN,
                                             x.data_ptr(), _stream()))
        return x

    def from_output(self, y, cp, hy, wx):
        n, H, W, c = y.shape
        out = torch.empty((n, cp, hy, wx), dtype=torch.float32, device=self.dev)
        N.check(self.lib.so3g_nn_to_planes(y.data_ptr(), n, H, W, c, cp, hy, wx,
                                           out.data_ptr(), _stream()))
        return out

    def __call__(self, planes):
        """float32 (n, 3, hy, wx) -> float32 (n, 3, hy, wx)."""
        n, cp, hy, wx = planes.shape
        y, _ = self.forward(self.to_input(planes))
        return self.from_output(y, cp, hy, wx)

    def state(self):
        return {k: (self.w[k].cpu().numpy(), self.b[k].cpu().numpy()) for k in self.specs}
