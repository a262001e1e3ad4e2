"""Boresight quaternions computed on the GPU: the per-sample part of
CelestialSightLine.for_horizon (python/proj/coords.py:190-209) and
naive_az_el (:99-125), optionally premultiplied by a Projectionist's
celestial -> native rotation, so that an engine's q_native input is made
on the device (no host quaternion arithmetic for large or many
observations).  Returns an (n, 4) float64 torch tensor on the device.
"""
import numpy as np
import torch

from . import _native as N
from .engine import _device, _stream
from .proj import quat
from .proj.coords import CelestialSightLine, DEG, ERA_EPOCH, ERA_POLY


def _dev(x, n, dev):
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x, np.float64))
    t = t.to(dev, torch.float64).contiguous().reshape(-1)
    if t.numel() != n:
        raise RuntimeError(f'expected {n} samples, got {t.numel()}')
    return t


def _pre(q_pre):
    if q_pre is None:
        return None
    return np.ascontiguousarray(np.asarray(q_pre, np.float64).reshape(4))


def horizon(az, el, roll=None, q_pre=None):
    """q_pre * euler(2, -az) euler(1, pi/2 - el) euler(2, roll)."""
    dev = _device()
    n = int(np.size(az)) if not isinstance(az, torch.Tensor) else az.numel()
    a, e = _dev(az, n, dev), _dev(el, n, dev)
    r, r0 = None, 0.
    if roll is not None and np.ndim(roll if not isinstance(roll, torch.Tensor) else roll.cpu()):
        r = _dev(roll, n, dev)
    elif roll is not None:
        r0 = float(roll)
    out = torch.empty((n, 4), dtype=torch.float64, device=dev)
    pre = _pre(q_pre)
    N.check(N.load().so3g_boresight(0, None, a.data_ptr(), e.data_ptr(),
                                    None if r is None else r.data_ptr(), r0, n, None,
                                    None if pre is None else pre.ctypes.data, out.data_ptr(),
                                    _stream()))
    return out


def naive_az_el(t, az, el, roll=0., site=None, q_pre=None):
    """q_pre * euler(2, lst) euler(1, pi/2 - lat) euler(2, pi) euler(2, -az)
    euler(1, pi/2 - el) euler(2, roll), lst from the linear Earth rotation
    angle of naive_az_el."""
    dev = _device()
    site = CelestialSightLine.decode_site(site)
    n = int(np.size(az)) if not isinstance(az, torch.Tensor) else az.numel()
    tt, a, e = _dev(t, n, dev), _dev(az, n, dev), _dev(el, n, dev)
    r, r0 = None, 0.
    if np.ndim(roll if not isinstance(roll, torch.Tensor) else roll.cpu()):
        r = _dev(roll, n, dev)
    else:
        r0 = float(roll)
    s = np.array([ERA_POLY[0], ERA_POLY[1], ERA_EPOCH, site.lon * DEG, site.lat * DEG],
                 np.float64)
    out = torch.empty((n, 4), dtype=torch.float64, device=dev)
    pre = _pre(q_pre)
    N.check(N.load().so3g_boresight(1, tt.data_ptr(), a.data_ptr(), e.data_ptr(),
                                    None if r is None else r.data_ptr(), r0, n, s.ctypes.data,
                                    None if pre is None else pre.ctypes.data, out.data_ptr(),
                                    _stream()))
    return out
