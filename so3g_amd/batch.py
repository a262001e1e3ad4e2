"""Batched projection: many observations, one map, one launch.

No counterpart in the reference, where a caller loops over observations and
calls Projectionist.to_map / from_map (python/proj/wcs.py:415-476) once per
observation.  ObservationBatch gathers the observations' device arrays once
and projects all of them with one persistent kernel launch, optionally
captured in a HIP graph and replayed, with a launch-shape autotuner.

    batch = ObservationBatch(engine, [dict(bore=q_native, dets=q_dets,
                                           resp=resp, signal=sig), ...])
    batch.tune('to_map', scratch_map)
    batch.to_map(map)            # map += P^T signal, all observations
    batch.from_map(map)          # signal = P map, all observations

`engine` is a ProjEng_* instance (not Precomp); its pixelization is the
map's.  bore is the native-frame boresight of each observation (what
Projectionist passes its engine: q_celestial_to_native * Q).
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from .engine import _Call, _device, _stream

MODES = {'eager': N.BATCH_EAGER, 'graph': N.BATCH_GRAPH, 'per_obs': N.BATCH_PER_OBS}
OPS = {'from_map': 0, 'to_map': 1}


def _dev_tensor(x, dtype, dev, shape=None):
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
    if t.dtype != dtype:
        raise ValueError(f'expected {dtype}, got {t.dtype}')
    t = t.to(dev).contiguous()
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise RuntimeError(f'expected shape {tuple(shape)}, got {tuple(t.shape)}')
    return t


class ObservationBatch:
    def __init__(self, engine, observations):
        self.engine = engine
        dev = _device()
        self._keep = []
        descs = (N.Obs * max(1, len(observations)))()
        self.n_obs = len(observations)
        self.signals = []
        for i, o in enumerate(observations):
            bore = _dev_tensor(o['bore'], torch.float64, dev)
            dets = _dev_tensor(o['dets'], torch.float64, dev)
            if bore.ndim != 2 or bore.shape[1] != 4 or dets.ndim != 2 or dets.shape[1] != 4:
                raise RuntimeError(f'observation {i}: quaternion arrays must be (n, 4)')
            n_t, n_det = bore.shape[0], dets.shape[0]
            resp = _dev_tensor(o['resp'], torch.float32, dev, (n_det, 2))
            sig = o.get('signal')
            sig = (torch.zeros((n_det, n_t), dtype=torch.float32, device=dev) if sig is None
                   else _dev_tensor(sig, torch.float32, dev, (n_det, n_t)))
            dw = o.get('det_weights')
            dw = None if dw is None else _dev_tensor(dw, torch.float32, dev, (n_det,))
            self._keep += [bore, dets, resp, sig, dw]
            self.signals.append(sig)
            descs[i] = N.Obs(bore.data_ptr(), n_t, dets.data_ptr(), n_det, resp.data_ptr(),
                             sig.data_ptr(), sig.data_ptr(),
                             None if dw is None else dw.data_ptr())
        h = ctypes.c_void_p()
        N.check(N.load().so3g_batch_create(engine._g, descs, self.n_obs, ctypes.byref(h)))
        self._h = h
        self._staged = None

    # -- map staging (cached for torch maps so that graphs are reused) -----
    def _map_args(self, map, writeback):
        eng = self.engine
        if not eng._tiled and isinstance(map, torch.Tensor):
            call = _Call()
            m, _, _ = eng._map(call, map, [eng.comp_count], writeback=writeback)
            return call, m, None, None
        if eng._tiled and isinstance(map, (list, tuple)) and \
                all(t is None or isinstance(t, torch.Tensor) for t in map):
            key = tuple(None if t is None else t.data_ptr() for t in map)
            if self._staged is None or self._staged[0] != key:
                call = _Call()
                _, tp, dp = eng._map(call, map, [eng.comp_count], writeback=False)
                self._staged = (key, call, tp, dp)
            return None, None, self._staged[2], self._staged[3]
        call = _Call()
        m, tp, dp = eng._map(call, map, [eng.comp_count], writeback=writeback)
        return call, m, tp, dp

    def to_map(self, map, mode='graph'):
        """map += sum over observations of P^T signal (so3g's to_map
        arithmetic: float32 products, float64 sums)."""
        call, m, tp, dp = self._map_args(map, True)
        N.check(N.load().so3g_batch_to_map(self._h, m, tp, dp, MODES[mode], _stream()))
        if call is not None:
            call.finish()
        return map

    def from_map(self, map, mode='graph', accumulate=False):
        """signal (each observation's, in place) = P map (+ signal if
        accumulate).  Returns the list of signal tensors."""
        call, m, tp, dp = self._map_args(map, False)
        N.check(N.load().so3g_batch_from_map(self._h, m, tp, dp, int(bool(accumulate)),
                                             MODES[mode], _stream()))
        return self.signals

    def set_shape(self, op, dpb, blocks_per_cu):
        N.check(N.load().so3g_batch_set_shape(self._h, OPS[op], int(dpb), int(blocks_per_cu)))

    def set_table(self, slots):
        """to_map through an LDS table of `slots` slots per work item (512
        or 1024), global atomics (0), or (flat maps, nearest neighbour)
        segments of -slots columns of each detector group's home map row in
        LDS."""
        N.check(N.load().so3g_batch_set_table(self._h, int(slots)))

    def tune(self, op, map, reps=1):
        """Time every launch shape on this batch and keep the fastest for op.
        Tuning to_map adds into `map` (pass a scratch map).  Returns a list
        of (dets per block, blocks per CU (0: one per item), LDS table slots
        (0: global atomics; -w: row segments of w columns), ms)."""
        call, m, tp, dp = self._map_args(map, False)
        res = (ctypes.c_double * 400)()
        n = ctypes.c_int()
        N.check(N.load().so3g_batch_tune(self._h, OPS[op], m, tp, dp, int(reps), res, 100,
                                         ctypes.byref(n), _stream()))
        return [(int(res[4 * i]), int(res[4 * i + 1]), int(res[4 * i + 2]), float(res[4 * i + 3]))
                for i in range(min(n.value, 100))]

    def info(self):
        a = (ctypes.c_int64 * 8)()
        N.check(N.load().so3g_batch_info(self._h, a))
        return {'n_obs': a[0], 'det_samples': a[1], 'from_map_shape': (a[2], a[3]),
                'to_map_shape': (a[4], a[5]), 'n_cu': a[6], 'to_map_slots': a[7]}

    def close(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            N.load().so3g_batch_destroy(h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
