"""ProjEng_* classes: the host-side mirror of so3g's boost-python
ProjectionEngine<C,P,S> and ProjEng_Precomp classes (src/Projection.cxx:
2258-2358), dispatching to the HIP kernels of libso3g_hip.so.

Each class takes the same constructor arguments (the pixelizor argument
tuple of wcs.py:681-695 / 837-854) and exposes the same methods with the
same argument meaning, return values and exception classes:

  coords, pixels, pointing_matrix, from_map, to_map, to_weight_map, zeros,
  pixel_ranges, tile_hits, tile_ranges; properties index_count, comp_count.

Arrays may be numpy arrays (staged to the GPU and written back in place,
as the reference's in-place semantics require) or torch tensors on the
ROCm device (used in place, zero-copy).  Signal-like outputs are returned
as lists of per-detector rows for numpy inputs, as the reference does;
when the caller passes torch tensors, the tensor itself is returned.
"""
import ctypes

import numpy as np

from . import _native as N
from .proj.ranges import Ranges, RangesMatrix

try:
    import torch
except ImportError:  # pragma: no cover - torch is a hard requirement on the GPU
    torch = None

__all__ = ['ProjEngBase', 'ProjEngPrecompBase', 'ENGINE_NAMES', 'make_engine_classes']

_NP2T = {np.dtype('float64'): 'float64', np.dtype('float32'): 'float32',
         np.dtype('int32'): 'int32'}


def _torch_dtype(name):
    return {'float64': torch.float64, 'float32': torch.float32, 'int32': torch.int32,
            'int64': torch.int64}[name]


def _is_torch(x):
    return torch is not None and isinstance(x, torch.Tensor)


def _device():
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError('so3g_amd projection needs a ROCm GPU (torch.cuda is unavailable); '
                           'there is no CPU fallback.')
    N.init()
    return torch.device('cuda', torch.cuda.current_device())


def _stream():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    return None if v in (0, None) else v


def _shape_str(shape):
    return '(' + ', '.join('*' if s == -1 else str(s) for s in shape) + ')'


def _check_shape(name, actual, spec):
    ok = len(actual) == len(spec) and all(s == -1 or s == a for a, s in zip(actual, spec))
    if not ok:
        raise RuntimeError(f"Buffer '{name}' has incompatible shape: Expected "
                           f"{_shape_str(spec)} but got {_shape_str(tuple(actual))}.")


def _type_name(dtype):
    return {'float64': 'float64', 'float32': 'float32', 'int32': 'int32'}[dtype]


class _Call:
    """Collects device staging buffers for one engine call and writes
    staged outputs back to their host arrays afterwards."""

    def __init__(self):
        self.dev = _device()
        self.keep = []
        self.writeback = []
        self.any_torch = False

    # -- generic arrays ---------------------------------------------------
    def array(self, x, dtype, name, spec=None, optional=False, out=False):
        """Validate x like so3g's BufferWrapper (dtype -> ValueError, shape ->
        RuntimeError, no buffer -> TypeError) and return a contiguous device
        tensor.  If out, the tensor's final contents are copied back to x."""
        if x is None:
            if optional:
                return None
            raise TypeError(f"Argument '{name}' does not expose buffer protocol, "
                            "is not contiguous, or does not export a format.")
        if _is_torch(x):
            self.any_torch = True
            if x.dtype != _torch_dtype(dtype):
                raise ValueError(f"Expected buffer '{name}' to contain items of type "
                                 f"{_type_name(dtype)}.")
            if spec is not None:
                _check_shape(name, tuple(x.shape), spec)
            t = x
            if t.device != self.dev or not t.is_contiguous():
                t = t.to(self.dev).contiguous()
                if out:
                    self.writeback.append((t, x))
            self.keep.append(t)
            return t
        if isinstance(x, (list, tuple)) or not hasattr(x, '__array_interface__') and \
                not isinstance(x, memoryview):
            raise TypeError(f"Argument '{name}' does not expose buffer protocol, "
                            "is not contiguous, or does not export a format.")
        a = np.asarray(x)
        if a.dtype != np.dtype(dtype):
            raise ValueError(f"Expected buffer '{name}' to contain items of type "
                             f"{_type_name(dtype)}.")
        if spec is not None:
            _check_shape(name, a.shape, spec)
        t = torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)
        self.keep.append(t)
        if out:
            self.writeback.append((t, a))
        return t

    def alloc(self, shape, dtype):
        t = torch.zeros(shape, dtype=_torch_dtype(dtype), device=self.dev)
        self.keep.append(t)
        return t

    def finish(self):
        for t, host in self.writeback:
            if _is_torch(host):
                host.copy_(t)
            else:
                host[...] = t.cpu().numpy()
        self.writeback = []


def _ptr(t):
    return None if t is None else t.data_ptr()


# -- signal-like objects (SignalSpace, Projection.cxx:1212-1298) -------------

def _signal_space(call, obj, name, dtype, n_det, n_t, third=None, zero_if_none=True,
                  out=True):
    """Return (device tensor (n_det, n_t[, third]), result object, fresh).

    obj may be None (zeros allocated; result is a list of rows), a list of
    1-d arrays, a 2-d (or 3-d) numpy array (rows written back in place) or
    a torch tensor (used in place).  out=False marks an input that the call
    only reads: a staged copy is then not written back to the caller's
    arrays."""
    tail = (n_t,) if third is None else (n_t, third)
    if obj is None:
        t = call.alloc((n_det,) + tail, dtype)
        return t, None, True
    if _is_torch(obj):
        spec = [n_det] + list(tail)
        spec = [s if (s is not None and s >= 0) else -1 for s in spec]
        t = call.array(obj, dtype, name, spec=spec, out=out)
        return t, obj, False
    if isinstance(obj, (list, tuple)):
        rows = list(obj)
        if n_det >= 0 and len(rows) != n_det:
            raise RuntimeError(f"Buffer '{name}' has incompatible shape: "
                               "must contain (n_det) vectors.")
        if len(rows) == 0:
            raise RuntimeError(f"Buffer '{name}' has incompatible shape: has not been "
                               "tested on shape 0 objects.")
        arrs = []
        for r in rows:
            if _is_torch(r):
                r = r.detach().cpu().numpy()
            a = np.asarray(r)
            if a.dtype != np.dtype(dtype):
                raise ValueError(f"Expected buffer '{name}' to contain items of type "
                                 f"{_type_name(dtype)}.")
            arrs.append(a)
        shapes = {a.shape for a in arrs}
        if len(shapes) != 1:
            raise RuntimeError(f"Buffer '{name}' has incompatible shape: "
                               "[all elements must have same stride].")
        stacked = np.stack(arrs)
        spec = [len(arrs)] + [s if (s is not None and s >= 0) else -1 for s in tail]
        _check_shape(name, stacked.shape, spec)
        t = torch.from_numpy(stacked).to(call.dev)
        call.keep.append(t)

        class _RowsWriter:
            def __init__(self, rows):
                self.rows = rows

            def __setitem__(self, key, val):
                for r, v in zip(self.rows, val):
                    r[...] = v

        if out:
            call.writeback.append((t, _RowsWriter(arrs)))
        return t, rows, False
    a = np.asarray(obj)
    spec = [n_det] + [s if (s is not None and s >= 0) else -1 for s in tail]
    if a.dtype != np.dtype(dtype):
        raise ValueError(f"Expected buffer '{name}' to contain items of type "
                         f"{_type_name(dtype)}.")
    _check_shape(name, a.shape, spec)
    t = call.array(a, dtype, name, out=out)
    return t, a, False


def _signal_result(call, t, obj):
    """Return value of a signal-like output, as the reference does: a list
    of per-detector rows (or the torch tensor for torch callers)."""
    if obj is None:
        if call.any_torch:
            return t
        return list(t.cpu().numpy())
    if _is_torch(obj):
        return obj
    if isinstance(obj, list):
        return obj
    return list(obj)


# -- thread_intervals (derive_ranges, Projection.cxx:1856-1922) -------------

def _is_ranges(x):
    return isinstance(x, Ranges)


def _flatten_intervals(ti, n_det, n_t, name='thread_intervals'):
    """Flatten None / RangesMatrix(n_det) / [thread][det] /
    [bunch][thread][det] into an int32 (n, 3) array of (det, t0, t1); None
    means full coverage."""
    if ti is None:
        return None
    if _is_ranges(ti[0]):
        bunches = [[ti]]
    elif _is_ranges(ti[0][0]):
        bunches = [[ti[i] for i in range(len(ti))]]
    elif _is_ranges(ti[0][0][0]):
        bunches = [[ti[i][j] for j in range(len(ti[i]))] for i in range(len(ti))]
    else:
        raise RuntimeError(f"Buffer '{name}' has incompatible shape: unrecognised "
                           "thread interval structure.")
    segs = []
    for bunch in bunches:
        for ival in bunch:
            if len(ival) != n_det:
                raise RuntimeError(f"Buffer '{name}' has incompatible shape: Expected "
                                   f"RangesMatrix with n_det={n_det} but encountered "
                                   f"n_det={len(ival)}.")
            for d in range(n_det):
                r = ival[d]
                if r.count != n_t:
                    raise RuntimeError(f"Buffer '{name}' has incompatible shape: Expected "
                                       f"RangesMatrix with n_time={n_t} but encountered "
                                       f"n_time={r.count}.")
                rr = r.ranges()
                if len(rr):
                    seg = np.empty((len(rr), 3), np.int32)
                    seg[:, 0] = d
                    seg[:, 1:] = rr
                    segs.append(seg)
    if not segs:
        return np.zeros((0, 3), np.int32)
    return np.ascontiguousarray(np.concatenate(segs).astype(np.int32))


def _rle_ranges(dom, n_domain, n_t):
    """Group per-sample domains (n_det, n_t) into the reference's
    [2][n_domain|1][n_det] Ranges structure (Projection.cxx:1507-1523)."""
    n_det = dom.shape[0]
    par = [[Ranges(n_t) for _ in range(n_det)] for _ in range(n_domain)]
    ser = [[Ranges(n_t) for _ in range(n_det)]]
    for d in range(n_det):
        row = dom[d]
        if n_t == 0:
            continue
        change = np.flatnonzero(row[1:] != row[:-1]) + 1
        starts = np.concatenate(([0], change))
        ends = np.concatenate((change, [n_t]))
        vals = row[starts]
        for s, e, v in zip(starts, ends, vals):
            if v < 0:
                continue
            target = par[v][d] if v < n_domain else ser[0][d]
            target.append_interval_no_check(int(s), int(e))
    return [par, ser]


# ---------------------------------------------------------------------------
# Engine classes
# ---------------------------------------------------------------------------

class ProjEngBase:
    """Common implementation of the ProjEng_<PROJ>_<SPIN>_<TILING>[_Bilinear]
    classes.  Subclasses set _proj, _comps, _tiled, _interp, _healpix."""
    _proj = 'CAR'
    _comps = 'TQU'
    _tiled = False
    _interp = 0
    _healpix = False
    # Opt-in for nside == 8192 (the reference rejects it, Projection.cxx:513).
    allow_nside_max = False

    def __init__(self, pix_args, allow_nside_max=None):
        if allow_nside_max is not None:
            self.allow_nside_max = bool(allow_nside_max)
        g = N.Geom()
        g.proj = N.PROJ[self._proj] if not self._healpix else N.PROJ['CAR']
        g.pixelizor = N.PIX_HEALPIX if self._healpix else N.PIX_FLAT
        g.tiled = int(self._tiled)
        g.interp = int(self._interp)
        g.ncomp = len(self._comps)
        args = tuple(pix_args)
        self._populate = None
        if self._healpix:
            # (nside, isnest[, nside_tile[, active_tiles]]), Projection.cxx:534-542, 640-675
            g.nside = int(args[0])
            g.nest = int(bool(args[1]))
            g.allow_nside_max = int(self.allow_nside_max)
            if self._tiled:
                g.nside_tile = int(args[2])
                if len(args) >= 4 and args[3] is not None:
                    n_tile = 12 * g.nside_tile * g.nside_tile
                    pop = np.zeros(n_tile, bool)
                    for i in args[3]:
                        i = int(i)
                        if 0 <= i < n_tile:
                            pop[i] = True
                    self._populate = pop
        else:
            # (ny, nx, dy, dx, iy0, ix0[, ty, tx[, active]]), Projection.cxx:821-829, 971-994
            g.naxis[0], g.naxis[1] = int(args[0]), int(args[1])
            g.cdelt[0], g.cdelt[1] = float(args[2]), float(args[3])
            g.crpix[0], g.crpix[1] = float(args[4]), float(args[5])
            if self._tiled:
                g.tile_shape[0], g.tile_shape[1] = int(args[6]), int(args[7])
        N.check(N.load().so3g_geom_init(g))
        if not self._healpix and self._tiled and len(args) >= 9:
            pop = np.zeros(g.n_tile, bool)
            for i in args[8]:
                i = int(i)
                if 0 <= i < g.n_tile:
                    pop[i] = True
            self._populate = pop
        self._g = g

    # -- properties --------------------------------------------------------
    @property
    def index_count(self):
        return int(self._g.index_count)

    @property
    def comp_count(self):
        return len(self._comps)

    # -- geometry helpers ----------------------------------------------------
    def _tile_dims(self):
        g = self._g
        if self._healpix:
            npt = int(g.npix // g.n_tile)
            return np.tile(np.array([1, npt], np.int32), (g.n_tile, 1))
        ty, tx = g.tile_shape[0], g.tile_shape[1]
        ny, nx = g.naxis[0], g.naxis[1]
        n_tx = (nx + tx - 1) // tx
        out = np.zeros((g.n_tile, 2), np.int32)
        for t in range(g.n_tile):
            iy, ix = t // n_tx, t % n_tx
            out[t] = (min(ty, ny - iy * ty), min(tx, nx - ix * tx))
        return out

    def _plane_shape(self):
        g = self._g
        if self._healpix:
            return (int(g.npix),)
        return (g.naxis[0], g.naxis[1])

    def zeros(self, shape):
        """Projection.cxx:1707-1726 with the pixelizors' zeros()."""
        if isinstance(shape, (int, np.integer)):
            lead = (int(shape),)
        elif isinstance(shape, tuple):
            lead = tuple(int(s) for s in shape)
        else:
            return None
        if not self._tiled:
            return np.zeros(lead + self._plane_shape())
        if self._populate is None:
            raise RuntimeError("Buffer 'zeros' has incompatible shape: Cannot create blank "
                               "tiled map unless user has specified what tiles to populate..")
        dims = self._tile_dims()
        out = []
        for t in range(self._g.n_tile):
            if self._populate[t]:
                tshape = (int(dims[t, 1]),) if self._healpix else (int(dims[t, 0]),
                                                                    int(dims[t, 1]))
                out.append(np.zeros(lead + tshape))
            else:
                out.append(None)
        return out

    # -- input staging ---------------------------------------------------------
    def _pointing(self, call, pbore, pofs):
        """Pointer<C>::TestInputs (Projection.cxx:199-226)."""
        bore = call.array(pbore, 'float64', 'boresight', spec=[-1, 4])
        dets = call.array(pofs, 'float64', 'detectors', spec=[-1, 4])
        # NaN scan of the detector offsets -> ValueError.
        nan_rows = torch.isnan(dets).any(dim=1)
        if bool(nan_rows.any()):
            i = int(torch.nonzero(nan_rows)[0, 0])
            raise ValueError(f'Pointing offset error: nan found at index {i}.')
        return bore, dets, bore.shape[0], dets.shape[0]

    def _map(self, call, m, lead, name='map', writeback=True, dtype='float64'):
        """Stage a (lead..., plane) map or a tile list; returns
        (map_ptr, tiles_ptr, dims_ptr)."""
        if not self._tiled:
            spec = list(lead) + list(self._plane_shape())
            t = call.array(m, dtype, name, spec=spec, out=writeback)
            return t.data_ptr(), None, None
        if not isinstance(m, (list, tuple)):
            raise RuntimeError(f"Buffer '{name}' has incompatible shape: tiled maps must be "
                               "given as a list of tiles.")
        g = self._g
        dims = self._tile_dims()
        ptrs = np.zeros(g.n_tile, np.int64)
        if len(m) != g.n_tile:
            raise RuntimeError(f"Buffer '{name}' has incompatible shape: expected "
                               f"{g.n_tile} tiles, got {len(m)}.")
        for i, tile in enumerate(m):
            if tile is None:
                if self._populate is not None and self._populate[i]:
                    raise RuntimeError(f'Tiling problem (index {i}): Projector expects tile '
                                       'but it is missing.')
                continue
            tshape = [int(dims[i, 1])] if self._healpix else [int(dims[i, 0]),
                                                              int(dims[i, 1])]
            t = call.array(tile, dtype, name, spec=list(lead) + tshape, out=writeback)
            ptrs[i] = t.data_ptr()
        ptr_t = torch.from_numpy(ptrs).to(call.dev)
        dims_t = torch.from_numpy(np.ascontiguousarray(dims)).to(call.dev)
        call.keep += [ptr_t, dims_t]
        return None, ptr_t.data_ptr(), dims_t.data_ptr()

    # -- engine methods --------------------------------------------------------
    def coords(self, pbore, pofs, coord=None):
        """Projection.cxx:1316-1349."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        out, obj, _ = _signal_space(call, coord, 'coord', 'float64', n_det, n_t, 4)
        N.check(N.load().so3g_coords(self._g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                     out.data_ptr(), _stream()))
        call.finish()
        return _signal_result(call, out, obj)

    def pixels(self, pbore, pofs, pixel=None):
        """Projection.cxx:1351-1385."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        out, obj, _ = _signal_space(call, pixel, 'pixel', 'int32', n_det, n_t,
                                    self.index_count)
        N.check(N.load().so3g_pixels(self._g, bore.data_ptr(), n_t, dets.data_ptr(), n_det,
                                     out.data_ptr(), _stream()))
        call.finish()
        return _signal_result(call, out, obj)

    def pointing_matrix(self, pbore, pofs, response, pixel=None, proj=None):
        """Projection.cxx:1391-1438."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        pix, pobj, _ = _signal_space(call, pixel, 'pixel', 'int32', n_det, n_t,
                                     self.index_count)
        sp, sobj, _ = _signal_space(call, proj, 'proj', 'float32', n_det, n_t,
                                    self.comp_count)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        N.check(N.load().so3g_pointing_matrix(self._g, bore.data_ptr(), n_t, dets.data_ptr(),
                                              n_det, resp.data_ptr(), pix.data_ptr(),
                                              sp.data_ptr(), _stream()))
        call.finish()
        return (_signal_result(call, pix, pobj), _signal_result(call, sp, sobj))

    def from_map(self, map, pbore, pofs, response, signal=None):
        """Projection.cxx:1728-1770."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        mp, tp, dp = self._map(call, map, [self.comp_count], writeback=False)
        sig, obj, fresh = _signal_space(call, signal, 'signal', 'float32', n_det, n_t)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        N.check(N.load().so3g_from_map(self._g, mp, tp, dp, bore.data_ptr(), n_t,
                                       dets.data_ptr(), n_det, resp.data_ptr(),
                                       sig.data_ptr(), int(not fresh), _stream()))
        call.finish()
        return _signal_result(call, sig, obj)

    def to_map(self, map, pbore, pofs, response, signal, det_weights=None,
               thread_intervals=None):
        """Projection.cxx:1924-1966."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        if map is None:
            map = self.zeros((self.comp_count,))
        mp, tp, dp = self._map(call, map, [self.comp_count])
        sig, _, _ = _signal_space(call, signal, 'signal', 'float32', n_det, n_t, out=False)
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        segs = _flatten_intervals(thread_intervals, n_det, n_t)
        seg_ptr, n_seg = (None, 0) if segs is None else (segs.ctypes.data, len(segs))
        if segs is not None and n_seg == 0:
            seg_ptr = np.zeros((1, 3), np.int32).ctypes.data
        N.check(N.load().so3g_to_map(self._g, mp, tp, dp, bore.data_ptr(), n_t,
                                     dets.data_ptr(), n_det, resp.data_ptr(), sig.data_ptr(),
                                     _ptr(dw), seg_ptr, n_seg, _stream()))
        call.finish()
        return map

    def to_weight_map(self, map, pbore, pofs, response, det_weights=None,
                      thread_intervals=None):
        """Projection.cxx:1968-2010."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        nc = self.comp_count
        if map is None:
            map = self.zeros((nc, nc))
        mp, tp, dp = self._map(call, map, [nc, nc])
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        segs = _flatten_intervals(thread_intervals, n_det, n_t)
        seg_ptr, n_seg = (None, 0) if segs is None else (segs.ctypes.data, len(segs))
        if segs is not None and n_seg == 0:
            seg_ptr = np.zeros((1, 3), np.int32).ctypes.data
        N.check(N.load().so3g_to_weights(self._g, mp, tp, dp, bore.data_ptr(), n_t,
                                         dets.data_ptr(), n_det, resp.data_ptr(), _ptr(dw),
                                         seg_ptr, n_seg, _stream()))
        call.finish()
        return map

    def to_map_prec(self, map, pbore, pofs, response, signal, det_weights=None, prec=0):
        """to_map in one of three arithmetics (no counterpart in the
        reference, whose arithmetic is prec=0): 0 float32 products added to a
        float64 map (so3g's), 1 float32 products and a float32 map, 2 float64
        products and a float64 map.  `map` may be None."""
        dt = 'float32' if prec == 1 else 'float64'
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        if map is None:
            map = self._zeros_dtype((self.comp_count,), dt)
        mp, tp, dp = self._map(call, map, [self.comp_count], dtype=dt)
        sig, _, _ = _signal_space(call, signal, 'signal', 'float32', n_det, n_t, out=False)
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        N.check(N.load().so3g_to_map_prec(self._g, mp, tp, dp, bore.data_ptr(), n_t,
                                          dets.data_ptr(), n_det, resp.data_ptr(),
                                          sig.data_ptr(), _ptr(dw), int(prec), _stream()))
        call.finish()
        return map

    def to_weight_map_prec(self, map, pbore, pofs, response, det_weights=None, prec=0):
        """to_weight_map in the arithmetic selected by prec (see to_map_prec)."""
        dt = 'float32' if prec == 1 else 'float64'
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        nc = self.comp_count
        if map is None:
            map = self._zeros_dtype((nc, nc), dt)
        mp, tp, dp = self._map(call, map, [nc, nc], dtype=dt)
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        resp = call.array(response, 'float32', 'response', spec=[n_det, 2])
        N.check(N.load().so3g_to_weights_prec(self._g, mp, tp, dp, bore.data_ptr(), n_t,
                                              dets.data_ptr(), n_det, resp.data_ptr(), _ptr(dw),
                                              int(prec), _stream()))
        call.finish()
        return map

    def _zeros_dtype(self, lead, dtype):
        z = self.zeros(lead)
        if isinstance(z, list):
            return [None if t is None else t.astype(dtype) for t in z]
        return z.astype(dtype)

    def tile_hits(self, pbore, pofs):
        """Projection.cxx:1543-1601; returns a list of per-tile hit counts."""
        if not self._tiled:
            raise RuntimeError('No tiles in this pixelization.')
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        hits = call.alloc((self._g.n_tile,), 'int32')
        N.check(N.load().so3g_tile_hits(self._g, bore.data_ptr(), n_t, dets.data_ptr(),
                                        n_det, hits.data_ptr(), _stream()))
        return [int(h) for h in hits.cpu().numpy()]

    def _domains(self, call, bore, dets, n_t, n_det, mode, n_domain, tmap=None,
                 tile_dom=None):
        out = call.alloc((n_det, n_t), 'int32')
        mp = tp = dp = None
        if tmap is not None:
            mp, tp, dp = self._map(call, tmap, [1], name='map', writeback=False)
        td = None
        if tile_dom is not None:
            td = torch.from_numpy(np.ascontiguousarray(tile_dom, np.int32)).to(call.dev)
            call.keep.append(td)
        N.check(N.load().so3g_sample_domains(self._g, bore.data_ptr(), n_t, dets.data_ptr(),
                                             n_det, mode, int(n_domain), mp, tp, dp, _ptr(td),
                                             out.data_ptr(), _stream()))
        return out.cpu().numpy()

    def pixel_ranges(self, pbore, pofs, map=None, n_domain=-1):
        """Projection.cxx:1440-1541."""
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        if n_domain is None or n_domain <= 0:
            from . import useful_info
            n_domain = useful_info()['omp_num_threads']
        mode = N.DOMAIN_MAP if map is not None else N.DOMAIN_STRIPE
        dom = self._domains(call, bore, dets, n_t, n_det, mode, n_domain, tmap=map)
        return _rle_ranges(dom, n_domain, n_t)

    def tile_ranges(self, pbore, pofs, tile_lists):
        """Projection.cxx:1606-1705."""
        if not self._tiled:
            raise RuntimeError('No tiles in this pixelization.')
        call = _Call()
        bore, dets, n_t, n_det = self._pointing(call, pbore, pofs)
        n_domain = len(tile_lists)
        tdom = np.full(self._g.n_tile, -1, np.int32)
        for i, tl in enumerate(tile_lists):
            for t in tl:
                tdom[int(t)] = i
        dom = self._domains(call, bore, dets, n_t, n_det, N.DOMAIN_TILE, max(n_domain, 1),
                            tile_dom=tdom)
        return _rle_ranges(dom, n_domain, n_t)


class ProjEngPrecompBase:
    """ProjEng_Precomp_{NonTiled,Tiled} (Projection.cxx:2012-2253): projection
    through a cached pixel_index / spin_proj pointing matrix.

    When pixel_index and spin_proj are torch tensors on the GPU, the engine
    builds a PrecompPlan from them on first use and keeps it (plan_cache,
    default True), so that repeated calls with the same pointing matrix (a
    CG loop) run the planned kernels: same from_map results bit for bit,
    to_map / to_weight_map the same float64 sums of the same float32 products
    in another order.  The plan is keyed on both tensors' storage, shape,
    strides and version counters, so an in-place edit of either rebuilds it;
    it holds device memory (about 20 B per sample) until clear_plan() or
    another pointing matrix replaces it.  thread_intervals, a tile list with
    inactive (None) tiles, numpy inputs (restaged every call anyway) and
    to_weight_map on a pixel-run plan use the direct kernels."""
    _tiled = False
    plan_cache = True

    def __init__(self):
        self._plan = None
        self._plan_key = None

    def clear_plan(self):
        """Free the cached plan."""
        self._plan = None
        self._plan_key = None

    def _cached_plan(self, pixel_index, spin_proj, pix, spin, map):
        """The PrecompPlan for (pix, spin) and this map's geometry, or None
        when the direct kernels are to be used."""
        if not (self.plan_cache and _is_torch(pixel_index) and _is_torch(spin_proj)):
            return None
        if pix.data_ptr() != pixel_index.data_ptr() or spin.data_ptr() != spin_proj.data_ptr():
            return None      # staged copies: not the caller's tensors
        n_det, n_t, k = pix.shape
        if n_det * n_t == 0 or n_det * n_t >= 2 ** 32 or n_det >= 2 ** 22 or n_t >= 2 ** 30:
            return None
        if k == 3:
            if not isinstance(map, (list, tuple)) or any(t is None for t in map):
                return None
            geo = ('tiles', tuple(tuple(int(x) for x in t.shape[-2:]) for t in map))
        else:
            if isinstance(map, (list, tuple)) or not hasattr(map, 'shape'):
                return None
            geo = ('map', tuple(int(x) for x in map.shape[-k:]))
        key = (pixel_index.data_ptr(), tuple(pixel_index.shape), tuple(pixel_index.stride()),
               pixel_index._version, spin_proj.data_ptr(), tuple(spin_proj.shape),
               tuple(spin_proj.stride()), spin_proj._version, geo, pixel_index.device)
        if key == self._plan_key:
            return self._plan
        self._plan = None
        self._plan_key = None
        try:
            if k == 3:
                plan = PrecompPlan(pix, spin, tile_dims=np.array(geo[1], np.int32))
            else:
                plan = PrecompPlan(pix, spin, geo[1])
        except RuntimeError as e:
            if 'out of device memory' in str(e):
                return None
            raise
        self._plan, self._plan_key = plan, key
        return plan

    def _pmat(self, call, pixel_index, spin_proj):
        pix, _, _ = _signal_space(call, pixel_index, 'pixel_index', 'int32', -1, -1, -1,
                                  out=False)
        if pix.dim() != 3:
            raise RuntimeError("Buffer 'pixel_index' has incompatible shape: Expected "
                               "(*, *, *).")
        n_det, n_t, k = pix.shape
        if not _is_torch(pixel_index) and not isinstance(pixel_index, (list, tuple)):
            a = np.asarray(pixel_index)
            if a.ndim == 3 and a.strides[2] != a.itemsize and a.shape[2] > 1:
                raise RuntimeError("Buffer 'pixel_index' has incompatible shape: Fast "
                                   "dimension of pixel indices must be close-packed..")
        spin, _, _ = _signal_space(call, spin_proj, 'spin_proj', 'float32', n_det, n_t, -1,
                                   out=False)
        return pix, spin, n_det, n_t, k, spin.shape[2]

    @staticmethod
    def _check_lead(m, nlead, nc, k):
        """The map's leading (component) axes for a planned call: (nc,) for
        maps, (nc, nc) for weights, as the direct kernels index them."""
        tiles = m if isinstance(m, (list, tuple)) else [m]
        for t in tiles:
            shape = tuple(t.shape)
            want = (nc,) * nlead
            if len(shape) != nlead + (1 if k == 1 else 2) or shape[:nlead] != want:
                raise RuntimeError(f"Buffer 'map' has incompatible shape: Expected leading "
                                   f"axes {want} but got {shape}.")

    def _map(self, call, m, nlead, k, writeback):
        """Returns (map_ptr, shape_ptr, tiles_ptr, dims_ptr)."""
        if not self._tiled:
            t = call.array(m, 'float64', 'map', out=writeback)
            if t.dim() != nlead + (1 if k == 1 else 2):
                raise RuntimeError("Buffer 'map' has incompatible shape: Expected (" +
                                   ', '.join(['*'] * (nlead + 2)) + ").")
            if k == 1:
                shape = np.array([t.shape[-1], 1], np.int64)
            else:
                shape = np.array([t.shape[-2], t.shape[-1]], np.int64)
            call.keep.append(shape)
            return t.data_ptr(), shape.ctypes.data, None, None
        if not isinstance(m, (list, tuple)):
            raise RuntimeError("Buffer 'map' has incompatible shape: tiled maps must be given "
                               "as a list of tiles.")
        ptrs = np.zeros(len(m), np.int64)
        dims = np.zeros((len(m), 2), np.int32)
        for i, tile in enumerate(m):
            if tile is None:
                continue
            t = call.array(tile, 'float64', 'map', out=writeback)
            ptrs[i] = t.data_ptr()
            dims[i] = (t.shape[-2], t.shape[-1])
        ptr_t = torch.from_numpy(ptrs).to(call.dev)
        dims_t = torch.from_numpy(dims).to(call.dev)
        shape = np.array([len(m), 1], np.int64)
        call.keep += [ptr_t, dims_t, shape]
        return None, shape.ctypes.data, ptr_t.data_ptr(), dims_t.data_ptr()

    def from_map(self, map, pixel_index, spin_proj, signal=None):
        call = _Call()
        pix, spin, n_det, n_t, k, nc = self._pmat(call, pixel_index, spin_proj)
        plan = self._cached_plan(pixel_index, spin_proj, pix, spin, map)
        if plan is not None:
            self._check_lead(map, 1, nc, k)
            if signal is None:
                # a fresh signal comes back as a device tensor, as on the
                # direct path for torch pointing matrices
                out = torch.empty((n_det, n_t), dtype=torch.float32, device=pix.device)
                plan.from_map(map, out, fresh=True)
                return out
            return plan.from_map(map, signal)
        mp, sp, tp, dp = self._map(call, map, 1, k, writeback=False)
        sig, obj, fresh = _signal_space(call, signal, 'signal', 'float32', n_det, n_t)
        N.check(N.load().so3g_precomp_from_map(mp, sp, tp, dp, pix.data_ptr(), k,
                                               spin.data_ptr(), nc, n_det, n_t,
                                               sig.data_ptr(), int(not fresh), _stream()))
        call.finish()
        return _signal_result(call, sig, obj)

    def to_map(self, map, pixel_index, spin_proj, signal, weights=None,
               thread_intervals=None):
        call = _Call()
        pix, spin, n_det, n_t, k, nc = self._pmat(call, pixel_index, spin_proj)
        if thread_intervals is None:
            plan = self._cached_plan(pixel_index, spin_proj, pix, spin, map)
            if plan is not None:
                self._check_lead(map, 1, nc, k)
                return plan.to_map(map, signal, weights)
        mp, sp, tp, dp = self._map(call, map, 1, k, writeback=True)
        sig, _, _ = _signal_space(call, signal, 'signal', 'float32', n_det, n_t, out=False)
        dw = call.array(weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        segs = _flatten_intervals(thread_intervals, n_det, n_t)
        seg_ptr, n_seg = (None, 0) if segs is None else (segs.ctypes.data, len(segs))
        if segs is not None and n_seg == 0:
            seg_ptr = np.zeros((1, 3), np.int32).ctypes.data
        N.check(N.load().so3g_precomp_to_map(mp, sp, tp, dp, pix.data_ptr(), k,
                                             spin.data_ptr(), nc, n_det, n_t, sig.data_ptr(),
                                             _ptr(dw), seg_ptr, n_seg, _stream()))
        call.finish()
        return map

    def to_weight_map(self, map, pixel_index, spin_proj, weights=None,
                      thread_intervals=None):
        call = _Call()
        pix, spin, n_det, n_t, k, nc = self._pmat(call, pixel_index, spin_proj)
        if thread_intervals is None:
            plan = self._cached_plan(pixel_index, spin_proj, pix, spin, map)
            if plan is not None:
                self._check_lead(map, 2, nc, k)
                return plan.to_weights(map, weights)
        mp, sp, tp, dp = self._map(call, map, 2, k, writeback=True)
        dw = call.array(weights, 'float32', 'det_weights', spec=[n_det], optional=True)
        segs = _flatten_intervals(thread_intervals, n_det, n_t)
        seg_ptr, n_seg = (None, 0) if segs is None else (segs.ctypes.data, len(segs))
        if segs is not None and n_seg == 0:
            seg_ptr = np.zeros((1, 3), np.int32).ctypes.data
        N.check(N.load().so3g_precomp_to_weights(mp, sp, tp, dp, pix.data_ptr(), k,
                                                 spin.data_ptr(), nc, n_det, n_t, _ptr(dw),
                                                 seg_ptr, n_seg, _stream()))
        call.finish()
        return map


FLAT_PROJS = ['Flat', 'Quat', 'CAR', 'CEA', 'ARC', 'SIN', 'TAN', 'ZEA']
SPINS = ['T', 'QU', 'TQU']


def make_engine_classes():
    """Build the named engine classes (Projection.cxx:2283-2356)."""
    out = {}
    for proj in FLAT_PROJS:
        for comps in SPINS:
            for tiling in ['Tiled', 'NonTiled']:
                for interp, suffix in [(0, ''), (1, '_Bilinear')]:
                    name = f'ProjEng_{proj}_{comps}_{tiling}{suffix}'
                    out[name] = type(name, (ProjEngBase,), dict(
                        _proj=proj, _comps=comps, _tiled=(tiling == 'Tiled'),
                        _interp=interp, _healpix=False))
    for comps in SPINS:
        for tiling in ['NonTiled', 'Tiled']:
            name = f'ProjEng_HP_{comps}_{tiling}'
            out[name] = type(name, (ProjEngBase,), dict(
                _proj='CAR', _comps=comps, _tiled=(tiling == 'Tiled'), _interp=0,
                _healpix=True))
    out['ProjEng_Precomp_NonTiled'] = type('ProjEng_Precomp_NonTiled',
                                           (ProjEngPrecompBase,), dict(_tiled=False))
    out['ProjEng_Precomp_Tiled'] = type('ProjEng_Precomp_Tiled', (ProjEngPrecompBase,),
                                        dict(_tiled=True))
    return out


ENGINE_NAMES = sorted(make_engine_classes())


class PrecompPlan:
    """Sorted plan for repeated cached-pointing projection with one pointing
    matrix (CG-style loops).  No counterpart in the reference: built once
    from a ProjEng_Precomp pixel_index (k = 2, non-tiled flat map of
    `map_shape` = (ny, nx); or k = 1, a non-tiled HEALPix map of `map_shape`
    = (npix,); or k = 3, a tiled flat map, with `tile_dims` the (n_tile, 2)
    tile shapes, maps then given as tile lists) and spin_proj (both copied
    into the plan).  It numbers the
    visited pixels compactly and sorts the on-map samples so that to_map
    needs no per-sample atomics: layout 1 (the default from 2^25 samples up;
    layout=None chooses by size) sorts them by map
    tile, keeping sample order inside a tile, and sums each work item in the
    tile held in LDS; layout 0 sorts them by pixel and sums each pixel's run
    with a wave-level segmented reduction; layout 2 (the default below 2^25
    samples where it applies, else layout 0) keeps time order in work items
    of one detector's consecutive samples whose visited pixels span at most
    2048 compact values, sums each item in that window in LDS and adds the
    items' window sums per pixel in a second pass, and gathers from_map from
    the window loaded into LDS.  Layouts 0 and 1 read a compact time-ordered
    copy in from_map.  to_map adds what ProjEng_Precomp_NonTiled.to_map adds
    over every sample (float64 sums of the same float32 products, in a
    different order); from_map returns what ProjEng_Precomp_NonTiled.from_map
    returns, bit for bit.  t_block_mb: the time block of the sort (None: the
    layout's default, 0 for tiles, 64 MiB of signal for pixel runs);
    time_copy=False: a tiled plan without the time-ordered copy (from_map
    then gathers from the tiles; 7.4 vs 16 GB for 7.2e8 samples)."""

    def __init__(self, pixel_index, spin_proj, map_shape=None, t_block_mb=None, layout=None,
                 time_copy=None, tile_dims=None):
        import ctypes
        call = _Call()
        pix = call.array(pixel_index, 'int32', 'pixel_index', spec=[-1, -1, -1])
        n_det, n_t, k = pix.shape
        spin = call.array(spin_proj, 'float32', 'spin_proj', spec=[n_det, n_t, -1])
        self.n_det, self.n_t, self.k = int(n_det), int(n_t), int(k)
        self.nc = int(spin.shape[2])
        if k == 3:
            # tiled map: pixel = (tile, row, column); tile_dims (n_tile, 2)
            if tile_dims is None:
                raise ValueError('plans for k = 3 (tiled) pixels need tile_dims')
            self.tile_dims = np.ascontiguousarray(tile_dims, np.int32).reshape(-1, 2)
            self.map_shape = None
        else:
            if k not in (1, 2) or map_shape is None or len(map_shape) != k:
                raise ValueError('plans take k = 2 pixels with map_shape (ny, nx), k = 1 '
                                 'pixels with map_shape (npix,) or k = 3 pixels with '
                                 'tile_dims')
            self.tile_dims = None
            self.map_shape = tuple(int(x) for x in map_shape)
            self._shape = np.array(self.map_shape + (1,) * (2 - k), np.int64)
        h = ctypes.c_void_p()
        opts = {'plan_layout': layout, 'plan_time_copy': time_copy}
        old = {n: N.get_option(n) for n in opts}
        try:
            for n, v in opts.items():
                if v is not None:
                    N.set_option(n, int(v))
            tb = -1 if t_block_mb is None else int(t_block_mb)
            if k == 3:
                N.check(N.load().so3g_pc_plan_create_tiled(
                    pix.data_ptr(), spin.data_ptr(), self.nc, self.n_det, self.n_t,
                    self.tile_dims.ctypes.data, len(self.tile_dims), tb, ctypes.byref(h),
                    _stream()))
            else:
                N.check(N.load().so3g_pc_plan_create(pix.data_ptr(), k, spin.data_ptr(),
                                                     self.nc, self.n_det, self.n_t,
                                                     self._shape.ctypes.data, tb,
                                                     ctypes.byref(h), _stream()))
        finally:
            for n, v in old.items():
                N.set_option(n, v)
        self._h = h

    def info(self):
        """n_on (on-map samples), runs (pixel runs in sorted order),
        t_const (T factor stored once per detector), stored_spin (factors
        stored per sample), t_block (samples per time block), bytes, n_vis
        (visited pixels), layout (0 pixel-sorted runs, 1 map tiles, 2
        time-ordered windows), for the tiled layout stretches (starts of runs
        of consecutive samples within a tile, counting each 64-position
        chunk's first position), items (work items of the tiled and windowed
        layouts); wide_groups: 512-sample groups of the time-ordered copy
        whose compact pixels span more than 65535 values; window_max and
        window_slots: the windowed layout's largest window and the windows'
        total size."""
        a = (ctypes.c_int64 * 13)()
        N.check(N.load().so3g_pc_plan_stats(self._h, a))
        return {'n_on': a[0], 'runs': a[1], 't_const': bool(a[2]), 'stored_spin': a[3],
                't_block': a[4], 'bytes': a[5], 'n_vis': a[6], 'layout': a[7],
                'stretches': a[8], 'items': a[9], 'wide_groups': a[10],
                'window_max': a[11], 'window_slots': a[12]}

    def _tiles(self, call, map, writeback):
        """Host array of the tile list's device pointers (0 for None)."""
        if not isinstance(map, (list, tuple)) or len(map) != len(self.tile_dims):
            raise RuntimeError(f"Buffer 'map' has incompatible shape: expected a list of "
                               f"{len(self.tile_dims)} tiles.")
        ptrs = np.zeros(len(map), np.int64)
        for i, tile in enumerate(map):
            if tile is not None:
                spec = [self.nc] + [int(x) for x in self.tile_dims[i]]
                ptrs[i] = call.array(tile, 'float64', 'map', spec=spec, out=writeback).data_ptr()
        call.keep.append(ptrs)
        return ptrs.ctypes.data

    def to_map(self, map, signal, det_weights=None, overwrite=False):
        """map += P^T signal; map is an (nc, ...) array, or a list of tiles
        (None for inactive ones) for a k = 3 plan.  overwrite=True (non-tiled
        maps): map = P^T signal, whatever map held before."""
        call = _Call()
        sig = call.array(signal, 'float32', 'signal', spec=[self.n_det, self.n_t])
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[self.n_det],
                        optional=True)
        if self.k == 3:
            N.check(N.load().so3g_pc_plan_to_map_tiles(self._h, self._tiles(call, map, True),
                                                       sig.data_ptr(), _ptr(dw), _stream()))
        else:
            m = call.array(map, 'float64', 'map', spec=[self.nc] + list(self.map_shape),
                           out=True)
            f = (N.load().so3g_pc_plan_to_map_set if overwrite else
                 N.load().so3g_pc_plan_to_map)
            N.check(f(self._h, m.data_ptr(), sig.data_ptr(), _ptr(dw), _stream()))
        call.finish()
        return map

    def to_weights(self, map, det_weights=None):
        """map += P^T diag(det_weights) P as ProjEng_Precomp.to_weight_map
        fills it: the upper triangle of an (nc, nc, ...) array, or of each
        tile of a tile list (k = 3)."""
        call = _Call()
        dw = call.array(det_weights, 'float32', 'det_weights', spec=[self.n_det],
                        optional=True)
        if self.k == 3:
            if not isinstance(map, (list, tuple)) or len(map) != len(self.tile_dims):
                raise RuntimeError(f"Buffer 'map' has incompatible shape: expected a list of "
                                   f"{len(self.tile_dims)} tiles.")
            ptrs = np.zeros(len(map), np.int64)
            for i, tile in enumerate(map):
                if tile is not None:
                    spec = [self.nc, self.nc] + [int(x) for x in self.tile_dims[i]]
                    ptrs[i] = call.array(tile, 'float64', 'map', spec=spec,
                                         out=True).data_ptr()
            call.keep.append(ptrs)
            N.check(N.load().so3g_pc_plan_to_weights_tiles(self._h, ptrs.ctypes.data, _ptr(dw),
                                                           _stream()))
        else:
            m = call.array(map, 'float64', 'map',
                           spec=[self.nc, self.nc] + list(self.map_shape), out=True)
            N.check(N.load().so3g_pc_plan_to_weights(self._h, m.data_ptr(), _ptr(dw),
                                                     _stream()))
        call.finish()
        return map

    def from_map(self, map, signal=None, fresh=None):
        """signal (+)= P map; a given signal is accumulated into, as
        ProjEng_Precomp.from_map does (fresh=True overwrites it instead);
        None returns fresh rows."""
        call = _Call()
        sig, obj, new = _signal_space(call, signal, 'signal', 'float32', self.n_det,
                                      self.n_t)
        fresh = new if fresh is None else bool(fresh)
        if self.k == 3:
            N.check(N.load().so3g_pc_plan_from_map_tiles(self._h, self._tiles(call, map, False),
                                                         sig.data_ptr(), 0 if fresh else 1,
                                                         _stream()))
        else:
            m = call.array(map, 'float64', 'map', spec=[self.nc] + list(self.map_shape))
            N.check(N.load().so3g_pc_plan_from_map(self._h, m.data_ptr(), sig.data_ptr(),
                                                   0 if fresh else 1, _stream()))
        call.finish()
        return _signal_result(call, sig, obj)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                N.load().so3g_pc_plan_destroy(h)
            except Exception:
                pass
            self._h = None
