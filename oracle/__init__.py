"""CPU oracle for the projection hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so (built from oracle/so3g_oracle.cpp by
oracle/Makefile) and of oracle/_ref/libso3g_ref.so (the reference's own
healpix_bare.c and so_linterp.h compiled in this container, when present).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
use this module, and only as the checker or the CPU baseline -- never as
the product path.

Layouts match the C ABI of the HIP library: bore (n_t, 4) f64, dets
(n_det, 4) f64, resp (n_det, 2) f32, signal (n_det, n_t) f32, maps
(lead..., plane) f64 or a list of tiles.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')
REF_LIB = os.path.join(HERE, '_ref', 'libso3g_ref.so')

PROJ = {'Flat': 0, 'Quat': 1, 'CAR': 2, 'CEA': 3, 'ARC': 4, 'SIN': 5, 'TAN': 6, 'ZEA': 7}


class Geom(ctypes.Structure):
    _fields_ = [('proj', ctypes.c_int32), ('pixelizor', ctypes.c_int32),
                ('tiled', ctypes.c_int32), ('interp', ctypes.c_int32),
                ('ncomp', ctypes.c_int32), ('ny', ctypes.c_int32), ('nx', ctypes.c_int32),
                ('cdelt_y', ctypes.c_double), ('cdelt_x', ctypes.c_double),
                ('crpix_y', ctypes.c_double), ('crpix_x', ctypes.c_double),
                ('tile_y', ctypes.c_int32), ('tile_x', ctypes.c_int32),
                ('nside', ctypes.c_int32), ('nest', ctypes.c_int32),
                ('nside_tile', ctypes.c_int32)]


def build(ref=True):
    """Compile liboracle.so (and, if the reference tree exists, _ref)."""
    targets = ['all']
    if ref and os.path.isdir('/root/reference'):
        targets.append('ref')
    subprocess.run(['make', '-s', '-C', HERE] + targets, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        _lib = ctypes.CDLL(LIB)
    return _lib


def ref_lib():
    """The reference's own healpix_bare.c / so_linterp.h, compiled here;
    None where the reference tree was not available to build it."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = ctypes.CDLL(REF_LIB)
    return _ref


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def geom(proj='CAR', ncomp=3, ny=1, nx=1, cdelt=(1., 1.), crpix=(1., 1.), tiled=False,
         tile_shape=(1, 1), interp=0, healpix=False, nside=0, nest=1, nside_tile=0):
    g = Geom()
    g.proj = PROJ[proj]
    g.pixelizor = 1 if healpix else 0
    g.tiled = int(tiled)
    g.interp = int(interp)
    g.ncomp = int(ncomp)
    g.ny, g.nx = int(ny), int(nx)
    g.cdelt_y, g.cdelt_x = float(cdelt[0]), float(cdelt[1])
    g.crpix_y, g.crpix_x = float(crpix[0]), float(crpix[1])
    g.tile_y, g.tile_x = int(tile_shape[0]), int(tile_shape[1])
    g.nside, g.nest, g.nside_tile = int(nside), int(nest), int(nside_tile)
    return g


def geom_from_args(eng_kind, comps, pix_args, interp=0):
    """Oracle geometry from an engine name's parts and its pixelizor args."""
    nc = len(comps)
    if eng_kind == 'HP':
        tiled = len(pix_args) > 2 and pix_args[2] is not None
        return geom('CAR', nc, healpix=True, nside=pix_args[0], nest=int(bool(pix_args[1])),
                    tiled=tiled, nside_tile=(pix_args[2] if tiled else 0))
    tiled = len(pix_args) >= 8
    return geom(eng_kind, nc, ny=pix_args[0], nx=pix_args[1], cdelt=pix_args[2:4],
                crpix=pix_args[4:6], tiled=tiled,
                tile_shape=pix_args[6:8] if tiled else (1, 1), interp=interp)


def index_count(g):
    if g.pixelizor == 1:
        return 2 if g.tiled else 1
    return 3 if g.tiled else 2


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _tiles(tiles):
    """(array of tile pointers, keepalive list) for a list of tiles."""
    ptrs = (ctypes.c_void_p * len(tiles))()
    keep = []
    for i, t in enumerate(tiles):
        if t is None:
            ptrs[i] = None
        else:
            assert t.dtype == np.float64 and t.flags.c_contiguous
            ptrs[i] = t.ctypes.data
            keep.append(t)
    return ptrs, keep


def _map_args(m):
    if isinstance(m, list):
        ptrs, keep = _tiles(m)
        return None, ptrs, keep
    assert m.dtype == np.float64 and m.flags.c_contiguous
    return _p(m), None, [m]


def _segs(segs):
    if segs is None:
        return None, 0
    s = _c(segs, np.int32).reshape(-1, 3)
    return s, len(s)


def coords(g, bore, dets):
    bore, dets = _c(bore, np.float64), _c(dets, np.float64)
    out = np.zeros((len(dets), len(bore), 4))
    f = lib().oracle_coords
    f(ctypes.byref(g), _p(bore), ctypes.c_int64(len(bore)), _p(dets),
      ctypes.c_int64(len(dets)), _p(out))
    return out


def pixels(g, bore, dets):
    bore, dets = _c(bore, np.float64), _c(dets, np.float64)
    out = np.zeros((len(dets), len(bore), index_count(g)), np.int32)
    lib().oracle_pixels(ctypes.byref(g), _p(bore), ctypes.c_int64(len(bore)), _p(dets),
                        ctypes.c_int64(len(dets)), _p(out))
    return out


def pointing_matrix(g, bore, dets, resp):
    bore, dets, resp = _c(bore, np.float64), _c(dets, np.float64), _c(resp, np.float32)
    pix = np.zeros((len(dets), len(bore), index_count(g)), np.int32)
    spin = np.zeros((len(dets), len(bore), g.ncomp), np.float32)
    lib().oracle_pointing_matrix(ctypes.byref(g), _p(bore), ctypes.c_int64(len(bore)),
                                 _p(dets), ctypes.c_int64(len(dets)), _p(resp), _p(pix),
                                 _p(spin))
    return pix, spin


def from_map(g, m, bore, dets, resp, signal=None):
    bore, dets, resp = _c(bore, np.float64), _c(dets, np.float64), _c(resp, np.float32)
    sig = np.zeros((len(dets), len(bore)), np.float32) if signal is None else \
        _c(signal, np.float32).copy()
    mp, tp, keep = _map_args(m)
    f = lib().oracle_from_map
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(bore), ctypes.c_int64(len(bore)), _p(dets),
            ctypes.c_int64(len(dets)), _p(resp), _p(sig))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return sig


def to_map(g, m, bore, dets, resp, signal, det_weights=None, segs=None):
    """Accumulates into m (array or tile list) in place; returns m."""
    bore, dets, resp = _c(bore, np.float64), _c(dets, np.float64), _c(resp, np.float32)
    sig = _c(signal, np.float32)
    dw = None if det_weights is None else _c(det_weights, np.float32)
    s, ns = _segs(segs)
    mp, tp, keep = _map_args(m)
    f = lib().oracle_to_map
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(bore), ctypes.c_int64(len(bore)), _p(dets),
            ctypes.c_int64(len(dets)), _p(resp), _p(sig), _p(dw), _p(s), ctypes.c_int64(ns))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return m


def to_weights(g, m, bore, dets, resp, det_weights=None, segs=None):
    bore, dets, resp = _c(bore, np.float64), _c(dets, np.float64), _c(resp, np.float32)
    dw = None if det_weights is None else _c(det_weights, np.float32)
    s, ns = _segs(segs)
    mp, tp, keep = _map_args(m)
    f = lib().oracle_to_weights
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(bore), ctypes.c_int64(len(bore)), _p(dets),
            ctypes.c_int64(len(dets)), _p(resp), _p(dw), _p(s), ctypes.c_int64(ns))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return m


def to_map_threaded(g, m, bore, dets, resp, signal, thread_segs, det_weights=None):
    """OpenMP to_map over per-thread pixel-disjoint segment lists."""
    bore, dets, resp = _c(bore, np.float64), _c(dets, np.float64), _c(resp, np.float32)
    sig = _c(signal, np.float32)
    dw = None if det_weights is None else _c(det_weights, np.float32)
    offs = np.zeros(len(thread_segs) + 1, np.int64)
    for i, s in enumerate(thread_segs):
        offs[i + 1] = offs[i] + len(s)
    allseg = _c(np.concatenate([np.asarray(s, np.int32).reshape(-1, 3)
                                for s in thread_segs] + [np.zeros((1, 3), np.int32)]),
                np.int32)
    mp, tp, keep = _map_args(m)
    f = lib().oracle_to_map_threaded
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(bore), ctypes.c_int64(len(bore)), _p(dets),
            ctypes.c_int64(len(dets)), _p(resp), _p(sig), _p(dw), _p(allseg), _p(offs),
            ctypes.c_int(len(thread_segs)))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return m


def sample_domains(g, bore, dets, mode, n_domain, tmap=None, tile_dom=None):
    bore, dets = _c(bore, np.float64), _c(dets, np.float64)
    out = np.zeros((len(dets), len(bore)), np.int32)
    mp = tp = None
    keep = []
    if tmap is not None:
        mp, tp, keep = _map_args(tmap)
    td = None if tile_dom is None else _c(tile_dom, np.int32)
    lib().oracle_sample_domains(ctypes.byref(g), _p(bore), ctypes.c_int64(len(bore)),
                                _p(dets), ctypes.c_int64(len(dets)), ctypes.c_int(mode),
                                ctypes.c_int(n_domain), mp, tp, _p(td), _p(out))
    return out


def tile_hits(g, bore, dets, n_tile):
    bore, dets = _c(bore, np.float64), _c(dets, np.float64)
    hits = np.zeros(n_tile, np.int32)
    lib().oracle_tile_hits(ctypes.byref(g), _p(bore), ctypes.c_int64(len(bore)), _p(dets),
                           ctypes.c_int64(len(dets)), _p(hits))
    return hits


def precomp_from_map(g, m, pix, spin, signal=None):
    pix, spin = _c(pix, np.int32), _c(spin, np.float32)
    n_det, n_t, k = pix.shape
    nc = spin.shape[2]
    sig = np.zeros((n_det, n_t), np.float32) if signal is None else \
        _c(signal, np.float32).copy()
    mp, tp, keep = _map_args(m)
    f = lib().oracle_precomp_from_map
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(pix), ctypes.c_int(k), _p(spin), ctypes.c_int(nc),
            ctypes.c_int64(n_det), ctypes.c_int64(n_t), _p(sig))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return sig


def precomp_to_map(g, m, pix, spin, signal, det_weights=None, segs=None):
    pix, spin, sig = _c(pix, np.int32), _c(spin, np.float32), _c(signal, np.float32)
    n_det, n_t, k = pix.shape
    nc = spin.shape[2]
    dw = None if det_weights is None else _c(det_weights, np.float32)
    s, ns = _segs(segs)
    mp, tp, keep = _map_args(m)
    f = lib().oracle_precomp_to_map
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(pix), ctypes.c_int(k), _p(spin), ctypes.c_int(nc),
            ctypes.c_int64(n_det), ctypes.c_int64(n_t), _p(sig), _p(dw), _p(s),
            ctypes.c_int64(ns))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return m


def precomp_to_weights(g, m, pix, spin, det_weights=None, segs=None):
    pix, spin = _c(pix, np.int32), _c(spin, np.float32)
    n_det, n_t, k = pix.shape
    nc = spin.shape[2]
    dw = None if det_weights is None else _c(det_weights, np.float32)
    s, ns = _segs(segs)
    mp, tp, keep = _map_args(m)
    f = lib().oracle_precomp_to_weights
    f.restype = ctypes.c_int
    err = f(ctypes.byref(g), mp, tp, _p(pix), ctypes.c_int(k), _p(spin), ctypes.c_int(nc),
            ctypes.c_int64(n_det), ctypes.c_int64(n_t), _p(dw), _p(s), ctypes.c_int64(ns))
    if err:
        raise RuntimeError('oracle: inactive tile hit')
    return m


# -- primitives ---------------------------------------------------------------

def tables():
    a = np.zeros(1 << 14)
    b = np.zeros(1 << 14)
    lib().oracle_tables(_p(a), _p(b))
    return a, b


def asin_lookup(x):
    x = _c(x, np.float64)
    y = np.zeros_like(x)
    lib().oracle_asin_lookup(_p(x), _p(y), ctypes.c_int64(x.size))
    return y


def atan2_lookup(y, x):
    y, x = np.broadcast_arrays(_c(y, np.float64), _c(x, np.float64))
    y, x = _c(y, np.float64), _c(x, np.float64)
    out = np.zeros_like(x)
    lib().oracle_atan2_lookup(_p(y), _p(x), _p(out), ctypes.c_int64(x.size))
    return out


def ang2pix(nside, nest, theta, phi):
    theta, phi = _c(theta, np.float64), _c(phi, np.float64)
    out = np.zeros(theta.shape, np.int64)
    lib().oracle_ang2pix(ctypes.c_int64(nside), ctypes.c_int(int(nest)), _p(theta), _p(phi),
                         _p(out), ctypes.c_int64(theta.size))
    return out


def qmul(l, r):
    l, r = np.broadcast_arrays(_c(l, np.float64), _c(r, np.float64))
    l, r = _c(l, np.float64), _c(r, np.float64)
    out = np.zeros_like(l)
    lib().oracle_qmul(_p(l), _p(r), _p(out), ctypes.c_int64(l.size // 4))
    return out


def num_threads():
    return lib().oracle_num_threads()


# -- ML cut handling (cuts_oracle.cpp) -----------------------------------------

CUT_OPS = {'insert': 0, 'extract': 1, 'clear': 2}
CUT_MODELS = {'full': 0, 'poly': 1}


def cut_segments(ranges):
    """(n, 3) int32 (det, start, end) of a list of Ranges objects, in the
    order the reference walks them."""
    segs = []
    for d, r in enumerate(ranges):
        for a, b in np.asarray(r.ranges()).reshape(-1, 2):
            segs.append((d, a, b))
    return np.asarray(segs, np.int32).reshape(-1, 3)


def cuts_measure(segs, model, resolution=0, nmax=0):
    f = lib().oracle_cuts_measure
    f.restype = ctypes.c_int64
    s = _c(segs, np.int32)
    return int(f(_p(s), ctypes.c_int64(len(s)), ctypes.c_int(CUT_MODELS[model]),
                 ctypes.c_int(resolution), ctypes.c_int(nmax)))


def process_cuts(segs, operation, model, tod, vals, resolution=0, nmax=0):
    """In place on tod (insert, clear) or vals (extract); float32 or float64."""
    s = _c(segs, np.int32)
    assert tod.flags['C_CONTIGUOUS'] and vals.flags['C_CONTIGUOUS']
    lib().oracle_process_cuts(_p(s), ctypes.c_int64(len(s)), ctypes.c_int(CUT_OPS[operation]),
                              ctypes.c_int(CUT_MODELS[model]), ctypes.c_int(resolution),
                              ctypes.c_int(nmax), _p(tod), ctypes.c_int64(tod.shape[1]),
                              _p(vals), ctypes.c_int(int(tod.dtype == np.float64)))


def translate_cuts(isegs, osegs, model, insamp, onsamp, ivals, ovals, resolution=0, nmax=0):
    i, o = _c(isegs, np.int32), _c(osegs, np.int32)
    lib().oracle_translate_cuts(_p(i), _p(o), ctypes.c_int64(len(i)),
                                ctypes.c_int(CUT_MODELS[model]), ctypes.c_int(resolution),
                                ctypes.c_int(nmax), ctypes.c_int64(insamp),
                                ctypes.c_int64(onsamp), _p(ivals), _p(ovals),
                                ctypes.c_int(int(ivals.dtype == np.float64)))


# -- G3SuperTimestream decoding (g3super_oracle.cpp) ---------------------------

G3S_TYPES = {'int32': 0, 'int64': 1, 'float32': 2, 'float64': 3}
G3S_FLAC, G3S_BZ, G3S_CONST = 1, 2, 4


def crc8(data):
    b = np.frombuffer(bytes(data), np.uint8)
    f = lib().oracle_crc8
    f.restype = ctypes.c_uint8
    return int(f(_p(b), ctypes.c_int64(len(b))))


def crc16(data):
    b = np.frombuffer(bytes(data), np.uint8)
    f = lib().oracle_crc16
    f.restype = ctypes.c_uint16
    return int(f(_p(b), ctypes.c_int64(len(b))))


def flac_decode(stream, cap):
    """int32 samples of one mono FLAC stream (bytes); raises ValueError with
    the oracle's status on a malformed stream."""
    b = np.frombuffer(bytes(stream), np.uint8)
    out = np.zeros(max(int(cap), 1), np.int32)
    f = lib().oracle_flac_decode
    f.restype = ctypes.c_int64
    n = int(f(_p(b), ctypes.c_int64(len(b)), _p(out), ctypes.c_int64(cap)))
    if n < 0:
        raise ValueError(f'FLAC stream rejected by the oracle (status {n})')
    return out[:n]


def g3s_split(data, quanta=None):
    """(y int32, r int64, branches int32) per channel: the quantised data
    (float types) split into a 24-bit part and a multiple of 2^24, as
    G3SuperTimestream.cxx:72-92, 218-231."""
    data = np.ascontiguousarray(data)
    n_det, n = data.shape
    q = np.ascontiguousarray(np.ones(n_det) if quanta is None else quanta, np.float64)
    y = np.zeros((n_det, n), np.int32)
    r = np.zeros((n_det, n), np.int64)
    br = np.zeros(n_det, np.int32)
    lib().oracle_g3s_split(_p(data), ctypes.c_int(G3S_TYPES[data.dtype.name]),
                           ctypes.c_int64(n_det), ctypes.c_int64(n), _p(q), _p(y), _p(r),
                           _p(br))
    return y, r, br


def g3s_extract(payload, offsets, dtype, n_samp, quanta, dest, dest_indices=None,
                src_indices=None, start=0, stop=None):
    """Restatement of G3SuperTimestream::Extract (G3SuperTimestream.cxx:664-812)
    over an encoded array blob: per extracted channel the algo byte, FLAC
    block (oracle decoder), bz2 offsets (Python's bz2 = libbz2), CONST datum,
    then the float conversion.  Writes dest in place (numpy)."""
    import bz2
    dtype = np.dtype(dtype)
    payload = bytes(payload)
    n_det = len(offsets) - 1
    stop = n_samp if stop is None or stop < 0 else stop
    wide = dtype.itemsize == 8
    itype = np.int64 if wide else np.int32
    if src_indices is not None:
        n_ex = len(src_indices)
    else:
        n_ex = n_det
    t = G3S_TYPES[dtype.name]
    for i in range(n_ex):
        si = int(src_indices[i]) if src_indices is not None else i
        if si < 0 or si >= n_det:
            continue
        di = int(dest_indices[i]) if dest_indices is not None else i
        if di < 0 or di >= dest.shape[0]:
            continue
        at = offsets[si]
        algo = np.frombuffer(payload[at:at + 1], np.int8)[0]
        at += 1
        flac = raw = offs = cval = None
        if algo == 0:
            raw = np.frombuffer(payload[at:at + n_samp * dtype.itemsize], itype).copy()
        if algo & G3S_FLAC:
            n = int(np.frombuffer(payload[at:at + 4], np.int32)[0])
            flac = flac_decode(payload[at + 4:at + 4 + n], n_samp)
            at += 4 + n
        if algo & G3S_BZ:
            n = int(np.frombuffer(payload[at:at + 4], np.int32)[0])
            offs = np.frombuffer(bz2.decompress(payload[at + 4:at + 4 + n]), itype).copy()
            at += 4 + n
        if algo & G3S_CONST:
            cval = np.frombuffer(payload[at:at + dtype.itemsize], itype).copy()
        row = np.ascontiguousarray(dest[di])
        lib().oracle_g3s_combine(ctypes.c_int(t), ctypes.c_int64(start), ctypes.c_int64(stop),
                                 _p(flac), _p(raw), _p(offs), _p(cval),
                                 ctypes.c_double(float(quanta[si]) if quanta is not None
                                                 else 1.0), _p(row))
        dest[di] = row
    return dest
