"""G3SuperTimestream: compressed detector timestreams, decoded on the GPU.

Drop-in for the array half of so3g's G3SuperTimestream
(src/G3SuperTimestream.cxx, include/G3SuperTimestream.h): the same
properties (names, times, data, quanta, dtype), the same methods (encode,
decode, extract, calibrate, options) with the same argument meanings and
ValueError behaviour, and the same array blob (per channel an algo byte,
then a FLAC block of the 24-bit part, a bz2 block of full-width offsets or
a CONST datum, or the raw data; G3SuperTimestream.cxx:136-181).

decode() and extract() run on the GPU (so3g_g3s_decode: frame-parallel FLAC
decoding, then the offset and quantum arithmetic in the same kernels); a
torch CUDA tensor passed to extract() is filled in place, so decoded data
need not return to the host.  The host walks the per-channel blocks and
inflates the rare bz2 blocks (Python's bz2 module is libbz2, the
reference's compressor).  encode() runs on the host (the library's FLAC
encoder and quantise/rebranch step), as the reference's does.

Serialisation into G3 frames (spt3g's archive format, the times vector's
bz2 packing) is outside this module: the blob is available as
``payload`` / ``offsets`` and a blob made elsewhere is attached with
``from_blob``.
"""
import bz2
import concurrent.futures
import ctypes
import os

import numpy as np

from . import _native as N

ALGO_NONE, ALGO_DO_FLAC, ALGO_DO_BZ, ALGO_DO_CONST = 0, 1, 2, 4
TYPE_CODE = {'int32': 0, 'int64': 1, 'float32': 2, 'float64': 3}
BZ2_BLOCKSIZE = 5   # blockSize100k of the reference's bz2 calls (SO3G_BZ2_BLOCKSIZE)


class Chan(ctypes.Structure):
    """so3g_g3s_chan of include/so3g_hip.h."""
    _fields_ = [('flac_pos', ctypes.c_int64), ('flac_end', ctypes.c_int64),
                ('raw_pos', ctypes.c_int64), ('off_row', ctypes.c_int64),
                ('const_bits', ctypes.c_int64), ('quantum', ctypes.c_double),
                ('dest_row', ctypes.c_int64), ('block', ctypes.c_int32),
                ('bps', ctypes.c_int32), ('has_const', ctypes.c_int32),
                ('pad_', ctypes.c_int32)]


CHAN_DTYPE = np.dtype([(name, np.int64 if ct is ctypes.c_int64 else np.float64
                         if ct is ctypes.c_double else np.int32) for name, ct in Chan._fields_])
assert CHAN_DTYPE.itemsize == ctypes.sizeof(Chan)


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('G3SuperTimestream decoding needs a ROCm GPU (torch.cuda is '
                           'unavailable); there is no CPU fallback.')
    N.init()
    return torch


def _stream(torch):
    s = torch.cuda.current_stream().cuda_stream
    return None if s in (0, None) else s


def flac_encode(x, level=1, bps=24, capacity=-1):
    """FLAC stream (bytes) of int32 samples x, or None if it would exceed
    capacity bytes (capacity < 0: unbounded)."""
    x = np.ascontiguousarray(x, np.int32)
    lib = N.load()
    nb = ctypes.c_int64()
    N.check(lib.so3g_flac_encode(x.ctypes.data, len(x), bps, level, None, -1,
                                 ctypes.byref(nb)))
    if 0 <= capacity < nb.value:
        return None
    out = np.zeros(max(nb.value, 1), np.uint8)
    N.check(lib.so3g_flac_encode(x.ctypes.data, len(x), bps, level, out.ctypes.data,
                                 nb.value, ctypes.byref(nb)))
    return out[:nb.value].tobytes()


def split(data, quanta=None):
    """(y int32, r int64, changes int32): quantised data split into a 24-bit
    part and a slowly varying multiple of 2^24 (G3SuperTimestream.cxx:72-92,
    218-231)."""
    data = np.ascontiguousarray(data)
    n_det, n = data.shape
    q = None if quanta is None else np.ascontiguousarray(quanta, np.float64)
    y = np.empty((n_det, n), np.int32)
    r = np.empty((n_det, n), np.int64)
    ch = np.empty(n_det, np.int32)
    N.check(N.load().so3g_g3s_split(data.ctypes.data, TYPE_CODE[data.dtype.name], n_det, n,
                                    None if q is None else q.ctypes.data, y.ctypes.data,
                                    r.ctypes.data, ch.ctypes.data))
    return y, r, ch


def last_decode_stats():
    """{'frames': FLAC frames decoded frame-parallel, 'serial': channels
    decoded frame after frame, 'why': reason bits (1 variable block size,
    2 missing sync, 4 malformed frame, 8 false sync, 16 unexpected header)}
    of the last decode / extract call."""
    a = np.zeros(3, np.int64)
    N.check(N.load().so3g_g3s_stats(a.ctypes.data))
    return {'frames': int(a[0]), 'serial': int(a[1]), 'why': int(a[2])}


def streaminfo(stream):
    """(first frame offset, max block size, bits per sample) of a FLAC
    stream held in a bytes-like object, or None if it is not one."""
    if len(stream) < 8 or bytes(stream[:4]) != b'fLaC':
        return None
    at, block, bps = 4, None, None
    while True:
        if at + 4 > len(stream):
            return None
        last, btype = stream[at] >> 7, stream[at] & 0x7F
        blen = int.from_bytes(bytes(stream[at + 1:at + 4]), 'big')
        if btype == 0 and blen >= 34:
            b = bytes(stream[at + 4:at + 4 + 34])
            block = int.from_bytes(b[2:4], 'big')
            bps = ((b[12] & 0x01) << 4 | b[13] >> 4) + 1
        at += 4 + blen
        if last:
            break
    if block is None:
        return None
    return at, block, bps


class G3SuperTimestream:
    """names, times, data (2-d int32/int64/float32/float64), quanta (float
    mode); encode/decode/extract/calibrate/options as the reference's."""

    def __init__(self, names=None, times=None, data=None, quanta=None):
        self._names = []
        self._times = np.zeros(0, np.int64)
        self._quanta = []
        self._float_mode = False
        self._dataful = False
        self._array = None
        self._blob = None            # (payload bytes, offsets list)
        self._dev_payload = None     # device copy of the payload (torch uint8)
        self._table = None           # parsed channel table of the blob
        self._offset_rows = []       # inflated bz2 offsets, rows of the table
        self._dev_offsets = None
        self._dtype = None
        self._shape = None
        self._opts = {}
        self.options(1, 1, 30, -1, -1)
        if names is not None:
            self.names = names
        if times is not None:
            self.times = times
        if quanta is not None:
            self.quanta = quanta
        if data is not None:
            self.data = data

    # -- properties (G3SuperTimestream.cxx:961-1110) ----------------------
    @property
    def names(self):
        return list(self._names)

    @names.setter
    def names(self, v):
        v = list(v)
        if len(v) != len(self._names) and len(self._names) != 0:
            raise ValueError('Cannot set .names because it conflicts with the established '
                             f'number of channels ({len(self._names)}).')
        self._names = v

    @property
    def times(self):
        return self._times

    @times.setter
    def times(self, v):
        v = np.asarray(v)
        if len(v) != len(self._times) and len(self._times) != 0:
            raise ValueError('Cannot set .times because it conflicts with the established '
                             f'number of samples ({len(self._times)}).')
        self._times = v

    @property
    def data(self):
        if self._array is None:
            self.decode()
        return self._array

    @data.setter
    def data(self, obj):
        a = np.asarray(obj)
        if a.ndim != 2:
            raise ValueError('Bad ndim.')
        if a.shape[0] != len(self._names):
            raise ValueError('Bad shape[0].')
        if a.shape[1] != len(self._times):
            raise ValueError('Bad shape[1].')
        if a.dtype.byteorder == '>' or (a.dtype.byteorder == '=' and
                                        np.little_endian is False):
            raise ValueError('Bad endianness.')
        if a.dtype.name not in TYPE_CODE:
            raise ValueError('Forbidden dtype.')
        floaty = a.dtype.kind == 'f'
        if floaty:
            if len(self._quanta) != a.shape[0]:
                raise ValueError('User must set .quanta before loading float array.')
        elif len(self._quanta) != 0:
            raise ValueError('The .quanta must be empty when loading integer array.')
        self._array = a
        self._blob = None
        self._dev_payload = None
        self._table = None
        self._offset_rows = []
        self._dev_offsets = None
        self._dataful = True
        self._float_mode = floaty
        self._dtype = a.dtype
        self._shape = a.shape

    @property
    def quanta(self):
        if not self._float_mode:
            return None
        return np.array(self._quanta, np.float64)

    @quanta.setter
    def quanta(self, v):
        if self._dataful:
            raise ValueError('The .quanta cannot be set directly once .data is set.  '
                             'Use .calibrate().')
        self.calibrate(v)

    @property
    def dtype(self):
        return self._array.dtype if self._array is not None else self._dtype

    @property
    def payload(self):
        """The encoded blob's bytes (None unless encoded)."""
        return None if self._blob is None else self._blob[0]

    @property
    def offsets(self):
        """Start of each channel's block in the payload, and its end."""
        return None if self._blob is None else list(self._blob[1])

    # -- options / calibrate (:817-913) ------------------------------------
    def options(self, enable=-1, flac_level=-1, bz2_workFactor=-1, data_algo=-1,
                times_algo=-1):
        if enable == 0:
            self._opts['data_algo'] = 0
            self._opts['times_algo'] = 0
        elif enable > 0:
            self._opts['data_algo'] = ALGO_DO_FLAC | ALGO_DO_BZ
            self._opts['times_algo'] = ALGO_DO_BZ
        if data_algo >= 0:
            self._opts['data_algo'] = data_algo
        if times_algo >= 0:
            self._opts['times_algo'] = times_algo
        if flac_level >= 0:
            self._opts['flac_level'] = flac_level
        if bz2_workFactor >= 0:
            self._opts['bz2_workFactor'] = bz2_workFactor
        return 0

    def calibrate(self, rescale):
        rescale = [float(x) for x in rescale]
        if len(rescale) != len(self._names):
            raise ValueError('Rescale vector has unexpected length.')
        cal = np.asarray(rescale, np.float64)[:, None]
        if self._float_mode:
            if self._array is not None:
                a = self._array
                a[...] = (a.astype(np.float64) * cal).astype(a.dtype)
            self._quanta = [q * c for q, c in zip(self._quanta, rescale)]
        else:
            if self._dataful:
                ftype = {'int32': np.float32, 'int64': np.float64}.get(self._dtype.name)
                if ftype is None:
                    raise ValueError('Unexpected dtype!')
                self._dtype = np.dtype(ftype)
                if self._array is not None:
                    a = self._array.astype(ftype)
                    a[...] = (a.astype(np.float64) * cal).astype(ftype)
                    self._array = a
            self._float_mode = True
            self._quanta = rescale

    # -- encode (:122-314, :522-535) -----------------------------------------
    def encode(self):
        if self._array is None or self._opts['data_algo'] == ALGO_NONE:
            return False
        self._blob = self._encode_array(self._array)
        self._array = None
        self._dev_payload = None
        self._table = None
        self._offset_rows = []
        self._dev_offsets = None
        return True

    def _encode_array(self, a):
        a = np.ascontiguousarray(a)
        n_det, n = a.shape
        item = a.dtype.itemsize
        wide = item == 8
        itype = np.int64 if wide else np.int32
        quanta = np.asarray(self._quanta, np.float64) if self._float_mode else None
        algo_opt = self._opts['data_algo']
        size = a.nbytes + n_det          # the reference's buffer capacity
        buf = bytearray()
        offsets = [0]
        y = r = changes = None
        if algo_opt & ALGO_DO_FLAC or self._float_mode:
            y, r, changes = split(a, quanta)
        streams = [None] * n_det
        if algo_opt & ALGO_DO_FLAC:
            # Channels encode independently (the library call releases the
            # GIL); a stream longer than the space left when its turn comes
            # fails as the reference's write callback does.
            level = self._opts['flac_level']
            with concurrent.futures.ThreadPoolExecutor(_workers()) as ex:
                streams = list(ex.map(lambda row: flac_encode(row, level, 24), y))
        for i in range(n_det):
            block_start = len(buf)
            buf.append(0)
            block_limit = len(buf) + n * item
            try_flac = bool(algo_opt & ALGO_DO_FLAC)
            try_bz = bool(algo_opt & ALGO_DO_BZ)
            try_const = False
            algo = 0
            ok = True
            rr = None
            if try_flac:
                if size - len(buf) < 4:
                    ok = False
                else:
                    stream = streams[i]
                    if len(stream) > size - len(buf) - 4:
                        del buf[block_start + 1:]
                    else:
                        buf += np.int32(len(stream)).tobytes() + stream
                        algo |= ALGO_DO_FLAC
                        rr = r[i].astype(itype)
                        if changes[i] <= 1:
                            try_bz = False
                            try_const = True
            # The full-width integers: the data, or the quantised data (r + y
            # with the integer type's wrap-around).
            full = (r[i] + y[i]).astype(itype) if self._float_mode else a[i].astype(itype)
            if not (algo & ALGO_DO_FLAC):
                rr = full
            if ok and try_const:
                n_copy = item if (rr[0] != 0 or algo == 0) else 0
                if len(buf) + n_copy < block_limit:
                    if n_copy:
                        buf += rr[:1].tobytes()
                        algo |= ALGO_DO_CONST
                else:
                    ok = False
            if ok and try_bz:
                if size - len(buf) < 4:
                    ok = False
                else:
                    algo |= ALGO_DO_BZ
                    at = len(buf)
                    buf += b'\0\0\0\0'
                    if len(buf) < block_limit:
                        z = bz2.compress(rr.tobytes(), BZ2_BLOCKSIZE)
                        if len(z) > block_limit - len(buf):
                            ok = False
                        else:
                            buf += z
                            buf[at:at + 4] = np.int32(len(z)).tobytes()
                    else:
                        ok = False
            if ok and len(buf) < block_limit:
                buf[block_start] = algo
            else:
                # Raw block.  The decoder reads it as full-width integers
                # times the quantum, so float channels store their quantised
                # integers (the reference copies the float bits here, which
                # its own decoder then misreads).
                del buf[block_start:]
                buf.append(ALGO_NONE)
                buf += full.tobytes()
            offsets.append(len(buf))
        return bytes(buf), offsets

    def from_blob(self, payload, offsets, dtype, quanta=None):
        """Attach an encoded blob (payload bytes, n_det + 1 offsets) of the
        given dtype, as loaded from a frame; names and times must be set."""
        dtype = np.dtype(dtype)
        if dtype.name not in TYPE_CODE:
            raise ValueError('Forbidden dtype.')
        if len(offsets) != len(self._names) + 1:
            raise ValueError('offsets must hold n_det + 1 entries')
        self._float_mode = dtype.kind == 'f'
        if self._float_mode:
            if quanta is None or len(quanta) != len(self._names):
                raise ValueError('float data need one quantum per channel')
            self._quanta = [float(q) for q in quanta]
        self._dtype = dtype
        self._shape = (len(self._names), len(self._times))
        self._dataful = True
        self._array = None
        self._blob = (bytes(payload), [int(o) for o in offsets])
        self._dev_payload = None
        self._table = None
        self._offset_rows = []
        self._dev_offsets = None

    # -- decode / extract (:636-812) -----------------------------------------
    def decode(self):
        if self._blob is None:
            return False
        if self._opts['data_algo'] == ALGO_NONE:
            raise ValueError('Decode called with ablob buffer but data_algo=0.')
        out = np.zeros(self._shape, self._dtype)
        self.extract(out, None, None, 0, self._shape[1])
        self._array = out
        self._blob = None
        self._dev_payload = None
        self._table = None
        self._offset_rows = []
        self._dev_offsets = None
        return True

    def extract(self, dest, dest_indices=None, src_indices=None, start=0, stop=-1):
        n_det, n_samp = self._shape if self._shape is not None else (0, 0)
        if src_indices is not None:
            src_indices = np.asarray(src_indices, np.int64)
            if src_indices.ndim != 1:
                raise ValueError('src_indices must be 1-d')
            n_ex = len(src_indices)
        else:
            n_ex = n_det
        if dest_indices is not None:
            dest_indices = np.asarray(dest_indices, np.int64)
            if dest_indices.shape != (n_ex,):
                raise ValueError(f'dest_indices must have shape ({n_ex},)')
        if start < 0 or start > n_samp:
            raise ValueError('sample start index out of bounds')
        if stop < 0:
            stop = n_samp
        if stop < start or stop > n_samp:
            raise ValueError('sample stop index out of bounds')
        m = stop - start
        is_torch = type(dest).__module__.startswith('torch')
        if is_torch:
            torch = _torch()
            if not dest.is_cuda:
                raise ValueError('Destination tensor must be on the GPU.')
            if str(dest.dtype).replace('torch.', '') != self._dtype.name:
                raise ValueError('Destination array not of correct type.')
            shape, contiguous_last = tuple(dest.shape), dest.ndim == 2 and dest.stride(1) == 1
        else:
            if not isinstance(dest, np.ndarray):
                raise ValueError('Destination array must be ndarray.')
            if dest.dtype != self._dtype:
                raise ValueError('Destination array not of correct type.')
            shape = dest.shape
            contiguous_last = dest.ndim == 2 and dest.strides[1] == dest.itemsize
        if len(shape) != 2 or shape[0] < n_ex or shape[1] != m:
            raise ValueError('Destination array does not have correct shape.')
        if not contiguous_last:
            raise ValueError('Destination array should be strictly packed on last dimension.')
        if self._blob is None:
            raise ValueError('These data are already fully decoded; use .data.')
        torch = _torch()
        table = self._channel_table()
        si = np.arange(n_ex) if src_indices is None else src_indices
        di = np.arange(n_ex) if dest_indices is None else dest_indices
        keep = (si >= 0) & (si < n_det) & (di >= 0) & (di < shape[0])
        chans = table[si[keep]]
        chans['dest_row'] = di[keep]
        if len(chans) == 0:
            return True
        payload = self._device_payload(torch)
        d_offs = self._device_offsets(torch)
        if is_torch:
            d_dest = dest
            stride = dest.stride(0)
        else:
            tdt = getattr(torch, self._dtype.name)
            d_dest = torch.from_numpy(np.ascontiguousarray(dest)).to(payload.device, tdt)
            stride = m
        chans = np.ascontiguousarray(chans)
        N.check(N.load().so3g_g3s_decode(
            payload.data_ptr(), self._payload_len, chans.ctypes.data, len(chans),
            TYPE_CODE[self._dtype.name], n_samp, start, stop,
            None if d_offs is None else d_offs.data_ptr(), len(self._offset_rows),
            d_dest.data_ptr(), shape[0], stride, _stream(torch)))
        if not is_torch:
            dest[...] = d_dest.cpu().numpy()
        return True

    # -- helpers -------------------------------------------------------------
    def _device_payload(self, torch):
        if self._dev_payload is None:
            p = np.frombuffer(self._blob[0], np.uint8)
            t = torch.zeros(len(p) + 16, dtype=torch.uint8, device='cuda')
            if len(p):
                t[:len(p)] = torch.from_numpy(p.copy()).to('cuda')
            self._dev_payload = t
            self._payload_len = len(p)
        return self._dev_payload

    def _device_offsets(self, torch):
        if self._dev_offsets is None and self._offset_rows:
            self._dev_offsets = torch.from_numpy(np.stack(self._offset_rows)).to('cuda')
        return self._dev_offsets

    def _channel_table(self):
        """so3g_g3s_chan entries of every channel of the blob (dest_row
        unset), parsed once per blob: the algo byte and block lengths, each
        FLAC stream's STREAMINFO, the inflated bz2 offset rows
        (self._offset_rows) and CONST data."""
        if self._table is not None:
            return self._table
        payload, offsets = self._blob
        n_det, n_samp = self._shape
        item = self._dtype.itemsize
        itype = np.int64 if item == 8 else np.int32
        t = np.zeros(n_det, CHAN_DTYPE)
        t['flac_pos'] = t['flac_end'] = t['raw_pos'] = t['off_row'] = -1
        t['quantum'] = np.asarray(self._quanta, np.float64) if self._float_mode else 1.0
        rows = []
        mv = memoryview(payload)
        for si in range(n_det):
            at = offsets[si]
            algo = payload[at] if payload[at] < 128 else payload[at] - 256
            at += 1
            if algo == ALGO_NONE:
                t['raw_pos'][si] = at
            if algo & ALGO_DO_FLAC:
                n = int.from_bytes(payload[at:at + 4], 'little', signed=True)
                info = streaminfo(mv[at + 4:at + 4 + min(n, 1 << 16)])
                if info is None:
                    raise ValueError(f'channel {si}: FLAC block without a stream header')
                t['flac_pos'][si] = at + 4 + info[0]
                t['flac_end'][si] = at + 4 + n
                t['block'][si], t['bps'][si] = info[1], info[2]
                at += 4 + n
            if algo & ALGO_DO_BZ:
                n = int.from_bytes(payload[at:at + 4], 'little', signed=True)
                row = np.frombuffer(bz2.decompress(payload[at + 4:at + 4 + n]), itype)
                if len(row) < n_samp:
                    raise ValueError(f'channel {si}: bz2 block shorter than the channel')
                t['off_row'][si] = len(rows)
                rows.append(row[:n_samp].copy())
                at += 4 + n
            if algo & ALGO_DO_CONST:
                t['const_bits'][si] = int(np.frombuffer(payload, itype, 1, at)[0])
                t['has_const'][si] = 1
        self._table = t
        self._offset_rows = rows
        return t
