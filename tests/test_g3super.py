"""G3SuperTimestream: oracle pins (CPU), encoder round trips (CPU), GPU
decoding against the oracle, and the reference's own test cases
(test/test_g3super.py) against the GPU path."""
import numpy as np
import pytest

import oracle

g3 = pytest.importorskip('so3g_amd.g3super')


# -- a FLAC bit writer for hand-assembled streams (RFC 9639 field layout) -----

class Bits:
    def __init__(self):
        self.bits = []

    def put(self, v, n):
        for i in range(n - 1, -1, -1):
            self.bits.append((int(v) >> i) & 1)

    def sput(self, v, n):
        self.put(int(v) & ((1 << n) - 1), n)

    def unary(self, q):
        self.bits += [0] * q + [1]

    def rice(self, v, k):
        u = 2 * v if v >= 0 else -2 * v - 1
        self.unary(u >> k)
        self.put(u & ((1 << k) - 1), k)

    def align(self):
        while len(self.bits) % 8:
            self.bits.append(0)

    def tobytes(self):
        assert len(self.bits) % 8 == 0
        return bytes(int(''.join(map(str, self.bits[i:i + 8])), 2)
                     for i in range(0, len(self.bits), 8))


def stream_header(block, bps, total):
    b = Bits()
    for ch in b'fLaC':
        b.put(ch, 8)
    b.put(1, 1); b.put(0, 7); b.put(34, 24)
    b.put(block, 16); b.put(block, 16); b.put(0, 24); b.put(0, 24)
    b.put(44100, 20); b.put(0, 3); b.put(bps - 1, 5); b.put(total, 36)
    b.put(0, 128)
    return b.tobytes()


def frame(number, block, bps, body, variable=False):
    """One frame: header (block size as an 8- or 16-bit field), the
    subframe written by body(Bits), padding, CRC-16."""
    h = Bits()
    h.put(0x3FFE, 14); h.put(0, 1); h.put(1 if variable else 0, 1)
    h.put(6 if block <= 256 else 7, 4)
    h.put(9, 4); h.put(0, 4)
    h.put({8: 1, 12: 2, 16: 4, 20: 5, 24: 6, 32: 7}[bps], 3); h.put(0, 1)
    assert number < 128
    h.put(number, 8)
    h.put(block - 1, 8 if block <= 256 else 16)
    hb = h.tobytes()
    h.put(oracle.crc8(hb), 8)
    body(h)
    h.align()
    fb = h.tobytes()
    return fb + oracle.crc16(fb).to_bytes(2, 'big')


def sub_header(b, type_code, wasted=0):
    b.put(0, 1); b.put(type_code, 6)
    if wasted:
        b.put(1, 1); b.unary(wasted - 1)
    else:
        b.put(0, 1)


def residual(b, res, k_list, porder=0, method=0, escape_bits=None):
    """Residual section: res holds the 2^porder partitions' values."""
    b.put(method, 2); b.put(porder, 4)
    for p in range(1 << porder):
        if escape_bits is not None and escape_bits[p] is not None:
            b.put(31 if method else 15, 5 if method else 4)
            b.put(escape_bits[p], 5)
            for v in res[p]:
                b.sput(v, escape_bits[p])
        else:
            b.put(k_list[p], 5 if method else 4)
            for v in res[p]:
                b.rice(v, k_list[p])


# -- oracle pins ---------------------------------------------------------------

def test_crc_catalogue_values():
    """FLAC's CRC-8 (poly 0x07) and CRC-16 (poly 0x8005), initial 0, are the
    catalogue's CRC-8/SMBUS and CRC-16/UMTS: check values over "123456789"
    0xF4 and 0xFEE8."""
    assert oracle.crc8(b'123456789') == 0xF4
    assert oracle.crc16(b'123456789') == 0xFEE8


def test_oracle_handmade_subframes():
    """Streams assembled field by field: CONSTANT, VERBATIM, FIXED order 2
    with two Rice partitions and an escape partition, LPC order 2, wasted
    bits; a short last frame."""
    rng = np.random.default_rng(3)
    frames, want = [], []
    # CONSTANT, 20 samples of -5, 16 bits
    frames.append(frame(0, 20, 16, lambda b: (sub_header(b, 0), b.sput(-5, 16))))
    want += [-5] * 20
    # VERBATIM
    v = rng.integers(-30000, 30000, 20)
    frames.append(frame(1, 20, 16, lambda b: (sub_header(b, 1), [b.sput(x, 16) for x in v])))
    want += list(v)
    # FIXED order 2: x = warm-up then x[i] = r[i] + 2 x[i-1] - x[i-2]
    r = rng.integers(-40, 40, 20)
    x = [100, 130]
    for i in range(2, 20):
        x.append(int(r[i]) + 2 * x[-1] - x[-2])

    def fixed2(b):
        sub_header(b, 8 + 2)
        b.sput(x[0], 16); b.sput(x[1], 16)
        residual(b, [list(r[2:10]), list(r[10:20])], [3, 0], porder=1,
                 escape_bits=[None, 8])
    frames.append(frame(2, 20, 16, fixed2))
    want += x
    # LPC order 2, precision 5, shift 2, coefficients (7, -3): x[i] = r[i] +
    # ((7 x[i-1] - 3 x[i-2]) >> 2)
    r = rng.integers(-20, 20, 20)
    x = [11, -4]
    for i in range(2, 20):
        x.append(int(r[i]) + ((7 * x[-1] - 3 * x[-2]) >> 2))

    def lpc2(b):
        sub_header(b, 32 + 1)
        b.sput(x[0], 16); b.sput(x[1], 16)
        b.put(5 - 1, 4); b.sput(2, 5); b.sput(7, 5); b.sput(-3, 5)
        residual(b, [list(r[2:])], [14], porder=0, method=1)
    frames.append(frame(3, 20, 16, lpc2))
    want += x
    # Wasted bits: VERBATIM of 16-bit samples that are multiples of 8,
    # coded as 13-bit values; a short last frame of 7 samples.
    w = rng.integers(-1000, 1000, 7) * 8
    frames.append(frame(4, 7, 16, lambda b: (sub_header(b, 1, wasted=3),
                                             [b.sput(q // 8, 13) for q in w])))
    want += list(w)
    s = stream_header(20, 16, len(want)) + b''.join(frames)
    got = oracle.flac_decode(s, len(want))
    np.testing.assert_array_equal(got, np.array(want, np.int32))
    # A flipped bit breaks the frame CRC.
    bad = bytearray(s)
    bad[len(stream_header(20, 16, 0)) + 40] ^= 0x10
    with pytest.raises(ValueError):
        oracle.flac_decode(bytes(bad), len(want))


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_oracle_fixed_predictors_continue_polynomials(order):
    """A FIXED order-o subframe with zero residuals continues the unique
    polynomial of degree o - 1 through its warm-up samples."""
    t = np.arange(30)
    coef = [7, -3, 2, 1][:order]
    poly = sum(c * t ** j for j, c in enumerate(coef))

    def body(b):
        sub_header(b, 8 + order)
        for i in range(order):
            b.sput(int(poly[i]), 24)
        residual(b, [[0] * (30 - order)], [0])
    s = stream_header(30, 24, 30) + frame(0, 30, 24, body)
    np.testing.assert_array_equal(oracle.flac_decode(s, 30), poly.astype(np.int32))


# -- encoder (product, host) against the oracle --------------------------------

def _signals(rng):
    n = 5000
    return {
        'white': rng.normal(0, 300, n).astype(np.int32),
        'ramp': (np.arange(n) * 3 - 7000).astype(np.int32),
        'const': np.full(n, -12345, np.int32),
        'one': np.array([5], np.int32),
        'wasted': (rng.integers(-1000, 1000, n) * 16).astype(np.int32),
        'sine': (1e6 * np.sin(np.arange(n) * 0.01) + rng.normal(0, 5, n)).astype(np.int32),
        'full_scale': rng.choice([-(1 << 23), (1 << 23) - 1, 0], n).astype(np.int32),
        'spiky': np.where(rng.random(n) < 0.01, rng.integers(-(1 << 23), 1 << 23, n),
                          rng.integers(-3, 4, n)).astype(np.int32),
    }


@pytest.mark.parametrize('level', [0, 1, 5, 8])
def test_encoder_roundtrip(level):
    for name, x in _signals(np.random.default_rng(level)).items():
        s = g3.flac_encode(x, level)
        np.testing.assert_array_equal(oracle.flac_decode(s, len(x)), x, err_msg=name)


def test_encoder_capacity():
    x = np.random.default_rng(0).integers(-(1 << 23), 1 << 23, 4000).astype(np.int32)
    assert g3.flac_encode(x, 1, capacity=100) is None
    assert g3.flac_encode(x, 1) is not None


def _round_half_away(v):
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def test_split_against_oracle_and_closed_form():
    rng = np.random.default_rng(5)
    n = 3000
    i32 = rng.integers(-(1 << 31), (1 << 31) - 1, (3, n), dtype=np.int64).astype(np.int32)
    i32[0, :10] = [2**31 - 1, 2**31 - 2**23, -2**31, 0, 2**23, -2**23, 2**24, -2**24 - 1, 5, -5]
    i64 = rng.integers(-(1 << 40), 1 << 40, (3, n), dtype=np.int64)
    f64 = rng.normal(0, 1e4, (3, n))
    f64[0, :4] = [0.5, -0.5, 1.5, -2.5]                 # halves round away from zero
    f32 = (rng.normal(0, 1e3, (3, n))).astype(np.float32)
    q = np.array([1.0, 0.01, 3.0])
    for data, quanta in ((i32, None), (i64, None), (f32, q), (f64, q)):
        y, r, ch = g3.split(data, quanta)
        yo, ro, cho = oracle.g3s_split(data, quanta)
        np.testing.assert_array_equal(y, yo)
        np.testing.assert_array_equal(r, ro)
        np.testing.assert_array_equal(ch, cho)
        # Invariants: y in [-2^23, 2^23), r a multiple of 2^24, r + y = the
        # (quantised) value with the type's wrap-around, changes counted.
        assert y.min() >= -(1 << 23) and y.max() < (1 << 23)
        assert np.all(r % (1 << 24) == 0)
        if data.dtype == np.float64:
            v = _round_half_away(data / quanta[:, None]).astype(np.int64)
        elif data.dtype == np.float32:
            v = _round_half_away((data / quanta[:, None]).astype(np.float32)).astype(np.int64)
        else:
            v = data.astype(np.int64)
        wide = data.dtype.itemsize == 8
        total = (r + y) if wide else (r + y).astype(np.int32).astype(np.int64)
        np.testing.assert_array_equal(total, v if wide else v.astype(np.int32))
        np.testing.assert_array_equal(ch, 1 + (np.diff(r, axis=1) != 0).sum(axis=1))


def _ts(n_det, n, sigma=20, dtype='int32', seed=0, offset=0, quanta=None):
    rng = np.random.default_rng(seed)
    d = np.round(rng.normal(0, sigma, (n_det, n)) + offset)
    q = None
    if dtype.startswith('float'):
        q = np.full(n_det, 0.01) if quanta is None else quanta
        d = d * q[:, None]
    d = d.astype(dtype)
    ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n), None, q)
    ts.data = d
    return ts, d


def _oracle_decode(ts):
    n_det, n = ts._shape
    out = np.zeros((n_det, n), ts.dtype)
    return oracle.g3s_extract(ts.payload, ts.offsets, ts.dtype, n, ts.quanta, out)


@pytest.mark.parametrize('dtype', ['int32', 'int64', 'float32', 'float64'])
def test_encode_roundtrip_oracle(dtype):
    """The reference's encode cases (test_10, test_20, test_40 offsets),
    encoded by the library and decoded by the oracle's Extract."""
    offsets = {'int32': [0, 2**25, 2**26 / 3., -1.78 * 2**27],
               'int64': [0, 2**25, 2**26 / 3., -1.78 * 2**27],
               'float32': [0],
               'float64': [0, 2**25, 2**26 / 3., -1.78 * 2**27, 1.8 * 2**35]}[dtype]
    for off in offsets:
        ts, d = _ts(9, 1290, sigma=100, dtype=dtype, offset=int(off))
        ts.encode()
        got = _oracle_decode(ts)
        if dtype.startswith('int'):
            np.testing.assert_array_equal(got, d)
        else:
            np.testing.assert_allclose(got, d, atol=0.01 * 1e-3, rtol=0)
        algos = {ts.payload[o] for o in ts.offsets[:-1]}
        assert algos <= {g3.ALGO_DO_FLAC, g3.ALGO_DO_FLAC | g3.ALGO_DO_CONST,
                         g3.ALGO_DO_FLAC | g3.ALGO_DO_BZ}


def test_encode_fallbacks_oracle():
    """test_11: short segments, random int64 data (raw fallback), one
    sample per channel; bz2 offsets when the branch moves."""
    for n in range(1, 20):
        ts, d = _ts(1, n, sigma=0)
        ts.encode()
        np.testing.assert_array_equal(_oracle_decode(ts), d)
    rng = np.random.default_rng(0)
    for n_det, n in ((1, 200), (7, 1), (19, 1)):
        ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n))
        d = (rng.uniform(size=n_det * n * 8) * 256).astype(np.uint8).view(np.int64)
        ts.data = d.reshape(n_det, n)
        ts.encode()
        np.testing.assert_array_equal(_oracle_decode(ts), ts._blob and d.reshape(n_det, n))
    # A ramp crossing many 2^24 branches: FLAC + bz2 offsets.
    ts = g3.G3SuperTimestream(['a', 'b'], np.arange(4000))
    d = np.stack([np.arange(4000) * 40000, -np.arange(4000) * 12345]).astype(np.int64)
    ts.data = d
    ts.encode()
    assert ts.payload[0] == g3.ALGO_DO_FLAC | g3.ALGO_DO_BZ
    np.testing.assert_array_equal(_oracle_decode(ts), d)
    # float raw fallback stores quantised integers (decodes correctly)
    ts = g3.G3SuperTimestream(['a'], np.arange(300), None, [1e-15])
    d = rng.normal(0, 1, (1, 300))
    ts.data = d
    ts.encode()
    assert ts.payload[0] == g3.ALGO_NONE
    np.testing.assert_allclose(_oracle_decode(ts), d, atol=1e-14, rtol=0)


def test_api_rules_cpu():
    """test_01 / test_02 / test_04 rules that need no decoding."""
    names, times = [f'c{i}' for i in range(5)], np.arange(1000)
    di = np.zeros((5, 1000), np.int32)
    df = np.zeros((5, 1000), np.float32)
    ts = g3.G3SuperTimestream()
    ts.times = times
    ts.names = names
    with pytest.raises(ValueError):
        ts.times = times[:-1]
    with pytest.raises(ValueError):
        ts.names = names[:-1]
    for setup in (lambda t: None, lambda t: setattr(t, 'names', names),
                  lambda t: setattr(t, 'times', times)):
        t = g3.G3SuperTimestream()
        setup(t)
        with pytest.raises(ValueError):
            t.data = di
    t = g3.G3SuperTimestream(names, times)
    with pytest.raises(ValueError):
        t.data = di[1:]
    with pytest.raises(ValueError):
        t.data = di[:, :-1]
    with pytest.raises(ValueError):
        t.data = di.astype('>i4')
    t = g3.G3SuperTimestream(names, times)
    with pytest.raises(ValueError):
        t.data = df                       # float data need quanta first
    t = g3.G3SuperTimestream(names, times)
    t.quanta = np.ones(5)
    t.data = df
    with pytest.raises(ValueError):
        t.data = di                       # no int data in float mode
    t = g3.G3SuperTimestream(names, times)
    with pytest.raises(ValueError):
        t.quanta = np.ones(4)
    t = g3.G3SuperTimestream(names, times)
    t.data = di + 2
    t.calibrate(np.ones(5))
    assert t.data.dtype == np.float32
    t.calibrate(np.ones(5) * 3)
    assert t.quanta[0] == 3.
    np.testing.assert_array_equal(t.data, 6.)
    with pytest.raises(ValueError):
        t.quanta = np.ones(5)             # quanta locked once data are set
    with pytest.raises(ValueError):
        g3.G3SuperTimestream(names, times, df)   # float data without quanta
    assert g3.G3SuperTimestream(names, times, df, np.ones(5)).dtype == np.float32


def test_decode_needs_gpu_or_fails_loudly():
    """Without a GPU decode raises instead of falling back to the CPU."""
    torch = pytest.importorskip('torch')
    if torch.cuda.is_available():
        pytest.skip('GPU present')
    ts, _ = _ts(2, 100)
    ts.encode()
    with pytest.raises(RuntimeError):
        ts.decode()


# -- GPU decoding ---------------------------------------------------------------

def _all_kinds_blob(dtype, n=2500, seed=0):
    """A blob whose channels exercise every block kind: FLAC, FLAC+CONST,
    FLAC+bz2, raw (random int64 / tiny-quantum floats), constant zero."""
    rng = np.random.default_rng(seed)
    n_det = 6
    if dtype.startswith('int'):
        big = 1.78 * 2**27 if dtype == 'int32' else 1.8 * 2**35
        d = np.stack([rng.normal(0, 50, n), rng.normal(0, 50, n) + 2**25,
                      np.arange(n) * 40000.0 - 5e7, rng.normal(0, 50, n) - big,
                      np.zeros(n), rng.normal(0, 1e4, n)])
        d = np.round(d).astype(dtype)
        if dtype == 'int64':
            d[5] = (rng.uniform(size=n * 8) * 256).astype(np.uint8).view(np.int64)
        else:
            d[5] = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
        ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n))
    else:
        q = np.array([0.01, 0.01, 0.5, 0.01, 1.0, 1e-15])
        d = np.stack([rng.normal(0, 0.5, n), rng.normal(0, 0.5, n) + 2**25 * 0.01,
                      np.arange(n) * 20000.0, rng.normal(0, 0.5, n) - 3e6,
                      np.zeros(n), rng.normal(0, 1, n)]).astype(dtype)
        if dtype == 'float32':
            d[1] = rng.normal(0, 0.5, n)
        ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n), None, q)
    ts.data = d
    ts.encode()
    if ts.dtype.kind == 'f':
        # Realistic float data always compress; append a raw block (quantised
        # integers, read back times the quantum) as a blob made elsewhere.
        item = ts.dtype.itemsize
        ints = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(f'i{item}')
        payload = ts.payload + bytes([g3.ALGO_NONE]) + ints.tobytes()
        t2 = g3.G3SuperTimestream([f'c{i}' for i in range(n_det + 1)], np.arange(n))
        t2.from_blob(payload, ts.offsets + [len(payload)], dtype, list(q) + [0.25])
        ts = t2
    return ts


def _gpu():
    torch = pytest.importorskip('torch')
    if not torch.cuda.is_available():
        pytest.skip('needs a GPU')
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', ['int32', 'int64', 'float32', 'float64'])
def test_gpu_extract_matches_oracle(dtype):
    """GPU extract == the oracle's Extract bit for bit: whole arrays, sample
    ranges crossing frame boundaries, index maps with skipped entries,
    accumulation into a non-zero destination for channels without FLAC or
    raw data, and a torch destination filled in place."""
    torch = _gpu()
    ts = _all_kinds_blob(dtype)
    n_det, n = ts._shape
    q = ts.quanta
    kinds = {ts.payload[o] for o in ts.offsets[:-1]}
    assert g3.ALGO_DO_FLAC | g3.ALGO_DO_BZ in kinds and g3.ALGO_NONE in kinds
    cases = [(None, None, 0, -1), (None, None, 1151, 1153), (None, None, 7, 2500),
             ([0, 4, 2, -1, 9], [2, 1, 3, 5, 0], 100, 2000),
             (None, [5, 4, 3, 2, 1, 0], 1152, 2304)]
    for di, si, start, stop in cases:
        stop_ = n if stop < 0 else stop
        n_ex = n_det if si is None else len(si)
        rows = max(n_ex, n_det)
        base = (np.arange(rows * (stop_ - start)) % 97).reshape(rows, -1).astype(dtype)
        want = oracle.g3s_extract(ts.payload, ts.offsets, dtype, n, q, base.copy(),
                                  None if di is None else np.array(di),
                                  None if si is None else np.array(si), start, stop)
        got = base.copy()
        ts.extract(got, None if di is None else np.array(di),
                   None if si is None else np.array(si), start, stop)
        np.testing.assert_array_equal(got.view(f'u{got.itemsize}'),
                                      want.view(f'u{want.itemsize}'))
    dest = torch.zeros((n_det, n), dtype=getattr(torch, dtype), device='cuda')
    ts.extract(dest)
    want = oracle.g3s_extract(ts.payload, ts.offsets, dtype, n, q, np.zeros((n_det, n), dtype))
    np.testing.assert_array_equal(dest.cpu().numpy(), want)


@pytest.mark.gpu
@pytest.mark.parametrize('level', [0, 1, 3, 5, 8])
def test_gpu_decode_levels(level):
    """Streams of every encoder preset (FIXED-only 1152-sample frames,
    LPC up to order 12 in 4096-sample frames)."""
    _gpu()
    rng = np.random.default_rng(level)
    n_det, n = 8, 20000
    d = np.stack([rng.normal(0, 300, n), 1e5 * np.sin(np.arange(n) * 0.003) + rng.normal(0, 3, n),
                  np.cumsum(rng.normal(0, 10, n)), np.zeros(n) + 7, rng.normal(0, 3, n) * 64,
                  rng.normal(0, 1e6, n), np.arange(n) % 5, rng.normal(0, 1, n)])
    d = np.round(d).astype(np.int32)
    ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n))
    ts.data = d
    ts.options(flac_level=level)
    ts.encode()
    np.testing.assert_array_equal(ts.data, d)
    st = g3.last_decode_stats()
    assert st['serial'] == 0 and st['frames'] == n_det * -(-n // (1152 if level < 3 else 4096))


def _fake_sync_stream():
    """Two VERBATIM frames whose first frame's samples spell a complete,
    CRC-valid header of frame 1: the sync scan records the fake offset for
    frame 1 and the decoder must notice and re-decode serially."""
    block, bps = 32, 24
    h = Bits()
    h.put(0x3FFE, 14); h.put(0, 1); h.put(0, 1); h.put(6, 4); h.put(9, 4); h.put(0, 4)
    h.put(6, 3); h.put(0, 1); h.put(1, 8); h.put(block - 1, 8)
    hb = h.tobytes()
    fake = hb + bytes([oracle.crc8(hb)])
    fake = fake + b'\0' * ((-len(fake)) % 3)
    vals = [int.from_bytes(fake[i:i + 3], 'big') for i in range(0, len(fake), 3)]
    vals = [v - (1 << 24) if v >= (1 << 23) else v for v in vals]
    rng = np.random.default_rng(1)
    x0 = list(rng.integers(-1000, 1000, 5)) + vals
    x0 += list(rng.integers(-1000, 1000, block - len(x0)))
    x1 = list(rng.integers(-1000, 1000, block))
    s = stream_header(block, bps, 2 * block)
    s += frame(0, block, bps, lambda b: (sub_header(b, 1), [b.sput(v, bps) for v in x0]))
    s += frame(1, block, bps, lambda b: (sub_header(b, 1), [b.sput(v, bps) for v in x1]))
    return s, np.array(x0 + x1, np.int32)


def _blob_from_streams(streams, n):
    payload = b''
    offsets = [0]
    for s in streams:
        payload += bytes([g3.ALGO_DO_FLAC]) + np.int32(len(s)).tobytes() + s
        offsets.append(len(payload))
    ts = g3.G3SuperTimestream([f'c{i}' for i in range(len(streams))], np.arange(n))
    ts.from_blob(payload, offsets, 'int32')
    return ts


@pytest.mark.gpu
def test_gpu_false_sync_and_serial_paths():
    """A fake frame header inside sample data, a variable-blocksize stream
    (decoded serially) and a corrupt stream (ValueError)."""
    _gpu()
    s, x = _fake_sync_stream()
    np.testing.assert_array_equal(oracle.flac_decode(s, len(x)), x)
    # the fake header really is inside frame 0's bytes
    assert s.find(bytes([0xFF, 0xF8]), len(stream_header(32, 24, 64)) + 2) < len(s) - 100
    # variable blocksize: blocking bit set, frame "number" = first sample
    block, bps = 32, 24
    rng = np.random.default_rng(2)
    y = rng.integers(-500, 500, 2 * block)
    sv = stream_header(block, bps, 2 * block)
    sv += frame(0, block, bps, lambda b: (sub_header(b, 1), [b.sput(v, bps) for v in y[:32]]),
                variable=True)
    sv += frame(32, block, bps, lambda b: (sub_header(b, 1), [b.sput(v, bps) for v in y[32:]]),
                variable=True)
    np.testing.assert_array_equal(oracle.flac_decode(sv, 64), y)
    good = g3.flac_encode(rng.integers(-100, 100, 64), 1)
    ts = _blob_from_streams([s, sv, good], 64)
    got = np.zeros((3, 64), np.int32)
    ts.extract(got)
    assert g3.last_decode_stats()['serial'] == 2
    np.testing.assert_array_equal(got[0], x)
    np.testing.assert_array_equal(got[1], y)
    np.testing.assert_array_equal(got[2], oracle.flac_decode(good, 64))
    bad = bytearray(good)
    bad[60] ^= 0x04
    ts = _blob_from_streams([bytes(bad)], 64)
    with pytest.raises(ValueError):
        ts.extract(np.zeros((1, 64), np.int32))


# -- the reference's test cases on the GPU path -----------------------------------

@pytest.mark.gpu
def test_gpu_reference_cases():
    """test_00 dtypes, test_03 idempotency, test_10/11/20 round trips,
    test_30 ones-and-zeros, test_60 extract semantics."""
    _gpu()
    for dtype in ['int32', 'int64', 'float32', 'float64']:
        ts, d = _ts(4, 100, sigma=0, dtype=dtype)
        assert ts.data.dtype == np.dtype(dtype)
        ts.encode()
        assert ts.dtype == np.dtype(dtype)
        ts.decode()
        assert ts.data.dtype == np.dtype(dtype)
        np.testing.assert_array_equal(ts.data, d)
    ts, d = _ts(10, 980, dtype='float32')
    a = ts.data
    assert ts.data is a
    ts.encode()
    b = ts.data
    assert b is not a
    ts.decode()
    assert ts.data is b
    for dtype in ['int32', 'int64']:
        for off in [0, 2**25, 2**26 / 3., -1.78 * 2**27]:
            ts, d = _ts(10, 980, dtype=dtype, offset=int(off))
            ts.encode()
            np.testing.assert_array_equal(ts.data, d)
    for nsamp in range(1, 20):
        ts, d = _ts(1, nsamp, sigma=0)
        ts.encode()
        np.testing.assert_array_equal(ts.data, d)
    for dtype in ['float32', 'float64']:
        ts, d = _ts(9, 1290, sigma=5., dtype=dtype)
        ts.encode()
        np.testing.assert_allclose(ts.data, d, atol=0.01 * 1e-3, rtol=0)
    # test_30: channel 0 = 77 over [first, second), others 0
    for first, second in ((0, 2000), (100, 1800)):
        buf = np.zeros((3, 2000), np.int32)
        buf[0, first:second] = 77
        ts = g3.G3SuperTimestream(['a', 'b', 'c'], np.arange(second - first))
        ts.data = buf[:, first:second].copy()
        ts.encode()
        assert ts.data.shape == (3, second - first)
        assert np.all(ts.data[0] == 77.) and np.all(ts.data[1:] == 0.)
    # test_60
    def fresh():
        t, _ = _ts(5, 100, seed=100, dtype='float32')
        t.encode()
        return t
    ts = fresh()
    dest = np.zeros((5, 100), np.float32)
    ts.extract(dest)
    np.testing.assert_array_equal(dest, ts.data)
    for dest in [np.zeros((4, 100), np.float32), np.zeros(100, np.float32),
                 np.zeros((5, 100), np.int32), ['blech']]:
        with pytest.raises(ValueError):
            fresh().extract(dest)
    idx = np.array([2, 1, 3])
    dest = np.zeros((3, 100), np.float32)
    ts = fresh()
    ts.extract(dest, None, idx)
    np.testing.assert_array_equal(dest, ts.data[idx])
    ts = fresh()
    dest = np.zeros((3, 100), np.float32)
    ts.extract(dest, src_indices=idx)
    np.testing.assert_array_equal(dest, ts.data[idx])
    with pytest.raises(ValueError):
        fresh().extract(np.zeros((1, 2, 100), np.float32), None, idx)
    with pytest.raises(ValueError):
        fresh().extract(np.zeros((3, 200), np.float32)[:, ::2], None, idx)
    ts = fresh()
    ts.decode()
    with pytest.raises(ValueError):
        ts.extract(np.zeros((3, 100), np.float32), None, idx)
    ts = fresh()
    ts.extract(np.zeros((5, 81), np.float32), start=10, stop=91)
    ts = fresh()
    src_i, dest_i = np.array([2, 1, 3]), np.array([0, 4, 2])
    dest = np.zeros((5, 81), np.float32)
    ts.extract(dest, dest_i, src_i, 10, 91)
    np.testing.assert_array_equal(dest[dest_i], ts.data[src_i, 10:91])
    assert np.all(dest[[1, 3]] == 0)


@pytest.mark.gpu
def test_gpu_decode_large_properties():
    """Full-size property check: 64 channels x 720000 samples of float32
    (1 h at 200 Hz) decode to the quantised input exactly."""
    _gpu()
    rng = np.random.default_rng(9)
    n_det, n = 64, 720000
    q = np.full(n_det, 1e-3)
    d = (np.round(rng.normal(0, 300, (n_det, n)) + 1e4 * np.sin(np.arange(n) * 1e-3))
         * 1e-3).astype(np.float32)
    ts = g3.G3SuperTimestream([f'c{i}' for i in range(n_det)], np.arange(n), None, q)
    ts.data = d
    ts.encode()
    y, r, _ = g3.split(d, q)
    want = ((r + y).astype(np.int32).astype(np.float32).astype(np.float64) * 1e-3).astype(
        np.float32)
    np.testing.assert_array_equal(ts.data, want)
