// flac.hip -- G3SuperTimestream decoding on the GPU.
//
// A G3SuperTimestream array blob holds, per channel, an algo byte and then
// a FLAC stream of the channel's 24-bit part, a bz2 block of slow
// full-width offsets or a CONST datum; float channels are finally scaled
// by their quantum (G3SuperTimestream.cxx:136-181 for the layout, :664-812
// Extract for the decode).  The host side (so3g_amd/g3super.py) walks the
// algo bytes, inflates the rare bz2 blocks and reads each FLAC stream's
// STREAMINFO; this file decodes the FLAC frames -- the bulk of the bytes and
// of the work -- and writes the final values.
//
// FLAC frames of one stream are independent once their byte offsets are
// known, and a fixed-blocksize stream (libFLAC's stream encoder writes
// those) numbers its frames.  So:
//   1. k_flac_scan lists every byte offset of every FLAC region where a
//      header of that stream could start (sync code, the first frame's
//      fixed header bytes); k_flac_headers parses the listed headers (valid
//      fields, CRC-8) and a header numbered k keeps the smallest such offset
//      in slot[k] (atomicMin).  Compressed data occasionally contain such a
//      header: it can only lower a slot.
//   2. k_flac_frames: one lane per (channel, frame) decodes frame k from
//      slot[k] if the header there carries number k and the expected block
//      size, writes its samples, checks the CRC-16 footer, records where the
//      frame ends (ends[k]) and whether that is where slot[k + 1] begins.
//      FIXED frames are decoded by k_flac_frames; LPC frames are listed
//      there and decoded by k_flac_frames_lpc, whose larger register
//      footprint then does not lower the occupancy of the common case.
//   3. k_flac_walk: channels with a broken chain (a false sync), a
//      variable-blocksize stream or an LPC order above kMaxLpcPar are walked
//      frame by frame from the first, one wave per channel: runs of frames
//      that start where their predecessor ended and decoded correctly in
//      step 2 are skipped 64 at a time, any other frame is decoded here
//      (overwriting what step 2 wrote for its samples).  A malformed frame
//      or a CRC mismatch is an error.
// Channels without FLAC (raw, bz2-only or CONST-only) go through
// k_g3s_plain.  Per sample the value is combined exactly as Extract does:
// FLAC (or raw) value, plus the inflated offsets, plus the CONST datum, in
// the full-width integer type with wrap-around, then (float32)((double)
// (float)v * quantum) or (double)v * quantum for float channels.
//
// The bit reader keeps the next 16-byte chunk in flight while the current
// one is consumed, so a lane's loads are not serialised behind its sample
// stores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/so3g_hip.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}

namespace {

#define FL_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return so3g::set_error(SO3G_ERR_HIP, std::string("HIP error in ") + #expr + ": " + \
                                                     hipGetErrorString(_e));               \
    } while (0)

constexpr int kFlBlock = 256;
constexpr int kScanBytes = 64;     // bytes tested per thread by the sync scan
constexpr int64_t kSlack = 64;     // bit-reader read-ahead allowed past a region
constexpr int kMaxLpcPar = 12;     // LPC orders decoded frame-parallel
constexpr uint32_t kNone = 0xFFFFFFFFu;
// Why a channel leaves the frame-parallel pass (bits of its flag).
constexpr int kWhyVariable = 1, kWhyNoSync = 2, kWhyBadFrame = 4, kWhyChain = 8,
              kWhyHeader = 16, kWhyLpc = 32;

// Per extracted channel, on the device.
struct Chan {
    int64_t flac_pos, flac_end;   // first frame and end of the FLAC block (-1: none)
    int64_t raw_pos;              // ALGO_NONE data (-1: none)
    int64_t off_row;              // row of the inflated offsets (-1: none)
    uint64_t cval;                // CONST datum (full-width bits)
    double quantum;
    int64_t dest_row;
    int64_t slot0;                // first frame slot of this channel
    int64_t tile0;                // first sync-scan tile of this channel
    int32_t n_frames, block, bps, has_const;
};

constexpr int kScanIters = 16;   // kFlBlock * kScanBytes slices per scan block
constexpr int64_t kTile = (int64_t)kFlBlock * kScanBytes * kScanIters;   // bytes per block

// Index j of the FLAC channel owning global index x of a per-channel prefix
// (first[j] = first index of the j-th FLAC channel, ascending): binary search.
__device__ __forceinline__ int32_t owner(const int64_t *first, int32_t n, int64_t x) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (first[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Where decoded samples go.
struct Out {
    const void *offs;             // inflated offsets, rows of n_samp
    void *dest;                   // rows of stride elements
    int64_t n_samp, start, stop, stride;
};

// -- bit reader over the payload (MSB first) ----------------------------------

__device__ __forceinline__ uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

struct Reader {
    const uint8_t *p;
    int64_t nbytes;     // readable payload bytes
    int64_t limit;      // a word beyond this byte is an error
    uint64_t buf;       // valid bits left-aligned, the rest zero
    int bits;
    uint32_t q0, q1, q2, q3;   // words of the current chunk (big-endian values)
    int nq;                    // of which not yet in buf
    uint4 pre;                 // the next chunk, in flight
    int64_t cpos;              // byte address of `pre`
    int64_t wpos;              // byte address of the next word to enter buf
    bool err;

    __device__ uint4 chunk(int64_t a) const {   // 16 bytes at a (a % 16 == 0)
        if (a + 16 <= nbytes) return *reinterpret_cast<const uint4 *>(p + a);
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t v = 0;
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                const int64_t b = a + 4 * j + i;
                v = (v << 8) | (b < nbytes ? p[b] : 0u);
            }
            w[j] = v;
        }
        return uint4{w[0], w[1], w[2], w[3]};
    }
    __device__ __forceinline__ uint32_t pop() {
        const uint32_t w = q0;
        q0 = q1; q1 = q2; q2 = q3;
        if (--nq == 0) {
            q0 = bswap32(pre.x); q1 = bswap32(pre.y); q2 = bswap32(pre.z); q3 = bswap32(pre.w);
            nq = 4;
            cpos += 16;
            pre = chunk(cpos);
        }
        wpos += 4;
        return w;
    }
    __device__ void start(int64_t pos) {
        const int64_t a = pos & ~(int64_t)15;
        const uint4 c0 = chunk(a);
        q0 = bswap32(c0.x); q1 = bswap32(c0.y); q2 = bswap32(c0.z); q3 = bswap32(c0.w);
        nq = 4;
        cpos = a + 16;
        pre = chunk(cpos);
        wpos = a;
        err = false;
        for (int i = (int)((pos & 15) >> 2); i > 0; i--) pop();
        const int skip = (int)(pos & 3) * 8;
        buf = ((uint64_t)pop() << 32) << skip;
        bits = 32 - skip;
    }
    __device__ __forceinline__ void refill() {
        if (bits <= 32) {
            if (wpos > limit) { err = true; buf |= 1ull << (63 - bits); }
            buf |= (uint64_t)pop() << (32 - bits);
            bits += 32;
        }
    }
    __device__ __forceinline__ uint32_t get(int n) {   // 0 <= n <= 32
        refill();
        if (n == 0) return 0;
        const uint32_t v = (uint32_t)(buf >> (64 - n));
        buf <<= n;
        bits -= n;
        return v;
    }
    __device__ __forceinline__ int32_t sget(int n) {   // two's complement, n <= 32
        if (n == 0) return 0;
        const uint32_t v = get(n);
        return n == 32 ? (int32_t)v : (int32_t)(v << (32 - n)) >> (32 - n);
    }
    __device__ __forceinline__ uint32_t unary() {      // zeros before the next one
        uint32_t q = 0;
        for (;;) {
            refill();
            if (buf != 0) {
                const int z = __clzll(buf);
                q += z;
                buf = (buf << z) << 1;
                bits -= z + 1;
                return q;
            }
            q += bits;
            bits = 0;
            if (err) return q;
        }
    }
    __device__ int64_t bitpos() const { return wpos * 8 - bits; }
    __device__ uint32_t byte(int64_t a) const { return a < nbytes ? p[a] : 0u; }
};

__device__ __forceinline__ uint8_t crc8_bytes(const Reader &r, int64_t a, int64_t n) {
    uint32_t c = 0;
    for (int64_t i = 0; i < n; i++) {
        c ^= r.byte(a + i);
        for (int b = 0; b < 8; b++) c = (c & 0x80) ? ((c << 1) ^ 0x07) & 0xFF : (c << 1) & 0xFF;
    }
    return (uint8_t)c;
}

// CRC-16 of payload bytes [a, b), 16-byte chunks at a time.
__device__ __forceinline__ uint32_t crc16_range(const Reader &r, int64_t a, int64_t b, const uint16_t *tab) {
    uint32_t crc = 0;
    for (int64_t c = a & ~(int64_t)15; c < b; c += 16) {
        const uint4 v = r.chunk(c);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int64_t at = c + i;
            if (at >= a && at < b) {
                const uint32_t byte = (w[i >> 2] >> (8 * (i & 3))) & 0xFF;
                crc = ((crc << 8) ^ tab[((crc >> 8) ^ byte) & 0xFF]) & 0xFFFF;
            }
        }
    }
    return crc;
}

struct Frame {
    int block, bps;
    uint64_t number;
    int blocking;
};

__device__ __forceinline__ int sample_size_bits(int code, int stream_bps) {
    switch (code) {
    case 0: return stream_bps;
    case 1: return 8;
    case 2: return 12;
    case 4: return 16;
    case 5: return 20;
    case 6: return 24;
    case 7: return 32;
    default: return -1;
    }
}

// Frame header at byte `at`: returns its length in bytes (CRC-8 included)
// or -1.  stream_bps: bits per sample of STREAMINFO.
__device__ __forceinline__ int frame_header(Reader &r, int64_t at, int stream_bps, Frame &f) {
    r.start(at);
    if (r.get(14) != 0x3FFE || r.get(1) != 0) return -1;
    f.blocking = (int)r.get(1);
    const int bs_code = (int)r.get(4), rate = (int)r.get(4), chan = (int)r.get(4);
    const int ss = (int)r.get(3);
    if (r.get(1) != 0 || chan != 0 || bs_code == 0 || rate == 15 || ss == 3) return -1;
    const uint32_t b0 = r.get(8);
    int extra;
    uint64_t num;
    if (b0 < 0x80) { num = b0; extra = 0; }
    else if ((b0 & 0xE0) == 0xC0) { num = b0 & 0x1F; extra = 1; }
    else if ((b0 & 0xF0) == 0xE0) { num = b0 & 0x0F; extra = 2; }
    else if ((b0 & 0xF8) == 0xF0) { num = b0 & 0x07; extra = 3; }
    else if ((b0 & 0xFC) == 0xF8) { num = b0 & 0x03; extra = 4; }
    else if ((b0 & 0xFE) == 0xFC) { num = b0 & 0x01; extra = 5; }
    else if (b0 == 0xFE) { num = 0; extra = 6; }
    else return -1;
    if (f.blocking == 0 && extra > 5) return -1;
    for (int i = 0; i < extra; i++) {
        const uint32_t c = r.get(8);
        if ((c & 0xC0) != 0x80) return -1;
        num = (num << 6) | (c & 0x3F);
    }
    f.number = num;
    if (bs_code == 1) f.block = 192;
    else if (bs_code <= 5) f.block = 576 << (bs_code - 2);
    else if (bs_code == 6) f.block = (int)r.get(8) + 1;
    else if (bs_code == 7) f.block = (int)r.get(16) + 1;
    else f.block = 256 << (bs_code - 8);
    if (rate == 12) r.get(8);
    else if (rate == 13 || rate == 14) r.get(16);
    f.bps = sample_size_bits(ss, stream_bps);
    const int64_t hl = r.bitpos() / 8 - at;   // header bytes before the CRC
    const uint32_t c = r.get(8);
    if (r.err || f.bps <= 0 || c != crc8_bytes(r, at, hl)) return -1;
    return (int)(hl + 1);
}

// The samples of one frame, frame-relative index i (sample t0 + i): only
// i in [lo, hi) -- the requested range -- is written, combined as Extract
// does (value, plus the channel's offsets, plus its CONST datum, in the
// full-width integer type; float types scaled by the quantum).  When the
// frame's row is 16-byte aligned, samples are gathered in registers and
// written 16 bytes at a time: on this GPU a load's wait also waits for the
// lane's earlier stores, so fewer, larger stores keep the bit reader's
// prefetches from queuing behind them.
template <int TYPE>
struct Sink {
    using E = typename std::conditional<TYPE == 1 || TYPE == 3, uint64_t, uint32_t>::type;
    static constexpr int G = 32 / sizeof(E);   // samples per 32-byte group
    uintptr_t base;     // address of sample t0 (i = 0) in its row
    const E *off;       // offsets of sample t0 + lo, or NULL
    E cval;
    double quantum;
    int lo, hi;
    bool vec;
    E g[G];             // the group's samples (g[G - 1] the newest)

    __device__ __forceinline__ Sink(const Chan &c, const Out &o, int64_t t0, int block) {
        const int64_t a = max(o.start, t0), b = min(o.stop, t0 + block);
        lo = (int)(a - t0);
        hi = (int)max(b - t0, a - t0);
        base = reinterpret_cast<uintptr_t>(o.dest) +
               (uintptr_t)((c.dest_row * o.stride + (t0 - o.start)) * (int64_t)sizeof(E));
        off = c.off_row >= 0 ? static_cast<const E *>(o.offs) + c.off_row * o.n_samp + a : nullptr;
        cval = c.has_const ? (E)c.cval : (E)0;
        quantum = c.quantum;
        vec = (base & 15) == 0;
#pragma unroll
        for (int j = 0; j < G; j++) g[j] = 0;
    }
    __device__ __forceinline__ E value(int i, int64_t x) const {
        E v = (E)x;
        if (off && i >= lo && i < hi) v += off[i - lo];
        v += cval;
        if constexpr (TYPE == 3) return (E)__double_as_longlong((double)(int64_t)v * quantum);
        else if constexpr (TYPE == 2)
            return (E)__float_as_uint((float)((double)(float)(int32_t)v * quantum));
        else return v;
    }
    __device__ __forceinline__ void store1(int i, E v) const {
        if (i >= lo && i < hi) reinterpret_cast<E *>(base)[i] = v;
    }
    // The group ending at sample `last` (g[G - 1]), of which samples
    // i0 = last - G + 1 + skip .. last are valid.
    __device__ __forceinline__ void flush(int last) const {
        const int i0 = last - (G - 1);
        if (i0 >= lo && last < hi && (i0 & (G - 1)) == 0) {
            uint4 *p = reinterpret_cast<uint4 *>(base + (uintptr_t)i0 * sizeof(E));
            if constexpr (G == 8) {
                p[0] = uint4{g[0], g[1], g[2], g[3]};
                p[1] = uint4{g[4], g[5], g[6], g[7]};
            } else {
                p[0] = uint4{(uint32_t)g[0], (uint32_t)(g[0] >> 32), (uint32_t)g[1],
                             (uint32_t)(g[1] >> 32)};
                p[1] = uint4{(uint32_t)g[2], (uint32_t)(g[2] >> 32), (uint32_t)g[3],
                             (uint32_t)(g[3] >> 32)};
            }
        } else {
#pragma unroll
            for (int j = 0; j < G; j++) store1(i0 + j, g[j]);
        }
    }
    __device__ __forceinline__ void put(int i, int64_t x) {
        const E v = value(i, x);
        if (!vec) { store1(i, v); return; }
#pragma unroll
        for (int j = 0; j < G - 1; j++) g[j] = g[j + 1];
        g[G - 1] = v;
        // A group closes at its last sample or at the last written sample;
        // a partial group's stale entries fall outside [lo, hi) or are
        // rewritten with the same values.
        if ((i & (G - 1)) == G - 1 || i == hi - 1) flush(i);
    }
};

// Writes sample t of a channel (the value v) into dest; the per-sample form
// of Sink::put for k_g3s_plain.
template <int TYPE>
__device__ __forceinline__ void emit(const Chan &c, const Out &o, int64_t t, uint64_t v_bits) {
    if (t < o.start || t >= o.stop) return;
    const int64_t at = c.dest_row * o.stride + (t - o.start);
    if constexpr (TYPE == 1 || TYPE == 3) {
        uint64_t v = v_bits;
        if (c.off_row >= 0) v += static_cast<const uint64_t *>(o.offs)[c.off_row * o.n_samp + t];
        if (c.has_const) v += c.cval;
        if constexpr (TYPE == 3) static_cast<double *>(o.dest)[at] = (double)(int64_t)v * c.quantum;
        else static_cast<uint64_t *>(o.dest)[at] = v;
    } else {
        uint32_t v = (uint32_t)v_bits;
        if (c.off_row >= 0) v += static_cast<const uint32_t *>(o.offs)[c.off_row * o.n_samp + t];
        if (c.has_const) v += (uint32_t)c.cval;
        if constexpr (TYPE == 2)
            static_cast<float *>(o.dest)[at] = (float)((double)(float)(int32_t)v * c.quantum);
        else static_cast<uint32_t *>(o.dest)[at] = v;
    }
}

// One Rice code with parameter k (<= 30): zigzag-decoded residual.  Codes
// of at most 32 bits (nearly all) are taken from the bit buffer after a
// single refill.
__device__ __forceinline__ int32_t rice(Reader &r, uint32_t k) {
    r.refill();
    const int z = __clzll(r.buf);
    uint32_t u;
    if (r.buf != 0 && z + 1 + (int)k <= 32) {
        const uint64_t t = (r.buf << z) << 1;
        const uint32_t rem = k ? (uint32_t)(t >> (64 - k)) : 0u;
        r.buf = t << k;
        r.bits -= z + 1 + (int)k;
        u = ((uint32_t)z << k) | rem;
    } else {
        const uint32_t q = r.unary();
        u = (q << k) | r.get((int)k);
    }
    return (int32_t)(u >> 1) ^ -(int32_t)(u & 1);
}

// Residual partitions of a subframe; pred(residual) returns the sample and
// updates the predictor's state.  One loop over the samples (a partition's
// parameter is read when its first sample comes up): lanes of a wave decode
// different frames, and per-partition loops would run out of step whenever
// their partition orders differ.
template <typename S, typename P>
__device__ __forceinline__ bool residuals(Reader &r, S &sink, int block, int order, int wasted,
                                          P &&pred) {
    const int method = (int)r.get(2);
    if (method > 1) return false;
    const int porder = (int)r.get(4);
    const int psize = block >> porder;
    if ((block & ((1 << porder) - 1)) != 0 || psize < order) return false;
    const uint32_t escape = method ? 31u : 15u;
    int boundary = order;   // first sample of the next partition (the first: `order`)
    uint32_t k = 0;
    int nb = -1;            // raw width of an escaped partition, -1 for Rice
    for (int i = order; i < block; i++) {
        if (i == boundary) {
            k = r.get(method ? 5 : 4);
            nb = k == escape ? (int)r.get(5) : -1;
            boundary = (i < psize ? psize : boundary + psize);
        }
        const int32_t res = nb >= 0 ? r.sget(nb) : rice(r, k);
        sink.put(i, (int64_t)pred(res) << wasted);
    }
    return !r.err;
}

// FIXED predictor of order 0..4 as the coefficients of its differences
// (x = r + c1 x1 + c2 x2 + c3 x3 + c4 x4), in 32-bit arithmetic when the
// samples are narrow enough that no term overflows, else in 64-bit.
template <typename S>
__device__ __forceinline__ bool fixed_body(Reader &r, S &sink, int block, int sbps, int wasted,
                                           int order) {
    int32_t x1 = 0, x2 = 0, x3 = 0, x4 = 0;
    for (int i = 0; i < order; i++) {
        const int32_t x = r.sget(sbps);
        x4 = x3; x3 = x2; x2 = x1; x1 = x;
        sink.put(i, (int64_t)x << wasted);
    }
    const int c1 = order == 0 ? 0 : order == 1 ? 1 : order == 2 ? 2 : order == 3 ? 3 : 4;
    const int c2 = order <= 1 ? 0 : order == 2 ? -1 : order == 3 ? -3 : -6;
    const int c3 = order <= 2 ? 0 : order == 3 ? 1 : 4;
    const int c4 = order <= 3 ? 0 : -1;
    if (sbps <= 26) {
        return residuals(r, sink, block, order, wasted, [&](int32_t res) {
            const int32_t x = res + c1 * x1 + c2 * x2 + c3 * x3 + c4 * x4;
            x4 = x3; x3 = x2; x2 = x1; x1 = x;
            return x;
        });
    }
    return residuals(r, sink, block, order, wasted, [&](int32_t res) {
        const int32_t x = (int32_t)((int64_t)res + (int64_t)c1 * x1 + (int64_t)c2 * x2 +
                                    (int64_t)c3 * x3 + (int64_t)c4 * x4);
        x4 = x3; x3 = x2; x2 = x1; x1 = x;
        return x;
    });
}

template <int MAXO, typename S>
__device__ __forceinline__ bool lpc_body(Reader &r, S &sink, int block, int sbps, int wasted,
                                         int order) {
    int32_t h[MAXO];     // h[j] = x[i - 1 - j]
    int32_t coef[MAXO];
#pragma unroll
    for (int j = 0; j < MAXO; j++) { h[j] = 0; coef[j] = 0; }
    for (int i = 0; i < order; i++) {
        const int32_t x = r.sget(sbps);
#pragma unroll
        for (int j = MAXO - 1; j > 0; j--) h[j] = h[j - 1];
        h[0] = x;
        sink.put(i, (int64_t)x << wasted);
    }
    const int prec = (int)r.get(4) + 1;
    const int shift = r.sget(5);
    if (prec == 16 || shift < 0) return false;
    for (int j = 0; j < order; j++) {
        const int32_t q = r.sget(prec);
#pragma unroll
        for (int m = 0; m < MAXO; m++)
            if (m == j) coef[m] = q;
    }
    return residuals(r, sink, block, order, wasted, [&](int32_t res) {
        int64_t sum = 0;
#pragma unroll
        for (int m = 0; m < MAXO; m++) sum += (int64_t)coef[m] * h[m];
        const int32_t x = res + (int32_t)(sum >> shift);
#pragma unroll
        for (int m = MAXO - 1; m > 0; m--) h[m] = h[m - 1];
        h[0] = x;
        return x;
    });
}

// Decodes one frame at byte `at` of channel c, emitting samples from t0 on.
// want_num >= 0: the header must carry that number and want_block samples
// (checked before anything is written).  Returns the byte offset just past
// the frame (CRC-16 included), or -1 with a reason in why.
template <int TYPE, int MAXO>
__device__ __forceinline__ int64_t decode_frame(Reader &r, const Chan &c, const Out &o,
                                                int64_t at, int64_t want_num, int want_block,
                                                int64_t t0, const uint16_t *crc_tab, Frame &f,
                                                int &why) {
    why = kWhyBadFrame;
    const int hl = frame_header(r, at, c.bps, f);
    if (hl < 0) return -1;
    if (want_num >= 0 && (f.number != (uint64_t)want_num || f.block != want_block)) {
        why = kWhyHeader;
        return -1;
    }
    if (r.get(1) != 0) return -1;
    const int type = (int)r.get(6);
    int wasted = 0;
    if (r.get(1)) wasted = (int)r.unary() + 1;
    const int sbps = f.bps - wasted;
    if (sbps <= 0 || sbps > 32) return -1;
    Sink<TYPE> sink(c, o, t0, f.block);
    bool ok;
    if (type == 0) {
        const int64_t v = (int64_t)r.sget(sbps) << wasted;
        for (int i = 0; i < f.block; i++) sink.put(i, v);
        ok = true;
    } else if (type == 1) {
        for (int i = 0; i < f.block; i++) sink.put(i, (int64_t)r.sget(sbps) << wasted);
        ok = true;
    } else if (type >= 8 && type <= 12) {
        const int order = type - 8;
        ok = order <= f.block && fixed_body(r, sink, f.block, sbps, wasted, order);
    } else if (type >= 32) {
        const int order = (type & 31) + 1;
        if (order > MAXO) { why = kWhyLpc; return -1; }
        if constexpr (MAXO > 0)
            ok = order <= f.block && lpc_body<MAXO>(r, sink, f.block, sbps, wasted, order);
        else
            ok = false;
    } else {
        ok = false;
    }
    if (!ok || r.err) return -1;
    const int64_t body_end = (r.bitpos() + 7) / 8;
    if (body_end + 2 > c.flac_end) return -1;
    const uint32_t want = (r.byte(body_end) << 8) | r.byte(body_end + 1);
    if (crc16_range(r, at, body_end, crc_tab) != want) return -1;
    why = 0;
    return body_end + 2;
}

__device__ void load_crc_table(uint16_t *tab) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        uint32_t c = (uint32_t)i << 8;
        for (int b = 0; b < 8; b++) c = (c & 0x8000) ? ((c << 1) ^ 0x8005) : (c << 1);
        tab[i] = (uint16_t)c;
    }
    __syncthreads();
}

// Sync scan, first stage: a block scans one kTile-byte tile (a channel's
// tiles are consecutive from its tile0), a thread kScanBytes byte offsets
// per pass, for the bytes a frame header of this stream starts with -- the
// sync code, then header bytes 2 and 3 as in the stream's first frame
// (block size / rate codes and channels / sample size codes; every frame
// of a fixed-blocksize stream repeats them, the last may change its block
// size code) -- and lists the offsets found in the channel's region of the
// candidate list (one counter per channel: appends rarely contend).
__global__ __launch_bounds__(kFlBlock) void k_flac_scan(const uint8_t *__restrict__ payload,
                                                        int64_t nbytes, const Chan *chans,
                                                        const int32_t *flac_ch,
                                                        const int64_t *tile0,
                                                        const int64_t *cand0, int32_t n_fch,
                                                        int64_t n_tiles, uint64_t *cand,
                                                        uint32_t *n_cand) {
    const int64_t tile = blockIdx.x;
    if (tile >= n_tiles) return;
    __shared__ int32_t s_j;
    if (threadIdx.x == 0) s_j = owner(tile0, n_fch, tile);
    __syncthreads();
    const int32_t j = s_j, ci = flac_ch[j];
    // The channel's candidate region: cand[cand0[j] .. cand0[j + 1]).
    uint64_t *mine_list = cand + cand0[j];
    const uint32_t cap = (uint32_t)(cand0[j + 1] - cand0[j]);
    const int64_t a0 = chans[ci].flac_pos + (tile - chans[ci].tile0) * kTile;
    const int64_t a1 = min(chans[ci].flac_end, a0 + kTile);
    const int64_t fpos = chans[ci].flac_pos, fend = chans[ci].flac_end;
    const uint32_t b2 = fpos + 2 < nbytes ? payload[fpos + 2] : 0u;
    const uint32_t b3 = fpos + 3 < nbytes ? payload[fpos + 3] : 0u;
    for (int it = 0; it < kScanIters; it++) {
        const int64_t lo = a0 + ((int64_t)it * kFlBlock + threadIdx.x) * kScanBytes;
        const int64_t hi = min(a1, lo + kScanBytes);
        // Offsets base .. base + kScanBytes + 15 in 16-byte chunks (the range
        // [lo, hi) starts up to 15 bytes past base), each tested with the first
        // word of the next chunk at hand for headers that cross into it; hits
        // are kept as a bit mask (bit j: offset base + j).
        constexpr int kChunks = kScanBytes / 16 + 1;
        const int64_t base = lo & ~(int64_t)15;
        auto ld = [&](int64_t c) {
            if (lo >= hi || c >= nbytes) return uint4{0u, 0u, 0u, 0u};
            if (c + 16 <= nbytes) return *reinterpret_cast<const uint4 *>(payload + c);
            uint32_t t[4];
            for (int j = 0; j < 4; j++) {
                uint32_t x = 0;
                for (int i = 3; i >= 0; i--)
                    x = (x << 8) | (c + 4 * j + i < nbytes ? payload[c + 4 * j + i] : 0u);
                t[j] = x;
            }
            return uint4{t[0], t[1], t[2], t[3]};
        };
        uint64_t mlo = 0;
        uint32_t mhi = 0;
        // Bounds relative to base, in 32 bits.
        const int rl = (int)(lo - base), rh = (int)(hi - base);
        const int rf = (int)min(fend - 3 - base, (int64_t)(1 << 20));
        const int rp = (fpos >= base && fpos < base + 16 * kChunks) ? (int)(fpos - base) : -1;
        uint4 cur = ld(base);
    #pragma unroll 1
        for (int q = 0; q < kChunks; q++) {
            const uint4 nxt = ld(base + 16 * (q + 1));
            const uint32_t c[5] = {cur.x, cur.y, cur.z, cur.w, nxt.x};
            uint32_t m = 0;
    #pragma unroll
            for (int b = 0; b < 16; b++) {
                auto B = [&](int k) { return (c[k >> 2] >> (8 * (k & 3))) & 0xFFu; };
                const int j = 16 * q + b;
                const bool h = B(b) == 0xFF && (B(b + 1) & 0xFE) == 0xF8 && j >= rl && j < rh &&
                               j < rf &&
                               (j == rp || (B(b + 3) == b3 && (B(b + 2) & 0x0F) == (b2 & 0x0F)));
                m |= (uint32_t)h << b;
            }
            if (q < 4) mlo |= (uint64_t)m << (16 * q);
            else mhi |= m;
            cur = nxt;
        }
        const uint32_t mine = (uint32_t)(__popcll(mlo) + __popc(mhi));
        // One atomic per wave: lanes take consecutive slots after a wave prefix sum.
        const int lane = threadIdx.x & 63;
        uint32_t incl = mine;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (lane >= d) incl += y;
        }
        const uint32_t total = __shfl(incl, 63);
        uint32_t first = 0;
        if (lane == 63 && total) first = atomicAdd(n_cand + j, total);
        first = __shfl(first, 63) + incl - mine;
        for (; mlo; mlo &= mlo - 1, first++)
            if (first < cap) mine_list[first] = (uint64_t)(base + __ffsll((long long)mlo) - 1);
        for (; mhi; mhi &= mhi - 1, first++)
            if (first < cap) mine_list[first] = (uint64_t)(base + 64 + __ffs(mhi) - 1);
    }
}

// Sync scan, second stage: one lane per candidate slot parses the listed
// header (valid fields, CRC-8); a header numbered k keeps the smallest
// offset in slot[k].  The first frame decides whether the stream uses
// variable block sizes; elsewhere a variable-blocksize header can only be a
// false match.  A channel whose region overflowed is sent to the walk.
__global__ __launch_bounds__(kFlBlock) void k_flac_headers(
    const uint8_t *__restrict__ payload, int64_t nbytes, const Chan *chans,
    const int32_t *flac_ch, const int64_t *cand0, int32_t n_fch, const uint64_t *cand,
    const uint32_t *n_cand, uint32_t *slots, int32_t *flags) {
    const int64_t idx = (int64_t)blockIdx.x * kFlBlock + threadIdx.x;
    if (idx >= cand0[n_fch]) return;
    const int32_t j = owner(cand0, n_fch, idx);
    const int32_t ci = flac_ch[j];
    const uint32_t cap = (uint32_t)(cand0[j + 1] - cand0[j]);
    const uint32_t local = (uint32_t)(idx - cand0[j]);
    if (local == 0 && n_cand[j] > cap) atomicOr(flags + ci, kWhyNoSync);
    if (local >= min(n_cand[j], cap)) return;
    const int64_t at = (int64_t)cand[idx];
    const Chan c = chans[ci];
    Reader r{payload, nbytes, c.flac_end + kSlack};
    Frame f;
    if (frame_header(r, at, c.bps, f) < 0) return;
    if (at == c.flac_pos && f.blocking) atomicOr(flags + ci, kWhyVariable);
    const bool last = f.number + 1 == (uint64_t)c.n_frames;
    if (!f.blocking && f.number < (uint64_t)c.n_frames &&
        (last || at == c.flac_pos || r.byte(at + 2) == r.byte(c.flac_pos + 2)))
        atomicMin(slots + c.slot0 + f.number, (uint32_t)(at - c.flac_pos));
}

// The frame-parallel task of global frame g (frames of a channel are
// consecutive from its slot0; flac_ch lists the FLAC channels in order).  MAXO = 0: FIXED frames only,
// an LPC frame is appended to lpc_list for the LPC-capable pass.
template <int TYPE, int MAXO>
__device__ __forceinline__ void frame_task(int64_t g, const uint8_t *__restrict__ payload,
                                           int64_t nbytes, const Chan *chans,
                                           const int32_t *flac_ch, const int64_t *slot0,
                                           int32_t n_fch, const uint32_t *slots,
                                           const Out &o, uint32_t *ends, int32_t *flags,
                                           uint32_t *lpc_list, uint32_t *lpc_count,
                                           const uint16_t *crc_tab) {
    const int32_t ci = flac_ch[owner(slot0, n_fch, g)];
    const Chan c = chans[ci];
    if (flags[ci] & kWhyVariable) return;
    const int64_t k = g - c.slot0;
    const int64_t t0 = k * c.block;
    // Frames after the requested samples are not needed; the frame before
    // them is decoded (emitting nothing) so that the first needed frame's
    // start is checked against its predecessor's end.
    if (t0 >= o.stop || (k + 2) * c.block <= o.start) return;
    const uint32_t s = slots[g];
    if (s == kNone) { atomicOr(flags + ci, kWhyNoSync); return; }
    Reader r{payload, nbytes, c.flac_end + kSlack};
    Frame f;
    int why;
    const int want_block = (int)min((int64_t)c.block, o.n_samp - t0);
    const int64_t end = decode_frame<TYPE, MAXO>(r, c, o, c.flac_pos + s, k, want_block, t0,
                                                 crc_tab, f, why);
    if (end >= 0) {
        ends[g] = (uint32_t)(end - c.flac_pos);
        int64_t want_end = c.flac_end;
        if (k + 1 < c.n_frames) want_end = slots[g + 1] == kNone ? -1 : c.flac_pos + slots[g + 1];
        if (end != want_end) why = kWhyChain;
    } else if (MAXO == 0 && why == kWhyLpc) {
        lpc_list[atomicAdd(lpc_count, 1u)] = (uint32_t)g;
        return;
    }
    if (why) atomicOr(flags + ci, why);
}

// Frame-parallel pass over every frame, FIXED predictors only (LPC frames
// are listed for k_flac_frames_lpc): one lane per frame.
template <int TYPE>
__global__ __launch_bounds__(kFlBlock) void k_flac_frames(
    const uint8_t *__restrict__ payload, int64_t nbytes, const Chan *chans,
    const int32_t *flac_ch, const int64_t *slot0, int32_t n_fch, int64_t n_frames,
    const uint32_t *slots, Out o, uint32_t *ends, int32_t *flags, uint32_t *lpc_list,
    uint32_t *lpc_count) {
    __shared__ uint16_t crc_tab[256];
    load_crc_table(crc_tab);
    const int64_t g = (int64_t)blockIdx.x * kFlBlock + threadIdx.x;
    if (g >= n_frames) return;
    frame_task<TYPE, 0>(g, payload, nbytes, chans, flac_ch, slot0, n_fch, slots, o, ends, flags,
                        lpc_list, lpc_count, crc_tab);
}

// The listed LPC frames (orders up to kMaxLpcPar), one lane each.
template <int TYPE>
__global__ __launch_bounds__(kFlBlock) void k_flac_frames_lpc(
    const uint8_t *__restrict__ payload, int64_t nbytes, const Chan *chans,
    const int32_t *flac_ch, const int64_t *slot0, int32_t n_fch, const uint32_t *slots, Out o,
    uint32_t *ends, int32_t *flags, const uint32_t *lpc_list, const uint32_t *lpc_count) {
    __shared__ uint16_t crc_tab[256];
    const uint32_t n = *lpc_count;
    if ((int64_t)blockIdx.x * kFlBlock >= n) return;   // whole block idle
    load_crc_table(crc_tab);
    const int64_t idx = (int64_t)blockIdx.x * kFlBlock + threadIdx.x;
    if (idx >= n) return;
    frame_task<TYPE, kMaxLpcPar>(lpc_list[idx], payload, nbytes, chans, flac_ch, slot0, n_fch,
                                 slots, o, ends, flags, nullptr, nullptr, crc_tab);
}

// One wave per channel: the frames in sequence from the first.  Runs of
// frames that start where their predecessor ended and decoded correctly in
// the parallel passes (ends[] set) are skipped 64 at a time (ballot); any
// other frame is decoded here by lane 0.  err[i] = 1 on a malformed stream.
template <int TYPE>
__global__ __launch_bounds__(64) void k_flac_walk(
    const uint8_t *__restrict__ payload, int64_t nbytes, const Chan *chans,
    const int32_t *which, int64_t n, const uint32_t *slots, const uint32_t *ends,
    const int32_t *flags, Out o, int32_t *err) {
    __shared__ uint16_t crc_tab[256];
    load_crc_table(crc_tab);
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const int32_t ci = which[i];
    const Chan c = chans[ci];
    const bool fixed = !(flags[ci] & kWhyVariable);
    int64_t at = c.flac_pos, t0 = 0, k = 0;
    while (at < c.flac_end && t0 < o.stop) {
        if (fixed && k < c.n_frames) {
            const int64_t kk = k + lane;
            bool ok = false;
            if (kk < c.n_frames) {
                const int64_t g = c.slot0 + kk;
                const uint32_t want = kk == k ? (uint32_t)(at - c.flac_pos) : ends[g - 1];
                ok = ends[g] != kNone && slots[g] == want;
            }
            const uint64_t m = __ballot(ok);
            const int run = m == ~0ull ? 64 : __ffsll((long long)~m) - 1;
            if (run > 0) {
                k += run;
                at = c.flac_pos + ends[c.slot0 + k - 1];
                t0 = min(k * (int64_t)c.block, o.n_samp);
                continue;
            }
        }
        int64_t next = -1;
        int blk = 0;
        if (lane == 0) {
            Reader r{payload, nbytes, c.flac_end + kSlack};
            Frame f;
            int why;
            next = decode_frame<TYPE, 32>(r, c, o, at, -1, 0, t0, crc_tab, f, why);
            blk = f.block;
        }
        next = __shfl(next, 0);
        blk = __shfl(blk, 0);
        if (next < 0) {
            if (lane == 0) err[i] = 1;
            return;
        }
        at = next;
        t0 += blk;
        k++;
    }
}

// Channels without FLAC: raw data or the caller's contents, plus offsets
// and the CONST datum.
template <int TYPE>
__global__ __launch_bounds__(kFlBlock) void k_g3s_plain(const uint8_t *__restrict__ payload,
                                                        const Chan *chans, const int32_t *which,
                                                        int64_t n, Out o) {
    constexpr bool W = TYPE == 1 || TYPE == 3;
    const int64_t m = o.stop - o.start;
    for (int64_t e = (int64_t)blockIdx.x * kFlBlock + threadIdx.x; e < n * m;
         e += (int64_t)gridDim.x * kFlBlock) {
        const Chan c = chans[which[e / m]];
        const int64_t t = o.start + e % m;
        const int64_t at = c.dest_row * o.stride + (t - o.start);
        uint64_t v;
        if (c.raw_pos >= 0) {
            const uint8_t *s = payload + c.raw_pos + t * (W ? 8 : 4);
            uint64_t b = 0;
            for (int q = (W ? 7 : 3); q >= 0; q--) b = (b << 8) | s[q];
            v = b;
        } else {
            v = W ? static_cast<const uint64_t *>(o.dest)[at]
                  : static_cast<const uint32_t *>(o.dest)[at];
        }
        emit<TYPE>(c, o, t, v);
    }
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Counts of the last decode call: frames decoded in parallel, channels
// walked frame by frame, OR of their reasons.
int64_t g_fl_stats[3] = {0, 0, 0};

// Grow-only device workspace of the decoder (one per process).
void *g_fl_ws = nullptr;
size_t g_fl_ws_cap = 0;

int workspace(size_t bytes, void **p) {
    if (bytes > g_fl_ws_cap) {
        if (g_fl_ws) (void)hipFree(g_fl_ws);
        g_fl_ws = nullptr;
        g_fl_ws_cap = 0;
        FL_TRY(hipMalloc(&g_fl_ws, bytes));
        g_fl_ws_cap = bytes;
    }
    *p = g_fl_ws;
    return SO3G_OK;
}

unsigned blocks_for(int64_t n, int per) {
    return (unsigned)std::max<int64_t>(1, (n + per - 1) / per);
}

template <int TYPE>
int decode_typed(const uint8_t *payload, int64_t nbytes, const std::vector<Chan> &ch, Out o,
                 hipStream_t s) {
    const int64_t n = (int64_t)ch.size();
    g_fl_stats[0] = g_fl_stats[1] = g_fl_stats[2] = 0;
    std::vector<Chan> chans = ch;
    std::vector<int32_t> flac_ch, plain;
    std::vector<int64_t> firsts;   // slot0 of each FLAC channel, then tile0 of each
    int64_t n_frames = 0, n_tiles = 0;
    for (int64_t i = 0; i < n; i++) {
        Chan &c = chans[i];
        if (c.flac_pos < 0) { plain.push_back((int32_t)i); c.n_frames = 0; continue; }
        c.slot0 = n_frames;
        c.n_frames = (int32_t)((o.n_samp + c.block - 1) / c.block);
        n_frames += c.n_frames;
        c.tile0 = n_tiles;
        n_tiles += (c.flac_end - c.flac_pos + kTile - 1) / kTile;
        flac_ch.push_back((int32_t)i);
    }
    const int32_t n_fch = (int32_t)flac_ch.size();
    for (int32_t j = 0; j < n_fch; j++) firsts.push_back(chans[flac_ch[j]].slot0);
    for (int32_t j = 0; j < n_fch; j++) firsts.push_back(chans[flac_ch[j]].tile0);
    // Candidate regions: a channel's frames twice over plus a margin for
    // false matches.
    int64_t n_cand_total = 0;
    for (int32_t j = 0; j < n_fch; j++) {
        firsts.push_back(n_cand_total);
        n_cand_total += 2 * (int64_t)chans[flac_ch[j]].n_frames + 64;
    }
    firsts.push_back(n_cand_total);
    // Workspace: chans | flac channels | slots | ends | lpc list | flags |
    // which | err | counters | header candidates
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_ch = up(n * sizeof(Chan)), b_fc = up(flac_ch.size() * 4),
                 b_fi = up(firsts.size() * 8),
                 b_sl = up(n_frames * 4), b_en = b_sl, b_ll = b_sl, b_fl = up(n * 4),
                 b_wh = up(n * 4), b_er = up(n * 4), b_lc = up(4 + 4 * (size_t)n_fch);
    const size_t b_ca = up((size_t)n_cand_total * 8);
    void *ws = nullptr;
    int rc = workspace(b_ch + b_fc + b_fi + b_sl + b_en + b_ll + b_fl + b_wh + b_er + b_lc + b_ca,
                       &ws);
    if (rc != SO3G_OK) return rc;
    char *w = static_cast<char *>(ws);
    Chan *d_ch = reinterpret_cast<Chan *>(w);
    w += b_ch;
    int32_t *d_fc = reinterpret_cast<int32_t *>(w);
    w += b_fc;
    int64_t *d_s0 = reinterpret_cast<int64_t *>(w), *d_t0 = d_s0 + n_fch, *d_c0 = d_t0 + n_fch;
    w += b_fi;
    uint32_t *d_sl = reinterpret_cast<uint32_t *>(w);
    w += b_sl;
    uint32_t *d_en = reinterpret_cast<uint32_t *>(w);
    w += b_en;
    uint32_t *d_ll = reinterpret_cast<uint32_t *>(w);
    w += b_ll;
    int32_t *d_fl = reinterpret_cast<int32_t *>(w);
    w += b_fl;
    int32_t *d_wh = reinterpret_cast<int32_t *>(w);
    w += b_wh;
    int32_t *d_er = reinterpret_cast<int32_t *>(w);
    w += b_er;
    uint32_t *d_lc = reinterpret_cast<uint32_t *>(w);   // [0] lpc count, [1..] candidates
    uint32_t *d_nc = d_lc + 1;                          // per FLAC channel
    w += b_lc;
    uint64_t *d_ca = reinterpret_cast<uint64_t *>(w);
    FL_TRY(hipMemcpyAsync(d_ch, chans.data(), n * sizeof(Chan), hipMemcpyHostToDevice, s));
    if (n_fch) {
        FL_TRY(hipMemcpyAsync(d_fc, flac_ch.data(), flac_ch.size() * 4, hipMemcpyHostToDevice, s));
        FL_TRY(hipMemcpyAsync(d_s0, firsts.data(), firsts.size() * 8, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, d_sl, n_frames, kNone);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, d_en, n_frames, kNone);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, s, reinterpret_cast<uint32_t *>(d_fl),
                       n, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, s, d_lc, (int64_t)1 + n_fch, 0u);
    if (!plain.empty()) {
        FL_TRY(hipMemcpyAsync(d_wh, plain.data(), plain.size() * 4, hipMemcpyHostToDevice, s));
        const int64_t work = (int64_t)plain.size() * (o.stop - o.start);
        hipLaunchKernelGGL(k_g3s_plain<TYPE>,
                           dim3(std::min<unsigned>(blocks_for(work, kFlBlock), 1u << 16)),
                           dim3(kFlBlock), 0, s, payload, d_ch, d_wh, (int64_t)plain.size(), o);
        // The walk below reuses d_wh.
        FL_TRY(hipStreamSynchronize(s));
    }
    if (n_frames > 0) {
        hipLaunchKernelGGL(k_flac_scan, dim3((unsigned)n_tiles), dim3(kFlBlock), 0, s, payload,
                           nbytes, d_ch, d_fc, d_t0, d_c0, n_fch, n_tiles, d_ca, d_nc);
        hipLaunchKernelGGL(k_flac_headers, dim3(blocks_for(n_cand_total, kFlBlock)),
                           dim3(kFlBlock), 0, s, payload, nbytes, d_ch, d_fc, d_c0, n_fch, d_ca,
                           d_nc, d_sl, d_fl);
        hipLaunchKernelGGL(k_flac_frames<TYPE>, dim3(blocks_for(n_frames, kFlBlock)),
                           dim3(kFlBlock), 0, s, payload, nbytes, d_ch, d_fc, d_s0, n_fch,
                           n_frames, d_sl, o, d_en, d_fl, d_ll, d_lc);
        hipLaunchKernelGGL(k_flac_frames_lpc<TYPE>, dim3(blocks_for(n_frames, kFlBlock)),
                           dim3(kFlBlock), 0, s, payload, nbytes, d_ch, d_fc, d_s0, n_fch, d_sl, o,
                           d_en, d_fl, d_ll, d_lc);
        FL_TRY(hipGetLastError());
        std::vector<int32_t> flags(n);
        FL_TRY(hipMemcpyAsync(flags.data(), d_fl, n * 4, hipMemcpyDeviceToHost, s));
        FL_TRY(hipStreamSynchronize(s));
        std::vector<int32_t> redo;
        for (int64_t i = 0; i < n; i++)
            if (flags[i] && chans[i].flac_pos >= 0) {
                redo.push_back((int32_t)i);
                g_fl_stats[2] |= flags[i];
            }
        g_fl_stats[0] = n_frames;
        g_fl_stats[1] = (int64_t)redo.size();
        if (!redo.empty()) {
            const int64_t nr = (int64_t)redo.size();
            FL_TRY(hipMemcpyAsync(d_wh, redo.data(), nr * 4, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, s,
                               reinterpret_cast<uint32_t *>(d_er), nr, 0u);
            hipLaunchKernelGGL(k_flac_walk<TYPE>, dim3((unsigned)nr), dim3(64), 0, s, payload,
                               nbytes, d_ch, d_wh, nr, d_sl, d_en, d_fl, o, d_er);
            FL_TRY(hipGetLastError());
            std::vector<int32_t> err(nr);
            FL_TRY(hipMemcpyAsync(err.data(), d_er, nr * 4, hipMemcpyDeviceToHost, s));
            FL_TRY(hipStreamSynchronize(s));
            for (int64_t i = 0; i < nr; i++)
                if (err[i])
                    return so3g::set_error(SO3G_ERR_VALUE,
                                           "corrupt FLAC stream in extracted channel " +
                                               std::to_string(redo[i]));
        }
    }
    FL_TRY(hipGetLastError());
    FL_TRY(hipStreamSynchronize(s));   // the workspace is reused by the next call
    return SO3G_OK;
}

}  // namespace

extern "C" int so3g_g3s_decode(const uint8_t *payload, int64_t payload_bytes,
                               const so3g_g3s_chan *chans, int64_t n_chan, int32_t type,
                               int64_t n_samp, int64_t start, int64_t stop, const void *offsets,
                               int64_t n_offset_rows, void *dest, int64_t dest_rows,
                               int64_t dest_stride, void *stream) {
    if (type < 0 || type > 3) return so3g::set_error(SO3G_ERR_VALUE, "bad type code");
    if (n_chan < 0 || n_samp < 0) return so3g::set_error(SO3G_ERR_VALUE, "negative shape");
    if (start < 0 || stop < start || stop > n_samp)
        return so3g::set_error(SO3G_ERR_VALUE, "sample range out of bounds");
    if (n_chan == 0 || stop == start) return SO3G_OK;
    if (payload == nullptr || (reinterpret_cast<uintptr_t>(payload) & 15) != 0)
        return so3g::set_error(SO3G_ERR_VALUE, "payload must be a 16-byte aligned device buffer");
    if (dest == nullptr || dest_stride < stop - start)
        return so3g::set_error(SO3G_ERR_VALUE, "bad destination");
    const int W = (type == 1 || type == 3) ? 8 : 4;
    std::vector<Chan> ch((size_t)n_chan);
    for (int64_t i = 0; i < n_chan; i++) {
        const so3g_g3s_chan &a = chans[i];
        Chan &c = ch[i];
        if (a.dest_row < 0 || a.dest_row >= dest_rows)
            return so3g::set_error(SO3G_ERR_VALUE, "destination row out of range");
        if (a.off_row >= n_offset_rows || (a.off_row >= 0 && offsets == nullptr))
            return so3g::set_error(SO3G_ERR_VALUE, "offset row out of range");
        if (a.flac_pos >= 0) {
            if (a.flac_end > payload_bytes || a.flac_pos >= a.flac_end)
                return so3g::set_error(SO3G_ERR_VALUE, "FLAC block outside the payload");
            if (a.block < 16 || a.block > 65535 || a.bps < 4 || a.bps > 32)
                return so3g::set_error(SO3G_ERR_VALUE, "bad FLAC stream parameters");
            if (a.flac_end - a.flac_pos >= (int64_t)UINT32_MAX)
                return so3g::set_error(SO3G_ERR_VALUE, "FLAC block too large");
        }
        if (a.raw_pos >= 0 && a.raw_pos + n_samp * W > payload_bytes)
            return so3g::set_error(SO3G_ERR_VALUE, "raw block outside the payload");
        c.flac_pos = a.flac_pos;
        c.flac_end = a.flac_pos >= 0 ? a.flac_end : -1;
        c.raw_pos = a.flac_pos >= 0 ? -1 : a.raw_pos;
        c.off_row = a.off_row;
        c.cval = (uint64_t)a.const_bits;
        c.has_const = a.has_const;
        c.quantum = a.quantum;
        c.dest_row = a.dest_row;
        c.slot0 = 0;
        c.tile0 = 0;
        c.n_frames = 0;
        c.block = a.block;
        c.bps = a.bps;
    }
    hipStream_t s = (hipStream_t)stream;
    const Out o{offsets, dest, n_samp, start, stop, dest_stride};
    switch (type) {
    case 0: return decode_typed<0>(payload, payload_bytes, ch, o, s);
    case 1: return decode_typed<1>(payload, payload_bytes, ch, o, s);
    case 2: return decode_typed<2>(payload, payload_bytes, ch, o, s);
    default: return decode_typed<3>(payload, payload_bytes, ch, o, s);
    }
}

extern "C" int so3g_g3s_stats(int64_t *out) {
    if (out == nullptr) return so3g::set_error(SO3G_ERR_VALUE, "out is NULL");
    out[0] = g_fl_stats[0];
    out[1] = g_fl_stats[1];
    out[2] = g_fl_stats[2];
    return SO3G_OK;
}
