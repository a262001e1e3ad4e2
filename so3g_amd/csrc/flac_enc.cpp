// flac_enc.cpp -- host FLAC encoder for G3SuperTimestream channels.
//
// G3SuperTimestream.encode() compresses each channel's 24-bit part with a
// mono FLAC stream (G3SuperTimestream.cxx:217-268 drives libFLAC at 24 bits
// per sample and the user's compression level).  libFLAC is not part of
// this build, so this is an encoder of our own, written from the FLAC
// format specification (RFC 9639); any conforming decoder (libFLAC's
// included) reads its streams, and the GPU decoder reads libFLAC's.
//
// Per frame it tries CONSTANT, FIXED orders 0..4, LPC (levels >= 3:
// Levinson-Durbin on a Welch-windowed autocorrelation, quantised
// coefficients) and VERBATIM, after removing wasted low bits, and keeps the
// smallest.  Residuals are Rice coded over the best partition order with a
// per-partition parameter or an escape (raw) partition.  Levels follow
// libFLAC's presets for block size (1152 for 0..2, else 4096), maximum LPC
// order and maximum partition order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/so3g_hip.h"

namespace so3g {
int set_error(int code, const std::string &msg);
}

namespace {

struct BitWriter {
    std::vector<uint8_t> buf;
    uint64_t acc = 0;
    int n = 0;   // bits in acc
    void put(uint64_t v, int k) {   // k <= 32
        if (k == 0) return;
        acc = (acc << k) | (v & ((k == 64) ? ~0ull : ((1ull << k) - 1)));
        n += k;
        while (n >= 8) {
            buf.push_back((uint8_t)(acc >> (n - 8)));
            n -= 8;
        }
    }
    void put_signed(int64_t v, int k) { put((uint64_t)v, k); }
    void unary(uint64_t q) {        // q zeros then a one
        while (q >= 32) { put(0, 32); q -= 32; }
        put(1, (int)q + 1);
    }
    void align() { if (n) put(0, 8 - n); }
    size_t bytes() const { return buf.size(); }
};

uint8_t crc8(const uint8_t *p, size_t n) {
    uint8_t c = 0;
    for (size_t i = 0; i < n; i++) {
        c ^= p[i];
        for (int b = 0; b < 8; b++) c = (c & 0x80) ? (uint8_t)((c << 1) ^ 0x07) : (uint8_t)(c << 1);
    }
    return c;
}

uint16_t crc16(const uint8_t *p, size_t n) {
    uint16_t c = 0;
    for (size_t i = 0; i < n; i++) {
        c ^= (uint16_t)(p[i] << 8);
        for (int b = 0; b < 8; b++)
            c = (c & 0x8000) ? (uint16_t)((c << 1) ^ 0x8005) : (uint16_t)(c << 1);
    }
    return c;
}

struct Preset { int block, max_lpc, max_porder; };
Preset preset(int level) {
    static const Preset p[9] = {{1152, 0, 3}, {1152, 0, 3}, {1152, 0, 3}, {4096, 6, 4},
                                {4096, 8, 4}, {4096, 8, 5}, {4096, 8, 6}, {4096, 12, 6},
                                {4096, 12, 6}};
    return p[std::max(0, std::min(8, level))];
}

inline uint64_t zigzag(int64_t v) { return v >= 0 ? (uint64_t)v << 1 : ((uint64_t)(-v) << 1) - 1; }

int bits_signed(int64_t lo, int64_t hi) {   // two's complement width holding [lo, hi]
    int k = 1;
    while (k < 64 && (lo < -(int64_t)(1ull << (k - 1)) || hi > (int64_t)(1ull << (k - 1)) - 1)) k++;
    return k;
}

// Best Rice coding of r[order..block): partition order, parameters (31/15
// = escape with raw width in raw_bits) and the total bits.
struct RiceChoice {
    int porder = 0, method = 0;
    std::vector<int> k, raw_bits;
    uint64_t bits = ~0ull;
};

RiceChoice choose_rice(const int64_t *r, int block, int order, int max_porder) {
    RiceChoice best;
    for (int po = 0; po <= max_porder; po++) {
        if (block % (1 << po) != 0 || (block >> po) <= order) break;
        const int parts = 1 << po, psize = block >> po;
        RiceChoice c;
        c.porder = po;
        c.k.resize(parts);
        c.raw_bits.assign(parts, 0);
        uint64_t total = 2 + 4;
        bool wide = false;
        for (int p = 0; p < parts; p++) {
            const int a = p == 0 ? order : p * psize, b = (p + 1) * psize;
            uint64_t sum = 0;
            int64_t lo = 0, hi = 0;
            for (int i = a; i < b; i++) {
                sum += zigzag(r[i]);
                lo = std::min(lo, r[i]);
                hi = std::max(hi, r[i]);
            }
            const int cnt = b - a;
            int k0 = 0;
            if (cnt > 0) {
                const uint64_t mean = sum / (uint64_t)cnt;
                while (k0 < 30 && (1ull << (k0 + 1)) <= mean) k0++;
            }
            uint64_t pbest = ~0ull;
            int kbest = 0;
            for (int k = std::max(0, k0 - 1); k <= std::min(30, k0 + 1); k++) {
                uint64_t bits = 0;
                for (int i = a; i < b; i++) bits += (zigzag(r[i]) >> k) + 1 + k;
                if (bits < pbest) { pbest = bits; kbest = k; }
            }
            const int rb = cnt > 0 ? bits_signed(lo, hi) : 0;
            const uint64_t esc = 5 + (uint64_t)cnt * rb;
            if (esc < pbest && rb <= 31) {
                c.k[p] = -1;
                c.raw_bits[p] = (lo == 0 && hi == 0) ? 0 : rb;
                pbest = 5 + (uint64_t)cnt * c.raw_bits[p];
            } else {
                c.k[p] = kbest;
                if (kbest >= 15) wide = true;
            }
            total += pbest;
        }
        c.method = wide ? 1 : 0;
        total += (uint64_t)parts * (wide ? 5 : 4);
        c.bits = total;
        if (total < best.bits) best = c;
    }
    return best;
}

void write_residual(BitWriter &w, const int64_t *r, int block, int order, const RiceChoice &c) {
    w.put(c.method, 2);
    w.put(c.porder, 4);
    const int psize = block >> c.porder, pbits = c.method ? 5 : 4;
    const uint64_t escape = c.method ? 31 : 15;
    for (int p = 0; p < (1 << c.porder); p++) {
        const int a = p == 0 ? order : p * psize, b = (p + 1) * psize;
        if (c.k[p] < 0) {
            w.put(escape, pbits);
            w.put(c.raw_bits[p], 5);
            for (int i = a; i < b; i++) w.put_signed(r[i], c.raw_bits[p]);
        } else {
            const int k = c.k[p];
            w.put(k, pbits);
            for (int i = a; i < b; i++) {
                const uint64_t u = zigzag(r[i]);
                w.unary(u >> k);
                w.put(u & ((1ull << k) - 1), k);
            }
        }
    }
}

// Quantised LPC coefficients of the given order (Levinson-Durbin on a
// Welch-windowed autocorrelation); false if the signal gives none.
bool lpc_coefs(const int64_t *x, int n, int order, int precision, int32_t *q, int *shift) {
    std::vector<double> wx(n), ac(order + 1, 0.0);
    for (int i = 0; i < n; i++) {
        const double t = (2.0 * i - (n - 1)) / (n + 1);
        wx[i] = (double)x[i] * (1.0 - t * t);
    }
    for (int l = 0; l <= order; l++)
        for (int i = l; i < n; i++) ac[l] += wx[i] * wx[i - l];
    if (ac[0] <= 0) return false;
    std::vector<double> a(order + 1, 0.0), tmp(order + 1);
    double err = ac[0];
    for (int i = 1; i <= order; i++) {
        double acc = ac[i];
        for (int j = 1; j < i; j++) acc -= a[j] * ac[i - j];
        const double k = acc / err;
        tmp = a;
        a[i] = k;
        for (int j = 1; j < i; j++) a[j] = tmp[j] - k * tmp[i - j];
        err *= (1 - k * k);
        if (err <= 0) return false;
    }
    double cmax = 0;
    for (int j = 1; j <= order; j++) cmax = std::max(cmax, std::fabs(a[j]));
    if (cmax <= 0) return false;
    int log2c = 0;
    std::frexp(cmax, &log2c);
    int s = precision - 1 - log2c;
    s = std::max(0, std::min(15, s));
    const int32_t qmax = (1 << (precision - 1)) - 1, qmin = -(1 << (precision - 1));
    double carry = 0;
    for (int j = 0; j < order; j++) {
        const double v = a[j + 1] * (double)(1 << s) + carry;
        long qi = std::lround(v);
        qi = std::max<long>(qmin, std::min<long>(qmax, qi));
        carry = v - (double)qi;
        q[j] = (int32_t)qi;
    }
    *shift = s;
    return true;
}

struct Subframe {
    int type = 1;       // 0 constant, 1 verbatim, 8..12 fixed, 32.. lpc
    int order = 0, precision = 0, shift = 0;
    int32_t coef[32];
    RiceChoice rice;
    uint64_t bits = ~0ull;
    std::vector<int64_t> res;
};

void encode_frame(BitWriter &w, const int32_t *xin, int block, int bps, uint64_t number,
                  const Preset &ps) {
    const size_t start = w.bytes();
    // Header.
    w.put(0x3FFE, 14); w.put(0, 1); w.put(0, 1);
    int bs_code, bs_extra = 0;
    if (block == 192) bs_code = 1;
    else if (block == 576 || block == 1152 || block == 2304 || block == 4608) {
        bs_code = 2; while ((576 << (bs_code - 2)) != block) bs_code++;
    } else if (block >= 256 && block <= 32768 && (block & (block - 1)) == 0) {
        bs_code = 8; while ((256 << (bs_code - 8)) != block) bs_code++;
    } else if (block <= 256) { bs_code = 6; bs_extra = 8; }
    else { bs_code = 7; bs_extra = 16; }
    w.put(bs_code, 4);
    w.put(9, 4);                                   // 44.1 kHz, as in STREAMINFO
    w.put(0, 4);                                   // mono
    int ss_code = 0;
    switch (bps) { case 8: ss_code = 1; break; case 12: ss_code = 2; break;
                   case 16: ss_code = 4; break; case 20: ss_code = 5; break;
                   case 24: ss_code = 6; break; case 32: ss_code = 7; break; }
    w.put(ss_code, 3); w.put(0, 1);
    // Coded frame number.
    if (number < 0x80) w.put(number, 8);
    else {
        int nb = 2;
        while (nb < 6 && number >= (1ull << (5 * nb + 1))) nb++;
        w.put((0xFF00u >> nb) & 0xFF | (number >> (6 * (nb - 1))), 8);
        for (int i = nb - 2; i >= 0; i--) w.put(0x80 | ((number >> (6 * i)) & 0x3F), 8);
    }
    if (bs_extra) w.put(block - 1, bs_extra);
    w.put(crc8(w.buf.data() + start, w.bytes() - start), 8);

    // Wasted low bits.
    uint32_t orv = 0;
    for (int i = 0; i < block; i++) orv |= (uint32_t)xin[i];
    int wasted = 0;
    if (orv != 0) while (wasted < bps - 1 && !((orv >> wasted) & 1)) wasted++;
    const int sbps = bps - wasted;
    std::vector<int64_t> x(block);
    for (int i = 0; i < block; i++) x[i] = (int64_t)xin[i] >> wasted;

    Subframe best;
    best.type = 1;
    best.bits = (uint64_t)block * sbps;
    bool constant = true;
    for (int i = 1; i < block; i++) constant &= x[i] == x[0];
    if (constant) {
        best.type = 0;
        best.bits = sbps;
    } else {
        std::vector<int64_t> r(block);
        const int max_fixed = std::min(4, block - 1);
        for (int o = 0; o <= max_fixed; o++) {
            for (int i = o; i < block; i++) {
                int64_t p = 0;
                switch (o) {
                case 1: p = x[i - 1]; break;
                case 2: p = 2 * x[i - 1] - x[i - 2]; break;
                case 3: p = 3 * x[i - 1] - 3 * x[i - 2] + x[i - 3]; break;
                case 4: p = 4 * x[i - 1] - 6 * x[i - 2] + 4 * x[i - 3] - x[i - 4]; break;
                }
                r[i] = x[i] - p;
            }
            bool fits = true;
            for (int i = o; i < block; i++) fits &= r[i] >= INT32_MIN && r[i] <= INT32_MAX;
            if (!fits) continue;
            RiceChoice rc = choose_rice(r.data(), block, o, ps.max_porder);
            const uint64_t bits = (uint64_t)o * sbps + rc.bits;
            if (bits < best.bits) {
                best.type = 8 + o; best.order = o; best.rice = rc; best.bits = bits; best.res = r;
            }
        }
        const int precision = 12;
        for (int o = 1; o <= std::min(ps.max_lpc, block - 1); o++) {
            if (o != ps.max_lpc && o != 2 && o != 4) continue;
            int32_t q[32];
            int shift = 0;
            if (!lpc_coefs(x.data(), block, o, precision, q, &shift)) continue;
            bool fits = true;
            for (int i = o; i < block && fits; i++) {
                int64_t sum = 0;
                for (int j = 0; j < o; j++) sum += (int64_t)q[j] * x[i - 1 - j];
                r[i] = x[i] - (sum >> shift);
                fits = r[i] >= INT32_MIN && r[i] <= INT32_MAX;
            }
            if (!fits) continue;
            RiceChoice rc = choose_rice(r.data(), block, o, ps.max_porder);
            const uint64_t bits = (uint64_t)o * sbps + 4 + 5 + (uint64_t)o * precision + rc.bits;
            if (bits < best.bits) {
                best.type = 32 + o - 1; best.order = o; best.precision = precision;
                best.shift = shift; std::memcpy(best.coef, q, sizeof(int32_t) * o);
                best.rice = rc; best.bits = bits; best.res = r;
            }
        }
    }
    // Subframe.
    w.put(0, 1);
    w.put(best.type, 6);
    if (wasted) { w.put(1, 1); w.unary(wasted - 1); } else w.put(0, 1);
    if (best.type == 0) {
        w.put_signed(x[0], sbps);
    } else if (best.type == 1) {
        for (int i = 0; i < block; i++) w.put_signed(x[i], sbps);
    } else {
        for (int i = 0; i < best.order; i++) w.put_signed(x[i], sbps);
        if (best.type >= 32) {
            w.put(best.precision - 1, 4);
            w.put_signed(best.shift, 5);
            for (int j = 0; j < best.order; j++) w.put_signed(best.coef[j], best.precision);
        }
        write_residual(w, best.res.data(), block, best.order, best.rice);
    }
    w.align();
    const uint16_t c = crc16(w.buf.data() + start, w.bytes() - start);
    w.put(c, 16);
}

}  // namespace

extern "C" int so3g_flac_encode(const int32_t *x, int64_t n, int32_t bps, int32_t level,
                                uint8_t *out, int64_t capacity, int64_t *n_bytes) {
    if (x == nullptr && n > 0) return so3g::set_error(SO3G_ERR_VALUE, "samples are NULL");
    if (n < 0 || n >= (1ll << 36)) return so3g::set_error(SO3G_ERR_VALUE, "bad sample count");
    if (bps < 4 || bps > 32) return so3g::set_error(SO3G_ERR_VALUE, "bits per sample must be 4..32");
    if (n_bytes == nullptr) return so3g::set_error(SO3G_ERR_VALUE, "n_bytes is NULL");
    const Preset ps = preset(level);
    if (bps < 32) {
        const int64_t lo = -(1ll << (bps - 1)), hi = (1ll << (bps - 1)) - 1;
        for (int64_t i = 0; i < n; i++)
            if (x[i] < lo || x[i] > hi)
                return so3g::set_error(SO3G_ERR_VALUE, "sample outside the bits-per-sample range");
    }
    BitWriter w;
    w.buf.reserve((size_t)(n * bps / 8 / 2 + 1024));
    for (char ch : std::string("fLaC")) w.put((uint8_t)ch, 8);
    w.put(1, 1); w.put(0, 7); w.put(34, 24);       // last metadata block: STREAMINFO
    const int blk = ps.block;
    w.put(blk, 16); w.put(blk, 16);                // min / max block size
    w.put(0, 24); w.put(0, 24);                    // frame sizes unknown
    w.put(44100, 20); w.put(0, 3); w.put(bps - 1, 5);
    w.put((uint64_t)n >> 32, 4); w.put((uint64_t)n & 0xFFFFFFFFu, 32);
    for (int i = 0; i < 4; i++) w.put(0, 32);      // MD5 unknown
    uint64_t frame = 0;
    for (int64_t a = 0; a < n; a += blk, frame++) {
        const int b = (int)std::min<int64_t>(blk, n - a);
        encode_frame(w, x + a, b, bps, frame, ps);
        if (capacity >= 0 && (int64_t)w.bytes() > capacity)
            return so3g::set_error(SO3G_ERR_VALUE, "FLAC stream exceeds the capacity");
    }
    *n_bytes = (int64_t)w.bytes();
    if (capacity >= 0 && *n_bytes > capacity)
        return so3g::set_error(SO3G_ERR_VALUE, "FLAC stream exceeds the capacity");
    if (out != nullptr) std::memcpy(out, w.buf.data(), w.bytes());
    return SO3G_OK;
}

// Per channel: quantise float data (x / quantum, rounded half away from
// zero: roundf of the float-converted quotient for float32, round for
// float64), then split each value v = r + y with r a multiple of 2^24
// that changes only when v moves to another multiple (2^24 snap, 2^23
// step), y in [-2^23, 2^23): the 24-bit part FLAC encodes.  Matches
// G3SuperTimestream.cxx:72-92 and :218-231, including 32-bit wrap-around
// for int32/float32 data.  type: 0 int32, 1 int64, 2 float32, 3 float64.
// r is written as int64 (int32 forms sign-extended); changes[d] counts the
// branch updates (1 = constant branch).
extern "C" int so3g_g3s_split(const void *data, int32_t type, int64_t n_det, int64_t n,
                              const double *quanta, int32_t *y, int64_t *r,
                              int32_t *changes) {
    if (type < 0 || type > 3) return so3g::set_error(SO3G_ERR_VALUE, "bad type code");
    if (n_det < 0 || n < 0) return so3g::set_error(SO3G_ERR_VALUE, "negative shape");
    if ((type >= 2) && quanta == nullptr) return so3g::set_error(SO3G_ERR_VALUE, "quanta are NULL");
    constexpr int64_t kSnap = 1 << 24, kStep = 1 << 23;
    for (int64_t d = 0; d < n_det; d++) {
        int32_t cnt = 0;
        if (type == 0 || type == 2) {
            uint32_t cur = 0;
            for (int64_t i = 0; i < n; i++) {
                const int64_t at = d * n + i;
                const int32_t v = type == 0
                    ? static_cast<const int32_t *>(data)[at]
                    : (int32_t)roundf((float)(static_cast<const float *>(data)[at] / quanta[d]));
                const uint32_t snap = ((uint32_t)v + (uint32_t)(kSnap / 2)) & ~(uint32_t)(kSnap - 1);
                const int32_t jump = (int32_t)(snap - cur);
                if (cnt == 0 || jump >= (int32_t)kStep || jump < -(int32_t)kStep) { cur = snap; cnt++; }
                y[at] = (int32_t)((uint32_t)v - cur);
                r[at] = (int32_t)cur;
            }
        } else {
            uint64_t cur = 0;
            for (int64_t i = 0; i < n; i++) {
                const int64_t at = d * n + i;
                const int64_t v = type == 1
                    ? static_cast<const int64_t *>(data)[at]
                    : (int64_t)round(static_cast<const double *>(data)[at] / quanta[d]);
                const uint64_t snap = ((uint64_t)v + (uint64_t)(kSnap / 2)) & ~(uint64_t)(kSnap - 1);
                const int64_t jump = (int64_t)(snap - cur);
                if (cnt == 0 || jump >= kStep || jump < -kStep) { cur = snap; cnt++; }
                y[at] = (int32_t)((uint64_t)v - cur);
                r[at] = (int64_t)cur;
            }
        }
        changes[d] = cnt;
    }
    return SO3G_OK;
}
