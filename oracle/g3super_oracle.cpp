// g3super_oracle.cpp -- CPU oracle for G3SuperTimestream decoding.
// TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's cpu_baseline leg
// as the checker, never by the product path.
//
// Restates, sample by sample:
//   * the reference's quantise-and-rebranch step of encode_array
//     (G3SuperTimestream.cxx:72-92 rebranch, :218-231 quantisation):
//     oracle_g3s_split;
//   * the reference's Extract (G3SuperTimestream.cxx:664-812): per channel
//     the algo byte, a FLAC block decoded to int32, a bz2 block of
//     full-width offsets (inflated by the caller: Python's bz2 module is
//     libbz2) added in, a CONST datum added in, then the float conversion
//     with the channel's quantum: oracle_g3s_combine;
//   * a sequential FLAC decoder of one mono stream written from the FLAC
//     format specification (RFC 9639: stream marker and metadata blocks,
//     frame header with CRC-8, CONSTANT / VERBATIM / FIXED / LPC subframes,
//     wasted bits, partitioned Rice residuals with escape codes, CRC-16
//     footer).  The reference decodes with libFLAC (not in this image);
//     this follows the published format instead: oracle_flac_decode.
//
// Parity pins (tests/test_g3super.py): CRC-8 and CRC-16 check values of the
// CRC catalogue; bitstreams assembled field by field in the test from the
// specification; fixed predictors annihilating polynomials; round trips
// through the library's encoder; the reference's own test cases.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

// CRC-8, polynomial x^8 + x^2 + x + 1, initial 0 (frame headers).
uint8_t crc8(const uint8_t *p, int64_t n) {
    uint8_t c = 0;
    for (int64_t i = 0; i < n; i++) {
        c ^= p[i];
        for (int b = 0; b < 8; b++) c = (c & 0x80) ? (uint8_t)((c << 1) ^ 0x07) : (uint8_t)(c << 1);
    }
    return c;
}

// CRC-16, polynomial x^16 + x^15 + x^2 + 1, initial 0 (frame footers).
uint16_t crc16(const uint8_t *p, int64_t n) {
    uint16_t c = 0;
    for (int64_t i = 0; i < n; i++) {
        c ^= (uint16_t)(p[i] << 8);
        for (int b = 0; b < 8; b++)
            c = (c & 0x8000) ? (uint16_t)((c << 1) ^ 0x8005) : (uint16_t)(c << 1);
    }
    return c;
}

// MSB-first bit reader over [p, p + n); reading past the end sets err.
struct Bits {
    const uint8_t *p;
    int64_t n;
    int64_t pos = 0;   // bit position
    bool err = false;
    uint64_t get(int k) {   // 0 <= k <= 57
        uint64_t v = 0;
        for (int i = 0; i < k; i++) {
            const int64_t byte = pos >> 3;
            if (byte >= n) { err = true; return 0; }
            v = (v << 1) | ((p[byte] >> (7 - (pos & 7))) & 1u);
            pos++;
        }
        return v;
    }
    int64_t sget(int k) {   // two's complement, k bits
        if (k == 0) return 0;
        const uint64_t v = get(k);
        return (v & (1ull << (k - 1))) ? (int64_t)v - (int64_t)(1ull << k) : (int64_t)v;
    }
    uint64_t unary() {      // zeros before the next one bit
        uint64_t q = 0;
        while (!err && get(1) == 0) q++;
        return q;
    }
    void align() { pos = (pos + 7) & ~(int64_t)7; }
};

// Frame header fields.
struct Header {
    int blocking, block, rate_code, chan, bps;
    uint64_t number;
};

// Parse a frame header at byte offset `at` (sync included); returns the
// header length in bytes (CRC-8 included) or -1.
int parse_header(const uint8_t *s, int64_t len, int64_t at, int stream_bps, Header &h) {
    Bits b{s + at, len - at};
    if (b.get(14) != 0x3FFE) return -1;
    if (b.get(1) != 0) return -1;
    h.blocking = (int)b.get(1);
    const int bs_code = (int)b.get(4);
    h.rate_code = (int)b.get(4);
    h.chan = (int)b.get(4);
    const int ss_code = (int)b.get(3);
    if (b.get(1) != 0) return -1;
    // Coded frame (fixed blocking) or sample (variable) number: 1 to 7 bytes.
    const uint64_t b0 = b.get(8);
    int extra = 0;
    uint64_t num = 0;
    if (b0 < 0x80) { num = b0; }
    else if ((b0 & 0xE0) == 0xC0) { num = b0 & 0x1F; extra = 1; }
    else if ((b0 & 0xF0) == 0xE0) { num = b0 & 0x0F; extra = 2; }
    else if ((b0 & 0xF8) == 0xF0) { num = b0 & 0x07; extra = 3; }
    else if ((b0 & 0xFC) == 0xF8) { num = b0 & 0x03; extra = 4; }
    else if ((b0 & 0xFE) == 0xFC) { num = b0 & 0x01; extra = 5; }
    else if (b0 == 0xFE) { num = 0; extra = 6; }
    else return -1;
    for (int i = 0; i < extra; i++) {
        const uint64_t c = b.get(8);
        if ((c & 0xC0) != 0x80) return -1;
        num = (num << 6) | (c & 0x3F);
    }
    if (h.blocking == 0 && extra > 5) return -1;
    h.number = num;
    if (bs_code == 0) return -1;
    else if (bs_code == 1) h.block = 192;
    else if (bs_code <= 5) h.block = 576 << (bs_code - 2);
    else if (bs_code == 6) h.block = (int)b.get(8) + 1;
    else if (bs_code == 7) h.block = (int)b.get(16) + 1;
    else h.block = 256 << (bs_code - 8);
    if (h.rate_code == 12) b.get(8);
    else if (h.rate_code == 13 || h.rate_code == 14) b.get(16);
    else if (h.rate_code == 15) return -1;
    static const int ss_bits[8] = {0, 8, 12, -1, 16, 20, 24, 32};
    h.bps = ss_code == 0 ? stream_bps : ss_bits[ss_code];
    if (h.bps <= 0) return -1;
    const int64_t hdr_bytes = b.pos >> 3;
    const uint8_t c = (uint8_t)b.get(8);
    if (b.err || c != crc8(s + at, hdr_bytes)) return -1;
    return (int)(hdr_bytes + 1);
}

// Residual of one subframe after `order` warm-up samples into r[order..).
bool residual(Bits &b, int block, int order, int64_t *r) {
    const int method = (int)b.get(2);
    if (method > 1) return false;
    const int pbits = method == 0 ? 4 : 5;
    const uint64_t escape = method == 0 ? 15u : 31u;
    const int porder = (int)b.get(4);
    const int parts = 1 << porder;
    if ((block >> porder) < order || (block & (parts - 1)) != 0) return false;
    int i = order;
    for (int p = 0; p < parts; p++) {
        const int cnt = (block >> porder) - (p == 0 ? order : 0);
        const uint64_t k = b.get(pbits);
        if (k == escape) {
            const int nb = (int)b.get(5);
            for (int j = 0; j < cnt; j++) r[i++] = b.sget(nb);
        } else {
            for (int j = 0; j < cnt; j++) {
                const uint64_t q = b.unary();
                const uint64_t u = (q << k) | b.get((int)k);
                r[i++] = (int64_t)(u >> 1) ^ -(int64_t)(u & 1);
            }
        }
        if (b.err) return false;
    }
    return true;
}

}  // namespace

extern "C" {

uint8_t oracle_crc8(const uint8_t *p, int64_t n) { return crc8(p, n); }
uint16_t oracle_crc16(const uint8_t *p, int64_t n) { return crc16(p, n); }

// Decode one mono FLAC stream s[0..len) into out (int32, at most cap
// samples).  Returns the number of samples, or -1 (bad marker / metadata),
// -2 (bad frame header), -3 (bad subframe), -4 (CRC-16 mismatch), -5 (more
// than cap samples), -6 (not mono).
int64_t oracle_flac_decode(const uint8_t *s, int64_t len, int32_t *out, int64_t cap) {
    if (len < 4 || memcmp(s, "fLaC", 4) != 0) return -1;
    int64_t at = 4;
    int stream_bps = 0;
    for (;;) {
        if (at + 4 > len) return -1;
        const int last = s[at] >> 7, type = s[at] & 0x7F;
        const int64_t blen = ((int64_t)s[at + 1] << 16) | (s[at + 2] << 8) | s[at + 3];
        if (at + 4 + blen > len) return -1;
        if (type == 0) {
            if (blen < 34) return -1;
            Bits b{s + at + 4, blen};
            b.get(16); b.get(16); b.get(24); b.get(24); b.get(20);
            if (b.get(3) != 0) return -6;
            stream_bps = (int)b.get(5) + 1;
        }
        at += 4 + blen;
        if (last) break;
    }
    int64_t n_out = 0;
    int64_t *x = nullptr, *r = nullptr;
    int cap_block = 0;
    while (at < len) {
        Header h;
        const int hl = parse_header(s, len, at, stream_bps, h);
        if (hl < 0) { delete[] x; delete[] r; return -2; }
        if (h.chan != 0) { delete[] x; delete[] r; return -6; }
        if (h.block > cap_block) {
            delete[] x; delete[] r;
            cap_block = h.block;
            x = new int64_t[cap_block];
            r = new int64_t[cap_block];
        }
        Bits b{s + at + hl, len - at - hl};
        if (b.get(1) != 0) { delete[] x; delete[] r; return -3; }
        const int type = (int)b.get(6);
        int wasted = 0;
        if (b.get(1)) wasted = (int)b.unary() + 1;
        const int bps = h.bps - wasted;
        if (bps <= 0) { delete[] x; delete[] r; return -3; }
        if (type == 0) {                       // CONSTANT
            const int64_t v = b.sget(bps);
            for (int i = 0; i < h.block; i++) x[i] = v;
        } else if (type == 1) {                // VERBATIM
            for (int i = 0; i < h.block; i++) x[i] = b.sget(bps);
        } else if (type >= 8 && type <= 12) {  // FIXED, order 0..4
            const int order = type - 8;
            if (order > h.block) { delete[] x; delete[] r; return -3; }
            for (int i = 0; i < order; i++) x[i] = b.sget(bps);
            if (!residual(b, h.block, order, r)) { delete[] x; delete[] r; return -3; }
            for (int i = order; i < h.block; i++) {
                int64_t p = 0;
                switch (order) {
                case 0: p = 0; break;
                case 1: p = x[i - 1]; break;
                case 2: p = 2 * x[i - 1] - x[i - 2]; break;
                case 3: p = 3 * x[i - 1] - 3 * x[i - 2] + x[i - 3]; break;
                default: p = 4 * x[i - 1] - 6 * x[i - 2] + 4 * x[i - 3] - x[i - 4]; break;
                }
                x[i] = r[i] + p;
            }
        } else if (type >= 32) {               // LPC, order 1..32
            const int order = (type & 31) + 1;
            if (order > h.block) { delete[] x; delete[] r; return -3; }
            for (int i = 0; i < order; i++) x[i] = b.sget(bps);
            const int prec = (int)b.get(4) + 1;
            if (prec == 16) { delete[] x; delete[] r; return -3; }
            const int64_t shift = b.sget(5);
            if (shift < 0) { delete[] x; delete[] r; return -3; }
            int64_t coef[32];
            for (int j = 0; j < order; j++) coef[j] = b.sget(prec);
            if (!residual(b, h.block, order, r)) { delete[] x; delete[] r; return -3; }
            for (int i = order; i < h.block; i++) {
                int64_t sum = 0;
                for (int j = 0; j < order; j++) sum += coef[j] * x[i - 1 - j];
                x[i] = r[i] + (sum >> shift);
            }
        } else {
            delete[] x; delete[] r; return -3;
        }
        if (b.err) { delete[] x; delete[] r; return -3; }
        b.align();
        const int64_t body_end = at + hl + (b.pos >> 3);
        if (body_end + 2 > len) { delete[] x; delete[] r; return -4; }
        const uint16_t want = (uint16_t)((s[body_end] << 8) | s[body_end + 1]);
        if (crc16(s + at, body_end - at) != want) { delete[] x; delete[] r; return -4; }
        if (n_out + h.block > cap) { delete[] x; delete[] r; return -5; }
        for (int i = 0; i < h.block; i++) out[n_out + i] = (int32_t)(x[i] << wasted);
        n_out += h.block;
        at = body_end + 2;
    }
    delete[] x; delete[] r;
    return n_out;
}

// Quantise (float types) and split x = r + y per channel as encode_array
// does before FLAC (G3SuperTimestream.cxx:218-231, rebranch :72-92 with
// snap_to = 2^24, step_at = 2^23).  type: 0 int32, 1 int64, 2 float32,
// 3 float64.  y (n_det, n) int32; r (n_det, n) int64 (the int32 forms are
// sign-extended); branches[d] = number of distinct branch values.
void oracle_g3s_split(const void *data, int type, int64_t n_det, int64_t n,
                      const double *quanta, int32_t *y, int64_t *r, int32_t *branches) {
    const int64_t M = 1 << 24, step = M / 2;
    for (int64_t d = 0; d < n_det; d++) {
        int cnt = 0;
        if (type == 0 || type == 2) {
            // 32-bit arithmetic with two's complement wrap-around.
            uint32_t branch = 0;
            for (int64_t i = 0; i < n; i++) {
                int32_t xi;
                if (type == 0) {
                    xi = ((const int32_t *)data)[d * n + i];
                } else {
                    const float f = ((const float *)data)[d * n + i];
                    xi = (int32_t)roundf((float)(f / quanta[d]));
                }
                const uint32_t nb = ((uint32_t)xi + (uint32_t)(M / 2)) & ~(uint32_t)(M - 1);
                const int32_t diff = (int32_t)(nb - branch);
                if (diff >= (int32_t)step || diff < -(int32_t)step || cnt == 0) {
                    branch = nb;
                    cnt++;
                }
                y[d * n + i] = (int32_t)((uint32_t)xi - branch);
                r[d * n + i] = (int32_t)branch;
            }
        } else {
            uint64_t branch = 0;
            for (int64_t i = 0; i < n; i++) {
                int64_t xi;
                if (type == 1) {
                    xi = ((const int64_t *)data)[d * n + i];
                } else {
                    const double f = ((const double *)data)[d * n + i];
                    xi = (int64_t)round(f / quanta[d]);
                }
                const uint64_t nb = ((uint64_t)xi + (uint64_t)(M / 2)) & ~(uint64_t)(M - 1);
                const int64_t diff = (int64_t)(nb - branch);
                if (diff >= step || diff < -step || cnt == 0) {
                    branch = nb;
                    cnt++;
                }
                y[d * n + i] = (int32_t)((uint64_t)xi - branch);
                r[d * n + i] = (int64_t)branch;
            }
        }
        branches[d] = cnt;
    }
}

// One channel of Extract (G3SuperTimestream.cxx:762-807) over samples
// [start, stop): dest (stop - start elements of the output type) holds the
// caller's contents on entry.  flac: decoded FLAC samples of the whole
// channel or NULL; raw: the ALGO_NONE payload (full-width) or NULL;
// offs: inflated bz2 offsets of the whole channel (full-width) or NULL;
// cval: pointer to the CONST datum (full-width) or NULL.  type as above.
void oracle_g3s_combine(int type, int64_t start, int64_t stop, const int32_t *flac,
                        const void *raw, const void *offs, const void *cval, double quantum,
                        void *dest) {
    const int64_t m = stop - start;
    const bool wide = (type == 1 || type == 3);
    for (int64_t j = 0; j < m; j++) {
        const int64_t t = start + j;
        if (wide) {
            uint64_t v = ((uint64_t *)dest)[j];
            if (raw) v = ((const uint64_t *)raw)[t];
            if (flac) v = (uint64_t)(int64_t)flac[t];
            if (offs) v += ((const uint64_t *)offs)[t];
            if (cval) v += *(const uint64_t *)cval;
            if (type == 3) ((double *)dest)[j] = (double)(int64_t)v * quantum;
            else ((uint64_t *)dest)[j] = v;
        } else {
            uint32_t v = ((uint32_t *)dest)[j];
            if (raw) v = ((const uint32_t *)raw)[t];
            if (flac) v = (uint32_t)flac[t];
            if (offs) v += ((const uint32_t *)offs)[t];
            if (cval) v += *(const uint32_t *)cval;
            if (type == 2) ((float *)dest)[j] = (float)((double)(float)(int32_t)v * quantum);
            else ((uint32_t *)dest)[j] = v;
        }
    }
}

}  // extern "C"
