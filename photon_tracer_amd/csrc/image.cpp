// Texture images from file bytes without an image library: PNG (not
// interlaced; 8 or 16 bits a sample; gray, gray + alpha, RGB, RGBA, 8-bit
// palette; alpha dropped, 16-bit samples cut to their high byte) and PPM (P3,
// P6; maximum value up to 255), the formats and results of the reference's
// decode_image (image_io.hpp:109-297).
//
// The zlib stream inside PNG is inflated here (RFC 1950/1951): the bit
// stream is read from a 64-bit buffer, and a Huffman code of up to 10 bits
// (nearly every symbol) is decoded by one lookup in a table over the next 10
// bits rather than by a walk over its bits. The output size is known from
// the PNG header and reserved up front.
#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "host.hpp"

namespace pt {
namespace {

[[noreturn]] void bad(const char* what) { throw std::runtime_error(what); }

// ---- Inflate ---------------------------------------------------------------

class bit_source {
public:
    bit_source(const uint8_t* data, size_t size) : data_(data), size_(size) {}

    // The next n (<= 32) bits, least significant first, without consuming
    // them; bits past the end read as 0 until consumed.
    uint32_t peek(int n) {
        while (count_ <= 56) {
            const uint64_t byte = next_ < size_ ? data_[next_] : 0;
            buf_ |= byte << count_;
            ++next_;
            count_ += 8;
        }
        return static_cast<uint32_t>(buf_ & ((uint64_t(1) << n) - 1));
    }
    void drop(int n) {
        buf_ >>= n;
        count_ -= n;
        used_ += static_cast<uint64_t>(n);
        if (used_ > 8 * static_cast<uint64_t>(size_)) bad("inflate: input ended early");
    }
    uint32_t take(int n) {
        if (n == 0) return 0;
        const uint32_t v = peek(n);
        drop(n);
        return v;
    }
    // Stored blocks start on a byte boundary: the rest of the current byte is
    // skipped and reading restarts at the next whole byte.
    size_t align_to_byte() {
        used_ = (used_ + 7) & ~uint64_t(7);
        buf_ = 0;
        count_ = 0;
        next_ = static_cast<size_t>(used_ / 8);
        return next_;
    }
    void skip_bytes(size_t n) {
        used_ += 8 * static_cast<uint64_t>(n);
        next_ += n;
    }
    const uint8_t* data() const { return data_; }
    size_t size() const { return size_; }

private:
    const uint8_t* data_;
    size_t size_;
    size_t next_ = 0;      // next byte to load into buf_
    uint64_t buf_ = 0;
    int count_ = 0;        // bits in buf_
    uint64_t used_ = 0;    // bits consumed in total
};

// A canonical Huffman code, decoded by one lookup in a table over the next
// kFast bits of input (entry = symbol << 4 | code length; 0: a longer code
// starts with these bits) and, for the rare codes longer than that, by
// walking the code lengths from the first bit (RFC 1951 3.2.2's canonical
// order: codes of one length are consecutive, after all shorter ones).
constexpr int kMaxBits = 15, kFast = 10;

class huffman_table {
public:
    void build(const uint8_t* lengths, int n) {
        for (auto& c : count_) c = 0;
        for (int i = 0; i < n; ++i) ++count_[lengths[i]];
        count_[0] = 0;
        // Kraft: more codes of a length than the shorter ones leave room for is no prefix code.
        int left = 1;
        for (int len = 1; len <= kMaxBits; ++len) {
            left = 2 * left - count_[len];
            if (left < 0) bad("inflate: over-subscribed Huffman code");
        }
        int offset[kMaxBits + 2] = {}, next_code[kMaxBits + 2] = {};
        for (int len = 1, code = 0; len <= kMaxBits; ++len) {
            offset[len + 1] = offset[len] + count_[len];
            code = (code + count_[len - 1]) << 1;
            next_code[len] = code;
        }
        symbol_.assign(static_cast<size_t>(offset[kMaxBits + 1]), 0);
        for (auto& e : fast_) e = 0;
        for (int sym = 0; sym < n; ++sym) {
            const int len = lengths[sym];
            if (!len) continue;
            symbol_[static_cast<size_t>(offset[len]++)] = static_cast<uint16_t>(sym);
            const uint32_t c = static_cast<uint32_t>(next_code[len]++);
            if (len > kFast) continue;
            // Deflate sends codes most significant bit first; the table is
            // indexed by bits in arrival order, so the code is reversed.
            uint32_t r = 0;
            for (int b = 0; b < len; ++b) r = (r << 1) | ((c >> b) & 1);
            const uint32_t entry = static_cast<uint32_t>(sym) << 4 | static_cast<uint32_t>(len);
            for (uint32_t k = r; k < (1u << kFast); k += 1u << len) fast_[k] = entry;
        }
    }
    int decode(bit_source& in) const {
        const uint32_t bits = in.peek(kMaxBits);
        const uint32_t e = fast_[bits & ((1u << kFast) - 1)];
        if (e) {
            in.drop(static_cast<int>(e & 15));
            return static_cast<int>(e >> 4);
        }
        int code = 0, first = 0, index = 0;
        for (int len = 1; len <= kMaxBits; ++len) {
            code |= static_cast<int>((bits >> (len - 1)) & 1);
            const int cnt = count_[len];
            if (code - first < cnt) {
                in.drop(len);
                return symbol_[static_cast<size_t>(index + code - first)];
            }
            index += cnt;
            first = (first + cnt) << 1;
            code <<= 1;
        }
        bad("inflate: bad Huffman code");
    }

private:
    uint16_t count_[kMaxBits + 1] = {};
    std::vector<uint16_t> symbol_;   // symbols in code order
    uint32_t fast_[1u << kFast] = {};
};

// RFC 1951 3.2.5: base lengths and extra bits of length symbols 257-285, and
// base distances and extra bits of distance symbols 0-29.
constexpr uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                    193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

// Output grows without a bound from the stream itself; this one stops a
// corrupt or hostile stream before it takes the host's memory.
constexpr size_t kMaxInflated = size_t(1) << 36;

void inflate_codes(bit_source& in, std::vector<uint8_t>& out, const huffman_table& lit, const huffman_table& dist) {
    for (;;) {
        const int sym = lit.decode(in);
        if (sym < 256) {
            out.push_back(static_cast<uint8_t>(sym));
            continue;
        }
        if (sym == 256) return;
        if (sym > 285) bad("inflate: bad length symbol");
        const size_t len = kLenBase[sym - 257] + in.take(kLenExtra[sym - 257]);
        const int d = dist.decode(in);
        if (d > 29) bad("inflate: bad distance symbol");
        const size_t back = kDistBase[d] + in.take(kDistExtra[d]);
        if (back > out.size()) bad("inflate: distance before the start");
        if (out.size() + len > kMaxInflated) bad("inflate: output too large");
        const size_t from = out.size() - back;
        // The copy may overlap what it writes (back < len repeats the last back bytes).
        for (size_t k = 0; k < len; ++k) out.push_back(out[from + k]);
    }
}

// A zlib stream (RFC 1950): two header bytes, deflate blocks; the Adler-32
// after them is not checked (nor by the reference).
std::vector<uint8_t> inflate_zlib(const uint8_t* data, size_t size, size_t expect) {
    if (size < 2 || (data[0] & 0x0f) != 8 || ((data[0] << 8) | data[1]) % 31 != 0) bad("zlib: bad header");
    bit_source in{data + 2, size - 2};
    std::vector<uint8_t> out;
    out.reserve(std::min(expect, kMaxInflated));
    huffman_table lit, dist;
    bool last;
    do {
        last = in.take(1) != 0;
        const uint32_t type = in.take(2);
        if (type == 0) {   // stored
            const size_t at = in.align_to_byte();
            if (at + 4 > in.size()) bad("inflate: stored block cut short");
            const size_t len = in.data()[at] | (size_t(in.data()[at + 1]) << 8);
            if (at + 4 + len > in.size()) bad("inflate: stored block cut short");
            if (out.size() + len > kMaxInflated) bad("inflate: output too large");
            out.insert(out.end(), in.data() + at + 4, in.data() + at + 4 + len);
            in.skip_bytes(4 + len);
        } else if (type == 1) {   // fixed codes (RFC 1951 3.2.6)
            uint8_t lengths[288];
            std::fill(lengths, lengths + 144, 8);
            std::fill(lengths + 144, lengths + 256, 9);
            std::fill(lengths + 256, lengths + 280, 7);
            std::fill(lengths + 280, lengths + 288, 8);
            lit.build(lengths, 288);
            std::fill(lengths, lengths + 30, 5);
            dist.build(lengths, 30);
            inflate_codes(in, out, lit, dist);
        } else if (type == 2) {   // dynamic codes, their lengths sent first under a code of their own
            const int nlen = static_cast<int>(in.take(5)) + 257, ndist = static_cast<int>(in.take(5)) + 1;
            const int ncode = static_cast<int>(in.take(4)) + 4;
            static constexpr uint8_t kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t lengths[320] = {};
            for (int i = 0; i < ncode; ++i) lengths[kOrder[i]] = static_cast<uint8_t>(in.take(3));
            huffman_table cl;
            cl.build(lengths, 19);
            std::fill(lengths, lengths + 19, 0);
            for (int i = 0; i < nlen + ndist;) {
                const int sym = cl.decode(in);
                if (sym < 16) {
                    lengths[i++] = static_cast<uint8_t>(sym);
                    continue;
                }
                uint8_t value = 0;
                int repeat;
                if (sym == 16) {
                    if (i == 0) bad("inflate: repeat with nothing before it");
                    value = lengths[i - 1];
                    repeat = 3 + static_cast<int>(in.take(2));
                } else if (sym == 17) {
                    repeat = 3 + static_cast<int>(in.take(3));
                } else {
                    repeat = 11 + static_cast<int>(in.take(7));
                }
                if (i + repeat > nlen + ndist) bad("inflate: too many code lengths");
                std::fill(lengths + i, lengths + i + repeat, value);
                i += repeat;
            }
            lit.build(lengths, nlen);
            dist.build(lengths + nlen, ndist);
            inflate_codes(in, out, lit, dist);
        } else {
            bad("inflate: bad block type");
        }
    } while (!last);
    return out;
}

// ---- PNG -------------------------------------------------------------------

uint32_t be32(const uint8_t* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
}

// The PNG filter type 4 predictor (PNG specification 9.4).
int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

constexpr uint32_t kMaxSide = 65536;

image_rec decode_png(const uint8_t* d, size_t size) {
    static constexpr uint8_t kSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (size < 8 || std::memcmp(d, kSignature, 8) != 0) bad("png: not a PNG");
    image_rec img;
    int depth = 0, type = 0;
    bool header = false;
    std::vector<uint8_t> idat, palette;
    for (size_t pos = 8; pos + 12 <= size;) {
        const uint32_t len = be32(d + pos);
        const uint8_t* tag = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if (pos + 12 + uint64_t(len) > size) bad("png: chunk runs past the end");
        if (!std::memcmp(tag, "IHDR", 4)) {
            if (len < 13) bad("png: IHDR cut short");
            img.width = be32(body);
            img.height = be32(body + 4);
            depth = body[8];
            type = body[9];
            header = true;
            if (body[12] != 0) bad("png: interlaced images are not supported");
            if (depth != 8 && depth != 16) bad("png: only 8 and 16 bit images are supported");
        } else if (!std::memcmp(tag, "PLTE", 4)) {
            palette.assign(body, body + len);
        } else if (!std::memcmp(tag, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(tag, "IEND", 4)) {
            break;
        }
        pos += 12 + size_t(len);
    }
    if (!header) bad("png: no IHDR");
    static constexpr int kChannels[7] = {1, 0, 3, 1, 2, 0, 4};   // by color type; 0: none
    if (type > 6 || kChannels[type] == 0) bad("png: unknown color type");
    if (type == 3 && depth != 8) bad("png: only 8 bit palettes are supported");
    if (img.width < 1 || img.height < 1 || img.width > kMaxSide || img.height > kMaxSide)
        bad("png: image size out of range (1 to 65536 a side)");
    const int channels = kChannels[type];
    const size_t bpp = size_t(channels) * depth / 8;   // bytes a pixel
    const size_t stride = bpp * img.width;
    const size_t need = (stride + 1) * img.height;
    const std::vector<uint8_t> raw = inflate_zlib(idat.data(), idat.size(), need);
    if (raw.size() < need) bad("png: image data cut short");

    // Each row's filter predicted every byte from its left (a), upper (b)
    // and upper-left (c) neighbours a pixel back; undo it in place of a copy.
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    img.rgb.resize(3 * size_t(img.width) * img.height);
    const int step = depth / 8;   // bytes a sample; the high byte comes first
    for (uint32_t y = 0; y < img.height; ++y) {
        const uint8_t* row = &raw[y * (stride + 1)];
        const uint8_t* in = row + 1;
        switch (row[0]) {
            case 0: std::memcpy(cur.data(), in, stride); break;
            case 1:
                for (size_t x = 0; x < stride; ++x) cur[x] = uint8_t(in[x] + (x >= bpp ? cur[x - bpp] : 0));
                break;
            case 2:
                for (size_t x = 0; x < stride; ++x) cur[x] = uint8_t(in[x] + prev[x]);
                break;
            case 3:
                for (size_t x = 0; x < stride; ++x) cur[x] = uint8_t(in[x] + ((x >= bpp ? cur[x - bpp] : 0) + prev[x]) / 2);
                break;
            case 4:
                for (size_t x = 0; x < stride; ++x)
                    cur[x] = uint8_t(in[x] + (x >= bpp ? paeth(cur[x - bpp], prev[x], prev[x - bpp]) : paeth(0, prev[x], 0)));
                break;
            default: bad("png: bad filter");
        }
        uint8_t* o = &img.rgb[3 * size_t(y) * img.width];
        for (uint32_t x = 0; x < img.width; ++x, o += 3) {
            const uint8_t* px = &cur[x * bpp];
            if (type == 3) {
                const size_t k = 3 * size_t(px[0]);
                if (k + 2 >= palette.size()) bad("png: palette index out of range");
                o[0] = palette[k];
                o[1] = palette[k + 1];
                o[2] = palette[k + 2];
            } else if (channels >= 3) {
                o[0] = px[0];
                o[1] = px[step];
                o[2] = px[2 * step];
            } else {
                o[0] = o[1] = o[2] = px[0];
            }
        }
        std::swap(prev, cur);
    }
    return img;
}

// ---- PPM -------------------------------------------------------------------

class ppm_reader {
public:
    ppm_reader(const uint8_t* d, size_t n) : d_(d), n_(n) {}

    static bool space(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }
    // The magic: the first run of non-space bytes.
    std::string token() {
        skip_space();
        const size_t from = at_;
        while (at_ < n_ && !space(d_[at_])) ++at_;
        return std::string(reinterpret_cast<const char*>(d_ + from), at_ - from);
    }
    // The next decimal number, skipping white space and # comments to the
    // end of their line.
    long number() {
        for (;;) {
            skip_space();
            if (at_ < n_ && d_[at_] == '#') {
                while (at_ < n_ && d_[at_] != '\n') ++at_;
                if (at_ < n_) ++at_;
                continue;
            }
            break;
        }
        bool neg = false;
        if (at_ < n_ && (d_[at_] == '+' || d_[at_] == '-')) neg = d_[at_++] == '-';
        if (at_ >= n_ || d_[at_] < '0' || d_[at_] > '9') bad("ppm: bad header");
        long v = 0;
        while (at_ < n_ && d_[at_] >= '0' && d_[at_] <= '9') {
            if (v > (std::numeric_limits<long>::max() - 9) / 10) bad("ppm: bad header");
            v = 10 * v + (d_[at_++] - '0');
        }
        return neg ? -v : v;
    }
    size_t at() const { return at_; }
    void skip(size_t k) { at_ += k; }
    size_t left() const { return at_ < n_ ? n_ - at_ : 0; }
    const uint8_t* here() const { return d_ + at_; }

private:
    void skip_space() {
        while (at_ < n_ && space(d_[at_])) ++at_;
    }
    const uint8_t* d_;
    size_t n_, at_ = 0;
};

image_rec decode_ppm(const uint8_t* d, size_t size) {
    ppm_reader in{d, size};
    const std::string magic = in.token();
    if (magic != "P3" && magic != "P6") bad("ppm: only P3 and P6 are supported");
    const long w = in.number(), h = in.number();
    if (w < 1 || h < 1 || w > long(kMaxSide) || h > long(kMaxSide)) bad("ppm: image size out of range (1 to 65536 a side)");
    const long max = in.number();
    if (max <= 0 || max > 255) bad("ppm: only 8 bit images are supported");
    image_rec img;
    img.width = static_cast<uint32_t>(w);
    img.height = static_cast<uint32_t>(h);
    const size_t n = 3 * size_t(w) * size_t(h);
    img.rgb.resize(n);
    if (magic == "P6") {
        in.skip(1);   // the one white-space byte after the header
        if (in.left() < n) bad("ppm: pixel data cut short");
        std::memcpy(img.rgb.data(), in.here(), n);
    } else {
        for (auto& v : img.rgb) v = static_cast<uint8_t>(in.number());   // kept modulo 256, as a byte
    }
    if (max != 255)
        for (auto& v : img.rgb) v = static_cast<uint8_t>(v * 255 / max);
    return img;
}

}  // namespace

image_rec decode_image(const uint8_t* bytes, size_t size) {
    if (size >= 2 && bytes[0] == 'P' && (bytes[1] == '3' || bytes[1] == '6')) return decode_ppm(bytes, size);
    return decode_png(bytes, size);
}

}  // namespace pt
