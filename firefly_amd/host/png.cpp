// png.cpp -- PNG decoder for textures embedded in glTF files (the reference
// decodes them through assimp / stb_image, CRT/Texture.cpp:79-112).
//
// Written from the PNG specification (chunks, colour types, the five row
// filters, Paeth predictor), RFC 1950 (zlib wrapper, Adler-32) and RFC 1951
// (stored, fixed-Huffman and dynamic-Huffman deflate blocks; canonical
// Huffman codes decoded length by length).  Supported: bit depth 8 (and 16,
// reduced to its high byte), colour types 0, 2, 3, 4, 6, no interlace.
// Output: 8-bit samples, top row first, 1-4 channels as stored (palette
// images expand to RGB or RGBA).
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "firefly.h"

namespace firefly {
namespace {

[[noreturn]] void bad(const std::string& what) { throw Error(FF_ERR_INVALID, "PNG: " + what); }

struct Bits {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t buf = 0;
    int cnt = 0;
    Bits(const uint8_t* d, size_t len) : p(d), n(len) {}
    uint32_t get(int k) {  // k <= 24 bits, least significant first
        while (cnt < k) {
            if (pos >= n) bad("deflate stream ends early");
            buf |= (uint32_t)p[pos++] << cnt;
            cnt += 8;
        }
        uint32_t v = buf & ((1u << k) - 1u);
        buf >>= k;
        cnt -= k;
        return v;
    }
    void align() {
        buf = 0;
        cnt = 0;
    }
};

// canonical Huffman code: symbols ordered by (length, value)
struct Huff {
    uint16_t count[16] = {};
    std::vector<uint16_t> sym;
    void build(const uint8_t* len, int n) {
        std::memset(count, 0, sizeof count);
        for (int i = 0; i < n; i++) count[len[i]]++;
        count[0] = 0;
        int left = 1;
        for (int l = 1; l < 16; l++) {
            left = (left << 1) - count[l];
            if (left < 0) bad("over-subscribed Huffman code");
        }
        uint16_t offs[16] = {};
        for (int l = 1; l < 15; l++) offs[l + 1] = (uint16_t)(offs[l] + count[l]);
        sym.assign((size_t)n, 0);
        for (int i = 0; i < n; i++)
            if (len[i]) sym[offs[len[i]]++] = (uint16_t)i;
    }
    int decode(Bits& b) const {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; l++) {
            code |= (int)b.get(1);
            int c = count[l];
            if (code - first < c) return sym[(size_t)(index + code - first)];
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        bad("invalid Huffman code");
    }
};

// RFC 1951 3.2.5: length and distance bases and extra bits
const uint16_t LBASE[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                            35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t LEXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t DBASE[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                            257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t DEXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflate_block(Bits& b, const Huff& lit, const Huff& dist, std::vector<uint8_t>& out, size_t limit) {
    for (;;) {
        int s = lit.decode(b);
        if (s < 256) {
            out.push_back((uint8_t)s);
        } else if (s == 256) {
            return;
        } else {
            s -= 257;
            if (s >= 29) bad("bad length code");
            size_t len = LBASE[s] + b.get(LEXTRA[s]);
            int ds = dist.decode(b);
            if (ds >= 30) bad("bad distance code");
            size_t d = DBASE[ds] + b.get(DEXTRA[ds]);
            if (d > out.size()) bad("distance beyond the output");
            size_t from = out.size() - d;
            for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);
        }
        if (out.size() > limit) bad("more image data than the header declares");
    }
}

std::vector<uint8_t> inflate_zlib(const std::vector<uint8_t>& z, size_t limit) {
    if (z.size() < 6) bad("short zlib stream");
    if ((z[0] & 0x0f) != 8 || ((z[0] << 8) | z[1]) % 31 != 0 || (z[1] & 0x20)) bad("bad zlib header");
    Bits b(z.data() + 2, z.size() - 2);
    std::vector<uint8_t> out;
    bool last = false;
    while (!last) {
        last = b.get(1) != 0;
        uint32_t type = b.get(2);
        if (type == 0) {
            b.align();
            if (b.pos + 4 > b.n) bad("short stored block");
            uint32_t len = b.p[b.pos] | (uint32_t)b.p[b.pos + 1] << 8;
            uint32_t nlen = b.p[b.pos + 2] | (uint32_t)b.p[b.pos + 3] << 8;
            if ((len ^ 0xffffu) != nlen) bad("stored block length check");
            b.pos += 4;
            if (b.pos + len > b.n) bad("short stored block");
            out.insert(out.end(), b.p + b.pos, b.p + b.pos + len);
            b.pos += len;
            if (out.size() > limit) bad("more image data than the header declares");
        } else if (type == 1) {
            uint8_t l[288], d[30];
            for (int i = 0; i < 144; i++) l[i] = 8;
            for (int i = 144; i < 256; i++) l[i] = 9;
            for (int i = 256; i < 280; i++) l[i] = 7;
            for (int i = 280; i < 288; i++) l[i] = 8;
            for (int i = 0; i < 30; i++) d[i] = 5;
            Huff lit, dist;
            lit.build(l, 288);
            dist.build(d, 30);
            inflate_block(b, lit, dist, out, limit);
        } else if (type == 2) {
            int hlit = (int)b.get(5) + 257, hdist = (int)b.get(5) + 1, hclen = (int)b.get(4) + 4;
            static const uint8_t ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (int i = 0; i < hclen; i++) cl[ORDER[i]] = (uint8_t)b.get(3);
            Huff ch;
            ch.build(cl, 19);
            uint8_t lens[320] = {};
            int k = 0;
            while (k < hlit + hdist) {
                int s = ch.decode(b);
                if (s < 16) {
                    lens[k++] = (uint8_t)s;
                } else {
                    int rep, val = 0;
                    if (s == 16) {
                        if (k == 0) bad("repeat without a previous length");
                        val = lens[k - 1];
                        rep = 3 + (int)b.get(2);
                    } else if (s == 17) {
                        rep = 3 + (int)b.get(3);
                    } else {
                        rep = 11 + (int)b.get(7);
                    }
                    if (k + rep > hlit + hdist) bad("code lengths overflow");
                    while (rep--) lens[k++] = (uint8_t)val;
                }
            }
            Huff lit, dist;
            lit.build(lens, hlit);
            dist.build(lens + hlit, hdist);
            inflate_block(b, lit, dist, out, limit);
        } else {
            bad("reserved deflate block type");
        }
    }
    // Adler-32 of the inflated data (RFC 1950)
    size_t end = b.pos;  // the checksum follows the last whole byte
    if (end + 4 > b.n) bad("missing Adler-32");
    uint32_t want = (uint32_t)b.p[end] << 24 | (uint32_t)b.p[end + 1] << 16 | (uint32_t)b.p[end + 2] << 8 | b.p[end + 3];
    uint32_t a = 1, s2 = 0;
    for (uint8_t v : out) {
        a = (a + v) % 65521u;
        s2 = (s2 + a) % 65521u;
    }
    if (((s2 << 16) | a) != want) bad("Adler-32 mismatch");
    return out;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

}  // namespace

void decode_png(const uint8_t* data, size_t n, int& width, int& height, int& channels, std::vector<uint8_t>& pixels) {
    static const uint8_t SIG[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (n < 8 || std::memcmp(data, SIG, 8) != 0) bad("not a PNG file");
    size_t off = 8;
    int depth = 0, ctype = -1;
    std::vector<uint8_t> idat, plte, trns;
    bool ended = false;
    while (off + 8 <= n && !ended) {
        uint32_t len = be32(data + off);
        const uint8_t* type = data + off + 4;
        if (off + 12 + (size_t)len > n) bad("chunk outside the file");
        const uint8_t* body = data + off + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) bad("bad IHDR");
            width = (int)be32(body);
            height = (int)be32(body + 4);
            depth = body[8];
            ctype = body[9];
            if (body[10] != 0 || body[11] != 0) bad("unknown compression or filter method");
            if (body[12] != 0) bad("interlaced images are not supported");
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            ended = true;
        }
        off += 12 + (size_t)len;
    }
    if (ctype < 0) bad("no IHDR");
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768) bad("image size");
    if (depth != 8 && depth != 16) bad("bit depth " + std::to_string(depth) + " is not supported");
    int spp;  // samples per pixel as stored
    switch (ctype) {
        case 0: spp = 1; break;
        case 2: spp = 3; break;
        case 3: spp = 1; if (depth != 8) bad("16-bit palette"); break;
        case 4: spp = 2; break;
        case 6: spp = 4; break;
        default: bad("colour type " + std::to_string(ctype));
    }
    size_t bpp = (size_t)spp * (size_t)(depth / 8);  // bytes per pixel
    size_t stride = bpp * (size_t)width;
    std::vector<uint8_t> raw = inflate_zlib(idat, (stride + 1) * (size_t)height);
    if (raw.size() != (stride + 1) * (size_t)height) bad("image data size");
    std::vector<uint8_t> img(stride * (size_t)height);
    for (int y = 0; y < height; y++) {
        const uint8_t* src = raw.data() + (size_t)y * (stride + 1);
        uint8_t f = src[0];
        src++;
        uint8_t* dst = img.data() + (size_t)y * stride;
        const uint8_t* up = y ? dst - stride : nullptr;
        for (size_t i = 0; i < stride; i++) {
            int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            int pred;
            switch (f) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: {  // Paeth
                    int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
                    pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: bad("row filter " + std::to_string(f));
            }
            dst[i] = (uint8_t)(src[i] + pred);
        }
    }
    // 8-bit output; palette entries expand to RGB(A)
    bool alpha_pal = ctype == 3 && !trns.empty();
    channels = ctype == 3 ? (alpha_pal ? 4 : 3) : spp;
    pixels.assign((size_t)width * (size_t)height * (size_t)channels, 0);
    for (size_t px = 0; px < (size_t)width * (size_t)height; px++) {
        if (ctype == 3) {
            size_t e = img[px];
            if (3 * e + 2 >= plte.size()) bad("palette index out of range");
            for (int k = 0; k < 3; k++) pixels[px * (size_t)channels + (size_t)k] = plte[3 * e + (size_t)k];
            if (alpha_pal) pixels[px * 4 + 3] = e < trns.size() ? trns[e] : 255;
        } else {
            for (int k = 0; k < spp; k++) pixels[px * (size_t)spp + (size_t)k] = img[px * bpp + (size_t)k * (size_t)(depth / 8)];
        }
    }
}

}  // namespace firefly
