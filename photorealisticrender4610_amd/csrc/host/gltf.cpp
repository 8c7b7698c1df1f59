// gltf.cpp — glTF 2.0 (.glb / .gltf) scene input: GLTFLoader::load
// (GLTFLoader.cpp:140-281) with its material and texture rules
// (GLTFLoader.cpp:64-138, TextureUtils.hpp:9-19), behind prr_gltf_load.
//
// Own parsers, no third-party code: a small JSON reader, the GLB container,
// buffers (GLB BIN chunk, files next to the asset, base64 data: URIs),
// accessors, a PNG decoder on zlib's inflate and a JPEG decoder (jpeg.cpp).
// Images are decoded to four channels as the reference's loader requests them
// (req_comp 4, 16-bit images kept at 16 bits, then reduced as
// GLTFLoader.cpp:77-137 does per texture slot).
#include <zlib.h>

#include "jpeg.hpp"

#include <charconv>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "prr_abi.h"

int fail(prr_context *c, int code, const std::string &msg);

namespace {

// ---------------------------------------------------------------------------
// JSON
// ---------------------------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json *get(const char *k) const {
        if (kind != Obj) return nullptr;
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    double numOr(const char *k, double d) const {
        const Json *v = get(k);
        return v && v->kind == Num ? v->num : d;
    }
    int intOr(const char *k, int d) const {
        const double x = numOr(k, d);
        // an out-of-range or non-integral index reads as "none" (-1)
        return (std::isfinite(x) && x >= -2147483648.0 && x <= 2147483647.0 && x == std::floor(x)) ? (int) x : -1;
    }
    // A byte count or offset: absent = 0; a negative, non-integral or
    // non-finite value, or one above `limit`, is an error (the value is
    // converted to size_t only once it is known to fit).
    size_t sizeOr0(const char *k, size_t limit) const {
        const double x = numOr(k, 0.0);
        if (!(std::isfinite(x) && x >= 0.0 && x == std::floor(x) && x <= (double) limit))
            throw std::runtime_error(std::string("glTF: bad ") + k);
        return (size_t) x;
    }
    const Json &at(size_t i) const {
        if (kind != Arr || i >= arr.size()) throw std::runtime_error("glTF: index out of range");
        return arr[i];
    }
    size_t size() const { return kind == Arr ? arr.size() : 0; }
};

struct JsonReader {
    const char *p, *e;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    [[noreturn]] void bad(const char *what) { throw std::runtime_error(std::string("glTF JSON: ") + what); }
    void expect(char c) {
        ws();
        if (p >= e || *p != c) bad("unexpected character");
        p++;
    }
    static void utf8(std::string &s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char) cp;
        } else if (cp < 0x800) {
            s += (char) (0xC0 | (cp >> 6));
            s += (char) (0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char) (0xE0 | (cp >> 12));
            s += (char) (0x80 | ((cp >> 6) & 0x3F));
            s += (char) (0x80 | (cp & 0x3F));
        } else {
            s += (char) (0xF0 | (cp >> 18));
            s += (char) (0x80 | ((cp >> 12) & 0x3F));
            s += (char) (0x80 | ((cp >> 6) & 0x3F));
            s += (char) (0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (e - p < 4) bad("short \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; i++) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t) (c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t) (c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t) (c - 'A' + 10);
            else bad("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        expect('"');
        std::string s;
        while (true) {
            if (p >= e) bad("unterminated string");
            const char c = *p++;
            if (c == '"') break;
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p >= e) bad("unterminated escape");
            const char x = *p++;
            switch (x) {
                case '"': s += '"'; break;
                case '\\': s += '\\'; break;
                case '/': s += '/'; break;
                case 'b': s += '\b'; break;
                case 'f': s += '\f'; break;
                case 'n': s += '\n'; break;
                case 'r': s += '\r'; break;
                case 't': s += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(s, cp);
                    break;
                }
                default: bad("bad escape");
            }
        }
        return s;
    }
    Json value(int depth) {
        if (depth > 256) bad("nesting too deep");
        ws();
        if (p >= e) bad("unexpected end");
        Json v;
        const char c = *p;
        if (c == '{') {
            p++;
            v.kind = Json::Obj;
            ws();
            if (p < e && *p == '}') {
                p++;
                return v;
            }
            while (true) {
                std::string k = string();
                expect(':');
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    p++;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            p++;
            v.kind = Json::Arr;
            ws();
            if (p < e && *p == ']') {
                p++;
                return v;
            }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') {
                    p++;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            v.kind = Json::Str;
            v.str = string();
            return v;
        }
        if (e - p >= 4 && !strncmp(p, "true", 4)) {
            p += 4;
            v.kind = Json::Bool;
            v.b = true;
            return v;
        }
        if (e - p >= 5 && !strncmp(p, "false", 5)) {
            p += 5;
            v.kind = Json::Bool;
            return v;
        }
        if (e - p >= 4 && !strncmp(p, "null", 4)) {
            p += 4;
            return v;
        }
        // number: JSON's grammar characters only (a NUL byte ends it), read by
        // std::from_chars, which does not depend on the process's locale
        const char *s = p;
        while (p < e && *p != '\0' && strchr("+-0123456789.eE", *p) != nullptr) p++;
        if (p == s) bad("unexpected character");
        v.kind = Json::Num;
        const char *b = (*s == '+') ? s + 1 : s;  // from_chars takes no leading '+'
        const auto r = std::from_chars(b, p, v.num);
        if (r.ec != std::errc() || r.ptr != p) bad("bad number");
        return v;
    }
};

Json parseJson(const char *p, size_t n) {
    JsonReader r{p, p + n};
    Json v = r.value(0);
    r.ws();
    if (r.p != r.e && *r.p != '\0') r.bad("trailing characters");
    return v;
}

// ---------------------------------------------------------------------------
// files, base64
// ---------------------------------------------------------------------------
std::vector<uint8_t> readFile(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("glTF: cannot open " + path);
    std::vector<uint8_t> out;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + k);
    fclose(f);
    return out;
}

std::vector<uint8_t> base64(const std::string &s, size_t from) {
    static int8_t tab[256];
    static bool init = false;
    if (!init) {
        memset(tab, -1, sizeof(tab));
        const char *a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; i++) tab[(uint8_t) a[i]] = (int8_t) i;
        init = true;
    }
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); i++) {
        const uint8_t c = (uint8_t) s[i];
        if (c == '=') break;
        if (tab[c] < 0) {
            if (c == '\n' || c == '\r' || c == ' ') continue;
            throw std::runtime_error("glTF: bad base64 data");
        }
        acc = (acc << 6) | (uint32_t) tab[c];
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t) (acc >> bits));
        }
    }
    return out;
}

std::string dirOf(const std::string &path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

std::string uriDecode(const std::string &u) {
    std::string s;
    for (size_t i = 0; i < u.size(); i++) {
        if (u[i] == '%' && i + 2 < u.size()) {
            s += (char) strtol(u.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            s += u[i];
        }
    }
    return s;
}

std::vector<uint8_t> loadUri(const std::string &uri, const std::string &dir) {
    if (uri.compare(0, 5, "data:") == 0) {
        const size_t k = uri.find(";base64,");
        if (k == std::string::npos) throw std::runtime_error("glTF: data URI is not base64");
        return base64(uri, k + 8);
    }
    return readFile(dir + uriDecode(uri));
}

// ---------------------------------------------------------------------------
// PNG (ISO/IEC 15948): every colour type, bit depths 1-16, no interlace;
// output RGBA at 8 or 16 bits per channel as stb_image returns it for a
// 4-channel request (grey g -> g,g,g,max; tRNS key colours -> alpha 0).
// ---------------------------------------------------------------------------
uint32_t be32(const uint8_t *p) { return (uint32_t) p[0] << 24 | (uint32_t) p[1] << 16 | (uint32_t) p[2] << 8 | p[3]; }

struct Image {
    int w = 0, h = 0, bits = 8;
    std::vector<uint16_t> px;  // w * h * 4
};

Image decodePng(const uint8_t *p, size_t n) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || memcmp(p, sig, 8)) throw std::runtime_error("glTF: unknown image format (PNG or JPEG expected)");
    size_t off = 8;
    int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    while (off + 8 <= n) {
        const uint32_t len = be32(p + off);
        const char *type = (const char *) p + off + 4;
        if (off + 12 + (size_t) len > n) throw std::runtime_error("PNG: truncated chunk");
        const uint8_t *d = p + off + 8;
        if (!memcmp(type, "IHDR", 4)) {
            if (len < 13) throw std::runtime_error("PNG: bad IHDR");
            w = (int) be32(d);
            h = (int) be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        off += 12 + (size_t) len;
    }
    const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool okDepth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                         (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                         ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (w <= 0 || h <= 0 || !chans || !okDepth) throw std::runtime_error("PNG: unsupported header");
    if ((int64_t) w * h > (int64_t) 1 << 28) throw std::runtime_error("PNG: image too large");
    if (interlace) throw std::runtime_error("PNG: interlaced images are not supported");
    if (ctype == 3 && plte.empty()) throw std::runtime_error("PNG: palette missing");
    const size_t bitsPx = (size_t) chans * depth;
    const size_t rowBytes = ((size_t) w * bitsPx + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bitsPx / 8);
    std::vector<uint8_t> raw((rowBytes + 1) * (size_t) h);
    {
        z_stream zs{};
        if (inflateInit(&zs) != Z_OK) throw std::runtime_error("PNG: inflateInit failed");
        zs.next_in = idat.data();
        zs.avail_in = (uInt) idat.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt) raw.size();
        const int r = inflate(&zs, Z_FINISH);
        const size_t got = raw.size() - zs.avail_out;
        inflateEnd(&zs);
        if ((r != Z_STREAM_END && r != Z_BUF_ERROR && r != Z_OK) || got != raw.size())
            throw std::runtime_error("PNG: corrupt image data");
    }
    // unfilter in place
    std::vector<uint8_t> img(rowBytes * (size_t) h);
    for (int y = 0; y < h; y++) {
        const uint8_t ft = raw[(size_t) y * (rowBytes + 1)];
        const uint8_t *src = raw.data() + (size_t) y * (rowBytes + 1) + 1;
        uint8_t *dst = img.data() + (size_t) y * rowBytes;
        const uint8_t *up = y ? dst - rowBytes : nullptr;
        for (size_t i = 0; i < rowBytes; i++) {
            const int a = i >= bpp ? dst[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= bpp) ? up[i - bpp] : 0;
            int v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {
                    const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: throw std::runtime_error("PNG: bad filter type");
            }
            dst[i] = (uint8_t) v;
        }
    }
    Image out;
    out.w = w;
    out.h = h;
    out.bits = depth == 16 ? 16 : 8;
    out.px.resize((size_t) w * h * 4);
    const int maxv = depth == 16 ? 65535 : 255;
    auto sample = [&](const uint8_t *row, int x, int c) -> int {  // channel c of pixel x, native depth
        if (depth == 16) {
            const uint8_t *q = row + ((size_t) x * chans + c) * 2;
            return q[0] << 8 | q[1];
        }
        if (depth == 8) return row[(size_t) x * chans + c];
        const size_t bit = (size_t) x * depth;  // one channel (grey or palette index)
        return (row[bit >> 3] >> (8 - depth - (int) (bit & 7))) & ((1 << depth) - 1);
    };
    // grey of depth < 8 scales to 8 bits (stb: x 0xff / 0x55 / 0x11)
    const int greyScale = (ctype == 0 && depth < 8) ? (depth == 1 ? 0xFF : depth == 2 ? 0x55 : 0x11) : 1;
    int key[3] = {-1, -1, -1};
    if (!trns.empty() && ctype == 0 && trns.size() >= 2) key[0] = (trns[0] << 8 | trns[1]) * greyScale;
    if (!trns.empty() && ctype == 2 && trns.size() >= 6)
        for (int c = 0; c < 3; c++) key[c] = trns[2 * c] << 8 | trns[2 * c + 1];
    if (depth == 8 && ctype != 3)
        for (int c = 0; c < 3; c++)
            if (key[c] >= 0) key[c] &= 0xFF;
    for (int y = 0; y < h; y++) {
        const uint8_t *row = img.data() + (size_t) y * rowBytes;
        uint16_t *o = out.px.data() + (size_t) y * w * 4;
        for (int x = 0; x < w; x++, o += 4) {
            int r, g, b, a = maxv;
            if (ctype == 3) {
                const int i = sample(row, x, 0);
                if ((size_t) i * 3 + 2 >= plte.size()) throw std::runtime_error("PNG: palette index out of range");
                r = plte[3 * i];
                g = plte[3 * i + 1];
                b = plte[3 * i + 2];
                a = (size_t) i < trns.size() ? trns[i] : 255;
            } else if (ctype == 0 || ctype == 4) {
                r = g = b = sample(row, x, 0) * greyScale;
                if (ctype == 4) a = sample(row, x, 1);
                else if (key[0] >= 0 && r == key[0]) a = 0;
            } else {
                r = sample(row, x, 0);
                g = sample(row, x, 1);
                b = sample(row, x, 2);
                if (ctype == 6) a = sample(row, x, 3);
                else if (key[0] >= 0 && r == key[0] && g == key[1] && b == key[2]) a = 0;
            }
            o[0] = (uint16_t) r;
            o[1] = (uint16_t) g;
            o[2] = (uint16_t) b;
            o[3] = (uint16_t) a;
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// accessors (GLTFLoader.cpp:13-62)
// ---------------------------------------------------------------------------
struct Model {
    Json j;
    std::vector<std::vector<uint8_t>> buffers;
};

int numComponents(const std::string &t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    throw std::runtime_error("glTF: bad accessor type " + t);
}

int componentSize(int ct) {
    switch (ct) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
        default: throw std::runtime_error("glTF: bad component type");
    }
}

struct View {
    const uint8_t *base;
    size_t count, stride;
    int nc, ct;
};

View accessorView(const Model &m, int idx) {
    const Json *accs = m.j.get("accessors");
    if (!accs || idx < 0 || (size_t) idx >= accs->size()) throw std::runtime_error("glTF: accessor out of range");
    const Json &acc = accs->at((size_t) idx);
    const int bvi = acc.intOr("bufferView", -1);
    const Json *bvs = m.j.get("bufferViews");
    if (bvi < 0 || !bvs || (size_t) bvi >= bvs->size())
        throw std::runtime_error("glTF: accessor without a buffer view (sparse accessors are not supported)");
    const Json &bv = bvs->at((size_t) bvi);
    const int bi = bv.intOr("buffer", -1);
    if (bi < 0 || (size_t) bi >= m.buffers.size()) throw std::runtime_error("glTF: buffer out of range");
    const Json *type = acc.get("type");
    View v;
    v.nc = numComponents(type && type->kind == Json::Str ? type->str : "");
    v.ct = acc.intOr("componentType", 0);
    const int cs = componentSize(v.ct);
    const auto &buf = m.buffers[(size_t) bi];
    // every offset, length and count is checked against the buffer before
    // any arithmetic, and the arithmetic itself for overflow
    const size_t bvOff = bv.sizeOr0("byteOffset", buf.size());
    const size_t bvLen = bv.sizeOr0("byteLength", buf.size() - bvOff);
    const size_t accOff = acc.sizeOr0("byteOffset", bvLen);
    v.count = acc.sizeOr0("count", buf.size());
    const size_t bvStride = bv.sizeOr0("byteStride", 255);
    v.stride = bvStride ? bvStride : (size_t) v.nc * cs;
    const size_t start = bvOff + accOff, bvEnd = bvOff + bvLen;  // both <= buf.size()
    size_t need = start, span = 0;
    if (v.count) {
        if (__builtin_mul_overflow(v.count - 1, v.stride, &span) || __builtin_add_overflow(span, start, &need) ||
            __builtin_add_overflow(need, (size_t) v.nc * cs, &need))
            throw std::runtime_error("glTF: accessor exceeds its buffer");
    }
    if (need > buf.size() || need > bvEnd) throw std::runtime_error("glTF: accessor exceeds its buffer");
    v.base = buf.data() + start;
    return v;
}

// tg_readFloats: float as is, unsigned short / 65535, unsigned byte / 255,
// any other component type reads as 0.
std::vector<float> readFloats(const Model &m, int idx) {
    const View v = accessorView(m, idx);
    std::vector<float> out;
    out.reserve(v.count * v.nc);
    for (size_t i = 0; i < v.count; i++) {
        const uint8_t *e = v.base + i * v.stride;
        for (int c = 0; c < v.nc; c++) {
            float x = 0;
            if (v.ct == 5126) {
                memcpy(&x, e + c * 4, 4);
            } else if (v.ct == 5123) {
                uint16_t u;
                memcpy(&u, e + c * 2, 2);
                x = u / 65535.0f;
            } else if (v.ct == 5121) {
                x = e[c] / 255.0f;
            }
            out.push_back(x);
        }
    }
    return out;
}

// tg_readIndices: unsigned int, unsigned short, anything else as bytes
// (tightly packed: the reference ignores the view's stride for indices).
std::vector<uint32_t> readIndices(const Model &m, int idx) {
    View v = accessorView(m, idx);
    std::vector<uint32_t> out(v.count);
    for (size_t i = 0; i < v.count; i++) {
        if (v.ct == 5125) {
            uint32_t x;
            memcpy(&x, v.base + i * 4, 4);
            out[i] = x;
        } else if (v.ct == 5123) {
            uint16_t x;
            memcpy(&x, v.base + i * 2, 2);
            out[i] = x;
        } else {
            out[i] = v.base[i];
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// loader
// ---------------------------------------------------------------------------
struct Loaded {
    prr_gltf_data d{};
    std::vector<int32_t> meshFirst, meshCount, meshMat;
    std::vector<float> pos, nrm, uv, tan;
    std::vector<uint8_t> flags;
    std::vector<prr_material_desc> mats;
    std::vector<int32_t> matTex;
    std::vector<std::vector<uint8_t>> texData;
    std::vector<prr_texture_desc> tex;
};

Model openModel(const std::string &path) {
    Model m;
    std::vector<uint8_t> file = readFile(path);
    const std::string dir = dirOf(path);
    const bool glb = path.size() >= 4 && path.compare(path.size() - 4, 4, ".glb") == 0;
    std::vector<uint8_t> bin;
    bool haveBin = false;
    if (glb) {
        if (file.size() < 20 || memcmp(file.data(), "glTF", 4)) throw std::runtime_error("glTF: not a GLB file");
        uint32_t ver, total;
        memcpy(&ver, file.data() + 4, 4);
        memcpy(&total, file.data() + 8, 4);
        if (ver != 2) throw std::runtime_error("glTF: GLB version 2 expected");
        size_t off = 12;
        bool haveJson = false;
        const size_t end = std::min<size_t>(total, file.size());
        while (off + 8 <= end) {
            uint32_t len, type;
            memcpy(&len, file.data() + off, 4);
            memcpy(&type, file.data() + off + 4, 4);
            if (off + 8 + (size_t) len > end) throw std::runtime_error("glTF: truncated GLB chunk");
            const uint8_t *d = file.data() + off + 8;
            if (type == 0x4E4F534Au && !haveJson) {
                m.j = parseJson((const char *) d, len);
                haveJson = true;
            } else if (type == 0x004E4942u && !haveBin) {
                bin.assign(d, d + len);
                haveBin = true;
            }
            off += 8 + (size_t) len;
        }
        if (!haveJson) throw std::runtime_error("glTF: GLB without a JSON chunk");
    } else {
        m.j = parseJson((const char *) file.data(), file.size());
    }
    if (m.j.kind != Json::Obj) throw std::runtime_error("glTF: top level is not an object");
    if (const Json *bufs = m.j.get("buffers")) {
        for (size_t i = 0; i < bufs->size(); i++) {
            const Json &b = bufs->at(i);
            const Json *uri = b.get("uri");
            std::vector<uint8_t> data;
            if (uri && uri->kind == Json::Str) {
                data = loadUri(uri->str, dir);
            } else if (glb && i == 0 && haveBin) {
                data = bin;
            } else {
                throw std::runtime_error("glTF: buffer without data");
            }
            const size_t len = b.sizeOr0("byteLength", std::numeric_limits<uint32_t>::max());
            if (data.size() < len) throw std::runtime_error("glTF: buffer shorter than its byteLength");
            m.buffers.push_back(std::move(data));
        }
    }
    (void) dir;
    return m;
}

// PNG, or JPEG (jpeg.cpp: 8 bits, four channels, alpha 255).
Image decodeImage(const uint8_t *p, size_t n) {
    if (n >= 2 && p[0] == 0xFF && p[1] == 0xD8) {
        Image img;
        std::vector<uint8_t> rgba;
        prr_host::decodeJpeg(p, n, img.w, img.h, rgba);
        img.bits = 8;
        img.px.assign(rgba.begin(), rgba.end());
        return img;
    }
    return decodePng(p, n);
}

// image index -> texture row, 8-bit RGBA. A 16-bit image becomes 8-bit by
// >> 8 (GLTFLoader.cpp:81-88) in the colour, metallic-roughness and normal
// slots; the emissive slot keeps the decoded image's bytes as they are
// (GLTFLoader.cpp:131-137: two little-endian bytes per channel, read with the
// default four one-byte channels per texel), of which texel fetches reach the
// first width * height * 4, the row kept here.
int textureOf(const Model &m, const std::string &dir, Loaded &L, std::map<int, int> &cache, int texIdx,
              bool emissive = false) {
    const Json *texs = m.j.get("textures");
    if (texIdx < 0 || !texs || (size_t) texIdx >= texs->size()) return -1;
    const int src = texs->at((size_t) texIdx).intOr("source", -1);
    const Json *imgs = m.j.get("images");
    if (src < 0 || !imgs || (size_t) src >= imgs->size()) return -1;
    const int key = 2 * src + (emissive ? 1 : 0);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const Json &im = imgs->at((size_t) src);
    std::vector<uint8_t> bytes;
    const Json *uri = im.get("uri");
    if (uri && uri->kind == Json::Str) {
        bytes = loadUri(uri->str, dir);
    } else {
        const int bvi = im.intOr("bufferView", -1);
        const Json *bvs = m.j.get("bufferViews");
        if (bvi < 0 || !bvs || (size_t) bvi >= bvs->size()) throw std::runtime_error("glTF: image without data");
        const Json &bv = bvs->at((size_t) bvi);
        const int bi = bv.intOr("buffer", -1);
        if (bi < 0 || (size_t) bi >= m.buffers.size()) throw std::runtime_error("glTF: buffer out of range");
        const auto &buf = m.buffers[(size_t) bi];
        const size_t o = bv.sizeOr0("byteOffset", buf.size());
        const size_t n = bv.sizeOr0("byteLength", buf.size() - o);  // o + n <= size, no overflow
        bytes.assign(buf.begin() + (std::ptrdiff_t) o, buf.begin() + (std::ptrdiff_t) (o + n));
    }
    const Image img = decodeImage(bytes.data(), bytes.size());
    std::vector<uint8_t> t(img.px.size());
    if (img.bits == 16 && emissive) {
        for (size_t i = 0; i < t.size(); i++) {
            const uint32_t v = img.px[i / 2];
            t[i] = (uint8_t) ((i & 1) ? v >> 8 : v & 0xFFu);
        }
    } else {
        for (size_t i = 0; i < t.size(); i++) t[i] = (uint8_t) (img.bits == 16 ? img.px[i] >> 8 : img.px[i]);
    }
    L.texData.push_back(std::move(t));
    prr_texture_desc td{img.w, img.h, 4, nullptr};
    L.tex.push_back(td);
    const int row = (int) L.tex.size() - 1;
    cache[key] = row;
    return row;
}

struct V3f {
    float x, y, z;
};
V3f nrmz(V3f v) {  // Vector3f::normalized (1 / sqrtf)
    const float n = v.x * v.x + v.y * v.y + v.z * v.z;
    if (n > 0) {
        const float inv = 1 / sqrtf(n);
        return {v.x * inv, v.y * inv, v.z * inv};
    }
    return v;
}

void load(const std::string &path, int matType, Loaded &L) {
    const Model m = openModel(path);
    const std::string dir = dirOf(path);
    std::map<int, int> texCache;
    const Json *meshes = m.j.get("meshes");
    const Json *materials = m.j.get("materials");
    for (size_t mi = 0; meshes && mi < meshes->size(); mi++) {
        const Json *prims = meshes->at(mi).get("primitives");
        for (size_t pi = 0; prims && pi < prims->size(); pi++) {
            const Json &prim = prims->at(pi);
            if (prim.intOr("mode", 4) != 4) continue;  // TINYGLTF_MODE_TRIANGLES only
            const Json *attrs = prim.get("attributes");
            const Json *aPos = attrs ? attrs->get("POSITION") : nullptr;
            if (!aPos || aPos->kind != Json::Num) throw std::runtime_error("glTF: primitive without POSITION");
            auto attr = [&](const char *k) -> int {
                const Json *a = attrs->get(k);
                return a && a->kind == Json::Num ? (int) a->num : -1;
            };
            const std::vector<float> pos = readFloats(m, (int) aPos->num);
            const std::vector<float> nrm = attr("NORMAL") >= 0 ? readFloats(m, attr("NORMAL")) : std::vector<float>{};
            const std::vector<float> uvs =
                attr("TEXCOORD_0") >= 0 ? readFloats(m, attr("TEXCOORD_0")) : std::vector<float>{};
            const std::vector<float> tan = attr("TANGENT") >= 0 ? readFloats(m, attr("TANGENT")) : std::vector<float>{};
            std::vector<uint32_t> idx;
            const int ii = prim.intOr("indices", -1);
            if (ii >= 0) {
                idx = readIndices(m, ii);
            } else {
                idx.resize(pos.size() / 3);
                for (size_t k = 0; k < idx.size(); k++) idx[k] = (uint32_t) k;
            }
            const size_t nv = pos.size() / 3;
            for (uint32_t k : idx)
                if (k >= nv || (!nrm.empty() && (size_t) k * 3 + 2 >= nrm.size()) ||
                    (!uvs.empty() && (size_t) k * 2 + 1 >= uvs.size()) ||
                    (!tan.empty() && (size_t) k * 4 + 3 >= tan.size()))
                    throw std::runtime_error("glTF: vertex index out of range");

            // material: DiffuseMaterial defaults, then the glTF factors and
            // textures (GLTFLoader.cpp:64-138), then the requested type's copy
            // rules (GLTFLoader.cpp:191-224)
            prr_material_desc md{};
            md.type = PRR_MAT_DIFFUSE;
            md.base_color[0] = md.base_color[1] = md.base_color[2] = 0.8f;
            md.roughness = 0.5f;
            md.metallic = 0.0f;
            md.ior = 1.5f;
            int tx[4] = {-1, -1, -1, -1};
            const int pm = prim.intOr("material", -1);
            if (pm >= 0) {
                if (!materials || (size_t) pm >= materials->size()) throw std::runtime_error("glTF: material out of range");
                const Json &gm = materials->at((size_t) pm);
                const Json *pbr = gm.get("pbrMetallicRoughness");
                const Json none;
                const Json &P = pbr ? *pbr : none;
                const Json *bc = P.get("baseColorFactor");
                for (int c = 0; c < 3; c++)
                    md.base_color[c] = bc && bc->size() > (size_t) c ? (float) bc->at((size_t) c).num : 1.0f;
                md.metallic = (float) P.numOr("metallicFactor", 1.0);
                md.roughness = (float) P.numOr("roughnessFactor", 1.0);
                const Json *ef = gm.get("emissiveFactor");
                for (int c = 0; c < 3; c++)
                    md.emission[c] = ef && ef->size() > (size_t) c ? (float) ef->at((size_t) c).num : 0.0f;
                auto texIndex = [&](const Json *o, const char *k) {
                    const Json *t = o ? o->get(k) : nullptr;
                    return t ? t->intOr("index", -1) : -1;
                };
                tx[0] = textureOf(m, dir, L, texCache, texIndex(&P, "baseColorTexture"));
                tx[1] = textureOf(m, dir, L, texCache, texIndex(&P, "metallicRoughnessTexture"));
                tx[2] = textureOf(m, dir, L, texCache, texIndex(&gm, "normalTexture"));
                tx[3] = textureOf(m, dir, L, texCache, texIndex(&gm, "emissiveTexture"), true);
            }
            if (matType == PRR_MAT_GLASS) {  // baseColor, emission, every texture; GlassMaterial defaults
                md.type = PRR_MAT_GLASS;
                md.roughness = 0.5f;
                md.metallic = 0.0f;
                md.ior = 1.5f;
            } else if (matType == PRR_MAT_MIRROR) {  // baseColor, emission, emissive texture
                md.type = PRR_MAT_MIRROR;
                md.roughness = 0.5f;
                md.metallic = 0.0f;
                md.ior = 1.5f;
                tx[0] = tx[1] = tx[2] = -1;
            } else if (matType == PRR_MAT_EMISSIVE) {  // a fresh EmissiveMaterial (its emission self-assigned: 0)
                md = prr_material_desc{};
                md.type = PRR_MAT_EMISSIVE;
                md.base_color[0] = md.base_color[1] = md.base_color[2] = 0.8f;
                md.roughness = 0.5f;
                md.ior = 1.5f;
                tx[0] = tx[1] = tx[2] = tx[3] = -1;
            }
            L.mats.push_back(md);
            L.matTex.insert(L.matTex.end(), tx, tx + 4);

            const int first = (int) L.flags.size();
            for (size_t i = 0; i + 2 < idx.size(); i += 3) {
                const uint32_t v[3] = {idx[i], idx[i + 1], idx[i + 2]};
                uint8_t fl = 0;
                for (int k = 0; k < 3; k++)
                    for (int c = 0; c < 3; c++) L.pos.push_back(pos[(size_t) v[k] * 3 + c]);
                if (!nrm.empty()) {
                    fl |= PRR_TRI_SMOOTH_NORMALS;
                    for (int k = 0; k < 3; k++) {
                        const V3f n = nrmz({nrm[(size_t) v[k] * 3], nrm[(size_t) v[k] * 3 + 1], nrm[(size_t) v[k] * 3 + 2]});
                        L.nrm.insert(L.nrm.end(), {n.x, n.y, n.z});
                    }
                } else {
                    L.nrm.insert(L.nrm.end(), 9, 0.0f);
                }
                if (!uvs.empty()) {
                    fl |= PRR_TRI_UVS;
                    for (int k = 0; k < 3; k++) L.uv.insert(L.uv.end(), {uvs[(size_t) v[k] * 2], uvs[(size_t) v[k] * 2 + 1]});
                } else {
                    L.uv.insert(L.uv.end(), 6, 0.0f);
                }
                if (!tan.empty()) {
                    fl |= PRR_TRI_TANGENTS;
                    for (int k = 0; k < 3; k++)
                        for (int c = 0; c < 4; c++) L.tan.push_back(tan[(size_t) v[k] * 4 + c]);
                } else {
                    L.tan.insert(L.tan.end(), 12, 0.0f);
                }
                L.flags.push_back(fl);
            }
            L.meshFirst.push_back(first);
            L.meshCount.push_back((int) L.flags.size() - first);
            L.meshMat.push_back((int) L.mats.size() - 1);
        }
    }
    if (L.flags.size() > (size_t) std::numeric_limits<int32_t>::max()) throw std::runtime_error("glTF: too many triangles");
}

}  // namespace

extern "C" {

int prr_gltf_load(const char *path, int32_t material_type, prr_gltf_data **out) {
    if (!path || !out) return fail(nullptr, PRR_ERR_INVALID, "prr_gltf_load: null argument");
    *out = nullptr;
    if (material_type < PRR_MAT_DIFFUSE || material_type > PRR_MAT_EMISSIVE)
        return fail(nullptr, PRR_ERR_INVALID, "prr_gltf_load: bad material type");
    std::unique_ptr<Loaded> L(new Loaded());
    try {
        load(path, material_type, *L);
    } catch (const std::exception &e) {
        return fail(nullptr, PRR_ERR_INVALID, std::string("prr_gltf_load: ") + path + ": " + e.what());
    }
    for (size_t i = 0; i < L->tex.size(); i++) L->tex[i].data = L->texData[i].data();
    prr_gltf_data &d = L->d;
    d.num_meshes = (int32_t) L->meshFirst.size();
    d.mesh_first = L->meshFirst.data();
    d.mesh_count = L->meshCount.data();
    d.mesh_material = L->meshMat.data();
    d.num_triangles = (int32_t) L->flags.size();
    d.positions = L->pos.data();
    d.normals = L->nrm.data();
    d.uvs = L->uv.data();
    d.tangents = L->tan.data();
    d.tri_flags = L->flags.data();
    d.num_materials = (int32_t) L->mats.size();
    d.materials = L->mats.data();
    d.material_textures = L->matTex.data();
    d.num_textures = (int32_t) L->tex.size();
    d.textures = L->tex.data();
    d.impl = L.get();
    *out = &L.release()->d;
    return PRR_OK;
}

void prr_gltf_free(prr_gltf_data *d) {
    if (d) delete static_cast<Loaded *>(d->impl);
}

}  // extern "C"
