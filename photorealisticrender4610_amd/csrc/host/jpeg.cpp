// jpeg.cpp — JPEG (ITU-T T.81) decoder for glTF textures: baseline,
// extended sequential and progressive Huffman coding, 8-bit samples, one,
// three or four components with any sampling factors from 1 to 4, restart
// intervals.
// Output is RGBA8, as the reference's loader requests images (four channels).
//
// The reference decodes JPEG through stb_image (tinygltf). To give the same
// texels this decoder follows the same arithmetic: the integer LLM IDCT with
// 16-bit dequantised coefficients, stb_image's chroma resampling rules and
// its fixed-point YCbCr conversion (each restated below). Arithmetic-coded,
// lossless, hierarchical and 12-bit files are refused with an error.
//
// Every scan decodes into per-block coefficient arrays (zigzag order, not yet
// dequantised); one IDCT pass over all blocks follows the last scan, so
// sequential and progressive files share the output path.
#include "jpeg.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace prr_host {
namespace {

[[noreturn]] void bad(const std::string &why) { throw std::runtime_error("JPEG: " + why); }

// zigzag position k -> natural (row-major) index in the 8x8 block
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 Annex C): per code length the smallest and
// largest code and the index of its first value.
struct Huffman {
    bool defined = false;
    uint8_t vals[256] = {};
    int mincode[17] = {}, maxcode[18] = {}, valptr[17] = {};
    void build(const uint8_t *counts, const uint8_t *v, int nv) {
        std::memcpy(vals, v, (size_t) nv);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; len++) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            if (code > (1 << len)) bad("bad Huffman table");
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
        defined = true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;      // Huffman tables of the current scan
    int bw = 0, bh = 0;      // blocks across / down (padded to whole MCUs)
    int pw = 0, ph = 0;      // plane size in samples (bw * 8, bh * 8)
    int cw = 0, ch = 0;      // samples that hold image data
    int dc = 0;              // DC predictor
    bool latched = false;    // quantisation table copied at its first scan
    uint16_t q[64] = {};     // that table, zigzag order
    std::vector<int16_t> coef;  // bw * bh blocks of 64, zigzag order
    std::vector<uint8_t> plane;
    int16_t *block(int bx, int by) { return &coef[((size_t) by * bw + bx) * 64]; }
};

// Entropy-coded segment reader: removes stuffed zero bytes and stops at a
// marker (the following bits read as zeros).
struct Bits {
    const uint8_t *p;
    size_t n, pos;
    uint32_t acc = 0;
    int cnt = 0;
    bool marker = false;
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!marker && pos < n) {
                b = p[pos];
                if (b == 0xFF) {
                    const uint32_t nx = pos + 1 < n ? p[pos + 1] : 0;
                    if (nx == 0x00) {
                        pos += 2;
                    } else {  // a marker: leave it for the caller
                        marker = true;
                        b = 0;
                    }
                } else {
                    pos++;
                }
            }
            acc |= b << (24 - cnt);
            cnt += 8;
        }
    }
    int bit() {
        if (cnt < 1) fill();
        const int r = (int) (acc >> 31);
        acc <<= 1;
        cnt--;
        return r;
    }
    int bits(int k) {  // k <= 16
        if (k == 0) return 0;
        if (cnt < k) fill();
        const int r = (int) (acc >> (32 - k));
        acc <<= k;
        cnt -= k;
        return r;
    }
    int decode(const Huffman &t) {
        int code = 0;
        for (int len = 1; len <= 16; len++) {
            code = (code << 1) | bit();
            if (code <= t.maxcode[len] && t.maxcode[len] >= 0) {
                const int k = t.valptr[len] + code - t.mincode[len];
                if (k < 0 || k > 255) bad("bad Huffman code");
                return t.vals[k];
            }
        }
        bad("bad Huffman code");
    }
    // T.81 F.2.2.1: a value of s bits, negative when its top bit is 0
    int extend(int s) {
        if (s == 0) return 0;
        if (s > 16) bad("bad coefficient size");
        const int v = bits(s);
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
    // after an RSTn: align to the next byte, skip the marker
    void restart() {
        acc = 0;
        cnt = 0;
        marker = false;
        while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) pos++;
        if (pos + 1 >= n) bad("missing restart marker");
        pos += 2;
    }
};

// Inverse DCT: the Loeffler-Ligtenberg-Moschytz factorisation in 12-bit fixed
// point (the IJG "islow" integer method), with the scaling stb_image uses:
// a column pass that keeps 2 extra bits (+512, >> 10), then a row pass that
// removes the remaining 2^17 with rounding and the +128 level shift. Integer
// throughout, so its samples are those of the reference's decoder.
struct Llm {
    // round(c * 4096) of the method's rotation constants
    static constexpr int kC0541 = 2217, kCm1847 = -7567, kC0765 = 3135, kC1175 = 4816, kC0298 = 1223,
                         kC2053 = 8410, kC3072 = 12586, kC1501 = 6149, kCm0899 = -3685, kCm2562 = -10497,
                         kCm1961 = -8034, kCm0390 = -1597;
    // in: 8 samples at stride; out[0..7] scaled by 4096. 64-bit, so that
    // extreme coefficients of a malformed file cannot overflow
    static void pass(const int64_t *in, int stride, int64_t out[8]) {
        const int64_t s0 = in[0], s1 = in[stride], s2 = in[2 * stride], s3 = in[3 * stride], s4 = in[4 * stride],
                  s5 = in[5 * stride], s6 = in[6 * stride], s7 = in[7 * stride];
        // even half: s0, s2, s4, s6
        const int64_t rot = (s2 + s6) * kC0541;
        const int64_t e2 = rot + s6 * kCm1847, e3 = rot + s2 * kC0765;
        const int64_t e0 = (s0 + s4) * 4096, e1 = (s0 - s4) * 4096;
        const int64_t a0 = e0 + e3, a3 = e0 - e3, a1 = e1 + e2, a2 = e1 - e2;
        // odd half: s7, s5, s3, s1
        const int64_t z73 = s7 + s3, z51 = s5 + s1, z71 = s7 + s1, z53 = s5 + s3;
        const int64_t zr = (z73 + z51) * kC1175;
        const int64_t m71 = zr + z71 * kCm0899, m53 = zr + z53 * kCm2562;
        const int64_t m73 = z73 * kCm1961, m51 = z51 * kCm0390;
        const int64_t b0 = s7 * kC0298 + m71 + m73, b1 = s5 * kC2053 + m53 + m51;
        const int64_t b2 = s3 * kC3072 + m53 + m73, b3 = s1 * kC1501 + m71 + m51;
        out[0] = a0 + b3, out[7] = a0 - b3;
        out[1] = a1 + b2, out[6] = a1 - b2;
        out[2] = a2 + b1, out[5] = a2 - b1;
        out[3] = a3 + b0, out[4] = a3 - b0;
    }
};

uint8_t clamp8(long v) { return (uint8_t) (v < 0 ? 0 : (v > 255 ? 255 : v)); }

// coef: dequantised (16-bit, as stored), natural (row-major) order
void idct8x8(const int16_t *coef, uint8_t *out, int stride) {
    int64_t in[64], mid[64], r[8];
    for (int i = 0; i < 64; i++) in[i] = coef[i];
    for (int x = 0; x < 8; x++) {  // columns
        Llm::pass(in + x, 8, r);
        for (int y = 0; y < 8; y++) mid[y * 8 + x] = (r[y] + 512) >> 10;
    }
    for (int y = 0; y < 8; y++) {  // rows
        Llm::pass(mid + y * 8, 1, r);
        for (int x = 0; x < 8; x++) out[y * stride + x] = clamp8((r[x] + 65536 + (128 << 17)) >> 17);
    }
}

// One component plane at full resolution (w x h), resampled as stb_image
// does: ratio 1 copies; 2:1 vertical, horizontal or both uses the triangle
// filter (3/4 nearer sample, 1/4 farther, rounded, edge samples repeated);
// any other ratio repeats samples. In the horizontal-only 2:1 case the
// second-to-last output sample weights the next-to-last input by 3/4 (not the
// last one), as stb_image's row resampler does.
std::vector<uint8_t> upsample(const Component &c, int w, int h, int hmax, int vmax) {
    std::vector<uint8_t> out((size_t) w * h);
    const int rx = hmax / c.h, ry = vmax / c.v, lw = c.cw;
    std::vector<uint8_t> row((size_t) 2 * lw + 8);
    for (int y = 0; y < h; y++) {
        const bool tri = (rx == 1 && ry == 2) || (rx == 2 && ry <= 2);
        int ny = y / ry, fy = ny;
        if (tri && ry == 2) {
            ny = y >> 1;
            fy = (y & 1) ? std::min(ny + 1, c.ch - 1) : std::max(ny - 1, 0);
        }
        const uint8_t *near = &c.plane[(size_t) ny * c.pw], *far = &c.plane[(size_t) fy * c.pw];
        uint8_t *o = &out[(size_t) y * w];
        if (rx == 1 && ry == 1) {
            std::memcpy(o, near, (size_t) w);
        } else if (rx == 1 && ry == 2) {
            for (int x = 0; x < w; x++) o[x] = (uint8_t) ((3 * near[x] + far[x] + 2) >> 2);
        } else if (rx == 2 && ry == 1) {
            uint8_t *r = row.data();
            if (lw == 1) {
                r[0] = r[1] = near[0];
            } else {
                r[0] = near[0];
                r[1] = (uint8_t) ((3 * near[0] + near[1] + 2) >> 2);
                for (int i = 1; i < lw - 1; i++) {
                    r[2 * i] = (uint8_t) ((3 * near[i] + near[i - 1] + 2) >> 2);
                    r[2 * i + 1] = (uint8_t) ((3 * near[i] + near[i + 1] + 2) >> 2);
                }
                r[2 * lw - 2] = (uint8_t) ((3 * near[lw - 2] + near[lw - 1] + 2) >> 2);
                r[2 * lw - 1] = near[lw - 1];
            }
            std::memcpy(o, r, (size_t) w);
        } else if (rx == 2 && ry == 2) {
            uint8_t *r = row.data();
            int t1 = 3 * near[0] + far[0];
            r[0] = (uint8_t) ((t1 + 2) >> 2);
            for (int i = 1; i < lw; i++) {
                const int t0 = t1;
                t1 = 3 * near[i] + far[i];
                r[2 * i - 1] = (uint8_t) ((3 * t0 + t1 + 8) >> 4);
                r[2 * i] = (uint8_t) ((3 * t1 + t0 + 8) >> 4);
            }
            r[2 * lw - 1] = (uint8_t) ((t1 + 2) >> 2);
            std::memcpy(o, r, (size_t) w);
        } else {
            for (int x = 0; x < w; x++) o[x] = near[x / rx];
        }
    }
    return out;
}

uint16_t be16(const uint8_t *p) { return (uint16_t) (p[0] << 8 | p[1]); }

// One block of a scan (T.81 F.2.2 sequential, G.1.2 progressive), into blk
// (zigzag order, not dequantised). eobrun counts the blocks left in a
// progressive AC end-of-band run.
struct ScanDecoder {
    Bits &br;
    const Huffman *dc, *ac;
    int ss, se, ah, al;
    bool progressive;
    int eobrun = 0;

    void sequential(Component &c, int16_t *blk) {
        c.dc += br.extend(br.decode(dc[c.td]));
        blk[0] = (int16_t) c.dc;
        for (int k = 1; k < 64;) {
            const int rs = br.decode(ac[c.ta]);
            const int r = rs >> 4, sz = rs & 15;
            if (sz == 0) {
                if (r != 15) break;  // EOB
                k += 16;             // ZRL
                continue;
            }
            k += r;
            if (k > 63) bad("coefficient index past 63");
            blk[k] = (int16_t) br.extend(sz);
            k++;
        }
    }
    void dcFirst(Component &c, int16_t *blk) {
        c.dc += br.extend(br.decode(dc[c.td]));
        blk[0] = (int16_t) (c.dc * (1 << al));
    }
    void dcRefine(int16_t *blk) {
        if (br.bit()) blk[0] = (int16_t) (blk[0] | (1 << al));
    }
    void acFirst(const Component &c, int16_t *blk) {
        if (eobrun > 0) {
            eobrun--;
            return;
        }
        for (int k = ss; k <= se;) {
            const int rs = br.decode(ac[c.ta]);
            const int r = rs >> 4, sz = rs & 15;
            if (sz == 0) {
                if (r < 15) {  // EOBr: this block and 2^r - 1 + (r bits) more end here
                    eobrun = (1 << r) - 1 + br.bits(r);
                    return;
                }
                k += 16;  // ZRL
                continue;
            }
            k += r;
            if (k > se) bad("coefficient index past the band");
            blk[k] = (int16_t) (br.extend(sz) * (1 << al));
            k++;
        }
    }
    // a correction bit for a coefficient already nonzero
    void refineNonzero(int16_t &v) {
        if (br.bit() && (v & (1 << al)) == 0) v = (int16_t) (v >= 0 ? v + (1 << al) : v - (1 << al));
    }
    void acRefine(const Component &c, int16_t *blk) {
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se; k++) {
                const int rs = br.decode(ac[c.ta]);
                int r = rs >> 4;
                int val = 0;
                if (rs & 15) {
                    if ((rs & 15) != 1) bad("bad refinement coefficient size");
                    val = br.bit() ? (1 << al) : -(1 << al);
                } else if (r != 15) {  // EOBr: the rest of this band, then more blocks
                    eobrun = (1 << r) + br.bits(r);
                    break;
                }
                // skip r zero coefficients (refining the nonzero ones passed),
                // then place val at the next zero one
                for (; k <= se; k++) {
                    if (blk[k] != 0)
                        refineNonzero(blk[k]);
                    else if (--r < 0)
                        break;
                }
                if (val && k <= se) blk[k] = (int16_t) val;  // past the band: dropped
            }
        }
        if (eobrun > 0) {
            for (; k <= se; k++)
                if (blk[k] != 0) refineNonzero(blk[k]);
            eobrun--;
        }
    }
    void operator()(Component &c, int16_t *blk) {
        if (!progressive)
            sequential(c, blk);
        else if (ss == 0)
            ah ? dcRefine(blk) : dcFirst(c, blk);
        else
            ah ? acRefine(c, blk) : acFirst(c, blk);
    }
};

}  // namespace

void decodeJpeg(const uint8_t *p, size_t n, int &width, int &height, std::vector<uint8_t> &rgba) {
    if (n < 4 || p[0] != 0xFF || p[1] != 0xD8) bad("not a JPEG file");
    size_t pos = 2;
    uint16_t quant[4][64];
    bool qdef[4] = {false, false, false, false};
    Huffman dc[4], ac[4];
    std::vector<Component> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restartInterval = 0, adobeTransform = -1;
    bool frame = false, progressive = false, jfif = false;
    for (;;) {
        while (pos < n && p[pos] != 0xFF) pos++;  // skip fill up to a marker
        while (pos < n && p[pos] == 0xFF) pos++;
        if (pos >= n) {
            if (frame) break;  // a missing EOI after the last scan is tolerated
            bad("truncated file");
        }
        const int m = p[pos++];
        if (m == 0xD9) break;                            // EOI
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;  // TEM, stray RSTn
        if (pos + 2 > n) bad("truncated segment");
        const size_t len = be16(p + pos);
        if (len < 2 || pos + len > n) bad("bad segment length");
        const uint8_t *s = p + pos + 2;
        const size_t sl = len - 2;
        pos += len;
        if (m == 0xDB) {  // DQT
            size_t k = 0;
            while (k < sl) {
                const int pq = s[k] >> 4, tq = s[k] & 15;
                if (tq > 3 || pq > 1 || k + 1 + (size_t) 64 * (pq + 1) > sl) bad("bad quantisation table");
                for (int i = 0; i < 64; i++) quant[tq][i] = pq ? be16(s + k + 1 + 2 * i) : s[k + 1 + i];
                qdef[tq] = true;
                k += 1 + (size_t) 64 * (pq + 1);
            }
        } else if (m == 0xC4) {  // DHT
            size_t k = 0;
            while (k < sl) {
                if (k + 17 > sl) bad("bad Huffman table");
                const int tc = s[k] >> 4, th = s[k] & 15;
                if (tc > 1 || th > 3) bad("bad Huffman table");
                int nv = 0;
                for (int i = 0; i < 16; i++) nv += s[k + 1 + i];
                if (nv > 256 || k + 17 + (size_t) nv > sl) bad("bad Huffman table");
                (tc ? ac[th] : dc[th]).build(s + k + 1, s + k + 17, nv);
                k += 17 + (size_t) nv;
            }
        } else if (m == 0xDD) {  // DRI
            if (sl < 2) bad("bad restart interval");
            restartInterval = be16(s);
        } else if (m == 0xE0) {  // APP0: a JFIF file is YCbCr whatever else it says
            if (sl >= 5 && !memcmp(s, "JFIF", 5)) jfif = true;
        } else if (m == 0xEE) {  // APP14 (Adobe): colour transform flag
            if (sl >= 12 && !memcmp(s, "Adobe", 6)) adobeTransform = s[11];
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2) {  // SOF0-2: Huffman, sequential or progressive
            if (frame) bad("more than one frame");
            if (sl < 6 || s[0] != 8) bad("only 8-bit samples are supported");
            H = be16(s + 1);
            W = be16(s + 3);
            const int nc = s[5];
            if (W <= 0 || H <= 0) bad("image without size (DNL) is not supported");
            if (nc != 1 && nc != 3 && nc != 4) bad("only one, three or four components are supported");
            if (sl < 6 + (size_t) 3 * nc) bad("bad frame header");
            if ((int64_t) W * H > ((int64_t) 1 << 28)) bad("image too large");
            comps.resize((size_t) nc);
            for (int i = 0; i < nc; i++) {
                Component &c = comps[(size_t) i];
                c.id = s[6 + 3 * i];
                c.h = s[7 + 3 * i] >> 4;
                c.v = s[7 + 3 * i] & 15;
                c.tq = s[8 + 3 * i];
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) bad("bad component");
                hmax = std::max(hmax, c.h);
                vmax = std::max(vmax, c.v);
            }
            const int mcuw = 8 * hmax, mcuh = 8 * vmax;
            const int mx = (W + mcuw - 1) / mcuw, my = (H + mcuh - 1) / mcuh;
            for (Component &c : comps) {
                if (hmax % c.h || vmax % c.v) bad("unsupported sampling factors");
                c.bw = mx * c.h;
                c.bh = my * c.v;
                c.pw = c.bw * 8;
                c.ph = c.bh * 8;
                c.cw = (W * c.h + hmax - 1) / hmax;
                c.ch = (H * c.v + vmax - 1) / vmax;
                c.coef.assign((size_t) c.bw * c.bh * 64, 0);
            }
            frame = true;
            progressive = m == 0xC2;
        } else if (m >= 0xC3 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            bad("lossless, hierarchical and arithmetic-coded images are not supported");
        } else if (m == 0xDA) {  // SOS
            if (!frame) bad("scan before frame");
            if (sl < 1) bad("bad scan header");
            const int ns = s[0];
            if (ns < 1 || ns > (int) comps.size() || sl < 4 + (size_t) 2 * ns) bad("bad scan header");
            const int ss = s[1 + 2 * ns], se = s[2 + 2 * ns], ah = s[3 + 2 * ns] >> 4, al = s[3 + 2 * ns] & 15;
            if (progressive) {
                const bool dcScan = ss == 0;
                if (se > 63 || ss > se || (dcScan && se != 0) || (!dcScan && ns != 1) || al > 13 || ah > 13 ||
                    (ah && ah != al + 1))
                    bad("bad progressive scan parameters");
            }
            std::vector<Component *> sc;
            for (int i = 0; i < ns; i++) {
                const int cid = s[1 + 2 * i], tt = s[2 + 2 * i];
                Component *c = nullptr;
                for (Component &x : comps)
                    if (x.id == cid) c = &x;
                if (!c) bad("scan names an unknown component");
                c->td = tt >> 4;
                c->ta = tt & 15;
                if (c->td > 3 || c->ta > 3) bad("bad Huffman table selector");
                const bool needDc = !progressive || (ss == 0 && ah == 0), needAc = !progressive || ss > 0;
                if ((needDc && !dc[c->td].defined) || (needAc && !ac[c->ta].defined)) bad("missing Huffman table");
                if (!c->latched) {
                    if (!qdef[c->tq]) bad("missing quantisation table");
                    std::memcpy(c->q, quant[c->tq], sizeof c->q);
                    c->latched = true;
                }
                c->dc = 0;
                sc.push_back(c);
            }
            Bits br{p, n, pos};
            ScanDecoder dec{br, dc, ac, ss, se, ah, al, progressive};
            // MCUs: all components of an interleaved scan, one block of a
            // single-component scan (its data blocks only)
            const bool single = ns == 1;
            const int mcux = single ? (sc[0]->cw + 7) / 8 : comps[0].bw / comps[0].h;
            const int mcuy = single ? (sc[0]->ch + 7) / 8 : comps[0].bh / comps[0].v;
            int64_t mcu = 0;
            for (int my = 0; my < mcuy; my++)
                for (int mx = 0; mx < mcux; mx++, mcu++) {
                    if (restartInterval && mcu > 0 && mcu % restartInterval == 0) {
                        br.restart();
                        for (Component *c : sc) c->dc = 0;
                        dec.eobrun = 0;
                    }
                    for (Component *c : sc) {
                        const int nbx = single ? 1 : c->h, nby = single ? 1 : c->v;
                        for (int by = 0; by < nby; by++)
                            for (int bx = 0; bx < nbx; bx++)
                                dec(*c, single ? c->block(mx, my) : c->block(mx * c->h + bx, my * c->v + by));
                    }
                }
            pos = br.pos;
            if (br.marker) {  // back to the marker the reader stopped at
                while (pos > 0 && p[pos] != 0xFF) pos--;
            }
        }
        // other segments (APPn, COM, ...) are skipped
    }
    if (!frame) bad("no image");
    for (Component &c : comps) {
        if (!c.latched) bad("a component has no scan");
        c.plane.assign((size_t) c.pw * c.ph, 0);
        int16_t nat[64];
        for (int by = 0; by < c.bh; by++)
            for (int bx = 0; bx < c.bw; bx++) {
                const int16_t *blk = c.block(bx, by);
                for (int k = 0; k < 64; k++) nat[kZigzag[k]] = (int16_t) (blk[k] * c.q[k]);
                idct8x8(nat, &c.plane[(size_t) by * 8 * c.pw + (size_t) bx * 8], c.pw);
            }
        std::vector<int16_t>().swap(c.coef);
    }
    width = W;
    height = H;
    rgba.assign((size_t) W * H * 4, 255);
    if (comps.size() == 1) {
        const std::vector<uint8_t> g = upsample(comps[0], W, H, hmax, vmax);
        for (size_t i = 0; i < (size_t) W * H; i++) rgba[4 * i] = rgba[4 * i + 1] = rgba[4 * i + 2] = g[i];
        return;
    }
    const std::vector<uint8_t> c0 = upsample(comps[0], W, H, hmax, vmax), c1 = upsample(comps[1], W, H, hmax, vmax),
                               c2 = upsample(comps[2], W, H, hmax, vmax);
    const bool four = comps.size() == 4;
    const std::vector<uint8_t> c3 = four ? upsample(comps[3], W, H, hmax, vmax) : std::vector<uint8_t>();
    // three components: RGB when the component ids spell R, G, B, or an Adobe
    // marker says so (transform 0) in a file without a JFIF marker; YCbCr
    // otherwise. Four: CMYK for Adobe transform 0, YCCK for transform 2,
    // YCbCr with the fourth ignored otherwise (stb_image's three cases).
    const bool rgb = !four && ((comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B') ||
                               (adobeTransform == 0 && !jfif));
    // YCbCr -> RGB in fixed point as stb_image computes it: the JFIF
    // coefficients at 12 bits, scaled to 20 bits with the rounding half in the
    // Y term; the Cb contribution to G is truncated to a multiple of 2^16
    constexpr int kCrR = 5743 << 8, kCrG = 2925 << 8, kCbG = 1410 << 8, kCbB = 7258 << 8;
    // x * k / 255 rounded, by the (t + t / 256) / 256 identity
    auto scale = [](int x, int k) -> uint8_t {
        const unsigned t = (unsigned) (x * k) + 128u;
        return (uint8_t) ((t + (t >> 8)) >> 8);
    };
    for (size_t i = 0; i < (size_t) W * H; i++) {
        uint8_t *o = &rgba[4 * i];
        if (rgb || (four && adobeTransform == 0)) {
            o[0] = c0[i], o[1] = c1[i], o[2] = c2[i];
        } else {
            const int yf = (c0[i] << 20) + (1 << 19), cb = c1[i] - 128, cr = c2[i] - 128;
            o[0] = clamp8((yf + cr * kCrR) >> 20);
            o[1] = clamp8((yf - cr * kCrG + ((-cb * kCbG) & ~0xFFFF)) >> 20);
            o[2] = clamp8((yf + cb * kCbB) >> 20);
        }
        if (four && adobeTransform == 0) {
            for (int k = 0; k < 3; k++) o[k] = scale(o[k], c3[i]);
        } else if (four && adobeTransform == 2) {
            for (int k = 0; k < 3; k++) o[k] = scale(255 - o[k], c3[i]);
        }
    }
}

}  // namespace prr_host
