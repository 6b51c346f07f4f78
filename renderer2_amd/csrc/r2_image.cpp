// Image files of CImageWriter::writeToFile (visualisation/src/image_writer.cpp:8-21):
// the post-processed 8-bit RGB frame, whose rows run bottom to top, written
// vertically flipped (stbi_flip_vertically_on_write(true)) as
//   * an uncompressed PNG (stbi_write_png_compression_level = 0): stored deflate
//     blocks, filter 0 on every row;
//   * a baseline sequential JPEG in a JFIF wrapper (stbi_write_jpg, which the
//     pipeline calls with quality 100, raytracer/pipeline.cu:72,84).  The encoder
//     follows ITU-T T.81: YCbCr (JFIF), 8x8 float DCT, the Annex K example
//     quantisation tables scaled by the IJG quality rule, the Annex K Huffman
//     tables, and byte stuffing.  Above quality 90 all three components are at full
//     resolution (4:4:4).  At 90 and below, chroma is averaged over 2x2 pixels
//     (4:2:0).
// Host only.
#include "r2_internal.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------------- PNG ---
uint32_t crc32(const uint8_t* d, size_t n, uint32_t c = 0xFFFFFFFFu) {
  struct Table {
    uint32_t v[256];
    Table() {
      for (uint32_t i = 0; i < 256; ++i) {
        uint32_t x = i;
        for (int k = 0; k < 8; ++k) x = (x & 1) ? 0xEDB88320u ^ (x >> 1) : x >> 1;
        v[i] = x;
      }
    }
  };
  static const Table table;  // initialised once, thread-safe
  for (size_t i = 0; i < n; ++i) c = table.v[(c ^ d[i]) & 0xFF] ^ (c >> 8);
  return c;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

void chunk(std::ofstream& f, const char* type, const std::vector<uint8_t>& data) {
  std::vector<uint8_t> b;
  put32(b, (uint32_t)data.size());
  b.insert(b.end(), type, type + 4);
  b.insert(b.end(), data.begin(), data.end());
  const uint32_t c = crc32(b.data() + 4, b.size() - 4) ^ 0xFFFFFFFFu;
  put32(b, c);
  f.write((const char*)b.data(), (std::streamsize)b.size());
}

// ------------------------------------------------------------------ JPEG ---
// T.81 Annex K, Tables K.1 and K.2, in row-major order
const uint8_t kLumQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                           14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                           18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                           49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kChrQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99,
                           99, 99, 47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// T.81 Annex K.3: code counts per length 1..16, then the symbols in code order
const uint8_t kDcLumBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChrBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChrBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChrVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// canonical Huffman codes of a table (T.81 Annex C): code and length per symbol
struct Huff {
  uint16_t code[256];
  uint8_t len[256];
  Huff(const uint8_t* bits, const uint8_t* vals, int nvals) {
    memset(len, 0, sizeof(len));
    memset(code, 0, sizeof(code));
    int k = 0;
    uint32_t c = 0;
    for (int l = 1; l <= 16; ++l) {
      for (int i = 0; i < bits[l - 1]; ++i, ++k) {
        if (k >= nvals) throw std::logic_error("jpeg: Huffman table counts exceed its symbols");
        code[vals[k]] = (uint16_t)c++;
        len[vals[k]] = (uint8_t)l;
      }
      c <<= 1;
    }
    if (k != nvals) throw std::logic_error("jpeg: Huffman table counts do not match its symbols");
  }
};

// zig-zag sequence (T.81 Figure 5): position k -> row-major index
struct ZigZag {
  int nat[64];
  ZigZag() {
    int k = 0;
    for (int s = 0; s < 15; ++s) {
      const int lo = std::max(0, s - 7), hi = std::min(s, 7);
      if (s & 1)
        for (int r = lo; r <= hi; ++r) nat[k++] = r * 8 + (s - r);
      else
        for (int r = hi; r >= lo; --r) nat[k++] = r * 8 + (s - r);
    }
  }
};

// quantisation table of a quality: IJG scaling of the Annex K tables, clamped to 1..255
void scale_table(const uint8_t* base, int quality, uint8_t* out) {
  const int q = std::min(100, std::max(1, quality));
  const int s = q < 50 ? 5000 / q : 200 - 2 * q;
  for (int i = 0; i < 64; ++i) out[i] = (uint8_t)std::min(255, std::max(1, (base[i] * s + 50) / 100));
}

class BitWriter {
 public:
  explicit BitWriter(std::vector<uint8_t>& out) : out_(out) {}
  void put(uint32_t bits, int n) {  // n <= 16, most significant bit first
    acc_ = (acc_ << n) | (bits & ((1u << n) - 1));
    cnt_ += n;
    while (cnt_ >= 8) {
      const uint8_t b = (uint8_t)(acc_ >> (cnt_ - 8));
      out_.push_back(b);
      if (b == 0xFF) out_.push_back(0);  // byte stuffing
      cnt_ -= 8;
    }
    acc_ &= (1u << cnt_) - 1;
  }
  void flush() {  // pad with 1-bits to a byte boundary
    if (cnt_ > 0) put(0x7F, 8 - cnt_);
  }

 private:
  std::vector<uint8_t>& out_;
  uint32_t acc_ = 0;
  int cnt_ = 0;
};

struct JpegEncoder {
  double cosm[8][8];  // cosm[u][x] = C(u)/2 cos((2x+1)u pi/16)
  ZigZag zz;
  Huff dc_lum{kDcLumBits, kDcVals, 12}, dc_chr{kDcChrBits, kDcVals, 12};
  Huff ac_lum{kAcLumBits, kAcLumVals, 162}, ac_chr{kAcChrBits, kAcChrVals, 162};

  JpegEncoder() {
    for (int u = 0; u < 8; ++u)
      for (int x = 0; x < 8; ++x)
        cosm[u][x] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * M_PI / 16.0);
  }

  static int category(int v) {
    int a = v < 0 ? -v : v, n = 0;
    while (a) {
      ++n;
      a >>= 1;
    }
    return n;
  }
  static void emit(BitWriter& bw, const Huff& h, int sym) {
    if (!h.len[sym]) throw std::logic_error("jpeg: symbol without a Huffman code");
    bw.put(h.code[sym], h.len[sym]);
  }

  // one level-shifted 8x8 block (row-major): DCT, quantise, entropy-code; returns its DC
  int block(BitWriter& bw, const double* f, const uint8_t* qt, int prev_dc, const Huff& dc, const Huff& ac) const {
    double t[64];
    for (int y = 0; y < 8; ++y)
      for (int u = 0; u < 8; ++u) {
        double s = 0;
        for (int x = 0; x < 8; ++x) s += cosm[u][x] * f[y * 8 + x];
        t[y * 8 + u] = s;
      }
    int q[64];
    for (int v = 0; v < 8; ++v)
      for (int u = 0; u < 8; ++u) {
        double s = 0;
        for (int y = 0; y < 8; ++y) s += cosm[v][y] * t[y * 8 + u];
        const long c = std::lround(s / qt[v * 8 + u]);
        const long lim = (v | u) ? 1023 : 2047;
        q[v * 8 + u] = (int)std::min(lim, std::max(-lim, c));
      }
    const int d = q[0] - prev_dc;
    const int dn = category(d);
    emit(bw, dc, dn);
    if (dn) bw.put((uint32_t)(d < 0 ? d + (1 << dn) - 1 : d), dn);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
      const int c = q[zz.nat[k]];
      if (c == 0) {
        ++run;
        continue;
      }
      while (run > 15) {
        emit(bw, ac, 0xF0);  // ZRL: 16 zeros
        run -= 16;
      }
      const int n = category(c);
      emit(bw, ac, (run << 4) | n);
      bw.put((uint32_t)(c < 0 ? c + (1 << n) - 1 : c), n);
      run = 0;
    }
    if (run) emit(bw, ac, 0x00);  // EOB
    return q[0];
  }
};

void marker(std::vector<uint8_t>& o, uint8_t m) {
  o.push_back(0xFF);
  o.push_back(m);
}
void put16(std::vector<uint8_t>& o, int v) {
  o.push_back((uint8_t)(v >> 8));
  o.push_back((uint8_t)v);
}

}  // namespace

namespace r2host {

void write_png(const std::string& path, int w, int h, const uint8_t* rgb) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  f.write((const char*)sig, 8);
  std::vector<uint8_t> ihdr;
  put32(ihdr, (uint32_t)w);
  put32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  chunk(f, "IHDR", ihdr);
  std::vector<uint8_t> raw;
  raw.reserve((size_t)h * (3 * (size_t)w + 1));
  for (int y = h - 1; y >= 0; --y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb + (size_t)y * w * 3, rgb + (size_t)(y + 1) * w * 3);
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  uint32_t a = 1, b = 0;
  for (uint8_t v : raw) {
    a = (a + v) % 65521u;
    b = (b + a) % 65521u;
  }
  size_t off = 0;
  do {
    const size_t n = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + n == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((uint8_t)(n & 0xFF));
    z.push_back((uint8_t)(n >> 8));
    z.push_back((uint8_t)(~n & 0xFF));
    z.push_back((uint8_t)((~n >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
    off += n;
  } while (off < raw.size());
  put32(z, (b << 16) | a);
  chunk(f, "IDAT", z);
  chunk(f, "IEND", {});
  if (!f) throw std::runtime_error("write failed: " + path);
}

void write_jpg(const std::string& path, int w, int h, const uint8_t* rgb, int quality) {
  if (w < 1 || h < 1 || w > 65535 || h > 65535) throw std::invalid_argument("jpeg: size outside 1..65535");
  static const JpegEncoder enc;
  const bool sub = quality <= 90;  // 4:2:0 at 90 and below
  uint8_t qy[64], qc[64];
  scale_table(kLumQ, quality, qy);
  scale_table(kChrQ, quality, qc);

  std::vector<uint8_t> o;
  o.reserve((size_t)w * h * 3 / 2 + 1024);
  marker(o, 0xD8);  // SOI
  marker(o, 0xE0);  // APP0 JFIF 1.01, aspect 1:1
  put16(o, 16);
  o.insert(o.end(), {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0});
  marker(o, 0xDB);  // DQT: tables 0 and 1, 8-bit, zig-zag order
  put16(o, 2 + 2 * 65);
  for (int t = 0; t < 2; ++t) {
    o.push_back((uint8_t)t);
    for (int k = 0; k < 64; ++k) o.push_back((t ? qc : qy)[enc.zz.nat[k]]);
  }
  marker(o, 0xC0);  // SOF0: baseline, 8-bit, 3 components
  put16(o, 17);
  o.push_back(8);
  put16(o, h);
  put16(o, w);
  o.push_back(3);
  o.insert(o.end(), {1, (uint8_t)(sub ? 0x22 : 0x11), 0, 2, 0x11, 1, 3, 0x11, 1});
  marker(o, 0xC4);  // DHT: DC0, AC0, DC1, AC1
  put16(o, 2 + 4 * 17 + 12 + 12 + 162 + 162);
  const struct {
    uint8_t id;
    const uint8_t *bits, *vals;
    int n;
  } tabs[4] = {{0x00, kDcLumBits, kDcVals, 12},
               {0x10, kAcLumBits, kAcLumVals, 162},
               {0x01, kDcChrBits, kDcVals, 12},
               {0x11, kAcChrBits, kAcChrVals, 162}};
  for (const auto& t : tabs) {
    o.push_back(t.id);
    o.insert(o.end(), t.bits, t.bits + 16);
    o.insert(o.end(), t.vals, t.vals + t.n);
  }
  marker(o, 0xDA);  // SOS: all 3 components, full spectrum
  put16(o, 12);
  o.insert(o.end(), {3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});

  // image row r is frame row h-1-r; pixels beyond the edge repeat the last row/column
  auto px = [&](int x, int y, double* ycc) {
    const uint8_t* p = rgb + ((size_t)(h - 1 - std::min(y, h - 1)) * w + std::min(x, w - 1)) * 3;
    const double r = p[0], g = p[1], b = p[2];
    ycc[0] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
    ycc[1] = -0.168736 * r - 0.331264 * g + 0.5 * b;
    ycc[2] = 0.5 * r - 0.418688 * g - 0.081312 * b;
  };
  BitWriter bw(o);
  int dc[3] = {0, 0, 0};
  const int mcu = sub ? 16 : 8;
  double by[4][64], bc[2][64], ycc[3];
  for (int my = 0; my < h; my += mcu)
    for (int mx = 0; mx < w; mx += mcu) {
      if (!sub) {
        for (int y = 0; y < 8; ++y)
          for (int x = 0; x < 8; ++x) {
            px(mx + x, my + y, ycc);
            by[0][y * 8 + x] = ycc[0];
            bc[0][y * 8 + x] = ycc[1];
            bc[1][y * 8 + x] = ycc[2];
          }
        dc[0] = enc.block(bw, by[0], qy, dc[0], enc.dc_lum, enc.ac_lum);
      } else {
        for (int i = 0; i < 64; ++i) bc[0][i] = bc[1][i] = 0;
        for (int y = 0; y < 16; ++y)
          for (int x = 0; x < 16; ++x) {
            px(mx + x, my + y, ycc);
            by[(y >> 3) * 2 + (x >> 3)][(y & 7) * 8 + (x & 7)] = ycc[0];
            bc[0][(y >> 1) * 8 + (x >> 1)] += 0.25 * ycc[1];
            bc[1][(y >> 1) * 8 + (x >> 1)] += 0.25 * ycc[2];
          }
        for (int k = 0; k < 4; ++k) dc[0] = enc.block(bw, by[k], qy, dc[0], enc.dc_lum, enc.ac_lum);
      }
      dc[1] = enc.block(bw, bc[0], qc, dc[1], enc.dc_chr, enc.ac_chr);
      dc[2] = enc.block(bw, bc[1], qc, dc[2], enc.dc_chr, enc.ac_chr);
    }
  bw.flush();
  marker(o, 0xD9);  // EOI

  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write((const char*)o.data(), (std::streamsize)o.size());
  if (!f) throw std::runtime_error("write failed: " + path);
}

}  // namespace r2host

extern "C" int r2_image_write(const char* path, int32_t w, int32_t h, const uint8_t* rgb, int32_t format,
                              int32_t quality, char* err, int32_t err_len) {
  auto fail = [&](int rc, const char* msg) {
    if (err && err_len > 0) {
      strncpy(err, msg, (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return rc;
  };
  if (!path || !rgb) return fail(R2_ERR_INVALID, "null path or pixels");
  if (w < 1 || h < 1 || (int64_t)w * h > (1ll << 28)) return fail(R2_ERR_INVALID, "image size out of range");
  if (format != R2_IMAGE_PNG && format != R2_IMAGE_JPG) return fail(R2_ERR_INVALID, "unknown image format");
  if (format == R2_IMAGE_JPG && (quality < 1 || quality > 100 || w > 65535 || h > 65535))
    return fail(R2_ERR_INVALID, "jpeg quality outside 1..100 or size above 65535");
  try {
    if (format == R2_IMAGE_PNG)
      r2host::write_png(path, w, h, rgb);
    else
      r2host::write_jpg(path, w, h, rgb, quality);
  } catch (const std::exception& e) {
    return fail(R2_ERR_IO, e.what());
  }
  return R2_OK;
}
