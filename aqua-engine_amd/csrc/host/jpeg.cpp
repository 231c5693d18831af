// JPEG decoder for `{"Image": "textures\\*.jpg"}` colours (scenes/room.json:6,
// 194, 382 and 13 others): ITU-T T.81 Huffman-coded DCT images, baseline /
// extended sequential (SOF0, SOF1) and progressive (SOF2), 8-bit samples,
// 1 or 3 components with any 1..2 sampling factors, restart intervals;
// YCbCr -> RGB per JFIF (or plain RGB when an Adobe APP14 segment says so).
// 4 of the 17 room textures are progressive, one is 4:2:0.
//
// Coefficients of every block are kept until the last scan (progressive
// scans refine them in place), then dequantised and inverse-transformed with
// a separable float IDCT. Subsampled chroma is upsampled with the triangle
// ("fancy") filter that libjpeg applies by default, so decoded texels agree
// with common decoders to within a few levels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace aqua {
namespace {

// Zig-zag scan position k -> natural (row-major) index.
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
  bool defined = false;
  uint8_t vals[256] = {};
  int32_t maxcode[18] = {};   // largest code of each length, -1 if none
  int32_t mincode[17] = {};
  int32_t valptr[17] = {};
  // 9-bit lookahead: (length << 8) | value for codes of length <= 9, 0 = longer
  uint16_t fast[512] = {};

  bool build(const uint8_t counts[16], const uint8_t* v, int nv) {
    memcpy(vals, v, size_t(nv));
    int code = 0, k = 0;
    memset(fast, 0, sizeof fast);
    for (int len = 1; len <= 16; ++len) {
      int n = counts[len - 1];
      valptr[len] = k;
      mincode[len] = code;
      for (int i = 0; i < n; ++i, ++k, ++code) {
        if (len <= 9) {
          int lo = code << (9 - len), hi = (code + 1) << (9 - len);
          if (hi > 512) return false;
          for (int c = lo; c < hi; ++c) fast[c] = uint16_t((len << 8) | vals[k]);
        }
      }
      maxcode[len] = n ? code - 1 : -1;
      if (code > (1 << len)) return false;  // over-subscribed
      code <<= 1;
    }
    maxcode[17] = 0x7FFFFFFF;
    defined = true;
    return true;
  }
};

struct Component {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;         // table selectors of the current scan
  int bw = 0, bh = 0;         // blocks per row / column (padded to whole MCUs)
  int pred = 0;               // DC predictor
  std::vector<int16_t> coef;  // bw * bh * 64, natural order
  std::vector<uint8_t> plane; // bw * 8 x bh * 8 decoded samples
};

class Decoder {
 public:
  Decoder(const uint8_t* d, size_t n) : data_(d), size_(n) {}

  bool run(uint32_t* w, uint32_t* h, std::vector<uint8_t>* rgba, std::string* err) {
    if (!decode()) {
      *err = err_;
      return false;
    }
    *w = uint32_t(width_);
    *h = uint32_t(height_);
    output(rgba);
    return true;
  }

 private:
  const uint8_t* data_;
  size_t size_;
  size_t pos_ = 0;
  std::string err_;
  uint16_t qt_[4][64] = {};  // natural order
  Huffman dc_[4], ac_[4];
  std::vector<Component> comp_;
  int width_ = 0, height_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
  bool progressive_ = false, frame_ = false;
  int restart_ = 0;
  int adobe_transform_ = -1;  // -1: no APP14 segment
  bool jfif_ = false;
  // entropy decoder state
  uint64_t acc_ = 0;
  int nbits_ = 0;
  bool marker_ = false;  // a marker stopped the bit stream
  int eobrun_ = 0;

  bool fail(const char* m) {
    err_ = m;
    return false;
  }
  bool need(size_t n) const { return pos_ + n <= size_; }
  int u8() { return pos_ < size_ ? data_[pos_++] : 0; }
  int u16() {
    int a = u8();
    return (a << 8) | u8();
  }

  // ---- bit stream (0xFF00 = a literal 0xFF; any other marker ends the data)
  void fill() {
    while (nbits_ <= 56) {
      int b = 0;
      if (!marker_ && pos_ < size_) {
        b = data_[pos_];
        if (b == 0xFF) {
          int nx = pos_ + 1 < size_ ? data_[pos_ + 1] : 0;
          if (nx == 0x00) {
            pos_ += 2;
          } else {
            marker_ = true;  // leave pos_ on the marker; feed zeros from here
            b = 0;
          }
        } else {
          ++pos_;
        }
      }
      acc_ |= uint64_t(b) << (56 - nbits_);
      nbits_ += 8;
    }
  }
  uint32_t bits(int n) {
    if (n == 0) return 0;
    if (nbits_ < n) fill();
    uint32_t v = uint32_t(acc_ >> (64 - n));
    acc_ <<= n;
    nbits_ -= n;
    return v;
  }
  int bit() { return int(bits(1)); }
  // T.81 F.2.2.1 EXTEND: an s-bit value with its top bit clear is negative.
  static int extend(uint32_t v, int s) { return (s > 0 && v < (1u << (s - 1))) ? int(v) - (1 << s) + 1 : int(v); }
  int receive_extend(int s) { return extend(bits(s), s); }

  int decode(const Huffman& t) {
    if (nbits_ < 16) fill();
    uint32_t peek = uint32_t(acc_ >> (64 - 9));
    uint16_t f = t.fast[peek];
    if (f) {
      int len = f >> 8;
      acc_ <<= len;
      nbits_ -= len;
      return f & 0xFF;
    }
    int32_t code = int32_t(peek);
    int len = 9;
    acc_ <<= 9;
    nbits_ -= 9;
    while (len < 16 && code > t.maxcode[len]) {
      code = (code << 1) | bit();
      ++len;
    }
    if (code > t.maxcode[len]) return -1;  // no such code
    return t.vals[t.valptr[len] + code - t.mincode[len]];
  }

  void reset_bits() {
    acc_ = 0;
    nbits_ = 0;
  }

  // Expects RSTn next (skipping fill bytes); resets predictors and EOB runs.
  bool restart() {
    reset_bits();
    marker_ = false;
    while (pos_ + 1 < size_ && !(data_[pos_] == 0xFF && data_[pos_ + 1] >= 0xD0 && data_[pos_ + 1] <= 0xD7)) {
      if (data_[pos_] == 0xFF && data_[pos_ + 1] != 0x00 && data_[pos_ + 1] != 0xFF) break;
      ++pos_;
    }
    if (pos_ + 1 < size_ && data_[pos_] == 0xFF && data_[pos_ + 1] >= 0xD0 && data_[pos_ + 1] <= 0xD7) pos_ += 2;
    for (Component& c : comp_) c.pred = 0;
    eobrun_ = 0;
    return true;
  }

  // ---- segments
  bool read_dqt(int len) {
    size_t end = pos_ + size_t(len);
    while (pos_ < end) {
      int pq_tq = u8();
      int pq = pq_tq >> 4, tq = pq_tq & 15;
      if (tq > 3) return fail("jpeg: bad quantisation table id");
      for (int k = 0; k < 64; ++k) qt_[tq][kZigzag[k]] = uint16_t(pq ? u16() : u8());
    }
    return pos_ == end || fail("jpeg: bad DQT length");
  }

  bool read_dht(int len) {
    size_t end = pos_ + size_t(len);
    while (pos_ < end) {
      int tc_th = u8();
      int tc = tc_th >> 4, th = tc_th & 15;
      if (tc > 1 || th > 3) return fail("jpeg: bad Huffman table id");
      uint8_t counts[16];
      int total = 0;
      for (int i = 0; i < 16; ++i) {
        counts[i] = uint8_t(u8());
        total += counts[i];
      }
      if (total > 256 || !need(size_t(total))) return fail("jpeg: bad Huffman table");
      uint8_t v[256];
      for (int i = 0; i < total; ++i) v[i] = uint8_t(u8());
      if (!(tc == 0 ? dc_[th] : ac_[th]).build(counts, v, total)) return fail("jpeg: bad Huffman code lengths");
    }
    return pos_ == end || fail("jpeg: bad DHT length");
  }

  bool read_sof(int len, int marker) {
    if (frame_) return fail("jpeg: more than one frame");
    if (u8() != 8) return fail("jpeg: only 8-bit samples are supported");
    height_ = u16();
    width_ = u16();
    int nc = u8();
    if (width_ <= 0 || height_ <= 0) return fail("jpeg: image size 0 (DNL not supported)");
    if (width_ > 32768 || height_ > 32768) return fail("jpeg: image too large");
    if (nc != 1 && nc != 3) return fail("jpeg: only 1- and 3-component images are supported");
    if (len != 6 + 3 * nc) return fail("jpeg: bad SOF length");
    comp_.resize(size_t(nc));
    for (Component& c : comp_) {
      c.id = u8();
      int hv = u8();
      c.h = hv >> 4;
      c.v = hv & 15;
      c.tq = u8();
      if (c.h < 1 || c.h > 2 || c.v < 1 || c.v > 2 || c.tq > 3) return fail("jpeg: unsupported sampling factors");
      hmax_ = std::max(hmax_, c.h);
      vmax_ = std::max(vmax_, c.v);
    }
    mcux_ = (width_ + 8 * hmax_ - 1) / (8 * hmax_);
    mcuy_ = (height_ + 8 * vmax_ - 1) / (8 * vmax_);
    for (Component& c : comp_) {
      c.bw = mcux_ * c.h;
      c.bh = mcuy_ * c.v;
      c.coef.assign(size_t(c.bw) * c.bh * 64, 0);
    }
    progressive_ = marker == 0xC2;
    frame_ = true;
    return true;
  }

  // ---- block decoders (one 8x8 block of component c at block (bx, by))
  bool block_baseline(Component& c, int16_t* b) {
    int t = decode(dc_[c.td]);
    if (t < 0 || t > 11) return fail("jpeg: bad DC code");
    c.pred += receive_extend(t);
    b[0] = int16_t(c.pred);
    for (int k = 1; k < 64;) {
      int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("jpeg: bad AC code");
      int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r != 15) break;  // end of block
        k += 16;
        continue;
      }
      k += r;
      if (k > 63) return fail("jpeg: AC run past the block");
      b[kZigzag[k]] = int16_t(receive_extend(s));
      ++k;
    }
    return true;
  }

  bool block_dc_first(Component& c, int16_t* b, int al) {
    int t = decode(dc_[c.td]);
    if (t < 0 || t > 11) return fail("jpeg: bad DC code");
    c.pred += receive_extend(t);
    b[0] = int16_t(c.pred * (1 << al));
    return true;
  }

  void block_dc_refine(int16_t* b, int al) {
    if (bit()) b[0] = int16_t(b[0] | (1 << al));
  }

  bool block_ac_first(Component& c, int16_t* b, int ss, int se, int al) {
    if (eobrun_ > 0) {
      --eobrun_;
      return true;
    }
    for (int k = ss; k <= se; ++k) {
      int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("jpeg: bad AC code");
      int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (r < 15) {  // EOB run of 2^r + extra bits blocks, this one included
          eobrun_ = (1 << r) - 1;
          if (r) eobrun_ += int(bits(r));
          break;
        }
        k += 15;  // 16 zeros (the loop adds the 16th)
        continue;
      }
      k += r;
      if (k > 63) return fail("jpeg: AC run past the block");
      b[kZigzag[k]] = int16_t(receive_extend(s) * (1 << al));
    }
    return true;
  }

  // Successive-approximation refinement of AC coefficients (T.81 G.1.2.3):
  // a coefficient already non-zero receives one correction bit when the scan
  // passes it; a new coefficient (magnitude 1 << al) lands after r still-zero
  // positions.
  bool block_ac_refine(Component& c, int16_t* b, int ss, int se, int al) {
    const int p1 = 1 << al, m1 = -(1 << al);
    auto correct = [&](int16_t* x) {
      if (bit() && (*x & p1) == 0) *x = int16_t(*x + (*x >= 0 ? p1 : m1));
    };
    int k = ss;
    if (eobrun_ == 0) {
      for (; k <= se; ++k) {
        int rs = decode(ac_[c.ta]);
        if (rs < 0) return fail("jpeg: bad AC code");
        int r = rs >> 4, s = rs & 15;
        int val = 0;
        if (s != 0) {
          if (s != 1) return fail("jpeg: bad refinement magnitude");
          val = bit() ? p1 : m1;
        } else if (r != 15) {
          eobrun_ = 1 << r;
          if (r) eobrun_ += int(bits(r));
          break;  // the rest of this block is an EOB run block
        }
        // skip r zero positions (correcting the non-zero ones met on the way)
        for (; k <= se; ++k) {
          int16_t* x = &b[kZigzag[k]];
          if (*x != 0) {
            correct(x);
          } else {
            if (r == 0) break;
            --r;
          }
        }
        if (val != 0) {
          if (k > se) return fail("jpeg: refinement past the band");
          b[kZigzag[k]] = int16_t(val);
        }
      }
    }
    if (eobrun_ > 0) {
      for (; k <= se; ++k) {
        int16_t* x = &b[kZigzag[k]];
        if (*x != 0) correct(x);
      }
      --eobrun_;
    }
    return true;
  }

  bool read_scan(int len) {
    int ns = u8();
    if (ns < 1 || ns > int(comp_.size()) || len != 4 + 2 * ns) return fail("jpeg: bad SOS");
    std::vector<Component*> sc;
    for (int i = 0; i < ns; ++i) {
      int id = u8(), t = u8();
      Component* c = nullptr;
      for (Component& k : comp_)
        if (k.id == id) c = &k;
      if (!c) return fail("jpeg: scan names an unknown component");
      c->td = t >> 4;
      c->ta = t & 15;
      if (c->td > 3 || c->ta > 3) return fail("jpeg: bad table selector");
      sc.push_back(c);
    }
    int ss = u8(), se = u8(), a = u8();
    int ah = a >> 4, al = a & 15;
    if (!progressive_) {
      ss = 0;
      se = 63;
      ah = al = 0;
    } else {
      if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && ns != 1) || al > 13)
        return fail("jpeg: bad progressive scan parameters");
    }
    // the tables this scan uses must exist
    for (Component* c : sc) {
      bool dc = !progressive_ || (ss == 0 && ah == 0);
      bool ac = !progressive_ || ss > 0;
      if ((dc && !dc_[c->td].defined) || (ac && !ac_[c->ta].defined)) return fail("jpeg: missing Huffman table");
    }
    reset_bits();
    marker_ = false;
    eobrun_ = 0;
    for (Component& c : comp_) c.pred = 0;
    auto one_block = [&](Component& c, int bx, int by) -> bool {
      int16_t* b = &c.coef[(size_t(by) * c.bw + bx) * 64];
      if (!progressive_) return block_baseline(c, b);
      if (ss == 0) {
        if (ah == 0) return block_dc_first(c, b, al);
        block_dc_refine(b, al);
        return true;
      }
      return ah == 0 ? block_ac_first(c, b, ss, se, al) : block_ac_refine(c, b, ss, se, al);
    };
    int todo = restart_;
    if (ns == 1) {
      // non-interleaved: the component's own blocks in raster order, only
      // those covering the image (not the MCU padding)
      Component& c = *sc[0];
      int cw = (width_ * c.h + hmax_ - 1) / hmax_, ch = (height_ * c.v + vmax_ - 1) / vmax_;
      int nbx = (cw + 7) / 8, nby = (ch + 7) / 8;
      for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
          if (restart_ && todo == 0) {
            restart();
            todo = restart_;
          }
          if (!one_block(c, bx, by)) return false;
          --todo;
        }
    } else {
      for (int my = 0; my < mcuy_; ++my)
        for (int mx = 0; mx < mcux_; ++mx) {
          if (restart_ && todo == 0) {
            restart();
            todo = restart_;
          }
          for (Component* c : sc)
            for (int v = 0; v < c->v; ++v)
              for (int h = 0; h < c->h; ++h)
                if (!one_block(*c, mx * c->h + h, my * c->v + v)) return false;
          --todo;
        }
    }
    // continue at the next marker
    reset_bits();
    marker_ = false;
    while (pos_ + 1 < size_ && !(data_[pos_] == 0xFF && data_[pos_ + 1] != 0x00 &&
                                  !(data_[pos_ + 1] >= 0xD0 && data_[pos_ + 1] <= 0xD7)))
      ++pos_;
    return true;
  }

  bool decode() {
    if (size_ < 4 || data_[0] != 0xFF || data_[1] != 0xD8) return fail("jpeg: missing SOI (not a JPEG file)");
    pos_ = 2;
    bool scanned = false;
    while (true) {
      // next marker (skipping fill bytes)
      while (pos_ < size_ && data_[pos_] != 0xFF) ++pos_;
      while (pos_ < size_ && data_[pos_] == 0xFF) ++pos_;
      if (pos_ >= size_) break;
      int m = data_[pos_++];
      if (m == 0xD9) break;                    // EOI
      if (m >= 0xD0 && m <= 0xD7) continue;    // stray RST
      if (m == 0x01 || m == 0xD8) continue;    // TEM, repeated SOI
      if (!need(2)) return fail("jpeg: truncated segment");
      int len = u16() - 2;
      if (len < 0 || !need(size_t(len))) return fail("jpeg: truncated segment");
      size_t next = pos_ + size_t(len);
      switch (m) {
        case 0xC0:
        case 0xC1:
        case 0xC2:
          if (!read_sof(len, m)) return false;
          break;
        case 0xC3:
        case 0xC5:
        case 0xC6:
        case 0xC7:
        case 0xC9:
        case 0xCA:
        case 0xCB:
        case 0xCD:
        case 0xCE:
        case 0xCF:
          return fail("jpeg: lossless, hierarchical and arithmetic-coded JPEG are not supported");
        case 0xC4:
          if (!read_dht(len)) return false;
          break;
        case 0xDB:
          if (!read_dqt(len)) return false;
          break;
        case 0xDD:
          if (len != 2) return fail("jpeg: bad DRI");
          restart_ = u16();
          break;
        case 0xDA:
          if (!frame_) return fail("jpeg: scan before frame header");
          if (!read_scan(len)) return false;
          scanned = true;
          continue;  // read_scan left pos_ at the next marker
        case 0xE0:
          if (len >= 5 && memcmp(data_ + pos_, "JFIF", 4) == 0) jfif_ = true;
          break;
        case 0xEE:
          if (len >= 12 && memcmp(data_ + pos_, "Adobe", 5) == 0) adobe_transform_ = data_[pos_ + 11];
          break;
        default:
          break;  // APPn, COM, ...: skipped
      }
      pos_ = next;
    }
    if (!frame_ || !scanned) return fail("jpeg: no image data");
    finish();
    return true;
  }

  // ---- dequantise + IDCT of every block into the component planes
  void finish() {
    // C(u) cos((2x + 1) u pi / 16) / 2, C(0) = 1 / sqrt(2)
    float cs[8][8];
    for (int x = 0; x < 8; ++x)
      for (int u = 0; u < 8; ++u)
        cs[x][u] = float((u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0) * 0.5);
    for (Component& c : comp_) {
      const uint16_t* q = qt_[c.tq];
      c.plane.assign(size_t(c.bw) * 8 * size_t(c.bh) * 8, 0);
      const size_t stride = size_t(c.bw) * 8;
      for (int by = 0; by < c.bh; ++by)
        for (int bx = 0; bx < c.bw; ++bx) {
          const int16_t* b = &c.coef[(size_t(by) * c.bw + bx) * 64];
          float f[64], tmp[64];
          for (int i = 0; i < 64; ++i) f[i] = float(int(b[i]) * int(q[i]));
          // rows: tmp[v][x] = sum_u cs[x][u] f[v][u]
          for (int v = 0; v < 8; ++v)
            for (int x = 0; x < 8; ++x) {
              float s = 0.0f;
              for (int u = 0; u < 8; ++u) s += cs[x][u] * f[v * 8 + u];
              tmp[v * 8 + x] = s;
            }
          // columns: out[y][x] = sum_v cs[y][v] tmp[v][x]
          for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
              float s = 0.0f;
              for (int v = 0; v < 8; ++v) s += cs[y][v] * tmp[v * 8 + x];
              int p = int(std::lrint(s + 128.0f));
              c.plane[(size_t(by) * 8 + y) * stride + size_t(bx) * 8 + x] = uint8_t(p < 0 ? 0 : (p > 255 ? 255 : p));
            }
        }
      c.coef.clear();
      c.coef.shrink_to_fit();
    }
  }

  // Component c's samples upsampled to full resolution (width_ x height_).
  std::vector<uint8_t> full_plane(const Component& c) const {
    const int W = width_, H = height_;
    const int sw = c.bw * 8;
    std::vector<uint8_t> out(size_t(W) * H);
    const int fx = hmax_ / c.h, fy = vmax_ / c.v;
    auto at = [&](int x, int y) -> int { return c.plane[size_t(y) * sw + x]; };
    // component extent in its own samples
    const int cw = (W * c.h + hmax_ - 1) / hmax_, ch = (H * c.v + vmax_ - 1) / vmax_;
    if (fx == 1 && fy == 1) {
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) out[size_t(y) * W + x] = uint8_t(at(x, y));
      return out;
    }
    // Triangle filter: each output sample sits 1/4 of an input sample from
    // its nearest input, weights 3/4 and 1/4 per upsampled axis.
    std::vector<int> col(static_cast<size_t>(cw));
    for (int y = 0; y < H; ++y) {
      // vertical pass (x 4 when fy == 2, x 1 otherwise)
      int yi = y / fy;
      if (fy == 2) {
        int yn = (y & 1) ? std::min(yi + 1, ch - 1) : std::max(yi - 1, 0);
        for (int x = 0; x < cw; ++x) col[size_t(x)] = 3 * at(x, yi) + at(x, yn);
      } else {
        for (int x = 0; x < cw; ++x) col[size_t(x)] = at(x, yi);
      }
      uint8_t* row = &out[size_t(y) * W];
      if (fx == 1) {
        for (int x = 0; x < W; ++x) row[x] = uint8_t(fy == 2 ? (col[size_t(x)] + 1 + (y & 1)) >> 2 : col[size_t(x)]);
        continue;
      }
      for (int x = 0; x < W; ++x) {
        int xi = x / 2;
        int xn = (x & 1) ? std::min(xi + 1, cw - 1) : std::max(xi - 1, 0);
        int v;
        if (fy == 2)
          v = (3 * col[size_t(xi)] + col[size_t(xn)] + ((x & 1) ? 7 : 8)) >> 4;
        else
          v = (3 * col[size_t(xi)] + col[size_t(xn)] + ((x & 1) ? 2 : 1)) >> 2;
        row[x] = uint8_t(v);
      }
    }
    return out;
  }

  void output(std::vector<uint8_t>* rgba) const {
    const size_t n = size_t(width_) * height_;
    rgba->assign(4 * n, 255);
    if (comp_.size() == 1) {
      std::vector<uint8_t> g = full_plane(comp_[0]);
      for (size_t i = 0; i < n; ++i) (*rgba)[4 * i] = (*rgba)[4 * i + 1] = (*rgba)[4 * i + 2] = g[i];
      return;
    }
    std::vector<uint8_t> p0 = full_plane(comp_[0]), p1 = full_plane(comp_[1]), p2 = full_plane(comp_[2]);
    bool rgb = adobe_transform_ == 0 ||
               (adobe_transform_ < 0 && !jfif_ && comp_[0].id == 'R' && comp_[1].id == 'G' && comp_[2].id == 'B');
    for (size_t i = 0; i < n; ++i) {
      if (rgb) {
        (*rgba)[4 * i] = p0[i];
        (*rgba)[4 * i + 1] = p1[i];
        (*rgba)[4 * i + 2] = p2[i];
        continue;
      }
      // JFIF (ITU-R BT.601 full range)
      float y = float(p0[i]), cb = float(p1[i]) - 128.0f, cr = float(p2[i]) - 128.0f;
      float r = y + 1.402f * cr, g = (y - 0.344136f * cb) - 0.714136f * cr, b = y + 1.772f * cb;
      auto clamp8 = [](float v) -> uint8_t {
        int k = int(std::lrint(v));
        return uint8_t(k < 0 ? 0 : (k > 255 ? 255 : k));
      };
      (*rgba)[4 * i] = clamp8(r);
      (*rgba)[4 * i + 1] = clamp8(g);
      (*rgba)[4 * i + 2] = clamp8(b);
    }
  }
};

}  // namespace

int decode_jpeg(const std::vector<uint8_t>& data, uint32_t* w, uint32_t* h, std::vector<uint8_t>* rgba) {
  std::string err;
  Decoder d(data.data(), data.size());
  if (!d.run(w, h, rgba, &err)) return fail(AQUA_ERR_PARSE, "%s", err.c_str());
  return AQUA_OK;
}

}  // namespace aqua
