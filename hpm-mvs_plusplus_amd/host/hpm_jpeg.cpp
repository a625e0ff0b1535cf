// hpm_jpeg.cpp -- baseline JPEG decoding for images/%08d.jpg, producing the
// pixels cv::imread(..., IMREAD_COLOR) produces through libjpeg(-turbo) with
// its defaults: Huffman-coded sequential or progressive DCT (SOF0 / SOF1 /
// SOF2), 8-bit samples,
// 1 or 3 components, any sampling factors up to 4, restart intervals,
// interleaved or per-component scans; the accurate integer IDCT (the
// Loeffler-Ligtenberg-Moschytz factorisation with 13-bit constants and two
// fraction bits between the passes), "fancy" triangle upsampling of
// 2x / 1x2 / 2x2 subsampled chroma, and the 16-bit fixed-point YCbCr -> RGB
// of ITU-R BT.601 (JFIF).  Arithmetic-coded, lossless and hierarchical files
// are rejected with a message.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hpm_host.hpp"

namespace hpmhost {
namespace {

constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
  // canonical code: for each length L, the first code, the count, and the
  // index of its first symbol
  int mincode[17], maxcode[18], valptr[17];
  uint8_t vals[256];
  bool defined = false;
};

struct Component {
  int id, h, v, tq;
  int td = 0, ta = 0;
  int bw = 0, bh = 0;          // blocks per row / column (padded to whole MCUs)
  int dw = 0, dh = 0;          // downsampled width / height (real samples)
  std::vector<int16_t> coef;   // bw * bh blocks of 64 coefficients, natural order
  int pred = 0;
};

class Decoder {
 public:
  Decoder(const uint8_t* p, size_t n) : d_(p), n_(n) {}
  bool run(MapU8& out, std::string& err);
  bool size_only(int& w, int& h, std::string& err);

 private:
  const uint8_t* d_;
  size_t n_, pos_ = 0;
  uint16_t q_[4][64] = {};
  bool qdef_[4] = {};
  Huffman dc_[4], ac_[4];
  std::vector<Component> comp_;
  int W_ = 0, H_ = 0, hmax_ = 1, vmax_ = 1, restart_ = 0, mcux_ = 0, mcuy_ = 0;
  bool frame_ = false, progressive_ = false;
  int eobrun_ = 0;
  // the current scan's spectral band and successive-approximation bits
  int ss_ = 0, se_ = 63, ah_ = 0, al_ = 0;
  // bit reader
  uint32_t acc_ = 0;
  int bits_ = 0;
  bool marker_hit_ = false;
  std::string err_;

  bool fail(const std::string& m) {
    if (err_.empty()) err_ = m;
    return false;
  }
  int byte() { return pos_ < n_ ? d_[pos_++] : -1; }
  int u16() {
    const int a = byte(), b = byte();
    return (a < 0 || b < 0) ? -1 : (a << 8) | b;
  }
  bool segment(int marker);
  bool sof(size_t len, int marker);
  bool dqt(size_t len);
  bool dht(size_t len);
  bool sos(size_t len);
  void reset_bits() { acc_ = 0; bits_ = 0; marker_hit_ = false; }
  void fill() {
    while (bits_ <= 24) {
      int b = 0;
      if (!marker_hit_ && pos_ < n_) {
        b = d_[pos_];
        if (b == 0xFF) {
          const int nb = pos_ + 1 < n_ ? d_[pos_ + 1] : -1;
          if (nb == 0x00) {
            pos_ += 2;
          } else {  // a marker: feed zeros from here on, as libjpeg does
            marker_hit_ = true;
            b = 0;
          }
        } else {
          ++pos_;
        }
      }
      acc_ |= (uint32_t)b << (24 - bits_);
      bits_ += 8;
    }
  }
  int getbits(int n) {
    if (n == 0) return 0;
    fill();
    const int v = (int)(acc_ >> (32 - n));
    acc_ <<= n;
    bits_ -= n;
    return v;
  }
  int decode(const Huffman& h) {
    fill();
    int code = 0;
    for (int l = 1; l <= 16; ++l) {
      code = (code << 1) | (int)(acc_ >> 31);
      acc_ <<= 1;
      --bits_;
      if (code <= h.maxcode[l]) return h.vals[h.valptr[l] + code - h.mincode[l]];
    }
    return -1;  // corrupt data: no code of 16 bits or fewer
  }
  static int extend(int v, int s) { return (s > 0 && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }
  bool block(Component& c, int16_t* out);
  bool dc_first(Component& c, int16_t* out);
  bool dc_refine(int16_t* out);
  bool ac_first(Component& c, int16_t* out);
  bool ac_refine(Component& c, int16_t* out);
  bool unit(Component& c, int16_t* out);
  bool next_restart();
  void idct_and_upsample(MapU8& out);
};

bool Decoder::sof(size_t len, int marker) {
  if (marker != 0xC0 && marker != 0xC1 && marker != 0xC2)
    return fail("only Huffman-coded sequential or progressive JPEG (SOF0 / SOF1 / SOF2) is supported");
  progressive_ = marker == 0xC2;
  if (frame_) return fail("more than one frame");
  const int P = byte();
  H_ = u16();
  W_ = u16();
  const int nc = byte();
  if (P != 8) return fail("only 8-bit samples are supported");
  if (H_ <= 0 || W_ <= 0) return fail("bad image size (zero height: DNL is not supported)");
  // the engine takes views of at most 2^28 pixels; a larger header (up to
  // 65535 x 65535) is refused before its planes are allocated
  if ((int64_t)W_ * H_ > ((int64_t)1 << 28)) return fail("image larger than 2^28 pixels");
  if (nc != 1 && nc != 3) return fail("only 1 or 3 components are supported");
  if (len != (size_t)(8 + 3 * nc)) return fail("bad SOF length");
  comp_.resize(nc);
  for (auto& c : comp_) {
    c.id = byte();
    const int hv = byte();
    c.h = hv >> 4;
    c.v = hv & 15;
    c.tq = byte();
    if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq < 0 || c.tq > 3) return fail("bad component parameters");
    hmax_ = std::max(hmax_, c.h);
    vmax_ = std::max(vmax_, c.v);
  }
  mcux_ = (W_ + 8 * hmax_ - 1) / (8 * hmax_);
  mcuy_ = (H_ + 8 * vmax_ - 1) / (8 * vmax_);
  for (auto& c : comp_) {
    c.bw = mcux_ * c.h;
    c.bh = mcuy_ * c.v;
    c.dw = (int)(((long long)W_ * c.h + hmax_ - 1) / hmax_);
    c.dh = (int)(((long long)H_ * c.v + vmax_ - 1) / vmax_);
    c.coef.assign((size_t)c.bw * c.bh * 64, 0);
  }
  frame_ = true;
  return true;
}

bool Decoder::dqt(size_t len) {
  size_t used = 2;
  while (used < len) {
    const int pq = byte();
    if (pq < 0) return fail("truncated DQT");
    const int prec = pq >> 4, t = pq & 15;
    if (t > 3 || prec > 1) return fail("bad DQT");
    for (int k = 0; k < 64; ++k) {
      const int v = prec ? u16() : byte();
      if (v < 0) return fail("truncated DQT");
      q_[t][kZigzag[k]] = (uint16_t)v;
    }
    qdef_[t] = true;
    used += 1 + 64 * (prec ? 2 : 1);
  }
  return used == len || fail("bad DQT length");
}

bool Decoder::dht(size_t len) {
  size_t used = 2;
  while (used < len) {
    const int tc_th = byte();
    if (tc_th < 0) return fail("truncated DHT");
    const int tc = tc_th >> 4, th = tc_th & 15;
    if (tc > 1 || th > 3) return fail("bad DHT");
    Huffman& h = tc ? ac_[th] : dc_[th];
    int counts[17] = {0}, total = 0;
    for (int l = 1; l <= 16; ++l) {
      counts[l] = byte();
      if (counts[l] < 0) return fail("truncated DHT");
      total += counts[l];
    }
    if (total > 256) return fail("bad DHT");
    for (int k = 0; k < total; ++k) {
      const int v = byte();
      if (v < 0) return fail("truncated DHT");
      h.vals[k] = (uint8_t)v;
    }
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
      h.valptr[l] = k;
      h.mincode[l] = code;
      code += counts[l];
      k += counts[l];
      h.maxcode[l] = counts[l] ? code - 1 : -1;
      code <<= 1;
    }
    h.maxcode[17] = 0x7fffffff;
    h.defined = true;
    used += 17 + total;
  }
  return used == len || fail("bad DHT length");
}

bool Decoder::block(Component& c, int16_t* out) {
  const Huffman& dc = dc_[c.td];
  const Huffman& ac = ac_[c.ta];
  const int s = decode(dc);
  if (s < 0 || s > 11) return fail("corrupt JPEG data (DC)");
  c.pred += extend(getbits(s), s);
  out[0] = (int16_t)c.pred;  // quantised: the IDCT dequantises in int arithmetic
  for (int k = 1; k < 64;) {
    const int rs = decode(ac);
    if (rs < 0) return fail("corrupt JPEG data (AC)");
    const int r = rs >> 4, sz = rs & 15;
    if (sz == 0) {
      if (r != 15) break;  // end of block
      k += 16;
      continue;
    }
    k += r;
    if (k > 63) return fail("corrupt JPEG data (AC run)");
    out[kZigzag[k]] = (int16_t)extend(getbits(sz), sz);
    ++k;
  }
  return true;
}

// Progressive scans (ITU T.81 G.1.2): DC first / refinement, AC first /
// refinement over the band ss_..se_ with end-of-band runs.
bool Decoder::dc_first(Component& c, int16_t* out) {
  const int s = decode(dc_[c.td]);
  if (s < 0 || s > 11) return fail("corrupt JPEG data (DC)");
  c.pred += extend(getbits(s), s);
  out[0] = (int16_t)(c.pred * (1 << al_));
  return true;
}

bool Decoder::dc_refine(int16_t* out) {
  if (getbits(1)) out[0] = (int16_t)(out[0] | (1 << al_));
  return true;
}

bool Decoder::ac_first(Component& c, int16_t* out) {
  if (eobrun_ > 0) {
    --eobrun_;
    return true;
  }
  for (int k = ss_; k <= se_; ++k) {
    const int rs = decode(ac_[c.ta]);
    if (rs < 0) return fail("corrupt JPEG data (AC)");
    const int r = rs >> 4, sz = rs & 15;
    if (sz) {
      k += r;
      if (k > se_) return fail("corrupt JPEG data (AC run)");
      out[kZigzag[k]] = (int16_t)(extend(getbits(sz), sz) * (1 << al_));
    } else if (r == 15) {
      k += 15;  // sixteen zeros
    } else {
      eobrun_ = (1 << r) - 1;
      if (r) eobrun_ += getbits(r);
      break;
    }
  }
  return true;
}

bool Decoder::ac_refine(Component& c, int16_t* out) {
  const int p1 = 1 << al_, m1 = -(1 << al_);
  auto refine = [&](int16_t& v) {  // a correction bit for a coefficient already nonzero
    if (getbits(1) && (v & p1) == 0) v = (int16_t)(v + (v >= 0 ? p1 : m1));
  };
  int k = ss_;
  if (eobrun_ == 0) {
    for (; k <= se_; ++k) {
      const int rs = decode(ac_[c.ta]);
      if (rs < 0) return fail("corrupt JPEG data (AC refinement)");
      int r = rs >> 4;
      const int sz = rs & 15;
      int value = 0;
      if (sz) {
        if (sz != 1) return fail("corrupt JPEG data (AC refinement size)");
        value = getbits(1) ? p1 : m1;
      } else if (r != 15) {
        eobrun_ = 1 << r;
        if (r) eobrun_ += getbits(r);
        break;  // the band's rest is refined below
      }
      // pass over coefficients already nonzero (refining them) and r zeros
      do {
        int16_t& v = out[kZigzag[k]];
        if (v != 0) {
          refine(v);
        } else if (--r < 0) {
          break;  // the zero that takes the new value
        }
        ++k;
      } while (k <= se_);
      if (sz) {
        if (k > se_) return fail("corrupt JPEG data (AC refinement run)");
        out[kZigzag[k]] = (int16_t)value;
      }
    }
  }
  if (eobrun_ > 0) {
    for (; k <= se_; ++k) {
      int16_t& v = out[kZigzag[k]];
      if (v != 0) refine(v);
    }
    --eobrun_;
  }
  return true;
}

bool Decoder::unit(Component& c, int16_t* out) {
  if (!progressive_) return block(c, out);
  if (ss_ == 0) return ah_ == 0 ? dc_first(c, out) : dc_refine(out);
  return ah_ == 0 ? ac_first(c, out) : ac_refine(c, out);
}

bool Decoder::next_restart() {
  // discard the remaining bits, then expect RSTn
  reset_bits();
  while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)) ++pos_;
  if (pos_ + 1 >= n_) return fail("missing restart marker");
  pos_ += 2;
  for (auto& c : comp_) c.pred = 0;
  eobrun_ = 0;
  return true;
}

bool Decoder::sos(size_t len) {
  if (!frame_) return fail("scan before frame");
  const int ns = byte();
  if (ns < 1 || ns > (int)comp_.size() || len != (size_t)(6 + 2 * ns)) return fail("bad SOS");
  std::vector<Component*> sc;
  for (int i = 0; i < ns; ++i) {
    const int cid = byte(), t = byte();
    Component* found = nullptr;
    for (auto& c : comp_)
      if (c.id == cid) found = &c;
    if (!found) return fail("scan names an unknown component");
    found->td = t >> 4;
    found->ta = t & 15;
    if (found->td > 3 || found->ta > 3) return fail("bad Huffman table selector");
    if (!qdef_[found->tq]) return fail("component uses an undefined quantisation table");
    sc.push_back(found);
  }
  ss_ = byte();
  se_ = byte();
  const int ahal = byte();
  ah_ = ahal >> 4;
  al_ = ahal & 15;
  if (!progressive_) {
    if (ss_ != 0 || se_ != 63 || ahal != 0) return fail("not a sequential scan");
  } else if (ss_ > se_ || se_ > 63 || al_ > 13 || (ss_ == 0 && se_ != 0) || (ss_ > 0 && ns != 1)) {
    return fail("bad progressive scan parameters");
  }
  // tables the scan decodes with: DC in sequential and DC-first scans, AC in
  // sequential and AC scans
  const bool uses_dc = !progressive_ || (ss_ == 0 && ah_ == 0), uses_ac = !progressive_ || ss_ > 0;
  for (auto* c : sc)
    if ((uses_dc && !dc_[c->td].defined) || (uses_ac && !ac_[c->ta].defined))
      return fail("scan uses an undefined Huffman table");
  for (auto* c : sc) c->pred = 0;
  eobrun_ = 0;
  reset_bits();
  int rst = restart_;
  if (ns == 1) {  // non-interleaved: the component's own block grid, real blocks only
    Component& c = *sc[0];
    const int bx = (c.dw + 7) / 8, by = (c.dh + 7) / 8;
    for (int y = 0; y < by; ++y)
      for (int x = 0; x < bx; ++x) {
        if (restart_ && rst == 0) {
          if (!next_restart()) return false;
          rst = restart_;
        }
        if (!unit(c, &c.coef[((size_t)y * c.bw + x) * 64])) return false;
        if (restart_) --rst;
      }
  } else {
    for (int my = 0; my < mcuy_; ++my)
      for (int mx = 0; mx < mcux_; ++mx) {
        if (restart_ && rst == 0) {
          if (!next_restart()) return false;
          rst = restart_;
        }
        for (auto* c : sc)
          for (int v = 0; v < c->v; ++v)
            for (int h = 0; h < c->h; ++h) {
              const size_t b = (size_t)(my * c->v + v) * c->bw + (mx * c->h + h);
              if (!unit(*c, &c->coef[b * 64])) return false;
            }
        if (restart_) --rst;
      }
  }
  // resynchronise on the next marker
  reset_bits();
  while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] != 0x00 && !(d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)))
    ++pos_;
  return true;
}

bool Decoder::segment(int marker) {
  const int len = u16();
  if (len < 2 || pos_ + len - 2 > n_) return fail("truncated segment");
  const size_t start = pos_;
  bool ok = true;
  if ((marker >= 0xC0 && marker <= 0xCF) && marker != 0xC4 && marker != 0xC8 && marker != 0xCC) ok = sof(len, marker);
  else if (marker == 0xC4) ok = dht(len);
  else if (marker == 0xDB) ok = dqt(len);
  else if (marker == 0xDD) { restart_ = u16(); ok = restart_ >= 0; }
  else if (marker == 0xDA) return sos(len);
  else if (marker == 0xCC) ok = fail("arithmetic-coded JPEG is not supported");
  pos_ = start + len - 2;  // APPn, COM and the rest are skipped
  return ok;
}

bool Decoder::size_only(int& w, int& h, std::string& err) {
  if (u16() != 0xFFD8) { err = "not a JPEG file"; return false; }
  while (pos_ < n_) {
    int b = byte();
    if (b != 0xFF) continue;
    int m = byte();
    while (m == 0xFF) m = byte();
    if (m < 0) break;
    if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) continue;
    if (m == 0xD9) break;
    const int len = u16();
    if (len < 2) break;
    if (m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
      byte();
      h = u16();
      w = u16();
      return h > 0 && w > 0;
    }
    pos_ += len - 2;
  }
  err = "no frame header";
  return false;
}

// Accurate integer IDCT of one block (natural order, dequantised), written
// as samples clamped by the post-IDCT range table (values wrap modulo 1024
// around the centre before clamping, as a corrupt block does in libjpeg).
inline uint8_t range_idct(int x) {
  const int k = x & 1023;
  if (k < 128) return (uint8_t)(k + 128);
  if (k < 512) return 255;
  if (k < 896) return 0;
  return (uint8_t)(k - 896);
}

void idct_islow(const int16_t* in, const uint16_t* q, uint8_t* out, int stride) {
  constexpr int CB = 13, P1 = 2;
  auto mul = [](int64_t a, int64_t b) { return a * b; };
  auto descale = [](int64_t x, int n) { return (x + ((int64_t)1 << (n - 1))) >> n; };
  constexpr int64_t F0_298 = 2446, F0_390 = 3196, F0_541 = 4433, F0_765 = 6270, F0_899 = 7373, F1_175 = 9633,
                    F1_501 = 12299, F1_847 = 15137, F1_961 = 16069, F2_053 = 16819, F2_562 = 20995, F3_072 = 25172;
  int ws[64];
  for (int c = 0; c < 8; ++c) {
    const int16_t* ip = in + c;
    const uint16_t* qp = q + c;
    auto dq = [&](int r) { return (int64_t)ip[8 * r] * qp[8 * r]; };
    int* wp = ws + c;
    if (!ip[8] && !ip[16] && !ip[24] && !ip[32] && !ip[40] && !ip[48] && !ip[56]) {
      const int dc = (int)(dq(0) * (1 << P1));
      for (int r = 0; r < 8; ++r) wp[8 * r] = dc;
      continue;
    }
    int64_t z2 = dq(2), z3 = dq(6);
    int64_t z1 = mul(z2 + z3, F0_541);
    int64_t t2 = z1 + mul(z3, -F1_847), t3 = z1 + mul(z2, F0_765);
    z2 = dq(0);
    z3 = dq(4);
    int64_t t0 = (z2 + z3) * (1 << CB), t1 = (z2 - z3) * (1 << CB);
    const int64_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    t0 = dq(7);
    t1 = dq(5);
    t2 = dq(3);
    t3 = dq(1);
    z1 = t0 + t3;
    z2 = t1 + t2;
    z3 = t0 + t2;
    int64_t z4 = t1 + t3;
    const int64_t z5 = mul(z3 + z4, F1_175);
    t0 = mul(t0, F0_298);
    t1 = mul(t1, F2_053);
    t2 = mul(t2, F3_072);
    t3 = mul(t3, F1_501);
    z1 = mul(z1, -F0_899);
    z2 = mul(z2, -F2_562);
    z3 = mul(z3, -F1_961);
    z4 = mul(z4, -F0_390);
    z3 += z5;
    z4 += z5;
    t0 += z1 + z3;
    t1 += z2 + z4;
    t2 += z2 + z3;
    t3 += z1 + z4;
    wp[0] = (int)descale(t10 + t3, CB - P1);
    wp[56] = (int)descale(t10 - t3, CB - P1);
    wp[8] = (int)descale(t11 + t2, CB - P1);
    wp[48] = (int)descale(t11 - t2, CB - P1);
    wp[16] = (int)descale(t12 + t1, CB - P1);
    wp[40] = (int)descale(t12 - t1, CB - P1);
    wp[24] = (int)descale(t13 + t0, CB - P1);
    wp[32] = (int)descale(t13 - t0, CB - P1);
  }
  for (int r = 0; r < 8; ++r) {
    const int* wp = ws + 8 * r;
    uint8_t* op = out + (size_t)r * stride;
    if (!wp[1] && !wp[2] && !wp[3] && !wp[4] && !wp[5] && !wp[6] && !wp[7]) {
      const uint8_t v = range_idct((int)descale(wp[0], P1 + 3));
      for (int c = 0; c < 8; ++c) op[c] = v;
      continue;
    }
    int64_t z2 = wp[2], z3 = wp[6];
    int64_t z1 = mul(z2 + z3, F0_541);
    int64_t t2 = z1 + mul(z3, -F1_847), t3 = z1 + mul(z2, F0_765);
    int64_t t0 = ((int64_t)wp[0] + wp[4]) * (1 << CB), t1 = ((int64_t)wp[0] - wp[4]) * (1 << CB);
    const int64_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    t0 = wp[7];
    t1 = wp[5];
    t2 = wp[3];
    t3 = wp[1];
    z1 = t0 + t3;
    z2 = t1 + t2;
    z3 = t0 + t2;
    int64_t z4 = t1 + t3;
    const int64_t z5 = mul(z3 + z4, F1_175);
    t0 = mul(t0, F0_298);
    t1 = mul(t1, F2_053);
    t2 = mul(t2, F3_072);
    t3 = mul(t3, F1_501);
    z1 = mul(z1, -F0_899);
    z2 = mul(z2, -F2_562);
    z3 = mul(z3, -F1_961);
    z4 = mul(z4, -F0_390);
    z3 += z5;
    z4 += z5;
    t0 += z1 + z3;
    t1 += z2 + z4;
    t2 += z2 + z3;
    t3 += z1 + z4;
    constexpr int S = CB + P1 + 3;
    op[0] = range_idct((int)descale(t10 + t3, S));
    op[7] = range_idct((int)descale(t10 - t3, S));
    op[1] = range_idct((int)descale(t11 + t2, S));
    op[6] = range_idct((int)descale(t11 - t2, S));
    op[2] = range_idct((int)descale(t12 + t1, S));
    op[5] = range_idct((int)descale(t12 - t1, S));
    op[3] = range_idct((int)descale(t13 + t0, S));
    op[4] = range_idct((int)descale(t13 - t0, S));
  }
}

// Upsampling of one component plane (dw x dh real samples, row stride ps) to
// the full size, by libjpeg's rules: fancy (triangle) filters for 2x1, 1x2
// and 2x2 when the plane is wider than 2 samples, else sample replication.
std::vector<uint8_t> upsample(const std::vector<uint8_t>& p, int ps, int dw, int dh, int hx, int vx, int W, int H) {
  const int ow = dw * hx, oh = dh * vx;
  std::vector<uint8_t> o((size_t)ow * oh);
  auto at = [&](int x, int y) -> int { return p[(size_t)y * ps + x]; };
  if (hx == 1 && vx == 1) {
    for (int y = 0; y < dh; ++y) std::memcpy(&o[(size_t)y * ow], &p[(size_t)y * ps], dw);
  } else if (hx == 2 && vx == 1 && dw > 2) {
    for (int y = 0; y < dh; ++y) {
      uint8_t* op = &o[(size_t)y * ow];
      op[0] = (uint8_t)at(0, y);
      op[1] = (uint8_t)((at(0, y) * 3 + at(1, y) + 2) >> 2);
      for (int x = 1; x < dw - 1; ++x) {
        const int v = at(x, y) * 3;
        op[2 * x] = (uint8_t)((v + at(x - 1, y) + 1) >> 2);
        op[2 * x + 1] = (uint8_t)((v + at(x + 1, y) + 2) >> 2);
      }
      op[2 * dw - 2] = (uint8_t)((at(dw - 1, y) * 3 + at(dw - 2, y) + 1) >> 2);
      op[2 * dw - 1] = (uint8_t)at(dw - 1, y);
    }
  } else if (hx == 1 && vx == 2) {
    for (int y = 0; y < dh; ++y)
      for (int v = 0; v < 2; ++v) {
        const int yn = v == 0 ? (y > 0 ? y - 1 : 0) : (y < dh - 1 ? y + 1 : dh - 1);
        const int bias = v == 0 ? 1 : 2;
        uint8_t* op = &o[(size_t)(2 * y + v) * ow];
        for (int x = 0; x < dw; ++x) op[x] = (uint8_t)((at(x, y) * 3 + at(x, yn) + bias) >> 2);
      }
  } else if (hx == 2 && vx == 2 && dw > 2) {
    for (int y = 0; y < dh; ++y)
      for (int v = 0; v < 2; ++v) {
        const int yn = v == 0 ? (y > 0 ? y - 1 : 0) : (y < dh - 1 ? y + 1 : dh - 1);
        uint8_t* op = &o[(size_t)(2 * y + v) * ow];
        auto col = [&](int x) { return at(x, y) * 3 + at(x, yn); };
        int last = col(0), cur = col(0), next = col(1);
        op[0] = (uint8_t)((cur * 4 + 8) >> 4);
        op[1] = (uint8_t)((cur * 3 + next + 7) >> 4);
        last = cur;
        cur = next;
        for (int x = 1; x < dw - 1; ++x) {
          next = col(x + 1);
          op[2 * x] = (uint8_t)((cur * 3 + last + 8) >> 4);
          op[2 * x + 1] = (uint8_t)((cur * 3 + next + 7) >> 4);
          last = cur;
          cur = next;
        }
        op[2 * dw - 2] = (uint8_t)((cur * 3 + last + 8) >> 4);
        op[2 * dw - 1] = (uint8_t)((cur * 4 + 7) >> 4);
      }
  } else {  // replication
    for (int y = 0; y < oh; ++y)
      for (int x = 0; x < ow; ++x) o[(size_t)y * ow + x] = (uint8_t)at(x / hx, y / vx);
  }
  // crop to W x H
  std::vector<uint8_t> c((size_t)W * H);
  for (int y = 0; y < H; ++y) std::memcpy(&c[(size_t)y * W], &o[(size_t)y * ow], W);
  return c;
}

void Decoder::idct_and_upsample(MapU8& out) {
  std::vector<std::vector<uint8_t>> planes;
  for (auto& c : comp_) {
    const int ps = c.bw * 8;
    std::vector<uint8_t> p((size_t)ps * c.bh * 8);
    for (int by = 0; by < c.bh; ++by)
      for (int bx = 0; bx < c.bw; ++bx)
        idct_islow(&c.coef[((size_t)by * c.bw + bx) * 64], q_[c.tq], &p[(size_t)by * 8 * ps + bx * 8], ps);
    planes.push_back(upsample(p, ps, c.dw, c.dh, hmax_ / c.h, vmax_ / c.v, W_, H_));
  }
  if (comp_.size() == 1) {
    out = MapU8(W_, H_, 1);
    out.data.assign(planes[0].begin(), planes[0].end());
    return;
  }
  // YCbCr -> RGB with 16-bit fixed-point BT.601 tables
  constexpr int SB = 16;
  constexpr int64_t HALF = (int64_t)1 << (SB - 1);
  auto fix = [](double x) { return (int64_t)(x * (1 << SB) + 0.5); };
  int cr_r[256], cb_b[256];
  int64_t cr_g[256], cb_g[256];
  for (int i = 0; i < 256; ++i) {
    const int64_t x = i - 128;
    cr_r[i] = (int)((fix(1.40200) * x + HALF) >> SB);
    cb_b[i] = (int)((fix(1.77200) * x + HALF) >> SB);
    cr_g[i] = -fix(0.71414) * x;
    cb_g[i] = -fix(0.34414) * x + HALF;
  }
  auto clamp = [](int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); };
  out = MapU8(W_, H_, 3);
  const auto &Y = planes[0], &Cb = planes[1], &Cr = planes[2];
  for (size_t i = 0, n = (size_t)W_ * H_; i < n; ++i) {
    const int y = Y[i], cb = Cb[i], cr = Cr[i];
    out.data[3 * i + 0] = clamp(y + cr_r[cr]);
    out.data[3 * i + 1] = clamp(y + (int)((cb_g[cb] + cr_g[cr]) >> SB));
    out.data[3 * i + 2] = clamp(y + cb_b[cb]);
  }
}

bool Decoder::run(MapU8& out, std::string& err) {
  if (u16() != 0xFFD8) { err = "not a JPEG file"; return false; }
  bool scanned = false;
  while (pos_ < n_) {
    int b = byte();
    if (b != 0xFF) continue;  // garbage between segments
    int m = byte();
    while (m == 0xFF) m = byte();
    if (m < 0) break;
    if (m == 0xD9) break;  // EOI
    if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) continue;
    if (m == 0xDA) scanned = true;
    if (!segment(m)) { err = err_; return false; }
  }
  if (!frame_ || !scanned) { err = err_.empty() ? "no frame or no scan" : err_; return false; }
  idct_and_upsample(out);
  return true;
}

}  // namespace

bool DecodeJpeg(const std::vector<uint8_t>& data, MapU8& out, std::string* err) {
  std::string e;
  Decoder d(data.data(), data.size());
  const bool ok = d.run(out, e);
  if (!ok && err) *err = e;
  return ok;
}

bool ReadJpeg(const std::string& path, MapU8& out, std::string* err) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  std::vector<uint8_t> data;
  uint8_t buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), fp)) > 0) data.insert(data.end(), buf, buf + k);
  std::fclose(fp);
  return DecodeJpeg(data, out, err);
}

bool ReadJpegSize(const std::string& path, int& w, int& h) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  std::vector<uint8_t> data(1 << 16);
  data.resize(std::fread(data.data(), 1, data.size(), fp));
  std::fclose(fp);
  std::string e;
  Decoder d(data.data(), data.size());
  return d.size_only(w, h, e);
}

}  // namespace hpmhost
