// 8-bit material texture images for the G-buffer producer: the files
// EightBitTexture::loadTexture (Renderer/textures.cpp:102-127) reads through
// stb_image as RGBA8 and stores with row 0 at the bottom of the picture
// (stbi rows are mirrored there).  Formats decoded here: PNG (every colour
// type and bit depth, palette and tRNS transparency, Adam7 interlacing;
// 16-bit samples keep their high byte, 1/2/4-bit grey is scaled to 0..255),
// TGA (true-colour and grey, raw or RLE), BMP (24/32-bit uncompressed) and
// binary PNM (P5 / P6).  Missing channels follow stb_image's RGBA
// expansion: grey g -> (g, g, g), no alpha -> 255.  JPEG and other formats
// are refused with NWPRT_ERR_UNSUPPORTED.
//
// The PNG decoder carries its own inflate (RFC 1950 / 1951: stored, fixed
// and dynamic Huffman blocks) so the library has no image dependency.
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "nwprt_internal.hpp"

namespace nwprt {
namespace {

// Largest image any reader accepts (the texture limit): checked before a
// pixel buffer is allocated, so a corrupt header cannot ask for gigabytes.
constexpr int64_t kMaxPixels = 1ll << 28;
void check_size(int64_t w, int64_t h, const char *what) {
  require(w > 0 && h > 0 && w * h <= kMaxPixels, NWPRT_ERR_FORMAT,
          std::string(what) + ": bad size (1 to 2^28 pixels)");
}

// ------------------------------------------------------------- inflate --

class BitReader {
 public:
  BitReader(const uint8_t *p, size_t n) : p_(p), n_(n) {}
  int bit() {
    require(pos_ < n_, NWPRT_ERR_FORMAT, "png: compressed data ends early");
    const int b = (p_[pos_] >> bit_) & 1;
    if (++bit_ == 8) {
      bit_ = 0;
      ++pos_;
    }
    return b;
  }
  uint32_t bits(int count) {  // LSB first
    uint32_t v = 0;
    for (int i = 0; i < count; ++i) v |= (uint32_t)bit() << i;
    return v;
  }
  void align() {
    if (bit_) {
      bit_ = 0;
      ++pos_;
    }
  }
  size_t pos() const { return pos_; }
  void skip(size_t k) { pos_ += k; }
  const uint8_t *at() const { return p_ + pos_; }
  size_t left() const { return pos_ <= n_ ? n_ - pos_ : 0; }

 private:
  const uint8_t *p_;
  size_t n_, pos_ = 0;
  int bit_ = 0;
};

// Canonical Huffman code from code lengths: symbols ordered by (length,
// symbol); a code of length L is decoded bit by bit, most significant
// first, against the first code and count of each length.
struct Huffman {
  int count[16] = {};
  std::vector<int> symbol;
  void build(const uint8_t *lengths, int n) {
    std::memset(count, 0, sizeof(count));
    for (int i = 0; i < n; ++i) ++count[lengths[i]];
    count[0] = 0;
    int offs[16] = {}, left = 1;
    for (int len = 1; len < 16; ++len) {
      left = 2 * left - count[len];
      require(left >= 0, NWPRT_ERR_FORMAT, "png: over-subscribed Huffman code");
      offs[len] = len == 1 ? 0 : offs[len - 1] + count[len - 1];
    }
    symbol.assign(n, 0);
    for (int i = 0; i < n; ++i)
      if (lengths[i]) symbol[offs[lengths[i]]++] = i;
  }
  int decode(BitReader &br) const {
    int code = 0, first = 0, index = 0;
    for (int len = 1; len < 16; ++len) {
      code |= br.bit();
      const int c = count[len];
      if (code - first < c) return symbol[index + (code - first)];
      index += c;
      first = (first + c) << 1;
      code <<= 1;
    }
    fail(NWPRT_ERR_FORMAT, "png: invalid Huffman code");
  }
};

constexpr int kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                              31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
constexpr int kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                               2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
constexpr int kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,
                               33,  49,  65,  97,  129, 193,  257,  385,  513,  769,
                               1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
constexpr int kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflate_codes(BitReader &br, const Huffman &lit, const Huffman &dist,
                   std::vector<uint8_t> &out) {
  for (;;) {
    const int s = lit.decode(br);
    if (s < 256) {
      out.push_back((uint8_t)s);
    } else if (s == 256) {
      return;
    } else {
      const int li = s - 257;
      require(li < 29, NWPRT_ERR_FORMAT, "png: invalid length code");
      const int len = kLenBase[li] + (int)br.bits(kLenExtra[li]);
      const int di = dist.decode(br);
      require(di < 30, NWPRT_ERR_FORMAT, "png: invalid distance code");
      const size_t d = (size_t)kDistBase[di] + br.bits(kDistExtra[di]);
      require(d <= out.size(), NWPRT_ERR_FORMAT, "png: distance before the start of the data");
      const size_t from = out.size() - d;
      for (int i = 0; i < len; ++i) out.push_back(out[from + i]);
    }
  }
}

std::vector<uint8_t> zlib_inflate(const std::vector<uint8_t> &z, size_t expect) {
  require(z.size() >= 6, NWPRT_ERR_FORMAT, "png: zlib stream too short");
  require((z[0] & 15) == 8 && ((z[0] << 8) | z[1]) % 31 == 0 && !(z[1] & 0x20),
          NWPRT_ERR_FORMAT, "png: bad zlib header");
  BitReader br(z.data() + 2, z.size() - 2);
  std::vector<uint8_t> out;
  out.reserve(expect);
  int final_block = 0;
  do {
    final_block = br.bit();
    const int type = (int)br.bits(2);
    if (type == 0) {
      br.align();
      require(br.left() >= 4, NWPRT_ERR_FORMAT, "png: stored block header cut");
      const uint8_t *h = br.at();
      const int len = h[0] | (h[1] << 8), nlen = h[2] | (h[3] << 8);
      require((len ^ 0xffff) == nlen, NWPRT_ERR_FORMAT, "png: stored block length check");
      br.skip(4);
      require(br.left() >= (size_t)len, NWPRT_ERR_FORMAT, "png: stored block cut");
      out.insert(out.end(), br.at(), br.at() + len);
      br.skip(len);
    } else if (type == 1) {
      // The fixed codes of RFC 1951 3.2.6 (built once; thread-safe static).
      struct Fixed {
        Huffman lit, dist;
        Fixed() {
          uint8_t l[288], d[30];
          for (int i = 0; i < 288; ++i) l[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
          for (int i = 0; i < 30; ++i) d[i] = 5;
          lit.build(l, 288);
          dist.build(d, 30);
        }
      };
      static const Fixed fixed;
      inflate_codes(br, fixed.lit, fixed.dist, out);
    } else if (type == 2) {
      const int hlit = (int)br.bits(5) + 257, hdist = (int)br.bits(5) + 1;
      const int hclen = (int)br.bits(4) + 4;
      static constexpr int kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5,
                                         11, 4,  12, 3, 13, 2, 14, 1, 15};
      uint8_t cl[19] = {};
      for (int i = 0; i < hclen; ++i) cl[kOrder[i]] = (uint8_t)br.bits(3);
      Huffman clh;
      clh.build(cl, 19);
      uint8_t lens[320] = {};
      int n = 0;
      while (n < hlit + hdist) {
        const int s = clh.decode(br);
        if (s < 16) {
          lens[n++] = (uint8_t)s;
        } else {
          int rep = 0, val = 0;
          if (s == 16) {
            require(n > 0, NWPRT_ERR_FORMAT, "png: repeat with no previous length");
            val = lens[n - 1];
            rep = 3 + (int)br.bits(2);
          } else if (s == 17) {
            rep = 3 + (int)br.bits(3);
          } else {
            rep = 11 + (int)br.bits(7);
          }
          require(n + rep <= hlit + hdist, NWPRT_ERR_FORMAT, "png: too many code lengths");
          while (rep--) lens[n++] = (uint8_t)val;
        }
      }
      require(lens[256] != 0, NWPRT_ERR_FORMAT, "png: no end-of-block code");
      Huffman lit, dist;
      lit.build(lens, hlit);
      dist.build(lens + hlit, hdist);
      inflate_codes(br, lit, dist, out);
    } else {
      fail(NWPRT_ERR_FORMAT, "png: invalid deflate block type");
    }
  } while (!final_block);
  return out;
}

// ----------------------------------------------------------------- PNG --

uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return pa <= pb && pa <= pc ? a : pb <= pc ? b : c;
}

// Undoes the scanline filters of a (sub-)image of w x h pixels in place;
// returns the filtered-away bytes: h rows of `row` bytes.
std::vector<uint8_t> unfilter(const uint8_t *src, size_t avail, int w, int h, int bits_pp,
                              size_t *used) {
  const size_t row = ((size_t)w * bits_pp + 7) / 8;
  const int bpp = std::max(1, bits_pp / 8);
  require(avail >= (row + 1) * h, NWPRT_ERR_FORMAT, "png: image data too short");
  std::vector<uint8_t> out(row * h);
  for (int y = 0; y < h; ++y) {
    const uint8_t f = src[y * (row + 1)];
    const uint8_t *in = src + y * (row + 1) + 1;
    uint8_t *cur = out.data() + y * row;
    const uint8_t *prev = y ? cur - row : nullptr;
    for (size_t i = 0; i < row; ++i) {
      const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0;
      const int c = prev && i >= (size_t)bpp ? prev[i - bpp] : 0;
      int v = in[i];
      switch (f) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: fail(NWPRT_ERR_FORMAT, "png: invalid filter type");
      }
      cur[i] = (uint8_t)v;
    }
  }
  *used = (row + 1) * h;
  return out;
}

Image read_png(const std::vector<uint8_t> &f) {
  require(f.size() >= 8 + 25 && std::memcmp(f.data(), "\x89PNG\r\n\x1a\n", 8) == 0,
          NWPRT_ERR_FORMAT, "png: bad signature");
  size_t p = 8;
  int w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  bool have_ihdr = false;
  while (p + 12 <= f.size()) {
    const uint32_t len = be32(&f[p]);
    require(p + 12 + (size_t)len <= f.size(), NWPRT_ERR_FORMAT, "png: chunk runs past the file");
    const uint8_t *t = &f[p + 4], *d = &f[p + 8];
    if (!std::memcmp(t, "IHDR", 4)) {
      require(len == 13, NWPRT_ERR_FORMAT, "png: bad IHDR");
      w = (int)be32(d);
      h = (int)be32(d + 4);
      depth = d[8];
      ctype = d[9];
      require(d[10] == 0 && d[11] == 0, NWPRT_ERR_FORMAT, "png: unknown compression or filter");
      interlace = d[12];
      have_ihdr = true;
    } else if (!std::memcmp(t, "PLTE", 4)) {
      plte.assign(d, d + len);
    } else if (!std::memcmp(t, "tRNS", 4)) {
      trns.assign(d, d + len);
    } else if (!std::memcmp(t, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!std::memcmp(t, "IEND", 4)) {
      break;
    }
    p += 12 + (size_t)len;
  }
  require(have_ihdr, NWPRT_ERR_FORMAT, "png: missing IHDR");
  check_size(w, h, "png");
  int channels = 0;
  switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: fail(NWPRT_ERR_FORMAT, "png: invalid colour type");
  }
  const bool depth_ok = ctype == 0   ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 ||
                                        depth == 16)
                        : ctype == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                     : (depth == 8 || depth == 16);
  require(depth_ok && interlace <= 1, NWPRT_ERR_FORMAT, "png: invalid bit depth or interlace");
  require(ctype != 3 || (!plte.empty() && plte.size() % 3 == 0), NWPRT_ERR_FORMAT,
          "png: palette image without PLTE");
  const int bits_pp = channels * depth;
  const std::vector<uint8_t> raw = zlib_inflate(idat, (size_t)h * (((size_t)w * bits_pp + 7) / 8 + 1));

  Image img;
  img.width = w;
  img.height = h;
  img.rgba.assign((size_t)w * h * 4, 0);
  // Sample s of pixel x in a row of unfiltered bytes.
  auto sample = [&](const uint8_t *row, int x, int s) -> int {
    if (depth == 8) return row[(size_t)x * channels + s];
    if (depth == 16) return row[((size_t)x * channels + s) * 2];  // high byte
    const int per_byte = 8 / depth, idx = x;  // one channel at depths < 8
    const int shift = 8 - depth * (idx % per_byte + 1);
    return (row[idx / per_byte] >> shift) & ((1 << depth) - 1);
  };
  auto raw16 = [&](const uint8_t *row, int x, int s) -> int {
    return (row[((size_t)x * channels + s) * 2] << 8) | row[((size_t)x * channels + s) * 2 + 1];
  };
  auto put = [&](const uint8_t *row, int x, int X, int Y) {
    uint8_t *o = &img.rgba[(((size_t)(h - 1 - Y)) * w + X) * 4];  // row 0 = bottom
    int r = 0, g = 0, b = 0, a = 255;
    if (ctype == 3) {
      const int i = sample(row, x, 0);
      require((size_t)i * 3 + 2 < plte.size(), NWPRT_ERR_FORMAT, "png: palette index out of range");
      r = plte[i * 3];
      g = plte[i * 3 + 1];
      b = plte[i * 3 + 2];
      if ((size_t)i < trns.size()) a = trns[i];
    } else if (ctype == 0 || ctype == 4) {
      int v = sample(row, x, 0);
      const int raw_v = depth == 16 ? raw16(row, x, 0) : v;
      if (depth < 8) v = v * 255 / ((1 << depth) - 1);
      r = g = b = v;
      if (ctype == 4) a = sample(row, x, 1);
      else if (trns.size() >= 2 && raw_v == ((trns[0] << 8) | trns[1])) a = 0;
    } else {
      r = sample(row, x, 0);
      g = sample(row, x, 1);
      b = sample(row, x, 2);
      if (ctype == 6) {
        a = sample(row, x, 3);
      } else if (trns.size() >= 6) {
        const bool key = depth == 16
            ? raw16(row, x, 0) == ((trns[0] << 8) | trns[1]) &&
              raw16(row, x, 1) == ((trns[2] << 8) | trns[3]) &&
              raw16(row, x, 2) == ((trns[4] << 8) | trns[5])
            : r == trns[1] && g == trns[3] && b == trns[5];
        if (key) a = 0;
      }
    }
    o[0] = (uint8_t)r;
    o[1] = (uint8_t)g;
    o[2] = (uint8_t)b;
    o[3] = (uint8_t)a;
  };
  if (!interlace) {
    size_t used = 0;
    const std::vector<uint8_t> px = unfilter(raw.data(), raw.size(), w, h, bits_pp, &used);
    const size_t row = ((size_t)w * bits_pp + 7) / 8;
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) put(px.data() + y * row, x, x, y);
  } else {
    // Adam7: seven sub-images, each filtered on its own.
    static constexpr int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static constexpr int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    size_t off = 0;
    for (int pass = 0; pass < 7; ++pass) {
      const int pw = (w - x0[pass] + dx[pass] - 1) / dx[pass];
      const int ph = (h - y0[pass] + dy[pass] - 1) / dy[pass];
      if (pw <= 0 || ph <= 0) continue;
      size_t used = 0;
      require(off <= raw.size(), NWPRT_ERR_FORMAT, "png: interlaced data too short");
      const std::vector<uint8_t> px =
          unfilter(raw.data() + off, raw.size() - off, pw, ph, bits_pp, &used);
      const size_t row = ((size_t)pw * bits_pp + 7) / 8;
      for (int y = 0; y < ph; ++y)
        for (int x = 0; x < pw; ++x)
          put(px.data() + y * row, x, x0[pass] + x * dx[pass], y0[pass] + y * dy[pass]);
      off += used;
    }
  }
  return img;
}

// ----------------------------------------------------------------- TGA --

Image read_tga(const std::vector<uint8_t> &f) {
  require(f.size() >= 18, NWPRT_ERR_FORMAT, "tga: file too short");
  const int id_len = f[0], cmap_type = f[1], type = f[2];
  const int cmap_len = f[5] | (f[6] << 8), cmap_bits = f[7];
  const int w = f[12] | (f[13] << 8), h = f[14] | (f[15] << 8), bits = f[16], desc = f[17];
  require(cmap_type == 0 && (type == 2 || type == 3 || type == 10 || type == 11),
          NWPRT_ERR_UNSUPPORTED, "tga: only true-colour or grey images, raw or RLE");
  const bool grey = type == 3 || type == 11;
  require(grey ? bits == 8 : (bits == 24 || bits == 32), NWPRT_ERR_UNSUPPORTED,
          "tga: 8-bit grey or 24/32-bit colour only");
  check_size(w, h, "tga");
  size_t p = 18 + id_len + (size_t)cmap_len * ((cmap_bits + 7) / 8);
  const int bpp = bits / 8;
  std::vector<uint8_t> px((size_t)w * h * bpp);
  if (type == 2 || type == 3) {
    require(p + px.size() <= f.size(), NWPRT_ERR_FORMAT, "tga: pixel data too short");
    std::memcpy(px.data(), &f[p], px.size());
  } else {
    size_t o = 0;
    while (o < px.size()) {
      require(p < f.size(), NWPRT_ERR_FORMAT, "tga: RLE data too short");
      const int hdr = f[p++], n = (hdr & 127) + 1;
      require(o + (size_t)n * bpp <= px.size(), NWPRT_ERR_FORMAT, "tga: RLE run past the image");
      if (hdr & 128) {
        require(p + bpp <= f.size(), NWPRT_ERR_FORMAT, "tga: RLE data too short");
        for (int i = 0; i < n; ++i) std::memcpy(&px[o + (size_t)i * bpp], &f[p], bpp);
        p += bpp;
      } else {
        require(p + (size_t)n * bpp <= f.size(), NWPRT_ERR_FORMAT, "tga: RLE data too short");
        std::memcpy(&px[o], &f[p], (size_t)n * bpp);
        p += (size_t)n * bpp;
      }
      o += (size_t)n * bpp;
    }
  }
  Image img;
  img.width = w;
  img.height = h;
  img.rgba.resize((size_t)w * h * 4);
  const bool top_down = (desc & 0x20) != 0, right_left = (desc & 0x10) != 0;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const uint8_t *s = &px[((size_t)y * w + x) * bpp];
      const int Y = top_down ? h - 1 - y : y, X = right_left ? w - 1 - x : x;
      uint8_t *o = &img.rgba[((size_t)Y * w + X) * 4];
      if (grey) {
        o[0] = o[1] = o[2] = s[0];
        o[3] = 255;
      } else {
        o[0] = s[2];
        o[1] = s[1];
        o[2] = s[0];
        o[3] = bpp == 4 ? s[3] : 255;
      }
    }
  return img;
}

// ----------------------------------------------------------------- BMP --

Image read_bmp(const std::vector<uint8_t> &f) {
  auto le32 = [&](size_t p) { return (uint32_t)f[p] | (f[p + 1] << 8) | (f[p + 2] << 16) | ((uint32_t)f[p + 3] << 24); };
  require(f.size() >= 54 && f[0] == 'B' && f[1] == 'M', NWPRT_ERR_FORMAT, "bmp: bad header");
  const size_t off = le32(10);
  const int64_t w = (int32_t)le32(18), hs = (int32_t)le32(22);
  const int bits = f[28] | (f[29] << 8);
  const uint32_t comp = le32(30);
  require((bits == 24 || bits == 32) && (comp == 0 || (comp == 3 && bits == 32)),
          NWPRT_ERR_UNSUPPORTED, "bmp: 24/32-bit uncompressed only");
  const int64_t h = hs < 0 ? -hs : hs;
  check_size(w, h, "bmp");
  const size_t stride = (((size_t)w * bits / 8) + 3) & ~(size_t)3;
  require(off + stride * h <= f.size(), NWPRT_ERR_FORMAT, "bmp: pixel data too short");
  Image img;
  img.width = (int)w;
  img.height = (int)h;
  img.rgba.resize((size_t)w * h * 4);
  for (int y = 0; y < h; ++y) {
    const int Y = hs > 0 ? y : (int)h - 1 - y;  // positive height: stored bottom-up
    for (int x = 0; x < w; ++x) {
      const uint8_t *s = &f[off + (size_t)y * stride + (size_t)x * (bits / 8)];
      uint8_t *o = &img.rgba[((size_t)Y * w + x) * 4];
      o[0] = s[2];
      o[1] = s[1];
      o[2] = s[0];
      o[3] = 255;  // stb_image ignores BMP alpha unless a mask says otherwise
    }
  }
  return img;
}

// ----------------------------------------------------------------- PNM --

Image read_pnm(const std::vector<uint8_t> &f) {
  size_t p = 2;
  auto token = [&]() -> int {
    for (;;) {
      require(p < f.size(), NWPRT_ERR_FORMAT, "pnm: header cut");
      if (f[p] == '#') {
        while (p < f.size() && f[p] != '\n') ++p;
      } else if (std::isspace(f[p])) {
        ++p;
      } else {
        break;
      }
    }
    long v = 0;
    require(std::isdigit(f[p]), NWPRT_ERR_FORMAT, "pnm: bad header number");
    while (p < f.size() && std::isdigit(f[p])) {
      v = v * 10 + (f[p++] - '0');
      require(v <= (1l << 30), NWPRT_ERR_FORMAT, "pnm: header number too large");
    }
    return (int)v;
  };
  const bool colour = f[1] == '6';
  const int w = token(), h = token(), maxv = token();
  ++p;  // the single whitespace byte before the raster
  check_size(w, h, "pnm");
  require(maxv > 0 && maxv < 65536, NWPRT_ERR_FORMAT, "pnm: bad maxval");
  const int ch = colour ? 3 : 1, bps = maxv > 255 ? 2 : 1;
  require(p + (size_t)w * h * ch * bps <= f.size(), NWPRT_ERR_FORMAT, "pnm: raster too short");
  Image img;
  img.width = w;
  img.height = h;
  img.rgba.resize((size_t)w * h * 4);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      uint8_t *o = &img.rgba[((size_t)(h - 1 - y) * w + x) * 4];
      for (int c = 0; c < 3; ++c) {
        const size_t s = p + (((size_t)y * w + x) * ch + (colour ? c : 0)) * bps;
        const int v = bps == 2 ? (f[s] << 8) | f[s + 1] : f[s];
        o[c] = (uint8_t)(maxv == 255 ? v : (v * 255 + maxv / 2) / maxv);
      }
      o[3] = 255;
    }
  return img;
}

}  // namespace

Image read_image_rgba8(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  require(in.good(), NWPRT_ERR_IO, "image: cannot open " + path);
  std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  require(f.size() >= 4, NWPRT_ERR_FORMAT, "image: file too short: " + path);
  if (f[0] == 0x89 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return read_png(f);
  if (f[0] == 'B' && f[1] == 'M') return read_bmp(f);
  if (f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return read_pnm(f);
  if (f[0] == 0xFF && f[1] == 0xD8) fail(NWPRT_ERR_UNSUPPORTED, "image: JPEG is not supported: " + path);
  const std::string ext = path.size() >= 4 ? path.substr(path.size() - 4) : "";
  if (ext == ".tga" || ext == ".TGA") return read_tga(f);
  fail(NWPRT_ERR_UNSUPPORTED, "image: unknown format: " + path);
}

}  // namespace nwprt
