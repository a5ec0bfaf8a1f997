// On-disk formats of the relighting path: .npy arrays (weights, wavelet
// selections), Radiance .hdr cubemap crosses, and config_comp.json.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "nwprt_internal.hpp"

namespace nwprt {

// ------------------------------------------------------------------- npy --
// Format: "\x93NUMPY", major, minor, header length (u16 for v1, u32 for
// v2/v3), then a Python dict literal with 'descr', 'fortran_order', 'shape'.
// The reference reads these with Renderer/npy.h (LoadArrayFromNumpy) and uses
// the payload flat, in file order (Renderer/tinycudann/tcnn.cu:30-38,
// Renderer/EnvironmentMap.cpp:56-67); so does this reader.

size_t NpyArray::count() const {
  size_t n = 1;
  for (int64_t s : shape) n *= (size_t)s;
  return n;
}

template <typename T>
static T load_le(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

std::vector<float> NpyArray::as_f32() const {
  const size_t n = count();
  std::vector<float> out(n);
  const uint8_t *p = raw.data();
  for (size_t i = 0; i < n; ++i, p += itemsize) {
    double v = 0;
    if (kind == 'f' && itemsize == 4) v = load_le<float>(p);
    else if (kind == 'f' && itemsize == 8) v = load_le<double>(p);
    else if (kind == 'f' && itemsize == 2) v = f16_bits_to_f32(load_le<uint16_t>(p));
    else if (kind == 'i' && itemsize == 4) v = load_le<int32_t>(p);
    else if (kind == 'i' && itemsize == 8) v = (double)load_le<int64_t>(p);
    else if ((kind == 'u' || kind == 'b') && itemsize == 1) v = *p;
    else fail(NWPRT_ERR_FORMAT, "npy: unsupported dtype for float conversion");
    out[i] = (float)v;
  }
  return out;
}

std::vector<int32_t> NpyArray::as_i32() const {
  const size_t n = count();
  std::vector<int32_t> out(n);
  const uint8_t *p = raw.data();
  for (size_t i = 0; i < n; ++i, p += itemsize) {
    int64_t v = 0;
    if (kind == 'i' && itemsize == 4) v = load_le<int32_t>(p);
    else if (kind == 'i' && itemsize == 8) v = load_le<int64_t>(p);
    else if ((kind == 'u' || kind == 'b') && itemsize == 1) v = *p;
    else if (kind == 'u' && itemsize == 4) v = load_le<uint32_t>(p);
    else fail(NWPRT_ERR_FORMAT, "npy: unsupported dtype for int32 conversion");
    require(v >= INT32_MIN && v <= INT32_MAX, NWPRT_ERR_FORMAT,
            "npy: integer out of int32 range");
    out[i] = (int32_t)v;
  }
  return out;
}

static std::string dict_value(const std::string &hdr, const std::string &key) {
  const std::string k1 = "'" + key + "'";
  size_t pos = hdr.find(k1);
  if (pos == std::string::npos) fail(NWPRT_ERR_FORMAT, "npy: header lacks " + key);
  pos = hdr.find(':', pos + k1.size());
  if (pos == std::string::npos) fail(NWPRT_ERR_FORMAT, "npy: malformed header");
  ++pos;
  while (pos < hdr.size() && std::isspace((unsigned char)hdr[pos])) ++pos;
  size_t end = pos;
  if (hdr[pos] == '(') {
    end = hdr.find(')', pos);
    if (end == std::string::npos) fail(NWPRT_ERR_FORMAT, "npy: malformed shape");
    return hdr.substr(pos, end - pos + 1);
  }
  if (hdr[pos] == '\'') {
    end = hdr.find('\'', pos + 1);
    if (end == std::string::npos) fail(NWPRT_ERR_FORMAT, "npy: malformed descr");
    return hdr.substr(pos + 1, end - pos - 1);
  }
  while (end < hdr.size() && hdr[end] != ',' && hdr[end] != '}') ++end;
  std::string v = hdr.substr(pos, end - pos);
  while (!v.empty() && std::isspace((unsigned char)v.back())) v.pop_back();
  return v;
}

NpyArray npy_read(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(NWPRT_ERR_IO, "npy: cannot open " + path);
  char magic[8];
  f.read(magic, 8);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0)
    fail(NWPRT_ERR_FORMAT, "npy: bad magic in " + path);
  const int major = (unsigned char)magic[6];
  uint32_t hlen = 0;
  if (major == 1) {
    uint8_t b[2];
    f.read((char *)b, 2);
    hlen = b[0] | (b[1] << 8);
  } else if (major == 2 || major == 3) {
    uint8_t b[4];
    f.read((char *)b, 4);
    hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
  } else {
    fail(NWPRT_ERR_FORMAT, "npy: unsupported version in " + path);
  }
  if (!f || hlen > (1u << 20)) fail(NWPRT_ERR_FORMAT, "npy: bad header length");
  std::string hdr(hlen, '\0');
  f.read(&hdr[0], hlen);
  if (!f) fail(NWPRT_ERR_FORMAT, "npy: truncated header in " + path);

  NpyArray a;
  const std::string descr = dict_value(hdr, "descr");
  if (descr.size() < 3) fail(NWPRT_ERR_FORMAT, "npy: bad descr " + descr);
  const char order = descr[0];
  a.kind = descr[1];
  a.itemsize = std::atoi(descr.c_str() + 2);
  if (!(order == '<' || order == '|' || (order == '=' )))
    fail(NWPRT_ERR_UNSUPPORTED, "npy: big-endian data not supported: " + descr);
  if (!(a.itemsize == 1 || a.itemsize == 2 || a.itemsize == 4 || a.itemsize == 8))
    fail(NWPRT_ERR_FORMAT, "npy: bad item size " + descr);
  const std::string fo = dict_value(hdr, "fortran_order");
  if (fo == "True") a.fortran_order = true;
  else if (fo == "False") a.fortran_order = false;
  else fail(NWPRT_ERR_FORMAT, "npy: bad fortran_order");
  const std::string sh = dict_value(hdr, "shape");
  for (size_t i = 1; i < sh.size();) {
    while (i < sh.size() && !std::isdigit((unsigned char)sh[i])) {
      if (sh[i] == '-') fail(NWPRT_ERR_FORMAT, "npy: negative dimension");
      ++i;
    }
    if (i >= sh.size()) break;
    int64_t v = 0;
    while (i < sh.size() && std::isdigit((unsigned char)sh[i])) {
      v = v * 10 + (sh[i] - '0');
      if (v > (1ll << 40)) fail(NWPRT_ERR_FORMAT, "npy: dimension too large");
      ++i;
    }
    a.shape.push_back(v);
  }
  const size_t bytes = a.count() * (size_t)a.itemsize;
  a.raw.resize(bytes);
  if (bytes) f.read((char *)a.raw.data(), (std::streamsize)bytes);
  if (!f) fail(NWPRT_ERR_FORMAT, "npy: truncated payload in " + path);
  return a;
}

void npy_write(const std::string &path, const void *data, char kind,
               int itemsize, const std::vector<int64_t> &shape) {
  std::ostringstream h;
  h << "{'descr': '<" << kind << itemsize << "', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t i = 0; i < shape.size(); ++i) {
    h << shape[i] << (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    n *= (size_t)shape[i];
  }
  h << "), }";
  std::string hdr = h.str();
  const size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr.push_back('\n');
  std::ofstream f(path, std::ios::binary);
  if (!f) fail(NWPRT_ERR_IO, "npy: cannot write " + path);
  f.write("\x93NUMPY\x01\x00", 8);
  const uint16_t hl = (uint16_t)hdr.size();
  const uint8_t hb[2] = {(uint8_t)(hl & 0xff), (uint8_t)(hl >> 8)};
  f.write((const char *)hb, 2);
  f.write(hdr.data(), (std::streamsize)hdr.size());
  f.write((const char *)data, (std::streamsize)(n * (size_t)itemsize));
  if (!f) fail(NWPRT_ERR_IO, "npy: write failed for " + path);
}

// ------------------------------------------------------------------- hdr --
// Radiance RGBE.  Decoding follows the reference's arithmetic
// (Renderer/HDRloader.cpp:101-118): value = byte/256 * 2^(E-128), no
// half-unit offset; scanlines kept in file order (top first).  Scanlines may
// be flat, old-style run-length (1,1,1,count) or new-style per-component RLE.

namespace {

struct Reader {
  const std::vector<uint8_t> &b;
  size_t pos = 0;
  explicit Reader(const std::vector<uint8_t> &buf) : b(buf) {}
  int get() { return pos < b.size() ? b[pos++] : -1; }
  bool eof() const { return pos >= b.size(); }
};

void read_flat_or_old(Reader &r, uint8_t *scan, int w, int first[4], bool have_first) {
  int i = 0, shift = 0;
  auto put = [&](const int px[4]) {
    for (int c = 0; c < 4; ++c) scan[4 * i + c] = (uint8_t)px[c];
    ++i;
  };
  if (have_first) {
    put(first);
    shift = 0;
  }
  while (i < w) {
    int px[4];
    for (int c = 0; c < 4; ++c) {
      px[c] = r.get();
      if (px[c] < 0) fail(NWPRT_ERR_FORMAT, "hdr: truncated scanline");
    }
    if (px[0] == 1 && px[1] == 1 && px[2] == 1) {  // old-style run
      if (i == 0) fail(NWPRT_ERR_FORMAT, "hdr: run without previous pixel");
      const int count = px[3] << shift;
      if (i + count > w) fail(NWPRT_ERR_FORMAT, "hdr: run overflows scanline");
      for (int k = 0; k < count; ++k) {
        std::memcpy(scan + 4 * i, scan + 4 * (i - 1), 4);
        ++i;
      }
      shift += 8;
    } else {
      put(px);
      shift = 0;
    }
  }
}

void read_scanline(Reader &r, uint8_t *scan, int w) {
  int first[4];
  for (int c = 0; c < 4; ++c) {
    first[c] = r.get();
    if (first[c] < 0) fail(NWPRT_ERR_FORMAT, "hdr: truncated scanline");
  }
  const bool new_rle = w >= 8 && w < 0x8000 && first[0] == 2 && first[1] == 2 &&
                       (first[2] & 0x80) == 0;
  if (!new_rle) {
    read_flat_or_old(r, scan, w, first, true);
    return;
  }
  if (((first[2] << 8) | first[3]) != w)
    fail(NWPRT_ERR_FORMAT, "hdr: scanline width mismatch");
  for (int c = 0; c < 4; ++c) {
    int i = 0;
    while (i < w) {
      int code = r.get();
      if (code < 0) fail(NWPRT_ERR_FORMAT, "hdr: truncated RLE scanline");
      if (code > 128) {
        const int count = code - 128;
        const int v = r.get();
        if (v < 0 || i + count > w) fail(NWPRT_ERR_FORMAT, "hdr: bad RLE run");
        for (int k = 0; k < count; ++k) scan[4 * (i++) + c] = (uint8_t)v;
      } else {
        if (code == 0 || i + code > w) fail(NWPRT_ERR_FORMAT, "hdr: bad RLE dump");
        for (int k = 0; k < code; ++k) {
          const int v = r.get();
          if (v < 0) fail(NWPRT_ERR_FORMAT, "hdr: truncated RLE dump");
          scan[4 * (i++) + c] = (uint8_t)v;
        }
      }
    }
  }
}

}  // namespace

HdrImage hdr_read(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) fail(NWPRT_ERR_IO, "hdr: cannot open " + path);
  std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)),
                           std::istreambuf_iterator<char>());
  Reader r(buf);
  auto line = [&]() {
    std::string s;
    int c;
    while ((c = r.get()) >= 0 && c != '\n') s.push_back((char)c);
    if (c < 0) fail(NWPRT_ERR_FORMAT, "hdr: truncated header in " + path);
    return s;
  };
  const std::string magic = line();
  if (magic.rfind("#?RADIANCE", 0) != 0 && magic.rfind("#?RGBE", 0) != 0)
    fail(NWPRT_ERR_FORMAT, "hdr: not a Radiance file: " + path);
  for (int guard = 0;; ++guard) {
    if (guard > 4096) fail(NWPRT_ERR_FORMAT, "hdr: header too long");
    const std::string s = line();
    if (s.empty()) break;
    if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe")
      fail(NWPRT_ERR_UNSUPPORTED, "hdr: unsupported " + s);
  }
  const std::string reso = line();
  int w = 0, h = 0;
  if (std::sscanf(reso.c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0 ||
      (long long)w * h > (1ll << 28))
    fail(NWPRT_ERR_UNSUPPORTED, "hdr: unsupported resolution line '" + reso + "'");
  HdrImage img;
  img.width = w;
  img.height = h;
  img.rgb.resize((size_t)w * h * 3);
  std::vector<uint8_t> scan((size_t)w * 4);
  for (int y = 0; y < h; ++y) {
    read_scanline(r, scan.data(), w);
    float *row = img.rgb.data() + (size_t)y * w * 3;
    for (int x = 0; x < w; ++x) {
      const int e = (int)scan[4 * x + 3] - 128;
      for (int c = 0; c < 3; ++c)
        row[3 * x + c] = std::ldexp((float)scan[4 * x + c] / 256.0f, e);
    }
  }
  return img;
}

void hdr_write(const std::string &path, const float *data, int width,
               int height, int channels) {
  require(width > 0 && height > 0 && (channels == 3 || channels == 4),
          NWPRT_ERR_INVALID_ARGUMENT, "hdr: bad image shape");
  std::ofstream f(path, std::ios::binary);
  if (!f) fail(NWPRT_ERR_IO, "hdr: cannot write " + path);
  f << "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " << height << " +X " << width << "\n";
  std::vector<uint8_t> scan((size_t)width * 4);
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      const float *px = data + ((size_t)y * width + x) * channels;
      const float r = std::fmax(px[0], 0.f), g = std::fmax(px[1], 0.f),
                  b = std::fmax(px[2], 0.f);
      const float v = std::fmax(r, std::fmax(g, b));
      uint8_t *o = &scan[4 * (size_t)x];
      if (!(v >= 1e-32f) || !std::isfinite(v)) {
        o[0] = o[1] = o[2] = o[3] = 0;
        continue;
      }
      int e;
      const float m = std::frexp(v, &e);  // v = m * 2^e, m in [0.5, 1)
      const float s = m * 256.0f / v;
      o[0] = (uint8_t)(r * s);
      o[1] = (uint8_t)(g * s);
      o[2] = (uint8_t)(b * s);
      o[3] = (uint8_t)(e + 128);
    }
    f.write((const char *)scan.data(), (std::streamsize)scan.size());
  }
  if (!f) fail(NWPRT_ERR_IO, "hdr: write failed for " + path);
}

// 4N x 3N horizontal cross (rows in file order, top first) to six faces
// [+X, -X, +Y, -Y, -Z, +Z], each N x N with rows bottom-up.  The reference
// flips rows on load (Renderer/HDRloader.h:21-25), then cuts (col, row)
// blocks (2,1) (0,1) (1,2) (1,0) (1,1) (3,1) counted from the bottom
// (Renderer/EnvironmentMap.cpp:112-126, Renderer/textures.cpp:72-83).
void cross_to_faces(const float *cross, int width, int height, float *faces) {
  require(height % 3 == 0 && width == 4 * (height / 3) && height > 0,
          NWPRT_ERR_FORMAT, "cubemap: cross must be 4N x 3N");
  const int n = height / 3;
  static const int block[6][2] = {{2, 1}, {0, 1}, {1, 2}, {1, 0}, {1, 1}, {3, 1}};
  for (int f = 0; f < 6; ++f) {
    const int c0 = block[f][0] * n, r0 = block[f][1] * n;
    for (int i = 0; i < n; ++i) {
      const int file_row = height - 1 - (r0 + i);
      for (int j = 0; j < n; ++j)
        for (int c = 0; c < 3; ++c)
          faces[(((size_t)f * n + i) * n + j) * 3 + c] =
              cross[((size_t)file_row * width + c0 + j) * 3 + c];
    }
  }
}

// ------------------------------------------------------------------ json --

const Json *Json::get(const std::string &key) const {
  if (type != Object) return nullptr;
  for (auto &kv : obj)
    if (kv.first == key) return &kv.second;
  return nullptr;
}
double Json::number_or(const std::string &key, double dflt) const {
  const Json *v = get(key);
  return (v && v->type == Number) ? v->num : dflt;
}
std::string Json::string_or(const std::string &key, const std::string &dflt) const {
  const Json *v = get(key);
  return (v && v->type == String) ? v->str : dflt;
}

namespace {
struct JsonParser {
  const std::string &s;
  size_t i = 0;
  int depth = 0;
  explicit JsonParser(const std::string &t) : s(t) {}
  [[noreturn]] void bad(const char *what) {
    fail(NWPRT_ERR_FORMAT, std::string("json: ") + what + " at offset " + std::to_string(i));
  }
  void ws() {
    while (i < s.size()) {
      if (std::isspace((unsigned char)s[i])) ++i;
      else if (s.compare(i, 2, "//") == 0) { while (i < s.size() && s[i] != '\n') ++i; }
      else break;
    }
  }
  std::string str() {
    if (s[i] != '"') bad("expected string");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\') {
        if (++i >= s.size()) bad("bad escape");
        const char e = s[i];
        out.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
      } else {
        out.push_back(s[i]);
      }
      ++i;
    }
    if (i >= s.size()) bad("unterminated string");
    ++i;
    return out;
  }
  Json value() {
    if (++depth > 64) bad("nesting too deep");
    ws();
    if (i >= s.size()) bad("unexpected end");
    Json v;
    const char c = s[i];
    if (c == '{') {
      v.type = Json::Object;
      ++i;
      ws();
      if (s[i] == '}') { ++i; --depth; return v; }
      for (;;) {
        ws();
        std::string k = str();
        ws();
        if (s[i] != ':') bad("expected ':'");
        ++i;
        Json child = value();
        v.obj.emplace_back(std::move(k), std::move(child));
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == '}') { ++i; break; }
        bad("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.type = Json::Array;
      ++i;
      ws();
      if (s[i] == ']') { ++i; --depth; return v; }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (s[i] == ',') { ++i; continue; }
        if (s[i] == ']') { ++i; break; }
        bad("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.type = Json::String;
      v.str = str();
    } else if (s.compare(i, 4, "true") == 0) {
      v.type = Json::Bool; v.b = true; i += 4;
    } else if (s.compare(i, 5, "false") == 0) {
      v.type = Json::Bool; v.b = false; i += 5;
    } else if (s.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      char *end = nullptr;
      v.num = std::strtod(s.c_str() + i, &end);
      if (end == s.c_str() + i) bad("unexpected character");
      v.type = Json::Number;
      i = (size_t)(end - s.c_str());
    }
    --depth;
    return v;
  }
};
}  // namespace

Json json_parse(const std::string &text) {
  JsonParser p(text);
  Json v = p.value();
  p.ws();
  if (p.i != text.size()) p.bad("trailing characters");
  return v;
}

}  // namespace nwprt
