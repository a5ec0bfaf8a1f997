// Seeded synthetic inputs standing in for the OptiX G-buffer pass, the
// trained network and the environment maps (none ship with the reference).
//
// Every value is drawn from a counter-based stream: u64 = mix(seed, stream,
// index) with SplitMix64's finaliser, so any element can be generated
// independently and in parallel, and the same (seed, stream, index) always
// yields the same bits.  Uniforms take the top 53 bits; normals use
// Box-Muller in double on the uniform pair (2i, 2i+1).
#include <algorithm>
#include <cmath>
#include <thread>

#include "nwprt_internal.hpp"

namespace nwprt {
namespace {

// Static split of [0, n) over up to 16 host threads.
template <class F>
void parallel_for(long long n, F &&fn) {
  const long long min_chunk = 1 << 16;
  int threads = (int)std::min<long long>(16, std::max<long long>(1, n / min_chunk));
  threads = std::min<int>(threads, std::max(1u, std::thread::hardware_concurrency()));
  if (threads <= 1) {
    for (long long i = 0; i < n; ++i) fn(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const long long b = n * t / threads, e = n * (t + 1) / threads;
    pool.emplace_back([&fn, b, e] {
      for (long long i = b; i < e; ++i) fn(i);
    });
  }
  for (auto &th : pool) th.join();
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t key;
  Stream(uint64_t seed, uint64_t stream) : key(mix64(seed + kGolden * (stream + 1))) {}
  uint64_t bits(uint64_t i) const { return mix64(key + kGolden * (i + 1)); }
  double uniform(uint64_t i) const {  // [0, 1)
    return (double)(bits(i) >> 11) * (1.0 / 9007199254740992.0);
  }
  double normal(uint64_t i) const {
    const double u1 = uniform(2 * i), u2 = uniform(2 * i + 1);
    const double r = std::sqrt(-2.0 * std::log(1.0 - u1));
    return r * std::cos(6.283185307179586 * u2);
  }
};

// Stream ids (fixed; part of the generator's definition).
enum : uint64_t {
  kPos = 1, kNormalView = 2, kKd = 3, kKs = 4, kRough = 5, kDirect = 6,
  kFlag = 7, kCubeTexel = 10, kCubeSun = 11, kHash = 20, kWaveletFeat = 21,
  kProj = 22, kLayer0 = 23
};

}  // namespace

void synth_scene(const nwprt_scene_desc &d, uint64_t seed, float hash_std,
                 float feature_std, void *hash_table, float *wfeat,
                 float *proj, float *layers) {
  validate_scene_desc(d);
  const size_t F = (size_t)d.grid.n_features_per_level;
  const size_t entries = scene_hash_entries(d);
  if (hash_table) {
    Stream s(seed, kHash);
    const size_t n = entries * F;
    parallel_for((long long)n, [&](long long i) {
      const float v = (float)(hash_std * s.normal((uint64_t)i));
      if (d.hash_dtype == NWPRT_DTYPE_F16)
        ((uint16_t *)hash_table)[i] = f32_to_f16_bits(v);
      else
        ((float *)hash_table)[i] = v;
    });
  }
  const size_t coefs = 6ull * d.face_size * d.face_size;
  if (wfeat) {
    Stream s(seed, kWaveletFeat);
    const size_t n = (size_t)d.feature_dim * coefs;
    parallel_for((long long)n, [&](long long i) {
      wfeat[i] = (float)(feature_std * s.normal((uint64_t)i));
    });
  }
  if (proj && d.has_projection) {  // LeCun-normal: std = 1/sqrt(fan_in)
    Stream s(seed, kProj);
    const size_t Fd = (size_t)d.feature_dim;
    const double std_ = 1.0 / std::sqrt((double)Fd);
    for (size_t i = 0; i < Fd * Fd; ++i) proj[i] = (float)(std_ * s.normal(i));
  }
  if (layers) {  // He-normal: std = sqrt(2/fan_in), one stream per layer
    const int H = d.hidden_width;
    size_t off = 0;
    for (int l = 0; l <= d.n_hidden_layers; ++l) {
      const int fan_in = l == 0 ? d.feature_dim : H;
      const int fan_out = l == d.n_hidden_layers ? 3 : H;
      Stream s(seed, kLayer0 + (uint64_t)l);
      const double std_ = std::sqrt(2.0 / fan_in);
      for (size_t i = 0; i < (size_t)fan_in * fan_out; ++i)
        layers[off + i] = (float)(std_ * s.normal(i));
      off += (size_t)fan_in * fan_out;
    }
    if (d.aux_dims) {  // second layer's 16 aux columns, He-normal over its 64 + 16 inputs
      Stream s(seed, kLayer0 + (uint64_t)d.n_hidden_layers + 1);
      const double std_ = std::sqrt(2.0 / (H + 16));
      for (size_t i = 0; i < (size_t)H * 16; ++i) layers[off + i] = (float)(std_ * s.normal(i));
    }
  }
}

// G-buffer of `w x h` pixels with the reference encodings
// (Renderer/device_functions/camera.cu:249, 260, 262):
//  position U[0,1)^3 (already bbox-normalised), normal uniform on S^2,
//  view uniform on the normal's hemisphere, reflect = 2(n.v)n - v,
//  normal/reflect stored as v/2 + 0.5, kd U[0,1)^3, ks U[0,1)^3,
//  roughness U[0.05, 1), direct U[0,1)^3, env_flag = 0 with probability
//  `background_fraction`.  w components are 1.
void synth_gbuffer(uint64_t seed, int w, int h, float background_fraction,
                   const nwprt_gbuffer &o) {
  require(w > 0 && h > 0, NWPRT_ERR_INVALID_ARGUMENT, "gbuffer: bad size");
  const Stream sp(seed, kPos), snv(seed, kNormalView), skd(seed, kKd),
      sks(seed, kKs), sr(seed, kRough), sd(seed, kDirect), sf(seed, kFlag);
  const long long P = (long long)w * h;
  parallel_for(P, [&](long long p) {
    const uint64_t q = (uint64_t)p;
    if (o.position)
      for (int c = 0; c < 3; ++c) o.position[4 * p + c] = (float)sp.uniform(3 * q + c);
    // unit normal, uniform on the sphere
    const double z = 1.0 - 2.0 * snv.uniform(4 * q + 0);
    const double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    const double phi = 6.283185307179586 * snv.uniform(4 * q + 1);
    const double n[3] = {rr * std::cos(phi), rr * std::sin(phi), z};
    // view direction uniform on the normal's hemisphere
    const double vz = 1.0 - 2.0 * snv.uniform(4 * q + 2);
    const double vr = std::sqrt(std::fmax(0.0, 1.0 - vz * vz));
    const double vphi = 6.283185307179586 * snv.uniform(4 * q + 3);
    double v[3] = {vr * std::cos(vphi), vr * std::sin(vphi), vz};
    double ndv = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
    if (ndv < 0) {
      for (double &x : v) x = -x;
      ndv = -ndv;
    }
    double r[3];
    for (int c = 0; c < 3; ++c) r[c] = 2.0 * ndv * n[c] - v[c];
    for (int c = 0; c < 3; ++c) {
      if (o.normal) o.normal[4 * p + c] = (float)(n[c] * 0.5 + 0.5);
      if (o.reflect) o.reflect[4 * p + c] = (float)(r[c] * 0.5 + 0.5);
      if (o.kd) o.kd[4 * p + c] = (float)skd.uniform(3 * q + c);
      if (o.ks) o.ks[4 * p + c] = (float)sks.uniform(3 * q + c);
      if (o.direct) o.direct[4 * p + c] = (float)sd.uniform(3 * q + c);
    }
    float *f4[] = {o.position, o.normal, o.reflect, o.kd, o.ks, o.direct};
    for (float *b : f4)
      if (b) b[4 * p + 3] = 1.0f;
    if (o.roughness) o.roughness[p] = (float)(0.05 + 0.95 * sr.uniform(q));
    if (o.env_flag)
      o.env_flag[p] = sf.uniform(q) < (double)background_fraction ? 0 : 1;
  });
}

// Cubemap faces [6][N][N][3]: per texel and channel exp(1.5 * N(0,1)), then
// one 3x3-texel sun at 1e3 (all channels) centred at a seeded texel at least
// one texel from the face border.
void synth_cubemap(uint64_t seed, int n, float *faces) {
  require(n >= 4 && n <= NWPRT_MAX_FACE_SIZE && (n & (n - 1)) == 0,
          NWPRT_ERR_INVALID_ARGUMENT,
          "cubemap: face size must be a power of two in [4, 128]");
  const Stream st(seed, kCubeTexel), ss(seed, kCubeSun);
  const long long total = 6ll * n * n * 3;
  parallel_for(total, [&](long long i) { faces[i] = (float)std::exp(1.5 * st.normal((uint64_t)i)); });
  const int face = (int)(ss.uniform(0) * 6.0);
  const int row = 1 + (int)(ss.uniform(1) * (n - 2));
  const int col = 1 + (int)(ss.uniform(2) * (n - 2));
  for (int dr = -1; dr <= 1; ++dr)
    for (int dc = -1; dc <= 1; ++dc)
      for (int c = 0; c < 3; ++c)
        faces[(((size_t)face * n + (row + dr)) * n + (col + dc)) * 3 + c] = 1e3f;
}

}  // namespace nwprt
