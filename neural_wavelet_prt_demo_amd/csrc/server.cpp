// Relight request server: continuous batching of relight requests onto the
// persistent fused kernel (no counterpart in the reference, whose viewer and
// offline driver relight one frame per call: Renderer/OutputVisualizer.cpp:
// 61-126, Renderer/offlineRenderer.cpp:125-142).
//
// Driven by the caller (no threads): each poll retires finished launches,
// admits queued requests in arrival order against the memory budget and the
// slot count, prepares admitted requests on the preparation stream, and
// keeps up to kInFlight persistent launches ("ticks") queued on the launch
// stream.  A tick's work items are 64-pixel groups taken in arrival order
// from every admitted request until the tick's pair budget is spent, so a
// request that arrives while others run joins the next tick and the GPU
// never idles between requests.  The launch stream waits for a request's
// preparation event before the first tick that holds its items.
//
// Device state per admitted request is one block from a pool of exact-size
// blocks (selection tiles and the K strengths; per-pixel features and aux
// inputs only for scenes the persistent kernel cannot encode itself; for the
// FP8 form the 32x32 tiles, calibration sample and maxima; and for rotated /
// fresh maps the cubemap, coefficients, scores and indices), so steady-state
// serving allocates nothing.  The budget counts the blocks admitted requests
// hold.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "engine_state.hpp"
#include "kernels.hpp"

namespace nwprt {
namespace {

constexpr int kTickSlots = 4;         // tick buffers in the ring
constexpr int kInFlight = 2;          // ticks queued on the launch stream
constexpr int kTickItems = 1 << 16;   // work items per tick at most
// About 1.7 requests of 1080p x 300 wavelets: with items in arrival order a
// larger tick delays no one, and every launch ends in a partly idle tail
// (measured: 2^28 -> 2^30 lowers a lone request's latency by ~1 ms).
constexpr int64_t kDefaultTickPairs = 1ll << 30;
constexpr int kDefaultVariant = 5;    // pipe.w4reg.occ2

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Exact-size block pool over hipMalloc.
struct Pool {
  std::multimap<size_t, void *> free_blocks;
  ~Pool() {
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
  }
  void *get(size_t n) {
    auto it = free_blocks.find(n);
    if (it != free_blocks.end()) {
      void *p = it->second;
      free_blocks.erase(it);
      return p;
    }
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return p;
  }
  void put(void *p, size_t n) {
    if (p) free_blocks.emplace(n, p);
  }
  // Returns the idle blocks to the device (when a new size does not fit).
  void trim() {
    for (auto &kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
  }
};

struct Block {
  void *p = nullptr;
  size_t n = 0;
};

// Byte offsets of a request's buffers inside its block (npos: absent).
struct Layout {
  static constexpr size_t npos = ~(size_t)0;
  size_t features = npos, aux = npos, mk = 0, strength = 0;
  size_t mk32 = npos, calib = npos, amax = npos;             // FP8 form
  size_t faces = npos, coef = npos, scores = npos, idx = npos;  // built maps
  size_t total = 0;
};

struct Request {
  int64_t id = 0;
  nwprt_request r{};
  double t_submit = 0.0;
  int P = 0, K = 0, N = 0;
  bool fused = false;  // encoded inside the persistent kernel
  bool aux = false;    // the scene has aux inputs
  int slot = -1;
  int next_base = 0;          // first pixel of the next unissued work item
  bool waited = false;        // launch stream already waits for its preparation
  Layout lay;
  Block block;                // pooled device state
  size_t bytes = 0;
  template <class T>
  T *at(size_t off) const { return (T *)((char *)block.p + off); }
};

struct Tick {
  BatchItem *items_host = nullptr;  // pinned
  int *n_host = nullptr;            // pinned
  DevBuf items, meta;               // meta: int n_items | uint counter
  hipEvent_t done = nullptr;
  bool in_flight = false;
  std::vector<Request *> completing;
};

}  // namespace

struct Server {
  nwprt_ctx ctx = nullptr;
  nwprt_server_config cfg{};
  int variant = kDefaultVariant;
  hipStream_t prep = nullptr, launch = nullptr;
  bool softplus_set = false, softplus = true;
  Pool pool;
  DevBuf jobs;                         // RelightArgs[max_active]
  RelightArgs *jobs_host = nullptr;    // pinned staging, one per slot
  std::vector<hipEvent_t> prep_done;   // per slot
  std::vector<int> free_slots;
  Tick ticks[kTickSlots];
  int next_tick = 0;
  std::deque<Tick *> in_flight;
  std::map<int64_t, std::unique_ptr<Request>> requests;
  std::deque<Request *> queue;
  std::vector<Request *> active;       // admission order
  std::deque<std::pair<int64_t, double>> done;  // (id, latency ms), not yet reported
  int64_t next_id = 1;
  nwprt_server_stats stats{};

  ~Server() {
    if (launch) (void)hipStreamSynchronize(launch);
    if (prep) (void)hipStreamSynchronize(prep);
    for (Tick &t : ticks) {
      if (t.items_host) (void)hipHostFree(t.items_host);
      if (t.n_host) (void)hipHostFree(t.n_host);
      if (t.done) (void)hipEventDestroy(t.done);
    }
    for (hipEvent_t e : prep_done)
      if (e) (void)hipEventDestroy(e);
    if (jobs_host) (void)hipHostFree(jobs_host);
    for (auto &kv : requests) {
      pool.put(kv.second->block.p, kv.second->block.n);
      // The context destroys its servers before its scenes and maps.
      const nwprt_request &r = kv.second->r;
      auto sc = ctx->scenes.find(r.scene_id);
      if (sc != ctx->scenes.end()) --sc->second->server_refs;
      if (r.env_source == NWPRT_ENV_RESIDENT) {
        auto e = ctx->envs.find(r.env_id);
        if (e != ctx->envs.end()) --e->second->server_refs;
      }
    }
    if (prep) (void)hipStreamDestroy(prep);
    if (launch) (void)hipStreamDestroy(launch);
  }
};

void destroy_server(Server *s) { delete s; }

namespace {

Server &get_server(nwprt_ctx ctx, int id) {
  auto it = ctx->servers.find(id);
  require(it != ctx->servers.end(), NWPRT_ERR_NOT_FOUND, "unknown server id " + std::to_string(id));
  return *it->second;
}

bool env_built(const nwprt_request &r) { return r.env_source != NWPRT_ENV_RESIDENT; }

// The request's buffers, each 256-byte aligned, in one block.
Layout request_layout(const Server &sv, const Scene &sc, const Request &q) {
  Layout l;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t at = o;
    o += (n + 255) & ~(size_t)255;
    return at;
  };
  const bool fp8 = sv.cfg.precision == NWPRT_PREC_FP8;
  if (!q.fused) l.features = take((size_t)q.P * 64);
  if (q.aux) l.aux = take((size_t)q.P * 32);
  l.mk = take((size_t)q.K * 4 * 64 * 16);
  l.strength = take((size_t)q.K * 12);
  if (fp8) {
    l.mk32 = take(fp8_mk_bytes(q.K));
    l.calib = take(fp8_calib_bytes(sc, q.P));
    l.amax = take(16);
  }
  if (env_built(q.r)) {
    const size_t m = 6ull * q.N * q.N;
    l.faces = take(m * 12);
    l.coef = take(m * 12);
    l.scores = take(envmap_work_bytes(q.N, 1));
    l.idx = take((size_t)q.K * 4);
  }
  l.total = o;
  return l;
}

void validate_request(Server &sv, const nwprt_request &r, Request &q) {
  const Scene &sc = get_scene(sv.ctx, r.scene_id);
  check_frame(&r.gbuffer, r.width, r.height, true);
  check_outputs(r.out);
  q.aux = sc.desc.aux_dims != 0;
  if (q.aux) {  // checks the G-buffer's aux inputs
    EncodeArgs e{};
    set_aux_encode(e, sc, &r.gbuffer, nullptr);
  }
  const bool sp = sc.desc.output_softplus != 0;
  require(!sv.softplus_set || sp == sv.softplus, NWPRT_ERR_UNSUPPORTED,
          "server: all requests must share the output activation");
  q.P = r.width * r.height;
  q.fused = fuse_encode(sc, sv.cfg.precision);
  switch (r.env_source) {
    case NWPRT_ENV_RESIDENT: {
      const EnvMap &e = get_env(sv.ctx, r.env_id);
      q.K = e.k;
      q.N = e.n;
      break;
    }
    case NWPRT_ENV_ROTATED:
    case NWPRT_ENV_FRESH: {
      const int n = r.face_size;
      require(n >= 2 && n <= NWPRT_MAX_FACE_SIZE && (n & (n - 1)) == 0,
              NWPRT_ERR_INVALID_ARGUMENT, "server: face size must be a power of two in [2, 128]");
      require(r.k >= 1 && r.k <= NWPRT_MAX_K && r.k <= 6 * n * n, NWPRT_ERR_INVALID_ARGUMENT,
              "server: K must be in [1, min(1024, 6 N^2)]");
      require(r.faces != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "server: null faces");
      if (r.env_source == NWPRT_ENV_ROTATED) {
        require(std::isfinite(r.yaw), NWPRT_ERR_INVALID_ARGUMENT, "server: yaw must be finite");
      } else {
        require(r.faces_location == NWPRT_HOST || r.faces_location == NWPRT_DEVICE,
                NWPRT_ERR_INVALID_ARGUMENT, "server: bad faces location");
      }
      q.K = r.k;
      q.N = n;
      break;
    }
    default: fail(NWPRT_ERR_INVALID_ARGUMENT, "server: unknown environment source");
  }
  require(q.N == sc.desc.face_size, NWPRT_ERR_INVALID_ARGUMENT,
          "server: cubemap face size does not match the scene's wavelet feature table");
  q.lay = request_layout(sv, sc, q);
  const size_t bytes = q.lay.total;
  require(!sv.cfg.memory_budget_bytes || bytes <= sv.cfg.memory_budget_bytes,
          NWPRT_ERR_INVALID_ARGUMENT,
          "server: the request needs " + std::to_string(bytes) +
              " device bytes, more than the whole memory budget");
  sv.softplus = sp;
  sv.softplus_set = true;
}

// Preparation of an admitted request on the preparation stream: its map's
// selection, the hash-grid encode (unless the persistent kernel encodes the
// pixels itself), the selection tiles (and for the FP8 form the calibration)
// and its job-table entry; then the slot's event.
void prepare(Server &sv, Request &q) {
  const hipStream_t s = sv.prep;
  const nwprt_precision prec = sv.cfg.precision;
  const Elem elem = elem_of(prec);
  const Scene &sc = get_scene(sv.ctx, q.r.scene_id);
  const Layout &l = q.lay;
  float *strength = q.at<float>(l.strength);
  const int32_t *idx = nullptr;
  if (!env_built(q.r)) {
    EnvMap &e = get_env(sv.ctx, q.r.env_id);
    hip_check(hipStreamWaitEvent(s, e.ready, 0), "hipStreamWaitEvent");
    idx = e.idx.as<int32_t>();
    hip_check(hipMemcpyAsync(strength, e.strength.p, (size_t)q.K * 12, hipMemcpyDeviceToDevice, s),
              "hipMemcpyAsync strengths");
  } else {
    const size_t m = 6ull * q.N * q.N;
    float *faces = q.at<float>(l.faces), *coef = q.at<float>(l.coef);
    int32_t *bidx = q.at<int32_t>(l.idx);
    const float *src = q.r.faces;
    if (q.r.env_source == NWPRT_ENV_ROTATED) {
      launch_rotate_cubemap(q.r.faces, faces, q.N, q.r.yaw, s);
      src = faces;
    } else if (q.r.faces_location == NWPRT_HOST) {
      hip_check(hipMemcpyAsync(faces, q.r.faces, m * 12, hipMemcpyHostToDevice, s),
                "hipMemcpyAsync faces");
      src = faces;
    }
    launch_haar(src, m * 3, coef, q.N, 1, q.at<void>(l.scores), s);
    launch_topk(coef, q.N, q.K, q.at<void>(l.scores), bidx, strength, s);
    idx = bidx;
  }
  void *features = q.fused ? nullptr : q.at<void>(l.features);
  void *aux = q.aux ? q.at<void>(l.aux) : nullptr;
  if (!q.fused) {
    EncodeArgs e = encode_args(sc, &q.r.gbuffer, q.P, elem, features);
    if (q.aux) set_aux_encode(e, sc, &q.r.gbuffer, aux);
    launch_encode(e, s);
  }
  RelightArgs &a = sv.jobs_host[q.slot];
  a = RelightArgs{};
  a.n_pixels = q.P;
  a.k = q.K;
  a.features = features;
  a.features_planar = planar_features(sc, elem);
  if (q.fused) set_fused_encode(a, sc, &q.r.gbuffer);
  if (q.aux) set_aux_relight(a, sc, prec, aux);
  a.mk = q.at<uint16_t>(l.mk);
  a.light = strength;
  a.w_tiles = (const uint16_t *)scene_tiles(sc, prec == NWPRT_PREC_FP8 ? NWPRT_PREC_BF16 : prec);
  a.env_flag = q.r.gbuffer.env_flag;
  a.direct = q.r.gbuffer.direct;
  a.indirect = q.r.out.indirect;
  a.with_direct = q.r.out.indirect_with_direct;
  a.rgba8 = q.r.out.rgba8;
  a.tonemap_source = (int)sv.cfg.tonemap;
  a.softplus = sv.softplus;
  if (prec == NWPRT_PREC_FP8) {
    Fp8Work w;
    w.mk16 = q.at<uint16_t>(l.mk);
    w.mk32 = q.at<uint16_t>(l.mk32);
    w.calib = q.at<uint16_t>(l.calib);
    w.amax = q.at<float>(l.amax);
    prepare_fp8(sc, idx, q.K, a, w, s);
    set_fp8_relight(a, sc, w);
  } else {
    launch_prepare_selection(sc.w_f32.as<float>(), sc.wfeat.as<float>(), 6 * q.N * q.N, idx, q.K,
                             elem, q.at<uint16_t>(l.mk), nullptr, s, nullptr,
                             planar_features(sc, elem), scene_projection(sc));
  }
  hip_check(hipMemcpyAsync(sv.jobs.as<RelightArgs>() + q.slot, &a, sizeof(RelightArgs),
                           hipMemcpyHostToDevice, s),
            "hipMemcpyAsync job");
  hip_check(hipEventRecord(sv.prep_done[q.slot], s), "hipEventRecord");
}

void release(Server &sv, Request &q) {
  sv.pool.put(q.block.p, q.block.n);
  q.block = Block{};
  sv.stats.bytes_in_use -= q.bytes;
  q.bytes = 0;
  if (q.slot >= 0) sv.free_slots.push_back(q.slot);
  q.slot = -1;
}

// The scene and resident map a request reads stay alive (destroy and
// in-place updates refuse) from submit until it completes or fails.
void add_refs(Server &sv, const nwprt_request &r, int d) {
  get_scene(sv.ctx, r.scene_id).server_refs += d;
  if (!env_built(r)) get_env(sv.ctx, r.env_id).server_refs += d;
}

// A request whose admission or preparation failed: its resources are
// returned, and it is reported done with latency -status.
void fail_request(Server &sv, Request &q, nwprt_status st) {
  release(sv, q);
  add_refs(sv, q.r, -1);
  sv.done.emplace_back(q.id, -(double)st);
  ++sv.stats.failed;
  sv.requests.erase(q.id);
}

void retire(Server &sv, bool block) {
  while (!sv.in_flight.empty()) {
    Tick *t = sv.in_flight.front();
    if (block) {
      hip_check(hipEventSynchronize(t->done), "hipEventSynchronize");
      block = false;
    } else {
      const hipError_t e = hipEventQuery(t->done);
      if (e == hipErrorNotReady) break;
      hip_check(e, "hipEventQuery");
    }
    const double tnow = now_s();
    for (Request *q : t->completing) {
      sv.done.emplace_back(q->id, (tnow - q->t_submit) * 1e3);
      sv.active.erase(std::find(sv.active.begin(), sv.active.end(), q));
      release(sv, *q);
      add_refs(sv, q->r, -1);
      ++sv.stats.completed;
      sv.requests.erase(q->id);
    }
    t->completing.clear();
    t->in_flight = false;
    sv.in_flight.pop_front();
  }
}

// Admits queued requests in arrival order while a slot is free and the
// budget allows.  Admission is transactional: a request whose device state
// cannot be allocated (with nothing admitted that could free some) or whose
// preparation fails gives back its blocks, slot and budget and is reported
// as failed, so the queue always moves on.
void admit(Server &sv) {
  while (!sv.queue.empty() && !sv.free_slots.empty()) {
    Request &q = *sv.queue.front();
    const size_t bytes = q.lay.total;
    if (sv.cfg.memory_budget_bytes && sv.stats.bytes_in_use + bytes > sv.cfg.memory_budget_bytes)
      break;
    try {
      void *p = sv.pool.get(bytes);
      if (!p) {  // return idle blocks of other sizes to the device and retry
        sv.pool.trim();
        p = sv.pool.get(bytes);
      }
      if (!p) {
        // Device memory exhausted: wait for admitted requests to complete;
        // with none admitted nothing would ever free.
        require(!sv.active.empty(), NWPRT_ERR_OUT_OF_MEMORY,
                "server: device memory for a request's state is not available");
        return;
      }
      q.block = Block{p, bytes};
      q.bytes = bytes;
      sv.stats.bytes_in_use += bytes;
      q.slot = sv.free_slots.back();
      sv.free_slots.pop_back();
      prepare(sv, q);
    } catch (const Error &e) {
      sv.queue.pop_front();
      fail_request(sv, q, e.status);
      continue;
    }
    sv.stats.peak_bytes = std::max(sv.stats.peak_bytes, sv.stats.bytes_in_use);
    sv.active.push_back(&q);
    sv.queue.pop_front();
    sv.stats.peak_active = std::max(sv.stats.peak_active, (int32_t)sv.active.size());
  }
}

// Fills and launches ticks while fewer than kInFlight are queued.
void launch_ticks(Server &sv) {
  const int64_t budget = sv.cfg.tick_pairs > 0 ? sv.cfg.tick_pairs : kDefaultTickPairs;
  while ((int)sv.in_flight.size() < kInFlight) {
    Tick &t = sv.ticks[sv.next_tick];
    if (t.in_flight) break;  // the ring is full (cannot happen with kInFlight < kTickSlots)
    int n = 0;
    int64_t pairs = 0;
    bool has_aux = false;
    std::vector<Request *> touched;
    for (Request *q : sv.active) {
      if (q->next_base >= q->P) continue;
      touched.push_back(q);
      has_aux = has_aux || q->aux;
      while (q->next_base < q->P && n < kTickItems && pairs < budget) {
        t.items_host[n++] = BatchItem{q->slot, q->next_base};
        q->next_base += 64;
        pairs += 64ll * q->K;
      }
      if (q->next_base >= q->P) t.completing.push_back(q);
      if (n >= kTickItems || pairs >= budget) break;
    }
    if (n == 0) break;
    for (Request *q : touched)
      if (!q->waited) {
        hip_check(hipStreamWaitEvent(sv.launch, sv.prep_done[q->slot], 0), "hipStreamWaitEvent");
        q->waited = true;
      }
    *t.n_host = n;
    hip_check(hipMemcpyAsync(t.items.p, t.items_host, (size_t)n * sizeof(BatchItem),
                             hipMemcpyHostToDevice, sv.launch),
              "hipMemcpyAsync items");
    hip_check(hipMemcpyAsync(t.meta.p, t.n_host, sizeof(int), hipMemcpyHostToDevice, sv.launch),
              "hipMemcpyAsync count");
    unsigned int *counter = (unsigned int *)(t.meta.as<int>() + 1);
    launch_reset_counter(counter, sv.launch);
    if (sv.cfg.precision == NWPRT_PREC_FP8)
      launch_relight_fp8_batch(sv.jobs.as<RelightArgs>(), t.items.as<BatchItem>(),
                               t.meta.as<int>(), counter, sv.softplus, has_aux, sv.variant,
                               sv.ctx->n_cu, sv.launch);
    else
      launch_relight_batch(sv.jobs.as<RelightArgs>(), t.items.as<BatchItem>(), t.meta.as<int>(),
                           counter, elem_of(sv.cfg.precision), sv.softplus, has_aux, sv.variant,
                           sv.ctx->n_cu, sv.launch);
    hip_check(hipEventRecord(t.done, sv.launch), "hipEventRecord");
    t.in_flight = true;
    sv.in_flight.push_back(&t);
    sv.next_tick = (sv.next_tick + 1) % kTickSlots;
    ++sv.stats.ticks;
    sv.stats.items += n;
  }
}

void advance(Server &sv, bool block_on_tick) {
  retire(sv, block_on_tick);
  admit(sv);
  launch_ticks(sv);
}

}  // namespace
}  // namespace nwprt

using namespace nwprt;

extern "C" {

nwprt_status nwprt_server_create(nwprt_ctx ctx, const nwprt_server_config *config,
                                 int32_t *server_id) {
  return guarded([&] {
    bind(ctx);
    require(config && server_id, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    const nwprt_server_config &c = *config;
    require(c.precision == NWPRT_PREC_BF16 || c.precision == NWPRT_PREC_FP16 ||
                c.precision == NWPRT_PREC_FP8,
            NWPRT_ERR_UNSUPPORTED, "server: BF16, FP16 or FP8 precision only");
    require(c.max_active >= 1 && c.max_active <= 4096, NWPRT_ERR_INVALID_ARGUMENT,
            "server: max_active must be in [1, 4096]");
    require(c.tick_pairs >= 0, NWPRT_ERR_INVALID_ARGUMENT, "server: negative tick_pairs");
    require(c.tonemap == NWPRT_TONEMAP_FULL || c.tonemap == NWPRT_TONEMAP_INDIRECT ||
                c.tonemap == NWPRT_TONEMAP_NONE,
            NWPRT_ERR_INVALID_ARGUMENT, "unknown tonemap source");
    const bool fp8 = c.precision == NWPRT_PREC_FP8;
    const int variant = c.variant >= 0 ? c.variant : fp8 ? kDefaultFp8Variant : kDefaultVariant;
    require(fp8 ? variant < relight_fp8_variant_count() : relight_batch_variant_ok(variant),
            NWPRT_ERR_INVALID_ARGUMENT, "server: tile variant has no persistent form");
    ServerPtr sv(new Server(), destroy_server);
    sv->ctx = ctx;
    sv->cfg = c;
    sv->variant = variant;
    hip_check(hipStreamCreateWithFlags(&sv->prep, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&sv->launch, hipStreamNonBlocking), "hipStreamCreate");
    sv->jobs.ensure((size_t)c.max_active * sizeof(RelightArgs));
    hip_check(hipHostMalloc((void **)&sv->jobs_host, (size_t)c.max_active * sizeof(RelightArgs)),
              "hipHostMalloc");
    sv->prep_done.assign(c.max_active, nullptr);
    for (int i = 0; i < c.max_active; ++i) {
      hip_check(hipEventCreateWithFlags(&sv->prep_done[i], hipEventDisableTiming),
                "hipEventCreate");
      sv->free_slots.push_back(c.max_active - 1 - i);
    }
    for (Tick &t : sv->ticks) {
      hip_check(hipHostMalloc((void **)&t.items_host, kTickItems * sizeof(BatchItem)),
                "hipHostMalloc");
      hip_check(hipHostMalloc((void **)&t.n_host, sizeof(int)), "hipHostMalloc");
      t.items.ensure(kTickItems * sizeof(BatchItem));
      t.meta.ensure(2 * sizeof(int));
      hip_check(hipEventCreateWithFlags(&t.done, hipEventDisableTiming), "hipEventCreate");
    }
    const int id = ctx->next_id++;
    ctx->servers.emplace(id, std::move(sv));
    *server_id = id;
  });
}

nwprt_status nwprt_server_submit(nwprt_ctx ctx, int32_t server_id, const nwprt_request *request,
                                 int64_t *request_id) {
  return guarded([&] {
    bind(ctx);
    require(request != nullptr, NWPRT_ERR_INVALID_ARGUMENT, "null request");
    Server &sv = get_server(ctx, server_id);
    auto q = std::make_unique<Request>();
    q->r = *request;
    validate_request(sv, *request, *q);
    add_refs(sv, q->r, +1);
    q->id = sv.next_id++;
    q->t_submit = now_s();
    sv.queue.push_back(q.get());
    const int64_t id = q->id;
    sv.requests.emplace(id, std::move(q));
    ++sv.stats.submitted;
    advance(sv, false);
    if (request_id) *request_id = id;
  });
}

nwprt_status nwprt_server_poll(nwprt_ctx ctx, int32_t server_id, int32_t wait, int64_t *done_ids,
                               double *latency_ms, int32_t max_done, int32_t *n_done) {
  return guarded([&] {
    bind(ctx);
    require(wait >= 0 && wait <= 2 && max_done >= 0 && (max_done == 0 || done_ids),
            NWPRT_ERR_INVALID_ARGUMENT, "server poll: bad argument");
    Server &sv = get_server(ctx, server_id);
    advance(sv, false);
    auto pending = [&] { return !sv.queue.empty() || !sv.active.empty(); };
    if (wait == 1) {
      while (sv.done.empty() && pending()) advance(sv, !sv.in_flight.empty());
    } else if (wait == 2) {
      while (pending()) advance(sv, !sv.in_flight.empty());
    }
    int n = 0;
    while (n < max_done && !sv.done.empty()) {
      done_ids[n] = sv.done.front().first;
      if (latency_ms) latency_ms[n] = sv.done.front().second;
      sv.done.pop_front();
      ++n;
    }
    if (n_done) *n_done = n;
  });
}

nwprt_status nwprt_server_get_stats(nwprt_ctx ctx, int32_t server_id, nwprt_server_stats *stats) {
  return guarded([&] {
    require(ctx && stats, NWPRT_ERR_INVALID_ARGUMENT, "null argument");
    Server &sv = get_server(ctx, server_id);
    *stats = sv.stats;
    stats->queued = (int64_t)sv.queue.size();
    stats->active = (int64_t)sv.active.size();
  });
}

nwprt_status nwprt_server_destroy(nwprt_ctx ctx, int32_t server_id) {
  return guarded([&] {
    bind(ctx);
    get_server(ctx, server_id);
    ctx->servers.erase(server_id);
  });
}

}  // extern "C"
