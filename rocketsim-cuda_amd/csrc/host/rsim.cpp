// The C ABI of librsim.so (include/rsim.h): device, buffers, uploads with the AoSoA-64 pack, the decision launch and the
// parity kernels. Mirrors the reference's SimDevice (crates/rocketsim-cuda/src/device.rs:30-172): one stream per device,
// every call ordered on it, host data returned only after a synchronisation. Misuse returns an error code instead of the
// reference's assert! panics; nothing here throws across the ABI.
#include "rsim.h"

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "device/cost_model.h"
#include "device/launch.h"

static_assert(RSIM_COST_ORDER_NOALLOC_MAX == rs::dev::ORDER_SCRATCH_MAX_N, "the header's limit is the workspace's");

using rs::ArenaState;
using rs::Controls;
using rs::StaticData;

struct rsim_device {
  int ordinal;
  hipStream_t stream;
  hipEvent_t t0, t1;
  int32_t* status;  // per-arena status words of the parity kernels
  uint32_t status_cap;
  uint64_t live;  // worlds, state and word buffers (views included) and graphs not yet freed: rsim_close refuses while > 0
  uint32_t* order_scratch;  // this GPU's copy of the allocation-free order's fixed workspace (rsim_cost_order_noalloc)
  uint32_t* queue;          // the block queue's counter of rsim_run_queued (one word, reset on the stream by each launch)
};
struct rsim_world {
  rsim_device* dev;
  StaticData* d;   // count StaticData back to back
  uint32_t count;
};
struct rsim_states {
  rsim_device* dev;
  uint32_t* d;
  uint32_t n;
};
struct rsim_words {
  rsim_device* dev;
  uint32_t* d;
  uint64_t n;
  bool view = false;  // a range of another buffer: freeing it leaves the memory alone
};
struct rsim_graph {
  rsim_device* dev;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

thread_local std::string g_error;

int fail(int code, const char* what) {
  g_error = what;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  return RSIM_ERR_HIP;
}
#define RSIM_HIP(call, what)                         \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Makes dev's GPU current for this thread (HIP's current device is per thread).
int use(rsim_device* dev) {
  if (!dev) return fail(RSIM_ERR_INVALID, "null device");
  RSIM_HIP(hipSetDevice(dev->ordinal), "hipSetDevice");
  return RSIM_OK;
}
// A call that allocates, frees, copies to or from the host or synchronises cannot run while the device's stream records a
// graph: it would record a memset or a copy into the graph, or its synchronisation would end the capture with an error.
// Such calls refuse (RSIM_ERR_INVALID) and leave everything as it was.
int use_outside_capture(rsim_device* dev, const char* what) {
  int rc = use(dev);
  if (rc) return rc;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  RSIM_HIP(hipStreamIsCapturing(dev->stream, &st), "hipStreamIsCapturing");
  if (st != hipStreamCaptureStatusNone)
    return fail(RSIM_ERR_INVALID, (std::string(what) + ": not allowed while the device's stream records a graph").c_str());
  return RSIM_OK;
}

void pack(const ArenaState* states, uint32_t n, uint32_t* out) {
  const size_t words = rs::dev::aosoa_words(n);
  const uint32_t L = rs::dev::AOSOA_LANES;
  for (uint32_t b = 0; b * L < n || (n == 0 && false); ++b) {
    uint32_t* blk = out + (size_t)b * rs::ARENA_WORDS * L;
    for (uint32_t l = 0; l < L; ++l) {
      uint32_t a = b * L + l;
      if (a < n) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&states[a]);
        for (int w = 0; w < rs::ARENA_WORDS; ++w) blk[(size_t)w * L + l] = src[w];
      } else {
        for (int w = 0; w < rs::ARENA_WORDS; ++w) blk[(size_t)w * L + l] = 0;
      }
    }
  }
  (void)words;
}
void unpack(const uint32_t* in, uint32_t n, ArenaState* states) {
  const uint32_t L = rs::dev::AOSOA_LANES;
  for (uint32_t a = 0; a < n; ++a) {
    const uint32_t* blk = in + (size_t)(a / L) * rs::ARENA_WORDS * L + (a % L);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&states[a]);
    for (int w = 0; w < rs::ARENA_WORDS; ++w) dst[w] = blk[(size_t)w * L];
  }
}

int ensure_status(rsim_device* dev, uint32_t n) {
  if (dev->status_cap >= n && dev->status) return RSIM_OK;
  if (dev->status) (void)hipFree(dev->status);
  dev->status = nullptr;
  dev->status_cap = 0;
  uint32_t cap = n < 64 ? 64 : n;
  RSIM_HIP(hipMalloc((void**)&dev->status, (size_t)cap * sizeof(int32_t)), "hipMalloc(status)");
  dev->status_cap = cap;
  return RSIM_OK;
}
// The status words of the last parity launch back to the host (synchronous).
int finish_status(rsim_device* dev, uint32_t n, int32_t* status_out) {
  RSIM_HIP(hipStreamSynchronize(dev->stream), "parity kernel");
  if (status_out && n)
    RSIM_HIP(hipMemcpy(status_out, dev->status, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost), "status download");
  return RSIM_OK;
}

}  // namespace

extern "C" {

// ---- layout ------------------------------------------------------------------------------------------------------------------
uint64_t rsim_arena_words(void) { return rs::ARENA_WORDS; }
uint64_t rsim_tick_words(void) { return rs::TICK_WORDS; }
uint64_t rsim_static_words(void) { return rs::STATIC_WORDS; }
uint64_t rsim_aosoa_words(uint32_t n) { return rs::dev::aosoa_words(n); }
uint64_t rsim_aosoa_index(uint32_t a, uint32_t w) { return rs::dev::aosoa_index(a, w); }
int rsim_aosoa_pack(const void* states, uint32_t n, uint32_t* aosoa) {
  if ((!states || !aosoa) && n) return fail(RSIM_ERR_INVALID, "rsim_aosoa_pack: null buffer");
  pack(static_cast<const ArenaState*>(states), n, aosoa);
  return RSIM_OK;
}
int rsim_aosoa_unpack(const uint32_t* aosoa, uint32_t n, void* states) {
  if ((!states || !aosoa) && n) return fail(RSIM_ERR_INVALID, "rsim_aosoa_unpack: null buffer");
  unpack(aosoa, n, static_cast<ArenaState*>(states));
  return RSIM_OK;
}
const char* rsim_last_error(void) { return g_error.c_str(); }

// ---- device ------------------------------------------------------------------------------------------------------------------
int rsim_open(int ordinal, rsim_device** out) {
  if (!out) return fail(RSIM_ERR_INVALID, "rsim_open: null out");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(RSIM_ERR_NO_DEVICE, "rsim_open: no GPU");
  if (ordinal < 0 || ordinal >= count) return fail(RSIM_ERR_NO_DEVICE, "rsim_open: no such GPU");
  RSIM_HIP(hipSetDevice(ordinal), "hipSetDevice");
  rsim_device* d = new (std::nothrow) rsim_device{};
  if (!d) return fail(RSIM_ERR_INVALID, "rsim_open: out of host memory");
  d->ordinal = ordinal;
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&d->t0);
  if (e == hipSuccess) e = hipEventCreate(&d->t1);
  // the device address of this GPU's fixed order workspace, looked up once with the GPU current
  if (e == hipSuccess) e = rs::dev::order_scratch_address(&d->order_scratch);
  size_t static_lds = 0;
  if (e == hipSuccess) e = rs::dev::stepping_static_lds(&static_lds);
  if (e == hipSuccess) e = hipMalloc((void**)&d->queue, 256);
  auto discard = [d]() {
    if (d->queue) (void)hipFree(d->queue);
    if (d->t1) (void)hipEventDestroy(d->t1);
    if (d->t0) (void)hipEventDestroy(d->t0);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
  };
  if (e != hipSuccess) {
    discard();
    return hip_fail(e, "rsim_open");
  }
  if (static_lds != 0) {  // a build whose stepping kernels gained static LDS: their LDS addressing would overlap it
    discard();
    return fail(RSIM_ERR_HIP, "rsim_open: the stepping kernels declare static LDS; their dynamic LDS would not start at 0");
  }
  *out = d;
  return RSIM_OK;
}
int rsim_close(rsim_device* dev) {
  if (!dev) return RSIM_OK;
  int rc = use_outside_capture(dev, "rsim_close");
  if (rc) return rc;
  if (dev->live)
    return fail(RSIM_ERR_INVALID, "rsim_close: worlds, buffers or graphs of this device are still open (free them first)");
  (void)hipStreamSynchronize(dev->stream);
  if (dev->status) (void)hipFree(dev->status);
  if (dev->queue) (void)hipFree(dev->queue);
  (void)hipEventDestroy(dev->t0);
  (void)hipEventDestroy(dev->t1);
  (void)hipStreamDestroy(dev->stream);
  delete dev;
  return RSIM_OK;
}
void* rsim_stream(rsim_device* dev) { return dev ? (void*)dev->stream : nullptr; }
int rsim_sync(rsim_device* dev) {
  int rc = use_outside_capture(dev, "rsim_sync");
  if (rc) return rc;
  RSIM_HIP(hipStreamSynchronize(dev->stream), "rsim_sync");
  return RSIM_OK;
}
int rsim_timer_begin(rsim_device* dev) {
  int rc = use_outside_capture(dev, "rsim_timer_begin");
  if (rc) return rc;
  RSIM_HIP(hipEventRecord(dev->t0, dev->stream), "hipEventRecord");
  return RSIM_OK;
}
int rsim_timer_end(rsim_device* dev, float* ms) {
  int rc = use_outside_capture(dev, "rsim_timer_end");
  if (rc) return rc;
  if (!ms) return fail(RSIM_ERR_INVALID, "rsim_timer_end: null ms");
  RSIM_HIP(hipEventRecord(dev->t1, dev->stream), "hipEventRecord");
  RSIM_HIP(hipEventSynchronize(dev->t1), "hipEventSynchronize");
  RSIM_HIP(hipEventElapsedTime(ms, dev->t0, dev->t1), "hipEventElapsedTime");
  return RSIM_OK;
}

// ---- world -------------------------------------------------------------------------------------------------------------------
int rsim_upload_worlds(rsim_device* dev, const void* static_data, uint32_t count, rsim_world** out) {
  int rc = use_outside_capture(dev, "rsim_upload_worlds");
  if (rc) return rc;
  if (!static_data || !out || count == 0) return fail(RSIM_ERR_INVALID, "rsim_upload_worlds: null argument or no world");
  if (count > RSIM_MAX_WORLDS) return fail(RSIM_ERR_INVALID, "rsim_upload_worlds: more than RSIM_MAX_WORLDS worlds");
  rsim_world* w = new (std::nothrow) rsim_world{dev, nullptr, count};
  if (!w) return fail(RSIM_ERR_INVALID, "out of host memory");
  const size_t bytes = sizeof(StaticData) * count;
  hipError_t e = hipMalloc((void**)&w->d, bytes);
  if (e == hipSuccess) e = hipMemcpyAsync(w->d, static_data, bytes, hipMemcpyHostToDevice, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  if (e != hipSuccess) {
    if (w->d) (void)hipFree(w->d);
    delete w;
    return hip_fail(e, "rsim_upload_worlds");
  }
  ++dev->live;
  *out = w;
  return RSIM_OK;
}
int rsim_upload_world(rsim_device* dev, const void* static_data, rsim_world** out) {
  return rsim_upload_worlds(dev, static_data, 1, out);
}
uint32_t rsim_world_count(const rsim_world* w) { return w ? w->count : 0; }
const void* rsim_world_device_ptr(const rsim_world* w) { return w ? w->d : nullptr; }
int rsim_world_free(rsim_world* w) {
  if (!w) return RSIM_OK;
  int rc = use_outside_capture(w->dev, "rsim_world_free");
  if (rc) return rc;
  (void)hipStreamSynchronize(w->dev->stream);
  (void)hipFree(w->d);
  --w->dev->live;
  delete w;
  return RSIM_OK;
}

// ---- states ------------------------------------------------------------------------------------------------------------------
int rsim_alloc_states(rsim_device* dev, uint32_t n, rsim_states** out) {
  int rc = use_outside_capture(dev, "rsim_alloc_states");
  if (rc) return rc;
  if (!out) return fail(RSIM_ERR_INVALID, "rsim_alloc_states: null out");
  size_t bytes = rs::dev::aosoa_words(n > 0 ? n : 1) * sizeof(uint32_t);
  rsim_states* s = new (std::nothrow) rsim_states{dev, nullptr, n};
  if (!s) return fail(RSIM_ERR_INVALID, "out of host memory");
  hipError_t e = hipMalloc((void**)&s->d, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(s->d, 0, bytes, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  if (e != hipSuccess) {
    if (s->d) (void)hipFree(s->d);
    delete s;
    return hip_fail(e, "rsim_alloc_states");
  }
  ++dev->live;
  *out = s;
  return RSIM_OK;
}
int rsim_write_states(rsim_device* dev, rsim_states* dst, const void* states) {
  int rc = use_outside_capture(dev, "rsim_write_states");
  if (rc) return rc;
  if (!dst || (!states && dst->n)) return fail(RSIM_ERR_INVALID, "rsim_write_states: null argument");
  std::vector<uint32_t> host;
  try {
    host.resize(rs::dev::aosoa_words(dst->n));
  } catch (...) {
    return fail(RSIM_ERR_INVALID, "rsim_write_states: out of host memory");
  }
  pack(static_cast<const ArenaState*>(states), dst->n, host.data());
  RSIM_HIP(hipMemcpyAsync(dst->d, host.data(), host.size() * 4, hipMemcpyHostToDevice, dev->stream), "state upload");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "state upload");
  return RSIM_OK;
}
int rsim_upload_states(rsim_device* dev, const void* states, uint32_t n, rsim_states** out) {
  int rc = rsim_alloc_states(dev, n, out);
  if (rc) return rc;
  rc = rsim_write_states(dev, *out, states);
  if (rc) {
    rsim_states_free(*out);
    *out = nullptr;
  }
  return rc;
}
int rsim_copy_states(rsim_device* dev, rsim_states* dst, const rsim_states* src) {
  int rc = use(dev);
  if (rc) return rc;
  if (!dst || !src || dst->n != src->n) return fail(RSIM_ERR_INVALID, "rsim_copy_states: buffers of different sizes");
  if (!src->n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_copy_words(dev->stream, src->d, dst->d, rs::dev::aosoa_words(src->n)), "state copy");
  return RSIM_OK;
}
int rsim_states_free(rsim_states* s) {
  if (!s) return RSIM_OK;
  int rc = use_outside_capture(s->dev, "rsim_states_free");
  if (rc) return rc;
  (void)hipStreamSynchronize(s->dev->stream);
  (void)hipFree(s->d);
  --s->dev->live;
  delete s;
  return RSIM_OK;
}
uint32_t rsim_states_count(const rsim_states* s) { return s ? s->n : 0; }
void* rsim_states_device_ptr(rsim_states* s) { return s ? s->d : nullptr; }
int rsim_download_words(rsim_device* dev, const rsim_states* s, uint32_t* out) {
  int rc = use_outside_capture(dev, "rsim_download_words");
  if (rc) return rc;
  if (!s || (!out && s->n)) return fail(RSIM_ERR_INVALID, "rsim_download_words: null argument");
  if (!s->n) return RSIM_OK;
  RSIM_HIP(hipMemcpyAsync(out, s->d, rs::dev::aosoa_words(s->n) * 4, hipMemcpyDeviceToHost, dev->stream), "state download");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "state download");
  return RSIM_OK;
}
int rsim_download_states(rsim_device* dev, const rsim_states* s, void* out) {
  if (!s) return fail(RSIM_ERR_INVALID, "rsim_download_states: null states");
  std::vector<uint32_t> host;
  try {
    host.resize(rs::dev::aosoa_words(s->n));
  } catch (...) {
    return fail(RSIM_ERR_INVALID, "rsim_download_states: out of host memory");
  }
  int rc = rsim_download_words(dev, s, host.data());
  if (rc) return rc;
  unpack(host.data(), s->n, static_cast<ArenaState*>(out));
  return RSIM_OK;
}

// ---- word buffers ------------------------------------------------------------------------------------------------------------
int rsim_alloc_words(rsim_device* dev, uint64_t words, rsim_words** out) {
  int rc = use_outside_capture(dev, "rsim_alloc_words");
  if (rc) return rc;
  if (!out) return fail(RSIM_ERR_INVALID, "rsim_alloc_words: null out");
  rsim_words* b = new (std::nothrow) rsim_words{dev, nullptr, words};
  if (!b) return fail(RSIM_ERR_INVALID, "out of host memory");
  size_t bytes = (words > 0 ? words : 1) * 4;
  hipError_t e = hipMalloc((void**)&b->d, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(b->d, 0, bytes, dev->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(dev->stream);
  if (e != hipSuccess) {
    if (b->d) (void)hipFree(b->d);
    delete b;
    return hip_fail(e, "rsim_alloc_words");
  }
  ++dev->live;
  *out = b;
  return RSIM_OK;
}
int rsim_write_words(rsim_device* dev, rsim_words* dst, const uint32_t* host, uint64_t words) {
  int rc = use_outside_capture(dev, "rsim_write_words");
  if (rc) return rc;
  if (!dst || (!host && words)) return fail(RSIM_ERR_INVALID, "rsim_write_words: null argument");
  if (words > dst->n) return fail(RSIM_ERR_SHORT, "rsim_write_words: buffer too short");
  if (!words) return RSIM_OK;
  RSIM_HIP(hipMemcpyAsync(dst->d, host, words * 4, hipMemcpyHostToDevice, dev->stream), "word upload");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "word upload");
  return RSIM_OK;
}
int rsim_upload_words(rsim_device* dev, const uint32_t* host, uint64_t words, rsim_words** out) {
  int rc = rsim_alloc_words(dev, words, out);
  if (rc) return rc;
  rc = rsim_write_words(dev, *out, host, words);
  if (rc) {
    rsim_words_free(*out);
    *out = nullptr;
  }
  return rc;
}
int rsim_read_words(rsim_device* dev, const rsim_words* src, uint32_t* host, uint64_t words) {
  int rc = use_outside_capture(dev, "rsim_read_words");
  if (rc) return rc;
  if (!src || (!host && words)) return fail(RSIM_ERR_INVALID, "rsim_read_words: null argument");
  if (words > src->n) return fail(RSIM_ERR_SHORT, "rsim_read_words: buffer too short");
  if (!words) return RSIM_OK;
  RSIM_HIP(hipMemcpyAsync(host, src->d, words * 4, hipMemcpyDeviceToHost, dev->stream), "word download");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "word download");
  return RSIM_OK;
}
int rsim_upload_controls(rsim_device* dev, const void* controls, uint32_t n, rsim_words** out) {
  static_assert(sizeof(Controls) == 32, "Controls is eight words");
  return rsim_upload_words(dev, static_cast<const uint32_t*>(controls), (uint64_t)n * rs::NUM_CARS * 8, out);
}
int rsim_words_free(rsim_words* b) {
  if (!b) return RSIM_OK;
  if (b->view) {
    --b->dev->live;
    delete b;
    return RSIM_OK;
  }
  int rc = use_outside_capture(b->dev, "rsim_words_free");
  if (rc) return rc;
  (void)hipStreamSynchronize(b->dev->stream);
  (void)hipFree(b->d);
  --b->dev->live;
  delete b;
  return RSIM_OK;
}
int rsim_words_view(rsim_words* parent, uint64_t offset, uint64_t words, rsim_words** out) {
  if (!parent || !out) return fail(RSIM_ERR_INVALID, "rsim_words_view: null argument");
  if (offset > parent->n || words > parent->n - offset) return fail(RSIM_ERR_SHORT, "rsim_words_view: range past the buffer");
  rsim_words* v = new (std::nothrow) rsim_words{parent->dev, parent->d + offset, words, true};
  if (!v) return fail(RSIM_ERR_INVALID, "out of host memory");
  ++parent->dev->live;
  *out = v;
  return RSIM_OK;
}
uint64_t rsim_words_count(const rsim_words* b) { return b ? b->n : 0; }
void* rsim_words_device_ptr(rsim_words* b) { return b ? b->d : nullptr; }

// ---- kernels -----------------------------------------------------------------------------------------------------------------
int rsim_sim8_timed(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* result, uint32_t block, rsim_words* trace, rsim_words* cycles) {
  int rc = use(dev);
  if (rc) return rc;
  if (!states || !world || !controls || !result) return fail(RSIM_ERR_INVALID, "rsim_sim8: null argument");
  if (block < 1 || block > RSIM_MAX_BLOCK) return fail(RSIM_ERR_INVALID, "rsim_sim8: block outside 1..RSIM_MAX_BLOCK");
  const uint32_t n = states->n;
  if (controls->n < (uint64_t)n * rs::NUM_CARS * 8 || result->n < n) return fail(RSIM_ERR_SHORT, "rsim_sim8: buffers too short");
  if (variant && variant->n < n) return fail(RSIM_ERR_SHORT, "rsim_sim8: variant buffer too short");
  if (cycles && cycles->n < 2ull * n) return fail(RSIM_ERR_SHORT, "rsim_sim8: cycles buffer too short");
  if (trace && trace->n < (uint64_t)RSIM_TICKS_PER_DECISION * rs::dev::aosoa_words(n))
    return fail(RSIM_ERR_SHORT, "rsim_sim8: trace too short");
  if (!n) return RSIM_OK;
  rs::dev::Sim8Args p{states->d, world->d, world->count, variant ? variant->d : nullptr,
                      reinterpret_cast<const Controls*>(controls->d), result->d, n, trace ? trace->d : nullptr, nullptr};
  if (cycles) p.cycles = reinterpret_cast<uint64_t*>(cycles->d);
  RSIM_HIP(rs::dev::launch_sim8(dev->stream, p, block), "sim8 launch");
  return RSIM_OK;
}
int rsim_sim8_variants(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                       const rsim_words* controls, rsim_words* result, uint32_t block, rsim_words* trace) {
  return rsim_sim8_timed(dev, states, world, variant, controls, result, block, trace, nullptr);
}
int rsim_sim8(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* controls, rsim_words* result,
              uint32_t block, rsim_words* trace) {
  return rsim_sim8_timed(dev, states, world, nullptr, controls, result, block, trace, nullptr);
}
int rsim_run(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant, const rsim_words* controls,
             rsim_words* results, uint32_t decisions, uint32_t block) {
  int rc = use(dev);
  if (rc) return rc;
  if (!states || !world || !controls || !results) return fail(RSIM_ERR_INVALID, "rsim_run: null argument");
  if (block < 1 || block > RSIM_MAX_BLOCK) return fail(RSIM_ERR_INVALID, "rsim_run: block outside 1..RSIM_MAX_BLOCK");
  const uint32_t n = states->n;
  if (controls->n < (uint64_t)decisions * n * rs::NUM_CARS * 8 || results->n < (uint64_t)decisions * n)
    return fail(RSIM_ERR_SHORT, "rsim_run: buffers too short");
  if (variant && variant->n < n) return fail(RSIM_ERR_SHORT, "rsim_run: variant buffer too short");
  if (!n || !decisions) return RSIM_OK;
  rs::dev::Sim8Args p{states->d, world->d, world->count, variant ? variant->d : nullptr,
                      reinterpret_cast<const Controls*>(controls->d), results->d, n, nullptr, nullptr};
  RSIM_HIP(rs::dev::launch_run(dev->stream, p, decisions, block), "run launch");
  return RSIM_OK;
}
int rsim_run_queued(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* results, uint32_t decisions, uint32_t block, const rsim_words* order,
                    uint32_t workgroups) {
  int rc = use(dev);
  if (rc) return rc;
  if (!states || !world || !controls || !results) return fail(RSIM_ERR_INVALID, "rsim_run_queued: null argument");
  if (block < 64 || block > RSIM_MAX_BLOCK || block % 64)
    return fail(RSIM_ERR_INVALID, "rsim_run_queued: block must be 64, 128, 192 or 256 (whole waves)");
  const uint32_t n = states->n, nblocks = (n + 63) / 64;
  if (controls->n < (uint64_t)decisions * n * rs::NUM_CARS * 8 || results->n < (uint64_t)decisions * n)
    return fail(RSIM_ERR_SHORT, "rsim_run_queued: buffers too short");
  if (variant && variant->n < n) return fail(RSIM_ERR_SHORT, "rsim_run_queued: variant buffer too short");
  if (order && order->n < nblocks) return fail(RSIM_ERR_SHORT, "rsim_run_queued: order shorter than the block count");
  if (!n || !decisions) return RSIM_OK;
  if (!workgroups) workgroups = rs::dev::run_queue_capacity(block);
  const uint32_t waves = block / 64;
  const uint32_t useful = (nblocks + waves - 1) / waves;  // more work-groups than that would find the queue empty
  if (!workgroups || workgroups > useful) workgroups = useful;
  rs::dev::Sim8Args p{states->d, world->d, world->count, variant ? variant->d : nullptr,
                      reinterpret_cast<const Controls*>(controls->d), results->d, n, nullptr, nullptr};
  RSIM_HIP(rs::dev::launch_run_queue(dev->stream, p, decisions, block, order ? order->d : nullptr, dev->queue, workgroups),
           "run_queued launch");
  return RSIM_OK;
}
int rsim_sim8_split(rsim_device* dev, rsim_states* states, const rsim_world* world, const rsim_words* variant,
                    const rsim_words* controls, rsim_words* result, uint32_t block, uint32_t ticks_per_launch) {
  int rc = use(dev);
  if (rc) return rc;
  if (!states || !world || !controls || !result) return fail(RSIM_ERR_INVALID, "rsim_sim8_split: null argument");
  if (block < 1 || block > RSIM_MAX_BLOCK) return fail(RSIM_ERR_INVALID, "rsim_sim8_split: block outside 1..RSIM_MAX_BLOCK");
  if (ticks_per_launch < 1 || ticks_per_launch > RSIM_TICKS_PER_DECISION)
    return fail(RSIM_ERR_INVALID, "rsim_sim8_split: ticks_per_launch outside 1..8");
  const uint32_t n = states->n;
  if (controls->n < (uint64_t)n * rs::NUM_CARS * 8 || result->n < n) return fail(RSIM_ERR_SHORT, "rsim_sim8_split: buffers too short");
  if (variant && variant->n < n) return fail(RSIM_ERR_SHORT, "rsim_sim8_split: variant buffer too short");
  if (!n) return RSIM_OK;
  rs::dev::Sim8Args p{states->d, world->d, world->count, variant ? variant->d : nullptr,
                      reinterpret_cast<const Controls*>(controls->d), result->d, n, nullptr, nullptr};
  RSIM_HIP(rs::dev::launch_sim8_ticks(dev->stream, p, block, ticks_per_launch), "sim8_split launch");
  return RSIM_OK;
}
uint32_t rsim_run_queue_capacity(rsim_device* dev, uint32_t block) {
  if (use(dev) || block < 64 || block > RSIM_MAX_BLOCK || block % 64) return 0;
  return rs::dev::run_queue_capacity(block);
}
// ---- the cost model and regrouping -------------------------------------------------------------------------------------------
int rsim_cost_features(rsim_device* dev, const rsim_states* states, rsim_words* features) {
  int rc = use(dev);
  if (rc) return rc;
  if (!states || !features) return fail(RSIM_ERR_INVALID, "rsim_cost_features: null argument");
  if (features->n < (uint64_t)states->n * RSIM_COST_FEATURES) return fail(RSIM_ERR_SHORT, "rsim_cost_features: buffer too short");
  if (!states->n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_features(dev->stream, states->d, states->n, reinterpret_cast<float*>(features->d)), "features");
  return RSIM_OK;
}
int rsim_cost_predict(rsim_device* dev, const rsim_words* params, const rsim_words* features, uint32_t n, rsim_words* out) {
  int rc = use(dev);
  if (rc) return rc;
  if (!params || !features || !out) return fail(RSIM_ERR_INVALID, "rsim_cost_predict: null argument");
  if (params->n < RSIM_COST_PARAMS || features->n < (uint64_t)n * RSIM_COST_FEATURES || out->n < 2ull * n)
    return fail(RSIM_ERR_SHORT, "rsim_cost_predict: buffers too short");
  if (!n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_mlp_infer(dev->stream, reinterpret_cast<const float*>(params->d), reinterpret_cast<const float*>(features->d),
                                     n, reinterpret_cast<float*>(out->d)),
           "cost inference");
  return RSIM_OK;
}
int rsim_cost_gradients(rsim_device* dev, const rsim_words* params, const rsim_words* features, const rsim_words* target,
                        const rsim_words* contact, uint32_t batch, rsim_words* scratch, rsim_words* loss, rsim_words* grad) {
  int rc = use(dev);
  if (rc) return rc;
  if (!params || !features || !target || !contact || !scratch || !loss || !grad) return fail(RSIM_ERR_INVALID, "rsim_cost_gradients: null argument");
  if (params->n < RSIM_COST_PARAMS || grad->n < RSIM_COST_PARAMS || features->n < (uint64_t)batch * RSIM_COST_FEATURES ||
      target->n < batch || contact->n < batch || loss->n < batch || scratch->n < (uint64_t)batch * rs::dev::COST_ACT_WORDS)
    return fail(RSIM_ERR_SHORT, "rsim_cost_gradients: buffers too short");
  if (!batch) return fail(RSIM_ERR_INVALID, "rsim_cost_gradients: empty batch");
  RSIM_HIP(rs::dev::launch_mlp_train(dev->stream, reinterpret_cast<const float*>(params->d), reinterpret_cast<const float*>(features->d),
                                     reinterpret_cast<const float*>(target->d), reinterpret_cast<const float*>(contact->d), batch,
                                     reinterpret_cast<float*>(scratch->d), reinterpret_cast<float*>(loss->d),
                                     reinterpret_cast<float*>(grad->d)),
           "cost gradients");
  return RSIM_OK;
}
int rsim_cost_adam(rsim_device* dev, rsim_words* params, rsim_words* m, rsim_words* v, const rsim_words* grad, float lr, uint32_t step) {
  int rc = use(dev);
  if (rc) return rc;
  if (!params || !m || !v || !grad || step == 0) return fail(RSIM_ERR_INVALID, "rsim_cost_adam: null argument or step 0");
  if (params->n < RSIM_COST_PARAMS || m->n < RSIM_COST_PARAMS || v->n < RSIM_COST_PARAMS || grad->n < RSIM_COST_PARAMS)
    return fail(RSIM_ERR_SHORT, "rsim_cost_adam: buffers too short");
  float bc1 = 1.0f - powf(0.9f, (float)step), bc2 = 1.0f - powf(0.999f, (float)step);
  RSIM_HIP(rs::dev::launch_adam(dev->stream, reinterpret_cast<float*>(params->d), reinterpret_cast<float*>(m->d),
                                reinterpret_cast<float*>(v->d), reinterpret_cast<const float*>(grad->d), lr, bc1, bc2),
           "adam");
  return RSIM_OK;
}
namespace {
bool is_permutation(const uint32_t* perm, uint32_t n) {
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    if (perm[i] >= n || seen[perm[i]]) return false;
    seen[perm[i]] = 1;
  }
  return true;
}
}  // namespace
int rsim_set_permutation(rsim_device* dev, const uint32_t* perm, uint32_t n, rsim_words* dst) {
  if (!perm && n) return fail(RSIM_ERR_INVALID, "rsim_set_permutation: null permutation");
  if (!is_permutation(perm, n)) return fail(RSIM_ERR_INVALID, "rsim_set_permutation: not a permutation of 0..n-1");
  return rsim_write_words(dev, dst, perm, n);
}
int rsim_gather_states(rsim_device* dev, const rsim_states* src, const rsim_words* perm, rsim_states* dst) {
  int rc = use(dev);
  if (rc) return rc;
  if (!src || !perm || !dst || src == dst) return fail(RSIM_ERR_INVALID, "rsim_gather_states: bad buffers");
  if (perm->n < dst->n) return fail(RSIM_ERR_SHORT, "rsim_gather_states: index list too short");
  if (!dst->n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_gather_states(dev->stream, src->d, src->n, perm->d, dst->d, dst->n), "gather states");
  return RSIM_OK;
}
uint64_t rsim_cost_order_work_words(uint32_t n) { return rs::dev::order_work_words(n); }
int rsim_cost_order(rsim_device* dev, const rsim_words* pred, uint32_t n, int mode, rsim_words* work, rsim_words* perm) {
  int rc = use(dev);
  if (rc) return rc;
  if (!pred || !work || !perm) return fail(RSIM_ERR_INVALID, "rsim_cost_order: null argument");
  if (mode != RSIM_ORDER_SORTED && mode != RSIM_ORDER_BALANCED) return fail(RSIM_ERR_INVALID, "rsim_cost_order: unknown mode");
  if (pred->n < 2ull * n || perm->n < n || work->n < rs::dev::order_work_words(n))
    return fail(RSIM_ERR_SHORT, "rsim_cost_order: buffers too short");
  if (!n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_cost_order(dev->stream, reinterpret_cast<const float*>(pred->d), n, mode, work->d, perm->d),
           "cost order");
  return RSIM_OK;
}
uint64_t rsim_cost_block_order_work_words(uint32_t n) { return rs::dev::block_order_work_words(n); }
int rsim_cost_block_order(rsim_device* dev, const rsim_words* pred, uint32_t n, rsim_words* work, rsim_words* order) {
  int rc = use(dev);
  if (rc) return rc;
  if (!pred || !work || !order) return fail(RSIM_ERR_INVALID, "rsim_cost_block_order: null argument");
  if (pred->n < 2ull * n || order->n < (n + 63) / 64 || work->n < rs::dev::block_order_work_words(n))
    return fail(RSIM_ERR_SHORT, "rsim_cost_block_order: buffers too short");
  if (!n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_block_order(dev->stream, reinterpret_cast<const float*>(pred->d), n, work->d, order->d), "block order");
  return RSIM_OK;
}
int rsim_cost_order_noalloc(rsim_device* dev, const rsim_words* pred, uint32_t n, int mode, rsim_words* perm) {
  int rc = use(dev);
  if (rc) return rc;
  if (!pred || !perm) return fail(RSIM_ERR_INVALID, "rsim_cost_order_noalloc: null argument");
  if (mode != RSIM_ORDER_SORTED && mode != RSIM_ORDER_BALANCED)
    return fail(RSIM_ERR_INVALID, "rsim_cost_order_noalloc: unknown mode");
  if (n > RSIM_COST_ORDER_NOALLOC_MAX)
    return fail(RSIM_ERR_INVALID, "rsim_cost_order_noalloc: n is above RSIM_COST_ORDER_NOALLOC_MAX (262,144 arenas)");
  if (pred->n < 2ull * n || perm->n < n) return fail(RSIM_ERR_SHORT, "rsim_cost_order_noalloc: buffers too short");
  if (!n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_cost_order(dev->stream, reinterpret_cast<const float*>(pred->d), n, mode, dev->order_scratch,
                                      perm->d),
           "cost order (fixed workspace)");
  return RSIM_OK;
}
int rsim_permute_words(rsim_device* dev, const rsim_words* src, const rsim_words* perm, uint32_t n, uint32_t words_per_item,
                       int scatter, rsim_words* dst) {
  int rc = use(dev);
  if (rc) return rc;
  if (!src || !perm || !dst || src == dst) return fail(RSIM_ERR_INVALID, "rsim_permute_words: bad buffers");
  if (perm->n < n || src->n < (uint64_t)n * words_per_item || dst->n < (uint64_t)n * words_per_item)
    return fail(RSIM_ERR_SHORT, "rsim_permute_words: buffers too short");
  if (!n || !words_per_item) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_gather_words(dev->stream, src->d, perm->d, dst->d, n, words_per_item, scatter != 0), "permute words");
  return RSIM_OK;
}
// ---- HIP graphs ------------------------------------------------------------------------------------------------------------
int rsim_capture_begin(rsim_device* dev) {
  int rc = use(dev);
  if (rc) return rc;
  RSIM_HIP(hipStreamBeginCapture(dev->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return RSIM_OK;
}
int rsim_capture_end(rsim_device* dev, rsim_graph** out) {
  int rc = use(dev);
  if (rc) return rc;
  if (!out) return fail(RSIM_ERR_INVALID, "rsim_capture_end: null argument");
  *out = nullptr;
  hipGraph_t g = nullptr;
  RSIM_HIP(hipStreamEndCapture(dev->stream, &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  rsim_graph* gr = new (std::nothrow) rsim_graph{dev, g, x};
  if (!gr) {
    (void)hipGraphExecDestroy(x);
    (void)hipGraphDestroy(g);
    return fail(RSIM_ERR_INVALID, "out of host memory");
  }
  ++dev->live;
  *out = gr;
  return RSIM_OK;
}
int rsim_graph_launch(rsim_device* dev, const rsim_graph* graph) {
  int rc = use(dev);
  if (rc) return rc;
  if (!graph || graph->dev != dev) return fail(RSIM_ERR_INVALID, "rsim_graph_launch: graph of another device");
  RSIM_HIP(hipGraphLaunch(graph->exec, dev->stream), "hipGraphLaunch");
  return RSIM_OK;
}
int rsim_graph_free(rsim_graph* graph) {
  if (!graph) return RSIM_OK;
  int rc = use_outside_capture(graph->dev, "rsim_graph_free");
  if (rc) return rc;
  (void)hipStreamSynchronize(graph->dev->stream);
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  --graph->dev->live;
  delete graph;
  return RSIM_OK;
}
int rsim_identity(rsim_device* dev, const rsim_states* in, rsim_states* out) {
  int rc = use_outside_capture(dev, "rsim_identity");
  if (rc) return rc;
  if (!in || !out || in->n != out->n) return fail(RSIM_ERR_INVALID, "rsim_identity: buffers of different sizes");
  if (!in->n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_identity(dev->stream, in->d, out->d, in->n, false), "identity launch");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "identity");
  return RSIM_OK;
}
int rsim_identity_tick(rsim_device* dev, const rsim_states* in, rsim_states* out) {
  int rc = use_outside_capture(dev, "rsim_identity_tick");
  if (rc) return rc;
  if (!in || !out || in->n != out->n) return fail(RSIM_ERR_INVALID, "rsim_identity_tick: buffers of different sizes");
  if (!in->n) return RSIM_OK;
  RSIM_HIP(rs::dev::launch_identity(dev->stream, in->d, out->d, in->n, true), "identity_tick launch");
  RSIM_HIP(hipStreamSynchronize(dev->stream), "identity_tick");
  return RSIM_OK;
}
int rsim_phase(rsim_device* dev, int k, rsim_states* states, const rsim_world* world, int32_t* status_out) {
  int rc = use_outside_capture(dev, "rsim_phase");
  if (rc) return rc;
  if (!states || !world) return fail(RSIM_ERR_INVALID, "rsim_phase: null argument");
  if (!states->n) return RSIM_OK;
  if ((rc = ensure_status(dev, states->n))) return rc;
  RSIM_HIP(rs::dev::launch_phase(dev->stream, k, states->d, world->d, states->n, dev->status), "phase launch");
  return finish_status(dev, states->n, status_out);
}
int rsim_step_tick(rsim_device* dev, rsim_states* states, const rsim_world* world, int32_t* status_out) {
  int rc = use_outside_capture(dev, "rsim_step_tick");
  if (rc) return rc;
  if (!states || !world) return fail(RSIM_ERR_INVALID, "rsim_step_tick: null argument");
  if (!states->n) return RSIM_OK;
  if ((rc = ensure_status(dev, states->n))) return rc;
  RSIM_HIP(rs::dev::launch_step_tick(dev->stream, states->d, world->d, states->n, dev->status), "step_tick launch");
  return finish_status(dev, states->n, status_out);
}

}  // extern "C"

namespace rs {
namespace host {
// The calling thread's rsim_last_error() message, for the host-only parts of the library in other files.
int set_error(int code, const char* what) { return fail(code, what); }
}  // namespace host
}  // namespace rs
