// vlhip_runtime.hip -- context, scratch arenas, copies and kernel timing.
#include "vlhip_runtime.h"
#include "vlhip.h"

extern "C" {
typedef int (*vl_printf_func_t)(char const* format, ...);
vl_printf_func_t vl_get_printf_func(void);
}

namespace vlhip {

static Context g_ctx;

Context& ctx() { return g_ctx; }

[[noreturn]] void fatal(const char* what, const char* file, int line, const char* detail) {
  vl_printf_func_t pf = vl_get_printf_func();
  if (pf) {
    pf("vlhip: fatal: %s failed at %s:%d: %s\n", what, file, line, detail ? detail : "");
    std::fflush(stdout);  // the hook's text must not stay in a buffer at abort()
  }
  std::fprintf(stderr, "vlhip: fatal: %s failed at %s:%d: %s\n", what, file, line, detail ? detail : "");
  std::fflush(stderr);
  std::abort();
}

void ensure_init() {
  Context& c = g_ctx;
  if (c.initialised) return;
  std::lock_guard<std::recursive_mutex> lk(c.mu);
  if (c.initialised) return;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    fatal("hipGetDeviceCount", __FILE__, __LINE__,
          "no HIP device available: this library runs its GMM/Fisher/VLAD kernels on an AMD GPU "
          "and has no CPU fallback");
  }
  VLHIP_CHECK(hipSetDevice(c.device));
  hipDeviceProp_t prop;
  VLHIP_CHECK(hipGetDeviceProperties(&prop, c.device));
  c.num_cus = prop.multiProcessorCount;
  c.initialised = true;
}

void* scratch(Slot s, size_t bytes) {
  Context& c = g_ctx;
  // a captured graph keeps the pointers it was recorded with, and a library
  // buffer may be freed and reallocated by any later call
  if (c.capturing)
    fatal("scratch", __FILE__, __LINE__, "a library work buffer was requested while an encode plan is captured");
  if (bytes == 0) bytes = 256;
  if (c.cap[s] < bytes) {
    if (c.buf[s]) {
      VLHIP_CHECK(hipStreamSynchronize(c.stream));
      VLHIP_CHECK(hipFree(c.buf[s]));
      c.buf[s] = nullptr;
      c.cap[s] = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom so slowly growing sizes do not thrash
    want = (want + 4095) & ~size_t(4095);
    VLHIP_CHECK(hipMalloc(&c.buf[s], want));
    c.cap[s] = want;
  }
  return c.buf[s];
}

void upload(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  VLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_ctx.stream));
}

void download(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return;
  VLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
  VLHIP_CHECK(hipStreamSynchronize(g_ctx.stream));
}

void sync() { VLHIP_CHECK(hipStreamSynchronize(g_ctx.stream)); }

static hipEvent_t take_event() {
  Context& c = g_ctx;
  if (!c.event_pool.empty()) {
    hipEvent_t e = c.event_pool.back();
    c.event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  VLHIP_CHECK(hipEventCreate(&e));
  return e;
}

// Resolve every recorded start/stop pair into the per-kernel totals.
static void resolve_pending() {
  Context& c = g_ctx;
  for (auto& p : c.pending) {
    VLHIP_CHECK(hipEventSynchronize(p.stop));
    float ms = 0.f;
    VLHIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
    c.timings[p.timing_index].total_ms += ms;
    c.timings[p.timing_index].count += 1;
    c.event_pool.push_back(p.start);
    c.event_pool.push_back(p.stop);
  }
  c.pending.clear();
}

int timing_begin(const char* name, hipEvent_t* start_out) {
  Context& c = g_ctx;
  if (!c.timing) return -1;
  int idx = -1;
  for (size_t i = 0; i < c.timings.size(); ++i)
    if (c.timings[i].name == name) { idx = (int)i; break; }
  if (idx < 0) {
    c.timings.push_back(KernelTiming{name, 0.0, 0});
    idx = (int)c.timings.size() - 1;
  }
  hipEvent_t s = take_event();
  VLHIP_CHECK(hipEventRecord(s, c.stream));
  *start_out = s;
  if (c.pending.size() > 4096) resolve_pending();
  return idx;
}

void timing_end(int index, hipEvent_t start) {
  Context& c = g_ctx;
  hipEvent_t e = take_event();
  VLHIP_CHECK(hipEventRecord(e, c.stream));
  c.pending.push_back(PendingEvent{index, start, e});
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

int vlhip_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

void vlhip_init(void) { ensure_init(); }

void vlhip_set_stream(void* stream) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  ctx().stream = (hipStream_t)stream;
}

void* vlhip_malloc(size_t bytes) {
  ensure_init();
  void* p = nullptr;
  VLHIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
  return p;
}

void vlhip_free(void* ptr) {
  if (!ptr) return;
  ensure_init();
  VLHIP_CHECK(hipStreamSynchronize(ctx().stream));
  VLHIP_CHECK(hipFree(ptr));
}

void vlhip_memcpy_h2d(void* dst, void const* src, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  upload(dst, src, bytes);
  sync();
}

void vlhip_memcpy_d2h(void* dst, void const* src, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  download(dst, src, bytes);
}

void vlhip_memcpy_h2d_async(void* dst, void const* src, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  upload(dst, src, bytes);
}

void vlhip_memcpy_d2h_async(void* dst, void const* src, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (bytes == 0) return;
  VLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
}

void vlhip_memcpy_d2d(void* dst, void const* src, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (bytes == 0) return;
  VLHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx().stream));
}

void vlhip_memset(void* dst, int value, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_CHECK(hipMemsetAsync(dst, value, bytes, ctx().stream));
}

void vlhip_synchronize(void) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  sync();
}

void* vlhip_buffer(int which, size_t bytes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  static const Slot map[] = {SLOT_DATA, SLOT_PARAMS, SLOT_POST, SLOT_ENC, SLOT_MISC, SLOT_MISC2};
  if (which < 0 || which >= (int)(sizeof(map) / sizeof(map[0])))
    fatal("vlhip_buffer", __FILE__, __LINE__, "unknown buffer index");
  return scratch(map[which], bytes);
}

void vlhip_lock(void) {
  ensure_init();
  ctx().mu.lock();
}

void vlhip_unlock(void) { ctx().mu.unlock(); }

int vlhip_num_compute_units(void) {
  ensure_init();
  return ctx().num_cus;
}

void vlhip_timing_enable(int on) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  ctx().timing = (on != 0);
}

void vlhip_timing_reset(void) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  resolve_pending();
  for (auto& t : ctx().timings) { t.total_ms = 0.0; t.count = 0; }
}

int vlhip_timing_count(void) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  resolve_pending();
  return (int)ctx().timings.size();
}

int vlhip_timing_get(int index, char* name, size_t cap, double* total_ms, unsigned long long* count) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  resolve_pending();
  if (index < 0 || index >= (int)ctx().timings.size()) return -1;
  const KernelTiming& t = ctx().timings[index];
  if (name && cap) {
    std::snprintf(name, cap, "%s", t.name.c_str());
  }
  if (total_ms) *total_ms = t.total_ms;
  if (count) *count = t.count;
  return 0;
}

}  // extern "C"
