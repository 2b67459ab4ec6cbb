// vlhip_runtime.h -- device context, error handling, scratch arenas and
// per-kernel event timing shared by every translation unit of the shim.
//
// One process-wide context per device: a stream (the legacy null stream unless
// the caller installs another one), a set of grow-only scratch buffers, and an
// optional timing log that brackets every kernel launch with hipEvents.  All
// entry points take the context mutex, so calls from different host threads
// serialise (the reference library is thread-safe per call, gmm.c/fisher.c
// keep no global state).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace vlhip {

[[noreturn]] void fatal(const char* what, const char* file, int line, const char* detail);

#define VLHIP_CHECK(expr)                                                        \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) ::vlhip::fatal(#expr, __FILE__, __LINE__, hipGetErrorString(_e)); \
  } while (0)

#define VLHIP_REQUIRE(cond, msg)                                                 \
  do { if (!(cond)) ::vlhip::fatal(#cond, __FILE__, __LINE__, (msg)); } while (0)

// Scratch slots: each is a grow-only device buffer owned by the context.
// SLOT_DATA, SLOT_PARAMS, SLOT_POST, SLOT_ENC, SLOT_MISC and SLOT_MISC2 are the
// buffers vlhip_buffer hands to callers, who pass them back as kernel
// arguments: kernels' own work buffers must never use them (growing one frees
// it under the caller's pointer).
enum Slot {
  SLOT_DATA = 0,      // caller: uploaded descriptors (host entry points)
  SLOT_PARAMS,        // caller: uploaded means / covariances / priors
  SLOT_PREP,          // per-component derived terms (inverse covariances, constants)
  SLOT_POST,          // caller: N x K posteriors
  SLOT_ROWSTAT,       // per-descriptor (max logit, sum) and argmax
  SLOT_PARTIAL,       // per-block partial sums (LL, M-step moments)
  SLOT_REDUCED,       // reduced M-step moments (FP64)
  SLOT_ENC,           // caller: encodings
  SLOT_MISC,          // caller: offsets, counters, assignments
  SLOT_MISC2,         // caller
  SLOT_SEGMENTS,      // per-segment sums of the FP32 Fisher encoder's parallel-segment mode
  SLOT_EMCHUNKS,      // chunk offsets of the matrix-core EM step
  SLOT_EMLL,          // log-likelihood partials of the matrix-core EM step
  SLOT_QTAB,          // split centre table of the matrix-core-screened quantizer
  SLOT_GATE,          // learned-gate activations / candidate lists
  SLOT_GATE2,
  SLOT_FISHER_WORK,   // per-batch work buffer of the Fisher encoders (posteriors or operand records)
  SLOT_GROUPS,        // per-image first group index of the matrix-core Fisher encoder
  SLOT_GROUPTAB,      // per-group first row and image end of the matrix-core Fisher encoder
  SLOT_GATELIST,      // posteriors of gate-served groups (component, q) of the matrix-core Fisher encoder
  SLOT_FUSED,         // partner exchange slots and abort flag of the fused matrix-core encoder
  SLOT_STREAMS,       // per-stream image lists of the fused matrix-core encoder
  SLOT_COUNT
};

struct KernelTiming {
  std::string name;
  double total_ms = 0.0;
  uint64_t count = 0;
};

struct PendingEvent {
  int timing_index;
  hipEvent_t start, stop;
};

struct Context {
  std::recursive_mutex mu;
  bool initialised = false;
  int device = 0;
  hipStream_t stream = nullptr;     // nullptr = legacy null stream
  void* buf[SLOT_COUNT] = {};
  size_t cap[SLOT_COUNT] = {};
  bool timing = false;
  bool capturing = false;  // an encode plan's graph is being captured on its stream
  std::vector<KernelTiming> timings;
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> event_pool;
  int num_cus = 0;
  void* gate = nullptr;  // learned gate of VLHIP_MODE_GATED
  int gate_m = 16;
  float gate_need = 0.99f;
  unsigned long long gate_groups = 0;  // descriptor groups seen by gated encodings
  unsigned long long gate_served = 0;  // of which served by the gate
};

Context& ctx();
void ensure_init();
void* scratch(Slot s, size_t bytes);
void upload(void* dst, const void* src, size_t bytes);
void download(void* dst, const void* src, size_t bytes);
void sync();

// Timing: begin/end bracket one kernel launch on the context stream.
int timing_begin(const char* name, hipEvent_t* start_out);
void timing_end(int index, hipEvent_t start);

}  // namespace vlhip

// Launch helper: checks the launch and, when timing is on, records events.
#define VLHIP_LAUNCH(name, kernel, grid, block, shmem, ...)                      \
  do {                                                                           \
    hipEvent_t _ts = nullptr;                                                    \
    int _ti = ::vlhip::timing_begin(name, &_ts);                                 \
    hipLaunchKernelGGL(kernel, grid, block, shmem, ::vlhip::ctx().stream, __VA_ARGS__); \
    VLHIP_CHECK(hipGetLastError());                                              \
    if (_ti >= 0) ::vlhip::timing_end(_ti, _ts);                                 \
  } while (0)
