// Handle, workspace arena, per-kernel event timer and error reporting.
#include <mutex>

#include "common.hpp"

namespace remifa {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

// Workspace: one device buffer owned by the handle, grown on demand (the
// stream is drained first, since queued kernels may still use the old one),
// or a caller-supplied buffer that is never reallocated.
int ensure_workspace(remifa_handle_t h, size_t bytes) {
  if (bytes <= h->ws_bytes) return 0;
  if (h->ws_user) {
    set_last_error("workspace too small: need " + std::to_string(bytes) + " bytes, have " +
                   std::to_string(h->ws_bytes));
    return REMIFA_ERR_WORKSPACE;
  }
  if (h->ws) {
    RF_HIP(hipStreamSynchronize(h->stream));
    RF_HIP(hipFree(h->ws));
    h->ws = nullptr;
    h->ws_bytes = 0;
  }
  size_t sz = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
  hipError_t e = hipMalloc(&h->ws, sz);
  if (e != hipSuccess) {
    h->ws = nullptr;
    set_last_error("hipMalloc of " + std::to_string(sz) + " bytes failed: " + hipGetErrorString(e));
    return REMIFA_ERR_ALLOC;
  }
  h->ws_bytes = sz;
  return 0;
}

// -------------------------------------------------------------- timer
static int timer_event(remifa_handle_t h) {
  KernelTimer& t = h->timer;
  if (t.next == t.pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    t.pool.push_back(e);
  }
  return (int)t.next++;
}

void timer_begin(remifa_handle_t h, int cls, double work, int variant) {
  if (!h->timer.enabled) return;
  int e0 = timer_event(h), e1 = timer_event(h);
  if (e0 < 0 || e1 < 0) return;
  hipEventRecord(h->timer.pool[e0], h->stream);
  h->timer.recs.push_back({cls, e0, e1, work, variant});
}

void timer_end(remifa_handle_t h) {
  if (!h->timer.enabled || h->timer.recs.empty()) return;
  hipEventRecord(h->timer.pool[h->timer.recs.back().ev1], h->stream);
}

}  // namespace remifa

using namespace remifa;

extern "C" {

const char* remifa_last_error(void) { return g_last_error.c_str(); }

int remifa_create(remifa_handle_t* handle, void* stream) {
  if (!handle) return -1;
  *handle = nullptr;
  auto* h = new (std::nothrow) remifa_handle_st();
  if (!h) return REMIFA_ERR_ALLOC;
  h->stream = (hipStream_t)stream;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    set_last_error("hipGetDevice failed");
    return REMIFA_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) h->num_cu = prop.multiProcessorCount;
  // the no-allocation solve's scratch on this device (a failure is reported
  // by that entry point, not here)
  if (noalloc_scratch(h->device, &h->scratch) != 0) h->scratch = nullptr;
  *handle = h;
  return 0;
}

int remifa_destroy(remifa_handle_t h) {
  if (!h) return -1;
  if (h->stream) hipStreamSynchronize(h->stream);
  else hipDeviceSynchronize();
  if (h->ws && !h->ws_user) hipFree(h->ws);
  for (auto e : h->timer.pool) hipEventDestroy(e);
  if (h->side) {
    hipStreamSynchronize(h->side);
    hipStreamDestroy(h->side);
  }
  if (h->ev_main) hipEventDestroy(h->ev_main);
  if (h->ev_side) hipEventDestroy(h->ev_side);
  delete h;
  return 0;
}

int remifa_set_stream(remifa_handle_t h, void* stream) {
  if (!h) return -1;
  h->stream = (hipStream_t)stream;
  return 0;
}

int remifa_get_stream(remifa_handle_t h, void** stream) {
  if (!h) return -1;
  if (!stream) return -2;
  *stream = (void*)h->stream;
  return 0;
}

int remifa_set_workspace(remifa_handle_t h, void* d_ws, size_t bytes) {
  if (!h) return -1;
  if (!d_ws && bytes) return -2;
  if (h->ws && !h->ws_user) {
    RF_HIP(hipStreamSynchronize(h->stream));
    RF_HIP(hipFree(h->ws));
  }
  h->ws = d_ws;
  h->ws_bytes = d_ws ? bytes : 0;
  h->ws_user = d_ws != nullptr;
  return 0;
}

int remifa_set_tuning(remifa_handle_t h, int panel_width, int split_k) {
  if (!h) return -1;
  if (panel_width < 0) return -2;
  if (split_k < 0 || split_k > 64) return -3;
  h->panel_width = panel_width;
  h->split_k = split_k;
  return 0;
}

int remifa_set_block_kernels(remifa_handle_t h, int one_launch, int units_per_wave,
                             int64_t wait_polls) {
  if (!h) return -1;
  if (one_launch != 0 && one_launch != 1) return -2;
  if (units_per_wave < 0 || units_per_wave > 64) return -3;
  if (wait_polls < 0) return -4;
  h->blk_fused = one_launch;
  h->blk_upw = units_per_wave;
  h->blk_wait_polls = wait_polls;
  return 0;
}

int remifa_set_gemm_rules(remifa_handle_t h, int count, const int64_t* rules) {
  if (!h) return -1;
  if (count < 0 || count > 4096) return -2;
  if (count > 0 && !rules) return -3;
  std::vector<std::array<int64_t, 7>> r((size_t)count);
  for (int i = 0; i < count; ++i) {
    for (int j = 0; j < 7; ++j) r[i][j] = rules[7 * i + j];
    if (r[i][5] < -1 || r[i][5] > 7 || r[i][6] < 0 || r[i][6] > 64) return -3;
  }
  h->gemm_rules.swap(r);
  return 0;
}

int remifa_set_gemm_tile(remifa_handle_t h, int tile) {
  if (!h) return -1;
  if (tile < -1 || tile > 7) return -2;
  h->gemm_tile = tile;
  return 0;
}

int remifa_set_lookahead(remifa_handle_t h, int on, int reserve_cus) {
  if (!h) return -1;
  if (on != 0 && on != 1) return -2;
  // the mask (drivers.hip, side_stream) takes reserve / 8 CUs of each of the
  // 8 XCDs of the 256-CU layout: at most 8 per XCD
  if (reserve_cus < 0 || reserve_cus > 64 || reserve_cus % 8) return -3;
  if (reserve_cus > 0 && h->num_cu != 256) return -3;
  h->lookahead = on;
  h->side_reserve = reserve_cus;
  return 0;
}

int remifa_get_lookahead(remifa_handle_t h, int* on, int* reserve_cus) {
  if (!h) return -1;
  if (on) *on = h->lookahead;
  if (reserve_cus) *reserve_cus = h->side_reserve;
  return 0;
}

int remifa_timing_enable(remifa_handle_t h, int enable) {
  if (!h) return -1;
  h->timer.enabled = enable != 0;
  return 0;
}

int remifa_timing_reset(remifa_handle_t h) {
  if (!h) return -1;
  RF_HIP(hipStreamSynchronize(h->stream));
  h->timer.recs.clear();
  h->timer.next = 0;
  return 0;
}

int remifa_timing_query(remifa_handle_t h, int cls, double* ms, int64_t* launches, double* work) {
  if (!h) return -1;
  if (cls < 0 || cls >= KC_COUNT) return -2;
  RF_HIP(hipStreamSynchronize(h->stream));
  double tot = 0.0, w = 0.0;
  int64_t cnt = 0;
  for (auto& r : h->timer.recs) {
    if (r.cls != cls) continue;
    float t = 0.0f;
    RF_HIP(hipEventElapsedTime(&t, h->timer.pool[r.ev0], h->timer.pool[r.ev1]));
    tot += t;
    w += r.work;
    ++cnt;
  }
  if (ms) *ms = tot;
  if (launches) *launches = cnt;
  if (work) *work = w;
  return 0;
}

int remifa_timing_records(remifa_handle_t h, int64_t max, int* cls, int* variant, double* ms,
                          double* work, int64_t* count) {
  if (!h) return -1;
  if (max < 0) return -2;
  if (!count) return -8;
  RF_HIP(hipStreamSynchronize(h->stream));
  int64_t k = 0;
  for (auto& r : h->timer.recs) {
    if (k < max) {
      float t = 0.0f;
      RF_HIP(hipEventElapsedTime(&t, h->timer.pool[r.ev0], h->timer.pool[r.ev1]));
      if (cls) cls[k] = r.cls;
      if (variant) variant[k] = r.variant;
      if (ms) ms[k] = t;
      if (work) work[k] = r.work;
    }
    ++k;
  }
  *count = k;
  return 0;
}

}  // extern "C"
