// Runtime entry points of libragnar_hip.so: init, errors, device buffers and
// the per-device scratch cache.
#include "rgn_core.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace rgn {

namespace {
thread_local std::string g_last_error;
std::mutex g_init_mutex;
bool g_initialized = false;
int g_device = 0;
int g_num_cus = 0;
std::string g_arch;

struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_scratch_mutex;
std::vector<Scratch> g_scratch;  // indexed by slot

// ScratchUse state: the event the last scratch user recorded and its stream.
std::recursive_mutex g_order_mutex;
hipEvent_t g_order_event = nullptr;
hipStream_t g_order_stream = nullptr;
bool g_order_pending = false;
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  const int code = (e == hipErrorOutOfMemory) ? RGN_ERR_NOMEM : RGN_ERR_HIP;
  throw Error(code, std::string("HIP error '") + hipGetErrorString(e) + "' in " + what +
                      " at " + file + ":" + std::to_string(line));
}

static void do_init(int device) {
  int count = 0;
  RGN_HIP(hipGetDeviceCount(&count));
  RGN_REQUIRE(count > 0, RGN_ERR_HIP, "no HIP device visible");
  RGN_REQUIRE(device >= 0 && device < count, RGN_ERR_INVALID,
              "device index " + std::to_string(device) + " out of range");
  RGN_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  RGN_HIP(hipGetDeviceProperties(&prop, device));
  const std::string arch = prop.gcnArchName;
  RGN_REQUIRE(arch.rfind("gfx950", 0) == 0, RGN_ERR_UNSUPPORTED,
              "libragnar_hip is built for gfx950; device " + std::to_string(device) +
                " is " + arch);
  g_device = device;
  g_num_cus = prop.multiProcessorCount;
  g_arch = arch;
  g_initialized = true;
}

void ensure_init() {
  std::lock_guard<std::mutex> lk(g_init_mutex);
  if (!g_initialized) do_init(0);
}

int num_cus() { return g_num_cus > 0 ? g_num_cus : 256; }

std::string device_arch() { return g_arch; }

int env_int(const char* name, int def) {
  const char* v = std::getenv(name);
  if (!v || !*v) return def;
  const int x = std::atoi(v);
  return x > 0 ? x : def;
}

void DeviceBuffer::reset(size_t n) {
  release();
  if (n == 0) return;
  RGN_HIP(hipMalloc(&ptr, n));
  bytes = n;
}

void DeviceBuffer::release() {
  if (ptr) {
    (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
}

namespace {
thread_local StageScope* t_stages = nullptr;
}

StageScope::StageScope(hipStream_t s) : st(s) {
  RGN_REQUIRE(t_stages == nullptr, RGN_ERR_INVALID, "nested stage timing");
  for (auto& e : ev) RGN_HIP(hipEventCreate(&e));
  RGN_HIP(hipEventRecord(ev[0], st));
  t_stages = this;
}

StageScope::~StageScope() {
  t_stages = nullptr;
  for (auto& e : ev) {
    if (e) (void)hipEventDestroy(e);
  }
}

int StageScope::finish(float* ms, const char** out_names, int max_stages) {
  t_stages = nullptr;
  if (n > 0) RGN_HIP(hipEventSynchronize(ev[n]));
  const int m = std::min(n, max_stages);
  for (int i = 0; i < m; ++i) {
    RGN_HIP(hipEventElapsedTime(ms + i, ev[i], ev[i + 1]));
    if (out_names) out_names[i] = names[i];
  }
  return n;
}

void stage_mark(hipStream_t st, const char* name) {
  StageScope* s = t_stages;
  if (!s || s->n >= kMaxStages) return;
  RGN_REQUIRE(st == s->st, RGN_ERR_INVALID, "stage timing across streams");
  s->names[s->n] = name;
  RGN_HIP(hipEventRecord(s->ev[s->n + 1], st));
  ++s->n;
}

ScratchUse::ScratchUse(hipStream_t st) : lock_(g_order_mutex), st_(st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  RGN_HIP(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return;
  if (!g_order_event) RGN_HIP(hipEventCreateWithFlags(&g_order_event, hipEventDisableTiming));
  if (g_order_pending && g_order_stream != st) RGN_HIP(hipStreamWaitEvent(st, g_order_event, 0));
  active_ = true;
}

ScratchUse::~ScratchUse() {
  if (!active_) return;
  if (hipEventRecord(g_order_event, st_) == hipSuccess) {
    g_order_stream = st_;
    g_order_pending = true;
  } else {
    // the next call on another stream cannot be ordered after this one
    (void)hipStreamSynchronize(st_);
    g_order_pending = false;
  }
}

void* scratch(size_t bytes, int slot) { return scratch(bytes, slot, nullptr); }

void* scratch(size_t bytes, int slot, bool* fresh) {
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (fresh) *fresh = false;
  if (static_cast<size_t>(slot) >= g_scratch.size()) g_scratch.resize(slot + 1);
  Scratch& s = g_scratch[slot];
  if (s.bytes < bytes) {
    if (s.ptr) {
      // the previous contents may still be read by queued work
      RGN_HIP(hipDeviceSynchronize());
      RGN_HIP(hipFree(s.ptr));
      s.ptr = nullptr;
      s.bytes = 0;
    }
    const size_t grow = bytes + bytes / 4 + 4096;
    RGN_HIP(hipMalloc(&s.ptr, grow));
    s.bytes = grow;
    if (fresh) *fresh = true;
  }
  return s.ptr;
}

}  // namespace rgn

using namespace rgn;

extern "C" {

int rgn_init(int device) {
  try {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if (g_initialized) {
      RGN_REQUIRE(device == g_device, RGN_ERR_INVALID,
                  "already initialised on device " + std::to_string(g_device));
      return RGN_OK;
    }
    do_init(device);
    return RGN_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

int rgn_finalize(void) {
  try {
    std::lock_guard<std::mutex> lk(g_init_mutex);
    if (!g_initialized) return RGN_OK;
    RGN_HIP(hipDeviceSynchronize());
    {
      std::lock_guard<std::mutex> lk2(g_scratch_mutex);
      for (auto& s : g_scratch) {
        if (s.ptr) (void)hipFree(s.ptr);
        s = Scratch{};
      }
    }
    {
      std::lock_guard<std::recursive_mutex> lk3(g_order_mutex);
      if (g_order_event) (void)hipEventDestroy(g_order_event);
      g_order_event = nullptr;
      g_order_stream = nullptr;
      g_order_pending = false;
    }
    g_initialized = false;
    return RGN_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  }
}

int rgn_is_initialized(void) {
  std::lock_guard<std::mutex> lk(g_init_mutex);
  return g_initialized ? 1 : 0;
}

const char* rgn_last_error(void) { return g_last_error.c_str(); }

const char* rgn_version(void) { return "ragnar_hip 0.1 (gfx950)"; }

int rgn_sync(rgn_stream_t stream) {
  return guarded([&] { RGN_HIP(hipStreamSynchronize(as_stream(stream))); });
}

int rgn_malloc(void** d_ptr, size_t bytes) {
  return guarded([&] {
    *d_ptr = nullptr;
    if (bytes) RGN_HIP(hipMalloc(d_ptr, bytes));
  });
}

int rgn_free(void* d_ptr) {
  return guarded([&] {
    if (d_ptr) RGN_HIP(hipFree(d_ptr));
  });
}

int rgn_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes) {
  return guarded([&] {
    if (bytes) RGN_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  });
}

int rgn_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes) {
  return guarded([&] {
    if (bytes) RGN_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  });
}

int rgn_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes) {
  return guarded([&] {
    if (bytes) RGN_HIP(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
  });
}

int rgn_memset(void* d_dst, int value, size_t bytes) {
  return guarded([&] {
    if (bytes) RGN_HIP(hipMemset(d_dst, value, bytes));
  });
}

int rgn_host_malloc(void** h_ptr, size_t bytes) {
  return guarded([&] {
    *h_ptr = nullptr;
    if (bytes) RGN_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
  });
}

int rgn_host_free(void* h_ptr) {
  return guarded([&] {
    if (h_ptr) RGN_HIP(hipHostFree(h_ptr));
  });
}

int rgn_memcpy_h2d_async(void* d_dst, const void* h_src, size_t bytes, rgn_stream_t stream) {
  return guarded([&] {
    if (bytes) RGN_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  });
}

int rgn_stream_create(rgn_stream_t* stream) {
  return guarded([&] {
    hipStream_t s = nullptr;
    RGN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}

int rgn_stream_destroy(rgn_stream_t stream) {
  return guarded([&] {
    if (stream) RGN_HIP(hipStreamDestroy(as_stream(stream)));
  });
}

int rgn_event_create(void** event) {
  return guarded([&] {
    hipEvent_t e = nullptr;
    RGN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = e;
  });
}

int rgn_event_destroy(void* event) {
  return guarded([&] {
    if (event) RGN_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  });
}

int rgn_event_record(void* event, rgn_stream_t stream) {
  return guarded([&] { RGN_HIP(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream))); });
}

int rgn_event_sync(void* event) {
  return guarded([&] { RGN_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event))); });
}

}  // extern "C"
