// Internal helpers shared by the HIP translation units of libragnar_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "ragnar_hip.h"

namespace rgn {

// An error carrying the C-ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string& msg);
void ensure_init();

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define RGN_HIP(call)                                          \
  do {                                                         \
    hipError_t rgn_e_ = (call);                                \
    if (rgn_e_ != hipSuccess) {                                \
      ::rgn::throw_hip(rgn_e_, #call, __FILE__, __LINE__);     \
    }                                                          \
  } while (0)

#define RGN_REQUIRE(cond, code, msg)                           \
  do {                                                         \
    if (!(cond)) throw ::rgn::Error((code), (msg));            \
  } while (0)

// Runs f() and converts any exception to a status code + last-error message.
template <class F>
int guarded(F&& f) {
  try {
    ensure_init();
    f();
    return RGN_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc& e) {
    set_last_error(std::string("host allocation failed: ") + e.what());
    return RGN_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RGN_ERR_INVALID;
  }
}

// A device buffer owned by the library (plans, tables).
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { reset(n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void reset(size_t n);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Scratch of the library's device, one buffer per slot, reused by every call
// (grows on demand; never shrinks).  The mutex only guards the slot table;
// calls that use the buffers are ordered by ScratchUse below.
void* scratch(size_t bytes, int slot);
// The same; *fresh is set when the slot was (re)allocated by this call, so its
// previous contents are gone.
void* scratch(size_t bytes, int slot, bool* fresh);

// Orders the library calls that use the scratch buffers, whatever streams they
// are issued on.  While alive it holds a process-wide (recursive) lock, so
// host threads issue such calls one at a time.  On entry, the stream waits for
// the event that the previous such call recorded, when that call ran on
// another stream; on exit (also after an error) the event is recorded on this
// stream.  So a call never starts on the device before the previous call's
// kernels are done with the shared buffers, and calls on one stream pay
// nothing but the record.  A stream under graph capture is left alone: a
// graph's replays are ordered by the stream that launches them.
class ScratchUse {
 public:
  explicit ScratchUse(hipStream_t st);
  ~ScratchUse();
  ScratchUse(const ScratchUse&) = delete;
  ScratchUse& operator=(const ScratchUse&) = delete;

 private:
  std::unique_lock<std::recursive_mutex> lock_;
  hipStream_t st_;
  bool active_ = false;
};

// Number of compute units of the current device.
int num_cus();
// Architecture name of the library's device (e.g. "gfx950:sramecc+:xnack-").
std::string device_arch();

// Integer from the environment (launch-shape knobs), or def when unset.
int env_int(const char* name, int def);

inline hipStream_t as_stream(rgn_stream_t s) { return static_cast<hipStream_t>(s); }

// Per-launch timing of one library call (the *_stages diagnostics entry
// points).  While a StageScope is alive on this thread, stage_mark records a
// HIP event on the launch stream after each kernel launch; finish() waits for
// the last event and returns the elapsed milliseconds of every launch.
constexpr int kMaxStages = 16;
struct StageScope {
  explicit StageScope(hipStream_t st);
  ~StageScope();
  int finish(float* ms, const char** names, int max_stages);
  hipStream_t st;
  hipEvent_t ev[kMaxStages + 1] = {};
  const char* names[kMaxStages] = {};
  int n = 0;
};
void stage_mark(hipStream_t st, const char* name);

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace rgn
