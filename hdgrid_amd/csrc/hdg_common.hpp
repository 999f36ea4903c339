// Shared host-side helpers for the HDGrid MI355X library: errors, HIP checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace hdg {

// Library error carrying the C-ABI status code it maps to.
struct Error : std::runtime_error {
  int status;
  Error(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

enum Status : int {
  kOk = 0,
  kInvalidArgument = 1,
  kConfigError = 2,
  kIoError = 3,
  kHipError = 4,
  kOutOfMemory = 5,
  kBadState = 6,
  kUnsupported = 7,
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
    throw Error(e == hipErrorOutOfMemory ? kOutOfMemory : kHipError, buf);
  }
}
#define HDG_HIP(x) ::hdg::hip_check((x), #x, __FILE__, __LINE__)

[[noreturn]] inline void fail(int st, const std::string& msg) { throw Error(st, msg); }

// Counts device (re)allocations; a captured launch graph holds raw pointers and is valid only while the
// count it was captured at still holds.
inline uint64_t& devbuf_generation() {
  static uint64_t g = 0;
  return g;
}

// Owning device allocation (hipMalloc / hipFree), grows on demand and never shrinks.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }

  void reserve(size_t bytes) {
    if (bytes <= bytes_) return;
    release();
    size_t b = bytes < 256 ? 256 : bytes;
    HDG_HIP(hipMalloc(&p_, b));
    bytes_ = b;
    ++devbuf_generation();
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p_);
  }
  void* data() const { return p_; }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

template <class T>
inline void upload(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.reserve(v.size() * sizeof(T));
  if (!v.empty()) {
    HDG_HIP(hipMemcpyAsync(b.data(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    HDG_HIP(hipStreamSynchronize(s));  // the host vector may be released right after
  }
}

// f32 -> u64 as GCC emits it for x86-64 host code: x < 2^63 -> cvttss2si(x), else
// cvttss2si(x - 2^63) ^ 2^63, where cvttss2si yields 0x8000000000000000 for NaN / out of range.
inline uint64_t f2u_host(float x) {
  auto cvt = [](float v) -> int64_t {
    if (v != v || v >= 9223372036854775808.0f || v < -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)v;
  };
  if (x < 9223372036854775808.0f) return (uint64_t)cvt(x);
  return (uint64_t)cvt(x - 9223372036854775808.0f) ^ 0x8000000000000000ull;
}

void log_line(int level, const char* fmt, ...);  // level 0=error 1=info 2=trace
int log_level();

}  // namespace hdg
