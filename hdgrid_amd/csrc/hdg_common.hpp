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

// Bump allocator over a device region the caller provides (the allocation-free query path carves its whole
// workspace out of one fixed array).  While an arena is active on this thread, DevBuf::reserve carves from it
// instead of calling hipMalloc, and the carved buffers own nothing.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
};
inline Arena*& active_arena() {
  static thread_local Arena* a = nullptr;
  return a;
}
// Thrown by DevBuf::reserve when the active arena cannot hold the request; `need` = bytes the arena would need.
struct ArenaOverflow : Error {
  size_t need;
  ArenaOverflow(size_t n, const std::string& msg) : Error(kOutOfMemory, msg), need(n) {}
};

// Device buffer that grows on demand and never shrinks: an owning hipMalloc allocation, or (while an arena is
// active) a non-owning piece of the arena.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_), owned_(o.owned_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
    o.owned_ = false;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      owned_ = o.owned_;
      o.p_ = nullptr;
      o.bytes_ = 0;
      o.owned_ = false;
    }
    return *this;
  }
  ~DevBuf() { release(); }

  void reserve(size_t bytes) {
    if (bytes <= bytes_) return;
    if (Arena* a = active_arena()) {
      if (owned_) fail(kBadState, "an owning device buffer would be reallocated inside an allocation-free call");
      const size_t b = ((bytes < 256 ? 256 : bytes) + 255) & ~(size_t)255;
      if (b > a->cap - a->used)
        throw ArenaOverflow(a->used + b, "workspace needs " + std::to_string(a->used + b) + " bytes of the " +
                                             std::to_string(a->cap) + "-byte scratch array");
      p_ = a->base + a->used;
      a->used += b;
      bytes_ = b;
      return;
    }
    release();
    size_t b = bytes < 256 ? 256 : bytes;
    HDG_HIP(hipMalloc(&p_, b));
    bytes_ = b;
    owned_ = true;
    ++devbuf_generation();
  }
  void release() {
    if (p_ && owned_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
    owned_ = false;
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
  bool owned_ = false;
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
