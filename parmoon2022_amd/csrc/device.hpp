// HIP runtime helpers shared by the host-side orchestration and the kernel
// launch wrappers.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include "common.hpp"

namespace pmcd {

#define PMCD_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if(e_ != hipSuccess) PMCD_FAIL("HIP error ", hipGetErrorString(e_), " in ", #call); \
  } while(0)

// Device bytes held by one owner (a system, a batch, a solver, an
// assembler): DevBuf::alloc charges the tally that is current on the calling
// thread and release credits the same tally. Sizes are charged rounded up to
// the 4 KiB page the device allocator hands out at least.
struct DevTally {
  std::atomic<long long> bytes{0};
  std::atomic<long long> peak{0};
  void add(long long b) {
    const long long v = bytes.fetch_add(b) + b;
    long long p = peak.load();
    while(v > p && !peak.compare_exchange_weak(p, v)) {}
  }
};
inline DevTally*& current_tally() {
  static thread_local DevTally* t = nullptr;
  return t;
}
// Makes `t` the current tally of this thread for the scope's lifetime.
struct TallyScope {
  DevTally* prev;
  explicit TallyScope(DevTally* t) : prev(current_tally()) { current_tally() = t; }
  ~TallyScope() { current_tally() = prev; }
  TallyScope(const TallyScope&) = delete;
  TallyScope& operator=(const TallyScope&) = delete;
};
inline long long charged_bytes(size_t b) { return (long long)((b + 4095) & ~(size_t)4095); }

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevTally* tally = nullptr;  // owner charged for this allocation
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), tally(o.tally) { o.p = nullptr; o.n = 0; o.tally = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if(this != &o) {
      release();
      p = o.p; n = o.n; tally = o.tally;
      o.p = nullptr; o.n = 0; o.tally = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if(!count) return;
    PMCD_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
    tally = current_tally();
    if(tally) tally->add(charged_bytes(bytes()));
  }
  void release() {
    if(p) {
      (void)hipFree(p);
      if(tally) tally->add(-charged_bytes(bytes()));
    }
    p = nullptr; n = 0; tally = nullptr;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    if(count != n) alloc(count);
    if(count) PMCD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    PMCD_CHECK(count <= n, "download of ", count, " > ", n);
    if(count) PMCD_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace pmcd
