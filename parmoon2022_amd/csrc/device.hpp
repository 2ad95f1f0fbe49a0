// HIP runtime helpers shared by the host-side orchestration and the kernel
// launch wrappers.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace pmcd {

#define PMCD_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if(e_ != hipSuccess) PMCD_FAIL("HIP error ", hipGetErrorString(e_), " in ", #call); \
  } while(0)

// Owning device allocation.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if(this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if(count) PMCD_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void release() {
    if(p) (void)hipFree(p);
    p = nullptr; n = 0;
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
