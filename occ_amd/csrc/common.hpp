// common.hpp -- error handling and device-buffer helpers for the occxc
// library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "occxc.h"

namespace occxc {

// Errors inside the library are C++ exceptions carrying a status; every C
// entry point catches them and turns them into a status + message.
struct Error : std::runtime_error {
  occxc_status status;
  Error(occxc_status s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string &msg);
void clear_last_error();

#define OCCXC_HIP(call)                                                        \
  do {                                                                         \
    hipError_t err_ = (call);                                                  \
    if (err_ != hipSuccess)                                                    \
      throw ::occxc::Error(err_ == hipErrorOutOfMemory ? OCCXC_ERR_OUT_OF_MEMORY \
                                                       : OCCXC_ERR_DEVICE,     \
                           std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

#define OCCXC_REQUIRE(cond, status, msg)                                       \
  do {                                                                         \
    if (!(cond)) throw ::occxc::Error(status, msg);                            \
  } while (0)

// Owning device allocation.
template <typename T> class DevBuf {
public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t n) {
    release();
    if (n == 0) return;
    OCCXC_HIP(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)));
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void upload(const T *h, size_t n, hipStream_t s = nullptr) {
    if (n > n_) throw Error(OCCXC_ERR_INTERNAL, "DevBuf upload overflow");
    if (n) OCCXC_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *h, size_t n, hipStream_t s = nullptr) const {
    if (n > n_) throw Error(OCCXC_ERR_INTERNAL, "DevBuf download overflow");
    if (n) OCCXC_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

private:
  T *p_ = nullptr;
  size_t n_ = 0;
};

// Make `device` current for the lifetime of the guard.
class DeviceGuard {
public:
  explicit DeviceGuard(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw Error(OCCXC_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
      throw Error(OCCXC_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    OCCXC_HIP(hipGetDevice(&prev_));
    OCCXC_HIP(hipSetDevice(device));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev_); }

private:
  int prev_ = 0;
};

} // namespace occxc
