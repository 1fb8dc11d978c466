// Error mapping and RAII device/pinned buffers for the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/cp3_hip.h"

namespace cp3 {
namespace capi {

// An error carrying its C ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(e == hipErrorOutOfMemory ? CP3_ERR_NOMEM : CP3_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
  }
}

// Runs f, turning every exception into a status code and a message.
template <class Fn>
int guard(Fn&& f) {
  try {
    f();
    return CP3_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::invalid_argument& e) {
    last_error() = e.what();
    return CP3_ERR_INVALID;
  } catch (const std::bad_alloc&) {
    last_error() = "host allocation failed";
    return CP3_ERR_NOMEM;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return CP3_ERR_INVALID;
  } catch (...) {
    last_error() = "unknown error";
    return CP3_ERR_INVALID;
  }
}

// Makes `dev` current for the scope and restores the previous device.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
      (void)hipSetDevice(prev);
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
};

// Device memory owned by one object.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int dev = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      DeviceScope ds(dev);
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
  }
  void alloc(size_t n, int device) {
    if (p && bytes >= n && dev == device) return;
    release();
    dev = device;
    if (n == 0) return;
    DeviceScope ds(device);
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
  }
  // Synchronous copy of n host bytes (grows the buffer when needed).
  void upload(const void* src, size_t n, int device) {
    alloc(n, device);
    if (n == 0) return;
    DeviceScope ds(device);
    hip_check(hipMemcpy(p, src, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
};

// Pinned host memory (hipHostMalloc; never hipHostRegister).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) {
      (void)hipHostFree(p);
      p = nullptr;
      bytes = 0;
    }
  }
  void alloc(size_t n) {
    if (p && bytes >= n) return;
    release();
    if (n == 0) return;
    hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
    bytes = n;
  }
};

}  // namespace capi
}  // namespace cp3
