// Host-side helpers shared by the runtime files (sim.hip, batch.hip): error
// codes carried by exceptions up to the C ABI, device buffers, and the guard
// that turns an exception into a status code and the thread's last error.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <new>
#include <string>

#include "../../include/espp_amd.h"

namespace espp_host {

// last error message of the calling thread (espp_last_error)
extern thread_local std::string g_err;

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(ESPP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer; resize() discards contents, grow_keep() preserves them.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t m) {
    if (m <= n && p) return;
    release();
    if (m == 0) m = 1;
    if (hipMalloc(&p, m * sizeof(T)) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(ESPP_ERR_OOM, "device allocation of " + std::to_string(m * sizeof(T)) + " bytes failed");
    }
    n = m;
  }
  void grow_keep(size_t m, size_t keep, hipStream_t s) {
    if (m <= n && p) return;
    T* q = nullptr;
    if (m == 0) m = 1;
    if (hipMalloc(&q, m * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      fail(ESPP_ERR_OOM, "device allocation of " + std::to_string(m * sizeof(T)) + " bytes failed");
    }
    if (p && keep) hip_check(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s), "grow copy");
    hip_check(hipStreamSynchronize(s), "grow sync");
    release();
    p = q;
    n = m;
  }
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return ESPP_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host out of memory";
    return ESPP_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return ESPP_ERR_STATE;
  }
}

inline void need(bool ok, const char* msg) {
  if (!ok) fail(ESPP_ERR_ARG, msg);
}

}  // namespace espp_host
