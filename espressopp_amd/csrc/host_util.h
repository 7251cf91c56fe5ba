// Host-side helpers shared by the runtime files (sim.hip, batch.hip): error
// codes carried by exceptions up to the C ABI, device buffers, and the guard
// that turns an exception into a status code and the thread's last error.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
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

// Scale 1 / delta of the list builds' 16-bit position copies (q16_pack) for a
// cubic box of length L with n cells a side, or 0 where they do not apply (a box
// too small for one period, or a prefilter top `top` above the cell bound).
// delta = L / (2^16 m) keeps a box length a whole number of 2^16-unit periods,
// so a 16-bit difference is that of the periodic image within 2^15 units.  m is
// the largest whole number for which every candidate of a particle's rows stays
// below 32000 units per coordinate (in x the clipped stretch: one cell, an
// eighth of slack and the drift terms; in y and z the two cells across the row
// plus twice the drift) and the three squares sum below 46000^2 < 2^31.
inline double q16_scale_cubic(double L, int n, double disp, double top) {
  const double cs = L / n;
  if (!(top <= 1.001 * cs)) return 0.0;
  const double bx = cs * 1.0012 + cs / 8.0 + 4.0 * disp;
  const double byz = 2.0 * cs * 1.0001 + 2.0 * disp;
  const double dmin = std::max(std::max(bx, byz) / 32000.0, std::sqrt(bx * bx + 2.0 * byz * byz) / 46000.0);
  const double m = std::floor(L / (65536.0 * dmin));
  return m >= 1.0 ? 65536.0 * m / L : 0.0;
}

// Thresholds of r^2 in 16-bit units for the cut (rc+skin) at scale q16 (> 0):
// each coordinate of a copy is within half a unit of the position, so a
// difference is off by at most one unit and r^2 by at most 2 sqrt(3) |d| + 3
// units^2.  Below tlo the pair is surely within the cut, at or above tlo + w
// surely not.
inline void q16_thresholds(double cut, double q16, int& tlo, int& w) {
  const double U = cut * q16, U2 = U * U;
  const double lo = std::floor(U2 - 2.0 * std::sqrt(3.0) * U * 1.0001 - 8.0);
  const double hi = std::ceil(U2 + 2.0 * std::sqrt(3.0) * U * 1.01 + 16.0);
  tlo = (int)std::max(lo, 0.0);
  w = (int)(hi - tlo);
}

}  // namespace espp_host
