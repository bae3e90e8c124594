// Host-side helpers shared by the translation units of libcfft_hip.so:
// status/error plumbing of the C ABI, device buffers, exact phase factors.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/cfft_hip.h"
#include "plan.hpp"

#include <mutex>

namespace cfft {

extern thread_local std::string g_last_error;

// Serialises hipGraph capture (stream capture in one thread) with the
// library's set-up calls that use the null stream, allocate or free memory
// (plan and context creation, table updates, destruction, every device
// allocation) in other threads: such calls during another thread's capture
// invalidate that capture.  Recursive: set-up calls allocate while holding it.
std::recursive_mutex& setup_mutex();


inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error{CFFT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

template <typename F>
inline int guarded(F&& f) {
  try {
    f();
    return CFFT_OK;
  } catch (const Error& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return CFFT_ERR_HIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CFFT_ERR_INVALID;
  }
}

inline void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw Error{code, msg};
}

// exp(2 pi i j / P), argument reduced in extended precision.
inline void cis2pi(long long j, long long P, double& re, double& im) {
  long long r = ((j % P) + P) % P;
  long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)r / (long double)P;
  re = (double)cosl(a);
  im = (double)sinl(a);
  // exact values on the axes
  if (4 * r == 0) { re = 1.0; im = 0.0; }
  else if (4 * r == P) { re = 0.0; im = 1.0; }
  else if (4 * r == 2 * P) { re = -1.0; im = 0.0; }
  else if (4 * r == 3 * P) { re = 0.0; im = -1.0; }
}

constexpr int kFftLens[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64};
inline bool fft_supported(int n) {
  for (int v : kFftLens)
    if (v == n) return true;
  return false;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void alloc(size_t n) {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    release();
    if (n == 0) return;
    hip_check(hipMalloc(&p, n), "hipMalloc");
    bytes = n;
  }
  void release() {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
};

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace cfft
