// In-register mixed-radix FFT building blocks for gfx950.
//
// One lane owns one length-LEN transform held entirely in VGPRs.  The
// transform is an iterative Stockham autosort (radix 4, then 2, then 3 passes)
// whose loops are fully unrolled at compile time, so every register index and
// every twiddle index is a constant: twiddles are uniform loads from a small
// FP64-accurate table (exp(-2 pi i m/LEN), m < LEN) that the compiler turns
// into scalar loads.  Forward = exp(-2 pi i jk/n) unnormalised, inverse =
// exp(+2 pi i jk/n) unnormalised (scaling is folded into the callers).
#pragma once

#include <hip/hip_runtime.h>

namespace cfft {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }

// Radix choice of the Stockham pass that starts at sub-length NS.
__host__ __device__ constexpr int fft_radix(int len, int ns) {
  return ((len / ns) % 4 == 0) ? 4 : (((len / ns) % 2 == 0) ? 2 : 3);
}

__host__ __device__ constexpr bool fft_len_ok(int len) {
  int n = len;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

// R-point DFT in place; DIR = -1 forward, +1 inverse.
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_small(cx<T>* v) {
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cx<T> t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    // forward: y1 = t1 - i t3, y3 = t1 + i t3 ; inverse swaps the signs
    cx<T> mi_t3 = {t3.y, -t3.x};  // -i * t3
    if constexpr (DIR < 0) {
      v[1] = cadd(t1, mi_t3);
      v[3] = csub(t1, mi_t3);
    } else {
      v[1] = csub(t1, mi_t3);
      v[3] = cadd(t1, mi_t3);
    }
  } else {  // R == 3
    const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
    cx<T> a = v[0];
    cx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cx<T> m = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
    // forward: y1 = m - i h d, y2 = m + i h d
    cx<T> ihd = {-h * d.y, h * d.x};  // i * h * d
    v[0] = cadd(a, s);
    if constexpr (DIR < 0) {
      v[1] = csub(m, ihd);
      v[2] = cadd(m, ihd);
    } else {
      v[1] = cadd(m, ihd);
      v[2] = csub(m, ihd);
    }
  }
}

// Stockham pass(es) starting at sub-transform length NS; tw[m] = exp(-2 pi i m/LEN).
template <typename T, int LEN, int DIR, int NS>
__device__ __forceinline__ void stockham(cx<T> (&a)[LEN], const cx<T>* __restrict__ tw) {
  if constexpr (NS < LEN) {
    constexpr int R = fft_radix(LEN, NS);
    constexpr int J = LEN / R;
    cx<T> b[LEN];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k = j % NS;
      cx<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v[r] = a[j + r * J];
        if (r > 0 && k > 0) {
          const int m = (r * k * (LEN / (NS * R))) % LEN;
          v[r] = (DIR < 0) ? cmul(v[r], tw[m]) : cmulc(v[r], tw[m]);
        }
      }
      dft_small<R, DIR>(v);
#pragma unroll
      for (int r = 0; r < R; ++r) b[(j / NS) * NS * R + k + r * NS] = v[r];
    }
#pragma unroll
    for (int i = 0; i < LEN; ++i) a[i] = b[i];
    stockham<T, LEN, DIR, NS * R>(a, tw);
  }
}

template <typename T, int LEN, int DIR>
__device__ __forceinline__ void fft_reg(cx<T> (&a)[LEN], const cx<T>* __restrict__ tw) {
  static_assert(fft_len_ok(LEN), "FFT length must be 2^a 3^b");
  stockham<T, LEN, DIR, 1>(a, tw);
}

}  // namespace cfft
