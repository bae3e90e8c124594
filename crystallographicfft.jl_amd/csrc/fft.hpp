// In-register mixed-radix FFT building blocks for gfx950.
//
// One lane owns one length-LEN transform held entirely in VGPRs.  The
// transform is an iterative Stockham autosort (radix 4, then 2, then 3 passes)
// whose loops are fully unrolled at compile time, so every register index and
// every twiddle index is a constant: twiddles come from compile-time FP64
// tables (twiddles.hpp) and fold into literal operands, with no memory
// traffic.  Forward = exp(-2 pi i jk/n) unnormalised, inverse =
// exp(+2 pi i jk/n) unnormalised (scaling is folded into the callers).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "twiddles.hpp"

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

// exp(-2 pi i m / LEN), 0 <= m < LEN, for LEN dividing 64 or 48.  With m a
// compile-time constant after unrolling, both parts are literals.
template <int LEN, typename TW>
__device__ __forceinline__ cx<TW> tw_c(int m) {
  static_assert(64 % LEN == 0 || 48 % LEN == 0, "FFT length must divide 64 or 48");
  if constexpr (64 % LEN == 0)
    return {(TW)kTwRe64[m * (64 / LEN)], (TW)kTwIm64[m * (64 / LEN)]};
  else
    return {(TW)kTwRe48[m * (48 / LEN)], (TW)kTwIm48[m * (48 / LEN)]};
}

// x * w (DIR < 0) or x * conj(w) (DIR > 0), the product formed in the
// table precision TW and rounded once to the data precision T.
template <typename T, typename TW, int DIR>
__device__ __forceinline__ cx<T> twmul(cx<T> x, cx<TW> w) {
  const TW xr = (TW)x.x, xi = (TW)x.y;
  if constexpr (DIR < 0)
    return {(T)(xr * w.x - xi * w.y), (T)(xr * w.y + xi * w.x)};
  else
    return {(T)(xr * w.x + xi * w.y), (T)(xi * w.x - xr * w.y)};
}

// R-point DFT in place; DIR = -1 forward, +1 inverse.  The radix-3 constant
// is applied in the table precision TW.
template <int R, int DIR, typename T, typename TW = T>
__device__ __forceinline__ void dft_small(cx<T>* v) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
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
    const TW h = TW(0.86602540378443864676372317075294);  // sqrt(3)/2
    cx<T> a = v[0];
    cx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cx<T> m = {a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
    // forward: y1 = m - i h d, y2 = m + i h d
    cx<T> ihd = {(T)(-h * (TW)d.y), (T)(h * (TW)d.x)};  // i * h * d
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

// Stockham pass(es) starting at sub-transform length NS on LEN values held in
// registers; twiddle exp(-2 pi i m/LEN).
template <typename T, typename TW, int LEN, int DIR, int NS>
__device__ __forceinline__ void stockham(cx<T> (&a)[LEN]) {
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
          v[r] = twmul<T, TW, DIR>(v[r], tw_c<LEN, TW>(m));
        }
      }
      dft_small<R, DIR, T, TW>(v);
#pragma unroll
      for (int r = 0; r < R; ++r) b[(j / NS) * NS * R + k + r * NS] = v[r];
    }
#pragma unroll
    for (int i = 0; i < LEN; ++i) a[i] = b[i];
    stockham<T, TW, LEN, DIR, NS * R>(a);
  }
}

template <typename T, int LEN, int DIR>
__device__ __forceinline__ void fft_reg(cx<T> (&a)[LEN]) {
  static_assert(fft_len_ok(LEN), "FFT length must be 2^a 3^b");
  stockham<T, T, LEN, DIR, 1>(a);
}

// v[k] *= w^(r k), k = 1 .. R2 - 1, for the lane index r < R1 (the twiddles
// between the two stages of the multi-lane FFT).  With FP64 twiddle
// arithmetic the lane selects w^r once among R1 literals and forms the powers
// by successive products (about k ulps from the literal; a literal select
// per power costs several times the arithmetic it feeds); FP32 twiddles are
// each selected among the literals.
template <typename T, typename TW, int LEN, int DIR, int R1, int R2>
__device__ __forceinline__ void lane_twiddles(cx<T> (&v)[R2], int r) {
  if constexpr (std::is_same<TW, double>::value) {
    cx<TW> wr = tw_c<LEN, TW>(0);
#pragma unroll
    for (int s = 1; s < R1; ++s)
      if (r == s) wr = tw_c<LEN, TW>(s);
    cx<TW> w = wr;
#pragma unroll
    for (int k = 1; k < R2; ++k) {
      v[k] = twmul<T, TW, DIR>(v[k], w);
      if (k + 1 < R2) w = cmul(w, wr);
    }
  } else {
#pragma unroll
    for (int k = 1; k < R2; ++k) {
      cx<TW> w = tw_c<LEN, TW>(0);
#pragma unroll
      for (int s = 1; s < R1; ++s)
        if (r == s) w = tw_c<LEN, TW>((s * k) % LEN);
      v[k] = twmul<T, TW, DIR>(v[k], w);
    }
  }
}

// Split of a length-LEN transform over R1 cooperating lanes with R2 = LEN/R1
// values each (R2 divisible by R1): the in-lane R2-point Stockham passes
// (radices 4, 2, 3 in that order) followed by one cross-lane radix-R1 pass.
// LANES = 0 picks R1 = 4 for LEN divisible by 16, 2 for LEN divisible by 4,
// else 1; LANES > 0 sets R1 (a radix schedule the tuner may choose).
__host__ __device__ constexpr int ml_r1(int len, int lanes) {
  return lanes ? lanes : ((len % 16 == 0) ? 4 : ((len % 4 == 0) ? 2 : 1));
}
__host__ __device__ constexpr int ml_r2(int len, int lanes) { return len / ml_r1(len, lanes); }

template <int LEN, int LANES = 0>
struct MLS {
  static constexpr int R1 = ml_r1(LEN, LANES);
  static constexpr int R2 = ml_r2(LEN, LANES);
  static_assert(R1 * R2 == LEN && R2 % R1 == 0, "lane split must divide the length, R2 divisible by R1");
};

// Multi-lane FFT.  A group of R1 lanes (lane index r) owns one transform; on
// entry v[m] = x[r + R1 m] (m < R2).  Each lane runs an R2-point FFT in
// registers, twiddles by w^(r k'), the group exchanges through LDS, and each
// lane runs R2/R1 R1-point DFTs.  On exit v[j R1 + k2] = X[(r + R1 j) + R2 k2].
// xch is the workgroup's exchange area of R2 * R1 * GP entries (GP = groups per
// workgroup + 1); every lane of the workgroup must call this function.
template <typename T, typename TW, int LEN, int DIR, int GP>
__device__ __forceinline__ void fft_ml(cx<T> (&v)[MLS<LEN>::R2], int r, int g, cx<T>* xch) {
  constexpr int R1 = MLS<LEN>::R1, R2 = MLS<LEN>::R2;
  stockham<T, TW, R2, DIR, 1>(v);
  if constexpr (R1 > 1) {
    lane_twiddles<T, TW, LEN, DIR, R1, R2>(v, r);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R2; ++k) xch[(k * R1 + r) * GP + g] = v[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R2 / R1; ++j) {
      const int kp = r + R1 * j;
      cx<T> u[R1];
#pragma unroll
      for (int s = 0; s < R1; ++s) u[s] = xch[(kp * R1 + s) * GP + g];
      dft_small<R1, DIR, T, TW>(u);
#pragma unroll
      for (int k2 = 0; k2 < R1; ++k2) v[j * R1 + k2] = u[k2];
    }
  }
}

// Multi-lane FFT of one transform that already lives in LDS (element p at
// slot[p * es]): the exchange between the R1 lanes of the group goes through
// the transform's own slots.  The R1 lanes are consecutive lanes of one
// wavefront, so wavefront-scope fences order the exchange; no workgroup
// barrier and no separate exchange buffer.  Same entry/exit layout as fft_ml.
template <typename T, typename TW, int LEN, int DIR, int LANES = 0>
__device__ __forceinline__ void fft_ml_lds(cx<T> (&v)[ml_r2(LEN, LANES)], int r, cx<T>* slot, int es) {
  constexpr int R1 = MLS<LEN, LANES>::R1, R2 = MLS<LEN, LANES>::R2;
  stockham<T, TW, R2, DIR, 1>(v);
  if constexpr (R1 > 1) {
    lane_twiddles<T, TW, LEN, DIR, R1, R2>(v, r);
    // every lane of the group has its inputs in registers: the slots are free
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < R2; ++k) slot[(k * R1 + r) * es] = v[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int j = 0; j < R2 / R1; ++j) {
      const int kp = r + R1 * j;
      cx<T> u[R1];
#pragma unroll
      for (int s = 0; s < R1; ++s) u[s] = slot[(kp * R1 + s) * es];
      dft_small<R1, DIR, T, TW>(u);
#pragma unroll
      for (int k2 = 0; k2 < R1; ++k2) v[j * R1 + k2] = u[k2];
    }
    // the group's reads of the exchange precede the caller's stores
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Output position of register p after fft_ml.
template <int LEN, int LANES = 0>
__device__ __forceinline__ int ml_out_pos(int r, int p) {
  constexpr int R1 = MLS<LEN, LANES>::R1, R2 = MLS<LEN, LANES>::R2;
  return (r + R1 * (p / R1)) + R2 * (p % R1);
}

// A 32-point transform over eight lanes of four values (lane r < 8, on entry
// v[m] = x[r + 8 m]): in-lane 4-point DFT, twiddles w32^(r k), exchange
// through the transform's own LDS slots; then lane r takes the outputs
// X[k + 4 k2] of k = r / 2 with k2 = h + 2 j, h = r % 2:
//   X[k + 4 (h + 2 j)] = sum_{s<4} w8^(s h) w4^(s j) (y_s + (-1)^h y_{s+4})
// with y_s the twiddled value of lane s for k (the parity split of the
// 8-point DFT across the lanes).  On exit v[j] = X[l32_out_pos(r, j)].
// The R1 = 8 lanes are consecutive lanes of one wavefront.
__device__ __forceinline__ int l32_out_pos(int r, int j) { return (r >> 1) + 4 * ((r & 1) + 2 * j); }

template <typename T, typename TW, int DIR>
__device__ __forceinline__ void fft32_l8_lds(cx<T> (&v)[4], int r, cx<T>* slot, int es) {
  stockham<T, TW, 4, DIR, 1>(v);
  lane_twiddles<T, TW, 32, DIR, 8, 4>(v, r);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int k = 0; k < 4; ++k) slot[(k * 8 + r) * es] = v[k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int k = r >> 1;
  const bool h = (r & 1) != 0;
  cx<T> y[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) y[s] = slot[(k * 8 + s) * es];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    v[s] = h ? csub(y[s], y[s + 4]) : cadd(y[s], y[s + 4]);
    if (s > 0) {
      const cx<TW> w8 = tw_c<8, TW>(s);
      const cx<TW> w = {h ? w8.x : TW(1), h ? w8.y : TW(0)};
      v[s] = twmul<T, TW, DIR>(v[s], w);
    }
  }
  dft_small<4, DIR, T, TW>(v);
  // the group's reads of the exchange precede the caller's stores
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Line transforms of the axis-split chain: length 32 over eight lanes of four
// values, other lengths the default multi-lane split.
template <int LEN>
struct LineSplit {
  static constexpr int R1 = LEN == 32 ? 8 : MLS<LEN>::R1;
  static constexpr int R2 = LEN / R1;
};

template <int LEN>
__device__ __forceinline__ int line_out_pos(int r, int j) {
  if constexpr (LEN == 32) return l32_out_pos(r, j);
  else return ml_out_pos<LEN>(r, j);
}

template <typename T, typename TW, int LEN, int DIR>
__device__ __forceinline__ void fft_line_lds(cx<T> (&v)[LineSplit<LEN>::R2], int r, cx<T>* slot, int es) {
  if constexpr (LEN == 32) fft32_l8_lds<T, TW, DIR>(v, r, slot, es);
  else fft_ml_lds<T, TW, LEN, DIR>(v, r, slot, es);
}

}  // namespace cfft
