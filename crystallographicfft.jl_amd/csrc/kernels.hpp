// gfx950 kernels of the KRFFT / diffusion hot path.
//
// Data layout (HBM), column-major, batch slowest:
//   f0   : batch x M0 x M1 x M2 real               (stride-L subgrid field)
//   ch   : batch x n_ch x H0 x H1 x H2 complex      (folded channels, H = M/2)
//   F    : batch x n_spec complex                   (spectral-ASU coefficients)
// The M-grid spectrum G0 is never materialised: G0(2k + off_c) lives in
// channel c at k, and every other parity is zero (extinct by centring).
#pragma once

#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace cfft {

struct FoldP {
  int M0, M1, M2, H0, H1, H2;
  int n_ch;
  int offs[8];  // parity offset of channel c as a bit mask (o0 | o1<<1 | o2<<2)
  int batch;
  // exp(-w dt/2) of each batch entry (nullptr: none).  Entries beyond
  // kMaxEPtr use E[0] + b * M^3 when E_contig is set.
  int E_contig;
  const void* E[64];
};
constexpr int kMaxEPtr = 64;

struct ReconP {
  int d;
  int R[8][9];  // representative ops, R row-major
  int t[8][3];
  int M[3], H[3], N[3];
  int P;      // lcm(N)
  int Pn[3];  // P / N_d
  int chmap[8];
  int n_ch, n_spec, batch;
};

struct IreconP {
  int d;
  long n_alive;  // n_ch * H0*H1*H2
  int n_spec, batch;
};

// 3-stage Walsh-Hadamard transform over the 8 corners (bit i = dimension i):
// v[o] <- sum_e (-1)^popcount(o & e) v[e].
template <typename T>
__device__ __forceinline__ void wht8(T* v) {
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      if (!(e & bit)) {
        T a = v[e], c = v[e | bit];
        v[e] = a + c;
        v[e | bit] = a - c;
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ cx<T> fold_twiddle(int o, const cx<T>* __restrict__ twM, int x, int y,
                                              int z, int M0, int M1) {
  cx<T> t = {T(1), T(0)};
  if (o & 1) t = twM[x];
  if (o & 2) t = cmul(t, twM[M0 + y]);
  if (o & 4) t = cmul(t, twM[M0 + M1 + z]);
  return t;
}

// Centring fold, src/centering_fold.jl:140-257 (optionally fused with the
// interaction factor exp(-w dt/2) of step_diffusion!).
template <typename T>
__global__ void __launch_bounds__(256) k_fold(const T* __restrict__ f0, cx<T>* __restrict__ ch,
                                              const cx<T>* __restrict__ twM, FoldP p) {
  const long H3 = (long)p.H0 * p.H1 * p.H2;
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= H3 * p.batch) return;
  const int b = (int)(tid / H3);
  const long n = tid - b * H3;
  const int x = (int)(n % p.H0);
  const int y = (int)((n / p.H0) % p.H1);
  const int z = (int)(n / ((long)p.H0 * p.H1));
  const long M3 = (long)p.M0 * p.M1 * p.M2;
  const T* fb = f0 + b * M3;
  const T* Eb = p.E_contig ? (p.E[0] ? static_cast<const T*>(p.E[0]) + b * M3 : nullptr)
                           : static_cast<const T*>(p.E[b]);
  T v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const long idx = (x + (e & 1) * p.H0) +
                     (long)p.M0 * ((y + ((e >> 1) & 1) * p.H1) + (long)p.M1 * (z + ((e >> 2) & 1) * p.H2));
    v[e] = Eb ? fb[idx] * Eb[idx] : fb[idx];
  }
  wht8(v);
  cx<T>* cb = ch + (long)b * p.n_ch * H3 + n;
  for (int c = 0; c < p.n_ch; ++c) {
    const int o = p.offs[c];
    const cx<T> tw = fold_twiddle(o, twM, x, y, z, p.M0, p.M1);
    T val = T(0);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (e == o) val = v[e];
    cb[c * H3] = {val * tw.x, val * tw.y};
  }
}

// Centring unfold, src/centering_fold.jl:481-580 (+ interaction factor); the
// IFFT's 1/H^3 and the inverse WHT's 1/8 are applied here as one scale.
template <typename T>
__global__ void __launch_bounds__(256) k_unfold(const cx<T>* __restrict__ ch, T* __restrict__ f0,
                                                const cx<T>* __restrict__ twM, FoldP p, T scale) {
  const long H3 = (long)p.H0 * p.H1 * p.H2;
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= H3 * p.batch) return;
  const int b = (int)(tid / H3);
  const long n = tid - b * H3;
  const int x = (int)(n % p.H0);
  const int y = (int)((n / p.H0) % p.H1);
  const int z = (int)(n / ((long)p.H0 * p.H1));
  const cx<T>* cb = ch + (long)b * p.n_ch * H3 + n;
  T v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = T(0);
  for (int c = 0; c < p.n_ch; ++c) {
    const int o = p.offs[c];
    const cx<T> tw = fold_twiddle(o, twM, x, y, z, p.M0, p.M1);
    const cx<T> g = cb[c * H3];
    const T val = g.x * tw.x + g.y * tw.y;  // Re(g * conj(tw))
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (e == o) v[e] = val;
  }
  wht8(v);
  const long M3 = (long)p.M0 * p.M1 * p.M2;
  T* fb = f0 + b * M3;
  const T* Eb = p.E_contig ? (p.E[0] ? static_cast<const T*>(p.E[0]) + b * M3 : nullptr)
                           : static_cast<const T*>(p.E[b]);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const long idx = (x + (e & 1) * p.H0) +
                     (long)p.M0 * ((y + ((e >> 1) & 1) * p.H1) + (long)p.M1 * (z + ((e >> 2) & 1) * p.H2));
    T r = v[e] * scale;
    fb[idx] = Eb ? r * Eb[idx] : r;
  }
}

// Batched 1-D FFT along one axis: transform i reads elements
// base + j*stride (j < LEN), base = (i % inner) + (i / inner) * outer.
template <typename T, int LEN, int DIR>
__global__ void __launch_bounds__(128) k_fft(cx<T>* __restrict__ data, long n_tr, long inner,
                                             long outer, long stride, const cx<T>* __restrict__ tw) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= n_tr) return;
  const long base = (tid % inner) + (tid / inner) * outer;
  cx<T> a[LEN];
#pragma unroll
  for (int j = 0; j < LEN; ++j) a[j] = data[base + j * stride];
  fft_reg<T, LEN, DIR>(a, tw);
#pragma unroll
  for (int j = 0; j < LEN; ++j) data[base + j * stride] = a[j];
}

__device__ __forceinline__ int pmodi(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Forward butterfly reconstruction + spectral K multiply,
// src/krfft.jl:529-545 and src/q_fused_krfft.jl:1013-1016:
// F(h) = K(h) sum_a exp(-2 pi i h.t_a/N) G0(R_a^T h mod M).
template <typename T>
__global__ void __launch_bounds__(256) k_recon(const cx<T>* __restrict__ ch, cx<T>* __restrict__ F,
                                               const int* __restrict__ hs, const T* __restrict__ K,
                                               const cx<T>* __restrict__ W, ReconP p) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= (long)p.n_spec * p.batch) return;
  const int b = (int)(tid / p.n_spec);
  const int i = (int)(tid - (long)b * p.n_spec);
  const int h0 = hs[3 * i], h1 = hs[3 * i + 1], h2 = hs[3 * i + 2];
  const long H3 = (long)p.H[0] * p.H[1] * p.H[2];
  const cx<T>* cb = ch + (long)b * p.n_ch * H3;
  cx<T> acc = {T(0), T(0)};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    if (a < p.d) {
      const int q0 = pmodi(p.R[a][0] * h0 + p.R[a][3] * h1 + p.R[a][6] * h2, p.M[0]);
      const int q1 = pmodi(p.R[a][1] * h0 + p.R[a][4] * h1 + p.R[a][7] * h2, p.M[1]);
      const int q2 = pmodi(p.R[a][2] * h0 + p.R[a][5] * h1 + p.R[a][8] * h2, p.M[2]);
      const int c = p.chmap[(q0 & 1) | ((q1 & 1) << 1) | ((q2 & 1) << 2)];
      if (c >= 0) {
        const int j = h0 * p.t[a][0] * p.Pn[0] + h1 * p.t[a][1] * p.Pn[1] + h2 * p.t[a][2] * p.Pn[2];
        const cx<T> w = W[pmodi(-j, p.P)];
        const long k = (q0 >> 1) + (long)p.H[0] * ((q1 >> 1) + (long)p.H[1] * (q2 >> 1));
        acc = cadd(acc, cmul(w, cb[c * H3 + k]));
      }
    }
  }
  const T kk = K ? K[i] : T(1);
  F[tid] = {acc.x * kk, acc.y * kk};
}

// Inverse reconstruction on the alive parities, src/krfft.jl:1285-1311 and
// src/centering_fold.jl:782-813: G0(q) = (1/d) sum_alpha F(q + M alpha), the
// full-grid coefficients read through the spectral reverse lookup.
// Table entry: bits 0-19 spectral index (0xFFFFF = extinct), 20-30 phase
// exponent j (phase exp(2 pi i j/P)), bit 31 conjugate flag.
template <typename T>
__global__ void __launch_bounds__(256) k_irecon(const cx<T>* __restrict__ F, cx<T>* __restrict__ ch,
                                                const uint32_t* __restrict__ tab,
                                                const cx<T>* __restrict__ W, IreconP p, T scale) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= p.n_alive * p.batch) return;
  const int b = (int)(tid / p.n_alive);
  const long e = tid - (long)b * p.n_alive;
  const cx<T>* Fb = F + (long)b * p.n_spec;
  cx<T> acc = {T(0), T(0)};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    if (a < p.d) {
      const uint32_t ent = tab[a * p.n_alive + e];
      const uint32_t idx = ent & 0xFFFFFu;
      if (idx != 0xFFFFFu) {
        cx<T> v = Fb[idx];
        if (ent >> 31) v.y = -v.y;
        acc = cadd(acc, cmul(v, W[(ent >> 20) & 0x7FFu]));
      }
    }
  }
  ch[tid] = {acc.x * scale, acc.y * scale};
}

// Q-fused multiply on the alive parities, src/q_fused_krfft.jl:581-592:
// Y'(q) = sum_b Q_b(q) Y(R_b^T q mod M).  gidx holds the channel-space index
// (c * H^3 + k) of each gather source, or -1 for an extinct parity.
template <typename T>
__global__ void __launch_bounds__(256) k_qmul(const cx<T>* __restrict__ ch, cx<T>* __restrict__ out,
                                              const cx<T>* __restrict__ Q, const int* __restrict__ gidx,
                                              int d, long n_alive, int batch) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= n_alive * batch) return;
  const int b = (int)(tid / n_alive);
  const long e = tid - (long)b * n_alive;
  const cx<T>* cb = ch + (long)b * n_alive;
  cx<T> acc = {T(0), T(0)};
  for (int a = 0; a < d; ++a) {
    const int g = gidx[a * n_alive + e];
    if (g >= 0) acc = cadd(acc, cmul(Q[a * n_alive + e], cb[g]));
  }
  out[tid] = acc;
}

template <typename T>
__global__ void k_exp(const T* __restrict__ w, T* __restrict__ E, T factor, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) E[i] = exp(w[i] * factor);
}

template <typename T>
__global__ void k_full_to_sub(const T* __restrict__ u, T* __restrict__ f0, int M0, int M1, int M2,
                              int L0, int L1, int L2, int N0, int N1) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long M3 = (long)M0 * M1 * M2;
  if (i >= M3) return;
  const int x = (int)(i % M0), y = (int)((i / M0) % M1), z = (int)(i / ((long)M0 * M1));
  f0[i] = u[(long)x * L0 + (long)N0 * ((long)y * L1 + (long)N1 * ((long)z * L2))];
}

template <typename T>
__global__ void k_sub_to_full(const T* __restrict__ f0, T* __restrict__ u, const int* __restrict__ fmap,
                              long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) u[i] = f0[fmap[i]];
}

// G0 assembly (src/centering_fold.jl:277-296) for inspection: M-grid
// spectrum from the channels, zeros at extinct parities.
template <typename T>
__global__ void k_assemble(const cx<T>* __restrict__ ch, cx<T>* __restrict__ G0, FoldP p,
                           const int* __restrict__ chmap) {
  const long M3 = (long)p.M0 * p.M1 * p.M2;
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= M3 * p.batch) return;
  const int b = (int)(tid / M3);
  const long q = tid - b * M3;
  const int q0 = (int)(q % p.M0), q1 = (int)((q / p.M0) % p.M1), q2 = (int)(q / ((long)p.M0 * p.M1));
  const int c = chmap[(q0 & 1) | ((q1 & 1) << 1) | ((q2 & 1) << 2)];
  const long H3 = (long)p.H0 * p.H1 * p.H2;
  cx<T> v = {T(0), T(0)};
  if (c >= 0) v = ch[((long)b * p.n_ch + c) * H3 + (q0 >> 1) + (long)p.H0 * ((q1 >> 1) + (long)p.H1 * (q2 >> 1))];
  G0[tid] = v;
}

}  // namespace cfft
