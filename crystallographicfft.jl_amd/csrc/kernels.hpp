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

// exp(-2 pi i off.(x,y,z)/M) of the fold, in the table precision.
template <typename TW>
__device__ __forceinline__ cx<TW> fold_twiddle(int o, const cx<TW>* __restrict__ twM, int x, int y,
                                               int z, int M0, int M1) {
  cx<TW> t = {TW(1), TW(0)};
  if (o & 1) t = twM[x];
  if (o & 2) t = cmul(t, twM[M0 + y]);
  if (o & 4) t = cmul(t, twM[M0 + M1 + z]);
  return t;
}

__device__ __forceinline__ int pmodi(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ const T* fold_E(const FoldP& p, int b, long M3) {
  return p.E_contig ? (p.E[0] ? static_cast<const T*>(p.E[0]) + b * M3 : nullptr)
                    : static_cast<const T*>(p.E[b]);
}

// Workgroup geometry of the multi-lane kernels: R1 lanes per transform.
template <int LEN>
struct MLGeom {
  static constexpr int WG = (MLS<LEN>::R2 >= 16) ? 128 : 256;
  static constexpr int G = WG / MLS<LEN>::R1;  // transforms per workgroup
  static constexpr int GP = G + 1;              // padded LDS row
  static constexpr int XCH = MLS<LEN>::R1 > 1 ? MLS<LEN>::R2 * MLS<LEN>::R1 * GP : 1;
};

// Batched multi-lane FFT along y or z: transform i reads elements
// base + j*stride (j < LEN), base = (i % inner) + (i / inner) * outer.
template <typename T, typename TW, int LEN, int DIR>
__global__ void __launch_bounds__(MLGeom<LEN>::WG)
    k_fft_ml(cx<T>* __restrict__ data, long n_tr, long inner, long outer, long stride) {
  constexpr int R1 = MLS<LEN>::R1, R2 = MLS<LEN>::R2, G = MLGeom<LEN>::G, GP = MLGeom<LEN>::GP;
  __shared__ cx<T> xch[MLGeom<LEN>::XCH];
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  const long gid = (long)blockIdx.x * G + g;
  const bool valid = gid < n_tr;
  const long base = valid ? (gid % inner) + (gid / inner) * outer : 0;
  cx<T> a[R2];
#pragma unroll
  for (int m = 0; m < R2; ++m) a[m] = valid ? data[base + (long)(r + R1 * m) * stride] : cx<T>{T(0), T(0)};
  fft_ml<T, TW, LEN, DIR, GP>(a, r, g, xch);
  if (valid) {
#pragma unroll
    for (int q = 0; q < R2; ++q) data[base + (long)ml_out_pos<LEN>(r, q) * stride] = a[q];
  }
}

// Sum over the 8 lanes of an aligned lane octet.
template <typename T>
__device__ __forceinline__ cx<T> octet_sum(cx<T> v) {
#pragma unroll
  for (int off = 4; off >= 1; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
  }
  return v;
}

// Inverse reconstruction on the alive parities, src/krfft.jl:1285-1311 and
// src/centering_fold.jl:782-813: G0(q) = (1/d) sum_alpha F(q + M alpha), the
// full-grid coefficients read through the spectral reverse lookup.  Table
// layout [alive q][8 aliases]; entry bits 0-19 spectral index (0xFFFFF =
// extinct), 20-30 phase exponent j (phase exp(2 pi i j/P)), bit 31 conjugate.
template <typename T, typename TW>
__global__ void __launch_bounds__(256) k_irecon(const cx<T>* __restrict__ F, cx<T>* __restrict__ ch,
                                                const uint32_t* __restrict__ tab,
                                                const cx<TW>* __restrict__ W, IreconP p, TW scale) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int a = (int)(tid & 7);
  const long q = tid >> 3;
  const bool valid = q < p.n_alive * p.batch;
  const int b = valid ? (int)(q / p.n_alive) : 0;
  const long e = valid ? q - (long)b * p.n_alive : 0;
  cx<T> v = {T(0), T(0)};
  if (valid) {
    const uint32_t ent = tab[e * 8 + a];
    const uint32_t idx = ent & 0xFFFFFu;
    if (idx != 0xFFFFFu) {
      cx<T> f = F[(long)b * p.n_spec + idx];
      if (ent >> 31) f.y = -f.y;
      v = twmul<T, TW, -1>(f, W[(ent >> 20) & 0x7FFu]);
    }
  }
  v = octet_sum(v);
  if (valid && a == 0) ch[q] = {(T)((TW)v.x * scale), (T)((TW)v.y * scale)};
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
__global__ void k_exp(const T* __restrict__ w, double* __restrict__ E, double factor, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) E[i] = exp((double)w[i] * factor);
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
// spectrum from the channels (layout [b][c][x][z][y]), zeros at extinct
// parities.
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
  if (c >= 0)
    v = ch[((long)b * p.n_ch + c) * H3 + ((long)(q0 >> 1) * p.H2 + (q2 >> 1)) * p.H1 + (q1 >> 1)];
  G0[tid] = v;
}

}  // namespace cfft
