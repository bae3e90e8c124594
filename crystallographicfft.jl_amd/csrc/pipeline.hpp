// Five-phase KRFFT / diffusion pipeline on LDS tiles and planes (gfx950).
//
// Channel array layout ("XZY"): ch[b][c][x][z][y], y fastest, over the H-grid
// (H = M/2).  A (c, x) slice is one contiguous H2 x H1 plane, so the y and z
// transforms of a plane run in one workgroup out of LDS.
//
//   P1  fold + E + FFT x         tile (b, z, y-block)      f0 -> ch
//   P2  FFT y, FFT z             plane (b, c, x)           ch -> ch
//   P3  reconstruction + K       8 lanes per point         ch -> F
//   P4  inverse recon + IFFT z, IFFT y   plane (b, c, x)   F -> ch
//   P5  IFFT x + unfold + E      tile (b, z, y-block)      ch -> f0
//
// Each phase is a device function over one work item, used both by one
// kernel per phase and by the persistent sweep kernel (grid-wide barriers
// between phases).
#pragma once

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "fft.hpp"
#include "kernels.hpp"

namespace cfft {

constexpr int kNT = 256;  // threads per workgroup of the phase kernels

struct PipeP {
  int M[3], H[3];
  int n_ch, offs[8], chmap[8];
  int batch, YB;  // y rows per P1/P5 tile: a power of two dividing H1
  int ybs;        // log2(YB)
  int n_spec, d, P, Pn[3];
  int R[8][9], t[8][3];
  // interaction factors: entry b uses E1[b] for steps < sw[b], else E2[b]
  int E_contig;
  const void* E1[kMaxEPtr];
  const void* E2[kMaxEPtr];
  int sw[kMaxEPtr];
  double scale_unfold;  // 1 / (8 H0 H1 H2)
  double scale_irecon;  // 1 / d
  long n_rep;           // orbit representatives (path QR)
  int qs_items;         // plane-pair work items per batch entry (path QS)
  int qs_sep;           // path QS: t_{s,d} depends on bit d of s only
};

struct PipeBufs {
  const void* f0_in;  // P1 input
  void* f0_out;       // P5 output; P51 writes q(s + 1) there when it is non-null
  void* ch;           // channels
  void* F;            // spectral coefficients
  const int* hs;      // signed spectral frequencies (3 per point)
  const double* K;    // kernel (nullptr: 1)
  const void* W;      // exp(2 pi i j / P)
  const void* twM;    // fold twiddles (M0 + M1 + M2)
  const uint32_t* btab;  // inverse reconstruction table [e][8]
  const void* grep;      // path QR: Q-multiplied values at the orbit representatives
  const uint32_t* qmap;  // path QR: element -> rep (bits 0-19) | phase exponent << 20
  const void* qsep;      // path QS: axis factors F[d][s][q] (cx<double>, 3 x 8 x M0)
};

template <int LEN, int NT = kNT, int LANES = 0>
struct PipeG {
  static constexpr int R1 = MLS<LEN, LANES>::R1;
  static constexpr int G = NT / R1;  // transforms per pass of a workgroup of NT threads
};

// LDS row stride (in entries) of the transform tiles of the plane-pair and
// unfold/fold kernels for L a multiple of 16: L + 4 for 8-byte entries, L + 2
// for 16-byte entries (the rows and columns that neighbouring transform
// groups read then start on different banks; each measured fastest among
// L + 1, L + 2, L + 4); L + 1 otherwise.
template <typename T, int L>
constexpr int kRow = (L % 16 == 0) ? (sizeof(T) == 4 ? L + 4 : L + 2) : L + 1;

// FFT of n1 * n2 length-LEN transforms held in LDS: element j of transform
// t = t1 + n1 t2 is buf[t1 * ts1 + t2 * ts2 + j * es].  Every thread of the
// workgroup must call.
template <typename T, typename TW, int LEN, int DIR, int NT = kNT, int LANES = 0>
__device__ __forceinline__ void lds_fft2(cx<T>* buf, int n1, int ts1, int n2, int ts2, int es) {
  if constexpr (LEN > 1) {
    constexpr int R1 = MLS<LEN, LANES>::R1, R2 = MLS<LEN, LANES>::R2, G = PipeG<LEN, NT, LANES>::G;
    const int r = threadIdx.x % R1, g = threadIdx.x / R1;
    const int n_tr = n1 * n2;
    for (int t0 = 0; t0 < n_tr; t0 += G) {
      const int t = t0 + g;
      // groups past the last transform work on a copy of transform 0's
      // values in registers and store nothing: whole groups only, so the
      // in-slot exchange of the real groups is not disturbed
      const bool ok = t < n_tr;
      const int t2 = ok ? t / n1 : 0;
      cx<T>* slot = buf + (ok ? (t - t2 * n1) * ts1 + t2 * ts2 : 0);
      cx<T> v[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) v[m] = slot[(r + R1 * m) * es];
      if (ok) {
        fft_ml_lds<T, TW, LEN, DIR, LANES>(v, r, slot, es);
#pragma unroll
        for (int q = 0; q < R2; ++q) slot[ml_out_pos<LEN, LANES>(r, q) * es] = v[q];
      }
    }
  }
  __syncthreads();
}

// FFT of n_tr length-LEN transforms held in LDS: element j of transform t is
// buf[t * ts + j * es].
template <typename T, typename TW, int LEN, int DIR, int NT = kNT>
__device__ __forceinline__ void lds_fft(cx<T>* buf, int n_tr, int ts, int es) {
  lds_fft2<T, TW, LEN, DIR, NT>(buf, n_tr, ts, 1, 0, es);
}

// Interaction factors E = exp(-w ds/2) of batch entry b at contour step
// `step`: FP64 in both precisions (an FP32 rounding of E would be the same
// relative error at a point on every step, compounding along the contour
// instead of averaging out).  Every candidate is read before the choice
// (selects, no branches), so the reads issue together.
__device__ __forceinline__ const double* pipe_E(const PipeP& p, int b, int step, long M3) {
  const void* c0 = p.E1[0];
  const void* e1 = p.E1[b];
  const void* e2 = p.E2[b];
  const int sw = p.sw[b];
  const double* contig = c0 ? static_cast<const double*>(c0) + b * M3 : nullptr;
  const double* split = static_cast<const double*>(step < sw ? e1 : e2);
  return p.E_contig ? contig : split;
}

__device__ __forceinline__ long corner_index(const PipeP& p, int x, int y, int z, int e) {
  return (x + (e & 1) * p.H[0]) +
         (long)p.M[0] * ((y + ((e >> 1) & 1) * p.H[1]) + (long)p.M[1] * (z + ((e >> 2) & 1) * p.H[2]));
}

// Plane of L1 x L2 entries (y fastest, contiguous) into LDS rows of stride
// L1 + 1.  The trip count is a compile-time constant, so every global load is
// issued before the first LDS write.
template <typename T, int L1, int L2, int NT = kNT>
__device__ __forceinline__ void plane_to_lds(const cx<T>* __restrict__ src, cx<T>* lds) {
  constexpr int NP = L1 * L2, NIT = (NP + NT - 1) / NT;
  cx<T> r[NIT];
  // loads are unconditional (a clamped index on the ragged last pass) so that
  // none of them waits behind a branch
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * NT;
    r[it] = src[(NP % NT == 0 || i < NP) ? i : NP - 1];
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * NT;
    if (NP % NT == 0 || i < NP) lds[(i / L1) * (L1 + 1) + i % L1] = r[it];
  }
}

template <typename T, int L1, int L2, int NT = kNT>
__device__ __forceinline__ void lds_to_plane(const cx<T>* lds, cx<T>* __restrict__ dst) {
  constexpr int NP = L1 * L2, NIT = (NP + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * NT;
    if (NP % NT == 0 || i < NP) dst[i] = lds[(i / L1) * (L1 + 1) + i % L1];
  }
}

// Tile (b, z, y0 .. y0 + YB) of every channel into LDS A[c][yy][x] (row
// stride L0 + 1); n_ch * YB <= 64 bounds the compile-time trip count.
template <typename T, int L0>
__device__ __forceinline__ void tile_to_lds(const PipeP& p, const cx<T>* __restrict__ ch, int b, int z,
                                            int y0, cx<T>* lds) {
  constexpr int MAXIT = (64 * L0 + kNT - 1) / kNT;
  const int YB = p.YB, n = p.n_ch * L0 * YB;
  const long HH = (long)p.H[2] * p.H[1];
  cx<T> r[MAXIT];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int i0 = threadIdx.x + it * kNT;
    const int i = i0 < n ? i0 : n - 1;  // unconditional load, clamped index
    const int yy = i & (YB - 1), x = (i >> p.ybs) % L0, c = (i >> p.ybs) / L0;
    r[it] = ch[(((long)b * p.n_ch + c) * L0 + x) * HH + (long)z * p.H[1] + y0 + yy];
  }
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < n) {
      const int yy = i & (YB - 1), x = (i >> p.ybs) % L0, c = (i >> p.ybs) / L0;
      lds[(c * YB + yy) * (L0 + 1) + x] = r[it];
    }
  }
}

template <typename T, int L0>
__device__ __forceinline__ void lds_to_tile(const PipeP& p, const cx<T>* lds, cx<T>* __restrict__ ch, int b,
                                            int z, int y0) {
  constexpr int MAXIT = (64 * L0 + kNT - 1) / kNT;
  const int YB = p.YB, n = p.n_ch * L0 * YB;
  const long HH = (long)p.H[2] * p.H[1];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < n) {
      const int yy = i & (YB - 1), x = (i >> p.ybs) % L0, c = (i >> p.ybs) / L0;
      ch[(((long)b * p.n_ch + c) * L0 + x) * HH + (long)z * p.H[1] + y0 + yy] = lds[(c * YB + yy) * (L0 + 1) + x];
    }
  }
}

// Per-thread tile points (x, yy) of a (b, z, y-block) tile: NPT covers
// L0 * YB <= L0 * 8 points.  The interaction factors and the three fold
// twiddle factors of every point are loaded up front, so their latency
// overlaps the tile loads instead of following them.
template <int L0>
struct TilePts {
  static constexpr int NPT = (L0 * 8 + kNT - 1) / kNT;
};

template <typename TW>
struct FoldTw {
  cx<TW> x, y, z;
};

// exp(-2 pi i off.(x,y,z)/M) from the prefetched axis factors (same products
// as fold_twiddle).
template <typename TW>
__device__ __forceinline__ cx<TW> fold_tw(int o, const FoldTw<TW>& f) {
  cx<TW> t = {TW(1), TW(0)};
  if (o & 1) t = f.x;
  if (o & 2) t = cmul(t, f.y);
  if (o & 4) t = cmul(t, f.z);
  return t;
}

// FP64 table entry in the arithmetic precision TW
template <typename TW>
__device__ __forceinline__ cx<TW> tab(const cx<double>* __restrict__ t, long i) {
  const cx<double> v = t[i];
  return {(TW)v.x, (TW)v.y};
}

// The fold twiddle table is FP64 in both precisions; TW is the precision of
// the products formed with it.
template <typename TW>
__device__ __forceinline__ FoldTw<TW> load_fold_tw(const cx<double>* __restrict__ twM, int x, int y, int z,
                                                   const PipeP& p) {
  return {tab<TW>(twM, x), tab<TW>(twM, p.M[0] + y), tab<TW>(twM, p.M[0] + p.M[1] + z)};
}

// Channel values of one tile point (LDS entry base + c * cstride of channel c)
// into its 8 corner parities: v[e] = Re(g_c conj(tw_e)) for the channel c alive
// at parity e (chmap), 0 at extinct parities.  The parity loop is unrolled, so
// the twiddle products are static; the 8 LDS reads are unconditional (extinct
// parities read channel 0 and discard it) so that all of them issue at once.
template <typename T, typename TW>
__device__ __forceinline__ void gather_parities(const PipeP& p, const cx<T>* lds, int base, int cstride,
                                                const FoldTw<TW>& ft, T (&v)[8]) {
  cx<T> g[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int c = p.chmap[e];
    g[e] = lds[base + (c >= 0 ? c : 0) * cstride];
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const cx<TW> tw = fold_tw(e, ft);
    const T val = (T)((TW)g[e].x * tw.x + (TW)g[e].y * tw.y);
    v[e] = p.chmap[e] >= 0 ? val : T(0);
  }
}

// The alive parities of v into the channel entries of one tile point:
// g_c = v[e] tw_e for the channel c alive at parity e.
template <typename T, typename TW>
__device__ __forceinline__ void scatter_parities(const PipeP& p, cx<T>* lds, int base, int cstride,
                                                 const FoldTw<TW>& ft, const T (&v)[8]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int c = p.chmap[e];
    if (c >= 0) {
      const cx<TW> tw = fold_tw(e, ft);
      lds[base + c * cstride] = {(T)((TW)v[e] * tw.x), (T)((TW)v[e] * tw.y)};
    }
  }
}

// Tile item -> (y-block, z, batch entry): 32-bit unsigned arithmetic (item
// counts stay far below 2^31; 64-bit division would be emulated in scalar
// code ahead of the tile's first load).
__device__ __forceinline__ void tile_decode(unsigned item, unsigned nyb, unsigned H2, int& yb, int& z, int& b) {
  const unsigned q = item / nyb;
  yb = (int)(item - q * nyb);
  const unsigned bb = q / H2;
  z = (int)(q - bb * H2);
  b = (int)bb;
}

// ---- P1: fold + E + FFT x --------------------------------------------------
// tile item = (b, z, y-block); LDS A[c][yy][x] with row stride H0 + 1.
template <typename T, typename TW, int L0>
__device__ void ph_fold_fftx(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  const int nyb = H1 / YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)nyb, (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  const T* fb = static_cast<const T*>(B.f0_in) + b * M3;
  const double* Eb = pipe_E(p, b, step, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  const int rs = H0 + 1;
  constexpr int NPT = TilePts<L0>::NPT;
  T v[NPT][8];
  FoldTw<TW> ft[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long idx = corner_index(p, x, y, z, e);
      v[k][e] = Eb ? (T)((double)fb[idx] * Eb[idx]) : fb[idx];
    }
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0;
    wht8(v[k]);
    scatter_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v[k]);
  }
  __syncthreads();
  lds_fft<T, TW, L0, -1>(lds, p.n_ch * YB, rs, 1);
  // store to ch[b][c][x][z][y] (y fastest within a tile row)
  lds_to_tile<T, L0>(p, lds, static_cast<cx<T>*>(B.ch), b, z, y0);
  __syncthreads();
}

// Row transforms of one or two planes straight from / to global memory.
// Transform t < n_rows is row z = t mod NR of plane t / NR (p0 or p1, NR rows
// of L entries, contiguous); its lanes hold x[r + R1 m] on entry, exchange
// through the transform's own LDS row (stride L + 1, plane stride NR (L + 1))
// and leave the result there (from_global) or store it (to_global).  Groups
// past the last row work on row 0 and store nothing.
template <typename T, typename TW, int L, int DIR, int NT, int NR = L, int RS = L + 1, int LANES = 0>
__device__ __forceinline__ void rows_fft_from_global(const cx<T>* __restrict__ p0, const cx<T>* __restrict__ p1,
                                                     int n_rows, cx<T>* lds) {
  constexpr int R1 = MLS<L, LANES>::R1, R2 = MLS<L, LANES>::R2, G = PipeG<L, NT, LANES>::G, rs = RS;
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  for (int t0 = 0; t0 < n_rows; t0 += G) {
    const int t = t0 + g;
    const bool ok = t < n_rows;
    const int tt = ok ? t : 0;
    const int pl = tt / NR, z = tt - pl * NR;
    const cx<T>* src = (pl ? p1 : p0) + z * L;
    cx<T> v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = src[r + R1 * m];
    if (ok) {
      cx<T>* slot = lds + pl * (NR * rs) + z * rs;
      fft_ml_lds<T, TW, L, DIR, LANES>(v, r, slot, 1);
#pragma unroll
      for (int q = 0; q < R2; ++q) slot[ml_out_pos<L, LANES>(r, q)] = v[q];
    }
  }
}

template <typename T, typename TW, int L, int DIR, int NT, int NR = L, int RS = L + 1, int LANES = 0>
__device__ __forceinline__ void rows_fft_to_global(cx<T>* lds, cx<T>* __restrict__ p0, cx<T>* __restrict__ p1,
                                                   int n_rows) {
  constexpr int R1 = MLS<L, LANES>::R1, R2 = MLS<L, LANES>::R2, G = PipeG<L, NT, LANES>::G, rs = RS;
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  for (int t0 = 0; t0 < n_rows; t0 += G) {
    const int t = t0 + g;
    const bool ok = t < n_rows;
    const int tt = ok ? t : 0;
    const int pl = tt / NR, z = tt - pl * NR;
    cx<T>* slot = lds + pl * (NR * rs) + z * rs;
    cx<T> v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = slot[r + R1 * m];
    if (ok) {
      fft_ml_lds<T, TW, L, DIR, LANES>(v, r, slot, 1);
      cx<T>* dst = (pl ? p1 : p0) + z * L;
#pragma unroll
      for (int q = 0; q < R2; ++q) dst[ml_out_pos<L, LANES>(r, q)] = v[q];
    }
  }
}

// ---- P2: FFT y then FFT z of one (b, c, x) plane ----------------------------
template <typename T, typename TW, int L1, int L2>
__device__ void ph_fft_yz(const PipeP& p, const PipeBufs& B, long item, cx<T>* lds) {
  constexpr int rs = L1 + 1;
  cx<T>* plane = static_cast<cx<T>*>(B.ch) + item * ((long)L1 * L2);
  rows_fft_from_global<T, TW, L1, -1, kNT, L2>(plane, plane, L2, lds);  // FFT y, rows from the plane
  __syncthreads();
  lds_fft<T, TW, L2, -1>(lds, L1, 1, rs);
  lds_to_plane<T, L1, L2>(lds, plane);
  __syncthreads();
}

// ---- P3: forward reconstruction + K (one lane octet per spectral point) -----
// F(h) = K(h) sum_a exp(-2 pi i h.t_a/N) G0(R_a^T h mod M)   (src/krfft.jl:529-545)
template <typename T, typename TW>
__device__ __forceinline__ void ph_recon_point(const PipeP& p, const PipeBufs& B, long pt, bool valid) {
  const int a = threadIdx.x & 7;
  const int b = valid ? (int)(pt / p.n_spec) : 0;
  const int i = valid ? (int)(pt - (long)b * p.n_spec) : 0;
  cx<T> v = {T(0), T(0)};
  if (valid && a < p.d) {
    const int* hs = B.hs;
    const int h0 = hs[3 * i], h1 = hs[3 * i + 1], h2 = hs[3 * i + 2];
    const int q0 = pmodi(p.R[a][0] * h0 + p.R[a][3] * h1 + p.R[a][6] * h2, p.M[0]);
    const int q1 = pmodi(p.R[a][1] * h0 + p.R[a][4] * h1 + p.R[a][7] * h2, p.M[1]);
    const int q2 = pmodi(p.R[a][2] * h0 + p.R[a][5] * h1 + p.R[a][8] * h2, p.M[2]);
    const int c = p.chmap[(q0 & 1) | ((q1 & 1) << 1) | ((q2 & 1) << 2)];
    if (c >= 0) {
      const int j = h0 * p.t[a][0] * p.Pn[0] + h1 * p.t[a][1] * p.Pn[1] + h2 * p.t[a][2] * p.Pn[2];
      const long H3 = (long)p.H[0] * p.H[1] * p.H[2];
      const long k = ((long)(q0 >> 1) * p.H[2] + (q2 >> 1)) * p.H[1] + (q1 >> 1);
      const cx<T>* ch = static_cast<const cx<T>*>(B.ch);
      v = twmul<T, TW, +1>(ch[((long)b * p.n_ch + c) * H3 + k],
                           static_cast<const cx<TW>*>(B.W)[pmodi(j, p.P)]);
    }
  }
  v = octet_sum(v);
  if (valid && a == 0) {
    const TW kk = B.K ? (TW)B.K[i] : TW(1);
    static_cast<cx<T>*>(B.F)[pt] = {(T)((TW)v.x * kk), (T)((TW)v.y * kk)};
  }
}

// ---- P4: inverse reconstruction, IFFT z, IFFT y of one (b, c, x) plane ------
// G0(q) = (1/d) sum_alpha F(q + M alpha) through the reverse-lookup table
// (src/krfft.jl:1285-1311, src/centering_fold.jl:782-813).
template <typename T, typename TW, int L1, int L2>
__device__ void ph_irecon_ifft_zy(const PipeP& p, const PipeBufs& B, long item, cx<T>* lds) {
  constexpr int rs = L1 + 1, NP = L1 * L2, NIT = (NP + kNT - 1) / kNT;
  const long H3 = (long)p.H[0] * NP;
  const long n_alive = (long)p.n_ch * H3;
  const int b = (int)(item / ((long)p.n_ch * p.H[0]));
  const long e0 = item * (long)NP - (long)b * n_alive;  // channel-space start
  const cx<T>* Fb = static_cast<const cx<T>*>(B.F) + (long)b * p.n_spec;
  const cx<TW>* W = static_cast<const cx<TW>*>(B.W);
  const uint4* tab4 = reinterpret_cast<const uint4*>(B.btab);
  // all table rows of this thread first, then the gathers
  uint4 ta[NIT], tb[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < NP) {
      ta[it] = tab4[2 * (e0 + i)];
      tb[it] = tab4[2 * (e0 + i) + 1];
    }
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i >= NP) continue;
    const uint32_t ent[8] = {ta[it].x, ta[it].y, ta[it].z, ta[it].w, tb[it].x, tb[it].y, tb[it].z, tb[it].w};
    cx<T> acc = {T(0), T(0)};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const uint32_t idx = ent[a] & 0xFFFFFu;
      const bool ok = idx != 0xFFFFFu;
      cx<T> f = Fb[ok ? idx : 0];
      if (ent[a] >> 31) f.y = -f.y;
      const cx<T> prod = twmul<T, TW, -1>(f, W[(ent[a] >> 20) & 0x7FFu]);
      if (ok) acc = cadd(acc, prod);
    }
    lds[(i / L1) * rs + (i % L1)] = {(T)((TW)acc.x * p.scale_irecon), (T)((TW)acc.y * p.scale_irecon)};
  }
  __syncthreads();
  lds_fft<T, TW, L2, +1>(lds, L1, 1, rs);
  {  // IFFT y, each row stored straight from its transform lanes
    cx<T>* plane = static_cast<cx<T>*>(B.ch) + item * (long)NP;
    rows_fft_to_global<T, TW, L1, +1, kNT, L2>(lds, plane, plane, L2);
  }
  __syncthreads();
}

// Orbit expansion G0(e) = exp(2 pi i j_e / P) G0'(rep_e), then IFFT z, IFFT y
// of one (b, c, x) plane (path QR).
template <typename T, typename TW, int L1, int L2>
__device__ void ph_expand_ifft_zy(const PipeP& p, const PipeBufs& B, long item, cx<T>* lds) {
  constexpr int rs = L1 + 1, NP = L1 * L2, NIT = (NP + kNT - 1) / kNT;
  const long n_alive = (long)p.n_ch * p.H[0] * NP;
  const int b = (int)(item / ((long)p.n_ch * p.H[0]));
  const long e0 = item * (long)NP - (long)b * n_alive;
  const cx<T>* G = static_cast<const cx<T>*>(B.grep) + (long)b * p.n_rep;
  const cx<TW>* W = static_cast<const cx<TW>*>(B.W);
  uint32_t ent[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < NP) ent[it] = B.qmap[e0 + i];
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < NP)
      lds[(i / L1) * rs + (i % L1)] = twmul<T, TW, -1>(G[ent[it] & 0xFFFFFu], W[ent[it] >> 20]);
  }
  __syncthreads();
  lds_fft<T, TW, L2, +1>(lds, L1, 1, rs);
  {  // IFFT y, each row stored straight from its transform lanes
    cx<T>* plane = static_cast<cx<T>*>(B.ch) + item * (long)NP;
    rows_fft_to_global<T, TW, L1, +1, kNT, L2>(lds, plane, plane, L2);
  }
  __syncthreads();
}

// ---- P2Q: FFT y, z + Q multiply + IFFT z, y on plane pairs (path QS) -------
// Lines of work unit u along an axis of length L whose points q = 2k + off map
// to kbar = (L - k) mod L (off = 0) or L - 1 - k (off = 1): a pair (a, b) of
// distinct lines (pair = true) or self-images (pair = false).  For even L and
// off = 0 the two self-paired lines 0 and L/2 share unit 0; for odd L the one
// self-paired line is its own unit (a = b).  Units 0 .. (L + 1) / 2 - 1 cover
// every line once.
// c ? a : b on the components (a select of two array elements written as
// `c ? arr[i] : arr[j]` can become a select of addresses, which moves the
// array to scratch memory)
template <typename U>
__device__ __forceinline__ cx<U> csel(bool c, cx<U> a, cx<U> b) {
  return {c ? a.x : b.x, c ? a.y : b.y};
}

template <int L>
__device__ __forceinline__ void unit_lines(int u, int off, int& a, int& b, bool& pair) {
  if (off == 0) {
    if (u == 0) {
      a = 0;
      b = (L % 2 == 0) ? L / 2 : 0;
      pair = false;
    } else {
      a = u;
      b = L - u;
      pair = true;
    }
  } else if (L % 2 == 1 && u == (L - 1) / 2) {
    a = b = u;
    pair = false;
  } else {
    a = u;
    b = L - 1 - u;
    pair = true;
  }
}

// Work item (b, c, x0): the planes x0 and x1, the image of x0 under
// q0 -> -q0 mod M0 (one plane when x1 = x0).  The d = 8 representative ops are
// diagonal, R_s = diag(+-1) with flip pattern s (bit i: axis i negated), so the
// Q-fused multiply G0'(q) = sum_s Q_s(q) G0(R_s q) (src/q_fused_krfft.jl:
// 518-667) reads only the 8 sign images of q: same channel, planes {x0, x1},
// rows {z, zbar}, columns {y, ybar} -- the octet of q.  For an orthogonal cell
// Q_s(q) = F0[s][q0] F1[s][q1] F2[s][q2] with the axis factors
// F_d[s][q] = sum_{a in {0,1}} K_d(q + M_d a) exp(-2 pi i (q + M_d a) t_{s,d} / N_d)
// (1/8 folded into F0).  Octets partition the item's elements and each thread
// owns whole octets, so the multiply runs in place between the forward and
// the inverse transforms.  The multiply is FP64 in both precisions: factors
// rounded to FP32 would scale every mode by the same wrong amount on every
// contour step (about 1e-7 per step at the k = 0 mode, so 1e-5 in Q after
// 100 steps); in FP64 only the rounding of each result to the data type
// remains, which changes from step to step.
template <typename T, typename TW, int L, int NT, bool SEP, int LANES = 0>
__device__ __forceinline__ void ph_plane_pair_q(const PipeP& p, const PipeBufs& B, long item, cx<T>* lds, cx<double>* fq) {
  using TQ = double;  // arithmetic of the Q multiply
  constexpr int rs = kRow<T, L>, NP = L * L, SL = L * rs, NF = 2 + 2 * L;
  constexpr int NFT = 8 * NF, FIT = (NFT + NT - 1) / NT;
  // item -> (b, c, x0) in 32-bit arithmetic; the channel walk reads the
  // offsets at static indices, so they load together ahead of it
  const unsigned it = (unsigned)item, nq = (unsigned)p.qs_items;
  const int b = (int)(it / nq);
  int r = (int)(it - (unsigned)b * nq);
  // selects, no branches and no indexing by c afterwards (an offset array
  // indexed at run time would be copied to scratch memory)
  int c = 0, off = p.offs[0];
#pragma unroll
  for (int cc = 0; cc < 7; ++cc) {
    const int nc = (p.offs[cc] & 1) ? (L + 1) / 2 : L / 2 + 1;  // orbits of x -> xbar
    const bool adv = c == cc && cc < p.n_ch - 1 && r >= nc;
    r = adv ? r - nc : r;
    c = adv ? cc + 1 : c;
    off = adv ? p.offs[cc + 1] : off;
  }
  const int o0 = off & 1, o1 = (off >> 1) & 1, o2 = (off >> 2) & 1;
  const int x0 = r, x1 = o0 ? L - 1 - r : (L - r) % L;
  const int ns = x1 == x0 ? 1 : 2;
  cx<T>* chan = static_cast<cx<T>*>(B.ch) + ((long)b * p.n_ch + c) * ((long)L * NP);
  // this item's axis factors: fq[s * NF + j], j = 0, 1: x0, x1; 2 + y; 2 + L + z
  const cx<double>* F = static_cast<const cx<double>*>(B.qsep);  // FP64 in both precisions
  const int Ms = p.M[0];
  // The item's axis factors go straight from global memory into LDS
  // (global_load_lds: no registers held across the row transforms, nothing
  // to spill or to wait for ahead of the row loads); the barrier after the
  // row transforms waits for them.  fq holds FIT * NT entries: the lanes past
  // the last factor reload the last one into the padding.
  typedef __attribute__((address_space(3))) unsigned char lds_byte;
  static_assert(sizeof(cx<double>) == 16, "one 16-byte LDS DMA load per factor");
#pragma unroll
  for (int k = 0; k < FIT; ++k) {
    const int i0 = threadIdx.x + k * NT;
    const int i = i0 < NFT ? i0 : NFT - 1;  // clamped source index
    const int sf = i / NF, j = i - sf * NF;
    const int d = j < 2 ? 0 : (j < 2 + L ? 1 : 2);
    const int q = j < 2 ? 2 * (j ? x1 : x0) + o0 : (j < 2 + L ? 2 * (j - 2) + o1 : 2 * (j - 2 - L) + o2);
    const int slot0 = k * NT + ((int)threadIdx.x / 64) * 64;  // the wave's first slot
    lds_byte* dst = (lds_byte*)fq + 16 * slot0;
    __builtin_amdgcn_global_load_lds((const void*)(F + (d * 8 + sf) * Ms + q), dst, 16, 0, 0);
  }
  // FFT y: the rows of both planes, loaded straight into the transform lanes
  rows_fft_from_global<T, TW, L, -1, NT, L, rs, LANES>(chan + (long)x0 * NP, chan + (long)x1 * NP, ns * L, lds);
  __syncthreads();
  lds_fft2<T, TW, L, -1, NT, LANES>(lds, L, 1, ns, SL, rs);      // FFT z: columns of both planes
  if constexpr (SEP) {
    // t_{s,d} depends on bit d of s only: Q_s(q) = prod_d g_d(s_d, q_d) with
    // g_d(s_d, .) = F_d[s_d << d][.], and the octet sum factors into three
    // two-point stages (z, y, x).  Work units pair rows {z, zbar} and columns
    // {y, ybar}; with offset 0 the two self-paired lines 0 and L/2 share one
    // unit (each is its own partner), so an item has ((L + 1) / 2)^2 units
    constexpr int NU = (L + 1) / 2;
    for (int un = threadIdx.x; un < NU * NU; un += NT) {
      int ya, yb, za, zb;
      bool yp, zp;
      unit_lines<L>(un % NU, o1, ya, yb, yp);
      unit_lines<L>(un / NU, o2, za, zb, zp);
      int pos[8];
      cx<TQ> u[8];
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        pos[o] = ((o & 1) && ns == 2 ? SL : 0) + ((o & 4) ? zb : za) * rs + ((o & 2) ? yb : ya);
        const cx<T> g = lds[pos[o]];
        u[o] = {(TQ)g.x, (TQ)g.y};
      }
      const cx<TQ> gz[2][2] = {{fq[2 + L + za], fq[4 * NF + 2 + L + za]}, {fq[2 + L + zb], fq[4 * NF + 2 + L + zb]}};
      const cx<TQ> gy[2][2] = {{fq[2 + ya], fq[2 * NF + 2 + ya]}, {fq[2 + yb], fq[2 * NF + 2 + yb]}};
      const cx<TQ> gx[2][2] = {{fq[0], fq[NF]}, {fq[1], fq[NF + 1]}};
      cx<TQ> w[8];
#pragma unroll
      for (int o = 0; o < 8; ++o)
        w[o] = cadd(cmul(gz[o >> 2][0], u[o]), cmul(gz[o >> 2][1], csel(zp, u[o ^ 4], u[o])));
#pragma unroll
      for (int o = 0; o < 8; ++o)
        u[o] = cadd(cmul(gy[(o >> 1) & 1][0], w[o]), cmul(gy[(o >> 1) & 1][1], csel(yp, w[o ^ 2], w[o])));
      // a single plane (x0 = x1) loaded its entries twice: u[o ^ 1] is then u[o]'s twin
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const cx<TQ> a = cadd(cmul(gx[o & 1][0], u[o]), cmul(gx[o & 1][1], u[o ^ 1]));
        lds[pos[o]] = {(T)a.x, (T)a.y};
      }
    }
  } else {
    // the same work units; a self-paired line is its own image under the flip
    // of its axis, so the terms of sign patterns that differ only in that axis
    // read the same entry
    constexpr int NU = (L + 1) / 2;
    for (int un = threadIdx.x; un < NU * NU; un += NT) {
      int ya, yb, za, zb;
      bool yp, zp;
      unit_lines<L>(un % NU, o1, ya, yb, yp);
      unit_lines<L>(un / NU, o2, za, zb, zp);
      int pos[8];
      cx<T> v[8];
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        pos[o] = ((o & 1) && ns == 2 ? SL : 0) + ((o & 4) ? zb : za) * rs + ((o & 2) ? yb : ya);
        v[o] = lds[pos[o]];
      }
      cx<TQ> acc[8];
#pragma unroll
      for (int o = 0; o < 8; ++o) acc[o] = {TQ(0), TQ(0)};
      // sign patterns in four classes by their y and z flips; within a class
      // the entries src[k] = v[k with the flips of self-paired axes undone]
      // are read at the static index o ^ sf
#pragma unroll
      for (int cls = 0; cls < 4; ++cls) {
        cx<T> src[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (cls == 0) {
            src[k] = v[k];
          } else if (cls == 1) {
            src[k] = csel(yp, v[k], v[k ^ 2]);
          } else if (cls == 2) {
            src[k] = csel(zp, v[k], v[k ^ 4]);
          } else {
            src[k] = csel(zp, csel(yp, v[k], v[k ^ 2]), csel(yp, v[k ^ 4], v[k ^ 6]));
          }
        }
#pragma unroll
        for (int xb = 0; xb < 2; ++xb) {
          const int sf = 2 * cls + xb;
          const cx<TQ>* f = fq + sf * NF;
          const cx<TQ> fx0 = f[0], fx1 = f[1];
          const cx<TQ> fy0 = f[2 + ya], fy1 = f[2 + yb];
          const cx<TQ> fz0 = f[2 + L + za], fz1 = f[2 + L + zb];
          const cx<TQ> fxy[4] = {cmul(fx0, fy0), cmul(fx1, fy0), cmul(fx0, fy1), cmul(fx1, fy1)};
#pragma unroll
          for (int o = 0; o < 8; ++o) {
            const cx<TQ> qc = cmul(fxy[o & 3], (o & 4) ? fz1 : fz0);
            const cx<T> g = src[o ^ sf];
            acc[o].x += qc.x * (TQ)g.x - qc.y * (TQ)g.y;
            acc[o].y += qc.x * (TQ)g.y + qc.y * (TQ)g.x;
          }
        }
      }
      // coinciding positions (self-images) receive identical values
#pragma unroll
      for (int o = 0; o < 8; ++o) lds[pos[o]] = {(T)acc[o].x, (T)acc[o].y};
    }
  }
  __syncthreads();
  lds_fft2<T, TW, L, +1, NT, LANES>(lds, L, 1, ns, SL, rs);  // IFFT z
  // IFFT y, each row stored straight from its transform lanes
  rows_fft_to_global<T, TW, L, +1, NT, L, rs, LANES>(lds, chan + (long)x0 * NP, chan + (long)x1 * NP, ns * L);
  __syncthreads();
}

// IFFT z, IFFT y of one plane of `src` written to `dst` (Q-fused path).
template <typename T, typename TW, int L1, int L2>
__device__ void ph_ifft_zy(const PipeP& p, const PipeBufs& B, const cx<T>* src, long item,
                           cx<T>* lds) {
  constexpr int rs = L1 + 1;
  plane_to_lds<T, L1, L2>(src + item * ((long)L1 * L2), lds);
  __syncthreads();
  lds_fft<T, TW, L2, +1>(lds, L1, 1, rs);
  {  // IFFT y, each row stored straight from its transform lanes
    cx<T>* plane = static_cast<cx<T>*>(B.ch) + item * ((long)L1 * L2);
    rows_fft_to_global<T, TW, L1, +1, kNT, L2>(lds, plane, plane, L2);
  }
  __syncthreads();
}

// ---- P5: IFFT x + unfold + E (tile (b, z, y-block)) ------------------------
// f0(x + H eps) = scale * sum_c (-1)^(off_c.eps) Re(g_c(x) conj(tw_c))
// (src/centering_fold.jl:481-580); scale = 1/(8 H^3) (IFFT and inverse WHT).
template <typename T, typename TW, int L0>
__device__ void ph_ifftx_unfold(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  const int nyb = H1 / YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)nyb, (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  const int rs = H0 + 1;
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  T* fb = static_cast<T*>(B.f0_out) + b * M3;
  const double* Eb = pipe_E(p, b, step, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  constexpr int NPT = TilePts<L0>::NPT;
  double ev[NPT][8];
  FoldTw<TW> ft[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) ev[k][e] = Eb ? Eb[corner_index(p, x, y, z, e)] : 1.0;
  }
  tile_to_lds<T, L0>(p, static_cast<const cx<T>*>(B.ch), b, z, y0, lds);
  __syncthreads();
  lds_fft<T, TW, L0, +1>(lds, p.n_ch * YB, rs, 1);
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0, y = y0 + yy;
    T v[8];
    gather_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);  // Re(g conj(tw))
    wht8(v);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const T r = (T)((TW)v[e] * p.scale_unfold);
      fb[corner_index(p, x, y, z, e)] = Eb ? (T)((double)r * ev[k][e]) : r;
    }
  }
  __syncthreads();
}

// x transforms of the tile rows (c, yy) of tile (b, z, y0 .. y0 + YB) straight
// from / to the channel array: transform t < n_ch * YB is row t = c * YB + yy
// of the LDS tile A[c][yy][x] (row stride L0 + 1); its lanes hold x = r + R1 m
// (a wave's loads cover the same 128-byte row segments as tile_to_lds),
// exchange through the row and leave the result there (from_global) or
// store it (to_global).  Groups past the last row work on row 0 and store
// nothing.
// Wavefront-scope ordering of LDS accesses between the lanes of one wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transform of group g for the tile rows: WAVE = false, row t = g + t0 of the
// whole tile; WAVE = true, the rows of the wave's own tile points (wave w holds
// the points of rows RPW w .. RPW w + RPW - 1, RPW = 64 / L0): its group gi
// takes channel gi mod n_ch of row RPW w + gi / n_ch.
template <int L0, bool WAVE, int LANES = 0>
__device__ __forceinline__ bool tile_row_of(const PipeP& p, int t0, int g, int& tt) {
  if constexpr (WAVE) {
    constexpr int GW = 64 / MLS<L0, LANES>::R1, RPW = 64 / L0;  // groups and tile rows per wave
    const int w = g / GW, gi = g - w * GW;
    const int c = gi % p.n_ch, yy = RPW * w + gi / p.n_ch;
    const bool ok = gi < RPW * p.n_ch && yy < p.YB;
    tt = ok ? c * p.YB + yy : 0;
    return ok;
  } else {
    const int t = t0 + g;
    const bool ok = t < p.n_ch * p.YB;
    tt = ok ? t : 0;
    return ok;
  }
}

template <typename T, typename TW, int L0, int DIR, bool WAVE = false, int RS = L0 + 1, int LANES = 0>
__device__ __forceinline__ void tile_rows_fft_from_global(const PipeP& p, const cx<T>* __restrict__ ch, int b, int z,
                                                          int y0, cx<T>* lds) {
  constexpr int R1 = MLS<L0, LANES>::R1, R2 = MLS<L0, LANES>::R2, G = PipeG<L0, kNT, LANES>::G, rs = RS;
  const int n_rows = WAVE ? 1 : p.n_ch * p.YB;
  const long HH = (long)p.H[2] * p.H[1];
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  for (int t0 = 0; t0 < n_rows; t0 += G) {
    int tt;
    const bool ok = tile_row_of<L0, WAVE, LANES>(p, t0, g, tt);
    const int c = tt >> p.ybs, yy = tt & (p.YB - 1);
    const cx<T>* src = ch + ((long)b * p.n_ch + c) * L0 * HH + (long)z * p.H[1] + y0 + yy;
    cx<T> v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = src[(long)(r + R1 * m) * HH];
    if (ok) {
      cx<T>* slot = lds + tt * rs;
      fft_ml_lds<T, TW, L0, DIR, LANES>(v, r, slot, 1);
#pragma unroll
      for (int q = 0; q < R2; ++q) slot[ml_out_pos<L0, LANES>(r, q)] = v[q];
    }
  }
}

template <typename T, typename TW, int L0, int DIR, bool WAVE = false, int RS = L0 + 1, int LANES = 0>
__device__ __forceinline__ void tile_rows_fft_to_global(const PipeP& p, cx<T>* lds, cx<T>* __restrict__ ch, int b,
                                                        int z, int y0) {
  constexpr int R1 = MLS<L0, LANES>::R1, R2 = MLS<L0, LANES>::R2, G = PipeG<L0, kNT, LANES>::G, rs = RS;
  const int n_rows = WAVE ? 1 : p.n_ch * p.YB;
  const long HH = (long)p.H[2] * p.H[1];
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  for (int t0 = 0; t0 < n_rows; t0 += G) {
    int tt;
    const bool ok = tile_row_of<L0, WAVE, LANES>(p, t0, g, tt);
    cx<T>* slot = lds + tt * rs;
    cx<T> v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = slot[r + R1 * m];
    if (ok) {
      fft_ml_lds<T, TW, L0, DIR, LANES>(v, r, slot, 1);
      const int c = tt >> p.ybs, yy = tt & (p.YB - 1);
      cx<T>* dst = ch + ((long)b * p.n_ch + c) * L0 * HH + (long)z * p.H[1] + y0 + yy;
#pragma unroll
      for (int q = 0; q < R2; ++q) dst[(long)ml_out_pos<L0, LANES>(r, q) * HH] = v[q];
    }
  }
}

// ---- P5 of step s fused with P1 of step s + 1 ------------------------------
// IFFT x + unfold + E_s gives q(s + 1), written to f0_out when STORE (a sweep
// that keeps every contour point, or the caller's q; between two steps of a
// plain multi-step sweep q(s + 1) has no reader and stays in registers); the
// same thread then folds E_{s+1} q(s + 1) at the same tile points, so q(s + 1)
// is never read back.
template <typename T, typename TW, int L0, bool STORE, int LANES = 0>
__device__ void ph_unfold_fold(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  const int nyb = H1 / YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)nyb, (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  constexpr int rs = kRow<T, L0>;  // H0 = L0
  cx<T>* ch = static_cast<cx<T>*>(B.ch);
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  T* fb = STORE ? static_cast<T*>(B.f0_out) + b * M3 : nullptr;
  const double* Eo = pipe_E(p, b, step, M3);
  const double* En = pipe_E(p, b, step + 1, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  constexpr int NPT = TilePts<L0>::NPT;
  double eo[NPT][8], en[NPT][8];
  FoldTw<TW> ft[NPT];
  const bool same_E = En == Eo;
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long idx = corner_index(p, x, y, z, e);
      eo[k][e] = Eo ? Eo[idx] : 1.0;
      en[k][e] = same_E ? eo[k][e] : (En ? En[idx] : 1.0);
    }
  }
  // With one tile point per thread and 64 % H0 = 0 each wave's points are
  // whole tile rows; when the wave's groups can also take the x transforms of
  // those rows (every channel), the wave only ever touches its own LDS rows
  // and wavefront-scope ordering replaces the workgroup barriers.
  constexpr bool ROWS_PER_WAVE = (64 % L0 == 0) && (L0 * 8 <= kNT);
  const bool wave = ROWS_PER_WAVE && (64 / L0) * p.n_ch <= 64 / MLS<L0, LANES>::R1;
  // IFFT x of the tile rows, loaded straight into the transform lanes
  if (wave) {
    tile_rows_fft_from_global<T, TW, L0, +1, true, rs, LANES>(p, ch, b, z, y0, lds);
    wave_sync();
  } else {
    tile_rows_fft_from_global<T, TW, L0, +1, false, rs, LANES>(p, ch, b, z, y0, lds);
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0, y = y0 + yy;
    T v[8];
    gather_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);
    wht8(v);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      T r = (T)((TW)v[e] * p.scale_unfold);
      if (Eo) r = (T)((double)r * eo[k][e]);
      // a stored sweep's slice is read back only by the density pass, after
      // the sweep: a streaming store keeps it out of the caches the chain uses
      if constexpr (STORE) __builtin_nontemporal_store(r, fb + corner_index(p, x, y, z, e));
      v[e] = En ? (T)((double)r * en[k][e]) : r;
    }
    wht8(v);
    // the entries of this thread's point only: no barrier needed before reuse
    scatter_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);
  }
  // FFT x of the tile rows, each stored straight from its transform lanes
  if (wave) {
    wave_sync();
    tile_rows_fft_to_global<T, TW, L0, -1, true, rs, LANES>(p, lds, ch, b, z, y0);
  } else {
    __syncthreads();
    tile_rows_fft_to_global<T, TW, L0, -1, false, rs, LANES>(p, lds, ch, b, z, y0);
    __syncthreads();
  }
}

// ---- axis-split chain (path QS, separable multiply) ------------------------
// With an orthogonal cell and translations separable in the sign pattern the
// Q multiply is a product of three 2 x 2 mixings, one per axis: along axis d
// the value at channel frequency k (M-grid q = 2k + o_d) becomes
//   F_d[0][q] g(k) + F_d[1 << d][q] g(kbar),  kbar the index of -q mod M,
// with the axis factors F of the plane-pair kernel (1/8 folded into the x
// factors).  With the line transforms the step is then
//   E U (A_z A_y A_x) F E,   A_d = IFFT_d Q_d FFT_d  (F fold, U unfold),
// the A_d acting on different axes and commuting.  The tile kernel applies A_x
// to its x rows after the fold; the plane kernel applies A_y A_z to single
// x-planes (twice the work items of the plane pairs the octet multiply
// needs).  Between kernels the channels hold real-space values on every axis.

// Index of -q mod M on an axis of L channel indices with parity offset o.
template <int L>
__device__ __forceinline__ int axis_partner(int k, int o) {
  return o ? L - 1 - k : (k == 0 ? 0 : L - k);
}

// a u + b w formed in FP64 and rounded once to the data type.
template <typename T>
__device__ __forceinline__ cx<T> axis_mix(cx<double> a, cx<T> u, cx<double> b, cx<T> w) {
  const cx<double> ud = {(double)u.x, (double)u.y}, wd = {(double)w.x, (double)w.y};
  const cx<double> v = cadd(cmul(a, ud), cmul(b, wd));
  return {(T)v.x, (T)v.y};
}

// Parity offset of channel c (selects: the offsets are not indexed at run time).
__device__ __forceinline__ int chan_offset(const PipeP& p, int c) {
  int o = p.offs[0];
#pragma unroll
  for (int cc = 1; cc < 8; ++cc) o = c == cc ? p.offs[cc] : o;
  return o;
}

// Tile (b, z, y0 .. y0 + YB) of every channel into LDS A[c][yy][x], row stride RS.
template <typename T, int L0, int RS>
__device__ __forceinline__ void tile_to_lds_rs(const PipeP& p, const cx<T>* __restrict__ ch, int b, int z, int y0,
                                               cx<T>* lds) {
  constexpr int MAXIT = (64 * L0 + kNT - 1) / kNT;
  const int YB = p.YB, n = p.n_ch * L0 * YB;
  const long HH = (long)p.H[2] * p.H[1];
  cx<T> r[MAXIT];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int i0 = threadIdx.x + it * kNT;
    const int i = i0 < n ? i0 : n - 1;  // unconditional load, clamped index
    const int yy = i & (YB - 1), x = (i >> p.ybs) % L0, c = (i >> p.ybs) / L0;
    r[it] = ch[(((long)b * p.n_ch + c) * L0 + x) * HH + (long)z * p.H[1] + y0 + yy];
  }
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int i = threadIdx.x + it * kNT;
    if (i < n) {
      const int yy = i & (YB - 1), x = (i >> p.ybs) % L0, c = (i >> p.ybs) / L0;
      lds[(c * YB + yy) * RS + x] = r[it];
    }
  }
}

// A_x on the tile rows (c, yy) held in LDS (row stride RS): FFT x, the x
// mixing, IFFT x, each row stored from its transform lanes to ch.  Groups
// past the last row take no part (no workgroup barrier inside).
template <typename T, typename TW, int L0, int RS>
__device__ __forceinline__ void tile_rows_ax_to_global(const PipeP& p, const PipeBufs& B, cx<T>* lds,
                                                       cx<T>* __restrict__ ch, int b, int z, int y0) {
  constexpr int R1 = LineSplit<L0>::R1, R2 = LineSplit<L0>::R2, G = kNT / R1;
  const int n_rows = p.n_ch * p.YB;
  const long HH = (long)p.H[2] * p.H[1];
  const cx<double>* F = static_cast<const cx<double>*>(B.qsep);
  const int Ms = p.M[0];
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;
  for (int t0 = 0; t0 < n_rows; t0 += G) {
    const int t = t0 + g;
    if (t >= n_rows) continue;
    const int c = t >> p.ybs, yy = t & (p.YB - 1);
    const int o0 = chan_offset(p, c) & 1;
    cx<double> f0[R2], f1[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      const int q = 2 * (r + R1 * m) + o0;
      f0[m] = F[q];       // s = 0
      f1[m] = F[Ms + q];  // s = 1 (x flipped)
    }
    cx<T>* slot = lds + t * RS;
    cx<T> v[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) v[m] = slot[r + R1 * m];
    fft_line_lds<T, TW, L0, -1>(v, r, slot, 1);
#pragma unroll
    for (int j = 0; j < R2; ++j) slot[line_out_pos<L0>(r, j)] = v[j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      const int k = r + R1 * m;
      v[m] = axis_mix(f0[m], slot[k], f1[m], slot[axis_partner<L0>(k, o0)]);
    }
    fft_line_lds<T, TW, L0, +1>(v, r, slot, 1);
    cx<T>* dst = ch + ((long)b * p.n_ch + c) * L0 * HH + (long)z * p.H[1] + y0 + yy;
#pragma unroll
    for (int j = 0; j < R2; ++j) dst[(long)line_out_pos<L0>(r, j) * HH] = v[j];
  }
}

// First step: fold + E of slice 0 (as P1), then A_x; channels stored in real space.
template <typename T, typename TW, int L0>
__device__ void ph_fold_ax(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)(H1 / YB), (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  constexpr int rs = kRow<T, L0>;
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  const T* fb = static_cast<const T*>(B.f0_in) + b * M3;
  const double* Eb = pipe_E(p, b, step, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  constexpr int NPT = TilePts<L0>::NPT;
  T v[NPT][8];
  FoldTw<TW> ft[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long idx = corner_index(p, x, y, z, e);
      v[k][e] = Eb ? (T)((double)fb[idx] * Eb[idx]) : fb[idx];
    }
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0;
    wht8(v[k]);
    scatter_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v[k]);
  }
  __syncthreads();
  tile_rows_ax_to_global<T, TW, L0, rs>(p, B, lds, static_cast<cx<T>*>(B.ch), b, z, y0);
}

// Between two steps: unfold + E of step `step` (q(step + 1) written to
// f0_out when STORE), E of step + 1, fold, then A_x.
template <typename T, typename TW, int L0, bool STORE>
__device__ void ph_unfold_fold_ax(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)(H1 / YB), (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  constexpr int rs = kRow<T, L0>;
  cx<T>* ch = static_cast<cx<T>*>(B.ch);
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  T* fb = STORE ? static_cast<T*>(B.f0_out) + b * M3 : nullptr;
  const double* Eo = pipe_E(p, b, step, M3);
  const double* En = pipe_E(p, b, step + 1, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  constexpr int NPT = TilePts<L0>::NPT;
  double eo[NPT][8], en[NPT][8];
  FoldTw<TW> ft[NPT];
  const bool same_E = En == Eo;
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long idx = corner_index(p, x, y, z, e);
      eo[k][e] = Eo ? Eo[idx] : 1.0;
      en[k][e] = same_E ? eo[k][e] : (En ? En[idx] : 1.0);
    }
  }
  tile_to_lds_rs<T, L0, rs>(p, ch, b, z, y0, lds);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0, y = y0 + yy;
    T v[8];
    gather_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);
    wht8(v);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      T r = (T)((TW)v[e] * p.scale_unfold);
      if (Eo) r = (T)((double)r * eo[k][e]);
      if constexpr (STORE) __builtin_nontemporal_store(r, fb + corner_index(p, x, y, z, e));
      v[e] = En ? (T)((double)r * en[k][e]) : r;
    }
    wht8(v);
    // the entries of this thread's point only
    scatter_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);
  }
  __syncthreads();
  tile_rows_ax_to_global<T, TW, L0, rs>(p, B, lds, ch, b, z, y0);
}

// Last step: unfold + E of real-space channels into f0_out.
template <typename T, typename TW, int L0>
__device__ void ph_unfold_ax(const PipeP& p, const PipeBufs& B, long item, int step, cx<T>* lds) {
  const int H0 = p.H[0], H1 = p.H[1], H2 = p.H[2], YB = p.YB;
  int yb, z, b;
  tile_decode((unsigned)item, (unsigned)(H1 / YB), (unsigned)H2, yb, z, b);
  const int y0 = yb * YB;
  constexpr int rs = kRow<T, L0>;
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  T* fb = static_cast<T*>(B.f0_out) + b * M3;
  const double* Eb = pipe_E(p, b, step, M3);
  const cx<double>* twM = static_cast<const cx<double>*>(B.twM);
  constexpr int NPT = TilePts<L0>::NPT;
  double ev[NPT][8];
  FoldTw<TW> ft[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt0 = threadIdx.x + k * kNT;
    const int pt = pt0 < H0 * YB ? pt0 : H0 * YB - 1;  // unconditional loads
    const int x = pt % H0, y = y0 + pt / H0;
    ft[k] = load_fold_tw<TW>(twM, x, y, z, p);
#pragma unroll
    for (int e = 0; e < 8; ++e) ev[k][e] = Eb ? Eb[corner_index(p, x, y, z, e)] : 1.0;
  }
  tile_to_lds_rs<T, L0, rs>(p, static_cast<const cx<T>*>(B.ch), b, z, y0, lds);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int pt = threadIdx.x + k * kNT;
    if (pt >= H0 * YB) continue;
    const int x = pt % H0, yy = pt / H0, y = y0 + yy;
    T v[8];
    gather_parities<T, TW>(p, lds, yy * rs + x, YB * rs, ft[k], v);
    wht8(v);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const T r = (T)((TW)v[e] * p.scale_unfold);
      fb[corner_index(p, x, y, z, e)] = Eb ? (T)((double)r * ev[k][e]) : r;
    }
  }
}

// A_y A_z on the x-plane `item` = (b, c, x) of the channels (L x L entries,
// y fastest): rows (z) FFT y into A; columns (y) the y mixing on read, FFT z,
// the z mixing, IFFT z into B (each column group in its own column); rows
// IFFT y from B, stored.  L * R1 threads: one line per lane group in each pass.
template <typename T, typename TW, int L>
__device__ void ph_plane_ax(const PipeP& p, const PipeBufs& B, long item, cx<T>* A, cx<T>* Bv) {
  constexpr int R1 = LineSplit<L>::R1, R2 = LineSplit<L>::R2, rs = kRow<T, L>;
  const int c = (int)((unsigned)(item / L) % (unsigned)p.n_ch);
  const int off = chan_offset(p, c), o1 = (off >> 1) & 1, o2 = (off >> 2) & 1;
  cx<T>* plane = static_cast<cx<T>*>(B.ch) + item * (long)(L * L);
  const cx<double>* F = static_cast<const cx<double>*>(B.qsep);
  const int Ms = p.M[0];
  const int r = threadIdx.x % R1, g = threadIdx.x / R1;  // g < L
  const int gb = axis_partner<L>(g, o1);
  const cx<double> fy0 = F[(8 + 0) * Ms + 2 * g + o1], fy1 = F[(8 + 2) * Ms + 2 * g + o1];
  cx<double> fz0[R2], fz1[R2];
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int q = 2 * (r + R1 * m) + o2;
    fz0[m] = F[(16 + 0) * Ms + q];
    fz1[m] = F[(16 + 4) * Ms + q];
  }
  cx<T> v[R2];
  // FFT y: row z = g, straight from global memory
#pragma unroll
  for (int m = 0; m < R2; ++m) v[m] = plane[g * L + r + R1 * m];
  fft_line_lds<T, TW, L, -1>(v, r, A + g * rs, 1);
#pragma unroll
  for (int j = 0; j < R2; ++j) A[g * rs + line_out_pos<L>(r, j)] = v[j];
  __syncthreads();
  // column y = g: the y mixing on read, FFT z, the z mixing, IFFT z
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int z = r + R1 * m;
    v[m] = axis_mix(fy0, A[z * rs + g], fy1, A[z * rs + gb]);
  }
  fft_line_lds<T, TW, L, -1>(v, r, Bv + g, rs);
#pragma unroll
  for (int j = 0; j < R2; ++j) Bv[line_out_pos<L>(r, j) * rs + g] = v[j];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int k = r + R1 * m;
    v[m] = axis_mix(fz0[m], Bv[k * rs + g], fz1[m], Bv[axis_partner<L>(k, o2) * rs + g]);
  }
  fft_line_lds<T, TW, L, +1>(v, r, Bv + g, rs);
#pragma unroll
  for (int j = 0; j < R2; ++j) Bv[line_out_pos<L>(r, j) * rs + g] = v[j];
  __syncthreads();
  // IFFT y: row z = g, stored from its transform lanes
#pragma unroll
  for (int m = 0; m < R2; ++m) v[m] = Bv[g * rs + r + R1 * m];
  fft_line_lds<T, TW, L, +1>(v, r, Bv + g * rs, 1);
#pragma unroll
  for (int j = 0; j < R2; ++j) plane[g * L + line_out_pos<L>(r, j)] = v[j];
}

// ---- kernels: one per phase -------------------------------------------------
template <typename T, typename TW, int L0>
__global__ void __launch_bounds__(kNT) k_p1(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * (L0 + 1)];
  ph_fold_fftx<T, TW, L0>(p, B, blockIdx.x, step, lds);
}

template <typename T, typename TW, int L1, int L2>
__global__ void __launch_bounds__(kNT) k_p2(PipeP p, PipeBufs B) {
  __shared__ cx<T> lds[L2 * (L1 + 1)];
  ph_fft_yz<T, TW, L1, L2>(p, B, blockIdx.x, lds);
}

template <typename T, typename TW>
__global__ void __launch_bounds__(kNT) k_p3(PipeP p, PipeBufs B) {
  const long pt = ((long)blockIdx.x * kNT + threadIdx.x) >> 3;
  ph_recon_point<T, TW>(p, B, pt, pt < (long)p.n_spec * p.batch);
}

template <typename T, typename TW, int L1, int L2>
__global__ void __launch_bounds__(kNT) k_p4(PipeP p, PipeBufs B) {
  __shared__ cx<T> lds[L2 * (L1 + 1)];
  ph_irecon_ifft_zy<T, TW, L1, L2>(p, B, blockIdx.x, lds);
}

template <typename T, typename TW, int L1, int L2>
__global__ void __launch_bounds__(kNT) k_p4q(PipeP p, PipeBufs B, const cx<T>* src) {
  __shared__ cx<T> lds[L2 * (L1 + 1)];
  ph_ifft_zy<T, TW, L1, L2>(p, B, src, blockIdx.x, lds);
}

template <typename T, typename TW, int L0>
__global__ void __launch_bounds__(kNT) k_p5(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * (L0 + 1)];
  ph_ifftx_unfold<T, TW, L0>(p, B, blockIdx.x, step, lds);
}

template <typename T, typename TW, int L1, int L2>
__global__ void __launch_bounds__(kNT) k_p4r(PipeP p, PipeBufs B) {
  __shared__ cx<T> lds[L2 * (L1 + 1)];
  ph_expand_ifft_zy<T, TW, L1, L2>(p, B, blockIdx.x, lds);
}

// Q multiply at the orbit representatives (path QR):
// G0'(rep) = sum_b Q_b(rep) G0(R_b^T rep), coefficients in FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_qrep(const cx<T>* __restrict__ ch, cx<T>* __restrict__ grep,
                                              const cx<double>* __restrict__ Q, const int* __restrict__ gidx,
                                              int d, long n_rep, long n_alive, int batch) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= n_rep * batch) return;
  const int b = (int)(tid / n_rep);
  const long r = tid - (long)b * n_rep;
  const cx<T>* cb = ch + (long)b * n_alive;
  // d <= 8 (L in {1,2}^3): all index and coefficient loads first, then the gathers
  int g[8];
  cx<double> q[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    g[a] = a < d ? gidx[a * n_rep + r] : -1;
    q[a] = a < d ? Q[a * n_rep + r] : cx<double>{0.0, 0.0};
  }
  cx<T> v[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) v[a] = g[a] >= 0 ? cb[g[a]] : cx<T>{T(0), T(0)};
  double ax = 0.0, ay = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    ax += q[a].x * (double)v[a].x - q[a].y * (double)v[a].y;
    ay += q[a].x * (double)v[a].y + q[a].y * (double)v[a].x;
  }
  grep[tid] = {(T)ax, (T)ay};
}

// Orbit expansion on the channel array (non-cubic path QR).
template <typename T>
__global__ void __launch_bounds__(256) k_expand(const cx<T>* __restrict__ grep, cx<T>* __restrict__ ch,
                                                const uint32_t* __restrict__ qmap,
                                                const cx<double>* __restrict__ W, long n_rep, long n_alive,
                                                int batch) {
  const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (tid >= n_alive * batch) return;
  const int b = (int)(tid / n_alive);
  const uint32_t ent = qmap[tid - (long)b * n_alive];
  ch[tid] = twmul<T, double, -1>(grep[(long)b * n_rep + (ent & 0xFFFFFu)], W[ent >> 20]);
}

// Threads per workgroup of the plane-pair kernel by default (512, a schedule
// the tuner may choose, gained 3 % at 2 fields per launch and lost a third
// at 16: one workgroup per CU instead of two).
constexpr int kQNT = 256;

// SEP: the separable multiply (p.qs_sep) or the general one, compiled apart so
// that each gets its own register allocation.  NT threads per workgroup;
// LANES: the lane split of the row and column transforms (0: the default).
template <typename T, typename TW, int L, bool SEP, int NT = kQNT, int LANES = 0>
__global__ void __launch_bounds__(NT) k_p2q(PipeP p, PipeBufs B) {
  constexpr int NFQ = (8 * (2 + 2 * L) + NT - 1) / NT * NT;  // whole loads of NT factors
  __shared__ cx<T> lds[2 * L * kRow<T, L>];
  __shared__ cx<double> fq[NFQ];  // axis factors, FP64 in both precisions
  ph_plane_pair_q<T, TW, L, NT, SEP, LANES>(p, B, blockIdx.x, lds, fq);
}

// Tile item of workgroup `blk` in a grid of `nblk` (one workgroup per item).
// Items 2j and 2j + 1 are neighbouring y-blocks of one (b, z) and share the
// 128-byte lines of the channel rows when a row of YB values is shorter than
// a line (`half_rows`: FP32 with YB = 8).  Workgroups go to the 8 XCDs round
// robin and each XCD has its own L2, so blocks i and i + 8 of every 16 take
// items 2j and 2j + 1: the pair then shares an XCD and its lines are fetched
// once.  A bijection on [0, nblk) when nblk % 16 == 0 and the y-block count is
// even, else identity; identity too for rows of whole lines (FP64 measured
// 1 % slower with the pairing, 4-6 % faster for FP32).
__device__ inline long p51_item(long blk, long nblk, int nyb, bool half_rows) {
  if (!half_rows || (nyb & 1) || (nblk & 15)) return blk;
  const int r = (int)(blk & 15);
  return (blk & ~15L) + ((r & 7) << 1) + (r >> 3);
}

// STORE: q(s + 1) written to B.f0_out, or (between two steps of a plain
// sweep) not written at all.
template <typename T, typename TW, int L0, bool STORE, int LANES = 0>
__global__ void __launch_bounds__(kNT) k_p51(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * kRow<T, L0>];
  const long item = p51_item(blockIdx.x, gridDim.x, p.H[1] / p.YB, p.YB * (int)sizeof(cx<T>) < 128);
  ph_unfold_fold<T, TW, L0, STORE, LANES>(p, B, item, step, lds);
}

// Axis-split chain kernels (path QS with CFFT_OPT_AXIS_SPLIT).
template <typename T, typename TW, int L0>
__global__ void __launch_bounds__(kNT) k_p1x(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * kRow<T, L0>];
  ph_fold_ax<T, TW, L0>(p, B, blockIdx.x, step, lds);
}

template <typename T, typename TW, int L>
__global__ void __launch_bounds__(L * LineSplit<L>::R1) k_p2x(PipeP p, PipeBufs B) {
  __shared__ cx<T> A[L * kRow<T, L>];
  __shared__ cx<T> Bv[L * kRow<T, L>];
  ph_plane_ax<T, TW, L>(p, B, blockIdx.x, A, Bv);
}

template <typename T, typename TW, int L0, bool STORE>
__global__ void __launch_bounds__(kNT) k_p51x(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * kRow<T, L0>];
  // neighbouring y-blocks of FP32 rows share 128-byte lines: same XCD
  const long item = p51_item(blockIdx.x, gridDim.x, p.H[1] / p.YB, p.YB * (int)sizeof(cx<T>) < 128);
  ph_unfold_fold_ax<T, TW, L0, STORE>(p, B, item, step, lds);
}

template <typename T, typename TW, int L0>
__global__ void __launch_bounds__(kNT) k_p5x(PipeP p, PipeBufs B, int step) {
  __shared__ cx<T> lds[64 * kRow<T, L0>];
  ph_unfold_ax<T, TW, L0>(p, B, blockIdx.x, step, lds);
}

// ---- persistent sweep: nsteps Strang steps in one cooperative launch ---------
// f0 is advanced in place (P1 reads it, P5 overwrites it after the barrier
// that follows P4).  Work items are distributed grid-stride over the resident
// workgroups; every workgroup reaches every barrier.
template <typename T, typename TW, int L>
__global__ void __launch_bounds__(kNT) k_sweep(PipeP p, PipeBufs B, int step0, int nsteps) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  constexpr int TILE = 64 * (L + 1);
  constexpr int PLANE = L * (L + 1);
  __shared__ cx<T> lds[TILE > PLANE ? TILE : PLANE];
  const long n_tiles = (long)p.batch * p.H[2] * (p.H[1] / p.YB);
  const long n_planes = (long)p.batch * p.n_ch * p.H[0];
  const long n_pts = (long)p.n_spec * p.batch;
  const long n_oct = (n_pts * 8 + kNT - 1) / kNT;  // workgroup-chunks of P3
  for (int s = 0; s < nsteps; ++s) {
    const int step = step0 + s;
    for (long it = blockIdx.x; it < n_tiles; it += gridDim.x) ph_fold_fftx<T, TW, L>(p, B, it, step, lds);
    grid.sync();
    for (long it = blockIdx.x; it < n_planes; it += gridDim.x) ph_fft_yz<T, TW, L, L>(p, B, it, lds);
    grid.sync();
    for (long it = blockIdx.x; it < n_oct; it += gridDim.x) {
      const long pt = (it * kNT + threadIdx.x) >> 3;
      ph_recon_point<T, TW>(p, B, pt, pt < n_pts);
    }
    grid.sync();
    for (long it = blockIdx.x; it < n_planes; it += gridDim.x) ph_irecon_ifft_zy<T, TW, L, L>(p, B, it, lds);
    grid.sync();
    for (long it = blockIdx.x; it < n_tiles; it += gridDim.x) ph_ifftx_unfold<T, TW, L>(p, B, it, step, lds);
    grid.sync();
  }
}

}  // namespace cfft
