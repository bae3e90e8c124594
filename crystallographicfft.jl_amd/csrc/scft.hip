// SCFT driver of an incompressible AB diblock melt on the stride-L subgrid:
// propagator sweeps kept on the device (q and q+ as one batch of two fields,
// every contour point stored), Simpson densities, free energy, the field
// update and Anderson mixing.  FP64 everywhere except the propagator slices,
// which have the plan's precision.
//
// No counterpart in the reference, which stops at the diffusion step
// (src/diffusion_solver.jl:254-282).  Textbook equations (Matsen & Schick,
// PRL 72, 2660 (1994)); Anderson mixing after Matsen, Eur. Phys. J. E 30,
// 361 (2009), eqs. 31-35.  Built on the public entry points of the plan
// (cfft_plan_info, cfft_sweep_hist_dev) only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <vector>

#include "../../include/cfft_hip.h"
#include "common.hpp"

using namespace cfft;

namespace {

constexpr int kRB = 256;      // threads per reduction block
constexpr int kMaxRed = 1024;  // reduction blocks
constexpr int kMaxHist = 32;   // Anderson history slots

// Sum NV values over the workgroup; thread 0 writes part[blockIdx.x * NV + v].
template <int NV>
__device__ void block_sum(double (&v)[NV], double* __restrict__ part) {
  __shared__ double red[kRB / 64][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off >= 1; off >>= 1) v[k] += __shfl_down(v[k], off);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = 0.0;
      for (int w = 0; w < kRB / 64; ++w) a += red[w][k];
      part[blockIdx.x * NV + k] = a;
    }
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, T v, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// E = exp(factor * w) from FP64 w into the plan's precision.
template <typename T>
__global__ void k_expw(const double* __restrict__ w, T* __restrict__ E, double factor, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) E[i] = (T)exp(factor * w[i]);
}

// Unnormalised densities from the stored propagators (slice j = [q(j), q+(j)]):
//   uA = sum_{j<=nA} wa_j q(j) q+(ns-j),  uB = sum_{j>=nA} wb_{j-nA} q(j) q+(ns-j);
// partial sums of q(ns) and q+(ns) for Q.
template <typename T>
__global__ void __launch_bounds__(kRB) k_density(const T* __restrict__ hist, long M3, int ns, int nA,
                                                 const double* __restrict__ wa, const double* __restrict__ wb,
                                                 double* __restrict__ u, double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (long x = blockIdx.x * (long)kRB + threadIdx.x; x < M3; x += (long)gridDim.x * kRB) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j <= ns; ++j) {
      const double pr = (double)hist[(2L * j) * M3 + x] * (double)hist[(2L * (ns - j) + 1) * M3 + x];
      if (j <= nA) a += wa[j] * pr;
      if (j >= nA) b += wb[j - nA] * pr;
    }
    u[x] = a;
    u[M3 + x] = b;
    acc[0] += (double)hist[(2L * ns) * M3 + x];
    acc[1] += (double)hist[(2L * ns + 1) * M3 + x];
  }
  block_sum<2>(acc, part);
}

// Field update and the newest row of the Anderson Gram matrix in one pass.
// Q = qtot[0] / M3 (the total of q(ns)); phi <- (ds / Q) u; d = t - w; the
// history slot `cur` of the rings receives (w, d).  With xi = (wA + wB - chiN)/2
// the targets are tA = chiN phiB + xi, tB = chiN phiA + xi.  Partial sums per
// block: phiA phiB, wA phiA, wB phiB, phiA + phiB, |d|^2, |w|^2, phiA, then
// <d, D_k> over the ring slots gs.slot (for the newest, cur, that is <d, d>).
struct GramSlots {
  int n;
  int cur;
  int slot[kMaxHist];
};
constexpr int kNU = 7 + kMaxHist;  // partial sums per block of k_update_gram
__global__ void __launch_bounds__(kRB) k_update_gram(const double* __restrict__ w, double* __restrict__ phi,
                                                     double* __restrict__ d, const double* __restrict__ qtot,
                                                     double ds, double chiN, long M3,
                                                     double* __restrict__ Wr, double* __restrict__ Dr, GramSlots gs,
                                                     double* __restrict__ part) {
  const double c = ds / (qtot[0] / (double)M3);
  const long n2 = 2 * M3;
  double acc[kNU];
#pragma unroll
  for (int k = 0; k < kNU; ++k) acc[k] = 0.0;
  for (long x = blockIdx.x * (long)kRB + threadIdx.x; x < M3; x += (long)gridDim.x * kRB) {
    const double pA = c * phi[x], pB = c * phi[M3 + x];
    phi[x] = pA;
    phi[M3 + x] = pB;
    const double wA = w[x], wB = w[M3 + x];
    const double xi = 0.5 * (wA + wB - chiN);
    const double dA = chiN * pB + xi - wA, dB = chiN * pA + xi - wB;
    d[x] = dA;
    d[M3 + x] = dB;
    Wr[gs.cur * n2 + x] = wA;
    Wr[gs.cur * n2 + M3 + x] = wB;
    Dr[gs.cur * n2 + x] = dA;
    Dr[gs.cur * n2 + M3 + x] = dB;
    acc[0] += pA * pB;
    acc[1] += wA * pA;
    acc[2] += wB * pB;
    acc[3] += pA + pB;
    acc[4] += dA * dA + dB * dB;
    acc[5] += wA * wA + wB * wB;
    acc[6] += pA;
#pragma unroll
    for (int k = 0; k < kMaxHist; ++k)
      if (k < gs.n) {
        const int sl = gs.slot[k];
        acc[7 + k] += sl == gs.cur ? dA * dA + dB * dB
                                   : dA * Dr[sl * n2 + x] + dB * Dr[sl * n2 + M3 + x];
      }
  }
  block_sum<kNU>(acc, part);
}

// Totals of NV partial sums over nb <= kMaxRed blocks: workgroup k sums
// quantity k, each thread a strided share of the blocks, then a fixed tree
// over the workgroup (the same association on every call).
template <int NV>
__global__ void __launch_bounds__(kRB) k_total(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[kRB];
  const int k = blockIdx.x;
  double a = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRB) a += part[b * NV + k];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int h = kRB / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

struct MixP {
  int k;              // history entries used (besides the newest)
  int cur;            // slot of the newest entry
  int slot[kMaxHist];  // slot of entry n - i, i = 1..k
  double c[kMaxHist];
  double lam;
};

// w = w_n + sum c_i (W_{n-i} - w_n) + lam (d_n + sum c_i (D_{n-i} - d_n)).
__global__ void k_mix(double* __restrict__ w, const double* __restrict__ W, const double* __restrict__ D,
                      MixP m, long n) {
  const long x = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double wn = W[m.cur * n + x], dn = D[m.cur * n + x];
  double wp = wn, dp = dn;
  for (int i = 0; i < m.k; ++i) {
    wp += m.c[i] * (W[m.slot[i] * n + x] - wn);
    dp += m.c[i] * (D[m.slot[i] * n + x] - dn);
  }
  w[x] = wp + m.lam * dp;
}

// Simpson weights on n + 1 points (trapezoid on the last interval when n is odd).
std::vector<double> simpson(int n) {
  std::vector<double> w(n + 1, 0.0);
  if (n == 0) return w;
  if (n % 2 == 0) {
    for (int j = 0; j <= n; ++j) w[j] = (j == 0 || j == n) ? 1.0 : (j % 2 ? 4.0 : 2.0);
    for (auto& v : w) v /= 3.0;
    return w;
  }
  std::vector<double> a = simpson(n - 1);
  for (int j = 0; j < n; ++j) w[j] = a[j];
  w[n - 1] += 0.5;
  w[n] += 0.5;
  return w;
}

// Least squares U c = V by Gaussian elimination with partial pivoting; false
// when U is numerically singular.
bool solve_small(std::vector<double> U, std::vector<double> V, int k, std::vector<double>& c) {
  double scale = 0.0;
  for (int i = 0; i < k; ++i) scale = std::max(scale, std::fabs(U[i * k + i]));
  if (scale == 0.0) return false;
  for (int col = 0; col < k; ++col) {
    int piv = col;
    for (int r = col + 1; r < k; ++r)
      if (std::fabs(U[r * k + col]) > std::fabs(U[piv * k + col])) piv = r;
    if (std::fabs(U[piv * k + col]) < 1e-14 * scale) return false;
    if (piv != col) {
      for (int j = 0; j < k; ++j) std::swap(U[col * k + j], U[piv * k + j]);
      std::swap(V[col], V[piv]);
    }
    for (int r = col + 1; r < k; ++r) {
      const double f = U[r * k + col] / U[col * k + col];
      for (int j = col; j < k; ++j) U[r * k + j] -= f * U[col * k + j];
      V[r] -= f * V[col];
    }
  }
  c.assign(k, 0.0);
  for (int r = k - 1; r >= 0; --r) {
    double a = V[r];
    for (int j = r + 1; j < k; ++j) a -= U[r * k + j] * c[j];
    c[r] = a / U[r * k + r];
  }
  return true;
}

}  // namespace

struct cfft_scft {
  cfft_plan* plan = nullptr;
  int prec = 0, ns = 0, nA = 0, m = 0;
  long M3 = 0;
  size_t esz = 8;
  double f = 0.0, chiN = 0.0, ds = 0.0;
  double lam = 1.0, lam_simple = 0.1, simple_above = 1e-2;
  hipStream_t stream = nullptr;
  DevBuf hist, E, w, phi, d, Wr, Dr, wa, wb, part, dpart, tot;
  double* tot_h = nullptr;  // pinned: totals of the density and update/Gram partials
  // Anderson history: ring slots (newest last), Gram matrix G[slot][slot] = <d_i, d_j>/n
  std::vector<int> order;
  std::vector<double> G;
  int next_slot = 0;
  int iters = 0;
  double last[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  ~cfft_scft() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (tot_h) (void)hipHostFree(tot_h);
  }
};

namespace {

unsigned red_blocks(long n) { return (unsigned)std::min<long>(kMaxRed, std::max<long>(1, nblk(n, kRB))); }

// Sweep and unnormalised densities; the partial sums of q(ns), q+(ns) stay on
// the device (dpart, *dnb blocks).
template <typename T>
void sweep_and_densities(cfft_scft* c, int* dnb) {
  const long M3 = c->M3;
  const long n2 = 2 * M3;
  T* E = static_cast<T*>(c->E.p);
  hipLaunchKernelGGL(k_expw<T>, dim3(nblk(n2, 256)), dim3(256), 0, c->stream,
                     static_cast<const double*>(c->w.p), E, -0.5 * c->ds, n2);
  hip_check(hipGetLastError(), "k_expw");
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n2, 256)), dim3(256), 0, c->stream, static_cast<T*>(c->hist.p),
                     T(1), n2);
  hip_check(hipGetLastError(), "k_fill");
  // q (entry 0) uses E_A on its first nA steps, q+ (entry 1) E_B on its first ns - nA
  const void* E1[2] = {E, E + M3};
  const void* E2[2] = {E + M3, E};
  const int32_t sw[2] = {c->nA, c->ns - c->nA};
  const int rc = cfft_sweep_hist_dev(c->plan, c->hist.p, E1, E2, sw, 0, c->ns, c->stream);
  if (rc != CFFT_OK) throw Error{rc, g_last_error};
  const unsigned nb = red_blocks(M3);
  hipLaunchKernelGGL(k_density<T>, dim3(nb), dim3(kRB), 0, c->stream, static_cast<const T*>(c->hist.p), M3,
                     c->ns, c->nA, static_cast<const double*>(c->wa.p), static_cast<const double*>(c->wb.p),
                     static_cast<double*>(c->phi.p), static_cast<double*>(c->dpart.p));
  hip_check(hipGetLastError(), "k_density");
  *dnb = (int)nb;
}

// One SCFT iteration; out = {Q, Q+, F, residual, <phiA + phiB>, <phiA>, mixing, k}.
// One host synchronisation per iteration (the totals of the partial sums);
// the mixing kernel is left in flight on the context's stream, ahead of the
// next iteration's sweep.
void iterate(cfft_scft* c, double* out) {
  const long M3 = c->M3, n2 = 2 * M3;
  int dnb = 0;
  if (c->prec == CFFT_FP64) sweep_and_densities<double>(c, &dnb);
  else sweep_and_densities<float>(c, &dnb);
  // ---- field update + Anderson history slot + Gram row, one pass ----
  const int slot = c->next_slot;
  c->next_slot = (c->next_slot + 1) % (c->m + 1);
  c->order.erase(std::remove(c->order.begin(), c->order.end(), slot), c->order.end());
  c->order.push_back(slot);
  const int ns = (int)c->order.size();
  GramSlots gs{};
  gs.n = ns;
  gs.cur = slot;
  for (int i = 0; i < ns; ++i) gs.slot[i] = c->order[i];
  double* tot = static_cast<double*>(c->tot.p);
  hipLaunchKernelGGL(k_total<2>, dim3(2), dim3(kRB), 0, c->stream, static_cast<const double*>(c->dpart.p), dnb,
                     tot);
  hip_check(hipGetLastError(), "k_total");
  const unsigned nb = red_blocks(M3);
  hipLaunchKernelGGL(k_update_gram, dim3(nb), dim3(kRB), 0, c->stream, static_cast<const double*>(c->w.p),
                     static_cast<double*>(c->phi.p), static_cast<double*>(c->d.p), static_cast<const double*>(tot),
                     c->ds, c->chiN, M3, static_cast<double*>(c->Wr.p), static_cast<double*>(c->Dr.p), gs,
                     static_cast<double*>(c->part.p));
  hip_check(hipGetLastError(), "k_update_gram");
  hipLaunchKernelGGL(k_total<kNU>, dim3(kNU), dim3(kRB), 0, c->stream, static_cast<const double*>(c->part.p),
                     (int)nb, tot + 2);
  hip_check(hipGetLastError(), "k_total");
  hip_check(hipMemcpyAsync(c->tot_h, tot, sizeof(double) * (2 + kNU), hipMemcpyDeviceToHost, c->stream),
            "D2H totals");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  const double Q[2] = {c->tot_h[0] / M3, c->tot_h[1] / M3};
  if (!(Q[0] > 0.0 && std::isfinite(Q[0]))) {
    c->order.pop_back();  // the newest slot holds no valid entry
    throw Error{CFFT_ERR_INVALID, "partition function is not positive"};
  }
  double s[7];
  for (int k = 0; k < 7; ++k) s[k] = c->tot_h[2 + k] / M3;
  const double F = -std::log(Q[0]) + c->chiN * s[0] - s[1] - s[2];
  const double res = std::sqrt(s[4] / std::max(s[5], 1e-300)) + std::fabs(s[3] - 1.0);
  {
    const int S = c->m + 1;
    for (int i = 0; i < ns; ++i) {
      const double g = c->tot_h[2 + 7 + i] / n2;
      c->G[slot * S + c->order[i]] = g;
      c->G[c->order[i] * S + slot] = g;
    }
  }
  // ---- mixing ----
  MixP mp{};
  mp.cur = slot;
  mp.lam = c->lam_simple;
  mp.k = 0;
  double mode = 0.0;
  const int k = ns - 1;
  if (res <= c->simple_above && k > 0) {
    const int S = c->m + 1;
    const double* G = c->G.data();
    auto g = [&](int a, int b) { return G[a * S + b]; };
    std::vector<double> U(k * k), V(k), cc;
    for (int i = 0; i < k; ++i) {
      const int si = c->order[ns - 2 - i];
      V[i] = g(slot, slot) - g(slot, si);
      for (int j = 0; j < k; ++j) {
        const int sj = c->order[ns - 2 - j];
        U[i * k + j] = g(slot, slot) - g(slot, si) - g(slot, sj) + g(si, sj);
      }
    }
    if (solve_small(U, V, k, cc)) {
      mp.k = k;
      mp.lam = c->lam;
      for (int i = 0; i < k; ++i) {
        mp.slot[i] = c->order[ns - 2 - i];
        mp.c[i] = cc[i];
      }
      mode = 1.0;
    }
  }
  hipLaunchKernelGGL(k_mix, dim3(nblk(n2, 256)), dim3(256), 0, c->stream, static_cast<double*>(c->w.p),
                     static_cast<const double*>(c->Wr.p), static_cast<const double*>(c->Dr.p), mp, n2);
  hip_check(hipGetLastError(), "k_mix");
  ++c->iters;
  const double o[8] = {Q[0], Q[1], F, res, s[3], s[6], mode, (double)mp.k};
  std::memcpy(c->last, o, sizeof(o));
  if (out) std::memcpy(out, o, sizeof(o));
}

}  // namespace

extern "C" {

int cfft_scft_create(cfft_plan* plan, double f, double chiN, int32_t ns, int32_t m_hist, cfft_scft** out) {
  return guarded([&] {
    require(plan && out, CFFT_ERR_INVALID, "null argument");
    require(ns >= 2 && f > 0.0 && f < 1.0, CFFT_ERR_INVALID, "need ns >= 2 and 0 < f < 1");
    require(m_hist >= 0 && m_hist < kMaxHist, CFFT_ERR_INVALID, "Anderson history must be < 32");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    int32_t info[18];
    int rc = cfft_plan_info(plan, info);
    if (rc != CFFT_OK) throw Error{rc, g_last_error};
    require(info[12] == 2, CFFT_ERR_INVALID, "the SCFT plan must have batch = 2 (q and q+)");
    std::unique_ptr<cfft_scft> c(new cfft_scft());
    c->plan = plan;
    c->prec = info[11];
    c->M3 = (long)info[0] * info[1] * info[2];
    c->esz = c->prec == CFFT_FP64 ? 8 : 4;
    c->f = f;
    c->chiN = chiN;
    c->ns = ns;
    c->nA = (int)std::lround(f * ns);
    require(c->nA >= 1 && c->nA < ns, CFFT_ERR_INVALID, "f * ns must round into [1, ns - 1]");
    c->ds = 1.0 / ns;
    c->m = m_hist;
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    const long n2 = 2 * c->M3;
    c->hist.alloc(c->esz * n2 * (ns + 1));
    c->E.alloc(c->esz * n2);
    c->w.alloc(8 * n2);
    c->phi.alloc(8 * n2);
    c->d.alloc(8 * n2);
    c->Wr.alloc(8 * n2 * (m_hist + 1));
    c->Dr.alloc(8 * n2 * (m_hist + 1));
    c->part.alloc(8 * kMaxRed * kNU);
    c->dpart.alloc(8 * kMaxRed * 2);
    c->tot.alloc(8 * (2 + kNU));
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->tot_h), sizeof(double) * (2 + kNU), hipHostMallocDefault),
              "hipHostMalloc");
    c->G.assign((m_hist + 1) * (m_hist + 1), 0.0);
    const std::vector<double> wa = simpson(c->nA), wb = simpson(ns - c->nA);
    c->wa.alloc(8 * (size_t)(ns + 1));
    c->wb.alloc(8 * (size_t)(ns + 1));
    hip_check(hipMemcpy(c->wa.p, wa.data(), 8 * wa.size(), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(c->wb.p, wb.data(), 8 * wb.size(), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(c->w.p, 0, 8 * n2), "memset");
    *out = c.release();
  });
}

int cfft_scft_set_params(cfft_scft* c, double f, double chiN) {
  return guarded([&] {
    require(c != nullptr, CFFT_ERR_INVALID, "null context");
    require(f > 0.0 && f < 1.0, CFFT_ERR_INVALID, "need 0 < f < 1");
    const int nA = (int)std::lround(f * c->ns);
    require(nA >= 1 && nA < c->ns, CFFT_ERR_INVALID, "f * ns must round into [1, ns - 1]");
    c->f = f;
    c->chiN = chiN;
    c->nA = nA;
    // the weight buffers hold ns + 1 entries each (sized at creation)
    const std::vector<double> wa = simpson(c->nA), wb = simpson(c->ns - c->nA);
    hip_check(hipMemcpyAsync(c->wa.p, wa.data(), 8 * wa.size(), hipMemcpyHostToDevice, c->stream), "upload");
    hip_check(hipMemcpyAsync(c->wb.p, wb.data(), 8 * wb.size(), hipMemcpyHostToDevice, c->stream), "upload");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

int cfft_scft_set_mixing(cfft_scft* c, double lam, double lam_simple, double simple_above) {
  return guarded([&] {
    require(c != nullptr, CFFT_ERR_INVALID, "null context");
    c->lam = lam;
    c->lam_simple = lam_simple;
    c->simple_above = simple_above;
  });
}

int cfft_scft_set_fields(cfft_scft* c, const double* wA, const double* wB) {
  return guarded([&] {
    require(c && wA && wB, CFFT_ERR_INVALID, "null argument");
    double* w = static_cast<double*>(c->w.p);
    hip_check(hipMemcpyAsync(w, wA, 8 * c->M3, hipMemcpyHostToDevice, c->stream), "H2D");
    hip_check(hipMemcpyAsync(w + c->M3, wB, 8 * c->M3, hipMemcpyHostToDevice, c->stream), "H2D");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    c->order.clear();
    c->next_slot = 0;
    std::fill(c->G.begin(), c->G.end(), 0.0);
  });
}

int cfft_scft_iterate(cfft_scft* c, double* out) {
  return guarded([&] {
    require(c != nullptr, CFFT_ERR_INVALID, "null context");
    iterate(c, out);
  });
}

int cfft_scft_solve(cfft_scft* c, double tol, int32_t max_iter, int32_t* iters_out, double* out) {
  return guarded([&] {
    require(c != nullptr && max_iter >= 0, CFFT_ERR_INVALID, "bad argument");
    int it = 0;
    double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    while (it < max_iter) {
      iterate(c, o);
      ++it;
      if (o[3] < tol) break;
    }
    if (iters_out) *iters_out = it;
    if (out) std::memcpy(out, o, sizeof(o));
  });
}

int cfft_scft_get(cfft_scft* c, int32_t which, double* out) {
  return guarded([&] {
    require(c && out, CFFT_ERR_INVALID, "null argument");
    const double* src = nullptr;
    switch (which) {
      case CFFT_SCFT_W: src = static_cast<const double*>(c->w.p); break;
      case CFFT_SCFT_PHI: src = static_cast<const double*>(c->phi.p); break;
      case CFFT_SCFT_DEV: src = static_cast<const double*>(c->d.p); break;
      default: throw Error{CFFT_ERR_INVALID, "unknown quantity"};
    }
    hip_check(hipMemcpyAsync(out, src, 16 * c->M3, hipMemcpyDeviceToHost, c->stream), "D2H");
    hip_check(hipStreamSynchronize(c->stream), "sync");
  });
}

int cfft_scft_propagators(cfft_scft* c, void** hist_out) {
  return guarded([&] {
    require(c && hist_out, CFFT_ERR_INVALID, "null argument");
    hip_check(hipStreamSynchronize(c->stream), "sync");  // the last iteration's work is complete
    *hist_out = c->hist.p;
  });
}

int cfft_scft_destroy(cfft_scft* c) {
  return guarded([&] {
    if (!c) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
  });
}

}  // extern "C"
