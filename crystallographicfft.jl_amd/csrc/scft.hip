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
#include <limits>
#include <memory>
#include <vector>

#include "../../include/cfft_hip.h"
#include "common.hpp"
#include "ensemble.hpp"
#include "lstsq.hpp"

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
__global__ void k_expw(const double* __restrict__ w, double* __restrict__ E, double factor, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) E[i] = exp(factor * w[i]);
}

// Unnormalised densities of one instance from the stored propagators (slice j
// holds nf fields; the instance's q(j) is field qf, its q+(j) field qf + 1):
//   uA = sum_{j<=nA} wa_j q(j) q+(ns-j),  uB = sum_{j>=nA} wb_{j-nA} q(j) q+(ns-j);
// partial sums of q(ns) and q+(ns) for Q.
template <typename T>
__global__ void __launch_bounds__(kRB) k_density(const T* __restrict__ hist, long M3, int nf, int qf, int ns,
                                                 int nA, const double* __restrict__ wa,
                                                 const double* __restrict__ wb, double* __restrict__ u,
                                                 double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  const T* hq = hist + (long)qf * M3;
  const T* hp = hist + (long)(qf + 1) * M3;
  const long sl = (long)nf * M3;
  for (long x = blockIdx.x * (long)kRB + threadIdx.x; x < M3; x += (long)gridDim.x * kRB) {
    double a = 0.0, b = 0.0;
    for (int j = 0; j <= ns; ++j) {
      const double pr = (double)hq[j * sl + x] * (double)hp[(ns - j) * sl + x];
      if (j <= nA) a += wa[j] * pr;
      if (j >= nA) b += wb[j - nA] * pr;
    }
    u[x] = a;
    u[M3 + x] = b;
    acc[0] += (double)hq[ns * sl + x];
    acc[1] += (double)hp[ns * sl + x];
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
// NSL = gs.n (the ring slots in use), a compile-time count so that every
// slot's loads are issued together; sums past NSL stay zero.
template <int NSL>
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
    for (int k = 0; k < NSL; ++k) {
      const int sl = gs.slot[k];
      acc[7 + k] += sl == gs.cur ? dA * dA + dB * dB : dA * Dr[sl * n2 + x] + dB * Dr[sl * n2 + M3 + x];
    }
  }
  block_sum<kNU>(acc, part);
}

// k_update_gram<gs.n>, 1 <= gs.n <= kMaxHist.
template <int N = 1>
void launch_update_gram(dim3 grid, hipStream_t st, const double* w, double* phi, double* d, const double* qtot,
                        double ds, double chiN, long M3, double* Wr, double* Dr, const GramSlots& gs,
                        double* part) {
  if constexpr (N > kMaxHist) {
    throw Error{CFFT_ERR_INVALID, "Anderson history slots out of range"};
  } else if (gs.n == N) {
    hipLaunchKernelGGL(k_update_gram<N>, grid, dim3(kRB), 0, st, w, phi, d, qtot, ds, chiN, M3, Wr, Dr, gs, part);
  } else {
    launch_update_gram<N + 1>(grid, st, w, phi, d, qtot, ds, chiN, M3, Wr, Dr, gs, part);
  }
}

// Totals of NV partial sums over nb <= kMaxRed blocks: workgroup k sums
// quantity k, each thread a strided share of the blocks, then a fixed tree
// over the workgroup (the same association on every call).
template <int NV>
__global__ void __launch_bounds__(kRB) k_total(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[kRB];
  const int k = blockIdx.x, tid = threadIdx.x;
  double a = 0.0;
  for (int b = tid; b < nb; b += kRB) a += part[b * NV + k];
  red[tid] = a;
  __syncthreads();
  for (int h = kRB / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) out[k] = red[0];
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


}  // namespace

// One melt of a context: its parameters, fields and mixing history.
struct ScftInst {
  double f = 0.0, chiN = 0.0;
  int nA = 0;
  DevBuf w, phi, d, Wr, Dr, wa, wb;
  // Anderson history: ring slots (newest last), Gram matrix G[slot][slot] = <d_i, d_j>/n
  std::vector<int> order;
  std::vector<double> G;
  int next_slot = 0;
  int iters = 0;
  bool failed = false;  // stopped by a failure of its own (see iterate)
  double last[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct cfft_scft {
  cfft_plan* plan = nullptr;
  int prec = 0, ns = 0, m = 0;
  long M3 = 0;
  size_t esz = 8;
  double ds = 0.0;
  double lam = 1.0, lam_simple = 0.1, simple_above = 1e-2;
  hipStream_t stream = nullptr;
  // Densities projected onto the symmetric fields before the field update
  // (mixed precision on the per-element Q paths: there FP32 rounding of each
  // element is independent, the propagators leave the symmetric subspace by
  // rounding-sized amounts and the fields would carry that from iteration to
  // iteration; the other paths pass every step through the spectral ASU).
  bool project = false;
  DevBuf phi_raw;  // unprojected densities of one instance (2 M^3)
  // shared by the instances: propagator slices [ns + 1][2 n_inst fields],
  // interaction factors [n_inst][E_A, E_B], reduction partials (reused in
  // stream order), per-instance totals and their pinned host copy
  DevBuf hist, E, part, dpart, tot;
  double* tot_h = nullptr;
  std::vector<std::unique_ptr<ScftInst>> in;
  ~cfft_scft() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (tot_h) (void)hipHostFree(tot_h);
  }
};

namespace {

constexpr int kTot = 2 + kNU;  // totals per instance: q(ns), q+(ns), then the k_update_gram sums

unsigned red_blocks(long n) { return (unsigned)std::min<long>(kMaxRed, std::max<long>(1, nblk(n, kRB))); }

int block_steps(double f, int ns) { return (int)std::lround(f * ns); }

void set_weights(cfft_scft* c, ScftInst& s, bool sync) {
  const std::vector<double> wa = simpson(s.nA), wb = simpson(c->ns - s.nA);
  hip_check(hipMemcpyAsync(s.wa.p, wa.data(), 8 * wa.size(), hipMemcpyHostToDevice, c->stream), "upload");
  hip_check(hipMemcpyAsync(s.wb.p, wb.data(), 8 * wb.size(), hipMemcpyHostToDevice, c->stream), "upload");
  if (sync) hip_check(hipStreamSynchronize(c->stream), "sync");  // the host vectors go out of scope
}

// Before a sweep of every instance (q, q+ of instance i are fields 2i,
// 2i + 1): E = exp(-ds w / 2) of the active instances and slice 0 = 1, on
// stream st; `part` receives the sweep's per-field factors and switch steps.
template <typename T>
void sweep_prepare(cfft_scft* c, const std::vector<char>& active, hipStream_t st, EnsPart& part) {
  const long M3 = c->M3, n2 = 2 * M3;
  const int ni = (int)c->in.size();
  double* E = static_cast<double*>(c->E.p);  // interaction factors: FP64 in both precisions
  for (int i = 0; i < ni; ++i) {
    if (!active[i]) continue;  // an idle instance keeps its last factors
    hipLaunchKernelGGL(k_expw, dim3(nblk(n2, 256)), dim3(256), 0, st,
                       static_cast<const double*>(c->in[i]->w.p), E + i * n2, -0.5 * c->ds, n2);
    hip_check(hipGetLastError(), "k_expw");
  }
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(ni * n2, 256)), dim3(256), 0, st, static_cast<T*>(c->hist.p), T(1),
                     ni * n2);
  hip_check(hipGetLastError(), "k_fill");
  // q of instance i uses E_A on its first nA steps, q+ E_B on its first ns - nA
  part.plan = c->plan;
  part.hist = c->hist.p;
  part.E1.assign(2 * ni, nullptr);
  part.E2.assign(2 * ni, nullptr);
  part.sw.assign(2 * ni, 0);
  for (int i = 0; i < ni; ++i) {
    const double* Ea = E + i * n2;
    const double* Eb = Ea + M3;
    part.E1[2 * i] = Ea;
    part.E2[2 * i] = Eb;
    part.sw[2 * i] = c->in[i]->nA;
    part.E1[2 * i + 1] = Eb;
    part.E2[2 * i + 1] = Ea;
    part.sw[2 * i + 1] = c->ns - c->in[i]->nA;
  }
}

// Sweep of every instance on the context's own stream.
template <typename T>
void sweep_all(cfft_scft* c, const std::vector<char>& active) {
  EnsPart part;
  sweep_prepare<T>(c, active, c->stream, part);
  const int rc = cfft_sweep_hist_dev(c->plan, c->hist.p, part.E1.data(), part.E2.data(), part.sw.data(), 0, c->ns,
                                     c->stream);
  if (rc != CFFT_OK) throw Error{rc, g_last_error};
}

// Densities, field update and Gram row of instance i (device only; totals to
// tot + i kTot).
template <typename T>
void update_inst(cfft_scft* c, int i, hipStream_t st) {
  ScftInst& s = *c->in[i];
  const long M3 = c->M3;
  const unsigned nb = red_blocks(M3);
  double* tot = static_cast<double*>(c->tot.p) + (long)i * kTot;
  double* phi = static_cast<double*>(c->project ? c->phi_raw.p : s.phi.p);
  hipLaunchKernelGGL(k_density<T>, dim3(nb), dim3(kRB), 0, st, static_cast<const T*>(c->hist.p), M3,
                     2 * (int)c->in.size(), 2 * i, c->ns, s.nA, static_cast<const double*>(s.wa.p),
                     static_cast<const double*>(s.wb.p), phi, static_cast<double*>(c->dpart.p));
  hip_check(hipGetLastError(), "k_density");
  if (c->project) {  // the projection keeps every sum over the grid (each op permutes the points)
    const int rc = cfft_symmetrize_f64_dev(c->plan, phi, s.phi.p, 2, st);
    if (rc != CFFT_OK) throw Error{rc, g_last_error};
  }
  hipLaunchKernelGGL(k_total<2>, dim3(2), dim3(kRB), 0, st, static_cast<const double*>(c->dpart.p),
                     (int)nb, tot);
  hip_check(hipGetLastError(), "k_total");
  // the newest history slot and the Gram row over the ring
  const int slot = s.next_slot;
  s.next_slot = (s.next_slot + 1) % (c->m + 1);
  s.order.erase(std::remove(s.order.begin(), s.order.end(), slot), s.order.end());
  s.order.push_back(slot);
  GramSlots gs{};
  gs.n = (int)s.order.size();
  gs.cur = slot;
  for (int k = 0; k < gs.n; ++k) gs.slot[k] = s.order[k];
  launch_update_gram(dim3(nb), st, static_cast<const double*>(s.w.p), static_cast<double*>(s.phi.p),
                     static_cast<double*>(s.d.p), static_cast<const double*>(tot), c->ds, s.chiN, M3,
                     static_cast<double*>(s.Wr.p), static_cast<double*>(s.Dr.p), gs, static_cast<double*>(c->part.p));
  hip_check(hipGetLastError(), "k_update_gram");
  hipLaunchKernelGGL(k_total<kNU>, dim3(kNU), dim3(kRB), 0, st, static_cast<const double*>(c->part.p),
                     (int)nb, tot + 2);
  hip_check(hipGetLastError(), "k_total");
}

// Host part of instance i's iteration from its totals: Q, F, residual, Gram
// row; then the mixing kernel (left in flight on the context's stream).
// out = {Q, Q+, F, residual, <phiA + phiB>, <phiA>, mixing, k}.
void finish_inst(cfft_scft* c, int i, double* out, hipStream_t st) {
  ScftInst& s = *c->in[i];
  const long M3 = c->M3, n2 = 2 * M3;
  const double* t = c->tot_h + (long)i * kTot;
  const double Q[2] = {t[0] / M3, t[1] / M3};
  const int slot = s.order.back();
  if (!(Q[0] > 0.0 && std::isfinite(Q[0]))) {
    s.order.pop_back();  // the newest slot holds no valid entry
    throw Error{CFFT_ERR_INVALID, "partition function is not positive"};
  }
  double v[7];
  for (int k = 0; k < 7; ++k) v[k] = t[2 + k] / M3;
  const double F = -std::log(Q[0]) + s.chiN * v[0] - v[1] - v[2];
  const double res = std::sqrt(v[4] / std::max(v[5], 1e-300)) + std::fabs(v[3] - 1.0);
  const int S = c->m + 1;
  const int ns = (int)s.order.size();
  for (int k = 0; k < ns; ++k) {
    const double g = t[2 + 7 + k] / n2;
    s.G[slot * S + s.order[k]] = g;
    s.G[s.order[k] * S + slot] = g;
  }
  // ---- mixing ----
  MixP mp{};
  mp.cur = slot;
  mp.lam = c->lam_simple;
  mp.k = 0;
  double mode = 0.0;
  const int k = ns - 1;
  if (res <= c->simple_above && k > 0) {
    const double* G = s.G.data();
    auto g = [&](int a, int b) { return G[a * S + b]; };
    std::vector<double> U(k * k), V(k), cc;
    for (int a = 0; a < k; ++a) {
      const int sa = s.order[ns - 2 - a];
      V[a] = g(slot, slot) - g(slot, sa);
      for (int b = 0; b < k; ++b) {
        const int sb = s.order[ns - 2 - b];
        U[a * k + b] = g(slot, slot) - g(slot, sa) - g(slot, sb) + g(sa, sb);
      }
    }
    if (lstsq_sym(U, V, k, cc)) {
      mp.k = k;
      mp.lam = c->lam;
      for (int a = 0; a < k; ++a) {
        mp.slot[a] = s.order[ns - 2 - a];
        mp.c[a] = cc[a];
      }
      mode = 1.0;
    }
  }
  hipLaunchKernelGGL(k_mix, dim3(nblk(n2, 256)), dim3(256), 0, st, static_cast<double*>(s.w.p),
                     static_cast<const double*>(s.Wr.p), static_cast<const double*>(s.Dr.p), mp, n2);
  hip_check(hipGetLastError(), "k_mix");
  ++s.iters;
  const double o[8] = {Q[0], Q[1], F, res, v[3], v[6], mode, (double)mp.k};
  std::memcpy(s.last, o, sizeof(o));
  if (out) std::memcpy(out, o, sizeof(o));
}

// Host part of the iteration of the active instances of c once their totals
// are in c->tot_h.  An instance whose iteration fails (partition function
// not positive and finite) is marked failed, its out[] reads {NaN, NaN, NaN,
// +inf, NaN, NaN, 0, 0} and the others go on; n_act and n_failed count the
// instances finished and failed, `first` is the first failure.
void finish_active(cfft_scft* c, const std::vector<char>& active, double* out, hipStream_t st, int& n_act,
                   int& n_failed, Error& first) {
  const int ni = (int)c->in.size();
  for (int i = 0; i < ni; ++i) {
    if (!active[i]) continue;
    ++n_act;
    try {
      finish_inst(c, i, out ? out + 8 * i : nullptr, st);
    } catch (const Error& e) {
      if (n_failed == 0) first = e;
      ++n_failed;
      ScftInst& s = *c->in[i];
      s.failed = true;
      const double nan = std::numeric_limits<double>::quiet_NaN();
      const double o[8] = {nan, nan, nan, std::numeric_limits<double>::infinity(), nan, nan, 0.0, 0.0};
      std::memcpy(s.last, o, sizeof(o));
      if (out) std::memcpy(out + 8 * i, o, sizeof(o));
    }
  }
}

// One SCFT iteration of the active instances: one sweep of all of them, then
// per instance densities, update and Gram row on the device; one host
// synchronisation for all their totals; host mixing coefficients and the
// mixing kernels.  out[8 i ..] receives instance i's results (active only);
// the call fails only when every instance it iterated failed.
void iterate(cfft_scft* c, const std::vector<char>& active, double* out) {
  const int ni = (int)c->in.size();
  if (c->prec == CFFT_FP64) sweep_all<double>(c, active);
  else sweep_all<float>(c, active);
  for (int i = 0; i < ni; ++i) {
    if (!active[i]) continue;
    if (c->prec == CFFT_FP64) update_inst<double>(c, i, c->stream);
    else update_inst<float>(c, i, c->stream);
  }
  hip_check(hipMemcpyAsync(c->tot_h, c->tot.p, sizeof(double) * kTot * ni, hipMemcpyDeviceToHost, c->stream),
            "D2H totals");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  int n_act = 0, n_failed = 0;
  Error first{CFFT_OK, ""};
  finish_active(c, active, out, c->stream, n_act, n_failed, first);
  if (n_failed > 0) g_last_error = first.msg;
  if (n_act > 0 && n_failed == n_act) throw first;
}

ScftInst& instance(cfft_scft* c, int32_t i) {
  require(c != nullptr, CFFT_ERR_INVALID, "null context");
  require(i >= 0 && i < (int)c->in.size(), CFFT_ERR_INVALID, "instance index out of range");
  return *c->in[i];
}

void set_fields_inst(cfft_scft* c, ScftInst& s, const double* wA, const double* wB) {
  require(wA && wB, CFFT_ERR_INVALID, "null argument");
  double* w = static_cast<double*>(s.w.p);
  hip_check(hipMemcpyAsync(w, wA, 8 * c->M3, hipMemcpyHostToDevice, c->stream), "H2D");
  hip_check(hipMemcpyAsync(w + c->M3, wB, 8 * c->M3, hipMemcpyHostToDevice, c->stream), "H2D");
  hip_check(hipStreamSynchronize(c->stream), "sync");
  s.failed = false;
  s.order.clear();
  s.next_slot = 0;
  std::fill(s.G.begin(), s.G.end(), 0.0);
}

void get_inst(cfft_scft* c, ScftInst& s, int32_t which, double* out) {
  require(out != nullptr, CFFT_ERR_INVALID, "null argument");
  const double* src = nullptr;
  switch (which) {
    case CFFT_SCFT_W: src = static_cast<const double*>(s.w.p); break;
    case CFFT_SCFT_PHI: src = static_cast<const double*>(s.phi.p); break;
    case CFFT_SCFT_DEV: src = static_cast<const double*>(s.d.p); break;
    default: throw Error{CFFT_ERR_INVALID, "unknown quantity"};
  }
  hip_check(hipMemcpyAsync(out, src, 16 * c->M3, hipMemcpyDeviceToHost, c->stream), "D2H");
  hip_check(hipStreamSynchronize(c->stream), "sync");
}

void set_params_inst(cfft_scft* c, ScftInst& s, double f, double chiN) {
  require(f > 0.0 && f < 1.0, CFFT_ERR_INVALID, "need 0 < f < 1");
  const int nA = block_steps(f, c->ns);
  require(nA >= 1 && nA < c->ns, CFFT_ERR_INVALID, "f * ns must round into [1, ns - 1]");
  s.f = f;
  s.chiN = chiN;
  s.nA = nA;
  set_weights(c, s, true);  // the weight buffers hold ns + 1 entries each (sized at creation)
}

// Iterate until every instance has residual < tol or max_iter iterations; an
// instance stops (and its fields stay as its last mixing left them) as soon
// as it meets either.
void solve(cfft_scft* c, double tol, int32_t max_iter, int32_t* iters, double* out) {
  const int ni = (int)c->in.size();
  std::vector<char> active(ni, 0);
  for (int i = 0; i < ni; ++i) active[i] = max_iter > 0 && !c->in[i]->failed;
  std::vector<int> it(ni, 0);
  std::vector<double> o(8 * ni, 0.0);
  for (;;) {
    bool any = false;
    for (int i = 0; i < ni; ++i) any = any || active[i];
    if (!any) break;
    iterate(c, active, o.data());
    for (int i = 0; i < ni; ++i) {
      if (!active[i]) continue;
      ++it[i];
      if (c->in[i]->failed || o[8 * i + 3] < tol || it[i] >= max_iter) active[i] = 0;
    }
  }
  for (int i = 0; i < ni; ++i) {
    if (iters) iters[i] = it[i];
    if (out) std::memcpy(out + 8 * i, o.data() + 8 * i, 8 * sizeof(double));
  }
}

cfft_scft* create(cfft_plan* plan, int32_t n_inst, const double* f, const double* chiN, int32_t ns,
                  int32_t m_hist) {
  require(plan && f && chiN, CFFT_ERR_INVALID, "null argument");
  // the sweep takes per-field interaction factors for at most 64 fields
  require(n_inst >= 1 && n_inst <= 32, CFFT_ERR_INVALID, "need 1 to 32 instances per context");
  require(ns >= 2, CFFT_ERR_INVALID, "need ns >= 2");
  require(m_hist >= 0 && m_hist < kMaxHist, CFFT_ERR_INVALID, "Anderson history must be < 32");
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  int32_t info[18];
  int rc = cfft_plan_info(plan, info);
  if (rc != CFFT_OK) throw Error{rc, g_last_error};
  require(info[12] == 2 * n_inst, CFFT_ERR_INVALID,
          "the SCFT plan must have batch = 2 x instances (q and q+ of each)");
  std::unique_ptr<cfft_scft> c(new cfft_scft());
  c->plan = plan;
  c->prec = info[11];
  c->project = c->prec == CFFT_MIXED &&
               (info[10] == CFFT_PATH_QS || info[10] == CFFT_PATH_M2Q || info[10] == CFFT_PATH_M7Q);
  c->M3 = (long)info[0] * info[1] * info[2];
  c->esz = c->prec == CFFT_FP64 ? 8 : 4;
  c->ns = ns;
  c->ds = 1.0 / ns;
  c->m = m_hist;
  hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  const long n2 = 2 * c->M3;
  c->hist.alloc(c->esz * n2 * n_inst * (ns + 1));
  c->E.alloc(sizeof(double) * n2 * n_inst);
  c->part.alloc(8 * kMaxRed * kNU);
  c->dpart.alloc(8 * kMaxRed * 2);
  c->tot.alloc(8 * (size_t)kTot * n_inst);
  if (c->project) c->phi_raw.alloc(8 * n2);
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->tot_h), sizeof(double) * kTot * n_inst,
                          hipHostMallocDefault),
            "hipHostMalloc");
  for (int i = 0; i < n_inst; ++i) {
    require(f[i] > 0.0 && f[i] < 1.0, CFFT_ERR_INVALID, "need 0 < f < 1");
    const int nA = block_steps(f[i], ns);
    require(nA >= 1 && nA < ns, CFFT_ERR_INVALID, "f * ns must round into [1, ns - 1]");
    std::unique_ptr<ScftInst> s(new ScftInst());
    s->f = f[i];
    s->chiN = chiN[i];
    s->nA = nA;
    s->w.alloc(8 * n2);
    s->phi.alloc(8 * n2);
    s->d.alloc(8 * n2);
    s->Wr.alloc(8 * n2 * (m_hist + 1));
    s->Dr.alloc(8 * n2 * (m_hist + 1));
    s->wa.alloc(8 * (size_t)(ns + 1));
    s->wb.alloc(8 * (size_t)(ns + 1));
    s->G.assign((m_hist + 1) * (m_hist + 1), 0.0);
    set_weights(c.get(), *s, true);
    hip_check(hipMemsetAsync(s->w.p, 0, 8 * n2, c->stream), "memset");
    c->in.push_back(std::move(s));
  }
  hip_check(hipStreamSynchronize(c->stream), "sync");
  return c.release();
}

cfft_scft* single(cfft_scft* c) {
  require(c != nullptr, CFFT_ERR_INVALID, "null context");
  require(c->in.size() == 1, CFFT_ERR_INVALID, "a context of several instances takes the _inst/_multi calls");
  return c;
}

}  // namespace

// Several contexts advanced together: per iteration one multi-plan sweep of
// all of them (per class of transform length, one launch per chain stage and
// contour step), then each instance's densities, update and mixing as in its
// own context; each context's work goes to its class's stream, and one host
// synchronisation per iteration covers all of them.
struct cfft_scft_ensemble {
  std::vector<cfft_scft*> ctx;
  std::vector<EnsPart> parts;
  EnsSweep* sweep = nullptr;
  int prec = 0;
  ~cfft_scft_ensemble() {
    if (sweep) ens_sweep_destroy(sweep);  // waits for its streams
  }
};

namespace {

// One iteration of the active instances of every context (act[k][i]);
// out[8 j ..] receives the results of instance j in context order (active
// only).  Work stays in flight on the class streams after the return (the
// mixing kernels).
void ens_iterate(cfft_scft_ensemble* e, const std::vector<std::vector<char>>& act, double* out) {
  const int nk = (int)e->ctx.size();
  std::vector<char> any(nk, 0);
  for (int k = 0; k < nk; ++k)
    for (char a : act[k]) any[k] = any[k] || a;
  for (int k = 0; k < nk; ++k) {
    if (!any[k]) continue;
    hipStream_t st = ens_part_stream(e->sweep, k);
    if (e->prec == CFFT_FP64) sweep_prepare<double>(e->ctx[k], act[k], st, e->parts[k]);
    else sweep_prepare<float>(e->ctx[k], act[k], st, e->parts[k]);
  }
  std::vector<std::vector<char>> fields(nk);  // q, q+ of instance i are fields 2i, 2i + 1
  for (int k = 0; k < nk; ++k) {
    fields[k].assign(2 * act[k].size(), 0);
    for (size_t i = 0; i < act[k].size(); ++i) fields[k][2 * i] = fields[k][2 * i + 1] = act[k][i];
  }
  ens_sweep_run(e->sweep, e->parts, fields);
  for (int k = 0; k < nk; ++k) {
    if (!any[k]) continue;
    cfft_scft* c = e->ctx[k];
    hipStream_t st = ens_part_stream(e->sweep, k);
    for (int i = 0; i < (int)c->in.size(); ++i) {
      if (!act[k][i]) continue;
      if (e->prec == CFFT_FP64) update_inst<double>(c, i, st);
      else update_inst<float>(c, i, st);
    }
    hip_check(hipMemcpyAsync(c->tot_h, c->tot.p, sizeof(double) * kTot * c->in.size(), hipMemcpyDeviceToHost, st),
              "D2H totals");
  }
  ens_sweep_sync(e->sweep);
  int n_act = 0, n_failed = 0;
  Error first{CFFT_OK, ""};
  long off = 0;
  for (int k = 0; k < nk; ++k) {
    cfft_scft* c = e->ctx[k];
    if (any[k])
      finish_active(c, act[k], out ? out + 8 * off : nullptr, ens_part_stream(e->sweep, k), n_act, n_failed, first);
    off += (long)c->in.size();
  }
  if (n_failed > 0) g_last_error = first.msg;
  if (n_act > 0 && n_failed == n_act) throw first;
}

// Work of the contexts' own streams completes before the ensemble's stream
// touches their buffers.
void ens_enter(cfft_scft_ensemble* e) {
  for (cfft_scft* c : e->ctx) hip_check(hipStreamSynchronize(c->stream), "sync");
}

void ens_solve(cfft_scft_ensemble* e, double tol, int32_t max_iter, int32_t* iters, double* out) {
  const int nk = (int)e->ctx.size();
  std::vector<std::vector<char>> act(nk);
  std::vector<std::vector<int>> it(nk);
  long ntot = 0;
  for (int k = 0; k < nk; ++k) {
    const int ni = (int)e->ctx[k]->in.size();
    act[k].assign(ni, 0);
    it[k].assign(ni, 0);
    for (int i = 0; i < ni; ++i) act[k][i] = max_iter > 0 && !e->ctx[k]->in[i]->failed;
    ntot += ni;
  }
  std::vector<double> o(8 * ntot, 0.0);
  for (;;) {
    bool go = false;
    for (int k = 0; k < nk; ++k)
      for (char a : act[k]) go = go || a;
    if (!go) break;
    ens_iterate(e, act, o.data());
    long off = 0;
    for (int k = 0; k < nk; ++k) {
      cfft_scft* c = e->ctx[k];
      for (int i = 0; i < (int)c->in.size(); ++i) {
        if (act[k][i]) {
          ++it[k][i];
          if (c->in[i]->failed || o[8 * (off + i) + 3] < tol || it[k][i] >= max_iter) act[k][i] = 0;
        }
      }
      off += (long)c->in.size();
    }
  }
  long off = 0;
  for (int k = 0; k < nk; ++k) {
    for (int i = 0; i < (int)e->ctx[k]->in.size(); ++i) {
      if (iters) iters[off + i] = it[k][i];
      if (out) std::memcpy(out + 8 * (off + i), o.data() + 8 * (off + i), 8 * sizeof(double));
    }
    off += (long)e->ctx[k]->in.size();
  }
}

cfft_scft_ensemble* ens_create(cfft_scft* const* ctx, int32_t n) {
  require(ctx != nullptr && n >= 1 && n <= kEnsMax, CFFT_ERR_INVALID, "an ensemble holds 1 to 64 contexts");
  std::unique_ptr<cfft_scft_ensemble> e(new cfft_scft_ensemble());
  for (int k = 0; k < n; ++k) {
    cfft_scft* c = ctx[k];
    require(c != nullptr, CFFT_ERR_INVALID, "null context");
    for (int j = 0; j < k; ++j) require(ctx[j] != c, CFFT_ERR_INVALID, "a context appears twice");
    if (k == 0) e->prec = c->prec;
    require(c->prec == e->prec && c->ns == ctx[0]->ns, CFFT_ERR_INVALID,
            "the contexts of an ensemble share precision and contour length");
    e->ctx.push_back(c);
  }
  // the sweep's factor pointers and switch steps, as sweep_prepare sets them
  e->parts.resize(n);
  for (int k = 0; k < n; ++k) {
    cfft_scft* c = ctx[k];
    const int ni = (int)c->in.size();
    EnsPart& p = e->parts[k];
    p.plan = c->plan;
    p.hist = c->hist.p;
    p.E1.assign(2 * ni, c->E.p);
    p.E2.assign(2 * ni, c->E.p);
    p.sw.assign(2 * ni, 0);
  }
  e->sweep = ens_sweep_create(e->parts, ctx[0]->ns);
  return e.release();
}

}  // namespace

extern "C" {

int cfft_scft_create(cfft_plan* plan, double f, double chiN, int32_t ns, int32_t m_hist, cfft_scft** out) {
  return guarded([&] {
    require(out != nullptr, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    *out = create(plan, 1, &f, &chiN, ns, m_hist);
  });
}

int cfft_scft_create_multi(cfft_plan* plan, int32_t n_inst, const double* f, const double* chiN, int32_t ns,
                           int32_t m_hist, cfft_scft** out) {
  return guarded([&] {
    require(out != nullptr, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    *out = create(plan, n_inst, f, chiN, ns, m_hist);
  });
}

int cfft_scft_set_params(cfft_scft* c, double f, double chiN) {
  return guarded([&] { set_params_inst(single(c), *c->in[0], f, chiN); });
}

int cfft_scft_set_params_inst(cfft_scft* c, int32_t inst, double f, double chiN) {
  return guarded([&] { set_params_inst(c, instance(c, inst), f, chiN); });
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
  return guarded([&] { set_fields_inst(single(c), *c->in[0], wA, wB); });
}

int cfft_scft_set_fields_inst(cfft_scft* c, int32_t inst, const double* wA, const double* wB) {
  return guarded([&] { set_fields_inst(c, instance(c, inst), wA, wB); });
}

int cfft_scft_iterate(cfft_scft* c, double* out) {
  return guarded([&] { iterate(single(c), std::vector<char>(1, 1), out); });
}

int cfft_scft_iterate_multi(cfft_scft* c, double* out) {
  return guarded([&] {
    require(c != nullptr, CFFT_ERR_INVALID, "null context");
    // instances that failed earlier stay stopped until their fields are set again
    std::vector<char> active(c->in.size(), 0);
    bool any = false;
    for (size_t i = 0; i < c->in.size(); ++i) any = (active[i] = !c->in[i]->failed) || any;
    require(any, CFFT_ERR_INVALID, "every instance of the context has failed; set new fields");
    iterate(c, active, out);
  });
}

int cfft_scft_solve(cfft_scft* c, double tol, int32_t max_iter, int32_t* iters_out, double* out) {
  return guarded([&] {
    require(max_iter >= 0, CFFT_ERR_INVALID, "bad argument");
    solve(single(c), tol, max_iter, iters_out, out);
  });
}

int cfft_scft_solve_multi(cfft_scft* c, double tol, int32_t max_iter, int32_t* iters_out, double* out) {
  return guarded([&] {
    require(c != nullptr && max_iter >= 0, CFFT_ERR_INVALID, "bad argument");
    solve(c, tol, max_iter, iters_out, out);
  });
}

int cfft_scft_get(cfft_scft* c, int32_t which, double* out) {
  return guarded([&] { get_inst(single(c), *c->in[0], which, out); });
}

int cfft_scft_get_inst(cfft_scft* c, int32_t inst, int32_t which, double* out) {
  return guarded([&] { get_inst(c, instance(c, inst), which, out); });
}

int cfft_scft_copy_state(cfft_scft* dst, cfft_scft* src) {
  return guarded([&] {
    require(dst && src && dst != src, CFFT_ERR_INVALID, "two distinct contexts");
    require(dst->M3 == src->M3 && dst->in.size() == src->in.size() && dst->m == src->m, CFFT_ERR_INVALID,
            "the contexts must share the subgrid, the instance count and the history length");
    // the fields and the Anderson history (FP64 in both precisions): the
    // destination's next iteration continues the source's mixing
    hip_check(hipStreamSynchronize(src->stream), "sync");
    hip_check(hipStreamSynchronize(dst->stream), "sync");
    const size_t n2 = 2 * (size_t)dst->M3;
    for (size_t i = 0; i < dst->in.size(); ++i) {
      ScftInst& d = *dst->in[i];
      const ScftInst& s = *src->in[i];
      hip_check(hipMemcpyAsync(d.w.p, s.w.p, 8 * n2, hipMemcpyDeviceToDevice, dst->stream), "copy w");
      hip_check(hipMemcpyAsync(d.Wr.p, s.Wr.p, d.Wr.bytes, hipMemcpyDeviceToDevice, dst->stream), "copy W");
      hip_check(hipMemcpyAsync(d.Dr.p, s.Dr.p, d.Dr.bytes, hipMemcpyDeviceToDevice, dst->stream), "copy D");
      d.order = s.order;
      d.G = s.G;
      d.next_slot = s.next_slot;
      d.failed = s.failed;
    }
    hip_check(hipStreamSynchronize(dst->stream), "sync");
  });
}

int cfft_scft_propagators(cfft_scft* c, void** hist_out) {
  return guarded([&] {
    require(c && hist_out, CFFT_ERR_INVALID, "null argument");
    hip_check(hipStreamSynchronize(c->stream), "sync");  // the last iteration's work is complete
    *hist_out = c->hist.p;
  });
}

int cfft_scft_ensemble_create(cfft_scft* const* ctx, int32_t n_ctx, cfft_scft_ensemble** out) {
  return guarded([&] {
    require(out != nullptr, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    *out = ens_create(ctx, n_ctx);
  });
}

int cfft_scft_ensemble_iterate(cfft_scft_ensemble* e, double* out) {
  return guarded([&] {
    require(e != nullptr, CFFT_ERR_INVALID, "null ensemble");
    ens_enter(e);
    std::vector<std::vector<char>> act(e->ctx.size());
    bool any = false;
    for (size_t k = 0; k < e->ctx.size(); ++k) {
      const cfft_scft* c = e->ctx[k];
      act[k].assign(c->in.size(), 0);
      for (size_t i = 0; i < c->in.size(); ++i) any = (act[k][i] = !c->in[i]->failed) || any;
    }
    require(any, CFFT_ERR_INVALID, "every instance of the ensemble has failed; set new fields");
    ens_iterate(e, act, out);
    ens_sweep_sync(e->sweep);
  });
}

int cfft_scft_ensemble_solve(cfft_scft_ensemble* e, double tol, int32_t max_iter, int32_t* iters_out, double* out) {
  return guarded([&] {
    require(e != nullptr && max_iter >= 0, CFFT_ERR_INVALID, "bad argument");
    ens_enter(e);
    ens_solve(e, tol, max_iter, iters_out, out);
    ens_sweep_sync(e->sweep);
  });
}

int cfft_scft_ensemble_destroy(cfft_scft_ensemble* e) {
  return guarded([&] {
    if (!e) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    delete e;
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
