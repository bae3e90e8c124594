// Dense real-ASU diffusion operator (the reference's MatrixDiffusionSolver,
// src/diffusion_solver.jl:20-71) on the device: the host plans
// A = Re(M Q M^+) over the real-ASU points, the device applies it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>

#include "../../include/cfft_hip.h"
#include "common.hpp"

using namespace cfft;

struct cfft_dense {
  long n = 0;
  DevBuf A;     // n x n FP64, row-major
  DevBuf u, w;  // the field and the interaction factors of one call
  hipStream_t stream = nullptr;
  ~cfft_dense() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {

constexpr int kRowsPerBlock = 4;  // one wavefront per row, four rows per workgroup

// y = A x: wavefront r of the grid takes row r; its lanes read the row in
// consecutive 8-byte entries (coalesced) and reduce their partial sums with
// cross-lane shuffles.
__global__ void __launch_bounds__(64 * kRowsPerBlock) k_dense_gemv(const double* __restrict__ A,
                                                                   const double* __restrict__ x,
                                                                   double* __restrict__ y, long n) {
  const long row = (long)blockIdx.x * kRowsPerBlock + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (row >= n) return;  // whole wavefronts leave together
  const double* a = A + row * n;
  double s = 0.0;
  for (long j = lane; j < n; j += 64) s += a[j] * x[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[row] = s;
}

// u *= exp(-w factor)
__global__ void k_dense_interaction(double* __restrict__ u, const double* __restrict__ w, double factor, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) u[i] *= exp(-w[i] * factor);
}

void gemv(cfft_dense* h, const double* x, double* y) {
  hipLaunchKernelGGL(k_dense_gemv, dim3(nblk(h->n, kRowsPerBlock)), dim3(64 * kRowsPerBlock), 0, h->stream,
                     static_cast<const double*>(h->A.p), x, y, h->n);
  hip_check(hipGetLastError(), "k_dense_gemv");
}

void interaction(cfft_dense* h, double* u, const double* w, double factor) {
  hipLaunchKernelGGL(k_dense_interaction, dim3(nblk(h->n, 256)), dim3(256), 0, h->stream, u, w, factor, h->n);
  hip_check(hipGetLastError(), "k_dense_interaction");
}

}  // namespace

extern "C" {

int cfft_dense_create(int64_t n, const double* A, cfft_dense** out) {
  return guarded([&] {
    require(A && out && n > 0, CFFT_ERR_INVALID, "bad argument");
    require(n <= (int64_t(1) << 20), CFFT_ERR_SIZE, "dense operator larger than 2^20 points");
    *out = nullptr;
    std::unique_ptr<cfft_dense> h(new cfft_dense());
    h->n = n;
    {
      std::lock_guard<std::recursive_mutex> guard(setup_mutex());
      hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    }
    h->A.alloc(sizeof(double) * (size_t)n * (size_t)n);
    h->u.alloc(sizeof(double) * 2 * (size_t)n);  // u and A u
    h->w.alloc(sizeof(double) * (size_t)n);
    {
      std::lock_guard<std::recursive_mutex> guard(setup_mutex());
      hip_check(hipMemcpy(h->A.p, A, h->A.bytes, hipMemcpyHostToDevice), "upload A");
    }
    *out = h.release();
  });
}

int cfft_dense_apply(cfft_dense* h, double* u) {
  return guarded([&] {
    require(h && u, CFFT_ERR_INVALID, "bad argument");
    double* d = static_cast<double*>(h->u.p);
    const size_t b = sizeof(double) * (size_t)h->n;
    hip_check(hipMemcpyAsync(d, u, b, hipMemcpyHostToDevice, h->stream), "H2D");
    gemv(h, d, d + h->n);
    hip_check(hipMemcpyAsync(u, d + h->n, b, hipMemcpyDeviceToHost, h->stream), "D2H");
    hip_check(hipStreamSynchronize(h->stream), "sync");
  });
}

int cfft_dense_step(cfft_dense* h, double* u, const double* w, double dt) {
  return guarded([&] {
    require(h && u && w, CFFT_ERR_INVALID, "bad argument");
    double* d = static_cast<double*>(h->u.p);
    double* dw = static_cast<double*>(h->w.p);
    const size_t b = sizeof(double) * (size_t)h->n;
    hip_check(hipMemcpyAsync(d, u, b, hipMemcpyHostToDevice, h->stream), "H2D");
    hip_check(hipMemcpyAsync(dw, w, b, hipMemcpyHostToDevice, h->stream), "H2D");
    interaction(h, d, dw, dt / 2.0);
    gemv(h, d, d + h->n);
    interaction(h, d + h->n, dw, dt / 2.0);
    hip_check(hipMemcpyAsync(u, d + h->n, b, hipMemcpyDeviceToHost, h->stream), "D2H");
    hip_check(hipStreamSynchronize(h->stream), "sync");
  });
}

int cfft_dense_destroy(cfft_dense* h) {
  return guarded([&] {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    delete h;
  });
}

}  // extern "C"
