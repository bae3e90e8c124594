// Multilayer perceptron for the SCFT warm start: training (forward, backward,
// Adam) and inference on the GPU.  Weights are kept as an FP32 master copy for
// the optimiser and a BF16 copy that feeds the matrix cores; every GEMM takes
// BF16 operands (activations are rounded to BF16 when loaded) and accumulates
// in FP32 with v_mfma_f32_16x16x16_bf16.  SiLU hidden activations, linear
// output, mean-squared-error loss.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "../../include/cfft_hip.h"
#include "common.hpp"

using namespace cfft;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Operand view: element (r, c) at p[r * rs + c * cs]; FP32 values are rounded
// to BF16 on load, BF16 values are used as stored.
struct OpF {
  const float* p;
  long rs, cs;
};
struct OpB {
  const uint16_t* p;
  long rs, cs;
};
__device__ inline short ld(const OpF& o, int r, int c) { return (short)f2bf(o.p[r * o.rs + c * o.cs]); }
__device__ inline short ld(const OpB& o, int r, int c) { return (short)o.p[r * o.rs + c * o.cs]; }

// Epilogues of C = A B (M x N):
//   0 plain store                     out0 = C
//   1 bias + SiLU                      out0 = Z = C + b, out1 = silu(Z)
//   2 bias, linear                     out0 = C + b
//   3 times SiLU'(Z) (backward)        out0 = C * silu'(aux)
enum { EP_STORE = 0, EP_BIAS_SILU = 1, EP_BIAS = 2, EP_DSILU = 3 };

struct Ep {
  int kind;
  float* out0;
  float* out1;
  const float* bias;  // N
  const float* aux;   // M x N, row-major (EP_DSILU)
  long ldo;           // row stride of out0 / out1 / aux
};

// One wavefront per 16 x 16 output tile; the K loop runs in steps of 16 with
// one MFMA each.  Lane l holds A(row l%16, k 4(l/16)+j), B(k 4(l/16)+j, col
// l%16) and D(row 4(l/16)+j, col l%16), j < 4.  Out-of-range rows, columns
// and k are fed as zeros.
template <typename TA, typename TB>
__global__ void __launch_bounds__(256) k_gemm(TA A, TB B, int M, int N, int K, Ep ep) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = (N + 15) / 16;
  const long tile = (long)blockIdx.x * 4 + wave;
  if (tile >= (long)((M + 15) / 16) * tiles_n) return;
  const int m0 = (int)(tile / tiles_n) * 16, n0 = (int)(tile % tiles_n) * 16;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + kq + j;
      a[j] = (m0 + r < M && k < K) ? ld(A, m0 + r, k) : (short)0;
      b[j] = (n0 + r < N && k < K) ? ld(B, k, n0 + r) : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int col = n0 + r;
  if (col >= N) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = m0 + kq + j;
    if (row >= M) continue;
    const long o = (long)row * ep.ldo + col;
    float v = acc[j];
    switch (ep.kind) {
      case EP_STORE: ep.out0[o] = v; break;
      case EP_BIAS_SILU:
        v += ep.bias[col];
        ep.out0[o] = v;
        ep.out1[o] = silu(v);
        break;
      case EP_BIAS: ep.out0[o] = v + ep.bias[col]; break;
      default: ep.out0[o] = v * dsilu(ep.aux[o]); break;
    }
  }
}

// Loss and output gradient: dY = 2 (Y - T) / (n_out * rows); partial sums of
// (Y - T)^2 per block.
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, long n, float scale,
                                             double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    acc += (double)d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Column sums of a rows x cols row-major matrix (bias gradients).
__global__ void k_colsum(const float* __restrict__ X, float* __restrict__ out, int rows, int cols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += X[(long)r * cols + c];
  out[c] = s;
}

// Adam on the FP32 master parameters; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, float* __restrict__ dst,
                              int rows, int cols) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  dst[i] = src[(long)idx[r] * cols + c];
}

}  // namespace

struct cfft_mlp {
  int L = 0;                 // weight layers
  std::vector<int> dims;     // L + 1
  std::vector<long> woff, boff;
  long n_params = 0;         // weights then biases, all layers
  int max_rows = 0;
  DevBuf w, wb, g, m, v;     // FP32 master, BF16 copy, gradient, Adam moments
  DevBuf Z, H;               // per layer pre/post activations (max_rows x dims[l+1])
  DevBuf dA, dB;             // backward ping-pong (max_rows x max dim)
  DevBuf X, T, idx, xb, tb, part;
  long n_data = 0;
  long step = 0;
  hipStream_t stream = nullptr;
  std::vector<long> zoff;
  ~cfft_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void ensure_rows(cfft_mlp* p, int rows) {
  if (rows <= p->max_rows) return;
  p->max_rows = rows;
  long tot = 0;
  p->zoff.assign(p->L + 1, 0);
  for (int l = 0; l < p->L; ++l) {
    p->zoff[l] = tot;
    tot += (long)rows * p->dims[l + 1];
  }
  p->zoff[p->L] = tot;
  p->Z.alloc(4 * tot);
  p->H.alloc(4 * tot);
  int dmax = 0;
  for (int d : p->dims) dmax = std::max(dmax, d);
  p->dA.alloc(4L * rows * dmax);
  p->dB.alloc(4L * rows * dmax);
  p->xb.alloc(4L * rows * p->dims[0]);
  p->tb.alloc(4L * rows * p->dims[p->L]);
  p->idx.alloc(4L * rows);
}

template <typename TA, typename TB>
void gemm(cfft_mlp* p, TA A, TB B, int M, int N, int K, Ep ep) {
  const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL((k_gemm<TA, TB>), dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, p->stream, A, B, M, N, K,
                     ep);
  hip_check(hipGetLastError(), "k_gemm");
}

float* Zl(cfft_mlp* p, int l) { return static_cast<float*>(p->Z.p) + p->zoff[l]; }
float* Hl(cfft_mlp* p, int l) { return static_cast<float*>(p->H.p) + p->zoff[l]; }
const uint16_t* Wb(cfft_mlp* p, int l) { return static_cast<const uint16_t*>(p->wb.p) + p->woff[l]; }
const float* Bias(cfft_mlp* p, int l) { return static_cast<const float*>(p->w.p) + p->boff[l]; }

// Forward on `rows` inputs at x (device, rows x dims[0]); the output is Z of
// the last layer.
void forward(cfft_mlp* p, const float* x, int rows) {
  const float* in = x;
  for (int l = 0; l < p->L; ++l) {
    const int K = p->dims[l], N = p->dims[l + 1];
    Ep ep{};
    ep.kind = l + 1 < p->L ? EP_BIAS_SILU : EP_BIAS;
    ep.out0 = Zl(p, l);
    ep.out1 = Hl(p, l);
    ep.bias = Bias(p, l);
    ep.ldo = N;
    // W of layer l is K x N row-major
    gemm(p, OpF{in, K, 1}, OpB{Wb(p, l), N, 1}, rows, N, K, ep);
    in = Hl(p, l);
  }
}

// Backward from dY (rows x dims[L]) at dA; fills the gradient buffer.
void backward(cfft_mlp* p, const float* x, int rows) {
  float* g = static_cast<float*>(p->g.p);
  float* cur = static_cast<float*>(p->dA.p);  // dZ of layer l
  float* nxt = static_cast<float*>(p->dB.p);
  for (int l = p->L - 1; l >= 0; --l) {
    const int K = p->dims[l], N = p->dims[l + 1];
    const float* in = l == 0 ? x : Hl(p, l - 1);
    // dW_l = in^T dZ  (K x N)
    Ep ew{};
    ew.kind = EP_STORE;
    ew.out0 = g + p->woff[l];
    ew.ldo = N;
    gemm(p, OpF{in, 1, K}, OpF{cur, N, 1}, K, N, rows, ew);
    hipLaunchKernelGGL(k_colsum, dim3(nblk(N, 256)), dim3(256), 0, p->stream, cur, g + p->boff[l], rows, N);
    hip_check(hipGetLastError(), "k_colsum");
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1})   (rows x K)
    Ep ed{};
    ed.kind = EP_DSILU;
    ed.out0 = nxt;
    ed.aux = Zl(p, l - 1);
    ed.ldo = K;
    gemm(p, OpF{cur, N, 1}, OpB{Wb(p, l), 1, N}, rows, K, N, ed);
    std::swap(cur, nxt);
  }
}

double loss_grad(cfft_mlp* p, const float* t, int rows) {
  const int out = p->dims[p->L];
  const long n = (long)rows * out;
  const unsigned nb = (unsigned)std::min<long>(256, std::max<long>(1, nblk(n, 256)));
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, p->stream, Zl(p, p->L - 1), t,
                     static_cast<float*>(p->dA.p), n, 1.0f / (float)n, static_cast<double*>(p->part.p));
  hip_check(hipGetLastError(), "k_mse");
  std::vector<double> h(nb);
  hip_check(hipMemcpyAsync(h.data(), p->part.p, 8 * nb, hipMemcpyDeviceToHost, p->stream), "D2H");
  hip_check(hipStreamSynchronize(p->stream), "sync");
  double s = 0.0;
  for (double v : h) s += v;
  return s / (double)n;
}

}  // namespace

extern "C" {

int cfft_mlp_create(int32_t n_layers, const int32_t* dims, uint64_t seed, cfft_mlp** out) {
  return guarded([&] {
    require(n_layers >= 1 && dims && out, CFFT_ERR_INVALID, "bad argument");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    std::unique_ptr<cfft_mlp> p(new cfft_mlp());
    p->L = n_layers;
    p->dims.assign(dims, dims + n_layers + 1);
    for (int d : p->dims) require(d >= 1, CFFT_ERR_INVALID, "layer widths must be positive");
    long off = 0;
    for (int l = 0; l < p->L; ++l) {
      p->woff.push_back(off);
      off += (long)p->dims[l] * p->dims[l + 1];
    }
    for (int l = 0; l < p->L; ++l) {
      p->boff.push_back(off);
      off += p->dims[l + 1];
    }
    p->n_params = off;
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    // He-style initialisation (normal, std sqrt(2 / fan_in)), zero biases
    std::vector<float> w(p->n_params, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < p->L; ++l) {
      std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / p->dims[l]));
      for (long i = 0; i < (long)p->dims[l] * p->dims[l + 1]; ++i) w[p->woff[l] + i] = nd(rng);
    }
    p->w.alloc(4 * p->n_params);
    p->wb.alloc(2 * p->n_params);
    p->g.alloc(4 * p->n_params);
    p->m.alloc(4 * p->n_params);
    p->v.alloc(4 * p->n_params);
    p->part.alloc(8 * 256);
    hip_check(hipMemcpy(p->w.p, w.data(), 4 * p->n_params, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(p->m.p, 0, 4 * p->n_params), "memset");
    hip_check(hipMemset(p->v.p, 0, 4 * p->n_params), "memset");
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(p->n_params, 256)), dim3(256), 0, p->stream,
                       static_cast<const float*>(p->w.p), static_cast<uint16_t*>(p->wb.p), p->n_params);
    hip_check(hipStreamSynchronize(p->stream), "sync");
    *out = p.release();
  });
}

int cfft_mlp_info(const cfft_mlp* p, int64_t* n_params) {
  return guarded([&] {
    require(p && n_params, CFFT_ERR_INVALID, "null argument");
    *n_params = p->n_params;
  });
}

int cfft_mlp_get_params(cfft_mlp* p, float* out) {
  return guarded([&] {
    require(p && out, CFFT_ERR_INVALID, "null argument");
    hip_check(hipMemcpyAsync(out, p->w.p, 4 * p->n_params, hipMemcpyDeviceToHost, p->stream), "D2H");
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_mlp_set_params(cfft_mlp* p, const float* in) {
  return guarded([&] {
    require(p && in, CFFT_ERR_INVALID, "null argument");
    hip_check(hipMemcpyAsync(p->w.p, in, 4 * p->n_params, hipMemcpyHostToDevice, p->stream), "H2D");
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(p->n_params, 256)), dim3(256), 0, p->stream,
                       static_cast<const float*>(p->w.p), static_cast<uint16_t*>(p->wb.p), p->n_params);
    hip_check(hipMemsetAsync(p->m.p, 0, 4 * p->n_params, p->stream), "memset");
    hip_check(hipMemsetAsync(p->v.p, 0, 4 * p->n_params, p->stream), "memset");
    p->step = 0;
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_mlp_forward(cfft_mlp* p, const float* x, int32_t rows, float* y) {
  return guarded([&] {
    require(p && x && y && rows >= 1, CFFT_ERR_INVALID, "bad argument");
    ensure_rows(p, rows);
    hip_check(hipMemcpyAsync(p->xb.p, x, 4L * rows * p->dims[0], hipMemcpyHostToDevice, p->stream), "H2D");
    forward(p, static_cast<const float*>(p->xb.p), rows);
    hip_check(hipMemcpyAsync(y, Zl(p, p->L - 1), 4L * rows * p->dims[p->L], hipMemcpyDeviceToHost, p->stream),
              "D2H");
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_mlp_gradient(cfft_mlp* p, const float* x, const float* t, int32_t rows, double* loss, float* grad) {
  return guarded([&] {
    require(p && x && t && rows >= 1, CFFT_ERR_INVALID, "bad argument");
    ensure_rows(p, rows);
    hip_check(hipMemcpyAsync(p->xb.p, x, 4L * rows * p->dims[0], hipMemcpyHostToDevice, p->stream), "H2D");
    hip_check(hipMemcpyAsync(p->tb.p, t, 4L * rows * p->dims[p->L], hipMemcpyHostToDevice, p->stream), "H2D");
    forward(p, static_cast<const float*>(p->xb.p), rows);
    const double l = loss_grad(p, static_cast<const float*>(p->tb.p), rows);
    backward(p, static_cast<const float*>(p->xb.p), rows);
    if (loss) *loss = l;
    if (grad) {
      hip_check(hipMemcpyAsync(grad, p->g.p, 4 * p->n_params, hipMemcpyDeviceToHost, p->stream), "D2H");
    }
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_mlp_train(cfft_mlp* p, const float* X, const float* T, int64_t n, int32_t epochs, int32_t batch,
                   float lr, uint64_t seed, double* loss_per_epoch) {
  return guarded([&] {
    require(p && X && T && n >= 1 && epochs >= 0 && batch >= 1, CFFT_ERR_INVALID, "bad argument");
    const int din = p->dims[0], dout = p->dims[p->L];
    ensure_rows(p, batch);
    p->X.alloc(4L * n * din);
    p->T.alloc(4L * n * dout);
    hip_check(hipMemcpyAsync(p->X.p, X, 4L * n * din, hipMemcpyHostToDevice, p->stream), "H2D");
    hip_check(hipMemcpyAsync(p->T.p, T, 4L * n * dout, hipMemcpyHostToDevice, p->stream), "H2D");
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    for (long i = 0; i < n; ++i) perm[i] = (int)i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    DevBuf perm_dev;
    perm_dev.alloc(4L * n);
    for (int ep = 0; ep < epochs; ++ep) {
      // the previous epoch's gathers have read the device copy before it is replaced
      hip_check(hipStreamSynchronize(p->stream), "sync");
      std::shuffle(perm.begin(), perm.end(), rng);
      hip_check(hipMemcpyAsync(perm_dev.p, perm.data(), 4L * n, hipMemcpyHostToDevice, p->stream), "H2D");
      double lsum = 0.0;
      long seen = 0;
      for (long s0 = 0; s0 < n; s0 += batch) {
        const int rows = (int)std::min<long>(batch, n - s0);
        const int* idx = static_cast<const int*>(perm_dev.p) + s0;
        hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long)rows * din, 256)), dim3(256), 0, p->stream,
                           static_cast<const float*>(p->X.p), idx, static_cast<float*>(p->xb.p), rows, din);
        hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long)rows * dout, 256)), dim3(256), 0, p->stream,
                           static_cast<const float*>(p->T.p), idx, static_cast<float*>(p->tb.p), rows, dout);
        hip_check(hipGetLastError(), "k_gather_rows");
        forward(p, static_cast<const float*>(p->xb.p), rows);
        lsum += loss_grad(p, static_cast<const float*>(p->tb.p), rows) * rows;
        seen += rows;
        backward(p, static_cast<const float*>(p->xb.p), rows);
        ++p->step;
        const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(p->n_params, 256)), dim3(256), 0, p->stream,
                           static_cast<float*>(p->w.p), static_cast<uint16_t*>(p->wb.p),
                           static_cast<const float*>(p->g.p), static_cast<float*>(p->m.p),
                           static_cast<float*>(p->v.p), p->n_params, lr, b1, b2, eps, c1, c2);
        hip_check(hipGetLastError(), "k_adam");
      }
      if (loss_per_epoch) loss_per_epoch[ep] = lsum / (double)seen;
    }
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_mlp_destroy(cfft_mlp* p) {
  return guarded([&] {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    delete p;
  });
}

}  // extern "C"
