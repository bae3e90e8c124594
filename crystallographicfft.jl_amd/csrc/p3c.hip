// Sparse-stage KRFFT families on the device.  P3c (the reference's
// recursive, sparse, selective-G0 and G0-ASU forward plans,
// src/recursive_blocks.jl:64-1240): for a cubic group whose shifted ops hold
// the body-diagonal 3-fold C (x, y, z) -> (z, x, y) with an even
// translation, the stride-2 subgrid's spectrum G0 follows from the FFTs of
// the stride-4 sub-sub-grids of one offset per orbit of the 8 offsets under
// C, and the spectral-ASU values from G0 through the 8 representative ops.
// Fractal (src/fractal_krfft.jl:670-901): leaves are the stride-2^d
// sub-grids of one offset per orbit under the whole group, combined level by
// level through the parity split.  The host builds each family's tables as
// CSR matrices (leaf FFT outputs -> ... -> spectral ASU); the device packs
// the leaves, runs their FFTs and the sparse products.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/cfft_hip.h"
#include "common.hpp"
#include "kernels.hpp"

using namespace cfft;

constexpr int kMaxStages = 8;

struct cfft_p3c {
  int N = 0, S = 0, Ml = 0, n_rep = 0, n_stage = 0;  // grid, leaf stride, leaf length N / S
  long n_out[kMaxStages] = {};  // rows of each stage; the last stage's rows are the spectral points
  long n_in0 = 0;               // n_rep Ml^3
  DevBuf offs;                  // n_rep x 3 leaf offsets (u at S k + e)
  DevBuf rowptr[kMaxStages], cols[kMaxStages], vals[kMaxStages];
  DevBuf u, g, mid[2], F;       // full grid, leaf FFT outputs, intermediate levels, spectral values
  hipStream_t stream = nullptr;
  ~cfft_p3c() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

namespace {

// y = A x (complex CSR): eight lanes per row, partial sums reduced across the lane octet
__global__ void k_p3c_spmv(const long* __restrict__ rowptr, const int* __restrict__ cols,
                           const cx<double>* __restrict__ vals, const cx<double>* __restrict__ x,
                           cx<double>* __restrict__ y, long n_rows) {
  const long row = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 3;
  const int lane = threadIdx.x & 7;
  const bool ok = row < n_rows;
  const long b = ok ? rowptr[row] : 0, e = ok ? rowptr[row + 1] : 0;
  cx<double> s = {0.0, 0.0};
  for (long j = b + lane; j < e; j += 8) s = cadd(s, cmul(vals[j], x[cols[j]]));
  s = octet_sum(s);
  if (ok && lane == 0) y[row] = s;
}

// Batched multi-lane FFT of WG-thread workgroups (k_fft_ml's transform
// addressing): the leaf batches are small (4 leaves of 32^3 at 128^3 are
// 4096 transforms per axis), so 64-thread workgroups spread one pass over
// four times as many CUs as k_fft_ml's 256 threads.
template <typename T, typename TW, int LEN, int DIR, int WG>
__global__ void __launch_bounds__(WG) k_leaf_fft(cx<T>* __restrict__ data, long n_tr, long inner, long outer,
                                                 long stride) {
  constexpr int R1 = MLS<LEN>::R1, R2 = MLS<LEN>::R2, G = WG / R1, GP = G + 1;
  __shared__ cx<T> xch[R1 > 1 ? R2 * R1 * GP : 1];
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

// The leaf pack fused with the x pass: row i = (leaf, ky, kz) of the leaves
// (x fastest) read from the full grid, g_r(k) = u(S k + e_r), transformed and
// stored as complex rows of g.
template <int LEN, int WG>
__global__ void __launch_bounds__(WG) k_leaf_pack_fftx(const double* __restrict__ u, cx<double>* __restrict__ g,
                                                       long n_tr, int N, int S, const int* __restrict__ offs) {
  constexpr int R1 = MLS<LEN>::R1, R2 = MLS<LEN>::R2, G = WG / R1, GP = G + 1;
  __shared__ cx<double> xch[R1 > 1 ? R2 * R1 * GP : 1];
  const int r = threadIdx.x % R1, gl = threadIdx.x / R1;
  const long gid = (long)blockIdx.x * G + gl;
  const bool valid = gid < n_tr;
  const long row = valid ? gid : 0;
  const int ky = (int)(row % LEN), kz = (int)((row / LEN) % LEN), leaf = (int)(row / (LEN * LEN));
  const long y = (long)S * ky + offs[3 * leaf + 1], z = (long)S * kz + offs[3 * leaf + 2];
  const double* src = u + (long)N * (y + (long)N * z) + offs[3 * leaf];
  cx<double> a[R2];
#pragma unroll
  for (int m = 0; m < R2; ++m) a[m] = {src[(long)S * (r + R1 * m)], 0.0};
  fft_ml<double, double, LEN, -1, GP>(a, r, gl, xch);
  if (valid) {
    cx<double>* dst = g + row * LEN;
#pragma unroll
    for (int q = 0; q < R2; ++q) dst[ml_out_pos<LEN>(r, q)] = a[q];
  }
}

// The leaf pack fused with the x and y transforms of one z-plane (leaf, kz)
// of a leaf: group g of R1 lanes takes x-row g straight from the full grid,
// leaves it transformed in LDS (row stride LEN + 1), then column g, whose
// y transform is stored to the plane in g.  LEN * R1 threads.
template <int LEN>
__global__ void __launch_bounds__(LEN * MLS<LEN>::R1) k_leaf_pack_fftxy(const double* __restrict__ u,
                                                                      cx<double>* __restrict__ g, int N, int S,
                                                                      const int* __restrict__ offs) {
  constexpr int R1 = MLS<LEN>::R1, R2 = MLS<LEN>::R2, rs = LEN + 1;
  __shared__ cx<double> lds[LEN * rs];
  const int r = threadIdx.x % R1, gi = threadIdx.x / R1;  // gi < LEN
  const int kz = blockIdx.x % LEN, leaf = blockIdx.x / LEN;
  const long y = (long)S * gi + offs[3 * leaf + 1], z = (long)S * kz + offs[3 * leaf + 2];
  const double* src = u + (long)N * (y + (long)N * z) + offs[3 * leaf];
  cx<double> v[R2];
#pragma unroll
  for (int m = 0; m < R2; ++m) v[m] = {src[(long)S * (r + R1 * m)], 0.0};
  cx<double>* row = lds + gi * rs;
  fft_ml_lds<double, double, LEN, -1>(v, r, row, 1);
#pragma unroll
  for (int q = 0; q < R2; ++q) row[ml_out_pos<LEN>(r, q)] = v[q];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R2; ++m) v[m] = lds[(r + R1 * m) * rs + gi];
  fft_ml_lds<double, double, LEN, -1>(v, r, lds + gi, rs);
  cx<double>* dst = g + (long)blockIdx.x * LEN * LEN + gi;
#pragma unroll
  for (int q = 0; q < R2; ++q) dst[(long)ml_out_pos<LEN>(r, q) * LEN] = v[q];
}

template <int LEN>
void fft3(const double* u, cx<double>* g, const cfft_p3c* h, hipStream_t s) {
  constexpr int WG = 64, G = WG / MLS<LEN>::R1;
  const long M2 = LEN, n = M2 * M2 * M2;
  const long n_tr = h->n_rep * M2 * M2;
  const dim3 grid(nblk(n_tr, G)), block(WG);
  const int* offs = static_cast<const int*>(h->offs.p);
  if constexpr (LEN >= 16) {
    // x and y in one pass over the leaves' z-planes, then z (stride M2^2)
    hipLaunchKernelGGL((k_leaf_pack_fftxy<LEN>), dim3(h->n_rep * LEN), dim3(LEN * MLS<LEN>::R1), 0, s, u, g, h->N,
                       h->S, offs);
  } else {
    // x (from the full grid), y (stride M2), z (stride M2^2); transform i
    // starts at (i % inner) + (i / inner) outer
    hipLaunchKernelGGL((k_leaf_pack_fftx<LEN, WG>), grid, block, 0, s, u, g, n_tr, h->N, h->S, offs);
    hipLaunchKernelGGL((k_leaf_fft<double, double, LEN, -1, WG>), grid, block, 0, s, g, n_tr, M2, M2 * M2, M2);
  }
  hipLaunchKernelGGL((k_leaf_fft<double, double, LEN, -1, WG>), grid, block, 0, s, g, n_tr, M2 * M2, n, M2 * M2);
}

void run(cfft_p3c* h, const double* u, double* F, hipStream_t s) {
  cx<double>* g = static_cast<cx<double>*>(h->g.p);
  switch (h->Ml) {
    case 2: fft3<2>(u, g, h, s); break;
    case 3: fft3<3>(u, g, h, s); break;
    case 4: fft3<4>(u, g, h, s); break;
    case 6: fft3<6>(u, g, h, s); break;
    case 8: fft3<8>(u, g, h, s); break;
    case 12: fft3<12>(u, g, h, s); break;
    case 16: fft3<16>(u, g, h, s); break;
    case 24: fft3<24>(u, g, h, s); break;
    case 32: fft3<32>(u, g, h, s); break;
    case 48: fft3<48>(u, g, h, s); break;
    case 64: fft3<64>(u, g, h, s); break;
    default: throw Error{CFFT_ERR_UNSUPPORTED, "leaf FFT length not supported"};
  }
  hip_check(hipGetLastError(), "k_leaf_fft");
  const cx<double>* x = g;
  for (int st = 0; st < h->n_stage; ++st) {
    cx<double>* y = st + 1 == h->n_stage ? reinterpret_cast<cx<double>*>(F)
                                         : static_cast<cx<double>*>(h->mid[st & 1].p);
    hipLaunchKernelGGL(k_p3c_spmv, dim3(nblk(8 * h->n_out[st], 256)), dim3(256), 0, s,
                       static_cast<const long*>(h->rowptr[st].p), static_cast<const int*>(h->cols[st].p),
                       static_cast<const cx<double>*>(h->vals[st].p), x, y, h->n_out[st]);
    hip_check(hipGetLastError(), "k_p3c_spmv");
    x = y;
  }
}

// validated build shared by the P3c and the strided (fractal) creators
cfft_p3c* create(int N, int S, int n_rep, const int32_t* offsets, int n_stage, const int64_t* n_rows,
                 const int64_t* const* rowptr, const int32_t* const* cols, const double* const* vals) {
  require(N > 0 && S >= 1 && N % S == 0, CFFT_ERR_NOT_DIVISIBLE, "grid size not divisible by the leaf stride");
  const int Ml = N / S;
  require(Ml >= 2 && Ml <= 64 && fft_supported(Ml), CFFT_ERR_UNSUPPORTED, "leaf FFT length not supported");
  require(n_rep >= 1 && (long)n_rep <= (long)S * S * S, CFFT_ERR_INVALID, "1 to S^3 leaves");
  require(n_stage >= 1 && n_stage <= kMaxStages, CFFT_ERR_INVALID, "one to eight sparse stages");
  for (int i = 0; i < 3 * n_rep; ++i)
    require(offsets[i] >= 0 && offsets[i] < S, CFFT_ERR_INVALID, "leaf offsets lie in [0, S)");
  std::unique_ptr<cfft_p3c> h(new cfft_p3c());
  h->N = N;
  h->S = S;
  h->Ml = Ml;
  h->n_rep = n_rep;
  h->n_stage = n_stage;
  h->n_in0 = (long)n_rep * Ml * Ml * Ml;
  long n_in = h->n_in0, n_mid = 1;
  for (int st = 0; st < n_stage; ++st) {
    const long nr = n_rows[st];
    require(nr >= 1 && rowptr[st] && cols[st] && vals[st], CFFT_ERR_INVALID, "bad stage");
    require(rowptr[st][0] == 0, CFFT_ERR_INVALID, "row pointers start at 0");
    for (long i = 0; i < nr; ++i)
      require(rowptr[st][i + 1] >= rowptr[st][i], CFFT_ERR_INVALID, "row pointers must not decrease");
    const long nnz = rowptr[st][nr];
    for (long j = 0; j < nnz; ++j)
      require(cols[st][j] >= 0 && cols[st][j] < n_in, CFFT_ERR_BOUNDS, "column index out of range");
    h->n_out[st] = nr;
    if (st + 1 < n_stage) n_mid = std::max(n_mid, nr);
    h->rowptr[st].alloc(sizeof(int64_t) * (nr + 1));
    h->cols[st].alloc(sizeof(int32_t) * std::max<long>(nnz, 1));
    h->vals[st].alloc(16 * std::max<long>(nnz, 1));
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    hip_check(hipMemcpy(h->rowptr[st].p, rowptr[st], sizeof(int64_t) * (nr + 1), hipMemcpyHostToDevice), "upload");
    if (nnz > 0) {
      hip_check(hipMemcpy(h->cols[st].p, cols[st], sizeof(int32_t) * nnz, hipMemcpyHostToDevice), "upload");
      hip_check(hipMemcpy(h->vals[st].p, vals[st], 16 * nnz, hipMemcpyHostToDevice), "upload");
    }
    n_in = nr;
  }
  h->offs.alloc(sizeof(int32_t) * 3 * n_rep);
  h->u.alloc(sizeof(double) * (size_t)N * N * N);
  h->g.alloc(16 * (size_t)h->n_in0);
  if (n_stage >= 2) h->mid[0].alloc(16 * (size_t)n_mid);
  if (n_stage >= 3) h->mid[1].alloc(16 * (size_t)n_mid);
  h->F.alloc(16 * (size_t)h->n_out[n_stage - 1]);
  {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    hip_check(hipMemcpy(h->offs.p, offsets, sizeof(int32_t) * 3 * n_rep, hipMemcpyHostToDevice), "upload");
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  return h.release();
}

}  // namespace

extern "C" {

int cfft_p3c_create(int32_t N, int32_t n_rep, const int32_t* offsets, int32_t n_stage, const int64_t* n_rows,
                    const int64_t* const* rowptr, const int32_t* const* cols, const double* const* vals,
                    cfft_p3c** out) {
  return guarded([&] {
    require(out && offsets && n_rows && rowptr && cols && vals, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    require(N > 0 && N % 4 == 0, CFFT_ERR_NOT_DIVISIBLE, "M must be divisible by 2 for P3c");
    require(n_rep >= 1 && n_rep <= 8, CFFT_ERR_INVALID, "1 to 8 sub-sub-grid representatives");
    require(n_stage == 1 || n_stage == 2, CFFT_ERR_INVALID, "one or two sparse stages");
    std::vector<int32_t> raw(3 * n_rep);
    for (int i = 0; i < 3 * n_rep; ++i) {
      require(offsets[i] == 0 || offsets[i] == 1, CFFT_ERR_INVALID, "offsets are 0 or 1");
      raw[i] = 2 * offsets[i];
    }
    *out = create(N, 4, n_rep, raw.data(), n_stage, n_rows, rowptr, cols, vals);
  });
}

int cfft_p3c_create_strided(int32_t N, int32_t stride, int32_t n_leaf, const int32_t* offsets, int32_t n_stage,
                            const int64_t* n_rows, const int64_t* const* rowptr, const int32_t* const* cols,
                            const double* const* vals, cfft_p3c** out) {
  return guarded([&] {
    require(out && offsets && n_rows && rowptr && cols && vals, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    *out = create(N, stride, n_leaf, offsets, n_stage, n_rows, rowptr, cols, vals);
  });
}

int cfft_p3c_forward(cfft_p3c* h, const double* u, double* F) {
  return guarded([&] {
    require(h && u && F, CFFT_ERR_INVALID, "null argument");
    const size_t nu = sizeof(double) * (size_t)h->N * h->N * h->N;
    const size_t nf = 16 * (size_t)h->n_out[h->n_stage - 1];
    hip_check(hipMemcpyAsync(h->u.p, u, nu, hipMemcpyHostToDevice, h->stream), "H2D");
    run(h, static_cast<const double*>(h->u.p), static_cast<double*>(h->F.p), h->stream);
    hip_check(hipMemcpyAsync(F, h->F.p, nf, hipMemcpyDeviceToHost, h->stream), "D2H");
    hip_check(hipStreamSynchronize(h->stream), "sync");
  });
}

int cfft_p3c_forward_dev(cfft_p3c* h, const double* u, double* F, void* stream) {
  return guarded([&] {
    require(h && u && F, CFFT_ERR_INVALID, "null argument");
    run(h, u, F, static_cast<hipStream_t>(stream));
  });
}

int cfft_p3c_destroy(cfft_p3c* h) {
  return guarded([&] {
    if (!h) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    delete h;
  });
}

}  // extern "C"
