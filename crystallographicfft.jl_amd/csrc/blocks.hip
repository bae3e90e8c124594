// Real-space ASU / block CFFT (src/asu.jl, src/pack_asu.jl, src/cfft_plan.jl,
// src/execution.jl) and the full-grid centring pre-fold KRFFT
// (src/centering_prefold.jl) behind the C ABI.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/cfft_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

using namespace cfft;

namespace {

void check_dim(int32_t dim) { require(dim == 2 || dim == 3, CFFT_ERR_INVALID, "dim must be 2 or 3"); }

// exp(i * sgn * 2 pi m / n), argument reduced exactly on the host side of the
// product j*k mod n.
__device__ __forceinline__ cx<double> cis_frac(long m, long n, int sgn) {
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  return {c, sgn * s};
}

// Batched direct DFT along one axis of every packed block: element e of block
// b at local (l0, l1, l2); out[e] = scale * sum_j in[line j] exp(dir 2 pi i j k/n).
__global__ void k_block_dft(const cx<double>* __restrict__ in, cx<double>* __restrict__ out,
                            const long* __restrict__ boff, const int* __restrict__ bshape,
                            const int* __restrict__ bid, long total, int axis, int dir) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int b = bid[e];
  const long loc = e - boff[b];
  const int s0 = bshape[3 * b], s1 = bshape[3 * b + 1], s2 = bshape[3 * b + 2];
  const int c[3] = {(int)(loc % s0), (int)((loc / s0) % s1), (int)(loc / ((long)s0 * s1))};
  const int n = axis == 0 ? s0 : (axis == 1 ? s1 : s2);
  const long stride = axis == 0 ? 1 : (axis == 1 ? s0 : (long)s0 * s1);
  const int k = c[axis];
  const long base = e - (long)k * stride;
  cx<double> acc = {0.0, 0.0};
  for (int j = 0; j < n; ++j) {
    const cx<double> w = cis_frac(((long)j * k) % n, n, dir);
    acc = cadd(acc, cmul(in[base + j * stride], w));
  }
  const double sc = dir > 0 ? 1.0 / n : 1.0;
  out[e] = {acc.x * sc, acc.y * sc};
}

struct PrefoldP {
  int N[3], Ns[3], Nh[3];
  int cent;
  long S3;
};

// Centring pre-fold of the full grid into n_ch channels on the sub-grid,
// src/centering_prefold.jl:64-131.
__global__ void k_prefold(const double* __restrict__ u, cx<double>* __restrict__ ch, PrefoldP p) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= p.S3) return;
  const int ix = (int)(i % p.Ns[0]), iy = (int)((i / p.Ns[0]) % p.Ns[1]),
            iz = (int)(i / ((long)p.Ns[0] * p.Ns[1]));
  auto U = [&](int x, int y, int z) { return u[x + (long)p.N[0] * (y + (long)p.N[1] * z)]; };
  // exp(-2 pi i sum_d par_d i_d / N_d) = prod_d exp(-2 pi i par_d i_d / N_d)
  auto tw = [&](int px, int py, int pz) {
    cx<double> t = {1.0, 0.0};
    if (px) t = cmul(t, cis_frac(ix, p.N[0], -1));
    if (py) t = cmul(t, cis_frac(iy, p.N[1], -1));
    if (pz) t = cmul(t, cis_frac(iz, p.N[2], -1));
    return t;
  };
  if (p.cent == CENT_I) {
    const double v0 = U(ix, iy, iz), vx = U(ix + p.Nh[0], iy, iz), vy = U(ix, iy + p.Nh[1], iz),
                 vz = U(ix, iy, iz + p.Nh[2]);
    const double t0 = v0 + vx + vy + vz, t1 = v0 - vx - vy + vz, t2 = v0 - vx + vy - vz,
                 t3 = v0 + vx - vy - vz;
    ch[i] = {2.0 * t0, 0.0};
    const cx<double> w1 = tw(1, 1, 0), w2 = tw(1, 0, 1), w3 = tw(0, 1, 1);
    ch[p.S3 + i] = {2.0 * t1 * w1.x, 2.0 * t1 * w1.y};
    ch[2 * p.S3 + i] = {2.0 * t2 * w2.x, 2.0 * t2 * w2.y};
    ch[3 * p.S3 + i] = {2.0 * t3 * w3.x, 2.0 * t3 * w3.y};
  } else if (p.cent == CENT_F) {
    const double v0 = U(ix, iy, iz), vs = U(ix + p.Nh[0], iy, iz);
    const cx<double> w = tw(1, 1, 1);
    ch[i] = {4.0 * (v0 + vs), 0.0};
    ch[p.S3 + i] = {4.0 * (v0 - vs) * w.x, 4.0 * (v0 - vs) * w.y};
  } else if (p.cent == CENT_C) {
    const double v0 = U(ix, iy, iz), vs = U(ix + p.Nh[0], iy, iz);
    const cx<double> w = tw(1, 1, 0);
    ch[i] = {2.0 * (v0 + vs), 0.0};
    ch[p.S3 + i] = {2.0 * (v0 - vs) * w.x, 2.0 * (v0 - vs) * w.y};
  } else {  // CENT_A
    const double v0 = U(ix, iy, iz), vs = U(ix, iy + p.Nh[1], iz);
    const cx<double> w = tw(0, 1, 1);
    ch[i] = {2.0 * (v0 + vs), 0.0};
    ch[p.S3 + i] = {2.0 * (v0 - vs) * w.x, 2.0 * (v0 - vs) * w.y};
  }
}

// _merge_channels!, src/centering_prefold.jl:246-257 (-1: unmapped -> 0).
__global__ void k_merge(const cx<double>* __restrict__ ch, cx<double>* __restrict__ out,
                        const long* __restrict__ merge, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long m = merge[i];
  out[i] = m >= 0 ? ch[m] : cx<double>{0.0, 0.0};
}

}  // namespace

struct cfft_bplan {
  int dim = 3;
  std::vector<Block> blocks;
  std::vector<long> offs;
  long total = 0;
  DevBuf d_off, d_shape, d_bid, d_a, d_b;
  hipStream_t stream = nullptr;
  ~cfft_bplan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct cfft_pplan {
  int N[3], Ns[3], Nh[3];
  int cent = 0, n_ch = 0;
  std::vector<std::array<int, 3>> pars;
  SpecASU spec;
  std::vector<long> merge;
  DevBuf d_u, d_ch, d_F, d_merge;
  hipStream_t stream = nullptr;
  ~cfft_pplan() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void block_run(cfft_bplan* p, const double* in, double* out, int dir) {
  if (p->total == 0) return;
  hip_check(hipMemcpyAsync(p->d_a.p, in, 16 * p->total, hipMemcpyHostToDevice, p->stream), "H2D");
  DevBuf* src = &p->d_a;
  DevBuf* dst = &p->d_b;
  for (int axis = 0; axis < p->dim; ++axis) {
    hipLaunchKernelGGL(k_block_dft, dim3(nblk(p->total, 256)), dim3(256), 0, p->stream,
                       static_cast<const cx<double>*>(src->p), static_cast<cx<double>*>(dst->p),
                       static_cast<const long*>(p->d_off.p), static_cast<const int*>(p->d_shape.p),
                       static_cast<const int*>(p->d_bid.p), p->total, axis, dir);
    hip_check(hipGetLastError(), "k_block_dft");
    std::swap(src, dst);
  }
  hip_check(hipMemcpyAsync(out, src->p, 16 * p->total, hipMemcpyDeviceToHost, p->stream), "D2H");
  hip_check(hipStreamSynchronize(p->stream), "sync");
}

template <int LEN>
void fft_pass(cfft_pplan* p, cx<double>* data, long n_tr, long inner, long outer, long stride) {
  hipLaunchKernelGGL((k_fft_ml<double, double, LEN, -1>), dim3(nblk(n_tr, MLGeom<LEN>::G)),
                     dim3(MLGeom<LEN>::WG), 0, p->stream, data, n_tr, inner, outer, stride);
  hip_check(hipGetLastError(), "k_fft_ml");
}

void fft_axis_pre(cfft_pplan* p, int axis) {
  const long S0 = p->Ns[0], S1 = p->Ns[1], S2 = p->Ns[2], nb = p->n_ch;
  long n_tr, inner, outer, stride;
  if (axis == 0) {
    n_tr = nb * S1 * S2; inner = 1; outer = S0; stride = 1;
  } else if (axis == 1) {
    n_tr = nb * S0 * S2; inner = S0; outer = S0 * S1; stride = S0;
  } else {
    n_tr = nb * S0 * S1; inner = S0 * S1; outer = S0 * S1 * S2; stride = S0 * S1;
  }
  cx<double>* data = static_cast<cx<double>*>(p->d_ch.p);
  switch (p->Ns[axis]) {
    case 1: return;
    case 2: fft_pass<2>(p, data, n_tr, inner, outer, stride); break;
    case 3: fft_pass<3>(p, data, n_tr, inner, outer, stride); break;
    case 4: fft_pass<4>(p, data, n_tr, inner, outer, stride); break;
    case 6: fft_pass<6>(p, data, n_tr, inner, outer, stride); break;
    case 8: fft_pass<8>(p, data, n_tr, inner, outer, stride); break;
    case 12: fft_pass<12>(p, data, n_tr, inner, outer, stride); break;
    case 16: fft_pass<16>(p, data, n_tr, inner, outer, stride); break;
    case 24: fft_pass<24>(p, data, n_tr, inner, outer, stride); break;
    case 32: fft_pass<32>(p, data, n_tr, inner, outer, stride); break;
    case 48: fft_pass<48>(p, data, n_tr, inner, outer, stride); break;
    case 64: fft_pass<64>(p, data, n_tr, inner, outer, stride); break;
    default: throw Error{CFFT_ERR_UNSUPPORTED, "FFT length not supported"};
  }
}

}  // namespace

extern "C" {

int cfft_calc_asu(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops, int32_t* idx_out,
                  int32_t* depth_out, int32_t* mult_out, int64_t cap, int64_t* n_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && n_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = N[d];
    std::vector<ASUPoint> pts = calc_asu(dim, n, unpack_ops(ops, nops));
    *n_out = (int64_t)pts.size();
    const int64_t m = std::min<int64_t>(cap, (int64_t)pts.size());
    for (int64_t i = 0; i < m; ++i)
      for (int d = 0; d < 3; ++d) {
        if (idx_out) idx_out[3 * i + d] = pts[i].idx[d];
        if (depth_out) depth_out[3 * i + d] = pts[i].depth[d];
        if (d == 0 && mult_out) mult_out[i] = pts[i].mult;
      }
  });
}

int cfft_pack_asu(int32_t dim, const int32_t* idx, const int32_t* depth, int64_t n_points,
                  int32_t* blocks_out, int64_t cap, int64_t* n_out) {
  return guarded([&] {
    check_dim(dim);
    require(n_out && n_points >= 0 && (n_points == 0 || (idx && depth)), CFFT_ERR_INVALID,
            "null argument");
    std::vector<ASUPoint> pts((size_t)n_points);
    for (int64_t i = 0; i < n_points; ++i)
      for (int d = 0; d < 3; ++d) {
        pts[i].idx[d] = d < dim ? idx[3 * i + d] : 0;
        pts[i].depth[d] = d < dim ? depth[3 * i + d] : 0;
        pts[i].mult = 1;
      }
    std::vector<Block> bl = pack_asu(dim, pts);
    *n_out = (int64_t)bl.size();
    const int64_t m = std::min<int64_t>(cap, (int64_t)bl.size());
    for (int64_t i = 0; i < m && blocks_out; ++i)
      for (int d = 0; d < 3; ++d) {
        blocks_out[9 * i + d] = bl[i].start[d];
        blocks_out[9 * i + 3 + d] = bl[i].shape[d];
        blocks_out[9 * i + 6 + d] = bl[i].depth[d];
      }
  });
}

int cfft_block_plan_create(int32_t dim, const int32_t* blocks, int64_t n_blocks, cfft_bplan** out) {
  return guarded([&] {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    check_dim(dim);
    require(out && n_blocks >= 0 && (n_blocks == 0 || blocks), CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<cfft_bplan> p(new cfft_bplan());
    p->dim = dim;
    std::vector<int32_t> shape, bid;
    long off = 0;
    for (int64_t b = 0; b < n_blocks; ++b) {
      Block bl;
      for (int d = 0; d < 3; ++d) {
        bl.start[d] = blocks[9 * b + d];
        bl.shape[d] = d < dim ? blocks[9 * b + 3 + d] : 1;
        bl.depth[d] = blocks[9 * b + 6 + d];
        require(bl.shape[d] >= 1, CFFT_ERR_INVALID, "block shape must be positive");
      }
      p->blocks.push_back(bl);
      p->offs.push_back(off);
      const long sz = 1L * bl.shape[0] * bl.shape[1] * bl.shape[2];
      require(off + sz < (1L << 31), CFFT_ERR_SIZE, "block data exceeds 2^31 entries");
      for (int d = 0; d < 3; ++d) shape.push_back(bl.shape[d]);
      for (long e = 0; e < sz; ++e) bid.push_back((int32_t)b);
      off += sz;
    }
    p->total = off;
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    auto up = [&](DevBuf& buf, const void* src, size_t n) {
      buf.alloc(n);
      if (n) hip_check(hipMemcpy(buf.p, src, n, hipMemcpyHostToDevice), "upload");
    };
    up(p->d_off, p->offs.data(), p->offs.size() * sizeof(long));
    up(p->d_shape, shape.data(), shape.size() * 4);
    up(p->d_bid, bid.data(), bid.size() * 4);
    p->d_a.alloc(16 * std::max(1L, p->total));
    p->d_b.alloc(16 * std::max(1L, p->total));
    *out = p.release();
  });
}

int cfft_block_plan_info(const cfft_bplan* p, int64_t* n_blocks, int64_t* total, int32_t* blocks_out) {
  return guarded([&] {
    require(p != nullptr, CFFT_ERR_INVALID, "null plan");
    if (n_blocks) *n_blocks = (int64_t)p->blocks.size();
    if (total) *total = p->total;
    if (blocks_out)
      for (size_t i = 0; i < p->blocks.size(); ++i)
        for (int d = 0; d < 3; ++d) {
          blocks_out[9 * i + d] = p->blocks[i].start[d];
          blocks_out[9 * i + 3 + d] = p->blocks[i].shape[d];
          blocks_out[9 * i + 6 + d] = p->blocks[i].depth[d];
        }
  });
}

int cfft_block_forward(cfft_bplan* p, const double* in, double* out) {
  return guarded([&] {
    require(p && in && out, CFFT_ERR_INVALID, "null argument");
    block_run(p, in, out, -1);
  });
}

int cfft_block_inverse(cfft_bplan* p, const double* in, double* out) {
  return guarded([&] {
    require(p && in && out, CFFT_ERR_INVALID, "null argument");
    block_run(p, in, out, +1);
  });
}

int cfft_block_plan_destroy(cfft_bplan* p) {
  return guarded([&] {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    delete p;
  });
}

int cfft_prefold_plan_create(const int32_t* N, const int32_t* ops, int32_t nops, cfft_pplan** out) {
  return guarded([&] {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    require(N && ops && out && nops > 0, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<cfft_pplan> p(new cfft_pplan());
    const std::vector<Op> o = unpack_ops(ops, nops);
    for (int d = 0; d < 3; ++d) p->N[d] = N[d];
    p->cent = detect_centering(3, p->N, o);
    require(p->cent != CENT_P, CFFT_ERR_CENTERING, "CentP has no pre-fold");
    bool halve[3];
    if (p->cent == CENT_I || p->cent == CENT_F) {
      halve[0] = halve[1] = halve[2] = true;
    } else if (p->cent == CENT_C) {
      halve[0] = halve[1] = true;
      halve[2] = false;
    } else {
      halve[0] = false;
      halve[1] = halve[2] = true;
    }
    for (int d = 0; d < 3; ++d) {
      require(!halve[d] || N[d] % 2 == 0, CFFT_ERR_NOT_DIVISIBLE, "pre-fold needs even N along halved axes");
      p->Nh[d] = N[d] / 2;
      p->Ns[d] = halve[d] ? p->Nh[d] : N[d];
      require(fft_supported(p->Ns[d]), CFFT_ERR_UNSUPPORTED, "pre-fold sub-grid FFT length not supported");
    }
    // src/centering_prefold.jl:36-57
    if (p->cent == CENT_I) p->pars = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    else if (p->cent == CENT_F) p->pars = {{0, 0, 0}, {1, 1, 1}};
    else if (p->cent == CENT_C) p->pars = {{0, 0, 0}, {1, 1, 0}};
    else p->pars = {{0, 0, 0}, {0, 1, 1}};
    p->n_ch = (int)p->pars.size();
    p->spec = calc_spectral_asu(3, p->N, o);
    const long S3 = 1L * p->Ns[0] * p->Ns[1] * p->Ns[2];
    // _build_merge_table, src/centering_prefold.jl:207-230
    for (const auto& h : p->spec.idx) {
      int par[3];
      for (int d = 0; d < 3; ++d) par[d] = halve[d] ? ((h[d] % 2) + 2) % 2 : 0;
      int c = -1;
      for (int k = 0; k < p->n_ch; ++k)
        if (p->pars[k][0] == par[0] && p->pars[k][1] == par[1] && p->pars[k][2] == par[2]) c = k;
      if (c < 0) {
        p->merge.push_back(-1);
        continue;
      }
      long k3[3];
      for (int d = 0; d < 3; ++d) {
        const long v = halve[d] ? (h[d] - par[d]) / 2 : h[d];
        k3[d] = ((v % p->Ns[d]) + p->Ns[d]) % p->Ns[d];
      }
      p->merge.push_back(c * S3 + k3[0] + p->Ns[0] * (k3[1] + 1L * p->Ns[1] * k3[2]));
    }
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    p->d_u.alloc(8L * N[0] * N[1] * N[2]);
    p->d_ch.alloc(16 * S3 * p->n_ch);
    p->d_F.alloc(16 * std::max<size_t>(1, p->spec.idx.size()));
    p->d_merge.alloc(8 * std::max<size_t>(1, p->merge.size()));
    if (!p->merge.empty())
      hip_check(hipMemcpy(p->d_merge.p, p->merge.data(), 8 * p->merge.size(), hipMemcpyHostToDevice), "upload");
    *out = p.release();
  });
}

int cfft_prefold_plan_info(const cfft_pplan* p, int32_t* info, int32_t* idx_out) {
  return guarded([&] {
    require(p != nullptr, CFFT_ERR_INVALID, "null plan");
    if (info) {
      int32_t v[6] = {(int32_t)p->spec.idx.size(), p->n_ch, p->cent, p->Ns[0], p->Ns[1], p->Ns[2]};
      std::memcpy(info, v, sizeof(v));
    }
    if (idx_out)
      for (size_t i = 0; i < p->spec.idx.size(); ++i)
        for (int d = 0; d < 3; ++d) idx_out[3 * i + d] = p->spec.idx[i][d];
  });
}

int cfft_prefold_forward(cfft_pplan* p, const double* u, double* F) {
  return guarded([&] {
    require(p && u && F, CFFT_ERR_INVALID, "null argument");
    const long N3 = 1L * p->N[0] * p->N[1] * p->N[2];
    hip_check(hipMemcpyAsync(p->d_u.p, u, 8 * N3, hipMemcpyHostToDevice, p->stream), "H2D");
    PrefoldP pp{};
    for (int d = 0; d < 3; ++d) {
      pp.N[d] = p->N[d];
      pp.Ns[d] = p->Ns[d];
      pp.Nh[d] = p->Nh[d];
    }
    pp.cent = p->cent;
    pp.S3 = 1L * p->Ns[0] * p->Ns[1] * p->Ns[2];
    hipLaunchKernelGGL(k_prefold, dim3(nblk(pp.S3, 256)), dim3(256), 0, p->stream,
                       static_cast<const double*>(p->d_u.p), static_cast<cx<double>*>(p->d_ch.p), pp);
    hip_check(hipGetLastError(), "k_prefold");
    for (int axis = 0; axis < 3; ++axis) fft_axis_pre(p, axis);
    const int n = (int)p->spec.idx.size();
    if (n > 0) {
      hipLaunchKernelGGL(k_merge, dim3(nblk(n, 256)), dim3(256), 0, p->stream,
                         static_cast<const cx<double>*>(p->d_ch.p), static_cast<cx<double>*>(p->d_F.p),
                         static_cast<const long*>(p->d_merge.p), n);
      hip_check(hipGetLastError(), "k_merge");
      hip_check(hipMemcpyAsync(F, p->d_F.p, 16L * n, hipMemcpyDeviceToHost, p->stream), "D2H");
    }
    hip_check(hipStreamSynchronize(p->stream), "sync");
  });
}

int cfft_prefold_plan_destroy(cfft_pplan* p) {
  return guarded([&] {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    delete p;
  });
}

}  // extern "C"
