// C ABI of libcfft_hip.so: plan objects with device-resident tables and
// buffers, and the launch sequences of the KRFFT / diffusion pipelines.
// The entry points and their reference counterparts are listed in
// include/cfft_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cfft_hip.h"
#include "kernels.hpp"
#include "pipeline.hpp"
#include "ensemble.hpp"
#include "plan.hpp"
#include "common.hpp"

using namespace cfft;

thread_local std::string cfft::g_last_error;

std::recursive_mutex& cfft::setup_mutex() {
  static std::recursive_mutex m;
  return m;
}

struct cfft_plan {
  int N[3], M[3], L[3], H[3];
  int device = 0;  // device current when the plan was created
  int d = 1, n_spec = 0, cent = 0, n_ch = 0, path = 0, prec = 0, batch = 1, P = 1;
  double ds = 0.0, lattice[9];
  std::vector<Op> ops;
  std::vector<int> rep;
  SpecASU spec;
  std::vector<double> K;
  std::vector<std::array<int, 3>> offs;
  int chmap[8];
  long H3 = 0, M3 = 0, n_alive = 0, N3 = 0;
  long n_rep = 0;  // orbit representatives of the alive M-grid (path QR)
  int qs_items = 0;  // plane-pair items per batch entry (path QS)
  int qs_sep = 0;    // path QS: translations separable in the sign pattern
  size_t esz = 8;  // element size of the plan's real type
  hipStream_t stream = nullptr;
  // device tables
  DevBuf d_hs, d_K, d_W, d_twM, d_btab, d_Q, d_gidx, d_fmap, d_chmap;
  // device work buffers
  DevBuf d_f0, d_ch, d_ch2, d_F, d_E, d_u;
  DevBuf d_qmap, d_grep;  // path QR: element -> (rep, phase); rep values
  DevBuf d_qsep;          // path QS: axis factors F[d][s][q]
  DevBuf d_symops;        // ops that keep the subgrid, acting on it (12 int32 each)
  int n_sym = 0;
  std::vector<int32_t> fmap_host;
  // options (cfft_plan_set_option)
  int persistent = 0;  // contour sweeps as one cooperative launch
  int fuse = 1;        // fuse P5 of a step with P1 of the next
  int graph = 1;       // replay multi-step chains from captured hipGraphs
  int tile_rows = 0;   // y rows per fold/unfold tile (0: the default rule)
  int arith32 = 0;     // mixed precision on path QS: FP32 twiddle and fold-factor arithmetic (option)
  int lanes = 0;       // path QS: lane split of the plane-pair and tile transforms (0: the default)
  int pair_threads = 0;  // path QS: threads per plane-pair workgroup (0: 256)
  int axis_split = 0;    // path QS: the per-axis chain (x mixing in the tile kernel, single planes)
  hipStream_t cap_stream = nullptr;  // capture stream of the chains
  struct Cached {
    std::string key;
    hipGraphExec_t exec;
  };
  std::vector<Cached> graphs;  // most recently used last
  // per-stage hipEvent timer (cfft_profile / cfft_profile_read)
  bool prof = false;
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  ~cfft_plan() {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    for (auto& m : marks) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

bool is_qr(const cfft_plan* p) { return p->path == CFFT_PATH_QR; }
bool is_qs(const cfft_plan* p) { return p->path == CFFT_PATH_QS; }
bool is_q_path(const cfft_plan* p) {
  return p->path == CFFT_PATH_M2Q || p->path == CFFT_PATH_M7Q || p->path == CFFT_PATH_QR ||
         p->path == CFFT_PATH_QS;
}
// paths whose steps are the fused launch chain (P51 between steps)
bool chain_path(const cfft_plan* p) { return !is_q_path(p) || is_qr(p) || is_qs(p); }

// Mixed precision on path QS: the fold, unfold, transform and Q-multiply
// arithmetic of its kernels in FP32 (tables stay FP64 in memory and are
// rounded where they are used); the other paths form twiddle products in FP64.
bool fp32_arith(const cfft_plan* p) { return p->prec == CFFT_MIXED && p->path == CFFT_PATH_QS && p->arith32; }


// Uploads hold the set-up lock (allocation and a blocking copy); the host
// work of plan building between them runs unlocked, so several threads can
// build plans at once.
template <typename T>
void upload_complex(DevBuf& buf, const std::vector<double>& re, const std::vector<double>& im) {
  std::vector<cx<T>> h(re.size());
  for (size_t i = 0; i < re.size(); ++i) h[i] = {(T)re[i], (T)im[i]};
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  buf.alloc(h.size() * sizeof(cx<T>));
  if (!h.empty()) hip_check(hipMemcpy(buf.p, h.data(), buf.bytes, hipMemcpyHostToDevice), "upload");
}

template <typename T>
void upload_real(DevBuf& buf, const std::vector<double>& v) {
  std::vector<T> h(v.begin(), v.end());
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  buf.alloc(h.size() * sizeof(T));
  if (!h.empty()) hip_check(hipMemcpy(buf.p, h.data(), buf.bytes, hipMemcpyHostToDevice), "upload");
}

void upload_bytes(DevBuf& buf, const void* src, size_t n) {
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  buf.alloc(n);
  if (n) hip_check(hipMemcpy(buf.p, src, n, hipMemcpyHostToDevice), "upload");
}

// Path QS axis factors F[d][s][q], q < M_d, s = sign pattern of the
// representative op (bit i: R_ii = -1):
//   F_d[s][q] = sum_{a in {0,1}} K_d(q + M_d a) exp(-2 pi i (q + M_d a) t_{s,d} / N_d),
// K_d(h) = exp(-(2 pi hc / A_dd)^2 ds), hc the centred frequency; 1/8 is folded
// into F_0.  Their product over d is Q_s(q) of the closed form used by the
// other Q paths (first row of B^-1 diag(K) B, src/q_fused_krfft.jl:182-413)
// when the cell is orthogonal and every R_s is diagonal.
bool orthogonal_cell(const double* A) {  // column-major 3 x 3
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && A[i + 3 * j] != 0.0) return false;
  return true;
}

void build_qs_table(cfft_plan* p) {
  const double* A = p->lattice;  // column-major 3 x 3
  require(orthogonal_cell(A), CFFT_ERR_UNSUPPORTED, "path QS needs an orthogonal cell");
  int tsel[8][3];
  bool seen[8] = {false, false, false, false, false, false, false, false};
  for (int b = 0; b < p->d; ++b) {
    const Op& g = p->ops[p->rep[b]];
    int sf = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const int v = g.R[i][j];
        require(i == j ? (v == 1 || v == -1) : v == 0, CFFT_ERR_UNSUPPORTED,
                "path QS needs diagonal representative ops");
      }
    for (int i = 0; i < 3; ++i) sf |= (g.R[i][i] < 0 ? 1 : 0) << i;
    require(!seen[sf], CFFT_ERR_UNSUPPORTED, "path QS needs 8 distinct sign patterns");
    seen[sf] = true;
    for (int i = 0; i < 3; ++i) tsel[sf][i] = g.t[i];
  }
  // translation component d a function of sign bit d alone (mirrors with the
  // magic shift: t_d = -1 on flipped axes): the octet sum factors per axis
  p->qs_sep = 1;
  for (int sf = 0; sf < 8; ++sf)
    for (int d = 0; d < 3; ++d)
      if ((((long)tsel[sf][d] - tsel[sf & (1 << d)][d]) % p->N[d]) != 0) p->qs_sep = 0;
  const int Ms = p->M[0];
  std::vector<double> re((size_t)3 * 8 * Ms), im(re.size());
  for (int d = 0; d < 3; ++d) {
    const double bd = 2.0 * M_PI / A[d + 3 * d];
    for (int sf = 0; sf < 8; ++sf)
      for (int q = 0; q < p->M[d]; ++q) {
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < 2; ++a) {
          const long h = q + (long)p->M[d] * a;
          const double kv = bd * k_centered((int)h, p->N[d]);
          const double kd = std::exp(-kv * kv * p->ds);
          const long long j = (long long)h * tsel[sf][d] * (p->P / p->N[d]);
          double cr, ci;
          cis2pi(-j, p->P, cr, ci);
          sr += kd * cr;
          si += kd * ci;
        }
        const double sc = d == 0 ? 0.125 : 1.0;
        re[((size_t)d * 8 + sf) * Ms + q] = sr * sc;
        im[((size_t)d * 8 + sf) * Ms + q] = si * sc;
      }
  }
  if (!p->d_qsep.p) {
    upload_complex<double>(p->d_qsep, re, im);
  } else {
    // same allocation: captured chains keep pointing at it
    std::vector<cx<double>> h(re.size());
    for (size_t i = 0; i < re.size(); ++i) h[i] = {re[i], im[i]};
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    hip_check(hipMemcpy(p->d_qsep.p, h.data(), p->d_qsep.bytes, hipMemcpyHostToDevice), "upload QS factors");
  }
  p->qs_items = 0;
  // orbits of x -> xbar: xbar = (H - x) mod H (offset 0) or H - 1 - x (offset 1)
  for (int c = 0; c < p->n_ch; ++c) p->qs_items += p->offs[c][0] ? (p->H[0] + 1) / 2 : p->H[0] / 2 + 1;
}

// Phase table W[j] = exp(2 pi i j/P) and fold twiddle tables (the FFT
// twiddles are compile-time constants).
template <typename T>
void build_tables(cfft_plan* p) {
  std::vector<double> re, im;
  re.resize(p->P);
  im.resize(p->P);
  for (int j = 0; j < p->P; ++j) cis2pi(j, p->P, re[j], im[j]);
  upload_complex<T>(p->d_W, re, im);
  // twM: exp(-2 pi i n / M_d), n < M_d, d = 0..2 concatenated
  re.clear();
  im.clear();
  for (int d = 0; d < 3; ++d)
    for (int n = 0; n < p->M[d]; ++n) {
      double a, b;
      cis2pi(-n, p->M[d], a, b);
      re.push_back(a);
      im.push_back(b);
    }
  upload_complex<T>(p->d_twM, re, im);
  upload_real<T>(p->d_K, p->K);
}

void build_plan(cfft_plan* p, const int* N, const std::vector<Op>& ops, double ds,
                const double* lattice, int path, int prec, int batch) {
  require(batch >= 1, CFFT_ERR_INVALID, "batch must be >= 1");
  require(prec == CFFT_FP64 || prec == CFFT_MIXED, CFFT_ERR_INVALID, "unknown precision");
  require(path >= CFFT_PATH_AUTO && path <= CFFT_PATH_QS, CFFT_ERR_UNKNOWN_METHOD,
          "Unknown method: path " + std::to_string(path));
  for (int d = 0; d < 3; ++d) require(N[d] >= 1, CFFT_ERR_INVALID, "grid size must be positive");
  p->ops = ops;
  p->ds = ds;
  std::memcpy(p->lattice, lattice, sizeof(double) * 9);
  p->prec = prec;
  p->esz = prec == CFFT_FP64 ? 8 : 4;
  p->batch = batch;
  for (int d = 0; d < 3; ++d) p->N[d] = N[d];
  auto_L(3, ops, p->L);
  for (int d = 0; d < 3; ++d) {
    p->M[d] = N[d] / p->L[d];
    require(p->M[d] * p->L[d] == N[d], CFFT_ERR_NOT_DIVISIBLE,
            "Grid size N=(" + std::to_string(N[0]) + "," + std::to_string(N[1]) + "," +
                std::to_string(N[2]) + ") not divisible by auto L.");
  }
  p->rep = select_rep_ops(3, ops, p->L, N);
  p->d = (int)p->rep.size();
  require(p->d <= 8, CFFT_ERR_UNSUPPORTED, "more than 8 subgrid cosets (L beyond 2 per axis)");
  p->cent = detect_centering(3, N, ops);
  bool allL2 = p->L[0] == 2 && p->L[1] == 2 && p->L[2] == 2;
  if (path == CFFT_PATH_AUTO) path = (p->cent != CENT_P && allL2) ? CFFT_PATH_M7 : CFFT_PATH_M2;
  if (path == CFFT_PATH_M7 || path == CFFT_PATH_M7Q) {
    require(p->cent != CENT_P, CFFT_ERR_CENTERING,
            "M7 requires a centred lattice (I/F/C/A); use the M2 path for P-centring.");
    require(allL2, CFFT_ERR_CENTERING, "M7 requires L = (2,2,2)");
  }
  if (path == CFFT_PATH_QS) require(allL2 && p->d == 8, CFFT_ERR_UNSUPPORTED, "path QS needs L = (2,2,2)");
  p->path = path;
  for (int d = 0; d < 3; ++d) {
    require(p->M[d] % 2 == 0, CFFT_ERR_ODD_M,
            "Subgrid dimensions M must be even for the radix-2 fold (M" + std::to_string(d) +
                " = " + std::to_string(p->M[d]) + ")");
    p->H[d] = p->M[d] / 2;
    require(fft_supported(p->H[d]), CFFT_ERR_UNSUPPORTED,
            "FFT length " + std::to_string(p->H[d]) + " not supported (need 2^a 3^b <= 64)");
  }
  const bool centred_fold = path == CFFT_PATH_M7 || path == CFFT_PATH_M7Q ||
                            ((path == CFFT_PATH_QR || path == CFFT_PATH_QS) && p->cent != CENT_P && allL2);
  p->offs = centred_fold ? alive_offsets(p->cent) : alive_offsets(CENT_P);
  p->n_ch = (int)p->offs.size();
  for (int i = 0; i < 8; ++i) p->chmap[i] = -1;
  for (int c = 0; c < p->n_ch; ++c) {
    const auto& o = p->offs[c];
    p->chmap[o[0] | (o[1] << 1) | (o[2] << 2)] = c;
  }
  p->H3 = 1L * p->H[0] * p->H[1] * p->H[2];
  p->M3 = 1L * p->M[0] * p->M[1] * p->M[2];
  p->N3 = 1L * N[0] * N[1] * N[2];
  p->n_alive = p->n_ch * p->H3;
  long long P = lcm3(N[0], N[1], N[2]);
  require(P <= 2048, CFFT_ERR_UNSUPPORTED, "lcm(N) > 2048 not supported by the phase table");
  p->P = (int)P;

  p->spec = calc_spectral_asu(3, N, ops);
  p->n_spec = (int)p->spec.idx.size();
  require(p->n_spec < 0xFFFFF, CFFT_ERR_UNSUPPORTED, "spectral ASU too large");
  p->K = kernel_spec(p->spec, ds, lattice);

  {
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
  }
  build_tables<double>(p);  // twiddles, phases and K stay FP64 in both precisions

  // signed spectral frequencies (get_k_vector)
  std::vector<int32_t> hs(3 * (size_t)p->n_spec);
  for (int i = 0; i < p->n_spec; ++i)
    for (int d = 0; d < 3; ++d) hs[3 * i + d] = k_signed(p->spec.idx[i][d], N[d]);
  upload_bytes(p->d_hs, hs.data(), hs.size() * 4);
  upload_bytes(p->d_chmap, p->chmap, sizeof(p->chmap));

  // alive parity index (c, k) -> q = 2k + off_c
  // channel-space index e = c H^3 + (x H2 + z) H1 + y  (layout [c][x][z][y])
  auto q_of = [&](long e, int* q) {
    int c = (int)(e / p->H3);
    long k = e - c * p->H3;
    int k1 = (int)(k % p->H[1]), k2 = (int)((k / p->H[1]) % p->H[2]), k0 = (int)(k / (1L * p->H[1] * p->H[2]));
    q[0] = 2 * k0 + p->offs[c][0];
    q[1] = 2 * k1 + p->offs[c][1];
    q[2] = 2 * k2 + p->offs[c][2];
  };
  // cosets alpha in Iterators.product order (first dimension fastest)
  std::vector<std::array<int, 3>> alphas;
  for (int a2 = 0; a2 < p->L[2]; ++a2)
    for (int a1 = 0; a1 < p->L[1]; ++a1)
      for (int a0 = 0; a0 < p->L[0]; ++a0) alphas.push_back({a0, a1, a2});

  if (is_qs(p)) {
    require(p->H[0] == p->H[1] && p->H[1] == p->H[2], CFFT_ERR_UNSUPPORTED, "path QS needs a cubic grid");
    build_qs_table(p);
  } else if (!is_q_path(p)) {
    ReverseLookup rl = reverse_lookup(p->spec, ops);
    std::vector<uint32_t> tab((size_t)8 * p->n_alive, 0xFFFFFu);
    for (long e = 0; e < p->n_alive; ++e) {
      int q[3];
      q_of(e, q);
      for (int a = 0; a < p->d; ++a) {
        long h[3];
        for (int dd = 0; dd < 3; ++dd) h[dd] = q[dd] + 1L * p->M[dd] * alphas[a][dd];
        long li = h[0] + N[0] * (h[1] + 1L * N[1] * h[2]);
        uint32_t ent = 0xFFFFFu;
        if (rl.idx[li] >= 0)
          ent = (uint32_t)rl.idx[li] | ((uint32_t)rl.j[li] << 20) | ((uint32_t)rl.conj[li] << 31);
        tab[(size_t)e * 8 + a] = ent;
      }
    }
    upload_bytes(p->d_btab, tab.data(), tab.size() * 4);
  } else {
    // Q_b(q) = (1/d) sum_a K(h_a) exp(-2 pi i h_a.t_b/N), h_a = q + M alpha_a
    // (first row of B^-1 diag(K) B, src/q_fused_krfft.jl:182-413), and the
    // channel-space gather index of R_b^T q mod M.
    double A[3][3], c[3][3], B[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = lattice[i + 3 * j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        c[i][j] = A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
      }
    double det = A[0][0] * c[0][0] + A[0][1] * c[0][1] + A[0][2] * c[0][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) B[i][j] = 2.0 * M_PI * (c[i][j] / det);
    // path QR: orbits of the alive M-grid points under the ops that keep the
    // stride-L subgrid (t = 0 mod L, s = t/L): G0(R^T q) = exp(2 pi i q.s/M) G0(q)
    // (the orbit expansion of src/centering_fold.jl:630-731).  Q is evaluated at
    // the representatives only; every element keeps (rep, phase exponent).
    std::vector<long> reps;
    std::vector<uint32_t> qmap;
    if (is_qr(p)) {
      std::vector<const Op*> rem;
      for (const Op& g : ops) {
        bool ok = true;
        for (int dd = 0; dd < 3; ++dd) ok = ok && (((g.t[dd] % p->L[dd]) + p->L[dd]) % p->L[dd] == 0);
        if (ok) rem.push_back(&g);
      }
      std::vector<int32_t> rep_of(p->n_alive, -1), jph(p->n_alive, 0);
      auto elem_of = [&](const int* q) -> long {
        const int c = p->chmap[(q[0] & 1) | ((q[1] & 1) << 1) | ((q[2] & 1) << 2)];
        if (c < 0) return -1;
        return c * p->H3 + ((q[0] >> 1) * 1L * p->H[2] + (q[2] >> 1)) * p->H[1] + (q[1] >> 1);
      };
      std::vector<long> queue;
      for (long e = 0; e < p->n_alive; ++e) {
        if (rep_of[e] >= 0) continue;
        const int r = (int)reps.size();
        reps.push_back(e);
        rep_of[e] = r;
        jph[e] = 0;
        queue.assign(1, e);
        for (size_t qi = 0; qi < queue.size(); ++qi) {
          const long cur = queue[qi];
          int q[3];
          q_of(cur, q);
          for (const Op* g : rem) {
            int rq[3];
            long long jj = 0;
            for (int d1 = 0; d1 < 3; ++d1) {
              long v = 0;
              for (int d2 = 0; d2 < 3; ++d2) v += 1L * g->R[d2][d1] * q[d2];
              rq[d1] = (int)(((v % p->M[d1]) + p->M[d1]) % p->M[d1]);
              jj += 1LL * q[d1] * (g->t[d1] / p->L[d1]) * (P / p->M[d1]);
            }
            const long e2 = elem_of(rq);
            require(e2 >= 0, CFFT_ERR_INVALID, "orbit leaves the alive parities");
            if (rep_of[e2] >= 0) continue;
            rep_of[e2] = r;
            jph[e2] = (int32_t)((((jph[cur] + jj) % P) + P) % P);
            queue.push_back(e2);
          }
        }
      }
      require((long)reps.size() < (1L << 20), CFFT_ERR_UNSUPPORTED, "too many orbit representatives");
      qmap.resize(p->n_alive);
      for (long e = 0; e < p->n_alive; ++e) qmap[e] = (uint32_t)rep_of[e] | ((uint32_t)jph[e] << 20);
    }
    const long n_q = is_qr(p) ? (long)reps.size() : p->n_alive;
    std::vector<double> qr((size_t)p->d * n_q), qi((size_t)p->d * n_q);
    std::vector<int32_t> gidx((size_t)p->d * n_q);
    std::vector<double> Kh(p->d);
    for (long ei = 0; ei < n_q; ++ei) {
      const long e = is_qr(p) ? reps[ei] : ei;
      int q[3];
      q_of(e, q);
      for (int a = 0; a < p->d; ++a) {
        double hc[3];
        for (int dd = 0; dd < 3; ++dd)
          hc[dd] = k_centered(q[dd] + p->M[dd] * alphas[a][dd], N[dd]);
        double k2 = 0;
        for (int i = 0; i < 3; ++i) {
          double kv = B[i][0] * hc[0] + B[i][1] * hc[1] + B[i][2] * hc[2];
          k2 += kv * kv;
        }
        Kh[a] = std::exp(-k2 * ds);
      }
      for (int b = 0; b < p->d; ++b) {
        const Op& g = ops[p->rep[b]];
        double sr = 0, si = 0;
        for (int a = 0; a < p->d; ++a) {
          long long j = 0;
          for (int dd = 0; dd < 3; ++dd)
            j += 1LL * (q[dd] + p->M[dd] * alphas[a][dd]) * g.t[dd] * (P / N[dd]);
          double cr, ci;
          cis2pi(-j, P, cr, ci);
          sr += Kh[a] * cr;
          si += Kh[a] * ci;
        }
        qr[(size_t)b * n_q + ei] = sr / p->d;
        qi[(size_t)b * n_q + ei] = si / p->d;
        int rq[3];
        for (int d1 = 0; d1 < 3; ++d1) {
          long v = 0;
          for (int d2 = 0; d2 < 3; ++d2) v += 1L * g.R[d2][d1] * q[d2];
          rq[d1] = (int)(((v % p->M[d1]) + p->M[d1]) % p->M[d1]);
        }
        int cc = p->chmap[(rq[0] & 1) | ((rq[1] & 1) << 1) | ((rq[2] & 1) << 2)];
        gidx[(size_t)b * n_q + ei] =
            cc < 0 ? -1
                   : (int32_t)(cc * p->H3 + ((rq[0] >> 1) * 1L * p->H[2] + (rq[2] >> 1)) * p->H[1] + (rq[1] >> 1));
      }
    }
    if (prec == CFFT_FP64 || is_qr(p))
      upload_complex<double>(p->d_Q, qr, qi);  // QR: FP64 coefficients in both precisions
    else
      upload_complex<float>(p->d_Q, qr, qi);
    upload_bytes(p->d_gidx, gidx.data(), gidx.size() * 4);
    if (is_qr(p)) {
      p->n_rep = (long)reps.size();
      upload_bytes(p->d_qmap, qmap.data(), qmap.size() * 4);
      p->d_grep.alloc(2 * p->esz * (size_t)p->n_rep * batch);
    }
  }

  // The ops that map the stride-L subgrid onto itself, as maps of the
  // subgrid: x'_i = sum_j R_ij L_j / L_i x_j + t_i / L_i (mod M_i), one entry
  // per distinct map.  A field symmetric under the group is invariant under
  // each; averaging over them projects onto those fields.
  {
    std::vector<std::array<int, 12>> sym;
    for (const Op& g : ops) {
      bool keeps = true;
      std::array<int, 12> e{};
      for (int i = 0; i < 3 && keeps; ++i) {
        keeps = g.t[i] % p->L[i] == 0;
        for (int j = 0; j < 3 && keeps; ++j) {
          keeps = (g.R[i][j] * p->L[j]) % p->L[i] == 0;
          e[3 * i + j] = g.R[i][j] * p->L[j] / p->L[i];
        }
        e[9 + i] = ((g.t[i] / p->L[i]) % p->M[i] + p->M[i]) % p->M[i];
      }
      if (keeps && std::find(sym.begin(), sym.end(), e) == sym.end()) sym.push_back(e);
    }
    p->n_sym = (int)sym.size();
    std::vector<int32_t> flat;
    for (const auto& e : sym) flat.insert(flat.end(), e.begin(), e.end());
    upload_bytes(p->d_symops, flat.data(), flat.size() * sizeof(int32_t));
  }

  const size_t es = p->esz;
  p->d_f0.alloc(es * p->M3 * batch);
  p->d_ch.alloc(2 * es * p->n_alive * batch);
  if (is_q_path(p) && !is_qr(p)) p->d_ch2.alloc(2 * es * p->n_alive * batch);
  p->d_F.alloc(2 * es * (size_t)std::max(p->n_spec, 1) * batch);
  p->d_E.alloc(sizeof(double) * p->M3 * batch);  // interaction factors: FP64 in both precisions
}


// Brackets one kernel launch with hipEvents when the plan's stage timer is on.
struct StageTimer {
  cfft_plan* p;
  hipStream_t s;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(cfft_plan* p_, hipStream_t s_, int stage_) : p(p_), s(s_), stage(stage_) {
    if (!p->prof) return;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventCreate(&b), "hipEventCreate");
    hip_check(hipEventRecord(a, s), "hipEventRecord");
  }
  ~StageTimer() {
    if (!p->prof || !a) return;
    if (hipEventRecord(b, s) == hipSuccess) p->marks.push_back({stage, a, b});
  }
};

template <typename T>
void launch_check(const char* what) {
  hip_check(hipGetLastError(), what);
}

#define CFFT_LEN_SWITCH(len, MACRO)                                                 \
  switch (len) {                                                                    \
    MACRO(1) MACRO(2) MACRO(3) MACRO(4) MACRO(6) MACRO(8) MACRO(12) MACRO(16)       \
    MACRO(24) MACRO(32) MACRO(48) MACRO(64)                                         \
    default: throw Error{CFFT_ERR_UNSUPPORTED, "FFT length not supported"};        \
  }

// Multi-lane FFT along y (axis 1) or z (axis 2) of a channel array in the
// [b][c][x][z][y] layout (generic path for non-cubic H-grids).
template <typename T, int DIR>
void fft_axis(cfft_plan* p, cx<T>* data, int axis, hipStream_t s) {
  const long H0 = p->H[0], H1 = p->H[1], H2 = p->H[2];
  const long nb = (long)p->batch * p->n_ch;
  const int len = p->H[axis];
  if (len == 1) return;
  long n_tr, inner, outer, stride;
  if (axis == 1) {
    n_tr = nb * H0 * H2; inner = 1; outer = H1; stride = 1;
  } else {
    n_tr = nb * H0 * H1; inner = H1; outer = H1 * H2; stride = H1;
  }
  StageTimer st(p, s, (DIR < 0 ? CFFT_STAGE_FFT_X : CFFT_STAGE_IFFT_X) + axis);
#define CFFT_FFT_CASE(LEN)                                                                      \
  case LEN:                                                                                     \
    hipLaunchKernelGGL((k_fft_ml<T, double, LEN, DIR>),                                          \
                       dim3(nblk(n_tr, MLGeom<LEN>::G)), dim3(MLGeom<LEN>::WG), 0, s, data, n_tr, \
                       inner, outer, stride);                                                   \
    break;
  CFFT_LEN_SWITCH(len, CFFT_FFT_CASE)
#undef CFFT_FFT_CASE
  launch_check<T>("k_fft_ml");
}

template <typename T>
FoldP fold_params(cfft_plan* p, const T* const* E, bool contig) {
  FoldP fp{};
  fp.M0 = p->M[0]; fp.M1 = p->M[1]; fp.M2 = p->M[2];
  fp.H0 = p->H[0]; fp.H1 = p->H[1]; fp.H2 = p->H[2];
  fp.n_ch = p->n_ch;
  for (int c = 0; c < p->n_ch; ++c) fp.offs[c] = p->offs[c][0] | (p->offs[c][1] << 1) | (p->offs[c][2] << 2);
  fp.batch = p->batch;
  fp.E_contig = contig ? 1 : 0;
  for (int b = 0; b < kMaxEPtr; ++b) fp.E[b] = nullptr;
  if (E) {
    if (contig)
      fp.E[0] = E[0];
    else
      for (int b = 0; b < p->batch && b < kMaxEPtr; ++b) fp.E[b] = E[b];
  }
  return fp;
}

// Interaction factors of one launch: none, one contiguous batch x M^3 array,
// or per-entry pointers switching from E1 to E2 at step sw[b].
struct ESel {
  int mode = 0;  // 0 none, 1 contiguous, 2 per entry
  const void* contig = nullptr;
  const void* const* E1 = nullptr;
  const void* const* E2 = nullptr;
  const int32_t* sw = nullptr;
};

PipeP pipe_params(cfft_plan* p, const ESel& es) {
  PipeP pp;
  std::memset(&pp, 0, sizeof(pp));  // padding too: the bytes key the graph cache
  for (int d = 0; d < 3; ++d) {
    pp.M[d] = p->M[d];
    pp.H[d] = p->H[d];
    pp.Pn[d] = p->P / p->N[d];
  }
  pp.n_ch = p->n_ch;
  for (int c = 0; c < 8; ++c) pp.offs[c] = 0;
  for (int c = 0; c < p->n_ch; ++c) pp.offs[c] = p->offs[c][0] | (p->offs[c][1] << 1) | (p->offs[c][2] << 2);
  for (int i = 0; i < 8; ++i) pp.chmap[i] = p->chmap[i];
  pp.batch = p->batch;
  // y rows per tile: a power of two (shift/mask tile indexing), by default the
  // largest <= 8 dividing H1; n_ch * YB <= 64 rows fit the tile's LDS
  int yb = p->tile_rows ? p->tile_rows : 8;
  while (p->H[1] % yb) yb >>= 1;
  pp.YB = yb;
  pp.ybs = yb == 8 ? 3 : (yb == 4 ? 2 : (yb == 2 ? 1 : 0));
  pp.n_spec = p->n_spec;
  pp.d = p->d;
  pp.P = p->P;
  for (int a = 0; a < p->d; ++a) {
    const Op& g = p->ops[p->rep[a]];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) pp.R[a][i * 3 + j] = g.R[i][j];
    for (int i = 0; i < 3; ++i) pp.t[a][i] = g.t[i];
  }
  pp.E_contig = es.mode == 2 ? 0 : 1;
  for (int b = 0; b < kMaxEPtr; ++b) {
    pp.E1[b] = nullptr;
    pp.E2[b] = nullptr;
    pp.sw[b] = 0;
  }
  if (es.mode == 1) pp.E1[0] = es.contig;
  if (es.mode == 2) {
    require(p->batch <= kMaxEPtr, CFFT_ERR_INVALID, "per-entry factors need batch <= 64");
    for (int b = 0; b < p->batch; ++b) {
      pp.E1[b] = es.E1[b];
      pp.E2[b] = es.E2 ? es.E2[b] : es.E1[b];
      pp.sw[b] = es.sw ? es.sw[b] : INT32_MAX;
    }
  }
  pp.scale_unfold = 1.0 / (8.0 * (double)p->H3);
  pp.scale_irecon = 1.0 / p->d;
  pp.n_rep = p->n_rep;
  pp.qs_items = p->qs_items;
  pp.qs_sep = p->qs_sep;
  return pp;
}

// ch: the channel buffer of the call (nullptr: the plan's own).
PipeBufs pipe_bufs(cfft_plan* p, const void* f0_in, void* f0_out, void* F, bool withK, void* ch = nullptr) {
  PipeBufs B{};
  B.f0_in = f0_in;
  B.f0_out = f0_out;
  B.ch = ch ? ch : p->d_ch.p;
  B.F = F;
  B.hs = static_cast<const int*>(p->d_hs.p);
  B.K = withK ? static_cast<const double*>(p->d_K.p) : nullptr;
  B.W = p->d_W.p;
  B.twM = p->d_twM.p;
  B.btab = static_cast<const uint32_t*>(p->d_btab.p);
  B.grep = p->d_grep.p;
  B.qmap = static_cast<const uint32_t*>(p->d_qmap.p);
  B.qsep = p->d_qsep.p;
  return B;
}

bool cubic(const cfft_plan* p) { return p->H[0] == p->H[1] && p->H[1] == p->H[2]; }
long n_tiles(const cfft_plan* p, const PipeP& pp) { return (long)p->batch * p->H[2] * (p->H[1] / pp.YB); }
long n_planes(const cfft_plan* p) { return (long)p->batch * p->n_ch * p->H[0]; }

template <typename T>
void launch_p1(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_FOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
  bool f32 = false;  // FP32-arithmetic kernels are instantiated for FP32 data only
  if constexpr (std::is_same<T, float>::value) {
    if (fp32_arith(p)) {
#define CFFT_P1(L) \
  case L: hipLaunchKernelGGL((k_p1<T, float, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
      CFFT_LEN_SWITCH(p->H[0], CFFT_P1)
#undef CFFT_P1
      f32 = true;
    }
  }
  if (!f32) {
#define CFFT_P1(L) \
  case L: hipLaunchKernelGGL((k_p1<T, double, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
    CFFT_LEN_SWITCH(p->H[0], CFFT_P1)
#undef CFFT_P1
  }
  launch_check<T>("k_p1");
}

template <typename T>
void launch_p2(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  if (!cubic(p)) {
    fft_axis<T, -1>(p, static_cast<cx<T>*>(B.ch), 1, s);
    fft_axis<T, -1>(p, static_cast<cx<T>*>(B.ch), 2, s);
    return;
  }
  StageTimer st(p, s, CFFT_STAGE_FFT_Y);
  const unsigned g = (unsigned)n_planes(p);
#define CFFT_P2(L) \
  case L: hipLaunchKernelGGL((k_p2<T, double, L, L>), dim3(g), dim3(kNT), 0, s, pp, B); break;
  CFFT_LEN_SWITCH(p->H[1], CFFT_P2)
#undef CFFT_P2
  launch_check<T>("k_p2");
}

template <typename T>
void launch_p3(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  if (p->n_spec == 0) return;
  StageTimer st(p, s, CFFT_STAGE_RECON);
  hipLaunchKernelGGL((k_p3<T, double>), dim3(nblk(8L * p->n_spec * p->batch, kNT)), dim3(kNT), 0, s, pp, B);
  launch_check<T>("k_p3");
}

template <typename T>
void launch_p4(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  if (!cubic(p)) {
    IreconP ip{};
    ip.d = p->d;
    ip.n_alive = p->n_alive;
    ip.n_spec = p->n_spec;
    ip.batch = p->batch;
    {
      StageTimer st(p, s, CFFT_STAGE_IRECON);
      hipLaunchKernelGGL((k_irecon<T, double>), dim3(nblk(8L * p->n_alive * p->batch, 256)), dim3(256), 0, s,
                         static_cast<const cx<T>*>(B.F), static_cast<cx<T>*>(B.ch), B.btab,
                         static_cast<const cx<double>*>(B.W), ip, 1.0 / p->d);
      launch_check<T>("k_irecon");
    }
    fft_axis<T, +1>(p, static_cast<cx<T>*>(B.ch), 2, s);
    fft_axis<T, +1>(p, static_cast<cx<T>*>(B.ch), 1, s);
    return;
  }
  StageTimer st(p, s, CFFT_STAGE_IRECON);
  const unsigned g = (unsigned)n_planes(p);
#define CFFT_P4(L) \
  case L: hipLaunchKernelGGL((k_p4<T, double, L, L>), dim3(g), dim3(kNT), 0, s, pp, B); break;
  CFFT_LEN_SWITCH(p->H[1], CFFT_P4)
#undef CFFT_P4
  launch_check<T>("k_p4");
}

template <typename T>
void launch_p5(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_UNFOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
  bool f32 = false;  // FP32-arithmetic kernels are instantiated for FP32 data only
  if constexpr (std::is_same<T, float>::value) {
    if (fp32_arith(p)) {
#define CFFT_P5(L) \
  case L: hipLaunchKernelGGL((k_p5<T, float, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
      CFFT_LEN_SWITCH(p->H[0], CFFT_P5)
#undef CFFT_P5
      f32 = true;
    }
  }
  if (!f32) {
#define CFFT_P5(L) \
  case L: hipLaunchKernelGGL((k_p5<T, double, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
    CFFT_LEN_SWITCH(p->H[0], CFFT_P5)
#undef CFFT_P5
  }
  launch_check<T>("k_p5");
}

// Path QR middle of a step: Q at the representatives, orbit expansion and
// IFFT z, IFFT y (one plane kernel on cubic H-grids).
template <typename T>
void launch_qr(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  {
    StageTimer st(p, s, CFFT_STAGE_QMUL);
    hipLaunchKernelGGL(k_qrep<T>, dim3(nblk(p->n_rep * p->batch, 256)), dim3(256), 0, s,
                       static_cast<const cx<T>*>(B.ch), static_cast<cx<T>*>(p->d_grep.p),
                       static_cast<const cx<double>*>(p->d_Q.p), static_cast<const int*>(p->d_gidx.p), p->d,
                       p->n_rep, p->n_alive, p->batch);
    launch_check<T>("k_qrep");
  }
  if (!cubic(p)) {
    {
      StageTimer st(p, s, CFFT_STAGE_IRECON);
      hipLaunchKernelGGL(k_expand<T>, dim3(nblk(p->n_alive * p->batch, 256)), dim3(256), 0, s,
                         static_cast<const cx<T>*>(p->d_grep.p), static_cast<cx<T>*>(B.ch),
                         static_cast<const uint32_t*>(p->d_qmap.p), static_cast<const cx<double>*>(p->d_W.p),
                         p->n_rep, p->n_alive, p->batch);
      launch_check<T>("k_expand");
    }
    fft_axis<T, +1>(p, static_cast<cx<T>*>(B.ch), 2, s);
    fft_axis<T, +1>(p, static_cast<cx<T>*>(B.ch), 1, s);
    return;
  }
  StageTimer st(p, s, CFFT_STAGE_IRECON);
  const unsigned g = (unsigned)n_planes(p);
#define CFFT_P4R(L) \
  case L: hipLaunchKernelGGL((k_p4r<T, double, L, L>), dim3(g), dim3(kNT), 0, s, pp, B); break;
  CFFT_LEN_SWITCH(p->H[1], CFFT_P4R)
#undef CFFT_P4R
  launch_check<T>("k_p4r");
}

// Path QS middle of a step: one kernel per plane pair.
template <typename T>
void launch_p2q(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_PLANE_Q);
  const unsigned g = (unsigned)((long)p->qs_items * p->batch);
  bool f32 = false;  // FP32-arithmetic kernels are instantiated for FP32 data only
  if constexpr (std::is_same<T, float>::value) {
    if (fp32_arith(p)) {
#define CFFT_P2Q(L)                                                                          \
  case L:                                                                                    \
    if (pp.qs_sep) hipLaunchKernelGGL((k_p2q<T, float, L, true>), dim3(g), dim3(kQNT), 0, s, pp, B); \
    else hipLaunchKernelGGL((k_p2q<T, float, L, false>), dim3(g), dim3(kQNT), 0, s, pp, B);          \
    break;
      CFFT_LEN_SWITCH(p->H[1], CFFT_P2Q)
#undef CFFT_P2Q
      f32 = true;
    }
  }
  if (!f32 && (p->lanes || p->pair_threads)) {
    // a tuned schedule (set_option checked that it is instantiated)
    const int nt = p->pair_threads ? p->pair_threads : kQNT;
#define CFFT_P2Q_SCHED(L, NT, LN)                                                                  \
  if (p->H[1] == L && nt == NT && p->lanes == LN) {                                                \
    if (pp.qs_sep) hipLaunchKernelGGL((k_p2q<T, double, L, true, NT, LN>), dim3(g), dim3(NT), 0, s, pp, B); \
    else hipLaunchKernelGGL((k_p2q<T, double, L, false, NT, LN>), dim3(g), dim3(NT), 0, s, pp, B);          \
    launch_check<T>("k_p2q");                                                                      \
    return;                                                                                        \
  }
    CFFT_P2Q_SCHED(32, 256, 2) CFFT_P2Q_SCHED(32, 512, 0) CFFT_P2Q_SCHED(32, 512, 2)
    CFFT_P2Q_SCHED(16, 256, 2) CFFT_P2Q_SCHED(16, 512, 0) CFFT_P2Q_SCHED(16, 512, 2)
    CFFT_P2Q_SCHED(24, 512, 0)
#undef CFFT_P2Q_SCHED
    throw Error{CFFT_ERR_UNSUPPORTED, "no plane-pair kernel for this schedule"};
  }
  if (!f32) {
#define CFFT_P2Q(L)                                                                           \
  case L:                                                                                     \
    if (pp.qs_sep) hipLaunchKernelGGL((k_p2q<T, double, L, true>), dim3(g), dim3(kQNT), 0, s, pp, B); \
    else hipLaunchKernelGGL((k_p2q<T, double, L, false>), dim3(g), dim3(kQNT), 0, s, pp, B);          \
    break;
    CFFT_LEN_SWITCH(p->H[1], CFFT_P2Q)
#undef CFFT_P2Q
  }
  launch_check<T>("k_p2q");
}

template <typename T>
void launch_p51(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_UNFOLD_FOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
  bool f32 = false;  // FP32-arithmetic kernels are instantiated for FP32 data only
  if constexpr (std::is_same<T, float>::value) {
    if (fp32_arith(p)) {
#define CFFT_P51(L)                                                                             \
  case L:                                                                                       \
    if (B.f0_out) hipLaunchKernelGGL((k_p51<T, float, L, true>), dim3(g), dim3(kNT), 0, s, pp, B, step); \
    else hipLaunchKernelGGL((k_p51<T, float, L, false>), dim3(g), dim3(kNT), 0, s, pp, B, step);        \
    break;
      CFFT_LEN_SWITCH(p->H[0], CFFT_P51)
#undef CFFT_P51
      f32 = true;
    }
  }
  if (!f32 && p->lanes) {
    // a tuned lane split (set_option checked that it is instantiated)
#define CFFT_P51_SCHED(L, LN)                                                                              \
  if (p->H[0] == L && p->lanes == LN) {                                                                    \
    if (B.f0_out) hipLaunchKernelGGL((k_p51<T, double, L, true, LN>), dim3(g), dim3(kNT), 0, s, pp, B, step); \
    else hipLaunchKernelGGL((k_p51<T, double, L, false, LN>), dim3(g), dim3(kNT), 0, s, pp, B, step);        \
    launch_check<T>("k_p51");                                                                              \
    return;                                                                                                \
  }
    CFFT_P51_SCHED(32, 2) CFFT_P51_SCHED(16, 2)
#undef CFFT_P51_SCHED
    throw Error{CFFT_ERR_UNSUPPORTED, "no unfold/fold kernel for this lane split"};
  }
  if (!f32) {
#define CFFT_P51(L)                                                                              \
  case L:                                                                                        \
    if (B.f0_out) hipLaunchKernelGGL((k_p51<T, double, L, true>), dim3(g), dim3(kNT), 0, s, pp, B, step); \
    else hipLaunchKernelGGL((k_p51<T, double, L, false>), dim3(g), dim3(kNT), 0, s, pp, B, step);        \
    break;
    CFFT_LEN_SWITCH(p->H[0], CFFT_P51)
#undef CFFT_P51
  }
  launch_check<T>("k_p51");
}

// The axis-split chain (CFFT_OPT_AXIS_SPLIT): H = 8, 12, 16, 24 or 32, FP64
// factor and twiddle products in both precisions.
#define CFFT_AX_SWITCH(len, MACRO) \
  switch (len) {                   \
    MACRO(8)                       \
    MACRO(12)                      \
    MACRO(16)                      \
    MACRO(24)                      \
    MACRO(32)                      \
    default: throw Error{CFFT_ERR_UNSUPPORTED, "the axis-split chain needs H = 8, 12, 16, 24 or 32"}; \
  }

template <typename T>
void launch_p1x(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_FOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
#define CFFT_P1X(L) \
  case L: hipLaunchKernelGGL((k_p1x<T, double, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
  CFFT_AX_SWITCH(p->H[0], CFFT_P1X)
#undef CFFT_P1X
  launch_check<T>("k_p1x");
}

template <typename T>
void launch_p2x(cfft_plan* p, const PipeP& pp, const PipeBufs& B, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_PLANE_Q);
  const unsigned g = (unsigned)n_planes(p);
#define CFFT_P2X(L)                                                                                  \
  case L:                                                                                            \
    hipLaunchKernelGGL((k_p2x<T, double, L>), dim3(g), dim3(L * LineSplit<L>::R1), 0, s, pp, B); \
    break;
  CFFT_AX_SWITCH(p->H[1], CFFT_P2X)
#undef CFFT_P2X
  launch_check<T>("k_p2x");
}

template <typename T>
void launch_p51x(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_UNFOLD_FOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
#define CFFT_P51X(L)                                                                                       \
  case L:                                                                                                  \
    if (B.f0_out) hipLaunchKernelGGL((k_p51x<T, double, L, true>), dim3(g), dim3(kNT), 0, s, pp, B, step); \
    else hipLaunchKernelGGL((k_p51x<T, double, L, false>), dim3(g), dim3(kNT), 0, s, pp, B, step);        \
    break;
  CFFT_AX_SWITCH(p->H[0], CFFT_P51X)
#undef CFFT_P51X
  launch_check<T>("k_p51x");
}

template <typename T>
void launch_p5x(cfft_plan* p, const PipeP& pp, const PipeBufs& B, int step, hipStream_t s) {
  StageTimer st(p, s, CFFT_STAGE_UNFOLD);
  const unsigned g = (unsigned)n_tiles(p, pp);
#define CFFT_P5X(L) \
  case L: hipLaunchKernelGGL((k_p5x<T, double, L>), dim3(g), dim3(kNT), 0, s, pp, B, step); break;
  CFFT_AX_SWITCH(p->H[0], CFFT_P5X)
#undef CFFT_P5X
  launch_check<T>("k_p5x");
}
#undef CFFT_AX_SWITCH

// Work buffers of one diffusion call: channel spectra, the Q-multiplied
// channels (paths M2Q, M7Q) and the spectral ASU values (paths M2, M7).
struct Work {
  void* ch;
  void* ch2;
  void* F;
};

Work own_work(const cfft_plan* p) { return {p->d_ch.p, p->d_ch2.p, p->d_F.p}; }

// f0 -> channel spectra (P1, P2).
template <typename T>
void run_forward_channels(cfft_plan* p, const T* f0, const PipeP& pp, int step, hipStream_t s,
                          void* ch = nullptr) {
  PipeBufs B = pipe_bufs(p, f0, nullptr, nullptr, false, ch);
  launch_p1<T>(p, pp, B, step, s);
  launch_p2<T>(p, pp, B, s);
}

template <typename T>
void run_recon(cfft_plan* p, const PipeP& pp, cx<T>* F, bool withK, hipStream_t s, void* ch = nullptr) {
  PipeBufs B = pipe_bufs(p, nullptr, nullptr, F, withK, ch);
  launch_p3<T>(p, pp, B, s);
}

// F -> f0 (P4, P5).
template <typename T>
void run_backward(cfft_plan* p, const PipeP& pp, const cx<T>* F, T* f0, int step, hipStream_t s,
                  void* ch = nullptr) {
  PipeBufs B = pipe_bufs(p, nullptr, f0, const_cast<cx<T>*>(F), false, ch);
  launch_p4<T>(p, pp, B, s);
  launch_p5<T>(p, pp, B, step, s);
}

// One diffusion operator application per step, with the interaction factors
// of `pp` on both sides: out <- E D (E in).  `in` may equal `out`.  The work
// buffers are the plan's own unless `wk` names others (path QR: the plan's
// own always).
template <typename T>
void run_diffuse(cfft_plan* p, const T* in, T* out_f0, const PipeP& pp, int step, hipStream_t s,
                 const Work* wk = nullptr) {
  const Work w = wk ? *wk : own_work(p);
  if (is_qs(p)) {
    PipeBufs B = pipe_bufs(p, in, out_f0, nullptr, false, w.ch);
    if (p->axis_split) {
      launch_p1x<T>(p, pp, B, step, s);
      launch_p2x<T>(p, pp, B, s);
      launch_p5x<T>(p, pp, B, step, s);
      return;
    }
    launch_p1<T>(p, pp, B, step, s);
    launch_p2q<T>(p, pp, B, s);
    launch_p5<T>(p, pp, B, step, s);
    return;
  }
  if (is_qr(p)) {
    run_forward_channels<T>(p, in, pp, step, s);
    PipeBufs B = pipe_bufs(p, nullptr, out_f0, nullptr, false);
    launch_qr<T>(p, pp, B, s);
    launch_p5<T>(p, pp, B, step, s);
    return;
  }
  if (is_q_path(p)) {
    run_forward_channels<T>(p, in, pp, step, s, w.ch);
    cx<T>* out = static_cast<cx<T>*>(w.ch2);
    {
      StageTimer st(p, s, CFFT_STAGE_QMUL);
      hipLaunchKernelGGL(k_qmul<T>, dim3(nblk(p->n_alive * p->batch, 256)), dim3(256), 0, s,
                         static_cast<const cx<T>*>(w.ch), out, static_cast<const cx<T>*>(p->d_Q.p),
                         static_cast<const int*>(p->d_gidx.p), p->d, p->n_alive, p->batch);
      launch_check<T>("k_qmul");
    }
    PipeBufs B = pipe_bufs(p, nullptr, out_f0, nullptr, false, w.ch);
    if (cubic(p)) {
      StageTimer st(p, s, CFFT_STAGE_IFFT_Z);
      const unsigned g = (unsigned)n_planes(p);
#define CFFT_P4Q(L) \
  case L: hipLaunchKernelGGL((k_p4q<T, double, L, L>), dim3(g), dim3(kNT), 0, s, pp, B, out); break;
      CFFT_LEN_SWITCH(p->H[1], CFFT_P4Q)
#undef CFFT_P4Q
      launch_check<T>("k_p4q");
    } else {
      fft_axis<T, +1>(p, out, 2, s);
      fft_axis<T, +1>(p, out, 1, s);
      B.ch = out;
    }
    launch_p5<T>(p, pp, B, step, s);
    return;
  }
  cx<T>* F = static_cast<cx<T>*>(w.F);
  run_forward_channels<T>(p, in, pp, step, s, w.ch);
  run_recon<T>(p, pp, F, true, s, w.ch);
  run_backward<T>(p, pp, F, out_f0, step, s, w.ch);
}

// Persistent cooperative sweep of nsteps Strang steps (cubic H-grids, fwd+K+bwd
// paths); returns false when it cannot be used so the caller falls back to
// one launch per phase.
template <typename T>
bool run_sweep_persistent(cfft_plan* p, T* q, const PipeP& pp, int step0, int nsteps, hipStream_t s) {
  if (!p->persistent || !cubic(p) || is_q_path(p)) return false;
  const void* kern = nullptr;
#define CFFT_SW(L) \
  case L: kern = reinterpret_cast<const void*>(&k_sweep<T, double, L>); break;
  switch (p->H[0]) {
    CFFT_SW(2) CFFT_SW(3) CFFT_SW(4) CFFT_SW(6) CFFT_SW(8) CFFT_SW(12) CFFT_SW(16) CFFT_SW(24)
    CFFT_SW(32) CFFT_SW(48) CFFT_SW(64)
    default: return false;
  }
#undef CFFT_SW
  int dev = 0, ncu = 0, per_cu = 0, coop = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev), "attr");
  if (!coop) return false;
  hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "attr");
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kNT, 0), "occupancy");
  if (per_cu < 1) return false;
  long want = std::max(n_tiles(p, pp), n_planes(p));
  want = std::max(want, (8L * p->n_spec * p->batch + kNT - 1) / kNT);
  const long grid = std::min<long>((long)per_cu * ncu, want);
  PipeBufs B = pipe_bufs(p, q, q, p->d_F.p, true);
  PipeP ppc = pp;
  int s0 = step0, ns = nsteps;
  void* args[] = {&ppc, &B, &s0, &ns};
  StageTimer st(p, s, CFFT_STAGE_SWEEP);
  hip_check(hipLaunchCooperativeKernel(kern, dim3((unsigned)grid), dim3(kNT), args, 0, s),
            "hipLaunchCooperativeKernel");
  return true;
}

// Projection of subgrid fields onto those invariant under the ops that keep
// the subgrid (ops: n_sym x 12 int32, R' row-major then t'; the identity is
// one of them): out(x) = mean_g in(R' x + t' mod M).  Fields of M0 M1 M2
// entries (x fastest) one after another; n = total entries.
__global__ void k_symmetrize(const double* __restrict__ in, double* __restrict__ out,
                             const int32_t* __restrict__ ops, int n_sym, int M0, int M1, int M2, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long M3 = (long)M0 * M1 * M2;
  const long b = i / M3;
  const long r = i - b * M3;
  const int x = (int)(r % M0), y = (int)((r / M0) % M1), z = (int)(r / ((long)M0 * M1));
  const double* f = in + b * M3;
  double s = 0.0;
  for (int g = 0; g < n_sym; ++g) {
    const int32_t* o = ops + 12 * g;
    int xp = (o[0] * x + o[1] * y + o[2] * z + o[9]) % M0;
    int yp = (o[3] * x + o[4] * y + o[5] * z + o[10]) % M1;
    int zp = (o[6] * x + o[7] * y + o[8] * z + o[11]) % M2;
    xp += xp < 0 ? M0 : 0;
    yp += yp < 0 ? M1 : 0;
    zp += zp < 0 ? M2 : 0;
    s += f[xp + (long)M0 * (yp + (long)M1 * zp)];
  }
  out[i] = s / n_sym;
}

// The stream a *_dev call runs on: exactly the caller's, with NULL the null
// stream (the default stream of HIP and of PyTorch), so the call is ordered
// after the caller's earlier work on that stream and before its later work.
hipStream_t pick(cfft_plan*, void* s) { return static_cast<hipStream_t>(s); }

template <typename T>
void to_dev(cfft_plan* p, DevBuf& dst, const double* src, size_t n) {
  std::vector<T> h(src, src + n);
  hip_check(hipMemcpyAsync(dst.p, h.data(), n * sizeof(T), hipMemcpyHostToDevice, p->stream), "H2D");
  hip_check(hipStreamSynchronize(p->stream), "sync");
}

template <typename T>
void from_dev(cfft_plan* p, const DevBuf& src, double* dst, size_t n) {
  std::vector<T> h(n);
  hip_check(hipMemcpyAsync(h.data(), src.p, n * sizeof(T), hipMemcpyDeviceToHost, p->stream), "D2H");
  hip_check(hipStreamSynchronize(p->stream), "sync");
  for (size_t i = 0; i < n; ++i) dst[i] = (double)h[i];
}

void check_plan(const cfft_plan* p) { require(p != nullptr, CFFT_ERR_INVALID, "null plan"); }

template <typename T>
void forward_host(cfft_plan* p, const double* f0, double* F) {
  to_dev<T>(p, p->d_f0, f0, p->M3 * p->batch);
  PipeP pp = pipe_params(p, ESel{});
  run_forward_channels<T>(p, static_cast<const T*>(p->d_f0.p), pp, 0, p->stream);
  run_recon<T>(p, pp, static_cast<cx<T>*>(p->d_F.p), false, p->stream);
  from_dev<T>(p, p->d_F, F, 2 * (size_t)p->n_spec * p->batch);
}

template <typename T>
void backward_host(cfft_plan* p, const double* F, double* f0) {
  to_dev<T>(p, p->d_F, F, 2 * (size_t)p->n_spec * p->batch);
  PipeP pp = pipe_params(p, ESel{});
  run_backward<T>(p, pp, static_cast<const cx<T>*>(p->d_F.p), static_cast<T*>(p->d_f0.p), 0, p->stream);
  from_dev<T>(p, p->d_f0, f0, p->M3 * p->batch);
}

template <typename T>
void g0_host(cfft_plan* p, const double* f0, double* G0) {
  to_dev<T>(p, p->d_f0, f0, p->M3 * p->batch);
  PipeP pp = pipe_params(p, ESel{});
  run_forward_channels<T>(p, static_cast<const T*>(p->d_f0.p), pp, 0, p->stream);
  FoldP fp = fold_params<T>(p, nullptr, false);
  DevBuf g;
  g.alloc(2 * sizeof(T) * p->M3 * p->batch);
  hipLaunchKernelGGL(k_assemble<T>, dim3(nblk(p->M3 * p->batch, 256)), dim3(256), 0, p->stream,
                     static_cast<const cx<T>*>(p->d_ch.p), static_cast<cx<T>*>(g.p), fp,
                     static_cast<const int*>(p->d_chmap.p));
  launch_check<T>("k_assemble");
  from_dev<T>(p, g, G0, 2 * (size_t)p->M3 * p->batch);
}

template <typename T>
void diffuse_host(cfft_plan* p, double* f0) {
  to_dev<T>(p, p->d_f0, f0, p->M3 * p->batch);
  PipeP pp = pipe_params(p, ESel{});
  run_diffuse<T>(p, static_cast<T*>(p->d_f0.p), static_cast<T*>(p->d_f0.p), pp, 0, p->stream);
  from_dev<T>(p, p->d_f0, f0, p->M3 * p->batch);
}

template <typename T>
void exp_dev(cfft_plan* p, const T* w, double* E, double factor, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_exp<T>, dim3(nblk(n, 256)), dim3(256), 0, s, w, E, factor, n);
  launch_check<T>("k_exp");
}

// nsteps Strang steps of the fwd + K + bwd paths as one launch chain: P1, P2
// for the first step, then per step P3, P4 and either P5 fused with the next
// step's P1 (followed by its P2) or, after the last step, P5.  Step i reads
// `q + i * stride` and writes `q + (i + 1) * stride` (stride 0: in place).
// Buffers of a fused unfold/fold between two steps: without a history
// (stride 0) the intermediate q has no reader and is not written.
PipeBufs inner_bufs(PipeBufs B, long stride) {
  if (stride == 0) B.f0_out = nullptr;
  return B;
}

template <typename T>
void run_chain(cfft_plan* p, T* q, long stride, const PipeP& pp, int step0, int nsteps, hipStream_t s) {
  cx<T>* F = static_cast<cx<T>*>(p->d_F.p);
  if (is_qs(p) && p->axis_split) {
    // per step: the plane kernel (A_y A_z), then the unfold of the step fused
    // with the fold and A_x of the next
    launch_p1x<T>(p, pp, pipe_bufs(p, q, nullptr, nullptr, false), step0, s);
    for (int i = 0; i < nsteps; ++i) {
      PipeBufs B = pipe_bufs(p, q + (i + 1) * stride, q + (i + 1) * stride, F, false);
      launch_p2x<T>(p, pp, B, s);
      if (i + 1 < nsteps && p->fuse) {
        launch_p51x<T>(p, pp, inner_bufs(B, stride), step0 + i, s);
      } else {
        launch_p5x<T>(p, pp, B, step0 + i, s);
        if (i + 1 < nsteps) launch_p1x<T>(p, pp, B, step0 + i + 1, s);
      }
    }
    return;
  }
  if (is_qs(p)) {
    // per step: the plane-pair kernel, then P5 of the step fused with P1 of the next
    launch_p1<T>(p, pp, pipe_bufs(p, q, nullptr, nullptr, false), step0, s);
    for (int i = 0; i < nsteps; ++i) {
      PipeBufs B = pipe_bufs(p, q + (i + 1) * stride, q + (i + 1) * stride, F, false);
      launch_p2q<T>(p, pp, B, s);
      if (i + 1 < nsteps && p->fuse) {
        launch_p51<T>(p, pp, inner_bufs(B, stride), step0 + i, s);
      } else {
        launch_p5<T>(p, pp, B, step0 + i, s);
        if (i + 1 < nsteps) launch_p1<T>(p, pp, B, step0 + i + 1, s);
      }
    }
    return;
  }
  run_forward_channels<T>(p, q, pp, step0, s);
  for (int i = 0; i < nsteps; ++i) {
    PipeBufs B = pipe_bufs(p, nullptr, q + (i + 1) * stride, F, false);
    if (is_qr(p)) {
      launch_qr<T>(p, pp, B, s);
    } else {
      run_recon<T>(p, pp, F, true, s);
      launch_p4<T>(p, pp, B, s);
    }
    if (i + 1 < nsteps && p->fuse) {
      launch_p51<T>(p, pp, inner_bufs(B, stride), step0 + i, s);
      launch_p2<T>(p, pp, B, s);
    } else {
      launch_p5<T>(p, pp, B, step0 + i, s);
      if (i + 1 < nsteps) run_forward_channels<T>(p, q + (i + 1) * stride, pp, step0 + i + 1, s);
    }
  }
}

// Launch sequence of nsteps steps (fused chain, or one diffusion per step on
// the Q paths).
template <typename T>
void launch_steps(cfft_plan* p, T* q, long stride, const PipeP& pp, int step0, int nsteps, hipStream_t s) {
  if (chain_path(p)) {
    run_chain<T>(p, q, stride, pp, step0, nsteps, s);
    return;
  }
  for (int i = 0; i < nsteps; ++i) run_diffuse<T>(p, q + i * stride, q + (i + 1) * stride, pp, step0 + i, s);
}

constexpr size_t kMaxGraphs = 16;

// The launch sequence of (q, stride, factors, step0, nsteps) is captured once
// into a hipGraph on the plan's capture stream and replayed on `s`.  Kernel
// arguments are baked into the graph, so the cache key holds every pointer
// and scalar the sequence depends on.  The chain contains kernel nodes only.
template <typename T>
void run_steps_graph(cfft_plan* p, T* q, long stride, const PipeP& pp, int step0, int nsteps, hipStream_t s) {
  std::string key(reinterpret_cast<const char*>(&pp), sizeof(PipeP));
  const long hdr[9] = {reinterpret_cast<long>(q), stride, step0, nsteps, p->fuse, p->arith32, p->lanes,
                       p->pair_threads, p->axis_split};
  key.append(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  for (size_t i = 0; i < p->graphs.size(); ++i)
    if (p->graphs[i].key == key) {
      cfft_plan::Cached c = p->graphs[i];
      p->graphs.erase(p->graphs.begin() + i);
      p->graphs.push_back(c);
      hip_check(hipGraphLaunch(c.exec, s), "hipGraphLaunch");
      return;
    }
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  if (!p->cap_stream) hip_check(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  hipGraph_t g = nullptr;
  hip_check(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    launch_steps<T>(p, q, stride, pp, step0, nsteps, p->cap_stream);
  } catch (...) {
    (void)hipStreamEndCapture(p->cap_stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(p->cap_stream, &g), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  hip_check(e, "hipGraphInstantiate");
  if (p->graphs.size() >= kMaxGraphs) {
    (void)hipGraphExecDestroy(p->graphs.front().exec);
    p->graphs.erase(p->graphs.begin());
  }
  p->graphs.push_back({key, exec});
  hip_check(hipGraphLaunch(exec, s), "hipGraphLaunch");
}

template <typename T>
void dispatch_steps(cfft_plan* p, T* q, long stride, const PipeP& pp, int step0, int nsteps, hipStream_t s) {
  if (p->graph && !p->prof && nsteps > 1)
    run_steps_graph<T>(p, q, stride, pp, step0, nsteps, s);
  else
    launch_steps<T>(p, q, stride, pp, step0, nsteps, s);
}

// nsteps Strang steps with the interaction factors of `es`.
template <typename T>
void run_steps(cfft_plan* p, T* q, const ESel& es, int step0, int nsteps, hipStream_t s) {
  PipeP pp = pipe_params(p, es);
  if (nsteps <= 0) return;
  if (run_sweep_persistent<T>(p, q, pp, step0, nsteps, s)) return;
  dispatch_steps<T>(p, q, 0, pp, step0, nsteps, s);
}

// nsteps Strang steps keeping every contour point: step i reads slice i of
// `hist` and writes slice i + 1 (slices of batch * M3 entries).
template <typename T>
void run_steps_hist(cfft_plan* p, T* hist, const ESel& es, int step0, int nsteps, hipStream_t s) {
  PipeP pp = pipe_params(p, es);
  if (nsteps <= 0) return;
  dispatch_steps<T>(p, hist, p->M3 * p->batch, pp, step0, nsteps, s);
}

// Average duration of one launch of a chain kernel, measured live: one
// untimed launch, then `reps` identical launches back to back on `s` between
// two hipEvents (the same launches a kernel trace lists one by one).  The
// unfold/fold kernel runs as it does between two steps of a plain sweep
// (q(s + 1) kept in registers).
template <typename T>
double time_stage(cfft_plan* p, T* q, const ESel& es, int stage, int reps, hipStream_t s) {
  PipeP pp = pipe_params(p, es);
  PipeBufs B = pipe_bufs(p, nullptr, q, p->d_F.p, false);
  const bool ax = is_qs(p) && p->axis_split;
  auto launch = [&] {
    if (stage == CFFT_STAGE_UNFOLD_FOLD) {
      if (ax) launch_p51x<T>(p, pp, inner_bufs(B, 0), 0, s);
      else launch_p51<T>(p, pp, inner_bufs(B, 0), 0, s);  // as between two steps of a sweep
    } else if (stage == CFFT_STAGE_PLANE_Q) {
      if (ax) launch_p2x<T>(p, pp, B, s);
      else launch_p2q<T>(p, pp, B, s);
    } else {
      launch_p2<T>(p, pp, B, s);
    }
  };
  launch();
  hipEvent_t a = nullptr, b = nullptr;
  hip_check(hipEventCreate(&a), "hipEventCreate");
  hipError_t e = hipEventCreate(&b);
  if (e != hipSuccess) {
    (void)hipEventDestroy(a);
    hip_check(e, "hipEventCreate");
  }
  float ms = 0.f;
  e = hipEventRecord(a, s);
  if (e == hipSuccess) {
    try {
      for (int i = 0; i < reps; ++i) launch();
    } catch (...) {
      (void)hipStreamSynchronize(s);
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      throw;
    }
    e = hipEventRecord(b, s);
  }
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  hip_check(e, "stage timing");
  return (double)ms / reps;
}

template <typename T>
void step_host(cfft_plan* p, double* q, const double* w, double dt, int nsteps) {
  to_dev<T>(p, p->d_f0, q, p->M3 * p->batch);
  to_dev<double>(p, p->d_E, w, p->M3 * p->batch);  // w in FP64, E in place
  exp_dev<double>(p, static_cast<const double*>(p->d_E.p), static_cast<double*>(p->d_E.p), -dt / 2.0,
                  p->M3 * p->batch, p->stream);
  ESel es;
  es.mode = 1;
  es.contig = p->d_E.p;
  run_steps<T>(p, static_cast<T*>(p->d_f0.p), es, 0, nsteps, p->stream);
  from_dev<T>(p, p->d_f0, q, p->M3 * p->batch);
}

}  // namespace

// ---- what a multi-plan sweep needs of a plan ---------------------------------
namespace cfft {

EnsPlanInfo ens_plan_info(cfft_plan* p) {
  check_plan(p);
  EnsPlanInfo i{};
  i.path = p->path;
  i.prec = p->prec;
  i.arith32 = fp32_arith(p) ? 1 : 0;
  i.device = p->device;
  i.L = p->H[0];
  i.cubic = cubic(p) ? 1 : 0;
  i.batch = p->batch;
  i.sep = p->qs_sep ? 1 : 0;
  i.lanes = p->lanes;
  i.axis_split = p->axis_split;
  i.pair_items = (long)p->qs_items * p->batch;
  i.tile_items = n_tiles(p, pipe_params(p, ESel{}));
  i.M3 = p->M3;
  return i;
}

// The same launch parameters and buffers as the plan's own history sweep
// (cfft_sweep_hist_dev), so an ensemble computes what the plan computes.
void ens_slot_params(cfft_plan* p, const EnsPart& part, EnsSlot* out) {
  check_plan(p);
  require((int)part.E1.size() == p->batch && (int)part.E2.size() == p->batch && (int)part.sw.size() == p->batch,
          CFFT_ERR_INVALID, "one pair of interaction factors and one switch step per field");
  ESel es;
  es.mode = 2;
  es.E1 = part.E1.data();
  es.E2 = part.E2.data();
  es.sw = part.sw.data();
  std::memset(static_cast<void*>(out), 0, sizeof(EnsSlot));
  out->p = pipe_params(p, es);
  out->B = pipe_bufs(p, nullptr, nullptr, p->d_F.p, false);
  out->q = part.hist;
  out->stride = p->M3 * p->batch;
  out->L = p->H[0];
}

}  // namespace cfft

// ============================================================================
// extern "C" entry points
// ============================================================================

extern "C" {

int cfft_last_error(char* buf, size_t n) {
  if (buf && n) {
    size_t m = std::min(n - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), m);
    buf[m] = 0;
  }
  return (int)g_last_error.size();
}

static void check_dim(int32_t dim) { require(dim == 2 || dim == 3, CFFT_ERR_INVALID, "dim must be 2 or 3"); }

int cfft_find_optimal_shift(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                            double* shift_out, int32_t* ops_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && shift_out && ops_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = N[d];
    std::vector<Op> sh;
    double s[3] = {0, 0, 0};
    find_optimal_shift(dim, n, unpack_ops(ops, nops), s, sh);
    for (int d = 0; d < dim; ++d) shift_out[d] = s[d];
    pack_ops(sh, ops_out);
  });
}

int cfft_count_special_positions(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                                 int64_t* count_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && count_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = N[d];
    *count_out = count_special_positions(dim, n, unpack_ops(ops, nops));
  });
}

int cfft_calc_spectral_asu(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                           int32_t* idx_out, int32_t* mult_out, int64_t cap, int64_t* n_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && n_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = N[d];
    SpecASU s = calc_spectral_asu(dim, n, unpack_ops(ops, nops));
    *n_out = (int64_t)s.idx.size();
    int64_t m = std::min<int64_t>(cap, (int64_t)s.idx.size());
    for (int64_t i = 0; i < m; ++i) {
      if (idx_out)
        for (int d = 0; d < 3; ++d) idx_out[3 * i + d] = s.idx[i][d];
      if (mult_out) mult_out[i] = s.mult[i];
    }
  });
}

int cfft_auto_L(int32_t dim, const int32_t* ops, int32_t nops, int32_t* L_out) {
  return guarded([&] {
    check_dim(dim);
    require(ops && L_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int L[3];
    auto_L(dim, unpack_ops(ops, nops), L);
    for (int d = 0; d < dim; ++d) L_out[d] = L[d];
  });
}

int cfft_select_rep_ops(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                        const int32_t* L, int32_t* rep_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && L && rep_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1}, l[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
      n[d] = N[d];
      l[d] = L[d];
    }
    std::vector<int> r = select_rep_ops(dim, unpack_ops(ops, nops), l, n);
    for (size_t i = 0; i < r.size(); ++i) rep_out[i] = r[i];
  });
}

int cfft_detect_centering(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                          int32_t* cent_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && cent_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) n[d] = N[d];
    *cent_out = detect_centering(dim, n, unpack_ops(ops, nops));
  });
}

int cfft_fill_map(int32_t dim, const int32_t* N, const int32_t* ops, int32_t nops,
                  const int32_t* L, int32_t* map_out) {
  return guarded([&] {
    check_dim(dim);
    require(N && ops && L && map_out && nops > 0, CFFT_ERR_INVALID, "null argument");
    int n[3] = {1, 1, 1}, l[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) {
      n[d] = N[d];
      l[d] = L[d];
      require(n[d] % l[d] == 0, CFFT_ERR_NOT_DIVISIBLE, "N not divisible by L");
    }
    std::vector<int32_t> m = fill_map(dim, n, unpack_ops(ops, nops), l);
    std::memcpy(map_out, m.data(), m.size() * 4);
  });
}

int cfft_plan_create(const int32_t* N, const int32_t* ops, int32_t nops, double ds,
                     const double* lattice, int32_t path, int32_t precision, int32_t batch,
                     cfft_plan** out) {
  return guarded([&] {
    require(N && ops && lattice && out && nops > 0, CFFT_ERR_INVALID, "null argument");
    *out = nullptr;
    // host table building runs unlocked; allocations and uploads take the
    // set-up lock one at a time
    std::unique_ptr<cfft_plan> p(new cfft_plan());
    int n[3] = {N[0], N[1], N[2]};
    build_plan(p.get(), n, unpack_ops(ops, nops), ds, lattice, path, precision, batch);
    hip_check(hipGetDevice(&p->device), "hipGetDevice");  // the device build_plan allocated on
    *out = p.release();
  });
}

int cfft_plan_info(const cfft_plan* p, int32_t* info) {
  return guarded([&] {
    check_plan(p);
    require(info != nullptr, CFFT_ERR_INVALID, "null argument");
    int32_t v[18] = {p->M[0], p->M[1], p->M[2], p->L[0], p->L[1], p->L[2], p->n_spec, p->n_ch,
                     p->cent, p->d,    p->path, p->prec, p->batch, p->H[0], p->H[1], p->H[2],
                     (int32_t)p->n_rep, (int32_t)p->n_alive};
    std::memcpy(info, v, sizeof(v));
  });
}

int cfft_plan_tables(const cfft_plan* p, int32_t* idx_out, int32_t* mult_out, double* K_out,
                     int32_t* rep_out) {
  return guarded([&] {
    check_plan(p);
    for (int i = 0; i < p->n_spec; ++i) {
      if (idx_out)
        for (int d = 0; d < 3; ++d) idx_out[3 * i + d] = p->spec.idx[i][d];
      if (mult_out) mult_out[i] = p->spec.mult[i];
      if (K_out) K_out[i] = p->K[i];
    }
    if (rep_out)
      for (int a = 0; a < p->d; ++a) rep_out[a] = p->rep[a];
  });
}

int cfft_update_kernel(cfft_plan* p, double ds, const double* lattice) {
  return guarded([&] {
    check_plan(p);
    require(!is_q_path(p) || is_qs(p), CFFT_ERR_INVALID,
            "update_kernel applies to the fwd+K+bwd paths and path QS; replan the other Q-fused paths");
    const double* lat = lattice ? lattice : p->lattice;
    // every check before any plan state changes: a rejected update leaves the
    // plan as it was
    if (is_qs(p))
      require(orthogonal_cell(lat), CFFT_ERR_UNSUPPORTED, "path QS needs an orthogonal cell");
    std::vector<double> K = kernel_spec(p->spec, ds, lat);
    require(p->d_K.bytes == K.size() * sizeof(double), CFFT_ERR_INVALID, "kernel table size changed");
    // work of this plan may be in flight on any stream; the new values go into
    // the same allocation, so captured chains keep pointing at valid memory
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    p->K = std::move(K);
    p->ds = ds;
    if (lattice) std::memcpy(p->lattice, lattice, sizeof(double) * 9);
    if (!p->K.empty())
      hip_check(hipMemcpy(p->d_K.p, p->K.data(), p->d_K.bytes, hipMemcpyHostToDevice), "upload K");
    if (is_qs(p)) build_qs_table(p);
  });
}

int cfft_plan_destroy(cfft_plan* p) {
  return guarded([&] {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> guard(setup_mutex());
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    delete p;
  });
}

#define CFFT_DISPATCH(p, FN, ...) \
  ((p)->prec == CFFT_FP64 ? FN<double>(__VA_ARGS__) : FN<float>(__VA_ARGS__))

int cfft_forward(cfft_plan* p, const double* f0, double* F) {
  return guarded([&] {
    check_plan(p);
    require(f0 && F, CFFT_ERR_INVALID, "null argument");
    CFFT_DISPATCH(p, forward_host, p, f0, F);
  });
}

int cfft_backward(cfft_plan* p, const double* F, double* f0) {
  return guarded([&] {
    check_plan(p);
    require(f0 && F, CFFT_ERR_INVALID, "null argument");
    require(!is_q_path(p), CFFT_ERR_INVALID, "backward transform needs a fwd+K+bwd path plan");
    CFFT_DISPATCH(p, backward_host, p, F, f0);
  });
}

int cfft_forward_g0(cfft_plan* p, const double* f0, double* G0) {
  return guarded([&] {
    check_plan(p);
    require(f0 && G0, CFFT_ERR_INVALID, "null argument");
    CFFT_DISPATCH(p, g0_host, p, f0, G0);
  });
}

int cfft_diffuse(cfft_plan* p, double* f0) {
  return guarded([&] {
    check_plan(p);
    require(f0 != nullptr, CFFT_ERR_INVALID, "null argument");
    CFFT_DISPATCH(p, diffuse_host, p, f0);
  });
}

int cfft_step(cfft_plan* p, double* q, const double* w, double dt, int32_t nsteps) {
  return guarded([&] {
    check_plan(p);
    require(q && w && nsteps >= 0, CFFT_ERR_INVALID, "bad argument");
    CFFT_DISPATCH(p, step_host, p, q, w, dt, nsteps);
  });
}

int cfft_sub_to_full(cfft_plan* p, const double* f0, double* u) {
  return guarded([&] {
    check_plan(p);
    require(f0 && u, CFFT_ERR_INVALID, "null argument");
    if (p->fmap_host.empty()) {
      p->fmap_host = fill_map(3, p->N, p->ops, p->L);
      upload_bytes(p->d_fmap, p->fmap_host.data(), p->fmap_host.size() * 4);
    }
    if (p->d_u.bytes < p->esz * p->N3) p->d_u.alloc(p->esz * p->N3);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      to_dev<T>(p, p->d_f0, f0, p->M3);
      hipLaunchKernelGGL(k_sub_to_full<T>, dim3(nblk(p->N3, 256)), dim3(256), 0, p->stream,
                         static_cast<const T*>(p->d_f0.p), static_cast<T*>(p->d_u.p),
                         static_cast<const int*>(p->d_fmap.p), p->N3);
      launch_check<T>("k_sub_to_full");
      from_dev<T>(p, p->d_u, u, p->N3);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_full_to_sub(cfft_plan* p, const double* u, double* f0) {
  return guarded([&] {
    check_plan(p);
    require(f0 && u, CFFT_ERR_INVALID, "null argument");
    if (p->d_u.bytes < p->esz * p->N3) p->d_u.alloc(p->esz * p->N3);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      to_dev<T>(p, p->d_u, u, p->N3);
      hipLaunchKernelGGL(k_full_to_sub<T>, dim3(nblk(p->M3, 256)), dim3(256), 0, p->stream,
                         static_cast<const T*>(p->d_u.p), static_cast<T*>(p->d_f0.p), p->M[0],
                         p->M[1], p->M[2], p->L[0], p->L[1], p->L[2], p->N[0], p->N[1]);
      launch_check<T>("k_full_to_sub");
      from_dev<T>(p, p->d_f0, f0, p->M3);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_forward_dev(cfft_plan* p, const void* f0, void* F, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(f0 && F, CFFT_ERR_INVALID, "null argument");
    hipStream_t s = pick(p, stream);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      PipeP pp = pipe_params(p, ESel{});
      run_forward_channels<T>(p, static_cast<const T*>(f0), pp, 0, s);
      run_recon<T>(p, pp, static_cast<cx<T>*>(F), false, s);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_backward_dev(cfft_plan* p, const void* F, void* f0, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(f0 && F, CFFT_ERR_INVALID, "null argument");
    require(!is_q_path(p), CFFT_ERR_INVALID, "backward transform needs a fwd+K+bwd path plan");
    hipStream_t s = pick(p, stream);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      PipeP pp = pipe_params(p, ESel{});
      run_backward<T>(p, pp, static_cast<const cx<T>*>(F), static_cast<T*>(f0), 0, s);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_diffuse_dev(cfft_plan* p, void* f0, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(f0 != nullptr, CFFT_ERR_INVALID, "null argument");
    hipStream_t s = pick(p, stream);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      PipeP pp = pipe_params(p, ESel{});
      run_diffuse<T>(p, static_cast<T*>(f0), static_cast<T*>(f0), pp, 0, s);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_symmetrize_f64_dev(cfft_plan* p, const void* in, void* out, int64_t n_fields, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(in && out && n_fields >= 0, CFFT_ERR_INVALID, "bad argument");
    require(in != out, CFFT_ERR_INVALID, "symmetrize reads and writes different arrays");
    const long n = p->M3 * n_fields;
    if (n == 0) return;
    hipLaunchKernelGGL(k_symmetrize, dim3(nblk(n, 256)), dim3(256), 0, pick(p, stream),
                       static_cast<const double*>(in), static_cast<double*>(out),
                       static_cast<const int32_t*>(p->d_symops.p), p->n_sym, p->M[0], p->M[1], p->M[2], n);
    launch_check<double>("k_symmetrize");
  });
}

// ---- diffusion without allocation ------------------------------------------
// The work buffers of cfft_diffuse_noalloc_dev live in one statically sized
// device array, loaded with the code object on every device that runs the
// library and held there until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress, taken once
// per device with that device current, and only that address is used.
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::atomic<void*> g_scratch_ptr[kScratchDevices];

void* scratch_ptr(int dev) {
  require(dev >= 0 && dev < kScratchDevices, CFFT_ERR_UNSUPPORTED, "device index beyond 63");
  void* ptr = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (ptr) return ptr;
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  ptr = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (!ptr) {
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    require(ptr != nullptr, CFFT_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev].store(ptr, std::memory_order_release);
  }
  return ptr;
}

size_t round256(size_t n) { return (n + 255) & ~size_t(255); }
}  // namespace

int cfft_diffuse_noalloc_dev(cfft_plan* p, void* f0, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(f0 != nullptr, CFFT_ERR_INVALID, "null argument");
    require(!is_qr(p), CFFT_ERR_UNSUPPORTED,
            "no-allocation diffusion supports paths M2, M7, M2Q, M7Q and QS (path QR keeps plan buffers)");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    require(dev == p->device, CFFT_ERR_INVALID, "the plan's device is not the current device");
    // the work buffers the path reads and writes, each 256-byte aligned
    const size_t ch = round256(2 * p->esz * (size_t)p->n_alive * p->batch);
    const size_t ch2 = (is_q_path(p) && !is_qs(p)) ? ch : 0;
    const size_t F = is_q_path(p) ? 0 : round256(2 * p->esz * (size_t)std::max(p->n_spec, 1) * p->batch);
    const size_t need = ch + ch2 + F;
    if (need > kScratchBytes)
      throw Error{CFFT_ERR_SIZE, "no-allocation diffusion needs " + std::to_string(need) +
                                     " bytes of work memory; the maximum is " + std::to_string(kScratchBytes) +
                                     " (16 MiB): use cfft_diffuse_dev"};
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(dev));
    // every work buffer is written in full before it is read, so the
    // previous call's contents need no clearing
    const Work w{base, ch2 ? base + ch : nullptr, F ? base + ch + ch2 : nullptr};
    hipStream_t s = pick(p, stream);
    auto run = [&](auto tag) {
      using T = decltype(tag);
      PipeP pp = pipe_params(p, ESel{});
      run_diffuse<T>(p, static_cast<T*>(f0), static_cast<T*>(f0), pp, 0, s, &w);
    };
    if (p->prec == CFFT_FP64) run(double{}); else run(float{});
  });
}

int cfft_step_dev(cfft_plan* p, void* q, const void* E, int32_t nsteps, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(q && E && nsteps >= 0, CFFT_ERR_INVALID, "bad argument");
    hipStream_t s = pick(p, stream);
    ESel es;
    es.mode = 1;
    es.contig = E;
    if (p->prec == CFFT_FP64)
      run_steps<double>(p, static_cast<double*>(q), es, 0, nsteps, s);
    else
      run_steps<float>(p, static_cast<float*>(q), es, 0, nsteps, s);
  });
}

int cfft_step_ptrs_dev(cfft_plan* p, void* q, const void* const* E, int32_t nsteps,
                       void* stream) {
  return guarded([&] {
    check_plan(p);
    require(q && E && nsteps >= 0, CFFT_ERR_INVALID, "bad argument");
    hipStream_t s = pick(p, stream);
    ESel es;
    es.mode = 2;
    es.E1 = E;
    if (p->prec == CFFT_FP64)
      run_steps<double>(p, static_cast<double*>(q), es, 0, nsteps, s);
    else
      run_steps<float>(p, static_cast<float*>(q), es, 0, nsteps, s);
  });
}

int cfft_sweep_dev(cfft_plan* p, void* q, const void* const* E1, const void* const* E2,
                   const int32_t* sw, int32_t step0, int32_t nsteps, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(q && E1 && E2 && sw && nsteps >= 0, CFFT_ERR_INVALID, "bad argument");
    hipStream_t s = pick(p, stream);
    ESel es;
    es.mode = 2;
    es.E1 = E1;
    es.E2 = E2;
    es.sw = sw;
    if (p->prec == CFFT_FP64)
      run_steps<double>(p, static_cast<double*>(q), es, step0, nsteps, s);
    else
      run_steps<float>(p, static_cast<float*>(q), es, step0, nsteps, s);
  });
}

int cfft_time_stage_dev(cfft_plan* p, int32_t stage, void* q, const void* const* E, int32_t reps,
                        void* stream, double* ms_per_launch) {
  return guarded([&] {
    check_plan(p);
    require(q && E && ms_per_launch && reps >= 1, CFFT_ERR_INVALID, "bad argument");
    require(stage == CFFT_STAGE_UNFOLD_FOLD || stage == CFFT_STAGE_FFT_Y ||
                (stage == CFFT_STAGE_PLANE_Q && is_qs(p)),
            CFFT_ERR_INVALID, "stage must be CFFT_STAGE_UNFOLD_FOLD, CFFT_STAGE_FFT_Y or (path QS) CFFT_STAGE_PLANE_Q");
    require(!p->prof, CFFT_ERR_INVALID, "per-launch event profiling is enabled on this plan");
    require(chain_path(p), CFFT_ERR_UNSUPPORTED, "plan path has no fused chain kernels");
    require(cubic(p), CFFT_ERR_UNSUPPORTED, "stage timing needs a cubic grid");
    hipStream_t s = pick(p, stream);
    ESel es;
    es.mode = 2;
    es.E1 = E;
    if (p->prec == CFFT_FP64)
      *ms_per_launch = time_stage<double>(p, static_cast<double*>(q), es, stage, reps, s);
    else
      *ms_per_launch = time_stage<float>(p, static_cast<float*>(q), es, stage, reps, s);
  });
}

int cfft_sweep_hist_dev(cfft_plan* p, void* hist, const void* const* E1, const void* const* E2,
                        const int32_t* sw, int32_t step0, int32_t nsteps, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(hist && E1 && E2 && sw && nsteps >= 0, CFFT_ERR_INVALID, "bad argument");
    hipStream_t s = pick(p, stream);
    ESel es;
    es.mode = 2;
    es.E1 = E1;
    es.E2 = E2;
    es.sw = sw;
    if (p->prec == CFFT_FP64)
      run_steps_hist<double>(p, static_cast<double*>(hist), es, step0, nsteps, s);
    else
      run_steps_hist<float>(p, static_cast<float*>(hist), es, step0, nsteps, s);
  });
}

int cfft_plan_set_option(cfft_plan* p, int32_t key, int64_t value) {
  return guarded([&] {
    check_plan(p);
    switch (key) {
      case CFFT_OPT_PERSISTENT: p->persistent = value != 0; break;
      case CFFT_OPT_FUSE: p->fuse = value != 0; break;
      case CFFT_OPT_GRAPH: p->graph = value != 0; break;
      case CFFT_OPT_TILE_ROWS:
        require(value == 0 || ((value == 1 || value == 2 || value == 4 || value == 8) && p->H[1] % value == 0),
                CFFT_ERR_INVALID, "tile rows must be 0 or 1, 2, 4, 8 dividing H1");
        p->tile_rows = (int)value;
        break;
      case CFFT_OPT_FP32_TWIDDLES:
        require(value == 0 || !p->axis_split, CFFT_ERR_UNSUPPORTED,
                "the axis-split chain forms its twiddle products in FP64");
        p->arith32 = value != 0;
        break;
      case CFFT_OPT_FFT_LANES: {
        // 0 or the default split: the default kernels; 2 on H = 16, 32 (path QS)
        const int def = (p->H[0] % 16 == 0) ? 4 : ((p->H[0] % 4 == 0) ? 2 : 1);
        require(value == 0 || value == def || (value == 2 && is_qs(p) && (p->H[0] == 16 || p->H[0] == 32)),
                CFFT_ERR_UNSUPPORTED, "lane split not available for this plan");
        require(value == 0 || value == def || !p->axis_split, CFFT_ERR_UNSUPPORTED,
                "the axis-split chain has its own line transforms");
        p->lanes = (value == 0 || value == def) ? 0 : (int)value;
        break;
      }
      case CFFT_OPT_PAIR_THREADS:
        require(value == 0 || value == 256 || (value == 512 && is_qs(p) && (p->H[1] == 16 || p->H[1] == 24 ||
                                                                            p->H[1] == 32)),
                CFFT_ERR_UNSUPPORTED, "plane-pair workgroup size not available for this plan");
        require(value != 512 || !fp32_arith(p), CFFT_ERR_UNSUPPORTED,
                "the FP32-twiddle plane-pair kernel runs 256 threads");
        p->pair_threads = value == 512 ? 512 : 0;
        break;
      case CFFT_OPT_AXIS_SPLIT:
        require(value == 0 || (value == 1 && is_qs(p) && p->qs_sep && cubic(p) &&
                               (p->H[0] == 8 || p->H[0] == 12 || p->H[0] == 16 || p->H[0] == 24 ||
                                p->H[0] == 32) &&
                               !p->arith32 && !p->lanes),
                CFFT_ERR_UNSUPPORTED,
                "the axis-split chain needs path QS, a separable multiply, H = 8, 12, 16, 24 or 32, FP64 "
                "twiddle products and the default lane split");
        p->axis_split = value != 0;
        break;
      default: throw Error{CFFT_ERR_INVALID, "unknown option " + std::to_string(key)};
    }
  });
}

int cfft_profile(cfft_plan* p, int32_t enable) {
  return guarded([&] {
    check_plan(p);
    p->prof = enable != 0;
  });
}

int cfft_profile_read(cfft_plan* p, double* ms_out, int64_t* count_out) {
  return guarded([&] {
    check_plan(p);
    require(ms_out && count_out, CFFT_ERR_INVALID, "null argument");
    for (int i = 0; i < CFFT_N_STAGES; ++i) {
      ms_out[i] = 0.0;
      count_out[i] = 0;
    }
    for (auto& m : p->marks) {
      hip_check(hipEventSynchronize(m.b), "hipEventSynchronize");
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, m.a, m.b), "hipEventElapsedTime");
      ms_out[m.stage] += ms;
      count_out[m.stage] += 1;
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    p->marks.clear();
  });
}

int cfft_exp_dev(cfft_plan* p, const void* w, void* E, double factor, int64_t n, void* stream) {
  return guarded([&] {
    check_plan(p);
    require(w && E && n >= 0, CFFT_ERR_INVALID, "bad argument");
    hipStream_t s = pick(p, stream);
    if (p->prec == CFFT_FP64)
      exp_dev<double>(p, static_cast<const double*>(w), static_cast<double*>(E), factor, n, s);
    else
      exp_dev<float>(p, static_cast<const float*>(w), static_cast<double*>(E), factor, n, s);
  });
}

int cfft_sync(cfft_plan* p) {
  return guarded([&] {
    check_plan(p);
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  });
}

}  // extern "C"
