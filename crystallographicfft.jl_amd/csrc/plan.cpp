// Host-side plan builder (see plan.hpp).  Each function cites the reference
// function whose algorithm it implements.
#include "plan.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

namespace cfft {

namespace {

inline int pmod(long long a, int n) {
  long long r = a % n;
  return static_cast<int>(r < 0 ? r + n : r);
}

long long gcdll(long long a, long long b) {
  while (b) {
    long long t = a % b;
    a = b;
    b = t;
  }
  return a < 0 ? -a : a;
}

// exp(-2 pi i x) for a phase x given as a double (cispi(-2x) of the reference).
inline void cis_m2pi(double x, double& re, double& im) {
  double r = std::fmod(2.0 * x, 2.0);  // angle / pi in (-2, 2)
  re = std::cos(M_PI * r);
  im = -std::sin(M_PI * r);
}

// Integer inverse of a unimodular matrix (dim 2 or 3), via the adjugate.
void int_inverse(int dim, const int R[3][3], int Ri[3][3]) {
  std::memset(Ri, 0, sizeof(int) * 9);
  if (dim == 2) {
    int det = R[0][0] * R[1][1] - R[0][1] * R[1][0];
    Ri[0][0] = R[1][1] / det;
    Ri[0][1] = -R[0][1] / det;
    Ri[1][0] = -R[1][0] / det;
    Ri[1][1] = R[0][0] / det;
    Ri[2][2] = 1;
    return;
  }
  int c[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      c[i][j] = R[i1][j1] * R[i2][j2] - R[i1][j2] * R[i2][j1];
    }
  int det = R[0][0] * c[0][0] + R[0][1] * c[0][1] + R[0][2] * c[0][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ri[i][j] = c[j][i] / det;
}

// check_shift_invariance, src/symmetry_ops.jl:81-94.
bool shift_deltas(int dim, const std::vector<Op>& ops, const double* shift, const int* N,
                  std::vector<std::array<int, 3>>& deltas) {
  deltas.assign(ops.size(), {0, 0, 0});
  for (size_t g = 0; g < ops.size(); ++g) {
    for (int i = 0; i < dim; ++i) {
      double s = 0.0;
      for (int j = 0; j < dim; ++j) s += ops[g].R[i][j] * shift[j];
      double df = (s - shift[i]) * N[i];
      double di = std::nearbyint(df);
      if (std::fabs(df - di) > 1e-5) return false;
      deltas[g][i] = static_cast<int>(di);
    }
  }
  return true;
}

std::vector<Op> add_deltas(const std::vector<Op>& ops, const std::vector<std::array<int, 3>>& d) {
  std::vector<Op> out = ops;
  for (size_t g = 0; g < ops.size(); ++g)
    for (int i = 0; i < 3; ++i) out[g].t[i] += d[g][i];
  return out;
}

int lprod_odd(int dim, const std::vector<Op>& ops) {
  int p = 1;
  for (int d = 0; d < dim; ++d) {
    bool odd = false;
    for (const Op& op : ops)
      if (op.t[d] % 2 != 0) odd = true;
    if (odd) p *= 2;
  }
  return p;
}

bool rotations_closed(int dim, const std::vector<std::array<int, 9>>& Rs) {
  std::set<std::array<int, 9>> s(Rs.begin(), Rs.end());
  std::array<int, 9> I{};
  for (int i = 0; i < 3; ++i) I[i * 3 + i] = 1;
  if (!s.count(I)) return false;
  for (const auto& a : Rs)
    for (const auto& b : Rs) {
      std::array<int, 9> c{};
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          int v = 0;
          for (int k = 0; k < 3; ++k) v += a[i * 3 + k] * b[k * 3 + j];
          c[i * 3 + j] = v;
        }
      if (!s.count(c)) return false;
    }
  (void)dim;
  return true;
}

}  // namespace

long long lcm3(int a, int b, int c) {
  long long l = a;
  l = l / gcdll(l, b) * b;
  l = l / gcdll(l, c) * c;
  return l;
}

std::vector<Op> unpack_ops(const int32_t* ops, int nops) {
  std::vector<Op> out(nops);
  for (int g = 0; g < nops; ++g) {
    const int32_t* p = ops + 12 * g;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) out[g].R[i][j] = p[i * 3 + j];
    for (int i = 0; i < 3; ++i) out[g].t[i] = p[9 + i];
  }
  return out;
}

void pack_ops(const std::vector<Op>& ops, int32_t* out) {
  for (size_t g = 0; g < ops.size(); ++g) {
    int32_t* p = out + 12 * g;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) p[i * 3 + j] = ops[g].R[i][j];
    for (int i = 0; i < 3; ++i) p[9 + i] = ops[g].t[i];
  }
}

// _count_special_positions_fast, src/asu.jl:148-214: orbits (worklist closure)
// with multiplicity < |G|.
int count_special_positions(int dim, const int* N, const std::vector<Op>& ops) {
  int n0 = N[0], n1 = dim > 1 ? N[1] : 1, n2 = dim > 2 ? N[2] : 1;
  long long n_total = 1LL * n0 * n1 * n2;
  std::vector<uint8_t> visited(n_total, 0);
  std::vector<long long> buf;
  int sp = 0;
  const int Nn[3] = {n0, n1, n2};
  for (long long lin = 0; lin < n_total; ++lin) {
    if (visited[lin]) continue;
    buf.clear();
    buf.push_back(lin);
    visited[lin] = 1;
    for (size_t wi = 0; wi < buf.size(); ++wi) {
      long long c = buf[wi];
      int x[3] = {static_cast<int>(c % n0), static_cast<int>((c / n0) % n1),
                  static_cast<int>(c / (1LL * n0 * n1))};
      for (const Op& op : ops) {
        long long li = 0, stride = 1;
        for (int d = 0; d < dim; ++d) {
          long long s = op.t[d];
          for (int j = 0; j < dim; ++j) s += 1LL * op.R[d][j] * x[j];
          li += pmod(s, Nn[d]) * stride;
          stride *= Nn[d];
        }
        if (!visited[li]) {
          visited[li] = 1;
          buf.push_back(li);
        }
      }
    }
    if (static_cast<long long>(buf.size()) < static_cast<long long>(ops.size())) ++sp;
  }
  return sp;
}

// find_optimal_shift, src/asu.jl:57-123.
void find_optimal_shift(int dim, const int* N, const std::vector<Op>& ops, double* shift,
                        std::vector<Op>& shifted) {
  int Np[3] = {12, 12, 12};
  // _rescale_ops, src/asu.jl:130-138: t_new = (t * 12) ÷ N (truncating).
  std::vector<Op> opsP = ops;
  for (Op& op : opsP)
    for (int d = 0; d < dim; ++d) op.t[d] = (op.t[d] * Np[d]) / N[d];
  const double cands[4] = {0.0, 0.5, 1.0 / 3.0, 0.25};
  int best = 0;
  int min_sp = INT_MAX;
  std::vector<std::array<int, 3>> deltas;
  for (int c = 0; c < 4; ++c) {
    double s[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) s[d] = cands[c] / Np[d];
    if (!shift_deltas(dim, opsP, s, Np, deltas)) continue;
    int sp = count_special_positions(dim, Np, add_deltas(opsP, deltas));
    if (sp < min_sp) {
      min_sp = sp;
      best = c;
    }
    if (min_sp == 0) break;
  }
  double s[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) s[d] = cands[best] / N[d];
  std::vector<Op> best_ops;
  if (shift_deltas(dim, ops, s, N, deltas)) {
    best_ops = add_deltas(ops, deltas);
  } else {
    for (int d = 0; d < 3; ++d) s[d] = 0.0;
    best_ops = ops;
  }
  double sh[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) sh[d] = 0.5 / N[d];
  if (shift_deltas(dim, ops, sh, N, deltas)) {
    std::vector<Op> half = add_deltas(ops, deltas);
    if (lprod_odd(dim, half) > lprod_odd(dim, best_ops)) {
      for (int d = 0; d < 3; ++d) s[d] = sh[d];
      best_ops = half;
    }
  }
  for (int d = 0; d < dim; ++d) shift[d] = s[d];
  shifted = best_ops;
}

// calc_spectral_asu, src/spectral_indexing.jl:22-176.
SpecASU calc_spectral_asu(int dim, const int* N, const std::vector<Op>& ops) {
  SpecASU out;
  out.dim = dim;
  for (int d = 0; d < 3; ++d) out.N[d] = d < dim ? N[d] : 1;
  const int n0 = out.N[0], n1 = out.N[1], n2 = out.N[2];
  const long long n_total = 1LL * n0 * n1 * n2;
  const int nops = static_cast<int>(ops.size());
  // reciprocal point group R* = (R^-1)^T (dual_ops)
  std::vector<std::array<int, 9>> Rs(nops);
  for (int g = 0; g < nops; ++g) {
    int Ri[3][3];
    int_inverse(dim, ops[g].R, Ri);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Rs[g][i * 3 + j] = Ri[j][i];
  }
  const bool closed = rotations_closed(dim, Rs);
  std::vector<uint8_t> visited(n_total, 0);
  std::vector<long long> work;
  auto image = [&](int g, const int* k) {
    long long li = 0, stride = 1;
    for (int d = 0; d < dim; ++d) {
      long long s = 0;
      for (int j = 0; j < dim; ++j) s += 1LL * Rs[g][d * 3 + j] * k[j];
      li += pmod(s, out.N[d]) * stride;
      stride *= out.N[d];
    }
    return li;
  };
  for (long long lin = 0; lin < n_total; ++lin) {
    if (visited[lin]) continue;
    visited[lin] = 1;
    long long min_lin = lin;
    int orbit = 1;
    int k[3] = {static_cast<int>(lin % n0), static_cast<int>((lin / n0) % n1),
                static_cast<int>(lin / (1LL * n0 * n1))};
    if (closed) {
      for (int g = 0; g < nops; ++g) {
        long long li = image(g, k);
        if (!visited[li]) {
          visited[li] = 1;
          ++orbit;
          if (li < min_lin) min_lin = li;
        }
      }
    } else {
      work.clear();
      work.push_back(lin);
      for (size_t wi = 0; wi < work.size(); ++wi) {
        long long c = work[wi];
        int kc[3] = {static_cast<int>(c % n0), static_cast<int>((c / n0) % n1),
                     static_cast<int>(c / (1LL * n0 * n1))};
        for (int g = 0; g < nops; ++g) {
          long long li = image(g, kc);
          if (!visited[li]) {
            visited[li] = 1;
            ++orbit;
            work.push_back(li);
            if (li < min_lin) min_lin = li;
          }
        }
      }
    }
    int kr[3] = {static_cast<int>(min_lin % n0), static_cast<int>((min_lin / n0) % n1),
                 static_cast<int>(min_lin / (1LL * n0 * n1))};
    // extinction filter: |sum over Stab(k) of exp(-2 pi i k.t_g/N)| > 1e-5
    double sr = 0.0, si = 0.0;
    for (int g = 0; g < nops; ++g) {
      bool stab = true;
      for (int d = 0; d < dim && stab; ++d) {
        long long s = 0;
        for (int j = 0; j < dim; ++j) s += 1LL * Rs[g][d * 3 + j] * kr[j];
        if ((s - kr[d]) % out.N[d] != 0) stab = false;
      }
      if (!stab) continue;
      double phase = 0.0;
      for (int d = 0; d < dim; ++d)
        phase += static_cast<double>(1LL * kr[d] * ops[g].t[d]) / out.N[d];
      double re, im;
      cis_m2pi(phase, re, im);
      sr += re;
      si += im;
    }
    if (std::hypot(sr, si) > 1e-5) {
      out.idx.push_back({kr[0], kr[1], kr[2]});
      out.mult.push_back(orbit);
    }
  }
  // sort lexicographically by (k1, k2, k3)
  std::vector<size_t> order(out.idx.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return out.idx[a] < out.idx[b]; });
  SpecASU sorted = out;
  for (size_t i = 0; i < order.size(); ++i) {
    sorted.idx[i] = out.idx[order[i]];
    sorted.mult[i] = out.mult[order[i]];
  }
  return sorted;
}

// auto_L, src/krfft.jl:225-279.
void auto_L(int dim, const std::vector<Op>& ops, int* L) {
  int Lmax[3] = {1, 1, 1};
  for (const Op& op : ops)
    for (int d = 0; d < dim; ++d)
      if (pmod(op.t[d], 2) == 1) Lmax[d] = 2;
  int n_sub = Lmax[0] * Lmax[1] * Lmax[2];
  auto reach = [&](const int* Lt) {
    std::set<std::array<int, 3>> r;
    for (const Op& op : ops) r.insert({pmod(op.t[0], Lt[0]), pmod(op.t[1], Lt[1]), pmod(op.t[2], Lt[2])});
    return static_cast<int>(r.size());
  };
  if (n_sub <= 1 || reach(Lmax) == n_sub) {
    for (int d = 0; d < 3; ++d) L[d] = Lmax[d];
    return;
  }
  int best[3] = {1, 1, 1};
  for (int mask = 1; mask < (1 << dim); ++mask) {
    int Lt[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d)
      if (((mask >> d) & 1) && Lmax[d] == 2) Lt[d] = 2;
    int n = Lt[0] * Lt[1] * Lt[2];
    if (reach(Lt) == n && n > best[0] * best[1] * best[2])
      for (int d = 0; d < 3; ++d) best[d] = Lt[d];
  }
  for (int d = 0; d < 3; ++d) L[d] = best[d];
}

// _select_rep_ops, src/krfft.jl:958-987.
std::vector<int> select_rep_ops(int dim, const std::vector<Op>& ops, const int* L, const int* N) {
  std::map<std::array<int, 3>, int> rep, qual;
  for (size_t g = 0; g < ops.size(); ++g) {
    const Op& op = ops[g];
    std::array<int, 3> x0 = {pmod(op.t[0], L[0]), pmod(op.t[1], L[1]), pmod(op.t[2], L[2])};
    bool diag = true;
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j)
        if (i != j && op.R[i][j] != 0) diag = false;
    bool simple = true;
    for (int d = 0; d < dim; ++d) {
      int m = pmod(op.t[d], N[d]);
      if (m != 0 && m != N[d] - 1) simple = false;
    }
    int q = diag ? (simple ? 2 : 1) : 0;
    auto it = qual.find(x0);
    if (it == qual.end() || q > it->second) {
      rep[x0] = static_cast<int>(g);
      qual[x0] = q;
    }
  }
  std::vector<int> out;
  for (int a2 = 0; a2 < L[2]; ++a2)
    for (int a1 = 0; a1 < L[1]; ++a1)
      for (int a0 = 0; a0 < L[0]; ++a0) {
        auto it = rep.find({a0, a1, a2});
        if (it == rep.end())
          throw Error{4, "Subgrid x0=(" + std::to_string(a0) + "," + std::to_string(a1) + "," +
                             std::to_string(a2) + ") not reachable"};
        out.push_back(it->second);
      }
  return out;
}

// detect_centering_type, src/symmetry_ops.jl:133-179.
int detect_centering(int dim, const int* N, const std::vector<Op>& ops) {
  std::vector<std::array<bool, 3>> pure;
  for (const Op& op : ops) {
    bool ident = true;
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j)
        if (op.R[i][j] != (i == j ? 1 : 0)) ident = false;
    if (!ident) continue;
    bool zero = true;
    for (int d = 0; d < dim; ++d)
      if (std::abs(op.t[d]) >= 0.1) zero = false;
    if (zero) continue;
    std::array<bool, 3> h = {false, false, false};
    for (int d = 0; d < dim; ++d) h[d] = std::fabs(std::abs(op.t[d]) / double(N[d]) - 0.5) < 0.01;
    pure.push_back(h);
  }
  if (pure.empty() || dim < 3) return CENT_P;
  bool xyz = false, yz = false, xz = false, xy = false;
  for (const auto& p : pure) {
    if (p[0] && p[1] && p[2]) xyz = true;
    if (!p[0] && p[1] && p[2]) yz = true;
    if (p[0] && !p[1] && p[2]) xz = true;
    if (p[0] && p[1] && !p[2]) xy = true;
  }
  if (int(yz) + int(xz) + int(xy) >= 2) return CENT_F;
  if (xyz) return CENT_I;
  if (xy) return CENT_C;
  if (yz) return CENT_A;
  if (xz) return CENT_A;
  return CENT_P;
}

// _alive_offsets, src/centering_fold.jl:111-127 (P: all eight parities).
std::vector<std::array<int, 3>> alive_offsets(int cent) {
  switch (cent) {
    case CENT_I: return {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    case CENT_F: return {{0, 0, 0}, {1, 1, 1}};
    case CENT_C: return {{0, 0, 0}, {1, 1, 0}, {0, 0, 1}, {1, 1, 1}};
    case CENT_A: return {{0, 0, 0}, {0, 1, 1}, {1, 0, 0}, {1, 1, 1}};
    default: {
      std::vector<std::array<int, 3>> all;
      for (int e = 0; e < 8; ++e) all.push_back({e & 1, (e >> 1) & 1, (e >> 2) & 1});
      return all;
    }
  }
}

// _build_spectral_reverse_lookup, src/krfft.jl:1183-1255 (first-come order).
ReverseLookup reverse_lookup(const SpecASU& spec, const std::vector<Op>& ops) {
  const int* N = spec.N;
  ReverseLookup rl;
  long long P = lcm3(N[0], N[1], N[2]);
  rl.P = static_cast<int>(P);
  const long long n_total = 1LL * N[0] * N[1] * N[2];
  rl.idx.assign(n_total, -1);
  rl.j.assign(n_total, 0);
  rl.conj.assign(n_total, 0);
  auto lin = [&](const long long* h) { return h[0] + N[0] * (h[1] + 1LL * N[1] * h[2]); };
  auto put = [&](long long li, int s, long long j, bool c) {
    if (rl.idx[li] < 0) {
      rl.idx[li] = s;
      rl.j[li] = static_cast<int32_t>(((j % P) + P) % P);
      rl.conj[li] = c ? 1 : 0;
    }
  };
  for (size_t s = 0; s < spec.idx.size(); ++s) {
    const auto& h = spec.idx[s];
    long long hh[3] = {h[0], h[1], h[2]};
    put(lin(hh), static_cast<int>(s), 0, false);
    for (const Op& op : ops) {
      long long hr[3], hn[3];
      for (int d = 0; d < 3; ++d) {
        long long v = 0;
        for (int d2 = 0; d2 < 3; ++d2) v += 1LL * op.R[d2][d] * h[d2];
        hr[d] = pmod(v, N[d]);
        hn[d] = pmod(-hr[d], N[d]);
      }
      long long j = 0;
      for (int d = 0; d < 3; ++d) j += 1LL * h[d] * op.t[d] * (P / N[d]);
      put(lin(hr), static_cast<int>(s), j, false);
      put(lin(hn), static_cast<int>(s), -j, true);
    }
    long long hm[3] = {pmod(-h[0], N[0]), pmod(-h[1], N[1]), pmod(-h[2], N[2])};
    put(lin(hm), static_cast<int>(s), 0, true);
  }
  return rl;
}

// _build_fill_map, src/q_fused_krfft.jl:733-783 (0-based M-grid index).
std::vector<int32_t> fill_map(int dim, const int* N, const std::vector<Op>& ops, const int* L) {
  int n[3] = {1, 1, 1}, M[3] = {1, 1, 1}, Ll[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    n[d] = N[d];
    Ll[d] = L[d];
    M[d] = N[d] / L[d];
  }
  const long long n_total = 1LL * n[0] * n[1] * n[2];
  std::vector<int32_t> out(n_total, -1);
  for (long long li = 0; li < n_total; ++li) {
    int x[3] = {static_cast<int>(li % n[0]), static_cast<int>((li / n[0]) % n[1]),
                static_cast<int>(li / (1LL * n[0] * n[1]))};
    for (const Op& op : ops) {
      int y[3] = {0, 0, 0};
      bool on = true;
      for (int d = 0; d < dim && on; ++d) {
        long long s = op.t[d];
        for (int j = 0; j < dim; ++j) s += 1LL * op.R[d][j] * x[j];
        y[d] = pmod(s, n[d]);
        if (y[d] % Ll[d] != 0) on = false;
      }
      if (on) {
        out[li] = (y[0] / Ll[0]) + M[0] * ((y[1] / Ll[1]) + M[1] * (y[2] / Ll[2]));
        break;
      }
    }
    if (out[li] < 0)
      throw Error{5, "No symmetry operation maps grid point (" + std::to_string(x[0]) + "," +
                         std::to_string(x[1]) + "," + std::to_string(x[2]) + ") to the subgrid"};
  }
  return out;
}

// calc_asu, src/asu.jl:216-325.  The work list is processed last-in first-out
// as the reference's pop!; the final sort makes the result independent of it.
std::vector<ASUPoint> calc_asu(int dim, const int* N0, const std::vector<Op>& ops0) {
  struct State {
    std::array<int, 3> N, scale, offset, depth;
    std::array<bool, 3> gp;
    std::vector<Op> ops;
  };
  std::vector<ASUPoint> pts;
  std::vector<State> queue;
  State s0;
  for (int d = 0; d < 3; ++d) {
    s0.N[d] = d < dim ? N0[d] : 1;
    s0.scale[d] = 1;
    s0.offset[d] = 0;
    s0.depth[d] = 0;
    s0.gp[d] = d >= dim;  // padding dimensions behave as general positions
  }
  s0.ops = ops0;
  queue.push_back(s0);
  auto apply = [&](const Op& op, const int* x, const std::array<int, 3>& n, int* out) {
    for (int i = 0; i < dim; ++i) {
      long long s = op.t[i];
      for (int j = 0; j < dim; ++j) s += 1LL * op.R[i][j] * x[j];
      out[i] = pmod(s, n[i]);
    }
  };
  while (!queue.empty()) {
    State st = std::move(queue.back());
    queue.pop_back();
    bool zero = false, all_gp = true, all_le1 = true;
    for (int d = 0; d < dim; ++d) {
      if (st.N[d] == 0) zero = true;
      if (!st.gp[d]) all_gp = false;
      if (st.N[d] > 1) all_le1 = false;
    }
    if (zero) continue;
    if (all_gp || all_le1) {
      // leaf: one representative (lexicographic minimum) per local orbit
      const long long n_loc = 1LL * st.N[0] * st.N[1] * st.N[2];
      std::vector<uint8_t> visited(n_loc, 0);
      auto lin = [&](const int* x) { return x[0] + 1LL * st.N[0] * (x[1] + 1LL * st.N[1] * x[2]); };
      std::vector<std::array<int, 3>> orbit, stack;
      for (long long li = 0; li < n_loc; ++li) {
        if (visited[li]) continue;
        std::array<int, 3> p = {int(li % st.N[0]), int((li / st.N[0]) % st.N[1]),
                                int(li / (1LL * st.N[0] * st.N[1]))};
        orbit.assign(1, p);
        stack.assign(1, p);
        visited[li] = 1;
        while (!stack.empty()) {
          std::array<int, 3> c = stack.back();
          stack.pop_back();
          for (const Op& op : st.ops) {
            std::array<int, 3> y = {0, 0, 0};
            apply(op, c.data(), st.N, y.data());
            long long ly = lin(y.data());
            if (!visited[ly]) {
              visited[ly] = 1;
              orbit.push_back(y);
              stack.push_back(y);
            }
          }
        }
        std::array<int, 3> rep = *std::min_element(orbit.begin(), orbit.end());
        ASUPoint ap;
        for (int d = 0; d < 3; ++d) {
          ap.idx[d] = d < dim ? st.scale[d] * rep[d] + st.offset[d] : 0;
          ap.depth[d] = d < dim ? st.depth[d] : 0;
        }
        ap.mult = (int)orbit.size();
        pts.push_back(ap);
      }
      continue;
    }
    std::array<bool, 3> eff = st.gp;
    for (int d = 0; d < dim; ++d) {
      if (st.gp[d]) continue;
      int pe[3] = {0, 0, 0}, po[3] = {0, 0, 0}, out[3] = {0, 0, 0};
      po[d] = 1;
      bool ok = true;
      for (const Op& op : st.ops) {
        apply(op, pe, st.N, out);
        if (out[d] % 2 != 0) { ok = false; break; }
        apply(op, po, st.N, out);
        if (out[d] % 2 != 1) { ok = false; break; }
      }
      if (!ok) eff[d] = true;
    }
    int rng[3];
    for (int d = 0; d < 3; ++d) rng[d] = (d < dim && !eff[d] && st.N[d] > 1) ? 2 : 1;
    std::vector<State> children;
    for (int p2 = 0; p2 < rng[2]; ++p2)
      for (int p1 = 0; p1 < rng[1]; ++p1)
        for (int p0 = 0; p0 < rng[0]; ++p0) {
          const int par[3] = {p0, p1, p2};
          int S[3];
          for (int d = 0; d < 3; ++d) S[d] = (d >= dim || eff[d]) ? 1 : 2;
          State ch;
          bool valid = true;
          for (const Op& op : st.ops) {
            Op no = op;
            for (int i = 0; i < dim && valid; ++i) {
              long long tmp = op.t[i] - par[i];
              for (int j = 0; j < dim; ++j) tmp += 1LL * op.R[i][j] * par[j];
              for (int j = 0; j < dim; ++j) {
                long long rv = 1LL * op.R[i][j] * S[j];
                if (rv % S[i] != 0) valid = false;
                no.R[i][j] = (int)(rv / S[i]);
              }
              if (tmp % S[i] != 0) valid = false;
              no.t[i] = (int)(tmp / S[i]);
            }
            if (!valid) break;
            ch.ops.push_back(no);
          }
          if (!valid) continue;
          for (int d = 0; d < 3; ++d) {
            const bool e = d >= dim || eff[d];
            ch.N[d] = e ? st.N[d] : st.N[d] / 2;
            ch.gp[d] = e || par[d] == 1;
            ch.depth[d] = e ? st.depth[d] : st.depth[d] + (par[d] == 0 ? 1 : 0);
            ch.scale[d] = st.scale[d] * S[d];
            ch.offset[d] = st.scale[d] * par[d] + st.offset[d];
          }
          children.push_back(std::move(ch));
        }
    for (auto& ch : children) queue.push_back(std::move(ch));
  }
  std::sort(pts.begin(), pts.end(), [](const ASUPoint& a, const ASUPoint& b) { return a.idx < b.idx; });
  return pts;
}

// pack_asu, src/pack_asu.jl:6-102.
std::vector<Block> pack_asu(int dim, const std::vector<ASUPoint>& pts) {
  std::map<std::array<int, 3>, std::vector<std::array<int, 3>>> groups;
  for (const auto& p : pts) groups[p.depth].push_back(p.idx);
  std::vector<Block> blocks;
  for (auto& kv : groups) {
    auto& pool = kv.second;
    std::sort(pool.begin(), pool.end());
    std::set<std::array<int, 3>> in_pool(pool.begin(), pool.end());
    size_t head = 0;
    while (true) {
      while (head < pool.size() && !in_pool.count(pool[head])) ++head;
      if (head >= pool.size()) break;
      const std::array<int, 3> start = pool[head];
      std::array<int, 3> lo = start, hi = start;
      for (int d = 0; d < dim; ++d) {
        int steps = 0;
        while (true) {
          const int nxt = start[d] + steps + 1;
          std::array<int, 3> a = lo, b = hi;
          a[d] = b[d] = nxt;
          bool full = true;
          for (int z = a[2]; z <= b[2] && full; ++z)
            for (int y = a[1]; y <= b[1] && full; ++y)
              for (int x = a[0]; x <= b[0] && full; ++x)
                if (!in_pool.count({x, y, z})) full = false;
          if (!full) break;
          ++steps;
        }
        hi[d] = start[d] + steps;
      }
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) in_pool.erase({x, y, z});
      Block b;
      b.start = lo;
      for (int d = 0; d < 3; ++d) b.shape[d] = hi[d] - lo[d] + 1;
      b.depth = kv.first;
      blocks.push_back(b);
    }
  }
  return blocks;
}

// K_spec of plan_m2_scft / plan_centered_scft (src/q_fused_krfft.jl:966-973).
std::vector<double> kernel_spec(const SpecASU& spec, double ds, const double* lattice) {
  // A[i][j] = lattice[i + 3 j] (columns are lattice vectors); B = 2 pi inv(A)^T
  double A[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = lattice[i + 3 * j];
  double c[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      c[i][j] = A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1];
    }
  double det = A[0][0] * c[0][0] + A[0][1] * c[0][1] + A[0][2] * c[0][2];
  // inv(A)[i][j] = c[j][i]/det ; B[i][j] = 2 pi inv(A)[j][i] = 2 pi c[i][j]/det
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = 2.0 * M_PI * (c[i][j] / det);
  std::vector<double> K(spec.idx.size());
  for (size_t s = 0; s < spec.idx.size(); ++s) {
    double hc[3];
    for (int d = 0; d < 3; ++d) hc[d] = k_centered(spec.idx[s][d], spec.N[d]);
    double k2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      double kv = B[i][0] * hc[0] + B[i][1] * hc[1] + B[i][2] * hc[2];
      k2 += kv * kv;
    }
    K[s] = std::exp(-k2 * ds);
  }
  return K;
}

}  // namespace cfft
