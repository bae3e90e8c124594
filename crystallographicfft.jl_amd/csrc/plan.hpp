// Host-side plan builder for the KRFFT / diffusion hot path.
//
// Plan-time integer maps and FP64 tables: magic shift (src/asu.jl:57-214),
// spectral ASU (src/spectral_indexing.jl:22-176), auto_L and representative ops
// (src/krfft.jl:225-279, 958-987), centering detection (src/symmetry_ops.jl:
// 133-179), alive fold offsets (src/centering_fold.jl:111-127), the spectral
// reverse lookup (src/krfft.jl:1183-1255) and the fill map
// (src/q_fused_krfft.jl:733-783).  Pure C++17, no HIP dependency.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace cfft {

struct Op {
  int R[3][3];
  int t[3];
};

enum Centering { CENT_P = 0, CENT_C = 1, CENT_A = 2, CENT_I = 3, CENT_F = 4 };

// Error with a status code from include/cfft_hip.h.
struct Error {
  int code;
  std::string msg;
};

std::vector<Op> unpack_ops(const int32_t* ops, int nops);
void pack_ops(const std::vector<Op>& ops, int32_t* out);

// Magic shift: returns the fractional shift b/N (dim entries) and the shifted
// ops t' = t + (R b - b) N.
void find_optimal_shift(int dim, const int* N, const std::vector<Op>& ops,
                        double* shift, std::vector<Op>& shifted);

int count_special_positions(int dim, const int* N, const std::vector<Op>& ops);

struct SpecASU {
  int dim = 3;
  int N[3] = {1, 1, 1};
  std::vector<std::array<int, 3>> idx;  // 0-based, sorted lexicographically
  std::vector<int> mult;
};

SpecASU calc_spectral_asu(int dim, const int* N, const std::vector<Op>& ops);

// src/spectral_indexing.jl:195-238: signed frequency of a 0..N-1 index.
inline int k_signed(int v, int n) { return v >= n - n / 2 ? v - n : v; }
// Centred frequency of the diffusion kernels (h >= N/2 -> h - N).
inline int k_centered(int v, int n) { return v >= n / 2 ? v - n : v; }

void auto_L(int dim, const std::vector<Op>& ops, int* L);
std::vector<int> select_rep_ops(int dim, const std::vector<Op>& ops, const int* L, const int* N);
int detect_centering(int dim, const int* N, const std::vector<Op>& ops);
std::vector<std::array<int, 3>> alive_offsets(int cent);

// Reverse lookup over the full grid: for each column-major h in [0,N)^3 the
// spectral index (-1 = extinct), the phase exponent j (phase = exp(2 pi i j/P),
// P = lcm(N)) and a conjugation flag.  Insertion order follows the reference.
struct ReverseLookup {
  int P = 1;
  std::vector<int32_t> idx;
  std::vector<int32_t> j;
  std::vector<uint8_t> conj;
};
ReverseLookup reverse_lookup(const SpecASU& spec, const std::vector<Op>& ops);

std::vector<int32_t> fill_map(int dim, const int* N, const std::vector<Op>& ops, const int* L);

// Real-space ASU point (src/asu.jl:23-27).
struct ASUPoint {
  std::array<int, 3> idx;
  std::array<int, 3> depth;
  int mult;
};
// calc_asu, src/asu.jl:216-325 (recursive parity filter), sorted by idx.
std::vector<ASUPoint> calc_asu(int dim, const int* N, const std::vector<Op>& ops);

// A packed block (src/pack_asu.jl): start, shape (unit stride), depth.
struct Block {
  std::array<int, 3> start, shape, depth;
};
// pack_asu, src/pack_asu.jl:6-102: greedy boxes per depth group (groups in
// sorted depth order).
std::vector<Block> pack_asu(int dim, const std::vector<ASUPoint>& pts);

// K(h) = exp(-|2 pi A^-T h_c|^2 ds), lattice column-major (columns a, b, c).
std::vector<double> kernel_spec(const SpecASU& spec, double ds, const double* lattice);

long long lcm3(int a, int b, int c);

}  // namespace cfft
