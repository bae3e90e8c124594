// Multi-plan contour sweeps: the stored (history) sweeps of several plans on
// the plane-pair Q chain advanced together, one launch per chain stage and
// contour step for all of them.  Plans may differ in grid, space group and
// batch; they share precision and contour length.  No counterpart in the
// reference (its only parallelism is threads over blocks,
// src/execution.jl:20, 50).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pipeline.hpp"

struct cfft_plan;

namespace cfft {

constexpr int kEnsMax = 64;  // plans per ensemble

// One plan of an ensemble as its kernels see it: the launch parameters of
// its history sweep, its buffers, and where the sweep's slices live (slice s
// at q + s * stride elements of the data type).
struct EnsSlot {
  PipeP p;
  PipeBufs B;  // f0_in / f0_out are set per launch from q, stride and the step
  void* q;
  long stride;
  int L;       // channel transform length H (cubic)
};

// Work entry of one workgroup: a slot and the item of its plan (a plane pair
// or a tile).  A launch reads its entry count and entries from a work list
// the host fills per sweep: only the fields still active are listed, and
// workgroups past the count exit at once.
struct EnsWork {
  int slot;
  int item;
};

// Transform lengths an ensemble kernel is compiled for.
constexpr int kEnsLens[] = {8, 12, 16, 24, 32};
inline bool ens_len_ok(int L) {
  for (int v : kEnsLens)
    if (v == L) return true;
  return false;
}

// A plan's history sweep: per-field interaction factors (E1 on steps < sw,
// then E2) and the slice array [ns + 1][batch][M^3].
struct EnsPart {
  cfft_plan* plan;
  void* hist;
  std::vector<const void*> E1, E2;
  std::vector<int32_t> sw;
};

// Host side of a plan needed by an ensemble (defined with the plan).
struct EnsPlanInfo {
  int path, prec, arith32, device, L, cubic, batch;
  int sep;                      // path QS: the separable Q multiply
  int lanes;                    // path QS: a lane split other than the default (0: none)
  int axis_split;               // path QS: the per-axis chain (not run by ensembles)
  long pair_items, tile_items;  // plane-pair and tile work items of one launch
  long M3;
};
EnsPlanInfo ens_plan_info(cfft_plan* p);
void ens_slot_params(cfft_plan* p, const EnsPart& part, EnsSlot* out);

struct EnsSweep;
// Checks the parts (path QS, FP64 twiddle arithmetic, one precision, cubic
// channel grids of a supported length, all on the current device), sorts
// them into classes by transform length, one stream per class, and
// allocates the slot table and work lists.
EnsSweep* ens_sweep_create(const std::vector<EnsPart>& parts, int ns);
// The stream of the class that part k belongs to: work on part k's buffers
// before and after a sweep goes there.
hipStream_t ens_part_stream(EnsSweep* e, int k);
// The ns-step history sweep of the fields flagged in `active` (active[k][b]
// for batch entry b of part k; the others are left out of the work lists),
// slice 0 of each part already set on its class stream; the slot table is
// refreshed from `parts` (factor pointers and switch steps may change).  Per class one launch per chain stage and
// contour step covers all its listed parts; the launch sequence is captured
// once per class into a hipGraph and replayed on the class stream, so the
// classes run concurrently.
void ens_sweep_run(EnsSweep* e, const std::vector<EnsPart>& parts, const std::vector<std::vector<char>>& active);
// Waits for the work on every class stream.
void ens_sweep_sync(EnsSweep* e);
void ens_sweep_destroy(EnsSweep* e);

}  // namespace cfft
