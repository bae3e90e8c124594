// Multi-plan history sweeps on the plane-pair Q chain.  The plans of an
// ensemble fall into classes by channel transform length; per class, one
// launch per chain stage and contour step takes the plane pairs (or the
// unfold/fold tiles) of the active fields of every plan of the class from a
// work list, whatever the plan's grid, space group and batch.  Each class
// replays its launch sequence from its own hipGraph on its own stream, so
// the classes run concurrently.  Every work item runs the same phase function, with the same
// parameters and buffers, as the plan's own sweep (cfft_sweep_hist_dev), so
// the results are the plan's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "ensemble.hpp"

using namespace cfft;

namespace {

// The slot table read through the constant address space: its entries do
// not change during a launch, so uniform reads of them are scalar loads
// whatever the kernel stores meanwhile.
typedef __attribute__((address_space(4))) const EnsSlot ConstSlot;
__device__ __forceinline__ const EnsSlot& ens_slot(const EnsSlot* S, int k) {
  return *(const EnsSlot*)((ConstSlot*)S + k);
}

// Plane pairs (FFT y, z, Q multiply, IFFT z, y) of the listed plans.
template <typename T, int L, bool SEP>
__global__ void __launch_bounds__(kQNT) k_p2q_ens(const EnsSlot* __restrict__ S, const int* __restrict__ count,
                                                  const EnsWork* __restrict__ work) {
  __shared__ cx<T> lds[2 * L * kRow<T, L>];
  __shared__ cx<double> fq[(8 * (2 + 2 * L) + kQNT - 1) / kQNT * kQNT];
  if ((int)blockIdx.x >= *count) return;
  const EnsWork wk = work[blockIdx.x];
  const EnsSlot& s = ens_slot(S, wk.slot);
  ph_plane_pair_q<T, double, L, kQNT, SEP>(s.p, s.B, wk.item, lds, fq);
}

// Fold + E + FFT x of slice 0 (the first step of a sweep).
template <typename T, int L>
__global__ void __launch_bounds__(kNT) k_p1_ens(const EnsSlot* __restrict__ S, const int* __restrict__ count,
                                                const EnsWork* __restrict__ work) {
  __shared__ cx<T> lds[64 * (L + 1)];
  if ((int)blockIdx.x >= *count) return;
  const EnsWork wk = work[blockIdx.x];
  const EnsSlot& s = ens_slot(S, wk.slot);
  PipeBufs B = s.B;
  B.f0_in = s.q;
  ph_fold_fftx<T, double, L>(s.p, B, wk.item, 0, lds);
}

// IFFT x + unfold + E of step `step`, q(step + 1) stored to its slice, fused
// with the fold + E + FFT x of step + 1.
template <typename T, int L>
__global__ void __launch_bounds__(kNT) k_p51_ens(const EnsSlot* __restrict__ S, const int* __restrict__ count,
                                                 const EnsWork* __restrict__ work, int step) {
  __shared__ cx<T> lds[64 * kRow<T, L>];
  if ((int)blockIdx.x >= *count) return;
  const EnsWork wk = work[blockIdx.x];
  const EnsSlot& s = ens_slot(S, wk.slot);
  PipeBufs B = s.B;
  B.f0_out = static_cast<T*>(s.q) + (long)(step + 1) * s.stride;
  ph_unfold_fold<T, double, L, true>(s.p, B, wk.item, step, lds);
}

// IFFT x + unfold + E of the last step, q(step + 1) stored to its slice.
template <typename T, int L>
__global__ void __launch_bounds__(kNT) k_p5_ens(const EnsSlot* __restrict__ S, const int* __restrict__ count,
                                                const EnsWork* __restrict__ work, int step) {
  __shared__ cx<T> lds[64 * (L + 1)];
  if ((int)blockIdx.x >= *count) return;
  const EnsWork wk = work[blockIdx.x];
  const EnsSlot& s = ens_slot(S, wk.slot);
  PipeBufs B = s.B;
  B.f0_out = static_cast<T*>(s.q) + (long)(step + 1) * s.stride;
  ph_ifftx_unfold<T, double, L>(s.p, B, wk.item, step, lds);
}

// The work lists of a class: plane pairs of the separable-multiply plans, of
// the others, and the tiles of all.
enum { kListSep = 0, kListGen = 1, kListTile = 2, kListsPerClass = 3 };

// Where a class's lists live: entry counts at count[0..2], entries of list t
// at work + off[t], at most cap[t] of them (the grid size of its launches).
struct ClassLists {
  const int* count;
  const EnsWork* work;
  long off[kListsPerClass];
  unsigned cap[kListsPerClass];
};

template <typename T, int L>
void capture_class(const EnsSlot* S, const ClassLists& c, int ns, hipStream_t s) {
  const int* n_sep = c.count + kListSep;
  const int* n_gen = c.count + kListGen;
  const int* n_tile = c.count + kListTile;
  const EnsWork* sep = c.work + c.off[kListSep];
  const EnsWork* gen = c.work + c.off[kListGen];
  const EnsWork* tiles = c.work + c.off[kListTile];
  const unsigned g_sep = c.cap[kListSep], g_gen = c.cap[kListGen], g_tile = c.cap[kListTile];
  hipLaunchKernelGGL((k_p1_ens<T, L>), dim3(g_tile), dim3(kNT), 0, s, S, n_tile, tiles);
  for (int i = 0; i < ns; ++i) {
    if (g_sep) hipLaunchKernelGGL((k_p2q_ens<T, L, true>), dim3(g_sep), dim3(kQNT), 0, s, S, n_sep, sep);
    if (g_gen) hipLaunchKernelGGL((k_p2q_ens<T, L, false>), dim3(g_gen), dim3(kQNT), 0, s, S, n_gen, gen);
    if (i + 1 < ns)
      hipLaunchKernelGGL((k_p51_ens<T, L>), dim3(g_tile), dim3(kNT), 0, s, S, n_tile, tiles, i);
    else
      hipLaunchKernelGGL((k_p5_ens<T, L>), dim3(g_tile), dim3(kNT), 0, s, S, n_tile, tiles, i);
  }
  hip_check(hipGetLastError(), "ensemble sweep launch");
}

template <typename T>
void capture_class_len(int L, const EnsSlot* S, const ClassLists& c, int ns, hipStream_t s) {
  switch (L) {
    case 8: capture_class<T, 8>(S, c, ns, s); break;
    case 12: capture_class<T, 12>(S, c, ns, s); break;
    case 16: capture_class<T, 16>(S, c, ns, s); break;
    case 24: capture_class<T, 24>(S, c, ns, s); break;
    case 32: capture_class<T, 32>(S, c, ns, s); break;
    default: throw Error{CFFT_ERR_UNSUPPORTED, "no ensemble kernel for this transform length"};
  }
}

}  // namespace

namespace cfft {

struct EnsClass {
  int L = 0;
  std::vector<int> parts;  // part indices
  long off[kListsPerClass] = {0, 0, 0};      // first entry of each list in the work table
  unsigned cap[kListsPerClass] = {0, 0, 0};  // entries each list can hold (its launches' grid)
  hipStream_t st = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;  // end of the class's last sweep (its reads of the tables)
};

struct EnsSweep {
  int prec = 0, ns = 0, n = 0, device = 0;
  std::vector<EnsPlanInfo> info;
  std::vector<int> cls_of;  // class of each part
  std::vector<EnsClass> cls;
  long n_work = 0;
  DevBuf slots, counts, work;   // EnsSlot[n]; int[kListsPerClass] per class; EnsWork[n_work]
  EnsSlot* slots_h = nullptr;   // pinned staging of the uploads
  int* counts_h = nullptr;
  EnsWork* work_h = nullptr;
  hipStream_t up = nullptr;     // the uploads
  hipEvent_t staged = nullptr;  // completion of the last upload
  hipStream_t cap = nullptr;    // graph capture
  ~EnsSweep() {
    for (EnsClass& c : cls)
      if (c.st) (void)hipStreamSynchronize(c.st);
    if (staged) (void)hipEventSynchronize(staged);
    for (EnsClass& c : cls) {
      if (c.exec) (void)hipGraphExecDestroy(c.exec);
      if (c.done) (void)hipEventDestroy(c.done);
      if (c.st) (void)hipStreamDestroy(c.st);
    }
    if (cap) (void)hipStreamDestroy(cap);
    if (up) (void)hipStreamDestroy(up);
    if (staged) (void)hipEventDestroy(staged);
    if (slots_h) (void)hipHostFree(slots_h);
    if (counts_h) (void)hipHostFree(counts_h);
    if (work_h) (void)hipHostFree(work_h);
  }
};

EnsSweep* ens_sweep_create(const std::vector<EnsPart>& parts, int ns) {
  require(!parts.empty() && (int)parts.size() <= kEnsMax, CFFT_ERR_INVALID, "an ensemble holds 1 to 64 plans");
  require(ns >= 1, CFFT_ERR_INVALID, "need ns >= 1");
  std::unique_ptr<EnsSweep> e(new EnsSweep());
  e->ns = ns;
  e->n = (int)parts.size();
  hip_check(hipGetDevice(&e->device), "hipGetDevice");
  for (int k = 0; k < e->n; ++k) {
    const EnsPlanInfo i = ens_plan_info(parts[k].plan);
    require(i.path == CFFT_PATH_QS, CFFT_ERR_UNSUPPORTED, "an ensemble runs plans on path QS");
    require(!i.arith32 && !i.lanes && !i.axis_split, CFFT_ERR_UNSUPPORTED,
            "an ensemble runs the plane-pair kernels with FP64 twiddles and the default lane split");
    require(i.cubic && ens_len_ok(i.L), CFFT_ERR_UNSUPPORTED,
            "an ensemble needs cubic channel grids of length 8, 12, 16, 24 or 32");
    require(i.device == e->device, CFFT_ERR_INVALID, "the plans of an ensemble share the current device");
    require(parts[k].hist != nullptr, CFFT_ERR_INVALID, "null history buffer");
    require(i.batch >= 1 && i.pair_items % i.batch == 0 && i.tile_items % i.batch == 0, CFFT_ERR_INVALID,
            "work items must split evenly over the batch");
    if (k == 0) e->prec = i.prec;
    require(i.prec == e->prec, CFFT_ERR_INVALID, "the plans of an ensemble share one precision");
    e->info.push_back(i);
  }
  // classes by transform length, larger first
  std::vector<int> lens;
  for (const EnsPlanInfo& i : e->info) lens.push_back(i.L);
  std::sort(lens.begin(), lens.end(), [](int a, int b) { return a > b; });
  lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
  e->cls.resize(lens.size());
  e->cls_of.assign(e->n, 0);
  long off = 0;
  for (size_t c = 0; c < lens.size(); ++c) {
    EnsClass& C = e->cls[c];
    C.L = lens[c];
    long cap[kListsPerClass] = {0, 0, 0};
    for (int k = 0; k < e->n; ++k) {
      if (e->info[k].L != C.L) continue;
      C.parts.push_back(k);
      e->cls_of[k] = (int)c;
      cap[e->info[k].sep ? kListSep : kListGen] += e->info[k].pair_items;
      cap[kListTile] += e->info[k].tile_items;
    }
    for (int t = 0; t < kListsPerClass; ++t) {
      require(cap[t] < (1L << 30), CFFT_ERR_SIZE, "ensemble too large");
      C.off[t] = off;
      C.cap[t] = (unsigned)cap[t];
      off += cap[t];
    }
  }
  e->n_work = off;
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  const size_t nc = kListsPerClass * e->cls.size();
  e->slots.alloc(sizeof(EnsSlot) * e->n);
  e->counts.alloc(sizeof(int) * nc);
  e->work.alloc(sizeof(EnsWork) * std::max<long>(1, e->n_work));
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&e->slots_h), sizeof(EnsSlot) * e->n, hipHostMallocDefault),
            "hipHostMalloc");
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&e->counts_h), sizeof(int) * nc, hipHostMallocDefault),
            "hipHostMalloc");
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&e->work_h), sizeof(EnsWork) * std::max<long>(1, e->n_work),
                          hipHostMallocDefault),
            "hipHostMalloc");
  hip_check(hipEventCreateWithFlags(&e->staged, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipStreamCreateWithFlags(&e->up, hipStreamNonBlocking), "hipStreamCreate");
  for (EnsClass& C : e->cls) {
    hip_check(hipStreamCreateWithFlags(&C.st, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&C.done, hipEventDisableTiming), "hipEventCreate");
  }
  return e.release();
}

hipStream_t ens_part_stream(EnsSweep* e, int k) {
  require(e != nullptr && k >= 0 && k < e->n, CFFT_ERR_INVALID, "part index out of range");
  return e->cls[e->cls_of[k]].st;
}

void ens_sweep_run(EnsSweep* e, const std::vector<EnsPart>& parts, const std::vector<std::vector<char>>& active) {
  require(e != nullptr, CFFT_ERR_INVALID, "null ensemble sweep");
  require((int)parts.size() == e->n && (int)active.size() == e->n, CFFT_ERR_INVALID, "one entry per plan");
  for (int k = 0; k < e->n; ++k)
    require((int)active[k].size() == e->info[k].batch, CFFT_ERR_INVALID, "one flag per batch entry");
  // the staging buffers are free once the previous upload has completed
  hip_check(hipEventSynchronize(e->staged), "hipEventSynchronize");
  for (int k = 0; k < e->n; ++k) {
    require(parts[k].plan != nullptr, CFFT_ERR_INVALID, "null plan");
    ens_slot_params(parts[k].plan, parts[k], e->slots_h + k);
  }
  // the work lists: every item of every active field (items are field-major:
  // field b of a plan owns items b * per_field .. (b + 1) * per_field - 1)
  std::vector<char> run(e->cls.size(), 0);
  for (size_t c = 0; c < e->cls.size(); ++c) {
    const EnsClass& C = e->cls[c];
    int* cnt = e->counts_h + kListsPerClass * c;
    for (int t = 0; t < kListsPerClass; ++t) cnt[t] = 0;
    for (int k : C.parts) {
      const EnsPlanInfo& in = e->info[k];
      const int lp = in.sep ? kListSep : kListGen;
      const long pp = in.pair_items / in.batch, pt = in.tile_items / in.batch;
      for (int b = 0; b < in.batch; ++b) {
        if (!active[k][b]) continue;
        run[c] = 1;
        EnsWork* wp = e->work_h + C.off[lp] + cnt[lp];
        for (long j = 0; j < pp; ++j) wp[j] = {k, (int)(b * pp + j)};
        cnt[lp] += (int)pp;
        EnsWork* wt = e->work_h + C.off[kListTile] + cnt[kListTile];
        for (long j = 0; j < pt; ++j) wt[j] = {k, (int)(b * pt + j)};
        cnt[kListTile] += (int)pt;
      }
    }
  }
  const size_t nc = kListsPerClass * e->cls.size();
  // the device tables are rewritten only after every class's previous sweep
  // has read them
  for (EnsClass& C : e->cls) hip_check(hipStreamWaitEvent(e->up, C.done, 0), "hipStreamWaitEvent");
  hip_check(hipMemcpyAsync(e->slots.p, e->slots_h, sizeof(EnsSlot) * e->n, hipMemcpyHostToDevice, e->up), "upload");
  hip_check(hipMemcpyAsync(e->counts.p, e->counts_h, sizeof(int) * nc, hipMemcpyHostToDevice, e->up), "upload");
  for (size_t c = 0; c < e->cls.size(); ++c) {
    // each list's filled part
    const EnsClass& C = e->cls[c];
    for (int t = 0; t < kListsPerClass; ++t) {
      const int m = e->counts_h[kListsPerClass * c + t];
      if (m > 0)
        hip_check(hipMemcpyAsync(static_cast<EnsWork*>(e->work.p) + C.off[t], e->work_h + C.off[t],
                                 sizeof(EnsWork) * m, hipMemcpyHostToDevice, e->up),
                  "upload");
    }
  }
  hip_check(hipEventRecord(e->staged, e->up), "hipEventRecord");
  const EnsSlot* S = static_cast<const EnsSlot*>(e->slots.p);
  for (size_t c = 0; c < e->cls.size(); ++c) {
    EnsClass& C = e->cls[c];
    if (!run[c]) continue;
    if (!C.exec) {
      // kernel arguments are table addresses and the step: the graph stays
      // valid while the tables' contents change between runs
      std::lock_guard<std::recursive_mutex> guard(setup_mutex());
      if (!e->cap) hip_check(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking), "hipStreamCreate");
      ClassLists cl;
      cl.count = static_cast<const int*>(e->counts.p) + kListsPerClass * c;
      cl.work = static_cast<const EnsWork*>(e->work.p);
      for (int t = 0; t < kListsPerClass; ++t) {
        cl.off[t] = C.off[t];
        cl.cap[t] = C.cap[t];
      }
      hipGraph_t g = nullptr;
      hip_check(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        if (e->prec == CFFT_FP64)
          capture_class_len<double>(C.L, S, cl, e->ns, e->cap);
        else
          capture_class_len<float>(C.L, S, cl, e->ns, e->cap);
      } catch (...) {
        (void)hipStreamEndCapture(e->cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(e->cap, &g), "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_check(err, "hipGraphInstantiate");
      C.exec = exec;
    }
    // the class's slice-0 work is already on its stream; the tables first
    hip_check(hipStreamWaitEvent(C.st, e->staged, 0), "hipStreamWaitEvent");
    hip_check(hipGraphLaunch(C.exec, C.st), "hipGraphLaunch");
    hip_check(hipEventRecord(C.done, C.st), "hipEventRecord");
  }
}

void ens_sweep_sync(EnsSweep* e) {
  require(e != nullptr, CFFT_ERR_INVALID, "null ensemble sweep");
  for (EnsClass& C : e->cls) hip_check(hipStreamSynchronize(C.st), "sync");
}

void ens_sweep_destroy(EnsSweep* e) {
  std::lock_guard<std::recursive_mutex> guard(setup_mutex());
  delete e;
}

}  // namespace cfft
