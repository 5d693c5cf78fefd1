// persist.hip -- one persistent J/K launch for the register-shape classes.
//
// Instead of one launch per angular-momentum class, a grid sized to fill the
// GPU once takes work items from a device queue: the bra-anchored chunks of
// every covered class, sorted by class (heaviest classes first).  Each
// workgroup claims the next item with one atomic on a device counter and runs
// that class's chunk code; it leaves when the queue is drained.  The queue
// replaces the per-ShellSetQuartet launches of the reference's CUDA engine
// (src/device/cuda/engine/cuda_engine.cu:562-665) with a single launch.
//
// Covered: the lane-per-quartet classes whose chunk code runs in a 256-lane
// workgroup -- those with the bra unsplit (256 lanes, one quartet per lane)
// or split into 2 or 4 slices (groups of 2 or 4 wavefronts, kernel A
// merged) -- at the base register budget of 27 outputs per lane.  One kernel
// holds all their code, so it runs at the register and LDS footprint of the
// largest of them.
#include <algorithm>

#include "eri_kernels.hpp"
#include "launch.hpp"

namespace lai {
namespace {

constexpr int kPersistNT = 256;

template <int ID>
struct PersistClass {
  static constexpr int pb = cls_bra(ID), pk = cls_ket(ID);
  static constexpr int La = pc_La(pb), Lb = pc_Lb(pb), Lc = pc_La(pk), Ld = pc_Lb(pk);
  using P = ClassPolicy<La, Lb, Lc, Ld>;
  // (classes whose slices take the larger register budget keep their own
  // launches: their footprint would set the one kernel's)
  static constexpr bool covered = P::regs && P::CHA == P::CHA0 && (P::NCH == 1 || P::NCH == 2 || P::NCH == 4);
  // dynamic LDS in doubles: the merged slices' ket Coulomb parts, then the
  // exchange rows (na + nb rows of nbf)
  static constexpr int jpart = P::NCH > 1 ? (kPersistNT / 64) * ncart(Lc) * ncart(Ld) * 64 : 0;
  static constexpr int rows = ncart(La) + ncart(Lb);
};

// Runs chunk ch of class cls (a covered class).
template <int ID>
__device__ __forceinline__ void persist_chunk(const EriArgs& g, int cls, const JkChunk& ch, double* dyn,
                                              SortLds& srt) {
  if constexpr (ID < kNumClasses) {
    using C = PersistClass<ID>;
    if constexpr (C::covered) {
      if (cls == ID) {
        if constexpr (C::P::NCH == 1) {
          jk_chunk_regs<C::La, C::Lb, C::Lc, C::Ld, false, 0, ncart(C::La)>(g, ch, dyn, srt);
        } else {
          dispatch_merged<C::La, C::Lb, C::Lc, C::Ld, false, C::P::CHA, 0>(g, ch, dyn + C::jpart, srt, dyn,
                                                                         (threadIdx.x >> 6) % C::P::NCH);
        }
        return;
      }
    }
    persist_chunk<ID + 1>(g, cls, ch, dyn, srt);
  }
}

// Every workgroup claims items until the queue is drained: the counter only
// grows, so each workgroup reaches an index >= n_items and leaves.
__global__ __launch_bounds__(kPersistNT) void jk_kernel_persistent(EriArgs g, const PersistItem* __restrict__ items,
                                                                   int n_items, int* __restrict__ counter) {
  extern __shared__ double dyn[];
  __shared__ SortLds srt;
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1);
    __syncthreads();
    const int it = next;
    __syncthreads();  // every lane has read next before lane 0 claims again
    if (it >= n_items) break;
    const PersistItem w = items[it];
    persist_chunk<0>(g, w.cls, w.ch, dyn, srt);
  }
}

__global__ void zero_counter(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

template <int ID>
constexpr size_t max_dyn_doubles(size_t nbf, bool rows) {
  if constexpr (ID < kNumClasses) {
    using C = PersistClass<ID>;
    const size_t here = C::covered ? size_t(C::jpart) + (rows ? size_t(C::rows) * nbf : 0) : 0;
    const size_t rest = max_dyn_doubles<ID + 1>(nbf, rows);
    return here > rest ? here : rest;
  } else {
    return 0;
  }
}

template <int ID>
constexpr bool covered_at(int c) {
  if constexpr (ID < kNumClasses)
    return c == ID ? PersistClass<ID>::covered : covered_at<ID + 1>(c);
  else
    return false;
}

}  // namespace

bool persist_covers(int cls) { return covered_at<0>(cls); }

int launch_persistent(const EriArgs& a0, const PersistItem* items, int n_items, int* counter, hipStream_t s,
                      int n_cu) {
  if (n_items <= 0) return 0;
  EriArgs a = a0;
  const size_t nbf = size_t(a.nbf);
  static const size_t stat = [] {
    hipFuncAttributes at{};
    return hipFuncGetAttributes(&at, reinterpret_cast<const void*>(jk_kernel_persistent)) == hipSuccess
               ? at.sharedSizeBytes
               : size_t(160 * 1024);
  }();
  static const size_t dmax = [] {
    const size_t want = stat < 160 * 1024 ? 160 * 1024 - stat : 0;
    if (want > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(jk_kernel_persistent),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) != hipSuccess)
      return stat < 64 * 1024 ? 64 * 1024 - stat : size_t(0);
    return want;
  }();
  // the exchange rows of the widest covered class in LDS when they fit
  size_t dyn = max_dyn_doubles<0>(nbf, true) * sizeof(double);
  a.klds = 1;
  if (dyn > dmax) {
    dyn = max_dyn_doubles<0>(nbf, false) * sizeof(double);
    a.klds = 0;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(jk_kernel_persistent),
                                                   kPersistNT, dyn) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int grid = std::min(n_items, per_cu * n_cu);
  hipLaunchKernelGGL(zero_counter, dim3(1), dim3(64), 0, s, counter);
  hipLaunchKernelGGL(jk_kernel_persistent, dim3(grid), dim3(kPersistNT), dyn, s, a, items, n_items, counter);
  return 0;
}

}  // namespace lai
