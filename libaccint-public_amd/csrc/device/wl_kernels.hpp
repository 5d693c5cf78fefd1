// wl_kernels.hpp -- device passes over the implicit worklist: per-class
// statistics (quartet and ERI counts, the order-independent list checksum,
// the FLOP model of the reference recursion) with the same per-quartet
// selection the J/K kernels apply, and the batched J/K symmetrisation.
#pragma once

#include "eri_kernels.hpp"

namespace lai {

struct WlStatsArgs {
  const JkChunk* __restrict__ chunks;
  int nchunks;
  const JkSegs* __restrict__ segs;
  const int* __restrict__ kets;
  const PairRec* __restrict__ pairs;
  QuartetSel sel;
  const int* __restrict__ L;      // angular momentum per shell
  const double* __restrict__ fp;  // F_prim[La][Lb][Lc][Ld], 4^4 entries, reference orientation
  int cls;
  // outputs, accumulated: per class [count, fp32 count, ERIs] and
  // [model flops FP64, FP32, executed flops FP64, FP32]; the checksum
  unsigned long long* counts;  // [kNumClasses][3]
  double* flops;               // [kNumClasses][4]
  unsigned long long* checksum;
  // gradient work per primitive quartet, two models ([La][Lb][Lc][Ld], 4^4
  // each), and the per-class sums over the executed primitive pairs of every
  // selected quartet ([kNumClasses][2])
  const double* __restrict__ fg;
  double* gflops;
};

__global__ void worklist_stats_kernel(WlStatsArgs a);

// J[m] += (Jt[m] + Jt[m]^T)/2, K[m] likewise, for every molecule block m
// (blockIdx.y = molecule); with anti != 0 (an antisymmetric density) only
// K[m] += (Kt[m] - Kt[m]^T)/2
__global__ void symmetrize_add_batch_kernel(const double* __restrict__ Jt, const double* __restrict__ Kt,
                                            double* __restrict__ J, double* __restrict__ K,
                                            const long long* __restrict__ moff, const int* __restrict__ mnbf,
                                            int anti);

}  // namespace lai
