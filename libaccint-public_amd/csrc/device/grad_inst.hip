// grad_inst.hip -- instantiates the gradient kernels for the classes with
// every shell at most d (quartet classes 0..20).
#include "grad_kernels.hpp"
#include "launch.hpp"

namespace lai {
namespace {

// per-workgroup force accumulators in LDS up to this many atoms
constexpr int kGradLdsAtoms = 2048;

template <int ID>
void launch_grad(const GradArgs& a0, hipStream_t s, int max_blocks) {
  constexpr int pb = cls_bra(ID), pk = cls_ket(ID);
  constexpr int La = pc_La(pb), Lb = pc_Lb(pb), Lc = pc_La(pk), Ld = pc_Lb(pk);
  using P = GradPolicy<La, Lb, Lc, Ld>;
  if (a0.nchunks <= 0) return;
  GradArgs a = a0;
  a.glds = a.n_atoms <= kGradLdsAtoms ? 1 : 0;
  const size_t dyn = a.glds ? size_t(3) * a.n_atoms * sizeof(double) : 0;
  if constexpr (P::regs) {
    const int blocks = a.nchunks < max_blocks ? a.nchunks : max_blocks;
    hipLaunchKernelGGL((grad_kernel_regs<La, Lb, Lc, Ld>), dim3(blocks), dim3(256), dyn, s, a);
  } else {
    // the worklist chunks follow the J/K kernel shape of the class: long ones
    // for its register shape are split over gridDim.y workgroups
    constexpr int split = ClassPolicy<La, Lb, Lc, Ld>::regs ? 64 : 1;
    const int blocks = a.nchunks < max_blocks ? a.nchunks : max_blocks;
    hipLaunchKernelGGL((grad_kernel<La, Lb, Lc, Ld, P::NT>), dim3(blocks, split), dim3(P::NT), dyn, s, a);
  }
}

template <int ID>
void fill(GradLauncher* f) {
  f[ID] = &launch_grad<ID>;
  if constexpr (ID + 1 < kNumGradClasses) fill<ID + 1>(f);
}

}  // namespace

void fill_grad_launchers(GradLauncher* f) { fill<0>(f); }

}  // namespace lai
