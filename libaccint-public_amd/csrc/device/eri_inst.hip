// eri_inst.hip -- instantiates the per-class ERI/J-K kernels.  Compiled
// LAI_NPARTS times with -DLAI_PART=k so the 55 classes x 2 precisions build in
// parallel; part k registers the launchers of the classes with id % LAI_NPARTS == k.
#include "eri_kernels.hpp"
#include "launch.hpp"

#ifndef LAI_PART
#error "compile with -DLAI_PART=<k> -DLAI_NPARTS=<n>"
#endif

namespace lai {
namespace {

// LDS a workgroup may use for the anchored exchange rows: register-shape
// kernels keep several workgroups per CU; the LDS-shape kernels may take the
// whole 160 KB of gfx950
constexpr size_t kLdsBudget = 64 * 1024;
constexpr size_t kLdsBudgetB = 160 * 1024;

// static LDS of a kernel as compiled (queried once per instantiation)
template <typename K>
size_t static_lds(K kern) {
  hipFuncAttributes at{};
  if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kern)) != hipSuccess) return kLdsBudgetB;
  return at.sharedSizeBytes;
}

// Exchange rows in LDS when static + dynamic LDS fit the budget; above 64 KB
// the kernel is allowed the larger dynamic size first.
template <typename K>
int use_rows(K kern, size_t stat, size_t dyn, size_t budget) {
  if (stat + dyn > budget) return 0;
  if (stat + dyn > 64 * 1024 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                          int(dyn)) != hipSuccess)
    return 0;
  return 1;
}

// MIXED = false: the recursion in FP64 (and the raw-ERI mode); true: in
// float, launched with QuartetSel::keep == 2 so that it takes only the
// quartets the worklist marks for FP32.
template <int ID, bool MIXED>
void launch_class(const EriArgs& a0, hipStream_t s, const LaunchTune& t) {
  constexpr int pb = cls_bra(ID), pk = cls_ket(ID);
  constexpr int La = pc_La(pb), Lb = pc_Lb(pb), Lc = pc_La(pk), Ld = pc_Lb(pk);
  using P = ClassPolicy<La, Lb, Lc, Ld>;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  EriArgs a = a0;
  if (a.mode == 1) {  // raw ERIs, one launch over the quartet list
    if (MIXED || a.nq <= 0) return;
    const int max_blocks = t.n_cu * 16;  // grid-stride over the quartet list
    if constexpr (P::regs) {
      int blocks = (a.nq + 127) / 128;
      if (blocks > max_blocks) blocks = max_blocks;
      dim3 grid(blocks, P::NCH);
      hipLaunchKernelGGL((eri_kernel_regs<La, Lb, Lc, Ld, double, P::CHA>), grid, dim3(128), 0, s, a);
    } else {
      int blocks = a.nq < 8 * max_blocks ? a.nq : 8 * max_blocks;
      hipLaunchKernelGGL((eri_kernel_lds<La, Lb, Lc, Ld, double, P::NT>), dim3(blocks), dim3(P::NT), 0, s, a);
    }
    return;
  }
  // J/K digestion over the bra-anchored chunks
  if (a.nchunks <= 0) return;
  const size_t nbf = size_t(a.nbf);
  const int shape = (t.alt && P::alt >= 0) ? P::alt : P::shape;
  auto grid_cap = [&](int dflt) {
    const long long cap = (long long)t.n_cu * (t.blocks_per_cu > 0 ? t.blocks_per_cu : dflt);
    return int(a.nchunks < cap ? a.nchunks : cap);
  };
  if (shape == kShapeRegs) {
    if constexpr (P::regs && P::NCH > 1) {
      // bra cut into NCH slices: one wavefront per slice, all rows in LDS
      auto kern = jk_kernel_merged<La, Lb, Lc, Ld, MIXED, P::CHA>;
      static const size_t stat = static_lds(kern);
      const size_t dyn = size_t(na + nb) * nbf * sizeof(double);
      a.klds = use_rows(kern, stat, dyn, kLdsBudgetB);
      hipLaunchKernelGGL(kern, dim3(grid_cap(kBlocksPerCuRegs)), dim3(64 * P::NCH), a.klds ? dyn : 0, s, a);
    } else if constexpr (P::regs) {
      const size_t dyn = size_t(P::CHA + nb) * nbf * sizeof(double);
      a.klds = use_rows(jk_kernel_regs<La, Lb, Lc, Ld, MIXED, P::CHA>, 0, dyn, kLdsBudget);
      dim3 grid(grid_cap(kBlocksPerCuRegs), P::NCH);
      hipLaunchKernelGGL((jk_kernel_regs<La, Lb, Lc, Ld, MIXED, P::CHA>), grid, dim3(256), a.klds ? dyn : 0, s, a);
    }
  } else if (shape == kShapeWaves) {
    if constexpr (P::waves || P::alt == kShapeWaves) {
      auto kern = jk_kernel_waves<La, Lb, Lc, Ld, MIXED, P::NW>;
      static const size_t stat = static_lds(kern);
      const size_t dyn = (size_t(na + nb) * nbf + na * nb) * sizeof(double);
      a.klds = use_rows(kern, stat, dyn, kLdsBudgetB);  // else per-quartet blocks to HBM
      hipLaunchKernelGGL(kern, dim3(grid_cap(kBlocksPerCuLds)), dim3(64 * P::NW), a.klds ? dyn : 0, s, a);
    }
  } else {
    if constexpr (!P::regs) {
      auto kern = jk_kernel_lds<La, Lb, Lc, Ld, MIXED, P::NT>;
      static const size_t stat = static_lds(kern);
      const size_t dyn = (size_t(na + nb) * nbf + nc * nd + na * nb) * sizeof(double);
      a.klds = use_rows(kern, stat, dyn, kLdsBudgetB);  // else per-quartet blocks to HBM
      hipLaunchKernelGGL(kern, dim3(grid_cap(kBlocksPerCuLds)), dim3(P::NT), a.klds ? dyn : 0, s, a);
    }
  }
}

template <int ID>
void fill(ClassLauncher* f64, ClassLauncher* f32, ClassInfo* info) {
  if constexpr (ID % LAI_NPARTS == LAI_PART) {
    constexpr int pb = cls_bra(ID), pk = cls_ket(ID);
    constexpr int La = pc_La(pb), Lb = pc_Lb(pb), Lc = pc_La(pk), Ld = pc_Lb(pk);
    using P = ClassPolicy<La, Lb, Lc, Ld>;
    f64[ID] = &launch_class<ID, false>;
    f32[ID] = &launch_class<ID, true>;
    info[ID].La = La;
    info[ID].Lb = Lb;
    info[ID].Lc = Lc;
    info[ID].Ld = Ld;
    info[ID].regs = P::regs ? 1 : 0;
    info[ID].nchunk = P::regs ? P::NCH : 0;
    info[ID].nt = P::regs ? (P::NCH > 1 ? 64 * P::NCH : 256) : (P::waves ? 64 * P::NW : P::NT);
    info[ID].waves = P::waves ? P::NW : 0;
    info[ID].shape = P::shape;
    info[ID].alt = P::alt;
  }
  if constexpr (ID + 1 < kNumClasses) fill<ID + 1>(f64, f32, info);
}

}  // namespace

#define LAI_CAT2(a, b) a##b
#define LAI_CAT(a, b) LAI_CAT2(a, b)
void LAI_CAT(fill_launchers_part_, LAI_PART)(ClassLauncher* f64, ClassLauncher* f32,
                                              ClassInfo* info) {
  fill<0>(f64, f32, info);
}

}  // namespace lai
