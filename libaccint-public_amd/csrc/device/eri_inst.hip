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

// LDS a workgroup may use (gfx950: 160 KB per CU)
constexpr size_t kLdsBudget = 160 * 1024;

// static LDS of a kernel as compiled (queried once per instantiation)
template <typename K>
size_t static_lds(K kern) {
  hipFuncAttributes at{};
  if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kern)) != hipSuccess) return kLdsBudget;
  return at.sharedSizeBytes;
}

// Dynamic LDS the kernel may be launched with: all of the budget beyond its
// static LDS when the runtime allows more than the 64 KB default.
template <typename K>
size_t dyn_limit(K kern, size_t stat) {
  const size_t want = stat < kLdsBudget ? kLdsBudget - stat : 0;
  if (want > 64 * 1024 && hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, int(want)) != hipSuccess)
    return stat < 64 * 1024 ? 64 * 1024 - stat : 0;
  return want;
}

// Workgroup of a lane-per-quartet kernel: rep replicas of `base` lanes (at
// most 512 lanes) share one set of the bra's exchange rows (`rows` bytes of
// dynamic LDS) and `fixed` bytes, and carry `per_rep` bytes each.  Picks the
// rep that keeps the most wavefronts resident per CU, LDS and registers
// counted by the runtime's occupancy calculator; the rows stay in LDS whenever
// they fit at all (without them every exchange update is an HBM atomic).
struct RowsPick {
  int rep = 1;
  int rows = 0;    // 1: exchange rows in LDS
  size_t dyn = 0;  // dynamic LDS bytes
  int waves = -1;  // resident wavefronts per CU (-1: nothing fits)
};
template <typename K>
RowsPick pick_rows(K kern, size_t dmax, int base, size_t per_rep, size_t rows, size_t fixed = 0) {
  RowsPick p;
  p.dyn = fixed + per_rep;
  int best = -1;
  for (int pass = 0; pass < 2 && best < 0; ++pass) {  // with the rows, else without
    const size_t rb = pass == 0 ? rows : 0;
    for (int r = 1; base * r <= 512; r *= 2) {
      const size_t dyn = fixed + size_t(r) * per_rep + rb;
      if (dyn > dmax) break;
      int blocks = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(kern), base * r,
                                                       dyn) != hipSuccess)
        blocks = 0;
      const int waves = blocks * base * r / 64;
      if (blocks > 0 && waves > best) {
        best = waves;
        p.rep = r;
        p.rows = pass == 0 ? 1 : 0;
        p.dyn = dyn;
        p.waves = waves;
      }
    }
  }
  return p;
}

// MIXED = false: the recursion in FP64 (and the raw-ERI mode); true: in
// float, launched with QuartetSel::keep == 2 so that it takes only the
// quartets the worklist marks for FP32 (wavefront and workgroup shapes only;
// the host never sends a lane-per-quartet shape here).
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
    // (mixed builds run the lane-per-quartet classes in FP64: no float
    // instantiation of these kernels)
    if constexpr (MIXED) {
    } else if constexpr (P::regs && P::NCH > 1) {
      // bra cut into NCH slices: groups of NCH wavefronts, one per slice
      auto kern = jk_kernel_merged<La, Lb, Lc, Ld, MIXED, P::CHA, false>;
      static const size_t stat = static_lds(kern);
      static const size_t dmax = dyn_limit(kern, stat);
      const size_t per_rep = size_t(P::NCH) * nc * nd * 64 * sizeof(double);
      const size_t rows = size_t(na + nb) * nbf * sizeof(double);
      RowsPick rp = pick_rows(kern, dmax, 64 * P::NCH, per_rep, rows);
      // the Rys tables in LDS as well (one to four roots) when that keeps as
      // many wavefronts resident and the exchange rows still in LDS
      if constexpr (kLdsRysDoubles<La, Lb, Lc, Ld> > 0) {
        auto kl = jk_kernel_merged<La, Lb, Lc, Ld, MIXED, P::CHA, true>;
        static const size_t stat_l = static_lds(kl);
        static const size_t dmax_l = dyn_limit(kl, stat_l);
        const RowsPick rl =
            pick_rows(kl, dmax_l, 64 * P::NCH, per_rep, rows, size_t(kLdsRysDoubles<La, Lb, Lc, Ld>) * sizeof(double));
        if (rl.rows >= rp.rows && rl.waves >= rp.waves) {
          a.klds = rl.rows;
          hipLaunchKernelGGL(kl, dim3(grid_cap(kBlocksPerCuRegs)), dim3(64 * P::NCH * rl.rep), rl.dyn, s, a);
          return;
        }
      }
      a.klds = rp.rows;
      hipLaunchKernelGGL(kern, dim3(grid_cap(kBlocksPerCuRegs)), dim3(64 * P::NCH * rp.rep), rp.dyn, s, a);
    } else if constexpr (P::regs) {
      const size_t rows = size_t(P::CHA + nb) * nbf * sizeof(double);
      auto kern = jk_kernel_regs<La, Lb, Lc, Ld, MIXED, P::CHA, false>;
      static const size_t stat = static_lds(kern);
      static const size_t dmax = dyn_limit(kern, stat);
      RowsPick rp = pick_rows(kern, dmax, 256, 0, rows);
      bool lb = false;
      // one- to four-root classes: the Rys tables (Boys Taylor rows or
      // Chebyshev coefficients) in LDS as well, when that keeps as many
      // wavefronts resident with the exchange rows in LDS and at least two
      // workgroups per CU (LDS reads instead of L2 gathers in the primitive
      // loop; one workgroup per CU was measured slower)
      if constexpr (kLdsRysDoubles<La, Lb, Lc, Ld> > 0) {
        auto kl = jk_kernel_regs<La, Lb, Lc, Ld, MIXED, P::CHA, true>;
        static const size_t stat_l = static_lds(kl);
        static const size_t dmax_l = dyn_limit(kl, stat_l);
        const RowsPick rl =
            pick_rows(kl, dmax_l, 256, 0, rows, size_t(kLdsRysDoubles<La, Lb, Lc, Ld>) * sizeof(double));
        if (rl.rows >= rp.rows && rl.waves >= rp.waves && rl.waves >= 2 * 4 * rl.rep) {
          rp = rl;
          lb = true;
        }
      }
      a.klds = rp.rows;
      dim3 grid(grid_cap(kBlocksPerCuRegs), P::NCH);
      if constexpr (kLdsRysDoubles<La, Lb, Lc, Ld> > 0) {
        if (lb) {
          hipLaunchKernelGGL((jk_kernel_regs<La, Lb, Lc, Ld, MIXED, P::CHA, true>), grid, dim3(256 * rp.rep), rp.dyn,
                             s, a);
          return;
        }
      }
      hipLaunchKernelGGL(kern, grid, dim3(256 * rp.rep), rp.dyn, s, a);
    }
  } else if (shape == kShapeWaves) {
    if constexpr (P::waves || P::alt == kShapeWaves) {
      a.klds = 0;
      hipLaunchKernelGGL((jk_kernel_waves<La, Lb, Lc, Ld, MIXED, P::NW>), dim3(grid_cap(kBlocksPerCuLds)),
                         dim3(64 * P::NW), 0, s, a);
    }
  } else {
    if constexpr (!P::regs) {
      a.klds = 0;
      hipLaunchKernelGGL((jk_kernel_lds<La, Lb, Lc, Ld, MIXED, P::NT>), dim3(grid_cap(kBlocksPerCuLds)),
                         dim3(P::NT), 0, s, a);
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
