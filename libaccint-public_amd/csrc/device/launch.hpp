// launch.hpp -- per-class launcher table shared by the instantiation units and
// the C-ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

namespace lai {

struct EriArgs;
struct GradArgs;

using ClassLauncher = void (*)(const EriArgs&, hipStream_t, int max_blocks);

struct ClassInfo {
  int La, Lb, Lc, Ld;
  int regs;    // 1: register shape (Kernel A), 0: LDS shape (Kernel B)
  int nchunk;  // Kernel A chunks along the bra index
  int nt;      // lanes per workgroup
  int waves;   // LDS shape with one quartet per wavefront: wavefronts per workgroup, else 0
};

#define LAI_NPARTS_FIXED 8
void fill_launchers_part_0(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_1(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_2(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_3(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_4(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_5(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_6(ClassLauncher*, ClassLauncher*, ClassInfo*);
void fill_launchers_part_7(ClassLauncher*, ClassLauncher*, ClassInfo*);

// gradient kernels exist for the classes whose shells are at most d
inline constexpr int kNumGradClasses = 21;
using GradLauncher = void (*)(const GradArgs&, hipStream_t, int max_blocks);
void fill_grad_launchers(GradLauncher*);

}  // namespace lai
