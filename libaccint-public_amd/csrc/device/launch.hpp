// launch.hpp -- per-class launcher table shared by the instantiation units and
// the C-ABI implementation.
#pragma once

#include <hip/hip_runtime.h>

namespace lai {

struct EriArgs;
struct GradArgs;

// Kernel shapes of the J/K build: a lane per quartet (register tables), a
// wavefront per quartet, a workgroup per quartet (LDS tables).
enum : int { kShapeNone = -1, kShapeRegs = 0, kShapeWaves = 1, kShapeLds = 2 };

// Launch parameters of one class: which of its compiled shapes runs and the
// grid cap in workgroups per CU (0: the shape's default).
struct LaunchTune {
  int alt = 0;            // 0: the class's primary shape, 1: its alternative
  int blocks_per_cu = 0;  // grid = min(chunks, n_cu * blocks_per_cu)
  int n_cu = 256;
};

using ClassLauncher = void (*)(const EriArgs&, hipStream_t, const LaunchTune&);

// default grid caps, workgroups per CU: with up to 64 per CU a lane-per-quartet
// launch has one workgroup per chunk for all but the largest classes, which
// leaves the load balance to the hardware dispatcher
inline constexpr int kBlocksPerCuRegs = 64;
inline constexpr int kBlocksPerCuLds = 128;

struct ClassInfo {
  int La, Lb, Lc, Ld;
  int regs;    // 1: register shape (Kernel A), 0: LDS shape (Kernel B)
  int nchunk;  // Kernel A chunks along the bra index
  int nt;      // lanes per workgroup (lane-per-quartet shape: the unit a launch replicates, up to 512 lanes)
  int waves;   // LDS shape with one quartet per wavefront: wavefronts per workgroup, else 0
  int shape;   // primary shape (kShape*)
  int alt;     // alternative shape compiled for the autotuner, kShapeNone if none
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

// the persistent launch (persist.hip): whether class cls is among the classes
// it covers, and the launch over n_items queue items (counter: one device int,
// cleared on the stream first); returns 0
struct PersistItem;
bool persist_covers(int cls);
int launch_persistent(const EriArgs& a, const PersistItem* items, int n_items, int* counter, hipStream_t s,
                      int n_cu);

// gradient kernels exist for the classes whose shells are at most d
inline constexpr int kNumGradClasses = 21;
using GradLauncher = void (*)(const GradArgs&, hipStream_t, int max_blocks);
void fill_grad_launchers(GradLauncher*);

}  // namespace lai
