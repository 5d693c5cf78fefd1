// common.hpp -- shared device-side definitions for the gfx950 Rys ERI / J-K
// backend: Cartesian component tables, the HBM data layout of pairs and
// primitive pairs, Rys root/weight evaluation and atomics.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rys_table.inc"

namespace lai {

// ---------------------------------------------------------------------------
// Cartesian components (order of include/libaccint/math/cartesian_indices.hpp:
// lx descending, then ly descending)
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

__host__ __device__ constexpr int cart_comp(int L, int i, int axis) {
  int c = 0;
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) {
      if (c == i) return axis == 0 ? lx : (axis == 1 ? ly : L - lx - ly);
      ++c;
    }
  return 0;
}

// 1/sqrt((2lx-1)!!(2ly-1)!!(2lz-1)!!): per-component normalisation relative
// to the (L,0,0) component (include/libaccint/kernels/norm_correction.hpp:28-34)
__host__ __device__ constexpr int dfact_odd(int n) {
  int r = 1;
  for (int i = 1; i <= n; ++i) r *= (2 * i - 1);
  return r;
}
__host__ __device__ inline double norm_corr(int L, int i) {
  return 1.0 / sqrt(double(dfact_odd(cart_comp(L, i, 0)) * dfact_odd(cart_comp(L, i, 1)) *
                           dfact_odd(cart_comp(L, i, 2))));
}

// ---------------------------------------------------------------------------
// Angular-momentum classes.  Pair classes (La >= Lb) are numbered
// (0,0)=0 (1,0)=1 (1,1)=2 (2,0)=3 (2,1)=4 (2,2)=5 (3,0)=6 (3,1)=7 (3,2)=8 (3,3)=9;
// a quartet class holds bra pair class >= ket pair class and is numbered
// pb*(pb+1)/2 + pk.
// ---------------------------------------------------------------------------
inline constexpr int kMaxL = 3;
inline constexpr int kNumPairClasses = (kMaxL + 1) * (kMaxL + 2) / 2;
inline constexpr int kNumClasses = kNumPairClasses * (kNumPairClasses + 1) / 2;

__host__ __device__ constexpr int pair_class(int La, int Lb) { return La * (La + 1) / 2 + Lb; }
__host__ __device__ constexpr int pc_La(int p) {
  int L = 0;
  while ((L + 1) * (L + 2) / 2 <= p) ++L;
  return L;
}
__host__ __device__ constexpr int pc_Lb(int p) { return p - pc_La(p) * (pc_La(p) + 1) / 2; }
__host__ __device__ constexpr int cls_bra(int id) {
  int p = 0;
  while ((p + 1) * (p + 2) / 2 <= id) ++p;
  return p;
}
__host__ __device__ constexpr int cls_ket(int id) { return id - cls_bra(id) * (cls_bra(id) + 1) / 2; }

// ---------------------------------------------------------------------------
// HBM layout
// ---------------------------------------------------------------------------
// One significant shell pair in "class order": shell a has La >= Lb of shell
// b.  The same record serves as bra or ket.
struct PairRec {
  int sa, sb;         // shell indices (class order)
  int fa, fb;         // first basis function of each shell
  int pp_off;         // first primitive pair in the PrimPair array
  int npp;            // number of primitive pairs (zero coefficients dropped)
  int same;           // 1 when sa == sb
  int npp_sig;        // leading primitive pairs used by the J/K and gradient kernels
                      // (the rest are below the primitive screening bound)
  double ABx, ABy, ABz;  // A - B (horizontal transfer)
  double pad2;
};

// One primitive pair, 64 bytes (one cache line sector per lane):
// zeta, P, P - A and the folded coefficient
//   Kc = c_a c_b exp(-mu |AB|^2) * sqrt(2 pi^{5/2}) / zeta
// so that the ERI prefactor 2 pi^{5/2} K_AB K_CD/(zeta eta sqrt(zeta+eta))
// c_a c_b c_c c_d equals Kc_ab * Kc_cd / sqrt(zeta + eta).
struct PrimPair {
  double zeta, Px, Py, Pz;
  double PAx, PAy, PAz, Kc;
};

// Work item of a quartet class: bra and ket pair indices.
struct Quartet {
  int bra, ket;
};

// ---------------------------------------------------------------------------
// Rys roots/weights from the piecewise-Chebyshev table (device copy of
// lai_rys_cheb_h), large-T Gauss-Hermite limit beyond T_cut(n).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void rys_roots(double T, const double* __restrict__ tab, double* u,
                                          double* w) {
  static_assert(N >= 1 && N <= LAI_RYS_NMAX, "Rys order out of table range");
  if (T < double(lai_rys_tcut[N])) {
    const int k = int(T);
    const double x = 2.0 * (T - double(k)) - 1.0;
    const double x2 = 2.0 * x;
    const double* c = tab + lai_rys_off[N] + k * (2 * N * LAI_RYS_NTERMS);
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
      const double* cf = c + f * LAI_RYS_NTERMS;
      double d1 = 0.0, d2 = 0.0;
#pragma unroll
      for (int j = LAI_RYS_NTERMS - 1; j >= 1; --j) {
        const double dk = fma(x2, d1, cf[j] - d2);
        d2 = d1;
        d1 = dk;
      }
      const double v = fma(x, d1, cf[0] - d2);
      if (f < N)
        u[f] = v;
      else
        w[f - N] = v;
    }
  } else {
    const double it = 1.0 / T;
    const double isq = sqrt(it);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      u[i] = lai_rys_hx2[lai_rys_hoff[N] + i] * it;
      w[i] = lai_rys_hw[lai_rys_hoff[N] + i] * isq;
    }
  }
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

}  // namespace lai
