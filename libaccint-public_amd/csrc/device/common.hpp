// common.hpp -- shared device-side definitions for the gfx950 Rys ERI / J-K
// backend: Cartesian component tables, the HBM data layout of pairs and
// primitive pairs, Rys root/weight evaluation and atomics.
#pragma once

#include <type_traits>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "boys_taylor.inc"
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
// b.  The same record serves as bra or ket.  Both shells belong to one
// molecule of the basis; fa, fb are function offsets inside that molecule and
// (moff, mn) locate the molecule's mn x mn block of D, J and K.
struct PairRec {
  int sa, sb;         // shell indices (class order)
  int fa, fb;         // first basis function of each shell, within the molecule
  int pp_off;         // first primitive pair in the PrimPair array
  int npp;            // number of primitive pairs (zero coefficients dropped)
  int same;           // 1 when sa == sb
  int npp_sig;        // leading primitive pairs used by the J/K and gradient kernels
                      // (the rest are below the primitive screening bound)
  int mn;             // basis functions of the molecule
  int mol;            // molecule index
  long long moff;     // offset of the molecule's matrices
  double ABx, ABy, ABz;  // A - B (horizontal transfer)
  double bsum;           // sum of the primitive-pair bounds |Kc| zeta^(-1/4)
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

// Worklist of a quartet class.  For one bra pair, the kets that pass the
// Schwarz test form up to kMaxSegs contiguous runs ("segments") of the ket
// array, one per primitive-count bucket of the ket pair class: segment s
// holds kets[start[s] .. start[s] + cum[s+1] - cum[s]).  A chunk is the part
// [k0, k1) of that concatenated run handled by one workgroup.
inline constexpr int kMaxSegs = 8;
struct JkSegs {
  int n;
  int start[kMaxSegs];
  int cum[kMaxSegs + 1];
  int pad;
};
struct JkChunk {
  int bra, seg, k0, k1;
};

// one item of the persistent launch's queue: a chunk of class cls
struct PersistItem {
  int cls;
  JkChunk ch;
};

// ---------------------------------------------------------------------------
// Rys roots/weights.  The device table holds the piecewise-Chebyshev
// interpolants (lai_rys_cheb_h), the Gauss-Hermite limits, and from
// LAI_BT_OFF (128-byte aligned) the Boys Taylor rows (lai_boys_taylor_h).
// ---------------------------------------------------------------------------
#define LAI_BT_OFF (((LAI_RYS_TABLE_SIZE + 15) / 16) * 16)
// From LAI_B8_OFF: the one-root rows, 8 doubles (64 B) a row, half the cache
// footprint of the 128-byte rows: H_i = F_i(T_k) / i! for i = 0..7, the Taylor
// coefficients of F_0 about the grid point (boys01_taylor).
#define LAI_B8_OFF (LAI_BT_OFF + LAI_BT_ROWS * LAI_BT_M)
// From LAI_B10_OFF: F_0..F_9 of every row, 10 doubles (80 B) a row, for
// rys_roots<2>.
#define LAI_B10_OFF (LAI_B8_OFF + LAI_BT_ROWS * 8)
// From LAI_B9_OFF: the one-root rows on the coarser grid T_k = k/8 (k = 0 ..
// 8 T_cut(1)), 9 doubles a row: H_i = F_i(T_k) / i! for i = 0..8, the source
// of a workgroup's LDS copy (22.5 KB instead of 40 KB at step 1/16, so that
// the copy fits beside as many resident waves)
#define LAI_B9_OFF (LAI_B10_OFF + LAI_BT_ROWS * 10)
#define LAI_B9_ROWS (8 * 39 + 1)
#define LAI_DEVICE_TABLE_SIZE (LAI_B9_OFF + LAI_B9_ROWS * 9)

// Taylor step from the grid row r (F_m(T_k)) to F_m(T_k + d):
// F_m + e1 (F_{m+1} + e2 (F_{m+2} + ... + e6 F_{m+6})) with e_j = -d / j.
__device__ __forceinline__ double boys_step(const double* r, int m, const double (&e)[7]) {
  double s = r[m + 6];
#pragma unroll
  for (int j = 6; j >= 1; --j) s = fma(e[j], s, r[m + j - 1]);
  return s;
}

// 1/x for positive normal x: v_rcp_f64 and two Newton steps, each of which
// squares the relative error (about one ulp at the end)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Boys Taylor rows of the one- and two-root evaluations as the workgroup may
// hold them: in the device table, or copied to LDS (row k at rows + k W, W =
// 8 doubles for one root, 10 for two; rows 0 .. 16 T_cut(N)).
using lds_cdptr = const __attribute__((address_space(3))) double*;
template <int N>
inline constexpr int kBoysRowW = N == 1 ? 8 : 10;
// Row stride of the rows as read through P: packed in the device table; one
// double of padding per row in a workgroup's LDS copy (odd strides 9 and 11),
// so that the rows that neighbouring lanes read start in different banks
template <typename P, int W>
inline constexpr int kRowStride = std::is_same_v<P, lds_cdptr> ? W + 1 : W;
template <int N>
inline constexpr int kBoysRowsN = lai_rys_tcut[N] * LAI_BT_PER_UNIT + 1;

// Gauss-Hermite limit of the N-root rule (T at or above the cut)
template <int N>
__device__ __forceinline__ void rys_hermite(double T, double* u, double* w) {
  // 1/sqrt(T) by v_rsq_f64 and two Newton steps; 1/T is its square
  double y = __builtin_amdgcn_rsq(T);
  y = fma(0.5 * y, fma(-T * y, y, 1.0), y);
  y = fma(0.5 * y, fma(-T * y, y, 1.0), y);
  const double it = y * y;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    u[i] = lai_rys_hx2[lai_rys_hoff[N] + i] * it;
    w[i] = lai_rys_hw[lai_rys_hoff[N] + i] * y;
  }
}

// F_0 and F_1 below the one-root cut from the one-root rows (LAI_B8_OFF),
// which hold H_i = F_i(T_k) / i! of the grid point T_k = k/16 nearest T.
// Since dF_m/dT = -F_{m+1}, F_0(T) = P(x) = sum_i H_i x^i with x = T_k - T,
// and F_1 = -dF_0/dT = P'(x): one Horner pass gives both (13 FMAs).  The
// error of F_1 is the term F_8 x^7 / 7!, as for the seven-term series of F_1
// itself; F_0 carries one more term.
// A workgroup's LDS copy holds the coarser rows of LAI_B9_OFF (step 1/8,
// H_0..H_8, |x| <= 1/16): one more term keeps the same truncation order.
template <typename P>
__device__ __forceinline__ void boys01_taylor(double T, P rows, double& F0, double& F1) {
  constexpr bool lds = std::is_same_v<P, lds_cdptr>;
  constexpr int NH = lds ? 9 : 8;                   // coefficients H_0 .. H_{NH-1}
  constexpr double G = lds ? 8.0 : double(LAI_BT_PER_UNIT);  // grid points per unit of T
  const int k = int(fma(T, G, 0.5));
  const double x = fma(double(k), 1.0 / G, -T);
  double h[NH];
#pragma unroll
  for (int m = 0; m < NH; ++m) h[m] = rows[k * (lds ? 9 : 8) + m];
  double p = fma(h[NH - 1], x, h[NH - 2]), q = h[NH - 1];
#pragma unroll
  for (int i = NH - 3; i >= 0; --i) {
    q = fma(q, x, p);
    p = fma(p, x, h[i]);
  }
  F0 = p;
  F1 = q;
}

// F_0(T) and F_1(T): the one-root rows below the one-root cut, the
// Gauss-Hermite limit of the one-root rule above it (F_0 = w, F_1 = u w)
template <typename P>
__device__ __forceinline__ void boys01_at(double T, P rows, double& F0, double& F1) {
  if (T < double(lai_rys_tcut[1])) {
    boys01_taylor(T, rows, F0, F1);
  } else {
    double u, w;
    rys_hermite<1>(T, &u, &w);
    F0 = w;
    F1 = w * u;
  }
}
__device__ __forceinline__ void boys01(double T, const double* __restrict__ tab, double& F0, double& F1) {
  boys01_at(T, tab + LAI_B8_OFF, F0, F1);
}

// One- and two-root rules: moments F_0..F_{2N-1} from the Boys Taylor rows
// (T below the cut), then the Gauss rule of the moments in closed form
template <int N, typename P>
__device__ __forceinline__ void rys_roots12(double T, P rows, double* u, double* w) {
  static_assert(N == 1 || N == 2, "one or two roots");
  if constexpr (N == 1) {
    if (T < double(lai_rys_tcut[1])) {
      double F0, F1;
      boys01_taylor(T, rows, F0, F1);
      u[0] = F1 * rcp_nr(F0);
      w[0] = F0;
    } else {
      rys_hermite<1>(T, u, w);
    }
  } else if (T < double(lai_rys_tcut[N])) {
    const int k = int(fma(T, double(LAI_BT_PER_UNIT), 0.5));
    const double d = T - double(k) * (1.0 / LAI_BT_PER_UNIT);
    double e[7];
#pragma unroll
    for (int j = 1; j <= 6; ++j) e[j] = -d * (1.0 / j);
    double rr[2 * N + 6];  // F_0 .. F_{2N+5}: orders up to 2N-1 plus six Taylor terms
#pragma unroll
    for (int m = 0; m < 2 * N + 6; ++m) rr[m] = rows[k * kRowStride<P, kBoysRowW<N>> + m];
    const double F0 = boys_step(rr, 0, e), F1 = boys_step(rr, 1, e);
    {
      const double F2 = boys_step(rr, 2, e), F3 = boys_step(rr, 3, e);
      // x^2 + p x + q orthogonal to 1 and x under the moments
      const double idet = rcp_nr(fma(F0, F2, -F1 * F1));
      const double p = fma(F1, F2, -F0 * F3) * idet;
      const double q = fma(F1, F3, -F2 * F2) * idet;
      const double x2 = 0.5 * (sqrt(fma(p, p, -4.0 * q)) - p);
      const double x1 = q * rcp_nr(x2);
      const double w1 = fma(-x2, F0, F1) * rcp_nr(x1 - x2);
      u[0] = x1;
      u[1] = x2;
      w[0] = w1;
      w[1] = F0 - w1;
    }
  } else {
    rys_hermite<N>(T, u, w);
  }
}

// Chebyshev evaluation of the N-root rule (N >= 3) below its cut: interval
// k = floor(T) holds 2N series of LAI_RYS_NTERMS coefficients (the N roots,
// then the N weights) at c0 + k*stride, in the device table (stride 2N terms)
// or in a workgroup's LDS copy (padded stride); the Gauss-Hermite limit above.
template <int N, typename P>
__device__ __forceinline__ void rys_cheb(double T, P c0, int stride, double* u, double* w) {
  if (T < double(lai_rys_tcut[N])) {
    const int k = int(T);
    const double x = 2.0 * (T - double(k)) - 1.0;
    const double x2 = 2.0 * x;
    const P c = c0 + k * stride;
#pragma unroll
    for (int f = 0; f < 2 * N; ++f) {
      const P cf = c + f * LAI_RYS_NTERMS;
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
    rys_hermite<N>(T, u, w);
  }
}

// Stride of an interval's coefficients in the LDS copy: odd, so that lanes
// of neighbouring intervals start in different banks
template <int N>
inline constexpr int kChebStrideLds = 2 * N * LAI_RYS_NTERMS + 1;

template <int N>
__device__ __forceinline__ void rys_roots(double T, const double* __restrict__ tab, double* u, double* w) {
  static_assert(N >= 1 && N <= LAI_RYS_NMAX, "Rys order out of table range");
  if constexpr (N <= 2)
    rys_roots12<N>(T, tab + (N == 1 ? LAI_B8_OFF : LAI_B10_OFF), u, w);
  else
    rys_cheb<N>(T, tab + lai_rys_off[N], 2 * N * LAI_RYS_NTERMS, u, w);
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// Read-only tables seen through the constant address space: loads with a
// workgroup-uniform address become scalar loads.
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) {
  return (const_ptr<T>)p;
}
// primitive pair i of a read-only table, member by member through the
// constant address space
__device__ __forceinline__ PrimPair load_prim(const PrimPair* base, int i) {
  const_ptr<double> p = as_const(reinterpret_cast<const double*>(base + i));
  PrimPair r;
  r.zeta = p[0];
  r.Px = p[1];
  r.Py = p[2];
  r.Pz = p[3];
  r.PAx = p[4];
  r.PAy = p[5];
  r.PAz = p[6];
  r.Kc = p[7];
  return r;
}

// v[ax] of (x, y, z) for a per-lane axis 0..2, by bit masks: a select the
// compiler cannot turn into an indexed load from a private copy of the
// record the values come from (which would put that record in scratch)
__device__ __forceinline__ double pick3(int ax, double x, double y, double z) {
  const long long m0 = -(long long)(ax == 0), m1 = -(long long)(ax == 1), m2 = -(long long)(ax == 2);
  return __longlong_as_double((__double_as_longlong(x) & m0) | (__double_as_longlong(y) & m1) |
                              (__double_as_longlong(z) & m2));
}

// A pair record whose index is uniform over the wavefront, member by member
// through the constant address space (scalar loads into SGPRs, no private copy)
__device__ __forceinline__ PairRec load_pair_uniform(const PairRec* base, int i) {
  const __attribute__((address_space(4))) PairRec* p = (const __attribute__((address_space(4))) PairRec*)(base + i);
  PairRec r;
  r.sa = p->sa;
  r.sb = p->sb;
  r.fa = p->fa;
  r.fb = p->fb;
  r.pp_off = p->pp_off;
  r.npp = p->npp;
  r.same = p->same;
  r.npp_sig = p->npp_sig;
  r.mn = p->mn;
  r.mol = p->mol;
  r.moff = p->moff;
  r.ABx = p->ABx;
  r.ABy = p->ABy;
  r.ABz = p->ABz;
  r.bsum = p->bsum;
  return r;
}

// FP64 add to LDS (ds_add_f64)
__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Orders the LDS accesses of one wavefront without waiting for the others:
// what the wavefront wrote before is visible to all its lanes after (the
// workgroup-scope fences wait for the wavefront's own LDS operations and keep
// the compiler from moving accesses across; the lanes of a wavefront run in
// step).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace lai
