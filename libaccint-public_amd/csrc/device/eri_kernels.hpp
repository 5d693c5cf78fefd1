// eri_kernels.hpp -- Rys-quadrature ERI kernels with fused J/K digestion for
// gfx950.  Two execution shapes, both templated on the angular-momentum class
// (La >= Lb, Lc >= Ld, bra pair class >= ket pair class) and on the compute
// type R (double, or float for the mixed-precision path):
//
//  * Kernel A ("register" shape): one lane per contracted shell quartet.  The
//    2D Rys integrals, the horizontal transfer and the contraction live in
//    VGPRs; all loops over Cartesian components are unrolled at compile
//    time.  Large classes split the bra index a into chunks (gridDim.y) so the
//    per-lane accumulator stays small.
//  * Kernel B ("LDS" shape): one workgroup of NT lanes per shell quartet.  For
//    every primitive quartet, 3*NR lanes build the (root, axis) 2D tables into
//    LDS; every lane then forms its strided subset of the Cartesian products.
//
// The algorithm per primitive quartet follows the reference's Rys recursion
// (src/host/kernels/eri_kernel.cpp:104-164,247-368): roots u_r and weights
// w_r for T = rho |PQ|^2, B10/B01/B00 and the effective displacements, the
// vertical recursion in (a,0|c,0), horizontal transfer to b then d, and
// (ab|cd) = sum_r Ix Iy Iz with the weight and prefactor folded into Iz.
// Roots come from the piecewise-Chebyshev table (common.hpp), not from a
// per-quartet moment/QL solve.
//
// J/K digestion (mode 0) accumulates the unsymmetrised matrices Jt, Kt from
// which J = (Jt + Jt^T)/2 and K = (Kt + Kt^T)/2.  For a canonical quartet
// with degeneracy flags ij, kl, bk (bra shells equal, ket shells equal, bra
// pair equal ket pair) and symmetric D this reproduces the 8-permutation
// scatter of FockBuilder::accumulate_symmetric
// (src/shared/consumers/fock_builder.cpp:123-222):
//   Jt[ab] += s_ij s_kl g D[cd]          Jt[cd] += s_ij s_kl g D[ab]   (!bk)
//   Kt[ac] += s_bk g D[bd]               Kt[ad] += s_bk g D[bc]        (!kl)
//   Kt[bc] += s_bk g D[ad]   (!ij)       Kt[bd] += s_bk g D[ac]        (!ij && !kl)
// with s_x = 1 if the flag is set, else 2.
// Mode 1 writes the contracted block to an ERI buffer in the caller's
// original shell order (parity tests, Schwarz diagonal).
#pragma once

#include "common.hpp"

namespace lai {

// Per-quartet selection of the J/K and gradient kernels: the density screen
// and the FP32 split, decided from the Schwarz bounds of the two pairs and the
// shell-pair max|D| of the quartet's molecule.
struct QuartetSel {
  const double* __restrict__ Qpair;      // Schwarz bound of each pair
  const double* __restrict__ sdmax;      // per molecule: ns_m x ns_m shell-pair max|D|
  const long long* __restrict__ dm_off;  // per molecule: offset of its block
  const int* __restrict__ dm_ns;         // per molecule: shells
  const int* __restrict__ sh_loc;        // per shell: index within its molecule
  int dscreen;                           // require Dk Q_ij Q_kl >= thr
  double thr;
  int mixed;                             // FP32 recursion where Q_ij Q_kl max|D| < tau
  double tau;
  int keep;                              // 0: every selected quartet; 1: FP64 only; 2: FP32 only
};

// 0: removed by the density screen, 1: FP64, 2: FP32 recursion.  The same
// arithmetic as the host worklist rules (exchange-type Dk over (ik, il, jk,
// jl); the FP32 test adds D_ij and D_kl).  With keep set, quartets of the
// other precision return 0: a mixed build runs each class as two launches,
// one per precision.
__device__ __forceinline__ int quartet_select(const QuartetSel& s, int bp, const PairRec& bra, int kp,
                                              const PairRec& ket) {
  if (!s.dscreen && !s.mixed) return s.keep == 2 ? 0 : 1;
  const double prod = s.Qpair[bp] * s.Qpair[kp];
  const double* dm = s.sdmax + s.dm_off[bra.mol];
  const int ns = s.dm_ns[bra.mol];
  const int i = s.sh_loc[bra.sa], j = s.sh_loc[bra.sb], k = s.sh_loc[ket.sa], l = s.sh_loc[ket.sb];
  const double dk = fmax(fmax(dm[i * ns + k], dm[i * ns + l]), fmax(dm[j * ns + k], dm[j * ns + l]));
  if (s.dscreen && !(dk * prod >= s.thr)) return 0;
  const int mode = (s.mixed && prod * fmax(dk, fmax(dm[i * ns + j], dm[k * ns + l])) < s.tau) ? 2 : 1;
  return (s.keep == 0 || s.keep == mode) ? mode : 0;
}

// Ket pair of position k of a chunk's run; seg walks forward with k.
__device__ __forceinline__ int chunk_ket(const int* __restrict__ kets, const JkSegs* __restrict__ sg, int k,
                                         int& seg) {
  while (seg + 1 < sg->n && k >= sg->cum[seg + 1]) ++seg;
  return kets[sg->start[seg] + (k - sg->cum[seg])];
}

// Lane schedule of the lane-per-quartet kernels.  The quartets of a chunk
// that pass the per-quartet selection are counting-sorted in LDS by (FP32
// flag, primitive pairs of the ket), heaviest first, so that the 64 lanes of a
// wavefront run primitive loops of nearly equal length and one precision;
// quartets removed by the density screen take no lane.
constexpr int kSortRun = 2048;                 // quartets sorted at a time
constexpr int kSortBins = 256;                 // key: 128 * FP32 + min(npp_sig, 127)
constexpr unsigned kSortF32 = 0x80000000u;     // flag bit of a sorted entry
struct SortLds {
  unsigned kp[kSortRun];  // ket pair, kSortF32 set for the FP32 recursion
  int cnt[kSortBins];
  int n;
};

// Sorts positions [s0, s1) (at most kSortRun) of a chunk's run into s.kp and
// returns how many were kept.  All nt lanes of the workgroup call it (nt a
// multiple of 64, at least NTMIN); it ends on a barrier.
template <int NTMIN>
__device__ __forceinline__ int sort_run(const int* __restrict__ kets, const JkSegs* __restrict__ sg,
                                        const PairRec* __restrict__ pairs, const QuartetSel& sel, int bp,
                                        const PairRec& bra, int s0, int s1, SortLds& s, int nt) {
  constexpr int kSortPer = (kSortRun + NTMIN - 1) / NTMIN;  // positions per lane, at most
  const int tid = threadIdx.x;
  for (int i = tid; i < kSortBins; i += nt) s.cnt[i] = 0;
  __syncthreads();
  int kp[kSortPer], key[kSortPer], rk[kSortPer];
  int seg = 0;
#pragma unroll
  for (int j = 0; j < kSortPer; ++j) {
    const int k = s0 + tid + nt * j;
    key[j] = -1;
    kp[j] = 0;
    rk[j] = 0;
    if (k < s1) {
      kp[j] = chunk_ket(kets, sg, k, seg);
      const PairRec ket = pairs[kp[j]];
      const int mode = quartet_select(sel, bp, bra, kp[j], ket);
      if (mode != 0) {
        key[j] = (mode == 2 ? 128 : 0) + min(ket.npp_sig, 127);
        rk[j] = atomicAdd(&s.cnt[key[j]], 1);
      }
    }
  }
  __syncthreads();
  if (tid < 64) {  // offsets in decreasing key order; lane l owns keys 255-4l .. 252-4l
    int c[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      c[q] = s.cnt[255 - 4 * tid - q];
      sum += c[q];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (tid >= off) incl += t;
    }
    int run = incl - sum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      s.cnt[255 - 4 * tid - q] = run;
      run += c[q];
    }
    if (tid == 63) s.n = incl;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortPer; ++j)
    if (key[j] >= 0) s.kp[s.cnt[key[j]] + rk[j]] = unsigned(kp[j]) | (key[j] >= 128 ? kSortF32 : 0u);
  __syncthreads();
  return s.n;
}

// Sorted position of lane tid in round r of a workgroup of nw wavefronts (a
// power of two): the wavefronts take the nw 64-entry blocks of a round in an
// order that rotates with r, so that no wavefront always gets the heaviest
// block.
__device__ __forceinline__ int sorted_pos(int tid, int r, int nw) {
  const int w = tid >> 6;
  return 64 * (nw * r + ((w + r) & (nw - 1))) + (tid & 63);
}

struct EriArgs {
  const Quartet* __restrict__ q;
  int nq;
  const PairRec* __restrict__ pairs;
  const PrimPair* __restrict__ pp;
  const double* __restrict__ rys;
  // bounds |Kc| zeta^(-1/4) of the primitive pairs, and the relative cut of
  // the FP32 recursion: primitive quartets with bound below
  // pcut32 * bsum(bra) * bsum(ket) are skipped (0: none)
  const float* __restrict__ ppb;
  double pcut32;
  const double* __restrict__ ppih;  // 1 / (2 zeta) of each primitive pair
  // mode 0: digestion into D, Jt, Kt (molecule blocks at PairRec::moff)
  const double* __restrict__ D;
  double* Jt;
  double* Kt;
  int nbf;  // largest molecule (LDS rows)
  int mode;
  // mode 1: raw ERI output, [q][block] in the original order given by perm
  double* out;
  const int* __restrict__ perm;        // per quartet: packed original position of class positions
  const long long* __restrict__ obase;  // per quartet: offset of its block in out
  // mode 0 worklist: bra-anchored chunks over ket segments
  const JkChunk* __restrict__ chunks;
  int nchunks;
  const JkSegs* __restrict__ segs;
  const int* __restrict__ kets;
  int klds;  // 1: the bra rows of Kt are accumulated in LDS (dynamic shared memory)
  QuartetSel sel;
};

// --------------------------------------------------------------------------
// 2D recursion into registers (Kernel A)
// --------------------------------------------------------------------------
template <int La, int Lb, int Lc, int Ld, typename R>
__device__ __forceinline__ void build_2d(R (&I)[La + 1][Lb + 1][Lc + 1][Ld + 1], R PA, R QC, R AB,
                                         R CD, R B10, R B01, R B00, R s) {
  constexpr int LAB = La + Lb, LCD = Lc + Ld;
  R G[LAB + 1][LCD + 1];
  G[0][0] = s;
#pragma unroll
  for (int a = 0; a < LAB; ++a) {
    R v = PA * G[a][0];
    if (a > 0) v += R(a) * B10 * G[a - 1][0];
    G[a + 1][0] = v;
  }
#pragma unroll
  for (int c = 0; c < LCD; ++c) {
#pragma unroll
    for (int a = 0; a <= LAB; ++a) {
      R v = QC * G[a][c];
      if (c > 0) v += R(c) * B01 * G[a][c - 1];
      if (a > 0) v += R(a) * B00 * G[a - 1][c];
      G[a][c + 1] = v;
    }
  }
  R H[Lb + 1][LAB + 1][LCD + 1];
#pragma unroll
  for (int a = 0; a <= LAB; ++a)
#pragma unroll
    for (int k = 0; k <= LCD; ++k) H[0][a][k] = G[a][k];
#pragma unroll
  for (int b = 0; b < Lb; ++b)
#pragma unroll
    for (int a = 0; a < LAB - b; ++a)
#pragma unroll
      for (int k = 0; k <= LCD; ++k) H[b + 1][a][k] = H[b][a + 1][k] + AB * H[b][a][k];
#pragma unroll
  for (int a = 0; a <= La; ++a)
#pragma unroll
    for (int b = 0; b <= Lb; ++b) {
      R X[Ld + 1][LCD + 1];
#pragma unroll
      for (int k = 0; k <= LCD; ++k) X[0][k] = H[b][a][k];
#pragma unroll
      for (int d = 0; d < Ld; ++d)
#pragma unroll
        for (int c = 0; c < LCD - d; ++c) X[d + 1][c] = X[d][c + 1] + CD * X[d][c];
#pragma unroll
      for (int c = 0; c <= Lc; ++c)
#pragma unroll
        for (int d = 0; d <= Ld; ++d) I[a][b][c][d] = X[d][c];
    }
}

// --------------------------------------------------------------------------
// Digestion helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ int orig_offset(int perm, const int (&idx)[4], const int (&n)[4]) {
  // perm packs, for class position p (0..3), the original position in bits
  // [2p, 2p+1].  Original dims are the class dims permuted the same way.
  int oi[4], on[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int o = (perm >> (2 * p)) & 3;
    oi[o] = idx[p];
    on[o] = n[p];
  }
  return ((oi[0] * on[1] + oi[1]) * on[2] + oi[2]) * on[3] + oi[3];
}

// --------------------------------------------------------------------------
// Kernel A: one lane per shell quartet, bra index a in [A0, A1)
// --------------------------------------------------------------------------
// Contracted, normalised block gv[a - A0][b][c][d] of the quartet (bra|ket);
// all_prims selects every primitive pair (raw ERIs) or only those above the
// primitive screening bound (J/K).
// LB: the Boys Taylor rows (one and two roots) or the Chebyshev coefficients
// of the rule (three and four roots) from the workgroup's copy in LDS
// (lrows), else from the device table
template <int La, int Lb, int Lc, int Ld, typename R, int A0, int A1, bool LB = false>
__device__ __forceinline__ void quartet_block_regs(
    const EriArgs& g, const PairRec& bra, const PairRec& ket, bool all_prims,
    double (&gv)[A1 - A0][ncart(Lb)][ncart(Lc)][ncart(Ld)], lds_cdptr lrows = nullptr) {
  constexpr int nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NA = A1 - A0;
  constexpr int NR = (La + Lb + Lc + Ld) / 2 + 1;
  const double* grows = g.rys + (NR == 1 ? LAI_B8_OFF : LAI_B10_OFF);
  R acc[NA][nb][nc][nd];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b)
#pragma unroll
      for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) acc[a][b][c][d] = R(0);

  const R ABx = R(bra.ABx), ABy = R(bra.ABy), ABz = R(bra.ABz);
  const R CDx = R(ket.ABx), CDy = R(ket.ABy), CDz = R(ket.ABz);
  const int nb_pp = all_prims ? bra.npp : bra.npp_sig;
  const int nk_pp = all_prims ? ket.npp : ket.npp_sig;
  // FP32 part: primitive pairs are sorted by bound, so both loops stop at the
  // first primitive quartet below the relative cut
  constexpr bool fp32 = sizeof(R) == 4;
  const float qcut = fp32 && !all_prims ? float(g.pcut32 * bra.bsum * ket.bsum) : 0.0f;
  const float bb0 = fp32 ? g.ppb[bra.pp_off] : 0.0f;
  // ket primitive pairs outside (one load per lane and pair), bra primitive
  // pairs inside: the bra is the same for the whole workgroup in the J/K
  // kernels, so its records are uniform (scalar) loads
  for (int ik = 0; ik < nk_pp; ++ik) {
    const float bk = fp32 ? g.ppb[ket.pp_off + ik] : 0.0f;
    if (fp32 && bk * bb0 < qcut) break;
    const PrimPair Q = g.pp[ket.pp_off + ik];
    const double eta = Q.zeta;
    const double h_e = g.ppih[ket.pp_off + ik];  // 1 / (2 eta)
    for (int ib = 0; ib < nb_pp; ++ib) {
      if (fp32 && as_const(g.ppb)[bra.pp_off + ib] * bk < qcut) break;
      // the bra is uniform in the J/K kernels: scalar loads
      const PrimPair P = load_prim(g.pp, bra.pp_off + ib);
      const double zeta = P.zeta;
      const double h_z = as_const(g.ppih)[bra.pp_off + ib];  // 1 / (2 zeta)
      // 1/sqrt(zeta + eta) by v_rsq_f64 and two Newton steps; 1/(zeta + eta) = its square
      const double x = zeta + eta;
      double y = __builtin_amdgcn_rsq(x);
      y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
      y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
      const double izpe = y * y;
      const double rz = eta * izpe;   // rho / zeta
      const double re = zeta * izpe;  // rho / eta
      const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
      const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = P.Kc * Q.Kc * y;
      if constexpr (La + Lb + Lc + Ld == 1) {
        // (ps|ss), the one class of total angular momentum 1: with one root,
        // pref w (PA - (rho/zeta) u PQ) = pref (F0 PA - (rho/zeta) F1 PQ), so
        // F0 and F1 replace the root and weight
        double F0, F1;
        if constexpr (LB)
          boys01_at(T, lrows, F0, F1);
        else
          boys01_at(T, grows, F0, F1);
        const double s0 = pref * F0, s1 = -pref * F1 * rz;
        const double pa[3] = {P.PAx, P.PAy, P.PAz}, pq[3] = {PQx, PQy, PQz};
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          if constexpr (sizeof(R) == 8) {
            acc[a][0][0][0] = fma(s1, pq[A0 + a], fma(s0, pa[A0 + a], acc[a][0][0][0]));
          } else {
            acc[a][0][0][0] += R(fma(s1, pq[A0 + a], s0 * pa[A0 + a]));
          }
        }
      } else {
      double u[NR], w[NR];
      if constexpr (NR <= 2) {
        if constexpr (LB)
          rys_roots12<NR>(T, lrows, u, w);
        else
          rys_roots12<NR>(T, grows, u, w);
      } else if constexpr (LB) {
        rys_cheb<NR>(T, lrows, kChebStrideLds<NR>, u, w);
      } else {
        rys_roots<NR>(T, g.rys, u, w);
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double ur = u[r];
        const R B00 = R(0.5 * ur * izpe);
        const R B10 = R(h_z - h_z * rz * ur);
        const R B01 = R(h_e - h_e * re * ur);
        const double sz = rz * ur, se = re * ur;
        R Ix[La + 1][Lb + 1][Lc + 1][Ld + 1], Iy[La + 1][Lb + 1][Lc + 1][Ld + 1],
            Iz[La + 1][Lb + 1][Lc + 1][Ld + 1];
        build_2d<La, Lb, Lc, Ld, R>(Ix, R(P.PAx - sz * PQx), R(Q.PAx + se * PQx), ABx, CDx, B10, B01,
                                    B00, R(1));
        build_2d<La, Lb, Lc, Ld, R>(Iy, R(P.PAy - sz * PQy), R(Q.PAy + se * PQy), ABy, CDy, B10, B01,
                                    B00, R(1));
        build_2d<La, Lb, Lc, Ld, R>(Iz, R(P.PAz - sz * PQz), R(Q.PAz + se * PQz), ABz, CDz, B10, B01,
                                    B00, R(pref * w[r]));
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
              for (int d = 0; d < nd; ++d) {
                const int ai = A0 + a;
                acc[a][b][c][d] += Ix[cart_comp(La, ai, 0)][cart_comp(Lb, b, 0)][cart_comp(Lc, c, 0)]
                                     [cart_comp(Ld, d, 0)] *
                                   Iy[cart_comp(La, ai, 1)][cart_comp(Lb, b, 1)][cart_comp(Lc, c, 1)]
                                     [cart_comp(Ld, d, 1)] *
                                   Iz[cart_comp(La, ai, 2)][cart_comp(Lb, b, 2)][cart_comp(Lc, c, 2)]
                                     [cart_comp(Ld, d, 2)];
              }
      }
      }
    }
  }

  // contracted integrals with per-component normalisation
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b)
#pragma unroll
      for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d)
          gv[a][b][c][d] = double(acc[a][b][c][d]) * (norm_corr(La, A0 + a) * norm_corr(Lb, b) *
                                                        norm_corr(Lc, c) * norm_corr(Ld, d));
}

template <int La, int Lb, int Lc, int Ld, typename R, int A0, int A1>
__device__ __forceinline__ void quartet_regs(const EriArgs& g, int qi) {
  constexpr int nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NA = A1 - A0;
  const Quartet q = g.q[qi];
  const PairRec bra = g.pairs[q.bra];
  const PairRec ket = g.pairs[q.ket];
  double gv[NA][nb][nc][nd];
  quartet_block_regs<La, Lb, Lc, Ld, R, A0, A1>(g, bra, ket, true, gv);

  constexpr int na = ncart(La);
  const int perm = g.perm[qi];
  const int n4[4] = {na, nb, nc, nd};
  double* out = g.out + g.obase[qi];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b)
#pragma unroll
      for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) {
          const int idx[4] = {A0 + a, b, c, d};
          out[orig_offset(perm, idx, n4)] = gv[a][b][c][d];
        }
}

template <int La, int Lb, int Lc, int Ld, typename R, int CHA, int C>
__device__ __forceinline__ void dispatch_chunk(const EriArgs& g, int qi, int chunk) {
  constexpr int na = ncart(La);
  constexpr int A0 = C * CHA;
  constexpr int A1 = (A0 + CHA < na) ? A0 + CHA : na;
  if (chunk == C) {
    quartet_regs<La, Lb, Lc, Ld, R, A0, A1>(g, qi);
  } else if constexpr (A1 < na) {
    dispatch_chunk<La, Lb, Lc, Ld, R, CHA, C + 1>(g, qi, chunk);
  }
}

template <int La, int Lb, int Lc, int Ld, typename R, int CHA>
__global__ __launch_bounds__(128) void eri_kernel_regs(EriArgs g) {
  const int chunk = blockIdx.y;
  for (int qi = blockIdx.x * blockDim.x + threadIdx.x; qi < g.nq; qi += gridDim.x * blockDim.x)
    dispatch_chunk<La, Lb, Lc, Ld, R, CHA, 0>(g, qi, chunk);
}

// --------------------------------------------------------------------------
// Kernel B: one workgroup of NT lanes per shell quartet, 2D tables in LDS
// --------------------------------------------------------------------------
#ifndef LAI_LDS_TAB_KB
#define LAI_LDS_TAB_KB 32
#endif
template <int La, int Lb, int Lc, int Ld, typename R>
struct LdsShape {
  static constexpr int NR = (La + Lb + Lc + Ld) / 2 + 1;
  static constexpr int S = (La + 1) * (Lb + 1) * (Lc + 1) * (Ld + 1);
  static constexpr int LCD = Lc + Ld;
  static constexpr int HS = (La + 1) * (Lb + 1) * (LCD + 1);  // bra-transferred scratch
  static constexpr int NOUT = ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld);
  static constexpr int TS = NR * 3 * S;    // table entries of one primitive quartet
  static constexpr int HSS = NR * 3 * HS;  // transfer scratch of one primitive quartet
  // the d-major tables of the workgroup kernel: one (root, axis) table every
  // SP (odd) entries, TSP entries per primitive quartet
  static constexpr int SP = S | 1;
  static constexpr int TSP = NR * 3 * SP;
  // primitive quartets whose tables are built at once by NT lanes, within
  // LAI_LDS_TAB_KB of tables
  template <int NT>
  static constexpr int npq() {
    constexpr int by_lanes = NT / (3 * NR);
    constexpr int by_lds = int((LAI_LDS_TAB_KB * 1024) / (TSP * sizeof(R)));
    constexpr int n = by_lanes < by_lds ? by_lanes : by_lds;
    return n < 1 ? 1 : n;
  }
};

// Builds one (root, axis) table I[a][b][c][d] into tab (stride layout
// ((a*(Lb+1)+b)*(Lc+1)+c)*(Ld+1)+d), using hs as private LDS scratch.
template <int La, int Lb, int Lc, int Ld, typename R>
__device__ __forceinline__ void build_2d_lds(R* __restrict__ tab, R* __restrict__ hs, R PA, R QC,
                                             R AB, R CD, R B10, R B01, R B00, R s) {
  constexpr int LAB = La + Lb, LCD = Lc + Ld;
  R cm1[LAB + 1], c0[LAB + 1];
#pragma unroll
  for (int k = 0; k <= LCD; ++k) {
    R c1[LAB + 1];
    if (k == 0) {
      c1[0] = s;
#pragma unroll
      for (int a = 0; a < LAB; ++a) {
        R v = PA * c1[a];
        if (a > 0) v += R(a) * B10 * c1[a - 1];
        c1[a + 1] = v;
      }
    } else {
#pragma unroll
      for (int a = 0; a <= LAB; ++a) {
        R v = QC * c0[a];
        if (k > 1) v += R(k - 1) * B01 * cm1[a];
        if (a > 0) v += R(a) * B00 * c0[a - 1];
        c1[a] = v;
      }
    }
    // bra transfer on this column
    R h[Lb + 1][LAB + 1];
#pragma unroll
    for (int a = 0; a <= LAB; ++a) h[0][a] = c1[a];
#pragma unroll
    for (int b = 0; b < Lb; ++b)
#pragma unroll
      for (int a = 0; a < LAB - b; ++a) h[b + 1][a] = h[b][a + 1] + AB * h[b][a];
#pragma unroll
    for (int a = 0; a <= La; ++a)
#pragma unroll
      for (int b = 0; b <= Lb; ++b) hs[(a * (Lb + 1) + b) * (LCD + 1) + k] = h[b][a];
#pragma unroll
    for (int a = 0; a <= LAB; ++a) {
      cm1[a] = c0[a];
      c0[a] = c1[a];
    }
  }
#pragma unroll
  for (int a = 0; a <= La; ++a)
#pragma unroll
    for (int b = 0; b <= Lb; ++b) {
      R X[Ld + 1][LCD + 1];
#pragma unroll
      for (int k = 0; k <= LCD; ++k) X[0][k] = hs[(a * (Lb + 1) + b) * (LCD + 1) + k];
#pragma unroll
      for (int d = 0; d < Ld; ++d)
#pragma unroll
        for (int c = 0; c < LCD - d; ++c) X[d + 1][c] = X[d][c + 1] + CD * X[d][c];
#pragma unroll
      for (int c = 0; c <= Lc; ++c)
#pragma unroll
        for (int d = 0; d <= Ld; ++d) tab[((a * (Lb + 1) + b) * (Lc + 1) + c) * (Ld + 1) + d] = X[d][c];
    }
}

// The same table in the d-major layout of the wavefront kernel,
// I[a][b][c][d] at d*SABC + (a*(Lb+1)+b)*(Lc+1)+c with SABC =
// (La+1)(Lb+1)(Lc+1): the Ld+1 entries a product reads for one (a, b, c) lie
// SABC apart, so the rows read by the lanes of a wavefront spread over the
// LDS banks.  No scratch: the bra transfer of column k goes straight into
// slot k of the (a, b) row's (Lc+1)(Ld+1) slots (slot k = (d, c) with
// k = d*(Lc+1) + c; Lc+Ld+1 <= (Lc+1)(Ld+1)), and the ket transfer of each
// row then runs in place from registers (all of a row's slots are read
// before any is written; the row belongs to this lane).
template <int La, int Lb, int Lc, int Ld, typename R>
__device__ __forceinline__ void build_2d_inplace(R* __restrict__ tab, R PA, R QC, R AB, R CD, R B10, R B01, R B00,
                                                 R s) {
  constexpr int LAB = La + Lb, LCD = Lc + Ld, SABC = (La + 1) * (Lb + 1) * (Lc + 1);
  auto slot = [](int row, int k) { return (k / (Lc + 1)) * SABC + row + k % (Lc + 1); };
  R cm1[LAB + 1], c0[LAB + 1];
#pragma unroll
  for (int k = 0; k <= LCD; ++k) {
    R c1[LAB + 1];
    if (k == 0) {
      c1[0] = s;
#pragma unroll
      for (int a = 0; a < LAB; ++a) {
        R v = PA * c1[a];
        if (a > 0) v += R(a) * B10 * c1[a - 1];
        c1[a + 1] = v;
      }
    } else {
#pragma unroll
      for (int a = 0; a <= LAB; ++a) {
        R v = QC * c0[a];
        if (k > 1) v += R(k - 1) * B01 * cm1[a];
        if (a > 0) v += R(a) * B00 * c0[a - 1];
        c1[a] = v;
      }
    }
    R h[Lb + 1][LAB + 1];
#pragma unroll
    for (int a = 0; a <= LAB; ++a) h[0][a] = c1[a];
#pragma unroll
    for (int b = 0; b < Lb; ++b)
#pragma unroll
      for (int a = 0; a < LAB - b; ++a) h[b + 1][a] = h[b][a + 1] + AB * h[b][a];
#pragma unroll
    for (int a = 0; a <= La; ++a)
#pragma unroll
      for (int b = 0; b <= Lb; ++b) tab[slot((a * (Lb + 1) + b) * (Lc + 1), k)] = h[b][a];
#pragma unroll
    for (int a = 0; a <= LAB; ++a) {
      cm1[a] = c0[a];
      c0[a] = c1[a];
    }
  }
  if constexpr (Ld > 0) {  // with Ld = 0, slot k = (0, k) already holds I[a][b][k][0]
#pragma unroll
    for (int a = 0; a <= La; ++a)
#pragma unroll
      for (int b = 0; b <= Lb; ++b) {
        const int row = (a * (Lb + 1) + b) * (Lc + 1);
        R X[Ld + 1][LCD + 1];
#pragma unroll
        for (int k = 0; k <= LCD; ++k) X[0][k] = tab[slot(row, k)];
#pragma unroll
        for (int d = 0; d < Ld; ++d)
#pragma unroll
          for (int c = 0; c < LCD - d; ++c) X[d + 1][c] = X[d][c + 1] + CD * X[d][c];
#pragma unroll
        for (int c = 0; c <= Lc; ++c)
#pragma unroll
          for (int d = 0; d <= Ld; ++d) tab[d * SABC + row + c] = X[d][c];
      }
  }
}

// One Rys root/weight pair (index r of N) from the table.
template <int N>
__device__ __forceinline__ void rys_root_one(double T, const double* __restrict__ tab, int r,
                                             double& u, double& w) {
  if (T < double(lai_rys_tcut[N])) {
    const int k = int(T);
    const double x = 2.0 * (T - double(k)) - 1.0;
    const double x2 = 2.0 * x;
    const double* c = tab + lai_rys_off[N] + k * (2 * N * LAI_RYS_NTERMS);
    double v[2];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const double* cf = c + (f * N + r) * LAI_RYS_NTERMS;
      double d1 = 0.0, d2 = 0.0;
#pragma unroll
      for (int j = LAI_RYS_NTERMS - 1; j >= 1; --j) {
        const double dk = fma(x2, d1, cf[j] - d2);
        d2 = d1;
        d1 = dk;
      }
      v[f] = fma(x, d1, cf[0] - d2);
    }
    u = v[0];
    w = v[1];
  } else {
    const double it = 1.0 / T;
    u = tab[LAI_RYS_HX2_OFF + lai_rys_hoff[N] + r] * it;
    w = tab[LAI_RYS_HW_OFF + lai_rys_hoff[N] + r] * sqrt(it);
  }
}

// Products are tiled by task = (a, b, c): a lane owns tasks t = tid + NT*j
// and computes all nd outputs (a, b, c, d) of each from three rows of Ld+1
// table entries held in registers.  ofs[j] packs the row start of each axis
// (10 bits each): ((a*(Lb+1)+b)*(Lc+1)+c)*(Ld+1) in the d-minor layout
// (ROW_LD), (a*(Lb+1)+b)*(Lc+1)+c in the d-major layout of the wavefront
// kernel.
template <int La, int Lb, int Lc, int Ld>
struct LdsTasks {
  static constexpr int NTASK = ncart(La) * ncart(Lb) * ncart(Lc);
};
template <int La, int Lb, int Lc, int Ld, int NT, int NJ, bool ROW_LD = true>
__device__ __forceinline__ void lds_component_offsets(int tid, int (&ofs)[NJ]) {
  constexpr int nb = ncart(Lb), nc = ncart(Lc);
  constexpr int NTASK = LdsTasks<La, Lb, Lc, Ld>::NTASK;
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int t = tid + NT * j;
    if (t < NTASK) {
      const int c = t % nc, b = (t / nc) % nb, a = t / (nc * nb);
      int ax[3], bx[3], cx[3];
      // component exponents by a runtime walk (once per launch)
      int cnt = 0;
      for (int lx = La; lx >= 0; --lx)
        for (int ly = La - lx; ly >= 0; --ly) {
          if (cnt == a) { ax[0] = lx; ax[1] = ly; ax[2] = La - lx - ly; }
          ++cnt;
        }
      cnt = 0;
      for (int lx = Lb; lx >= 0; --lx)
        for (int ly = Lb - lx; ly >= 0; --ly) {
          if (cnt == b) { bx[0] = lx; bx[1] = ly; bx[2] = Lb - lx - ly; }
          ++cnt;
        }
      cnt = 0;
      for (int lx = Lc; lx >= 0; --lx)
        for (int ly = Lc - lx; ly >= 0; --ly) {
          if (cnt == c) { cx[0] = lx; cx[1] = ly; cx[2] = Lc - lx - ly; }
          ++cnt;
        }
      int p = 0;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        const int off = ((ax[x] * (Lb + 1) + bx[x]) * (Lc + 1) + cx[x]) * (ROW_LD ? Ld + 1 : 1);
        p |= off << (10 * x);
      }
      ofs[j] = p;
    } else {
      ofs[j] = 0;
    }
  }
}

// One shell quartet by the whole workgroup.  Primitive quartets are taken NPQ
// at a time: 3*NR*NPQ lanes build the (primitive quartet, root, axis) tables
// in LDS, then every lane accumulates its outputs over the batch.  Starts and
// ends with a workgroup barrier.
template <int La, int Lb, int Lc, int Ld, typename R, int NT, int NJ, int NPQ>
__device__ __forceinline__ void lds_quartet_acc(const EriArgs& g, const PairRec& bra, const PairRec& ket,
                                                bool all_prims, R* tab, const int (&ofs)[NJ],
                                                R (&acc)[NJ][ncart(Ld)]) {
  constexpr int nd = ncart(Ld);
  using Sh = LdsShape<La, Lb, Lc, Ld, R>;
  constexpr int NR = Sh::NR, S = Sh::S, SP = Sh::SP, TS = Sh::TSP, TASKS = 3 * NR;
  constexpr int SABC = (La + 1) * (Lb + 1) * (Lc + 1);
  static_assert(NPQ * TASKS <= NT, "not enough lanes for the table batch");
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int d = 0; d < nd; ++d) acc[j][d] = R(0);
  const int npp_b = all_prims ? bra.npp : bra.npp_sig, npp_k = all_prims ? ket.npp : ket.npp_sig;
  const int ntot = npp_b * npp_k;
  constexpr bool fp32 = sizeof(R) == 4;
  const float qcut = fp32 && !all_prims ? float(g.pcut32 * bra.bsum * ket.bsum) : 0.0f;
  __syncthreads();  // the previous user of tab has finished
  for (int t0 = 0; t0 < ntot; t0 += NPQ) {
    if (tid < NPQ * TASKS) {
      const int pq = tid / TASKS, task = tid - pq * TASKS;
      const int r = task / 3, ax = task - 3 * r;
      const int t = t0 + pq;
      R* tb = tab + pq * TS + task * SP;
      bool live = t < ntot;
      const int ib = live ? t / npp_k : 0, ik = live ? t - ib * npp_k : 0;
      if (live && fp32 && g.ppb[bra.pp_off + ib] * g.ppb[ket.pp_off + ik] < qcut) live = false;
      if (live) {
        const PrimPair P = g.pp[bra.pp_off + ib];
        const PrimPair Q = g.pp[ket.pp_off + ik];
        const double zeta = P.zeta, eta = Q.zeta;
        const double izpe = 1.0 / (zeta + eta);
        const double rz = eta * izpe, re = zeta * izpe;
        const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
        const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
        double ur, wr;
        rys_root_one<NR>(T, g.rys, r, ur, wr);
        const double h_z = g.ppih[bra.pp_off + ib], h_e = g.ppih[ket.pp_off + ik];
        const R B00 = R(0.5 * ur * izpe);
        const R B10 = R(h_z - h_z * rz * ur);
        const R B01 = R(h_e - h_e * re * ur);
        const double PQ = pick3(ax, PQx, PQy, PQz);
        const double PA = pick3(ax, P.PAx, P.PAy, P.PAz);
        const double QC = pick3(ax, Q.PAx, Q.PAy, Q.PAz);
        const double AB = pick3(ax, bra.ABx, bra.ABy, bra.ABz);
        const double CD = pick3(ax, ket.ABx, ket.ABy, ket.ABz);
        const double sc = ax == 2 ? P.Kc * Q.Kc * sqrt(izpe) * wr : 1.0;
        build_2d_inplace<La, Lb, Lc, Ld, R>(tb, R(PA - rz * ur * PQ), R(QC + re * ur * PQ), R(AB), R(CD), B10,
                                            B01, B00, R(sc));
      } else {
        for (int e = 0; e < S; ++e) tb[e] = R(0);  // contributes nothing
      }
    }
    __syncthreads();
    const int npq = ntot - t0 < NPQ ? ntot - t0 : NPQ;
#pragma unroll 1
    for (int pq = 0; pq < npq; ++pq) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const R* tx = tab + pq * TS + (r * 3) * SP;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int p = ofs[j];
          R X[Ld + 1], Y[Ld + 1], Z[Ld + 1];
#pragma unroll
          for (int e = 0; e <= Ld; ++e) {
            X[e] = tx[(p & 1023) + e * SABC];
            Y[e] = tx[SP + ((p >> 10) & 1023) + e * SABC];
            Z[e] = tx[2 * SP + (p >> 20) + e * SABC];
          }
#pragma unroll
          for (int d = 0; d < nd; ++d)
            acc[j][d] += X[cart_comp(Ld, d, 0)] * Y[cart_comp(Ld, d, 1)] * Z[cart_comp(Ld, d, 2)];
        }
      }
    }
    __syncthreads();  // products finished with tab before the next batch
  }
}

// Kernel B for raw ERIs (mode 1).
template <int La, int Lb, int Lc, int Ld, typename R, int NT>
__global__ __launch_bounds__(NT) void eri_kernel_lds(EriArgs g) {
  using Sh = LdsShape<La, Lb, Lc, Ld, R>;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NTASK = LdsTasks<La, Lb, Lc, Ld>::NTASK;
  constexpr int NJ = (NTASK + NT - 1) / NT;
  constexpr int NPQ = Sh::template npq<NT>();
  __shared__ R tab[NPQ * Sh::TSP];
  const int tid = threadIdx.x;
  int ofs[NJ];
  lds_component_offsets<La, Lb, Lc, Ld, NT, NJ, false>(tid, ofs);

  for (int qi = blockIdx.x; qi < g.nq; qi += gridDim.x) {
    const Quartet q = g.q[qi];
    const PairRec bra = g.pairs[q.bra];
    const PairRec ket = g.pairs[q.ket];
    R acc[NJ][nd];
    lds_quartet_acc<La, Lb, Lc, Ld, R, NT, NJ, NPQ>(g, bra, ket, true, tab, ofs, acc);
    const int perm = g.perm[qi];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int t = tid + NT * j;
      if (t >= NTASK) continue;
      const int c = t % nc, b = (t / nc) % nb, a = t / (nc * nb);
#pragma unroll
      for (int d = 0; d < nd; ++d) {
        const double v = double(acc[j][d]) * norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c) *
                         norm_corr(Ld, d);
        const int idx[4] = {a, b, c, d};
        const int n4[4] = {na, nb, nc, nd};
        g.out[g.obase[qi] + orig_offset(perm, idx, n4)] = v;
      }
    }
  }
}

// --------------------------------------------------------------------------
// Anchored J/K digestion (mode 0).  A workgroup takes one chunk: one bra pair
// (A, B) and a run of its kets.  Every exchange update of such a quartet lands
// in a row of A or B (Kt[ac], Kt[ad], Kt[bc], Kt[bd]), so the workgroup
// accumulates those rows (rows x mn doubles) in LDS and adds them to Kt once
// per chunk; the bra Coulomb block Jt[ab] is summed over the chunk as well.
// Only the ket-side Coulomb update Jt[cd] goes to HBM per quartet.  With
// klds == 0 (rows too long for LDS) the exchange updates go straight to HBM.
// MIXED: every quartet of the launch runs the recursion in float (the launch
// selects only the FP32 quartets, QuartetSel::keep == 2); otherwise FP64.
// Lanes take the chunk's quartets in the order of sort_run.
// --------------------------------------------------------------------------
template <int La, int Lb, int Lc, int Ld, bool MIXED, int A0, int A1, bool LB = false>
__device__ __forceinline__ void jk_chunk_regs(const EriArgs& g, const JkChunk& ch, double* krow, SortLds& srt,
                                              lds_cdptr lrows = nullptr) {
  constexpr int nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NA = A1 - A0;
  constexpr int ROWS = NA + nb;
  const int tid = threadIdx.x, nt = blockDim.x;
  const bool lds = g.klds != 0;
  const PairRec bra = g.pairs[ch.bra];
  const int n = bra.mn;
  const double* __restrict__ D = g.D + bra.moff;
  double* Jt = g.Jt + bra.moff;
  double* Kt = g.Kt + bra.moff;
  const int fa = bra.fa + A0, fb = bra.fb;
  const bool ij = bra.same;
  const JkSegs* sg = g.segs + ch.seg;
  if (lds) {
    for (int e = tid; e < ROWS * n; e += nt) krow[e] = 0.0;
    __syncthreads();
  }
  // row r of the chunk's exchange rows: bra function fa + r (r < NA) or fb + r - NA
  auto kadd = [&](int row, int col, double v) {
    if (lds)
      lds_add(&krow[row * n + col], v);
    else
      atomic_add(&Kt[size_t(row < NA ? fa + row : fb + row - NA) * n + col], v);
  };
  double jab[NA][nb];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b) jab[a][b] = 0.0;

  for (int s0 = ch.k0; s0 < ch.k1; s0 += kSortRun) {
  const int nsel = sort_run<256>(g.kets, sg, g.pairs, g.sel, ch.bra, bra, s0, min(ch.k1, s0 + kSortRun), srt, nt);
  for (int r = 0; nt * r < nsel; ++r) {
    const int i = sorted_pos(tid, r, nt >> 6);
    if (i >= nsel) continue;
    const unsigned e = srt.kp[i];
    const int kp = int(e & ~kSortF32);
    const PairRec ket = g.pairs[kp];
    double gv[NA][nb][nc][nd];
    if constexpr (MIXED)
      quartet_block_regs<La, Lb, Lc, Ld, float, A0, A1>(g, bra, ket, false, gv);
    else
      quartet_block_regs<La, Lb, Lc, Ld, double, A0, A1, LB>(g, bra, ket, false, gv, lrows);
    const int fc = ket.fa, fd = ket.fb;
    const bool kl = ket.same, bk = (kp == ch.bra);
    const double fJ = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0);
    const double fK = bk ? 1.0 : 2.0;
    // Jt[ab] += fJ sum_cd g D[cd]  (summed over the chunk)
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < nb; ++b) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
          for (int d = 0; d < nd; ++d) s += gv[a][b][c][d] * D[size_t(fc + c) * n + fd + d];
        jab[a][b] += fJ * s;
      }
    if (!bk) {  // Jt[cd] += fJ sum_ab g D[ab]
#pragma unroll
      for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < nb; ++b) s += gv[a][b][c][d] * D[size_t(fa + a) * n + fb + b];
          atomic_add(&Jt[size_t(fc + c) * n + fd + d], fJ * s);
        }
    }
    // Kt[ac] += fK sum_bd g D[bd]
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int c = 0; c < nc; ++c) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < nb; ++b)
#pragma unroll
          for (int d = 0; d < nd; ++d) s += gv[a][b][c][d] * D[size_t(fb + b) * n + fd + d];
        kadd(a, fc + c, fK * s);
      }
    if (!kl) {  // Kt[ad] += fK sum_bc g D[bc]
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int d = 0; d < nd; ++d) {
          double s = 0.0;
#pragma unroll
          for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int c = 0; c < nc; ++c) s += gv[a][b][c][d] * D[size_t(fb + b) * n + fc + c];
          kadd(a, fd + d, fK * s);
        }
    }
    if (!ij) {  // Kt[bc] += fK sum_ad g D[ad]
#pragma unroll
      for (int b = 0; b < nb; ++b)
#pragma unroll
        for (int c = 0; c < nc; ++c) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int d = 0; d < nd; ++d) s += gv[a][b][c][d] * D[size_t(fa + a) * n + fd + d];
          kadd(NA + b, fc + c, fK * s);
        }
    }
    if (!ij && !kl) {  // Kt[bd] += fK sum_ac g D[ac]
#pragma unroll
      for (int b = 0; b < nb; ++b)
#pragma unroll
        for (int d = 0; d < nd; ++d) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int c = 0; c < nc; ++c) s += gv[a][b][c][d] * D[size_t(fa + a) * n + fc + c];
          kadd(NA + b, fd + d, fK * s);
        }
    }
  }
  }
  // Jt[ab]: one atomic per wavefront and element
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b) {
      const double v = wave_sum(jab[a][b]);
      if ((tid & 63) == 0 && v != 0.0) atomic_add(&Jt[size_t(fa + a) * n + fb + b], v);
    }
  if (lds) {
    __syncthreads();
    for (int e = tid; e < ROWS * n; e += nt) {
      const double v = krow[e];
      if (v != 0.0) {
        const int row = e / n, col = e - row * n;
        atomic_add(&Kt[size_t(row < NA ? fa + row : fb + row - NA) * n + col], v);
      }
    }
    __syncthreads();  // krow is reused by the next chunk
  }
}

template <int La, int Lb, int Lc, int Ld, bool MIXED, int CHA, int C, bool LB>
__device__ __forceinline__ void dispatch_jk_chunk(const EriArgs& g, const JkChunk& ch, double* krow, SortLds& srt,
                                                  int chunk, lds_cdptr lrows) {
  constexpr int na = ncart(La);
  constexpr int A0 = C * CHA;
  constexpr int A1 = (A0 + CHA < na) ? A0 + CHA : na;
  if (chunk == C) {
    jk_chunk_regs<La, Lb, Lc, Ld, MIXED, A0, A1, LB>(g, ch, krow, srt, lrows);
  } else if constexpr (A1 < na) {
    dispatch_jk_chunk<La, Lb, Lc, Ld, MIXED, CHA, C + 1, LB>(g, ch, krow, srt, chunk, lrows);
  }
}

// The Rys tables a class may copy to LDS (doubles): for one root the coarser
// rows of LAI_B9_OFF (step 1/8, nine Taylor coefficients of F_0); for two
// roots the Boys Taylor rows 0 .. 16 T_cut of F_0..F_9 (row stride padded to
// kRowStride); for three and four roots the Chebyshev coefficients
// of the rule below its cut (34 and 51 KB, interval stride padded to
// kChebStrideLds); none for more roots.
constexpr int lds_rys_doubles(int NR) {
  return NR == 1   ? LAI_B9_ROWS * 9
         : NR == 2 ? kBoysRowsN<2> * kRowStride<lds_cdptr, kBoysRowW<2>>
         : NR == 3 ? lai_rys_tcut[3] * kChebStrideLds<3>
         : NR == 4 ? lai_rys_tcut[4] * kChebStrideLds<4>
                   : 0;
}
template <int La, int Lb, int Lc, int Ld>
inline constexpr int kLdsRysDoubles = lds_rys_doubles((La + Lb + Lc + Ld) / 2 + 1);

// Copies the class's Rys tables into dst (kLdsRysDoubles doubles) with the
// workgroup's lanes; ends with a workgroup barrier.
template <int NR>
__device__ __forceinline__ void copy_rys_lds(const double* __restrict__ rys, double* dst) {
  constexpr int NB = lds_rys_doubles(NR);
  if constexpr (NR == 1) {
    // the coarser one-root rows, packed (odd stride 9)
    const double* src = rys + LAI_B9_OFF;
    for (int i = threadIdx.x; i < NB; i += blockDim.x) dst[i] = src[i];
  } else if constexpr (NR == 2) {
    constexpr int W = kBoysRowW<NR>, ST = kRowStride<lds_cdptr, W>;
    const double* src = rys + LAI_B10_OFF;
    for (int i = threadIdx.x; i < NB; i += blockDim.x) {
      const int k = i / ST, r = i - k * ST;
      dst[i] = r < W ? src[k * W + r] : 0.0;
    }
  } else {
    constexpr int W = 2 * NR * LAI_RYS_NTERMS, ST = kChebStrideLds<NR>;
    const double* src = rys + lai_rys_off[NR];
    for (int i = threadIdx.x; i < NB; i += blockDim.x) {
      const int k = i / ST, r = i - k * ST;
      dst[i] = r < W ? src[k * W + r] : 0.0;
    }
  }
  __syncthreads();
}

// Kernel A, anchored: 256 or 512 lanes per chunk (the launch picks the
// workgroup that keeps the most wavefronts resident with the exchange rows in
// LDS), bra index chunk C = blockIdx.y.  Dynamic LDS: with LB, the class's
// Rys tables, copied once per workgroup, then the exchange rows [CHA + nb][mn].
template <int La, int Lb, int Lc, int Ld, bool MIXED, int CHA, bool LB>
__global__ __launch_bounds__(512) void jk_kernel_regs(EriArgs g) {
  extern __shared__ double dyn[];
  __shared__ SortLds srt;
  constexpr int NB = LB ? kLdsRysDoubles<La, Lb, Lc, Ld> : 0;
  static_assert(!LB || NB > 0, "LDS Rys tables are for one- to four-root classes");
  lds_cdptr lrows = nullptr;
  if constexpr (LB) {
    copy_rys_lds<(La + Lb + Lc + Ld) / 2 + 1>(g.rys, dyn);
    lrows = (lds_cdptr)(dyn);
  }
  double* krow = dyn + NB;
  for (int c = blockIdx.x; c < g.nchunks; c += gridDim.x) {
    const JkChunk ch = g.chunks[c];
    dispatch_jk_chunk<La, Lb, Lc, Ld, MIXED, CHA, 0, LB>(g, ch, krow, srt, blockIdx.y, lrows);
  }
}

// Kernel A with merged bra slices: for classes whose bra index is cut into
// NCH slices (to fit the registers), a group of NCH wavefronts takes 64
// quartets of a chunk at a time, wavefront C of the group computing slice C
// of the same sorted quartets in step.  A workgroup holds rep such groups
// (rep * NCH wavefronts, at most 512 lanes; the launch picks rep), which take
// consecutive 64-quartet blocks of the sorted run.  The exchange rows of A
// and B, the sort and the chunk's loads are shared by the whole workgroup;
// the slices' partial ket Coulomb blocks Jt[cd] meet in LDS and go to HBM
// with one atomic per element instead of one per slice.
template <int La, int Lb, int Lc, int Ld, bool MIXED, int CHA, int C, bool LB = false>
__device__ __forceinline__ void jk_chunk_merged(const EriArgs& g, const JkChunk& ch, double* krow, SortLds& srt,
                                                double* jpart, lds_cdptr lrows = nullptr) {
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NCH = (na + CHA - 1) / CHA;
  constexpr int A0 = C * CHA, A1 = (A0 + CHA < na) ? A0 + CHA : na;
  constexpr int NA = A1 - A0;
  constexpr int NCD = nc * nd;
  constexpr int ROWS = na + nb;
  const int tid = threadIdx.x, lane = tid & 63, NT = blockDim.x;
  const int rep = NT / (64 * NCH), grp = (tid >> 6) / NCH;
  jpart += grp * (NCH * NCD * 64);  // this group's slices
  const bool lds = g.klds != 0;
  const PairRec bra = g.pairs[ch.bra];
  const int n = bra.mn;
  const double* __restrict__ D = g.D + bra.moff;
  double* Jt = g.Jt + bra.moff;
  double* Kt = g.Kt + bra.moff;
  const int fa = bra.fa, fb = bra.fb;
  const bool ij = bra.same;
  const JkSegs* sg = g.segs + ch.seg;
  if (lds) {
    for (int e = tid; e < ROWS * n; e += NT) krow[e] = 0.0;
    __syncthreads();
  }
  // row r of the exchange rows: bra function fa + r (r < na) or fb + r - na
  auto kadd = [&](int row, int col, double v) {
    if (lds)
      lds_add(&krow[row * n + col], v);
    else
      atomic_add(&Kt[size_t(row < na ? fa + row : fb + row - na) * n + col], v);
  };
  double jab[NA][nb];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b) jab[a][b] = 0.0;

  for (int s0 = ch.k0; s0 < ch.k1; s0 += kSortRun) {
    const int nsel =
        sort_run<64 * NCH>(g.kets, sg, g.pairs, g.sel, ch.bra, bra, s0, min(ch.k1, s0 + kSortRun), srt, NT);
    for (int r = 0; 64 * rep * r < nsel; ++r) {
      const int i = 64 * (rep * r + grp) + lane;
      const bool valid = i < nsel;
      int fc = 0, fd = 0;
      bool bk = true;
      if (valid) {
        const unsigned e = srt.kp[i];
        const int kp = int(e & ~kSortF32);
        const PairRec ket = g.pairs[kp];
        double gv[NA][nb][nc][nd];
        if constexpr (MIXED)
          quartet_block_regs<La, Lb, Lc, Ld, float, A0, A1>(g, bra, ket, false, gv);
        else
          quartet_block_regs<La, Lb, Lc, Ld, double, A0, A1, LB>(g, bra, ket, false, gv, lrows);
        fc = ket.fa;
        fd = ket.fb;
        const bool kl = ket.same;
        bk = (kp == ch.bra);
        const double fJ = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0);
        const double fK = bk ? 1.0 : 2.0;
        // Jt[ab] += fJ sum_cd g D[cd]  (this slice's rows, summed over the chunk)
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < nb; ++b) {
            double t = 0.0;
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
              for (int d = 0; d < nd; ++d) t += gv[a][b][c][d] * D[size_t(fc + c) * n + fd + d];
            jab[a][b] += fJ * t;
          }
        // this slice's part of Jt[cd] += fJ sum_ab g D[ab], to LDS
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
          for (int d = 0; d < nd; ++d) {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
              for (int b = 0; b < nb; ++b) t += gv[a][b][c][d] * D[size_t(fa + A0 + a) * n + fb + b];
            jpart[(C * NCD + c * nd + d) * 64 + lane] = fJ * t;
          }
        // Kt[ac] += fK sum_bd g D[bd]
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int c = 0; c < nc; ++c) {
            double t = 0.0;
#pragma unroll
            for (int b = 0; b < nb; ++b)
#pragma unroll
              for (int d = 0; d < nd; ++d) t += gv[a][b][c][d] * D[size_t(fb + b) * n + fd + d];
            kadd(A0 + a, fc + c, fK * t);
          }
        if (!kl) {  // Kt[ad] += fK sum_bc g D[bc]
#pragma unroll
          for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int d = 0; d < nd; ++d) {
              double t = 0.0;
#pragma unroll
              for (int b = 0; b < nb; ++b)
#pragma unroll
                for (int c = 0; c < nc; ++c) t += gv[a][b][c][d] * D[size_t(fb + b) * n + fc + c];
              kadd(A0 + a, fd + d, fK * t);
            }
        }
        if (!ij) {  // Kt[bc] += fK sum_ad g D[ad]  (this slice's a)
#pragma unroll
          for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int c = 0; c < nc; ++c) {
              double t = 0.0;
#pragma unroll
              for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int d = 0; d < nd; ++d) t += gv[a][b][c][d] * D[size_t(fa + A0 + a) * n + fd + d];
              kadd(na + b, fc + c, fK * t);
            }
        }
        if (!ij && !kl) {  // Kt[bd] += fK sum_ac g D[ac]  (this slice's a)
#pragma unroll
          for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int d = 0; d < nd; ++d) {
              double t = 0.0;
#pragma unroll
              for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int c = 0; c < nc; ++c) t += gv[a][b][c][d] * D[size_t(fa + A0 + a) * n + fc + c];
              kadd(na + b, fd + d, fK * t);
            }
        }
      }
      __syncthreads();  // every slice's part of Jt[cd] is in LDS
      if (valid && !bk) {
        // element e of the ket block: wavefront e mod NCH adds the slices' sum
        for (int e = C; e < NCD; e += NCH) {
          double t = 0.0;
#pragma unroll
          for (int w = 0; w < NCH; ++w) t += jpart[(w * NCD + e) * 64 + lane];
          atomic_add(&Jt[size_t(fc + e / nd) * n + fd + e % nd], t);
        }
      }
      __syncthreads();  // jpart is rewritten by the next round
    }
  }
  // Jt[ab] of this slice: one atomic per wavefront and element
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b) {
      const double v = wave_sum(jab[a][b]);
      if (lane == 0 && v != 0.0) atomic_add(&Jt[size_t(fa + A0 + a) * n + fb + b], v);
    }
  if (lds) {
    __syncthreads();
    for (int e = tid; e < ROWS * n; e += NT) {
      const double v = krow[e];
      if (v != 0.0) {
        const int row = e / n, col = e - row * n;
        atomic_add(&Kt[size_t(row < na ? fa + row : fb + row - na) * n + col], v);
      }
    }
    __syncthreads();  // krow is reused by the next chunk
  }
}

template <int La, int Lb, int Lc, int Ld, bool MIXED, int CHA, int C, bool LB = false>
__device__ __forceinline__ void dispatch_merged(const EriArgs& g, const JkChunk& ch, double* krow, SortLds& srt,
                                                double* jpart, int w, lds_cdptr lrows = nullptr) {
  constexpr int na = ncart(La);
  if (w == C) {
    jk_chunk_merged<La, Lb, Lc, Ld, MIXED, CHA, C, LB>(g, ch, krow, srt, jpart, lrows);
  } else if constexpr ((C + 1) * CHA < na) {
    dispatch_merged<La, Lb, Lc, Ld, MIXED, CHA, C + 1, LB>(g, ch, krow, srt, jpart, w, lrows);
  }
}

// Kernel A, merged slices: rep groups of NCH wavefronts per chunk, wavefront
// w = slice w mod NCH of group w / NCH.  Dynamic LDS: with LB the class's Rys
// tables, then the groups' ket Coulomb parts [rep][NCH][nc*nd][64], then the
// exchange rows [na + nb][mn]; static: the sort.
template <int La, int Lb, int Lc, int Ld, bool MIXED, int CHA, bool LB = false>
__global__ __launch_bounds__(512) void jk_kernel_merged(EriArgs g) {
  constexpr int NCH = (ncart(La) + CHA - 1) / CHA;
  constexpr int NB = LB ? kLdsRysDoubles<La, Lb, Lc, Ld> : 0;
  static_assert(!LB || NB > 0, "LDS Rys tables are for one- to four-root classes");
  extern __shared__ double dyn[];
  __shared__ SortLds srt;
  lds_cdptr lrows = nullptr;
  if constexpr (LB) {
    copy_rys_lds<(La + Lb + Lc + Ld) / 2 + 1>(g.rys, dyn);
    lrows = (lds_cdptr)(dyn);
  }
  double* jpart = dyn + NB;
  double* krow = jpart + (blockDim.x / 64) * (ncart(Lc) * ncart(Ld) * 64);
  const int w = (threadIdx.x >> 6) % NCH;  // uniform over the wavefront
  for (int c = blockIdx.x; c < g.nchunks; c += gridDim.x) {
    const JkChunk ch = g.chunks[c];
    dispatch_merged<La, Lb, Lc, Ld, MIXED, CHA, 0, LB>(g, ch, krow, srt, jpart, w, lrows);
  }
}

// Split of a gathered contraction with O outputs and NU x NV terms each: S
// lanes per output take the u rows u = part, part + S, ...; the S*O output
// slots fill R rounds of NL lanes.  Chosen for the fewest serial terms per
// lane, a round counting as two terms (its index set-up and LDS add).
struct GatherSplit {
  int s, r;
};
constexpr GatherSplit gather_split(int O, int NU, int NV, int NL) {
  GatherSplit best{1, (O + NL - 1) / NL};
  int best_cost = 1 << 30;
  for (int s = 1; s <= NU; ++s) {
    const int r = (O * s + NL - 1) / NL;
    const int cost = r * (((NU + s - 1) / s) * NV + 2);
    if (cost < best_cost) {
      best_cost = cost;
      best = GatherSplit{s, r};
    }
  }
  return best;
}

// One contraction of a quartet by NL lanes (a wavefront or a workgroup),
// gathered by its outputs:
//   out[OO + p1*N2 + p2] += f * sum_{u<NU, v<NV} X[p1*X1 + p2*X2 + u*XU + v*XV]
//                                              * Dk[OD + u*DU + v*DV]
// for the N1*N2 outputs (p1, p2).  X holds the quartet's normalised integrals
// X[((a*nb + b)*nc + c)*nd + d]; Dk the staged density blocks.  An output's
// u range is split over S lanes (S-way LDS adds); S = 1 writes are exclusive.
template <int NL, int N1, int N2, int NU, int NV, int X1, int X2, int XU, int XV, int OD, int DU, int DV, int OO>
__device__ __forceinline__ void gather_contract(const double* X, const double* Dk, double* out, double f, int lane) {
  constexpr int O = N1 * N2;
  constexpr GatherSplit sp = gather_split(O, NU, NV, NL);
  constexpr int S = sp.s, R = sp.r, F = NU / S, REM = NU % S;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int slot = lane + NL * r;
    const int o = slot % O, part = slot / O;
    if (part < S) {
      const int p1 = o / N2, p2 = o - p1 * N2;
      const double* x = X + p1 * X1 + p2 * X2 + part * XU;
      const double* dk = Dk + OD + part * DU;
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < F; ++i)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc += x[S * i * XU + v * XV] * dk[S * i * DU + v * DV];
      if constexpr (REM > 0) {
        if (part < REM) {
#pragma unroll
          for (int v = 0; v < NV; ++v) acc += x[S * F * XU + v * XV] * dk[S * F * DU + v * DV];
        }
      }
      if constexpr (S == 1)
        out[OO + o] += f * acc;
      else
        lds_add(&out[OO + o], f * acc);
    }
  }
}

// The six contractions of a quartet's integrals X with its density blocks
// (at the offsets Oxy of the J/K blocks of the same shape), each gathered by
// output over NL lanes: J[ab] = sum_cd X D[cd], J[cd] = sum_ab X D[ab],
// K[ac] = sum_bd X D[bd], K[ad] = sum_bc X D[bc], K[bc] = sum_ad X D[ad],
// K[bd] = sum_ac X D[ac]; the same-pair and bra-equals-ket cases drop the
// duplicates.
template <int NL, int La, int Lb, int Lc, int Ld>
__device__ __forceinline__ void gather_all(const double* X, const double* Dk, double* out, double fJ, double fK,
                                           bool ij, bool kl, bool bk, int lane) {
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int OAB = 0, OCD = na * nb, OAC = OCD + nc * nd, OAD = OAC + na * nc, OBC = OAD + na * nd,
                OBD = OBC + nb * nc;
  constexpr int SA = nb * nc * nd, SB = nc * nd, SC = nd;  // strides of a, b, c in X
  gather_contract<NL, na, nb, nc, nd, SA, SB, SC, 1, OCD, nd, 1, OAB>(X, Dk, out, fJ, lane);
  gather_contract<NL, na, nc, nb, nd, SA, SC, SB, 1, OBD, nd, 1, OAC>(X, Dk, out, fK, lane);
  if (!ij) gather_contract<NL, nb, nc, na, nd, SB, SC, SA, 1, OAD, nd, 1, OBC>(X, Dk, out, fK, lane);
  if (!bk) gather_contract<NL, nc, nd, na, nb, SC, 1, SA, SB, OAB, nb, 1, OCD>(X, Dk, out, fJ, lane);
  if (!kl) gather_contract<NL, na, nd, nb, nc, SA, 1, SB, SC, OBC, nc, 1, OAD>(X, Dk, out, fK, lane);
  if (!ij && !kl) gather_contract<NL, nb, nd, na, nc, SB, 1, SA, SC, OAC, nc, 1, OBD>(X, Dk, out, fK, lane);
}

// Kernel B: one workgroup per chunk, the chunk's quartets one after another.
// Each quartet's ket Coulomb and exchange blocks go to HBM with one atomic per
// element; the bra Coulomb block Jt[ab] is summed in LDS over the chunk.
// Where the quartet's integrals fit in the table storage, all six
// contractions are gathered by output (gather_all).
// Wavefronts per SIMD that bound the registers of the workgroup kernel: three
// for at most 4 accumulators per lane, two up to the largest classes whose
// integrals fit in LDS, one for the rest.
template <int La, int Lb, int Lc, int Ld, int NT>
constexpr int lds_kernel_wpe() {
  constexpr int NJ = (LdsTasks<La, Lb, Lc, Ld>::NTASK + NT - 1) / NT;
  constexpr int nd = ncart(Ld);
  return NJ * nd <= 4 ? 3 : (LdsTasks<La, Lb, Lc, Ld>::NTASK * nd <= 6144 ? 2 : 1);
}

template <int La, int Lb, int Lc, int Ld, bool MIXED, int NT>
__global__ __launch_bounds__(NT, (lds_kernel_wpe<La, Lb, Lc, Ld, NT>())) void jk_kernel_lds(EriArgs g) {
  using Sh = LdsShape<La, Lb, Lc, Ld, double>;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NTASK = LdsTasks<La, Lb, Lc, Ld>::NTASK;
  constexpr int NJ = (NTASK + NT - 1) / NT;
  constexpr int NPQ = Sh::template npq<NT>();
  // FP64 tables (the FP32 recursion uses the same storage as float), then
  // the quartet's normalised integrals X[t*nd + d] by task t = (a*nb + b)*nc
  // + c for the gathered contractions: the storage grows to hold X up to 48
  // KiB (these classes' register demand leaves at most two workgroups per CU)
  constexpr int XN = NTASK * nd;
  constexpr int TABN = NPQ * Sh::TSP > XN || XN > 6144 ? NPQ * Sh::TSP : XN;
  __shared__ double tab[TABN];
  constexpr bool XLDS = XN <= TABN;
  // the quartet's six J/K blocks, summed in LDS and added to HBM once per
  // quartet (Jt[cd], Kt[ac], Kt[ad], Kt[bc], Kt[bd]) or once per chunk (Jt[ab])
  constexpr int OAB = 0, OCD = na * nb, OAC = OCD + nc * nd, OAD = OAC + na * nc, OBC = OAD + na * nd,
                OBD = OBC + nb * nc, NBLK = OBD + nb * nd;
  __shared__ double blk[NBLK];
  __shared__ double dblk[NBLK];  // the quartet's density blocks D[xy], laid out like blk
  const int tid = threadIdx.x;
  int ofs[NJ];
  lds_component_offsets<La, Lb, Lc, Ld, NT, NJ, false>(tid, ofs);
  // normalisation of this lane's (a, b, c) tasks, computed once
  double sabc[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int t = min(tid + NT * j, NTASK - 1);
    const int c = t % nc, b = (t / nc) % nb, a = t / (nc * nb);
    sabc[j] = norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c);
  }
  for (int e = tid; e < NBLK; e += NT) blk[e] = 0.0;  // ordered by the barriers of lds_quartet_acc
  // the density element this lane stages in each of its NDS slots o = tid +
  // NT*i: row fr[rs] + ro, column fr[cs] + co of the shells' first functions
  // fr = (fa, fb, fc, fd), packed rs | cs << 2 | ro << 4 | co << 8
  constexpr int NDS = (NBLK + NT - 1) / NT;
  int dsl[NDS];
#pragma unroll
  for (int i = 0; i < NDS; ++i) {
    const int o = tid + NT * i;
    int rs, cs, ro, co;
    if (o < OCD) { rs = 0; cs = 1; ro = (o - OAB) / nb; co = (o - OAB) % nb; }
    else if (o < OAC) { rs = 2; cs = 3; ro = (o - OCD) / nd; co = (o - OCD) % nd; }
    else if (o < OAD) { rs = 0; cs = 2; ro = (o - OAC) / nc; co = (o - OAC) % nc; }
    else if (o < OBC) { rs = 0; cs = 3; ro = (o - OAD) / nd; co = (o - OAD) % nd; }
    else if (o < OBD) { rs = 1; cs = 2; ro = (o - OBC) / nc; co = (o - OBC) % nc; }
    else { rs = 1; cs = 3; ro = (o - OBD) / nd; co = (o - OBD) % nd; }
    dsl[i] = o < NBLK ? (rs | cs << 2 | ro << 4 | co << 8) : -1;
  }

  for (int ci = blockIdx.x; ci < g.nchunks; ci += gridDim.x) {
    const JkChunk ch = g.chunks[ci];
    const PairRec bra = load_pair_uniform(g.pairs, ch.bra);
    const int n = bra.mn;
    const double* __restrict__ D = g.D + bra.moff;
    double* Jt = g.Jt + bra.moff;
    double* Kt = g.Kt + bra.moff;
    const int fa = bra.fa, fb = bra.fb;
    const bool ij = bra.same;
    const JkSegs* sg = g.segs + ch.seg;
    int seg = 0;
    for (int k = ch.k0; k < ch.k1; ++k) {
      const int kp = chunk_ket(g.kets, sg, k, seg);
      const PairRec ket = load_pair_uniform(g.pairs, kp);
      const int mode = quartet_select(g.sel, ch.bra, bra, kp, ket);
      if (mode == 0) continue;  // uniform over the workgroup
      const int fc = ket.fa, fd = ket.fb;
      // the quartet's six density blocks, loaded now and staged after the
      // products so that the loads' latency is hidden behind them
      double dpre[NDS];
#pragma unroll
      for (int i = 0; i < NDS; ++i) {
        const int pk = dsl[i];
        if (pk >= 0) {
          const int rs = pk & 3, cs = (pk >> 2) & 3;
          const int r = (rs == 0 ? fa : rs == 1 ? fb : fc) + ((pk >> 4) & 15);
          const int q = (cs == 1 ? fb : cs == 2 ? fc : fd) + (pk >> 8);
          dpre[i] = D[size_t(r) * n + q];
        }
      }
      double v[NJ][nd];
      if constexpr (MIXED) {
        float acc[NJ][nd];
        lds_quartet_acc<La, Lb, Lc, Ld, float, NT, NJ, NPQ>(g, bra, ket, false, reinterpret_cast<float*>(tab), ofs,
                                                            acc);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int d = 0; d < nd; ++d) v[j][d] = acc[j][d];
      } else {
        double acc[NJ][nd];
        lds_quartet_acc<La, Lb, Lc, Ld, double, NT, NJ, NPQ>(g, bra, ket, false, tab, ofs, acc);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int d = 0; d < nd; ++d) v[j][d] = acc[j][d];
      }
      const bool kl = ket.same, bk = (kp == ch.bra);
      const double fJ = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0);
      const double fK = bk ? 1.0 : 2.0;
      // the quartet's six density blocks D[xy] at the offsets Oxy of the J/K
      // blocks of the same shape (the previous quartet's readers have passed
      // the barriers of lds_quartet_acc)
#pragma unroll
      for (int i = 0; i < NDS; ++i)
        if (dsl[i] >= 0) dblk[tid + NT * i] = dpre[i];
      // the normalised integrals of this lane's tasks through the table
      // storage, free after the products
      if constexpr (XLDS) {
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (tid + NT * j < NTASK)
#pragma unroll
            for (int d = 0; d < nd; ++d) tab[(tid + NT * j) * nd + d] = v[j][d] * sabc[j] * norm_corr(Ld, d);
      }
      __syncthreads();
      if constexpr (XLDS) {
        gather_all<NT, La, Lb, Lc, Ld>(tab, dblk, blk, fJ, fK, ij, kl, bk, tid);
      } else {
        // (ff|ff): the integrals stay in registers, and each task's terms
        // are scattered with LDS adds (an output shared by up to nb*nc lanes)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int t = tid + NT * j;
          if (t >= NTASK) continue;
          const int c = t % nc, b = (t / nc) % nb, a = t / (nc * nb);
          double x[nd];
#pragma unroll
          for (int d = 0; d < nd; ++d) x[d] = v[j][d] * sabc[j] * norm_corr(Ld, d);
          double sj = 0.0, s1 = 0.0, s3 = 0.0;
#pragma unroll
          for (int d = 0; d < nd; ++d) {
            sj += x[d] * dblk[OCD + c * nd + d];
            s1 += x[d] * dblk[OBD + b * nd + d];
            s3 += x[d] * dblk[OAD + a * nd + d];
          }
          lds_add(&blk[OAB + a * nb + b], fJ * sj);
          lds_add(&blk[OAC + a * nc + c], fK * s1);
          if (!ij) lds_add(&blk[OBC + b * nc + c], fK * s3);
          const double dab = dblk[OAB + a * nb + b], dbc = dblk[OBC + b * nc + c], dac = dblk[OAC + a * nc + c];
#pragma unroll
          for (int d = 0; d < nd; ++d) {
            if (!bk) lds_add(&blk[OCD + c * nd + d], fJ * x[d] * dab);
            if (!kl) lds_add(&blk[OAD + a * nd + d], fK * x[d] * dbc);
            if (!ij && !kl) lds_add(&blk[OBD + b * nd + d], fK * x[d] * dac);
          }
        }
      }
      __syncthreads();
      // the quartet's ket and exchange blocks to HBM, one atomic per element
      for (int e = OCD + tid; e < NBLK; e += NT) {
        const double x = blk[e];
        if (x == 0.0) continue;
        blk[e] = 0.0;
        size_t r, q;
        double* M = Kt;
        if (e < OAC) { r = fc + (e - OCD) / nd; q = fd + (e - OCD) % nd; M = Jt; }
        else if (e < OAD) { r = fa + (e - OAC) / nc; q = fc + (e - OAC) % nc; }
        else if (e < OBC) { r = fa + (e - OAD) / nd; q = fd + (e - OAD) % nd; }
        else if (e < OBD) { r = fb + (e - OBC) / nc; q = fc + (e - OBC) % nc; }
        else { r = fb + (e - OBD) / nd; q = fd + (e - OBD) % nd; }
        atomic_add(&M[r * n + q], x);
      }
      __syncthreads();
    }
    // the chunk's bra Coulomb block
    for (int e = tid; e < OCD; e += NT) {
      const double x = blk[e];
      if (x == 0.0) continue;
      blk[e] = 0.0;
      atomic_add(&Jt[size_t(fa + e / nb) * n + fb + e % nb], x);
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Kernel W: NW wavefronts per workgroup, one quartet per wavefront.  Each
// wave takes every NW-th quartet of the chunk, builds its own (primitive
// quartet, root, axis) tables and accumulates its (a, b, c) tasks; the
// quartet's ket Coulomb and exchange blocks go to HBM with one atomic per
// element, the bra Coulomb block Jt[ab] once per chunk.  The waves run
// independently: no workgroup barrier, wavefront-level ordering of the
// wave's own LDS.  (Keeping the bra's exchange rows in LDS, as kernel A
// does, costs more in occupancy than it saves in atomics for these classes:
// measured 1.1-1.7x slower on 400-1400 basis functions.)
// --------------------------------------------------------------------------
template <int La, int Lb, int Lc, int Ld, typename R, int NJ, int NPQ>
__device__ __forceinline__ void wave_quartet_acc(const EriArgs& g, const PairRec& bra, const PairRec& ket, int ntot,
                                                 R* tab, int lane, const int (&ofs)[NJ], R (&acc)[NJ][ncart(Ld)]) {
  using Sh = LdsShape<La, Lb, Lc, Ld, R>;
  constexpr int nd = ncart(Ld);
  constexpr int NR = Sh::NR, S = Sh::S, TASKS = 3 * NR;
  // tables one (root, axis) apart by an odd number of entries (bank spread of
  // the table lanes' writes); SABC apart between the d entries of a row
  constexpr int SP = S | 1, TS = TASKS * SP, SABC = (La + 1) * (Lb + 1) * (Lc + 1);
  static_assert(NPQ * TASKS <= 64, "not enough lanes for the table batch");
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int d = 0; d < nd; ++d) acc[j][d] = R(0);
  const int npp_k = ket.npp_sig;
  constexpr bool fp32 = sizeof(R) == 4;
  const float qcut = fp32 ? float(g.pcut32 * bra.bsum * ket.bsum) : 0.0f;
  wave_sync();  // the previous user of tab and hs has finished
  for (int t0 = 0; t0 < ntot; t0 += NPQ) {
    if (lane < NPQ * TASKS) {
      const int pq = lane / TASKS, task = lane - pq * TASKS;
      const int r = task / 3, ax = task - 3 * r;
      const int t = t0 + pq;
      R* tb = tab + pq * TS + task * SP;
      bool live = t < ntot;
      const int ib = live ? t / npp_k : 0, ik = live ? t - ib * npp_k : 0;
      if (live && fp32 && g.ppb[bra.pp_off + ib] * g.ppb[ket.pp_off + ik] < qcut) live = false;
      if (live) {
        const PrimPair P = g.pp[bra.pp_off + ib];
        const PrimPair Q = g.pp[ket.pp_off + ik];
        const double zeta = P.zeta, eta = Q.zeta;
        const double izpe = 1.0 / (zeta + eta);
        const double rz = eta * izpe, re = zeta * izpe;
        const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
        const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
        double ur, wr;
        rys_root_one<NR>(T, g.rys, r, ur, wr);
        const double h_z = g.ppih[bra.pp_off + ib], h_e = g.ppih[ket.pp_off + ik];
        const R B00 = R(0.5 * ur * izpe);
        const R B10 = R(h_z - h_z * rz * ur);
        const R B01 = R(h_e - h_e * re * ur);
        const double PQ = pick3(ax, PQx, PQy, PQz);
        const double PA = pick3(ax, P.PAx, P.PAy, P.PAz);
        const double QC = pick3(ax, Q.PAx, Q.PAy, Q.PAz);
        const double AB = pick3(ax, bra.ABx, bra.ABy, bra.ABz);
        const double CD = pick3(ax, ket.ABx, ket.ABy, ket.ABz);
        const double sc = ax == 2 ? P.Kc * Q.Kc * sqrt(izpe) * wr : 1.0;
        build_2d_inplace<La, Lb, Lc, Ld, R>(tb, R(PA - rz * ur * PQ), R(QC + re * ur * PQ), R(AB), R(CD), B10,
                                            B01, B00, R(sc));
      } else {
        for (int e = 0; e < S; ++e) tb[e] = R(0);  // contributes nothing
      }
    }
    wave_sync();
    const int rem = ntot - t0;
    const int npq = rem < NPQ ? (rem > 0 ? rem : 0) : NPQ;
#pragma unroll 1
    for (int pq = 0; pq < npq; ++pq) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const R* tx = tab + pq * TS + (r * 3) * SP;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int p = ofs[j];
          R X[Ld + 1], Y[Ld + 1], Z[Ld + 1];
#pragma unroll
          for (int e = 0; e <= Ld; ++e) {
            X[e] = tx[(p & 1023) + e * SABC];
            Y[e] = tx[SP + ((p >> 10) & 1023) + e * SABC];
            Z[e] = tx[2 * SP + (p >> 20) + e * SABC];
          }
#pragma unroll
          for (int d = 0; d < nd; ++d)
            acc[j][d] += X[cart_comp(Ld, d, 0)] * Y[cart_comp(Ld, d, 1)] * Z[cart_comp(Ld, d, 2)];
        }
      }
    }
    wave_sync();  // products finished with tab before the next batch
  }
}

#ifndef LAI_WAVE_TAB_KB
#define LAI_WAVE_TAB_KB 12
#endif
// Tables per wave of the wavefront kernel, in KB: LAI_WAVE_TAB_KB, or 8 where
// the larger budget would leave the class's LDS holding fewer wavefronts per
// SIMD than its registers allow and the smaller one holds more (more resident
// waves for fewer primitive quartets per step; measured 1.08x on (fp|ds),
// (fs|dp) and (fd|ps), and slower on the classes it does not apply to).  The
// occupancy arithmetic is that of WaveLds below.
template <int La, int Lb, int Lc, int Ld>
constexpr int wave_tab_kb() {
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NR = (La + Lb + Lc + Ld) / 2 + 1;
  constexpr int S = (La + 1) * (Lb + 1) * (Lc + 1) * (Ld + 1);
  constexpr int TS = 3 * NR * (S | 1);
  constexpr int NJ = (na * nb * nc + 63) / 64;
  constexpr int NBLK = na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd;
  constexpr int wpe_regs = (NJ * nd <= 9 && NJ <= 3) ? 3 : 2;
  auto wpe_lds = [&](int kb) {
    int npq = kb * 1024 / (TS * 8);
    if (npq > 64 / (3 * NR)) npq = 64 / (3 * NR);
    if (npq < 1) npq = 1;
    const int nw = npq * TS * 8 <= kb * 1024 ? 4 : 2;
    const int tabn = npq * TS > NJ * nd * 64 ? npq * TS : NJ * nd * 64;
    return (160 * 1024 / (nw * (tabn + 2 * NBLK) * 8)) * nw / 4;
  };
  const int big = wpe_lds(LAI_WAVE_TAB_KB), small = wpe_lds(8);
  return (big < wpe_regs && small > big) ? 8 : LAI_WAVE_TAB_KB;
}

template <int La, int Lb, int Lc, int Ld>
struct WaveShape {
  static constexpr int NTASK = LdsTasks<La, Lb, Lc, Ld>::NTASK;
  static constexpr int NJ = (NTASK + 63) / 64;
  static constexpr int KB = wave_tab_kb<La, Lb, Lc, Ld>();
  // table entries of one primitive quartet (odd stride per (root, axis))
  static constexpr int TS = 3 * LdsShape<La, Lb, Lc, Ld, double>::NR * (LdsShape<La, Lb, Lc, Ld, double>::S | 1);
  // primitive quartets per wave and step, within KB of tables per wave
  static constexpr int NPQ = [] {
    using Sh = LdsShape<La, Lb, Lc, Ld, double>;
    constexpr int by_lanes = 64 / (3 * Sh::NR);
    constexpr int by_lds = int((KB * 1024) / (3 * Sh::NR * (Sh::S | 1) * sizeof(double)));
    constexpr int n = by_lanes < by_lds ? by_lanes : by_lds;
    return n < 1 ? 1 : n;
  }();
};

// LDS of the wavefront kernel: per wave the tables (reused for the quartet's
// integrals) and the J/K and density blocks; the wavefronts per SIMD that it
// leaves room for bound the registers the compiler may take.
template <int La, int Lb, int Lc, int Ld, int NW>
struct WaveLds {
  using Ws = WaveShape<La, Lb, Lc, Ld>;
  static constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  static constexpr int NBLK = na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd;
  static constexpr int TABN = Ws::NPQ * Ws::TS > Ws::NJ * nd * 64 ? Ws::NPQ * Ws::TS : Ws::NJ * nd * 64;
  static constexpr int bytes = NW * (TABN + 2 * NBLK) * 8;
  static constexpr int wpe_lds = (160 * 1024 / bytes) * NW / 4;
  // three wavefronts per SIMD (168 registers) for at most 9 accumulators per
  // lane in at most 3 task slots, two (256) otherwise: the products' register
  // demand grows with the accumulators
  static constexpr int wpe_regs = (Ws::NJ * nd <= 9 && Ws::NJ <= 3) ? 3 : 2;
  static constexpr int wpe = wpe_lds < wpe_regs ? wpe_lds : wpe_regs;
  static constexpr int WPE = wpe < 1 ? 1 : wpe;
};

template <int La, int Lb, int Lc, int Ld, bool MIXED, int NW>
__global__ __launch_bounds__(64 * NW, (WaveLds<La, Lb, Lc, Ld, NW>::WPE)) void jk_kernel_waves(EriArgs g) {
  using Ws = WaveShape<La, Lb, Lc, Ld>;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NTASK = Ws::NTASK, NJ = Ws::NJ, NPQ = Ws::NPQ;
  constexpr int OAB = 0, OCD = na * nb, OAC = OCD + nc * nd, OAD = OAC + na * nc, OBC = OAD + na * nd,
                OBD = OBC + nb * nc, NBLK = OBD + nb * nd;
  // per wave: the (primitive quartet, root, axis) tables, then the quartet's
  // normalised integrals for the digestion
  constexpr int TABN = WaveLds<La, Lb, Lc, Ld, NW>::TABN;
  __shared__ double tab[NW][TABN];
  __shared__ double blk[NW][NBLK];   // per wave: the quartet's J/K blocks; Jt[ab] summed over the chunk
  __shared__ double dblk[NW][NBLK];  // per wave: the quartet's density blocks, laid out like blk
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int ofs[NJ];
  lds_component_offsets<La, Lb, Lc, Ld, 64, NJ, false>(lane, ofs);
  // normalisation of this lane's (a, b, c) tasks, computed once
  double sabc[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int t = min(lane + 64 * j, NTASK - 1);
    const int c = t % nc, b = (t / nc) % nb, a = t / (nc * nb);
    sabc[j] = norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c);
  }
  double* bw = blk[wave];
  double* dw = dblk[wave];
  for (int o = lane; o < NBLK; o += 64) bw[o] = 0.0;  // ordered by the first wave_sync below
  // the density element this lane stages in each of its NDS slots o =
  // lane + 64*i: row fr[rs] + ro, column fr[cs] + co of the shells' first
  // functions fr = (fa, fb, fc, fd), packed rs | cs << 2 | ro << 4 | co << 8
  constexpr int NDS = (NBLK + 63) / 64;
  int dsl[NDS];
#pragma unroll
  for (int i = 0; i < NDS; ++i) {
    const int o = lane + 64 * i;
    int rs, cs, ro, co;
    if (o < OCD) { rs = 0; cs = 1; ro = (o - OAB) / nb; co = (o - OAB) % nb; }
    else if (o < OAC) { rs = 2; cs = 3; ro = (o - OCD) / nd; co = (o - OCD) % nd; }
    else if (o < OAD) { rs = 0; cs = 2; ro = (o - OAC) / nc; co = (o - OAC) % nc; }
    else if (o < OBC) { rs = 0; cs = 3; ro = (o - OAD) / nd; co = (o - OAD) % nd; }
    else if (o < OBD) { rs = 1; cs = 2; ro = (o - OBC) / nc; co = (o - OBC) % nc; }
    else { rs = 1; cs = 3; ro = (o - OBD) / nd; co = (o - OBD) % nd; }
    dsl[i] = o < NBLK ? (rs | cs << 2 | ro << 4 | co << 8) : -1;
  }

  for (int ci = blockIdx.x; ci < g.nchunks; ci += gridDim.x) {
    const JkChunk ch = g.chunks[ci];
    const PairRec bra = load_pair_uniform(g.pairs, ch.bra);
    const int n = bra.mn;
    const double* __restrict__ D = g.D + bra.moff;
    double* Jt = g.Jt + bra.moff;
    double* Kt = g.Kt + bra.moff;
    const int fa = bra.fa, fb = bra.fb;
    const bool ij = bra.same;
    const JkSegs* sg = g.segs + ch.seg;
    int seg = 0;
    for (int k = ch.k0 + wave; k < ch.k1; k += NW) {  // uniform over the wavefront
      const int kp = chunk_ket(g.kets, sg, k, seg);
      const PairRec ket = load_pair_uniform(g.pairs, kp);
      const int mode = quartet_select(g.sel, ch.bra, bra, kp, ket);
      if (mode == 0) continue;
      const int ntot = bra.npp_sig * ket.npp_sig;
      const int fc = ket.fa, fd = ket.fb;
      // the quartet's six density blocks, loaded now and staged after the
      // products so that the loads' latency is hidden behind them
      double dpre[NDS];
#pragma unroll
      for (int i = 0; i < NDS; ++i) {
        const int pk = dsl[i];
        if (pk >= 0) {
          const int rs = pk & 3, cs = (pk >> 2) & 3;
          const int r = (rs == 0 ? fa : rs == 1 ? fb : fc) + ((pk >> 4) & 15);
          const int q = (cs == 1 ? fb : cs == 2 ? fc : fd) + (pk >> 8);
          dpre[i] = D[size_t(r) * n + q];
        }
      }
      double v[NJ][nd];
      if constexpr (MIXED) {
        float acc[NJ][nd];
        wave_quartet_acc<La, Lb, Lc, Ld, float, NJ, NPQ>(g, bra, ket, ntot, reinterpret_cast<float*>(tab[wave]), lane,
                                                         ofs, acc);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int d = 0; d < nd; ++d) v[j][d] = acc[j][d];
      } else {
        double acc[NJ][nd];
        wave_quartet_acc<La, Lb, Lc, Ld, double, NJ, NPQ>(g, bra, ket, ntot, tab[wave], lane, ofs, acc);
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int d = 0; d < nd; ++d) v[j][d] = acc[j][d];
      }
      const bool kl = ket.same, bk = (kp == ch.bra);
      const double fJ = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0);
      const double fK = bk ? 1.0 : 2.0;
      // the six density blocks D[xy] at the offsets of the J/K blocks of the
      // same shape (Oxy)
#pragma unroll
      for (int i = 0; i < NDS; ++i)
        if (dsl[i] >= 0) dw[lane + 64 * i] = dpre[i];
      // the normalised integrals X[t*nd + d] of task t = (a*nb + b)*nc + c,
      // through the wave's table storage (free after the products)
      double* xw = tab[wave];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
        if (lane + 64 * j < NTASK)
#pragma unroll
          for (int d = 0; d < nd; ++d) xw[(lane + 64 * j) * nd + d] = v[j][d] * sabc[j] * norm_corr(Ld, d);
      wave_sync();
      gather_all<64, La, Lb, Lc, Ld>(xw, dw, bw, fJ, fK, ij, kl, bk, lane);
      wave_sync();  // the wave's quartet block is complete
      // this quartet's ket and exchange blocks to HBM, one atomic per element;
      // Jt[ab] stays in LDS for the rest of the chunk
      for (int o = OCD + lane; o < NBLK; o += 64) {
        const double x = bw[o];
        if (x == 0.0) continue;
        bw[o] = 0.0;
        size_t r, q;
        double* M = Kt;
        if (o < OAC) { r = fc + (o - OCD) / nd; q = fd + (o - OCD) % nd; M = Jt; }
        else if (o < OAD) { r = fa + (o - OAC) / nc; q = fc + (o - OAC) % nc; }
        else if (o < OBC) { r = fa + (o - OAD) / nd; q = fd + (o - OAD) % nd; }
        else if (o < OBD) { r = fb + (o - OBC) / nc; q = fc + (o - OBC) % nc; }
        else { r = fb + (o - OBD) / nd; q = fd + (o - OBD) % nd; }
        atomic_add(&M[r * n + q], x);
      }
      wave_sync();  // the blocks are clear for the wave's next quartet
    }
    // the wave's share of the chunk's bra Coulomb block
    for (int o = lane; o < OCD; o += 64) {
      const double x = bw[o];
      if (x == 0.0) continue;
      bw[o] = 0.0;
      atomic_add(&Jt[size_t(fa + o / nb) * n + fb + o % nb], x);
    }
    wave_sync();
  }
}

// --------------------------------------------------------------------------
// Per-class execution policy
// --------------------------------------------------------------------------
template <int La, int Lb, int Lc, int Ld>
struct ClassPolicy {
  static constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  static constexpr int rest = nb * nc * nd;
  // lane-per-quartet classes: at most 27 outputs per lane in registers
  // decide the shape; a class whose bra is cut into slices then takes up to
  // 54 per lane (half the slices, each of which repeats the roots and the 2D
  // recursion: measured faster despite scratch spills at the 256-register cap
  // of a 512-lane workgroup)
  static constexpr int CHA0 = rest >= 27 ? 1 : (27 / rest < na ? 27 / rest : na);
  static constexpr int NCH0 = (na + CHA0 - 1) / CHA0;
  static constexpr bool regs = rest <= 27 && NCH0 <= 6;
  static constexpr int CHA1 = regs && NCH0 > 1 ? (54 / rest < na ? 54 / rest : na) : CHA0;
  static constexpr int NCH = (na + CHA1 - 1) / CHA1;
  // the bra components spread evenly over the slices (a group of slice
  // wavefronts finishes with its largest slice)
  static constexpr int CHA = (na + NCH - 1) / NCH;
  static constexpr int NOUT = na * nb * nc * nd;
  static constexpr int NT = 256;  // lanes of the workgroup-per-quartet shape
  // LDS-shape classes of at most 6 (a, b, c) tasks per lane of a wavefront
  // run one quartet per wavefront, four wavefronts per workgroup
  static constexpr bool waves = !regs && LdsTasks<La, Lb, Lc, Ld>::NTASK <= 6 * 64;
  // four wavefronts, two where one wavefront's tables exceed the per-wave budget
  static constexpr int NW =
      WaveShape<La, Lb, Lc, Ld>::NPQ * WaveShape<La, Lb, Lc, Ld>::TS * 8 <= WaveShape<La, Lb, Lc, Ld>::KB * 1024 ? 4
                                                                                                        : 2;
  // shape codes: 0 lane per quartet, 1 wavefront per quartet, 2 workgroup per quartet
  static constexpr int shape = regs ? 0 : (waves ? 1 : 2);
  // the second shape compiled for the autotuner: register classes that split
  // the bra over several workgroups may run a wavefront per quartet instead,
  // wavefront classes a workgroup per quartet
  static constexpr int alt =
      (regs && NCH >= 2 && LdsTasks<La, Lb, Lc, Ld>::NTASK <= 6 * 64) ? 1 : (waves ? 2 : -1);
};

}  // namespace lai
