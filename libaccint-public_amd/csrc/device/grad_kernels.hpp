// grad_kernels.hpp -- nuclear gradient of E = 1/2 tr(DJ) - x/4 tr(DK) for a
// fixed AO density, contracted on the fly (derivative integrals are recomputed,
// never stored).
//
// For a canonical quartet q with orbit size deg(q) = s_ij s_kl s_bk, the
// contribution to dE/dR is
//   deg(q) sum_abcd d(ab|cd)/dR * W_abcd,
//   W_abcd = 1/2 D_ab D_cd - x/8 (D_ac D_bd + D_ad D_bc)
// (the orbit average of the J and K weights).  The derivative of a Cartesian
// Gaussian on centre A is d/dA_x phi_a = 2 alpha phi_{a+1x} - a_x phi_{a-1x},
// applied per primitive in the Rys 2D tables, which are built one unit higher
// in a, b and c.  The fourth centre follows from translational invariance,
// dE/dD = -(dE/dA + dE/dB + dE/dC).
// Shape: one workgroup per shell quartet, (root, axis) tables and the weights
// W in LDS, each lane a strided subset of the Cartesian components.
#pragma once

#include "eri_kernels.hpp"

namespace lai {

struct GradArgs {
  const Quartet* __restrict__ q;
  int nq;
  const PairRec* __restrict__ pairs;
  const PrimPair* __restrict__ pp;
  const double2* __restrict__ ppab;  // exponents (alpha_a, alpha_b) of each primitive pair
  const double* __restrict__ rys;
  const double* __restrict__ D;
  int nbf;
  double xscale;
  const int* __restrict__ shell_atom;
  double* grad;
  int n_atoms;
  int glds;  // 1: each workgroup sums its quartets' forces in LDS ([3 * n_atoms])
};

template <int La, int Lb, int Lc, int Ld, int NT>
__global__ __launch_bounds__(NT) void grad_kernel(GradArgs g) {
  constexpr int NR = (La + Lb + Lc + Ld + 1) / 2 + 1;
  constexpr int A1 = La + 2, B1 = Lb + 2, C1 = Lc + 2, D1 = Ld + 1;
  constexpr int S = A1 * B1 * C1 * D1;
  constexpr int LCD1 = Lc + 1 + Ld;
  constexpr int HS = A1 * B1 * (LCD1 + 1);
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NOUT = na * nb * nc * nd;
  constexpr int NJ = (NOUT + NT - 1) / NT;
  constexpr int SA = B1 * C1 * D1, SB = C1 * D1, SC = D1;
  __shared__ double tab[NR * 3 * S];
  __shared__ double hs[NR * 3 * HS];
  __shared__ double W[NOUT];
  __shared__ double red[NT / 64][9];
  extern __shared__ double gacc[];
  const int tid = threadIdx.x;
  const bool glds = g.glds != 0;
  if (glds)
    for (int e = tid; e < 3 * g.n_atoms; e += NT) gacc[e] = 0.0;

  // per-lane component data: base offsets (8 bits per axis) and exponents
  int base[NJ], expo[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int o = tid + NT * j;
    base[j] = 0;
    expo[j] = 0;
    if (o < NOUT) {
      const int d = o % nd, c = (o / nd) % nc, b = (o / (nd * nc)) % nb, a = o / (nd * nc * nb);
      int e = 0, bs = 0;
      for (int x = 0; x < 3; ++x) {
        const int ax = cart_comp(La, a, x), bx = cart_comp(Lb, b, x), cx = cart_comp(Lc, c, x),
                  dx = cart_comp(Ld, d, x);
        bs |= (((ax * B1 + bx) * C1 + cx) * D1 + dx) << (8 * x);
        e |= (ax | (bx << 2) | (cx << 4)) << (6 * x);
      }
      base[j] = bs;
      expo[j] = e;
    }
  }

  for (int qi = blockIdx.x; qi < g.nq; qi += gridDim.x) {
    const Quartet q = g.q[qi];
    const PairRec bra = g.pairs[q.bra];
    const PairRec ket = g.pairs[q.ket];
    const bool ij = bra.same, kl = ket.same, bk = (q.bra == q.ket);
    const double deg = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0) * (bk ? 1.0 : 2.0);
    const size_t n = size_t(g.nbf);
    __syncthreads();  // W of the previous quartet is no longer read
    for (int o = tid; o < NOUT; o += NT) {
      const int d = o % nd, c = (o / nd) % nc, b = (o / (nd * nc)) % nb, a = o / (nd * nc * nb);
      const size_t mu = bra.fa + a, nu = bra.fb + b, la = ket.fa + c, si = ket.fb + d;
      const double* D = g.D;
      const double w = 0.5 * D[mu * n + nu] * D[la * n + si] -
                       0.125 * g.xscale * (D[mu * n + la] * D[nu * n + si] + D[mu * n + si] * D[nu * n + la]);
      W[o] = deg * w * norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c) * norm_corr(Ld, d);
    }
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
    const int npp_b = bra.npp_sig, npp_k = ket.npp_sig;
    for (int t = 0; t < npp_b * npp_k; ++t) {
      const int ib = t / npp_k, ik = t - ib * npp_k;
      const PrimPair P = g.pp[bra.pp_off + ib];
      const PrimPair Q = g.pp[ket.pp_off + ik];
      const double2 eb = g.ppab[bra.pp_off + ib];
      const double2 ek = g.ppab[ket.pp_off + ik];
      __syncthreads();
      if (tid < 3 * NR) {
        const int r = tid / 3, ax = tid - 3 * (tid / 3);
        const double zeta = P.zeta, eta = Q.zeta;
        const double izpe = 1.0 / (zeta + eta);
        const double rz = eta * izpe, re = zeta * izpe;
        const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
        const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
        double ur, wr;
        rys_root_one<NR>(T, g.rys, r, ur, wr);
        const double h_z = 0.5 / zeta, h_e = 0.5 / eta;
        const double B00 = 0.5 * ur * izpe, B10 = h_z - h_z * rz * ur, B01 = h_e - h_e * re * ur;
        const double PQ = ax == 0 ? PQx : (ax == 1 ? PQy : PQz);
        const double PA = ax == 0 ? P.PAx : (ax == 1 ? P.PAy : P.PAz);
        const double QC = ax == 0 ? Q.PAx : (ax == 1 ? Q.PAy : Q.PAz);
        const double AB = ax == 0 ? bra.ABx : (ax == 1 ? bra.ABy : bra.ABz);
        const double CD = ax == 0 ? ket.ABx : (ax == 1 ? ket.ABy : ket.ABz);
        const double sc = ax == 2 ? P.Kc * Q.Kc * sqrt(izpe) * wr : 1.0;
        build_2d_lds<La + 1, Lb + 1, Lc + 1, Ld, double>(tab + (r * 3 + ax) * S, hs + (r * 3 + ax) * HS,
                                                         PA - rz * ur * PQ, QC + re * ur * PQ, AB, CD, B10, B01,
                                                         B00, sc);
      }
      __syncthreads();
      const double ta2 = 2.0 * eb.x, tb2 = 2.0 * eb.y, tc2 = 2.0 * ek.x;
#pragma unroll 1
      for (int r = 0; r < NR; ++r) {
        const double* T3[3] = {tab + (r * 3) * S, tab + (r * 3 + 1) * S, tab + (r * 3 + 2) * S};
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const int o = tid + NT * j;
          if (o >= NOUT) continue;
          const double w = W[o];
          int i0[3], ea[3], eb_[3], ec[3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            i0[x] = (base[j] >> (8 * x)) & 255;
            const int e = (expo[j] >> (6 * x)) & 63;
            ea[x] = e & 3;
            eb_[x] = (e >> 2) & 3;
            ec[x] = (e >> 4) & 3;
          }
          const double v0[3] = {T3[0][i0[0]], T3[1][i0[1]], T3[2][i0[2]]};
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            const double* tx = T3[x];
            const double rest = w * v0[(x + 1) % 3] * v0[(x + 2) % 3];
            double dA = ta2 * tx[i0[x] + SA];
            if (ea[x] > 0) dA -= ea[x] * tx[i0[x] - SA];
            double dB = tb2 * tx[i0[x] + SB];
            if (eb_[x] > 0) dB -= eb_[x] * tx[i0[x] - SB];
            double dC = tc2 * tx[i0[x] + SC];
            if (ec[x] > 0) dC -= ec[x] * tx[i0[x] - SC];
            gA[x] += dA * rest;
            gB[x] += dB * rest;
            gC[x] += dC * rest;
          }
        }
      }
    }
    // workgroup reduction of the 9 components, then atomics per atom
    double v[9] = {gA[0], gA[1], gA[2], gB[0], gB[1], gB[2], gC[0], gC[1], gC[2]};
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double x = v[k];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
      v[k] = x;
    }
    if ((tid & 63) == 0)
#pragma unroll
      for (int k = 0; k < 9; ++k) red[tid >> 6][k] = v[k];
    __syncthreads();
    if (tid < 9) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) s += red[w][tid];
      red[0][tid] = s;  // each lane owns its slot
    }
    __syncthreads();
    if (tid < 3) {
      const double a = red[0][tid], b = red[0][3 + tid], c = red[0][6 + tid];
      double* out = glds ? gacc : g.grad;
      const int ia = 3 * g.shell_atom[bra.sa] + tid, ib = 3 * g.shell_atom[bra.sb] + tid;
      const int ic = 3 * g.shell_atom[ket.sa] + tid, id = 3 * g.shell_atom[ket.sb] + tid;
      if (glds) {
        lds_add(&out[ia], a);
        lds_add(&out[ib], b);
        lds_add(&out[ic], c);
        lds_add(&out[id], -(a + b + c));
      } else {
        atomic_add(&out[ia], a);
        atomic_add(&out[ib], b);
        atomic_add(&out[ic], c);
        atomic_add(&out[id], -(a + b + c));
      }
    }
  }
  if (glds) {  // this workgroup's forces, once
    __syncthreads();
    for (int e = tid; e < 3 * g.n_atoms; e += NT) {
      const double v = gacc[e];
      if (v != 0.0) atomic_add(&g.grad[e], v);
    }
  }
}

template <int La, int Lb, int Lc, int Ld>
struct GradPolicy {
  static constexpr int NOUT = ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld);
  static constexpr int NT = NOUT <= 256 ? 64 : 128;
};

}  // namespace lai
