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
  // the J/K worklist: bra-anchored chunks over ket segments
  const JkChunk* __restrict__ chunks;
  int nchunks;
  const JkSegs* __restrict__ segs;
  const int* __restrict__ kets;
  QuartetSel sel;
  const double* __restrict__ ppih;  // 1 / (2 zeta) of each primitive pair
};

// Primitive quartets per step of the LDS-shape gradient kernel: 3 n_r lanes
// build the (root, axis) tables of one primitive quartet, as many primitive
// quartets at once as the lanes and 16 KB of tables allow (more would cost
// workgroups per CU).
template <int La, int Lb, int Lc, int Ld, int NT>
struct GradShape {
  static constexpr int NR = (La + Lb + Lc + Ld + 1) / 2 + 1;
  static constexpr int A1 = La + 2, B1 = Lb + 2, C1 = Lc + 2, D1 = Ld + 1;
  static constexpr int S = A1 * B1 * C1 * D1;
  static constexpr int HS = A1 * B1 * (Lc + 1 + Ld + 1);
  static constexpr int NPQ = [] {
    constexpr int by_lanes = NT / (3 * NR);
    constexpr int by_lds = int((16 * 1024) / (3 * NR * (S + HS) * sizeof(double)));
    constexpr int n = by_lanes < by_lds ? by_lanes : by_lds;
    return n < 1 ? 1 : n;
  }();
};

template <int La, int Lb, int Lc, int Ld, int NT>
__global__ __launch_bounds__(NT) void grad_kernel(GradArgs g) {
  using Sh = GradShape<La, Lb, Lc, Ld, NT>;
  constexpr int NR = Sh::NR, NPQ = Sh::NPQ;
  constexpr int B1 = Sh::B1, C1 = Sh::C1, D1 = Sh::D1;
  constexpr int S = Sh::S, HS = Sh::HS;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  constexpr int NOUT = na * nb * nc * nd;
  constexpr int SA = B1 * C1 * D1, SB = C1 * D1, SC = D1;
  __shared__ double tab[NPQ * NR * 3 * S];
  __shared__ double hs[NPQ * NR * 3 * HS];
  __shared__ double W[NOUT];
  __shared__ double ex2[NPQ][3];  // 2 alpha_a, 2 alpha_b, 2 alpha_c of each primitive quartet
  __shared__ int cbase[NOUT], cexpo[NOUT];
  __shared__ double nrm[NOUT];  // normalisation of each component, computed once
  __shared__ double red[NT / 64][9];
  extern __shared__ double gacc[];
  const int tid = threadIdx.x;
  const bool glds = g.glds != 0;
  if (glds)
    for (int e = tid; e < 3 * g.n_atoms; e += NT) gacc[e] = 0.0;

  // per component: table offsets (8 bits per axis) and exponents
  for (int o = tid; o < NOUT; o += NT) {
    const int d = o % nd, c = (o / nd) % nc, b = (o / (nd * nc)) % nb, a = o / (nd * nc * nb);
    int e = 0, bs = 0;
    for (int x = 0; x < 3; ++x) {
      const int ax = cart_comp(La, a, x), bx = cart_comp(Lb, b, x), cx = cart_comp(Lc, c, x),
                dx = cart_comp(Ld, d, x);
      bs |= (((ax * B1 + bx) * C1 + cx) * D1 + dx) << (8 * x);
      e |= (ax | (bx << 2) | (cx << 4)) << (6 * x);
    }
    cbase[o] = bs;
    cexpo[o] = e;
    nrm[o] = norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c) * norm_corr(Ld, d);
  }

  for (int ci = blockIdx.x; ci < g.nchunks; ci += gridDim.x) {
   const JkChunk ch = g.chunks[ci];
   const PairRec bra = g.pairs[ch.bra];
   const JkSegs* sg = g.segs + ch.seg;
   int seg = 0;
   // gridDim.y workgroups share a chunk (chunks of the register-shape J/K
   // classes hold up to 2048 quartets)
   const int len = ch.k1 - ch.k0;
   const int kb = ch.k0 + int((long long)len * blockIdx.y / gridDim.y);
   const int ke = ch.k0 + int((long long)len * (blockIdx.y + 1) / gridDim.y);
   for (int kq = kb; kq < ke; ++kq) {
    const int kp = chunk_ket(g.kets, sg, kq, seg);
    const PairRec ket = g.pairs[kp];
    if (quartet_select(g.sel, ch.bra, bra, kp, ket) == 0) continue;  // uniform over the workgroup
    const bool ij = bra.same, kl = ket.same, bk = (kp == ch.bra);
    const double deg = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0) * (bk ? 1.0 : 2.0);
    const size_t n = size_t(bra.mn);
    __syncthreads();  // W and the tables of the previous quartet are no longer read
    for (int o = tid; o < NOUT; o += NT) {
      const int d = o % nd, c = (o / nd) % nc, b = (o / (nd * nc)) % nb, a = o / (nd * nc * nb);
      const size_t mu = bra.fa + a, nu = bra.fb + b, la = ket.fa + c, si = ket.fb + d;
      const double* D = g.D + bra.moff;
      const double w = 0.5 * D[mu * n + nu] * D[la * n + si] -
                       0.125 * g.xscale * (D[mu * n + la] * D[nu * n + si] + D[mu * n + si] * D[nu * n + la]);
      W[o] = deg * w * nrm[o];
    }
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
    const int npp_b = bra.npp_sig, npp_k = ket.npp_sig;
    const int npt = npp_b * npp_k;
    for (int t0 = 0; t0 < npt; t0 += NPQ) {
      const int nq = min(NPQ, npt - t0);
      if (t0 > 0) __syncthreads();  // the previous step's tables are no longer read
      if (tid < 3 * NR * nq) {
        const int q = tid / (3 * NR), rr = tid - q * (3 * NR);
        const int r = rr / 3, ax = rr - 3 * r;
        const int t = t0 + q;
        const int ib = t / npp_k, ik = t - ib * npp_k;
        const PrimPair P = g.pp[bra.pp_off + ib];
        const PrimPair Q = g.pp[ket.pp_off + ik];
        const double zeta = P.zeta, eta = Q.zeta;
        const double izpe = 1.0 / (zeta + eta);
        const double rz = eta * izpe, re = zeta * izpe;
        const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
        const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
        double ur, wr;
        rys_root_one<NR>(T, g.rys, r, ur, wr);
        const double h_z = 0.5 / zeta, h_e = 0.5 / eta;
        const double B00 = 0.5 * ur * izpe, B10 = h_z - h_z * rz * ur, B01 = h_e - h_e * re * ur;
        const double PQ = pick3(ax, PQx, PQy, PQz);
        const double PA = pick3(ax, P.PAx, P.PAy, P.PAz);
        const double QC = pick3(ax, Q.PAx, Q.PAy, Q.PAz);
        const double AB = pick3(ax, bra.ABx, bra.ABy, bra.ABz);
        const double CD = pick3(ax, ket.ABx, ket.ABy, ket.ABz);
        const double sc = ax == 2 ? P.Kc * Q.Kc * sqrt(izpe) * wr : 1.0;
        build_2d_lds<La + 1, Lb + 1, Lc + 1, Ld, double>(tab + ((q * NR + r) * 3 + ax) * S,
                                                         hs + ((q * NR + r) * 3 + ax) * HS, PA - rz * ur * PQ,
                                                         QC + re * ur * PQ, AB, CD, B10, B01, B00, sc);
        if (rr == 0) {
          const double2 eb = g.ppab[bra.pp_off + ib];
          const double2 ek = g.ppab[ket.pp_off + ik];
          ex2[q][0] = 2.0 * eb.x;
          ex2[q][1] = 2.0 * eb.y;
          ex2[q][2] = 2.0 * ek.x;
        }
      }
      __syncthreads();
      // contraction: tasks (component o, primitive quartet q) over all lanes
      for (int tk = tid; tk < NOUT * nq; tk += NT) {
        const int q = tk / NOUT, o = tk - q * NOUT;
        const double w = W[o];
        const int bs = cbase[o], ee = cexpo[o];
        const double ta2 = ex2[q][0], tb2 = ex2[q][1], tc2 = ex2[q][2];
        int i0[3], ea[3], eb_[3], ec[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          i0[x] = (bs >> (8 * x)) & 255;
          const int e = (ee >> (6 * x)) & 63;
          ea[x] = e & 3;
          eb_[x] = (e >> 2) & 3;
          ec[x] = (e >> 4) & 3;
        }
#pragma unroll 1
        for (int r = 0; r < NR; ++r) {
          const double* T3[3] = {tab + ((q * NR + r) * 3) * S, tab + ((q * NR + r) * 3 + 1) * S,
                                 tab + ((q * NR + r) * 3 + 2) * S};
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
  }
  if (glds) {  // this workgroup's forces, once
    __syncthreads();
    for (int e = tid; e < 3 * g.n_atoms; e += NT) {
      const double v = gacc[e];
      if (v != 0.0) atomic_add(&g.grad[e], v);
    }
  }
}

// --------------------------------------------------------------------------
// Register shape for the light classes: one lane per shell quartet of a
// bra-anchored chunk, the 2D tables raised by one on A, B and C held in
// registers.  The forces on the bra atoms are summed over the lane's quartets
// and reduced over the wavefront; those on the ket atoms go to the
// workgroup's LDS force array (or to HBM atomics without one).
// --------------------------------------------------------------------------
template <int La, int Lb, int Lc, int Ld>
__device__ __forceinline__ void grad_quartet_regs(const GradArgs& g, const PairRec& bra, const PairRec& ket,
                                                  bool bk, double (&f)[9]) {
  constexpr int NR = (La + Lb + Lc + Ld + 1) / 2 + 1;
  constexpr int na = ncart(La), nb = ncart(Lb), nc = ncart(Lc), nd = ncart(Ld);
  const size_t n = size_t(bra.mn);
  const bool ij = bra.sa == bra.sb, kl = ket.sa == ket.sb;
  const double deg = (ij ? 1.0 : 2.0) * (kl ? 1.0 : 2.0) * (bk ? 1.0 : 2.0);
  const double* __restrict__ D = g.D + bra.moff;
  double W[na][nb][nc][nd];
#pragma unroll
  for (int a = 0; a < na; ++a)
#pragma unroll
    for (int b = 0; b < nb; ++b)
#pragma unroll
      for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) {
          const size_t mu = bra.fa + a, nu = bra.fb + b, la = ket.fa + c, si = ket.fb + d;
          const double w = 0.5 * D[mu * n + nu] * D[la * n + si] -
                           0.125 * g.xscale * (D[mu * n + la] * D[nu * n + si] + D[mu * n + si] * D[nu * n + la]);
          W[a][b][c][d] = deg * w * norm_corr(La, a) * norm_corr(Lb, b) * norm_corr(Lc, c) * norm_corr(Ld, d);
        }
#pragma unroll
  for (int x = 0; x < 9; ++x) f[x] = 0.0;
  const double ABx = bra.ABx, ABy = bra.ABy, ABz = bra.ABz;
  const double CDx = ket.ABx, CDy = ket.ABy, CDz = ket.ABz;
  for (int ik = 0; ik < ket.npp_sig; ++ik) {
    const PrimPair Q = g.pp[ket.pp_off + ik];
    const double2 ek = g.ppab[ket.pp_off + ik];
    const double eta = Q.zeta, h_e = g.ppih[ket.pp_off + ik];
    const double tc2 = 2.0 * ek.x;
    for (int ib = 0; ib < bra.npp_sig; ++ib) {
      // the bra is uniform over the workgroup: scalar loads
      const PrimPair P = load_prim(g.pp, bra.pp_off + ib);
      const double2 eb = make_double2(as_const(reinterpret_cast<const double*>(g.ppab))[2 * (bra.pp_off + ib)],
                                      as_const(reinterpret_cast<const double*>(g.ppab))[2 * (bra.pp_off + ib) + 1]);
      const double zeta = P.zeta, h_z = as_const(g.ppih)[bra.pp_off + ib];
      const double ta2 = 2.0 * eb.x, tb2 = 2.0 * eb.y;
      const double x = zeta + eta;
      double y = __builtin_amdgcn_rsq(x);
      y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
      y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
      const double izpe = y * y;
      const double rz = eta * izpe, re = zeta * izpe;
      const double PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;
      const double T = zeta * rz * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = P.Kc * Q.Kc * y;
      double u[NR], w[NR];
      rys_roots<NR>(T, g.rys, u, w);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double ur = u[r];
        const double B00 = 0.5 * ur * izpe, B10 = h_z - h_z * rz * ur, B01 = h_e - h_e * re * ur;
        const double sz = rz * ur, se = re * ur;
        double Ix[La + 2][Lb + 2][Lc + 2][Ld + 1], Iy[La + 2][Lb + 2][Lc + 2][Ld + 1],
            Iz[La + 2][Lb + 2][Lc + 2][Ld + 1];
        build_2d<La + 1, Lb + 1, Lc + 1, Ld, double>(Ix, P.PAx - sz * PQx, Q.PAx + se * PQx, ABx, CDx, B10, B01,
                                                     B00, 1.0);
        build_2d<La + 1, Lb + 1, Lc + 1, Ld, double>(Iy, P.PAy - sz * PQy, Q.PAy + se * PQy, ABy, CDy, B10, B01,
                                                     B00, 1.0);
        build_2d<La + 1, Lb + 1, Lc + 1, Ld, double>(Iz, P.PAz - sz * PQz, Q.PAz + se * PQz, ABz, CDz, B10, B01,
                                                     B00, pref * w[r]);
#pragma unroll
        for (int a = 0; a < na; ++a)
#pragma unroll
          for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
              for (int d = 0; d < nd; ++d) {
                const int ax = cart_comp(La, a, 0), ay = cart_comp(La, a, 1), az = cart_comp(La, a, 2);
                const int bx = cart_comp(Lb, b, 0), by = cart_comp(Lb, b, 1), bz = cart_comp(Lb, b, 2);
                const int cx = cart_comp(Lc, c, 0), cy = cart_comp(Lc, c, 1), cz = cart_comp(Lc, c, 2);
                const int dx = cart_comp(Ld, d, 0), dy = cart_comp(Ld, d, 1), dz = cart_comp(Ld, d, 2);
                const double wt = W[a][b][c][d];
                const double x0 = Ix[ax][bx][cx][dx], y0 = Iy[ay][by][cy][dy], z0 = Iz[az][bz][cz][dz];
                // d/dA: 2 alpha (a+1) - a (a-1), and likewise for B and C
                const double dAx = ta2 * Ix[ax + 1][bx][cx][dx] - (ax ? ax * Ix[ax ? ax - 1 : 0][bx][cx][dx] : 0.0);
                const double dAy = ta2 * Iy[ay + 1][by][cy][dy] - (ay ? ay * Iy[ay ? ay - 1 : 0][by][cy][dy] : 0.0);
                const double dAz = ta2 * Iz[az + 1][bz][cz][dz] - (az ? az * Iz[az ? az - 1 : 0][bz][cz][dz] : 0.0);
                const double dBx = tb2 * Ix[ax][bx + 1][cx][dx] - (bx ? bx * Ix[ax][bx ? bx - 1 : 0][cx][dx] : 0.0);
                const double dBy = tb2 * Iy[ay][by + 1][cy][dy] - (by ? by * Iy[ay][by ? by - 1 : 0][cy][dy] : 0.0);
                const double dBz = tb2 * Iz[az][bz + 1][cz][dz] - (bz ? bz * Iz[az][bz ? bz - 1 : 0][cz][dz] : 0.0);
                const double dCx = tc2 * Ix[ax][bx][cx + 1][dx] - (cx ? cx * Ix[ax][bx][cx ? cx - 1 : 0][dx] : 0.0);
                const double dCy = tc2 * Iy[ay][by][cy + 1][dy] - (cy ? cy * Iy[ay][by][cy ? cy - 1 : 0][dy] : 0.0);
                const double dCz = tc2 * Iz[az][bz][cz + 1][dz] - (cz ? cz * Iz[az][bz][cz ? cz - 1 : 0][dz] : 0.0);
                const double yz = wt * y0 * z0, xz = wt * x0 * z0, xy = wt * x0 * y0;
                f[0] += dAx * yz;
                f[1] += dAy * xz;
                f[2] += dAz * xy;
                f[3] += dBx * yz;
                f[4] += dBy * xz;
                f[5] += dBz * xy;
                f[6] += dCx * yz;
                f[7] += dCy * xz;
                f[8] += dCz * xy;
              }
      }
    }
  }
}

template <int La, int Lb, int Lc, int Ld>
__global__ __launch_bounds__(256) void grad_kernel_regs(GradArgs g) {
  extern __shared__ double gacc[];
  __shared__ SortLds srt;  // lane schedule of the chunk (sort_run)
  const int tid = threadIdx.x;
  const bool glds = g.glds != 0;
  if (glds) {
    for (int e = tid; e < 3 * g.n_atoms; e += blockDim.x) gacc[e] = 0.0;
    __syncthreads();
  }
  auto add = [&](int i, double v) {
    if (glds)
      lds_add(&gacc[i], v);
    else
      atomic_add(&g.grad[i], v);
  };
  for (int ci = blockIdx.x; ci < g.nchunks; ci += gridDim.x) {
    const JkChunk ch = g.chunks[ci];
    const PairRec bra = g.pairs[ch.bra];
    const int iA = 3 * g.shell_atom[bra.sa], iB = 3 * g.shell_atom[bra.sb];
    double fab[6] = {0, 0, 0, 0, 0, 0};
    const JkSegs* sg = g.segs + ch.seg;
    for (int s0 = ch.k0; s0 < ch.k1; s0 += kSortRun) {
    const int nsel = sort_run<256>(g.kets, sg, g.pairs, g.sel, ch.bra, bra, s0, min(ch.k1, s0 + kSortRun), srt, 256);
    for (int r = 0; 256 * r < nsel; ++r) {
      const int i = sorted_pos(tid, r, 4);
      if (i >= nsel) continue;
      const int kp = int(srt.kp[i] & ~kSortF32);
      const PairRec ket = g.pairs[kp];
      double f[9];
      grad_quartet_regs<La, Lb, Lc, Ld>(g, bra, ket, kp == ch.bra, f);
#pragma unroll
      for (int x = 0; x < 6; ++x) fab[x] += f[x];
      const int iC = 3 * g.shell_atom[ket.sa], iD = 3 * g.shell_atom[ket.sb];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        add(iC + x, f[6 + x]);
        add(iD + x, -(f[x] + f[3 + x] + f[6 + x]));  // translational invariance
      }
    }
    }
#pragma unroll
    for (int x = 0; x < 6; ++x) {
      const double v = wave_sum(fab[x]);
      if ((tid & 63) == 0 && v != 0.0) add((x < 3 ? iA : iB) + x % 3, v);
    }
  }
  if (glds) {  // this workgroup's forces, once
    __syncthreads();
    for (int e = tid; e < 3 * g.n_atoms; e += blockDim.x) {
      const double v = gacc[e];
      if (v != 0.0) atomic_add(&g.grad[e], v);
    }
  }
}

template <int La, int Lb, int Lc, int Ld>
struct GradPolicy {
  static constexpr int NOUT = ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld);
  static constexpr int NT = NOUT <= 256 ? 64 : 128;
  // register shape while the raised tables and outputs stay small: up to 36
  // table entries per axis (one wavefront per SIMD, at most a few bytes of
  // spill); measured faster than the LDS shape there, whose per-quartet
  // overhead dominates these light classes
  static constexpr bool regs = (La + 2) * (Lb + 2) * (Lc + 2) * (Ld + 1) <= 36 && NOUT <= 54;
};

}  // namespace lai
