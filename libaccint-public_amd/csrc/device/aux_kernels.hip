// aux_kernels.hip -- fills, J/K symmetrisation, Schwarz reduction, one-electron
// integrals and the Rys/Boys probes.
#include "aux_kernels.hpp"

namespace lai {

__global__ void fill_zero_kernel(double* __restrict__ p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0.0;
}

__global__ void symmetrize_add_kernel(const double* __restrict__ Jt, const double* __restrict__ Kt,
                                      double* __restrict__ J, double* __restrict__ K, int n) {
  const size_t nn = size_t(n) * n;
  for (size_t x = size_t(blockIdx.x) * blockDim.x + threadIdx.x; x < nn; x += size_t(gridDim.x) * blockDim.x) {
    const size_t i = x / n, j = x - i * n;
    const size_t t = j * n + i;
    J[x] += 0.5 * (Jt[x] + Jt[t]);
    K[x] += 0.5 * (Kt[x] + Kt[t]);
  }
}

__global__ void schwarz_reduce_kernel(const double* __restrict__ blocks,
                                      const long long* __restrict__ offs, const int* __restrict__ ni,
                                      const int* __restrict__ nj, double* __restrict__ Q, int npairs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const double* blk = blocks + offs[p];
  const int a_n = ni[p], b_n = nj[p];
  double mx = 0.0;
  for (int a = 0; a < a_n; ++a)
    for (int b = 0; b < b_n; ++b) mx = fmax(mx, fabs(blk[((a * b_n + b) * a_n + a) * b_n + b]));
  Q[p] = sqrt(mx);
}

// ---------------------------------------------------------------------------
// One-electron integrals: one lane per shell pair (i <= j), runtime angular
// momentum up to kMaxL.  Overlap by the Obara-Saika 1D recursion, kinetic as
// T = Tx Sy Sz + Sx Ty Sz + Sx Sy Tz with
// T_d(i,j) = -1/2 [4 b^2 S(i,j+2) - 2b(2j+1) S(i,j) + j(j-1) S(i,j-2)], and
// nuclear attraction by 1D Rys recursions over the point charges
// (reference: src/host/kernels/overlap_kernel.cpp:38-170,
// kinetic_kernel.cpp:80-215, nuclear_kernel.cpp:30-240).
// ---------------------------------------------------------------------------
namespace {
constexpr int D1 = kMaxL + 3;

__device__ void os1d(int La, int Lb, double PA, double PB, double o2z, double (&I)[D1][D1]) {
  I[0][0] = 1.0;
  for (int i = 0; i < La; ++i) I[i + 1][0] = PA * I[i][0] + (i > 0 ? i * o2z * I[i - 1][0] : 0.0);
  for (int j = 0; j < Lb; ++j)
    for (int i = 0; i <= La; ++i) {
      double v = PB * I[i][j];
      if (i > 0) v += i * o2z * I[i - 1][j];
      if (j > 0) v += j * o2z * I[i][j - 1];
      I[i][j + 1] = v;
    }
}

__device__ double kin1d(const double (&S)[D1][D1], int i, int j, double beta) {
  double v = 4.0 * beta * beta * S[i][j + 2] - 2.0 * beta * (2 * j + 1) * S[i][j];
  if (j >= 2) v += j * (j - 1) * S[i][j - 2];
  return -0.5 * v;
}

__device__ void cart_list(int L, int (&c)[15][3]) {
  int n = 0;
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) {
      c[n][0] = lx;
      c[n][1] = ly;
      c[n][2] = L - lx - ly;
      ++n;
    }
}

__device__ double ncorr(int lx, int ly, int lz) {
  return 1.0 / sqrt(double(dfact_odd(lx) * dfact_odd(ly) * dfact_odd(lz)));
}

template <int N>
__device__ void roots_n(double T, const double* tab, double* u, double* w) {
  rys_roots<N>(T, tab, u, w);
}

__device__ void roots_rt(int n, double T, const double* tab, double* u, double* w) {
  switch (n) {
    case 1: roots_n<1>(T, tab, u, w); break;
    case 2: roots_n<2>(T, tab, u, w); break;
    case 3: roots_n<3>(T, tab, u, w); break;
    case 4: roots_n<4>(T, tab, u, w); break;
    case 5: roots_n<5>(T, tab, u, w); break;
    case 6: roots_n<6>(T, tab, u, w); break;
    case 7: roots_n<7>(T, tab, u, w); break;
    case 8: roots_n<8>(T, tab, u, w); break;
    default: roots_n<9>(T, tab, u, w); break;
  }
}
}  // namespace

__global__ void onee_kernel(OneEArgs g) {
  const int npairs = g.n_shells * (g.n_shells + 1) / 2;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x) {
    // p -> (i, j), i <= j, row-major over i
    int i = 0, rem = p;
    while (rem >= g.n_shells - i) {
      rem -= g.n_shells - i;
      ++i;
    }
    const int j = i + rem;
    const int La = g.L[i], Lb = g.L[j];
    const int na = ncart(La), nb = ncart(Lb);
    int ca[15][3], cb[15][3];
    cart_list(La, ca);
    cart_list(Lb, cb);
    double buf[15 * 15];
    for (int x = 0; x < na * nb; ++x) buf[x] = 0.0;
    const double Ax = g.center[3 * i], Ay = g.center[3 * i + 1], Az = g.center[3 * i + 2];
    const double Bx = g.center[3 * j], By = g.center[3 * j + 1], Bz = g.center[3 * j + 2];
    const double dx = Ax - Bx, dy = Ay - By, dz = Az - Bz;
    const double AB2 = dx * dx + dy * dy + dz * dz;
    const double PI = 3.14159265358979323846;
    const int nr = (La + Lb) / 2 + 1;
    for (int pa = 0; pa < g.nprim[i]; ++pa) {
      const double al = g.exps[g.prim_off[i] + pa], cA = g.coefs[g.prim_off[i] + pa];
      if (cA == 0.0) continue;
      for (int pb = 0; pb < g.nprim[j]; ++pb) {
        const double be = g.exps[g.prim_off[j] + pb], cB = g.coefs[g.prim_off[j] + pb];
        if (cB == 0.0) continue;
        const double zeta = al + be, iz = 1.0 / zeta, mu = al * be * iz;
        const double Px = (al * Ax + be * Bx) * iz, Py = (al * Ay + be * By) * iz,
                     Pz = (al * Az + be * Bz) * iz;
        const double KAB = exp(-mu * AB2);
        const double PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az;
        const double PBx = Px - Bx, PBy = Py - By, PBz = Pz - Bz;
        double Ix[D1][D1], Iy[D1][D1], Iz[D1][D1];
        if (g.kind <= 1) {
          const double o2z = 0.5 * iz;
          const double pref = cA * cB * pow(PI * iz, 1.5) * KAB;
          const int lbx = g.kind == 1 ? Lb + 2 : Lb;
          os1d(La, lbx, PAx, PBx, o2z, Ix);
          os1d(La, lbx, PAy, PBy, o2z, Iy);
          os1d(La, lbx, PAz, PBz, o2z, Iz);
          for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
              const int ax = ca[a][0], ay = ca[a][1], az = ca[a][2];
              const int bx = cb[b][0], by = cb[b][1], bz = cb[b][2];
              double v;
              if (g.kind == 0) {
                v = Ix[ax][bx] * Iy[ay][by] * Iz[az][bz];
              } else {
                v = kin1d(Ix, ax, bx, be) * Iy[ay][by] * Iz[az][bz] +
                    Ix[ax][bx] * kin1d(Iy, ay, by, be) * Iz[az][bz] +
                    Ix[ax][bx] * Iy[ay][by] * kin1d(Iz, az, bz, be);
              }
              buf[a * nb + b] += pref * v;
            }
        } else {
          const double pref = cA * cB * (2.0 * PI * iz) * KAB;
          const double o2z = 0.5 * iz;
          for (int c = 0; c < g.n_charges; ++c) {
            const double Zc = g.Z[c];
            if (Zc == 0.0) continue;
            const double PCx = Px - g.cxyz[3 * c], PCy = Py - g.cxyz[3 * c + 1], PCz = Pz - g.cxyz[3 * c + 2];
            const double T = zeta * (PCx * PCx + PCy * PCy + PCz * PCz);
            double u[LAI_RYS_NMAX], w[LAI_RYS_NMAX];
            roots_rt(nr, T, g.rys, u, w);
            for (int r = 0; r < nr; ++r) {
              const double B00 = o2z * (1.0 - u[r]);
              os1d(La, Lb, PAx - u[r] * PCx, PBx - u[r] * PCx, B00, Ix);
              os1d(La, Lb, PAy - u[r] * PCy, PBy - u[r] * PCy, B00, Iy);
              os1d(La, Lb, PAz - u[r] * PCz, PBz - u[r] * PCz, B00, Iz);
              const double co = -Zc * pref * w[r];
              for (int a = 0; a < na; ++a)
                for (int b = 0; b < nb; ++b)
                  buf[a * nb + b] += co * Ix[ca[a][0]][cb[b][0]] * Iy[ca[a][1]][cb[b][1]] * Iz[ca[a][2]][cb[b][2]];
            }
          }
        }
      }
    }
    const size_t n = size_t(g.nbf);
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b) {
        const double v = buf[a * nb + b] * ncorr(ca[a][0], ca[a][1], ca[a][2]) * ncorr(cb[b][0], cb[b][1], cb[b][2]);
        const size_t mu = g.func_off[i] + a, nu = g.func_off[j] + b;
        g.out[mu * n + nu] = v;
        g.out[nu * n + mu] = v;
      }
  }
}

__global__ void rys_probe_kernel(int nroots, int n, const double* __restrict__ T,
                                 const double* __restrict__ tab, double* __restrict__ r,
                                 double* __restrict__ w) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double u[LAI_RYS_NMAX], ww[LAI_RYS_NMAX];
  roots_rt(nroots, T[t], tab, u, ww);
  for (int i = 0; i < nroots; ++i) {
    r[size_t(t) * nroots + i] = u[i];
    w[size_t(t) * nroots + i] = ww[i];
  }
}

// Boys function F_0..F_nmax: for T < 40 the top order from the
// positive-term series F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// followed by the downward recursion F_{m-1} = (2T F_m + exp(-T))/(2m-1);
// for T >= 40, F_0 = sqrt(pi/T)/2 (erfc(sqrt T) < 1e-18) and the upward
// recursion F_{m+1} = ((2m+1) F_m - exp(-T))/(2T).
__global__ void boys_probe_kernel(int nmax, int n, const double* __restrict__ T_in,
                                  double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double T = T_in[t];
  double* o = out + size_t(t) * (nmax + 1);
  const double e = exp(-T);
  if (T < 40.0) {
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * T / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    o[nmax] = e * sum;
    for (int m = nmax; m >= 1; --m) o[m - 1] = (2.0 * T * o[m] + e) / (2 * m - 1);
  } else {
    o[0] = 0.5 * sqrt(3.14159265358979323846 / T);
    for (int m = 0; m < nmax; ++m) o[m + 1] = ((2 * m + 1) * o[m] - e) / (2.0 * T);
  }
}

}  // namespace lai
