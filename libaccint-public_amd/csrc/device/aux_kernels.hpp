// aux_kernels.hpp -- small device kernels of the backend: fills, the J/K
// symmetrisation, the Schwarz diagonal reduction, one-electron integrals and
// the math probes.
#pragma once

#include "common.hpp"

namespace lai {

__global__ void fill_zero_kernel(double* __restrict__ p, size_t n);

// J += (Jt + Jt^T)/2, K += (Kt + Kt^T)/2
__global__ void symmetrize_add_kernel(const double* __restrict__ Jt, const double* __restrict__ Kt,
                                      double* __restrict__ J, double* __restrict__ K, int n);

// Q[p] = sqrt(max_ab |(ab|ab)|) for the (i,j,i,j) blocks of every pair
__global__ void schwarz_reduce_kernel(const double* __restrict__ blocks,
                                      const long long* __restrict__ offs, const int* __restrict__ ni,
                                      const int* __restrict__ nj, double* __restrict__ Q, int npairs);

struct OneEArgs {
  int n_shells;
  const int* L;
  const int* prim_off;
  const int* nprim;
  const int* func_off;
  const double* center;  // 3 per shell
  const double* exps;
  const double* coefs;
  int kind;  // 0 S, 1 T, 2 V
  int n_charges;
  const double* Z;
  const double* cxyz;
  const double* rys;
  double* out;
  int nbf;
};
__global__ void onee_kernel(OneEArgs a);

__global__ void rys_probe_kernel(int nroots, int n, const double* __restrict__ T,
                                 const double* __restrict__ tab, double* __restrict__ r,
                                 double* __restrict__ w);
__global__ void boys_probe_kernel(int nmax, int n, const double* __restrict__ T,
                                  double* __restrict__ out);

}  // namespace lai
