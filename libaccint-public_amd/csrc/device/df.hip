// df.hip -- density-fitted J/K on the device.
//
// Set-up (lai_dev_df_create): the auxiliary metric through the raw ERI
// kernels ((P u|Q u) with the unit function u) or the overlap kernel, its
// Cholesky factor and triangular inverse on the host (n_aux^3 / 3 flops each,
// once), the three-centre integrals (ab|P u) of every orbital shell pair in
// bounded batches through the raw ERI kernels, scattered into
// T[a][P][b], and B[a] = L^-1 T[a] for every orbital function a (one batched
// FP64 MFMA GEMM).  B stays in HBM in the layout [a][P][b]: row a of B is a
// contiguous n_aux x n matrix, so both exchange products are single GEMMs.
//
// Per build (lai_dev_df_jk): J in two HBM-bound passes over B (gamma_P, then
// J_ab = sum_P B[a][P][b] gamma_P); K = sum_P B_P D B_P^T as
// X[a][P][:] = B[a][P][:] D (one GEMM, rows (a, P)) and
// K[a][c] = sum_(P,b) X[a][P][b] B[c][P][b] (one GEMM over the summation
// index n_aux n, split into slices that are summed by a last kernel).
#include "lai_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "api_internal.hpp"

namespace {

using lai::set_error;

#define DF_HIP(call)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return set_error(LAI_ERR_BACKEND, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

constexpr int ncart_of(int L) { return (L + 1) * (L + 2) / 2; }

// ---------------------------------------------------------------------------
// FP64 GEMM on the matrix cores: C = A B with A row-major M x K and B either
// row-major K x N (bt = 0) or given as its transpose, row-major N x K
// (bt = 1).  A workgroup of four wavefronts computes a (64 BM) x (16 BN) tile
// of C, the wavefronts stacked along M, each a (16 BM) x (16 BN) strip as
// BM x BN blocks of v_mfma_f64_16x16x4f64.  The summation runs in steps of
// 16 through two LDS buffers: the next step's global loads are in flight
// (in registers) while the current step multiplies, then land in the other
// buffer; one workgroup barrier per step.  BN = 5 makes the 80-column tiles
// that divide the 400-function orbital dimension of the exchange products.
// blockIdx.z is a batch index (split = 0: A, B, C offset by sA, sB, sC) or a
// slice of the summation index (split = 1: columns [z kchunk, (z+1) kchunk)
// of A and rows of B, written to C + z sC).
// Lane layout of v_mfma_f64_16x16x4f64 on gfx950 (measured with one
// wavefront of marked inputs): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[l / 16 + 4 r][l % 16] in accumulator element r.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int KT = 16;  // summation step

struct GemmArgs {
  const double* A;
  const double* B;
  double* C;
  long long lda, ldb, ldc;
  long long sA, sB, sC;
  int M, N, K;
  int bt, split, kchunk;
  int skip_lower;  // 1: tiles whose every row is below every column of C are not computed
  // 1: one-dimensional grid in an XCD-aware order.  Workgroups go to the 8
  // XCDs (each with its own L2) round-robin by block id, so block b runs on
  // XCD b % 8: split = 1 puts every tile of slice z on XCD z % 8 (the slice's
  // rows of A and B are read into one L2), split = 0 every column tile of row
  // tile m on XCD m % 8 (the row tile of A is read once).  mt, nt: tiles
  // along M and N; nz: slices.
  int xcd, mt, nt, nz;
};

template <int BM, int BN>
__global__ __launch_bounds__(256) void dgemm_mfma(GemmArgs g) {
  constexpr int TM = 64 * BM, TN = 16 * BN;
  constexpr int NA = TM * KT / 256;              // A entries per thread per step
  constexpr int NB = (KT * TN + 255) / 256;      // B entries per thread per step
  __shared__ double As[2][TM][KT + 1];
  __shared__ double Bs[2][KT][TN + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int mi = blockIdx.x, ni = blockIdx.y, z = blockIdx.z;
  if (g.xcd) {  // every return below is uniform over the workgroup, before any barrier
    const int b = blockIdx.x;
    if (g.split) {
      const int T = g.mt * g.nt;
      const int grp = b / (8 * T), w = b - grp * 8 * T;
      const int t = w / 8;
      z = grp * 8 + (w - t * 8);
      if (z >= g.nz) return;
      mi = t / g.nt;
      ni = t - mi * g.nt;
    } else {
      const int grp = b / (8 * g.nt), w = b - grp * 8 * g.nt;
      ni = w / 8;
      mi = grp * 8 + (w - ni * 8);
      if (mi >= g.mt) return;
      z = 0;
    }
  }
  const int m0 = mi * TM, n0 = ni * TN;
  if (g.skip_lower && m0 > n0 + TN - 1) return;  // uniform over the workgroup, before any barrier
  const double* A = g.A;
  const double* B = g.B;
  double* C = g.C + z * g.sC;
  int kb = 0, ke = g.K;
  if (g.split) {
    kb = z * g.kchunk;
    ke = min(g.K, kb + g.kchunk);
  } else {
    A += z * g.sA;
    B += z * g.sB;
  }
  const int wm = wave * 16 * BM;
  v4d acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  double ra[NA], rb[NB];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {  // 16 lanes per row of A
      const int e = tid + 256 * r;
      const int gm = m0 + e / KT, gk = k0 + e % KT;
      ra[r] = (gm < g.M && gk < ke) ? A[(long long)gm * g.lda + gk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int e = tid + 256 * r;
      double x = 0.0;
      if (e < KT * TN) {
        if (g.bt) {  // rows of B^T: contiguous along the summation index
          const int gn = n0 + e / KT, gk = k0 + e % KT;
          if (gn < g.N && gk < ke) x = B[(long long)gn * g.ldb + gk];
        } else {  // rows of B: contiguous along N
          const int gk = k0 + e / TN, gn = n0 + e % TN;
          if (gn < g.N && gk < ke) x = B[(long long)gk * g.ldb + gn];
        }
      }
      rb[r] = x;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int e = tid + 256 * r;
      As[buf][e / KT][e % KT] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      const int e = tid + 256 * r;
      if (e < KT * TN) {
        if (g.bt)
          Bs[buf][e % KT][e / KT] = rb[r];
        else
          Bs[buf][e / TN][e % TN] = rb[r];
      }
    }
  };
  if (kb < ke) {
    load(kb);
    store(0);
  }
  __syncthreads();
  int cur = 0;
  for (int k0 = kb; k0 < ke; k0 += KT) {
    const bool more = k0 + KT < ke;
    if (more) load(k0 + KT);
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      double a[BM], b[BN];
#pragma unroll
      for (int i = 0; i < BM; ++i) a[i] = As[cur][wm + 16 * i + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < BN; ++j) b[j] = Bs[cur][kk + (lane >> 4)][16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < BM; ++i)
#pragma unroll
        for (int j = 0; j < BN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
        const int col = n0 + 16 * j + (lane & 15);
        if (row < g.M && col < g.N) C[(long long)row * g.ldc + col] = acc[i][j][r];
      }
}

// K[r][c] = K[c][r] for r > c (n x n, row-major)
__global__ __launch_bounds__(256) void mirror_lower(double* __restrict__ K, int n) {
  const long long nn = (long long)n * n;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < nn; i += 256ll * gridDim.x) {
    const int r = int(i / n), c = int(i - (long long)r * n);
    if (r > c) K[i] = K[(long long)c * n + r];
  }
}

// out[i] = sum over the ns slices part[z slice + i]
__global__ __launch_bounds__(256) void sum_slices(const double* __restrict__ part, long long slice, int ns,
                                                  long long n, double* __restrict__ out) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
    double s = 0.0;
    for (int z = 0; z < ns; ++z) s += part[z * slice + i];
    out[i] = s;
  }
}

// ---------------------------------------------------------------------------
// Coulomb passes over B[a][P][b] (both HBM-bound: one read of B each).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[a][P] = sum_b B[a][P][b] D[a][b]: workgroup (a, 32 rows of P), each
// wavefront 8 rows at a time (eight loads in flight per lane, lanes along b),
// then eight wavefront sums.
constexpr int kGammaRows = 32;
__global__ __launch_bounds__(256) void df_gamma_part(const double* __restrict__ B, const double* __restrict__ D,
                                                     int n, int naux, double* __restrict__ part) {
  const int a = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int P0 = blockIdx.y * kGammaRows + wave * 8;
  const double* Da = D + (long long)a * n;
  const double* Ba = B + ((long long)a * naux + P0) * n;
  double s[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) s[r] = 0.0;
  for (int b = lane; b < n; b += 64) {
    const double d = Da[b];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (P0 + r < naux) s[r] = fma(Ba[(long long)r * n + b], d, s[r]);
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) s[r] = wave_sum(s[r]);
  if (lane < 8 && P0 + lane < naux) {
    double v = s[0];
#pragma unroll
    for (int r = 1; r < 8; ++r)
      if (lane == r) v = s[r];
    part[(long long)a * naux + P0 + lane] = v;
  }
}

// gamma[P] = sum_a part[a][P]: workgroup of 64 P (lanes) and 16 wavefronts,
// each summing every 16th a, then a fixed-order sum of the 16 partials.
__global__ __launch_bounds__(1024) void df_gamma_sum(const double* __restrict__ part, int n, int naux,
                                                     double* __restrict__ gamma) {
  __shared__ double red[16][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int P = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (P < naux)
    for (int a = wave; a < n; a += 16) s += part[(long long)a * naux + P];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && P < naux) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += red[w][lane];
    gamma[P] = t;
  }
}

// gamma[P] = sum_a X[a][P][a] with X[a][P][c] = sum_b B[a][P][b] D[b][c]:
// each B_P is symmetric in (a, b), so sum_ab B_P,ab D_ab = sum_ab B_P,ab D_ba =
// sum_a (B_P D)_aa for any D.  Reads the n n_aux diagonal entries of X
// instead of all of B; workgroup shape and summation order as df_gamma_sum.
__global__ __launch_bounds__(1024) void df_gamma_diag(const double* __restrict__ X, int n, int naux,
                                                      double* __restrict__ gamma) {
  __shared__ double red[16][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int P = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (P < naux)
    for (int a = wave; a < n; a += 16) s += X[((long long)a * naux + P) * n + a];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && P < naux) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) t += red[w][lane];
    gamma[P] = t;
  }
}

// J[a][b] = sum_P B[a][P][b] gamma[P]: workgroup (64 columns b, a), the four
// wavefronts take every fourth P (lanes along b, two accumulators each) and
// meet in LDS.
__global__ __launch_bounds__(256) void df_coulomb(const double* __restrict__ B, const double* __restrict__ gamma,
                                                  int n, int naux, double* __restrict__ J) {
  __shared__ double red[4][64];
  const int a = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x * 64 + lane;
  double s0 = 0.0, s1 = 0.0;
  if (b < n) {
    const double* Ba = B + (long long)a * naux * n + b;
    int P = wave;
    for (; P + 4 < naux; P += 8) {
      s0 = fma(Ba[(long long)P * n], gamma[P], s0);
      s1 = fma(Ba[(long long)(P + 4) * n], gamma[P + 4], s1);
    }
    if (P < naux) s0 = fma(Ba[(long long)P * n], gamma[P], s0);
  }
  red[wave][lane] = s0 + s1;
  __syncthreads();
  if (wave == 0 && b < n) J[(long long)a * n + b] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// ---------------------------------------------------------------------------
// Three-centre scatter: raw blocks of the quartets (a b|P u) of one orbital
// shell pair, P over all auxiliary shells in order, start at base; the block
// of auxiliary shell P starts na nb ps(P) further (ps = the shell's first
// auxiliary function) and is row-major [na][nb][nP].  Writes T[a][P][b] and,
// for a != b, T[b][P][a].
// ---------------------------------------------------------------------------
struct Pair3 {
  int fa, fb, na, nb;
  long long base;
  int same, pad;
};

__global__ __launch_bounds__(256) void df_scatter3(const double* __restrict__ blk, const Pair3* __restrict__ prs,
                                                   const int* __restrict__ pstart, const int* __restrict__ pn,
                                                   int naux, int n, double* __restrict__ T) {
  const Pair3 q = prs[blockIdx.x];
  const int tot = q.na * q.nb * naux;
  for (int e = threadIdx.x; e < tot; e += 256) {
    const int p = e % naux, ij = e / naux;
    const int i = ij / q.nb, j = ij - i * q.nb;
    const int ps = pstart[p], np = pn[p];
    const double v = blk[q.base + (long long)q.na * q.nb * ps + (long long)(i * q.nb + j) * np + (p - ps)];
    T[((long long)(q.fa + i) * naux + p) * n + q.fb + j] = v;
    if (!q.same) T[((long long)(q.fb + j) * naux + p) * n + q.fa + i] = v;
  }
}

template <typename T>
struct Dev {
  T* p = nullptr;
  ~Dev() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  hipError_t alloc(size_t n) {
    release();
    return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (n ? n : 1));
  }
  T* take() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

// C tiles of 128 x 80 (BM = 2) for tall products, 64 x 80 (BM = 1) when M is
// small (the exchange product, M = n_orb).
int gemm(const GemmArgs& g0, int gz, hipStream_t st) {
  GemmArgs g = g0;
  if (g.M <= 0 || g.N <= 0 || gz <= 0) return LAI_OK;
  const bool tall = g.M >= 1024;
  const int TM = tall ? 128 : 64, TN = 80;
  dim3 grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN, gz);
  if (g.xcd) {
    if (!g.split && gz != 1)
      return set_error(LAI_ERR_INVALID_ARGUMENT, "density fitting: XCD order needs one batch or a split product");
    g.mt = int(grid.x);
    g.nt = int(grid.y);
    g.nz = gz;
    const long long blocks = g.split ? (long long)(gz + 7) / 8 * 8 * g.mt * g.nt
                                     : (long long)(g.mt + 7) / 8 * 8 * g.nt;
    if (blocks > 0x7fffffffll)
      return set_error(LAI_ERR_INVALID_ARGUMENT, "density fitting: GEMM grid beyond the launch limits");
    grid = dim3(unsigned(blocks), 1, 1);
  }
  if (grid.y > 65535u || grid.z > 65535u)
    return set_error(LAI_ERR_INVALID_ARGUMENT, "density fitting: GEMM grid beyond the launch limits");
  if (tall)
    hipLaunchKernelGGL((dgemm_mfma<2, 5>), grid, dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((dgemm_mfma<1, 5>), grid, dim3(256), 0, st, g);
  DF_HIP(hipGetLastError());
  return LAI_OK;
}

// Lower Cholesky factor of the row-major SPD matrix A in place (upper part
// zeroed): the row-oriented algorithm of math::cholesky_decompose
// (include/libaccint/math/cholesky.hpp), which fails when a pivot is <= 0.
bool cholesky_lower(std::vector<double>& A, int n) {
  for (int i = 0; i < n; ++i) {
    double* Ai = &A[size_t(i) * n];
    double s = Ai[i];
    for (int k = 0; k < i; ++k) s -= Ai[k] * Ai[k];
    if (!(s > 0.0)) return false;
    Ai[i] = std::sqrt(s);
    for (int j = i + 1; j < n; ++j) {
      double* Aj = &A[size_t(j) * n];
      double t = Aj[i];
      for (int k = 0; k < i; ++k) t -= Aj[k] * Ai[k];
      Aj[i] = t / Ai[i];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) A[size_t(i) * n + j] = 0.0;
  return true;
}

// X = L^-1 for lower-triangular L (row-major), by forward substitution row by
// row: X[i][j] = (delta_ij - sum_{k=j}^{i-1} L[i][k] X[k][j]) / L[i][i], the
// sum taken in increasing k as in math::triangular_inverse.
void lower_inverse(const std::vector<double>& L, std::vector<double>& X, int n) {
  X.assign(size_t(n) * n, 0.0);
  std::vector<double> acc(n);
  for (int i = 0; i < n; ++i) {
    const double* Li = &L[size_t(i) * n];
    std::fill(acc.begin(), acc.begin() + i + 1, 0.0);
    acc[i] = 1.0;
    for (int k = 0; k < i; ++k) {
      const double l = Li[k];
      const double* Xk = &X[size_t(k) * n];
      for (int j = 0; j <= k; ++j) acc[j] -= l * Xk[j];
    }
    double* Xi = &X[size_t(i) * n];
    for (int j = 0; j <= i; ++j) Xi[j] = acc[j] / Li[i];
  }
}

}  // namespace

struct lai_df {
  int n = 0, naux = 0;
  int nsplit = 1, kchunk = 0;          // slices of the exchange product
  int nsplit_sym = 1, kchunk_sym = 0;  // the same over the tiles a symmetric D needs
  double* B = nullptr;     // [n][naux][n]
  double* X = nullptr;     // [n][naux][n]
  double* part = nullptr;   // nsplit * n * n: slices of the exchange product
  double* gpart = nullptr;  // [n][naux]: partial gamma
  double* gamma = nullptr;  // [naux]
  hipEvent_t ev[6] = {};
  bool have_ev = false;
  // the Coulomb passes run beside the exchange GEMMs on this stream
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

namespace {
void df_release(lai_df* d) {
  if (!d) return;
  if (d->B) (void)hipFree(d->B);
  if (d->X) (void)hipFree(d->X);
  if (d->part) (void)hipFree(d->part);
  if (d->gpart) (void)hipFree(d->gpart);
  if (d->gamma) (void)hipFree(d->gamma);
  if (d->side) (void)hipStreamDestroy(d->side);
  if (d->fork) (void)hipEventDestroy(d->fork);
  if (d->join) (void)hipEventDestroy(d->join);
  if (d->have_ev)
    for (auto& e : d->ev) (void)hipEventDestroy(e);
  delete d;
}
}  // namespace

extern "C" int lai_dev_df_free(lai_df* df) {
  df_release(df);
  return LAI_OK;
}

extern "C" int lai_dev_df_dims(const lai_df* df, int* n_orb, int* n_aux) {
  if (!df) return set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_df_dims: null handle");
  if (n_orb) *n_orb = df->n;
  if (n_aux) *n_aux = df->naux;
  return LAI_OK;
}

extern "C" int lai_dev_df_create(lai_basis* b, int n_orb_shells, int n_aux_shells, int metric,
                                 double cond_hard_limit, double* cond_estimate, lai_df** out) {
  if (!b || !out || n_orb_shells < 1 || n_aux_shells < 1)
    return set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_df_create: invalid arguments");
  *out = nullptr;
  const lai::BasisView v = lai::basis_view(b);
  if (v.n_mol != 1 || v.ns != n_orb_shells + n_aux_shells + 1)
    return set_error(LAI_ERR_INVALID_ARGUMENT,
                     "lai_dev_df_create: the handle must hold one molecule of orbital shells, auxiliary shells "
                     "and one unit shell");
  const int u = v.ns - 1;
  if (v.L[u] != 0 || v.nprim[u] != 1 || v.exps[v.prim_off[u]] != 0.0 || v.coefs[v.prim_off[u]] != 1.0)
    return set_error(LAI_ERR_INVALID_ARGUMENT,
                     "lai_dev_df_create: the last shell must be the unit s function (exponent 0, coefficient 1)");
  if (metric == 2)
    return set_error(LAI_ERR_INVALID_ARGUMENT,
                     "AttenuatedCoulomb metric is not yet supported. Use Coulomb or Overlap metric instead.");
  if (metric != 0 && metric != 1) return set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_df_create: unknown metric");
  const int n = v.func_off[n_orb_shells];
  const int naux = v.func_off[u] - n;
  const int a0 = n_orb_shells;  // first auxiliary shell

  // ---- metric V (host, naux x naux) ----
  std::vector<double> V(size_t(naux) * naux, 0.0);
  if (metric == 0) {
    std::vector<int> s4;
    for (int P = 0; P < n_aux_shells; ++P)
      for (int Q = P; Q < n_aux_shells; ++Q) {
        const int q[4] = {a0 + P, u, a0 + Q, u};
        s4.insert(s4.end(), q, q + 4);
      }
    const int64_t nq = int64_t(s4.size() / 4);
    std::vector<int64_t> off(nq + 1);
    int64_t tot = 0;
    for (int P = 0; P < n_aux_shells; ++P)
      for (int Q = P; Q < n_aux_shells; ++Q) tot += int64_t(ncart_of(v.L[a0 + P])) * ncart_of(v.L[a0 + Q]);
    std::vector<double> blk(static_cast<size_t>(tot));
    int rc = lai_dev_eri_quartets(b, nq, s4.data(), blk.data(), off.data());
    if (rc != LAI_OK) return rc;
    int64_t t = 0;
    for (int P = 0; P < n_aux_shells; ++P)
      for (int Q = P; Q < n_aux_shells; ++Q, ++t) {
        const int fP = v.func_off[a0 + P] - n, fQ = v.func_off[a0 + Q] - n;
        const int nP = ncart_of(v.L[a0 + P]), nQ = ncart_of(v.L[a0 + Q]);
        for (int i = 0; i < nP; ++i)
          for (int j = 0; j < nQ; ++j) {
            const double x = blk[size_t(off[t]) + size_t(i) * nQ + j];
            V[size_t(fP + i) * naux + fQ + j] = x;
            V[size_t(fQ + j) * naux + fP + i] = x;
          }
      }
  } else {
    Dev<double> dS;
    DF_HIP(dS.alloc(size_t(v.nbf) * v.nbf));
    int rc = lai_dev_onee(b, 0, 0, nullptr, nullptr, dS.p, nullptr);
    if (rc != LAI_OK) return rc;
    DF_HIP(hipDeviceSynchronize());
    std::vector<double> S(size_t(v.nbf) * v.nbf);
    DF_HIP(hipMemcpy(S.data(), dS.p, sizeof(double) * S.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < naux; ++i)
      for (int j = 0; j < naux; ++j) V[size_t(i) * naux + j] = S[size_t(n + i) * v.nbf + n + j];
  }
  double dmax = 0.0, dmin = HUGE_VAL;
  for (int i = 0; i < naux; ++i) {
    const double d = V[size_t(i) * naux + i];
    dmax = std::max(dmax, d);
    if (d > 0.0) dmin = std::min(dmin, d);
  }
  const double cond = dmin < HUGE_VAL ? dmax / dmin : HUGE_VAL;
  if (cond_estimate) *cond_estimate = cond;
  if (cond_hard_limit > 0.0 && cond > cond_hard_limit)
    return set_error(LAI_ERR_INVALID_ARGUMENT,
                     "DF metric matrix is severely ill-conditioned (condition estimate = " + std::to_string(cond) +
                         ", hard limit = " + std::to_string(cond_hard_limit) + ")");
  if (!cholesky_lower(V, naux))
    return set_error(LAI_ERR_NUMERIC, "Cholesky decomposition failed: matrix is not positive definite");
  std::vector<double> Linv;
  lower_inverse(V, Linv, naux);
  V.clear();
  V.shrink_to_fit();

  // ---- three-centre integrals T[a][P][b] ----
  std::vector<int> pstart(naux), pn(naux);
  for (int P = 0; P < n_aux_shells; ++P) {
    const int f = v.func_off[a0 + P] - n, m = ncart_of(v.L[a0 + P]);
    for (int i = 0; i < m; ++i) {
      pstart[f + i] = f;
      pn[f + i] = m;
    }
  }
  const size_t nB = size_t(n) * naux * n;
  Dev<double> dT, dLinv, dB;
  Dev<int> dps, dpn;
  DF_HIP(dT.alloc(nB));
  DF_HIP(dps.alloc(naux));
  DF_HIP(dpn.alloc(naux));
  DF_HIP(hipMemcpy(dps.p, pstart.data(), sizeof(int) * naux, hipMemcpyHostToDevice));
  DF_HIP(hipMemcpy(dpn.p, pn.data(), sizeof(int) * naux, hipMemcpyHostToDevice));
  {
    // orbital shell pairs in batches of at most ~4M quartets / 64M doubles
    const int64_t kMaxQ = int64_t(1) << 22, kMaxD = int64_t(1) << 26;
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < n_orb_shells; ++i)
      for (int j = i; j < n_orb_shells; ++j) pairs.emplace_back(i, j);
    size_t p0 = 0;
    Dev<double> dblk;
    Dev<Pair3> dpr;
    int64_t blk_cap = 0, pr_cap = 0;
    while (p0 < pairs.size()) {
      size_t p1 = p0;
      int64_t nq = 0, nd = 0;
      while (p1 < pairs.size()) {
        const int64_t d = int64_t(ncart_of(v.L[pairs[p1].first])) * ncart_of(v.L[pairs[p1].second]) * naux;
        if (p1 > p0 && (nq + n_aux_shells > kMaxQ || nd + d > kMaxD)) break;
        nq += n_aux_shells;
        nd += d;
        ++p1;
      }
      std::vector<int> s4;
      s4.reserve(size_t(nq) * 4);
      std::vector<Pair3> prs;
      int64_t base = 0;
      for (size_t p = p0; p < p1; ++p) {
        const int i = pairs[p].first, j = pairs[p].second;
        Pair3 r{};
        r.fa = v.func_off[i];
        r.fb = v.func_off[j];
        r.na = ncart_of(v.L[i]);
        r.nb = ncart_of(v.L[j]);
        r.base = base;
        r.same = (i == j);
        prs.push_back(r);
        base += int64_t(r.na) * r.nb * naux;
        for (int P = 0; P < n_aux_shells; ++P) {
          const int q[4] = {i, j, a0 + P, u};
          s4.insert(s4.end(), q, q + 4);
        }
      }
      if (nd > blk_cap) {
        DF_HIP(dblk.alloc(size_t(nd)));
        blk_cap = nd;
      }
      if (int64_t(prs.size()) > pr_cap) {
        DF_HIP(dpr.alloc(prs.size()));
        pr_cap = int64_t(prs.size());
      }
      std::vector<int64_t> off(size_t(nq) + 1);
      int rc = lai_dev_eri_quartets_device(b, nq, s4.data(), dblk.p, off.data(), nullptr);
      if (rc != LAI_OK) return rc;
      if (off[size_t(nq)] != nd)
        return set_error(LAI_ERR_BACKEND, "lai_dev_df_create: three-centre block layout mismatch");
      DF_HIP(hipMemcpy(dpr.p, prs.data(), sizeof(Pair3) * prs.size(), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(df_scatter3, dim3(unsigned(prs.size())), dim3(256), 0, nullptr, dblk.p, dpr.p, dps.p, dpn.p,
                         naux, n, dT.p);
      DF_HIP(hipGetLastError());
      DF_HIP(hipDeviceSynchronize());
      p0 = p1;
    }
  }

  // ---- B[a] = L^-1 T[a] ----
  DF_HIP(dLinv.alloc(size_t(naux) * naux));
  DF_HIP(hipMemcpy(dLinv.p, Linv.data(), sizeof(double) * Linv.size(), hipMemcpyHostToDevice));
  DF_HIP(dB.alloc(nB));
  {
    GemmArgs g{};
    g.A = dLinv.p;
    g.B = dT.p;
    g.C = dB.p;
    g.lda = naux;
    g.ldb = n;
    g.ldc = n;
    g.sA = 0;
    g.sB = (long long)naux * n;
    g.sC = (long long)naux * n;
    g.M = naux;
    g.N = n;
    g.K = naux;
    int rc = gemm(g, n, nullptr);
    if (rc != LAI_OK) return rc;
    DF_HIP(hipDeviceSynchronize());
  }
  dT.release();
  dLinv.release();

  auto* d = new lai_df;
  d->n = n;
  d->naux = naux;
  // slices of the exchange GEMM: about 2048 workgroups (8 per CU) over its
  // tiles, all of them or those a symmetric D needs (gemm's tile rule)
  const int TMk = n >= 1024 ? 128 : 64, TNk = 80;
  long long tiles = 0, tiles_sym = 0;
  for (int mi = 0; mi < (n + TMk - 1) / TMk; ++mi)
    for (int ni = 0; ni < (n + TNk - 1) / TNk; ++ni) {
      ++tiles;
      if (!(mi * TMk > ni * TNk + TNk - 1)) ++tiles_sym;
    }
  const long long ksum = (long long)naux * n;
  auto slices = [&](long long t, int& ns_out, int& kc_out) {
    int ns = int(std::max(1ll, std::min(std::max(1ll, 2048ll / t), (ksum + 16 * KT - 1) / (16 * KT))));
    long long kc = (ksum + ns - 1) / ns;
    kc = (kc + KT - 1) / KT * KT;
    ns_out = int((ksum + kc - 1) / kc);
    kc_out = int(kc);
  };
  slices(tiles, d->nsplit, d->kchunk);
  slices(tiles_sym, d->nsplit_sym, d->kchunk_sym);
  const int ns = std::max(d->nsplit, d->nsplit_sym);
  Dev<double> dX, dpart, dgp, dg;
  hipError_t e;
  if ((e = dX.alloc(nB)) != hipSuccess || (e = dpart.alloc(size_t(ns) * n * n)) != hipSuccess ||
      (e = dgp.alloc(size_t(n) * naux)) != hipSuccess || (e = dg.alloc(size_t(naux))) != hipSuccess) {
    delete d;
    return set_error(LAI_ERR_MEMORY, std::string("lai_dev_df_create: ") + hipGetErrorString(e));
  }
  d->B = dB.take();
  d->X = dX.take();
  d->part = dpart.take();
  d->gpart = dgp.take();
  d->gamma = dg.take();
  if ((e = hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&d->fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&d->join, hipEventDisableTiming)) != hipSuccess) {
    df_release(d);
    return set_error(LAI_ERR_BACKEND, std::string("lai_dev_df_create: ") + hipGetErrorString(e));
  }
  *out = d;
  return LAI_OK;
}

namespace {
int df_run(lai_df* d, const double* D, double* J, double* K, int flags, hipStream_t st, double* ms) {
  const int n = d->n, naux = d->naux;
  auto mark = [&](int i) -> int {
    if (ms) DF_HIP(hipEventRecord(d->ev[i], st));
    return LAI_OK;
  };
  // ms[out] = time between events i0 and i1 (recorded on st, serial in the profiled runs)
  auto take = [&](int i0, int i1, int out) -> int {
    if (!ms) return LAI_OK;
    float t;
    DF_HIP(hipEventSynchronize(d->ev[i1]));
    DF_HIP(hipEventElapsedTime(&t, d->ev[i0], d->ev[i1]));
    ms[out] = t;
    return LAI_OK;
  };
  if (ms && !d->have_ev) {
    for (auto& e : d->ev) DF_HIP(hipEventCreate(&e));
    d->have_ev = true;
  }
  if (ms)
    for (int i = 0; i < 5; ++i) ms[i] = 0.0;
  const bool both = (flags & LAI_DF_J) && (flags & LAI_DF_K);
  auto coulomb = [&](hipStream_t s) -> int {
    hipLaunchKernelGGL(df_coulomb, dim3((n + 63) / 64, n), dim3(256), 0, s, d->B, d->gamma, n, naux, J);
    DF_HIP(hipGetLastError());
    return LAI_OK;
  };
  auto x_gemm = [&]() -> int {
    GemmArgs g{};
    g.A = d->B;  // rows (a, P), n columns
    g.B = D;
    g.C = d->X;
    g.lda = n;
    g.ldb = n;
    g.ldc = n;
    g.M = n * naux;
    g.N = n;
    g.K = n;
    g.xcd = 1;
    return gemm(g, 1, st);
  };
  auto k_product = [&]() -> int {  // K = X B^T over (P, b), then the slice sum
    int rc = mark(3);
    if (rc) return rc;
    GemmArgs h{};
    h.A = d->X;  // rows a, summation (P, b)
    h.B = d->B;  // given transposed: rows c, summation (P, b)
    h.C = d->part;
    h.lda = (long long)naux * n;
    h.ldb = (long long)naux * n;
    h.ldc = n;
    h.sC = (long long)n * n;
    h.M = n;
    h.N = n;
    h.K = naux * n;
    h.bt = 1;
    h.split = 1;
    h.skip_lower = (flags & LAI_DF_SYMMETRIC) ? 1 : 0;
    h.xcd = 1;
    h.kchunk = h.skip_lower ? d->kchunk_sym : d->kchunk;
    const int nsl = h.skip_lower ? d->nsplit_sym : d->nsplit;
    if ((rc = gemm(h, nsl, st))) return rc;
    if ((rc = mark(4))) return rc;
    const long long nn = (long long)n * n;
    hipLaunchKernelGGL(sum_slices, dim3(unsigned(std::min(2048ll, (nn + 255) / 256))), dim3(256), 0, st, d->part, nn,
                       nsl, nn, K);
    DF_HIP(hipGetLastError());
    if (flags & LAI_DF_SYMMETRIC) {  // the skipped tiles hold stale slices: take the upper triangle
      hipLaunchKernelGGL(mirror_lower, dim3(unsigned(std::min(2048ll, (nn + 255) / 256))), dim3(256), 0, st, K, n);
      DF_HIP(hipGetLastError());
    }
    if ((rc = mark(5))) return rc;
    if ((rc = take(3, 4, 3))) return rc;
    return take(4, 5, 4);
  };
  int rc;
  if (both) {
    // X = B D first: J's fitted density is the diagonal of X, so B is read
    // twice per build (X, Coulomb pass) plus once more with X by the K product.
    // Not profiled: the Coulomb pass (HBM-bound) runs on the side stream
    // beside the K product (matrix cores), joined back into st.
    const bool overlap = !ms && d->side;
    const hipStream_t sj = overlap ? d->side : st;
    if ((rc = mark(0))) return rc;
    if ((rc = x_gemm())) return rc;
    if ((rc = mark(1))) return rc;
    if ((rc = take(0, 1, 2))) return rc;
    if (overlap) {
      DF_HIP(hipEventRecord(d->fork, st));
      DF_HIP(hipStreamWaitEvent(d->side, d->fork, 0));
    }
    // from here on st is joined to the side stream on every return path, so
    // that a failed launch never leaves work writing J or gamma unordered
    // with the caller's stream
    auto side_part = [&]() -> int {
      int r;
      if ((r = mark(0))) return r;
      hipLaunchKernelGGL(df_gamma_diag, dim3((naux + 63) / 64), dim3(1024), 0, sj, d->X, n, naux, d->gamma);
      DF_HIP(hipGetLastError());
      if ((r = mark(1))) return r;
      if ((r = coulomb(sj))) return r;
      if ((r = mark(2))) return r;
      if ((r = take(0, 1, 0)) || (r = take(1, 2, 1))) return r;
      return k_product();
    };
    rc = side_part();
    if (overlap) {
      const hipError_t e1 = hipEventRecord(d->join, d->side);
      const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, d->join, 0) : e1;
      if (e2 != hipSuccess) {
        (void)hipStreamSynchronize(d->side);  // no event edge: order by waiting on the host
        if (rc == LAI_OK)
          rc = set_error(LAI_ERR_BACKEND, std::string("lai_dev_df_jk: join: ") + hipGetErrorString(e2));
      }
    }
    return rc;
  }
  if (flags & LAI_DF_J) {  // J alone: the fitted density from one pass over B
    if ((rc = mark(0))) return rc;
    hipLaunchKernelGGL(df_gamma_part, dim3(n, (naux + kGammaRows - 1) / kGammaRows), dim3(256), 0, st, d->B, D, n,
                       naux, d->gpart);
    DF_HIP(hipGetLastError());
    hipLaunchKernelGGL(df_gamma_sum, dim3((naux + 63) / 64), dim3(1024), 0, st, d->gpart, n, naux, d->gamma);
    DF_HIP(hipGetLastError());
    if ((rc = mark(1))) return rc;
    if ((rc = coulomb(st))) return rc;
    if ((rc = mark(2))) return rc;
    if ((rc = take(0, 1, 0)) || (rc = take(1, 2, 1))) return rc;
  }
  if (flags & LAI_DF_K) {
    if ((rc = mark(0))) return rc;
    if ((rc = x_gemm())) return rc;
    if ((rc = mark(1))) return rc;
    if ((rc = take(0, 1, 2))) return rc;
    if ((rc = k_product())) return rc;
  }
  return LAI_OK;
}
}  // namespace

extern "C" int lai_dev_df_jk(lai_df* df, const double* d_D, double* d_J, double* d_K, int flags, void* stream) {
  if (!df || !d_D || (flags & ~(LAI_DF_J | LAI_DF_K | LAI_DF_SYMMETRIC)) || ((flags & LAI_DF_J) && !d_J) ||
      ((flags & LAI_DF_K) && !d_K))
    return set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_df_jk: invalid arguments");
  return df_run(df, d_D, d_J, d_K, flags, hipStream_t(stream), nullptr);
}

extern "C" int lai_dev_df_jk_profile(lai_df* df, const double* d_D, double* d_J, double* d_K, int flags,
                                     void* stream, double* ms) {
  if (!df || !d_D || !ms || (flags & ~(LAI_DF_J | LAI_DF_K | LAI_DF_SYMMETRIC)) || ((flags & LAI_DF_J) && !d_J) ||
      ((flags & LAI_DF_K) && !d_K))
    return set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_df_jk_profile: invalid arguments");
  return df_run(df, d_D, d_J, d_K, flags, hipStream_t(stream), ms);
}
