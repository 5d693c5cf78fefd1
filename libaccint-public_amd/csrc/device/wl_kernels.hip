// wl_kernels.hip -- worklist statistics and batched symmetrisation.
#include "wl_kernels.hpp"

namespace lai {

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One lane per quartet over the chunks of one class, 256 lanes per chunk.
__global__ __launch_bounds__(256) void worklist_stats_kernel(WlStatsArgs a) {
  const int tid = threadIdx.x;
  unsigned long long n = 0, n32 = 0, neri = 0, sum = 0;
  double fm64 = 0.0, fm32 = 0.0, fx64 = 0.0, fx32 = 0.0, g0 = 0.0, g1 = 0.0;
  for (int ci = blockIdx.x; ci < a.nchunks; ci += gridDim.x) {
    const JkChunk ch = a.chunks[ci];
    const PairRec bra = a.pairs[ch.bra];
    const JkSegs* sg = a.segs + ch.seg;
    int seg = 0;
    for (int k = ch.k0 + tid; k < ch.k1; k += blockDim.x) {
      const int kp = chunk_ket(a.kets, sg, k, seg);
      const PairRec ket = a.pairs[kp];
      const int mode = quartet_select(a.sel, ch.bra, bra, kp, ket);
      if (mode == 0) continue;
      // reference orientation: the pair with the smaller index first, each
      // pair's shells in index order
      const PairRec& p0 = ch.bra <= kp ? bra : ket;
      const PairRec& p1 = ch.bra <= kp ? ket : bra;
      const int i = min(p0.sa, p0.sb), j = max(p0.sa, p0.sb), kk = min(p1.sa, p1.sb), l = max(p1.sa, p1.sb);
      const unsigned long long key =
          ((unsigned long long)(unsigned short)a.sel.sh_loc[i] << 48) |
          ((unsigned long long)(unsigned short)a.sel.sh_loc[j] << 32) |
          ((unsigned long long)(unsigned short)a.sel.sh_loc[kk] << 16) |
          (unsigned long long)(unsigned short)a.sel.sh_loc[l];
      sum += mix64(key);
      const int Li = a.L[i], Lj = a.L[j], Lk = a.L[kk], Ll = a.L[l];
      const double F = a.fp[((Li * 4 + Lj) * 4 + Lk) * 4 + Ll];
      const double fm = double(bra.npp) * ket.npp * F, fx = double(bra.npp_sig) * ket.npp_sig * F;
      const int gi = ((Li * 4 + Lj) * 4 + Lk) * 4 + Ll;
      g0 += double(bra.npp_sig) * ket.npp_sig * a.fg[gi];
      g1 += double(bra.npp_sig) * ket.npp_sig * a.fg[256 + gi];
      neri += (unsigned long long)(ncart(Li) * ncart(Lj)) * (ncart(Lk) * ncart(Ll));
      ++n;
      if (mode == 2) {
        ++n32;
        fm32 += fm;
        fx32 += fx;
      } else {
        fm64 += fm;
        fx64 += fx;
      }
    }
  }
  n = wave_sum_u64(n);
  n32 = wave_sum_u64(n32);
  neri = wave_sum_u64(neri);
  sum = wave_sum_u64(sum);
  fm64 = wave_sum(fm64);
  fm32 = wave_sum(fm32);
  fx64 = wave_sum(fx64);
  fx32 = wave_sum(fx32);
  g0 = wave_sum(g0);
  g1 = wave_sum(g1);
  if ((tid & 63) == 0 && n > 0) {
    atomic_add(&a.gflops[2 * a.cls], g0);
    atomic_add(&a.gflops[2 * a.cls + 1], g1);
    atomicAdd(&a.counts[3 * a.cls], n);
    atomicAdd(&a.counts[3 * a.cls + 1], n32);
    atomicAdd(&a.counts[3 * a.cls + 2], neri);
    atomicAdd(a.checksum, sum);  // modulo 2^64
    atomic_add(&a.flops[4 * a.cls], fm64);
    atomic_add(&a.flops[4 * a.cls + 1], fm32);
    atomic_add(&a.flops[4 * a.cls + 2], fx64);
    atomic_add(&a.flops[4 * a.cls + 3], fx32);
  }
}

__global__ void symmetrize_add_batch_kernel(const double* __restrict__ Jt, const double* __restrict__ Kt,
                                            double* __restrict__ J, double* __restrict__ K,
                                            const long long* __restrict__ moff, const int* __restrict__ mnbf,
                                            int anti) {
  const int m = blockIdx.y;
  const size_t n = size_t(mnbf[m]), nn = n * n;
  const size_t o = size_t(moff[m]);
  for (size_t x = size_t(blockIdx.x) * blockDim.x + threadIdx.x; x < nn; x += size_t(gridDim.x) * blockDim.x) {
    const size_t i = x / n, j = x - i * n;
    const size_t t = j * n + i;
    if (anti) {
      K[o + x] += 0.5 * (Kt[o + x] - Kt[o + t]);
    } else {
      J[o + x] += 0.5 * (Jt[o + x] + Jt[o + t]);
      K[o + x] += 0.5 * (Kt[o + x] + Kt[o + t]);
    }
  }
}

}  // namespace lai
