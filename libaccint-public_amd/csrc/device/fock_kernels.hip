// fock_kernels.hip -- Fock digestion of precomputed ERI blocks (the plain
// N^4 accumulate of GpuFockBuilder::accumulate and
// accumulate_device_eri_batch) and the stream ordering helper of the C ABI.
#include "lai_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "api_internal.hpp"
#include "common.hpp"

namespace lai {
namespace {

constexpr int kFockWaves = 4;      // wavefronts per workgroup, one block each
constexpr int kMaxCart = 10;       // Cartesian functions of an f shell
constexpr int kRows = kMaxCart * kMaxCart;

// One wavefront per block: the block's J rows (na x nb) and K rows (na x nc)
// are summed in LDS over the other two indices, then added to J and K with
// one atomic per element.  Blocks whose dimensions or function ranges do not
// fit are skipped (the host validates what it builds; this keeps a bad
// descriptor from writing outside the matrices).
__global__ __launch_bounds__(64 * kFockWaves) void fock_accumulate_kernel(
    const double* __restrict__ eri, const lai_eri_block* __restrict__ blocks, long long n_blocks,
    const double* __restrict__ D, double* J, double* K, int n) {
  __shared__ double jrow[kFockWaves][kRows];
  __shared__ double krow[kFockWaves][kRows];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* jr = jrow[w];
  double* kr = krow[w];
  for (long long q = (long long)blockIdx.x * kFockWaves + w; q < n_blocks;
       q += (long long)gridDim.x * kFockWaves) {
    const lai_eri_block bk = blocks[q];
    const int na = bk.na, nb = bk.nb, nc = bk.nc, nd = bk.nd;
    const bool ok = na >= 1 && na <= kMaxCart && nb >= 1 && nb <= kMaxCart && nc >= 1 && nc <= kMaxCart &&
                    nd >= 1 && nd <= kMaxCart && bk.off >= 0 && bk.fa >= 0 && bk.fb >= 0 && bk.fc >= 0 &&
                    bk.fd >= 0 && bk.fa + na <= n && bk.fb + nb <= n && bk.fc + nc <= n && bk.fd + nd <= n;
    if (!ok) continue;  // uniform over the wavefront
    for (int e = lane; e < kRows; e += 64) {
      jr[e] = 0.0;
      kr[e] = 0.0;
    }
    wave_sync();
    const int ncd = nc * nd, nbcd = nb * ncd, nel = na * nbcd;
    const double* g = eri + bk.off;
    for (int e = lane; e < nel; e += 64) {
      const int a = e / nbcd, r = e - a * nbcd;
      const int b = r / ncd, r2 = r - b * ncd;
      const int c = r2 / nd, d = r2 - c * nd;
      const double v = g[e];
      const size_t la = size_t(bk.fc + c), si = size_t(bk.fd + d), nu = size_t(bk.fb + b);
      lds_add(&jr[a * nb + b], v * D[la * n + si]);
      lds_add(&kr[a * nc + c], v * D[nu * n + si]);
    }
    wave_sync();
    for (int e = lane; e < na * nb; e += 64) {
      const int a = e / nb, b = e - a * nb;
      const double x = jr[e];
      if (x != 0.0) atomic_add(&J[size_t(bk.fa + a) * n + bk.fb + b], x);
    }
    for (int e = lane; e < na * nc; e += 64) {
      const int a = e / nc, c = e - a * nc;
      const double x = kr[e];
      if (x != 0.0) atomic_add(&K[size_t(bk.fa + a) * n + bk.fc + c], x);
    }
    wave_sync();  // the rows are read before the next block clears them
  }
}

}  // namespace
}  // namespace lai

extern "C" int lai_dev_fock_accumulate(const double* d_eri, const lai_eri_block* d_blocks, int64_t n_blocks,
                                       const double* d_D, double* d_J, double* d_K, int nbf, void* stream) {
  if (n_blocks < 0 || nbf <= 0 || (n_blocks > 0 && (!d_eri || !d_blocks || !d_D || !d_J || !d_K)))
    return lai::set_error(LAI_ERR_INVALID_ARGUMENT, "lai_dev_fock_accumulate: invalid arguments");
  if (n_blocks == 0) return LAI_OK;
  const long long groups = (n_blocks + lai::kFockWaves - 1) / lai::kFockWaves;
  const unsigned grid = unsigned(std::min<long long>(groups, 65536));
  hipLaunchKernelGGL(lai::fock_accumulate_kernel, dim3(grid), dim3(64 * lai::kFockWaves), 0, hipStream_t(stream),
                     d_eri, d_blocks, (long long)n_blocks, d_D, d_J, d_K, nbf);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return lai::set_error(LAI_ERR_BACKEND, std::string("lai_dev_fock_accumulate: ") + hipGetErrorString(e));
  return LAI_OK;
}

extern "C" int lai_dev_stream_wait(void* waiter, void* producer) {
  if (waiter == producer) return LAI_OK;
  hipEvent_t ev = nullptr;
  hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ev, hipStream_t(producer));
  if (e == hipSuccess) e = hipStreamWaitEvent(hipStream_t(waiter), ev, 0);
  if (ev) (void)hipEventDestroy(ev);  // released once the wait has been enqueued
  if (e != hipSuccess)
    return lai::set_error(LAI_ERR_BACKEND, std::string("lai_dev_stream_wait: ") + hipGetErrorString(e));
  return LAI_OK;
}
