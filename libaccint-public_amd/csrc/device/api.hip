// api.hip -- implementation of the C ABI in include/lai_device.h: context,
// device memory helpers, basis upload with primitive-pair tables, Schwarz
// bounds, the screened AM-sorted worklist, the fused J/K build (plain launch
// and hipGraph), raw ERI blocks, one-electron integrals and math probes.
#include "lai_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "api_internal.hpp"
#include "aux_kernels.hpp"
#include "eri_kernels.hpp"
#include "grad_kernels.hpp"
#include "launch.hpp"
#include "wl_kernels.hpp"

using namespace lai;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define LAI_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(LAI_ERR_BACKEND, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Launchers {
  ClassLauncher f64[kNumClasses] = {};
  ClassLauncher f32[kNumClasses] = {};
  ClassInfo info[kNumClasses] = {};
  Launchers() {
    fill_launchers_part_0(f64, f32, info);
    fill_launchers_part_1(f64, f32, info);
    fill_launchers_part_2(f64, f32, info);
    fill_launchers_part_3(f64, f32, info);
    fill_launchers_part_4(f64, f32, info);
    fill_launchers_part_5(f64, f32, info);
    fill_launchers_part_6(f64, f32, info);
    fill_launchers_part_7(f64, f32, info);
  }
};
const Launchers& launchers() {
  static Launchers L;
  return L;
}

struct GradLaunchers {
  GradLauncher f[kNumGradClasses] = {};
  GradLaunchers() { fill_grad_launchers(f); }
};
const GradLaunchers& grad_launchers() {
  static GradLaunchers G;
  return G;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t assign(const T* host, size_t count) {
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, sizeof(T) * count);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    n = count;
    return hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice);
  }
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, sizeof(T) * count);
    if (e != hipSuccess) p = nullptr;
    else n = count;
    return e;
  }
};

// primitive-quartet magnitude below which the J/K kernels skip a primitive pair
constexpr double kPrimCut = 1e-20;
// default relative primitive cut of the FP32 recursion (mixed precision)
constexpr double kMixedPrimCut = 1e-7;

inline int pair_index(int ns, int i, int j) {  // i <= j, reference loop order
  return i * ns - i * (i - 1) / 2 + (j - i);
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// F_dir of the reference recursion (src/host/kernels/eri_kernel.cpp:119-163):
// VRR terms 2t-1, two flops per horizontal-transfer entry (bra transfer on
// the d = 0 slice only).
double fdir(int La, int Lb, int Lc, int Ld) {
  const int LAB = La + Lb, LCD = Lc + Ld;
  double f = 0;
  for (int a = 0; a < LAB; ++a) f += 2 * (1 + (a > 0)) - 1;
  for (int c = 0; c < LCD; ++c)
    for (int a = 0; a <= LAB; ++a) f += 2 * (1 + (c > 0) + (a > 0)) - 1;
  for (int b = 0; b < Lb; ++b) f += 2.0 * (LAB - b) * (LCD + 1);
  for (int d = 0; d < Ld; ++d) f += 2.0 * (La + 1) * (Lb + 1) * (LCD - d);
  return f;
}
double fprim(int La, int Lb, int Lc, int Ld) {
  const int nr = (La + Lb + Lc + Ld) / 2 + 1;
  const double ncart4 = double(ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld));
  return 30.0 + nr * (30.0 + 20.0 + 3.0 * fdir(La, Lb, Lc, Ld) + 3.0 * ncart4);
}
// Gradient work per primitive quartet of a canonical quartet (La Lb|Lc Ld),
// the derivatives on three centres (the fourth by translational invariance):
// model 0 evaluates, per centre, the derivative classes with that centre's l
// raised and (l > 0) lowered by one, F_prim each, plus 9 N_cart contraction
// flops per root; model 1 builds one table per (root, axis) with the three
// centres raised together (what the adjoint kernels do):
// 30 + n_r (50 + 3 F_dir(La+1, Lb+1, Lc+1, Ld) + 12 N_cart), n_r of the raised
// total angular momentum.
double fgrad(int model, int La, int Lb, int Lc, int Ld) {
  const double ncart4 = double(ncart(La) * ncart(Lb) * ncart(Lc) * ncart(Ld));
  const int nrg = (La + Lb + Lc + Ld + 1) / 2 + 1;
  if (model == 1) return 30.0 + nrg * (50.0 + 3.0 * fdir(La + 1, Lb + 1, Lc + 1, Ld) + 12.0 * ncart4);
  double f = nrg * 9.0 * ncart4;
  const int l[4] = {La, Lb, Lc, Ld};
  for (int x = 0; x < 3; ++x)
    for (int d = -1; d <= 1; d += 2) {
      int m[4] = {l[0], l[1], l[2], l[3]};
      m[x] += d;
      if (m[x] < 0) continue;
      f += fprim(m[0], m[1], m[2], m[3]);
    }
  return f;
}

int hw_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return int(std::min(n, 32u));
}

// f(k, thread) for k in [0, n), dynamically scheduled in blocks of grain
template <typename F>
void parallel_for(int n, F&& f, int grain = 16) {
  const int nt = std::min(hw_threads(), std::max(1, n));
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (;;) {
        const int i = next.fetch_add(grain);
        if (i >= n) break;
        for (int k = i; k < std::min(n, i + grain); ++k) f(k, t);
      }
    });
  for (auto& x : th) x.join();
}
}  // namespace

// ---------------------------------------------------------------------------
// static staging array of lai_dev_eri_quartets_device_noalloc
// ---------------------------------------------------------------------------
// 2 MiB of device memory in the library's code object, one copy per GPU, held
// until the process ends.  Its name is only a host-side stand-in: host code
// uses the device address that hipGetSymbolAddress returns for the current
// device (kept per device in g_scratch).
constexpr size_t kScratchBytes = size_t(2) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;             // guards g_scratch and serialises the staging calls
void* g_scratch[kMaxDevices] = {};    // device address of scratch, per device

// Looks up the scratch address of `device` (current) once.
int init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(LAI_ERR_INVALID_ARGUMENT, "device index above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mtx);
  if (g_scratch[device]) return LAI_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(LAI_ERR_BACKEND, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = p;
  return LAI_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// opaque types
// ---------------------------------------------------------------------------
struct lai_ctx {
  int device = 0;
  int n_cu = 256;
  DevBuf<double> rys;
};

struct ClassList {
  DevBuf<JkChunk> chunks;  // bra-anchored chunks over the ket segments
  int64_t nchunks = 0;
  // one entry per (bra, segment list): k0 = 0, k1 = its length; chunks are
  // cut from these
  std::vector<JkChunk> runs;
  // chunks of the default rule when the tuned ones differ (the gradient
  // kernels are sized for those)
  DevBuf<JkChunk> gchunks;
  int64_t ngchunks = 0;
};

// Launch choice of one class (autotuner or caller table): alt selects the
// class's alternative kernel shape, chunk caps the quartets per chunk (0: the
// shape's default rule), bpc the workgroups per CU (0: the shape's default).
struct ClassTune {
  int alt = 0;
  int chunk = 0;
  int bpc = 0;
  bool is_default() const { return alt == 0 && chunk == 0 && bpc == 0; }
};

// class kernels of one J/K build run on the caller's stream and this many more
constexpr int kSideStreams = 3;

struct lai_basis {
  lai_ctx* ctx = nullptr;
  int ns = 0, nbf = 0;  // shells and basis functions of all molecules
  // molecule m: shells [mol_shell[m], mol_shell[m+1]), pairs
  // [mol_pair[m], mol_pair[m+1]), functions 0..mol_nbf[m]-1 and its D, J, K
  // blocks at mol_moff[m] in the batched matrices
  int n_mol = 1;
  std::vector<int> mol_shell, mol_pair, mol_nbf;
  std::vector<long long> mol_moff, mol_dmoff;
  std::vector<int> sh_mol, sh_loc;
  long long mat_size = 0;
  int max_nbf = 0;
  std::vector<int> L, nprim, prim_off, func_off;
  std::vector<double> center, exps, coefs;
  // pairs, molecule by molecule, each in reference order (i <= j)
  std::vector<PairRec> pairs_h;
  std::vector<int> pair_cls;  // pair class of each pair
  std::vector<int> pi, pj;    // shells of each pair
  DevBuf<PairRec> pairs;
  DevBuf<PrimPair> pp;
  DevBuf<double2> ppab;  // exponents of the two primitives of each primitive pair
  DevBuf<float> ppb;     // primitive-pair bounds |Kc| zeta^(-1/4)
  DevBuf<double> ppih;   // 1 / (2 zeta) of each primitive pair
  double pcut32 = kMixedPrimCut;
  DevBuf<int> shell_atom;
  int n_atoms = 0;
  int max_l = 0;
  // device copies of shell data for the one-electron kernel
  DevBuf<int> dL, dnprim, dprim_off, dfunc_off;
  DevBuf<double> dcenter, dexps, dcoefs;
  DevBuf<long long> dmoff;  // mol_moff on the device
  DevBuf<int> dmnbf;        // mol_nbf on the device
  DevBuf<int> dshloc;       // sh_loc on the device
  DevBuf<long long> ddmoff; // mol_dmoff on the device
  DevBuf<int> ddmns;        // shells per molecule on the device
  // Schwarz: packed per molecule, mol_pair[m] + j*(j+1)/2 + i (molecule-local shells)
  std::vector<double> Q;
  bool have_Q = false;
  DevBuf<double> dQ;  // per pair, pair order
  // worklist
  ClassList cls[kNumClasses];
  DevBuf<JkSegs> segs;
  DevBuf<int> kets;
  DevBuf<double> sdmax;  // per molecule: ns_m x ns_m shell-pair max|D|
  bool have_worklist = false;
  int wl_dscreen = 0, wl_mixed = 0;
  double wl_thr = 0.0, wl_tau = 0.0;
  int64_t counts[kNumClasses] = {};
  double cls_f64[kNumClasses] = {}, cls_f32[kNumClasses] = {};
  int64_t cls_n32[kNumClasses] = {};  // quartets of each class the worklist marks for FP32
  double cls_grad[2 * kNumClasses] = {};  // gradient work models 0 and 1 per class
  double flops64 = 0, flops32 = 0;
  int64_t n_eri = 0, n_fp32 = 0;
  uint64_t checksum = 0;
  // J/K scratch (unsymmetrised)
  DevBuf<double> Jt, Kt;
  // per-class launch choice
  ClassTune tune[kNumClasses];
  // queue of the persistent launch (LAI_JK_PERSISTENT), built from the
  // worklist on first use, and its claim counter
  DevBuf<PersistItem> pitems;
  int n_pitems = 0;
  bool pitems_valid = false;
  DevBuf<int> pcounter;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // side streams of the concurrent class launches (created before the first
  // J/K build) and the fork / join events that tie them to the caller's stream
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kSideStreams] = {};
};

// pair index of molecule-local shells il <= jl of molecule m
inline int mol_pair_index(const lai_basis* b, int m, int il, int jl) {
  const int ns = b->mol_shell[m + 1] - b->mol_shell[m];
  return b->mol_pair[m] + pair_index(ns, il, jl);
}
// packed Schwarz index of pair p
inline size_t packed_q_index(const lai_basis* b, int p) {
  const int m = b->sh_mol[b->pi[p]];
  const size_t i = size_t(b->sh_loc[b->pi[p]]), j = size_t(b->sh_loc[b->pj[p]]);
  return size_t(b->mol_pair[m]) + j * (j + 1) / 2 + i;
}

// the per-quartet selection of the current worklist
inline QuartetSel make_sel(const lai_basis* b) {
  QuartetSel s{};
  s.Qpair = b->dQ.p;
  s.sdmax = b->sdmax.p;
  s.dm_off = b->ddmoff.p;
  s.dm_ns = b->ddmns.p;
  s.sh_loc = b->dshloc.p;
  s.dscreen = b->wl_dscreen;
  s.thr = b->wl_thr;
  s.mixed = b->wl_mixed;
  s.tau = b->wl_tau;
  return s;
}

// ---------------------------------------------------------------------------
// context and memory
// ---------------------------------------------------------------------------
extern "C" const char* lai_dev_last_error(void) { return g_last_error.c_str(); }

extern "C" int lai_dev_device_count(int* n) {
  if (!n) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_device_count: null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return LAI_OK;
}

extern "C" int lai_dev_init(int device, lai_ctx** out) {
  if (!out) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_init: null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(LAI_ERR_BACKEND, "lai_dev_init: no HIP device available");
  if (device < 0 || device >= count)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_init: device index out of range");
  LAI_HIP(hipSetDevice(device));
  const int src = init_scratch(device);  // the device is current
  if (src != LAI_OK) return src;
  auto* ctx = new lai_ctx;
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
  std::vector<double> tab(LAI_DEVICE_TABLE_SIZE, 0.0);
  std::memcpy(tab.data(), lai_rys_cheb_h, sizeof(double) * LAI_RYS_CHEB_SIZE);
  for (int i = 0; i < int(sizeof(lai_rys_hx2) / sizeof(double)); ++i) {
    tab[LAI_RYS_HX2_OFF + i] = lai_rys_hx2[i];
    tab[LAI_RYS_HW_OFF + i] = lai_rys_hw[i];
  }
  std::memcpy(tab.data() + LAI_BT_OFF, lai_boys_taylor_h, sizeof(double) * LAI_BT_ROWS * LAI_BT_M);
  for (int k = 0; k < LAI_BT_ROWS; ++k) {
    // one-root rows: F_i / i!, the Taylor coefficients of F_0 (boys01_taylor)
    double fact = 1.0;
    for (int i = 0; i < 8; ++i) {
      if (i > 1) fact *= double(i);
      tab[LAI_B8_OFF + size_t(k) * 8 + i] = lai_boys_taylor_h[size_t(k) * LAI_BT_M + i] / fact;
    }
    std::memcpy(tab.data() + LAI_B10_OFF + size_t(k) * 10, lai_boys_taylor_h + size_t(k) * LAI_BT_M, sizeof(double) * 10);
  }
  // coarser one-root rows for the LDS copies: T_k = k/8 is row 2k of the
  // 1/16 grid; H_i = F_i / i! for i = 0..8
  static_assert(2 * (LAI_B9_ROWS - 1) < LAI_BT_ROWS && LAI_BT_M >= 9, "one-root rows out of the Taylor table");
  static_assert(LAI_B9_ROWS == 8 * lai_rys_tcut[1] + 1, "one-root rows must cover the one-root cut");
  for (int k = 0; k < LAI_B9_ROWS; ++k) {
    double fact = 1.0;
    for (int i = 0; i < 9; ++i) {
      if (i > 1) fact *= double(i);
      tab[LAI_B9_OFF + size_t(k) * 9 + i] = lai_boys_taylor_h[size_t(2 * k) * LAI_BT_M + i] / fact;
    }
  }
  hipError_t e = ctx->rys.assign(tab.data(), tab.size());
  if (e != hipSuccess) {
    delete ctx;
    return fail(LAI_ERR_BACKEND, std::string("lai_dev_init: table upload: ") + hipGetErrorString(e));
  }
  (void)launchers();
  *out = ctx;
  return LAI_OK;
}

extern "C" int lai_dev_fini(lai_ctx* ctx) {
  if (!ctx) return LAI_OK;
  ctx->rys.release();
  delete ctx;
  return LAI_OK;
}

extern "C" int lai_dev_malloc(lai_ctx*, size_t bytes, void** p) {
  if (!p) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_malloc: null output");
  *p = nullptr;
  if (bytes == 0) return LAI_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(LAI_ERR_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  return LAI_OK;
}

extern "C" int lai_dev_free(lai_ctx*, void* p) {
  if (p) LAI_HIP(hipFree(p));
  return LAI_OK;
}

extern "C" int lai_dev_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return LAI_OK;
  LAI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, hipStream_t(stream)));
  LAI_HIP(hipStreamSynchronize(hipStream_t(stream)));
  return LAI_OK;
}

extern "C" int lai_dev_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return LAI_OK;
  LAI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, hipStream_t(stream)));
  LAI_HIP(hipStreamSynchronize(hipStream_t(stream)));
  return LAI_OK;
}

extern "C" int lai_dev_zero(void* p, size_t n, void* stream) {
  if (n == 0) return LAI_OK;
  if (!p) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_zero: null pointer");
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(blocks)), dim3(256), 0, hipStream_t(stream),
                     static_cast<double*>(p), n);
  LAI_HIP(hipGetLastError());
  return LAI_OK;
}

extern "C" int lai_dev_stream_sync(void* stream) {
  LAI_HIP(hipStreamSynchronize(hipStream_t(stream)));
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// basis
// ---------------------------------------------------------------------------
extern "C" int lai_dev_upload_basis(lai_ctx* ctx, int ns, const int* L, const double* centers,
                                    const int* nprim, const double* exps, const double* coefs,
                                    lai_basis** out) {
  const int off[2] = {0, ns < 0 ? 0 : ns};
  return lai_dev_upload_batch(ctx, 1, off, ns, L, centers, nprim, exps, coefs, out);
}

extern "C" int lai_dev_upload_batch(lai_ctx* ctx, int n_mol, const int* mol_shell_off, int ns, const int* L,
                                    const double* centers, const int* nprim, const double* exps,
                                    const double* coefs, lai_basis** out) {
  if (!ctx || !out || ns < 0 || n_mol < 1 || !mol_shell_off ||
      (ns > 0 && (!L || !centers || !nprim || !exps || !coefs)))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_upload_batch: invalid arguments");
  if (ns > 65535) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_upload_batch: more than 65535 shells");
  if (mol_shell_off[0] != 0 || mol_shell_off[n_mol] != ns)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_upload_batch: molecule offsets must run from 0 to n_shells");
  for (int m = 0; m < n_mol; ++m)
    if (mol_shell_off[m + 1] < mol_shell_off[m])
      return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_upload_batch: decreasing molecule offsets");
  *out = nullptr;
  for (int s = 0; s < ns; ++s) {
    if (L[s] < 0 || L[s] > kMaxL)
      return fail(LAI_ERR_UNSUPPORTED, "lai_dev_upload_basis: angular momentum " + std::to_string(L[s]) +
                                           " of shell " + std::to_string(s) + " outside [0, 3]");
    if (nprim[s] <= 0) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_upload_basis: shell without primitives");
  }
  auto* b = new lai_basis;
  b->ctx = ctx;
  b->ns = ns;
  b->n_mol = n_mol;
  b->L.assign(L, L + ns);
  b->nprim.assign(nprim, nprim + ns);
  b->center.assign(centers, centers + 3 * ns);
  b->mol_shell.assign(mol_shell_off, mol_shell_off + n_mol + 1);
  b->sh_mol.resize(ns);
  b->sh_loc.resize(ns);
  int np = 0, nbf = 0;
  b->mol_pair.assign(1, 0);
  b->mol_moff.assign(1, 0);
  b->mol_dmoff.assign(1, 0);
  for (int m = 0; m < n_mol; ++m) {
    int mf = 0;
    for (int s = b->mol_shell[m]; s < b->mol_shell[m + 1]; ++s) {
      b->sh_mol[s] = m;
      b->sh_loc[s] = s - b->mol_shell[m];
      b->prim_off.push_back(np);
      b->func_off.push_back(mf);  // within the molecule
      np += nprim[s];
      mf += ncart(L[s]);
    }
    const long long nsm = b->mol_shell[m + 1] - b->mol_shell[m];
    b->mol_nbf.push_back(mf);
    b->mol_pair.push_back(b->mol_pair.back() + int(nsm * (nsm + 1) / 2));
    b->mol_moff.push_back(b->mol_moff.back() + (long long)mf * mf);
    b->mol_dmoff.push_back(b->mol_dmoff.back() + nsm * nsm);
    b->max_nbf = std::max(b->max_nbf, mf);
    nbf += mf;
  }
  b->nbf = nbf;
  b->mat_size = b->mol_moff.back();
  for (int s = 0; s < ns; ++s) b->max_l = std::max(b->max_l, L[s]);
  b->exps.assign(exps, exps + np);
  b->coefs.assign(coefs, coefs + np);

  // pair records and primitive-pair tables, molecule by molecule
  const double sq2pi52 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
  const int npairs = b->mol_pair.back();
  b->pairs_h.resize(npairs);
  b->pair_cls.resize(npairs);
  b->pi.resize(npairs);
  b->pj.resize(npairs);
  std::vector<PrimPair> pph;
  std::vector<double2> pabh;
  std::vector<double> pbound;
  double bmax = 0.0;
  int p = 0;
  for (int m = 0; m < n_mol; ++m)
  for (int i = b->mol_shell[m]; i < b->mol_shell[m + 1]; ++i)
    for (int j = i; j < b->mol_shell[m + 1]; ++j, ++p) {
      const int sa = (L[i] >= L[j]) ? i : j, sb = (sa == i) ? j : i;
      PairRec r{};
      r.sa = sa;
      r.sb = sb;
      r.fa = b->func_off[sa];
      r.fb = b->func_off[sb];
      r.same = (i == j);
      r.mn = b->mol_nbf[m];
      r.mol = m;
      r.moff = b->mol_moff[m];
      const double* A = &centers[3 * sa];
      const double* B = &centers[3 * sb];
      r.ABx = A[0] - B[0];
      r.ABy = A[1] - B[1];
      r.ABz = A[2] - B[2];
      const double AB2 = r.ABx * r.ABx + r.ABy * r.ABy + r.ABz * r.ABz;
      r.pp_off = int(pph.size());
      std::vector<std::pair<double, int>> order;
      std::vector<PrimPair> loc;
      std::vector<double2> locab;
      for (int x = 0; x < nprim[sa]; ++x) {
        const double al = exps[b->prim_off[sa] + x], ca = coefs[b->prim_off[sa] + x];
        if (ca == 0.0) continue;
        for (int y = 0; y < nprim[sb]; ++y) {
          const double be = exps[b->prim_off[sb] + y], cb = coefs[b->prim_off[sb] + y];
          if (cb == 0.0) continue;
          // two zero-exponent (unit) functions have no finite product; the
          // density-fitting basis pairs its unit function only with others
          if (al + be == 0.0) continue;
          const double zeta = al + be, iz = 1.0 / zeta, mu = al * be * iz;
          PrimPair q;
          q.zeta = zeta;
          q.Px = (al * A[0] + be * B[0]) * iz;
          q.Py = (al * A[1] + be * B[1]) * iz;
          q.Pz = (al * A[2] + be * B[2]) * iz;
          q.PAx = q.Px - A[0];
          q.PAy = q.Py - A[1];
          q.PAz = q.Pz - A[2];
          q.Kc = ca * cb * std::exp(-mu * AB2) * sq2pi52 * iz;
          // |prefactor| <= bound_ab * bound_cd since 1/sqrt(z+e) <= (z e)^(-1/4)
          const double bound = std::fabs(q.Kc) / std::pow(zeta, 0.25);
          order.emplace_back(bound, int(loc.size()));
          loc.push_back(q);
          locab.push_back(make_double2(al, be));
          bmax = std::max(bmax, bound);
        }
      }
      std::stable_sort(order.begin(), order.end(), [](const auto& u, const auto& v) { return u.first > v.first; });
      for (const auto& o : order) {
        pph.push_back(loc[o.second]);
        pabh.push_back(locab[o.second]);
        pbound.push_back(o.first);
      }
      r.npp = int(pph.size()) - r.pp_off;
      r.bsum = 0.0;
      for (int k = r.pp_off; k < r.pp_off + r.npp; ++k) r.bsum += pbound[k];
      b->pairs_h[p] = r;
      b->pair_cls[p] = pair_class(L[sa], L[sb]);
      b->pi[p] = i;
      b->pj[p] = j;
    }
  // primitive screening of the J/K kernels: a primitive pair whose bound times
  // the largest bound is below kPrimCut cannot contribute a primitive quartet
  // above that size; such pairs are sorted last and skipped
  for (auto& r : b->pairs_h) {
    int k = 0;
    while (k < r.npp && pbound[r.pp_off + k] * bmax >= kPrimCut) ++k;
    r.npp_sig = k;
  }
  std::vector<float> pbf(pbound.begin(), pbound.end());
  std::vector<int> dmns(n_mol);
  for (int m = 0; m < n_mol; ++m) dmns[m] = b->mol_shell[m + 1] - b->mol_shell[m];
  std::vector<double> pih(pph.size());
  for (size_t k = 0; k < pph.size(); ++k) pih[k] = 0.5 / pph[k].zeta;
  hipError_t e;
  if ((e = b->pairs.assign(b->pairs_h.data(), b->pairs_h.size())) != hipSuccess ||
      (e = b->pp.assign(pph.data(), pph.size())) != hipSuccess ||
      (e = b->ppb.assign(pbf.data(), pbf.size())) != hipSuccess ||
      (e = b->ppih.assign(pih.data(), pih.size())) != hipSuccess ||
      (e = b->ppab.assign(pabh.data(), pabh.size())) != hipSuccess ||
      (e = b->dL.assign(b->L.data(), ns)) != hipSuccess ||
      (e = b->dnprim.assign(b->nprim.data(), ns)) != hipSuccess ||
      (e = b->dprim_off.assign(b->prim_off.data(), ns)) != hipSuccess ||
      (e = b->dfunc_off.assign(b->func_off.data(), ns)) != hipSuccess ||
      (e = b->dcenter.assign(b->center.data(), b->center.size())) != hipSuccess ||
      (e = b->dexps.assign(b->exps.data(), b->exps.size())) != hipSuccess ||
      (e = b->dcoefs.assign(b->coefs.data(), b->coefs.size())) != hipSuccess ||
      (e = b->dmoff.assign(b->mol_moff.data(), b->mol_moff.size())) != hipSuccess ||
      (e = b->dmnbf.assign(b->mol_nbf.data(), b->mol_nbf.size())) != hipSuccess ||
      (e = b->dshloc.assign(b->sh_loc.data(), b->sh_loc.size())) != hipSuccess ||
      (e = b->ddmoff.assign(b->mol_dmoff.data(), b->mol_dmoff.size())) != hipSuccess ||
      (e = b->ddmns.assign(dmns.data(), dmns.size())) != hipSuccess ||
      (e = b->Jt.alloc(size_t(b->mat_size))) != hipSuccess ||
      (e = b->Kt.alloc(size_t(b->mat_size))) != hipSuccess) {
    lai_dev_basis_free(b);
    return fail(LAI_ERR_MEMORY, std::string("lai_dev_upload_basis: ") + hipGetErrorString(e));
  }
  *out = b;
  return LAI_OK;
}

extern "C" int lai_dev_basis_free(lai_basis* b) {
  if (!b) return LAI_OK;
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int i = 0; i < kSideStreams; ++i) {
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  b->pairs.release();
  b->pp.release();
  b->ppab.release();
  b->ppb.release();
  b->ppih.release();
  b->shell_atom.release();
  b->dL.release();
  b->dnprim.release();
  b->dprim_off.release();
  b->dfunc_off.release();
  b->dcenter.release();
  b->dexps.release();
  b->dcoefs.release();
  b->Jt.release();
  b->Kt.release();
  for (auto& c : b->cls) {
    c.chunks.release();
    c.gchunks.release();
  }
  b->segs.release();
  b->kets.release();
  b->sdmax.release();
  b->dQ.release();
  b->dmoff.release();
  b->dmnbf.release();
  b->dshloc.release();
  b->ddmoff.release();
  b->ddmns.release();
  delete b;
  return LAI_OK;
}

extern "C" int lai_dev_basis_nbf(const lai_basis* b) { return b ? b->nbf : -1; }

extern "C" int lai_dev_batch_layout(const lai_basis* b, int* n_mol, int* nbf, int64_t* mat_off) {
  if (!b) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_batch_layout: null basis");
  if (n_mol) *n_mol = b->n_mol;
  for (int m = 0; m < b->n_mol; ++m) {
    if (nbf) nbf[m] = b->mol_nbf[m];
    if (mat_off) mat_off[m] = b->mol_moff[m];
  }
  if (mat_off) mat_off[b->n_mol] = b->mol_moff[b->n_mol];
  return LAI_OK;
}

namespace lai {
int set_error(int code, const std::string& msg) { return fail(code, msg); }

BasisView basis_view(const lai_basis* b) {
  return BasisView{b->ctx, b->ns, b->nbf, b->n_mol, b->L.data(), b->nprim.data(), b->prim_off.data(),
                   b->func_off.data(), b->exps.data(), b->coefs.data()};
}
}  // namespace lai

// ---------------------------------------------------------------------------
// raw ERI blocks
// ---------------------------------------------------------------------------
namespace {
// Orders a shell quartet into class order and returns the class id, the
// (bra, ket) pair indices and the permutation code (class position p ->
// original position, 2 bits each).
// Returns -1 when a pair's two shells belong to different molecules.
int classify(const lai_basis* b, const int s[4], Quartet& q, int& perm) {
  auto pidx = [&](int x, int y) -> int {
    const int m = b->sh_mol[x];
    if (b->sh_mol[y] != m) return -1;
    const int lx = b->sh_loc[x], ly = b->sh_loc[y];
    return lx <= ly ? mol_pair_index(b, m, lx, ly) : mol_pair_index(b, m, ly, lx);
  };
  const int p0 = pidx(s[0], s[1]), p1 = pidx(s[2], s[3]);
  if (p0 < 0 || p1 < 0) return -1;
  const int c0 = b->pair_cls[p0], c1 = b->pair_cls[p1];
  const bool swap_bk = c0 < c1;
  const int pb = swap_bk ? p1 : p0, pk = swap_bk ? p0 : p1;
  const int ob = swap_bk ? 2 : 0, ok = swap_bk ? 0 : 2;  // original positions of bra/ket pairs
  const PairRec& rb = b->pairs_h[pb];
  const PairRec& rk = b->pairs_h[pk];
  // class position 0 = rb.sa: original position ob if s[ob] == rb.sa else ob+1
  const int o0 = (s[ob] == rb.sa) ? ob : ob + 1;
  const int o1 = (o0 == ob) ? ob + 1 : ob;
  const int o2 = (s[ok] == rk.sa) ? ok : ok + 1;
  const int o3 = (o2 == ok) ? ok + 1 : ok;
  perm = o0 | (o1 << 2) | (o2 << 4) | (o3 << 6);
  q.bra = pb;
  q.ket = pk;
  const int cb = std::max(c0, c1), ck = std::min(c0, c1);
  return cb * (cb + 1) / 2 + ck;
}

struct EriBatch {
  std::vector<Quartet> q[kNumClasses];
  std::vector<int> perm[kNumClasses];
  std::vector<long long> obase[kNumClasses];
};

// Runs mode-1 kernels for a classified batch writing into d_out (device).
int run_eri_batch(lai_basis* b, EriBatch& eb, double* d_out, hipStream_t st) {
  const auto& Ls = launchers();
  for (int c = 0; c < kNumClasses; ++c) {
    const int nq = int(eb.q[c].size());
    if (nq == 0) continue;
    DevBuf<Quartet> dq;
    DevBuf<int> dp;
    DevBuf<long long> dob;
    hipError_t e;
    if ((e = dq.assign(eb.q[c].data(), nq)) != hipSuccess || (e = dp.assign(eb.perm[c].data(), nq)) != hipSuccess ||
        (e = dob.assign(eb.obase[c].data(), nq)) != hipSuccess)
      return fail(LAI_ERR_MEMORY, std::string("eri batch: ") + hipGetErrorString(e));
    EriArgs a{};
    a.q = dq.p;
    a.nq = nq;
    a.pairs = b->pairs.p;
    a.pp = b->pp.p;
    a.rys = b->ctx->rys.p;
    a.ppb = b->ppb.p;
    a.ppih = b->ppih.p;
    a.pcut32 = b->pcut32;
    a.mode = 1;
    a.out = d_out;
    a.perm = dp.p;
    a.obase = dob.p;
    a.nbf = b->max_nbf;
    LaunchTune t;
    t.n_cu = b->ctx->n_cu;
    Ls.f64[c](a, st, t);
    LAI_HIP(hipGetLastError());
    LAI_HIP(hipStreamSynchronize(st));
  }
  return LAI_OK;
}
}  // namespace

namespace {
// Sorts n shell quartets into the class batches and fills offsets[n + 1] (the
// row-major block of quartet t starts at offsets[t]).
int classify_batch(lai_basis* b, int64_t n, const int* s4, int64_t* offsets, EriBatch& eb, const char* what) {
  int64_t off = 0;
  for (int64_t t = 0; t < n; ++t) {
    const int* s = s4 + 4 * t;
    for (int k = 0; k < 4; ++k)
      if (s[k] < 0 || s[k] >= b->ns) return fail(LAI_ERR_INVALID_ARGUMENT, std::string(what) + ": shell index out of range");
    Quartet q;
    int perm;
    const int c = classify(b, s, q, perm);
    if (c < 0)
      return fail(LAI_ERR_INVALID_ARGUMENT, std::string(what) + ": the shells of a pair must belong to one molecule");
    offsets[t] = off;
    eb.q[c].push_back(q);
    eb.perm[c].push_back(perm);
    eb.obase[c].push_back(off);
    off += int64_t(ncart(b->L[s[0]])) * ncart(b->L[s[1]]) * ncart(b->L[s[2]]) * ncart(b->L[s[3]]);
  }
  offsets[n] = off;
  return LAI_OK;
}
}  // namespace

extern "C" int lai_dev_eri_quartets(lai_basis* b, int64_t n, const int* s4, double* out, int64_t* offsets) {
  if (!b || n < 0 || (n > 0 && (!s4 || !out || !offsets)))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_eri_quartets: invalid arguments");
  EriBatch eb;
  int rc = classify_batch(b, n, s4, offsets, eb, "lai_dev_eri_quartets");
  if (rc != LAI_OK) return rc;
  const int64_t off = offsets[n];
  DevBuf<double> d;
  hipError_t e = d.alloc(size_t(off));
  if (e != hipSuccess) return fail(LAI_ERR_MEMORY, "lai_dev_eri_quartets: output allocation");
  rc = run_eri_batch(b, eb, d.p, nullptr);
  if (rc != LAI_OK) return rc;
  LAI_HIP(hipMemcpy(out, d.p, sizeof(double) * off, hipMemcpyDeviceToHost));
  d.release();
  return LAI_OK;
}

extern "C" int lai_dev_eri_quartets_device(lai_basis* b, int64_t n, const int* s4, double* d_out, int64_t* offsets,
                                           void* stream) {
  if (!b || n < 0 || !offsets || (n > 0 && !s4))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_eri_quartets_device: invalid arguments");
  EriBatch eb;
  int rc = classify_batch(b, n, s4, offsets, eb, "lai_dev_eri_quartets_device");
  if (rc != LAI_OK || !d_out) return rc;  // no output: offsets only
  rc = run_eri_batch(b, eb, d_out, hipStream_t(stream));
  if (rc != LAI_OK) return rc;
  LAI_HIP(hipStreamSynchronize(hipStream_t(stream)));
  return LAI_OK;
}

extern "C" int lai_dev_eri_quartets_device_noalloc(lai_basis* b, int64_t n, const int* s4, double* d_out,
                                                   int64_t* offsets, void* stream) {
  if (!b || n < 0 || !offsets || (n > 0 && !s4))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_eri_quartets_device_noalloc: invalid arguments");
  if (n > LAI_NOALLOC_MAX_QUARTETS)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_eri_quartets_device_noalloc: " + std::to_string(n) +
                                              " quartets, more than the staging limit of " +
                                              std::to_string(LAI_NOALLOC_MAX_QUARTETS));
  EriBatch eb;
  int rc = classify_batch(b, n, s4, offsets, eb, "lai_dev_eri_quartets_device_noalloc");
  if (rc != LAI_OK || !d_out || n == 0) return rc;  // no output: offsets only
  const int dev = b->ctx->device;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev])
    return fail(LAI_ERR_INVALID_STATE, "lai_dev_eri_quartets_device_noalloc: staging array not set up");
  unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
  // layout: per class with work, its Quartet, perm and obase arrays, each
  // 256-byte aligned; at most 20 bytes per quartet plus 3 * 256 per class
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t need = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    const size_t nq = eb.q[c].size();
    if (nq) need += up(nq * sizeof(Quartet)) + up(nq * sizeof(int)) + up(nq * sizeof(long long));
  }
  if (need > kScratchBytes)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_eri_quartets_device_noalloc: staging array too small");
  hipStream_t st = hipStream_t(stream);
  std::lock_guard<std::mutex> lk(g_scratch_mtx);  // one staging call at a time in this process
  const auto& Ls = launchers();
  size_t off = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    const int nq = int(eb.q[c].size());
    if (nq == 0) continue;
    auto* dq = reinterpret_cast<Quartet*>(base + off);
    off += up(size_t(nq) * sizeof(Quartet));
    auto* dp = reinterpret_cast<int*>(base + off);
    off += up(size_t(nq) * sizeof(int));
    auto* dob = reinterpret_cast<long long*>(base + off);
    off += up(size_t(nq) * sizeof(long long));
    LAI_HIP(hipMemcpyAsync(dq, eb.q[c].data(), sizeof(Quartet) * nq, hipMemcpyHostToDevice, st));
    LAI_HIP(hipMemcpyAsync(dp, eb.perm[c].data(), sizeof(int) * nq, hipMemcpyHostToDevice, st));
    LAI_HIP(hipMemcpyAsync(dob, eb.obase[c].data(), sizeof(long long) * nq, hipMemcpyHostToDevice, st));
    EriArgs a{};
    a.q = dq;
    a.nq = nq;
    a.pairs = b->pairs.p;
    a.pp = b->pp.p;
    a.rys = b->ctx->rys.p;
    a.ppb = b->ppb.p;
    a.ppih = b->ppih.p;
    a.pcut32 = b->pcut32;
    a.mode = 1;
    a.out = d_out;
    a.perm = dp;
    a.obase = dob;
    a.nbf = b->max_nbf;
    LaunchTune t;
    t.n_cu = b->ctx->n_cu;
    Ls.f64[c](a, st, t);
    LAI_HIP(hipGetLastError());
  }
  // the host descriptor arrays are released on return: the copies must be done
  LAI_HIP(hipStreamSynchronize(st));
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// Schwarz
// ---------------------------------------------------------------------------
extern "C" int lai_dev_schwarz(lai_basis* b, double* Qout) {
  if (!b) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_schwarz: null basis");
  const int npairs = b->mol_pair.back();
  EriBatch eb;
  std::vector<long long> offs(npairs);
  std::vector<int> ni(npairs), nj(npairs);
  long long off = 0;
  for (int p = 0; p < npairs; ++p) {
    const int i = b->pi[p], j = b->pj[p];
    const int s[4] = {i, j, i, j};
    Quartet q;
    int perm;
    const int c = classify(b, s, q, perm);
    eb.q[c].push_back(q);
    eb.perm[c].push_back(perm);
    eb.obase[c].push_back(off);
    offs[p] = off;
    ni[p] = ncart(b->L[i]);
    nj[p] = ncart(b->L[j]);
    off += (long long)(ni[p] * nj[p]) * (ni[p] * nj[p]);
  }
  DevBuf<double> d, dQ;
  DevBuf<long long> doff;
  DevBuf<int> dni, dnj;
  hipError_t e;
  if ((e = d.alloc(size_t(off))) != hipSuccess || (e = dQ.alloc(npairs)) != hipSuccess ||
      (e = doff.assign(offs.data(), npairs)) != hipSuccess || (e = dni.assign(ni.data(), npairs)) != hipSuccess ||
      (e = dnj.assign(nj.data(), npairs)) != hipSuccess)
    return fail(LAI_ERR_MEMORY, std::string("lai_dev_schwarz: ") + hipGetErrorString(e));
  int rc = run_eri_batch(b, eb, d.p, nullptr);
  if (rc != LAI_OK) return rc;
  hipLaunchKernelGGL(schwarz_reduce_kernel, dim3((npairs + 127) / 128), dim3(128), 0, nullptr, d.p, doff.p,
                     dni.p, dnj.p, dQ.p, npairs);
  LAI_HIP(hipGetLastError());
  std::vector<double> Qp(npairs);
  LAI_HIP(hipMemcpy(Qp.data(), dQ.p, sizeof(double) * npairs, hipMemcpyDeviceToHost));
  // reference packing j*(j+1)/2 + i, molecule by molecule
  b->Q.assign(npairs, 0.0);
  for (int p = 0; p < npairs; ++p) b->Q[packed_q_index(b, p)] = Qp[p];
  b->have_Q = true;
  if (Qout) std::memcpy(Qout, b->Q.data(), sizeof(double) * npairs);
  return LAI_OK;
}

extern "C" int lai_dev_set_schwarz(lai_basis* b, const double* Q) {
  if (!b || !Q) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_schwarz: invalid arguments");
  b->Q.assign(Q, Q + size_t(b->mol_pair.back()));
  b->have_Q = true;
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// worklist
// ---------------------------------------------------------------------------
extern "C" uint64_t lai_quartet_mix64(int i, int j, int k, int l) {
  const uint64_t key = (uint64_t(uint16_t(i)) << 48) | (uint64_t(uint16_t(j)) << 32) |
                       (uint64_t(uint16_t(k)) << 16) | uint64_t(uint16_t(l));
  return splitmix64(key);
}

namespace {
// Chunk cap of the default rule for a class run in shape `shape` with n
// quartets: lane-per-quartet kernels take up to 2048 quartets (a workgroup of
// nt to 512 lanes; the launch replicates the nt-lane unit), wavefront-per-
// quartet kernels up to 64, workgroup-per-quartet kernels up to 32; small
// classes get shorter chunks so that a launch still has a few workgroups per
// CU.
int default_chunk(int shape, size_t n, int n_cu, int nt = 256) {
  const size_t q = n / (4 * size_t(n_cu));
  if (shape == kShapeRegs) {
    const size_t lo = size_t(nt), hi = 2048;
    return int(std::clamp<size_t>(q * size_t(nt) / 256, lo, hi));
  }
  if (shape == kShapeWaves) return int(std::clamp<size_t>(q, 4, 64));
  return int(std::clamp<size_t>(q, 1, 32));
}

// Cuts each run into near-equal chunks of at most cmax quartets.
std::vector<JkChunk> cut_runs(const std::vector<JkChunk>& runs, int cmax) {
  std::vector<JkChunk> C;
  for (const auto& r : runs) {
    const int len = r.k1;
    const int pieces = (len + cmax - 1) / cmax;
    const int per = (len + pieces - 1) / pieces;
    for (int k0 = 0; k0 < len; k0 += per) C.push_back(JkChunk{r.bra, r.seg, k0, std::min(len, k0 + per)});
  }
  return C;
}

int class_shape(const ClassInfo& info, const ClassTune& t) {
  return (t.alt && info.alt != kShapeNone) ? info.alt : info.shape;
}

// Whether a mixed build runs class c's FP32 quartets through the float
// instantiation (a launch of its own).  Only the wavefront and workgroup
// shapes: there the float tables halve the LDS traffic of the products and the
// recursion is measured faster.  The lane-per-quartet classes are bound by
// memory latency, their float recursion is measured slower, so a mixed build
// runs their FP32-marked quartets in FP64, in the class's one launch.
bool class_split(const lai_basis* b, int c, bool mixed) {
  return mixed && class_shape(launchers().info[c], b->tune[c]) != kShapeRegs;
}

// Chunks of class c under its launch choice (and, when that is not the
// default, the default chunks for the gradient kernels) on the device.
int upload_chunks(lai_basis* b, int c) {
  const auto& Ls = launchers();
  ClassList& C = b->cls[c];
  size_t n = 0;
  for (const auto& r : C.runs) n += size_t(r.k1);
  const ClassTune& t = b->tune[c];
  const int shape = class_shape(Ls.info[c], t);
  const int cmax = t.chunk > 0 ? t.chunk : default_chunk(shape, n, b->ctx->n_cu, Ls.info[c].nt);
  const std::vector<JkChunk> ch = cut_runs(C.runs, cmax);
  hipError_t e;
  if ((e = C.chunks.assign(ch.data(), ch.size())) != hipSuccess)
    return fail(LAI_ERR_MEMORY, std::string("chunk upload: ") + hipGetErrorString(e));
  C.nchunks = int64_t(ch.size());
  C.gchunks.release();
  C.ngchunks = 0;
  if (!t.is_default()) {
    const std::vector<JkChunk> g = cut_runs(C.runs, default_chunk(Ls.info[c].shape, n, b->ctx->n_cu, Ls.info[c].nt));
    if ((e = C.gchunks.assign(g.data(), g.size())) != hipSuccess)
      return fail(LAI_ERR_MEMORY, std::string("chunk upload: ") + hipGetErrorString(e));
    C.ngchunks = int64_t(g.size());
  }
  return LAI_OK;
}
}  // namespace

extern "C" int lai_dev_build_worklist(lai_basis* b, double thr, const double* shell_dmax, int density_screen,
                                      double tau, int64_t* n_out, int64_t* n32_out, uint64_t* checksum) {
  if (!b) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_build_worklist: null basis");
  const bool screen = thr > 0.0;
  if (screen && !b->have_Q)
    return fail(LAI_ERR_INVALID_STATE, "lai_dev_build_worklist: Schwarz bounds not computed");
  if (screen && density_screen && !shell_dmax)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_build_worklist: density screening needs shell_dmax");
  const int npairs = b->mol_pair.back();
  std::vector<double> Qp(npairs, 1.0);
  if (b->have_Q)
    for (int p = 0; p < npairs; ++p) Qp[p] = b->Q[packed_q_index(b, p)];
  const bool mixed = tau > 0.0 && b->have_Q && shell_dmax != nullptr;
  const bool dscreen = screen && density_screen && shell_dmax != nullptr;
  const auto& Ls = launchers();
  const int nm = b->n_mol;
  constexpr int NPC = kNumPairClasses, NB = kMaxSegs;

  // Ket lists: per (molecule, pair class, primitive-count bucket), the pairs in
  // order of decreasing Schwarz bound.  Bucket k holds npp_sig in [2^k, 2^(k+1))
  // (the last one everything above), so the lanes of a wavefront run similar
  // primitive-loop trip counts.
  auto bucket = [&](int p) {
    const int n = b->pairs_h[p].npp_sig;
    int k = 0;
    while (k < NB - 1 && (2 << k) <= n) ++k;
    return k;
  };
  const size_t nlists = size_t(nm) * NPC * NB;
  std::vector<std::vector<int>> lists(nlists);
  std::vector<int> pbk(npairs), ppos(npairs);
  for (int p = 0; p < npairs; ++p) {
    pbk[p] = bucket(p);
    const int m = b->sh_mol[b->pi[p]];
    lists[(size_t(m) * NPC + b->pair_cls[p]) * NB + pbk[p]].push_back(p);
  }
  std::vector<int> loff(nlists + 1, 0);
  for (size_t l = 0; l < nlists; ++l) {
    auto& v = lists[l];
    std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return Qp[x] > Qp[y]; });
    for (int k = 0; k < int(v.size()); ++k) ppos[v[k]] = k;
    loff[l + 1] = loff[l] + int(v.size());
  }
  std::vector<int> kets_h(npairs);
  for (size_t l = 0; l < nlists; ++l) std::copy(lists[l].begin(), lists[l].end(), kets_h.begin() + loff[l]);

  // Per quartet class: for every bra pair (heaviest primitive bucket first,
  // then decreasing bound) the segments of kets that pass the Schwarz test.
  // Pairs of the same class count once: with the order (bucket, position), a
  // ket is taken when it does not precede the bra.
  struct Entry {
    int bra;
    JkSegs sg;
  };
  std::vector<std::vector<Entry>> E(kNumClasses);
  std::vector<int64_t> etot(kNumClasses, 0);
  parallel_for(kNumClasses, [&](int c, int) {
    const int pb = cls_bra(c), pk = cls_ket(c);
    int64_t tot = 0;
    for (int m = 0; m < nm; ++m)
      for (int bb = NB - 1; bb >= 0; --bb) {
        const size_t lb = (size_t(m) * NPC + pb) * NB + bb;
        for (int p : lists[lb]) {
          Entry e{};
          e.bra = p;
          int acc = 0;
          for (int kb = NB - 1; kb >= 0; --kb) {
            const size_t lk = (size_t(m) * NPC + pk) * NB + kb;
            const int len = loff[lk + 1] - loff[lk];
            if (len == 0) continue;
            int lo = 0;
            if (pb == pk) {
              if (kb < bb) continue;
              if (kb == bb) lo = ppos[p];
            }
            int hi = len;
            if (screen) {  // first position where Q_p Q_r >= thr fails
              const int* v = kets_h.data() + loff[lk];
              int a = 0, z = len;
              const double qp = Qp[p];
              while (a < z) {
                const int mid = (a + z) / 2;
                if (qp * Qp[v[mid]] >= thr)
                  a = mid + 1;
                else
                  z = mid;
              }
              hi = a;
            }
            if (hi <= lo) continue;
            e.sg.start[e.sg.n] = loff[lk] + lo;
            e.sg.cum[e.sg.n] = acc;
            acc += hi - lo;
            ++e.sg.n;
          }
          if (acc == 0) continue;
          e.sg.cum[e.sg.n] = acc;
          tot += acc;
          E[c].push_back(e);
        }
      }
    etot[c] = tot;
  }, 1);

  // Chunks: one run per (bra, segment list), cut by the class's chunk rule
  std::vector<JkSegs> segs_h;
  for (int c = 0; c < kNumClasses; ++c) {
    auto& R = b->cls[c].runs;
    R.clear();
    for (const auto& e : E[c]) {
      R.push_back(JkChunk{e.bra, int(segs_h.size()), 0, e.sg.cum[e.sg.n]});
      segs_h.push_back(e.sg);
    }
    std::vector<Entry>().swap(E[c]);
  }
  // shell-pair max|D| per molecule (density screen, FP32 split)
  std::vector<double> sdm;
  if (shell_dmax) sdm.assign(shell_dmax, shell_dmax + b->mol_dmoff.back());
  else sdm.assign(1, 0.0);
  hipError_t e;
  if ((e = b->kets.assign(kets_h.data(), kets_h.size())) != hipSuccess ||
      (e = b->segs.assign(segs_h.data(), segs_h.size())) != hipSuccess ||
      (e = b->dQ.assign(Qp.data(), Qp.size())) != hipSuccess ||
      (e = b->sdmax.assign(sdm.data(), sdm.size())) != hipSuccess)
    return fail(LAI_ERR_MEMORY, std::string("lai_dev_build_worklist: ") + hipGetErrorString(e));
  for (int c = 0; c < kNumClasses; ++c) {
    const int rc = upload_chunks(b, c);
    if (rc != LAI_OK) return rc;
  }
  b->wl_dscreen = dscreen ? 1 : 0;
  b->wl_mixed = mixed ? 1 : 0;
  b->wl_thr = thr;
  b->wl_tau = tau;

  // statistics on the device, with the kernels' own selection
  double fp_tab[4 * 4 * 4 * 4];
  for (int x = 0; x <= kMaxL; ++x)
    for (int y = 0; y <= kMaxL; ++y)
      for (int z = 0; z <= kMaxL; ++z)
        for (int w = 0; w <= kMaxL; ++w) fp_tab[((x * 4 + y) * 4 + z) * 4 + w] = fprim(x, y, z, w);
  double fg_tab[2 * 256];
  for (int m = 0; m < 2; ++m)
    for (int x = 0; x <= kMaxL; ++x)
      for (int y = 0; y <= kMaxL; ++y)
        for (int z = 0; z <= kMaxL; ++z)
          for (int w = 0; w <= kMaxL; ++w) fg_tab[256 * m + ((x * 4 + y) * 4 + z) * 4 + w] = fgrad(m, x, y, z, w);
  DevBuf<double> dfp, dfl, dfg, dgf;
  DevBuf<unsigned long long> dcnt, dsum;
  if ((e = dfp.assign(fp_tab, 256)) != hipSuccess || (e = dfl.alloc(4 * kNumClasses)) != hipSuccess ||
      (e = dfg.assign(fg_tab, 512)) != hipSuccess || (e = dgf.alloc(2 * kNumClasses)) != hipSuccess ||
      (e = dcnt.alloc(3 * kNumClasses)) != hipSuccess || (e = dsum.alloc(1)) != hipSuccess)
    return fail(LAI_ERR_MEMORY, std::string("lai_dev_build_worklist: ") + hipGetErrorString(e));
  LAI_HIP(hipMemset(dfl.p, 0, sizeof(double) * 4 * kNumClasses));
  LAI_HIP(hipMemset(dgf.p, 0, sizeof(double) * 2 * kNumClasses));
  LAI_HIP(hipMemset(dcnt.p, 0, sizeof(unsigned long long) * 3 * kNumClasses));
  LAI_HIP(hipMemset(dsum.p, 0, sizeof(unsigned long long)));
  WlStatsArgs sa{};
  sa.segs = b->segs.p;
  sa.kets = b->kets.p;
  sa.pairs = b->pairs.p;
  sa.sel = make_sel(b);
  sa.L = b->dL.p;
  sa.fp = dfp.p;
  sa.counts = dcnt.p;
  sa.flops = dfl.p;
  sa.checksum = dsum.p;
  sa.fg = dfg.p;
  sa.gflops = dgf.p;
  for (int c = 0; c < kNumClasses; ++c) {
    if (b->cls[c].nchunks == 0) continue;
    sa.cls = c;
    sa.chunks = b->cls[c].chunks.p;
    sa.nchunks = int(b->cls[c].nchunks);
    const int blocks = int(std::min<int64_t>(b->cls[c].nchunks, int64_t(b->ctx->n_cu) * 16));
    hipLaunchKernelGGL(worklist_stats_kernel, dim3(blocks), dim3(256), 0, nullptr, sa);
  }
  LAI_HIP(hipGetLastError());
  std::vector<unsigned long long> cnt(3 * kNumClasses);
  std::vector<double> fl(4 * kNumClasses);
  unsigned long long sum = 0;
  LAI_HIP(hipMemcpy(cnt.data(), dcnt.p, sizeof(unsigned long long) * cnt.size(), hipMemcpyDeviceToHost));
  LAI_HIP(hipMemcpy(fl.data(), dfl.p, sizeof(double) * fl.size(), hipMemcpyDeviceToHost));
  LAI_HIP(hipMemcpy(&sum, dsum.p, sizeof(sum), hipMemcpyDeviceToHost));
  LAI_HIP(hipMemcpy(b->cls_grad, dgf.p, sizeof(double) * 2 * kNumClasses, hipMemcpyDeviceToHost));
  int64_t total = 0, total32 = 0;
  b->flops64 = b->flops32 = 0.0;
  b->n_eri = 0;
  for (int c = 0; c < kNumClasses; ++c) {
    b->counts[c] = int64_t(cnt[3 * c]);
    total += int64_t(cnt[3 * c]);
    total32 += int64_t(cnt[3 * c + 1]);
    b->n_eri += int64_t(cnt[3 * c + 2]);
    b->flops64 += fl[4 * c];
    b->flops32 += fl[4 * c + 1];
    b->cls_f64[c] = fl[4 * c + 2];
    b->cls_f32[c] = fl[4 * c + 3];
    b->cls_n32[c] = int64_t(cnt[3 * c + 1]);
  }
  b->n_fp32 = total32;
  b->checksum = sum;
  b->have_worklist = true;
  b->pitems_valid = false;
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  if (n_out) *n_out = total;
  if (n32_out) *n32_out = total32;
  if (checksum) *checksum = sum;
  return LAI_OK;
}

extern "C" int lai_dev_worklist_stats(const lai_basis* b, int64_t* counts, double* f64, double* f32,
                                      int64_t* neri) {
  if (!b) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_worklist_stats: null basis");
  if (!b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_worklist_stats: no worklist");
  if (counts) std::memcpy(counts, b->counts, sizeof(b->counts));
  if (f64) *f64 = b->flops64;
  if (f32) *f32 = b->flops32;
  if (neri) *neri = b->n_eri;
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// J/K build
// ---------------------------------------------------------------------------
// Side streams and fork/join events of a basis handle (before any capture).
static int ensure_side_streams(lai_basis* b) {
  if (b->fork) return LAI_OK;
  for (int i = 0; i < kSideStreams; ++i) {
    LAI_HIP(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
    LAI_HIP(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
  }
  LAI_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  return LAI_OK;
}

// The persistent launch's queue: the default chunks of every covered class
// with work, class by class in launch order (heaviest first).
static int build_persist_queue(lai_basis* b) {
  if (b->pitems_valid) return LAI_OK;
  const auto& Ls = launchers();
  std::vector<PersistItem> items;
  for (int c = kNumClasses - 1; c >= 0; --c) {
    const ClassList& C = b->cls[c];
    if (!persist_covers(c) || C.runs.empty()) continue;
    size_t n = 0;
    for (const auto& r : C.runs) n += size_t(r.k1);
    for (const JkChunk& ch : cut_runs(C.runs, default_chunk(kShapeRegs, n, b->ctx->n_cu, Ls.info[c].nt)))
      items.push_back(PersistItem{c, ch});
  }
  hipError_t e;
  if ((e = b->pitems.assign(items.data(), items.size())) != hipSuccess ||
      (!b->pcounter.p && (e = b->pcounter.alloc(1)) != hipSuccess))
    return fail(LAI_ERR_MEMORY, std::string("persistent queue: ") + hipGetErrorString(e));
  b->n_pitems = int(items.size());
  b->pitems_valid = true;
  return LAI_OK;
}

// Class kernels over the bra-anchored chunks, heavy classes first.  A mixed
// build runs each class as two launches, the FP64 instantiation over the
// quartets selected for FP64 (QuartetSel::keep = 1) and the float one over
// those selected for FP32 (keep = 2), so that each has its own register
// budget; a part with no quartets is not launched (the worklist statistics
// count them with the same selection).  With ev (4 events per class) a
// hipEvent pair brackets each of the two parts, one after another on st.
// Without: the parts are spread over st and the side streams (each to the
// stream with the least work so far, by executed flops), forked from and
// joined back into st, so that the last workgroups of one class overlap the
// next classes (every class adds into Jt and Kt with atomics only, so the
// order is free).
static int launch_jk_classes(lai_basis* b, EriArgs a, int flags, hipStream_t st, hipEvent_t* ev) {
  const auto& Ls = launchers();
  const bool mixed = (flags & LAI_JK_MIXED) != 0 && b->wl_mixed;
  const bool persist = (flags & LAI_JK_PERSISTENT) != 0 && (flags & LAI_JK_MIXED) == 0;
  a.segs = b->segs.p;
  a.kets = b->kets.p;
  a.sel = make_sel(b);
  const int sel_mixed = a.sel.mixed;
  const bool fan = ev == nullptr && b->fork != nullptr;
  hipStream_t streams[kSideStreams + 1] = {st};
  double load[kSideStreams + 1] = {};
  if (fan) {
    LAI_HIP(hipEventRecord(b->fork, st));
    for (int i = 0; i < kSideStreams; ++i) {
      streams[i + 1] = b->side[i];
      LAI_HIP(hipStreamWaitEvent(b->side[i], b->fork, 0));
    }
  }
  if (persist) {  // the covered classes: one launch on st, first
    const int rc = build_persist_queue(b);
    if (rc != LAI_OK) return rc;
    a.sel.mixed = 0;
    for (int c = 0; c < kNumClasses; ++c)
      if (persist_covers(c)) load[0] += b->cls_f64[c] + b->cls_f32[c] + 1e6 * double(b->cls[c].nchunks);
    launch_persistent(a, b->pitems.p, b->n_pitems, b->pcounter.p, st, b->ctx->n_cu);
  }
  for (int c = kNumClasses - 1; c >= 0; --c) {
    const ClassList& C = b->cls[c];
    if (persist && persist_covers(c)) continue;  // in the persistent launch
    const bool split = class_split(b, c, mixed);
    a.sel.mixed = split ? sel_mixed : 0;
    for (int part = 0; part < 2; ++part) {  // 0: FP64, 1: FP32 (split classes only)
      hipEvent_t* e = ev ? ev + 4 * c + 2 * part : nullptr;
      if (e) (void)hipEventRecord(e[0], st);
      const int64_t n32 = split ? b->cls_n32[c] : 0;
      const bool run = C.nchunks > 0 && (part == 0 ? b->counts[c] - n32 > 0 : n32 > 0);
      if (run) {
        hipStream_t s = st;
        if (fan) {
          int k = 0;
          for (int i = 1; i <= kSideStreams; ++i)
            if (load[i] < load[k]) k = i;
          const double fl = part == 1 ? b->cls_f32[c] : split ? b->cls_f64[c] : b->cls_f64[c] + b->cls_f32[c];
          load[k] += fl + 1e6 * double(C.nchunks);
          s = streams[k];
        }
        a.chunks = C.chunks.p;
        a.nchunks = int(C.nchunks);
        a.sel.keep = split ? part + 1 : 0;
        LaunchTune t;
        t.alt = b->tune[c].alt;
        t.blocks_per_cu = b->tune[c].bpc;
        t.n_cu = b->ctx->n_cu;
        (part == 1 ? Ls.f32[c] : Ls.f64[c])(a, s, t);
      }
      if (e) (void)hipEventRecord(e[1], st);
      if (!split) {
        if (e) {  // the FP32 part is empty: a zero-length pair
          (void)hipEventRecord(e[2], st);
          (void)hipEventRecord(e[3], st);
        }
        break;
      }
    }
  }
  if (fan)
    for (int i = 0; i < kSideStreams; ++i) {
      LAI_HIP(hipEventRecord(b->join[i], b->side[i]));
      LAI_HIP(hipStreamWaitEvent(st, b->join[i], 0));
    }
  return LAI_OK;
}

// arguments of a J/K build into the scratch Jt, Kt
static EriArgs jk_args(lai_basis* b, const double* d_D) {
  EriArgs a{};
  a.pairs = b->pairs.p;
  a.pp = b->pp.p;
  a.rys = b->ctx->rys.p;
  a.ppb = b->ppb.p;
  a.ppih = b->ppih.p;
  a.pcut32 = b->pcut32;
  a.D = d_D;
  a.Jt = b->Jt.p;
  a.Kt = b->Kt.p;
  a.nbf = b->max_nbf;
  a.mode = 0;
  return a;
}

// J += (Jt + Jt^T)/2, K += (Kt + Kt^T)/2; with anti (antisymmetric D) only
// K += (Kt - Kt^T)/2
static int symmetrize(lai_basis* b, double* d_J, double* d_K, hipStream_t st, bool anti = false) {
  const size_t nn = size_t(b->max_nbf) * b->max_nbf;
  const unsigned blocks = unsigned(std::max<size_t>(1, std::min<size_t>((nn + 255) / 256, 2048)));
  hipLaunchKernelGGL(symmetrize_add_batch_kernel, dim3(blocks, unsigned(b->n_mol)), dim3(256), 0, st, b->Jt.p,
                     b->Kt.p, d_J, d_K, b->dmoff.p, b->dmnbf.p, anti ? 1 : 0);
  LAI_HIP(hipGetLastError());
  return LAI_OK;
}

extern "C" int lai_dev_jk(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags, void* stream) {
  if (!b || !d_D || !d_J || !d_K) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk: invalid arguments");
  if (!b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_jk: worklist not built");
  if ((flags & ~(LAI_JK_MIXED | LAI_JK_ANTISYMMETRIC | LAI_JK_PERSISTENT)) != 0)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk: unknown flags");
  hipStream_t st = hipStream_t(stream);
  const size_t nn = size_t(b->mat_size);
  int rc;
  if (!b->fork) {  // side streams are created outside any stream capture
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LAI_HIP(hipStreamIsCapturing(st, &cs));
    if (cs == hipStreamCaptureStatusNone && (rc = ensure_side_streams(b)) != LAI_OK) return rc;
  }
  if ((rc = lai_dev_zero(b->Jt.p, nn, stream)) != LAI_OK) return rc;
  if ((rc = lai_dev_zero(b->Kt.p, nn, stream)) != LAI_OK) return rc;
  if ((rc = launch_jk_classes(b, jk_args(b, d_D), flags, st, nullptr)) != LAI_OK) return rc;
  LAI_HIP(hipGetLastError());
  return symmetrize(b, d_J, d_K, st, (flags & LAI_JK_ANTISYMMETRIC) != 0);
}

extern "C" int lai_dev_jk_graph_capture(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags,
                                        void* stream) {
  if (!b) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk_graph_capture: null basis");
  if (!stream) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk_graph_capture: capture needs a non-default stream");
  hipStream_t st = hipStream_t(stream);
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  int rc = ensure_side_streams(b);  // not inside the capture
  if (rc != LAI_OK) return rc;
  if ((flags & LAI_JK_PERSISTENT) && b->have_worklist && (rc = build_persist_queue(b)) != LAI_OK) return rc;
  LAI_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rc = lai_dev_jk(b, d_D, d_J, d_K, flags, stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != LAI_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(LAI_ERR_BACKEND, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  LAI_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  return LAI_OK;
}

extern "C" int lai_dev_jk_graph_launch(lai_basis* b, void* stream) {
  if (!b || !b->graph_exec) return fail(LAI_ERR_INVALID_STATE, "lai_dev_jk_graph_launch: no captured graph");
  LAI_HIP(hipGraphLaunch(b->graph_exec, hipStream_t(stream)));
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// one-electron integrals
// ---------------------------------------------------------------------------
extern "C" int lai_dev_onee(lai_basis* b, int kind, int nch, const double* Z, const double* xyz, double* d_out,
                            void* stream) {
  if (b && b->n_mol != 1)
    return fail(LAI_ERR_UNSUPPORTED, "lai_dev_onee: one-electron matrices need a single-molecule basis handle");
  if (!b || !d_out || kind < 0 || kind > 2 || (kind == 2 && nch > 0 && (!Z || !xyz)))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_onee: invalid arguments");
  DevBuf<double> dZ, dxyz;
  hipError_t e;
  if (kind == 2 && nch > 0) {
    if ((e = dZ.assign(Z, nch)) != hipSuccess || (e = dxyz.assign(xyz, size_t(3) * nch)) != hipSuccess)
      return fail(LAI_ERR_MEMORY, "lai_dev_onee: charge upload");
  }
  OneEArgs a{};
  a.n_shells = b->ns;
  a.L = b->dL.p;
  a.prim_off = b->dprim_off.p;
  a.nprim = b->dnprim.p;
  a.func_off = b->dfunc_off.p;
  a.center = b->dcenter.p;
  a.exps = b->dexps.p;
  a.coefs = b->dcoefs.p;
  a.kind = kind;
  a.n_charges = kind == 2 ? nch : 0;
  a.Z = dZ.p;
  a.cxyz = dxyz.p;
  a.rys = b->ctx->rys.p;
  a.out = d_out;
  a.nbf = b->nbf;
  const int npairs = b->ns * (b->ns + 1) / 2;
  hipLaunchKernelGGL(onee_kernel, dim3((npairs + 63) / 64), dim3(64), 0, hipStream_t(stream), a);
  LAI_HIP(hipGetLastError());
  LAI_HIP(hipStreamSynchronize(hipStream_t(stream)));  // charge buffers are released on return
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// probes
// ---------------------------------------------------------------------------
static lai_ctx* probe_ctx() {
  static lai_ctx* c = nullptr;
  if (!c) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (lai_dev_init(dev, &c) != LAI_OK) c = nullptr;
  }
  return c;
}

extern "C" int lai_dev_rys(int nroots, int n, const double* d_T, double* d_r, double* d_w, void* stream) {
  if (nroots < 1 || nroots > LAI_RYS_NMAX || n < 0 || (n > 0 && (!d_T || !d_r || !d_w)))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_rys: invalid arguments");
  if (n == 0) return LAI_OK;
  lai_ctx* c = probe_ctx();
  if (!c) return fail(LAI_ERR_BACKEND, "lai_dev_rys: no device context");
  hipLaunchKernelGGL(rys_probe_kernel, dim3((n + 127) / 128), dim3(128), 0, hipStream_t(stream), nroots, n, d_T,
                     c->rys.p, d_r, d_w);
  LAI_HIP(hipGetLastError());
  return LAI_OK;
}

extern "C" int lai_dev_boys(int nmax, int n, const double* d_T, double* d_out, void* stream) {
  if (nmax < 0 || nmax > 40 || n < 0 || (n > 0 && (!d_T || !d_out)))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_boys: invalid arguments");
  if (n == 0) return LAI_OK;
  hipLaunchKernelGGL(boys_probe_kernel, dim3((n + 127) / 128), dim3(128), 0, hipStream_t(stream), nmax, n, d_T,
                     d_out);
  LAI_HIP(hipGetLastError());
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// nuclear gradient
// ---------------------------------------------------------------------------
extern "C" int lai_dev_set_mixed_prim_cut(lai_basis* b, double rel_cut) {
  if (!b || !(rel_cut >= 0.0) || rel_cut > 1e-3)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_mixed_prim_cut: cut must lie in [0, 1e-3]");
  if (b->pcut32 != rel_cut && b->graph_exec) {  // a captured graph holds the old value
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  b->pcut32 = rel_cut;
  return LAI_OK;
}

extern "C" int lai_dev_set_shell_atoms(lai_basis* b, const int* atom, int n_atoms) {
  if (!b || !atom || n_atoms <= 0) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_shell_atoms: invalid arguments");
  for (int s = 0; s < b->ns; ++s)
    if (atom[s] < 0 || atom[s] >= n_atoms)
      return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_shell_atoms: atom index out of range");
  hipError_t e = b->shell_atom.assign(atom, b->ns);
  if (e != hipSuccess) return fail(LAI_ERR_MEMORY, "lai_dev_set_shell_atoms: upload");
  b->n_atoms = n_atoms;
  return LAI_OK;
}

extern "C" int lai_dev_jk_grad(lai_basis* b, const double* d_D, double xscale, double* d_grad, int n_atoms,
                               void* stream) {
  if (!b || !d_D || !d_grad) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk_grad: invalid arguments");
  if (!b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_jk_grad: worklist not built");
  if (!b->shell_atom.p || n_atoms < b->n_atoms)
    return fail(LAI_ERR_INVALID_STATE, "lai_dev_jk_grad: shell atoms not set or gradient too short");
  if (b->max_l > 2) return fail(LAI_ERR_UNSUPPORTED, "lai_dev_jk_grad: gradient kernels cover shells up to d");
  hipStream_t st = hipStream_t(stream);
  const auto& G = grad_launchers();
  GradArgs a{};
  a.pairs = b->pairs.p;
  a.pp = b->pp.p;
  a.ppab = b->ppab.p;
  a.rys = b->ctx->rys.p;
  a.D = d_D;
  a.nbf = b->max_nbf;
  a.xscale = xscale;
  a.shell_atom = b->shell_atom.p;
  a.grad = d_grad;
  a.n_atoms = b->n_atoms;
  a.ppih = b->ppih.p;
  a.segs = b->segs.p;
  a.kets = b->kets.p;
  a.sel = make_sel(b);
  a.sel.mixed = 0;  // the gradient runs every quartet in FP64
  const int max_blocks = b->ctx->n_cu * 16;
  // the classes fan out over the side streams as in the J/K build (every
  // class adds into d_grad with atomics only)
  if (!b->fork) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LAI_HIP(hipStreamIsCapturing(st, &cs));
    if (cs == hipStreamCaptureStatusNone) {
      const int rc = ensure_side_streams(b);
      if (rc != LAI_OK) return rc;
    }
  }
  const bool fan = b->fork != nullptr;
  hipStream_t streams[kSideStreams + 1] = {st};
  double load[kSideStreams + 1] = {};
  if (fan) {
    LAI_HIP(hipEventRecord(b->fork, st));
    for (int i = 0; i < kSideStreams; ++i) {
      streams[i + 1] = b->side[i];
      LAI_HIP(hipStreamWaitEvent(b->side[i], b->fork, 0));
    }
  }
  for (int c = kNumGradClasses - 1; c >= 0; --c) {
    const ClassList& C = b->cls[c];
    if (C.nchunks <= 0) continue;
    const bool own = C.ngchunks > 0;  // the J/K chunks are tuned: default ones
    a.chunks = own ? C.gchunks.p : C.chunks.p;
    a.nchunks = int(own ? C.ngchunks : C.nchunks);
    int k = 0;
    for (int i = 1; fan && i <= kSideStreams; ++i)
      if (load[i] < load[k]) k = i;
    load[k] += b->cls_f64[c] + 1e6 * double(a.nchunks);
    G.f[c](a, streams[k], max_blocks);
  }
  if (fan)
    for (int i = 0; i < kSideStreams; ++i) {
      LAI_HIP(hipEventRecord(b->join[i], b->side[i]));
      LAI_HIP(hipStreamWaitEvent(st, b->join[i], 0));
    }
  LAI_HIP(hipGetLastError());
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// autotuner and dispatch table
// ---------------------------------------------------------------------------
namespace {
void drop_graph(lai_basis* b) {
  if (b->graph_exec) {
    (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
}

lai_class_tune tune_entry(const lai_basis* b, int c) {
  lai_class_tune e{};
  e.cls = c;
  e.strategy = class_shape(launchers().info[c], b->tune[c]);
  e.chunk = b->tune[c].chunk;
  e.blocks_per_cu = b->tune[c].bpc;
  return e;
}

// quartets of class c in the current worklist
size_t class_quartets(const lai_basis* b, int c) {
  size_t n = 0;
  for (const auto& r : b->cls[c].runs) n += size_t(r.k1);
  return n;
}
}  // namespace

extern "C" int lai_dev_class_strategies(int cls, int* strategy, int* alternative) {
  if (cls < 0 || cls >= kNumClasses || !strategy || !alternative)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_class_strategies: invalid arguments");
  const ClassInfo& info = launchers().info[cls];
  *strategy = info.shape;
  *alternative = info.alt;
  return LAI_OK;
}

extern "C" int lai_dev_get_tuning(const lai_basis* b, lai_class_tune* table, int cap, int* n_out) {
  if (!b || cap < 0 || (cap > 0 && !table)) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_get_tuning: invalid arguments");
  const int n = std::min(cap, kNumClasses);
  for (int c = 0; c < n; ++c) table[c] = tune_entry(b, c);
  if (n_out) *n_out = kNumClasses;
  return LAI_OK;
}

extern "C" int lai_dev_set_tuning(lai_basis* b, const lai_class_tune* table, int n) {
  if (!b || n < 0 || (n > 0 && !table)) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_tuning: invalid arguments");
  const auto& Ls = launchers();
  ClassTune next[kNumClasses];
  bool touched[kNumClasses] = {};
  if (n == 0) {
    for (int c = 0; c < kNumClasses; ++c) touched[c] = true;
  } else {
    for (int c = 0; c < kNumClasses; ++c) next[c] = b->tune[c];
  }
  for (int i = 0; i < n; ++i) {  // validate every entry before applying any
    const lai_class_tune& e = table[i];
    if (e.cls < 0 || e.cls >= kNumClasses)
      return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_tuning: class " + std::to_string(e.cls) + " out of range");
    const ClassInfo& info = Ls.info[e.cls];
    if (e.strategy != info.shape && (info.alt == kShapeNone || e.strategy != info.alt))
      return fail(LAI_ERR_UNSUPPORTED, "lai_dev_set_tuning: strategy " + std::to_string(e.strategy) +
                                           " is not compiled for class " + std::to_string(e.cls));
    if (e.chunk < 0 || e.chunk > (1 << 20) || e.blocks_per_cu < 0 || e.blocks_per_cu > 4096)
      return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_set_tuning: chunk or blocks_per_cu out of range");
    next[e.cls].alt = e.strategy == info.shape ? 0 : 1;
    next[e.cls].chunk = e.chunk;
    next[e.cls].bpc = e.blocks_per_cu;
    touched[e.cls] = true;
  }
  for (int c = 0; c < kNumClasses; ++c) {
    if (!touched[c]) continue;
    b->tune[c] = next[c];
    if (b->have_worklist) {
      const int rc = upload_chunks(b, c);
      if (rc != LAI_OK) return rc;
    }
  }
  drop_graph(b);
  return LAI_OK;
}

extern "C" int lai_dev_autotune(lai_basis* b, const double* d_D, int flags, int reps, lai_class_tune* table, int cap,
                                int* n_out, void* stream) {
  if (!b || !d_D || reps < 1 || reps > 100 || cap < 0 || (cap > 0 && !table))
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_autotune: invalid arguments");
  if (!b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_autotune: worklist not built");
  if (flags != LAI_JK_FP64 && flags != LAI_JK_MIXED)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_autotune: unknown flags");
  hipStream_t st = hipStream_t(stream);
  const auto& Ls = launchers();
  const int n_cu = b->ctx->n_cu;
  EriArgs a = jk_args(b, d_D);
  a.segs = b->segs.p;
  a.kets = b->kets.p;
  a.sel = make_sel(b);
  const bool mixed = flags == LAI_JK_MIXED && b->wl_mixed;
  const int sel_mixed = a.sel.mixed;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LAI_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(LAI_ERR_BACKEND, "lai_dev_autotune: hipEventCreate failed");
  }
  // the trial launches accumulate into the J/K scratch, which every build clears
  int rc = LAI_OK;
  int n_rep = 0;
  // milliseconds per launch of class c under b->tune[c] (chunks uploaded);
  // a mixed build times its FP64 and FP32 launches together, as it runs them
  auto time_class = [&](int c, double& ms) -> int {
    const ClassList& C = b->cls[c];
    a.chunks = C.chunks.p;
    a.nchunks = int(C.nchunks);
    LaunchTune t;
    t.alt = b->tune[c].alt;
    t.blocks_per_cu = b->tune[c].bpc;
    t.n_cu = n_cu;
    const bool split = class_split(b, c, mixed);
    const int64_t n32 = split ? b->cls_n32[c] : 0;
    const bool p64 = b->counts[c] - n32 > 0, p32 = n32 > 0;
    a.sel.mixed = split ? sel_mixed : 0;
    auto f = [&](hipStream_t s, const LaunchTune& tt) {
      if (p64) {
        a.sel.keep = split ? 1 : 0;
        Ls.f64[c](a, s, tt);
      }
      if (p32) {
        a.sel.keep = 2;
        Ls.f32[c](a, s, tt);
      }
    };
    f(st, t);  // warm-up
    LAI_HIP(hipGetLastError());
    LAI_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) f(st, t);
    LAI_HIP(hipEventRecord(e1, st));
    LAI_HIP(hipGetLastError());
    LAI_HIP(hipEventSynchronize(e1));
    float t_ms = 0.f;
    LAI_HIP(hipEventElapsedTime(&t_ms, e0, e1));
    ms = double(t_ms) / reps;
    return LAI_OK;
  };
  int failed_class = -1;
  ClassTune failed_saved{};
  for (int c = kNumClasses - 1; c >= 0 && rc == LAI_OK; --c) {
    const size_t nq = class_quartets(b, c);
    if (nq == 0) continue;
    const ClassInfo& info = Ls.info[c];
    const ClassTune saved = b->tune[c];
    auto eval = [&](const ClassTune& t, double& ms) -> int {
      b->tune[c] = t;
      const int r = upload_chunks(b, c);
      if (r != LAI_OK) return r;
      return time_class(c, ms);
    };
    // stage 1: shape x chunk cap at the default grid
    std::vector<std::pair<int, int>> tried;  // (shape, effective cap)
    ClassTune best{};
    double best_ms = 0.0, dflt_ms = 0.0;
    failed_class = c;
    failed_saved = saved;
    if ((rc = eval(best, dflt_ms)) != LAI_OK) break;
    best_ms = dflt_ms;
    tried.emplace_back(info.shape, default_chunk(info.shape, nq, n_cu, info.nt));
    for (int alt = 0; alt <= (info.alt != kShapeNone ? 1 : 0) && rc == LAI_OK; ++alt) {
      const int shape = alt ? info.alt : info.shape;
      std::vector<int> caps;
      // lane-per-quartet classes: up to whole ket runs per chunk (fewer
      // exchange-row flushes, fewer workgroups)
      if (shape == kShapeRegs) caps = {std::min(2 * info.nt, 2048), std::min(4 * info.nt, 2048), 2048, 4096, 8192};
      else if (shape == kShapeWaves) caps = {8, 16, 32, 64};
      else caps = {4, 8, 16, 32};
      if (alt) caps.insert(caps.begin(), 0);
      for (int cap_q : caps) {
        const int eff = cap_q > 0 ? cap_q : default_chunk(shape, nq, n_cu, shape == info.shape ? info.nt : 256);
        if (std::find(tried.begin(), tried.end(), std::make_pair(shape, eff)) != tried.end()) continue;
        tried.emplace_back(shape, eff);
        ClassTune t{};
        t.alt = alt;
        t.chunk = cap_q;
        double ms = 0.0;
        if ((rc = eval(t, ms)) != LAI_OK) break;
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    }
    if (rc != LAI_OK) break;
    // stage 2: grid cap of the best
    const int shape = class_shape(info, best);
    const std::vector<int> bpcs = shape == kShapeRegs ? std::vector<int>{4, 16} : std::vector<int>{16, 512};
    const ClassTune base = best;
    for (int bpc : bpcs) {
      ClassTune t = base;
      t.bpc = bpc;
      double ms = 0.0;
      if ((rc = eval(t, ms)) != LAI_OK) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    if (rc != LAI_OK) break;
    if (!(best_ms < 0.99 * dflt_ms)) {  // keep the default unless clearly faster
      best = ClassTune{};
      best_ms = dflt_ms;
    }
    b->tune[c] = best;
    if ((rc = upload_chunks(b, c)) != LAI_OK) break;
    if (n_rep < cap) {
      lai_class_tune& e = table[n_rep];
      e = tune_entry(b, c);
      e.ms = best_ms;
      e.ms_default = dflt_ms;
    }
    ++n_rep;
    failed_class = -1;
  }
  if (rc != LAI_OK && failed_class >= 0) {  // leave the failing class as it was
    b->tune[failed_class] = failed_saved;
    (void)upload_chunks(b, failed_class);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graph(b);
  if (rc != LAI_OK) return rc;
  if (n_out) *n_out = n_rep;
  return LAI_OK;
}

// ---------------------------------------------------------------------------
// profiled J/K build: one hipEvent pair around each class kernel
// ---------------------------------------------------------------------------
extern "C" int lai_dev_worklist_class_flops(const lai_basis* b, double* f64, double* f32) {
  if (!b || !b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_worklist_class_flops: no worklist");
  if (f64) std::memcpy(f64, b->cls_f64, sizeof(b->cls_f64));
  if (f32) std::memcpy(f32, b->cls_f32, sizeof(b->cls_f32));
  return LAI_OK;
}

extern "C" int lai_dev_worklist_grad_flops(const lai_basis* b, double* model0, double* model1) {
  if (!b || !b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_worklist_grad_flops: no worklist");
  for (int c = 0; c < kNumClasses; ++c) {
    if (model0) model0[c] = b->cls_grad[2 * c];
    if (model1) model1[c] = b->cls_grad[2 * c + 1];
  }
  return LAI_OK;
}

extern "C" int lai_dev_jk_profile(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags, void* stream,
                                  double* ms_fp64, double* ms_fp32) {
  if (!b || !d_D || !d_J || !d_K) return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk_profile: invalid arguments");
  if (!b->have_worklist) return fail(LAI_ERR_INVALID_STATE, "lai_dev_jk_profile: worklist not built");
  if ((flags & ~(LAI_JK_MIXED | LAI_JK_ANTISYMMETRIC)) != 0)
    return fail(LAI_ERR_INVALID_ARGUMENT, "lai_dev_jk_profile: unknown flags");
  hipStream_t st = hipStream_t(stream);
  const size_t nn = size_t(b->mat_size);
  int rc;
  if ((rc = lai_dev_zero(b->Jt.p, nn, stream)) != LAI_OK) return rc;
  if ((rc = lai_dev_zero(b->Kt.p, nn, stream)) != LAI_OK) return rc;
  std::vector<hipEvent_t> ev(4 * kNumClasses, nullptr);
  for (auto& e : ev) LAI_HIP(hipEventCreate(&e));
  launch_jk_classes(b, jk_args(b, d_D), flags, st, ev.data());
  LAI_HIP(hipGetLastError());
  if ((rc = symmetrize(b, d_J, d_K, st, (flags & LAI_JK_ANTISYMMETRIC) != 0)) != LAI_OK) return rc;
  LAI_HIP(hipStreamSynchronize(st));
  // the FP64 and FP32 launches of each class (the FP32 one is empty unless
  // the build is mixed)
  for (int c = 0; c < kNumClasses; ++c) {
    float t64 = 0.f, t32 = 0.f;
    LAI_HIP(hipEventElapsedTime(&t64, ev[4 * c], ev[4 * c + 1]));
    LAI_HIP(hipEventElapsedTime(&t32, ev[4 * c + 2], ev[4 * c + 3]));
    if (ms_fp64) ms_fp64[c] = t64;
    if (ms_fp32) ms_fp32[c] = t32;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return LAI_OK;
}
