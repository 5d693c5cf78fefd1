/*
 * lai_device.h -- C ABI of the MI355X (gfx950) Rys ERI / Fock-build backend.
 *
 * This is the boundary the C++ host layer (Engine, GpuFockBuilder, FockBuilder)
 * calls.  Plain pointers, sizes and status codes only; no C++ or HIP types.
 * Every function returns LAI_OK (0) or a negative status; the message of the
 * last failure on the calling thread is available from lai_dev_last_error().
 * No exception crosses this boundary.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository brycewestheimer/libaccint-public).
 *
 * Pointer conventions: "d_" arguments are device pointers (allocated with
 * lai_dev_malloc or any HIP allocator), everything else is host memory.
 * "stream" is a hipStream_t passed as void* (NULL = the default stream).
 * Matrices are row-major nbf x nbf doubles.  Shell order, function offsets
 * and Cartesian component order follow the reference BasisSet
 * (src/shared/basis/basis_set.cpp:161-185,
 *  include/libaccint/math/cartesian_indices.hpp:26-35).
 */
#ifndef LAI_DEVICE_H
#define LAI_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LAI_OK 0
#define LAI_ERR_INVALID_ARGUMENT (-1) /* -> InvalidArgumentException */
#define LAI_ERR_INVALID_STATE (-2)    /* -> InvalidStateException */
#define LAI_ERR_MEMORY (-3)           /* -> MemoryException / std::runtime_error */
#define LAI_ERR_BACKEND (-4)          /* -> BackendError (HIP runtime failure) */
#define LAI_ERR_UNSUPPORTED (-5)      /* -> std::invalid_argument (angular momentum) */
#define LAI_ERR_NUMERIC (-6)          /* -> std::runtime_error (metric not positive definite) */

/* J/K build flags */
#define LAI_JK_FP64 0  /* every quartet in FP64 */
/* quartets marked FP32 by the worklist run the FP32 recursion, in the classes
 * of wavefront and workgroup shape (a launch per precision); the
 * lane-per-quartet classes, bound by memory latency rather than by their
 * arithmetic, run them in FP64 */
#define LAI_JK_MIXED 1
/* flag bit, with either of the above: d_D is antisymmetric (D^T = -D); d_J is
 * left alone and only the antisymmetric part (Kt - Kt^T)/2 of the exchange
 * build is added to d_K.  K is linear in D, so a general D is the build of
 * (D + D^T)/2 followed by this build of (D - D^T)/2: the exchange matrix of
 * the reference's N^4 FockBuilder::accumulate for any D. */
#define LAI_JK_ANTISYMMETRIC 2
/* flag bit, with LAI_JK_FP64: the register-shape classes whose chunk code
 * fits a 256-lane workgroup run as one persistent launch over a device queue
 * of their chunks, sorted by class, heaviest first (workgroups claim items
 * with an atomic counter); the other classes keep their per-class launches,
 * concurrently.  Same J/K to the order of the FP64 atomic sums.  Ignored by
 * LAI_JK_MIXED builds.  The queue is allocated by the first such build after
 * a worklist change (lai_dev_jk_graph_capture does that before it captures;
 * inside a capture of the caller's own, a first build fails).  The
 * one-launch-over-a-queue form of the reference's
 * per-ShellSetQuartet launches (src/device/cuda/engine/cuda_engine.cu:562-665). */
#define LAI_JK_PERSISTENT 4

#define LAI_MAX_L 3 /* highest shell angular momentum of the energy kernels (f) */

typedef struct lai_ctx lai_ctx;
typedef struct lai_basis lai_basis;
typedef struct lai_df lai_df;

/* ---- context --------------------------------------------------------------
 * Replaces CudaEngine construction (src/device/cuda/engine/cuda_engine.cu:183-187:
 * BackendError on init failure) and the Boys/Rys constant uploads of
 * src/device/cuda/math/boys_device.cu. */
int lai_dev_init(int device, lai_ctx** out);
int lai_dev_fini(lai_ctx* ctx);
const char* lai_dev_last_error(void);
int lai_dev_device_count(int* n);

/* ---- memory helpers (replace include/libaccint/memory/device_memory.hpp) -- */
int lai_dev_malloc(lai_ctx* ctx, size_t bytes, void** d_ptr);
int lai_dev_free(lai_ctx* ctx, void* d_ptr);
int lai_dev_memcpy_h2d(void* d_dst, const void* src, size_t bytes, void* stream);
int lai_dev_memcpy_d2h(void* dst, const void* d_src, size_t bytes, void* stream);
/* zero n doubles with a fill kernel (graph-capture safe; replaces
 * matrix_zero_kernel, src/device/cuda/consumers/fock_builder_gpu.cu:99-105) */
int lai_dev_zero(void* d_ptr, size_t n_doubles, void* stream);
int lai_dev_stream_sync(void* stream);
/* Makes stream waiter wait for the work enqueued so far on stream producer
 * (an event edge; no host wait).  The ERI-producer to consumer ordering of
 * include/libaccint/engine/cuda_engine.hpp:889-894. */
int lai_dev_stream_wait(void* waiter, void* producer);

/* ---- basis ------------------------------------------------------------------
 * Uploads shells (normalised coefficients, src/shared/basis/shell.cpp:32-113)
 * and builds the primitive-pair tables of every shell pair.  Replaces
 * ShellSetDeviceCache::get_or_upload (src/device/cuda/basis/device_data.cu:42-90)
 * and ShellSetPair::build_pair_data_impl (src/shared/basis/shell_set_pair.cpp:128-181).
 * Zero-coefficient primitives are dropped from the pair tables (exact). */
int lai_dev_upload_basis(lai_ctx* ctx, int n_shells, const int* L, const double* centers,
                         const int* nprim, const double* exps, const double* coefs,
                         lai_basis** out);
int lai_dev_basis_free(lai_basis* b);
/* basis functions of all molecules of the handle */
int lai_dev_basis_nbf(const lai_basis* b);

/* ---- batches of molecules ------------------------------------------------------
 * One handle for several independent molecules (the uneven multi-molecule
 * workload): molecule m owns shells [mol_shell_off[m], mol_shell_off[m+1])
 * of the arrays (same layout as lai_dev_upload_basis).  Shell pairs, Schwarz
 * bounds and quartets never mix molecules; every J/K class kernel covers all
 * molecules in one launch.  D, J and K are the molecules' nbf_m x nbf_m
 * row-major matrices one after another (lai_dev_batch_layout); the Schwarz
 * table is the molecules' packed tables one after another; shell_dmax of
 * lai_dev_build_worklist is the molecules' n_shells_m x n_shells_m blocks one
 * after another.  One-electron matrices need a single-molecule handle.
 * No reference counterpart (the reference computes one BasisSet per Engine). */
int lai_dev_upload_batch(lai_ctx* ctx, int n_mol, const int* mol_shell_off, int n_shells, const int* L,
                         const double* centers, const int* nprim, const double* exps, const double* coefs,
                         lai_basis** out);
/* n_mol; per molecule its basis functions (nbf[n_mol]) and the offset of its
 * block in the batched matrices (mat_off[n_mol + 1], the last entry the total
 * length); any output may be NULL */
int lai_dev_batch_layout(const lai_basis* b, int* n_mol, int* nbf, int64_t* mat_off);

/* ---- raw ERIs ---------------------------------------------------------------
 * (ab|cd) blocks for n arbitrary shell quartets (shells4 = 4n shell indices),
 * block t written row-major [na][nb][nc][nd] at out + offsets[t] (host
 * buffers; offsets has n+1 entries and is filled by the call).
 * Replaces kernels::cuda::dispatch_eri_kernel + the D2H of
 * CudaEngine::compute_eri_batch_device (src/device/cuda/kernels/eri_kernel.cu:635-689,
 * src/device/cuda/engine/cuda_engine.cu:1233-1277) and, on the CPU side,
 * kernels::compute_eri (src/host/kernels/eri_kernel.cpp:168-381). */
int lai_dev_eri_quartets(lai_basis* b, int64_t n, const int* shells4, double* out,
                         int64_t* offsets);
/* The same blocks written to a device buffer d_out (offsets[n] doubles) on
 * stream; with d_out NULL only offsets is filled (to size the buffer).
 * Synchronises the stream.  Replaces CudaEngine::compute_eri_batch_device
 * without its D2H copy, the device-resident batch that
 * GpuFockBuilder::accumulate_device_eri_batch consumes
 * (src/device/cuda/engine/cuda_engine.cu:1233-1277,
 * src/device/cuda/consumers/fock_builder_gpu.cu:600-621). */
int lai_dev_eri_quartets_device(lai_basis* b, int64_t n, const int* shells4, double* d_out, int64_t* offsets,
                                void* stream);
/* The same blocks as lai_dev_eri_quartets_device, bit for bit, with no device
 * allocation: the per-class quartet descriptors are staged in a static 2 MiB
 * device array of the library (one copy per GPU, set up by lai_dev_init and
 * held until the process ends) instead of hipMalloc'd buffers.  At most
 * LAI_NOALLOC_MAX_QUARTETS quartets per call (LAI_ERR_INVALID_ARGUMENT
 * beyond).  d_out must hold offsets[n] doubles (d_out NULL fills offsets
 * only).  Every call on a GPU shares that GPU's staging array: the calls of
 * one process are serialised by the library, and calls that overlap on one
 * GPU from elsewhere would race.  Synchronises the stream.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
#define LAI_NOALLOC_MAX_QUARTETS 65536
int lai_dev_eri_quartets_device_noalloc(lai_basis* b, int64_t n, const int* shells4, double* d_out,
                                        int64_t* offsets, void* stream);

/* ---- Fock digestion of ERI blocks ---------------------------------------------
 * One block per shell quartet (a b | c d): na*nb*nc*nd doubles at d_eri + off,
 * row-major [a][b][c][d]; fa..fd are the shells' first basis functions. */
typedef struct lai_eri_block {
  int64_t off;
  int32_t fa, fb, fc, fd;
  int32_t na, nb, nc, nd; /* each in [1, 10] (s..f Cartesian) */
} lai_eri_block;
/* d_J[mu nu] += (mu nu|la si) D[la si] and d_K[mu la] += (mu nu|la si) D[nu si]
 * over every element of n_blocks blocks (device arrays d_eri, d_blocks; each
 * element counted once: the plain N^4 accumulate of
 * src/shared/consumers/fock_builder.cpp:98-117).  Every index must lie in
 * [0, nbf): a block that does not fit is skipped.  A wavefront per block sums
 * its J and K rows in LDS and adds them with one atomic per element.
 * Asynchronous on stream.  Replaces detail::launch_fock_accumulate_kernel and
 * launch_fock_accumulate_batch_kernel
 * (include/libaccint/consumers/fock_builder_gpu.hpp:178-214,
 * src/device/cuda/consumers/fock_builder_gpu.cu:55-96,116-180). */
int lai_dev_fock_accumulate(const double* d_eri, const lai_eri_block* d_blocks, int64_t n_blocks, const double* d_D,
                            double* d_J, double* d_K, int nbf, void* stream);

/* ---- Schwarz screening -------------------------------------------------------
 * Q_ij = sqrt(max_ab |(ab|ab)|) for every shell pair i <= j, packed at
 * j*(j+1)/2 + i (src/shared/screening/schwarz_bounds.cpp:20-59,
 * include/libaccint/screening/schwarz_bounds.hpp:110-113).  Computed on the
 * device and kept there; copied to Q_packed when it is not NULL. */
int lai_dev_schwarz(lai_basis* b, double* Q_packed);
/* use a caller-provided Q table instead (same packing) */
int lai_dev_set_schwarz(lai_basis* b, const double* Q_packed);

/* ---- worklist ---------------------------------------------------------------
 * Canonical quartets i<=j, k<=l, pair(ij) <= pair(kl) that pass
 * Q_ij*Q_kl >= thr (ScreeningOptions::passes_screening,
 * include/libaccint/screening/screening_options.hpp:147-149; loop
 * include/libaccint/engine/engine.hpp:1815-1819), sorted by angular-momentum
 * class.  thr <= 0 keeps every quartet (no Schwarz table needed).
 * shell_dmax (optional, n_shells*n_shells, row-major) holds max |D| over each
 * shell-pair block.  density_screen != 0 additionally requires
 * Dk * Q_ij * Q_kl >= thr with the exchange-type
 * Dk = max(Dmax_ik, Dmax_il, Dmax_jk, Dmax_jl)
 * (include/libaccint/screening/density_screening.hpp:80-104).
 * tau_fp32 > 0 (with shell_dmax) marks quartets with
 * Q_ij*Q_kl*max(Dmax_ij, Dmax_kl, Dmax_ik, Dmax_il, Dmax_jk, Dmax_jl) < tau_fp32
 * for the FP32 recursion of LAI_JK_MIXED.  Returns the quartet count, the
 * FP32 count and an order-independent checksum of the (i,j,k,l) list: the sum
 * over quartets of lai_quartet_mix64(i,j,k,l) modulo 2^64.  Replaces the
 * QuartetTask list of compute_and_consume_screened_parallel_impl
 * (engine.hpp:1877-1923). */
int lai_dev_build_worklist(lai_basis* b, double thr, const double* shell_dmax, int density_screen,
                           double tau_fp32, int64_t* n_quartets, int64_t* n_fp32, uint64_t* checksum);
uint64_t lai_quartet_mix64(int i, int j, int k, int l);
/* Per-class quartet counts and algorithmic work of the current worklist.
 * Classes: pair class p(La,Lb) = La*(La+1)/2 + Lb for La >= Lb; quartet class
 * pb*(pb+1)/2 + pk for bra pair class pb >= ket pair class pk (55 classes up
 * to f).  flops_fp64/flops_fp32 sum K_q * F_prim over the FP64 / FP32 parts,
 * with K_q the product of non-zero-coefficient primitive counts and
 * F_prim = 30 + n_r (50 + 3 F_dir + 3 N_cart) in the reference orientation
 * of each canonical quartet; n_eri sums na*nb*nc*nd. */
int lai_dev_worklist_stats(const lai_basis* b, int64_t* counts, double* flops_fp64,
                           double* flops_fp32, int64_t* n_eri);

/* per-class FLOPs of the work the kernels execute (55 entries each): as
 * flops_fp64/flops_fp32 of lai_dev_worklist_stats, but with K_q counting only
 * the primitive pairs above the primitive screening bound */
int lai_dev_worklist_class_flops(const lai_basis* b, double* flops_fp64, double* flops_fp32);

/* Per-class work of the gradient over the current worklist (55 entries each;
 * executed primitive pairs, the same selection as lai_dev_jk_grad).  model0:
 * per primitive quartet the derivative classes of three centres (each
 * centre's l raised and, when l > 0, lowered by one), F_prim each, plus
 * 9 N_cart contraction flops per root; model1: one table per (root, axis)
 * with the three centres raised together, 30 + n_r (50 + 3 F_dir(La+1, Lb+1,
 * Lc+1, Ld) + 12 N_cart).  Measurement aid; no reference counterpart (the
 * reference has no derivative integrals). */
int lai_dev_worklist_grad_flops(const lai_basis* b, double* model0, double* model1);

/* ---- J/K build ---------------------------------------------------------------
 * d_J += J[D], d_K += K[D] over the current worklist (D symmetric, device).
 * J_mn = sum (mn|ls) D_ls, K_ml = sum (mn|ls) D_ns, i.e. the result of
 * FockBuilder over the full basis (src/shared/consumers/fock_builder.cpp:64-222).
 * Replaces GpuFockBuilder::accumulate_device_eri_batch +
 * fock_accumulate_batch_kernel (src/device/cuda/consumers/fock_builder_gpu.cu:116-180,600-621)
 * with digestion fused into the ERI kernels.  Asynchronous on stream. */
int lai_dev_jk(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags, void* stream);

/* The same J/K build with a hipEvent pair around every class launch, one
 * class after another; writes the per-class milliseconds of the FP64 launch
 * and of the FP32 launch (55 entries each; a mixed build runs each class as
 * one launch per precision, the FP32 ones are 0 otherwise) and synchronises
 * the stream.  Measurement aid; no reference counterpart. */
int lai_dev_jk_profile(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags, void* stream,
                       double* ms_fp64, double* ms_fp32);

/* Relative primitive cut of the FP32 recursion of LAI_JK_MIXED: an FP32
 * quartet skips the primitive quartets whose bound |Kc_ab| zeta^(-1/4)
 * |Kc_cd| eta^(-1/4) is below rel_cut times the product of the two pairs'
 * summed bounds (0 = none; default 1e-7; at most 1e-3).  FP64 quartets are
 * unaffected.  No reference counterpart (the reference has no FP32 ERIs). */
int lai_dev_set_mixed_prim_cut(lai_basis* b, double rel_cut);

/* Capture the same J/K build into a hipGraph bound to these pointers, and
 * replay it (the per-SCF-iteration Fock build). */
int lai_dev_jk_graph_capture(lai_basis* b, const double* d_D, double* d_J, double* d_K, int flags,
                             void* stream);
int lai_dev_jk_graph_launch(lai_basis* b, void* stream);

/* ---- dispatch table and autotuner ---------------------------------------------
 * One entry per quartet class (class index as in lai_dev_worklist_stats): the
 * execution strategy of its J/K kernel -- LAI_STRATEGY_THREAD (a lane per
 * quartet, register tables), LAI_STRATEGY_WAVEFRONT (a wavefront per quartet)
 * or LAI_STRATEGY_BLOCK (a workgroup per quartet, LDS tables) -- the quartets
 * per chunk (a chunk is one bra pair with a run of its kets, taken by one
 * workgroup; 0 = the strategy's default rule) and the grid cap in workgroups
 * per CU (0 = the strategy's default).  Mirrors
 * OptimalDispatchTable / TwoElectronEntry::gpu_strategy
 * (include/libaccint/kernels/optimal_dispatch_table.hpp:29-33,64-67,83-149),
 * which the reference fills offline from bench_optimal_dispatch. */
#define LAI_STRATEGY_THREAD 0
#define LAI_STRATEGY_WAVEFRONT 1
#define LAI_STRATEGY_BLOCK 2
typedef struct lai_class_tune {
  int cls;
  int strategy;
  int chunk;
  int blocks_per_cu;
  double ms;         /* lai_dev_autotune: milliseconds per launch of the choice, else 0 */
  double ms_default; /* lai_dev_autotune: milliseconds per launch of the default, else 0 */
} lai_class_tune;

/* The strategy compiled as a class's default and its alternative (-1: none). */
int lai_dev_class_strategies(int cls, int* strategy, int* alternative);
/* Times candidate launches of every class of the current worklist on d_D --
 * each compiled strategy with several chunk caps, then grid caps for the best
 * -- reps launches each after one warm-up, and keeps per class the fastest
 * (the default unless a candidate is at least 1% faster).  Writes one entry
 * per class with work (at most cap) and their number to n_out.  Synchronises
 * the stream; drops a captured graph.  Takes the place of the reference's
 * offline bench_optimal_dispatch run (tests/benchmark/bench_optimal_dispatch.cpp)
 * and LIBACCINT_DISPATCH_TABLE loading (src/device/cuda/engine/cuda_engine.cu:201-206). */
int lai_dev_autotune(lai_basis* b, const double* d_D, int flags, int reps, lai_class_tune* table, int cap,
                     int* n_out, void* stream);
/* The current choice of the 55 classes (ms fields 0); n_out = 55. */
int lai_dev_get_tuning(const lai_basis* b, lai_class_tune* table, int cap, int* n_out);
/* Applies n entries (validated first; a strategy not compiled for the class
 * is LAI_ERR_UNSUPPORTED); n = 0 restores the defaults of every class.  The
 * choice holds across later worklist builds; drops a captured graph. */
int lai_dev_set_tuning(lai_basis* b, const lai_class_tune* table, int n);

/* ---- nuclear gradient (new capability; the reference has no derivative
 * integrals: every DerivOrder is static_assert(== 0), engine.hpp:169-170) ---
 * atom_of_shell[n_shells] maps shells to atoms for the gradient. */
int lai_dev_set_shell_atoms(lai_basis* b, const int* atom_of_shell, int n_atoms);
/* d_grad[3*n_atoms] += dE/dR for E = 1/2 tr(DJ) - x/4 tr(DK) with fixed AO
 * density d_D (device, symmetric) over the current worklist, x =
 * exchange_scale.  Derivative integrals are contracted with D (x) D on the fly;
 * the component layout per atom is (x, y, z), the centre-derivative
 * convention of include/libaccint/core/derivative_utils.hpp:69-73.
 * Shells up to d (LAI_ERR_UNSUPPORTED otherwise).  Asynchronous on stream. */
int lai_dev_jk_grad(lai_basis* b, const double* d_D, double exchange_scale, double* d_grad, int n_atoms,
                    void* stream);

/* ---- one-electron integrals ---------------------------------------------------
 * kind 0 = overlap S, 1 = kinetic T, 2 = nuclear attraction V over point
 * charges (Z, xyz[3*n_charges]).  Writes the full nbf x nbf matrix to d_out.
 * Replaces overlap/kinetic/nuclear/fused_1e kernels
 * (src/device/cuda/kernels/{overlap,kinetic,nuclear,fused_1e}_kernel.cu) and
 * CpuEngine::compute_1e_parallel_impl (src/host/engine/cpu_engine.cpp:482-593). */
int lai_dev_onee(lai_basis* b, int kind, int n_charges, const double* Z, const double* xyz,
                 double* d_out, void* stream);

/* ---- density-fitted J/K ------------------------------------------------------
 * The basis handle b holds, as one molecule, the orbital shells
 * [0, n_orb_shells), the auxiliary shells [n_orb_shells, n_orb_shells +
 * n_aux_shells) and one unit shell last (L = 0, one primitive, exponent 0,
 * coefficient 1: the constant function, so (ab|P u) = (ab|P)).  Coefficients
 * are normalised by the caller as for any shell.  lai_dev_df_create computes
 * the metric V_PQ (metric 0: Coulomb (P|Q); 1: overlap; 2, attenuated
 * Coulomb, is LAI_ERR_INVALID_ARGUMENT as in the reference), its diagonal
 * condition estimate max/min (returned in *cond_estimate when not NULL;
 * above cond_hard_limit > 0 the call fails with LAI_ERR_INVALID_ARGUMENT),
 * the Cholesky factor L (LAI_ERR_NUMERIC when V is not positive definite) and
 * L^-1, the three-centre integrals (ab|P) of every orbital pair and the
 * fitted tensor B = L^-1 (ab|P), kept in HBM as B[a][P][b] (n_orb^2 n_aux
 * doubles).  Replaces DFFockBuilder::initialize, compute_metric and
 * compute_b_tensor (src/shared/consumers/df_fock_builder.cpp:100-235) with
 * kernels::compute_two_center_metric / compute_three_center_tensor /
 * compute_B_tensor (src/host/kernels/two_center_coulomb_kernel.cpp,
 * three_center_eri_kernel.cpp) and math::cholesky_decompose /
 * triangular_inverse (include/libaccint/math/cholesky.hpp). */
int lai_dev_df_create(lai_basis* b, int n_orb_shells, int n_aux_shells, int metric, double cond_hard_limit,
                      double* cond_estimate, lai_df** out);
int lai_dev_df_free(lai_df* df);
/* orbital and auxiliary function counts */
int lai_dev_df_dims(const lai_df* df, int* n_orb, int* n_aux);
/* flags: LAI_DF_J and/or LAI_DF_K.  J = sum_P B_P gamma_P with
 * gamma_P = sum_ab B_P,ab D_ab; K = sum_P B_P D B_P^T (D, J, K device
 * n_orb x n_orb row-major; J, K overwritten; D need not be symmetric).
 * Asynchronous on stream.  Replaces DFFockBuilder::compute_j_internal /
 * compute_k_internal (src/shared/consumers/df_fock_builder.cpp:283-378). */
#define LAI_DF_J 1
#define LAI_DF_K 2
/* with LAI_DF_K: the caller states that D is symmetric, so K is symmetric;
 * the exchange product then skips the workgroup tiles below the diagonal and
 * copies the lower triangle of K from the upper one */
#define LAI_DF_SYMMETRIC 4
int lai_dev_df_jk(lai_df* df, const double* d_D, double* d_J, double* d_K, int flags, void* stream);
/* as lai_dev_df_jk, synchronous, with each kernel timed by HIP events on
 * stream: ms[0] the fitted density gamma (a pass over B for J alone; with J
 * and K together, the diagonal of X: gamma_P = sum_a (B_P D)_aa), ms[1] the
 * J pass, ms[2] the X = B D GEMM,
 * ms[3] the K = X B^T GEMM (split over the summation index), ms[4] the sum of
 * its slices (entries of kernels not run are 0).  No reference counterpart
 * (measurement). */
int lai_dev_df_jk_profile(lai_df* df, const double* d_D, double* d_J, double* d_K, int flags, void* stream,
                          double* ms);

/* ---- math probes (parity tests) ----------------------------------------------
 * Rys roots/weights for n_roots at n points T (device arrays), as used by
 * every kernel.  Replaces rys_quadrature_impl (src/device/cuda/math/rys_device.cuh:194-234). */
int lai_dev_rys(int n_roots, int n, const double* d_T, double* d_roots, double* d_weights,
                void* stream);
/* Boys F_0..F_nmax at n points (device arrays, out [n][nmax+1]); replaces
 * boys_evaluate_array_device (src/device/cuda/math/boys_device.cuh:223-300). */
int lai_dev_boys(int nmax, int n, const double* d_T, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
