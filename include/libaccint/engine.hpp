// engine.hpp -- the Engine facade: the reference's compute(...) /
// compute_and_consume(...) surface (include/libaccint/engine/engine.hpp:87-760,
// 826-930, 1578-1667) over the gfx950 backend.
//
// Every integral is computed on the GPU through the C ABI; there is no host
// integral path.  Without a visible HIP device the Engine still constructs
// (gpu_available() == false, as in the reference) but every compute call
// throws BackendError.
//
// Routing of two-electron work:
//  * FockBuilder / GpuFockBuilder over the full basis: the screened canonical
//    worklist (AM-sorted) and the fused J/K kernels; J and K are the full
//    matrices of the reference's N^4 FockBuilder path.
//  * any other TwoElectronConsumer: (ab|cd) blocks are computed on the device
//    in batches and handed to consumer.accumulate(...) in the reference's
//    N^4 loop order (cpu_engine.hpp:553-622), or to accumulate_symmetric over
//    canonical quartets when permutation symmetry is requested.
#pragma once

#include <libaccint/basis.hpp>
#include <libaccint/buffers.hpp>
#include <libaccint/consumers.hpp>
#include <libaccint/device.hpp>
#include <libaccint/operators.hpp>
#include <libaccint/screening.hpp>

#include <array>
#include <memory>
#include <span>
#include <string>
#include <string_view>
#include <type_traits>
#include <vector>

namespace libaccint {

// Options of the fused J/K build.
struct JKOptions {
    Real schwarz_threshold = 0.0;   // <= 0: every canonical quartet
    bool density_weighted = false;  // also require Dmax_K(q) Q_ij Q_kl >= thr
    bool mixed_precision = false;   // FP32 recursion for small-bound quartets
    Real fp32_threshold = 1e-7;     // FP32 iff Q_ij Q_kl max|D|(q) < fp32_threshold
    // FP32 quartets skip primitive quartets whose bound is below this fraction
    // of the quartet's summed primitive bound (0: none)
    Real fp32_prim_cut = 1e-7;
    // FP64 builds: the register-shape classes as one persistent launch over a
    // device queue of their chunks instead of one launch per class
    bool persistent = false;
};

// lai_dev_jk flags of a build with options o
inline int jk_flags(const JKOptions& o) {
    return (o.mixed_precision ? LAI_JK_MIXED : LAI_JK_FP64) | (o.persistent ? LAI_JK_PERSISTENT : 0);
}

struct WorklistInfo {
    Index n_quartets{0};
    Index n_fp32{0};
    std::uint64_t checksum{0};
    Real flops_fp64{0}, flops_fp32{0};
    Index n_eri{0};
    std::array<Index, 55> class_counts{};
};

// Execution strategy of a quartet class's J/K kernel, as in the reference's
// dispatch table (include/libaccint/kernels/optimal_dispatch_table.hpp:29-56);
// a warp is a 64-lane wavefront here.
enum class GpuExecutionStrategy : int { ThreadPerQuartet = 0, WarpPerQuartet = 1, BlockPerQuartet = 2 };
[[nodiscard]] std::string_view to_string(GpuExecutionStrategy s) noexcept;
// throws std::invalid_argument for an unknown name
[[nodiscard]] GpuExecutionStrategy gpu_execution_strategy_from_string(std::string_view s);

// Quartet class index c = pb (pb + 1) / 2 + pk of canonical angular momenta
// (la >= lb, lc >= ld, pair class p = la (la + 1) / 2 + lb, pb >= pk; 55
// classes up to f) and back; -1 for a non-canonical or out-of-range quartet.
[[nodiscard]] int quartet_class_index(const std::array<int, 4>& am) noexcept;
[[nodiscard]] std::array<int, 4> quartet_class_am(int cls);

// One quartet class of the dispatch table.
struct DispatchEntry {
    int cls{0};  // class index (the 55-entry numbering of WorklistInfo::class_counts)
    std::array<int, 4> am{};
    GpuExecutionStrategy strategy{GpuExecutionStrategy::ThreadPerQuartet};
    int chunk{0};          // quartets per chunk; 0 = the strategy's default rule
    int blocks_per_cu{0};  // grid cap; 0 = the strategy's default
    Real ms{0};            // autotune: milliseconds per launch of this choice
    Real ms_default{0};    // autotune: milliseconds per launch of the default
};

// The device build's counterpart of OptimalDispatchTable
// (optimal_dispatch_table.hpp:83-149): per-class kernel choices, produced by
// Engine::autotune or read from JSON; LIBACCINT_DISPATCH_TABLE names a file
// applied when an Engine creates its device basis (cuda_engine.cu:199-207).
struct DispatchTable {
    std::vector<DispatchEntry> entries;
    [[nodiscard]] static DispatchTable from_json(const std::string& path);
    void to_json(const std::string& path) const;
    [[nodiscard]] static DispatchTable parse_json(const std::string& text);
    [[nodiscard]] std::string dump_json() const;
};

class Engine {
public:
    explicit Engine(const BasisSet& basis, DispatchConfig config = {});
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    [[nodiscard]] const BasisSet& basis() const noexcept { return *basis_; }
    [[nodiscard]] int max_angular_momentum() const noexcept { return basis_->max_angular_momentum(); }
    [[nodiscard]] bool gpu_available() const noexcept { return gpu_available_; }
    void set_dispatch_config(DispatchConfig c) { config_ = c; }
    [[nodiscard]] const DispatchConfig& dispatch_config() const noexcept { return config_; }

    // ---- one-electron ------------------------------------------------------
    void compute_1e(const OneElectronOperator& op, std::vector<Real>& result, BackendHint hint = BackendHint::Auto);
    void compute_overlap_matrix(std::vector<Real>& S, BackendHint hint = BackendHint::Auto);
    void compute_kinetic_matrix(std::vector<Real>& T, BackendHint hint = BackendHint::Auto);
    void compute_nuclear_matrix(const PointChargeParams& charges, std::vector<Real>& V,
                                BackendHint hint = BackendHint::Auto);
    void compute_core_hamiltonian(const PointChargeParams& charges, std::vector<Real>& H,
                                  BackendHint hint = BackendHint::Auto);
    void compute(const OneElectronOperator& op, std::vector<Real>& result, BackendHint hint = BackendHint::Auto) {
        compute_1e(op, result, hint);
    }

    // ---- raw two-electron integrals ---------------------------------------
    void compute(const Operator& op, const Shell& a, const Shell& b, const Shell& c, const Shell& d,
                 TwoElectronBuffer<0>& buffer, BackendHint hint = BackendHint::Auto);
    // blocks for arbitrary shell quartets (indices into basis()), concatenated
    // row-major; offsets has n+1 entries
    void compute_eri_blocks(std::span<const std::array<int, 4>> quartets, std::vector<Real>& out,
                            std::vector<Index>& offsets);
    [[nodiscard]] std::vector<Real> compute_eri_tensor();
    // The blocks of these shell quartets computed into device memory on
    // stream, with their digestion descriptors; the reference's
    // CudaEngine::compute_eri_batch_device_handle (cuda_engine.hpp:355-356,
    // cuda_engine.cu:1233-1277) for an explicit quartet list.
    using DeviceEriBatch = libaccint::DeviceEriBatch;
    [[nodiscard]] DeviceEriBatch compute_eri_batch_device(std::span<const std::array<int, 4>> quartets,
                                                          hipStream_t stream = nullptr);
    // every shell quartet of a ShellSetQuartet, in its set_a x set_b x set_c x
    // set_d loop order
    [[nodiscard]] DeviceEriBatch compute_eri_batch_device_handle(const ShellSetQuartet& quartet,
                                                                 hipStream_t stream = nullptr);

    // ---- screening ----------------------------------------------------------
    void precompute_schwarz_bounds();
    [[nodiscard]] bool schwarz_bounds_precomputed() const noexcept { return schwarz_.is_initialized(); }
    [[nodiscard]] const screening::SchwarzBounds& schwarz_bounds() const;
    // use externally computed bounds (same packing)
    void set_schwarz_bounds(std::vector<Real> packed);
    void set_density_matrix(const Real* D, Size nbf);
    [[nodiscard]] bool density_matrix_set() const noexcept { return !density_.empty(); }

    // ---- fused J/K (device) -------------------------------------------------
    // Builds (or reuses) the screened AM-sorted worklist for these options.
    // D_host is only read for density-weighted screening or the FP32 split.
    WorklistInfo prepare_worklist(const JKOptions& opts, const Real* D_host = nullptr);
    [[nodiscard]] const WorklistInfo& worklist_info() const noexcept { return wl_info_; }
    // d_J += J[D], d_K += K[D] on device pointers, asynchronous on stream
    void build_jk(const JKOptions& opts, const Real* d_D, Real* d_J, Real* d_K, hipStream_t stream = nullptr,
                  const Real* D_host = nullptr);
    // host convenience: J, K (nbf*nbf each) for the host D
    void build_jk_host(const JKOptions& opts, const Real* D, Real* J, Real* K);
    // capture the J/K build into a hipGraph bound to these device pointers
    void capture_jk_graph(const JKOptions& opts, const Real* d_D, Real* d_J, Real* d_K, hipStream_t stream);
    void launch_jk_graph(hipStream_t stream);

    // ---- consumers -------------------------------------------------------------
    template <typename Consumer>
    void compute_and_consume(const Operator& op, Consumer& consumer, BackendHint hint = BackendHint::Auto) {
        compute_and_consume(op, consumer, screening::ScreeningOptions::none(), hint);
    }

    template <typename Consumer>
    void compute_and_consume(const Operator& op, Consumer& consumer, const screening::ScreeningOptions& options,
                             BackendHint hint = BackendHint::Auto) {
        static_assert(TwoElectronConsumer<Consumer>,
                      "Consumer must provide accumulate(const TwoElectronBuffer<0>&, Index, Index, Index, "
                      "Index, int, int, int, int)");
        (void)hint;
        require_two_electron(op, "compute_and_consume");
        if (options.density_weighted && !density_matrix_set() && !std::is_same_v<Consumer, consumers::FockBuilder> &&
            !std::is_same_v<Consumer, consumers::GpuFockBuilder> &&
            !std::is_same_v<Consumer, consumers::MixedPrecisionFockBuilder>)
            throw InvalidArgumentException(
                "density_weighted screening requires a density matrix; call set_density_matrix() first");
        JKOptions jo;
        jo.schwarz_threshold = options.effective_threshold();
        jo.density_weighted = options.density_weighted;
        if constexpr (std::is_same_v<Consumer, consumers::GpuFockBuilder>) {
            jk_into(consumer, jo);
        } else if constexpr (std::is_same_v<Consumer, consumers::FockBuilder>) {
            jk_into(consumer, jo);
        } else if constexpr (std::is_same_v<Consumer, consumers::MixedPrecisionFockBuilder>) {
            jo.mixed_precision = true;
            jo.fp32_threshold = consumer.fp32_threshold();
            jk_into(consumer, jo);
        } else {
            generic_full(consumer, options);
        }
    }

    template <typename Consumer>
    void compute(const Operator& op, Consumer& consumer, BackendHint hint = BackendHint::Auto) {
        compute_and_consume(op, consumer, hint);
    }

    template <typename Consumer>
    void compute_and_consume_parallel(const Operator& op, Consumer& consumer, int n_threads = 0,
                                      BackendHint hint = BackendHint::Auto) {
        (void)n_threads;
        compute_and_consume(op, consumer, hint);
    }

    template <typename Consumer>
    void compute_and_consume_screened_parallel(const Operator& op, Consumer& consumer,
                                               const screening::ScreeningOptions& options, int n_threads = 0) {
        (void)n_threads;
        compute_and_consume(op, consumer, options);
    }

    // a ShellSetQuartet: every shell quartet of set_a x set_b x set_c x set_d
    // (in that loop order), plain accumulate (cpu_engine.hpp:435-547).  A
    // GpuFockBuilder takes the device route of cuda_engine.hpp:866-896: the
    // ERIs stay in device memory and are digested on the builder's stream;
    // other consumers receive host blocks.
    template <typename Consumer>
    void compute(const Operator& op, const ShellSetQuartet& quartet, Consumer& consumer,
                 BackendHint hint = BackendHint::Auto) {
        (void)hint;
        require_two_electron(op, "compute(ShellSetQuartet)");
        const std::vector<std::array<int, 4>> q = shell_set_quartet_list(quartet);
        if constexpr (std::is_same_v<Consumer, consumers::GpuFockBuilder>) {
            if (consumer.nbf() != basis_->n_basis_functions())
                throw InvalidArgumentException("compute(ShellSetQuartet): GpuFockBuilder nbf does not match the basis");
            // batches of at most kDeviceBatchDoubles ERIs
            Size s = 0;
            while (s < q.size()) {
                Size e = s, cnt = 0;
                while (e < q.size()) {
                    const Size c = block_size(q[e]);
                    if (e > s && cnt + c > kDeviceBatchDoubles) break;
                    cnt += c;
                    ++e;
                }
                DeviceEriBatch batch =
                    compute_eri_batch_device(std::span<const std::array<int, 4>>(q.data() + s, e - s), consumer.stream());
                consumer.accumulate_device_eri_batch(batch, basis_->n_basis_functions());
                // the batch's memory is released on return: its digestion has to finish first
                consumer.synchronize();
                s = e;
            }
        } else {
            consume_blocks(q, consumer, false);
        }
    }

    template <typename Consumer>
    void compute(const Operator& op, std::span<const ShellSetQuartet> quartets, Consumer& consumer,
                 BackendHint hint = BackendHint::Auto) {
        for (const auto& q : quartets) compute(op, q, consumer, hint);
    }
    template <typename Consumer>
    void compute_and_consume(const Operator& op, std::span<const ShellSetQuartet> quartets, Consumer& consumer,
                             BackendHint hint = BackendHint::Auto) {
        compute(op, quartets, consumer, hint);
    }

    // ---- nuclear gradient of E = 1/2 tr(DJ) - x/4 tr(DK) (new capability) --
    // Returns dE/dR, 3 per atom, for the atoms of basis() (atom_index of the
    // shells); n_atoms defaults to 1 + the largest atom index.
    [[nodiscard]] std::vector<Real> compute_jk_gradient(const Real* D, Real exchange_scale = 1.0,
                                                        int n_atoms = -1);

    // ---- dispatch table and autotuner ------------------------------------------
    // Times the compiled kernel strategies, chunk caps and grid caps of every
    // class of the worklist for these options on d_D and keeps the fastest per
    // class; returns the classes with work.  With
    // DispatchConfig::enable_auto_tuning, build_jk does this once by itself
    // (the reference's ProfileOnce mode, dispatch_policy.hpp:75-93).
    DispatchTable autotune(const JKOptions& opts, const Real* d_D, hipStream_t stream = nullptr, int reps = 3,
                           const Real* D_host = nullptr);
    // the current choice of all 55 classes
    [[nodiscard]] DispatchTable dispatch_table();
    // applies the entries (an empty table restores the defaults)
    void set_dispatch_table(const DispatchTable& table);

    // backend handle (for advanced callers)
    [[nodiscard]] lai_basis* device_basis();

private:
    static constexpr Size kDeviceBatchDoubles = Size(1) << 24;
    void require_gpu(const char* what) const;
    static void require_two_electron(const Operator& op, const char* what);
    [[nodiscard]] std::vector<std::array<int, 4>> shell_set_quartet_list(const ShellSetQuartet& quartet) const;
    [[nodiscard]] Size block_size(const std::array<int, 4>& q) const;
    // J/K of a general density: the build of (D + D^T)/2 and, when D is not
    // symmetric, the exchange build of (D - D^T)/2 (K is linear in D; J of an
    // antisymmetric D is zero), i.e. the reference's N^4 accumulate for any D
    void build_jk_general(const JKOptions& o, const Real* D_host, const Real* d_D, Real* d_J, Real* d_K,
                          hipStream_t s);
    void jk_into(consumers::FockBuilder& fb, const JKOptions& o);
    void jk_into(consumers::GpuFockBuilder& fb, const JKOptions& o);
    void jk_into(consumers::MixedPrecisionFockBuilder& fb, const JKOptions& o);
    std::vector<Real> shell_dmax(const Real* D) const;

    template <typename Consumer>
    void consume_blocks(const std::vector<std::array<int, 4>>& q, Consumer& consumer, bool symmetric) {
        constexpr Size kBatch = 1u << 16;
        if constexpr (ParallelConsumer<Consumer>) consumer.prepare_parallel(1);
        TwoElectronBuffer<0> buf;
        std::vector<Real> out;
        std::vector<Index> offs;
        for (Size s = 0; s < q.size(); s += kBatch) {
            const Size n = std::min(kBatch, q.size() - s);
            compute_eri_blocks(std::span<const std::array<int, 4>>(q.data() + s, n), out, offs);
            for (Size t = 0; t < n; ++t) {
                const auto& x = q[s + t];
                const Shell& A = basis_->shell(x[0]);
                const Shell& B = basis_->shell(x[1]);
                const Shell& C = basis_->shell(x[2]);
                const Shell& D = basis_->shell(x[3]);
                buf.resize(A.n_functions(), B.n_functions(), C.n_functions(), D.n_functions());
                std::copy(out.begin() + offs[t], out.begin() + offs[t + 1], buf.data().begin());
                if constexpr (SymmetryAwareConsumer<Consumer>) {
                    if (symmetric) {
                        consumer.accumulate_symmetric(buf, A.function_index(), B.function_index(),
                                                      C.function_index(), D.function_index(), A.n_functions(),
                                                      B.n_functions(), C.n_functions(), D.n_functions(),
                                                      x[0] == x[1], x[2] == x[3], x[0] == x[2] && x[1] == x[3]);
                        continue;
                    }
                }
                consumer.accumulate(buf, A.function_index(), B.function_index(), C.function_index(),
                                    D.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                                    D.n_functions());
            }
        }
        if constexpr (ParallelConsumer<Consumer>) consumer.finalize_parallel();
    }

    template <typename Consumer>
    void generic_full(Consumer& consumer, const screening::ScreeningOptions& options) {
        require_gpu("compute_and_consume");
        const int n = static_cast<int>(basis_->n_shells());
        const Real thr = options.effective_threshold();
        if (thr > 0.0) precompute_schwarz_bounds();
        const bool sym = options.use_permutation_symmetry && SymmetryAwareConsumer<Consumer>;
        std::vector<std::array<int, 4>> q;
        auto pass = [&](int i, int j, int k, int l) {
            return thr <= 0.0 || schwarz_.passes_screening(i, j, k, l, thr);
        };
        if (sym) {
            for (int i = 0; i < n; ++i)
                for (int j = i; j < n; ++j)
                    for (int k = i; k < n; ++k)
                        for (int l = (k == i) ? j : k; l < n; ++l)
                            if (pass(i, j, k, l)) q.push_back({i, j, k, l});
        } else {
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k)
                        for (int l = 0; l < n; ++l)
                            if (pass(i, j, k, l)) q.push_back({i, j, k, l});
        }
        consume_blocks(q, consumer, sym);
    }

    const BasisSet* basis_;
    DispatchConfig config_;
    bool gpu_available_{false};
    lai_basis* dev_{nullptr};
    screening::SchwarzBounds schwarz_;
    std::vector<Real> density_;
    // cached worklist
    bool wl_valid_{false};
    JKOptions wl_opts_{};
    WorklistInfo wl_info_{};
    std::vector<Real> wl_dmax_;
    int tuned_flags_{-1};  // precision flags of the last autotune, -1: none
};

// Several independent molecules through one device basis handle: each J/K
// class kernel covers the quartets of every molecule in one launch, from one
// worklist ordered by angular-momentum class.  D, J and K are the molecules'
// nbf_m x nbf_m matrices one after another (matrix_offsets()).  The engine
// shares ownership of the bases.  No reference counterpart: the reference
// runs one Engine per BasisSet.
class BatchEngine {
public:
    explicit BatchEngine(std::vector<std::shared_ptr<const BasisSet>> bases);
    ~BatchEngine();
    BatchEngine(const BatchEngine&) = delete;
    BatchEngine& operator=(const BatchEngine&) = delete;

    [[nodiscard]] Size n_molecules() const noexcept { return bases_.size(); }
    [[nodiscard]] const BasisSet& basis(Size m) const { return *bases_.at(m); }
    [[nodiscard]] const std::vector<Size>& n_basis_functions() const noexcept { return nbf_; }
    // offsets of the molecules' blocks in the batched matrices, n_molecules() + 1 entries
    [[nodiscard]] const std::vector<Size>& matrix_offsets() const noexcept { return moff_; }
    [[nodiscard]] Size matrix_size() const noexcept { return moff_.back(); }

    void precompute_schwarz_bounds();
    // D_host: the batched density (matrix_size() doubles) for density-dependent options
    WorklistInfo prepare_worklist(const JKOptions& opts, const Real* D_host = nullptr);
    [[nodiscard]] const WorklistInfo& worklist_info() const noexcept { return wl_info_; }
    // d_J += J[D], d_K += K[D] for every molecule (device pointers, batched layout)
    void build_jk(const JKOptions& opts, const Real* d_D, Real* d_J, Real* d_K, hipStream_t stream = nullptr,
                  const Real* D_host = nullptr);
    void build_jk_host(const JKOptions& opts, const Real* D, Real* J, Real* K);
    void capture_jk_graph(const JKOptions& opts, const Real* d_D, Real* d_J, Real* d_K, hipStream_t stream);
    void launch_jk_graph(hipStream_t stream);
    // dispatch table of the batch (as Engine::autotune / dispatch_table /
    // set_dispatch_table, over the classes of every molecule)
    DispatchTable autotune(const JKOptions& opts, const Real* d_D, hipStream_t stream = nullptr, int reps = 3,
                           const Real* D_host = nullptr);
    [[nodiscard]] DispatchTable dispatch_table();
    void set_dispatch_table(const DispatchTable& table);
    [[nodiscard]] lai_basis* device_handle();

private:
    std::vector<Real> shell_dmax(const Real* D) const;
    std::vector<std::shared_ptr<const BasisSet>> bases_;
    std::vector<Size> nbf_, moff_;
    lai_basis* dev_{nullptr};
    bool have_q_{false};
    bool wl_valid_{false};
    JKOptions wl_opts_{};
    WorklistInfo wl_info_{};
};

}  // namespace libaccint
