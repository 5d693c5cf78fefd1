// consumers.hpp -- two-electron consumers: concepts, FockBuilder (host J/K)
// and GpuFockBuilder (device-resident J/K/D).
//
//  * Concepts follow include/libaccint/consumers/consumer_concepts.hpp:32-70.
//  * FockBuilder: J_mn += (mn|ls) D_ls, K_ml += (mn|ls) D_ns over full-N^4
//    quartets (accumulate) or canonical quartets with the 8-permutation
//    scatter (accumulate_symmetric); threading strategies Sequential /
//    Atomic / ThreadLocal with the LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES
//    fallback; F = H + J - x K; set_density keeps the caller's pointer
//    (src/shared/consumers/fock_builder.cpp:23-442).
//  * GpuFockBuilder: owns device J, K and D; set_density copies D to the
//    device on the builder's stream; getters copy back and synchronise
//    (include/libaccint/consumers/fock_builder_gpu.hpp:46-159).  Host ERI
//    blocks given to accumulate() are collected on the host and digested on
//    the device in batches (one upload and one kernel per batch); device
//    batches from Engine::compute_eri_batch_device are digested where they
//    are (accumulate_device_eri_batch).
#pragma once

#include <libaccint/buffers.hpp>
#include <libaccint/device.hpp>

#include <concepts>
#include <span>
#include <vector>

namespace libaccint {

template <typename C>
concept TwoElectronConsumer = requires(C& c, const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc,
                                       Index fd, int na, int nb, int nc, int nd) {
    c.accumulate(buf, fa, fb, fc, fd, na, nb, nc, nd);
};
template <typename C>
concept EriConsumer = TwoElectronConsumer<C>;
template <typename C>
concept SymmetryAwareConsumer = TwoElectronConsumer<C> &&
    requires(C& c, const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
             int nc, int nd, bool ij, bool kl, bool bk) {
        c.accumulate_symmetric(buf, fa, fb, fc, fd, na, nb, nc, nd, ij, kl, bk);
    };
template <typename C>
concept ParallelConsumer = TwoElectronConsumer<C> && requires(C& c, int n) {
    c.prepare_parallel(n);
    c.finalize_parallel();
};

namespace consumers {

enum class FockThreadingStrategy { Sequential, Atomic, ThreadLocal };

class FockBuilder {
public:
    explicit FockBuilder(Size nbf);
    void set_density(const Real* D, Size nbf);
    [[nodiscard]] const Real* density() const noexcept { return D_; }
    void accumulate(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                    int nc, int nd);
    void accumulate_symmetric(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na,
                              int nb, int nc, int nd, bool ij_same, bool kl_same, bool braket_same);
    // Adds complete J and K contributions (row-major nbf x nbf); used when the
    // device backend has already digested the integrals.
    void add_jk(const Real* J, const Real* K);
    [[nodiscard]] std::span<const Real> get_coulomb_matrix() const noexcept { return J_; }
    [[nodiscard]] std::span<const Real> get_exchange_matrix() const noexcept { return K_; }
    [[nodiscard]] std::vector<Real> get_fock_matrix(std::span<const Real> H_core, Real exchange_fraction = 1.0) const;
    void reset() noexcept;
    [[nodiscard]] Size nbf() const noexcept { return nbf_; }
    void set_threading_strategy(FockThreadingStrategy s);
    [[nodiscard]] FockThreadingStrategy threading_strategy() const noexcept { return strategy_; }
    void prepare_parallel(int n_threads = 0);
    void finalize_parallel();

private:
    void acc_plain(std::vector<Real>& J, std::vector<Real>& K, const TwoElectronBuffer<0>& buf, Index fa, Index fb,
                   Index fc, Index fd, int na, int nb, int nc, int nd);
    void acc_atomic(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                    int nc, int nd);
    Size nbf_;
    const Real* D_{nullptr};
    std::vector<Real> J_, K_;
    FockThreadingStrategy strategy_{FockThreadingStrategy::Sequential};
    bool auto_thread_local_{false}, auto_atomic_{false};
    int n_threads_{1};
    std::vector<std::vector<Real>> J_tl_, K_tl_;
};

// J and K in FP64 from FP32 or FP64 integral blocks
// (include/libaccint/consumers/mixed_precision_fock_builder.hpp:43-147).  As the
// consumer of Engine::compute_and_consume it receives the device build with
// the FP32 Rys recursion for the quartets whose Q_ij Q_kl max|D| is below
// fp32_threshold() (FP64 roots, weights and accumulation); the statistics
// then count the FP32 and FP64 quartets.
class MixedPrecisionFockBuilder {
public:
    explicit MixedPrecisionFockBuilder(Size nbf);
    void set_density(const Real* D, Size nbf);
    [[nodiscard]] const Real* density() const noexcept { return D_; }
    void accumulate(const TwoElectronBuffer<0, float>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                    int nc, int nd);
    void accumulate(const TwoElectronBuffer<0, double>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                    int nc, int nd);
    // complete J and K of a device build and its quartet counts
    void add_jk(const Real* J, const Real* K, Size n_fp32, Size n_fp64);
    [[nodiscard]] std::span<const Real> get_coulomb_matrix() const noexcept { return J_; }
    [[nodiscard]] std::span<const Real> get_exchange_matrix() const noexcept { return K_; }
    [[nodiscard]] std::vector<Real> get_fock_matrix(std::span<const Real> H_core, Real exchange_fraction = 1.0) const;
    void reset() noexcept;
    [[nodiscard]] Size nbf() const noexcept { return nbf_; }
    [[nodiscard]] MixedPrecisionMode mode() const noexcept { return MixedPrecisionMode::Compute32Accumulate64; }
    [[nodiscard]] Size n_float32_accumulations() const noexcept { return n32_; }
    [[nodiscard]] Size n_float64_accumulations() const noexcept { return n64_; }
    // FP32 selection bound of the device build (default 1e-7)
    void set_fp32_threshold(Real tau) noexcept { tau_ = tau; }
    [[nodiscard]] Real fp32_threshold() const noexcept { return tau_; }

private:
    template <typename R>
    void accumulate_impl(const TwoElectronBuffer<0, R>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                         int nc, int nd);
    Size nbf_;
    const Real* D_{nullptr};
    std::vector<Real> J_, K_;
    Size n32_{0}, n64_{0};
    Real tau_{1e-7};
};

class GpuFockBuilder {
public:
    explicit GpuFockBuilder(Size nbf, hipStream_t stream = nullptr);
    ~GpuFockBuilder() = default;
    GpuFockBuilder(GpuFockBuilder&&) noexcept = default;
    GpuFockBuilder& operator=(GpuFockBuilder&&) noexcept = default;
    GpuFockBuilder(const GpuFockBuilder&) = delete;
    GpuFockBuilder& operator=(const GpuFockBuilder&) = delete;

    // Copies D to the device on the builder's stream (blocks accumulated so
    // far are digested with the previous density first).
    void set_density(const Real* D, Size nbf);
    // host ERI block -> device J/K (plain accumulate semantics,
    // fock_builder_gpu.cu:475-530); the block is copied, so buf may be reused
    // at once.  Blocks are digested in batches on the builder's stream, at the
    // latest by the getters, synchronize(), d_J()/d_K() and set_density().
    void accumulate(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na, int nb,
                    int nc, int nd);
    // device ERI batch -> device J/K, on the builder's stream ordered after the
    // batch's stream (fock_builder_gpu.hpp:137, fock_builder_gpu.cu:600-621)
    void accumulate_device_eri_batch(const DeviceEriBatch& batch, Size nbf);
    [[nodiscard]] std::vector<Real> get_coulomb_matrix() const;
    [[nodiscard]] std::vector<Real> get_exchange_matrix() const;
    [[nodiscard]] std::vector<Real> get_fock_matrix(std::span<const Real> H_core, Real exchange_fraction = 1.0) const;
    // zeroes J and K (blocks not yet digested are dropped with them)
    void reset();
    void synchronize();
    [[nodiscard]] Size nbf() const noexcept { return nbf_; }
    [[nodiscard]] hipStream_t stream() const noexcept { return stream_; }
    // device J and K with every accumulated block digested (enqueued on stream())
    [[nodiscard]] double* d_J();
    [[nodiscard]] double* d_K();
    [[nodiscard]] const double* d_D() const noexcept { return D_.get(); }
    [[nodiscard]] bool density_set() const noexcept { return density_set_; }
    // host copy of the density given to set_density (nbf x nbf)
    [[nodiscard]] const Real* h_D() const noexcept { return density_set_ ? D_host_.data() : nullptr; }
    void prepare_parallel(int = 0) {}
    void finalize_parallel() {}

private:
    // digests the staged host blocks on the device
    void flush() const;
    Size nbf_;
    hipStream_t stream_;
    device::DeviceArray J_, K_, D_;
    std::vector<Real> D_host_;
    bool density_set_{false};
    // host blocks waiting for the device digestion, and their device staging
    mutable std::vector<Real> stage_eri_;
    mutable std::vector<lai_eri_block> stage_blk_;
    mutable device::DeviceArray d_stage_eri_;
    mutable device::DeviceBytes d_stage_blk_;
};

}  // namespace consumers
}  // namespace libaccint
