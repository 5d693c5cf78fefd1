// df.hpp -- density-fitted (resolution-of-identity) J/K on the device.
//
// Mirrors the reference's AuxiliaryBasisSet (include/libaccint/basis/auxiliary_basis_set.hpp)
// and DFFockBuilder (include/libaccint/consumers/df_fock_builder.hpp):
//   J_ab = sum_PQ (ab|P) [V^-1]_PQ (Q|cd) D_cd,  K_ac = sum_P B_P D B_P^T,
// with V the metric of the auxiliary functions, L its Cholesky factor and
// B_P,ab = sum_Q [L^-1]_PQ (Q|ab).  The metric, B and every build run on the
// device through lai_dev_df_create / lai_dev_df_jk (include/lai_device.h); B
// stays in HBM (n_orb^2 n_aux doubles), so the reference's block storage for
// B tensors beyond memory_limit_mb has no counterpart here.
#pragma once

#include <libaccint/basis.hpp>
#include <libaccint/device.hpp>

#include <array>
#include <memory>
#include <span>
#include <string>
#include <vector>

namespace libaccint {

/// Kind of fitting basis (auxiliary_basis_set.hpp:35-43).
enum class FittingType { RI, JKFIT, JFIT, RIFIT };

[[nodiscard]] const char* fitting_type_to_string(FittingType t) noexcept;

/// Shells of an auxiliary (fitting) basis, normalised like orbital shells,
/// with function offsets in shell order.
class AuxiliaryBasisSet {
public:
    AuxiliaryBasisSet() = default;
    explicit AuxiliaryBasisSet(std::vector<Shell> shells, FittingType type = FittingType::JKFIT,
                               std::string name = "");
    AuxiliaryBasisSet(const AuxiliaryBasisSet&) = delete;
    AuxiliaryBasisSet& operator=(const AuxiliaryBasisSet&) = delete;
    AuxiliaryBasisSet(AuxiliaryBasisSet&&) noexcept = default;
    AuxiliaryBasisSet& operator=(AuxiliaryBasisSet&&) noexcept = default;

    [[nodiscard]] Size n_shells() const noexcept { return shells_.size(); }
    [[nodiscard]] Size n_functions() const noexcept { return n_functions_; }
    [[nodiscard]] Size n_primitives() const noexcept { return n_primitives_; }
    [[nodiscard]] int max_angular_momentum() const noexcept { return max_am_; }
    [[nodiscard]] const Shell& shell(Size i) const;
    [[nodiscard]] std::span<const Shell> shells() const noexcept { return shells_; }
    [[nodiscard]] bool empty() const noexcept { return shells_.empty(); }
    [[nodiscard]] Size shell_to_function(Size shell_idx) const;
    [[nodiscard]] Size function_to_shell(Size func_idx) const;
    [[nodiscard]] std::span<const Size> function_offsets() const noexcept { return offsets_; }
    [[nodiscard]] FittingType fitting_type() const noexcept { return type_; }
    [[nodiscard]] const std::string& name() const noexcept { return name_; }

private:
    std::vector<Shell> shells_;
    std::vector<Size> offsets_;
    std::vector<Size> func_shell_;
    Size n_functions_{0}, n_primitives_{0};
    int max_am_{0};
    FittingType type_{FittingType::JKFIT};
    std::string name_;
};

/// Loads a fitting basis from the Basis Set Exchange records shipped with the
/// package (share/basis_sets): "cc-pVDZ-RI" (= cc-pVDZ-RIFIT) and
/// "def2-SVP-RI" (= def2-SVP-RIFIT), any case, with or without the "FIT"
/// suffix.  One shell list per atom, in atom order
/// (auxiliary_basis_set.cpp:224-246).
[[nodiscard]] AuxiliaryBasisSet load_auxiliary_basis(const std::string& basis_name,
                                                     std::span<const int> atomic_numbers,
                                                     std::span<const std::array<Real, 3>> centers);
[[nodiscard]] std::vector<std::string> available_auxiliary_bases();
[[nodiscard]] bool is_auxiliary_basis_available(const std::string& basis_name,
                                                std::span<const int> atomic_numbers);

namespace consumers {

/// df_fock_builder.hpp:37-41
enum class DFMetricType { Coulomb, Overlap, AttenuatedCoulomb };

/// df_fock_builder.hpp:48-80 (same fields and defaults).  use_symmetry,
/// incremental, memory_limit_mb, n_threads and range_separation_omega are
/// kept for source compatibility; the device build always uses the full B.
struct DFFockBuilderConfig {
    Real exchange_fraction{1.0};
    bool compute_coulomb{true};
    bool compute_exchange{true};
    bool use_symmetry{true};
    bool incremental{false};
    Size memory_limit_mb{4096};
    int n_threads{0};
    Real conditioning_threshold{1e8};
    Real conditioning_hard_limit{1e14};
    DFMetricType metric_type{DFMetricType::Coulomb};
    Real range_separation_omega{0.0};
};

class DFFockBuilder {
public:
    /// Non-owning: orbital and auxiliary must outlive the builder.
    DFFockBuilder(const BasisSet& orbital, const AuxiliaryBasisSet& auxiliary, DFFockBuilderConfig config = {});
    /// Takes ownership of the auxiliary basis.
    DFFockBuilder(const BasisSet& orbital, std::unique_ptr<AuxiliaryBasisSet> auxiliary,
                  DFFockBuilderConfig config = {});
    ~DFFockBuilder();
    DFFockBuilder(const DFFockBuilder&) = delete;
    DFFockBuilder& operator=(const DFFockBuilder&) = delete;
    DFFockBuilder(DFFockBuilder&&) noexcept;
    DFFockBuilder& operator=(DFFockBuilder&&) noexcept;

    /// Copies D (n_orb^2, row-major); InvalidArgumentException on a size mismatch.
    void set_density(std::span<const Real> D);
    /// UHF: J from D_alpha + D_beta, K = K_alpha + K_beta.
    void set_density_unrestricted(std::span<const Real> D_alpha, std::span<const Real> D_beta);
    /// Metric, its Cholesky inverse and B on the device (lazy on first compute).
    void initialize();
    [[nodiscard]] bool is_initialized() const noexcept { return df_ != nullptr; }

    /// F = J - exchange_fraction K (with J or K zero when not enabled).
    [[nodiscard]] std::vector<Real> compute();
    [[nodiscard]] std::vector<Real> compute_coulomb();
    [[nodiscard]] std::vector<Real> compute_exchange();
    void compute_accumulate(std::span<Real> F);

    [[nodiscard]] std::span<const Real> coulomb_matrix() const noexcept { return J_; }
    [[nodiscard]] std::span<const Real> exchange_matrix() const noexcept { return K_; }
    [[nodiscard]] std::span<const Real> exchange_matrix_alpha() const noexcept { return K_alpha_; }
    [[nodiscard]] std::span<const Real> exchange_matrix_beta() const noexcept { return K_beta_; }
    [[nodiscard]] bool is_uhf() const noexcept { return uhf_mode_; }
    [[nodiscard]] std::vector<Real> fock_matrix(std::span<const Real> H_core = {}) const;

    [[nodiscard]] Size n_orb() const noexcept { return n_orb_; }
    [[nodiscard]] Size n_aux() const noexcept { return n_aux_; }
    [[nodiscard]] Size b_tensor_memory() const noexcept { return n_orb_ * n_orb_ * n_aux_ * sizeof(Real); }
    [[nodiscard]] const DFFockBuilderConfig& config() const noexcept { return config_; }
    /// max/min of the metric diagonal (after initialize); above
    /// conditioning_threshold the builder is flagged ill-conditioned.
    [[nodiscard]] Real metric_condition_estimate() const noexcept { return cond_; }
    [[nodiscard]] bool metric_ill_conditioned() const noexcept { return cond_ > config_.conditioning_threshold; }

    /// Device handle of the fitted tensor (nullptr before initialize).
    [[nodiscard]] lai_df* device_handle() const noexcept { return df_; }

private:
    void release() noexcept;
    void run(int flags, const std::vector<Real>& D, std::vector<Real>* J, std::vector<Real>* K);
    void compute_j_internal();
    void compute_k_internal();

    const BasisSet* orbital_{nullptr};
    const AuxiliaryBasisSet* auxiliary_{nullptr};
    std::unique_ptr<AuxiliaryBasisSet> owned_auxiliary_;
    DFFockBuilderConfig config_;
    Size n_orb_{0}, n_aux_{0};
    bool uhf_mode_{false};
    std::vector<Real> D_, D_alpha_, D_beta_, J_, K_, K_alpha_, K_beta_;
    Real cond_{0.0};
    lai_basis* basis_{nullptr};
    lai_df* df_{nullptr};
    device::DeviceArray dD_, dJ_, dK_;  // n_orb^2 each, allocated by initialize()
};

/// df_fock_builder.cpp:476-509: without atom data, one s shell of four fixed
/// primitives per atom centre (a placeholder set, as in the reference);
/// InvalidArgumentException for an empty name.
[[nodiscard]] std::unique_ptr<DFFockBuilder> make_df_fock_builder(const BasisSet& orbital,
                                                                  const std::string& aux_basis_name = "");
/// df_fock_builder.cpp:511-531: the named fitting basis on the given atoms
/// (empty name: "cc-pVDZ-RI").
[[nodiscard]] std::unique_ptr<DFFockBuilder> make_df_fock_builder(const BasisSet& orbital,
                                                                  std::span<const int> atomic_numbers,
                                                                  std::span<const std::array<Real, 3>> centers,
                                                                  const std::string& aux_basis_name = "",
                                                                  DFFockBuilderConfig config = {});

}  // namespace consumers
}  // namespace libaccint
