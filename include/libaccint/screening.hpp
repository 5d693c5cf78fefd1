// screening.hpp -- Schwarz screening vocabulary of the reference
// (include/libaccint/screening/screening_options.hpp:16-160,
//  include/libaccint/screening/schwarz_bounds.hpp:40-124).
// The bounds themselves are computed by the gfx950 backend (lai_dev_schwarz).
#pragma once

#include <libaccint/core.hpp>

#include <string>
#include <vector>

namespace libaccint::screening {

enum class ScreeningPreset { None, Loose, Normal, Tight, Custom };

struct ScreeningOptions {
    Real threshold = 1e-12;
    bool enabled = true;
    bool density_weighted = false;
    bool use_permutation_symmetry = false;
    bool enable_statistics = false;

    [[nodiscard]] static ScreeningOptions none() { return ScreeningOptions{0.0, false}; }
    [[nodiscard]] static ScreeningOptions loose() { return ScreeningOptions{1e-10, true}; }
    [[nodiscard]] static ScreeningOptions normal() { return ScreeningOptions{1e-12, true}; }
    [[nodiscard]] static ScreeningOptions tight() { return ScreeningOptions{1e-14, true}; }
    [[nodiscard]] static ScreeningOptions from_preset(ScreeningPreset p) {
        switch (p) {
            case ScreeningPreset::None: return none();
            case ScreeningPreset::Loose: return loose();
            case ScreeningPreset::Tight: return tight();
            default: return normal();
        }
    }
    // 0 when disabled, negative thresholds clamp to 0 (screening_options.hpp:138-141)
    [[nodiscard]] Real effective_threshold() const noexcept {
        if (!enabled) return 0.0;
        return threshold < 0.0 ? 0.0 : threshold;
    }
    // passes iff bound >= effective threshold (screening_options.hpp:147-149)
    [[nodiscard]] bool passes_screening(Real schwarz_product) const noexcept {
        return schwarz_product >= effective_threshold();
    }
    [[nodiscard]] std::string preset_name() const {
        if (!enabled) return "None";
        if (threshold == 1e-10) return "Loose";
        if (threshold == 1e-12) return "Normal";
        if (threshold == 1e-14) return "Tight";
        return "Custom";
    }
};

// Q_ij = sqrt(max |(ij|ij)|), packed lower-triangular at j*(j+1)/2 + i.
class SchwarzBounds {
public:
    SchwarzBounds() = default;
    SchwarzBounds(Size n_shells, std::vector<Real> packed) : n_(n_shells), q_(std::move(packed)) {
        for (Real v : q_) max_ = v > max_ ? v : max_;
    }
    [[nodiscard]] Real operator()(Size i, Size j) const noexcept {
        if (i > j) std::swap(i, j);
        return q_[j * (j + 1) / 2 + i];
    }
    [[nodiscard]] Size n_shells() const noexcept { return n_; }
    [[nodiscard]] bool passes_screening(Size i, Size j, Size k, Size l, Real thr) const noexcept {
        return (*this)(i, j) * (*this)(k, l) >= thr;
    }
    [[nodiscard]] Real quartet_bound(Size i, Size j, Size k, Size l) const noexcept {
        return (*this)(i, j) * (*this)(k, l);
    }
    [[nodiscard]] Real max_bound() const noexcept { return max_; }
    [[nodiscard]] bool is_initialized() const noexcept { return n_ > 0; }
    [[nodiscard]] const std::vector<Real>& packed() const noexcept { return q_; }
    // canonical quartets (i<=j, k<=l, ij<=kl) passing Q_ij*Q_kl >= thr
    // (schwarz_bounds.cpp:62-99)
    [[nodiscard]] Size count_passing_quartets(Real thr) const;
    [[nodiscard]] Real estimate_pass_fraction(Real thr) const;

private:
    Size n_{0};
    std::vector<Real> q_;
    Real max_{0.0};
};

}  // namespace libaccint::screening
