// operators.hpp -- the operators of the Fock-build path: Coulomb (two-electron)
// and overlap / kinetic / nuclear attraction (one-electron), with the
// reference's names (include/libaccint/operators/operator_types.hpp:22-60,
// include/libaccint/operators/operator.hpp:17-80,
// include/libaccint/operators/one_electron_operator.hpp).
#pragma once

#include <libaccint/core.hpp>

#include <string_view>
#include <vector>

namespace libaccint {

enum class OperatorKind { Overlap, Kinetic, Nuclear, PointCharge, Coulomb };

[[nodiscard]] constexpr std::string_view to_string(OperatorKind k) noexcept {
    switch (k) {
        case OperatorKind::Overlap: return "Overlap";
        case OperatorKind::Kinetic: return "Kinetic";
        case OperatorKind::Nuclear: return "Nuclear";
        case OperatorKind::PointCharge: return "PointCharge";
        case OperatorKind::Coulomb: return "Coulomb";
    }
    return "Unknown";
}

[[nodiscard]] constexpr bool is_one_electron(OperatorKind k) noexcept { return k != OperatorKind::Coulomb; }
[[nodiscard]] constexpr bool is_two_electron(OperatorKind k) noexcept { return k == OperatorKind::Coulomb; }

struct PointChargeParams {
    std::vector<Real> x, y, z, charge;
    [[nodiscard]] Size n_centers() const noexcept { return charge.size(); }
};

class Operator {
public:
    [[nodiscard]] static Operator coulomb() { return Operator(OperatorKind::Coulomb); }
    [[nodiscard]] static Operator overlap() { return Operator(OperatorKind::Overlap); }
    [[nodiscard]] static Operator kinetic() { return Operator(OperatorKind::Kinetic); }
    [[nodiscard]] static Operator nuclear(PointChargeParams p) {
        Operator o(OperatorKind::Nuclear);
        o.charges_ = std::move(p);
        return o;
    }
    [[nodiscard]] static Operator point_charge(PointChargeParams p) {
        Operator o(OperatorKind::PointCharge);
        o.charges_ = std::move(p);
        return o;
    }
    [[nodiscard]] OperatorKind kind() const noexcept { return kind_; }
    [[nodiscard]] bool is_one_electron() const noexcept { return libaccint::is_one_electron(kind_); }
    [[nodiscard]] bool is_two_electron() const noexcept { return libaccint::is_two_electron(kind_); }
    [[nodiscard]] const PointChargeParams& point_charges() const noexcept { return charges_; }

private:
    explicit Operator(OperatorKind k) : kind_(k) {}
    OperatorKind kind_;
    PointChargeParams charges_;
};

// Sum of one-electron operators with scale factors (H_core = T + V).
class OneElectronOperator {
public:
    struct Term {
        Operator op;
        Real scale;
    };
    explicit OneElectronOperator(const Operator& op, Real scale = 1.0) { add(op, scale); }
    OneElectronOperator& add(const Operator& op, Real scale = 1.0) {
        if (!op.is_one_electron())
            throw InvalidArgumentException("OneElectronOperator: operator is not one-electron");
        terms_.push_back({op, scale});
        return *this;
    }
    [[nodiscard]] const std::vector<Term>& terms() const noexcept { return terms_; }

private:
    std::vector<Term> terms_;
};

}  // namespace libaccint
