// basis.hpp -- Shell, ShellSet, ShellSetPair, ShellSetQuartet and BasisSet.
//
// Semantics follow the reference basis model:
//  * Shell normalises its contraction coefficients on construction:
//    primitive norm (2a/pi)^(3/4) (4a)^(l/2) / sqrt((2l-1)!!), then the
//    contracted (L,0,0) self-overlap is scaled to 1
//    (src/shared/basis/shell.cpp:32-113).  Zero coefficients are kept.
//  * BasisSet assigns function offsets in shell order and groups shells into
//    ShellSets keyed by (L, n_primitives) in first-seen order
//    (src/shared/basis/basis_set.cpp:161-211).
//  * shell_set_pairs()/shell_set_quartets() are the upper-triangle lists of
//    src/shared/basis/basis_set.cpp:80-141.
#pragma once

#include <libaccint/core.hpp>

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <span>
#include <vector>

namespace libaccint {

class Shell {
public:
    struct PreNormalizedTag {};
    Shell(int am, Point3D center, std::vector<Real> exponents, std::vector<Real> coefficients);
    Shell(AngularMomentum am, Point3D center, std::vector<Real> exponents,
          std::vector<Real> coefficients)
        : Shell(to_int(am), center, std::move(exponents), std::move(coefficients)) {}
    Shell(PreNormalizedTag, int am, Point3D center, std::vector<Real> exponents,
          std::vector<Real> coefficients);

    [[nodiscard]] int angular_momentum() const noexcept { return am_; }
    [[nodiscard]] int n_functions() const noexcept { return n_cartesian(am_); }
    [[nodiscard]] Size n_primitives() const noexcept { return exponents_.size(); }
    [[nodiscard]] const Point3D& center() const noexcept { return center_; }
    [[nodiscard]] std::span<const Real> exponents() const noexcept { return exponents_; }
    [[nodiscard]] std::span<const Real> coefficients() const noexcept { return coefficients_; }
    [[nodiscard]] Real exponent(Size i) const;
    [[nodiscard]] Real coefficient(Size i) const;
    [[nodiscard]] Index shell_index() const noexcept { return shell_index_; }
    [[nodiscard]] Index function_index() const noexcept { return function_index_; }
    [[nodiscard]] Index atom_index() const noexcept { return atom_index_; }
    void set_shell_index(Index i) noexcept { shell_index_ = i; }
    void set_function_index(Index i) noexcept { function_index_ = i; }
    void set_atom_index(Index i) noexcept { atom_index_ = i; }
    void set_center(Point3D c) noexcept { center_ = c; }

private:
    int am_{0};
    Point3D center_{};
    std::vector<Real> exponents_;
    std::vector<Real> coefficients_;
    Index shell_index_{-1};
    Index function_index_{-1};
    Index atom_index_{-1};
};

struct ShellSetKey {
    int angular_momentum{0};
    int n_primitives{0};
    constexpr ShellSetKey() = default;
    constexpr ShellSetKey(int am, int k) : angular_momentum(am), n_primitives(k) {}
    constexpr bool operator==(const ShellSetKey&) const noexcept = default;
    constexpr bool operator<(const ShellSetKey& o) const noexcept {
        return angular_momentum != o.angular_momentum ? angular_momentum < o.angular_momentum
                                                      : n_primitives < o.n_primitives;
    }
};

// SoA view of a ShellSet (include/libaccint/basis/shell_set.hpp:78-98):
// exponents/coefficients flat [i*K + k].
struct ShellSetDataSoA {
    std::vector<Real> center_x, center_y, center_z;
    std::vector<Real> exponents, coefficients;
    std::vector<Index> shell_indices, atom_indices, function_offsets;
    [[nodiscard]] Size n_shells() const noexcept { return center_x.size(); }
    [[nodiscard]] Size n_total_primitives() const noexcept { return exponents.size(); }
};

class ShellSet {
public:
    ShellSet() = default;
    ShellSet(int am, int n_primitives) : am_(am), n_primitives_(n_primitives) {}
    void add_shell(const Shell& s);
    [[nodiscard]] int angular_momentum() const noexcept { return am_; }
    [[nodiscard]] int n_primitives_per_shell() const noexcept { return n_primitives_; }
    [[nodiscard]] Size n_shells() const noexcept { return shells_.size(); }
    [[nodiscard]] int n_functions_per_shell() const noexcept { return n_cartesian(am_); }
    [[nodiscard]] Size n_total_functions() const noexcept {
        return n_shells() * static_cast<Size>(n_functions_per_shell());
    }
    [[nodiscard]] const Shell& shell(Size i) const;
    [[nodiscard]] std::span<const Shell> shells() const noexcept { return shells_; }
    [[nodiscard]] ShellSetKey key() const noexcept { return {am_, n_primitives_}; }
    [[nodiscard]] bool empty() const noexcept { return shells_.empty(); }
    [[nodiscard]] ShellSetDataSoA soa_data() const;

private:
    int am_{0};
    int n_primitives_{0};
    std::vector<Shell> shells_;
};

class ShellSetPair {
public:
    ShellSetPair(const ShellSet& a, const ShellSet& b) : a_(&a), b_(&b) {}
    [[nodiscard]] const ShellSet& shell_set_a() const noexcept { return *a_; }
    [[nodiscard]] const ShellSet& shell_set_b() const noexcept { return *b_; }
    [[nodiscard]] int La() const noexcept { return a_->angular_momentum(); }
    [[nodiscard]] int Lb() const noexcept { return b_->angular_momentum(); }
    [[nodiscard]] int L_total() const noexcept { return La() + Lb(); }
    [[nodiscard]] Size n_pairs() const noexcept { return a_->n_shells() * b_->n_shells(); }

private:
    const ShellSet* a_;
    const ShellSet* b_;
};

class ShellSetQuartet {
public:
    ShellSetQuartet(const ShellSetPair& bra, const ShellSetPair& ket) noexcept
        : bra_(&bra), ket_(&ket) {}
    [[nodiscard]] const ShellSetPair& bra_pair() const noexcept { return *bra_; }
    [[nodiscard]] const ShellSetPair& ket_pair() const noexcept { return *ket_; }
    [[nodiscard]] int La() const noexcept { return bra_->La(); }
    [[nodiscard]] int Lb() const noexcept { return bra_->Lb(); }
    [[nodiscard]] int Lc() const noexcept { return ket_->La(); }
    [[nodiscard]] int Ld() const noexcept { return ket_->Lb(); }
    [[nodiscard]] int L_total() const noexcept { return La() + Lb() + Lc() + Ld(); }
    [[nodiscard]] Size n_quartets() const noexcept { return bra_->n_pairs() * ket_->n_pairs(); }

private:
    const ShellSetPair* bra_;
    const ShellSetPair* ket_;
};

class BasisSet {
public:
    BasisSet() = default;
    explicit BasisSet(std::vector<Shell> shells);
    BasisSet(const BasisSet&) = delete;
    BasisSet& operator=(const BasisSet&) = delete;
    BasisSet(BasisSet&&) noexcept;
    BasisSet& operator=(BasisSet&&) noexcept;

    [[nodiscard]] Size n_shells() const noexcept { return shells_.size(); }
    [[nodiscard]] Size n_basis_functions() const noexcept { return nbf_; }
    [[nodiscard]] Size n_basis_functions_spherical() const noexcept { return nbf_sph_; }
    [[nodiscard]] int max_angular_momentum() const noexcept { return max_am_; }
    [[nodiscard]] int max_n_primitives() const noexcept { return max_prim_; }
    [[nodiscard]] const Shell& shell(Size i) const;
    [[nodiscard]] std::span<const Shell> shells() const noexcept { return shells_; }
    [[nodiscard]] std::vector<const ShellSet*> shell_sets() const;
    [[nodiscard]] Size n_shell_sets() const noexcept { return sets_.size(); }
    [[nodiscard]] const ShellSet* shell_set(int am, int n_primitives) const;
    [[nodiscard]] std::vector<const ShellSet*> shell_sets_with_am(int am) const;
    [[nodiscard]] const std::vector<ShellSetPair>& shell_set_pairs() const;
    [[nodiscard]] const std::vector<ShellSetQuartet>& shell_set_quartets() const;
    [[nodiscard]] std::vector<const Shell*> shells_on_atom(Index atom) const;
    // Cartesian AO basis (the spherical flag of the reference is never
    // consumed on the J/K path, include/libaccint/basis/basis_set.hpp:236-249)
    [[nodiscard]] bool is_spherical() const noexcept { return false; }

private:
    void assign_indices();
    void organize();
    std::vector<Shell> shells_;
    std::vector<std::unique_ptr<ShellSet>> sets_;
    std::map<ShellSetKey, Size> set_index_;
    mutable std::vector<ShellSetPair> pairs_;
    mutable std::vector<ShellSetQuartet> quartets_;
    mutable bool pairs_done_{false}, quartets_done_{false};
    mutable std::unique_ptr<std::mutex> mtx_{std::make_unique<std::mutex>()};
    Size nbf_{0}, nbf_sph_{0};
    int max_am_{0}, max_prim_{0};
};

}  // namespace libaccint
