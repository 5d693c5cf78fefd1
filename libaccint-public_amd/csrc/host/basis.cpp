// basis.cpp -- Shell normalisation and BasisSet organisation (see basis.hpp).
#include <libaccint/basis.hpp>

#include <cmath>
#include <sstream>

namespace libaccint {

namespace {

constexpr Real kPi = 3.14159265358979323846;

int dfact_odd(int n) {  // (2n-1)!!
    int r = 1;
    for (int i = 1; i <= n; ++i) r *= (2 * i - 1);
    return r;
}

void validate(int am, const std::vector<Real>& e, const std::vector<Real>& c) {
    if (am < 0 || am > MAX_ANGULAR_MOMENTUM) {
        std::ostringstream o;
        o << "Shell: angular momentum " << am << " is outside valid range [0, " << MAX_ANGULAR_MOMENTUM << "]";
        throw InvalidArgumentException(o.str());
    }
    if (e.empty()) throw InvalidArgumentException("Shell: must have at least one primitive Gaussian");
    if (e.size() != c.size()) {
        std::ostringstream o;
        o << "Shell: exponents size (" << e.size() << ") does not match coefficients size (" << c.size() << ")";
        throw InvalidArgumentException(o.str());
    }
    for (Size i = 0; i < e.size(); ++i)
        if (!(e[i] > 0.0)) {
            std::ostringstream o;
            o << "Shell: exponent[" << i << "] = " << e[i] << " is not positive";
            throw InvalidArgumentException(o.str());
        }
}

// primitive normalisation, then scale the contracted (L,0,0) self-overlap to 1
std::vector<Real> normalize(int am, const std::vector<Real>& e, const std::vector<Real>& c) {
    const Size n = e.size();
    std::vector<Real> out(c);
    for (Size i = 0; i < n; ++i) {
        const Real pn = std::pow(2.0 * e[i] / kPi, 0.75) * std::pow(4.0 * e[i], am * 0.5) /
                        std::sqrt(static_cast<Real>(dfact_odd(am)));
        out[i] *= pn;
    }
    Real so = 0.0;
    for (Size i = 0; i < n; ++i)
        for (Size j = 0; j < n; ++j) {
            const Real p = e[i] + e[j];
            Real ov = std::pow(kPi / p, 1.5);
            for (int k = 0; k < am; ++k) ov *= 0.5 / p;
            so += out[i] * out[j] * ov;
        }
    const Real s = 1.0 / std::sqrt(so);
    for (auto& x : out) x *= s;
    return out;
}

}  // namespace

Shell::Shell(int am, Point3D center, std::vector<Real> exponents, std::vector<Real> coefficients)
    : am_(am), center_(center), exponents_(std::move(exponents)) {
    validate(am_, exponents_, coefficients);
    coefficients_ = normalize(am_, exponents_, coefficients);
}

Shell::Shell(PreNormalizedTag, int am, Point3D center, std::vector<Real> exponents,
             std::vector<Real> coefficients)
    : am_(am), center_(center), exponents_(std::move(exponents)), coefficients_(std::move(coefficients)) {
    validate(am_, exponents_, coefficients_);
}

Real Shell::exponent(Size i) const {
    if (i >= exponents_.size())
        throw InvalidArgumentException("Shell::exponent: index " + std::to_string(i) + " is out of bounds");
    return exponents_[i];
}

Real Shell::coefficient(Size i) const {
    if (i >= coefficients_.size())
        throw InvalidArgumentException("Shell::coefficient: index " + std::to_string(i) + " is out of bounds");
    return coefficients_[i];
}

// ---------------------------------------------------------------------------
void ShellSet::add_shell(const Shell& s) {
    if (s.angular_momentum() != am_ || static_cast<int>(s.n_primitives()) != n_primitives_)
        throw InvalidArgumentException("ShellSet::add_shell: shell does not match the set key");
    shells_.push_back(s);
}

const Shell& ShellSet::shell(Size i) const {
    if (i >= shells_.size()) throw InvalidArgumentException("ShellSet::shell: index out of range");
    return shells_[i];
}

ShellSetDataSoA ShellSet::soa_data() const {
    ShellSetDataSoA d;
    for (const auto& s : shells_) {
        d.center_x.push_back(s.center().x);
        d.center_y.push_back(s.center().y);
        d.center_z.push_back(s.center().z);
        for (Size k = 0; k < s.n_primitives(); ++k) {
            d.exponents.push_back(s.exponents()[k]);
            d.coefficients.push_back(s.coefficients()[k]);
        }
        d.shell_indices.push_back(s.shell_index());
        d.atom_indices.push_back(s.atom_index());
        d.function_offsets.push_back(s.function_index());
    }
    return d;
}

// ---------------------------------------------------------------------------
BasisSet::BasisSet(std::vector<Shell> shells) : shells_(std::move(shells)) {
    if (shells_.empty()) return;
    assign_indices();
    organize();
}

BasisSet::BasisSet(BasisSet&& o) noexcept { *this = std::move(o); }

BasisSet& BasisSet::operator=(BasisSet&& o) noexcept {
    if (this != &o) {
        shells_ = std::move(o.shells_);
        sets_ = std::move(o.sets_);
        set_index_ = std::move(o.set_index_);
        pairs_.clear();
        quartets_.clear();
        pairs_done_ = quartets_done_ = false;  // views point into sets_, rebuilt lazily
        nbf_ = o.nbf_;
        nbf_sph_ = o.nbf_sph_;
        max_am_ = o.max_am_;
        max_prim_ = o.max_prim_;
        mtx_ = std::make_unique<std::mutex>();
    }
    return *this;
}

void BasisSet::assign_indices() {
    Index off = 0, off_sph = 0;
    max_am_ = 0;
    max_prim_ = 0;
    for (Size i = 0; i < shells_.size(); ++i) {
        shells_[i].set_shell_index(static_cast<Index>(i));
        shells_[i].set_function_index(off);
        off += shells_[i].n_functions();
        off_sph += n_spherical(shells_[i].angular_momentum());
        max_am_ = std::max(max_am_, shells_[i].angular_momentum());
        max_prim_ = std::max(max_prim_, static_cast<int>(shells_[i].n_primitives()));
    }
    nbf_ = static_cast<Size>(off);
    nbf_sph_ = static_cast<Size>(off_sph);
}

void BasisSet::organize() {
    sets_.clear();
    set_index_.clear();
    for (const auto& s : shells_) {
        ShellSetKey k{s.angular_momentum(), static_cast<int>(s.n_primitives())};
        auto it = set_index_.find(k);
        if (it == set_index_.end()) {
            set_index_[k] = sets_.size();
            sets_.push_back(std::make_unique<ShellSet>(k.angular_momentum, k.n_primitives));
            sets_.back()->add_shell(s);
        } else {
            sets_[it->second]->add_shell(s);
        }
    }
}

const Shell& BasisSet::shell(Size i) const {
    if (i >= shells_.size())
        throw InvalidArgumentException("BasisSet::shell: index " + std::to_string(i) + " out of range [0, " +
                                       std::to_string(shells_.size()) + ")");
    return shells_[i];
}

std::vector<const ShellSet*> BasisSet::shell_sets() const {
    std::vector<const ShellSet*> r;
    for (const auto& p : sets_) r.push_back(p.get());
    return r;
}

const ShellSet* BasisSet::shell_set(int am, int k) const {
    auto it = set_index_.find({am, k});
    return it == set_index_.end() ? nullptr : sets_[it->second].get();
}

std::vector<const ShellSet*> BasisSet::shell_sets_with_am(int am) const {
    std::vector<const ShellSet*> r;
    for (const auto& p : sets_)
        if (p->angular_momentum() == am) r.push_back(p.get());
    return r;
}

const std::vector<ShellSetPair>& BasisSet::shell_set_pairs() const {
    std::lock_guard<std::mutex> lk(*mtx_);
    if (!pairs_done_) {
        pairs_.clear();
        const Size n = sets_.size();
        pairs_.reserve(n * (n + 1) / 2);
        for (Size i = 0; i < n; ++i)
            for (Size j = i; j < n; ++j) pairs_.emplace_back(*sets_[i], *sets_[j]);
        pairs_done_ = true;
    }
    return pairs_;
}

const std::vector<ShellSetQuartet>& BasisSet::shell_set_quartets() const {
    const auto& pr = shell_set_pairs();
    std::lock_guard<std::mutex> lk(*mtx_);
    if (!quartets_done_) {
        quartets_.clear();
        const Size n = pr.size();
        quartets_.reserve(n * (n + 1) / 2);
        for (Size i = 0; i < n; ++i)
            for (Size j = i; j < n; ++j) quartets_.emplace_back(pr[i], pr[j]);
        quartets_done_ = true;
    }
    return quartets_;
}

std::vector<const Shell*> BasisSet::shells_on_atom(Index atom) const {
    std::vector<const Shell*> r;
    for (const auto& s : shells_)
        if (s.atom_index() == atom) r.push_back(&s);
    return r;
}

}  // namespace libaccint
