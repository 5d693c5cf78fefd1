// df.cpp -- AuxiliaryBasisSet, the fitting-basis loader and the device
// DFFockBuilder (include/libaccint/df.hpp).
#include <libaccint/data.hpp>
#include <libaccint/df.hpp>

#include <algorithm>
#include <cctype>
#include <filesystem>
#include <map>

namespace libaccint {

const char* fitting_type_to_string(FittingType t) noexcept {
    switch (t) {
        case FittingType::RI: return "RI";
        case FittingType::JKFIT: return "JKFIT";
        case FittingType::JFIT: return "JFIT";
        case FittingType::RIFIT: return "RIFIT";
    }
    return "unknown";
}

AuxiliaryBasisSet::AuxiliaryBasisSet(std::vector<Shell> shells, FittingType type, std::string name)
    : shells_(std::move(shells)), type_(type), name_(std::move(name)) {
    Size f = 0;
    for (Size s = 0; s < shells_.size(); ++s) {
        offsets_.push_back(f);
        shells_[s].set_shell_index(static_cast<Index>(s));
        shells_[s].set_function_index(static_cast<Index>(f));
        for (int i = 0; i < shells_[s].n_functions(); ++i) func_shell_.push_back(s);
        f += static_cast<Size>(shells_[s].n_functions());
        n_primitives_ += shells_[s].n_primitives();
        max_am_ = std::max(max_am_, shells_[s].angular_momentum());
    }
    n_functions_ = f;
}

const Shell& AuxiliaryBasisSet::shell(Size i) const {
    if (i >= shells_.size())
        throw InvalidArgumentException("AuxiliaryBasisSet::shell: index " + std::to_string(i) + " out of range");
    return shells_[i];
}

Size AuxiliaryBasisSet::shell_to_function(Size s) const {
    if (s >= offsets_.size())
        throw InvalidArgumentException("AuxiliaryBasisSet::shell_to_function: index out of range");
    return offsets_[s];
}

Size AuxiliaryBasisSet::function_to_shell(Size f) const {
    if (f >= func_shell_.size())
        throw InvalidArgumentException("AuxiliaryBasisSet::function_to_shell: index out of range");
    return func_shell_[f];
}

namespace {
std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
    return s;
}

// fitting-basis name -> shipped BSE record, or "" when unknown
std::string aux_file(const std::string& name) {
    const std::string n = lower(name);
    if (n == "cc-pvdz-ri" || n == "cc-pvdz-rifit") return "cc-pvdz-rifit";
    if (n == "def2-svp-ri" || n == "def2-svp-rifit") return "def2-svp-rifit";
    return "";
}
}  // namespace

AuxiliaryBasisSet load_auxiliary_basis(const std::string& basis_name, std::span<const int> atomic_numbers,
                                       std::span<const std::array<Real, 3>> centers) {
    if (atomic_numbers.size() != centers.size())
        throw InvalidArgumentException("Mismatched sizes: atomic_numbers(" + std::to_string(atomic_numbers.size()) +
                                       ") vs centers(" + std::to_string(centers.size()) + ")");
    const std::string file = aux_file(basis_name);
    if (file.empty())
        throw InvalidArgumentException("Unknown auxiliary basis set: '" + basis_name +
                                       "'. Available: cc-pVDZ-RI, def2-SVP-RI");
    std::vector<data::Atom> atoms;
    for (Size i = 0; i < atomic_numbers.size(); ++i)
        atoms.push_back(data::Atom{atomic_numbers[i], Point3D{centers[i][0], centers[i][1], centers[i][2]}});
    BasisSet b = data::load_basis_set(file, atoms);
    std::vector<Shell> shells(b.shells().begin(), b.shells().end());
    return AuxiliaryBasisSet(std::move(shells), FittingType::RI, basis_name);
}

std::vector<std::string> available_auxiliary_bases() { return {"cc-pVDZ-RI", "def2-SVP-RI"}; }

bool is_auxiliary_basis_available(const std::string& basis_name, std::span<const int> atomic_numbers) {
    if (aux_file(basis_name).empty()) return false;
    std::vector<std::array<Real, 3>> c(atomic_numbers.size(), std::array<Real, 3>{0.0, 0.0, 0.0});
    try {
        (void)load_auxiliary_basis(basis_name, atomic_numbers, c);
    } catch (const std::exception&) {
        return false;
    }
    return true;
}

namespace consumers {

DFFockBuilder::DFFockBuilder(const BasisSet& orbital, const AuxiliaryBasisSet& auxiliary, DFFockBuilderConfig config)
    : orbital_(&orbital),
      auxiliary_(&auxiliary),
      config_(config),
      n_orb_(orbital.n_basis_functions()),
      n_aux_(auxiliary.n_functions()) {
    D_.assign(n_orb_ * n_orb_, 0.0);
    J_.assign(n_orb_ * n_orb_, 0.0);
    K_.assign(n_orb_ * n_orb_, 0.0);
}

DFFockBuilder::DFFockBuilder(const BasisSet& orbital, std::unique_ptr<AuxiliaryBasisSet> auxiliary,
                             DFFockBuilderConfig config)
    : orbital_(&orbital),
      auxiliary_(auxiliary.get()),
      owned_auxiliary_(std::move(auxiliary)),
      config_(config),
      n_orb_(orbital.n_basis_functions()),
      n_aux_(owned_auxiliary_ ? owned_auxiliary_->n_functions() : 0) {
    D_.assign(n_orb_ * n_orb_, 0.0);
    J_.assign(n_orb_ * n_orb_, 0.0);
    K_.assign(n_orb_ * n_orb_, 0.0);
}

DFFockBuilder::~DFFockBuilder() { release(); }

DFFockBuilder::DFFockBuilder(DFFockBuilder&& o) noexcept
    : orbital_(o.orbital_),
      auxiliary_(o.auxiliary_),
      owned_auxiliary_(std::move(o.owned_auxiliary_)),
      config_(o.config_),
      n_orb_(o.n_orb_),
      n_aux_(o.n_aux_),
      uhf_mode_(o.uhf_mode_),
      D_(std::move(o.D_)),
      D_alpha_(std::move(o.D_alpha_)),
      D_beta_(std::move(o.D_beta_)),
      J_(std::move(o.J_)),
      K_(std::move(o.K_)),
      K_alpha_(std::move(o.K_alpha_)),
      K_beta_(std::move(o.K_beta_)),
      cond_(o.cond_),
      basis_(o.basis_),
      df_(o.df_),
      dD_(std::move(o.dD_)),
      dJ_(std::move(o.dJ_)),
      dK_(std::move(o.dK_)) {
    o.basis_ = nullptr;
    o.df_ = nullptr;
}

DFFockBuilder& DFFockBuilder::operator=(DFFockBuilder&& o) noexcept {
    if (this != &o) {
        release();
        orbital_ = o.orbital_;
        auxiliary_ = o.auxiliary_;
        owned_auxiliary_ = std::move(o.owned_auxiliary_);
        config_ = o.config_;
        n_orb_ = o.n_orb_;
        n_aux_ = o.n_aux_;
        uhf_mode_ = o.uhf_mode_;
        D_ = std::move(o.D_);
        D_alpha_ = std::move(o.D_alpha_);
        D_beta_ = std::move(o.D_beta_);
        J_ = std::move(o.J_);
        K_ = std::move(o.K_);
        K_alpha_ = std::move(o.K_alpha_);
        K_beta_ = std::move(o.K_beta_);
        cond_ = o.cond_;
        basis_ = o.basis_;
        df_ = o.df_;
        dD_ = std::move(o.dD_);
        dJ_ = std::move(o.dJ_);
        dK_ = std::move(o.dK_);
        o.basis_ = nullptr;
        o.df_ = nullptr;
    }
    return *this;
}

void DFFockBuilder::release() noexcept {
    if (df_) (void)lai_dev_df_free(df_);
    if (basis_) (void)lai_dev_basis_free(basis_);
    df_ = nullptr;
    basis_ = nullptr;
}

void DFFockBuilder::set_density(std::span<const Real> D) {
    if (D.size() != n_orb_ * n_orb_)
        throw InvalidArgumentException("Density matrix size mismatch: expected " + std::to_string(n_orb_ * n_orb_) +
                                       ", got " + std::to_string(D.size()));
    uhf_mode_ = false;
    std::copy(D.begin(), D.end(), D_.begin());
}

void DFFockBuilder::set_density_unrestricted(std::span<const Real> Da, std::span<const Real> Db) {
    if (Da.size() != n_orb_ * n_orb_ || Db.size() != n_orb_ * n_orb_)
        throw InvalidArgumentException("Density matrix size mismatch");
    uhf_mode_ = true;
    D_alpha_.assign(Da.begin(), Da.end());
    D_beta_.assign(Db.begin(), Db.end());
    for (Size i = 0; i < n_orb_ * n_orb_; ++i) D_[i] = Da[i] + Db[i];
    K_alpha_.assign(n_orb_ * n_orb_, 0.0);
    K_beta_.assign(n_orb_ * n_orb_, 0.0);
}

void DFFockBuilder::initialize() {
    if (df_) return;
    if (!auxiliary_ || auxiliary_->empty()) throw InvalidStateException("DFFockBuilder: empty auxiliary basis");
    // one handle: orbital shells, auxiliary shells, the unit s function
    std::vector<int> L, np;
    std::vector<double> ctr, ex, co;
    auto add = [&](const Shell& s) {
        L.push_back(s.angular_momentum());
        np.push_back(static_cast<int>(s.n_primitives()));
        ctr.insert(ctr.end(), {s.center().x, s.center().y, s.center().z});
        ex.insert(ex.end(), s.exponents().begin(), s.exponents().end());
        co.insert(co.end(), s.coefficients().begin(), s.coefficients().end());
    };
    for (const Shell& s : orbital_->shells()) add(s);
    for (const Shell& s : auxiliary_->shells()) add(s);
    L.push_back(0);
    np.push_back(1);
    ctr.insert(ctr.end(), {0.0, 0.0, 0.0});
    ex.push_back(0.0);
    co.push_back(1.0);
    lai_basis* b = nullptr;
    device::check(lai_dev_upload_basis(device::context(), static_cast<int>(L.size()), L.data(), ctr.data(), np.data(),
                                       ex.data(), co.data(), &b),
                  "DFFockBuilder::initialize");
    basis_ = b;
    const int metric = config_.metric_type == DFMetricType::Coulomb   ? 0
                       : config_.metric_type == DFMetricType::Overlap ? 1
                                                                      : 2;
    lai_df* d = nullptr;
    double cond = 0.0;
    const int rc = lai_dev_df_create(basis_, static_cast<int>(orbital_->n_shells()),
                                     static_cast<int>(auxiliary_->n_shells()), metric, config_.conditioning_hard_limit,
                                     &cond, &d);
    cond_ = cond;
    if (rc != LAI_OK) {
        release();
        device::check(rc, "DFFockBuilder::initialize");
    }
    df_ = d;
    const Size nn = n_orb_ * n_orb_;
    dD_ = device::DeviceArray(nn);
    dJ_ = device::DeviceArray(nn);
    dK_ = device::DeviceArray(nn);
}

void DFFockBuilder::run(int flags, const std::vector<Real>& D, std::vector<Real>* J, std::vector<Real>* K) {
    const Size nn = n_orb_ * n_orb_;
    dD_.upload(D.data(), nn);
    device::check(lai_dev_df_jk(df_, dD_.get(), dJ_.get(), dK_.get(), flags, nullptr), "DFFockBuilder::compute");
    if (J) dJ_.download(J->data(), nn);
    if (K) dK_.download(K->data(), nn);
    device::check(lai_dev_stream_sync(nullptr), "DFFockBuilder::compute");
}

void DFFockBuilder::compute_j_internal() { run(LAI_DF_J, D_, &J_, nullptr); }

namespace {
// exact symmetry of a row-major n x n matrix (then K is symmetric too)
bool symmetric(const std::vector<Real>& M, Size n) {
    for (Size i = 0; i < n; ++i)
        for (Size j = i + 1; j < n; ++j)
            if (M[i * n + j] != M[j * n + i]) return false;
    return true;
}
}  // namespace

void DFFockBuilder::compute_k_internal() {
    auto kflags = [&](const std::vector<Real>& D) { return LAI_DF_K | (symmetric(D, n_orb_) ? LAI_DF_SYMMETRIC : 0); };
    if (uhf_mode_) {
        run(kflags(D_alpha_), D_alpha_, nullptr, &K_alpha_);
        run(kflags(D_beta_), D_beta_, nullptr, &K_beta_);
        for (Size i = 0; i < n_orb_ * n_orb_; ++i) K_[i] = K_alpha_[i] + K_beta_[i];
    } else {
        run(kflags(D_), D_, nullptr, &K_);
    }
}

std::vector<Real> DFFockBuilder::compute() {
    initialize();
    std::fill(J_.begin(), J_.end(), 0.0);
    std::fill(K_.begin(), K_.end(), 0.0);
    if (config_.compute_coulomb && config_.compute_exchange && !uhf_mode_) {
        // one device build of both: J's fitted density from the diagonal of X = B D
        run(LAI_DF_J | LAI_DF_K | (symmetric(D_, n_orb_) ? LAI_DF_SYMMETRIC : 0), D_, &J_, &K_);
    } else {
        if (config_.compute_coulomb) compute_j_internal();
        if (config_.compute_exchange) compute_k_internal();
    }
    std::vector<Real> F(n_orb_ * n_orb_);
    for (Size i = 0; i < F.size(); ++i) F[i] = J_[i] - config_.exchange_fraction * K_[i];
    return F;
}

std::vector<Real> DFFockBuilder::compute_coulomb() {
    initialize();
    compute_j_internal();
    return J_;
}

std::vector<Real> DFFockBuilder::compute_exchange() {
    initialize();
    compute_k_internal();
    return K_;
}

void DFFockBuilder::compute_accumulate(std::span<Real> F) {
    if (F.size() != n_orb_ * n_orb_) throw InvalidArgumentException("Fock matrix size mismatch");
    const auto r = compute();
    for (Size i = 0; i < F.size(); ++i) F[i] += r[i];
}

std::vector<Real> DFFockBuilder::fock_matrix(std::span<const Real> H_core) const {
    std::vector<Real> F(n_orb_ * n_orb_);
    for (Size i = 0; i < F.size(); ++i) {
        F[i] = J_[i] - config_.exchange_fraction * K_[i];
        if (!H_core.empty()) F[i] += H_core[i];
    }
    return F;
}

std::unique_ptr<DFFockBuilder> make_df_fock_builder(const BasisSet& orbital, const std::string& aux_basis_name) {
    if (aux_basis_name.empty())
        throw InvalidArgumentException(
            "make_df_fock_builder: cannot auto-select auxiliary basis without atom data. "
            "Use the overload that accepts atomic_numbers and centers, or provide an explicit auxiliary basis name.");
    std::map<Index, Point3D> atoms;
    for (const Shell& s : orbital.shells()) atoms[s.atom_index()] = s.center();
    std::vector<Shell> shells;
    for (const auto& [atom, c] : atoms) {
        Shell s(0, c, {10.0, 2.0, 0.5, 0.1}, {0.25, 0.25, 0.25, 0.25});
        s.set_atom_index(atom);
        shells.push_back(std::move(s));
    }
    auto aux = std::make_unique<AuxiliaryBasisSet>(std::move(shells), FittingType::JKFIT, aux_basis_name);
    return std::make_unique<DFFockBuilder>(orbital, std::move(aux));
}

std::unique_ptr<DFFockBuilder> make_df_fock_builder(const BasisSet& orbital, std::span<const int> atomic_numbers,
                                                    std::span<const std::array<Real, 3>> centers,
                                                    const std::string& aux_basis_name, DFFockBuilderConfig config) {
    const std::string name = aux_basis_name.empty() ? std::string("cc-pVDZ-RI") : aux_basis_name;
    auto aux = std::make_unique<AuxiliaryBasisSet>(load_auxiliary_basis(name, atomic_numbers, centers));
    return std::make_unique<DFFockBuilder>(orbital, std::move(aux), config);
}

}  // namespace consumers
}  // namespace libaccint
