// engine.cpp -- Engine over the gfx950 backend (see engine.hpp).
#include <libaccint/engine.hpp>

#include "dispatch_detail.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace libaccint {

namespace screening {
Size SchwarzBounds::count_passing_quartets(Real thr) const {
    if (n_ == 0 || thr <= 0.0) {
        const Size np = n_ * (n_ + 1) / 2;
        return np * (np + 1) / 2;
    }
    Size count = 0;
    for (Size i = 0; i < n_; ++i)
        for (Size j = i; j < n_; ++j) {
            const Real qij = (*this)(i, j);
            if (qij * max_ < thr) continue;
            for (Size k = i; k < n_; ++k)
                for (Size l = (k == i) ? j : k; l < n_; ++l)
                    if (qij * (*this)(k, l) >= thr) ++count;
        }
    return count;
}

Real SchwarzBounds::estimate_pass_fraction(Real thr) const {
    if (n_ == 0) return 1.0;
    const Size np = n_ * (n_ + 1) / 2;
    const Size total = np * (np + 1) / 2;
    return total == 0 ? 1.0 : static_cast<Real>(count_passing_quartets(thr)) / static_cast<Real>(total);
}
}  // namespace screening

Engine::Engine(const BasisSet& basis, DispatchConfig config) : basis_(&basis), config_(config) {
    gpu_available_ = device::gpu_visible();
}

Engine::~Engine() {
    if (dev_) (void)lai_dev_basis_free(dev_);
}

void Engine::require_gpu(const char* what) const {
    if (!gpu_available_)
        throw BackendError(std::string(what) + ": no HIP device is available (this library has no CPU integral path)");
}

void Engine::require_two_electron(const Operator& op, const char* what) {
    if (!op.is_two_electron())
        throw InvalidArgumentException(std::string(what) + " requires a two-electron operator, got: " +
                                       std::string(to_string(op.kind())));
}

lai_basis* Engine::device_basis() {
    require_gpu("Engine");
    if (!dev_) {
        const Size ns = basis_->n_shells();
        std::vector<int> L(ns), np(ns), atom(ns);
        std::vector<double> ctr(3 * ns), ex, cf;
        int n_atoms = 0;
        for (Size s = 0; s < ns; ++s) {
            const Shell& sh = basis_->shell(s);
            L[s] = sh.angular_momentum();
            np[s] = static_cast<int>(sh.n_primitives());
            ctr[3 * s] = sh.center().x;
            ctr[3 * s + 1] = sh.center().y;
            ctr[3 * s + 2] = sh.center().z;
            for (Size k = 0; k < sh.n_primitives(); ++k) {
                ex.push_back(sh.exponents()[k]);
                cf.push_back(sh.coefficients()[k]);
            }
            atom[s] = sh.atom_index() >= 0 ? static_cast<int>(sh.atom_index()) : static_cast<int>(s);
            n_atoms = std::max(n_atoms, atom[s] + 1);
        }
        lai_basis* b = nullptr;
        device::check(lai_dev_upload_basis(device::context(), static_cast<int>(ns), L.data(), ctr.data(), np.data(),
                                           ex.data(), cf.data(), &b),
                      "Engine: basis upload");
        dev_ = b;
        device::check(lai_dev_set_shell_atoms(dev_, atom.data(), n_atoms), "Engine: shell atoms");
        // a dispatch table file applies to every engine (cuda_engine.cu:199-207)
        detail::apply_dispatch_table_env(dev_);
    }
    return dev_;
}

// ---- one-electron --------------------------------------------------------------
void Engine::compute_1e(const OneElectronOperator& op, std::vector<Real>& result, BackendHint) {
    lai_basis* b = device_basis();
    const Size n = basis_->n_basis_functions();
    result.assign(n * n, 0.0);
    device::DeviceArray d(n * n);
    std::vector<Real> tmp(n * n);
    for (const auto& t : op.terms()) {
        const auto& o = t.op;
        int kind = 0;
        std::vector<Real> Z, xyz;
        switch (o.kind()) {
            case OperatorKind::Overlap: kind = 0; break;
            case OperatorKind::Kinetic: kind = 1; break;
            case OperatorKind::Nuclear:
            case OperatorKind::PointCharge: {
                kind = 2;
                const auto& c = o.point_charges();
                for (Size i = 0; i < c.n_centers(); ++i) {
                    Z.push_back(c.charge[i]);
                    xyz.push_back(c.x[i]);
                    xyz.push_back(c.y[i]);
                    xyz.push_back(c.z[i]);
                }
                break;
            }
            default: throw InvalidArgumentException("compute_1e: not a one-electron operator");
        }
        device::check(lai_dev_onee(b, kind, static_cast<int>(Z.size()), Z.data(), xyz.data(), d.get(), nullptr),
                      "compute_1e");
        d.download(tmp.data(), n * n);
        for (Size i = 0; i < n * n; ++i) result[i] += t.scale * tmp[i];
    }
}

void Engine::compute_overlap_matrix(std::vector<Real>& S, BackendHint h) {
    compute_1e(OneElectronOperator(Operator::overlap()), S, h);
}
void Engine::compute_kinetic_matrix(std::vector<Real>& T, BackendHint h) {
    compute_1e(OneElectronOperator(Operator::kinetic()), T, h);
}
void Engine::compute_nuclear_matrix(const PointChargeParams& c, std::vector<Real>& V, BackendHint h) {
    compute_1e(OneElectronOperator(Operator::nuclear(c)), V, h);
}
void Engine::compute_core_hamiltonian(const PointChargeParams& c, std::vector<Real>& H, BackendHint h) {
    OneElectronOperator op(Operator::kinetic());
    op.add(Operator::nuclear(c));
    compute_1e(op, H, h);
}

// ---- raw two-electron -------------------------------------------------------------
void Engine::compute_eri_blocks(std::span<const std::array<int, 4>> q, std::vector<Real>& out,
                                std::vector<Index>& offsets) {
    lai_basis* b = device_basis();
    offsets.assign(q.size() + 1, 0);
    Index total = 0;
    for (Size t = 0; t < q.size(); ++t) {
        for (int k = 0; k < 4; ++k)
            if (q[t][k] < 0 || static_cast<Size>(q[t][k]) >= basis_->n_shells())
                throw InvalidArgumentException("compute_eri_blocks: shell index out of range");
        total += Index(basis_->shell(q[t][0]).n_functions()) * basis_->shell(q[t][1]).n_functions() *
                 basis_->shell(q[t][2]).n_functions() * basis_->shell(q[t][3]).n_functions();
    }
    out.assign(static_cast<Size>(total), 0.0);
    if (q.empty()) return;
    static_assert(sizeof(std::array<int, 4>) == 4 * sizeof(int));
    device::check(lai_dev_eri_quartets(b, static_cast<int64_t>(q.size()), q.data()->data(), out.data(),
                                       reinterpret_cast<int64_t*>(offsets.data())),
                  "compute_eri_blocks");
}

void Engine::compute(const Operator& op, const Shell& a, const Shell& b, const Shell& c, const Shell& d,
                     TwoElectronBuffer<0>& buffer, BackendHint) {
    require_two_electron(op, "compute with shell quartet");
    auto owned = [&](const Shell& s) {
        return s.shell_index() >= 0 && static_cast<Size>(s.shell_index()) < basis_->n_shells() &&
               &basis_->shell(s.shell_index()) == &s;
    };
    buffer.resize(a.n_functions(), b.n_functions(), c.n_functions(), d.n_functions());
    std::vector<Real> out;
    std::vector<Index> offs;
    if (owned(a) && owned(b) && owned(c) && owned(d)) {
        const std::array<int, 4> q{int(a.shell_index()), int(b.shell_index()), int(c.shell_index()),
                                   int(d.shell_index())};
        compute_eri_blocks(std::span<const std::array<int, 4>>(&q, 1), out, offs);
    } else {
        // shells from elsewhere: a four-shell basis of their (already normalised) data
        std::vector<Shell> v;
        for (const Shell* s : {&a, &b, &c, &d})
            v.emplace_back(Shell::PreNormalizedTag{}, s->angular_momentum(), s->center(),
                           std::vector<Real>(s->exponents().begin(), s->exponents().end()),
                           std::vector<Real>(s->coefficients().begin(), s->coefficients().end()));
        BasisSet tmp(std::move(v));
        Engine e(tmp);
        const std::array<int, 4> q{0, 1, 2, 3};
        e.compute_eri_blocks(std::span<const std::array<int, 4>>(&q, 1), out, offs);
    }
    std::copy(out.begin(), out.end(), buffer.data().begin());
}

std::vector<Real> Engine::compute_eri_tensor() {
    const Size n = basis_->n_basis_functions();
    const int ns = static_cast<int>(basis_->n_shells());
    std::vector<std::array<int, 4>> q;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j)
            for (int k = 0; k < ns; ++k)
                for (int l = 0; l < ns; ++l) q.push_back({i, j, k, l});
    std::vector<Real> out, T(n * n * n * n);
    std::vector<Index> offs;
    compute_eri_blocks(q, out, offs);
    for (Size t = 0; t < q.size(); ++t) {
        const Shell& A = basis_->shell(q[t][0]);
        const Shell& B = basis_->shell(q[t][1]);
        const Shell& C = basis_->shell(q[t][2]);
        const Shell& D = basis_->shell(q[t][3]);
        Index o = offs[t];
        for (int a = 0; a < A.n_functions(); ++a)
            for (int b = 0; b < B.n_functions(); ++b)
                for (int c = 0; c < C.n_functions(); ++c)
                    for (int d = 0; d < D.n_functions(); ++d, ++o) {
                        const Size mu = A.function_index() + a, nu = B.function_index() + b;
                        const Size la = C.function_index() + c, si = D.function_index() + d;
                        T[((mu * n + nu) * n + la) * n + si] = out[o];
                    }
    }
    return T;
}

std::vector<std::array<int, 4>> Engine::shell_set_quartet_list(const ShellSetQuartet& quartet) const {
    std::vector<std::array<int, 4>> q;
    const auto& A = quartet.bra_pair().shell_set_a();
    const auto& B = quartet.bra_pair().shell_set_b();
    const auto& C = quartet.ket_pair().shell_set_a();
    const auto& D = quartet.ket_pair().shell_set_b();
    q.reserve(A.n_shells() * B.n_shells() * C.n_shells() * D.n_shells());
    for (const auto& a : A.shells())
        for (const auto& b : B.shells())
            for (const auto& c : C.shells())
                for (const auto& d : D.shells())
                    q.push_back({int(a.shell_index()), int(b.shell_index()), int(c.shell_index()),
                                 int(d.shell_index())});
    return q;
}

Size Engine::block_size(const std::array<int, 4>& q) const {
    Size n = 1;
    for (int k = 0; k < 4; ++k) {
        if (q[k] < 0 || static_cast<Size>(q[k]) >= basis_->n_shells())
            throw InvalidArgumentException("shell index out of range");
        n *= static_cast<Size>(basis_->shell(q[k]).n_functions());
    }
    return n;
}

DeviceEriBatch Engine::compute_eri_batch_device(std::span<const std::array<int, 4>> q, hipStream_t stream) {
    lai_basis* b = device_basis();
    DeviceEriBatch batch;
    batch.stream_ = stream;
    if (q.empty()) return batch;
    std::vector<lai_eri_block> blk(q.size());
    std::vector<int64_t> offs(q.size() + 1);
    Size total = 0;
    for (Size t = 0; t < q.size(); ++t) {
        const Size c = block_size(q[t]);
        const Shell& A = basis_->shell(q[t][0]);
        const Shell& B = basis_->shell(q[t][1]);
        const Shell& C = basis_->shell(q[t][2]);
        const Shell& D = basis_->shell(q[t][3]);
        auto& e = blk[t];
        e.off = static_cast<int64_t>(total);
        e.fa = static_cast<int32_t>(A.function_index());
        e.fb = static_cast<int32_t>(B.function_index());
        e.fc = static_cast<int32_t>(C.function_index());
        e.fd = static_cast<int32_t>(D.function_index());
        e.na = A.n_functions();
        e.nb = B.n_functions();
        e.nc = C.n_functions();
        e.nd = D.n_functions();
        total += c;
    }
    batch.eri_ = device::DeviceArray(total);
    static_assert(sizeof(std::array<int, 4>) == 4 * sizeof(int));
    device::check(lai_dev_eri_quartets_device(b, static_cast<int64_t>(q.size()), q.data()->data(), batch.eri_.get(),
                                              offs.data(), stream),
                  "compute_eri_batch_device");
    for (Size t = 0; t < q.size(); ++t)
        if (offs[t] != blk[t].off) throw BackendError("compute_eri_batch_device: block layout mismatch");
    batch.blk_ = device::DeviceBytes(q.size() * sizeof(lai_eri_block));
    batch.blk_.upload(blk.data(), q.size() * sizeof(lai_eri_block), stream);
    batch.count_ = total;
    batch.n_blocks_ = q.size();
    return batch;
}

DeviceEriBatch Engine::compute_eri_batch_device_handle(const ShellSetQuartet& quartet, hipStream_t stream) {
    const auto q = shell_set_quartet_list(quartet);
    return compute_eri_batch_device(q, stream);
}

// ---- screening ----------------------------------------------------------------------
void Engine::precompute_schwarz_bounds() {
    if (schwarz_.is_initialized()) return;
    lai_basis* b = device_basis();
    const Size ns = basis_->n_shells();
    std::vector<Real> q(ns * (ns + 1) / 2);
    device::check(lai_dev_schwarz(b, q.data()), "precompute_schwarz_bounds");
    schwarz_ = screening::SchwarzBounds(ns, std::move(q));
    wl_valid_ = false;
}

const screening::SchwarzBounds& Engine::schwarz_bounds() const {
    if (!schwarz_.is_initialized())
        throw InvalidStateException("Schwarz bounds not computed; call precompute_schwarz_bounds()");
    return schwarz_;
}

void Engine::set_schwarz_bounds(std::vector<Real> packed) {
    const Size ns = basis_->n_shells();
    if (packed.size() != ns * (ns + 1) / 2) throw InvalidArgumentException("set_schwarz_bounds: size mismatch");
    device::check(lai_dev_set_schwarz(device_basis(), packed.data()), "set_schwarz_bounds");
    schwarz_ = screening::SchwarzBounds(ns, std::move(packed));
    wl_valid_ = false;
}

void Engine::set_density_matrix(const Real* D, Size nbf) {
    if (nbf != basis_->n_basis_functions())
        throw InvalidArgumentException("set_density_matrix: nbf mismatch");
    density_.assign(D, D + nbf * nbf);
}

std::vector<Real> Engine::shell_dmax(const Real* D) const {
    const Size ns = basis_->n_shells(), n = basis_->n_basis_functions();
    std::vector<Real> m(ns * ns, 0.0);
    for (Size i = 0; i < ns; ++i) {
        const Shell& A = basis_->shell(i);
        for (Size j = 0; j < ns; ++j) {
            const Shell& B = basis_->shell(j);
            Real v = 0.0;
            for (int a = 0; a < A.n_functions(); ++a)
                for (int b = 0; b < B.n_functions(); ++b) {
                    const Size mu = A.function_index() + a, nu = B.function_index() + b;
                    v = std::max(v, std::max(std::abs(D[mu * n + nu]), std::abs(D[nu * n + mu])));
                }
            m[i * ns + j] = v;
        }
    }
    return m;
}

// ---- fused J/K ------------------------------------------------------------------------
WorklistInfo Engine::prepare_worklist(const JKOptions& o, const Real* D_host) {
    lai_basis* b = device_basis();
    const bool need_d = (o.density_weighted && o.schwarz_threshold > 0.0) || o.mixed_precision;
    const bool same_opts = wl_valid_ && wl_opts_.schwarz_threshold == o.schwarz_threshold &&
                           wl_opts_.density_weighted == o.density_weighted &&
                           wl_opts_.mixed_precision == o.mixed_precision && wl_opts_.fp32_threshold == o.fp32_threshold;
    // no density given: a worklist built for the same options is reused as is
    if (same_opts && (!need_d || (!D_host && density_.empty()))) return wl_info_;
    if (need_d && !D_host) {
        if (density_.empty())
            throw InvalidArgumentException("prepare_worklist: density-dependent options need a host density");
        D_host = density_.data();
    }
    if (o.schwarz_threshold > 0.0 || o.mixed_precision) precompute_schwarz_bounds();
    std::vector<Real> dm = need_d ? shell_dmax(D_host) : std::vector<Real>{};
    if (same_opts && dm == wl_dmax_) return wl_info_;
    WorklistInfo info;
    int64_t nq = 0, n32 = 0;
    uint64_t cs = 0;
    device::check(lai_dev_build_worklist(b, o.schwarz_threshold, dm.empty() ? nullptr : dm.data(),
                                         o.density_weighted ? 1 : 0, o.mixed_precision ? o.fp32_threshold : 0.0,
                                         &nq, &n32, &cs),
                  "prepare_worklist");
    info.n_quartets = nq;
    info.n_fp32 = n32;
    info.checksum = cs;
    int64_t neri = 0;
    device::check(lai_dev_worklist_stats(b, reinterpret_cast<int64_t*>(info.class_counts.data()), &info.flops_fp64,
                                         &info.flops_fp32, &neri),
                  "prepare_worklist");
    info.n_eri = neri;
    wl_valid_ = true;
    wl_opts_ = o;
    wl_info_ = info;
    wl_dmax_ = std::move(dm);
    return info;
}

void Engine::build_jk(const JKOptions& o, const Real* d_D, Real* d_J, Real* d_K, hipStream_t s, const Real* D_host) {
    prepare_worklist(o, D_host);
    const int flags = jk_flags(o);
    // ProfileOnce: the first build of a precision tunes the dispatch table
    if (config_.enable_auto_tuning && tuned_flags_ != flags &&
        static_cast<Size>(wl_info_.n_quartets) >= config_.auto_tune_min_batch)
        (void)autotune(o, d_D, s, 3, D_host);
    device::check(lai_dev_set_mixed_prim_cut(device_basis(), o.fp32_prim_cut), "build_jk");
    device::check(lai_dev_jk(device_basis(), d_D, d_J, d_K, jk_flags(o), s),
                  "build_jk");
}

void Engine::build_jk_host(const JKOptions& o, const Real* D, Real* J, Real* K) {
    const Size n = basis_->n_basis_functions();
    device::DeviceArray dD(n * n), dJ(n * n), dK(n * n);
    dD.upload(D, n * n);
    dJ.zero();
    dK.zero();
    build_jk(o, dD.get(), dJ.get(), dK.get(), nullptr, D);
    device::check(lai_dev_stream_sync(nullptr), "build_jk_host");
    dJ.download(J, n * n);
    dK.download(K, n * n);
}

void Engine::capture_jk_graph(const JKOptions& o, const Real* d_D, Real* d_J, Real* d_K, hipStream_t s) {
    prepare_worklist(o, nullptr);
    const int flags = jk_flags(o);
    if (config_.enable_auto_tuning && tuned_flags_ != flags &&
        static_cast<Size>(wl_info_.n_quartets) >= config_.auto_tune_min_batch)
        (void)autotune(o, d_D, s, 3, nullptr);
    device::check(lai_dev_set_mixed_prim_cut(device_basis(), o.fp32_prim_cut), "capture_jk_graph");
    device::check(
        lai_dev_jk_graph_capture(device_basis(), d_D, d_J, d_K, jk_flags(o), s),
        "capture_jk_graph");
}

void Engine::launch_jk_graph(hipStream_t s) { device::check(lai_dev_jk_graph_launch(device_basis(), s), "launch_jk_graph"); }

namespace {
bool is_symmetric(const Real* D, Size n) {
    for (Size i = 0; i < n; ++i)
        for (Size j = i + 1; j < n; ++j)
            if (D[i * n + j] != D[j * n + i]) return false;
    return true;
}
}  // namespace

void Engine::build_jk_general(const JKOptions& o, const Real* D_host, const Real* d_D, Real* d_J, Real* d_K,
                              hipStream_t s) {
    const Size n = basis_->n_basis_functions();
    if (is_symmetric(D_host, n)) {
        build_jk(o, d_D, d_J, d_K, s, D_host);
        return;
    }
    // the worklist of the full D: its shell-pair max|D| bounds both parts
    std::vector<Real> Ds(n * n), Da(n * n);
    for (Size i = 0; i < n; ++i)
        for (Size j = 0; j < n; ++j) {
            Ds[i * n + j] = 0.5 * (D_host[i * n + j] + D_host[j * n + i]);
            Da[i * n + j] = 0.5 * (D_host[i * n + j] - D_host[j * n + i]);
        }
    device::DeviceArray dDs(n * n), dDa(n * n);
    dDs.upload(Ds.data(), n * n, s);
    dDa.upload(Da.data(), n * n, s);
    build_jk(o, dDs.get(), d_J, d_K, s, D_host);
    device::check(lai_dev_jk(device_basis(), dDa.get(), d_J, d_K,
                             jk_flags(o) | LAI_JK_ANTISYMMETRIC, s),
                  "build_jk");
    device::check(lai_dev_stream_sync(s), "build_jk");  // dDs and dDa are released on return
}

void Engine::jk_into(consumers::FockBuilder& fb, const JKOptions& o) {
    if (fb.nbf() != basis_->n_basis_functions())
        throw InvalidArgumentException("compute_and_consume: FockBuilder nbf does not match the basis");
    if (!fb.density()) throw InvalidStateException("FockBuilder::accumulate called without density matrix");
    const Size n = fb.nbf();
    std::vector<Real> J(n * n), K(n * n);
    device::DeviceArray dD(n * n), dJ(n * n), dK(n * n);
    dD.upload(fb.density(), n * n);
    dJ.zero();
    dK.zero();
    build_jk_general(o, fb.density(), dD.get(), dJ.get(), dK.get(), nullptr);
    device::check(lai_dev_stream_sync(nullptr), "compute_and_consume");
    dJ.download(J.data(), n * n);
    dK.download(K.data(), n * n);
    fb.add_jk(J.data(), K.data());
}

void Engine::jk_into(consumers::MixedPrecisionFockBuilder& fb, const JKOptions& o) {
    if (fb.nbf() != basis_->n_basis_functions())
        throw InvalidArgumentException("compute_and_consume: MixedPrecisionFockBuilder nbf does not match the basis");
    if (!fb.density()) throw InvalidStateException("MixedPrecisionFockBuilder: density matrix not set");
    const Size n = fb.nbf();
    std::vector<Real> J(n * n), K(n * n);
    device::DeviceArray dD(n * n), dJ(n * n), dK(n * n);
    dD.upload(fb.density(), n * n);
    dJ.zero();
    dK.zero();
    build_jk_general(o, fb.density(), dD.get(), dJ.get(), dK.get(), nullptr);
    device::check(lai_dev_stream_sync(nullptr), "compute_and_consume");
    dJ.download(J.data(), n * n);
    dK.download(K.data(), n * n);
    const Size nq = static_cast<Size>(wl_info_.n_quartets), n32 = static_cast<Size>(wl_info_.n_fp32);
    fb.add_jk(J.data(), K.data(), n32, nq - n32);
}

void Engine::jk_into(consumers::GpuFockBuilder& fb, const JKOptions& o) {
    if (fb.nbf() != basis_->n_basis_functions())
        throw InvalidArgumentException("compute_and_consume: GpuFockBuilder nbf does not match the basis");
    if (!fb.density_set()) throw InvalidStateException("GpuFockBuilder: density matrix not set");
    build_jk_general(o, fb.h_D(), fb.d_D(), fb.d_J(), fb.d_K(), fb.stream());
}

// ---- gradient -----------------------------------------------------------------------
std::vector<Real> Engine::compute_jk_gradient(const Real* D, Real exchange_scale, int n_atoms) {
    lai_basis* b = device_basis();
    int na = 0;
    for (Size s = 0; s < basis_->n_shells(); ++s)
        na = std::max(na, static_cast<int>(std::max<Index>(basis_->shell(s).atom_index(), Index(s))) + 1);
    if (n_atoms < 0) {
        n_atoms = 0;
        for (Size s = 0; s < basis_->n_shells(); ++s)
            n_atoms = std::max(n_atoms, static_cast<int>(basis_->shell(s).atom_index()) + 1);
    }
    (void)na;
    const Size n = basis_->n_basis_functions();
    // the gradient uses every canonical quartet that the current options keep
    if (!wl_valid_) prepare_worklist(JKOptions{}, D);
    device::DeviceArray dD(n * n), dG(3 * static_cast<Size>(std::max(n_atoms, 1)));
    dD.upload(D, n * n);
    dG.zero();
    device::check(lai_dev_jk_grad(b, dD.get(), exchange_scale, dG.get(), n_atoms, nullptr), "compute_jk_gradient");
    device::check(lai_dev_stream_sync(nullptr), "compute_jk_gradient");
    std::vector<Real> g(3 * static_cast<Size>(n_atoms));
    dG.download(g.data(), g.size());
    return g;
}

}  // namespace libaccint
