// batch.cpp -- BatchEngine: many molecules through one device basis handle
// (lai_dev_upload_batch), one class kernel per angular-momentum class for all
// of them.
#include <algorithm>
#include <cmath>

#include "libaccint/engine.hpp"
#include "dispatch_detail.hpp"

namespace libaccint {

BatchEngine::BatchEngine(std::vector<std::shared_ptr<const BasisSet>> bases) : bases_(std::move(bases)) {
    if (bases_.empty()) throw InvalidArgumentException("BatchEngine: no molecules");
    moff_.assign(1, 0);
    for (const auto& b : bases_) {
        if (!b) throw InvalidArgumentException("BatchEngine: null basis");
        const Size n = b->n_basis_functions();
        nbf_.push_back(n);
        moff_.push_back(moff_.back() + n * n);
    }
}

BatchEngine::~BatchEngine() {
    if (dev_) (void)lai_dev_basis_free(dev_);
}

lai_basis* BatchEngine::device_handle() {
    if (dev_) return dev_;
    std::vector<int> L, np, off(1, 0);
    std::vector<double> ctr, ex, cf;
    for (const auto& bp : bases_) {
        const BasisSet& b = *bp;
        for (Size s = 0; s < b.n_shells(); ++s) {
            const Shell& sh = b.shell(s);
            L.push_back(sh.angular_momentum());
            np.push_back(static_cast<int>(sh.n_primitives()));
            ctr.insert(ctr.end(), {sh.center().x, sh.center().y, sh.center().z});
            for (Size k = 0; k < sh.n_primitives(); ++k) {
                ex.push_back(sh.exponents()[k]);
                cf.push_back(sh.coefficients()[k]);
            }
        }
        off.push_back(off.back() + static_cast<int>(b.n_shells()));
    }
    lai_basis* h = nullptr;
    device::check(lai_dev_upload_batch(device::context(), static_cast<int>(bases_.size()), off.data(),
                                       static_cast<int>(L.size()), L.data(), ctr.data(), np.data(), ex.data(),
                                       cf.data(), &h),
                  "BatchEngine");
    dev_ = h;
    detail::apply_dispatch_table_env(dev_);
    return dev_;
}

void BatchEngine::precompute_schwarz_bounds() {
    if (have_q_) return;
    device::check(lai_dev_schwarz(device_handle(), nullptr), "BatchEngine::precompute_schwarz_bounds");
    have_q_ = true;
}

// per molecule: n_shells x n_shells max |D| over each shell-pair block, one
// molecule after another
std::vector<Real> BatchEngine::shell_dmax(const Real* D) const {
    std::vector<Real> out;
    for (Size m = 0; m < bases_.size(); ++m) {
        const BasisSet& b = *bases_[m];
        const Size ns = b.n_shells(), n = nbf_[m];
        const Real* Dm = D + moff_[m];
        for (Size i = 0; i < ns; ++i) {
            const Shell& A = b.shell(i);
            for (Size j = 0; j < ns; ++j) {
                const Shell& B = b.shell(j);
                Real v = 0.0;
                for (int a = 0; a < A.n_functions(); ++a)
                    for (int c = 0; c < B.n_functions(); ++c) {
                        const Size mu = A.function_index() + a, nu = B.function_index() + c;
                        v = std::max(v, std::max(std::abs(Dm[mu * n + nu]), std::abs(Dm[nu * n + mu])));
                    }
                out.push_back(v);
            }
        }
    }
    return out;
}

WorklistInfo BatchEngine::prepare_worklist(const JKOptions& o, const Real* D_host) {
    lai_basis* h = device_handle();
    const bool need_d = (o.density_weighted && o.schwarz_threshold > 0.0) || o.mixed_precision;
    const bool same_opts = wl_valid_ && wl_opts_.schwarz_threshold == o.schwarz_threshold &&
                           wl_opts_.density_weighted == o.density_weighted &&
                           wl_opts_.mixed_precision == o.mixed_precision && wl_opts_.fp32_threshold == o.fp32_threshold;
    if (same_opts && (!need_d || !D_host)) return wl_info_;
    if (need_d && !D_host)
        throw InvalidArgumentException("BatchEngine::prepare_worklist: density-dependent options need a host density");
    if (o.schwarz_threshold > 0.0 || o.mixed_precision) precompute_schwarz_bounds();
    std::vector<Real> dm = need_d ? shell_dmax(D_host) : std::vector<Real>{};
    WorklistInfo info;
    int64_t nq = 0, n32 = 0, neri = 0;
    uint64_t cs = 0;
    device::check(lai_dev_build_worklist(h, o.schwarz_threshold, dm.empty() ? nullptr : dm.data(),
                                         o.density_weighted ? 1 : 0, o.mixed_precision ? o.fp32_threshold : 0.0,
                                         &nq, &n32, &cs),
                  "BatchEngine::prepare_worklist");
    device::check(lai_dev_worklist_stats(h, reinterpret_cast<int64_t*>(info.class_counts.data()), &info.flops_fp64,
                                         &info.flops_fp32, &neri),
                  "BatchEngine::prepare_worklist");
    info.n_quartets = nq;
    info.n_fp32 = n32;
    info.checksum = cs;
    info.n_eri = neri;
    wl_valid_ = true;
    wl_opts_ = o;
    wl_info_ = info;
    return info;
}

void BatchEngine::build_jk(const JKOptions& o, const Real* d_D, Real* d_J, Real* d_K, hipStream_t s,
                           const Real* D_host) {
    prepare_worklist(o, D_host);
    device::check(lai_dev_set_mixed_prim_cut(device_handle(), o.fp32_prim_cut), "BatchEngine::build_jk");
    device::check(lai_dev_jk(device_handle(), d_D, d_J, d_K, jk_flags(o), s),
                  "BatchEngine::build_jk");
}

void BatchEngine::build_jk_host(const JKOptions& o, const Real* D, Real* J, Real* K) {
    const Size n = matrix_size();
    device::DeviceArray dD(n), dJ(n), dK(n);
    dD.upload(D, n);
    dJ.zero();
    dK.zero();
    build_jk(o, dD.get(), dJ.get(), dK.get(), nullptr, D);
    device::check(lai_dev_stream_sync(nullptr), "BatchEngine::build_jk_host");
    dJ.download(J, n);
    dK.download(K, n);
}

void BatchEngine::capture_jk_graph(const JKOptions& o, const Real* d_D, Real* d_J, Real* d_K, hipStream_t s) {
    prepare_worklist(o, nullptr);
    device::check(lai_dev_set_mixed_prim_cut(device_handle(), o.fp32_prim_cut), "BatchEngine::capture_jk_graph");
    device::check(
        lai_dev_jk_graph_capture(device_handle(), d_D, d_J, d_K, jk_flags(o), s),
        "BatchEngine::capture_jk_graph");
}

void BatchEngine::launch_jk_graph(hipStream_t s) {
    device::check(lai_dev_jk_graph_launch(device_handle(), s), "BatchEngine::launch_jk_graph");
}

}  // namespace libaccint
