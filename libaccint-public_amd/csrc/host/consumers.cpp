// consumers.cpp -- FockBuilder (host J/K) and GpuFockBuilder (device J/K/D).
#include <libaccint/consumers.hpp>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <limits>
#include <string>
#include <thread>

namespace libaccint::consumers {

namespace {
constexpr std::size_t kDefaultMaxThreadLocalBytes = 512ull * 1024ull * 1024ull;

std::size_t max_thread_local_bytes() {
    const char* env = std::getenv("LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES");
    if (!env || !*env) return kDefaultMaxThreadLocalBytes;
    char* end = nullptr;
    const unsigned long long v = std::strtoull(env, &end, 10);
    if (end != env && *end == '\0' && v > 0) return static_cast<std::size_t>(v);
    return kDefaultMaxThreadLocalBytes;
}

thread_local int t_thread_id = 0;
}  // namespace

FockBuilder::FockBuilder(Size nbf) : nbf_(nbf), J_(nbf * nbf, 0.0), K_(nbf * nbf, 0.0) {}

void FockBuilder::set_density(const Real* D, Size nbf) {
    if (nbf != nbf_)
        throw InvalidArgumentException("FockBuilder::set_density: nbf mismatch (expected " + std::to_string(nbf_) +
                                       ", got " + std::to_string(nbf) + ")");
    D_ = D;
}

void FockBuilder::acc_plain(std::vector<Real>& J, std::vector<Real>& K, const TwoElectronBuffer<0>& buf, Index fa,
                            Index fb, Index fc, Index fd, int na, int nb, int nc, int nd) {
    const Size n = nbf_;
    for (int a = 0; a < na; ++a) {
        const Size mu = fa + a;
        for (int b = 0; b < nb; ++b) {
            const Size nu = fb + b;
            for (int c = 0; c < nc; ++c) {
                const Size la = fc + c;
                for (int d = 0; d < nd; ++d) {
                    const Size si = fd + d;
                    const Real g = buf(a, b, c, d);
                    J[mu * n + nu] += g * D_[la * n + si];
                    K[mu * n + la] += g * D_[nu * n + si];
                }
            }
        }
    }
}

void FockBuilder::acc_atomic(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na,
                             int nb, int nc, int nd) {
    const Size n = nbf_;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d) {
                    const Size mu = fa + a, nu = fb + b, la = fc + c, si = fd + d;
                    const Real g = buf(a, b, c, d);
                    std::atomic_ref<Real>(J_[mu * n + nu]).fetch_add(g * D_[la * n + si], std::memory_order_relaxed);
                    std::atomic_ref<Real>(K_[mu * n + la]).fetch_add(g * D_[nu * n + si], std::memory_order_relaxed);
                }
}

void FockBuilder::accumulate(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na,
                             int nb, int nc, int nd) {
    if (!D_) throw InvalidStateException("FockBuilder::accumulate called without density matrix");
    const FockThreadingStrategy eff = auto_atomic_ ? FockThreadingStrategy::Atomic
                                      : auto_thread_local_ ? FockThreadingStrategy::ThreadLocal
                                                           : strategy_;
    if (eff == FockThreadingStrategy::Atomic) {
        acc_atomic(buf, fa, fb, fc, fd, na, nb, nc, nd);
    } else if (eff == FockThreadingStrategy::ThreadLocal && !J_tl_.empty()) {
        const int tid = t_thread_id;
        if (tid < 0 || static_cast<Size>(tid) >= J_tl_.size())
            throw InvalidStateException("FockBuilder::accumulate without prepared thread-local buffers for thread " +
                                        std::to_string(tid));
        acc_plain(J_tl_[tid], K_tl_[tid], buf, fa, fb, fc, fd, na, nb, nc, nd);
    } else {
        acc_plain(J_, K_, buf, fa, fb, fc, fd, na, nb, nc, nd);
    }
}

// 8-permutation scatter of a canonical quartet (fock_builder.cpp:123-222)
void FockBuilder::accumulate_symmetric(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd,
                                       int na, int nb, int nc, int nd, bool ij, bool kl, bool bk) {
    if (!D_) throw InvalidStateException("FockBuilder::accumulate_symmetric called without density matrix");
    const Size n = nbf_;
    auto* J = J_.data();
    auto* K = K_.data();
    const Real* D = D_;
    for (int a = 0; a < na; ++a) {
        const Size mu = fa + a;
        for (int b = 0; b < nb; ++b) {
            const Size nu = fb + b;
            for (int c = 0; c < nc; ++c) {
                const Size la = fc + c;
                for (int d = 0; d < nd; ++d) {
                    const Size si = fd + d;
                    const Real g = buf(a, b, c, d);
                    const Real Dls = D[la * n + si], Dsl = D[si * n + la];
                    const Real Dmn = D[mu * n + nu], Dnm = D[nu * n + mu];
                    const Real Dns = D[nu * n + si], Dms = D[mu * n + si];
                    const Real Dnl = D[nu * n + la], Dml = D[mu * n + la];
                    J[mu * n + nu] += g * Dls;
                    K[mu * n + la] += g * Dns;
                    if (!kl) {
                        J[mu * n + nu] += g * Dsl;
                        K[mu * n + si] += g * Dnl;
                    }
                    if (!ij) {
                        J[nu * n + mu] += g * Dls;
                        K[nu * n + la] += g * Dms;
                    }
                    if (!ij && !kl) {
                        J[nu * n + mu] += g * Dsl;
                        K[nu * n + si] += g * Dml;
                    }
                    if (!bk) {
                        J[la * n + si] += g * Dmn;
                        K[la * n + mu] += g * Dns;
                        if (!ij) {
                            J[la * n + si] += g * Dnm;
                            K[la * n + nu] += g * Dms;
                        }
                        if (!kl) {
                            J[si * n + la] += g * Dmn;
                            K[si * n + mu] += g * Dnl;
                        }
                        if (!ij && !kl) {
                            J[si * n + la] += g * Dnm;
                            K[si * n + nu] += g * Dml;
                        }
                    }
                }
            }
        }
    }
}

// ---- MixedPrecisionFockBuilder -------------------------------------------------
MixedPrecisionFockBuilder::MixedPrecisionFockBuilder(Size nbf) : nbf_(nbf), J_(nbf * nbf, 0.0), K_(nbf * nbf, 0.0) {}

void MixedPrecisionFockBuilder::set_density(const Real* D, Size nbf) {
    if (nbf != nbf_)
        throw InvalidArgumentException("MixedPrecisionFockBuilder::set_density: nbf " + std::to_string(nbf) +
                                       " does not match " + std::to_string(nbf_));
    D_ = D;
}

template <typename R>
void MixedPrecisionFockBuilder::accumulate_impl(const TwoElectronBuffer<0, R>& buf, Index fa, Index fb, Index fc,
                                                Index fd, int na, int nb, int nc, int nd) {
    if (!D_) throw InvalidStateException("MixedPrecisionFockBuilder::accumulate called without density matrix");
    const Size n = nbf_;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d) {
                    const Size mu = fa + a, nu = fb + b, la = fc + c, si = fd + d;
                    const double g = static_cast<double>(buf(a, b, c, d));  // promoted before accumulation
                    J_[mu * n + nu] += g * D_[la * n + si];
                    K_[mu * n + la] += g * D_[nu * n + si];
                }
}

void MixedPrecisionFockBuilder::accumulate(const TwoElectronBuffer<0, float>& buf, Index fa, Index fb, Index fc,
                                           Index fd, int na, int nb, int nc, int nd) {
    accumulate_impl(buf, fa, fb, fc, fd, na, nb, nc, nd);
    ++n32_;
}

void MixedPrecisionFockBuilder::accumulate(const TwoElectronBuffer<0, double>& buf, Index fa, Index fb, Index fc,
                                           Index fd, int na, int nb, int nc, int nd) {
    accumulate_impl(buf, fa, fb, fc, fd, na, nb, nc, nd);
    ++n64_;
}

void MixedPrecisionFockBuilder::add_jk(const Real* J, const Real* K, Size n_fp32, Size n_fp64) {
    for (Size i = 0; i < nbf_ * nbf_; ++i) {
        J_[i] += J[i];
        K_[i] += K[i];
    }
    n32_ += n_fp32;
    n64_ += n_fp64;
}

std::vector<Real> MixedPrecisionFockBuilder::get_fock_matrix(std::span<const Real> H, Real x) const {
    if (H.size() != nbf_ * nbf_)
        throw InvalidArgumentException("MixedPrecisionFockBuilder::get_fock_matrix: H_core size mismatch");
    std::vector<Real> F(nbf_ * nbf_);
    for (Size i = 0; i < F.size(); ++i) F[i] = H[i] + J_[i] - x * K_[i];
    return F;
}

void MixedPrecisionFockBuilder::reset() noexcept {
    std::fill(J_.begin(), J_.end(), 0.0);
    std::fill(K_.begin(), K_.end(), 0.0);
    n32_ = n64_ = 0;
}

void FockBuilder::add_jk(const Real* J, const Real* K) {
    for (Size i = 0; i < nbf_ * nbf_; ++i) {
        J_[i] += J[i];
        K_[i] += K[i];
    }
}

std::vector<Real> FockBuilder::get_fock_matrix(std::span<const Real> H, Real x) const {
    if (H.size() != nbf_ * nbf_) throw InvalidArgumentException("FockBuilder::get_fock_matrix: H_core size mismatch");
    std::vector<Real> F(nbf_ * nbf_);
    for (Size i = 0; i < F.size(); ++i) F[i] = H[i] + J_[i] - x * K_[i];
    return F;
}

void FockBuilder::reset() noexcept {
    std::fill(J_.begin(), J_.end(), 0.0);
    std::fill(K_.begin(), K_.end(), 0.0);
    for (auto& v : J_tl_) std::fill(v.begin(), v.end(), 0.0);
    for (auto& v : K_tl_) std::fill(v.begin(), v.end(), 0.0);
}

void FockBuilder::set_threading_strategy(FockThreadingStrategy s) {
    strategy_ = s;
    auto_thread_local_ = false;
    auto_atomic_ = false;
    if (s != FockThreadingStrategy::ThreadLocal) {
        J_tl_.clear();
        K_tl_.clear();
    }
}

// prepare_parallel / finalize_parallel: thread-local buffers unless they
// exceed LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES, then atomics
// (fock_builder.cpp:276-371)
void FockBuilder::prepare_parallel(int n_threads) {
    if (n_threads <= 0) n_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    n_threads_ = std::max(1, n_threads);
    auto_atomic_ = false;
    auto_thread_local_ = (strategy_ == FockThreadingStrategy::Sequential && n_threads_ > 1);
    bool use_tl = strategy_ == FockThreadingStrategy::ThreadLocal || auto_thread_local_;
    if (use_tl) {
        const std::size_t mat = nbf_ * nbf_;
        if (mat != 0 && static_cast<std::size_t>(n_threads_) > std::numeric_limits<std::size_t>::max() / mat / 2 /
                                                                    sizeof(Real))
            throw MemoryException("FockBuilder::prepare_parallel: overflow while sizing thread-local buffers");
        const std::size_t bytes = static_cast<std::size_t>(n_threads_) * mat * 2 * sizeof(Real);
        if (bytes > max_thread_local_bytes()) {
            if (strategy_ == FockThreadingStrategy::ThreadLocal)
                throw MemoryException("FockBuilder::prepare_parallel: requested ThreadLocal strategy needs " +
                                      std::to_string(bytes) + " bytes, exceeding limit " +
                                      std::to_string(max_thread_local_bytes()) +
                                      ". Reduce threads, reduce basis size, or raise "
                                      "LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES.");
            auto_thread_local_ = false;
            auto_atomic_ = n_threads_ > 1;
            use_tl = false;
        } else {
            J_tl_.assign(n_threads_, std::vector<Real>(mat, 0.0));
            K_tl_.assign(n_threads_, std::vector<Real>(mat, 0.0));
        }
    }
    if (!use_tl) {
        J_tl_.clear();
        K_tl_.clear();
    }
}

void FockBuilder::finalize_parallel() {
    const bool use_tl = strategy_ == FockThreadingStrategy::ThreadLocal || auto_thread_local_;
    if (use_tl && !J_tl_.empty()) {
        for (Size t = 0; t < J_tl_.size(); ++t)
            for (Size i = 0; i < J_.size(); ++i) {
                J_[i] += J_tl_[t][i];
                K_[i] += K_tl_[t][i];
            }
        J_tl_.clear();
        K_tl_.clear();
    }
    auto_thread_local_ = false;
    auto_atomic_ = false;
}

// ---------------------------------------------------------------------------
GpuFockBuilder::GpuFockBuilder(Size nbf, hipStream_t stream)
    : nbf_(nbf), stream_(stream), J_(nbf * nbf), K_(nbf * nbf), D_(nbf * nbf) {
    J_.zero(stream_);
    K_.zero(stream_);
    D_.zero(stream_);
    device::check(lai_dev_stream_sync(stream_), "GpuFockBuilder");
}

namespace {
// doubles of host ERI blocks collected before one device digestion
constexpr Size kStageDoubles = Size(1) << 20;
}  // namespace

void GpuFockBuilder::set_density(const Real* D, Size nbf) {
    if (nbf != nbf_)
        throw InvalidArgumentException("GpuFockBuilder::set_density: nbf mismatch (expected " +
                                       std::to_string(nbf_) + ", got " + std::to_string(nbf) + ")");
    flush();  // staged blocks belong to the previous density
    D_host_.assign(D, D + nbf * nbf);
    D_.upload(D, nbf * nbf, stream_);
    density_set_ = true;
}

void GpuFockBuilder::accumulate(const TwoElectronBuffer<0>& buf, Index fa, Index fb, Index fc, Index fd, int na,
                                int nb, int nc, int nd) {
    if (!density_set_) throw InvalidStateException("GpuFockBuilder::accumulate called without density matrix");
    for (int x : {na, nb, nc, nd})
        if (x < 1 || x > 10) throw InvalidArgumentException("GpuFockBuilder::accumulate: block dimension out of range");
    const Index n = static_cast<Index>(nbf_);
    if (fa < 0 || fb < 0 || fc < 0 || fd < 0 || fa + na > n || fb + nb > n || fc + nc > n || fd + nd > n)
        throw InvalidArgumentException("GpuFockBuilder::accumulate: function offsets outside the basis");
    const Size cnt = static_cast<Size>(na) * nb * nc * nd;
    if (buf.na() < na || buf.nb() < nb || buf.nc() < nc || buf.nd() < nd)
        throw InvalidArgumentException("GpuFockBuilder::accumulate: buffer smaller than the block");
    if (stage_eri_.size() + cnt > kStageDoubles) flush();
    lai_eri_block blk{};
    blk.off = static_cast<int64_t>(stage_eri_.size());
    blk.fa = static_cast<int32_t>(fa);
    blk.fb = static_cast<int32_t>(fb);
    blk.fc = static_cast<int32_t>(fc);
    blk.fd = static_cast<int32_t>(fd);
    blk.na = na;
    blk.nb = nb;
    blk.nc = nc;
    blk.nd = nd;
    stage_blk_.push_back(blk);
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d) stage_eri_.push_back(buf(a, b, c, d));
}

void GpuFockBuilder::flush() const {
    if (stage_blk_.empty()) return;
    const Size nd = stage_eri_.size(), nbk = stage_blk_.size();
    if (d_stage_eri_.size() < nd) d_stage_eri_ = device::DeviceArray(std::max(nd, kStageDoubles));
    if (d_stage_blk_.bytes() < nbk * sizeof(lai_eri_block))
        d_stage_blk_ = device::DeviceBytes(std::max(nbk, kStageDoubles / 16) * sizeof(lai_eri_block));
    // synchronous copies on the builder's stream: the previous digestion has
    // read the staging buffers before they are overwritten
    d_stage_eri_.upload(stage_eri_.data(), nd, stream_);
    d_stage_blk_.upload(stage_blk_.data(), nbk * sizeof(lai_eri_block), stream_);
    auto* self = const_cast<GpuFockBuilder*>(this);
    device::check(lai_dev_fock_accumulate(d_stage_eri_.get(), static_cast<const lai_eri_block*>(d_stage_blk_.get()),
                                          static_cast<int64_t>(nbk), D_.get(), self->J_.get(), self->K_.get(),
                                          static_cast<int>(nbf_), stream_),
                  "GpuFockBuilder::accumulate");
    stage_eri_.clear();
    stage_blk_.clear();
}

void GpuFockBuilder::accumulate_device_eri_batch(const DeviceEriBatch& batch, Size nbf) {
    if (nbf != nbf_)
        throw InvalidArgumentException("GpuFockBuilder::accumulate_device_eri_batch: nbf mismatch (expected " +
                                       std::to_string(nbf_) + ", got " + std::to_string(nbf) + ")");
    if (!density_set_)
        throw InvalidStateException("GpuFockBuilder::accumulate_device_eri_batch called without density matrix");
    if (batch.empty()) return;
    flush();
    device::stream_wait(stream_, batch.stream());
    device::check(lai_dev_fock_accumulate(batch.data(), batch.blocks(), static_cast<int64_t>(batch.n_blocks()),
                                          D_.get(), J_.get(), K_.get(), static_cast<int>(nbf_), stream_),
                  "GpuFockBuilder::accumulate_device_eri_batch");
}

double* GpuFockBuilder::d_J() {
    flush();
    return J_.get();
}

double* GpuFockBuilder::d_K() {
    flush();
    return K_.get();
}

std::vector<Real> GpuFockBuilder::get_coulomb_matrix() const {
    flush();
    std::vector<Real> J(nbf_ * nbf_);
    J_.download(J.data(), J.size(), stream_);
    return J;
}

std::vector<Real> GpuFockBuilder::get_exchange_matrix() const {
    flush();
    std::vector<Real> K(nbf_ * nbf_);
    K_.download(K.data(), K.size(), stream_);
    return K;
}

std::vector<Real> GpuFockBuilder::get_fock_matrix(std::span<const Real> H, Real x) const {
    if (H.size() != nbf_ * nbf_)
        throw InvalidArgumentException("GpuFockBuilder::get_fock_matrix: H_core size mismatch");
    auto J = get_coulomb_matrix();
    auto K = get_exchange_matrix();
    for (Size i = 0; i < J.size(); ++i) J[i] = H[i] + J[i] - x * K[i];
    return J;
}

void GpuFockBuilder::reset() {
    stage_eri_.clear();
    stage_blk_.clear();
    J_.zero(stream_);
    K_.zero(stream_);
}

void GpuFockBuilder::synchronize() {
    flush();
    device::check(lai_dev_stream_sync(stream_), "GpuFockBuilder::synchronize");
}

}  // namespace libaccint::consumers
