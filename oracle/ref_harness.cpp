// ref_harness.cpp -- extern "C" driver over the reference LibAccInt CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own source files, read in place from the reference tree,
// into oracle/_ref/libref.so.  It is used to pin the C restatement
// (oracle.c) and, optionally, as the timed CPU baseline.  Only the loops that
// live in the reference's config-dependent engine headers are restated here
// (include/libaccint/engine/engine.hpp:1815-1861 canonical loop,
// include/libaccint/engine/cpu_engine.hpp:553-622 parallel N^4 loop); every
// integral and every J/K update is the reference's own code.
#include <libaccint/basis/auxiliary_basis_set.hpp>
#include <libaccint/basis/basis_set.hpp>
#include <libaccint/basis/shell.hpp>
#include <libaccint/buffers/one_electron_buffer.hpp>
#include <libaccint/buffers/two_electron_buffer.hpp>
#include <libaccint/consumers/df_fock_builder.hpp>
#include <libaccint/consumers/fock_builder.hpp>
#include <libaccint/kernels/eri_kernel.hpp>
#include <libaccint/kernels/three_center_eri_kernel.hpp>
#include <libaccint/kernels/two_center_coulomb_kernel.hpp>
#include <libaccint/kernels/kinetic_kernel.hpp>
#include <libaccint/kernels/nuclear_kernel.hpp>
#include <libaccint/kernels/overlap_kernel.hpp>
#include <libaccint/math/boys_function.hpp>
#include <libaccint/math/rys_quadrature.hpp>
#include <libaccint/screening/schwarz_bounds.hpp>

#include <omp.h>

#include <cstring>
#include <string>
#include <memory>
#include <vector>

using namespace libaccint;

namespace {
struct RefBasis {
    std::unique_ptr<BasisSet> basis;
};
std::string g_df_error;
}  // namespace

extern "C" {

void ref_boys(int nmax, double T, double* out) { math::boys_evaluate_array(nmax, T, out); }

void ref_rys(int n, double T, double* roots, double* weights) {
    math::rys_compute(n, T, roots, weights);
}

void* ref_basis_create(int n_shells, const int* L, const double* centers, const int* nprim,
                       const double* exps, const double* raw_coefs) {
    std::vector<Shell> shells;
    int off = 0;
    for (int s = 0; s < n_shells; ++s) {
        std::vector<Real> e(exps + off, exps + off + nprim[s]);
        std::vector<Real> c(raw_coefs + off, raw_coefs + off + nprim[s]);
        shells.emplace_back(L[s], Point3D{centers[3 * s], centers[3 * s + 1], centers[3 * s + 2]},
                            std::move(e), std::move(c));
        off += nprim[s];
    }
    auto* h = new RefBasis;
    h->basis = std::make_unique<BasisSet>(std::move(shells));
    return h;
}

void ref_basis_free(void* h) { delete static_cast<RefBasis*>(h); }

int ref_basis_nbf(void* h) {
    return static_cast<int>(static_cast<RefBasis*>(h)->basis->n_basis_functions());
}

void ref_basis_export(void* h, double* coefs, int* func_off) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    int off = 0;
    for (Size s = 0; s < b.n_shells(); ++s) {
        const auto& sh = b.shell(s);
        auto c = sh.coefficients();
        for (Size k = 0; k < c.size(); ++k) coefs[off + k] = c[k];
        off += static_cast<int>(c.size());
        func_off[s] = static_cast<int>(sh.function_index());
    }
}

void ref_eri_shell_quartet(void* h, int i, int j, int k, int l, double* out) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    TwoElectronBuffer<0> buf;
    kernels::compute_eri(b.shell(i), b.shell(j), b.shell(k), b.shell(l), buf);
    const int n = b.shell(i).n_functions() * b.shell(j).n_functions() * b.shell(k).n_functions() *
                  b.shell(l).n_functions();
    std::memcpy(out, buf.data().data(), sizeof(double) * n);
}

static void one_e(void* h, int kind, int nch, const double* Z, const double* xyz, double* M) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const Size n = b.n_basis_functions();
    PointChargeParams ch;
    for (int c = 0; c < nch; ++c) {
        ch.x.push_back(xyz[3 * c]);
        ch.y.push_back(xyz[3 * c + 1]);
        ch.z.push_back(xyz[3 * c + 2]);
        ch.charge.push_back(Z[c]);
    }
    OverlapBuffer buf;
    for (Size i = 0; i < b.n_shells(); ++i)
        for (Size j = i; j < b.n_shells(); ++j) {
            const auto& A = b.shell(i);
            const auto& B = b.shell(j);
            if (kind == 0) kernels::compute_overlap(A, B, buf);
            else if (kind == 1) kernels::compute_kinetic(A, B, buf);
            else kernels::compute_nuclear(A, B, ch, buf);
            for (int a = 0; a < A.n_functions(); ++a)
                for (int c = 0; c < B.n_functions(); ++c) {
                    const Size mu = A.function_index() + a, nu = B.function_index() + c;
                    M[mu * n + nu] = buf(a, c);
                    M[nu * n + mu] = buf(a, c);
                }
        }
}

void ref_overlap(void* h, double* S) { one_e(h, 0, 0, nullptr, nullptr, S); }
void ref_kinetic(void* h, double* T) { one_e(h, 1, 0, nullptr, nullptr, T); }
void ref_nuclear(void* h, int nch, const double* Z, const double* xyz, double* V) {
    one_e(h, 2, nch, Z, xyz, V);
}

void ref_schwarz(void* h, double* Q_packed) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    screening::SchwarzBounds sb(b);
    const Size n = b.n_shells();
    for (Size j = 0; j < n; ++j)
        for (Size i = 0; i <= j; ++i) Q_packed[j * (j + 1) / 2 + i] = sb(i, j);
}

// Canonical screened loop (engine.hpp:1815-1861) over the reference's
// accumulate_symmetric; one FockBuilder per thread, summed at the end.
void ref_jk_canonical(void* h, const double* D, const double* Q_packed, double thr, double* J,
                      double* K, int n_threads) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const int n = static_cast<int>(b.n_shells());
    const Size nbf = b.n_basis_functions();
    std::memset(J, 0, sizeof(double) * nbf * nbf);
    std::memset(K, 0, sizeof(double) * nbf * nbf);
    auto Q = [&](int i, int j) {
        if (i > j) std::swap(i, j);
        return Q_packed[static_cast<Size>(j) * (j + 1) / 2 + i];
    };
    const bool use_q = Q_packed != nullptr && thr > 0.0;
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) pairs.emplace_back(i, j);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
        consumers::FockBuilder fb(nbf);
        fb.set_density(D, nbf);
        TwoElectronBuffer<0> buf;
#pragma omp for schedule(dynamic, 1)
        for (Size p = 0; p < pairs.size(); ++p) {
            const int i = pairs[p].first, j = pairs[p].second;
            for (int k = i; k < n; ++k)
                for (int l = (k == i) ? j : k; l < n; ++l) {
                    if (use_q && !(Q(i, j) * Q(k, l) >= thr)) continue;
                    const auto& A = b.shell(i);
                    const auto& B = b.shell(j);
                    const auto& C = b.shell(k);
                    const auto& E = b.shell(l);
                    kernels::compute_eri(A, B, C, E, buf);
                    fb.accumulate_symmetric(buf, A.function_index(), B.function_index(),
                                            C.function_index(), E.function_index(), A.n_functions(),
                                            B.n_functions(), C.n_functions(), E.n_functions(), i == j,
                                            k == l, i == k && j == l);
                }
        }
        auto Jl = fb.get_coulomb_matrix();
        auto Kl = fb.get_exchange_matrix();
#pragma omp critical
        for (Size x = 0; x < nbf * nbf; ++x) {
            J[x] += Jl[x];
            K[x] += Kl[x];
        }
    }
}

// The canonical screened loop restricted to the quartets that have at least
// one shell s with mask[s] != 0: the part of J and K (and so of the energy
// 1/2 tr(DJ) - 1/4 tr(DK)) that depends on those shells' centres.  Finite
// differences of the energy in one atom's coordinates need only this part.
void ref_jk_canonical_touching(void* h, const double* D, const double* Q_packed, double thr, const int* mask,
                               double* J, double* K, int n_threads) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const int n = static_cast<int>(b.n_shells());
    const Size nbf = b.n_basis_functions();
    std::memset(J, 0, sizeof(double) * nbf * nbf);
    std::memset(K, 0, sizeof(double) * nbf * nbf);
    auto Q = [&](int i, int j) {
        if (i > j) std::swap(i, j);
        return Q_packed[static_cast<Size>(j) * (j + 1) / 2 + i];
    };
    std::vector<std::pair<int, int>> pairs;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) pairs.emplace_back(i, j);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
        consumers::FockBuilder fb(nbf);
        fb.set_density(D, nbf);
        TwoElectronBuffer<0> buf;
#pragma omp for schedule(dynamic, 1)
        for (Size p = 0; p < pairs.size(); ++p) {
            const int i = pairs[p].first, j = pairs[p].second;
            const bool bra_in = mask[i] || mask[j];
            for (int k = i; k < n; ++k)
                for (int l = (k == i) ? j : k; l < n; ++l) {
                    if (!bra_in && !mask[k] && !mask[l]) continue;
                    if (!(Q(i, j) * Q(k, l) >= thr)) continue;
                    const auto& A = b.shell(i);
                    const auto& B = b.shell(j);
                    const auto& C = b.shell(k);
                    const auto& E = b.shell(l);
                    kernels::compute_eri(A, B, C, E, buf);
                    fb.accumulate_symmetric(buf, A.function_index(), B.function_index(), C.function_index(),
                                            E.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                                            E.n_functions(), i == j, k == l, i == k && j == l);
                }
        }
        auto Jl = fb.get_coulomb_matrix();
        auto Kl = fb.get_exchange_matrix();
#pragma omp critical
        for (Size x = 0; x < nbf * nbf; ++x) {
            J[x] += Jl[x];
            K[x] += Kl[x];
        }
    }
}

// The same canonical screened loop over a sample of bra pairs (every
// stride-th pair of the reference order, starting at offset); counts the
// quartets and ERIs computed so a bounded sample can be timed.
void ref_jk_canonical_sample(void* h, const double* D, const double* Q_packed, double thr, int stride, int offset,
                             double* J, double* K, int n_threads, long long* n_quartets, long long* n_eri) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const int n = static_cast<int>(b.n_shells());
    const Size nbf = b.n_basis_functions();
    std::memset(J, 0, sizeof(double) * nbf * nbf);
    std::memset(K, 0, sizeof(double) * nbf * nbf);
    auto Q = [&](int i, int j) {
        if (i > j) std::swap(i, j);
        return Q_packed[static_cast<Size>(j) * (j + 1) / 2 + i];
    };
    std::vector<std::pair<int, int>> pairs;
    int p = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j, ++p)
            if (p % stride == offset) pairs.emplace_back(i, j);
    long long nq = 0, ne = 0;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads) reduction(+ : nq, ne)
    {
        consumers::FockBuilder fb(nbf);
        fb.set_density(D, nbf);
        TwoElectronBuffer<0> buf;
#pragma omp for schedule(dynamic, 1)
        for (Size t = 0; t < pairs.size(); ++t) {
            const int i = pairs[t].first, j = pairs[t].second;
            for (int k = i; k < n; ++k)
                for (int l = (k == i) ? j : k; l < n; ++l) {
                    if (!(Q(i, j) * Q(k, l) >= thr)) continue;
                    const auto& A = b.shell(i);
                    const auto& B = b.shell(j);
                    const auto& C = b.shell(k);
                    const auto& E = b.shell(l);
                    kernels::compute_eri(A, B, C, E, buf);
                    fb.accumulate_symmetric(buf, A.function_index(), B.function_index(), C.function_index(),
                                            E.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                                            E.n_functions(), i == j, k == l, i == k && j == l);
                    ++nq;
                    ne += static_cast<long long>(A.n_functions()) * B.n_functions() * C.n_functions() *
                          E.n_functions();
                }
        }
        auto Jl = fb.get_coulomb_matrix();
        auto Kl = fb.get_exchange_matrix();
#pragma omp critical
        for (Size x = 0; x < nbf * nbf; ++x) {
            J[x] += Jl[x];
            K[x] += Kl[x];
        }
    }
    *n_quartets = nq;
    *n_eri = ne;
}

// Default CPU path (cpu_engine.hpp:553-622): all n_shells^4 quartets,
// dynamic,8, the reference FockBuilder's thread-local strategy.  [i_begin,
// i_end) restricts the first shell so a bounded sample can be timed.
void ref_jk_full(void* h, const double* D, double* J, double* K, int n_threads, int i_begin,
                 int i_end) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const long long n = static_cast<long long>(b.n_shells());
    const Size nbf = b.n_basis_functions();
    if (i_begin < 0) i_begin = 0;
    if (i_end <= 0 || i_end > n) i_end = static_cast<int>(n);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    consumers::FockBuilder fb(nbf);
    fb.set_density(D, nbf);
    fb.set_threading_strategy(consumers::FockThreadingStrategy::ThreadLocal);
    fb.prepare_parallel(n_threads);
    const long long total = static_cast<long long>(i_end - i_begin) * n * n * n;
#pragma omp parallel num_threads(n_threads)
    {
        TwoElectronBuffer<0> buf;
#pragma omp for schedule(dynamic, 8)
        for (long long t = 0; t < total; ++t) {
            const int l = static_cast<int>(t % n);
            const int k = static_cast<int>((t / n) % n);
            const int j = static_cast<int>((t / (n * n)) % n);
            const int i = i_begin + static_cast<int>(t / (n * n * n));
            const auto& A = b.shell(i);
            const auto& B = b.shell(j);
            const auto& C = b.shell(k);
            const auto& E = b.shell(l);
            kernels::compute_eri(A, B, C, E, buf);
            fb.accumulate(buf, A.function_index(), B.function_index(), C.function_index(),
                          E.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                          E.n_functions());
        }
    }
    fb.finalize_parallel();
    auto Jm = fb.get_coulomb_matrix();
    auto Km = fb.get_exchange_matrix();
    std::memcpy(J, Jm.data(), sizeof(double) * nbf * nbf);
    std::memcpy(K, Km.data(), sizeof(double) * nbf * nbf);
}

// The same default path restricted to the given (i, j) first-shell pairs
// (n_ij pairs, ij[2t], ij[2t+1]) with every (k, l): a bounded sample of the
// N^4 loop for timing; J and K hold that part.
void ref_jk_full_pairs(void* h, const double* D, double* J, double* K, int n_threads, const int* ij, int n_ij) {
    const auto& b = *static_cast<RefBasis*>(h)->basis;
    const long long n = static_cast<long long>(b.n_shells());
    const Size nbf = b.n_basis_functions();
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    consumers::FockBuilder fb(nbf);
    fb.set_density(D, nbf);
    fb.set_threading_strategy(consumers::FockThreadingStrategy::ThreadLocal);
    fb.prepare_parallel(n_threads);
    const long long total = static_cast<long long>(n_ij) * n * n;
#pragma omp parallel num_threads(n_threads)
    {
        TwoElectronBuffer<0> buf;
#pragma omp for schedule(dynamic, 8)
        for (long long t = 0; t < total; ++t) {
            const int l = static_cast<int>(t % n);
            const int k = static_cast<int>((t / n) % n);
            const long long p = t / (n * n);
            const int i = ij[2 * p], j = ij[2 * p + 1];
            const auto& A = b.shell(i);
            const auto& B = b.shell(j);
            const auto& C = b.shell(k);
            const auto& E = b.shell(l);
            kernels::compute_eri(A, B, C, E, buf);
            fb.accumulate(buf, A.function_index(), B.function_index(), C.function_index(),
                          E.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                          E.n_functions());
        }
    }
    fb.finalize_parallel();
    auto Jm = fb.get_coulomb_matrix();
    auto Km = fb.get_exchange_matrix();
    std::memcpy(J, Jm.data(), sizeof(double) * nbf * nbf);
    std::memcpy(K, Km.data(), sizeof(double) * nbf * nbf);
}

// Density-fitted J and K of the reference DFFockBuilder
// (src/shared/consumers/df_fock_builder.cpp) for the orbital basis orb and the
// fitting basis aux (a basis handle of this file; its shells become the
// AuxiliaryBasisSet).  metric 0 = Coulomb, 1 = overlap.  Returns 0, or -1
// when the builder throws.  With Db != NULL: the UHF path
// (set_density_unrestricted), K = Ka + Kb and Ka, Kb returned as well.
int ref_df_jk(void* orb, void* aux, int metric, const double* Da, const double* Db, double* J, double* K,
              double* Ka, double* Kb) {
    const auto& ob = *static_cast<RefBasis*>(orb)->basis;
    const auto& ab = *static_cast<RefBasis*>(aux)->basis;
    try {
        AuxiliaryBasisSet as(ab.shells(), FittingType::JKFIT, "aux");
        consumers::DFFockBuilderConfig cfg;
        cfg.metric_type = metric == 0 ? consumers::DFMetricType::Coulomb : consumers::DFMetricType::Overlap;
        consumers::DFFockBuilder f(ob, as, cfg);
        const Size n = ob.n_basis_functions();
        if (Db)
            f.set_density_unrestricted(std::span<const Real>(Da, n * n), std::span<const Real>(Db, n * n));
        else
            f.set_density(std::span<const Real>(Da, n * n));
        const auto j = f.compute_coulomb();
        const auto k = f.compute_exchange();
        std::memcpy(J, j.data(), sizeof(double) * n * n);
        std::memcpy(K, k.data(), sizeof(double) * n * n);
        if (Db) {
            std::memcpy(Ka, f.exchange_matrix_alpha().data(), sizeof(double) * n * n);
            std::memcpy(Kb, f.exchange_matrix_beta().data(), sizeof(double) * n * n);
        }
    } catch (const std::exception& e) {
        g_df_error = e.what();
        return -1;
    }
    return 0;
}

const char* ref_df_last_error() { return g_df_error.c_str(); }

// The reference's two-centre metric (P|Q) of the shells of aux (n_aux^2,
// row-major) and three-centre tensor (ab|P) in the abP layout.
void ref_df_metric(void* aux, double* V) {
    const auto& ab = *static_cast<RefBasis*>(aux)->basis;
    kernels::compute_two_center_metric(ab.shells(), V, ab.n_basis_functions());
}

void ref_df_three_center(void* orb, void* aux, double* T) {
    const auto& ob = *static_cast<RefBasis*>(orb)->basis;
    const auto& ab = *static_cast<RefBasis*>(aux)->basis;
    kernels::compute_three_center_tensor(ob.shells(), ab.shells(), T, ob.n_basis_functions(), ab.n_basis_functions(),
                                         kernels::ThreeCenterStorageFormat::abP);
}

}  // extern "C"
