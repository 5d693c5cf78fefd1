// host_driver.cpp -- exercises the host-only code of libaccint_host under
// AddressSanitizer and UndefinedBehaviorSanitizer (built and run by
// tests/test_host_sanitize.py with g++; no GPU is touched): the BSE JSON basis
// reader and BasisSet organisation, the fitting-set loader, FockBuilder
// (plain and symmetric accumulate; Sequential, Atomic and ThreadLocal
// threading with the LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES fallback),
// MixedPrecisionFockBuilder, the dispatch table JSON reader and writer,
// DispatchPolicy, ScreeningOptions, and the Engine's error paths without a
// device.  Prints "host driver ok" on success; any sanitizer report fails it.
#include <libaccint/data.hpp>
#include <libaccint/df.hpp>
#include <libaccint/engine.hpp>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace libaccint;

namespace {

int g_fail = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++g_fail;
    }
}

template <typename E, typename F>
void expect_throw(F&& f, const char* what) {
    try {
        f();
    } catch (const E&) {
        return;
    } catch (...) {
        std::fprintf(stderr, "FAILED (other exception): %s\n", what);
        ++g_fail;
        return;
    }
    std::fprintf(stderr, "FAILED (no exception): %s\n", what);
    ++g_fail;
}

std::vector<data::Atom> waters(int n) {
    std::vector<data::Atom> a;
    for (int k = 0; k < n; ++k) {
        const double x = 5.5 * k;
        a.push_back({8, {x, 0.0, 0.0}});
        a.push_back({1, {x, 1.43233673, -1.10866041}});
        a.push_back({1, {x, -1.43233673, -1.10866041}});
    }
    return a;
}

void fill_block(TwoElectronBuffer<0>& buf, std::mt19937_64& rng) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& v : buf.data()) v = u(rng);
}

// J and K of random blocks over every shell quartet through one strategy
void fock_strategies(const BasisSet& b) {
    const Size n = b.n_basis_functions();
    std::vector<double> D(n * n);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-0.5, 0.5);
    for (Size i = 0; i < n; ++i)
        for (Size j = 0; j <= i; ++j) D[i * n + j] = D[j * n + i] = u(rng);
    std::vector<std::vector<double>> results;
    for (auto strat : {consumers::FockThreadingStrategy::Sequential, consumers::FockThreadingStrategy::Atomic,
                       consumers::FockThreadingStrategy::ThreadLocal}) {
        consumers::FockBuilder fb(n);
        fb.set_density(D.data(), n);
        fb.set_threading_strategy(strat);
        const int nt = strat == consumers::FockThreadingStrategy::Sequential ? 1 : 4;
        fb.prepare_parallel(nt);
        const Size ns = b.n_shells();
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                std::mt19937_64 r(100 + t);
                TwoElectronBuffer<0> buf;
                for (Size i = static_cast<Size>(t); i < ns; i += static_cast<Size>(nt))
                    for (Size j = 0; j < ns; ++j) {
                        const Shell& A = b.shell(i);
                        const Shell& B = b.shell(j);
                        const Shell& C = b.shell((i + j) % ns);
                        const Shell& E = b.shell((i * 7 + j) % ns);
                        buf.resize(A.n_functions(), B.n_functions(), C.n_functions(), E.n_functions());
                        r.seed(i * 1000003 + j);
                        fill_block(buf, r);
                        fb.accumulate(buf, A.function_index(), B.function_index(), C.function_index(),
                                      E.function_index(), A.n_functions(), B.n_functions(), C.n_functions(),
                                      E.n_functions());
                    }
            });
        for (auto& x : th) x.join();
        fb.finalize_parallel();
        auto J = fb.get_coulomb_matrix();
        results.emplace_back(J.begin(), J.end());
        std::vector<double> H(n * n, 0.25);
        auto F = fb.get_fock_matrix(H, 0.5);
        expect(F.size() == n * n, "Fock size");
    }
    double dmax = 0.0, jmax = 0.0;
    for (Size x = 0; x < n * n; ++x) {
        jmax = std::max(jmax, std::abs(results[0][x]));
        dmax = std::max(dmax, std::max(std::abs(results[1][x] - results[0][x]), std::abs(results[2][x] - results[0][x])));
    }
    expect(jmax > 0.0 && dmax <= 1e-12 * jmax, "threading strategies agree");

    // thread-local buffers above the byte limit: ThreadLocal refuses, the
    // automatic choice falls back to atomics
    setenv("LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES", "64", 1);
    consumers::FockBuilder tl(n);
    tl.set_density(D.data(), n);
    tl.set_threading_strategy(consumers::FockThreadingStrategy::ThreadLocal);
    expect_throw<MemoryException>([&]() { tl.prepare_parallel(4); }, "ThreadLocal above the byte limit");
    consumers::FockBuilder au(n);
    au.set_density(D.data(), n);
    au.prepare_parallel(4);
    TwoElectronBuffer<0> buf(1, 1, 1, 1);
    buf(0, 0, 0, 0) = 1.0;
    au.accumulate(buf, 0, 0, 0, 0, 1, 1, 1, 1);
    au.finalize_parallel();
    expect(au.get_coulomb_matrix()[0] == D[0], "atomic fallback accumulates");
    unsetenv("LIBACCINT_MAX_FOCK_THREADLOCAL_BYTES");

    // symmetric scatter of one canonical quartet vs the plain sum of its
    // eight permutations (symmetric block values)
    consumers::FockBuilder s1(n), s2(n);
    s1.set_density(D.data(), n);
    s2.set_density(D.data(), n);
    Size ia = 0, ib = 0;  // a p shell and a d shell (cc-pVDZ oxygen)
    for (Size s = b.n_shells(); s-- > 0;) {
        if (b.shell(s).angular_momentum() == 1) ia = s;
        if (b.shell(s).angular_momentum() == 2) ib = s;
    }
    const Shell& A = b.shell(ia);
    const Shell& B = b.shell(ib);
    const int na = A.n_functions(), nb = B.n_functions();
    TwoElectronBuffer<0> g(na, nb, na, nb);
    std::mt19937_64 r2(3);
    fill_block(g, r2);
    for (int a = 0; a < na; ++a)  // (ab|cd) = (cd|ab) for the bra == ket quartet
        for (int bb = 0; bb < nb; ++bb)
            for (int c = 0; c < na; ++c)
                for (int d = 0; d < nb; ++d) g(c, d, a, bb) = g(a, bb, c, d);
    s1.accumulate_symmetric(g, A.function_index(), B.function_index(), A.function_index(), B.function_index(), na,
                            nb, na, nb, false, false, true);
    TwoElectronBuffer<0> p(nb, na, nb, na);
    for (int a = 0; a < na; ++a)
        for (int bb = 0; bb < nb; ++bb)
            for (int c = 0; c < na; ++c)
                for (int d = 0; d < nb; ++d) p(bb, a, d, c) = g(a, bb, c, d);
    TwoElectronBuffer<0> p2(na, nb, nb, na), p3(nb, na, na, nb);
    for (int a = 0; a < na; ++a)
        for (int bb = 0; bb < nb; ++bb)
            for (int c = 0; c < na; ++c)
                for (int d = 0; d < nb; ++d) {
                    p2(a, bb, d, c) = g(a, bb, c, d);
                    p3(bb, a, c, d) = g(a, bb, c, d);
                }
    s2.accumulate(g, A.function_index(), B.function_index(), A.function_index(), B.function_index(), na, nb, na, nb);
    s2.accumulate(p, B.function_index(), A.function_index(), B.function_index(), A.function_index(), nb, na, nb, na);
    s2.accumulate(p2, A.function_index(), B.function_index(), B.function_index(), A.function_index(), na, nb, nb, na);
    s2.accumulate(p3, B.function_index(), A.function_index(), A.function_index(), B.function_index(), nb, na, na, nb);
    double e1 = 0.0, m1 = 0.0;
    for (Size x = 0; x < n * n; ++x) {
        e1 = std::max(e1, std::abs(s1.get_coulomb_matrix()[x] - s2.get_coulomb_matrix()[x]) +
                              std::abs(s1.get_exchange_matrix()[x] - s2.get_exchange_matrix()[x]));
        m1 = std::max(m1, std::abs(s2.get_coulomb_matrix()[x]));
    }
    expect(m1 > 0.0 && e1 <= 1e-12 * m1, "symmetric scatter equals the permutation sum");

    consumers::MixedPrecisionFockBuilder mp(n);
    mp.set_density(D.data(), n);
    TwoElectronBuffer<0, float> f32(1, 1, 1, 1);
    f32(0, 0, 0, 0) = 0.5f;
    mp.accumulate(f32, 0, 0, 0, 0, 1, 1, 1, 1);
    expect(mp.n_float32_accumulations() == 1, "mixed-precision counts");
    consumers::FockBuilder nod(n);
    expect_throw<InvalidStateException>([&]() { nod.accumulate(buf, 0, 0, 0, 0, 1, 1, 1, 1); }, "no density");
    expect_throw<InvalidArgumentException>([&]() { nod.set_density(D.data(), n + 1); }, "nbf mismatch");
}

void dispatch_checks() {
    DispatchTable t;
    for (int c = 0; c < 55; c += 7) {
        DispatchEntry e;
        e.cls = c;
        e.am = quartet_class_am(c);
        e.strategy = static_cast<GpuExecutionStrategy>(c % 3);
        e.chunk = 16 * c;
        e.blocks_per_cu = c % 5;
        t.entries.push_back(e);
    }
    const std::string js = t.dump_json();
    DispatchTable u = DispatchTable::parse_json(js);
    expect(u.entries.size() == t.entries.size(), "dispatch table round trip");
    for (Size k = 0; k < u.entries.size() && k < t.entries.size(); ++k)
        expect(u.entries[k].cls == t.entries[k].cls && u.entries[k].chunk == t.entries[k].chunk &&
                   u.entries[k].strategy == t.entries[k].strategy,
               "dispatch entry round trip");
    for (const char* bad : {"", "{", "[{\"class\": \"(ss|ss)\"", "{\"entries\": [1, 2", "[{\"am\": [0, 0, 0]}]",
                            "\"\\u12", "[{\"strategy\": \"nope\"}]"}) {
        try {
            (void)DispatchTable::parse_json(bad);
        } catch (const std::exception&) {
        }
    }
    for (int c = 0; c < 55; ++c) expect(quartet_class_index(quartet_class_am(c)) == c, "class index round trip");
    DispatchPolicy pol;
    for (int am = 0; am <= 12; ++am)
        for (Size batch : {Size(1), Size(16), Size(100000)})
            for (auto h : {BackendHint::Auto, BackendHint::ForceCPU, BackendHint::ForceGPU, BackendHint::PreferCPU,
                           BackendHint::PreferGPU})
                (void)pol.select_backend(WorkUnitType::ShellSetQuartet, batch, am, batch * 81, h, false);
    screening::ScreeningOptions so = screening::ScreeningOptions::normal();
    expect(so.passes_screening(1.0) && !so.passes_screening(0.0), "screening predicate");
}

void basis_checks() {
    const auto atoms = waters(2);
    for (const char* name : {"sto-3g", "cc-pvdz", "def2-svp", "cc-pvtz", "aug-cc-pvtz"}) {
        BasisSet b = data::load_basis_set(name, atoms);
        expect(b.n_shells() > 0 && b.n_basis_functions() > 0, name);
        const auto& q = b.shell_set_quartets();
        const Size np = b.shell_set_pairs().size();
        expect(q.size() == np * (np + 1) / 2, "shell set quartet count");
        if (std::string(name) == "cc-pvdz") fock_strategies(b);
    }
    expect_throw<std::exception>([&]() { (void)data::load_basis_set("no-such-basis", atoms); }, "unknown basis");
    for (const char* bad : {"", "{", "{\"elements\": {\"1\": {\"electron_shells\": [{\"angular_momentum\": [0]",
                            "{\"elements\": {\"1\": {\"electron_shells\": [{\"angular_momentum\": [9], "
                            "\"exponents\": [\"1.0\"], \"coefficients\": [[\"1.0\"]]}]}}}",
                            "{\"elements\": {\"1\": {\"electron_shells\": [{\"angular_momentum\": [0], "
                            "\"exponents\": [\"x\"], \"coefficients\": [[\"1.0\"]]}]}}}"}) {
        try {
            (void)data::BseJsonParser::parse(bad, atoms);
        } catch (const std::exception&) {
        }
    }
    std::vector<int> z;
    std::vector<std::array<Real, 3>> c;
    for (const auto& a : atoms) {
        z.push_back(a.atomic_number);
        c.push_back({a.position.x, a.position.y, a.position.z});
    }
    AuxiliaryBasisSet aux = load_auxiliary_basis("cc-pVDZ-RI", z, c);
    expect(aux.n_functions() > 0, "fitting set");
    for (Size f = 0; f < aux.n_functions(); f += 7)
        expect(aux.shell_to_function(aux.function_to_shell(f)) <= f, "fitting function offsets");
}

void engine_without_device() {
    const auto atoms = waters(1);
    BasisSet b = data::load_basis_set("sto-3g", atoms);
    Engine e(b);
    if (e.gpu_available()) return;  // this driver runs where no device is visible
    std::vector<Real> S;
    expect_throw<BackendError>([&]() { e.compute_overlap_matrix(S); }, "overlap without a device");
    consumers::FockBuilder fb(b.n_basis_functions());
    std::vector<Real> D(b.n_basis_functions() * b.n_basis_functions(), 0.0);
    fb.set_density(D.data(), b.n_basis_functions());
    expect_throw<BackendError>([&]() { e.compute_and_consume(Operator::coulomb(), fb); }, "J/K without a device");
    expect_throw<InvalidArgumentException>([&]() { e.compute_and_consume(Operator::overlap(), fb); },
                                           "one-electron operator to the two-electron route");
}

}  // namespace

int main() {
    basis_checks();
    dispatch_checks();
    engine_without_device();
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("host driver ok\n");
    return 0;
}
