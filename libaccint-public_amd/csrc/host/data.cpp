// data.cpp -- built-in STO-3G, a small JSON reader and the BSE basis loader.
#include <libaccint/data.hpp>

#include "json_reader.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace libaccint::data {

// ---------------------------------------------------------------------------
// STO-3G (Hehre, Stewart, Pople, J. Chem. Phys. 51, 2657 (1969)) as published
// by the Basis Set Exchange: a 1s shell and, for C-F, 2s and 2p shells sharing
// exponents.
// ---------------------------------------------------------------------------
namespace {
const std::vector<Real> kSto1s = {0.15432897, 0.53532814, 0.44463454};
const std::vector<Real> kSto2s = {-0.09996723, 0.39951283, 0.70011547};
const std::vector<Real> kSto2p = {0.15591627, 0.60768372, 0.39195739};

struct StoElement {
    std::vector<Real> e1s, e2sp;
};
const std::map<int, StoElement>& sto_table() {
    static const std::map<int, StoElement> t = {
        {1, {{3.42525091, 0.62391373, 0.16885540}, {}}},
        {6, {{71.6168370, 13.0450960, 3.5305122}, {2.9412494, 0.6834831, 0.2222899}}},
        {7, {{99.1061690, 18.0523120, 4.8856602}, {3.7804559, 0.8784966, 0.2857144}}},
        {8, {{130.7093200, 23.8088610, 6.4436083}, {5.0331513, 1.1695961, 0.3803890}}},
        {9, {{166.6791300, 30.3608120, 8.2168207}, {6.4648032, 1.5022812, 0.4885885}}},
    };
    return t;
}
}  // namespace

BasisSet create_sto3g(const std::vector<Atom>& atoms) {
    std::vector<Shell> shells;
    for (Size i = 0; i < atoms.size(); ++i) {
        const auto it = sto_table().find(atoms[i].atomic_number);
        if (it == sto_table().end())
            throw InvalidArgumentException("STO-3G basis not available for element Z=" +
                                           std::to_string(atoms[i].atomic_number) +
                                           ". Supported: H(1), C(6), N(7), O(8), F(9).");
        const auto& el = it->second;
        shells.emplace_back(0, atoms[i].position, el.e1s, kSto1s);
        shells.back().set_atom_index(static_cast<Index>(i));
        if (!el.e2sp.empty()) {
            shells.emplace_back(0, atoms[i].position, el.e2sp, kSto2s);
            shells.back().set_atom_index(static_cast<Index>(i));
            shells.emplace_back(1, atoms[i].position, el.e2sp, kSto2p);
            shells.back().set_atom_index(static_cast<Index>(i));
        }
    }
    return BasisSet(std::move(shells));
}

BasisSet create_builtin_basis(const std::string& name, const std::vector<Atom>& atoms) {
    std::string n = name;
    std::transform(n.begin(), n.end(), n.begin(), [](unsigned char c) { return std::tolower(c); });
    if (n == "sto-3g" || n == "sto3g") return create_sto3g(atoms);
    throw InvalidArgumentException("Unknown built-in basis set: '" + name + "'. Supported: sto-3g");
}

namespace {
using detail::JParser;
using detail::JValue;

Real to_real(const JValue& v) {
    if (v.kind == JValue::String) return std::stod(v.str);  // BSE stores numbers as strings
    if (v.kind == JValue::Number) return v.num;
    throw InvalidArgumentException("BSE JSON: expected a number");
}

// one shell per non-zero coefficient column (bse_json_parser.cpp:27-87)
std::vector<Shell> element_shells(const JValue& el, const Point3D& c) {
    std::vector<Shell> out;
    const JValue* shells = el.get("electron_shells");
    if (!shells) return out;
    for (const auto& sh : shells->arr) {
        std::vector<Real> exps;
        for (const auto& e : sh.get("exponents")->arr) exps.push_back(to_real(e));
        std::vector<int> ams;
        for (const auto& a : sh.get("angular_momentum")->arr) ams.push_back(static_cast<int>(to_real(a)));
        const auto& cols = sh.get("coefficients")->arr;
        auto add_col = [&](int am, const JValue& col) {
            std::vector<Real> cf;
            bool all_zero = true;
            for (const auto& x : col.arr) {
                cf.push_back(to_real(x));
                if (cf.back() != 0.0) all_zero = false;
            }
            if (!all_zero) out.emplace_back(am, c, exps, cf);
        };
        if (ams.size() == 1) {
            for (const auto& col : cols) add_col(ams[0], col);
        } else {
            for (Size k = 0; k < ams.size() && k < cols.size(); ++k) add_col(ams[k], cols[k]);
        }
    }
    return out;
}
}  // namespace

BasisSet BseJsonParser::parse(const std::string& text, const std::vector<Atom>& atoms) {
    JValue root = JParser(text, "BSE JSON").parse();
    if (const JValue* schema = root.get("molssi_bse_schema")) {
        if (const JValue* t = schema->get("schema_type")) {
            if (t->str != "complete" && t->str != "minimal")
                throw InvalidArgumentException("Unsupported BSE schema type: " + t->str +
                                               " (expected 'complete' or 'minimal')");
        }
    }
    const JValue* elements = root.get("elements");
    if (!elements) throw InvalidArgumentException("BSE JSON missing 'elements' key");
    std::vector<Shell> all;
    for (Size i = 0; i < atoms.size(); ++i) {
        const JValue* el = elements->get(std::to_string(atoms[i].atomic_number));
        if (!el)
            throw InvalidArgumentException("BSE JSON has no data for element Z=" +
                                           std::to_string(atoms[i].atomic_number));
        for (auto& s : element_shells(*el, atoms[i].position)) {
            s.set_atom_index(static_cast<Index>(i));
            all.push_back(std::move(s));
        }
    }
    return BasisSet(std::move(all));
}

BasisSet BseJsonParser::parse_file(const std::string& path, const std::vector<Atom>& atoms) {
    std::ifstream f(path);
    if (!f.is_open()) throw InvalidArgumentException("Cannot open BSE JSON file: " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), atoms);
}

std::string basis_data_dir() {
    if (const char* env = std::getenv("LIBACCINT_DATA_DIR"); env && *env) return env;
    Dl_info info{};
    if (dladdr(reinterpret_cast<void*>(&basis_data_dir), &info) && info.dli_fname) {
        namespace fs = std::filesystem;
        fs::path p = fs::path(info.dli_fname).parent_path().parent_path() / "share" / "basis_sets";
        return p.string();
    }
    return "share/basis_sets";
}

BasisSet load_basis_set(const std::string& name, const std::vector<Atom>& atoms) {
    std::string n = name;
    std::transform(n.begin(), n.end(), n.begin(), [](unsigned char c) { return std::tolower(c); });
    if (n == "sto-3g" || n == "sto3g") return create_sto3g(atoms);
    const std::string path = basis_data_dir() + "/" + n + ".json";
    if (!std::filesystem::exists(path))
        throw InvalidArgumentException("Basis set '" + name + "' not found (looked for " + path + ")");
    return BseJsonParser::parse_file(path, atoms);
}

std::vector<std::string> list_available_basis_sets() {
    std::vector<std::string> out = {"sto-3g"};
    namespace fs = std::filesystem;
    std::error_code ec;
    for (const auto& e : fs::directory_iterator(basis_data_dir(), ec))
        if (e.path().extension() == ".json") out.push_back(e.path().stem().string());
    std::sort(out.begin() + 1, out.end());
    return out;
}

}  // namespace libaccint::data
