// data.hpp -- atoms and basis-set construction.
//
//  * create_builtin_basis("sto-3g", atoms): STO-3G for H, C, N, O, F
//    (reference src/shared/data/builtin_basis.cpp:24-203).
//  * BseJsonParser: Basis Set Exchange JSON -> shells, one shell per
//    coefficient column, all-zero columns skipped, zero coefficients inside
//    a column kept (reference src/shared/data/bse_json_parser.cpp:27-140).
//  * load_basis_set(name, atoms): built-in names first, then
//    <data dir>/<lower-case name>.json; the data directory is
//    $LIBACCINT_DATA_DIR or the package's share/basis_sets
//    (reference src/shared/data/basis_parser.cpp:81).
#pragma once

#include <libaccint/basis.hpp>

#include <string>
#include <vector>

namespace libaccint::data {

struct Atom {
    int atomic_number{0};
    Point3D position{};
};

[[nodiscard]] BasisSet create_sto3g(const std::vector<Atom>& atoms);
[[nodiscard]] BasisSet create_builtin_basis(const std::string& name, const std::vector<Atom>& atoms);

class BseJsonParser {
public:
    [[nodiscard]] static BasisSet parse(const std::string& json_text, const std::vector<Atom>& atoms);
    [[nodiscard]] static BasisSet parse_file(const std::string& path, const std::vector<Atom>& atoms);
};

[[nodiscard]] std::string basis_data_dir();
[[nodiscard]] BasisSet load_basis_set(const std::string& name, const std::vector<Atom>& atoms);
[[nodiscard]] std::vector<std::string> list_available_basis_sets();

}  // namespace libaccint::data
