// buffers.hpp -- integral buffers with the reference layouts
// (include/libaccint/buffers/two_electron_buffer.hpp:41-307: row-major
// [a][b][c][d]; one_electron_buffer.hpp:44-223: row-major [a][b]).
#pragma once

#include <libaccint/core.hpp>

#include <algorithm>
#include <span>
#include <vector>

namespace libaccint {

template <int DerivOrder = 0>
class TwoElectronBuffer {
    static_assert(DerivOrder == 0, "only DerivOrder == 0 buffers are materialised");

public:
    TwoElectronBuffer() = default;
    TwoElectronBuffer(int na, int nb, int nc, int nd) { resize(na, nb, nc, nd); }
    void resize(int na, int nb, int nc, int nd) {
        na_ = na;
        nb_ = nb;
        nc_ = nc;
        nd_ = nd;
        data_.resize(static_cast<Size>(na) * nb * nc * nd);
    }
    void clear() { std::fill(data_.begin(), data_.end(), Real(0)); }
    [[nodiscard]] Real& operator()(int a, int b, int c, int d) {
        return data_[((static_cast<Size>(a) * nb_ + b) * nc_ + c) * nd_ + d];
    }
    [[nodiscard]] Real operator()(int a, int b, int c, int d) const {
        return data_[((static_cast<Size>(a) * nb_ + b) * nc_ + c) * nd_ + d];
    }
    [[nodiscard]] std::span<Real> data() noexcept { return data_; }
    [[nodiscard]] std::span<const Real> data() const noexcept { return data_; }
    [[nodiscard]] Size size() const noexcept { return data_.size(); }
    [[nodiscard]] int na() const noexcept { return na_; }
    [[nodiscard]] int nb() const noexcept { return nb_; }
    [[nodiscard]] int nc() const noexcept { return nc_; }
    [[nodiscard]] int nd() const noexcept { return nd_; }

private:
    int na_{0}, nb_{0}, nc_{0}, nd_{0};
    std::vector<Real> data_;
};

template <int DerivOrder = 0>
class OneElectronBuffer {
    static_assert(DerivOrder == 0, "only DerivOrder == 0 buffers are materialised");

public:
    OneElectronBuffer() = default;
    OneElectronBuffer(int na, int nb) { resize(na, nb); }
    void resize(int na, int nb) {
        na_ = na;
        nb_ = nb;
        data_.resize(static_cast<Size>(na) * nb);
    }
    void clear() { std::fill(data_.begin(), data_.end(), Real(0)); }
    [[nodiscard]] Real& operator()(int a, int b) { return data_[static_cast<Size>(a) * nb_ + b]; }
    [[nodiscard]] Real operator()(int a, int b) const { return data_[static_cast<Size>(a) * nb_ + b]; }
    [[nodiscard]] std::span<const Real> data() const noexcept { return data_; }
    [[nodiscard]] int na() const noexcept { return na_; }
    [[nodiscard]] int nb() const noexcept { return nb_; }

private:
    int na_{0}, nb_{0};
    std::vector<Real> data_;
};

using OverlapBuffer = OneElectronBuffer<0>;
using KineticBuffer = OneElectronBuffer<0>;
using NuclearBuffer = OneElectronBuffer<0>;

}  // namespace libaccint
