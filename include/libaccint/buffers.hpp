// buffers.hpp -- integral buffers with the reference layouts
// (include/libaccint/buffers/two_electron_buffer.hpp:41-307: row-major
// [a][b][c][d]; one_electron_buffer.hpp:44-223: row-major [a][b]).
#pragma once

#include <libaccint/core.hpp>

#include <algorithm>
#include <span>
#include <type_traits>
#include <vector>

namespace libaccint {

// Derivative components of an integral over n_centers centres: 1, 3 n, and
// the 3n(3n+1)/2 upper-triangle second derivatives
// (include/libaccint/core/types.hpp:102-114).
template <int DerivOrder, int NCenters>
[[nodiscard]] consteval int n_derivative_components() {
    if constexpr (DerivOrder == 0) return 1;
    else if constexpr (DerivOrder == 1) return 3 * NCenters;
    else return 3 * NCenters * (3 * NCenters + 1) / 2;
}

// Four-centre integral block in FP64 or FP32; derivative components are the
// slowest index, [deriv][a][b][c][d] with deriv = centre * 3 + direction for
// first derivatives (include/libaccint/core/derivative_utils.hpp:69-73).
template <int DerivOrder = 0, typename RealType = Real>
class TwoElectronBuffer {
    static_assert(DerivOrder >= 0 && DerivOrder <= 2, "derivative order 0, 1 or 2");
    static_assert(std::is_same_v<RealType, float> || std::is_same_v<RealType, double>, "float or double");

public:
    using Scalar = RealType;
    static constexpr int N_DERIV = n_derivative_components<DerivOrder, 4>();
    static constexpr bool is_single_precision = std::is_same_v<RealType, float>;
    static constexpr bool is_double_precision = std::is_same_v<RealType, double>;

    TwoElectronBuffer() = default;
    TwoElectronBuffer(int na, int nb, int nc, int nd) { resize(na, nb, nc, nd); }
    void resize(int na, int nb, int nc, int nd) {
        na_ = na;
        nb_ = nb;
        nc_ = nc;
        nd_ = nd;
        data_.resize(static_cast<Size>(N_DERIV) * na * nb * nc * nd);
    }
    void clear() noexcept { std::fill(data_.begin(), data_.end(), RealType(0)); }
    [[nodiscard]] RealType& operator()(int a, int b, int c, int d)
        requires(DerivOrder == 0)
    {
        return data_[index(a, b, c, d)];
    }
    [[nodiscard]] RealType operator()(int a, int b, int c, int d) const
        requires(DerivOrder == 0)
    {
        return data_[index(a, b, c, d)];
    }
    [[nodiscard]] RealType& operator()(int a, int b, int c, int d, int component)
        requires(DerivOrder > 0)
    {
        return data_[static_cast<Size>(component) * n_integrals() + index(a, b, c, d)];
    }
    [[nodiscard]] RealType operator()(int a, int b, int c, int d, int component) const
        requires(DerivOrder > 0)
    {
        return data_[static_cast<Size>(component) * n_integrals() + index(a, b, c, d)];
    }
    [[nodiscard]] std::span<RealType> data() noexcept { return data_; }
    [[nodiscard]] std::span<const RealType> data() const noexcept { return data_; }
    [[nodiscard]] RealType* data_ptr() noexcept { return data_.data(); }
    [[nodiscard]] const RealType* data_ptr() const noexcept { return data_.data(); }
    [[nodiscard]] Size size() const noexcept { return data_.size(); }
    [[nodiscard]] Size size_bytes() const noexcept { return data_.size() * sizeof(RealType); }
    [[nodiscard]] Size n_integrals() const noexcept { return static_cast<Size>(na_) * nb_ * nc_ * nd_; }
    [[nodiscard]] bool empty() const noexcept { return data_.empty(); }
    [[nodiscard]] int na() const noexcept { return na_; }
    [[nodiscard]] int nb() const noexcept { return nb_; }
    [[nodiscard]] int nc() const noexcept { return nc_; }
    [[nodiscard]] int nd() const noexcept { return nd_; }
    // element-wise conversion to another precision
    template <typename Target>
    [[nodiscard]] TwoElectronBuffer<DerivOrder, Target> to_precision() const {
        TwoElectronBuffer<DerivOrder, Target> r(na_, nb_, nc_, nd_);
        std::transform(data_.begin(), data_.end(), r.data().begin(), [](RealType v) { return Target(v); });
        return r;
    }
    template <typename Source>
    void copy_from(const TwoElectronBuffer<DerivOrder, Source>& src) {
        resize(src.na(), src.nb(), src.nc(), src.nd());
        std::transform(src.data().begin(), src.data().end(), data_.begin(), [](Source v) { return RealType(v); });
    }

private:
    [[nodiscard]] Size index(int a, int b, int c, int d) const noexcept {
        return ((static_cast<Size>(a) * nb_ + b) * nc_ + c) * nd_ + d;
    }
    int na_{0}, nb_{0}, nc_{0}, nd_{0};
    std::vector<RealType> data_;
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
