// core.hpp -- scalar types, geometry, exceptions and backend vocabulary of the
// host layer.  Names and meanings follow the reference's
// include/libaccint/core/types.hpp:24-33,141-170,
// include/libaccint/utils/error_handling.hpp:12-95,
// include/libaccint/core/backend.hpp:25-127 and
// include/libaccint/engine/dispatch_policy.hpp:41-93.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>

namespace libaccint {

using Real = double;
using Float = float;
using Index = std::int64_t;
using Size = std::size_t;

inline constexpr int MAX_ANGULAR_MOMENTUM = 3;  // f: the highest class of the gfx950 kernels

enum class AngularMomentum : int { S = 0, P = 1, D = 2, F = 3 };

[[nodiscard]] constexpr int to_int(AngularMomentum am) noexcept { return static_cast<int>(am); }
[[nodiscard]] constexpr int n_cartesian(int l) noexcept { return (l + 1) * (l + 2) / 2; }
[[nodiscard]] constexpr int n_spherical(int l) noexcept { return 2 * l + 1; }

struct Point3D {
    Real x{0.0}, y{0.0}, z{0.0};
    constexpr Point3D() = default;
    constexpr Point3D(Real x_, Real y_, Real z_) : x(x_), y(y_), z(z_) {}
    [[nodiscard]] constexpr Real distance_squared(const Point3D& o) const noexcept {
        const Real dx = x - o.x, dy = y - o.y, dz = z - o.z;
        return dx * dx + dy * dy + dz * dz;
    }
};

// ---- exceptions (error_handling.hpp) --------------------------------------
class Exception : public std::runtime_error {
public:
    explicit Exception(const std::string& m) : std::runtime_error(m) {}
};
class InvalidArgumentException : public Exception {
public:
    explicit InvalidArgumentException(const std::string& m) : Exception(m) {}
};
class InvalidStateException : public Exception {
public:
    explicit InvalidStateException(const std::string& m) : Exception(m) {}
};
class MemoryException : public Exception {
public:
    explicit MemoryException(const std::string& m) : Exception(m) {}
};
class NotImplementedException : public Exception {
public:
    explicit NotImplementedException(const std::string& m) : Exception(m) {}
};
class BackendError : public Exception {
public:
    explicit BackendError(const std::string& m) : Exception(m) {}
};

// ---- backends -------------------------------------------------------------
// The device backend of this library is HIP on gfx950; there is no host
// integral backend (the CPU value is kept for API compatibility and is never
// selected for integral work).
enum class BackendType { CPU, HIP };

[[nodiscard]] constexpr std::string_view backend_name(BackendType b) noexcept {
    return b == BackendType::CPU ? "CPU" : "HIP";
}

enum class BackendHint { Auto, ForceCPU, ForceGPU, PreferCPU, PreferGPU };

struct DispatchConfig {
    Size min_gpu_batch_size = 16;
    Size min_gpu_primitives = 1000;
    int high_am_threshold = 4;
    Size min_gpu_shells = 10;
    bool enable_auto_tuning = false;
    Size auto_tune_min_batch = 100;
    Size n_gpu_slots = 4;
};

// Kinds of integral work a dispatch decision is asked for
// (include/libaccint/engine/dispatch_policy.hpp:28-34).
enum class WorkUnitType { SingleShellPair, SingleShellQuartet, ShellSetPair, ShellSetQuartet, FullBasis };

// The reference's CPU/GPU routing rule (src/host/engine/dispatch_policy.cpp:12-90):
// force hints first, CPU without a GPU, then per work type thresholds of the
// config, moved by the Prefer hints.  In this library every integral runs on
// the device whatever the decision: the rule is kept so that callers that
// consult it see the reference's answer (HIP where the reference says CUDA).
class DispatchPolicy {
public:
    DispatchPolicy() = default;
    explicit DispatchPolicy(DispatchConfig config) : config_(config) {}
    [[nodiscard]] BackendType select_backend(WorkUnitType work_type, Size batch_size, int total_am,
                                             Size n_primitives, BackendHint hint, bool gpu_available) const;
    [[nodiscard]] const DispatchConfig& config() const noexcept { return config_; }
    void set_config(DispatchConfig config) noexcept { config_ = config; }
    void set_auto_tuning(bool enable) noexcept { config_.enable_auto_tuning = enable; }
    [[nodiscard]] bool auto_tuning_enabled() const noexcept { return config_.enable_auto_tuning; }

private:
    DispatchConfig config_;
};

// Precision vocabulary (include/libaccint/core/precision.hpp:310-352).
enum class Precision { Float32, Float64 };
enum class MixedPrecisionMode : std::uint8_t { Pure32 = 0, Pure64 = 1, Compute32Accumulate64 = 2, Adaptive = 3 };
[[nodiscard]] constexpr const char* mixed_precision_to_string(MixedPrecisionMode m) noexcept {
    switch (m) {
        case MixedPrecisionMode::Pure32: return "pure32";
        case MixedPrecisionMode::Pure64: return "pure64";
        case MixedPrecisionMode::Compute32Accumulate64: return "compute32_accumulate64";
        case MixedPrecisionMode::Adaptive: return "adaptive";
    }
    return "unknown";
}

}  // namespace libaccint
