// device.hpp -- host-side handle on the gfx950 backend behind the C ABI
// (include/lai_device.h).  Converts status codes into the reference's
// exception types (src/device/cuda/engine/cuda_engine.cu:183-187 BackendError,
// InvalidArgumentException, InvalidStateException, MemoryException).
#pragma once

#include <libaccint/core.hpp>

#include "../lai_device.h"

#include <string>

typedef struct ihipStream_t* hipStream_t;

namespace libaccint::device {

// Throws the exception type matching a non-zero lai_dev status.
void check(int status, const char* what);

// True when a HIP device is visible (no context is created).
[[nodiscard]] bool gpu_visible() noexcept;

// Process-wide context on device 0, created on first use; throws BackendError
// when no device is available.
[[nodiscard]] lai_ctx* context();

// RAII device allocation of doubles.
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(Size n_doubles);
    ~DeviceArray();
    DeviceArray(DeviceArray&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceArray& operator=(DeviceArray&& o) noexcept;
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    [[nodiscard]] double* get() const noexcept { return p_; }
    [[nodiscard]] Size size() const noexcept { return n_; }
    void upload(const double* host, Size n, hipStream_t s = nullptr);
    void download(double* host, Size n, hipStream_t s = nullptr) const;
    void zero(hipStream_t s = nullptr);

private:
    double* p_{nullptr};
    Size n_{0};
};

}  // namespace libaccint::device
