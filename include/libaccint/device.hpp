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

// RAII device allocation of raw bytes (descriptor arrays).
class DeviceBytes {
public:
    DeviceBytes() = default;
    explicit DeviceBytes(Size bytes);
    ~DeviceBytes();
    DeviceBytes(DeviceBytes&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceBytes& operator=(DeviceBytes&& o) noexcept;
    DeviceBytes(const DeviceBytes&) = delete;
    DeviceBytes& operator=(const DeviceBytes&) = delete;
    [[nodiscard]] void* get() const noexcept { return p_; }
    [[nodiscard]] Size bytes() const noexcept { return n_; }
    // synchronous on stream s (the host buffer may be reused on return)
    void upload(const void* host, Size bytes, hipStream_t s = nullptr);

private:
    void* p_{nullptr};
    Size n_{0};
};

// Makes stream waiter wait for the work enqueued so far on producer.
void stream_wait(hipStream_t waiter, hipStream_t producer);

}  // namespace libaccint::device

namespace libaccint {

// A device-resident batch of ERI blocks, one per shell quartet, with the
// device descriptors (function offsets, sizes, offset in the batch) that the
// Fock digestion reads; produced by Engine::compute_eri_batch_device and
// consumed by GpuFockBuilder::accumulate_device_eri_batch.  Owns its memory,
// so the device pointers stay valid while the object lives (the owning handle
// of CudaEngine::DeviceEriBatch, include/libaccint/engine/cuda_engine.hpp:96-135).
class DeviceEriBatch {
public:
    DeviceEriBatch() = default;
    DeviceEriBatch(DeviceEriBatch&&) noexcept = default;
    DeviceEriBatch& operator=(DeviceEriBatch&&) noexcept = default;
    DeviceEriBatch(const DeviceEriBatch&) = delete;
    DeviceEriBatch& operator=(const DeviceEriBatch&) = delete;

    [[nodiscard]] const double* data() const noexcept { return eri_.get(); }
    [[nodiscard]] double* data() noexcept { return eri_.get(); }
    // doubles in the batch
    [[nodiscard]] Size size() const noexcept { return count_; }
    // shell quartets (blocks) in the batch and their device descriptors
    [[nodiscard]] Size n_blocks() const noexcept { return n_blocks_; }
    [[nodiscard]] const lai_eri_block* blocks() const noexcept {
        return static_cast<const lai_eri_block*>(blk_.get());
    }
    // the stream the ERIs were computed on
    [[nodiscard]] hipStream_t stream() const noexcept { return stream_; }
    [[nodiscard]] bool empty() const noexcept { return n_blocks_ == 0 || count_ == 0; }
    explicit operator bool() const noexcept { return !empty(); }

private:
    friend class Engine;
    device::DeviceArray eri_;
    device::DeviceBytes blk_;
    Size count_{0}, n_blocks_{0};
    hipStream_t stream_{nullptr};
};

}  // namespace libaccint
