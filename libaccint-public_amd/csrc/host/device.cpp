// device.cpp -- status-code translation, the process-wide context and RAII
// device arrays over the C ABI.
#include <libaccint/device.hpp>

#include <mutex>
#include <string>

namespace libaccint::device {

void check(int status, const char* what) {
    if (status == LAI_OK) return;
    const std::string msg = std::string(what) + ": " + lai_dev_last_error();
    switch (status) {
        case LAI_ERR_INVALID_ARGUMENT: throw InvalidArgumentException(msg);
        case LAI_ERR_INVALID_STATE: throw InvalidStateException(msg);
        case LAI_ERR_MEMORY: throw MemoryException(msg);
        case LAI_ERR_UNSUPPORTED: throw std::invalid_argument(msg);
        case LAI_ERR_NUMERIC: throw std::runtime_error(msg);
        default: throw BackendError(msg);
    }
}

bool gpu_visible() noexcept {
    int n = 0;
    return lai_dev_device_count(&n) == LAI_OK && n > 0;
}

lai_ctx* context() {
    static std::mutex m;
    static lai_ctx* ctx = nullptr;
    std::lock_guard<std::mutex> lk(m);
    if (!ctx) {
        lai_ctx* c = nullptr;
        const int rc = lai_dev_init(0, &c);
        if (rc != LAI_OK) throw BackendError(std::string("HIP backend unavailable: ") + lai_dev_last_error());
        ctx = c;
    }
    return ctx;
}

DeviceArray::DeviceArray(Size n) : n_(n) {
    void* p = nullptr;
    check(lai_dev_malloc(context(), n * sizeof(double), &p), "DeviceArray");
    p_ = static_cast<double*>(p);
}

DeviceArray::~DeviceArray() {
    if (p_) (void)lai_dev_free(nullptr, p_);
}

DeviceArray& DeviceArray::operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
        if (p_) (void)lai_dev_free(nullptr, p_);
        p_ = o.p_;
        n_ = o.n_;
        o.p_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}

void DeviceArray::upload(const double* host, Size n, hipStream_t s) {
    if (n > n_) throw InvalidArgumentException("DeviceArray::upload: size exceeds allocation");
    check(lai_dev_memcpy_h2d(p_, host, n * sizeof(double), s), "DeviceArray::upload");
}

void DeviceArray::download(double* host, Size n, hipStream_t s) const {
    if (n > n_) throw InvalidArgumentException("DeviceArray::download: size exceeds allocation");
    check(lai_dev_memcpy_d2h(host, p_, n * sizeof(double), s), "DeviceArray::download");
}

void DeviceArray::zero(hipStream_t s) { check(lai_dev_zero(p_, n_, s), "DeviceArray::zero"); }

DeviceBytes::DeviceBytes(Size n) : n_(n) {
    check(lai_dev_malloc(context(), n, &p_), "DeviceBytes");
}

DeviceBytes::~DeviceBytes() {
    if (p_) (void)lai_dev_free(nullptr, p_);
}

DeviceBytes& DeviceBytes::operator=(DeviceBytes&& o) noexcept {
    if (this != &o) {
        if (p_) (void)lai_dev_free(nullptr, p_);
        p_ = o.p_;
        n_ = o.n_;
        o.p_ = nullptr;
        o.n_ = 0;
    }
    return *this;
}

void DeviceBytes::upload(const void* host, Size n, hipStream_t s) {
    if (n > n_) throw InvalidArgumentException("DeviceBytes::upload: size exceeds allocation");
    check(lai_dev_memcpy_h2d(p_, host, n, s), "DeviceBytes::upload");
}

void stream_wait(hipStream_t waiter, hipStream_t producer) {
    check(lai_dev_stream_wait(waiter, producer), "stream_wait");
}

}  // namespace libaccint::device
