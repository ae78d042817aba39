// Device plumbing of the drop-in layer: one dvm_ctx_t per process (device 0
// unless DVM_DEVICE is set) and an RAII device buffer.  The reference classes
// keep their data in host memory (callers write the public fields directly,
// src/DVMBase.cpp:179-180, tests/integration/test_BiotSavart.cpp:24-33), so each
// drop-in method stages its operands to the device, runs the HIP kernels and
// copies results back.
#ifndef DVMPP_AMD_DEVICE_HPP
#define DVMPP_AMD_DEVICE_HPP

#include <cstddef>
#include <cstdint>

#include "dvm_hip.h"

namespace dvm {

// The process-wide context (created on first use; destroyed at exit).
dvm_ctx_t device_context();

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { reset(bytes); }
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void reset(std::size_t bytes);
    template <class T>
    T* as() const { return static_cast<T*>(ptr_); }
    void upload(const void* src, std::size_t bytes, std::size_t offset = 0);
    void download(void* dst, std::size_t bytes, std::size_t offset = 0) const;

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

}  // namespace dvm

#endif
