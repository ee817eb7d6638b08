// Host-side error type and owning device allocation shared by the tracer and
// the device BVH builder.
#pragma once

#include "../../include/hdcx.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <new>
#include <stdexcept>
#include <string>

namespace hdcx {

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void Check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Failure(HDCX_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    throw Failure(HDCX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Thread-local message of the last failed C-ABI call (hdcx_last_error).
inline std::string& LastError()
{
    thread_local std::string message;
    return message;
}

// Runs one C-ABI entry point: exceptions become status codes and a message.
template <class F>
int Guard(F&& f)
{
    try {
        f();
        LastError().clear();
        return HDCX_OK;
    } catch (const Failure& e) {
        LastError() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        LastError() = "host allocation failed";
        return HDCX_ERR_OUT_OF_MEMORY;
    } catch (const std::invalid_argument& e) {
        LastError() = e.what();
        return HDCX_ERR_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        LastError() = e.what();
        return HDCX_ERR_INTERNAL;
    }
}

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_)
    {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            Reset();
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { Reset(); }
    void Reset()
    {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    // Exact size; keeps the allocation when the size is unchanged.
    void Allocate(std::size_t bytes)
    {
        if (bytes == bytes_ && ptr_) return;
        Reset();
        if (bytes == 0) return;
        Check(hipMalloc(&ptr_, bytes), "hipMalloc");
        bytes_ = bytes;
    }
    // At least `bytes`; keeps a larger allocation.
    void Reserve(std::size_t bytes)
    {
        if (bytes <= bytes_ && ptr_) return;
        Allocate(bytes);
    }
    void Upload(const void* host, std::size_t bytes, hipStream_t st)
    {
        Allocate(bytes);
        if (bytes) Check(hipMemcpyAsync(ptr_, host, bytes, hipMemcpyHostToDevice, st), "upload");
    }
    template <class T> T* as() const { return static_cast<T*>(ptr_); }
    [[nodiscard]] std::size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

} // namespace hdcx
