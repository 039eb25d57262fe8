// Owning, move-only device (HBM) allocation.
//
// Replaces CudaBuffer<T> (include/pbf/cuda_buffer.hpp:11-139) with the same
// semantics: allocate-before-release reallocation (strong guarantee),
// zero-sized buffers own nothing, fillBytes has memset byte semantics, and
// host copies check count and null pointers before touching the device.
#pragma once

#include <cstddef>
#include <limits>
#include <stdexcept>
#include <utility>

#include "pbf/types.hpp"

namespace pbf {

template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t count) { allocate(count); }

    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    DeviceBuffer(DeviceBuffer&& other) noexcept
        : ptr_(std::exchange(other.ptr_, nullptr)), count_(std::exchange(other.count_, 0)) {}

    DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
        if (this != &other) {
            release();
            ptr_ = std::exchange(other.ptr_, nullptr);
            count_ = std::exchange(other.count_, 0);
        }
        return *this;
    }

    ~DeviceBuffer() noexcept { release(); }

    void allocate(std::size_t count) {
        if (count == 0) {
            release();
            return;
        }
        if (count > std::numeric_limits<std::size_t>::max() / sizeof(T))
            throw std::length_error("Device buffer size overflows size_t");
        void* next = nullptr;
        check(pbf_device_alloc(&next, count * sizeof(T)));
        release();
        ptr_ = static_cast<T*>(next);
        count_ = count;
    }

    void fillBytes(int byteValue) {
        if (count_ == 0) return;
        check(pbf_memset(ptr_, byteValue, count_ * sizeof(T)));
    }

    void copyFromHostToDevice(const T* source, std::size_t count) {
        if (count > count_) throw std::out_of_range("Source is larger than device buffer");
        if (source == nullptr && count != 0) throw std::invalid_argument("Source cannot be null");
        if (count != 0) check(pbf_memcpy_htod(ptr_, source, count * sizeof(T)));
    }

    void copyFromDeviceToHost(T* destination, std::size_t count) const {
        if (count > count_) throw std::out_of_range("Copy exceeds device buffer size");
        if (destination == nullptr && count != 0)
            throw std::invalid_argument("Destination cannot be null");
        if (count != 0) check(pbf_memcpy_dtoh(destination, ptr_, count * sizeof(T)));
    }

    T* data() noexcept { return ptr_; }
    const T* data() const noexcept { return ptr_; }
    std::size_t size() const noexcept { return count_; }

private:
    void release() noexcept {
        if (ptr_ != nullptr) {
            pbf_device_free(ptr_);
            ptr_ = nullptr;
            count_ = 0;
        }
    }

    T* ptr_ = nullptr;
    std::size_t count_ = 0;
};

}  // namespace pbf
