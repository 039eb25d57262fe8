// Error reporting and device memory helpers of the C ABI.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>

#include "pbf/types.hpp"
#include "pbf_amd.h"
#include "pbf_launch.hpp"

namespace {
thread_local std::string g_last_error;
}

pbf_status pbf_fail(pbf_status status, const char* message) {
    g_last_error = message != nullptr ? message : "";
    return status;
}

pbf_status pbf_hip_check(hipError_t error) {
    if (error == hipSuccess) return PBF_OK;
    return pbf_fail(PBF_ERR_DEVICE, hipGetErrorString(error));
}

namespace pbf {

void throw_status(pbf_status status) {
    const std::string message = pbf_last_error();
    switch (status) {
        case PBF_OK:
            return;
        case PBF_ERR_INVALID_ARGUMENT:
            throw std::invalid_argument(message);
        case PBF_ERR_LOGIC:
            throw std::logic_error(message);
        case PBF_ERR_OUT_OF_RANGE:
            throw std::out_of_range(message);
        case PBF_ERR_LENGTH:
            throw std::length_error(message);
        case PBF_ERR_OVERFLOW:
            throw std::overflow_error(message);
        case PBF_ERR_NONCONVERGENCE:
            throw collision_nonconvergence(message);
        default:
            throw std::runtime_error(message);
    }
}

}  // namespace pbf

extern "C" {

const char* pbf_last_error(void) { return g_last_error.c_str(); }

const char* pbf_version(void) { return "pbf_amd 0.1.0 gfx950"; }

pbf_status pbf_device_alloc(void** ptr, size_t bytes) {
    if (ptr == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null output pointer");
    *ptr = nullptr;
    if (bytes == 0) return PBF_OK;
    return pbf_hip_check(hipMalloc(ptr, bytes));
}

pbf_status pbf_device_free(void* ptr) {
    if (ptr == nullptr) return PBF_OK;
    return pbf_hip_check(hipFree(ptr));
}

pbf_status pbf_host_alloc(void** ptr, size_t bytes) {
    if (ptr == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null output pointer");
    *ptr = nullptr;
    if (bytes == 0) return PBF_OK;
    return pbf_hip_check(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
}

pbf_status pbf_host_free(void* ptr) {
    if (ptr == nullptr) return PBF_OK;
    return pbf_hip_check(hipHostFree(ptr));
}

pbf_status pbf_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return PBF_OK;
    if (dst == nullptr || src == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null pointer");
    return pbf_hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

pbf_status pbf_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return PBF_OK;
    if (dst == nullptr || src == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null pointer");
    return pbf_hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

pbf_status pbf_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return PBF_OK;
    if (dst == nullptr || src == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null pointer");
    return pbf_hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}

pbf_status pbf_memset(void* dst, int byteValue, size_t bytes) {
    if (bytes == 0) return PBF_OK;
    if (dst == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null pointer");
    return pbf_hip_check(hipMemset(dst, byteValue, bytes));
}

pbf_status pbf_device_synchronize(void) { return pbf_hip_check(hipDeviceSynchronize()); }

pbf_status pbf_stream_create(void** stream) {
    if (stream == nullptr) return pbf_fail(PBF_ERR_INVALID_ARGUMENT, "null output pointer");
    hipStream_t s = nullptr;
    const pbf_status st = pbf_hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return st;
}

pbf_status pbf_stream_destroy(void* stream) {
    if (stream == nullptr) return PBF_OK;
    return pbf_hip_check(hipStreamDestroy(static_cast<hipStream_t>(stream)));
}

pbf_status pbf_stream_synchronize(void* stream) {
    return pbf_hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

}  // extern "C"
