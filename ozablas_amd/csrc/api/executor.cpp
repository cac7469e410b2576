// executor.cpp — Executor implementations (reference: src/core/executor.cpp:153-330).
// HipExecutor forwards to the C ABI; no HIP header is included here.
#include "ozablas/core/executor.hpp"

#include <cstdlib>
#include <cstring>
#include <new>

#include "status.hpp"

namespace ozablas {

// ------------------------------------------------------------------ CPU (host memory)
void CpuExecutor::allocate(void** ptr, size_t bytes) const {
    constexpr size_t kAlign = 64;
    *ptr = std::aligned_alloc(kAlign, (bytes + kAlign - 1) / kAlign * kAlign);
    if (!*ptr) throw std::bad_alloc();
}
void CpuExecutor::free(void* ptr) const { std::free(ptr); }
void CpuExecutor::copy_from_host(void* dst, const void* src, size_t bytes) const { std::memcpy(dst, src, bytes); }
void CpuExecutor::copy_to_host(void* dst, const void* src, size_t bytes) const { std::memcpy(dst, src, bytes); }
void CpuExecutor::copy_device_to_device(void* dst, const void* src, size_t bytes) const {
    std::memcpy(dst, src, bytes);
}
void CpuExecutor::synchronize() const {}

// ------------------------------------------------------------------ CUDA (not built)
CudaExecutor::CudaExecutor(int device_id) : device_id_(device_id) {
    throw std::runtime_error("OzaBLAS was not compiled with CUDA support.");
}
CudaExecutor::~CudaExecutor() = default;
void CudaExecutor::allocate(void**, size_t) const {}
void CudaExecutor::free(void*) const {}
void CudaExecutor::copy_from_host(void*, const void*, size_t) const {}
void CudaExecutor::copy_to_host(void*, const void*, size_t) const {}
void CudaExecutor::copy_device_to_device(void*, const void*, size_t) const {}
void CudaExecutor::synchronize() const {}
void* CudaExecutor::get_blas_handle() const { return nullptr; }

// ------------------------------------------------------------------ HIP (gfx950)
HipExecutor::HipExecutor(int device_id) : device_id_(device_id) {
    oza_ctx c = nullptr;
    const int st = oza_ctx_create(device_id, &c);
    if (st != OZA_SUCCESS) throw std::runtime_error("HIP: Failed to set device.");
    ctx_ = c;
}

HipExecutor::~HipExecutor() {
    if (ctx_) oza_ctx_destroy(ctx_);
}

void HipExecutor::allocate(void** ptr, size_t bytes) const {
    if (oza_malloc(ctx_, ptr, bytes) != OZA_SUCCESS) throw std::bad_alloc();
}

void HipExecutor::free(void* ptr) const { (void)oza_free(ctx_, ptr); }

void HipExecutor::copy_from_host(void* dst, const void* src, size_t bytes) const {
    if (oza_memcpy_h2d(ctx_, dst, src, bytes) != OZA_SUCCESS)
        throw std::runtime_error("HIP: Memcpy HostToDevice failed.");
}

void HipExecutor::copy_to_host(void* dst, const void* src, size_t bytes) const {
    if (oza_memcpy_d2h(ctx_, dst, src, bytes) != OZA_SUCCESS)
        throw std::runtime_error("HIP: Memcpy DeviceToHost failed.");
}

void HipExecutor::copy_device_to_device(void* dst, const void* src, size_t bytes) const {
    if (oza_memcpy_d2d(ctx_, dst, src, bytes) != OZA_SUCCESS)
        throw std::runtime_error("HIP: Memcpy DeviceToDevice failed.");
}

void HipExecutor::synchronize() const {
    if (oza_synchronize(ctx_) != OZA_SUCCESS) throw std::runtime_error("HIP: Device synchronize failed.");
}

void* HipExecutor::get_blas_handle() const {
    void* h = nullptr;
    detail::check(oza_blas_handle(ctx_, &h));
    return h;
}

}  // namespace ozablas
