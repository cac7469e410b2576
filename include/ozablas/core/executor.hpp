// ozablas/core/executor.hpp — execution contexts (reference: include/ozablas/core/executor.hpp:15-112).
//
// Executor is the allocation / copy / synchronization interface the workspaces allocate through.
// HipExecutor is the gfx950 backend: it owns a library context (device, stream, CRT tables resident
// on that device) created through the C ABI in <ozablas_hip.h>.  CpuExecutor allocates host memory
// (the GEMM entry points reject it, as the reference's dispatcher does).  CudaExecutor exists for
// source compatibility only and throws on construction: this build has no CUDA path.
#pragma once

#include <cstddef>
#include <memory>
#include <stdexcept>

struct oza_context;

namespace ozablas {

class Executor {
public:
    virtual ~Executor() = default;

    virtual void allocate(void** ptr, size_t bytes) const = 0;
    virtual void free(void* ptr) const = 0;

    virtual void copy_from_host(void* device_dst, const void* host_src, size_t bytes) const = 0;
    virtual void copy_to_host(void* host_dst, const void* device_src, size_t bytes) const = 0;
    virtual void copy_device_to_device(void* dst, const void* src, size_t bytes) const = 0;

    // Returns once all previously issued work on this executor's device has finished.
    virtual void synchronize() const = 0;

    // Vendor BLAS handle as an opaque pointer (rocblas_handle for HipExecutor), or nullptr.
    virtual void* get_blas_handle() const { return nullptr; }
};

class CpuExecutor : public Executor {
public:
    CpuExecutor() = default;
    ~CpuExecutor() override = default;

    void allocate(void** ptr, size_t bytes) const override;
    void free(void* ptr) const override;
    void copy_from_host(void* device_dst, const void* host_src, size_t bytes) const override;
    void copy_to_host(void* host_dst, const void* device_src, size_t bytes) const override;
    void copy_device_to_device(void* dst, const void* src, size_t bytes) const override;
    void synchronize() const override;
};

class CudaExecutor : public Executor {
public:
    explicit CudaExecutor(int device_id = 0);  // throws std::runtime_error
    ~CudaExecutor() override;

    void allocate(void** ptr, size_t bytes) const override;
    void free(void* ptr) const override;
    void copy_from_host(void* device_dst, const void* host_src, size_t bytes) const override;
    void copy_to_host(void* host_dst, const void* device_src, size_t bytes) const override;
    void copy_device_to_device(void* dst, const void* src, size_t bytes) const override;
    void synchronize() const override;
    void* get_blas_handle() const override;

private:
    int device_id_ = 0;
};

class HipExecutor : public Executor {
public:
    explicit HipExecutor(int device_id = 0);
    ~HipExecutor() override;

    HipExecutor(const HipExecutor&) = delete;
    HipExecutor& operator=(const HipExecutor&) = delete;

    void allocate(void** ptr, size_t bytes) const override;
    void free(void* ptr) const override;
    void copy_from_host(void* device_dst, const void* host_src, size_t bytes) const override;
    void copy_to_host(void* host_dst, const void* device_src, size_t bytes) const override;
    void copy_device_to_device(void* dst, const void* src, size_t bytes) const override;
    void synchronize() const override;
    void* get_blas_handle() const override;

    int device_id() const { return device_id_; }
    // The C-ABI context behind this executor (see <ozablas_hip.h>).
    oza_context* context() const { return ctx_; }

private:
    int device_id_;
    oza_context* ctx_ = nullptr;
};

}  // namespace ozablas
