// ozablas/core/workspace.hpp — per-shape device buffers (reference: include/ozablas/core/workspace.hpp:14-103).
//
// A workspace fixes (M, N, K, slices) and owns every intermediate buffer of the pipeline,
// allocated through its executor.  Buffer layouts of this build (see <ozablas_hip.h>,
// oza_workspace_bytes):
//   get_shift_A()  int32[M]   row shifts, reference semantics
//   get_shift_B()  int32[N]   column shifts, reference semantics
//   get_A_slices() int8 [S][roundup(M,256)][roundup(K,128)]   fragment-tiled, K zero padded
//   get_B_slices() int8 [S][roundup(N,256)][roundup(K,128)]   B transposed, fragment-tiled
//                  (rows past roundup(M,32), columns past roundup(N,64): padding, not written)
//   get_C_tc()     Scheme II: uint8 CRT digits in GEMM tile order; Scheme I: int32 [S][M][N]
// The reference's planar [s][row][col] views are available through unpack_slices().
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

#include "ozablas/core/executor.hpp"

namespace ozablas {

class WorkspaceScheme1 {
public:
    WorkspaceScheme1(std::shared_ptr<const Executor> exec, size_t M, size_t N, size_t K, int slices);
    ~WorkspaceScheme1();

    WorkspaceScheme1(const WorkspaceScheme1&) = delete;
    WorkspaceScheme1& operator=(const WorkspaceScheme1&) = delete;

    std::shared_ptr<const Executor> get_executor() const { return exec_; }

    size_t get_M() const { return M_; }
    size_t get_N() const { return N_; }
    size_t get_K() const { return K_; }
    int get_slices() const { return slices_; }

    int32_t* get_shift_A() const { return shift_A_; }
    int32_t* get_shift_B() const { return shift_B_; }
    int8_t* get_A_slices() const { return A_slices_; }
    int8_t* get_B_slices() const { return B_slices_; }
    int32_t* get_C_tc() const { return C_tc_; }

    // Copies the slices into the reference's planar device layouts A [S][M][K], B [S][K][N]
    // (either pointer may be null).  Asynchronous on the executor's stream.
    void unpack_slices(int8_t* A_planar, int8_t* B_planar) const;

private:
    void release() noexcept;

    std::shared_ptr<const Executor> exec_;
    size_t M_, N_, K_;
    int slices_;
    int32_t* shift_A_ = nullptr;
    int32_t* shift_B_ = nullptr;
    int8_t* A_slices_ = nullptr;
    int8_t* B_slices_ = nullptr;
    int32_t* C_tc_ = nullptr;
};

class WorkspaceScheme2 {
public:
    WorkspaceScheme2(std::shared_ptr<const Executor> exec, size_t M, size_t N, size_t K, int slices);
    ~WorkspaceScheme2();

    WorkspaceScheme2(const WorkspaceScheme2&) = delete;
    WorkspaceScheme2& operator=(const WorkspaceScheme2&) = delete;

    std::shared_ptr<const Executor> get_executor() const { return exec_; }

    size_t get_M() const { return M_; }
    size_t get_N() const { return N_; }
    size_t get_K() const { return K_; }
    int get_slices() const { return slices_; }

    int32_t* get_shift_A() const { return shift_A_; }
    int32_t* get_shift_B() const { return shift_B_; }
    int8_t* get_A_slices() const { return A_slices_; }
    int8_t* get_B_slices() const { return B_slices_; }
    int32_t* get_C_tc() const { return C_tc_; }

    void unpack_slices(int8_t* A_planar, int8_t* B_planar) const;

private:
    void release() noexcept;

    std::shared_ptr<const Executor> exec_;
    size_t M_, N_, K_;
    int slices_;
    int32_t* shift_A_ = nullptr;
    int32_t* shift_B_ = nullptr;
    int8_t* A_slices_ = nullptr;
    int8_t* B_slices_ = nullptr;
    int32_t* C_tc_ = nullptr;
};

// Stage times in milliseconds (Scheme I reports its fused GEMM + reconstruction as step3_ms).
struct OzaTimings {
    float step1_ms = 0.0f;
    float step2_ms = 0.0f;
    float step3_ms = 0.0f;
    float step4_ms = 0.0f;
    float total_ms = 0.0f;
};

}  // namespace ozablas
