// ozablas/core/workspace.hpp — per-shape device buffers (reference: include/ozablas/core/workspace.hpp:14-103).
//
// A workspace fixes (M, N, K, slices) and owns every intermediate buffer of the pipeline,
// allocated through its executor.  Buffer layouts of this build (see <ozablas_hip.h>,
// oza_workspace_bytes), with Mp = roundup(M,256), Np = roundup(N,256), Kp = roundup(K,128):
//   get_shift_A()  int32[M]   row shifts, reference semantics
//   get_shift_B()  int32[N]   column shifts, reference semantics
//   get_A_slices() int8 [S][Mp][Kp]   fragment-tiled, K zero padded
//   get_B_slices() int8 [S][Np][Kp]   B transposed, fragment-tiled
//                  (32-row blocks past the last logical row: padding, not written)
//   get_C_tc()     Scheme II: S planes of uint8 CRT digits r_s = C_s mod m_s, Mp x Np each, in
//                  32 x 32 element blocks (get_C_digits() is the same pointer with its true type),
//                  followed, for shapes whose GEMM runs in K parts (few tiles, long K), by the
//                  partial digits of parts 1.. (added into the S planes before step 4);
//                  Scheme I: S planes of int32 group products T_g (g = 2..S+1), Mp x Np each, in
//                  32 x 32 element blocks.
// get_C_tc() keeps the reference's type (int32_t*) so that callers compile unchanged; the buffer
// is not the reference's planar [S][M][N] INT32 C_s.  The reference's planar views come from the
// unpack_* members: unpack_slices (A [S][M][K], B [S][K][N] int8), unpack_products (the INT32
// products C_s of Scheme II or T_g of Scheme I, [S][M][N], recomputed from the slices that the last
// call left in the workspace, since the digits of Scheme II keep only C_s mod m_s) and, for Scheme
// II, unpack_digits ([S][M][N] uint8).  All of them are asynchronous on the executor's stream.
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
    // The group products T_g, g = 2..S+1, as the reference's planar int32 [S][M][N] (device), from
    // the slices in the workspace.
    void unpack_products(int32_t* T_planar) const;

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
    uint8_t* get_C_digits() const { return reinterpret_cast<uint8_t*>(C_tc_); }

    void unpack_slices(int8_t* A_planar, int8_t* B_planar) const;
    // The INT32 products C_s = A_s B_s as the reference's planar int32 [S][M][N] (device), from the
    // slices in the workspace.
    void unpack_products(int32_t* C_planar) const;
    // The CRT digits C_s mod m_s as planar uint8 [S][M][N] (device).
    void unpack_digits(uint8_t* digits_planar) const;

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

// Stage times in milliseconds.  Large calls overlap step 4 with step 3 in row bands: then
// step3_ms runs from the end of step 2 to the end of the last band's products (it includes the
// earlier bands' reconstruction beside them) and step4_ms is the last band's reconstruction.
struct OzaTimings {
    float step1_ms = 0.0f;
    float step2_ms = 0.0f;
    float step3_ms = 0.0f;
    float step4_ms = 0.0f;
    float total_ms = 0.0f;
};

}  // namespace ozablas
