// workspace.cpp — workspaces (reference: src/core/workspace.cpp:21-124).  Buffer sizes come from
// oza_workspace_bytes; the buffers are allocated through the executor, and a failed allocation
// releases what was already taken before rethrowing (reference: workspace.cpp:41-46, 86-89).
#include "ozablas/core/workspace.hpp"

#include "status.hpp"

namespace ozablas {

namespace {

template <class W>
void allocate_all(const Executor& ex, int scheme, size_t M, size_t N, size_t K, int S, int32_t*& sa, int32_t*& sb,
                  int8_t*& as, int8_t*& bs, int32_t*& ct) {
    size_t b[5];
    detail::check(oza_workspace_bytes(scheme, M, N, K, S, b));
    ex.allocate(reinterpret_cast<void**>(&sa), b[0]);
    ex.allocate(reinterpret_cast<void**>(&sb), b[1]);
    ex.allocate(reinterpret_cast<void**>(&as), b[2]);
    ex.allocate(reinterpret_cast<void**>(&bs), b[3]);
    ex.allocate(reinterpret_cast<void**>(&ct), b[4]);
}

oza_ctx context_of(const Executor& ex);

void unpack(const Executor& ex, size_t M, size_t N, size_t K, int S, const int8_t* as, const int8_t* bs,
            int8_t* ap, int8_t* bp) {
    detail::check(oza_unpack_slices(context_of(ex), M, N, K, S, as, bs, ap, bp));
}

oza_ctx context_of(const Executor& ex) {
    const auto* hip = dynamic_cast<const HipExecutor*>(&ex);
    if (!hip) throw std::runtime_error("OzaBLAS: No suitable hardware backend found for the provided Executor.");
    return hip->context();
}

}  // namespace

WorkspaceScheme1::WorkspaceScheme1(std::shared_ptr<const Executor> exec, size_t M, size_t N, size_t K, int slices)
    : exec_(std::move(exec)), M_(M), N_(N), K_(K), slices_(slices) {
    if (!exec_) throw std::invalid_argument("OzaBLAS: Executor cannot be null.");
    if (slices_ < 1) throw std::invalid_argument("OzaBLAS: Slices must be >= 1.");
    try {
        allocate_all<WorkspaceScheme1>(*exec_, 1, M, N, K, slices, shift_A_, shift_B_, A_slices_, B_slices_, C_tc_);
    } catch (...) {
        release();
        throw;
    }
}

WorkspaceScheme1::~WorkspaceScheme1() { release(); }

void WorkspaceScheme1::release() noexcept {
    if (!exec_) return;
    for (void* p : {(void*)shift_A_, (void*)shift_B_, (void*)A_slices_, (void*)B_slices_, (void*)C_tc_})
        if (p) exec_->free(p);
    shift_A_ = shift_B_ = C_tc_ = nullptr;
    A_slices_ = B_slices_ = nullptr;
}

void WorkspaceScheme1::unpack_slices(int8_t* A_planar, int8_t* B_planar) const {
    unpack(*exec_, M_, N_, K_, slices_, A_slices_, B_slices_, A_planar, B_planar);
}

void WorkspaceScheme1::unpack_products(int32_t* T_planar) const {
    detail::check(oza_stage3_s1_groups(context_of(*exec_), M_, N_, K_, slices_, A_slices_, B_slices_, T_planar));
}

WorkspaceScheme2::WorkspaceScheme2(std::shared_ptr<const Executor> exec, size_t M, size_t N, size_t K, int slices)
    : exec_(std::move(exec)), M_(M), N_(N), K_(K), slices_(slices) {
    if (!exec_) throw std::invalid_argument("OzaBLAS: Executor cannot be null.");
    if (slices_ < 1 || slices_ > 20) throw std::invalid_argument("OzaBLAS: Scheme II supports between 1 and 20 slices.");
    try {
        allocate_all<WorkspaceScheme2>(*exec_, 2, M, N, K, slices, shift_A_, shift_B_, A_slices_, B_slices_, C_tc_);
    } catch (...) {
        release();
        throw;
    }
}

WorkspaceScheme2::~WorkspaceScheme2() { release(); }

void WorkspaceScheme2::release() noexcept {
    if (!exec_) return;
    for (void* p : {(void*)shift_A_, (void*)shift_B_, (void*)A_slices_, (void*)B_slices_, (void*)C_tc_})
        if (p) exec_->free(p);
    shift_A_ = shift_B_ = C_tc_ = nullptr;
    A_slices_ = B_slices_ = nullptr;
}

void WorkspaceScheme2::unpack_slices(int8_t* A_planar, int8_t* B_planar) const {
    unpack(*exec_, M_, N_, K_, slices_, A_slices_, B_slices_, A_planar, B_planar);
}

void WorkspaceScheme2::unpack_products(int32_t* C_planar) const {
    detail::check(oza_stage3_s2_products(context_of(*exec_), M_, N_, K_, slices_, A_slices_, B_slices_, C_planar));
}

void WorkspaceScheme2::unpack_digits(uint8_t* digits_planar) const {
    detail::check(oza_unpack_digits(context_of(*exec_), M_, N_, slices_, C_tc_, digits_planar));
}

}  // namespace ozablas
