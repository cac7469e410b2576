// ozablas.cpp — the public entry points (reference dispatcher: src/ozablas.cpp:22-74).  The
// executor type selects the backend; only the gfx950 HIP backend exists in this build.
#include "ozablas/ozablas.hpp"

#include "ozablas/core/executor.hpp"
#include "ozablas/core/workspace.hpp"
#include "status.hpp"

namespace ozablas {

namespace {

const HipExecutor& backend_of(const std::shared_ptr<const Executor>& exec) {
    if (dynamic_cast<const CudaExecutor*>(exec.get()))
        throw std::runtime_error("OzaBLAS: CUDA backend requested but library was not built with CUDA support.");
    if (const auto* hip = dynamic_cast<const HipExecutor*>(exec.get())) return *hip;
    throw std::runtime_error("OzaBLAS: No suitable hardware backend found for the provided Executor.");
}

void fill(OzaTimings* t, const oza_timings& ct) {
    if (!t) return;
    t->step1_ms = ct.step1_ms;
    t->step2_ms = ct.step2_ms;
    t->step3_ms = ct.step3_ms;
    t->step4_ms = ct.step4_ms;
    t->total_ms = ct.total_ms;
}

}  // namespace

void ozaki_scheme1_gemm(WorkspaceScheme1& ws, const double* A, const double* B, double* C, OzaTimings* timings) {
    const HipExecutor& hip = backend_of(ws.get_executor());
    oza_timings ct{};
    detail::check(oza_scheme1_gemm(hip.context(), ws.get_M(), ws.get_N(), ws.get_K(), ws.get_slices(), A, B, C,
                                   ws.get_shift_A(), ws.get_shift_B(), ws.get_A_slices(), ws.get_B_slices(),
                                   ws.get_C_tc(), timings ? &ct : nullptr));
    fill(timings, ct);
}

void ozaki_scheme2_gemm(WorkspaceScheme2& ws, const double* A, const double* B, double* C, OzaTimings* timings) {
    const HipExecutor& hip = backend_of(ws.get_executor());
    oza_timings ct{};
    detail::check(oza_scheme2_gemm(hip.context(), ws.get_M(), ws.get_N(), ws.get_K(), ws.get_slices(), A, B, C,
                                   ws.get_shift_A(), ws.get_shift_B(), ws.get_A_slices(), ws.get_B_slices(),
                                   ws.get_C_tc(), timings ? &ct : nullptr));
    fill(timings, ct);
}

}  // namespace ozablas
