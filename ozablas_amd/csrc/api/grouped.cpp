// grouped.cpp — ozablas::GroupScheme2 over the C ABI's oza_group_* entry points.
#include "ozablas/grouped.hpp"

#include <stdexcept>

#include "ozablas/core/executor.hpp"
#include "ozablas/core/workspace.hpp"
#include "status.hpp"

namespace ozablas {

namespace {

const HipExecutor& hip_of(const std::shared_ptr<const Executor>& exec) {
    if (!exec) throw std::invalid_argument("OzaBLAS: Executor cannot be null.");
    if (const auto* hip = dynamic_cast<const HipExecutor*>(exec.get())) return *hip;
    throw std::runtime_error("OzaBLAS: grouped GEMMs need a HipExecutor.");
}

}  // namespace

GroupScheme2::GroupScheme2(std::shared_ptr<const Executor> executor, std::vector<GemmShape> shapes, int tile)
    : executor_(std::move(executor)), shapes_(std::move(shapes)) {
    const HipExecutor& hip = hip_of(executor_);
    std::vector<size_t> M, N, K;
    std::vector<int> S, T, P;
    for (const auto& s : shapes_) {
        M.push_back(s.M);
        N.push_back(s.N);
        K.push_back(s.K);
        S.push_back(s.slices);
        T.push_back(s.tile ? s.tile : tile);
        P.push_back(s.split);
    }
    detail::check(oza_group_create_split(hip.context(), (int)shapes_.size(), M.data(), N.data(), K.data(), S.data(),
                                         T.data(), P.data(), &group_));
}

GroupScheme2::~GroupScheme2() { oza_group_destroy(group_); }

size_t GroupScheme2::workspace_bytes() const {
    size_t b = 0;
    detail::check(oza_group_info(group_, &b, nullptr));
    return b;
}

std::vector<int> GroupScheme2::tiles() const {
    std::vector<int> t(shapes_.size());
    detail::check(oza_group_info(group_, nullptr, t.data()));
    return t;
}

std::vector<int> GroupScheme2::splits() const {
    std::vector<int> s(shapes_.size());
    detail::check(oza_group_splits(group_, s.data()));
    return s;
}

void GroupScheme2::run(const std::vector<const double*>& A, const std::vector<const double*>& B,
                       const std::vector<double*>& C, bool use_graph, OzaTimings* timings) {
    if (A.size() != shapes_.size() || B.size() != shapes_.size() || C.size() != shapes_.size())
        throw std::invalid_argument("OzaBLAS: one A, B and C pointer per grouped problem.");
    oza_timings ct{};
    detail::check(
        oza_group_scheme2_gemm(group_, A.data(), B.data(), C.data(), use_graph ? 1 : 0, timings ? &ct : nullptr));
    if (timings) {
        timings->step1_ms = ct.step1_ms;
        timings->step2_ms = ct.step2_ms;
        timings->step3_ms = ct.step3_ms;
        timings->step4_ms = ct.step4_ms;
        timings->total_ms = ct.total_ms;
    }
}

std::vector<double> autotune_grouped(const std::shared_ptr<const Executor>& executor) {
    std::vector<double> c(6);
    detail::check(oza_group_autotune(hip_of(executor).context(), c.data()));
    return c;
}

}  // namespace ozablas
