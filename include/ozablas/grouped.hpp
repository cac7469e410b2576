// ozablas/grouped.hpp — many independent Scheme II GEMMs of different shapes in one pass
// (gfx950 build).  No counterpart in the reference, whose API runs one dense GEMM per call
// (include/ozablas/ozablas.hpp:31); a caller with a batch would loop over ozaki_scheme2_gemm.
//
// A GroupScheme2 plans the batch once (tile shape per problem, work lists, one workspace) and
// runs it with one launch per pipeline stage; by default the launches are captured into a HIP
// graph on the first run and replayed afterwards.  Each problem's C is bit-identical to
// ozaki_scheme2_gemm's with a WorkspaceScheme2 of the same M, N, K and moduli count.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

struct oza_group;

namespace ozablas {

class Executor;
struct OzaTimings;

struct GemmShape {
    size_t M, N, K;
    int slices;     // Scheme II moduli count, 1..20
    int tile = 0;   // GEMM tile edge 256 / 128 / 64; 0: the constructor's `tile`, else the cost model's
    int split = 0;  // K splits of the GEMM, 1..64 (1: none); 0: the planner's choice.  Bits do not
                    // depend on it.
};

class GroupScheme2 {
public:
    // tile: 0 = chosen per problem by the executor's cost model, or 256 / 128 / 64 for all.
    GroupScheme2(std::shared_ptr<const Executor> executor, std::vector<GemmShape> shapes, int tile = 0);
    ~GroupScheme2();
    GroupScheme2(const GroupScheme2&) = delete;
    GroupScheme2& operator=(const GroupScheme2&) = delete;

    size_t size() const { return shapes_.size(); }
    const std::vector<GemmShape>& shapes() const { return shapes_; }
    std::shared_ptr<const Executor> get_executor() const { return executor_; }
    size_t workspace_bytes() const;
    std::vector<int> tiles() const;   // GEMM tile edge chosen for each problem
    std::vector<int> splits() const;  // K splits of each problem's GEMM (1: not split)

    // A[i], B[i], C[i]: device pointers of problem i (dense row-major, as ozaki_scheme2_gemm).
    // Asynchronous unless timings is given (then it blocks and times the whole batch, eagerly).
    void run(const std::vector<const double*>& A, const std::vector<const double*>& B,
             const std::vector<double*>& C, bool use_graph = true, OzaTimings* timings = nullptr);

private:
    std::shared_ptr<const Executor> executor_;
    std::vector<GemmShape> shapes_;
    oza_group* group_ = nullptr;
};

// Measures the grouped GEMM's tile cost model on the executor's device and keeps it in the
// executor's context for later GroupScheme2 plans; returns the six coefficients.
std::vector<double> autotune_grouped(const std::shared_ptr<const Executor>& executor);

}  // namespace ozablas
