// group_internal.hpp — device descriptors and launchers of the grouped Scheme II pipeline: many
// independent GEMMs (each with its own M, N, K and moduli count) run through the four stages,
// one launch per stage over the union of every problem's work items.
//
// A work item is one workgroup's worth of a stage for one problem, (problem, local index); each
// stage has its own item list, built once per group on the host.  Blocks of one GEMM tile shape
// go to one launch per shape (256, 128 or 64 square tiles), so a stage is at most three launches.
//
// Split K: a problem's GEMM may run as P items per (modulus, tile), item j over its own range of
// 128-byte K blocks.  Item j writes the digits of its partial product, (A_s B_s)_j mod m_s, split 0
// into the problem's digit planes and split j > 0 into partial planes; a reduction stage adds the
// partial digits into the digit planes mod m_s.  Since (sum_j x_j) mod m = (sum_j (x_j mod m)) mod m,
// the digits, and so every bit of C, are those of the unsplit product whatever P is.
#pragma once

#include <vector>

#include "oza_internal.hpp"
#include "slice_core.hpp"

namespace oza {

struct GroupProblem {
    int64_t M, N, K, Mp, Np, Kp;  // Mp, Np padded to the problem's tile edge, Kp to 128
    int32_t* shA;                 // exponent keys during step 1, shifts afterwards
    int32_t* shB;
    int8_t* Asl;                  // [S][Mp][Kp] fragment-tiled
    int8_t* Bsl;                  // [S][Np][Kp] fragment-tiled
    uint8_t* dg;                  // [S][Mp][Np] blocked digits
    uint8_t* dgp;                 // [split - 1][S][Mp][Np] partial digits of splits 1.. (split > 1)
    int32_t S, kparam, wide, tile;
    int32_t tiles_m, tiles_n;
    int32_t split, kbs;           // K splits of the GEMM (1: none), 128-byte K blocks per split
    int64_t chunksA, nblkA, colblkB, nblkB;  // step-1 blocks of A and B
    int64_t slA, slB;                        // step-2 blocks of A and B
    int64_t nfA;  // blocks of A with steps 1 and 2 fused (Kp <= core::kFuseMaxKp; nblkA = slA = 0 then)
};

struct GroupPtrs {
    const double* A;
    const double* B;
    double* C;
};

struct GItem {
    int32_t prob;
    int32_t local;
};

enum GroupStage { kGStats, kGFinalize, kGFuseA, kGSlice, kGGemm256, kGGemm128, kGGemm64, kGReduce, kGRecon, kGStages };

struct GroupLaunch {
    const GroupProblem* probs;
    const GroupPtrs* ptrs;
    const GItem* items;  // this stage's list
    int64_t n;
    const CrtTables* crt;
    core::ResidueConsts rc;
    double gemm_share;   // fraction of the resident GEMM workgroups launched (1: all CUs)
};

// Tile edges of the three GEMM shapes, in the order of kGGemm256.. .
constexpr int kGroupTiles[3] = {256, 128, 64};
// The tile cost model measured once on an MI355X (the contexts' default until oza_group_autotune).
constexpr double kDefaultGroupCost[6] = {6.3, 1.58, 1.5, 0.48, 0.66, 0.25};
// Largest K split the planner picks by itself, and the largest a caller may ask for.
constexpr int kAutoSplitMax = 32;
constexpr int kSplitMax = 64;

// What the planner knows of the device: the tile cost model (CU-microseconds per tile,
// cost[2c] + cost[2c+1] * 128-byte K blocks, for the tile edges kGroupTiles[c]), the CU count and,
// per tile shape, the workgroups resident at once (CUs x workgroups per CU).
struct GroupModel {
    double cost[6];
    int cus;
    int slots[3];
};

// ---- host-side planning (capi_group.hip)
// A plan: descriptors of every problem (workspace addresses as offsets from the workspace base
// until relocate()) and the stage work lists, concatenated in GroupStage order.
struct HostPlan {
    std::vector<GroupProblem> probs;
    std::vector<GItem> items;
    int64_t stage_off[kGStages] = {};
    int64_t stage_n[kGStages] = {};
    size_t ws_bytes = 0;
    int64_t nkeys = 0;  // 32-bit keys at the start of the workspace, cleared by every run
};

// Device copies of a plan.
struct DevPlan {
    const GroupProblem* probs;
    const GroupPtrs* ptrs;
    const GItem* items;
    uint32_t* keys;
    int64_t nkeys;
    int64_t stage_off[kGStages];
    int64_t stage_n[kGStages];
};

// Tile edge (256, 128, 64) with the least estimated CU time under `cost` (6 coefficients).
int choose_tile(const double* cost, int64_t M, int64_t N, int64_t Kp, int S);
// The K split the planner gives one problem planned alone (1: none).
int alone_split(const GroupModel& model, int64_t M, int64_t N, int64_t K, int S, int tile);
// Workspace bytes one problem takes in a plan (keys, operand planes, digit planes, partial digit
// planes of a K split, alignment).
size_t problem_bytes(size_t M, size_t N, size_t K, int S, int tile, int split = 1);
// Returns an oza_status (message via the C-ABI error slot).  Tile edge per problem: tiles[i] when
// tiles is given and tiles[i] != 0, else `tile` when nonzero, else the cost model's choice.  K split
// per problem: splits[i] when splits is given and splits[i] != 0 (clamped to the problem's
// 128-byte K blocks), else the planner's choice from `model` (split_planner).
int build_plan(const GroupModel& model, int n, const size_t* M, const size_t* N, const size_t* K, const int* S,
               int tile, HostPlan& out, const int* tiles = nullptr, const int* splits = nullptr);
void relocate(HostPlan& plan, char* base);
// Zero fill of the keys, then every stage; ev (optional, 5 events) brackets steps 1..4.
// gemm_share < 1 launches the persistent GEMM on that fraction of its resident workgroups, leaving
// CUs free for work on other streams (the request server's small batches).
hipError_t enqueue_plan(const DevPlan& d, const CrtTables* crt, hipStream_t st, hipEvent_t* ev,
                        double gemm_share = 1.0);

hipError_t launch_group_zero(uint32_t* keys, int64_t n, hipStream_t st);
// Resident workgroups of each grouped GEMM shape on the current device (CUs x per-CU occupancy).
hipError_t group_gemm_slots(int slots[3]);
hipError_t launch_group_stage(int stage, const GroupLaunch& g, hipStream_t st);
// Deterministic pseudo-random doubles in [-1, 1) (autotuning inputs).
hipError_t launch_fill_pattern(double* p, int64_t n, uint32_t seed, hipStream_t st);

}  // namespace oza
