// k_igemm_cfg.hpp — tile configurations of the INT8 MFMA GEMM (shared with the step-4 kernel,
// which reads the digit layout the GEMM epilogue writes).
#pragma once

#include <cstdint>

namespace oza {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// BM x BN output tile per workgroup, BK bytes of K per LDS stage, NWM x NWN waves, STAGES-deep
// LDS ring filled by direct global->LDS loads of whole 1 KiB fragment blocks (the operand planes are
// fragment-tiled, see frag_off16).  GM: tiles are visited in groups of GM tile-rows so that
// concurrently running workgroups share A and B panels in L2.
template <int BM_, int BN_, int BK_, int NWM_, int NWN_, int STAGES_, int GM_>
struct IgemmCfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, NWM = NWM_, NWN = NWN_, STAGES = STAGES_;
    static constexpr int GM = GM_;
    static constexpr int NWAVES = NWM * NWN;
    static constexpr int NT = 64 * NWAVES;
    static constexpr int WTM = BM / NWM, WTN = BN / NWN;
    static constexpr int FM = WTM / 32, FN = WTN / 32;
    static constexpr int KB = BK / 32;             // fragment blocks along K per stage
    static constexpr int ABLK = BM * BK / 1024;    // 1 KiB blocks (one wave-wide direct load) per stage
    static constexpr int BBLK = BN * BK / 1024;
    static constexpr int LPS = (ABLK + BBLK) / NWAVES;  // direct loads per thread per stage
    static constexpr int STAGE_BYTES = (BM + BN) * BK;
    static constexpr int LDS_BYTES = STAGES * STAGE_BYTES;
    static_assert(WTM % 32 == 0 && WTN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert(ABLK % NWAVES == 0 && BBLK % NWAVES == 0, "stage must split evenly over waves");
    static_assert(BK % 32 == 0, "BK must be a multiple of the MFMA K (32)");
    static_assert(256 % BM == 0 && 256 % BN == 0, "tile must divide the 256-row padding");
    static_assert(128 % BK == 0, "BK must divide the 128-byte K padding");
    static_assert(STAGES >= 2 && LDS_BYTES <= 160 * 1024, "LDS budget");
};

using IgemmCfgS2 = IgemmCfg<256, 256, 128, 4, 2, 2, 8>;
using IgemmCfgS1 = IgemmCfg<128, 128, 64, 2, 2, 2, 4>;

}  // namespace oza
