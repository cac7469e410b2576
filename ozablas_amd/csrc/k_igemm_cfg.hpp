// k_igemm_cfg.hpp — tile configurations of the INT8 MFMA GEMM (shared with the step-4 kernel,
// which reads the digit layout the GEMM epilogue writes).
#pragma once

#include <cstdint>

namespace oza {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// BM x BN output tile per workgroup, BK bytes of K per LDS stage, NWM x NWN waves.
// SWZ: LDS 16-byte chunk index is XORed with (row >> SWZ) to spread MFMA fragment reads over banks.
template <int BM_, int BN_, int BK_, int NWM_, int NWN_, int SWZ_>
struct IgemmCfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, NWM = NWM_, NWN = NWN_, SWZ = SWZ_;
    static constexpr int NT = 64 * NWM * NWN;
    static constexpr int WTM = BM / NWM, WTN = BN / NWN;
    static constexpr int FM = WTM / 32, FN = WTN / 32;
    static constexpr int KCH = BK / 16;
    static constexpr int A_LD = BM * KCH / NT;
    static constexpr int B_LD = BN * KCH / NT;
    static_assert(WTM % 32 == 0 && WTN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert(BM * KCH % NT == 0 && BN * KCH % NT == 0, "stage must split evenly over threads");
    static_assert(BK % 32 == 0, "BK must be a multiple of the MFMA K (32)");
    static_assert(256 % BM == 0 && 256 % BN == 0, "tile must divide the 256-row padding");
};

using IgemmCfgS2 = IgemmCfg<128, 128, 64, 2, 2, 1>;
using IgemmCfgS1 = IgemmCfg<128, 128, 64, 2, 2, 1>;

}  // namespace oza
