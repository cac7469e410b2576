// oza_internal.hpp — definitions shared by the gfx950 kernels and the C-ABI layer.
//
// HBM layout of one Ozaki workspace (all planes dense, device memory):
//   shift_A   int32[M]                   row shifts (reference semantics, step1_statistics.hpp:23-76)
//   shift_B   int32[N]                   column shifts
//   A_slices  int8 [S][Mp][Kp]           A residues/slices, K contiguous, fragment-tiled (frag_off16)
//   B_slices  int8 [S][Np][Kp]           B residues/slices TRANSPOSED (column n is a K-contiguous row)
//   C_tc      Scheme II: uint8 [S][Mp][Np] CRT digits C_s mod m_s in 32 x 32 element blocks
//             Scheme I : int32 [S][Mp][Np] group products T_g in 32 x 32 element blocks (the stage
//                        entry point writes the reference's planar int32 [S][M][N] instead)
// with Mp = roundup(M, 256), Np = roundup(N, 256), Kp = roundup(K, 128).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace oza {

constexpr int kMaxSlices = 20;
constexpr int kPadMN = 256;  // row padding of the operand planes (largest GEMM tile edge)
constexpr int kPadK = 128;   // K padding of the operand planes (largest K stage)
constexpr int kPadDigits = 16;

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// The Scheme II moduli, identical to the reference list (src/common/crt_tables.hpp:13-16).
inline constexpr uint8_t kModuli[kMaxSlices] = {251, 241, 239, 233, 229, 227, 223, 211, 199, 197,
                                                193, 191, 181, 179, 173, 167, 163, 157, 151, 149};

// Bit length of M(S) = m_0 * ... * m_{S-1}, at compile time.
constexpr int bitlen_M(int S) {
    uint64_t d[4] = {1, 0, 0, 0};  // 256-bit little-endian
    for (int s = 0; s < S; ++s) {
        unsigned __int128 carry = 0;
        for (int i = 0; i < 4; ++i) {
            const unsigned __int128 p = (unsigned __int128)d[i] * kModuli[s] + carry;
            d[i] = (uint64_t)p;
            carry = p >> 64;
        }
    }
    for (int i = 3; i >= 0; --i)
        if (d[i]) {
            int b = 64 * i;
            for (uint64_t v = d[i]; v; v >>= 1) ++b;
            return b;
        }
    return 0;
}
static_assert(bitlen_M(8) == 63 && bitlen_M(16) == 124 && bitlen_M(20) == 153, "moduli products");

// Per-S CRT constants used by the kernels.  Big integers are little-endian 16-bit limbs held
// in 32-bit words so that c * limb (c < 2^8) and sums of up to 20 such products fit in 32 bits.
constexpr int kLimbs = 12;  // 192 bits: M < 2^153, sum c_s*(M/m_s) < 20*2^8*M < 2^166
struct CrtSet {
    int S;
    int nlimbs;                       // limbs actually used for this S
    uint32_t M16[kLimbs];             // M = prod m_s
    uint32_t Mhalf16[kLimbs];         // floor(M/2)
    uint32_t Q16[kMaxSlices][kLimbs];  // Q_s = ((M/m_s)^-1 mod m_s) * (M/m_s) mod M
    // Byte k of Q8x4[j][l] = 8-bit limb l of Q_{4j+k} (0 past S): operand of v_dot4_u32_u8
    // against four digits of one element.
    uint32_t Q8x4[kMaxSlices / 4][2 * kLimbs];
    uint32_t m[kMaxSlices];
    uint32_t inv[kMaxSlices];         // (M/m_s)^-1 mod m_s
    double rcp[kMaxSlices];           // 1/m_s rounded
    double wl[kLimbs];                // 2^(16 l) / M rounded (quotient estimate from limbs)
    double qthr;                      // |frac - 1/2| below 1/M + margin: quotient may be ambiguous
    uint64_t M64[4], Mhalf64[4];      // the same, 64-bit limbs
};

// All S = 1..20, computed on the host from the moduli (closed forms).
struct CrtTables {
    CrtSet set[kMaxSlices];  // set[S-1]
};
const CrtTables& host_crt_tables();

// Host-side parameters (reference: step1_statistics.hpp:12-20, pipeline.hip:205-207).
int scheme1_beta(size_t K);
int scheme2_kparam(int S, size_t K);
int scheme2_slices_for_digits(size_t K, double digits);
double scheme2_spread_shortfall(double spread);
int scheme2_slices_for_accuracy(size_t K, double digits, double spread);

// ------------------------------------------------------------------------------------------
// Launchers (each enqueues on `stream`, returns hipError_t from hipGetLastError()).
// ------------------------------------------------------------------------------------------
struct Dims {
    int64_t M, N, K;
    int64_t Mp, Np, Kp, Nd;
    int S;
};
Dims make_dims(size_t M, size_t N, size_t K, int S);

// An FP64 operand as steps 1 and 2 read it: logical rows (M rows of op(A), or the N columns of
// op(B)) of K values each, element (r, k) at p[r * ld + k] when kcontig, else at p[k * ld + r].
// The reference's dense row-major A is {A, K, true}, its B is {B, N, false}.
struct Operand {
    const double* p;
    int64_t ld;
    bool kcontig;
};
// The output of step 4: C[i * ldc + j] = alpha * X_ij + beta * C[i * ldc + j], evaluated as
// fma(alpha, X, beta * C); with beta == 0, C is not read and alpha == 1 stores X itself.
struct OutC {
    double* p;
    int64_t ldc;
    double alpha, beta;
};

// C = beta * C over an M x N block with leading dimension ldc (beta == 0 writes zeros).
hipError_t launch_scale_c(const OutC& c, int64_t M, int64_t N, hipStream_t st);

// Step 1: shift_A[i] = offset - E_i, shift_B[j] = offset - F_j.
// The key-clearing launch also clears `nzero` words at `zero` (may be null): the work counters
// of the calls' step-3 GEMMs.
hipError_t launch_stats(const Dims& d, const Operand& A, const Operand& B, int32_t offset, int32_t* shift_A,
                        int32_t* shift_B, hipStream_t st, uint32_t* zero = nullptr, int64_t nzero = 0);

// Step 2.
hipError_t launch_slice_s2(const Dims& d, const Operand& A, const Operand& B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st);
hipError_t launch_slice_s1(const Dims& d, int beta, const Operand& A, const Operand& B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st);
// Unpack product-layout operand planes to the reference's planar layout [s][rows][cols].
hipError_t launch_unpack_A(const Dims& d, const int8_t* A_sl, int8_t* out, hipStream_t st);
hipError_t launch_unpack_B(const Dims& d, const int8_t* B_sl, int8_t* out, hipStream_t st);

// Step 3 (+ fused epilogues).
enum class S2Out { Digits, Int32 };
// Tile rows [tm0, tm0 + tms) of BM = 256 rows (tms < 0: to the last tile row).  queue: 8 device
// counters for the dynamic work order, zero when the launch starts, or null (static order).
// split > 1 (Digits only): the K steps run in `split` parts of ceil(KT / split) steps (every part
// non-empty), part j writing the digits of its partial product to plane set j of `out`
// ([split][S][Mp][Np]); launch_digits_reduce then adds sets 1.. into set 0.
hipError_t launch_igemm_s2(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, S2Out mode, void* out,
                           const CrtSet* crt_dev, hipStream_t st, int tm0 = 0, int tms = -1,
                           unsigned* queue = nullptr, int split = 1);
// Digit planes [split][S][Mp][Np]: set 0 += sets 1..split-1, mod m_s.
hipError_t launch_digits_reduce(const Dims& d, uint8_t* digits, int split, const CrtSet* crt_dev, hipStream_t st);
constexpr int kS2TileRows = 256;
// K split of a single Scheme II call (1: none): the grouped planner's choice for the problem alone,
// on a fixed description of an MI355X (256 CUs, the default tile cost model) and the 256-wide tile
// of the single-call GEMM, so that the workspace size and the call agree on it.
int scheme2_single_split(size_t M, size_t N, size_t K, int S);
// Scheme I: group products T_g, g = 2..S+1, into blocked INT32 planes [S][Mp][Np] (32 x 32 blocks).
// Tile rows [tm0, tm0 + tms) of 256 rows (tms < 0: to the last tile row).
hipError_t launch_igemm_s1_blocked(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, int32_t* T, hipStream_t st,
                                   unsigned* queue = nullptr, int tm0 = 0, int tms = -1);
// Scheme I: group products T_g only, int32 [S][M][N] (reference C_tc semantics per group).
hipError_t launch_igemm_s1_groups(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, int32_t* T,
                                  hipStream_t st);

// Step 4.
// Rows [row0, row0 + rows) of C (row0 a multiple of 32; rows < 0: to the last row).  narrow: one
// workgroup per CU walks the blocks, leaving room on each CU for a concurrent step-3 workgroup.
hipError_t launch_recon_s2_digits(const Dims& d, const uint8_t* digits, const int32_t* shift_A,
                                  const int32_t* shift_B, const CrtSet* crt_dev, const OutC& C, hipStream_t st,
                                  int64_t row0 = 0, int64_t rows = -1, bool narrow = false);
hipError_t launch_recon_s2_int32(const Dims& d, const int32_t* C_tc, const int32_t* shift_A,
                                 const int32_t* shift_B, const CrtSet* crt_dev, double* C, hipStream_t st);
// Blocked digit planes <-> planar [s][M][N] digits in [0, m_s) (pack reduces its input mod m_s).
hipError_t launch_digits_unpack(const Dims& d, const uint8_t* digits, uint8_t* out, hipStream_t st);
hipError_t launch_digits_pack(const Dims& d, const uint8_t* planar, const CrtSet* crt_dev, uint8_t* digits,
                              hipStream_t st);
// Scheme I step 4 from blocked planes: C = fma(T_g, scale_g, C) for g ascending, then the output rule.
hipError_t launch_recon_s1_blocked(const Dims& d, int beta, const int32_t* T, const int32_t* shift_A,
                                   const int32_t* shift_B, const OutC& C, hipStream_t st, int64_t row0 = 0,
                                   int64_t rows = -1, bool narrow = false);
hipError_t launch_recon_s1_groups(const Dims& d, int beta, const int32_t* T, const int32_t* shift_A,
                                  const int32_t* shift_B, double* C, hipStream_t st);

// Accuracy helpers: double-double GEMM (C_hi = the dd dot product rounded to FP64, optional
// C_abs = |A||B| in FP64) and error statistics of `actual` against `exact`.
hipError_t launch_dd_gemm(int64_t M, int64_t N, int64_t K, const double* A, const double* B, double* C_hi,
                          double* C_abs, hipStream_t st);
// stats (device, 4 doubles, zeroed by the launcher): max |d|, max |d|/|e| (|d| when e == 0),
// max |d|/absab (absab optional), count of mismatching bit patterns.
hipError_t launch_error_stats(int64_t n, const double* exact, const double* actual, const double* absab,
                              double* stats, hipStream_t st);
// INT8 MFMA rate with operands in registers (random residues or zeros), `grid` workgroups of 8 waves;
// acc4 (device, 4 x 64-bit, zeroed by the launcher) receives the summed shader-clock cycles and
// 100 MHz ticks of thread 0 of every workgroup.
hipError_t launch_mfma_ceiling(int iters, bool zero, unsigned grid, double* acc4, hipStream_t st);

}  // namespace oza
