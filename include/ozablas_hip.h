/*
 * ozablas_hip.h — C ABI of the MI355X (gfx950) Ozaki FP64-GEMM emulation library
 * (libozablas_hip.so).  Plain pointers and sizes only; every entry point returns an
 * oza_status (0 on success) and leaves a message for oza_last_error() on failure.
 *
 * The C++ drop-in API (include/ozablas/...) is implemented on top of these functions, and each
 * function names the reference interface it replaces (file:line into IlyaNyrkov/ozablas).
 *
 * Matrices are dense row-major device arrays: A is M x K, B is K x N, C is M x N (FP64).
 * Work is enqueued on the context's stream (a blocking HIP stream, so it is ordered with
 * null-stream work) and is asynchronous unless a timings pointer is passed, exactly like the
 * reference (src/hip/pipeline.hip:262-264).
 */
#ifndef OZABLAS_HIP_H
#define OZABLAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OZA_MAX_SLICES 20

/* Status codes; the C++ layer maps them to the reference's exception types. */
typedef enum oza_status {
    OZA_SUCCESS = 0,
    OZA_INVALID_ARGUMENT = 1, /* std::invalid_argument (src/core/workspace.cpp:24-25, 68-71) */
    OZA_ALLOC_FAILED = 2,     /* std::bad_alloc       (src/core/executor.cpp:154-157) */
    OZA_HIP_ERROR = 3,        /* std::runtime_error   (src/core/executor.cpp:138-144, 164-185) */
    OZA_NOT_SUPPORTED = 4     /* std::runtime_error   (src/ozablas.cpp:37-45) */
} oza_status;

/* Per-call stage times in milliseconds; mirrors struct OzaTimings
 * (include/ozablas/core/workspace.hpp:97-103). */
typedef struct oza_timings {
    float step1_ms, step2_ms, step3_ms, step4_ms, total_ms;
} oza_timings;

typedef struct oza_context* oza_ctx; /* device, stream, resident CRT tables, events */

const char* oza_version(void);
const char* oza_status_string(int status);
/* Message of the last failure on the calling thread ("" if none). */
const char* oza_last_error(void);

/* ---------------------------------------------------------------- execution context
 * Replaces HipExecutor (include/ozablas/core/executor.hpp:91-112, src/core/executor.cpp:137-190)
 * and the process-once constant upload (src/hip/pipeline.hip:30-37, src/core/workspace.cpp:94-114):
 * the CRT tables are uploaded per context, i.e. per device. */
int oza_get_device_count(int* count);
int oza_ctx_create(int device_id, oza_ctx* out);
int oza_ctx_destroy(oza_ctx ctx);
int oza_ctx_device(oza_ctx ctx, int* device_id);
int oza_ctx_stream(oza_ctx ctx, void** hip_stream);
/* HipExecutor::allocate / free / copy_* / synchronize (executor.cpp:154-186). */
int oza_malloc(oza_ctx ctx, void** ptr, size_t bytes);
int oza_free(oza_ctx ctx, void* ptr);
int oza_memcpy_h2d(oza_ctx ctx, void* dst, const void* src, size_t bytes);
int oza_memcpy_d2h(oza_ctx ctx, void* dst, const void* src, size_t bytes);
int oza_memcpy_d2d(oza_ctx ctx, void* dst, const void* src, size_t bytes);
/* Blocks until all work on the context's device is done (hipDeviceSynchronize, as the
 * reference's HipExecutor::synchronize, executor.cpp:182-186). */
int oza_synchronize(oza_ctx ctx);
/* Blocks until the work enqueued on the context's stream is done. */
int oza_stream_synchronize(oza_ctx ctx);
/* HipExecutor::get_blas_handle (executor.cpp:188-190): a rocBLAS handle bound to the context
 * stream, created on first use.  Comparator use only (native DGEMM); never on the Ozaki path. */
int oza_blas_handle(oza_ctx ctx, void** handle);

/* Stage events for a loop of pipeline calls, recorded on the context stream without
 * synchronizing (no counterpart in the reference, whose OzaTimings path blocks every call).
 * While enabled, each oza_scheme*_gemm call made without a timings pointer records five events
 * (up to 512 calls).  oza_ctx_event_times waits for the last of them and returns, summed over
 * the recorded calls, {step1, step2, step3, step4, total} in ms plus the number of calls, then
 * starts a new collection. */
/* Row bands of steps 3 and 4, both schemes (no counterpart in the reference, which runs the steps
 * in sequence, src/hip/pipeline.hip:101-145, 237-262).  With nb > 1 bands, band b's reconstruction
 * runs on a second stream of the context beside band b+1's integer products; the call still
 * completes on the context stream and its results do not depend on nb.  0 (default) picks nb per
 * call from the shape, 1 keeps the steps in sequence, 2..8 force that many bands (at most one per
 * 256 rows).  With nb > 1, step4_ms of oza_timings and of oza_ctx_event_times is the part of
 * step 4 left after the last products. */
int oza_ctx_set_bands(oza_ctx ctx, int bands);
/* The band count a call of this scheme and shape makes on this context (1 = steps in sequence);
 * each band is one launch of the step-3 kernel. */
int oza_ctx_bands_for(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int S, int* bands);
int oza_ctx_event_timing(oza_ctx ctx, int enable);
int oza_ctx_event_times(oza_ctx ctx, double sums[5], int* count);

/* ---------------------------------------------------------------- workspace sizing
 * Replaces the allocation list of WorkspaceScheme1/2 (src/core/workspace.cpp:27-39, 73-84).
 * scheme is 1 or 2.  bytes[5] = {shift_A, shift_B, A_slices, B_slices, C_tc}.  Layouts:
 *   shift_A int32[M], shift_B int32[N]          (reference semantics)
 *   A_slices int8 [S][Mp][Kp], B_slices int8 [S][Np][Kp] (B transposed: K contiguous)
 *   C_tc   Scheme II: S planes of uint8 CRT digits C_s mod m_s, Mp x Np each in 32 x 32 element
 *          blocks (oza_unpack_digits gives the planar view), followed, for shapes whose GEMM runs
 *          in K parts (few tiles, long K), by the partial digits of parts 1..; Scheme I: S planes of int32 group
 *          products T_g, Mp x Np each in 32 x 32 element blocks, and at least the reference's
 *          planar int32 [S][M][N] size, which the stage entry points write
 * with Mp = roundup(M,256), Np = roundup(N,256), Kp = roundup(K,128).
 * Fails with OZA_INVALID_ARGUMENT for S outside 1..20 or K beyond exact INT32 accumulation. */
int oza_workspace_bytes(int scheme, size_t M, size_t N, size_t K, int slices, size_t bytes[5]);

/* ---------------------------------------------------------------- whole pipelines
 * Replace hip::ozaki_scheme1_gemm (src/hip/pipeline.hip:48-176) and hip::ozaki_scheme2_gemm
 * (src/hip/pipeline.hip:182-275), reached from ozablas::ozaki_scheme{1,2}_gemm
 * (include/ozablas/ozablas.hpp:20,31).  Workspace buffers as sized by oza_workspace_bytes.
 * With timings != NULL the call blocks until the result is ready and fills the stage times.
 * Large calls overlap step 4 with step 3 in row bands (oza_ctx_set_bands): step3_ms then runs to
 * the end of the last band's products and includes the earlier bands' reconstruction beside them,
 * and step4_ms is the last band's reconstruction. */
int oza_scheme1_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const double* A, const double* B,
                     double* C, int32_t* shift_A, int32_t* shift_B, int8_t* A_slices, int8_t* B_slices,
                     void* C_tc, oza_timings* timings);
int oza_scheme2_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const double* A, const double* B,
                     double* C, int32_t* shift_A, int32_t* shift_B, int8_t* A_slices, int8_t* B_slices,
                     void* C_tc, oza_timings* timings);

/* ---------------------------------------------------------------- single stages
 * The reference's stage kernels one at a time, on the same workspace buffers, for inspection
 * and bit-exact parity checks.  Step 1: compute_row_shifts_A / compute_col_shifts_B
 * (src/pipeline/step1_statistics.hpp:23-99) with offset 0 (Scheme I) or K_param (Scheme II). */
int oza_stage1_shifts(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int slices, const double* A,
                      const double* B, int32_t* shift_A, int32_t* shift_B);
/* Step 2: slice_scheme{1,2}_{A,B} (src/pipeline/step2_slicing.hpp:44-206), product layout. */
int oza_stage2_slices(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int slices, const double* A,
                      const double* B, const int32_t* shift_A, const int32_t* shift_B, int8_t* A_slices,
                      int8_t* B_slices);
/* Product layout -> the reference's planar layout: A_planar int8 [S][M][K], B_planar int8 [S][K][N]
 * (what WorkspaceScheme*::get_{A,B}_slices point at in the reference, workspace.hpp:29-33).
 * Either output may be NULL. */
int oza_unpack_slices(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int8_t* A_slices,
                      const int8_t* B_slices, int8_t* A_planar, int8_t* B_planar);
/* Step 3, Scheme II: C_s = A_s B_s as int32 [S][M][N] (rocblas_gemm_strided_batched_ex,
 * src/hip/pipeline.hip:237-245), or as CRT digits in the workspace's C_tc layout. */
int oza_stage3_s2_products(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int8_t* A_slices,
                           const int8_t* B_slices, int32_t* C_planar);
int oza_stage3_s2_digits(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int8_t* A_slices,
                         const int8_t* B_slices, void* C_tc);
/* Digits in C_tc layout -> planar uint8 [S][M][N]. */
int oza_unpack_digits(oza_ctx ctx, size_t M, size_t N, int slices, const void* C_tc, uint8_t* planar);
/* The inverse: planar uint8 [S][M][N] digits (taken modulo m_s) -> C_tc layout, padding 0; with
 * oza_stage4_s2_digits it reconstructs C from any residues.  No counterpart in the reference. */
int oza_pack_digits(oza_ctx ctx, size_t M, size_t N, int slices, const uint8_t* planar, void* C_tc);
/* Step 3, Scheme I: group products T_g, g = 2..S+1, as int32 [S][M][N] (the rocblas_gemm_ex loop,
 * src/hip/pipeline.hip:110-135). */
int oza_stage3_s1_groups(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int8_t* A_slices,
                         const int8_t* B_slices, int32_t* T_planar);
/* Step 4: reconstruct_scheme2_leq7 / _gt7 (src/pipeline/step4_reconstruction.hpp:59-184) from
 * int32 planes or from digits, and reconstruct_scheme1_group over all groups
 * (step4_reconstruction.hpp:21-49, pipeline.hip:80,139-141) from int32 group planes. */
int oza_stage4_s2_products(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int32_t* C_planar,
                           const int32_t* shift_A, const int32_t* shift_B, double* C);
int oza_stage4_s2_digits(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const void* C_tc,
                         const int32_t* shift_A, const int32_t* shift_B, double* C);
int oza_stage4_s1_groups(oza_ctx ctx, size_t M, size_t N, size_t K, int slices, const int32_t* T_planar,
                         const int32_t* shift_A, const int32_t* shift_B, double* C);

/* ---------------------------------------------------------------- parameters and inputs
 * calculate_scheme1_beta / calculate_scheme2_k_param (src/pipeline/step1_statistics.hpp:12-20,
 * src/hip/pipeline.hip:205-207). */
int oza_scheme1_beta(size_t K, int* beta);
int oza_scheme2_kparam(int slices, size_t K, int* kparam);
/* Accuracy target -> moduli count (no counterpart in the reference, which takes S directly,
 * README.md:43): the smallest S whose K_param reaches ceil(digits * log2 10) bits, i.e. the
 * truncation bound |C_ij - C'_ij| <= K 2^(E_i+F_j) 10^-digits, with E_i, F_j the row/column
 * exponents of step 1. */
int oza_scheme2_slices_for_digits(size_t K, double digits, int* slices);
/* Accuracy target -> moduli count for componentwise digits (no counterpart in the reference): the
 * smallest S whose K_param reaches digits + s(spread) decimal digits, so that
 * max |C_ij - C'_ij| / (|A||B|)_ij <= 10^-digits for inputs whose magnitudes spread like
 * exp(spread * Z), Z standard normal (spread = the reference generator's phi; 0 for uniform data).
 * s(spread) is the worst componentwise shortfall measured against FP128 products (0.8, 1.3, 2.2,
 * 3.5, 4.8, 7.2, 9.3 digits at spread 0, 0.5, 1, 1.5, 2, 3, 4, linear in between, +2.5 per unit
 * beyond; outputs up to 8192 x 8192).  Fails when no S <= 20 reaches the target. */
int oza_scheme2_slices_for_accuracy(size_t K, double digits, double spread, int* slices);
/* The reference's input generator (utils/matrix_gen.hpp:15-34), host memory. */
int oza_generate_matrix(size_t rows, size_t cols, double phi, int seed, double* host_out);

/* ---------------------------------------------------------------- BLAS-style front end
 * C = alpha * op(A) op(B) + beta * C with row- or column-major storage, transposes and leading
 * dimensions, emulated with Scheme I or II.  No counterpart in the reference, whose entry points
 * take dense row-major A, B, C only (include/ozablas/ozablas.hpp:20,31).  With alpha = 1 and
 * beta = 0 the result is bit-identical to oza_scheme{1,2}_gemm on dense row-major copies of op(A)
 * and op(B) (a column-major call is the row-major product C^T = op(B)^T op(A)^T).  Otherwise each
 * entry is fma(alpha, X, beta * C) of the emulated product X; beta == 0 does not read C, and
 * alpha == 0 or K == 0 only scales C, as in BLAS.  workspace may be NULL: the context then keeps a
 * workspace of its own, grown on demand; otherwise it must hold oza_gemm_workspace_bytes.
 * Asynchronous on the context's stream. */
typedef enum oza_layout { OZA_ROW_MAJOR = 0, OZA_COL_MAJOR = 1 } oza_layout;
typedef enum oza_op { OZA_OP_N = 0, OZA_OP_T = 1 } oza_op;
int oza_gemm_workspace_bytes(int scheme, size_t M, size_t N, size_t K, int slices, size_t* bytes);
int oza_gemm(oza_ctx ctx, int scheme, int slices, int layout, int transA, int transB, size_t M, size_t N, size_t K,
             double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
             size_t ldc, void* workspace, size_t workspace_bytes);
/* oza_gemm that allocates nothing: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as oza_gemm, without the workspace: the call's shifts,
 * operand planes and digit (or group-product) planes live in one scratch array of
 * OZA_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on until the process
 * ends.  A problem whose oza_gemm_workspace_bytes exceeds that size is rejected with
 * OZA_INVALID_ARGUMENT before anything is enqueued.  All calls on one device share the array, so
 * calls that overlap on one device (two streams, two contexts, two threads) race: order them. */
#define OZA_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int oza_gemm_noalloc(oza_ctx ctx, int scheme, int slices, int layout, int transA, int transB, size_t M, size_t N,
                     size_t K, double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta,
                     double* C, size_t ldc);

/* ---------------------------------------------------------------- grouped Scheme II
 * A group is a planned batch of independent Scheme II GEMMs, each with its own M, N, K and moduli
 * count.  No counterpart in the reference, whose API runs one dense GEMM per call
 * (include/ozablas/ozablas.hpp:31, src/hip/pipeline.hip:182-275): a caller with many small or
 * uneven GEMMs would loop over ozaki_scheme2_gemm.  Every problem's C is bit-identical to
 * oza_scheme2_gemm's for the same M, N, K, S.
 *
 * oza_group_create plans the batch (per-problem GEMM tile shape, work lists, one workspace owned
 * by the group).  tile = 0 chooses each problem's tile edge from the context's cost model
 * (oza_group_autotune measures it on the device); 256, 128 or 64 forces one edge for all.
 * oza_group_scheme2_gemm runs it: A, B, C are host arrays of `count` device pointers (the same
 * layout rules as oza_scheme2_gemm).  With use_graph != 0 the stage launches are captured once
 * into a HIP graph and replayed on later calls (new pointers are uploaded before each replay);
 * timings (optional, blocks like oza_scheme2_gemm's) covers the whole batch and forces the
 * eager path.  Calls on one group are ordered on the context's stream; a group is not
 * thread-safe. */
typedef struct oza_group* oza_group_t;
int oza_group_create(oza_ctx ctx, int count, const size_t* M, const size_t* N, const size_t* K, const int* slices,
                     int tile, oza_group_t* out);
/* oza_group_create with the GEMM plan given per problem: tiles[i] in {0, 256, 128, 64} (0: the
 * cost model's choice) and splits[i] in 0..64, the number of K splits of the problem's GEMM (0:
 * the planner's choice, 1: none; clamped to the problem's 128-byte K blocks).  Either array may be
 * NULL (all automatic).  A split GEMM computes the digits of each part's product and adds them mod
 * m_s, so every result is bit-identical to the unsplit one.  The planner splits a problem when its
 * GEMM items are long against the whole batch's (a problem with few M x N tiles and a long K,
 * alone or among short ones), judged with the tile cost model of oza_group_autotune. */
int oza_group_create_split(oza_ctx ctx, int count, const size_t* M, const size_t* N, const size_t* K,
                           const int* slices, const int* tiles, const int* splits, oza_group_t* out);
int oza_group_destroy(oza_group_t group);
/* splits (count entries): the K splits of each problem's GEMM (1: not split). */
int oza_group_splits(oza_group_t group, int* splits);
/* The planner alone, on the host (no device needed): plans `count` problems for a device described
 * by the tile cost model cost[6] (as oza_group_costs), its CU count and slots[3], the workgroups of
 * the 256, 128 and 64 tile shapes resident at once.  tiles and splits (count entries each) are as
 * in oza_group_create_split on entry (0: planner's choice) and hold the plan's values on return;
 * workspace_bytes (may be NULL) receives the group's workspace size. */
int oza_group_plan(const double cost[6], int cus, const int slots[3], int count, const size_t* M, const size_t* N,
                   const size_t* K, const int* slices, int* tiles, int* splits, size_t* workspace_bytes);
/* workspace_bytes: device bytes the group holds; tiles (count entries, may be NULL): the GEMM
 * tile edge chosen for each problem. */
int oza_group_info(oza_group_t group, size_t* workspace_bytes, int* tiles);
int oza_group_scheme2_gemm(oza_group_t group, const double* const* A, const double* const* B, double* const* C,
                           int use_graph, oza_timings* timings);
/* The tile cost model: a GEMM tile of edge 256, 128, 64 (c = 0, 1, 2) over Kp/128 K steps costs
 * cost[2c] + cost[2c+1] * Kp/128 CU-microseconds.  oza_group_autotune measures it on the
 * context's device (a few GEMMs of 4096 x 4096, K = 128 and 1024), stores it in the context and
 * returns it; oza_group_costs reads (set == 0) or replaces (set != 0) it. */
int oza_group_autotune(oza_ctx ctx, double cost[6]);
int oza_group_costs(oza_ctx ctx, double cost[6], int set);

/* ---------------------------------------------------------------- request server (Scheme II)
 * Continuous batching of independent GEMM requests.  No counterpart in the reference (one
 * synchronous-or-async GEMM per call, src/hip/pipeline.hip:182-275).  The server owns two
 * resident workspace arenas of workspace_bytes / 2 each, each with its own HIP stream, so two
 * batches run on the device at once.  A worker thread admits queued requests into a free arena
 * (as many as fit its bytes, at most max_batch per batch); a completion thread per arena marks
 * requests done.  A request is large when its INT8 work 2 M N K S exceeds a threshold (default
 * 1e12, oza_server_config): a batch holds only large or only small requests, at most one large
 * batch runs at a time, and while it runs the other arena admits small requests only, in their
 * order of arrival, ahead of waiting large ones, so a small request never completes with a large
 * batch.  A large batch's GEMM can be held to a share of the CUs (default 1.0, all of them), which
 * leaves CUs to the small batches beside it at some cost in throughput.  Each result is
 * bit-identical to oza_scheme2_gemm for the same request alone.
 *
 * oza_server_submit is non-blocking and thread-safe; A, B, C must stay valid until the request
 * completes.  It rejects a request that can never fit one arena.  oza_server_wait blocks until
 * the request is complete, reports its latency (submission to observed completion, ms) and
 * releases the ticket.  oza_server_destroy drains everything queued.  stats[4] = {batches,
 * requests served, mean batch size, largest batch}.  oza_server_config sets the large-request
 * threshold (INT8 ops, >= 0; later submissions) and the large batches' GEMM CU share (0, 1]. */
typedef struct oza_server* oza_server_t;
int oza_server_create(oza_ctx ctx, size_t workspace_bytes, int max_batch, oza_server_t* out);
int oza_server_destroy(oza_server_t server);
int oza_server_submit(oza_server_t server, size_t M, size_t N, size_t K, int slices, const double* A,
                      const double* B, double* C, uint64_t* ticket);
int oza_server_wait(oza_server_t server, uint64_t ticket, double* latency_ms);
int oza_server_drain(oza_server_t server);
int oza_server_stats(oza_server_t server, double stats[4]);
int oza_server_config(oza_server_t server, double large_ops, double large_gemm_share);

/* ---------------------------------------------------------------- accuracy reporting
 * No counterpart in the reference, which compares on the host against an FP128 product
 * (utils/matrix_reference.hpp:22-51) with utils/matrix_compare.hpp:24-65. */
/* Double-double GEMM: C_ref = the dot products in ~106-bit arithmetic rounded to FP64;
 * C_abs (optional) = |A||B| in FP64. */
int oza_dd_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, const double* A, const double* B, double* C_ref,
                double* C_abs);
/* Native FP64 GEMM (rocBLAS DGEMM, row-major), the comparator of the reference's drivers
 * (e.g. performance_benchmarks/01_ozaki_scheme1_2_raw_performance.cpp:64-65). */
int oza_dgemm(oza_ctx ctx, size_t M, size_t N, size_t K, const double* A, const double* B, double* C);
/* The INT8 MFMA rate the device sustains with operands in registers and no memory traffic (one
 * workgroup of 8 waves per CU, v_mfma_i32_32x32x32_i8, eight accumulators per wave), on random
 * residues in [-125, 125] (zero_operands = 0) or on zeros: out[0] = TOP/s, out[1] = the mean shader
 * clock in GHz over the run (s_memtime against s_memrealtime).  Random operands draw more power, so
 * the clock and the rate fall below the nominal peak; this is the ceiling a GEMM on random residues
 * can approach.  Blocks until done (about 50 ms).  Measurement tooling, no counterpart in the
 * reference. */
int oza_mfma_ceiling(oza_ctx ctx, int zero_operands, double out[2]);
/* out[4] (host) = {max |d|, max |d|/|exact| (|d| where exact == 0), max |d|/absab (absab may be
 * NULL), number of entries whose bits differ}.  Blocks until done. */
int oza_error_stats(oza_ctx ctx, size_t n, const double* exact, const double* actual, const double* absab,
                    double out[4]);

#ifdef __cplusplus
}
#endif
#endif
