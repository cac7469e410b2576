// ozablas/ozablas.hpp — FP64 GEMM emulated on INT8 matrix cores (Ozaki schemes), gfx950 build.
//
// Drop-in for the reference's public entry points (include/ozablas/ozablas.hpp:20,31): same
// namespace, names, argument meaning and defaults.  Construct the workspace with
// <ozablas/core/workspace.hpp>; it fixes M, N, K and the slice/moduli count.
//
// A, B, C are device pointers to dense row-major FP64 matrices (M x K, K x N, M x N).  C is fully
// overwritten.  The call is asynchronous on the executor's stream unless `timings` is given, in
// which case it returns after the result is complete and fills the per-stage times.
#pragma once

namespace ozablas {

class WorkspaceScheme1;
class WorkspaceScheme2;
struct OzaTimings;

// Scheme I: mantissa slices, products summed group by group and scaled ("sum and scale").
void ozaki_scheme1_gemm(WorkspaceScheme1& ws, const double* A, const double* B, double* C,
                        OzaTimings* timings = nullptr);

// Scheme II: symmetric residues modulo up to 20 primes, CRT reconstruction.
void ozaki_scheme2_gemm(WorkspaceScheme2& ws, const double* A, const double* B, double* C,
                        OzaTimings* timings = nullptr);

}  // namespace ozablas
