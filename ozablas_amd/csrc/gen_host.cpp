// gen_host.cpp — the reference's synthetic input generator (utils/matrix_gen.hpp:15-34):
// a_ij = ((1 - U) - 0.5) * exp(phi * Z), U ~ uniform_real_distribution<double>(0, 1),
// Z ~ normal_distribution<double>(0, 1), both drawn from one std::mt19937(seed), U before Z,
// row-major.  Built by g++ for baseline x86-64 with -ffp-contract=off so the libstdc++
// distributions round exactly as written (no fused x*x + y*y in the polar method).
#include <cmath>
#include <cstddef>
#include <random>

#include "ozablas_hip.h"

extern "C" int oza_generate_matrix(size_t rows, size_t cols, double phi, int seed, double* out) {
    if (!out && rows != 0 && cols != 0) return OZA_INVALID_ARGUMENT;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    std::normal_distribution<double> norm(0.0, 1.0);
    const size_t n = rows * cols;
    for (size_t i = 0; i < n; ++i) {
        const double u = 1.0 - unif(rng);
        const double z = norm(rng);
        out[i] = (u - 0.5) * std::exp(phi * z);
    }
    return OZA_SUCCESS;
}
