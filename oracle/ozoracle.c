/*
 * ozoracle.c — CPU restatement of the OzaBLAS Ozaki Scheme I / II pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see ozoracle.h for
 * who may load it and how each function is pinned.  Built by oracle/Makefile with
 * -ffp-contract=off so that every floating-point operation below rounds exactly
 * where the C source says it does.
 *
 * Citations are file:line into the reference (IlyaNyrkov/ozablas).
 */
#include "ozoracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* The moduli list, src/common/crt_tables.hpp:13-16.  This list is the one piece of
 * data that has to be identical to the reference (it defines the residues). */
static const uint8_t kModuli[ORC_MAX_SLICES] = {
    251, 241, 239, 233, 229, 227, 223, 211, 199, 197,
    193, 191, 181, 179, 173, 167, 163, 157, 151, 149};

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------
 * Input generator: utils/matrix_gen.hpp:15-34 restated on top of a from-scratch
 * MT19937 (Matsumoto & Nishimura 1998) and the libstdc++ distribution algorithms:
 *   uniform_real_distribution<double>(0,1) = generate_canonical<double,53>: two 32-bit
 *     draws, (x0 + x1*2^32) / 2^64, clamped below 1 (bits/random.tcc generate_canonical);
 *   normal_distribution<double> = Marsaglia polar method with one cached deviate.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t mt[624];
    int idx;
} mt19937_t;

static void mt_seed(mt19937_t *g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(mt19937_t *g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static double canonical53(mt19937_t *g) {
    double sum = 0.0, tmp = 1.0;
    sum += (double)mt_next(g) * tmp;
    tmp *= 4294967296.0;
    sum += (double)mt_next(g) * tmp;
    tmp *= 4294967296.0;
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

typedef struct {
    int saved_available;
    double saved;
} polar_t;

static double normal_polar(polar_t *st, mt19937_t *g) {
    double ret;
    if (st->saved_available) {
        st->saved_available = 0;
        ret = st->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * canonical53(g) - 1.0;
            y = 2.0 * canonical53(g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        double mult = sqrt(-2.0 * log(r2) / r2);
        st->saved = x * mult;
        st->saved_available = 1;
        ret = y * mult;
    }
    return ret * 1.0 + 0.0;
}

void orc_generate_matrix(int64_t rows, int64_t cols, double phi, int seed, double *out) {
    mt19937_t g;
    polar_t st = {0, 0.0};
    mt_seed(&g, (uint32_t)seed);
    int64_t n = rows * cols;
    for (int64_t i = 0; i < n; ++i) {
        double u = 1.0 - canonical53(&g); /* matrix_gen.hpp:27 */
        double z = normal_polar(&st, &g); /* matrix_gen.hpp:28 */
        out[i] = (u - 0.5) * exp(phi * z); /* matrix_gen.hpp:30 */
    }
}

/* ------------------------------------------------------------------------------------
 * CRT tables from closed forms (the reference stores them as literals,
 * src/common/crt_tables.hpp:19-112).  256-bit values are 4 little-endian u64 limbs.
 * ---------------------------------------------------------------------------------- */
static void big_mul_small(uint64_t a[4], uint64_t m) {
    u128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 p = (u128)a[i] * m + carry;
        a[i] = (uint64_t)p;
        carry = p >> 64;
    }
}

static uint64_t big_div_small(const uint64_t a[4], uint64_t m, uint64_t q[4]) {
    u128 rem = 0;
    for (int i = 3; i >= 0; --i) {
        u128 cur = (rem << 64) | a[i];
        q[i] = (uint64_t)(cur / m);
        rem = cur % m;
    }
    return (uint64_t)rem;
}

void orc_moduli(uint8_t out[ORC_MAX_SLICES]) { memcpy(out, kModuli, ORC_MAX_SLICES); }

int orc_crt_tables(int S, uint64_t M[4], uint64_t M_half[4],
                   uint64_t partial[ORC_MAX_SLICES][4], uint64_t inv[ORC_MAX_SLICES]) {
    if (S < 1 || S > ORC_MAX_SLICES) return -1;
    uint64_t prod[4] = {1, 0, 0, 0};
    for (int i = 0; i < S; ++i) big_mul_small(prod, kModuli[i]);
    memcpy(M, prod, sizeof(prod));
    for (int i = 0; i < 4; ++i) /* floor(M/2) */
        M_half[i] = (prod[i] >> 1) | (i < 3 ? (prod[i + 1] << 63) : 0);
    for (int s = 0; s < ORC_MAX_SLICES; ++s) {
        memset(partial[s], 0, 4 * sizeof(uint64_t));
        inv[s] = 0;
    }
    for (int s = 0; s < S; ++s) {
        uint64_t q[4];
        uint64_t r = big_div_small(prod, kModuli[s], q); /* M / m_s, remainder 0 */
        if (r != 0) return -2;
        memcpy(partial[s], q, sizeof(q));
        uint64_t qmod[4];
        uint64_t qm = big_div_small(q, kModuli[s], qmod); /* (M/m_s) mod m_s */
        uint64_t found = 0;
        for (uint64_t c = 1; c < kModuli[s]; ++c)
            if ((qm * c) % kModuli[s] == 1) { found = c; break; }
        if (found == 0) return -3;
        inv[s] = found;
    }
    return 0;
}

/* src/common/crt_math.hpp:73-80 — limb-by-limb conversion, four roundings. */
double orc_to_double_256(const uint64_t a[4]) {
    double res = 0.0;
    res += (double)a[0];
    res += ldexp((double)a[1], 64);
    res += ldexp((double)a[2], 128);
    res += ldexp((double)a[3], 192);
    return res;
}

/* src/pipeline/step1_statistics.hpp:12-15 */
int orc_scheme1_beta(int64_t K) {
    int beta = (int)floor((31.0 - log2((double)K)) / 2.0);
    if (beta < 1) beta = 1;
    if (beta > 7) beta = 7;
    return beta;
}

/* src/pipeline/step1_statistics.hpp:17-20, src/hip/pipeline.hip:205-207 */
int orc_scheme2_kparam(int S, int64_t K) {
    uint64_t M[4], Mh[4], part[ORC_MAX_SLICES][4], inv[ORC_MAX_SLICES];
    if (orc_crt_tables(S, M, Mh, part, inv) != 0) return -1000000;
    double M_double = orc_to_double_256(M);
    double frac = (M_double / 2.0 - 1.0) / (double)K;
    return (int)floor(0.5 * log2(frac));
}

/* ------------------------------------------------------------------------------------
 * Step 1: src/pipeline/step1_statistics.hpp:23-99.  E = max(-1023, max frexp exponent of
 * the nonzero, non-NaN |entries|); shift = offset - E.
 * ---------------------------------------------------------------------------------- */
static inline int exp_of(double x, int cur) {
    double v = fabs(x);
    if (v > 0.0) {
        int e;
        frexp(v, &e);
        if (e > cur) cur = e;
    }
    return cur;
}

void orc_row_shifts(const double *A, int64_t rows, int64_t cols, int32_t offset, int32_t *shifts) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i) {
        int m = -1023;
        for (int64_t j = 0; j < cols; ++j) m = exp_of(A[i * cols + j], m);
        shifts[i] = offset - m;
    }
}

void orc_col_shifts(const double *B, int64_t rows, int64_t cols, int32_t offset, int32_t *shifts) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < cols; ++j) {
        int m = -1023;
        for (int64_t i = 0; i < rows; ++i) m = exp_of(B[i * cols + j], m);
        shifts[j] = offset - m;
    }
}

/* ------------------------------------------------------------------------------------
 * Step 2, Scheme I: src/pipeline/step2_slicing.hpp:20-25 (extract_fp64_rn) and :44-100.
 * ---------------------------------------------------------------------------------- */
static inline double extract_rn(double val, double sigma) {
    volatile double tmp = val + sigma;
    return tmp - sigma;
}

static void slice_s1_elem(double a, int e, int S, int beta, int8_t *out, int64_t plane) {
    double scale = ldexp(1.0, e + 1 - beta);
    double inv_scale = ldexp(1.0, -(e + 1 - beta));
    for (int s = 0; s < S; ++s) {
        double sigma = 0.75 * ldexp(scale, 53 - beta * s);
        double a_s = extract_rn(a, sigma);
        double normalized = a_s * inv_scale;
        double scaled = ldexp(normalized, beta * s);
        out[(int64_t)s * plane] = (int8_t)(int)scaled;
        a -= a_s;
    }
}

void orc_slice_s1_rows(const double *A, const int32_t *neg_exp, int64_t rows, int64_t cols,
                       int S, int beta, int8_t *out) {
    int64_t plane = rows * cols;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j)
            slice_s1_elem(A[i * cols + j], -neg_exp[i], S, beta, out + i * cols + j, plane);
}

void orc_slice_s1_cols(const double *B, const int32_t *neg_exp, int64_t rows, int64_t cols,
                       int S, int beta, int8_t *out) {
    int64_t plane = rows * cols;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j)
            slice_s1_elem(B[i * cols + j], -neg_exp[j], S, beta, out + i * cols + j, plane);
}

/* ------------------------------------------------------------------------------------
 * Step 2, Scheme II: src/pipeline/step2_slicing.hpp:29-38 and :110-206.
 * ---------------------------------------------------------------------------------- */
int orc_symmetric_mod(int64_t a, int m) {
    int64_t im = m;
    int64_t r = a % im;
    int64_t half = im >> 1;
    if (r > half) r -= im;
    else if (r < -half) r += im;
    return (int)r;
}

static void slice_s2_elem(double a, int shift, int S, int8_t *out, int64_t plane) {
    double scaled = ldexp(a, shift);
    if (fabs(scaled) < 9.0e18) {
        int64_t t = (int64_t)llrint(scaled);
        for (int s = 0; s < S; ++s) out[(int64_t)s * plane] = (int8_t)orc_symmetric_mod(t, kModuli[s]);
    } else {
        double rounded = rint(scaled);
        for (int s = 0; s < S; ++s) {
            int32_t m = kModuli[s];
            double rem_f = fmod(rounded, (double)m);
            int32_t r = (int32_t)rem_f;
            if (r < 0) r += m;
            int32_t half = m >> 1;
            if (r > half) r -= m;
            out[(int64_t)s * plane] = (int8_t)r;
        }
    }
}

void orc_slice_s2_rows(const double *A, const int32_t *shifts, int64_t rows, int64_t cols,
                       int S, int8_t *out) {
    int64_t plane = rows * cols;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j)
            slice_s2_elem(A[i * cols + j], shifts[i], S, out + i * cols + j, plane);
}

void orc_slice_s2_cols(const double *B, const int32_t *shifts, int64_t rows, int64_t cols,
                       int S, int8_t *out) {
    int64_t plane = rows * cols;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < cols; ++j)
            slice_s2_elem(B[i * cols + j], shifts[j], S, out + i * cols + j, plane);
}

/* ------------------------------------------------------------------------------------
 * Step 3: exact integer GEMM (the reference calls rocBLAS here, pipeline.hip:125-133,
 * 237-245).  |entries| <= 128 so int32 accumulation is exact for K < 2^17; beyond that
 * int64 accumulation with a count of entries outside int32.
 * ---------------------------------------------------------------------------------- */
static int64_t igemm_acc(const int8_t *A, const int8_t *B, int64_t M, int64_t N, int64_t K,
                         int32_t *C, int accumulate) {
    int64_t bad = 0;
    if (K < (1 << 17)) {
#pragma omp parallel
        {
            int32_t *acc = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
#pragma omp for schedule(static)
            for (int64_t i = 0; i < M; ++i) {
                for (int64_t j = 0; j < N; ++j) acc[j] = accumulate ? C[i * N + j] : 0;
                for (int64_t k = 0; k < K; ++k) {
                    int32_t a = A[i * K + k];
                    const int8_t *b = B + k * N;
                    for (int64_t j = 0; j < N; ++j) acc[j] += a * (int32_t)b[j];
                }
                for (int64_t j = 0; j < N; ++j) C[i * N + j] = acc[j];
            }
            free(acc);
        }
    } else {
#pragma omp parallel reduction(+ : bad)
        {
            int64_t *acc = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
#pragma omp for schedule(static)
            for (int64_t i = 0; i < M; ++i) {
                for (int64_t j = 0; j < N; ++j) acc[j] = accumulate ? C[i * N + j] : 0;
                for (int64_t k = 0; k < K; ++k) {
                    int64_t a = A[i * K + k];
                    const int8_t *b = B + k * N;
                    for (int64_t j = 0; j < N; ++j) acc[j] += a * (int64_t)b[j];
                }
                for (int64_t j = 0; j < N; ++j) {
                    if (acc[j] > INT32_MAX || acc[j] < INT32_MIN) ++bad;
                    C[i * N + j] = (int32_t)acc[j];
                }
            }
            free(acc);
        }
    }
    return bad;
}

int64_t orc_igemm(const int8_t *A, const int8_t *B, int64_t M, int64_t N, int64_t K, int32_t *C) {
    return igemm_acc(A, B, M, N, K, C, 0);
}

int64_t orc_s2_products(const int8_t *A_sl, const int8_t *B_sl, int64_t M, int64_t N, int64_t K,
                        int S, int32_t *C_tc) {
    int64_t bad = 0;
    for (int s = 0; s < S; ++s)
        bad += igemm_acc(A_sl + (int64_t)s * M * K, B_sl + (int64_t)s * K * N, M, N, K,
                         C_tc + (int64_t)s * M * N, 0);
    return bad;
}

/* src/hip/pipeline.hip:110-135: groups g = 2..S+1, pairs s + t = g (1-based). */
int64_t orc_s1_group_products(const int8_t *A_sl, const int8_t *B_sl, int64_t M, int64_t N,
                              int64_t K, int S, int32_t *T) {
    int64_t bad = 0;
    for (int g = 2; g <= S + 1; ++g) {
        int first = 1;
        int32_t *Tg = T + (int64_t)(g - 2) * M * N;
        for (int s = 1; s <= g - 1; ++s) {
            int t = g - s;
            if (s > S || t > S) continue;
            bad += igemm_acc(A_sl + (int64_t)(s - 1) * M * K, B_sl + (int64_t)(t - 1) * K * N,
                             M, N, K, Tg, first ? 0 : 1);
            first = 0;
        }
    }
    return bad;
}

/* ------------------------------------------------------------------------------------
 * Step 4, Scheme I: src/pipeline/step4_reconstruction.hpp:21-49, groups in ascending g on
 * a zeroed C (pipeline.hip:80, 110-144).  The reference kernel's `C += val * scale` is
 * compiled by hipcc with FP contraction, i.e. as one fused multiply-add; it is written
 * here as fma() (identical to mul-then-add whenever val*scale is exact, i.e. whenever the
 * scale is a normal power of two).
 * ---------------------------------------------------------------------------------- */
void orc_recon_s1(const int32_t *T, const int32_t *neg_exp_A, const int32_t *neg_exp_B,
                  int64_t M, int64_t N, int S, int beta, double *C) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        for (int64_t j = 0; j < N; ++j) {
            double c = 0.0;
            int e_i = -neg_exp_A[i];
            int e_j = -neg_exp_B[j];
            double u_i = ldexp(1.0, e_i + 1 - beta);
            double v_j = ldexp(1.0, e_j + 1 - beta);
            for (int g = 2; g <= S + 1; ++g) {
                int32_t val = T[(int64_t)(g - 2) * M * N + i * N + j];
                double scale = ldexp(u_i * v_j, -beta * g + 2 * beta);
                c = fma((double)val, scale, c);
            }
            C[i * N + j] = c;
        }
    }
}

/* ------------------------------------------------------------------------------------
 * Step 4, Scheme II: src/pipeline/step4_reconstruction.hpp:59-110 (S <= 7, u64) and
 * :116-184 (S > 7, four u64 limbs), with crt_math.hpp:38-80.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    uint64_t d[4];
} u256;

static void add256(u256 *a, const u256 *b) { /* crt_math.hpp:51-61 */
    uint64_t carry = 0;
    for (int i = 0; i < 4; ++i) {
        uint64_t s1 = a->d[i] + carry;
        uint64_t c1 = s1 < a->d[i];
        uint64_t s = s1 + b->d[i];
        uint64_t c2 = s < s1;
        a->d[i] = s;
        carry = c1 | c2;
    }
}

static void sub256(u256 *a, const u256 *b) { /* crt_math.hpp:38-48 */
    uint64_t borrow = 0;
    for (int i = 0; i < 4; ++i) {
        uint64_t s1 = a->d[i] - borrow;
        uint64_t b1 = a->d[i] < borrow;
        uint64_t diff = s1 - b->d[i];
        uint64_t b2 = s1 < b->d[i];
        a->d[i] = diff;
        borrow = b1 | b2;
    }
}

static int ge256(const u256 *a, const u256 *b) { /* crt_math.hpp:64-70 */
    for (int i = 3; i >= 0; --i) {
        if (a->d[i] > b->d[i]) return 1;
        if (a->d[i] < b->d[i]) return 0;
    }
    return 1;
}

void orc_crt_digits(const int32_t *C_tc, int64_t MN, int S, uint8_t *digits) {
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < MN; ++idx)
        for (int s = 0; s < S; ++s) {
            int32_t m = kModuli[s];
            int32_t rem = C_tc[(int64_t)s * MN + idx] % m;
            if (rem < 0) rem += m;
            digits[(int64_t)s * MN + idx] = (uint8_t)rem;
        }
}

void orc_recon_s2(const int32_t *C_tc, const int32_t *shift_A, const int32_t *shift_B,
                  int64_t M_, int64_t N_, int S, double *C) {
    uint64_t Mp[4], Mh[4], part[ORC_MAX_SLICES][4], inv[ORC_MAX_SLICES];
    orc_crt_tables(S, Mp, Mh, part, inv);
    int64_t MN = M_ * N_;
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < M_; ++row) {
        for (int64_t col = 0; col < N_; ++col) {
            int64_t idx = row * N_ + col;
            double f;
            if (S <= 7) {
                uint64_t Mm = Mp[0], half = Mh[0], acc = 0;
                for (int s = 0; s < S; ++s) {
                    int32_t m_i = kModuli[s];
                    int32_t val = C_tc[(int64_t)s * MN + idx];
                    int32_t rem = val % m_i;
                    if (rem < 0) rem += m_i;
                    uint64_t c_i = ((uint64_t)rem * inv[s]) % (uint64_t)m_i;
                    uint64_t term = c_i * part[s][0];
                    acc = (acc + term) % Mm;
                }
                f = acc >= half ? -(double)(Mm - acc) : (double)acc;
            } else {
                u256 Mm, half, acc = {{0, 0, 0, 0}};
                memcpy(Mm.d, Mp, 32);
                memcpy(half.d, Mh, 32);
                for (int s = 0; s < S; ++s) {
                    int32_t m_i = kModuli[s];
                    int32_t val = C_tc[(int64_t)s * MN + idx];
                    int32_t rem = val % m_i;
                    if (rem < 0) rem += m_i;
                    uint64_t c_i = ((uint64_t)rem * inv[s]) % (uint64_t)m_i;
                    u256 term;
                    uint64_t carry = 0;
                    for (int i = 0; i < 4; ++i) {
                        u128 p = (u128)part[s][i] * c_i + carry;
                        term.d[i] = (uint64_t)p;
                        carry = (uint64_t)(p >> 64);
                    }
                    add256(&acc, &term);
                    if (ge256(&acc, &Mm)) sub256(&acc, &Mm);
                }
                if (ge256(&acc, &half)) {
                    u256 diff = Mm;
                    sub256(&diff, &acc);
                    f = -orc_to_double_256(diff.d);
                } else {
                    f = orc_to_double_256(acc.d);
                }
            }
            int32_t e = shift_A[row] + shift_B[col];
            C[idx] = ldexp(f, -e);
        }
    }
}

/* ------------------------------------------------------------------------------------
 * Whole pipelines (src/hip/pipeline.hip:48-176 and :182-275).
 * ---------------------------------------------------------------------------------- */
void orc_s1_gemm(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K, int S) {
    int beta = orc_scheme1_beta(K);
    int32_t *sa = malloc(sizeof(int32_t) * M), *sb = malloc(sizeof(int32_t) * N);
    int8_t *as = malloc((size_t)S * M * K), *bs = malloc((size_t)S * K * N);
    int32_t *T = malloc(sizeof(int32_t) * (size_t)S * M * N);
    orc_row_shifts(A, M, K, 0, sa);
    orc_col_shifts(B, K, N, 0, sb);
    orc_slice_s1_rows(A, sa, M, K, S, beta, as);
    orc_slice_s1_cols(B, sb, K, N, S, beta, bs);
    orc_s1_group_products(as, bs, M, N, K, S, T);
    orc_recon_s1(T, sa, sb, M, N, S, beta, C);
    free(sa); free(sb); free(as); free(bs); free(T);
}

void orc_s2_gemm(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K, int S) {
    int kp = orc_scheme2_kparam(S, K);
    int32_t *sa = malloc(sizeof(int32_t) * M), *sb = malloc(sizeof(int32_t) * N);
    int8_t *as = malloc((size_t)S * M * K), *bs = malloc((size_t)S * K * N);
    int32_t *Ct = malloc(sizeof(int32_t) * (size_t)S * M * N);
    orc_row_shifts(A, M, K, kp, sa);
    orc_col_shifts(B, K, N, kp, sb);
    orc_slice_s2_rows(A, sa, M, K, S, as);
    orc_slice_s2_cols(B, sb, K, N, S, bs);
    orc_s2_products(as, bs, M, N, K, S, Ct);
    orc_recon_s2(Ct, sa, sb, M, N, S, C);
    free(sa); free(sb); free(as); free(bs); free(Ct);
}

/* ------------------------------------------------------------------------------------
 * End-to-end oracles.
 * ---------------------------------------------------------------------------------- */
/* utils/matrix_reference.hpp:22-51 */
void orc_gemm_fp128(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K) {
    double *Bt = malloc(sizeof(double) * (size_t)K * N);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < K; ++k)
        for (int64_t j = 0; j < N; ++j) Bt[j * K + k] = B[k * N + j];
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < M; ++i)
        for (int64_t j = 0; j < N; ++j) {
            __float128 sum = 0;
            for (int64_t k = 0; k < K; ++k) sum += (__float128)A[i * K + k] * (__float128)Bt[j * K + k];
            C[i * N + j] = (double)sum;
        }
    free(Bt);
}

/* Error-free transformations without FMA: Knuth TwoSum and Dekker/Veltkamp TwoProduct. */
static inline void two_sum(double a, double b, double *s, double *e) {
    double x = a + b;
    double z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

static inline void split(double a, double *hi, double *lo) {
    double c = 134217729.0 * a; /* 2^27 + 1 */
    double h = c - (c - a);
    *hi = h;
    *lo = a - h;
}

static inline void two_prod(double a, double b, double *p, double *e) {
    double ah, al, bh, bl;
    *p = a * b;
    split(a, &ah, &al);
    split(b, &bh, &bl);
    *e = al * bl - (((*p - ah * bh) - al * bh) - ah * bl);
}

void orc_gemm_dd(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K) {
    double *Bt = malloc(sizeof(double) * (size_t)K * N);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < K; ++k)
        for (int64_t j = 0; j < N; ++j) Bt[j * K + k] = B[k * N + j];
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < M; ++i)
        for (int64_t j = 0; j < N; ++j) {
            double s = 0.0, c = 0.0;
            for (int64_t k = 0; k < K; ++k) {
                double p, ep, es;
                two_prod(A[i * K + k], Bt[j * K + k], &p, &ep);
                two_sum(s, p, &s, &es);
                c += ep + es;
            }
            C[i * N + j] = s + c;
        }
    free(Bt);
}

void orc_gemm_abs(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < M; ++i) {
        for (int64_t j = 0; j < N; ++j) C[i * N + j] = 0.0;
        for (int64_t k = 0; k < K; ++k) {
            double a = fabs(A[i * K + k]);
            for (int64_t j = 0; j < N; ++j) C[i * N + j] += a * fabs(B[k * N + j]);
        }
    }
}

/* utils/matrix_compare.hpp:24-65 (sequential, same accumulation order). */
void orc_errors(const double *expected, const double *actual, int64_t n, double out[4]) {
    double max_abs = 0.0, max_rel = 0.0, ssd = 0.0, sse = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double e = expected[i], a = actual[i];
        double diff = a - e;
        double ad = fabs(diff);
        if (ad > max_abs) max_abs = ad;
        double rel = fabs(e) > 0.0 ? ad / fabs(e) : ad;
        if (rel > max_rel) max_rel = rel;
        ssd += diff * diff;
        sse += e * e;
    }
    double fe = sqrt(ssd), ef = sqrt(sse);
    out[0] = max_abs;
    out[1] = max_rel;
    out[2] = fe;
    out[3] = ef > 0.0 ? fe / ef : fe;
}
