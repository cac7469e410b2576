// crt_host.cpp — CRT constants and host-side pipeline parameters.
//
// The reference keeps the derived CRT tables as literals (src/common/crt_tables.hpp:19-112);
// here only the moduli list is data and everything else is computed once from closed forms:
// M = prod m_s, floor(M/2), M/m_s, (M/m_s)^-1 mod m_s and the CRT basis Q_s = inv_s * M/m_s mod M
// (Q_s = 1 mod m_s, 0 mod m_t for t != s), so that x = sum_s (x mod m_s) Q_s (mod M).
#include <cmath>
#include <cstring>
#include <mutex>

#include "oza_internal.hpp"

namespace oza {

namespace {

using u128 = unsigned __int128;

struct Big {  // 256-bit little-endian
    uint64_t d[4] = {0, 0, 0, 0};
};

void mul_small(Big& a, uint64_t m) {
    u128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        u128 p = (u128)a.d[i] * m + carry;
        a.d[i] = (uint64_t)p;
        carry = p >> 64;
    }
}

uint64_t div_small(const Big& a, uint64_t m, Big& q) {
    u128 rem = 0;
    for (int i = 3; i >= 0; --i) {
        u128 cur = (rem << 64) | a.d[i];
        q.d[i] = (uint64_t)(cur / m);
        rem = cur % m;
    }
    return (uint64_t)rem;
}

void to16(const Big& a, uint32_t out[kLimbs]) {
    for (int j = 0; j < kLimbs; ++j) out[j] = (uint32_t)((a.d[j / 4] >> (16 * (j % 4))) & 0xffffu);
}

int bitlen(const Big& a) {
    for (int i = 3; i >= 0; --i)
        if (a.d[i]) return 64 * i + (64 - __builtin_clzll(a.d[i]));
    return 0;
}

CrtTables build_tables() {
    CrtTables t;
    std::memset(&t, 0, sizeof(t));
    for (int S = 1; S <= kMaxSlices; ++S) {
        CrtSet& cs = t.set[S - 1];
        cs.S = S;
        Big M;
        M.d[0] = 1;
        for (int s = 0; s < S; ++s) mul_small(M, kModuli[s]);
        Big half;
        for (int i = 0; i < 4; ++i) half.d[i] = (M.d[i] >> 1) | (i < 3 ? (M.d[i + 1] << 63) : 0);
        to16(M, cs.M16);
        // X = sum_s r_s Q_s with r_s < 256: X/M < 20 * 256.  X mod M in {floor(M/2), floor(M/2)+1}
        // puts X/M within 1/(2M) of a half integer; the device estimates X/M from X's top limbs
        // (wl[l] = 2^(16 l)/M) to within 1e-10, so anything closer to 1/2 than this threshold is
        // re-checked exactly.
        double Md = 0.0;
        for (int i = 3; i >= 0; --i) Md = Md * 18446744073709551616.0 + (double)M.d[i];
        cs.qthr = 0.5 - 2.0 / Md - 1e-9;
        for (int l = 0; l < kLimbs; ++l) cs.wl[l] = std::ldexp(1.0, 16 * l) / Md;
        to16(half, cs.Mhalf16);
        std::memcpy(cs.M64, M.d, 32);
        std::memcpy(cs.Mhalf64, half.d, 32);
        // sum_s r_s * Q_s < S * 256 * M: limbs needed for that bound
        cs.nlimbs = (bitlen(M) + 8 + 5 + 15) / 16 + 1;
        if (cs.nlimbs > kLimbs) cs.nlimbs = kLimbs;
        for (int s = 0; s < S; ++s) {
            const uint64_t m = kModuli[s];
            Big q;
            div_small(M, m, q);
            Big qq;
            uint64_t r = div_small(q, m, qq);
            uint64_t inv = 0;
            for (uint64_t c = 1; c < m; ++c)
                if ((r * c) % m == 1) {
                    inv = c;
                    break;
                }
            // Q_s = inv * (M/m_s) mod M = inv * (M/m_s) - k M with k = floor(inv * (M/m_s) / M)
            //     = floor(inv / m_s) = 0 since inv < m_s: Q_s = inv * (M/m_s) exactly.
            Big Q = q;
            mul_small(Q, inv);
            to16(Q, cs.Q16[s]);
            cs.m[s] = (uint32_t)m;
            cs.inv[s] = (uint32_t)inv;
            cs.rcp[s] = 1.0 / (double)m;
        }
        for (int s = 0; s < S; ++s)
            for (int l = 0; l < 2 * kLimbs; ++l)
                cs.Q8x4[s / 4][l] |= ((cs.Q16[s][l / 2] >> (8 * (l % 2))) & 0xffu) << (8 * (s % 4));
    }
    return t;
}

}  // namespace

const CrtTables& host_crt_tables() {
    static const CrtTables tables = build_tables();
    return tables;
}

// Reference: src/pipeline/step1_statistics.hpp:12-15.
int scheme1_beta(size_t K) {
    int beta = (int)std::floor((31.0 - std::log2((double)K)) / 2.0);
    if (beta < 1) beta = 1;
    if (beta > 7) beta = 7;
    return beta;
}

// Reference: src/pipeline/step1_statistics.hpp:17-20 with M_double = to_double_256(M)
// (src/hip/pipeline.hip:205-206, src/common/crt_math.hpp:73-80).
int scheme2_kparam(int S, size_t K) {
    const CrtSet& cs = host_crt_tables().set[S - 1];
    double M_double = 0.0;
    M_double += (double)cs.M64[0];
    M_double += std::ldexp((double)cs.M64[1], 64);
    M_double += std::ldexp((double)cs.M64[2], 128);
    M_double += std::ldexp((double)cs.M64[3], 192);
    double frac = (M_double / 2.0 - 1.0) / (double)K;
    return (int)std::floor(0.5 * std::log2(frac));
}

// Smallest S with K_param(S, K) >= ceil(digits * log2(10)); 0 if no S <= 20 reaches it.
// Scheme II truncates every entry of row i of A to an integer multiple of 2^(E_i - k) (k = K_param)
// and likewise for columns of B, so |C_ij - C'_ij| <= K * 2^(E_i + F_j - k): `digits` decimal digits
// relative to the row/column scale K * 2^(E_i + F_j), which bounds sum_k |a_ik| |b_kj|.
int scheme2_slices_for_digits(size_t K, double digits) {
    const int bits = (int)std::ceil(digits * 3.321928094887362 - 1e-12);
    for (int S = 1; S <= kMaxSlices; ++S)
        if (scheme2_kparam(S, K) >= bits) return S;
    return 0;
}

// Componentwise digits lost against the row/column scale for inputs whose magnitudes spread like
// exp(spread * Z), Z standard normal (the reference generator's phi, utils/matrix_gen.hpp:15-34).
// With k = K_param, the worst entry of |C - C'| / (|A||B|) sits near 2^-k * 10^s(spread): rows and
// columns whose largest entry stands far above their typical one carry a truncation error of the
// largest entry's scale into products of typical size.  Table: the worst shortfall measured against
// the FP128 product over K = 64..4096, M = N = 256..1024, three seeds per point (0, 0.43, 1.41,
// 2.65, 3.93, 6.31, 8.42 digits at spread 0, 0.5, 1, 1.5, 2, 3, 4; tests/golden/scheme2_shortfall.py
// prints them), plus 0.8-1.05 digits for the deeper tails of outputs up to 8192 x 8192 (the worst
// entry grew by 0.5-0.9 digits from 64^2 to 1024^2 entries); linear in between, and 2.5 digits per
// unit of spread beyond 4.
double scheme2_spread_shortfall(double spread) {
    static const double x[] = {0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0};
    static const double y[] = {0.8, 1.3, 2.2, 3.5, 4.8, 7.2, 9.3};
    constexpr int n = 7;
    if (spread <= x[0]) return y[0];
    if (spread >= x[n - 1]) return y[n - 1] + 2.5 * (spread - x[n - 1]);
    int i = 0;
    while (spread > x[i + 1]) ++i;
    return y[i] + (y[i + 1] - y[i]) * (spread - x[i]) / (x[i + 1] - x[i]);
}

// Smallest S whose K_param covers `digits` componentwise digits (relative to |A||B|) plus the
// shortfall of inputs of the given spread; 0 if no S <= 20 reaches it.
int scheme2_slices_for_accuracy(size_t K, double digits, double spread) {
    const double need = digits + scheme2_spread_shortfall(spread);
    for (int S = 1; S <= kMaxSlices; ++S)
        if (scheme2_kparam(S, K) * 0.30102999566398120 >= need - 1e-12) return S;
    return 0;
}

Dims make_dims(size_t M, size_t N, size_t K, int S) {
    Dims d;
    d.M = (int64_t)M;
    d.N = (int64_t)N;
    d.K = (int64_t)K;
    d.S = S;
    d.Mp = (int64_t)round_up(M, kPadMN);
    d.Np = (int64_t)round_up(N, kPadMN);
    d.Kp = (int64_t)round_up(K, kPadK);
    d.Nd = (int64_t)round_up(N, kPadDigits);
    return d;
}

}  // namespace oza
