// test_api.cpp — exercises the C++ drop-in API exactly as the reference's drivers use it
// (e.g. precision_benchmarks/02_ozaki_scheme2_fp128_benchmark.cpp:53-118): HipExecutor,
// allocate / copy, WorkspaceScheme1/2, ozaki_scheme{1,2}_gemm with and without OzaTimings, the
// executor's rocBLAS handle, and the reference's exception behaviour.  Results are written to
// <outdir> for the Python side to compare bit for bit with the oracle.
//
// usage: test_api <outdir> M N K S1 S2
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ozablas/core/executor.hpp"
#include "ozablas/core/workspace.hpp"
#include "ozablas/grouped.hpp"
#include "ozablas/ozablas.hpp"
#include "ozablas_hip.h"

static int failures = 0;
#define EXPECT(cond, what)                                          \
    do {                                                            \
        if (!(cond)) {                                              \
            std::fprintf(stderr, "FAILED: %s (line %d)\n", what, __LINE__); \
            ++failures;                                             \
        }                                                           \
    } while (0)

template <class Ex, class F>
static bool throws(F f) {
    try {
        f();
    } catch (const Ex&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static void dump(const std::string& path, const void* p, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s outdir M N K S1 S2\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const size_t M = std::strtoul(argv[2], nullptr, 10), N = std::strtoul(argv[3], nullptr, 10),
                 K = std::strtoul(argv[4], nullptr, 10);
    const int S1 = std::atoi(argv[5]), S2 = std::atoi(argv[6]);

    std::vector<double> hA(M * K), hB(K * N), hC(M * N);
    oza_generate_matrix(M, K, 1.0, 42, hA.data());
    oza_generate_matrix(K, N, 1.0, 43, hB.data());

    std::shared_ptr<ozablas::Executor> exec = std::make_shared<ozablas::HipExecutor>(0);
    double *dA, *dB, *dC;
    exec->allocate((void**)&dA, M * K * sizeof(double));
    exec->allocate((void**)&dB, K * N * sizeof(double));
    exec->allocate((void**)&dC, M * N * sizeof(double));
    exec->copy_from_host(dA, hA.data(), M * K * sizeof(double));
    exec->copy_from_host(dB, hB.data(), K * N * sizeof(double));

    // Scheme II, async then synchronize (reference usage without timings)
    {
        ozablas::WorkspaceScheme2 ws(exec, M, N, K, S2);
        EXPECT(ws.get_M() == M && ws.get_N() == N && ws.get_K() == K && ws.get_slices() == S2, "getters");
        EXPECT(ws.get_shift_A() && ws.get_shift_B() && ws.get_A_slices() && ws.get_B_slices() && ws.get_C_tc(),
               "workspace buffers");
        ozablas::ozaki_scheme2_gemm(ws, dA, dB, dC);
        exec->synchronize();
        exec->copy_to_host(hC.data(), dC, M * N * sizeof(double));
        dump(out + "/c_s2.bin", hC.data(), hC.size() * sizeof(double));
        // timings path blocks and fills every field
        ozablas::OzaTimings t;
        ozablas::ozaki_scheme2_gemm(ws, dA, dB, dC, &t);
        EXPECT(t.total_ms > 0.f && t.step3_ms > 0.f, "scheme2 timings");
        std::vector<double> hC2(M * N);
        exec->copy_to_host(hC2.data(), dC, M * N * sizeof(double));
        EXPECT(hC2 == hC, "repeat call identical");
        // planar slices view
        std::vector<int8_t> ap(S2 * M * K), bp(S2 * K * N);
        int8_t *dap, *dbp;
        exec->allocate((void**)&dap, ap.size());
        exec->allocate((void**)&dbp, bp.size());
        ws.unpack_slices(dap, dbp);
        exec->synchronize();
        exec->copy_to_host(ap.data(), dap, ap.size());
        exec->copy_to_host(bp.data(), dbp, bp.size());
        dump(out + "/a_s2.bin", ap.data(), ap.size());
        dump(out + "/b_s2.bin", bp.data(), bp.size());
        exec->free(dap);
        exec->free(dbp);
        // planar INT32 products C_s (recomputed from the slices) and planar CRT digits
        std::vector<int32_t> cp(S2 * M * N);
        std::vector<uint8_t> dg(S2 * M * N);
        int32_t* dcp;
        uint8_t* ddg;
        exec->allocate((void**)&dcp, cp.size() * sizeof(int32_t));
        exec->allocate((void**)&ddg, dg.size());
        EXPECT((void*)ws.get_C_digits() == (void*)ws.get_C_tc(), "digits view");
        ws.unpack_products(dcp);
        ws.unpack_digits(ddg);
        exec->synchronize();
        exec->copy_to_host(cp.data(), dcp, cp.size() * sizeof(int32_t));
        exec->copy_to_host(dg.data(), ddg, dg.size());
        dump(out + "/c_tc_s2.bin", cp.data(), cp.size() * sizeof(int32_t));
        dump(out + "/digits_s2.bin", dg.data(), dg.size());
        exec->free(dcp);
        exec->free(ddg);
    }
    // Scheme I
    {
        ozablas::WorkspaceScheme1 ws(exec, M, N, K, S1);
        ozablas::OzaTimings t;
        ozablas::ozaki_scheme1_gemm(ws, dA, dB, dC, &t);
        EXPECT(t.total_ms > 0.f, "scheme1 timings");
        exec->copy_to_host(hC.data(), dC, M * N * sizeof(double));
        dump(out + "/c_s1.bin", hC.data(), hC.size() * sizeof(double));
        std::vector<int32_t> tp(S1 * M * N);
        int32_t* dtp;
        exec->allocate((void**)&dtp, tp.size() * sizeof(int32_t));
        ws.unpack_products(dtp);
        exec->synchronize();
        exec->copy_to_host(tp.data(), dtp, tp.size() * sizeof(int32_t));
        dump(out + "/t_s1.bin", tp.data(), tp.size() * sizeof(int32_t));
        exec->free(dtp);
    }
    // Grouped Scheme II: the full problem and two sub-blocks (leading rows of A / columns of B)
    // in one batch, the last with its tile edge and a K split of 3 fixed; each result must equal a
    // single ozaki_scheme2_gemm call, graph or not.
    {
        const size_t m2 = M / 2 + 1, n2 = N / 3 + 1;
        std::vector<ozablas::GemmShape> shapes = {{M, N, K, S2}, {m2, N, K, S1 + 1}, {M, n2, K, 3, 128, 3}};
        ozablas::GroupScheme2 g(exec, shapes);
        EXPECT(g.size() == 3 && g.workspace_bytes() > 0 && g.tiles().size() == 3, "group info");
        const size_t kb = (K + 127) / 128, kbs = (kb + 2) / 3;  // K split of 3 over kb 128-byte blocks
        EXPECT(g.tiles()[2] == 128 && g.splits()[2] == (int)((kb + kbs - 1) / kbs), "group fixed tile and split");
        double *dB2, *dC0, *dC1, *dC2, *dR;
        exec->allocate((void**)&dB2, K * n2 * sizeof(double));
        std::vector<double> hB2(K * n2);
        for (size_t k = 0; k < K; ++k)
            for (size_t j = 0; j < n2; ++j) hB2[k * n2 + j] = hB[k * N + j];
        exec->copy_from_host(dB2, hB2.data(), hB2.size() * sizeof(double));
        exec->allocate((void**)&dC0, M * N * sizeof(double));
        exec->allocate((void**)&dC1, m2 * N * sizeof(double));
        exec->allocate((void**)&dC2, M * n2 * sizeof(double));
        exec->allocate((void**)&dR, M * N * sizeof(double));
        for (int use_graph = 0; use_graph < 2; ++use_graph) {
            for (int rep = 0; rep < 2; ++rep) {
                g.run({dA, dA, dA}, {dB, dB, dB2}, {dC0, dC1, dC2}, use_graph != 0);
                exec->synchronize();
                const double* outs[3] = {dC0, dC1, dC2};
                const double* Bs[3] = {dB, dB, dB2};
                for (int i = 0; i < 3; ++i) {
                    const auto& sh = shapes[i];
                    ozablas::WorkspaceScheme2 ws(exec, sh.M, sh.N, sh.K, sh.slices);
                    ozablas::ozaki_scheme2_gemm(ws, dA, Bs[i], dR);
                    exec->synchronize();
                    std::vector<double> x(sh.M * sh.N), y(sh.M * sh.N);
                    exec->copy_to_host(x.data(), outs[i], x.size() * sizeof(double));
                    exec->copy_to_host(y.data(), dR, y.size() * sizeof(double));
                    EXPECT(x == y, use_graph ? "grouped (graph) equals single" : "grouped equals single");
                }
            }
        }
        ozablas::OzaTimings t;
        g.run({dA, dA, dA}, {dB, dB, dB2}, {dC0, dC1, dC2}, true, &t);
        EXPECT(t.total_ms > 0.f, "grouped timings");
        EXPECT(throws<std::invalid_argument>([&] { g.run({dA}, {dB}, {dC0}); }), "grouped pointer count");
        EXPECT(throws<std::invalid_argument>([&] { ozablas::GroupScheme2 b(exec, {{8, 8, 8, 0}}); }), "group S=0");
        exec->free(dB2);
        exec->free(dC0);
        exec->free(dC1);
        exec->free(dC2);
        exec->free(dR);
    }

    // native DGEMM through the executor's BLAS handle, as the reference drivers do
    EXPECT(exec->get_blas_handle() != nullptr, "blas handle");

    // error behaviour (src/core/workspace.cpp:24-25, 68-71; src/ozablas.cpp:42-45)
    EXPECT(throws<std::invalid_argument>([&] { ozablas::WorkspaceScheme2 w(exec, M, N, K, 0); }), "S=0");
    EXPECT(throws<std::invalid_argument>([&] { ozablas::WorkspaceScheme2 w(exec, M, N, K, 21); }), "S=21");
    EXPECT(throws<std::invalid_argument>([&] { ozablas::WorkspaceScheme1 w(exec, M, N, K, 0); }), "S1=0");
    EXPECT(throws<std::invalid_argument>([&] { ozablas::WorkspaceScheme2 w(nullptr, M, N, K, 4); }), "null exec");
    EXPECT(throws<std::runtime_error>([&] { ozablas::CudaExecutor c(0); }), "cuda executor");
    {
        auto cpu = std::make_shared<ozablas::CpuExecutor>();
        ozablas::WorkspaceScheme2 w(cpu, 8, 8, 8, 2);
        EXPECT(throws<std::runtime_error>([&] { ozablas::ozaki_scheme2_gemm(w, dA, dB, dC); }), "cpu executor");
    }
    EXPECT(throws<std::bad_alloc>([&] {
               void* p;
               exec->allocate(&p, (size_t)1 << 60);
           }),
           "huge allocation");

    exec->free(dA);
    exec->free(dB);
    exec->free(dC);
    std::printf("test_api: %s\n", failures ? "FAILED" : "OK");
    return failures ? 1 : 0;
}
