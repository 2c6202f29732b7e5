// backend_check.cpp -- exercises the header-only C++ class TdmaBackendGPU
// (include/tdma_backend_gpu.hpp) and the C ABI's hipGraph capture of a
// handle's first solve, as a program of its own linked against libftdma.so.
// Run by tests/test_gpu_cpp_backend.py, which writes the cases and checks
// the solutions against the oracle.
//
//   backend_check <cases.bin> <results.bin>
//
// cases.bin: int32 n_cases, then per case
//   int32 kind, n_sys, n_row, partitions, cyclic, options (bit 0: the guard,
//   bit 1: learned J); float64 eps_constant;
//   float64 A, B, C, D [n_row * n_sys] each (row-major, systems fastest).
// results.bin: per case
//   int32 J[partitions]; float64 gpu_ms (profiled solve), gpu_ms_off (after
//   set_profiling(false)), comm_ms; int32 parse_ok, graph_equal (3 replays
//   of the captured first solve of a fresh handle, each bit-identical to the
//   class's solve); float64 x [n_row * n_sys] (the class's solve); with
//   options, int32 records [partitions * n_sys] (last_J_systems).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "tdma_backend_gpu.hpp"

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void ftd_check(ftd_status s, const char* what) {
    if (s != FTD_OK) throw std::runtime_error(std::string(what) + ": " + ftd_last_error_string());
}

struct Dev {
    double* p = nullptr;
    explicit Dev(size_t n) { hip_check(hipMalloc(&p, n * sizeof(double)), "hipMalloc"); }
    ~Dev() { (void)hipFree(p); }
    void put(const std::vector<double>& h) {
        hip_check(hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "H2D");
    }
    void get(std::vector<double>& h) const {
        hip_check(hipMemcpy(h.data(), p, h.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    }
};

bool parse_table_ok() {
    using K = TdmaBackendGPU::Kind;
    return TdmaBackendGPU::parse("pascal") == K::PASCAL &&
           TdmaBackendGPU::parse("PaScaL_TDMA") == K::PASCAL &&
           TdmaBackendGPU::parse("filtered_v2") == K::FILTERED_V2 &&
           TdmaBackendGPU::parse("Filtered_V2") == K::FILTERED_V2 &&
           TdmaBackendGPU::parse("filtered") == K::FILTERED &&
           TdmaBackendGPU::parse("filtred") == K::FILTERED;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s cases.bin results.bin\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) {
        std::fprintf(stderr, "cannot open the case or result file\n");
        return 2;
    }
    try {
        int32_t n_cases = 0;
        if (std::fread(&n_cases, 4, 1, in) != 1) throw std::runtime_error("short case file");
        const int parse_ok = parse_table_ok() ? 1 : 0;
        for (int c = 0; c < n_cases; ++c) {
            int32_t hdr[6];
            double eps_constant = 0.0;
            if (std::fread(hdr, 4, 6, in) != 6 || std::fread(&eps_constant, 8, 1, in) != 1)
                throw std::runtime_error("short case header");
            const int kind = hdr[0], n_sys = hdr[1], n_row = hdr[2], P = hdr[3], cyclic = hdr[4];
            const int guard = hdr[5] & 1, learned = (hdr[5] >> 1) & 1;
            const size_t n = (size_t)n_sys * n_row;
            std::vector<double> h[4];
            for (auto& v : h) {
                v.resize(n);
                if (std::fread(v.data(), 8, n, in) != n) throw std::runtime_error("short arrays");
            }
            std::printf("[backend_check] case %d: kind %d n_sys %d n_row %d P %d cyclic %d "
                        "guard %d learned %d\n", c, kind, n_sys, n_row, P, cyclic, guard, learned);
            std::fflush(stdout);
            Dev A(n), B(n), C(n), D(n);
            auto load = [&]() {
                A.put(h[0]);
                B.put(h[1]);
                C.put(h[2]);
                D.put(h[3]);
            };

            // the reference's calling protocol on the C++ class
            // (apps/heat_gpu/solve_theta.cu:522-552): profiled by default
            std::vector<double> x(n), x_off(n);
            std::vector<int> J, rec;
            double gpu_ms = 0.0, gpu_ms_off = -1.0, comm_ms = -1.0;
            {
                TdmaBackendGPU be(static_cast<TdmaBackendGPU::Kind>(kind), n_sys, n_row, P,
                                  eps_constant);
                load();
                if (guard) be.set_guard(true);
                if (learned) be.set_j_mode(1);
                be.set_eps_constant(eps_constant);
                be.set_rho_device(A.p, B.p, C.p);
                if (cyclic)
                    be.solve_cyclic(A.p, B.p, C.p, D.p);
                else
                    be.solve(A.p, B.p, C.p, D.p);
                gpu_ms = be.last_gpu_ms();
                comm_ms = be.last_comm_ms();
                J = be.last_J();
                if (guard || learned) rec = be.last_J_systems();
                D.get(x);
                be.set_profiling(false);
                load();
                if (cyclic)
                    be.solve_cyclic(A.p, B.p, C.p, D.p);
                else
                    be.solve(A.p, B.p, C.p, D.p);
                hip_check(hipDeviceSynchronize(), "sync");
                gpu_ms_off = be.last_gpu_ms();
                D.get(x_off);
            }
            if (std::memcmp(x.data(), x_off.data(), n * sizeof(double)) != 0)
                throw std::runtime_error("profiled and unprofiled solves differ");

            // a fresh handle's first solve captured into a hipGraph (global
            // capture mode: an allocation inside the solve would fail it),
            // then replayed three times on fresh inputs
            int graph_equal = 1;
            {
                hipStream_t st;
                hip_check(hipStreamCreate(&st), "hipStreamCreate");
                ftd_handle* hd = nullptr;
                ftd_check(ftd_create(&hd, kind, n_sys, n_row, P, eps_constant, st), "ftd_create");
                if (guard) ftd_check(ftd_set_guard(hd, 1), "ftd_set_guard");
                if (learned) ftd_check(ftd_set_j_mode(hd, 1), "ftd_set_j_mode");
                load();
                ftd_check(ftd_set_rho(hd, A.p, B.p, C.p), "ftd_set_rho");
                hip_check(hipStreamSynchronize(st), "sync");
                hipGraph_t g;
                hipGraphExec_t ge;
                hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal), "begin capture");
                const ftd_status s = cyclic ? ftd_solve_cyclic(hd, A.p, B.p, C.p, D.p)
                                            : ftd_solve(hd, A.p, B.p, C.p, D.p);
                const hipError_t ec = hipStreamEndCapture(st, &g);
                ftd_check(s, "captured solve");
                hip_check(ec, "end capture");
                hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "instantiate");
                std::vector<double> xg(n);
                for (int r = 0; r < 3; ++r) {
                    load();
                    hip_check(hipGraphLaunch(ge, st), "graph launch");
                    hip_check(hipStreamSynchronize(st), "sync");
                    D.get(xg);
                    if (std::memcmp(xg.data(), x.data(), n * sizeof(double)) != 0) graph_equal = 0;
                }
                hip_check(hipGraphExecDestroy(ge), "exec destroy");
                hip_check(hipGraphDestroy(g), "graph destroy");
                ftd_check(ftd_destroy(hd), "ftd_destroy");
                hip_check(hipStreamDestroy(st), "hipStreamDestroy");
            }

            std::vector<int32_t> Jw(P, -1);
            for (int p = 0; p < P && p < (int)J.size(); ++p) Jw[p] = J[p];
            const double ms[3] = {gpu_ms, gpu_ms_off, comm_ms};
            const int32_t flags[2] = {parse_ok, graph_equal};
            std::fwrite(Jw.data(), 4, P, out);
            std::fwrite(ms, 8, 3, out);
            std::fwrite(flags, 4, 2, out);
            std::fwrite(x.data(), 8, n, out);
            if (guard || learned) {
                if (rec.size() != (size_t)P * n_sys) throw std::runtime_error("records size");
                std::fwrite(rec.data(), 4, rec.size(), out);
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "backend_check: %s\n", e.what());
        std::fclose(out);
        return 1;
    }
    std::fclose(in);
    std::fclose(out);
    std::printf("[backend_check] done\n");
    return 0;
}
