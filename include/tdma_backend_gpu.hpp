// tdma_backend_gpu.hpp -- header-only C++ mirror of the reference's
// TdmaBackendGPU (apps/heat_gpu/tdma_backend_gpu.hpp:21-70) over the C ABI
// of libftdma.so (ftdma.h). Same member names and meaning; the MPI
// constructor arguments (myrank, nprocs, comm, left_rank, right_rank) are
// replaced by the number of partitions each system's rows are split into on
// one GPU. Errors throw std::runtime_error (the reference exits).
//
// Like the reference, whose solve always runs the *_profile variants
// (tdma_backend_gpu.cu:35-46), every solve is timed by default: it waits for
// its HIP events, and last_gpu_ms() returns its device time.
// set_profiling(false) drops the events and the wait (last_gpu_ms() is then 0).
//
//   TdmaBackendGPU be(TdmaBackendGPU::parse("filtered_v2"), n_sys, n_row,
//                     /*partitions=*/8, /*eps_constant=*/dt);
//   be.set_eps_constant(dt);
//   be.set_rho_device(dA, dB, dC);
//   be.solve(dA, dB, dC, dD);
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "ftdma.h"

class TdmaBackendGPU {
public:
    enum class Kind { FILTERED = FTD_KIND_FILTERED, FILTERED_V2 = FTD_KIND_FILTERED_V2,
                      PASCAL = FTD_KIND_PASCAL };

    /// Case-insensitive: "pascal" | "pascal_tdma" -> PASCAL, "filtered_v2" ->
    /// FILTERED_V2, anything else -> FILTERED (tdma_backend_gpu.cu:6-13).
    static Kind parse(const std::string& s) {
        return static_cast<Kind>(ftd_kind_from_string(s.c_str()));
    }

    TdmaBackendGPU(Kind kind, int n_sys, int n_row, int partitions, double eps_constant,
                   void* stream = nullptr)
        : kind_(kind), partitions_(partitions), n_sys_(n_sys) {
        check(ftd_create(&h_, static_cast<int>(kind), n_sys, n_row, partitions, eps_constant,
                         stream));
        const ftd_status s = ftd_set_profiling(h_, 1);
        if (s != FTD_OK) {
            ftd_destroy(h_);
            check(s);
        }
    }
    ~TdmaBackendGPU() { ftd_destroy(h_); }
    TdmaBackendGPU(const TdmaBackendGPU&) = delete;
    TdmaBackendGPU& operator=(const TdmaBackendGPU&) = delete;

    void solve(double* d_A, double* d_B, double* d_C, double* d_D) {
        check(ftd_solve(h_, d_A, d_B, d_C, d_D));
    }
    void solve(double* d_A, double* d_B, double* d_C, double* d_D, int n_sys, int n_row) {
        check(ftd_solve_nr(h_, d_A, d_B, d_C, d_D, n_sys, n_row));
    }
    void solve_cyclic(double* d_A, double* d_B, double* d_C, double* d_D) {
        check(ftd_solve_cyclic(h_, d_A, d_B, d_C, d_D));
    }
    void set_rho_device(const double* d_A, const double* d_B, const double* d_C) {
        check(ftd_set_rho(h_, d_A, d_B, d_C));
    }
    void set_eps_constant(double eps_constant) { check(ftd_set_eps_constant(h_, eps_constant)); }

    /// Time every solve (default, as the reference) or not.
    void set_profiling(bool enable) { check(ftd_set_profiling(h_, enable ? 1 : 0)); }
    void set_stream(void* stream) { check(ftd_set_stream(h_, stream)); }

    Kind kind() const { return kind_; }
    double last_comm_ms() const { return ftd_last_comm_ms(h_); }
    double last_gpu_ms() const { return ftd_last_gpu_ms(h_); }

    /// FilteredTDMACUDA::last_J, one value per partition (blocking).
    std::vector<int> last_J() const {
        std::vector<int> J(partitions_);
        check(ftd_last_J(h_, J.data(), partitions_));
        return J;
    }

    // ---- extensions (no counterpart in the reference; ftdma.h) ----

    /// ftd_set_guard: the filtered kinds solve exactly the systems whose
    /// dropped terms could put x more than eps_ from the exact solve.
    void set_guard(bool enable) { check(ftd_set_guard(h_, enable ? 1 : 0)); }
    /// ftd_set_j_mode: 0 = the reference's cal_J, 1 = learned J per
    /// (partition, system), checked against eps_ (FILTERED kind).
    void set_j_mode(int mode) { check(ftd_set_j_mode(h_, mode)); }
    /// ftd_solve_complex: complex128 right-hand sides (interleaved re, im),
    /// one coefficient set for both parts.
    void solve_complex(double* d_A, double* d_B, double* d_C, double* d_Z) {
        check(ftd_solve_complex(h_, d_A, d_B, d_C, d_Z));
    }
    /// ftd_last_J_systems: J | fallback << 16 | guard << 17 per (partition,
    /// system) of the last learned or guarded solve, index p * n + i (n =
    /// n_sys, doubled for a complex solve). Blocking.
    std::vector<int> last_J_systems(bool complex_rhs = false) const {
        std::vector<int> out((size_t)partitions_ * n_sys_ * (complex_rhs ? 2 : 1));
        check(ftd_last_J_systems(h_, out.data(), (int64_t)out.size()));
        return out;
    }

    ftd_handle* handle() { return h_; }

private:
    static void check(ftd_status s) {
        if (s != FTD_OK) throw std::runtime_error(ftd_last_error_string());
    }
    Kind kind_;
    int partitions_;
    int n_sys_;
    ftd_handle* h_ = nullptr;
};
