// ref_driver.cpp -- drives the reference's own CPU solvers (compiled from the
// sources under /root/reference by oracle/Makefile.ref) to write golden
// vectors for the oracle. TEST INFRASTRUCTURE ONLY; it needs the reference
// tree to build, is not used where the GPU tests run, and no product code
// calls it.
//
// Three modes. `time` (run under `mpiexec -n P`) times the reference's CPU
// backend on a generated batch (run_time below); the other two read one
// input file and write the solved D:
//
//   backend : run under `mpiexec -n P`. Rank r takes rows
//             para_range(1, n_row, P, r) of every system (the z-slab split
//             of apps/heat_cpu), builds the reference CPU boundary object
//             TdmaBackend (apps/heat_cpu/tdma_backend.cpp:16-31), calls
//             set_rho (:33-44) when asked, then solve / solve_cyclic
//             (:51-69), and rank 0 gathers D.
//   steps   : one process, P partitions. Calls the reference's Filtered
//             step functions (libs/filtered_tdma/cpu/filtered_tdma_kernels.hpp)
//             in the order of FilteredTDMA::solve_filtered_v1/_v2
//             (filtered_tdma.cpp:170-258) with a caller-chosen J per
//             partition, so that the J-dependent sweeps can be pinned
//             independently of the J estimator.
//
// Input file (little endian):
//   char[4] "FTDI"; int32 n_sys, n_row, P, kind, cyclic, do_rho;
//   float64 eps_constant; int32 J[P]; float64 A,B,C,D [n_row*n_sys] each.
//   kind: 0 filtered, 1 filtered_v2, 2 pascal (backend mode);
//         1 or 2 selects v1 / v2 in steps mode (kind 0 -> v1).
// Output file: float64 D [n_row*n_sys].
#include <mpi.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tdma_backend.hpp"
#include "filtered_tdma_kernels.hpp"
#include "tdma_local.hpp"  // libs/filtered_tdma/cpu/tdma_local.hpp (ftdma_many)

namespace {

struct Input {
    int n_sys = 0, n_row = 0, P = 1, kind = 0, cyclic = 0, do_rho = 0;
    double eps_constant = 0.0;
    std::vector<int> J;
    std::vector<double> A, B, C, D;
};

bool read_input(const char* path, Input& in) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[4];
    int32_t hdr[6];
    bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "FTDI", 4) == 0 &&
              std::fread(hdr, sizeof(int32_t), 6, f) == 6;
    if (ok) {
        in.n_sys = hdr[0]; in.n_row = hdr[1]; in.P = hdr[2];
        in.kind = hdr[3]; in.cyclic = hdr[4]; in.do_rho = hdr[5];
        ok = std::fread(&in.eps_constant, sizeof(double), 1, f) == 1;
    }
    if (ok) {
        in.J.resize(in.P);
        ok = std::fread(in.J.data(), sizeof(int32_t), in.P, f) == (size_t)in.P;
    }
    const size_t n = (size_t)in.n_sys * in.n_row;
    for (std::vector<double>* v : {&in.A, &in.B, &in.C, &in.D}) {
        if (!ok) break;
        v->resize(n);
        ok = std::fread(v->data(), sizeof(double), n, f) == n;
    }
    std::fclose(f);
    return ok;
}

bool write_output(const char* path, const std::vector<double>& D) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    bool ok = std::fwrite(D.data(), sizeof(double), D.size(), f) == D.size();
    std::fclose(f);
    return ok;
}

// libs/pascal_tdma/cpu/para_range.cpp:4-9, 1-based rows.
void rows_of(int n_row, int P, int r, int& s0, int& m) {
    int len = n_row, base = len / P, rem = len % P;
    int ista = 1 + r * base + (r < rem ? r : rem);
    int iend = ista + base - 1 + (r < rem ? 1 : 0);
    s0 = ista - 1;
    m = iend - ista + 1;
}

int run_backend(const Input& in, const char* out_path) {
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (P != in.P) {
        if (rank == 0) std::fprintf(stderr, "ref_driver: file says P=%d, mpiexec gave %d\n", in.P, P);
        return 2;
    }
    int s0, m;
    rows_of(in.n_row, P, rank, s0, m);
    const size_t ns = (size_t)in.n_sys, nloc = (size_t)m * ns;
    std::vector<double> A(in.A.begin() + s0 * ns, in.A.begin() + s0 * ns + nloc);
    std::vector<double> B(in.B.begin() + s0 * ns, in.B.begin() + s0 * ns + nloc);
    std::vector<double> C(in.C.begin() + s0 * ns, in.C.begin() + s0 * ns + nloc);
    std::vector<double> D(in.D.begin() + s0 * ns, in.D.begin() + s0 * ns + nloc);

    int left = rank - 1, right = rank + 1;
    if (in.cyclic) {
        left = (rank + P - 1) % P;
        right = (rank + 1) % P;
    } else {
        if (left < 0) left = MPI_PROC_NULL;
        if (right >= P) right = MPI_PROC_NULL;
    }
    static const char* names[3] = {"filtered", "filtered_v2", "pascal"};
    TdmaBackend::Kind kind = TdmaBackend::parse(names[in.kind]);
    {
        TdmaBackend be(kind, in.n_sys, m, rank, P, MPI_COMM_WORLD, left, right, in.eps_constant);
        if (in.do_rho) be.set_rho(A.data(), B.data(), C.data());
        if (in.cyclic)
            be.solve_cyclic(A.data(), B.data(), C.data(), D.data());
        else
            be.solve(A.data(), B.data(), C.data(), D.data());
    }
    std::vector<int> counts(P), displs(P);
    for (int r = 0; r < P; ++r) {
        int sr, mr;
        rows_of(in.n_row, P, r, sr, mr);
        counts[r] = mr * in.n_sys;
        displs[r] = sr * in.n_sys;
    }
    std::vector<double> out(rank == 0 ? (size_t)in.n_row * ns : 0);
    MPI_Gatherv(D.data(), (int)nloc, MPI_DOUBLE, out.data(), counts.data(), displs.data(),
                MPI_DOUBLE, 0, MPI_COMM_WORLD);
    if (rank == 0 && !write_output(out_path, out)) return 3;
    return 0;
}

// Filtered solve with forced J, partitions emulated in one process, using the
// reference's step functions in the order of filtered_tdma.cpp:170-258.
int run_steps(Input& in, const char* out_path) {
    using namespace ftdma_kernel;
    const int P = in.P, ns = in.n_sys;
    const int version = in.kind == 1 ? 2 : 1;
    std::vector<int> s0(P), m(P);
    for (int p = 0; p < P; ++p) rows_of(in.n_row, P, p, s0[p], m[p]);
    auto at = [&](std::vector<double>& v, int p) { return v.data() + (size_t)s0[p] * ns; };
    if (P == 1) {
        ftdma_many(in.A.data(), in.B.data(), in.C.data(), in.D.data(), ns, in.n_row);
        return write_output(out_path, in.D) ? 0 : 3;
    }
    for (int p = 0; p < P; ++p) {
        double *A = at(in.A, p), *B = at(in.B, p), *C = at(in.C, p), *D = at(in.D, p);
        if (version == 1) {
            fwd_elim_v1(A, B, C, D, ns, m[p]);
            bwd_subst_v1(A, C, D, ns, m[p]);
        } else {
            const int J = in.J[p], lo = (m[p] - 2) - J;
            fwd_elim_v2(A, B, C, D, ns, m[p], J);
            bwd_subst_v2(A, C, D, ns, m[p], J, lo);
        }
        pack_reduced(A, C, D, ns);
    }
    // neighbour exchange (filtered_tdma.cpp:151-163): copies of the rows sent
    std::vector<double> cl(P * (size_t)ns), dl(P * (size_t)ns), ar(P * (size_t)ns), dr(P * (size_t)ns);
    for (int p = 0; p < P; ++p) {
        const double* CN = at(in.C, p) + (size_t)(m[p] - 1) * ns;
        const double* DN = at(in.D, p) + (size_t)(m[p] - 1) * ns;
        std::memcpy(cl.data() + (size_t)p * ns, CN, ns * sizeof(double));
        std::memcpy(dl.data() + (size_t)p * ns, DN, ns * sizeof(double));
        std::memcpy(ar.data() + (size_t)p * ns, at(in.A, p), ns * sizeof(double));
        std::memcpy(dr.data() + (size_t)p * ns, at(in.D, p), ns * sizeof(double));
    }
    for (int p = 0; p < P; ++p) {
        const bool has_left = in.cyclic || p > 0, has_right = in.cyclic || p < P - 1;
        const int pl = (p + P - 1) % P, pr = (p + 1) % P;
        solve_interfaces(at(in.A, p), at(in.C, p), at(in.D, p),
                         cl.data() + (size_t)pl * ns, dl.data() + (size_t)pl * ns,
                         ar.data() + (size_t)pr * ns, dr.data() + (size_t)pr * ns,
                         has_left, has_right, ns, m[p]);
    }
    for (int p = 0; p < P; ++p) {
        const int J = in.J[p];
        int left_end, right_beg;
        if (version == 1) {
            left_end = std::min(J, m[p] - 2);
            right_beg = std::max(1, (m[p] - 1) - J);
        } else {
            left_end = J;
            right_beg = std::max(1, (m[p] - 2) - J + 1);
        }
        final_solve(at(in.A, p), at(in.C, p), at(in.D, p), ns, m[p], left_end, right_beg);
    }
    return write_output(out_path, in.D) ? 0 : 3;
}

// splitmix64 -> uniform [0, 1) with 53 bits (the distribution of the GPU
// workloads; the timing does not depend on the exact values)
double uniform01(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Times the reference's CPU TdmaBackend (apps/heat_cpu/tdma_backend.cpp) on
// the batched workload: rank r owns rows para_range(1, n_row, P, r) of every
// system (the z-slab split), a = c = -r_s, b = 1 + 2 r_s with r_s ~ U[r_lo,
// r_hi] per system, D ~ U[-1, 1]. Each rep restores A, C, D (the solve
// clobbers them), then times set_rho + solve between barriers. Rank 0 prints
// one JSON line: the median and the minimum over reps, rows/s from the median.
int run_time(int argc, char** argv) {
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (argc != 9) {
        if (rank == 0)
            std::fprintf(stderr, "usage: ref_driver time kind n_sys n_row reps r_lo r_hi eps_\n");
        return 1;
    }
    const std::string kname = argv[2];
    const int n_sys = std::atoi(argv[3]), n_row = std::atoi(argv[4]), reps = std::atoi(argv[5]);
    const double r_lo = std::atof(argv[6]), r_hi = std::atof(argv[7]), eps = std::atof(argv[8]);
    int s0, m;
    rows_of(n_row, P, rank, s0, m);
    const size_t ns = (size_t)n_sys, nloc = (size_t)m * ns;
    std::vector<double> A0(nloc), B(nloc), C0(nloc), D0(nloc);
    for (size_t i = 0; i < ns; ++i) {
        const double r = r_lo + (r_hi - r_lo) * uniform01(2 * i);
        for (int j = 0; j < m; ++j) {
            const size_t k = (size_t)j * ns + i;
            A0[k] = -r;
            C0[k] = -r;
            B[k] = 1.0 + 2.0 * r;
            D0[k] = 2.0 * uniform01(((uint64_t)(s0 + j) << 32) ^ (2 * i + 1)) - 1.0;
        }
    }
    std::vector<double> A(nloc), C(nloc), D(nloc);
    int left = rank - 1 < 0 ? MPI_PROC_NULL : rank - 1;
    int right = rank + 1 >= P ? MPI_PROC_NULL : rank + 1;
    // eps_ = eps_constant / (n_row_local * nprocs)^2 (filtered_tdma.cpp:50-51)
    const double eps_constant = eps * (double)m * P * (double)m * P;
    TdmaBackend be(TdmaBackend::parse(kname), n_sys, m, rank, P, MPI_COMM_WORLD, left, right,
                   eps_constant);
    std::vector<double> t(reps);
    for (int k = 0; k < reps; ++k) {
        std::copy(A0.begin(), A0.end(), A.begin());
        std::copy(C0.begin(), C0.end(), C.begin());
        std::copy(D0.begin(), D0.end(), D.begin());
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        be.set_rho(A.data(), B.data(), C.data());
        be.solve(A.data(), B.data(), C.data(), D.data());
        MPI_Barrier(MPI_COMM_WORLD);
        t[k] = MPI_Wtime() - t0;
    }
    if (rank == 0) {
        std::vector<double> s(reps);
        std::copy(t.begin(), t.end(), s.begin());
        std::sort(s.begin(), s.end());
        const double med = s[reps / 2];
        std::printf("{\"kind\": \"%s\", \"ranks\": %d, \"n_sys\": %d, \"n_row\": %d, "
                    "\"reps\": %d, \"median_s\": %.6f, \"min_s\": %.6f, \"rows_per_s\": %.6e}\n",
                    kname.c_str(), P, n_sys, n_row, reps, med, s[0],
                    (double)n_sys * n_row / med);
        std::fflush(stdout);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "time") {
        MPI_Init(&argc, &argv);
        const int rc = run_time(argc, argv);
        MPI_Finalize();
        return rc;
    }
    if (argc != 4) {
        std::fprintf(stderr, "usage: ref_driver backend|steps in.bin out.bin\n"
                             "       ref_driver time kind n_sys n_row reps r_lo r_hi eps_\n");
        return 1;
    }
    const std::string mode = argv[1];
    Input in;
    if (!read_input(argv[2], in)) {
        std::fprintf(stderr, "ref_driver: cannot read %s\n", argv[2]);
        return 1;
    }
    if (mode == "steps") return run_steps(in, argv[3]);
    MPI_Init(&argc, &argv);
    int rc = run_backend(in, argv[3]);
    MPI_Finalize();
    return rc;
}
