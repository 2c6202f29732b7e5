// ftdma_api.cpp -- C ABI of the batched tridiagonal engine (include/ftdma.h).
//
// Host side of the drop-in for TdmaBackendGPU
// (apps/heat_gpu/tdma_backend_gpu.{hpp,cu}): it owns the device scratch of a
// handle, picks a kernel path per solve and orders every launch on the
// handle's stream. No MPI: the reference's z-slab ranks become partitions of
// each system on one GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ftdma.h"
#include "ftdma_kernels.h"

namespace {

thread_local std::string g_err;

ftd_status fail(ftd_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

ftd_status hip_fail(hipError_t e, const char* what) {
    return fail(FTD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define FTD_HIP(call)                                     \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

#define FTD_TRY(expr)                  \
    do {                               \
        ftd_status s_ = (expr);        \
        if (s_ != FTD_OK) return s_;   \
    } while (0)

void part_of(int n, int P, int p, int& s, int& m) {
    int base = n / P, rem = n % P;
    s = p * base + (p < rem ? p : rem);
    m = base + (p < rem ? 1 : 0);
}

int pow2_floor(int v) {
    int r = 1;
    while (r * 2 <= v) r *= 2;
    return r;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    size_t want = 0;  // largest size a sizing pass asked for
};

}  // namespace

struct ftd_handle {
    int kind = FTD_KIND_FILTERED;
    int n_sys = 0, n_row = 0, P = 1;
    double eps_value = 0.0;  // eps_constant, or eps_ itself when eps_direct
    int eps_direct = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int profiling = 0;
    double last_ms = 0.0;
    char path[96] = "none";
    int t_M = 0, t_T = 0, t_S = 0, t_Q = 0;  // forced tuning (0 = automatic)
    int force_general = 0;                   // use the general kernels only
    int no_stream = 0;                       // automatic choice without the streamed kernels
    int j_mode = 0;                          // 0: cal_J (reference), 1: learned J per system
    int guard = 0;                           // filtered kinds: exact solve where terms exceed eps_/4
    DevBuf jsys;                             // learned mode: J (| fallback << 16) per lane
    size_t jsys_count = 0;
    double* rho = nullptr;                   // A_rho[n_row+1] then C_rho[n_row+1]
    int* J = nullptr;                        // J per partition
    DevBuf partial, xbuf, R, E;
    // uneven batches: the plan of the last layout (host copy of the layout
    // it was made for, device lanes and long-system lists)
    std::vector<int64_t> b_off;
    std::vector<int> b_rows;
    std::vector<ftd::BatchLane> b_lanes;
    std::vector<int64_t> b_long_off;
    std::vector<int> b_long_n;
    int b_items = 0, b_chunked = 0, b_valid = 0;
    int64_t b_last = 0;  // last element of the planned layout
    int b_chunk = 0;     // rows per chunk asked for (0: the default)
    int b_plan_chunk = 0;  // rows per chunk of the planned layout
    DevBuf b_dlanes, b_dlong_off, b_dlong_n;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // sizing pass: the solve paths run without launching anything and only
    // record the scratch they would use (DevBuf::want)
    int sizing = 0;
};

namespace {

// Grows a buffer to `bytes` (contents are not kept). The solve paths call
// the handle form below, which only records the size during a sizing pass.
ftd_status ensure(DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return FTD_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree");
    }
    hipError_t e = hipMalloc(&b.p, bytes + 256);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(FTD_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return FTD_OK;
}

ftd_status ensure(ftd_handle* h, DevBuf& b, size_t bytes) {
    if (h->sizing) {
        b.want = std::max(b.want, bytes);
        return FTD_OK;
    }
    return ensure(b, bytes);
}

// a kernel launch of a solve path: skipped during a sizing pass
#define FTD_LAUNCH(call)               \
    do {                               \
        if (!h->sizing) FTD_HIP(call); \
    } while (0)

// ---------------- PaScaL / Thomas (exact) ----------------

struct FusedShape {
    int M = 0, T = 0, S = 0, Q = 0;
};

// Fused-kernel shape for an exact solve of n_row rows split into Q lanes per
// system (sub-partitions of <= M rows, >= 3 rows each), S systems per
// workgroup. With P > 1 partitions on the handle, Q = P whenever a compiled
// variant holds them: the result is then the reference's P-rank PaScaL
// statement for statement. Otherwise Q is chosen here (at least 2: one
// partition is the plain Thomas kernel).
bool shape_for(int M, int T, int Q, int n_row, int forced_S, FusedShape& out) {
    if (!ftd::pascal_fused_supported(M, T) || Q < 2 || Q > T) return false;
    int s_, m_min, m_max;
    part_of(n_row, Q, Q - 1, s_, m_min);
    part_of(n_row, Q, 0, s_, m_max);
    if (m_min < 3 || m_max > M) return false;
    const int S = forced_S ? forced_S : std::min(64, pow2_floor(T / Q));
    if (!is_pow2(S) || S > 64 || S * Q > T) return false;
    out.M = M;
    out.T = T;
    out.S = S;
    out.Q = Q;
    return true;
}

bool pick_pascal(const ftd_handle* h, int n_row, int cyclic, FusedShape& out, int csh = 0) {
    if (h->t_M) {
        int Q = h->t_Q ? h->t_Q : std::max(2, (n_row + h->t_M - 1) / h->t_M);
        return shape_for(h->t_M, h->t_T ? h->t_T : 256, Q, n_row, h->t_S, out);
    }
    // (M, T) in order of preference, fastest first at 262144 x 512
    static const int Ms[] = {64, 32, 32, 16, 8};
    static const int Ts[] = {256, 512, 256, 512, 512};
    const int P = h->P;
    if (P > 1) {
        int s_, m_max;
        part_of(n_row, P, 0, s_, m_max);
        for (int k = 0; k < 5; ++k) {
            if (Ms[k] < m_max) continue;
            FusedShape fs;
            if (shape_for(Ms[k], Ts[k], P, n_row, 0, fs) && fs.S >= 16) {
                // smallest M that holds the partitions
                out = fs;
                for (int k2 = 0; k2 < 5; ++k2)
                    if (Ms[k2] >= m_max && Ms[k2] < out.M && shape_for(Ms[k2], Ts[k2], P, n_row, 0, fs) &&
                        fs.S >= 16)
                        out = fs;
                return true;
            }
        }
    }
    (void)cyclic;
    if (csh || n_row > 8 * 64) {
        // complex right-hand sides, one caller partition: 32-row partitions
        // in 512-thread workgroups measured fastest (16384 x 129 and
        // 33024 x 256 channel modes: 9-17 % under 64-row ones). So do real
        // systems longer than eight 64-row partitions, where either shape
        // holds 16 systems per workgroup: 262144 systems of 600 / 768 / 1024
        // rows in 1.83 / 2.23 / 2.96 against 1.99 / 2.47 / 3.19 ms
        const int Q = std::max(2, (n_row + 31) / 32);
        FusedShape fs;
        if (shape_for(32, 512, Q, n_row, 0, fs) && fs.S >= 16) {
            out = fs;
            return true;
        }
    }
    for (int k = 0; k < 5; ++k) {
        const int Q = std::max(2, (n_row + Ms[k] - 1) / Ms[k]);
        FusedShape fs;
        if (shape_for(Ms[k], Ts[k], Q, n_row, 0, fs) && fs.S >= 16) {
            out = fs;
            return true;
        }
    }
    return false;
}

// csh = 1: D holds complex right-hand sides as interleaved (re, im) columns,
// n_sys counts those real columns and the coefficients have n_sys / 2
ftd_status solve_exact(ftd_handle* h, double* A, double* B, double* C, double* D, int n_sys,
                       int n_row, int cyclic, int csh = 0) {
    hipStream_t st = h->stream;
    if (cyclic && n_row < 3) return fail(FTD_ERR_UNSUPPORTED, "periodic solve needs n_row >= 3");
    FusedShape fs;
    if (!h->force_general && !h->t_M && h->P == 1 && n_row >= 3 && ftd::thomas_fused_rows(n_row)) {
        // one partition that fits a lane: plain (or Sherman-Morrison periodic)
        // Thomas in registers, the reference's single-rank path
        if (cyclic) {
            std::snprintf(h->path, sizeof h->path, "thomas_cyclic_fused M=%d",
                          ftd::thomas_fused_rows(n_row));
            FTD_LAUNCH(ftd::launch_thomas_cyclic_fused(A, B, C, D, n_sys, n_row, st));
        } else {
            std::snprintf(h->path, sizeof h->path, "thomas_fused M=%d", ftd::thomas_fused_rows(n_row));
            FTD_LAUNCH(ftd::launch_thomas_fused(A, B, C, D, n_sys, n_row, csh, st));
        }
        return FTD_OK;
    }
    if (!h->force_general && n_row >= 3 && pick_pascal(h, n_row, cyclic, fs, csh)) {
        // the streamed kernel runs the same partitioning (bit-identical
        // results) where a variant holds it and the arrays allow its DMA
        int s_, m_max;
        part_of(n_row, fs.Q, 0, s_, m_max);
        const int Ms = (h->t_M || h->no_stream || csh) ? 0 : ftd::stream_rows(m_max, fs.Q);
        if (Ms && ftd::stream_eligible(n_sys, A, B, C, D)) {
            std::snprintf(h->path, sizeof h->path, "pascal_stream M=%d Q=%d R=%d", Ms, fs.Q,
                          ftd::stream_ring_depth(Ms, fs.Q));
            FTD_LAUNCH(ftd::launch_pascal_stream(Ms, fs.Q, A, B, C, D, n_sys, n_row, cyclic, st));
            return FTD_OK;
        }
        std::snprintf(h->path, sizeof h->path, "pascal_fused M=%d T=%d S=%d Q=%d", fs.M, fs.T,
                      fs.S, fs.Q);
        FTD_LAUNCH(ftd::launch_pascal_fused(fs.M, fs.T, A, B, C, D, n_sys, n_row, fs.Q, fs.S, cyclic,
                                         csh, st));
        return FTD_OK;
    }
    if (h->t_M) return fail(FTD_ERR_UNSUPPORTED, "forced fused shape does not fit this system");
    if (csh) return fail(FTD_ERR_UNSUPPORTED, "complex right-hand sides need a fused kernel shape");
    int P = h->P;
    if (P > 1) {
        int s_, ml;
        part_of(n_row, P, P - 1, s_, ml);
        if (ml < 3) P = 1;
    }
    if (P == 1 && !cyclic) {
        std::snprintf(h->path, sizeof h->path, "thomas_general");
        FTD_LAUNCH(ftd::launch_gen_thomas(A, B, C, D, n_sys, n_row, st));
        return FTD_OK;
    }
    if (P == 1) {
        FTD_TRY(ensure(h, h->E, sizeof(double) * (size_t)n_row * n_sys));
        std::snprintf(h->path, sizeof h->path, "thomas_cyclic_general");
        FTD_LAUNCH(ftd::launch_gen_thomas_cyclic(A, B, C, D, (double*)h->E.p, n_sys, n_row, st));
        return FTD_OK;
    }
    FTD_TRY(ensure(h, h->R, sizeof(double) * (size_t)6 * P * n_sys));
    FTD_TRY(ensure(h, h->E, sizeof(double) * (size_t)2 * P * n_sys));
    std::snprintf(h->path, sizeof h->path, "pascal_general P=%d", P);
    FTD_LAUNCH(ftd::launch_gen_pascal(A, B, C, D, n_sys, n_row, P, cyclic, (double*)h->R.p,
                                   (double*)h->E.p, st));
    return FTD_OK;
}

// ---------------- Filtered ----------------

ftd_status solve_filtered(ftd_handle* h, double* A, double* B, double* C, double* D, int cyclic,
                          int csh = 0) {
    const int n_sys = h->n_sys << csh, n_row = h->n_row, P = h->P;
    const int version = (h->kind == FTD_KIND_FILTERED_V2) ? 2 : 1;
    hipStream_t st = h->stream;
    h->jsys_count = 0;  // J records belong to the last solve, if it kept them
    int s_, mmax, mmin;
    part_of(n_row, P, 0, s_, mmax);
    part_of(n_row, P, P - 1, s_, mmin);
    if (mmin < 3) return fail(FTD_ERR_UNSUPPORTED, "every partition needs at least 3 rows");
    // the guard: systems whose dropped terms exceed eps_/4 are solved exactly
    const int g = h->guard;
    const char* gs = g ? " guard" : "";

    // fused shape: smallest compiled M holding the largest partition
    FusedShape fs;
    if (!h->force_general) {
        static const int Ms[] = {8, 16, 32, 64};
        static const int Ts[] = {512, 512, 256, 256};
        for (int k = 0; k < 4; ++k) {
            if (h->t_M && Ms[k] != h->t_M) continue;
            if (Ms[k] < mmax) continue;
            const int T = (h->t_M && h->t_T) ? h->t_T : Ts[k];
            if (!ftd::filtered_fused_supported(Ms[k], T) || P > T) continue;
            const int S = h->t_S ? h->t_S : std::min(64, pow2_floor(T / P));
            if (!is_pow2(S) || S > 64 || S * P > T) continue;
            fs.M = Ms[k];
            fs.T = T;
            fs.S = S;
            break;
        }
    }
    if (h->t_M && !fs.M) return fail(FTD_ERR_UNSUPPORTED, "forced fused shape does not fit this system");

    int* Jp = h->J;
    const size_t xbytes = sizeof(double) * (size_t)2 * P * n_sys;
    const int Ms = (h->force_general || h->t_M || h->no_stream || csh ||
                    (g && !ftd::stream_guard_supported(P)))
                       ? 0
                       : ftd::stream_rows(mmax, P);
    const bool learned = h->j_mode == 1 && version == 1;
    // per (partition, system) records: learned J, and the guard's
    const size_t nj = (size_t)P * n_sys;
    const double ev = g ? h->eps_value : 0.0;
    const int ed = g ? h->eps_direct : 0;
    // learned J: the streamed kernel for 64-row partitions, not under the guard
    if (Ms && ftd::stream_eligible(n_sys, A, B, C, D) && (!learned || (Ms == 64 && !g))) {
        const int R = ftd::stream_ring_depth(Ms, P);
        if (learned || g) {
            FTD_TRY(ensure(h, h->jsys, sizeof(int) * nj));
            h->jsys_count = nj;
        }
        int* const Js = (learned || g) ? (int*)h->jsys.p : nullptr;
        if (learned) {
            // learned J: one streamed pass (the 64-row variants hold the
            // network's registers without spilling; shorter partitions and
            // guarded solves take the fused kernel below)
            std::snprintf(h->path, sizeof h->path, "filtered_learned_stream M=%d Q=%d R=%d%s", Ms,
                          P, R, gs);
            FTD_LAUNCH(ftd::launch_filtered_stream(Ms, P, 1, 3, A, B, C, D, n_sys, n_row, cyclic, Jp,
                                                   nullptr, nullptr, h->eps_value, h->eps_direct,
                                                   Js, g, st));
            return FTD_OK;
        }
        if (version == 2) {
            int nt = 0;
            FTD_TRY(ensure(h, h->partial, sizeof(double) * (size_t)((n_sys + 255) / 256) * (P << csh)));
            FTD_LAUNCH(ftd::launch_boundary_maxabs(D, n_sys, n_row, P, csh, (double*)h->partial.p, &nt, st));
            FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, nt, P, n_row, h->rho, h->eps_value,
                                             h->eps_direct, csh, Jp, st));
            std::snprintf(h->path, sizeof h->path, "filtered_v2_stream M=%d Q=%d R=%d%s", Ms, P, R,
                          gs);
            FTD_LAUNCH(ftd::launch_filtered_stream(Ms, P, 2, 0, A, B, C, D, n_sys, n_row, cyclic, Jp,
                                                   nullptr, nullptr, ev, ed, Js, g, st));
        } else {
            const int tiles = n_sys / 32;
            FTD_TRY(ensure(h, h->partial, sizeof(double) * (size_t)tiles * (P << csh)));
            FTD_TRY(ensure(h, h->xbuf, xbytes));
            std::snprintf(h->path, sizeof h->path, "filtered_v1_stream M=%d Q=%d R=%d%s", Ms, P, R,
                          gs);
            FTD_LAUNCH(ftd::launch_filtered_stream(Ms, P, 1, 1, A, B, C, D, n_sys, n_row, cyclic, Jp,
                                                   (double*)h->xbuf.p, (double*)h->partial.p, 0.0,
                                                   0, nullptr, 0, st));
            FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, tiles, P, n_row, h->rho,
                                             h->eps_value, h->eps_direct, csh, Jp, st));
            FTD_LAUNCH(ftd::launch_filtered_stream(Ms, P, 1, 2, A, B, C, D, n_sys, n_row, cyclic, Jp,
                                                   (double*)h->xbuf.p, nullptr, ev, ed, Js, g, st));
        }
        return FTD_OK;
    }
    if (fs.M && csh && fs.S < 2) fs.M = 0;  // the two parts of a system share a tile
    if ((learned || g) && fs.M) {
        FTD_TRY(ensure(h, h->jsys, sizeof(int) * nj));
        h->jsys_count = nj;
    }
    if (learned && fs.M) {
        // learned J: one pass (v1 sweeps, interface, J per system from the
        // MLP, checked against eps_, truncated correction)
        std::snprintf(h->path, sizeof h->path, "filtered_learned_fused M=%d T=%d S=%d%s", fs.M,
                      fs.T, fs.S, gs);
        FTD_LAUNCH(ftd::launch_filtered_fused(fs.M, fs.T, 1, 3, A, B, C, D, n_sys, n_row, P, fs.S,
                                              cyclic, csh, Jp, nullptr, nullptr, h->eps_value,
                                              h->eps_direct, (int*)h->jsys.p, g, st));
        return FTD_OK;
    }
    if (fs.M) {
        const int tiles = (n_sys + fs.S - 1) / fs.S;
        int* const Js = g ? (int*)h->jsys.p : nullptr;
        if (version == 2) {
            int nt = 0;
            FTD_TRY(ensure(h, h->partial, sizeof(double) * (size_t)((n_sys + 255) / 256) * (P << csh)));
            FTD_LAUNCH(ftd::launch_boundary_maxabs(D, n_sys, n_row, P, csh, (double*)h->partial.p, &nt, st));
            FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, nt, P, n_row, h->rho, h->eps_value,
                                             h->eps_direct, csh, Jp, st));
            std::snprintf(h->path, sizeof h->path, "filtered_v2_fused M=%d T=%d S=%d%s", fs.M, fs.T,
                          fs.S, gs);
            FTD_LAUNCH(ftd::launch_filtered_fused(fs.M, fs.T, 2, 0, A, B, C, D, n_sys, n_row, P, fs.S,
                                                  cyclic, csh, Jp, nullptr, nullptr, ev, ed, Js, g,
                                                  st));
        } else {
            FTD_TRY(ensure(h, h->partial, sizeof(double) * (size_t)tiles * (P << csh)));
            FTD_TRY(ensure(h, h->xbuf, xbytes));
            std::snprintf(h->path, sizeof h->path, "filtered_v1_fused M=%d T=%d S=%d%s", fs.M, fs.T,
                          fs.S, gs);
            FTD_LAUNCH(ftd::launch_filtered_fused(fs.M, fs.T, 1, 1, A, B, C, D, n_sys, n_row, P, fs.S,
                                                  cyclic, csh, Jp, (double*)h->xbuf.p,
                                                  (double*)h->partial.p, 0.0, 0, nullptr, 0, st));
            FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, tiles, P, n_row, h->rho,
                                             h->eps_value, h->eps_direct, csh, Jp, st));
            FTD_LAUNCH(ftd::launch_filtered_fused(fs.M, fs.T, 1, 2, A, B, C, D, n_sys, n_row, P, fs.S,
                                                  cyclic, csh, Jp, (double*)h->xbuf.p, nullptr, ev,
                                                  ed, Js, g, st));
        }
        return FTD_OK;
    }

    if (csh) return fail(FTD_ERR_UNSUPPORTED, "complex right-hand sides need a fused kernel shape");
    // general path: sweeps in place, interface, J, final. Learned J and the
    // guard take the v1 sweeps and a per-system J (check kernel).
    const int tiles = (n_sys + 255) / 256;
    FTD_TRY(ensure(h, h->partial, sizeof(double) * (size_t)tiles * (P << csh)));
    FTD_TRY(ensure(h, h->xbuf, xbytes));
    const bool per_sys = learned || g;
    int* Js = nullptr;
    if (per_sys) {
        FTD_TRY(ensure(h, h->jsys, sizeof(int) * nj));
        h->jsys_count = nj;
        Js = (int*)h->jsys.p;
    }
    if (g) FTD_TRY(ensure(h, h->R, sizeof(double) * (size_t)4 * P * n_sys));
    if (learned)
        std::snprintf(h->path, sizeof h->path, "filtered_learned_general%s", gs);
    else
        std::snprintf(h->path, sizeof h->path, "filtered_v%d_general%s", version, gs);
    int nt = 0;
    if (version == 2) {
        FTD_LAUNCH(ftd::launch_boundary_maxabs(D, n_sys, n_row, P, csh, (double*)h->partial.p, &nt, st));
        FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, nt, P, n_row, h->rho, h->eps_value,
                                         h->eps_direct, csh, Jp, st));
    }
    FTD_LAUNCH(ftd::launch_gen_filtered_sweep(per_sys ? 1 : version, A, B, C, D, n_sys, n_row, P,
                                              Jp, learned, h->eps_value, h->eps_direct, Js, st));
    FTD_LAUNCH(ftd::launch_gen_filtered_iface(A, C, D, n_sys, n_row, P, cyclic, (double*)h->xbuf.p,
                                              (double*)h->partial.p, &nt, st));
    if (version == 1 && !learned)
        FTD_LAUNCH(ftd::launch_compute_J((double*)h->partial.p, nt, P, n_row, h->rho, h->eps_value,
                                         h->eps_direct, csh, Jp, st));
    if (per_sys)
        FTD_LAUNCH(ftd::launch_gen_filtered_check(learned, g, A, C, D, n_sys, n_row, P, cyclic,
                                                  Jp, (double*)h->xbuf.p, Js, h->eps_value,
                                                  h->eps_direct, (double*)h->R.p, st));
    FTD_LAUNCH(ftd::launch_gen_filtered_final(A, C, D, n_sys, n_row, P, Jp, (double*)h->xbuf.p, Js,
                                              st));
    return FTD_OK;
}

ftd_status do_solve(ftd_handle* h, double* A, double* B, double* C, double* D, int n_sys,
                    int n_row, int cyclic) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (!A || !B || !C || !D) return fail(FTD_ERR_INVALID, "null array pointer");
    if (n_sys < 0 || n_row < 1) return fail(FTD_ERR_INVALID, "need n_sys >= 0 and n_row >= 1");
    if (n_sys == 0) return FTD_OK;
    const bool filtered = h->kind != FTD_KIND_PASCAL && h->P > 1;
    if (filtered && (n_sys != h->n_sys || n_row != h->n_row))
        return fail(FTD_ERR_INVALID, "filtered solves use the dimensions of ftd_create");
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    ftd_status s = filtered ? solve_filtered(h, A, B, C, D, cyclic)
                            : solve_exact(h, A, B, C, D, n_sys, n_row, cyclic);
    if (s != FTD_OK) return s;
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

// Sizes the handle's scratch for every solve that its dimensions, kind,
// tuning and J mode can take (ftd_solve, ftd_solve_cyclic and
// ftd_solve_complex, with or without 16-byte aligned arrays, which select
// the streamed or the fused kernels) and allocates it, so that those solves
// allocate nothing. Waits for the device before it replaces a buffer that
// queued work may still use.
ftd_status reserve(ftd_handle* h) {
    if (h->n_sys == 0) return FTD_OK;
    char path[sizeof h->path];
    std::memcpy(path, h->path, sizeof path);
    const size_t jcount = h->jsys_count;
    DevBuf* const bufs[] = {&h->partial, &h->xbuf, &h->R, &h->E, &h->jsys};
    for (DevBuf* b : bufs) b->want = 0;
    double* const aligned = reinterpret_cast<double*>(uintptr_t(4096));
    double* const unaligned = reinterpret_cast<double*>(uintptr_t(4104));
    const bool filtered = h->kind != FTD_KIND_PASCAL && h->P > 1;
    const bool complex_ok = h->n_sys <= (1 << 29);
    h->sizing = 1;
    for (double* p : {aligned, unaligned})
        for (int cyclic = 0; cyclic < 2; ++cyclic) {
            // sizing passes launch nothing; a shape they reject is rejected
            // again, with its message, by the solve itself
            if (filtered) {
                (void)solve_filtered(h, p, p, p, p, cyclic);
                if (!cyclic && complex_ok) (void)solve_filtered(h, p, p, p, p, 0, 1);
            } else {
                (void)solve_exact(h, p, p, p, p, h->n_sys, h->n_row, cyclic);
                if (!cyclic && complex_ok)
                    (void)solve_exact(h, p, p, p, p, 2 * h->n_sys, h->n_row, 0, 1);
            }
        }
    h->sizing = 0;
    std::memcpy(h->path, path, sizeof path);
    h->jsys_count = jcount;
    bool grow = false;
    for (DevBuf* b : bufs) grow = grow || (b->p && b->want > b->bytes);
    if (grow) FTD_HIP(hipDeviceSynchronize());
    for (DevBuf* b : bufs)
        if (b->want) FTD_TRY(ensure(*b, b->want));
    return FTD_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================

extern "C" {

int ftd_kind_from_string(const char* s) {
    if (!s) return FTD_KIND_FILTERED;
    std::string t;
    for (const char* c = s; *c; ++c) t.push_back((char)std::tolower((unsigned char)*c));
    if (t == "pascal" || t == "pascal_tdma") return FTD_KIND_PASCAL;
    if (t == "filtered_v2") return FTD_KIND_FILTERED_V2;
    return FTD_KIND_FILTERED;
}

ftd_status ftd_create(ftd_handle** out, int kind, int n_sys, int n_row, int partitions,
                      double eps_constant, void* stream) {
    if (!out) return fail(FTD_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (kind < FTD_KIND_FILTERED || kind > FTD_KIND_PASCAL)
        return fail(FTD_ERR_INVALID, "kind must be FTD_KIND_FILTERED, _FILTERED_V2 or _PASCAL");
    if (n_sys < 0 || n_row < 1 || partitions < 1)
        return fail(FTD_ERR_INVALID, "need n_sys >= 0, n_row >= 1, partitions >= 1");
    if (partitions > n_row) return fail(FTD_ERR_INVALID, "more partitions than rows");
    ftd_handle* h = new (std::nothrow) ftd_handle();
    if (!h) return fail(FTD_ERR_ALLOC, "out of host memory");
    h->kind = kind;
    h->n_sys = n_sys;
    h->n_row = n_row;
    h->P = partitions;
    h->eps_value = eps_constant;
    h->stream = (hipStream_t)stream;
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess) e = hipMalloc(&h->rho, sizeof(double) * 2 * (size_t)(n_row + 1));
    if (e == hipSuccess) e = hipMemset(h->rho, 0, sizeof(double) * 2 * (size_t)(n_row + 1));
    // J per partition; entries P..2P-1 hold the imaginary part's J after a
    // complex solve
    if (e == hipSuccess) e = hipMalloc(&h->J, sizeof(int) * 2 * (size_t)partitions);
    if (e == hipSuccess) {
        std::vector<int> neg(2 * partitions, -1);
        e = hipMemcpy(h->J, neg.data(), sizeof(int) * 2 * partitions, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e != hipSuccess) {
        ftd_destroy(h);
        return hip_fail(e, "ftd_create");
    }
    const ftd_status st = reserve(h);
    if (st != FTD_OK) {
        const std::string msg = g_err;
        ftd_destroy(h);
        return fail(st, msg);
    }
    *out = h;
    return FTD_OK;
}

ftd_status ftd_destroy(ftd_handle* h) {
    if (!h) return FTD_OK;
    // teardown: errors of the frees are not reported (the handle is gone)
    if (h->rho) (void)hipFree(h->rho);
    if (h->J) (void)hipFree(h->J);
    for (DevBuf* b : {&h->partial, &h->xbuf, &h->R, &h->E, &h->b_dlanes, &h->b_dlong_off,
                      &h->b_dlong_n, &h->jsys})
        if (b->p) (void)hipFree(b->p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    delete h;
    return FTD_OK;
}

ftd_status ftd_solve(ftd_handle* h, double* A, double* B, double* C, double* D) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    return do_solve(h, A, B, C, D, h->n_sys, h->n_row, 0);
}

ftd_status ftd_solve_nr(ftd_handle* h, double* A, double* B, double* C, double* D, int n_sys,
                        int n_row) {
    return do_solve(h, A, B, C, D, n_sys, n_row, 0);
}

ftd_status ftd_solve_cyclic(ftd_handle* h, double* A, double* B, double* C, double* D) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    return do_solve(h, A, B, C, D, h->n_sys, h->n_row, 1);
}

ftd_status ftd_set_rho(ftd_handle* h, const double* A, const double* B, const double* C) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (h->kind == FTD_KIND_PASCAL) return FTD_OK;
    if (!A || !B || !C) return fail(FTD_ERR_INVALID, "null array pointer");
    if (h->n_sys == 0) return FTD_OK;
    FTD_HIP(ftd::launch_set_rho(A, B, C, h->n_sys, h->n_row, h->rho, h->stream));
    return FTD_OK;
}

ftd_status ftd_set_eps_constant(ftd_handle* h, double eps_constant) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (h->kind == FTD_KIND_PASCAL) return FTD_OK;
    h->eps_value = eps_constant;
    h->eps_direct = 0;
    return FTD_OK;
}

ftd_status ftd_set_eps(ftd_handle* h, double eps) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    h->eps_value = eps;
    h->eps_direct = 1;
    return FTD_OK;
}

ftd_status ftd_set_stream(ftd_handle* h, void* stream) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    h->stream = (hipStream_t)stream;
    return FTD_OK;
}

ftd_status ftd_set_profiling(ftd_handle* h, int enable) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    h->profiling = enable ? 1 : 0;
    if (!enable) h->last_ms = 0.0;
    return FTD_OK;
}

ftd_status ftd_last_J(ftd_handle* h, int* J, int n) {
    if (!h || !J) return fail(FTD_ERR_INVALID, "null argument");
    if (n < 0 || n > 2 * h->P) return fail(FTD_ERR_INVALID, "n must be in [0, 2 * partitions]");
    if (n == 0) return FTD_OK;
    FTD_HIP(hipMemcpyAsync(J, h->J, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    FTD_HIP(hipStreamSynchronize(h->stream));
    return FTD_OK;
}

double ftd_last_gpu_ms(const ftd_handle* h) { return h ? h->last_ms : 0.0; }

double ftd_last_comm_ms(const ftd_handle* h) {
    (void)h;
    return 0.0;
}

ftd_status ftd_set_tuning(ftd_handle* h, int rows_per_lane, int max_threads, int systems_per_tile,
                          int lanes_per_system) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (rows_per_lane == 0 || rows_per_lane == -1 || rows_per_lane == -2) {
        h->t_M = h->t_T = h->t_S = h->t_Q = 0;
        h->force_general = rows_per_lane == -1;
        h->no_stream = rows_per_lane == -2;
        return reserve(h);
    }
    h->force_general = 0;
    h->no_stream = 0;
    if (!ftd::pascal_fused_supported(rows_per_lane, max_threads) &&
        !ftd::filtered_fused_supported(rows_per_lane, max_threads))
        return fail(FTD_ERR_UNSUPPORTED, "no compiled fused kernel with this rows_per_lane/max_threads");
    if (systems_per_tile && (!is_pow2(systems_per_tile) || systems_per_tile > 64))
        return fail(FTD_ERR_INVALID, "systems_per_tile must be a power of two <= 64");
    if (lanes_per_system < 0) return fail(FTD_ERR_INVALID, "lanes_per_system must be >= 0");
    h->t_M = rows_per_lane;
    h->t_T = max_threads;
    h->t_S = systems_per_tile;
    h->t_Q = lanes_per_system;
    return reserve(h);
}

// ---------------------------------------------------------------------------
// Autotuner: times the automatic choice and every compiled register-resident
// shape that holds this solve, each on a fresh copy of D, and keeps the
// fastest as the handle's tuning. PaScaL keeps P partitions when P > 1 (and
// the filtered kinds always keep theirs), so the tuned solve gives the same
// bits as the reference partitioning; with P == 1 the internal partition
// count is a tuning choice (results stay exact to rounding).
// ---------------------------------------------------------------------------
namespace {

struct TuneCand {
    int mode;  // 0 automatic, 1 automatic without the streamed kernels, 2 forced shape
    int M, T, S, Q;
};

static void tune_candidates(const ftd_handle* h, int n_row, std::vector<TuneCand>& out) {
    out.push_back({0, 0, 0, 0, 0});
    out.push_back({1, 0, 0, 0, 0});
    static const int Ms[] = {8, 16, 32, 64};
    static const int Ts[] = {256, 512};
    static const int Ss[] = {16, 32, 64};
    const bool filtered = h->kind != FTD_KIND_PASCAL && h->P > 1;
    for (int M : Ms)
        for (int T : Ts) {
            if (filtered) {
                if (!ftd::filtered_fused_supported(M, T)) continue;
                int s_, m_max;
                part_of(n_row, h->P, 0, s_, m_max);
                if (m_max > M) continue;
                for (int S : Ss)
                    if (S * h->P <= T) out.push_back({2, M, T, S, 0});
            } else {
                if (!ftd::pascal_fused_supported(M, T)) continue;
                const int Q = h->P > 1 ? h->P : std::max(2, (n_row + M - 1) / M);
                for (int S : Ss) {
                    FusedShape fs;
                    if (shape_for(M, T, Q, n_row, S, fs)) out.push_back({2, M, T, S, Q});
                }
            }
        }
}

static void apply_cand(ftd_handle* h, const TuneCand& c) {
    h->force_general = 0;
    h->no_stream = c.mode == 1;
    h->t_M = c.mode == 2 ? c.M : 0;
    h->t_T = c.mode == 2 ? c.T : 0;
    h->t_S = c.mode == 2 ? c.S : 0;
    h->t_Q = c.mode == 2 ? c.Q : 0;
}

}  // namespace

ftd_status ftd_autotune(ftd_handle* h, double* A, double* B, double* C, const double* D, int reps,
                        char* report, size_t report_len) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (!A || !B || !C || !D) return fail(FTD_ERR_INVALID, "null array pointer");
    if (reps < 1 || reps > 100) return fail(FTD_ERR_INVALID, "reps must be in [1, 100]");
    if (h->n_sys == 0 || h->n_row < 3) return fail(FTD_ERR_INVALID, "nothing to tune");
    const size_t bytes = sizeof(double) * (size_t)h->n_sys * h->n_row;
    const TuneCand saved = {h->no_stream ? 1 : (h->t_M ? 2 : 0), h->t_M, h->t_T, h->t_S, h->t_Q};
    const int saved_general = h->force_general, saved_prof = h->profiling;
    std::vector<TuneCand> cands;
    tune_candidates(h, h->n_row, cands);
    // every solve runs on fresh copies of A, C and D (the general kernels
    // use A and C as scratch), so the caller's arrays are left as they were
    double* work = nullptr;
    FTD_HIP(hipMalloc(&work, 3 * bytes));
    const size_t nel = bytes / sizeof(double);
    double* const wa = work;
    double* const wc = work + nel;
    double* const wd = work + 2 * nel;
    auto restore = [&]() {
        hipError_t e = hipMemcpyAsync(wa, A, bytes, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(wc, C, bytes, hipMemcpyDeviceToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(wd, D, bytes, hipMemcpyDeviceToDevice, h->stream);
        return e;
    };
    std::string rep = "{\"candidates\": [";
    int best = -1;
    double best_ms = 0.0;
    std::string best_path;
    h->profiling = 0;
    for (size_t k = 0; k < cands.size(); ++k) {
        apply_cand(h, cands[k]);
        hipError_t e = restore();
        ftd_status st = e == hipSuccess ? do_solve(h, wa, B, wc, wd, h->n_sys, h->n_row, 0)
                                        : FTD_ERR_HIP;
        if (st != FTD_OK) continue;  // this shape does not hold the solve
        const std::string path = h->path;
        float total = 0.0f;
        for (int r = 0; r < reps && st == FTD_OK; ++r) {
            e = restore();
            if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
            if (e != hipSuccess) break;
            st = do_solve(h, wa, B, wc, wd, h->n_sys, h->n_row, 0);
            if (st == FTD_OK) e = hipEventRecord(h->ev1, h->stream);
            if (e == hipSuccess && st == FTD_OK) e = hipEventSynchronize(h->ev1);
            float ms = 0.0f;
            if (e == hipSuccess && st == FTD_OK) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
            if (e != hipSuccess) break;
            total += ms;
        }
        if (e != hipSuccess || st != FTD_OK) {
            (void)hipFree(work);
            apply_cand(h, saved);
            h->force_general = saved_general;
            h->profiling = saved_prof;
            return e != hipSuccess ? hip_fail(e, "autotune") : st;
        }
        const double ms = total / reps;
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "%s{\"mode\": %d, \"M\": %d, \"T\": %d, \"S\": %d, \"Q\": %d, "
                      "\"ms\": %.6f, \"path\": \"%s\"}",
                      best < 0 && rep.back() == '[' ? "" : ", ", cands[k].mode, cands[k].M,
                      cands[k].T, cands[k].S, cands[k].Q, ms, path.c_str());
        rep += buf;
        if (best < 0 || ms < best_ms) {
            best = (int)k;
            best_ms = ms;
            best_path = path;
        }
    }
    const hipError_t ef = hipFree(work);
    h->profiling = saved_prof;
    h->force_general = saved_general;
    if (ef != hipSuccess) {
        apply_cand(h, saved);
        return hip_fail(ef, "hipFree");
    }
    if (best < 0) {
        apply_cand(h, saved);
        return fail(FTD_ERR_UNSUPPORTED, "no candidate shape holds this solve");
    }
    apply_cand(h, cands[best]);
    FTD_TRY(reserve(h));
    char buf[256];
    std::snprintf(buf, sizeof buf, "], \"best\": {\"index\": %d, \"ms\": %.6f, \"path\": \"%s\"}}",
                  best, best_ms, best_path.c_str());
    rep += buf;
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return FTD_OK;
}

static ftd_status solve_f32_common(ftd_handle* h, const float* A, const float* B, const float* C,
                                   const void* D, double* X, int d_is_float) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (!A || !B || !C || !D || !X) return fail(FTD_ERR_INVALID, "null array pointer");
    if (h->kind != FTD_KIND_PASCAL && h->P > 1)
        return fail(FTD_ERR_UNSUPPORTED, "FP32 storage is available for the exact (PaScaL) solve");
    if (h->n_sys == 0) return FTD_OK;
    FusedShape fs;
    if (h->n_row < 6 || !pick_pascal(h, h->n_row, 0, fs))
        return fail(FTD_ERR_UNSUPPORTED, "no fused kernel holds this system for FP32 storage");
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    // FP32 A, B, C, D: the streamed kernel runs the same partitions (same
    // bits) where a variant holds them and the arrays allow its DMA
    int s_, m_max;
    part_of(h->n_row, fs.Q, 0, s_, m_max);
    const int Ms =
        (d_is_float && !h->t_M && !h->no_stream) ? ftd::stream_rows_f32(m_max, fs.Q) : 0;
    const bool streamed =
        Ms && ftd::stream_eligible(h->n_sys, A, B, C, D) && !((uintptr_t)X & 15u);
    if (streamed) {
        std::snprintf(h->path, sizeof h->path, "pascal_stream_f32 M=%d Q=%d R=%d", Ms, fs.Q,
                      ftd::stream_ring_depth_f32(Ms, fs.Q));
        FTD_HIP(ftd::launch_pascal_stream_f32(Ms, fs.Q, A, B, C, (const float*)D, X, h->n_sys,
                                              h->n_row, h->stream));
    } else {
        std::snprintf(h->path, sizeof h->path, "pascal_fused_%s M=%d T=%d S=%d Q=%d",
                      d_is_float ? "f32" : "f32c", fs.M, fs.T, fs.S, fs.Q);
        if (d_is_float)
            FTD_HIP(ftd::launch_pascal_fused_f32(fs.M, fs.T, A, B, C, (const float*)D, X,
                                                 h->n_sys, h->n_row, fs.Q, fs.S, 0, h->stream));
        else
            FTD_HIP(ftd::launch_pascal_fused_f32c(fs.M, fs.T, A, B, C, X, h->n_sys, h->n_row,
                                                  fs.Q, fs.S, 0, h->stream));
    }
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

ftd_status ftd_solve_f32c(ftd_handle* h, const float* A, const float* B, const float* C, double* D) {
    return solve_f32_common(h, A, B, C, D, D, 0);
}

ftd_status ftd_solve_f32(ftd_handle* h, const float* A, const float* B, const float* C,
                         const float* D, double* X) {
    return solve_f32_common(h, A, B, C, D, X, 1);
}

ftd_status ftd_solve_f32_refine(ftd_handle* h, const float* A, const float* B, const float* C,
                                const float* D, double* X, int sweeps) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (!A || !B || !C || !D || !X) return fail(FTD_ERR_INVALID, "null array pointer");
    if (sweeps < 0 || sweeps > FTD_MAX_REFINE_SWEEPS)
        return fail(FTD_ERR_INVALID, "refinement sweeps must lie in [0, 4]");
    if (h->kind != FTD_KIND_PASCAL && h->P > 1)
        return fail(FTD_ERR_UNSUPPORTED, "FP32 storage is available for the exact (PaScaL) solve");
    if (h->n_sys == 0) return FTD_OK;
    FusedShape fs;
    if (h->n_row < 6 || !pick_pascal(h, h->n_row, 0, fs))
        return fail(FTD_ERR_UNSUPPORTED, "no fused kernel holds this system for FP32 storage");
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    std::snprintf(h->path, sizeof h->path, "pascal_ir M=%d T=%d S=%d Q=%d sweeps=%d", fs.M, fs.T,
                  fs.S, fs.Q, sweeps);
    for (int k = 0; k <= sweeps; ++k)
        FTD_HIP(ftd::launch_pascal_ir(fs.M, fs.T, A, B, C, D, X, h->n_sys, h->n_row, fs.Q, fs.S,
                                      k > 0, h->stream));
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

namespace {

// the plan of this layout on the handle (made and uploaded once per layout)
static ftd_status batch_layout(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                        int n_systems) {
    if (!row_offsets || !n_rows) return fail(FTD_ERR_INVALID, "null layout array");
    const size_t n = (size_t)n_systems;
    const int chunk = h->b_chunk ? h->b_chunk : ftd::batched_default_chunk();
    const bool same = h->b_valid && h->b_plan_chunk == chunk && h->b_off.size() == n &&
                      std::memcmp(h->b_off.data(), row_offsets, n * sizeof(int64_t)) == 0 &&
                      std::memcmp(h->b_rows.data(), n_rows, n * sizeof(int)) == 0;
    if (same) return FTD_OK;
    for (size_t s = 0; s < n; ++s) {
        if (n_rows[s] < 1) return fail(FTD_ERR_INVALID, "every system needs n_rows >= 1");
        if (row_offsets[s] < 0) return fail(FTD_ERR_INVALID, "row offsets must be >= 0");
    }
    // the previous plan may still be in use by work queued on the stream
    if (h->b_valid) FTD_HIP(hipStreamSynchronize(h->stream));
    h->b_valid = 0;
    h->b_items = ftd::batch_plan(n_systems, row_offsets, n_rows, chunk, h->b_lanes,
                                 h->b_long_off, h->b_long_n, h->b_chunked, h->b_last);
    if (h->b_items < 0) return fail(FTD_ERR_INVALID, "every system needs n_rows >= 1");
    const size_t nl = h->b_long_n.size();
    FTD_TRY(ensure(h->b_dlanes, sizeof(ftd::BatchLane) * std::max<size_t>(1, h->b_lanes.size())));
    FTD_TRY(ensure(h->b_dlong_off, sizeof(int64_t) * std::max<size_t>(1, nl)));
    FTD_TRY(ensure(h->b_dlong_n, sizeof(int) * std::max<size_t>(1, nl)));
    if (!h->b_lanes.empty())
        FTD_HIP(hipMemcpyAsync(h->b_dlanes.p, h->b_lanes.data(),
                               sizeof(ftd::BatchLane) * h->b_lanes.size(),
                               hipMemcpyHostToDevice, h->stream));
    if (nl) {
        FTD_HIP(hipMemcpyAsync(h->b_dlong_off.p, h->b_long_off.data(), sizeof(int64_t) * nl,
                               hipMemcpyHostToDevice, h->stream));
        FTD_HIP(hipMemcpyAsync(h->b_dlong_n.p, h->b_long_n.data(), sizeof(int) * nl,
                               hipMemcpyHostToDevice, h->stream));
    }
    // the host copies must outlive the asynchronous uploads
    FTD_HIP(hipStreamSynchronize(h->stream));
    h->b_off.assign(row_offsets, row_offsets + n);
    h->b_rows.assign(n_rows, n_rows + n);
    h->b_plan_chunk = chunk;
    h->b_valid = 1;
    return FTD_OK;
}

}  // namespace

ftd_status ftd_solve_batched(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                             int n_systems, double* A, double* B, double* C, double* D) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (n_systems < 0) return fail(FTD_ERR_INVALID, "n_systems must be >= 0");
    if (n_systems == 0) return FTD_OK;
    if (!A || !B || !C || !D) return fail(FTD_ERR_INVALID, "null array pointer");
    FTD_TRY(batch_layout(h, row_offsets, n_rows, n_systems));
    const int nl = (int)h->b_long_n.size();
    std::snprintf(h->path, sizeof h->path, "batched chunk=%d items=%d chunked=%d long=%d",
                  h->b_plan_chunk, h->b_items, h->b_chunked, nl);
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    FTD_HIP(ftd::launch_batched(A, B, C, D, h->b_plan_chunk,
                                (const ftd::BatchLane*)h->b_dlanes.p, h->b_chunked,
                                h->b_items, h->b_last,
                                (const int64_t*)h->b_dlong_off.p, (const int*)h->b_dlong_n.p, nl,
                                h->stream));
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

ftd_status ftd_solve_batched_f32(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                                 int n_systems, const float* A, const float* B, const float* C,
                                 const float* D, double* X) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (n_systems < 0) return fail(FTD_ERR_INVALID, "n_systems must be >= 0");
    if (n_systems == 0) return FTD_OK;
    if (!A || !B || !C || !D || !X) return fail(FTD_ERR_INVALID, "null array pointer");
    FTD_TRY(batch_layout(h, row_offsets, n_rows, n_systems));
    if (!h->b_long_n.empty())
        return fail(FTD_ERR_UNSUPPORTED,
                    "FP32 batches take systems of up to " +
                        std::to_string(ftd::batched_max_rows(h->b_plan_chunk)) + " rows");
    if (h->b_last < 3)
        return fail(FTD_ERR_UNSUPPORTED, "FP32 batches must span at least 4 elements");
    std::snprintf(h->path, sizeof h->path, "batched_f32 chunk=%d items=%d chunked=%d",
                  h->b_plan_chunk, h->b_items, h->b_chunked);
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    FTD_HIP(ftd::launch_batched_f32(A, B, C, D, X, h->b_plan_chunk,
                                    (const ftd::BatchLane*)h->b_dlanes.p, h->b_chunked,
                                    h->b_items, h->b_last, h->stream));
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

ftd_status ftd_set_batch_chunk(ftd_handle* h, int rows) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (rows != 0 && rows != 32 && rows != 48)
        return fail(FTD_ERR_INVALID, "rows per chunk must be 32, 48 or 0 (the default)");
    h->b_chunk = rows;
    return FTD_OK;
}

ftd_status ftd_autotune_batched(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                                int n_systems, double* A, double* B, const double* C,
                                const double* D, int reps, char* report, size_t report_len) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (n_systems < 1) return fail(FTD_ERR_INVALID, "nothing to tune");
    if (!A || !B || !C || !D) return fail(FTD_ERR_INVALID, "null array pointer");
    if (reps < 1 || reps > 100) return fail(FTD_ERR_INVALID, "reps must be in [1, 100]");
    const int saved_chunk = h->b_chunk, saved_prof = h->profiling;
    h->b_chunk = 0;
    FTD_TRY(batch_layout(h, row_offsets, n_rows, n_systems));
    // the solves run on copies of C (the lane-per-system kernel keeps c' there)
    // and D, so the caller's arrays are left as they were
    const size_t bytes = sizeof(double) * (size_t)(h->b_last + 1);
    double* work = nullptr;
    FTD_HIP(hipMalloc(&work, 2 * bytes));
    double* wc = work;
    double* wd = work + (h->b_last + 1);
    std::string rep = "{\"candidates\": [";
    int best = 0;
    double best_ms = 0.0;
    h->profiling = 0;
    ftd_status st = FTD_OK;
    hipError_t e = hipSuccess;
    const int chunks[2] = {48, 32};
    for (int k = 0; k < 2 && st == FTD_OK && e == hipSuccess; ++k) {
        h->b_chunk = chunks[k];
        float total = 0.0f;
        for (int r = 0; r <= reps && st == FTD_OK && e == hipSuccess; ++r) {  // r = 0: warm-up
            e = hipMemcpyAsync(wc, C, bytes, hipMemcpyDeviceToDevice, h->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(wd, D, bytes, hipMemcpyDeviceToDevice, h->stream);
            if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
            if (e != hipSuccess) break;
            st = ftd_solve_batched(h, row_offsets, n_rows, n_systems, A, B, wc, wd);
            if (st == FTD_OK) e = hipEventRecord(h->ev1, h->stream);
            if (st == FTD_OK && e == hipSuccess) e = hipEventSynchronize(h->ev1);
            float ms = 0.0f;
            if (st == FTD_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
            if (r > 0) total += ms;
        }
        if (st != FTD_OK || e != hipSuccess) break;
        const double ms = total / reps;
        char buf[200];
        std::snprintf(buf, sizeof buf, "%s{\"chunk\": %d, \"ms\": %.6f, \"path\": \"%s\"}",
                      k ? ", " : "", chunks[k], ms, h->path);
        rep += buf;
        if (k == 0 || ms < best_ms) {
            best = chunks[k];
            best_ms = ms;
        }
    }
    hipError_t ef = hipFree(work);
    h->profiling = saved_prof;
    if (st != FTD_OK || e != hipSuccess || ef != hipSuccess) {
        h->b_chunk = saved_chunk;
        return st != FTD_OK ? st : hip_fail(e != hipSuccess ? e : ef, "autotune_batched");
    }
    h->b_chunk = best;
    char buf[96];
    std::snprintf(buf, sizeof buf, "], \"best\": {\"chunk\": %d, \"ms\": %.6f}}", best,
                  best_ms);
    rep += buf;
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = '\0';
    }
    return FTD_OK;
}

ftd_status ftd_solve_complex(ftd_handle* h, double* A, double* B, double* C, double* Z) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (!A || !B || !C || !Z) return fail(FTD_ERR_INVALID, "null array pointer");
    if (h->n_sys == 0) return FTD_OK;
    if (h->n_sys > (1 << 29)) return fail(FTD_ERR_INVALID, "too many systems for complex mode");
    const bool filtered = h->kind != FTD_KIND_PASCAL && h->P > 1;
    if (h->profiling) FTD_HIP(hipEventRecord(h->ev0, h->stream));
    ftd_status s = filtered ? solve_filtered(h, A, B, C, Z, 0, 1)
                            : solve_exact(h, A, B, C, Z, 2 * h->n_sys, h->n_row, 0, 1);
    if (s != FTD_OK) return s;
    if (h->profiling) {
        FTD_HIP(hipEventRecord(h->ev1, h->stream));
        FTD_HIP(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        FTD_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

ftd_status ftd_set_j_mode(ftd_handle* h, int mode) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (mode != 0 && mode != 1) return fail(FTD_ERR_INVALID, "J mode must be 0 or 1");
    h->j_mode = mode;
    return reserve(h);
}

ftd_status ftd_set_guard(ftd_handle* h, int enable) {
    if (!h) return fail(FTD_ERR_INVALID, "null handle");
    if (enable != 0 && enable != 1) return fail(FTD_ERR_INVALID, "guard must be 0 or 1");
    h->guard = enable;
    return reserve(h);
}

ftd_status ftd_last_J_systems(ftd_handle* h, int* out, int64_t n) {
    if (!h || !out) return fail(FTD_ERR_INVALID, "null argument");
    if (n < 0 || (size_t)n > h->jsys_count)
        return fail(FTD_ERR_INVALID, "n exceeds the J values of the last learned solve");
    if (n == 0) return FTD_OK;
    FTD_HIP(hipMemcpyAsync(out, h->jsys.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                           h->stream));
    FTD_HIP(hipStreamSynchronize(h->stream));
    return FTD_OK;
}

// ---------------------------------------------------------------------------
// ftd_solve_static: the solve of a handle's kind with its scratch in a static
// device array instead of hipMalloc'd buffers, for callers that cannot
// allocate. The array is a host-side stand-in in host code: its device
// address is taken once per device with hipGetSymbolAddress, and only that
// address is used.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

constexpr size_t kStaticScratch = (size_t)FTD_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up once)
ftd_status static_scratch(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(FTD_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(FTD_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *out = (unsigned char*)g_scratch_ptr[dev];
    return FTD_OK;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

ftd_status ftd_solve_static(int kind, int n_sys, int n_row, int partitions, double eps_constant,
                            int cyclic, double* A, double* B, double* C, double* D,
                            void* stream) {
    if (kind < FTD_KIND_FILTERED || kind > FTD_KIND_PASCAL)
        return fail(FTD_ERR_INVALID, "kind must be FTD_KIND_FILTERED, _FILTERED_V2 or _PASCAL");
    if (n_sys < 0 || n_row < 1 || partitions < 1)
        return fail(FTD_ERR_INVALID, "need n_sys >= 0, n_row >= 1, partitions >= 1");
    if (partitions > n_row) return fail(FTD_ERR_INVALID, "more partitions than rows");
    if (!A || !B || !C || !D) return fail(FTD_ERR_INVALID, "null array pointer");
    if (n_sys == 0) return FTD_OK;
    // a handle on the stack: its buffers will point into the static array
    ftd_handle h;
    h.kind = kind;
    h.n_sys = n_sys;
    h.n_row = n_row;
    h.P = partitions;
    h.eps_value = eps_constant;
    h.stream = (hipStream_t)stream;
    const bool filtered = kind != FTD_KIND_PASCAL && partitions > 1;
    // what this solve would use, for these arrays (their alignment picks the
    // streamed or the fused kernels)
    h.sizing = 1;
    const ftd_status sz = filtered ? solve_filtered(&h, A, B, C, D, cyclic ? 1 : 0)
                                   : solve_exact(&h, A, B, C, D, n_sys, n_row, cyclic ? 1 : 0);
    h.sizing = 0;
    if (sz != FTD_OK) return sz;
    const size_t rho_b = round256(sizeof(double) * 2 * ((size_t)n_row + 1));
    const size_t J_b = round256(sizeof(int) * 2 * (size_t)partitions);
    DevBuf* const bufs[] = {&h.partial, &h.xbuf, &h.R, &h.E};
    size_t total = (filtered ? rho_b + J_b : 0);
    for (DevBuf* b : bufs) total += round256(b->want);
    if (total > kStaticScratch)
        return fail(FTD_ERR_UNSUPPORTED,
                    "ftd_solve_static needs " + std::to_string(total) +
                        " bytes of scratch for this solve; the static buffer holds " +
                        std::to_string(kStaticScratch) + " (FTD_STATIC_SCRATCH_BYTES)");
    int dev = 0;
    FTD_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    FTD_TRY(static_scratch(dev, &base));
    size_t pos = 0;
    if (filtered) {
        h.rho = (double*)(base + pos);
        pos += rho_b;
        h.J = (int*)(base + pos);
        pos += J_b;
    }
    for (DevBuf* b : bufs) {
        if (!b->want) continue;
        b->p = base + pos;
        b->bytes = b->want;
        pos += round256(b->want);
    }
    // ftd_set_rho, then the solve; every scratch value a kernel reads is
    // written earlier in the same call (column-0 bounds, J, tile maxima,
    // boundary values), so nothing needs clearing
    ftd_status st = FTD_OK;
    if (filtered) {
        const hipError_t e = ftd::launch_set_rho(A, B, C, n_sys, n_row, h.rho, h.stream);
        if (e != hipSuccess) st = hip_fail(e, "set_rho");
    }
    if (st == FTD_OK)
        st = filtered ? solve_filtered(&h, A, B, C, D, cyclic ? 1 : 0)
                      : solve_exact(&h, A, B, C, D, n_sys, n_row, cyclic ? 1 : 0);
    // the buffers are the static array's: the handle must not free them
    for (DevBuf* b : bufs) b->p = nullptr;
    h.rho = nullptr;
    h.J = nullptr;
    return st;
}

const char* ftd_last_path(const ftd_handle* h) { return h ? h->path : "none"; }

const char* ftd_last_error_string(void) { return g_err.c_str(); }

}  // extern "C"

namespace ftd {
int handle_filtered_state(ftd_handle* h, int** J, const double** rho, double* eps_value,
                          int* eps_direct) {
    if (!h || h->kind == FTD_KIND_PASCAL || h->P < 2 || !h->J || !h->rho)
        return fail(FTD_ERR_INVALID, "not a filtered handle over partitions");
    if (h->j_mode != 0 || h->guard)
        return fail(FTD_ERR_UNSUPPORTED, "learned J and the guard run on the handle's own solve");
    *J = h->J;
    *rho = h->rho;
    *eps_value = h->eps_value;
    *eps_direct = h->eps_direct;
    return FTD_OK;
}
}  // namespace ftd
