// ftdma_kernels.h -- launchers of the gfx950 kernels (internal to libftdma).
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace ftd {

hipError_t launch_set_rho(const double* A, const double* B, const double* C, int n_sys, int n_row,
                          double* rho, hipStream_t st);
// csh = 1: complex right-hand sides stored as interleaved (re, im) columns
// of n_sys = 2 x (complex systems) real columns; the coefficients have
// n_sys / 2 columns, and J is kept per parity (J[p], J[P + p])
hipError_t launch_boundary_maxabs(const double* D, int n_sys, int n_row, int P, int csh,
                                  double* partial, int* n_tiles, hipStream_t st);
hipError_t launch_compute_J(const double* partial, int n_tiles, int P, int n_row, const double* rho,
                            double eps_value, int eps_direct, int csh, int* J, hipStream_t st);

int pascal_fused_supported(int M, int maxt);
size_t pascal_fused_lds(int Q, int S, int cyclic);
hipError_t launch_pascal_fused(int M, int maxt, const double* A, const double* B, const double* C,
                               double* D, int n_sys, int n_row, int Q, int S, int cyclic, int csh,
                               hipStream_t st);

hipError_t launch_pascal_fused_f32c(int M, int maxt, const float* A, const float* B,
                                    const float* C, double* D, int n_sys, int n_row, int Q, int S,
                                    int cyclic, hipStream_t st);
hipError_t launch_pascal_fused_f32(int M, int maxt, const float* A, const float* B, const float* C,
                                   const float* D, double* X, int n_sys, int n_row, int Q, int S,
                                   int cyclic, hipStream_t st);

// FP32 elimination, FP64 reduced solve and update: res = 0 solves T x = D
// into X; res = 1 is one refinement sweep, X += T32^-1 (D - T X) with the
// residual in FP64. Same shapes (M, maxt, Q, S) as launch_pascal_fused.
hipError_t launch_pascal_ir(int M, int maxt, const float* A, const float* B, const float* C,
                            const float* D, double* X, int n_sys, int n_row, int Q, int S, int res,
                            hipStream_t st);

// LDS-streamed persistent kernels (S = 32 systems per tile, two partitions per
// wave): M rows per partition for Q partitions, or 0 when no variant holds
// them; eligible when n_sys % 32 == 0 and the arrays are 16-byte aligned
int stream_rows(int m_max, int Q);
// the guarded filtered variants of the streamed kernel exist for Q <= 8
bool stream_guard_supported(int Q);
int stream_ring_depth(int M, int Q);
bool stream_eligible(int n_sys, const void* A, const void* B, const void* C, const void* D);
// FP32 storage of A, B, C, D, x in FP64 to X: same partitions and bits as
// launch_pascal_fused_f32
int stream_rows_f32(int m_max, int Q);
int stream_ring_depth_f32(int M, int Q);
hipError_t launch_pascal_stream_f32(int M, int Q, const float* A, const float* B, const float* C,
                                    const float* D, double* X, int n_sys, int n_row,
                                    hipStream_t st);
hipError_t launch_pascal_stream(int M, int Q, const double* A, const double* B, const double* C,
                                double* D, int n_sys, int n_row, int cyclic, hipStream_t st);
// pass 3: learned J per (system, partition), eps_ from eps_value/eps_direct,
// J records to Jsys [Q x n_sys]. guard = 1 (non-periodic; version 2 pass 0,
// or version 1 pass 2 or 3): v1 sweeps, and the systems whose dropped terms
// exceed eps_/4 solved exactly; records (J | fallback << 16 | guard << 17)
// to Jsys
hipError_t launch_filtered_stream(int M, int Q, int version, int pass, const double* A,
                                  const double* B, const double* C, double* D, int n_sys,
                                  int n_row, int cyclic, const int* Jp, double* xbuf,
                                  double* partial, double eps_value, int eps_direct, int* Jsys,
                                  int guard, hipStream_t st);

// uneven batches: one lane per chunk of <= 64 rows of a system, 64 lanes per
// work item (a wave); kind 0 = single-chunk systems (plain Thomas), kind 1 =
// chunked systems (PaScaL with P = Q inside the wave)
struct alignas(16) BatchLane {
    int64_t off;     // element offset of the chunk's first row
    int32_t m;       // rows in the chunk (0: idle lane)
    uint8_t q;       // chunk index within its system
    uint8_t Q;       // chunks of its system
    uint8_t lead;    // lane of chunk 0 of its system within the item
    uint8_t kind;    // 1: chunked system, 0: single-chunk (uniform over the item)
};
static_assert(sizeof(BatchLane) == 16, "BatchLane is one dwordx4");

// Plans the batch in chunks of at most `chunk` rows (32 or 48; a system's
// chunks after the first start on 16-element boundaries where its length
// allows, else the para_range split): lanes (128 per item;
// the n_chunked items of chunked systems first, then single-chunk items, each
// range longest first), the systems longer than 128 * chunk rows
// (lane-per-system kernel), and the last element any system holds (the
// staging never reads past it). Returns the number of items, or -1 when
// some n_rows[s] < 1 or the chunk is not one of the two.
int batch_plan(int n_systems, const int64_t* row_offsets, const int* n_rows, int chunk,
               std::vector<BatchLane>& lanes, std::vector<int64_t>& long_off,
               std::vector<int>& long_n, int& n_chunked, int64_t& last);
hipError_t launch_batched(const double* A, const double* B, double* C, double* D, int chunk,
                          const BatchLane* lanes, int n_chunked, int n_items, int64_t last,
                          const int64_t* long_off, const int* long_n, int n_long,
                          hipStream_t st);
// FP32 storage: A, B, C, D float, elimination in FP64 on the widened values,
// x (FP64) into X at the same offsets. No long-system list: the caller
// rejects systems above batched_max_rows(chunk) and batches spanning fewer
// than 4 elements.
hipError_t launch_batched_f32(const float* A, const float* B, const float* C, const float* D,
                              double* X, int chunk, const BatchLane* lanes, int n_chunked,
                              int n_items, int64_t last, hipStream_t st);
int batched_default_chunk();
int batched_max_rows(int chunk);

int thomas_fused_rows(int n_row);
hipError_t launch_thomas_fused(const double* A, const double* B, const double* C, double* D,
                               int n_sys, int n_row, int csh, hipStream_t st);

hipError_t launch_thomas_cyclic_fused(const double* A, const double* B, const double* C, double* D,
                                      int n_sys, int n_row, hipStream_t st);

int filtered_fused_supported(int M, int maxt);
size_t filtered_fused_lds(int P, int S, int guard);
// guard: as launch_filtered_stream
hipError_t launch_filtered_fused(int M, int maxt, int version, int pass, const double* A,
                                 const double* B, const double* C, double* D, int n_sys, int n_row,
                                 int P, int S, int cyclic, int csh, const int* Jp, double* xbuf,
                                 double* partial, double eps_value, int eps_direct, int* Jsys,
                                 int guard, hipStream_t st);

hipError_t launch_gen_thomas(const double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, hipStream_t st);
hipError_t launch_gen_thomas_cyclic(const double* A, const double* B, double* C, double* D,
                                    double* E, int n_sys, int n_row, hipStream_t st);
hipError_t launch_gen_pascal(double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, int P, int cyclic, double* R, double* E, hipStream_t st);
// learned = 1 (version 1): the network's J per (partition, system) to Jsys
hipError_t launch_gen_filtered_sweep(int version, double* A, const double* B, double* C, double* D,
                                     int n_sys, int n_row, int P, const int* Jp, int learned,
                                     double eps_value, int eps_direct, int* Jsys, hipStream_t st);
// per-system J after the interface solve: the learned J's check against
// eps_ and/or the guard (exact reduced solve of flagged systems, scratch
// gbuf of 4P x n_sys doubles); records J | fallback << 16 | guard << 17 to
// Jsys, x0 and xN of guarded systems to xbuf
hipError_t launch_gen_filtered_check(int learned, int guard, const double* A, const double* C,
                                     const double* D, int n_sys, int n_row, int P, int cyclic,
                                     const int* Jp, double* xbuf, int* Jsys, double eps_value,
                                     int eps_direct, double* gbuf, hipStream_t st);
hipError_t launch_gen_filtered_iface(const double* A, const double* C, const double* D, int n_sys,
                                     int n_row, int P, int cyclic, double* xbuf, double* partial,
                                     int* n_tiles, hipStream_t st);
// Jsys != nullptr: J per (partition, system) from its low 16 bits
hipError_t launch_gen_filtered_final(const double* A, const double* C, double* D, int n_sys,
                                     int n_row, int P, const int* Jp, const double* xbuf,
                                     const int* Jsys, hipStream_t st);

}  // namespace ftd

// For a caller inside the library that runs a filtered kind's sweeps itself
// (the heat step's fused filtered z sweep): the handle's J per partition
// (device), its column-0 bounds as set by ftd_set_rho (device, A_rho then
// C_rho, n_row + 1 each) and its eps_ setting. An error for a handle that is
// not a filtered kind over more than one partition, or that has learned J or
// the guard switched on (those run on the handle's own solve).
struct ftd_handle;
namespace ftd {
int handle_filtered_state(ftd_handle* h, int** J, const double** rho, double* eps_value,
                          int* eps_direct);
}  // namespace ftd
