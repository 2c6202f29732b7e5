// ftdma_kernels.h -- launchers of the gfx950 kernels (internal to libftdma).
#pragma once

#include <hip/hip_runtime.h>

#include <stddef.h>

namespace ftd {

hipError_t launch_set_rho(const double* A, const double* B, const double* C, int n_sys, int n_row,
                          double* rho, hipStream_t st);
hipError_t launch_boundary_maxabs(const double* D, int n_sys, int n_row, int P, double* partial,
                                  int* n_tiles, hipStream_t st);
hipError_t launch_compute_J(const double* partial, int n_tiles, int P, int n_row, const double* rho,
                            double eps_value, int eps_direct, int* J, hipStream_t st);

int pascal_fused_supported(int M, int maxt);
size_t pascal_fused_lds(int Q, int S, int cyclic);
hipError_t launch_pascal_fused(int M, int maxt, const double* A, const double* B, const double* C,
                               double* D, int n_sys, int n_row, int Q, int S, int cyclic,
                               hipStream_t st);

hipError_t launch_pascal_fused_f32c(int M, int maxt, const float* A, const float* B,
                                    const float* C, double* D, int n_sys, int n_row, int Q, int S,
                                    int cyclic, hipStream_t st);
hipError_t launch_pascal_fused_f32(int M, int maxt, const float* A, const float* B, const float* C,
                                   const float* D, double* X, int n_sys, int n_row, int Q, int S,
                                   int cyclic, hipStream_t st);

// LDS-streamed persistent kernels (S = 32 systems per tile, two partitions per
// wave): M rows per partition for Q partitions, or 0 when no variant holds
// them; eligible when n_sys % 32 == 0 and the arrays are 16-byte aligned
int stream_rows(int m_max, int Q);
int stream_ring_depth(int M, int Q);
bool stream_eligible(int n_sys, const void* A, const void* B, const void* C, const void* D);
hipError_t launch_pascal_stream(int M, int Q, const double* A, const double* B, const double* C,
                                double* D, int n_sys, int n_row, int cyclic, hipStream_t st);
hipError_t launch_filtered_stream(int M, int Q, int version, int pass, const double* A,
                                  const double* B, const double* C, double* D, int n_sys,
                                  int n_row, int cyclic, const int* Jp, double* xbuf,
                                  double* partial, hipStream_t st);

int thomas_fused_rows(int n_row);
hipError_t launch_thomas_fused(const double* A, const double* B, const double* C, double* D,
                               int n_sys, int n_row, hipStream_t st);

hipError_t launch_thomas_cyclic_fused(const double* A, const double* B, const double* C, double* D,
                                      int n_sys, int n_row, hipStream_t st);

int filtered_fused_supported(int M, int maxt);
size_t filtered_fused_lds(int P, int S);
hipError_t launch_filtered_fused(int M, int maxt, int version, int pass, const double* A,
                                 const double* B, const double* C, double* D, int n_sys, int n_row,
                                 int P, int S, int cyclic, const int* Jp, double* xbuf,
                                 double* partial, hipStream_t st);

hipError_t launch_gen_thomas(const double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, hipStream_t st);
hipError_t launch_gen_thomas_cyclic(const double* A, const double* B, double* C, double* D,
                                    double* E, int n_sys, int n_row, hipStream_t st);
hipError_t launch_gen_pascal(double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, int P, int cyclic, double* R, double* E, hipStream_t st);
hipError_t launch_gen_filtered_sweep(int version, double* A, const double* B, double* C, double* D,
                                     int n_sys, int n_row, int P, const int* Jp, hipStream_t st);
hipError_t launch_gen_filtered_iface(const double* A, const double* C, const double* D, int n_sys,
                                     int n_row, int P, int cyclic, double* xbuf, double* partial,
                                     int* n_tiles, hipStream_t st);
hipError_t launch_gen_filtered_final(const double* A, const double* C, double* D, int n_sys,
                                     int n_row, int P, const int* Jp, const double* xbuf,
                                     hipStream_t st);

}  // namespace ftd
