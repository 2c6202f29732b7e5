/*
 * ftdma_oracle.h -- CPU restatement of the reference's batched tridiagonal
 * solvers. TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it as the checker (or as the timed CPU
 * baseline). No product code links or calls it.
 *
 * Layout (libs/pascal_tdma/include/pascal_tdma.hpp:13): row-major
 * [n_row x n_sys], row j of system i at j*n_sys + i. Every routine works in
 * place like the reference: D <- solution, A and C clobbered, B read-only.
 *
 * Partition p of P covers rows para_range(0, n_row-1, P, p)
 * (libs/pascal_tdma/cpu/para_range.cpp:4-9): the MPI z-slab split of the
 * reference, emulated in one address space.
 */
#ifndef FTDMA_ORACLE_H
#define FTDMA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Normalisation order / J policy of the reference path being restated. */
enum {
    FTDO_SEM_GPU = 0, /* the CUDA path behind TdmaBackendGPU (drop-in target) */
    FTDO_SEM_CPU = 1  /* the host path behind TdmaBackend (libs/.../cpu)      */
};

/* J estimator selector for ftdo_cal_J. */
enum {
    FTDO_J_GPU_RHS = 0, /* k_cal_J_rhs_bound, filtered_tdma_kernels.cuh:220-261 */
    FTDO_J_CPU_V1  = 1, /* FilteredTDMA::cal_J_v1, filtered_tdma.cpp:89-115      */
    FTDO_J_CPU_RHS = 2  /* FilteredTDMA::cal_J_rhs_bound, filtered_tdma.cpp:63-87 */
};

void ftdo_para_range(int start, int end, int nproc, int rank, int* ista, int* iend);

/* Plain Thomas, all systems (tdma_local.cpp:6-52 / tdma_local_cuda.cu:32-64). */
int ftdo_thomas(double* A, const double* B, double* C, double* D,
                int n_sys, int n_row, int sem, int nthreads);

/* Sherman-Morrison periodic Thomas (tdma_local.cpp:108-208). */
int ftdo_thomas_cyclic(double* A, const double* B, double* C, double* D,
                       int n_sys, int n_row, int sem, int nthreads);

/* PaScaL with P partitions (pascal_tdma.cpp:113-348, pascal_tdma_cuda.cu:264-398). */
int ftdo_pascal(double* A, const double* B, double* C, double* D,
                int n_sys, int n_row, int P, int cyclic, int sem, int nthreads);

/* PaScaL with P partitions whose 2P-row reduced system is solved by PaScaL
 * itself over G groups of rows (G = 1: Thomas, as ftdo_pascal); the order
 * of the uneven-batch kernel's reduced solve. Non-periodic. */
int ftdo_pascal_nested(double* A, const double* B, double* C, double* D,
                       int n_sys, int n_row, int P, int G, int sem, int nthreads);

/* ftdo_pascal_nested over the P >= 2 partitions starts[p] .. starts[p+1] - 1
 * (starts[0] = 0, starts[P] = n_row, each of at least 3 rows) in place of
 * the para_range split: the uneven-batch kernel's chunks, whose boundaries
 * fall on 128-byte lines of the batch arrays. */
int ftdo_pascal_starts(double* A, const double* B, double* C, double* D, int n_sys, int n_row,
                       int P, const int* starts, int G, int sem, int nthreads);

/* FP32 elimination with FP64 refinement (no counterpart in the reference):
 * float A, B, C, D [n_row x n_sys] (read only), x (double) to X. A solve is
 * PaScaL over Q partitions with the modified Thomas in float arithmetic (GPU
 * order: rows 0, 1 divided), the reduced system and the interior update in
 * double on the widened values; then `sweeps` times: r = D - T x in double
 * (row 0 without its A term, row n_row-1 without its C term), rounded to
 * float, solved the same way, added to x. */
int ftdo_pascal_ir(const float* A, const float* B, const float* C, const float* D, double* X,
                   int n_sys, int n_row, int Q, int sweeps, int nthreads);

/* Column-0 spectral bounds, n_row+1 entries each (tdma_backend.cpp:33-44). */
void ftdo_set_rho(const double* A, const double* B, const double* C,
                  int n_sys, int n_row, double* A_rho, double* C_rho);

/* J estimators. `bound` is max_b for the RHS bounds and max|x0|+max|xN|
 * for FTDO_J_CPU_V1. n_row is the partition's local row count. */
int ftdo_cal_J(double rho, double bound, double eps, int n_row, int which);

/* Filtered TDMA v1/v2 with P partitions.
 *   A_rho/C_rho : n_row+1 global column-0 bounds (ftdo_set_rho), or NULL (= 0,
 *                 the reference's state before any set_rho call).
 *   eps_value   : eps_constant, turned into eps_ = eps_constant/(m_p*P)^2 per
 *                 partition (filtered_tdma.hpp:60-63); if eps_direct != 0 it
 *                 is used as eps_ unchanged.
 *   J_force     : if non-NULL, P values of J used instead of the estimator.
 *   J_out       : if non-NULL, receives the P values of J used.           */
int ftdo_filtered(double* A, const double* B, double* C, double* D,
                  int n_sys, int n_row, int P, int version, int cyclic, int sem,
                  const double* A_rho, const double* C_rho,
                  double eps_value, int eps_direct,
                  const int* J_force, int* J_out, int nthreads);

/* Learned-J filtered mode (no counterpart in the reference): v1 sweeps and
 * interface, J per (system, partition) from the FP32 network `weights`
 * of hidden width H (W1[5][H], b1[H], W2[H][H], b2[H], W3[H], b3: H^2 + 8H + 1
 * floats, H <= 64) plus `margin`, checked against eps_ (J = m-1 where the first dropped term
 * exceeds it). J_out (optional) receives P*n_sys values, index p*n_sys+i,
 * J | fallback << 16. GPU semantics. */
int ftdo_filtered_learned(double* A, const double* B, double* C, double* D,
                          int n_sys, int n_row, int P, int cyclic,
                          double eps_value, int eps_direct,
                          const float* weights, int hidden, int margin, int* J_out,
                          int nthreads);

/* The network's J (before the clamp to [0, m-1]) for 5 features, and the
 * piecewise-linear log2 feature, exposed for tests. */
int ftdo_mlp_J(const float* weights, int hidden, int margin, const float* features);
float ftdo_lg(double x);

/* Guard mode of the filtered kinds (no counterpart in the reference): v1
 * sweeps and interface; per (partition, system) the terms the filtered
 * solve leaves out (|c0' xN|, |a'_{m-1} x0| and the first truncated term of
 * each side for J[p * n_sys + i]) against eps_/4; a system where any
 * exceeds it is solved exactly (2P-row reduced Thomas, full correction).
 * rec_out[p * n_sys + i] = J used | guard << 17. 2..512 partitions; with
 * cyclic the interfaces wrap and the exact reduced system is periodic. */
int ftdo_filtered_guard(double* A, const double* B, double* C, double* D,
                        int n_sys, int n_row, int P, int cyclic, const int* J,
                        double eps_value, int eps_direct, int* rec_out, int nthreads);

/* Heat ADI harness (apps/heat_cpu on a parts[0] x parts[1] x parts[2] process
 * grid): nx, ny, nz interior cells, kind 0/1/2 = filtered / filtered_v2 /
 * pascal (a direction with one part is plain Thomas), periodic[3] flags, Nt
 * steps of dt. theta_out (optional) receives the (nz+2)(ny+2)(nx+2) field
 * with ghosts, l2_out (optional) the L2 error of main.cpp:59-81. */
int ftdo_heat_run(int nx, int ny, int nz, const int* parts, int kind, int sem,
                  const int* periodic, double dt, int Nt, double* theta_out, double* l2_out,
                  int nthreads);

/* Error text of the last failing call in this thread. */
const char* ftdo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
