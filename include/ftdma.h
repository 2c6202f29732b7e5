/*
 * ftdma.h -- C ABI of the MI355X (gfx950) batched tridiagonal engine.
 *
 * Drop-in for the solve(A,B,C,D,n_sys,n_row) path behind the reference's
 * runtime backend switch, class TdmaBackendGPU
 * (apps/heat_gpu/tdma_backend_gpu.hpp:21-70, tdma_backend_gpu.cu:6-66).
 * Each entry point below names the reference member it replaces.
 *
 * Data contract (same as the reference):
 *   - A, B, C, D are device pointers to row-major [n_row x n_sys] float64
 *     arrays: row j of system i at j*n_sys + i;
 *   - D is overwritten with the solution; A and C may be clobbered; B is
 *     only read;
 *   - A[row 0] and C[row n_row-1] are not read by the non-periodic solves;
 *     the periodic solves read them as the wrap-around couplings.
 *
 * One MI355X replaces the reference's MPI z-slab split: `partitions` (the
 * reference's nprocs) splits every system's rows into P contiguous blocks
 * with para_range (libs/pascal_tdma/cpu/para_range.cpp:4-9). Filtered
 * results depend on P exactly as the reference's do on nprocs; PaScaL is
 * exact for any P, which then only affects rounding.
 *
 * Every call is ordered on the handle's stream and returns without a host
 * synchronisation, except ftd_last_J, solves with profiling enabled and
 * ftd_solve_batched with a layout the handle has not planned yet.
 * Calls on one handle must not overlap (the handle owns scratch memory).
 *
 * Scratch memory: ftd_create allocates all the device scratch that
 * ftd_solve, ftd_solve_cyclic, ftd_solve_complex, ftd_solve_f32 and
 * ftd_solve_f32c of the handle's dimensions can need, for any array
 * alignment; ftd_set_tuning, ftd_set_j_mode, ftd_set_guard and ftd_autotune size it again
 * for their new choice (and may wait for the device to replace a buffer).
 * Those solves therefore allocate nothing, and a handle's first solve may be
 * captured into a hipGraph. Only ftd_solve_nr with dimensions other than the
 * handle's and ftd_solve_batched with a layout not planned yet may allocate.
 */
#ifndef FTDMA_H
#define FTDMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ftd_handle ftd_handle;

typedef enum {
    FTD_OK = 0,
    FTD_ERR_INVALID = 1,     /* bad argument (null pointer, bad size, bad kind) */
    FTD_ERR_UNSUPPORTED = 2, /* a shape the engine cannot solve (e.g. < 3 rows per partition) */
    FTD_ERR_HIP = 3,         /* HIP runtime error (message in ftd_last_error_string) */
    FTD_ERR_ALLOC = 4        /* device allocation failed */
} ftd_status;

/* TdmaBackendGPU::Kind, same order (tdma_backend_gpu.hpp:23). */
enum { FTD_KIND_FILTERED = 0, FTD_KIND_FILTERED_V2 = 1, FTD_KIND_PASCAL = 2 };

/* TdmaBackendGPU::parse (tdma_backend_gpu.cu:6-13): case-insensitive;
 * "pascal" | "pascal_tdma" -> PASCAL, "filtered_v2" -> FILTERED_V2, any other
 * string (including NULL) -> FILTERED. */
int ftd_kind_from_string(const char* s);

/* TdmaBackendGPU::TdmaBackendGPU(kind, n_sys, n_row, myrank, nprocs, comm,
 * left_rank, right_rank, eps_constant) (tdma_backend_gpu.cu:15-33), with the
 * MPI arguments replaced by `partitions`; n_row is the full row count of a
 * system (the sum of the reference's per-rank n_row). eps_ of partition p is
 * eps_constant / (m_p * partitions)^2 (filtered_tdma_cuda.cu:52-53).
 * `stream` is a hipStream_t (NULL = the null stream). */
ftd_status ftd_create(ftd_handle** out, int kind, int n_sys, int n_row, int partitions,
                      double eps_constant, void* stream);

/* TdmaBackendGPU::~TdmaBackendGPU: frees the handle's device scratch. */
ftd_status ftd_destroy(ftd_handle* h);

/* TdmaBackendGPU::solve (tdma_backend_gpu.cu:35-46):
 * FILTERED -> solve_filtered_v1, FILTERED_V2 -> solve_filtered_v2,
 * PASCAL -> PaScaLTDMACUDA::solve; partitions == 1 -> plain Thomas. */
ftd_status ftd_solve(ftd_handle* h, double* A, double* B, double* C, double* D);

/* PaScaLTDMACUDA::solve(A, B, C, D, n_sys, n_row) (pascal_tdma_cuda.hpp:28-29):
 * dimensions given per call. PASCAL accepts any n_sys, n_row; the filtered
 * kinds require the dimensions the handle was created with. */
ftd_status ftd_solve_nr(ftd_handle* h, double* A, double* B, double* C, double* D, int n_sys,
                        int n_row);

/* TdmaBackendGPU::solve_cyclic (tdma_backend_gpu.cu:48-56): periodic systems,
 * row 0 coupled to row n_row-1 through A[0] and C[n_row-1]. */
ftd_status ftd_solve_cyclic(ftd_handle* h, double* A, double* B, double* C, double* D);

/* TdmaBackendGPU::set_rho_device (tdma_backend_gpu.cu:58-62,
 * filtered_tdma_kernels.cuh:46-67): |A/B| and |C/B| of system 0 per row,
 * used by the filtered J estimator. No-op for PASCAL. Until it is called the
 * bounds are 0, which selects full propagation (J = n_row_p - 1). */
ftd_status ftd_set_rho(ftd_handle* h, const double* A, const double* B, const double* C);

/* TdmaBackendGPU::set_eps_constant (tdma_backend_gpu.cu:64-66). No-op for PASCAL. */
ftd_status ftd_set_eps_constant(ftd_handle* h, double eps_constant);

/* Extension without a reference counterpart: set eps_ itself (the same value
 * for every partition) instead of eps_constant / (m_p P)^2. */
ftd_status ftd_set_eps(ftd_handle* h, double eps);

/* Extension: change the stream later calls are ordered on. */
ftd_status ftd_set_stream(ftd_handle* h, void* stream);

/* Extension: when enabled, every solve is bracketed by HIP events and waits
 * for them, so that ftd_last_gpu_ms reports its device time -- what the
 * reference's *_profile variants do on every TdmaBackendGPU::solve
 * (tdma_backend_gpu.cu:35-46). Off by default (no host synchronisation). */
ftd_status ftd_set_profiling(ftd_handle* h, int enable);

/* Extension without a reference counterpart (FP32 storage): A, B, C are
 * float [n_row x n_sys], D is float64 and receives x in place. The
 * elimination runs in FP64 on the exactly widened coefficients, so x equals
 * the FP64 PaScaL solve of the stored system; HBM traffic is 28 B/row
 * instead of 40. Exact kinds only (PASCAL, or one partition). */
ftd_status ftd_solve_f32c(ftd_handle* h, const float* A, const float* B, const float* C, double* D);

/* Extension (FP32 storage of coefficients and right-hand side): A, B, C, D
 * are float, x (float64) goes to X; 24 B/row. Same numerics as
 * ftd_solve_f32c with D widened. */
ftd_status ftd_solve_f32(ftd_handle* h, const float* A, const float* B, const float* C,
                         const float* D, double* X);

/* Extension without a reference counterpart (FP32 storage, mixed-precision
 * refinement): A, B, C, D float, x (float64) to X. The partitions are
 * eliminated in FP32 arithmetic; the reduced system and the interior update
 * run in FP64. Then `sweeps` (0..FTD_MAX_REFINE_SWEEPS) refinement sweeps,
 * each one launch: the residual D - T x in FP64, rounded to FP32, solved
 * the same way, added to x. HBM traffic 24 + 32 * sweeps B/row. Exact kinds
 * only (PASCAL, or one partition). ftd_solve_f32 reaches the FP64 solve of
 * the stored system in one 24 B/row pass; this entry point is the
 * FP32-elimination alternative, for comparison and for callers that want
 * it. */
#define FTD_MAX_REFINE_SWEEPS 4
ftd_status ftd_solve_f32_refine(ftd_handle* h, const float* A, const float* B, const float* C,
                                const float* D, double* X, int sweeps);

/* FilteredTDMACUDA::last_J (filtered_tdma_cuda.cu:83-87), one value per
 * partition: J of the most recent filtered solve (blocking device->host
 * copy of n <= 2 * partitions ints; entries from P on are the imaginary
 * part's J of a complex solve). -1 where no filtered solve ran yet. */
ftd_status ftd_last_J(ftd_handle* h, int* J, int n);

/* TdmaBackendGPU::last_gpu_ms (tdma_backend_gpu.hpp:60-63): device time of
 * the most recent solve in ms when profiling is on, else 0. */
double ftd_last_gpu_ms(const ftd_handle* h);

/* TdmaBackendGPU::last_comm_ms (tdma_backend_gpu.hpp:54-57): always 0 --
 * the partitions of one GPU exchange nothing over a network. */
double ftd_last_comm_ms(const ftd_handle* h);

/* Extension (autotuner hook): force the register-resident fused-kernel
 * shape of later solves. rows_per_lane in {8,16,32,64} (0 = automatic),
 * max_threads in {256,512}, systems_per_tile a power of two <= 64
 * (0 = automatic), lanes_per_system (PaScaL sub-partitions, 0 = automatic).
 * rows_per_lane == -1 selects the general kernels (partitions walked in
 * global memory, in place) for every later solve; -2 keeps the automatic
 * choice but without the LDS-streamed kernels (which the automatic choice
 * takes when n_sys % 32 == 0, the arrays are 16-byte aligned and a variant
 * holds the partitions; they give the same bits as the fused kernels).
 * Returns FTD_ERR_UNSUPPORTED for combinations that are not compiled. */
ftd_status ftd_set_tuning(ftd_handle* h, int rows_per_lane, int max_threads,
                          int systems_per_tile, int lanes_per_system);

/* Extension (autotuner; no counterpart in the reference): times the
 * automatic choice, the automatic choice without the streamed kernels and
 * every compiled register-resident shape (rows per lane, workgroup size,
 * systems per workgroup, PaScaL sub-partitions) that holds a solve of the
 * handle's dimensions, each `reps` times on fresh device copies of A, C and
 * D (the caller's arrays are not written), and keeps the fastest as the
 * handle's tuning (what ftd_set_tuning sets). PaScaL keeps the handle's P
 * partitions when P > 1 and the filtered kinds always keep theirs, so
 * results stay bit-identical; with P == 1 the sub-partition count is tuned.
 * Allocates three n_row x n_sys arrays for the duration and waits on the
 * host. `report` (optional) receives a JSON text of every candidate's mean
 * time and the choice. */
ftd_status ftd_autotune(ftd_handle* h, double* A, double* B, double* C, const double* D,
                        int reps, char* report, size_t report_len);

/* Extension (learned truncation depth; no counterpart in the reference):
 * mode 1 makes the FILTERED kind (v1 sweeps) choose J per (system,
 * partition) with a small MLP (filtered_tdma_amd/learned_j/train.py) from
 * the partition's coupling bound, right-hand-side magnitudes and eps_,
 * instead of one cal_J per partition; the first correction term it drops is
 * checked against eps_ and the full correction (J = m-1) is used where it
 * exceeds it. One pass instead of v1's two (no global J) on partitions of
 * at most 64 rows; longer partitions run on the general kernels (the J
 * records kept in device memory). Mode 0 (default) is the reference's
 * cal_J. Real solves, periodic or not; complex ones on partitions of at
 * most 64 rows. */
ftd_status ftd_set_j_mode(ftd_handle* h, int mode);

/* Extension (no counterpart in the reference): the guard of the filtered
 * kinds. Besides the truncation by J, the reference's filtered solve drops
 * row 0's coupling to its own partition's last unknown (pack_reduced,
 * libs/filtered_tdma/cpu/filtered_tdma_kernels.hpp:178-189) and row m-1's
 * to its own row 0 (solve_interfaces, :197-218), an error of order
 * q^(m-1) max|x| that eps_ does not bound. With enable = 1 every kind runs
 * the v1 (full) sweeps, evaluates per (partition, system) those dropped
 * terms and the first term the truncation by J drops on each side, and
 * solves a system exactly (its 2P-row reduced system by Thomas, then the
 * full correction) where a dropped coupling exceeds eps_/4 times
 * min(1, |1 - cN_left a0'|) (the determinant of the interface 2x2 solve that
 * passes it on) or a truncated term exceeds eps_/4, so that x stays within
 * about eps_ of the exact solve. Off by default (the reference's bits).
 * Periodic systems (ftd_solve_cyclic) wrap the interfaces and solve the
 * periodic reduced system (Sherman-Morrison order, as the periodic PaScaL
 * reduced solve). Real solves of any partition length (partitions longer
 * than 64 rows run on the general kernels, with global scratch of 4 P n_sys
 * doubles); complex ones on partitions of at most 64 rows, others return
 * FTD_ERR_UNSUPPORTED. ftd_last_J_systems reports which systems it solved
 * exactly. */
ftd_status ftd_set_guard(ftd_handle* h, int enable);

/* Extension: the J used per (partition, system) by the last learned-mode or
 * guarded solve, index p * n_sys + i (n_sys doubled for complex solves),
 * with bit 16 set where the learned J's eps_ check fell back to the full
 * correction and bit 17 where the guard solved the system exactly.
 * Blocking copy of n values. */
ftd_status ftd_last_J_systems(ftd_handle* h, int* out, int64_t n);

/* Extension (complex right-hand sides; the reference's channel pressure
 * solve, apps/channel_cpu/PressureSolver.cpp:412-470, solves the real and
 * imaginary parts as two real systems with copies of one coefficient set):
 * Z is a complex128 [n_row x n_sys] array, (re, im) interleaved, overwritten
 * with x; A, B, C are the shared real coefficients [n_row x n_sys]. Both parts
 * run in one pass that reads A, B, C once (56 B per complex row instead of
 * 80 B), with results identical to two ftd_solve calls on the parts. For the
 * filtered kinds each part keeps its own J: ftd_last_J entries 0..P-1 (real)
 * and P..2P-1 (imaginary). Non-periodic; register-resident kernels only
 * (FTD_ERR_UNSUPPORTED where a partition exceeds 64 rows). */
ftd_status ftd_solve_complex(ftd_handle* h, double* A, double* B, double* C, double* Z);

/* Extension (uneven batches; no counterpart in the reference, whose batches
 * share one n_row): n_systems independent systems, system s occupying
 * elements row_offsets[s] .. row_offsets[s] + n_rows[s] - 1 of each of A, B,
 * C, D (device arrays; row_offsets and n_rows are HOST arrays). Exact solve:
 * a system of <= 48 rows is plain Thomas, a longer one PaScaL over
 * partitions of at most 48 rows inside one 128-lane workgroup (48: the
 * default chunk, see ftd_set_batch_chunk). The partitions after a system's
 * first start on 16-element boundaries of the arrays (128-byte lines of
 * FP64 data), the first runs to 16 rows short of a full chunk past the
 * system's first boundary and a last one under 16 rows takes 16 from the
 * one before; a system too short for that split, or needing more than 128
 * partitions, takes ceil(n_rows/48) para_range partitions. All in one
 * persistent launch (systems above 6144 rows, and a batch that is one 1-row
 * system at offset 0, in a second, lane-per-system launch that uses C as
 * scratch). D is overwritten with x; A and B are read only. No element past
 * the last row of the highest system is read, so the arrays may end there.
 * The plan is made on the host once per layout and kept on the handle; a
 * call with a new layout waits for the handle's stream first. The handle's
 * kind and sizes are not used. */
ftd_status ftd_solve_batched(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                             int n_systems, double* A, double* B, double* C, double* D);

/* Extension (uneven batches with FP32 storage; no counterpart in the
 * reference): the layout and partitioning of ftd_solve_batched, with A, B,
 * C, D float and the elimination in FP64 on their exactly widened values;
 * x (FP64) goes to X at the same offsets, A..D are read only. Results equal
 * ftd_solve_batched on the widened arrays, bit for bit. FTD_ERR_UNSUPPORTED
 * for a system above 6144 rows or a batch spanning fewer than 4 elements. */
ftd_status ftd_solve_batched_f32(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                                 int n_systems, const float* A, const float* B, const float* C,
                                 const float* D, double* X);

/* Extension (uneven batches): rows per chunk of the batched plans, 32 or 48
 * (0: the default, 48). A system of up to `rows` rows is one lane's Thomas,
 * a longer one PaScaL with ceil(n_rows / rows) partitions (the path string
 * names the chunk), up to 128 * rows rows. Takes effect at the next batched
 * solve (which plans the layout again). */
ftd_status ftd_set_batch_chunk(ftd_handle* h, int rows);

/* Extension (uneven batches): times ftd_solve_batched on this layout with
 * each chunk length (reps timed solves each, after one untimed), on device
 * copies of C and D (the caller's arrays are left as they were), and keeps
 * the faster one on the handle (as ftd_set_batch_chunk). A JSON report of
 * the candidates goes to `report` (if not NULL). Waits for the handle's
 * stream; allocates two arrays of the layout's extent for the copies. */
ftd_status ftd_autotune_batched(ftd_handle* h, const int64_t* row_offsets, const int* n_rows,
                                int n_systems, double* A, double* B, const double* C,
                                const double* D, int reps, char* report, size_t report_len);

/* Size of the static device scratch array behind ftd_solve_static, per GPU. */
#define FTD_STATIC_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one solve with no handle and no allocation, equal bit
 * for bit to ftd_create(kind, n_sys, n_row, partitions, eps_constant,
 * stream) + ftd_set_rho(A, B, C) + ftd_solve (cyclic == 0) or
 * ftd_solve_cyclic (cyclic != 0) + ftd_destroy on the same arrays. The
 * solve's scratch (column-0 bounds, J per partition, per-tile maxima,
 * boundary values, reduced rows) lives in one static device array of
 * FTD_STATIC_SCRATCH_BYTES per GPU, held from the library's load to the
 * process's end; a solve needing more returns FTD_ERR_UNSUPPORTED with the
 * size in ftd_last_error_string. Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race
 * and give undefined results. Nothing in the array needs to start at zero
 * (every value a kernel reads is written earlier in the same call).
 * Ordered on `stream` (a hipStream_t, NULL = the null stream); no host
 * synchronisation. */
ftd_status ftd_solve_static(int kind, int n_sys, int n_row, int partitions, double eps_constant,
                            int cyclic, double* A, double* B, double* C, double* D,
                            void* stream);

/* Extension (diagnostics): the path the last solve took, as a short
 * static string, e.g. "pascal_stream M=64 Q=8 R=16",
 * "pascal_fused M=32 T=256 S=16 Q=16" or "filtered_v1_general". */
const char* ftd_last_path(const ftd_handle* h);

/* Text of the last error reported by any ftd_* call in this thread. */
const char* ftd_last_error_string(void);

/* ------------------------------------------------------------------------
 * Heat ADI harness: the reference's caller of the path, SolveTheta
 * (apps/heat_gpu/solve_theta.cu:391-650; apps/heat_cpu/solve_theta.cpp:18-334)
 * with MPISubdomain's ghost cells (apps/heat_cpu/mpi_subdomain.cpp:90-236)
 * and main.cpp's L2 error (:59-81), for the manufactured solution
 * sin(pi x)sin(pi y)sin(pi z)exp(-3 pi^2 t) + cos(pi x)cos(pi y)cos(pi z) on
 * [-1,1]^3. The field lives on the device with one ghost layer per face;
 * parts[3] = the reference's (npx, npy, npz) become partitions of the x, y
 * and z sweeps, each one batched solve of kind `kind`.
 * ------------------------------------------------------------------------ */
typedef struct ftd_heat ftd_heat;

/* GlobalParams + MPISubdomain set-up and the initial field (main.cpp:20-49);
 * nx, ny, nz interior cells, periodic[3] flags (NULL = Dirichlet), step dt,
 * stream NULL = a stream owned by the handle. */
ftd_status ftd_heat_create(ftd_heat** out, int nx, int ny, int nz, const int* parts, int kind,
                           const int* periodic, double dt, void* stream);
ftd_status ftd_heat_destroy(ftd_heat* h);

/* re-initialise the field to t = 0 (mpi_subdomain.cpp:164-173) */
ftd_status ftd_heat_init(ftd_heat* h);

/* nsteps time steps (solve_theta.cpp:92-334); use_graph != 0 replays one
 * step captured into a hipGraph. Waits for the last step. */
ftd_status ftd_heat_step(ftd_heat* h, int nsteps, int use_graph);

/* Extension: fused sweeps on (default) or off. Where a sweep's coefficients
 * are the stencil's (Dirichlet walls: the PaScaL kind, the filtered kinds in
 * directions of one caller partition, and the filtered kinds' y and z sweeps
 * over 2..16 caller partitions of at most 64 rows from the second step on), the
 * fused sweep generates them in the kernel and reads and writes the field in
 * place (16 B per row instead of build + solve + scatter); the field is
 * bit-identical either way. A captured step is dropped and re-captured by
 * the next graph step. */
ftd_status ftd_heat_set_fused(ftd_heat* h, int enable);

/* Extension (tuning hook): the fused sweep of direction dir (0 = x, 1 = y,
 * 2 = z) holds rows_per_lane (16, 32 or 64; 0 = automatic) rows per lane --
 * with one caller partition in that direction this sets the number of
 * internal PaScaL partitions -- and systems_per_tile systems per workgroup
 * (a power of two, 0 = automatic). The materialized sweep is tuned to the
 * same partitioning, so ftd_heat_set_fused(h, 0) keeps giving the same
 * field. A captured step is dropped. */
ftd_status ftd_heat_set_plan(ftd_heat* h, int dir, int rows_per_lane, int systems_per_tile);

/* the field with ghosts, (nz+2)(ny+2)(nx+2) doubles, x fastest, to host memory */
ftd_status ftd_heat_field(ftd_heat* h, double* host_out);

/* L2 error at t = steps * dt, as main.cpp:59-81 prints it */
ftd_status ftd_heat_l2(ftd_heat* h, double* l2);

/* device time of the last step of the most recent ftd_heat_step, in ms */
double ftd_heat_last_step_ms(const ftd_heat* h);

/* text of the last error reported by an ftd_heat_* call in this thread */
const char* ftd_heat_last_error_string(void);

/* Extension (diagnostics): how the last enqueued step ran each sweep, e.g.
 * "z:fused M=16 Q=4 S=64; y:fused M=64 Q=1 S=64; x:fused M=64 Q=1 S=16" --
 * fused (coefficients generated in the kernel), filtered_fused (the filtered
 * kinds over the caller's z partitions) or materialized (build, solve,
 * scatter). Owned by h; valid until its next step. */
const char* ftd_heat_last_plan(const ftd_heat* h);

/* ------------------------------------------------------------------------
 * Channel pressure Poisson solve: the reference's PressureSolver::solve
 * (apps/channel_cpu/PressureSolver.cpp:493-533) on one GPU. x and y are
 * periodic and uniform, z is wall-normal on the tanh-stretched grid of
 * Grid.cpp:16-30 (gamma_z; mirrored ghost centres). The wall-normal
 * systems are solved by an ftd_handle of `kind` and `partitions` (the
 * reference uses PaScaL) on the complex spectra in place.
 * ------------------------------------------------------------------------ */
typedef struct ftd_poisson ftd_poisson;

/* Metrics, modified wavenumbers and the wall-normal coefficients
 * (PressureSolver.cpp:36-104, 400-456) are built once; two hipFFT plans.
 * nx even. `stream` is a hipStream_t (NULL = the null stream), as for
 * ftd_create. */
ftd_status ftd_poisson_create(ftd_poisson** out, int nx, int ny, int nz, double Lx, double Ly,
                              double Lz, double gamma_z, int kind, int partitions,
                              double eps_constant, void* stream);
ftd_status ftd_poisson_destroy(ftd_poisson* p);

/* The solver of the wall-normal systems (n_sys = (nx/2+1) ny complex, n_row
 * = nz), e.g. for ftd_set_j_mode or ftd_set_tuning. Owned by p. */
ftd_handle* ftd_poisson_solver(ftd_poisson* p);

/* One pressure step (solve, :493-533): rhs = div(u*)/dt, FFT, wall-normal
 * solves with the (0,0) mode pinned, inverse FFT scaled by 1/(nx ny), then
 * P += dp and u -= dt grad dp (project_, :470-490). U, V, W, P are device
 * fields with one ghost layer, (nx+2)(ny+2)(nz+2) doubles, x fastest;
 * U(i) is the face between cells i-1 and i. Ordered on the handle's stream;
 * with profiling on it waits for the step and records its device time. */
ftd_status ftd_poisson_solve(ftd_poisson* p, double* U, double* V, double* W, double* P,
                             double dt);

/* The Poisson solve alone: dp [nz][ny][nx] from rhs [nz][ny][nx] (device
 * arrays, no ghosts). Waits. */
ftd_status ftd_poisson_solve_rhs(ftd_poisson* p, const double* rhs, double* dp);

/* time each ftd_poisson_solve (waits for it) or not (default) */
ftd_status ftd_poisson_set_profiling(ftd_poisson* p, int enable);

/* device time of the last ftd_poisson_solve with profiling on, in ms */
double ftd_poisson_last_ms(const ftd_poisson* p);

/* text of the last error reported by an ftd_poisson_* call in this thread */
const char* ftd_poisson_last_error_string(void);

#ifdef __cplusplus
}
#endif

#endif /* FTDMA_H */
