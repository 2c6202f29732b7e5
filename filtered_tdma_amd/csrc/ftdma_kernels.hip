// ftdma_kernels.hip -- gfx950 kernels of the batched tridiagonal engine.
//
// Arrays are row-major [n_row x n_sys] (row j of system i at j*n_sys + i), so
// the lanes of a wavefront that hold consecutive systems read consecutive
// addresses of one row: every row load of a wave is one or more contiguous
// 128..512-byte segments.
//
// Three families of kernels:
//
//  * streamed kernels (k_pascal_stream, k_filtered_stream): a resident grid
//    walks tiles of 32 systems x Q partitions; the rows of a tile come into
//    an LDS ring by direct global->LDS DMA, R rows ahead of their use, and
//    each lane keeps its partition's elimination state in registers. Used
//    where a compiled (M, Q) variant holds the partitions and n_sys % 32 ==
//    0 with 16-byte aligned arrays.
//
//  * fused kernels (k_pascal_fused, k_filtered_fused): one lane per
//    (system, partition). A lane keeps the rows of its partition in
//    registers (M rows x 3 doubles), the partitions of S systems meet in LDS
//    for the reduced solve, and D is written once. Any n_sys and alignment.
//
//   Both move the algorithmic minimum: read A,B,C,D once, write x once
//   (40 B/row).
//
//  * general kernels (k_gen_*): one lane per (system, partition) walking its
//    rows in global memory, in place in A, C, D as the reference contract
//    allows (A and C clobbered). Used where a partition has more rows than
//    the fused kernels hold in registers.
//
// Arithmetic follows the reference statement by statement (reciprocal then
// multiply where it multiplies, division where it divides); the library is
// built with -ffp-contract=off so that no product is fused into a sum, which
// makes the results bit-identical to the CPU restatement for the same
// partitioning.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <map>
#include <mutex>
#include <utility>

#include "ftdma_kernels.h"
#include "ftdma_mlpj_weights.h"

namespace ftd {

// para_range over [0, n) in P parts (libs/pascal_tdma/cpu/para_range.cpp:4-9)
__host__ __device__ inline void part_of(int n, int P, int p, int& s, int& m) {
    int base = n / P, rem = n % P;
    s = p * base + (p < rem ? p : rem);
    m = base + (p < rem ? 1 : 0);
}

// float -> int truncation that saturates instead of being undefined, NaN -> 0
__device__ inline long long trunc_sat(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647LL;
    if (x <= -2147483648.0) return -2147483648LL;
    return (long long)(int)x;
}

// k_cal_J_rhs_bound's formula (filtered_tdma_kernels.cuh:245-259) with the
// clamp to [0, n_row-1]
__device__ inline int cal_J_gpu(double rho, double max_b, double eps, int n_row) {
    if (rho == 0.0 || rho >= 0.5) return n_row - 1;
    double lambda_p = (1.0 + sqrt(1.0 - 4.0 * rho * rho)) * 0.5;
    double q = rho / lambda_p;
    double Bv = q * (2.0 + q) / ((1.0 - q) * (1.0 - 2.0 * rho)) * max_b;
    long long J = trunc_sat(log(eps / Bv) / log(q)) + 1;
    if (J > n_row - 1) J = n_row - 1;
    if (J < 0) J = 0;
    return (int)J;
}

// max over aligned groups of `width` lanes (width a power of two <= 64)
__device__ inline double group_max(double v, int width) {
    for (int off = width >> 1; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// ---------------------------------------------------------------------------
// column-0 spectral bounds (filtered_tdma_kernels.cuh:46-67)
// ---------------------------------------------------------------------------
__global__ void k_set_rho(const double* __restrict__ A, const double* __restrict__ B,
                          const double* __restrict__ C, int n_sys, int n_row,
                          double* __restrict__ A_rho, double* __restrict__ C_rho) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n_row) return;
    if (k < n_row) {
        double bk = B[(size_t)k * n_sys];
        A_rho[k] = (bk != 0.0) ? fabs(A[(size_t)k * n_sys] / bk) : 0.0;
        C_rho[k] = (bk != 0.0) ? fabs(C[(size_t)k * n_sys] / bk) : 0.0;
    } else {
        A_rho[k] = 0.0;
        C_rho[k] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// J machinery: per-tile partial maxima -> per-partition J
// ---------------------------------------------------------------------------

// max over systems of |D[row 0 of p]| and |D[row m_p-1 of p]|, one block of
// 256 systems per (tile, partition); partial[tile * P + p]
// (csh = 1: complex right-hand sides, even and odd columns reduce separately
// into partial[(parity * tiles + tile) * P + p])
__global__ __launch_bounds__(256) void k_boundary_maxabs(const double* __restrict__ D, int n_sys,
                                                         int n_row, int P, int csh,
                                                         double* __restrict__ partial) {
    __shared__ double red[2][4];
    const int p = blockIdx.y;
    int s, m;
    part_of(n_row, P, p, s, m);
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n_sys) {
        double a = fabs(D[(size_t)s * n_sys + i]);
        double b = fabs(D[(size_t)(s + m - 1) * n_sys + i]);
        v = fmax(a, b);
    }
    const int odd = csh ? (i & 1) : 0;
    double v0 = odd ? 0.0 : v, v1 = odd ? v : 0.0;  // |D| >= 0: 0 is neutral
    v0 = group_max(v0, 64);
    v1 = group_max(v1, 64);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = v0;
        red[1][threadIdx.x >> 6] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 1 + csh) {
        const double* r = red[threadIdx.x];
        double w = fmax(fmax(r[0], r[1]), fmax(r[2], r[3]));
        partial[((size_t)threadIdx.x * gridDim.x + blockIdx.x) * P + p] = w;
    }
}

// J_p from rho (column-0 bounds of the partition's rows 1..m_p-1), the
// reduced max_b and eps_p; one block per partition
__global__ __launch_bounds__(256) void k_compute_J(const double* __restrict__ partial, int n_tiles,
                                                   int P, int n_row,
                                                   const double* __restrict__ A_rho,
                                                   const double* __restrict__ C_rho,
                                                   double eps_value, int eps_direct,
                                                   int* __restrict__ J_out) {
    __shared__ double red_b[4], red_r[4];
    // block p + k * P: partition p of parity k (complex right-hand sides)
    const int p = blockIdx.x % P;
    const double* part = partial + (size_t)(blockIdx.x / P) * n_tiles * P;
    int s, m;
    part_of(n_row, P, p, s, m);
    double vb = 0.0, vr = 0.0;
    for (int t = threadIdx.x; t < n_tiles; t += 256) vb = fmax(vb, part[(size_t)t * P + p]);
    for (int k = 1 + threadIdx.x; k < m; k += 256)
        vr = fmax(vr, fmax(fabs(A_rho[s + k]), fabs(C_rho[s + k])));
    vb = group_max(vb, 64);
    vr = group_max(vr, 64);
    if ((threadIdx.x & 63) == 0) {
        red_b[threadIdx.x >> 6] = vb;
        red_r[threadIdx.x >> 6] = vr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mb = fmax(fmax(red_b[0], red_b[1]), fmax(red_b[2], red_b[3]));
        double rho = fmax(fmax(red_r[0], red_r[1]), fmax(red_r[2], red_r[3]));
        double Nl = (double)(m * P);
        double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        J_out[blockIdx.x] = cal_J_gpu(rho, mb, eps, m);
    }
}

// Row loads of the fused kernels: row j of the lane's partition, clamped to
// the partition's last row so that the load never depends on a branch (the
// compiler may then issue it PF rows ahead of its use). Rows past the count
// reload the last row, which is in cache, and are never used. Values are
// widened to double on load (exact for float storage). The coefficients are
// read at cbase / cld, the right-hand side at base / ld: the same place for
// real systems; for complex right-hand sides stored as interleaved (re, im)
// pairs, lanes 2k and 2k+1 solve the real and imaginary parts of system k
// and read the same coefficient column k (CSH = 1).
// Streamed operands are touched once: FTD_NONTEMPORAL=1 marks the row loads
// and the solution stores nontemporal (no reuse expected in the caches).
#ifndef FTD_NONTEMPORAL
#define FTD_NONTEMPORAL 1
#endif
#if FTD_NONTEMPORAL
#define FTD_LD(p) __builtin_nontemporal_load(p)
#define FTD_ST(v, p) __builtin_nontemporal_store((v), (p))
#else
#define FTD_LD(p) (*(p))
#define FTD_ST(v, p) (*(p) = (v))
#endif

#define FTD_LOAD_ROW(j)                                                      \
    do {                                                                     \
        const int jj_ = (j) < cnt ? (j) : cnt - 1;                           \
        const size_t o_ = base + (size_t)jj_ * ld;                           \
        const size_t oc_ = cbase + (size_t)jj_ * cld;                        \
        ra[(j)] = (double)FTD_LD(A + oc_);                                   \
        rb[(j)] = (double)FTD_LD(B + oc_);                                   \
        rc[(j)] = (double)FTD_LD(C + oc_);                                   \
        rd[(j)] = (double)FTD_LD(Din + o_);                                  \
    } while (0)

// ---------------------------------------------------------------------------
// plain Thomas with the whole system in registers (n_row <= M): the
// reference's single-rank path (tdma_many_kernel, tdma_local_cuda.cu:32-64),
// row 0 divided by b. One lane per system, D read once and written once.
// ---------------------------------------------------------------------------
template <int M, int PF, int CSH>
__global__ __launch_bounds__(256) void k_thomas_fused(const double* __restrict__ A,
                                                      const double* __restrict__ B,
                                                      const double* __restrict__ C,
                                                      double* __restrict__ D, int n_sys,
                                                      int n_row) {
    const int i_raw = blockIdx.x * blockDim.x + threadIdx.x;
    if (i_raw >= n_sys) return;
    const int i = i_raw;
    const int cnt = n_row;
    const double* Din = D;
    const size_t ld = (size_t)n_sys;
    const size_t base = (size_t)i;
    const size_t cld = ld >> CSH;
    const size_t cbase = (size_t)(i >> CSH);
    double c[M], d[M];
    double ra[M], rb[M], rc[M], rd[M];
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_LOAD_ROW(j);
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_LOAD_ROW(j + PF);
        if (j < cnt) {
            if (j == 0) {
                d[0] = rd[0] / rb[0];
                c[0] = rc[0] / rb[0];
            } else {
                const double r = 1.0 / (rb[j] - ra[j] * c[j - 1]);
                d[j] = r * (rd[j] - ra[j] * d[j - 1]);
                c[j] = r * rc[j];
            }
        }
    }
#pragma unroll
    for (int j = M - 2; j >= 0; --j)
        if (j <= cnt - 2) d[j] = d[j] - c[j] * d[j + 1];
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (j < cnt) FTD_ST(d[j], D + base + (size_t)j * ld);
}

// ---------------------------------------------------------------------------
// Sherman-Morrison periodic Thomas with the whole system in registers
// (n_row <= M): the reference's single-rank periodic path
// (tdma_cyclic_many_kernel, tdma_local_cuda.cu:67-128).
// ---------------------------------------------------------------------------
template <int M, int PF>
__global__ __launch_bounds__(256) void k_thomas_cyclic_fused(const double* __restrict__ A,
                                                             const double* __restrict__ B,
                                                             const double* __restrict__ C,
                                                             double* __restrict__ D, int n_sys,
                                                             int n_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sys) return;
    const int cnt = n_row;
    const double* Din = D;
    const size_t ld = (size_t)n_sys;
    const size_t base = (size_t)i;
    const size_t cld = ld, cbase = base;
    double c[M], d[M], e[M];
    double ra[M], rb[M], rc[M], rd[M];
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_LOAD_ROW(j);
    double a0 = 0.0, b0 = 0.0, c0 = 0.0, d0 = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_LOAD_ROW(j + PF);
        if (j < cnt) {
            if (j == 0) {
                a0 = ra[0];
                b0 = rb[0];
                c0 = rc[0];
                d0 = rd[0];
            } else if (j == 1) {
                const double inv_b = 1.0 / rb[1];
                // E[1] = -A[1] (E[n-1] = -C[n-1] joins at the last row)
                d[1] = rd[1] * inv_b;
                e[1] = (cnt - 1 == 1 ? -rc[1] : -ra[1]) * inv_b;
                c[1] = rc[1] * inv_b;
            } else {
                const double r = 1.0 / (rb[j] - ra[j] * c[j - 1]);
                const double ej = (j == cnt - 1) ? -rc[j] : 0.0;
                d[j] = r * (rd[j] - ra[j] * d[j - 1]);
                e[j] = r * (ej - ra[j] * e[j - 1]);
                c[j] = r * rc[j];
            }
        }
    }
#pragma unroll
    for (int j = M - 2; j >= 1; --j) {
        if (j <= cnt - 2) {
            d[j] = d[j] - c[j] * d[j + 1];
            e[j] = e[j] - c[j] * e[j + 1];
        }
    }
    double dN = 0.0, eN = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (j == cnt - 1) {
            dN = d[j];
            eN = e[j];
        }
    const double x0 = (d0 - a0 * dN - c0 * d[1]) / (b0 + a0 * eN + c0 * e[1]);
    FTD_ST(x0, D + base);
#pragma unroll
    for (int j = 1; j < M; ++j)
        if (j < cnt) FTD_ST(d[j] + x0 * e[j], D + base + (size_t)j * ld);
}

// Reduced system of NR rows (B == 1) in LDS, row k of this lane's system at
// index k * S, solved by Thomas in place (x to Dr). 4 rows at a time: their
// values are read into registers first and c', d' are carried in registers,
// so the serial chain is the arithmetic, not an LDS round trip per row.
__device__ __forceinline__ void reduced_thomas_lds(double* Ar, double* Cr, double* Dr, int NR,
                                                   int S) {
    double cp = Cr[0], dp = Dr[0];
    int k = 1;
    for (; k + 4 <= NR; k += 4) {
        double ak[4], ck[4], dk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ak[u] = Ar[(k + u) * S];
            ck[u] = Cr[(k + u) * S];
            dk[u] = Dr[(k + u) * S];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double inv = 1.0 / (1.0 - ak[u] * cp);
            dp = inv * (dk[u] - ak[u] * dp);
            cp = inv * ck[u];
            Dr[(k + u) * S] = dp;
            Cr[(k + u) * S] = cp;
        }
    }
    for (; k < NR; ++k) {
        const double ak = Ar[k * S];
        const double inv = 1.0 / (1.0 - ak * cp);
        dp = inv * (Dr[k * S] - ak * dp);
        cp = inv * Cr[k * S];
        Dr[k * S] = dp;
        Cr[k * S] = cp;
    }
    double dn = dp;
    k = NR - 2;
    for (; k >= 3; k -= 4) {
        double ck[4], dk[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ck[u] = Cr[(k - u) * S];
            dk[u] = Dr[(k - u) * S];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            dn = dk[u] - ck[u] * dn;
            Dr[(k - u) * S] = dn;
        }
    }
    for (; k >= 0; --k) {
        dn = Dr[k * S] - Cr[k * S] * dn;
        Dr[k * S] = dn;
    }
}

// ---------------------------------------------------------------------------
// fused PaScaL: modified Thomas per (system, sub-partition) in registers,
// reduced 2Q-row system in LDS, interior update, single store of x.
// Restates modified_thomas_kernel / tdma_many_kernel / tdma_cyclic_many_kernel
// / update_solution_kernel (libs/pascal_tdma/gpu/tdma_local_cuda.cu:67-224)
// for Q partitions of one GPU.
// ---------------------------------------------------------------------------
// TC: storage type of A, B, C; TD: of the right-hand side Din. The solution
// goes to X (float64), which may be Din itself when TD is double.
template <int M, int MAXT, bool CYCLIC, int PF, int WPE, typename TC, typename TD, int CSH = 0>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pascal_fused(
    const TC* __restrict__ A, const TC* __restrict__ B, const TC* __restrict__ C, const TD* Din,
    double* X, int n_sys, int n_row, int Q, int lgS) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int sl = t & (S - 1);
    const int q = t >> lgS;
    const int i_raw = blockIdx.x * S + sl;
    const bool live = i_raw < n_sys;
    // dead lanes shadow the last system (of their parity when CSH = 1)
    const int i = live ? i_raw : n_sys - (1 << CSH) + (i_raw & ((1 << CSH) - 1));
    int r0, cnt;
    part_of(n_row, Q, q, r0, cnt);
    const size_t ld = (size_t)n_sys;
    const size_t base = (size_t)r0 * ld + (size_t)i;
    const size_t cld = ld >> CSH;
    const size_t cbase = (size_t)r0 * cld + (size_t)(i >> CSH);

    double a[M], c[M], d[M];
    double ra[M], rb[M], rc[M], rd[M];  // raw rows; PF of them live at a time
    double aN = 0.0, cN = 0.0, dN = 0.0;

#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_LOAD_ROW(j);

    // forward sweep: rows 0,1 divided by b, then modified elimination
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_LOAD_ROW(j + PF);
        if (j < cnt) {
            const double aj = ra[j], bj = rb[j], cj = rc[j], dj = rd[j];
            if (j < 2) {
                a[j] = aj / bj;
                c[j] = cj / bj;
                d[j] = dj / bj;
            } else {
                const double inv = 1.0 / (bj - aj * c[j - 1]);
                d[j] = inv * (dj - aj * d[j - 1]);
                c[j] = inv * cj;
                a[j] = -inv * aj * a[j - 1];
            }
            if (j == cnt - 1) {
                aN = a[j];
                cN = c[j];
                dN = d[j];
            }
        }
    }
    // backward sweep rows cnt-3 .. 1
#pragma unroll
    for (int j = M - 3; j >= 1; --j) {
        if (j <= cnt - 3) {
            const double cj = c[j];
            d[j] = d[j] - cj * d[j + 1];
            a[j] = a[j] - cj * a[j + 1];
            c[j] = -cj * c[j + 1];
        }
    }
    // row-0 2x2 reduction
    const double r = 1.0 / (1.0 - a[1] * c[0]);
    const double d0 = r * (d[0] - c[0] * d[1]);
    const double a0 = r * a[0];
    const double c0 = -r * c[0] * c[1];

    const int NR = 2 * Q;
    double* rA = lds;
    double* rC = rA + NR * S;
    double* rD = rC + NR * S;
    double* rE = rD + NR * S;  // cyclic only
    rA[(2 * q) * S + sl] = a0;
    rC[(2 * q) * S + sl] = c0;
    rD[(2 * q) * S + sl] = d0;
    rA[(2 * q + 1) * S + sl] = aN;
    rC[(2 * q + 1) * S + sl] = cN;
    rD[(2 * q + 1) * S + sl] = dN;
    __syncthreads();

    // reduced solve, B == 1: one lane per system
    if (q == 0) {
        double* Ar = rA + sl;
        double* Cr = rC + sl;
        double* Dr = rD + sl;
        if (!CYCLIC) {
            reduced_thomas_lds(Ar, Cr, Dr, NR, S);
        } else {
            double* Er = rE + sl;
            for (int k = 0; k < NR; ++k) Er[k * S] = 0.0;
            Er[1 * S] = -Ar[1 * S];
            Er[(NR - 1) * S] = -Cr[(NR - 1) * S];
            for (int k = 2; k < NR; ++k) {
                const double inv = 1.0 / (1.0 - Ar[k * S] * Cr[(k - 1) * S]);
                Dr[k * S] = inv * (Dr[k * S] - Ar[k * S] * Dr[(k - 1) * S]);
                Er[k * S] = inv * (Er[k * S] - Ar[k * S] * Er[(k - 1) * S]);
                Cr[k * S] = inv * Cr[k * S];
            }
            for (int k = NR - 2; k >= 1; --k) {
                Dr[k * S] = Dr[k * S] - Cr[k * S] * Dr[(k + 1) * S];
                Er[k * S] = Er[k * S] - Cr[k * S] * Er[(k + 1) * S];
            }
            Dr[0] = (Dr[0] - Ar[0] * Dr[(NR - 1) * S] - Cr[0] * Dr[1 * S]) /
                    (1.0 + Ar[0] * Er[(NR - 1) * S] + Cr[0] * Er[1 * S]);
            // x_k = y_k + x_0 e_k: each partition forms its own two rows below
        }
    }
    __syncthreads();

    double x0 = rD[(2 * q) * S + sl];
    double xN = rD[(2 * q + 1) * S + sl];
    if constexpr (CYCLIC) {
        const double x0r = rD[sl];
        const double y0 = x0 + x0r * rE[(2 * q) * S + sl];
        xN = xN + x0r * rE[(2 * q + 1) * S + sl];
        x0 = q == 0 ? x0 : y0;  // row 0 of the reduced system is x_0 itself
    }
    if (live) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (j < cnt) {
                double x;
                if (j == 0)
                    x = x0;
                else if (j == cnt - 1)
                    x = xN;
                else
                    x = d[j] - (a[j] * x0 + c[j] * xN);
                FTD_ST(x, X + base + (size_t)j * ld);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP32 elimination with FP64 refinement (no counterpart in the reference;
// the mixed-precision mode of the FP32-storage path). A, B, C, D are float
// [n_row x n_sys], x float64 in X. One launch is one solve of the partitions
// of k_pascal_fused (same lanes, same LDS layout):
//   modified Thomas per partition in FP32 arithmetic (rows 0, 1 divided by b),
//   the 2Q-row reduced system widened to FP64 and solved in FP64, the interior
//   update in FP64 on the widened FP32 values.
// RES = false: the right-hand side is D, and X receives the solution.
// RES = true : one refinement sweep. The right-hand side is the residual
//   r_j = D_j - (A_j x_{j-1} + B_j x_j + C_j x_{j+1}) of the current X,
//   formed in FP64 on the widened operands (row 0 without its A term, row
//   n_row - 1 without its C term) and rounded to FP32; X_j += delta_j.
//   X is updated in place: a workgroup holds every partition of its systems,
//   reads all of their x (the halo rows r0 - 1 and r0 + cnt included) before
//   the barrier that publishes the reduced rows, and writes after it.
// ---------------------------------------------------------------------------
template <int M, int MAXT, int PF, int WPE, bool RES>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pascal_ir(
    const float* __restrict__ A, const float* __restrict__ B, const float* __restrict__ C,
    const float* __restrict__ D, double* X, int n_sys, int n_row, int Q, int lgS) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int sl = t & (S - 1);
    const int q = t >> lgS;
    const int i_raw = blockIdx.x * S + sl;
    const bool live = i_raw < n_sys;
    const int i = live ? i_raw : n_sys - 1;  // dead lanes shadow the last system
    int r0, cnt;
    part_of(n_row, Q, q, r0, cnt);
    const size_t ld = (size_t)n_sys;
    const size_t base = (size_t)r0 * ld + (size_t)i;

    float a[M], c[M], d[M];
    float ra[M], rb[M], rc[M], rd[M];  // raw rows; PF of them live at a time
    double xv[M];                      // RES: x of the partition's rows
    // RES: x of the rows just outside the partition (0 past the system's ends)
    double xlo = 0.0, xhi = 0.0;
    if (RES) {
        if (r0 > 0) xlo = X[base - ld];
        if (r0 + cnt < n_row) xhi = X[base + (size_t)cnt * ld];
    }
    float aN = 0.0f, cN = 0.0f, dN = 0.0f;

#define FTD_IR_LOAD(j)                                        \
    do {                                                      \
        const int jj_ = (j) < cnt ? (j) : cnt - 1;            \
        const size_t o_ = base + (size_t)jj_ * ld;            \
        ra[(j)] = FTD_LD(A + o_);                             \
        rb[(j)] = FTD_LD(B + o_);                             \
        rc[(j)] = FTD_LD(C + o_);                             \
        rd[(j)] = FTD_LD(D + o_);                             \
        if (RES) xv[(j)] = X[o_];                             \
    } while (0)

#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_IR_LOAD(j);

#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_IR_LOAD(j + PF);
        if (j < cnt) {
            const float aj = ra[j], bj = rb[j], cj = rc[j];
            float dj;
            if (RES) {
                const int g = r0 + j;  // global row
                const double xm = j == 0 ? xlo : xv[j > 0 ? j - 1 : 0];
                const double xp = j == cnt - 1 ? xhi : xv[j + 1 < M ? j + 1 : j];
                double tt = (double)bj * xv[j];
                if (g > 0) tt = (double)aj * xm + tt;
                if (g < n_row - 1) tt = tt + (double)cj * xp;
                dj = (float)((double)rd[j] - tt);
            } else {
                dj = rd[j];
            }
            if (j < 2) {
                a[j] = aj / bj;
                c[j] = cj / bj;
                d[j] = dj / bj;
            } else {
                const float inv = 1.0f / (bj - aj * c[j - 1]);
                d[j] = inv * (dj - aj * d[j - 1]);
                c[j] = inv * cj;
                a[j] = -inv * aj * a[j - 1];
            }
            if (j == cnt - 1) {
                aN = a[j];
                cN = c[j];
                dN = d[j];
            }
        }
    }
#undef FTD_IR_LOAD
#pragma unroll
    for (int j = M - 3; j >= 1; --j) {
        if (j <= cnt - 3) {
            const float cj = c[j];
            d[j] = d[j] - cj * d[j + 1];
            a[j] = a[j] - cj * a[j + 1];
            c[j] = -cj * c[j + 1];
        }
    }
    const float r = 1.0f / (1.0f - a[1] * c[0]);
    const float d0 = r * (d[0] - c[0] * d[1]);
    const float a0 = r * a[0];
    const float c0 = -r * c[0] * c[1];

    const int NR = 2 * Q;
    double* rA = lds;
    double* rC = rA + NR * S;
    double* rD = rC + NR * S;
    rA[(2 * q) * S + sl] = (double)a0;
    rC[(2 * q) * S + sl] = (double)c0;
    rD[(2 * q) * S + sl] = (double)d0;
    rA[(2 * q + 1) * S + sl] = (double)aN;
    rC[(2 * q + 1) * S + sl] = (double)cN;
    rD[(2 * q + 1) * S + sl] = (double)dN;
    __syncthreads();
    if (q == 0) reduced_thomas_lds(rA + sl, rC + sl, rD + sl, NR, S);
    __syncthreads();

    const double x0 = rD[(2 * q) * S + sl];
    const double xN = rD[(2 * q + 1) * S + sl];
    if (live) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (j < cnt) {
                double x;
                if (j == 0)
                    x = x0;
                else if (j == cnt - 1)
                    x = xN;
                else
                    x = (double)d[j] - ((double)a[j] * x0 + (double)c[j] * xN);
                if (RES) x = xv[j] + x;
                FTD_ST(x, X + base + (size_t)j * ld);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-streamed PaScaL (persistent): the arithmetic of k_pascal_fused, with
// the row operands brought in by direct global->LDS loads
// (global_load_lds_dwordx4) into a per-wave ring of R row-steps instead of
// through VGPRs. The ring keeps R rows of every lane in flight without
// spending registers on them (the M x 3 elimination state already fills the
// register file at one wave per SIMD), and a block walks tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... so that the first R rows of the next
// tile stream in while the current tile does its backward sweep, reduced
// solve and stores.
//
// Shape: S = 32 systems per tile, Q partitions, two partitions per wave
// (lanes 0..31 -> partition 2w, lanes 32..63 -> 2w+1), Q/2 waves per block.
// Requirements checked by the launcher: n_sys % 32 == 0 (no partial tiles),
// 16-byte aligned arrays, every partition has 2 <= m_p <= M rows.
//
// Ring slot layout (one wave, one row-step, 2 KB): two 1 KB blocks, each
// filled by one dwordx4 DMA of the wave (lane l -> bytes 16l..16l+15):
//   block 0: [A of 2w][A of 2w+1][B of 2w][B of 2w+1]  (32 doubles each)
//   block 1: [C of 2w][C of 2w+1][D of 2w][D of 2w+1]
// so the reading lane l finds a, b, c, d at doubles l, 64 + l of the blocks.
// Each wave loads exactly the rows it consumes: the ring needs no barrier,
// only the wave's own vmcnt.
//
// vmcnt bookkeeping. It relies only on a wave's loads completing in issue
// order among themselves (the counter's decrements for loads are in order);
// stores may complete in any order relative to them. Per row-step two DMA
// ops. At the wait for row j the ops issued after that row's DMAs are
//   j >= R, or the first tile : the DMAs of the next R-1 row-steps -> 2(R-1):
//     while any of row j's is pending all of those are too, so vmcnt(2(R-1))
//     means row j has landed, whatever stores are also in flight;
//   j <  R, later tiles       : no wait. Those rows were requested during the
//     previous tile's last R row-steps, and the previous tile waited for
//     vmcnt(0) after its reduced solve, before its stores.
// The last tile re-fetches its own rows into consumed slots so that the
// pattern stays uniform; the kernel drains vmcnt before it ends.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef const __attribute__((address_space(3))) double* lds_cdptr_t;
typedef double ftd_dbl2 __attribute__((ext_vector_type(2)));

// The four operands of one row-step from a ring slot (a, b at bytes 0 and
// 512 of block 0, c, d at 1024 and 1536), waited for before returning, so the
// slot may be refilled right after. Written as inline asm so that the
// compiler's own wait insertion (which cannot tell ring slots apart and would
// drain every DMA in flight) stays out of the ring; the waits the data needs
// are the explicit vmcnt waits above.
__device__ __forceinline__ void ring_read(lds_cdptr_t p, double& a, double& b, double& c,
                                          double& d) {
    ftd_dbl2 ab, cd;
    asm volatile(
        "ds_read2st64_b64 %0, %2 offset1:1\n\t"
        "ds_read2st64_b64 %1, %2 offset0:2 offset1:3\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(ab), "=&v"(cd)
        : "v"(p)
        : "memory");
    a = ab.x;
    b = ab.y;
    c = cd.x;
    d = cd.y;
}

// x0 and xN of lane l from a 1 KB block holding four 32-double rows
// [x0 of 2w][xN of 2w][x0 of 2w+1][xN of 2w+1] (p points at x0 of the lane)
__device__ __forceinline__ void xin_read(lds_cdptr_t p, double& x0, double& xN) {
    ftd_dbl2 v;
    asm volatile(
        "ds_read2_b64 %0, %1 offset1:32\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v)
        : "v"(p)
        : "memory");
    x0 = v.x;
    xN = v.y;
}

// workgroup barrier for LDS traffic only: the LDS writes are complete, the
// DMAs in flight are left alone (__syncthreads would wait for them as well)
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

#ifndef FTD_DMA_AUX
#define FTD_DMA_AUX 2  // nontemporal: the rows are read once
#endif

// The ring of one wave and its bookkeeping.
template <int M, int Q, int R>
struct StreamRing {
    static constexpr int S = 32;
    double (*slots)[2][128];  // this wave's R slots
    const double* src0;       // DMA sources of this lane (row 0 of its sub-block)
    const double* src1;
    size_t ld;

    __device__ __forceinline__ void init(double (*wave_slots)[2][128], const double* A,
                                         const double* B, const double* C, const double* D,
                                         int n_sys, int n_row, int w, int l, int& dcnt) {
        slots = wave_slots;
        ld = (size_t)n_sys;
        const int g = l >> 4;
        int dr0;
        part_of(n_row, Q, 2 * w + (g & 1), dr0, dcnt);
        src0 = (g < 2 ? A : B) + (size_t)dr0 * ld + 2 * (l & 15);
        src1 = (g < 2 ? C : D) + (size_t)dr0 * ld + 2 * (l & 15);
    }
    // row j (clamped to the sub-block's last row) of tile T into slot
    __device__ __forceinline__ void dma(int T, int j, int slot, int dcnt) const {
        const int jj = j < dcnt ? j : dcnt - 1;
        const size_t o = (size_t)jj * ld + (size_t)T * S;
        __builtin_amdgcn_global_load_lds((const void*)(src0 + o), (lds_ptr_t)&slots[slot][0][0], 16,
                                         0, FTD_DMA_AUX);
        __builtin_amdgcn_global_load_lds((const void*)(src1 + o), (lds_ptr_t)&slots[slot][1][0], 16,
                                         0, FTD_DMA_AUX);
    }
    __device__ __forceinline__ void prologue(int tile, int dcnt) const {
#pragma unroll
        for (int j = 0; j < R; ++j) dma(tile, j, j, dcnt);
    }
    // operands of row j of the current tile; refills the slot with row j + R
    // (of this tile, else of tile `ntile`)
    __device__ __forceinline__ void take(int j, bool first, int tile, int ntile, int dcnt, int l,
                                         double& a, double& b, double& c, double& d) const {
        // rows j < R of a later tile landed before the previous tile's stores
        if (j >= R || first) wait_vmcnt<2 * (R - 1)>();
        ring_read((lds_cdptr_t)&slots[j % R][0][l], a, b, c, d);
        if (j + R < M)
            dma(tile, j + R, j % R, dcnt);
        else
            dma(ntile, j + R - M, j % R, dcnt);
    }
};

template <int M, int Q, int R, bool CYCLIC, int WPE>
__global__ __launch_bounds__(Q * 32) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pascal_stream(
    const double* __restrict__ A, const double* __restrict__ B, const double* __restrict__ C,
    double* D, int n_sys, int n_row, int n_tiles) {
    constexpr int S = 32, NW = Q / 2, NR = 2 * Q;
    static_assert(Q % 2 == 0 && M % R == 0 && R >= 2, "stream shape");
    __shared__ __attribute__((aligned(16))) double ring[NW][R][2][128];
    __shared__ double red[CYCLIC ? 4 : 3][NR][S];

    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int l = t & 63;
    const int q = 2 * w + (l >> 5);
    const int sl = l & 31;
    const size_t ld = (size_t)n_sys;
    StreamRing<M, Q, R> rg;
    int dcnt0;
    rg.init(ring[w], A, B, C, D, n_sys, n_row, w, l, dcnt0);
    int r0, cnt0;
    part_of(n_row, Q, q, r0, cnt0);

    int tile = blockIdx.x;
    rg.prologue(tile, dcnt0);
    bool first = true;
    for (; tile < n_tiles; tile += gridDim.x) {
        // opaque per tile: keeps the row clamps from being hoisted out of the
        // tile loop as M live values
        int cnt = cnt0, dcnt = dcnt0;
        asm volatile("" : "+v"(cnt), "+v"(dcnt));
        const int next = tile + (int)gridDim.x;
        const int ntile = next < n_tiles ? next : tile;
        const size_t base = (size_t)r0 * ld + (size_t)tile * S + sl;

        double a[M], c[M], d[M];
        double aN = 0.0, cN = 0.0, dN = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double aj, bj, cj, dj;
            rg.take(j, first, tile, ntile, dcnt, l, aj, bj, cj, dj);
            if (j < cnt) {
                if (j < 2) {
                    a[j] = aj / bj;
                    c[j] = cj / bj;
                    d[j] = dj / bj;
                } else {
                    const double inv = 1.0 / (bj - aj * c[j - 1]);
                    d[j] = inv * (dj - aj * d[j - 1]);
                    c[j] = inv * cj;
                    a[j] = -inv * aj * a[j - 1];
                }
                if (j == cnt - 1) {
                    aN = a[j];
                    cN = c[j];
                    dN = d[j];
                }
            }
        }
#pragma unroll
        for (int j = M - 3; j >= 1; --j) {
            if (j <= cnt - 3) {
                const double cj = c[j];
                d[j] = d[j] - cj * d[j + 1];
                a[j] = a[j] - cj * a[j + 1];
                c[j] = -cj * c[j + 1];
            }
        }
        const double r = 1.0 / (1.0 - a[1] * c[0]);
        const double d0 = r * (d[0] - c[0] * d[1]);
        const double a0 = r * a[0];
        const double c0 = -r * c[0] * c[1];

        lds_barrier();  // the previous tile's readers of red[] are done
        red[0][2 * q][sl] = a0;
        red[1][2 * q][sl] = c0;
        red[2][2 * q][sl] = d0;
        red[0][2 * q + 1][sl] = aN;
        red[1][2 * q + 1][sl] = cN;
        red[2][2 * q + 1][sl] = dN;
        lds_barrier();
        if (q == 0) {
            // reduced system (B == 1), one lane per system; the running
            // c, d (and e) stay in registers, the eliminated rows go back to
            // LDS for the backward pass
            double* Ar = &red[0][0][sl];
            double* Cr = &red[1][0][sl];
            double* Dr = &red[2][0][sl];
            if (!CYCLIC) {
                double cp = Cr[0], dp = Dr[0];
#pragma unroll 2
                for (int k = 1; k < NR; ++k) {
                    const double ak = Ar[k * S];
                    const double inv = 1.0 / (1.0 - ak * cp);
                    dp = inv * (Dr[k * S] - ak * dp);
                    cp = inv * Cr[k * S];
                    Dr[k * S] = dp;
                    Cr[k * S] = cp;
                }
                double dn = dp;
#pragma unroll 2
                for (int k = NR - 2; k >= 0; --k) {
                    dn = Dr[k * S] - Cr[k * S] * dn;
                    Dr[k * S] = dn;
                }
            } else {
                double* Er = &red[3][0][sl];
                const double e_last = -Cr[(NR - 1) * S];
                double cp = Cr[1 * S], dp = Dr[1 * S], ep = -Ar[1 * S];
                Er[1 * S] = ep;
#pragma unroll 2
                for (int k = 2; k < NR; ++k) {
                    const double ak = Ar[k * S];
                    const double inv = 1.0 / (1.0 - ak * cp);
                    const double ek = (k == NR - 1) ? e_last : 0.0;
                    dp = inv * (Dr[k * S] - ak * dp);
                    ep = inv * (ek - ak * ep);
                    cp = inv * Cr[k * S];
                    Dr[k * S] = dp;
                    Er[k * S] = ep;
                    Cr[k * S] = cp;
                }
                double dn = dp, en = ep;
#pragma unroll 2
                for (int k = NR - 2; k >= 1; --k) {
                    const double ck = Cr[k * S];
                    dn = Dr[k * S] - ck * dn;
                    en = Er[k * S] - ck * en;
                    Dr[k * S] = dn;
                    Er[k * S] = en;
                }
                const double x0r = (Dr[0] - Ar[0] * dp - Cr[0] * Dr[1 * S]) /
                                   (1.0 + Ar[0] * ep + Cr[0] * Er[1 * S]);
                Dr[0] = x0r;
                Er[0] = 0.0;
                // x_k = y_k + x_0 e_k: each partition forms its own two rows below
            }
        }
        lds_barrier();
        // the next tile's first R rows have landed before this tile's stores
        // (requested during the last R row-steps, so this costs no wait)
        wait_vmcnt<0>();
        double x0 = red[2][2 * q][sl];
        double xN = red[2][2 * q + 1][sl];
        if constexpr (CYCLIC) {
            const double x0r = red[2][0][sl];
            const double y0 = x0 + x0r * red[3][2 * q][sl];
            xN = xN + x0r * red[3][2 * q + 1][sl];
            x0 = q == 0 ? x0 : y0;  // row 0 of the reduced system is x_0 itself
        }
        // rows past m_p store row m_p - 1 (x_N) again
        double* const xlast = D + base + (size_t)(cnt - 1) * ld;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double x;
            if (j == 0)
                x = x0;
            else if (j >= cnt - 1)
                x = xN;
            else
                x = d[j] - (a[j] * x0 + c[j] * xN);
            FTD_ST(x, j < cnt - 1 ? D + base + (size_t)j * ld : xlast);
        }
        first = false;
    }
    wait_vmcnt<0>();  // no DMA into LDS outlives the wave
}

// ---------------------------------------------------------------------------
// LDS-streamed PaScaL with FP32 storage of A, B, C and D and the FP64 x in
// its own array (a kernel of its own rather than a storage-type parameter of
// k_pascal_stream, so that the FP64 kernel's register allocation and wait
// counts stay as audited): the arithmetic of k_pascal_fused<..., float, float> (FP64
// elimination on the exactly widened values, same partitions, same bits),
// over a ring whose row-step is ONE DMA: a slot of 1 KB holds
// [A | B | C | D] x [partition 2w | 2w+1] x 32 systems as floats, so lane l
// loads 16 bytes of array l/16, partition 2w + (l/8)%2, systems 4(l%8) ..
// 4(l%8)+3. 24 algorithmic bytes per row.
// ---------------------------------------------------------------------------
typedef const __attribute__((address_space(3))) float* lds_cfptr_t;
typedef float ftd_flt2 __attribute__((ext_vector_type(2)));

// a, b, c, d of lane l of a slot (floats at bytes 4l, 256 + 4l, 512 + 4l,
// 768 + 4l), widened; waited for before returning
__device__ __forceinline__ void ring_read_f32(lds_cfptr_t p, double& a, double& b, double& c,
                                              double& d) {
    ftd_flt2 ab, cd;
    asm volatile(
        "ds_read2st64_b32 %0, %2 offset1:1\n\t"
        "ds_read2st64_b32 %1, %2 offset0:2 offset1:3\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(ab), "=&v"(cd)
        : "v"(p)
        : "memory");
    a = (double)ab.x;
    b = (double)ab.y;
    c = (double)cd.x;
    d = (double)cd.y;
}

template <int M, int Q, int R>
struct StreamRingF32 {
    static constexpr int S = 32;
    float (*slots)[256];  // this wave's R slots
    const float* src;     // DMA source of this lane (row 0 of its sub-block)
    size_t ld;

    __device__ __forceinline__ void init(float (*wave_slots)[256], const float* A, const float* B,
                                         const float* C, const float* D, int n_sys, int n_row,
                                         int w, int l, int& dcnt) {
        slots = wave_slots;
        ld = (size_t)n_sys;
        const int arr = l >> 4;
        int dr0;
        part_of(n_row, Q, 2 * w + ((l >> 3) & 1), dr0, dcnt);
        const float* base = arr == 0 ? A : arr == 1 ? B : arr == 2 ? C : D;
        src = base + (size_t)dr0 * ld + 4 * (l & 7);
    }
    __device__ __forceinline__ void dma(int T, int j, int slot, int dcnt) const {
        const int jj = j < dcnt ? j : dcnt - 1;
        const size_t o = (size_t)jj * ld + (size_t)T * S;
        __builtin_amdgcn_global_load_lds((const void*)(src + o), (lds_ptr_t)&slots[slot][0], 16, 0,
                                         FTD_DMA_AUX);
    }
    __device__ __forceinline__ void prologue(int tile, int dcnt) const {
#pragma unroll
        for (int j = 0; j < R; ++j) dma(tile, j, j, dcnt);
    }
    // one DMA per row-step: R - 1 in flight after this row's; rows j < R of
    // a later tile landed before the previous tile's stores
    __device__ __forceinline__ void take(int j, bool first, int tile, int ntile, int dcnt, int l,
                                         double& a, double& b, double& c, double& d) const {
        if (j >= R || first) wait_vmcnt<R - 1>();
        ring_read_f32((lds_cfptr_t)&slots[j % R][l], a, b, c, d);
        if (j + R < M)
            dma(tile, j + R, j % R, dcnt);
        else
            dma(ntile, j + R - M, j % R, dcnt);
    }
};

template <int M, int Q, int R, int WPE>
__global__ __launch_bounds__(Q * 32) __attribute__((amdgpu_waves_per_eu(WPE))) void k_pascal_stream_f32(
    const float* __restrict__ A, const float* __restrict__ B, const float* __restrict__ C,
    const float* __restrict__ Din, double* __restrict__ X, int n_sys, int n_row, int n_tiles) {
    constexpr int S = 32, NW = Q / 2, NR = 2 * Q;
    static_assert(Q % 2 == 0 && M % R == 0 && R >= 2, "stream shape");
    __shared__ __attribute__((aligned(16))) float ring[NW][R][256];
    __shared__ double red[3][NR][S];

    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int l = t & 63;
    const int q = 2 * w + (l >> 5);
    const int sl = l & 31;
    const size_t ld = (size_t)n_sys;
    StreamRingF32<M, Q, R> rg;
    int dcnt0;
    rg.init(ring[w], A, B, C, Din, n_sys, n_row, w, l, dcnt0);
    int r0, cnt0;
    part_of(n_row, Q, q, r0, cnt0);

    int tile = blockIdx.x;
    rg.prologue(tile, dcnt0);
    bool first = true;
    for (; tile < n_tiles; tile += gridDim.x) {
        int cnt = cnt0, dcnt = dcnt0;
        asm volatile("" : "+v"(cnt), "+v"(dcnt));
        const int next = tile + (int)gridDim.x;
        const int ntile = next < n_tiles ? next : tile;
        const size_t base = (size_t)r0 * ld + (size_t)tile * S + sl;

        double a[M], c[M], d[M];
        double aN = 0.0, cN = 0.0, dN = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double aj, bj, cj, dj;
            rg.take(j, first, tile, ntile, dcnt, l, aj, bj, cj, dj);
            if (j < cnt) {
                if (j < 2) {
                    a[j] = aj / bj;
                    c[j] = cj / bj;
                    d[j] = dj / bj;
                } else {
                    const double inv = 1.0 / (bj - aj * c[j - 1]);
                    d[j] = inv * (dj - aj * d[j - 1]);
                    c[j] = inv * cj;
                    a[j] = -inv * aj * a[j - 1];
                }
                if (j == cnt - 1) {
                    aN = a[j];
                    cN = c[j];
                    dN = d[j];
                }
            }
        }
#pragma unroll
        for (int j = M - 3; j >= 1; --j) {
            if (j <= cnt - 3) {
                const double cj = c[j];
                d[j] = d[j] - cj * d[j + 1];
                a[j] = a[j] - cj * a[j + 1];
                c[j] = -cj * c[j + 1];
            }
        }
        const double r = 1.0 / (1.0 - a[1] * c[0]);
        const double d0 = r * (d[0] - c[0] * d[1]);
        const double a0 = r * a[0];
        const double c0 = -r * c[0] * c[1];

        lds_barrier();  // the previous tile's readers of red[] are done
        red[0][2 * q][sl] = a0;
        red[1][2 * q][sl] = c0;
        red[2][2 * q][sl] = d0;
        red[0][2 * q + 1][sl] = aN;
        red[1][2 * q + 1][sl] = cN;
        red[2][2 * q + 1][sl] = dN;
        lds_barrier();
        if (q == 0) {
            double* Ar = &red[0][0][sl];
            double* Cr = &red[1][0][sl];
            double* Dr = &red[2][0][sl];
            double cp = Cr[0], dp = Dr[0];
#pragma unroll 2
            for (int k = 1; k < NR; ++k) {
                const double ak = Ar[k * S];
                const double inv = 1.0 / (1.0 - ak * cp);
                dp = inv * (Dr[k * S] - ak * dp);
                cp = inv * Cr[k * S];
                Dr[k * S] = dp;
                Cr[k * S] = cp;
            }
            double dn = dp;
#pragma unroll 2
            for (int k = NR - 2; k >= 0; --k) {
                dn = Dr[k * S] - Cr[k * S] * dn;
                Dr[k * S] = dn;
            }
        }
        lds_barrier();
        // the next tile's first R rows have landed before this tile's stores
        wait_vmcnt<0>();
        const double x0 = red[2][2 * q][sl];
        const double xN = red[2][2 * q + 1][sl];
        // rows past m_p store row m_p - 1 (x_N) again
        double* const xlast = X + base + (size_t)(cnt - 1) * ld;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double x;
            if (j == 0)
                x = x0;
            else if (j >= cnt - 1)
                x = xN;
            else
                x = d[j] - (a[j] * x0 + c[j] * xN);
            FTD_ST(x, j < cnt - 1 ? X + base + (size_t)j * ld : xlast);
        }
        first = false;
    }
    wait_vmcnt<0>();  // no DMA into LDS outlives the wave
}

// ---------------------------------------------------------------------------
// fused Filtered v1/v2 (GPU semantics, filtered_tdma_kernels.cuh:69-303):
// one lane per (system, partition) with its m_p <= M rows in registers.
//   PASS 0: v2 (J_p known) -> x stored.
//   PASS 1: v1 first pass  -> x0, xN of every (partition, system) to xbuf,
//           per-tile max of |x0|,|xN| to partial (J needs all systems).
//   PASS 2: v1 second pass -> recompute the sweeps, take x0, xN from xbuf and
//           J_p, store x.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Learned J (PASS 3 of the filtered kernel): one J per (system, partition)
// from a 5 -> 32 -> 32 -> 1 MLP (weights from learned_j/train.py), with the
// first dropped correction term checked against eps_ afterwards.
// ---------------------------------------------------------------------------

// piecewise-linear log2 of |x| from its float64 bit pattern: (e - 1) +
// (2 f - 1) for |x| = f 2^e, f in [0.5, 1); -1100 for 0 (exact, so every
// machine computes the same feature)
__device__ inline float lg_feature(double x) {
    x = fabs(x);
    if (x == 0.0) return -1100.0f;
    int e;
    const double f = frexp(x, &e);
    return (float)((double)(e - 1) + (2.0 * f - 1.0));
}

// the predicted J (before clamping); the operation order is fixed (the CPU
// restatement follows it), products and sums rounded separately
__device__ inline int mlp_J(float f0, float f1, float f2, float f3, float f4) {
    // 5 -> H -> H -> 1, ReLU, FP32; every multiply-add is one explicit fma
    // (correctly rounded), and every sum runs in index order
    constexpr int H = kMlpH;
    float acc[H];
#pragma unroll
    for (int l = 0; l < H; ++l) acc[l] = kMlpB2[l];
    for (int k = 0; k < H; ++k) {
        // unit k's weights are contiguous (b1, W1 column, W2 row): three
        // wide scalar loads and one wait per unit
        const float* w = kMlpK + k * (8 + H);
        float h = w[0];
        h = fmaf(w[1], f0, h);
        h = fmaf(w[2], f1, h);
        h = fmaf(w[3], f2, h);
        h = fmaf(w[4], f3, h);
        h = fmaf(w[5], f4, h);
        h = h > 0.0f ? h : 0.0f;
#pragma unroll
        for (int l = 0; l < H; ++l) acc[l] = fmaf(w[8 + l], h, acc[l]);
    }
    float out = kMlpB3[0];
#pragma unroll
    for (int l = 0; l < H; ++l) out = fmaf(kMlpW3[l], acc[l] > 0.0f ? acc[l] : 0.0f, out);
    const float cj = ceilf(out);
    // saturate before converting (the prediction of a wild input is still
    // clamped to [0, m-1] by the caller)
    const float lim = 1.0e6f;
    return (int)(cj > lim ? lim : (cj < -lim ? -lim : cj)) + kMlpMargin;
}

// ---------------------------------------------------------------------------
// Guard (GRD = 1; no counterpart in the reference): besides the truncation
// by J, the filtered solve drops row 0's coupling to its own partition's last
// unknown (c0', pack_reduced) and row m-1's to its own row 0 (a'_{m-1},
// k_solve_both). With the interface values x0, xN the guard evaluates those
// terms and the first truncated term of each side; where any exceeds eps_/4
// in any partition of a system, the system is solved exactly: its 2P-row
// reduced system (c0' kept) by Thomas in LDS, then the full correction.
// Sweeps are v1's (full) under the guard, for every kind.
// ---------------------------------------------------------------------------

// the guard's exact reduced solve of one system (column sl of the [P][S]
// LDS arrays): rows 2q = row 0 of partition q (a0', c0', d0' in gA, gC0,
// gD0), 2q + 1 = row m-1 (aN, cN, dN in gAN, gCN, gDN); x into gD0 / gDN
__device__ inline void guard_reduced_thomas(double* gA, double* gC0, double* gD0, double* gAN,
                                            double* gCN, double* gDN, int P, int S, int sl) {
    double cp = gC0[sl], dp = gD0[sl];
    for (int q = 0; q < P; ++q) {
        const int o = q * S + sl;
        if (q > 0) {
            const double ak = gA[o];
            const double inv = 1.0 / (1.0 - ak * cp);
            dp = inv * (gD0[o] - ak * dp);
            cp = inv * gC0[o];
            gC0[o] = cp;
            gD0[o] = dp;
        }
        const double ak = gAN[o];
        const double inv = 1.0 / (1.0 - ak * cp);
        dp = inv * (gDN[o] - ak * dp);
        cp = inv * gCN[o];
        gCN[o] = cp;
        gDN[o] = dp;
    }
    double xn = dp;
    for (int q = P - 1; q >= 0; --q) {
        const int o = q * S + sl;
        if (q < P - 1) {
            xn = gDN[o] - gCN[o] * xn;
            gDN[o] = xn;
        }
        xn = gD0[o] - gC0[o] * xn;
        gD0[o] = xn;
    }
}

// the same for periodic systems, whose reduced system is periodic too (row 0
// of partition 0 couples to row m-1 of partition P-1): Thomas over rows
// 1..2P-1 with a second right-hand side e (-a at row 1, -c at row 2P-1),
// row 0 from both, then x_k = y_k + x_0 e_k -- the Sherman-Morrison order of
// the periodic reduced solve, B == 1. Row k's e takes its a slot once the a
// has been read (row 0's a stays).
__device__ inline void guard_reduced_cyclic(double* gA, double* gC0, double* gD0, double* gAN,
                                            double* gCN, double* gDN, int P, int S, int sl) {
    const int n = 2 * P;
    double cp = gCN[sl], dp = gDN[sl], e = -gAN[sl];  // row 1: row m-1 of partition 0
    gAN[sl] = e;
    for (int k = 2; k < n; ++k) {
        const int o = (k >> 1) * S + sl;
        double* const pa = (k & 1) ? gAN : gA;
        double* const pc = (k & 1) ? gCN : gC0;
        double* const pd = (k & 1) ? gDN : gD0;
        const double ak = pa[o], ck = pc[o];
        const double e0 = k == n - 1 ? -ck : 0.0;
        const double r = 1.0 / (1.0 - ak * cp);
        dp = r * (pd[o] - ak * dp);
        e = r * (e0 - ak * e);
        cp = r * ck;
        pa[o] = e;
        pc[o] = cp;
        pd[o] = dp;
    }
    const double yl = dp, el = e;  // row n-1
    for (int k = n - 2; k >= 1; --k) {
        const int o = (k >> 1) * S + sl;
        double* const pa = (k & 1) ? gAN : gA;
        double* const pd = (k & 1) ? gDN : gD0;
        const double ck = ((k & 1) ? gCN : gC0)[o];
        dp = pd[o] - ck * dp;
        e = pa[o] - ck * e;
        pd[o] = dp;
        pa[o] = e;
    }
    const double x0 = (gD0[sl] - gA[sl] * yl - gC0[sl] * dp) / (1.0 + gA[sl] * el + gC0[sl] * e);
    gD0[sl] = x0;
    for (int k = 1; k < n; ++k) {
        const int o = (k >> 1) * S + sl;
        double* const pd = (k & 1) ? gDN : gD0;
        pd[o] = pd[o] + x0 * ((k & 1) ? gAN : gA)[o];
    }
}

// the guard's test of one (system, partition): the dropped couplings against
// eps_/4 min(1, |det|) of the interface they reach (c0' xN the left one,
// a'_{m-1} x0 the right one: the interface 2x2 solve divides them by its
// determinant 1 - cN_left a0'), the first truncated terms for J against eps_/4
template <int M>
__device__ __forceinline__ bool guard_flag(const double (&a)[M], const double (&c)[M], int m,
                                           int J, double c0p, double aN, double x0, double xN,
                                           double eps, double detL, double detR) {
    const double thr = eps * 0.25;
    const double thrL = thr * fmin(1.0, fabs(detL));
    const double thrR = thr * fmin(1.0, fabs(detR));
    const int kL = (J < m - 2 ? J : m - 2) + 1;
    const int kR = ((m - 1) - J > 1 ? (m - 1) - J : 1) - 1;
    double akL = 0.0, ckR = 0.0;
#pragma unroll
    for (int j = 1; j < M; ++j) {
        if (j == kL) akL = a[j];
        if (j == kR) ckR = c[j];
    }
    const double ta = kL <= m - 2 ? fabs(akL * x0) : 0.0;
    const double tc = kR >= 1 ? fabs(ckR * xN) : 0.0;
    return fabs(c0p * xN) > thrL || fabs(aN * x0) > thrR || ta > thr || tc > thr;
}

// The learned J of one (system, partition), checked (PASS 3): the network's
// Jl where the first term it drops on each side, |a'_{Jl+1} x0| and
// |c'_{m-2-Jl} xN|, is at most eps_; else the full correction, m - 1.
// Returns J | (Jl failed) << 16. dropped_exceeds picks the two coefficients
// by a select chain and multiplies once (products per row would be shared
// with the final pass and held live until then).
template <int M>
__device__ __forceinline__ bool dropped_exceeds(const double (&a)[M], const double (&c)[M], int m,
                                                int J, double x0, double xN, double eps) {
    const int kL = (J < m - 2 ? J : m - 2) + 1;              // first a-term J drops
    const int kR = ((m - 1) - J > 1 ? (m - 1) - J : 1) - 1;  // first c-term J drops
    double akL = 0.0, ckR = 0.0;
#pragma unroll
    for (int j = 1; j < M; ++j) {
        if (j == kL) akL = a[j];
        if (j == kR) ckR = c[j];
    }
    const double ta = kL <= m - 2 ? fabs(akL * x0) : 0.0;
    const double tc = kR >= 1 ? fabs(ckR * xN) : 0.0;
    return ta > eps || tc > eps;
}

template <int M>
__device__ __forceinline__ int learned_J_checked(const double (&a)[M], const double (&c)[M], int m,
                                                 int Jl, double x0, double xN, double eps) {
    return dropped_exceeds<M>(a, c, m, Jl, x0, xN, eps) ? (m - 1) | (1 << 16) : Jl;
}

template <int M, int MAXT, int VERSION, int PASS, int PF, int WPE, int CSH, int GRD>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_filtered_fused(const double* __restrict__ A,
                                                         const double* __restrict__ B,
                                                         const double* __restrict__ C,
                                                         double* __restrict__ D, int n_sys,
                                                         int n_row, int P, int lgS, int cyclic,
                                                         const int* __restrict__ Jp,
                                                         double* __restrict__ xbuf,
                                                         double* __restrict__ partial,
                                                         double eps_value, int eps_direct,
                                                         int* __restrict__ Jsys) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int sl = t & (S - 1);
    const int p = t >> lgS;
    const int i_raw = blockIdx.x * S + sl;
    const bool live = i_raw < n_sys;
    // dead lanes shadow the last system (of their parity when CSH = 1)
    const int i = live ? i_raw : n_sys - (1 << CSH) + (i_raw & ((1 << CSH) - 1));
    int r0, m;
    part_of(n_row, P, p, r0, m);
    const size_t ld = (size_t)n_sys;
    const size_t base = (size_t)r0 * ld + (size_t)i;
    const size_t cld = ld >> CSH;
    const size_t cbase = (size_t)r0 * cld + (size_t)(i >> CSH);
    // complex right-hand sides: the real and imaginary parts keep their own J
    // (the reference solves them as two real systems)
    const int par = CSH ? (i & 1) : 0;
    static_assert(!GRD || VERSION == 1, "the guard runs the v1 sweeps");
    const int J = (VERSION == 2 || PASS == 2 || (GRD && PASS == 0)) ? Jp[par * P + p] : m;

    double a[M], c[M], d[M];
    double ra[M], rb[M], rc[M], rd[M];  // raw rows; PF of them live at a time
    double aN = 0.0, cN = 0.0, dN = 0.0;
    double b1 = 1.0;
    const int cnt = m;
    const double* Din = D;
    // PASS 3 features: max |a/b|, |c/b| over rows 1..m-1, raw d of rows 0
    // and m-1, max |d|
    double f_rho = 0.0, f_d0 = 0.0, f_dN = 0.0, f_dmax = 0.0;

#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_LOAD_ROW(j);

    // forward sweep (k_fwd_pass_v1 / k_fwd_pass_v2)
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_LOAD_ROW(j + PF);
        if (j < m) {
            const double aj = ra[j], bj = rb[j], cj = rc[j], dj = rd[j];
            if (PASS == 3) {
                // max(|a/b|, |c/b|) as max(|a|, |c|) / |b|: the same value
                // (rounding is monotonic in the numerator), one division
                if (j >= 1) f_rho = fmax(f_rho, fmax(fabs(aj), fabs(cj)) / fabs(bj));
                if (j == 0) f_d0 = dj;
                if (j == m - 1) f_dN = dj;
                f_dmax = fmax(f_dmax, fabs(dj));
            }
            if (j < 2) {
                const double r = 1.0 / bj;
                a[j] = aj * r;
                c[j] = cj * r;
                d[j] = dj * r;
                if (j == 1) b1 = bj;
            } else if (VERSION == 1 || j <= J) {
                const double inv = 1.0 / (bj - aj * c[j - 1]);
                d[j] = inv * (dj - aj * d[j - 1]);
                c[j] = inv * cj;
                a[j] = -inv * aj * a[j - 1];
            } else {
                const double inv = 1.0 / (bj - aj * c[j - 1]);
                d[j] = inv * (dj - aj * d[j - 1]);
                c[j] = inv * cj;
                a[j] = aj;  // left raw, as the skip-A loop leaves A[j]
            }
            if (VERSION == 2 && j == 1 && J == 0) {
                // k_fwd_pass_v2 with J == 0 starts its skip-A loop at row 1 and
                // eliminates the normalised row 1 against itself
                const double inv = 1.0 / (b1 - a[1] * c[1]);
                d[1] = inv * (d[1] - a[1] * d[1]);
                c[1] = inv * c[1];
            }
            if (j == m - 1) {
                aN = a[j];
                cN = c[j];
                dN = d[j];
            }
        }
    }
    // backward sweep (k_bwd_pass_v1 / k_bwd_pass_v2)
    const int lo = (m - 2) - J;
#pragma unroll
    for (int j = M - 3; j >= 1; --j) {
        if (j <= m - 3) {
            const double cj = c[j];
            d[j] = d[j] - cj * d[j + 1];
            if (VERSION == 1) {
                a[j] = a[j] - cj * a[j + 1];
                c[j] = -cj * c[j + 1];
            } else {
                if (J >= 2 && j < J) {
                    // seed A[J]; with J == m-1 the first update (j = m-3)
                    // reads A[m-1], as k_bwd_pass_v2 does
                    const double an = (J == m - 1 && j == m - 3) ? aN : a[j + 1];
                    a[j] = a[j] - cj * an;
                }
                if (j > lo) c[j] = -cj * c[j + 1];
            }
        }
    }
    // pack_reduced: row 0 keeps d0, a0 (c0 coupling dropped)
    const double rr = 1.0 / (1.0 - a[1] * c[0]);
    const double d0p = rr * (d[0] - c[0] * d[1]);
    const double a0p = rr * a[0];

    double x0, xN;
    double detL = 1.0, detR = 1.0;  // the interfaces' 2x2 determinants (the guard's)
    // neighbour exchange through LDS: [C|D] of row m-1 to the right, [A|D]
    // of row 0 to the left (filtered_tdma_cuda.cu:93-101)
    double* sC = lds;
    double* sDr = sC + P * S;
    double* sA = sDr + P * S;
    double* sDl = sA + P * S;
    const bool has_left = cyclic || p > 0;
    const bool has_right = cyclic || p < P - 1;
    const int pl = (p + P - 1) % P, pr = (p + 1) % P;
    if (PASS != 2 || GRD) {
        sC[p * S + sl] = cN;
        sDr[p * S + sl] = dN;
        sA[p * S + sl] = a0p;
        sDl[p * S + sl] = d0p;
        __syncthreads();
    }
    // both interface solves on every lane, then selected (no lane-divergent
    // branch while the rows are held in registers; as k_filtered_stream)
    if (PASS == 2) {
        x0 = xbuf[(size_t)(2 * p) * ld + i];
        xN = xbuf[(size_t)(2 * p + 1) * ld + i];
        if (GRD) {
            const double dl = 1.0 - sC[pl * S + sl] * a0p;
            const double dr = 1.0 - sA[pr * S + sl] * cN;
            detL = has_left ? dl : 1.0;
            detR = has_right ? dr : 1.0;
        }
    } else {
        // k_solve_both (filtered_tdma_kernels.cuh:284-303)
        const double C_left = sC[pl * S + sl], D_left = sDr[pl * S + sl];
        const double dl = 1.0 - C_left * a0p;
        const double xl = (d0p - a0p * D_left) / dl;
        const double A_right = sA[pr * S + sl], D0_right = sDl[pr * S + sl];
        const double dr = 1.0 - A_right * cN;
        const double xr = (dN - cN * D0_right) / dr;
        x0 = has_left ? xl : d0p;
        xN = has_right ? xr : dN;
        detL = has_left ? dl : 1.0;
        detR = has_right ? dr : 1.0;
    }

    if (PASS == 1) {
        if (live) {
            xbuf[(size_t)(2 * p) * ld + i] = x0;
            xbuf[(size_t)(2 * p + 1) * ld + i] = xN;
        }
        // dead lanes shadow a live system, so they do not change the max;
        // with CSH = 1 the two parities reduce separately (xor offsets >= 2)
        double v = fmax(fabs(x0), fabs(xN));
        const int width = S < 64 ? S : 64;
        for (int off = width >> 1; off >= (1 << CSH); off >>= 1)
            v = fmax(v, __shfl_xor(v, off, 64));
        if (sl < (1 << CSH)) partial[((size_t)sl * gridDim.x + blockIdx.x) * P + p] = v;
        return;
    }

    int Jf = J;
    if (PASS == 3) {
        // learned J of this (system, partition), then the first dropped
        // term of each side against eps_; a larger one falls back to the
        // full correction (J = m-1)
        const double Nl = (double)(m * P);
        const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        int Jl = mlp_J((float)fmin(f_rho, 1.0), lg_feature(f_d0), lg_feature(f_dN),
                       lg_feature(f_dmax), lg_feature(eps));
        Jl = Jl < 0 ? 0 : (Jl > m - 1 ? m - 1 : Jl);
        const int rec = learned_J_checked<M>(a, c, m, Jl, x0, xN, eps);
        if (!GRD && Jsys && live) Jsys[(size_t)p * ld + i] = rec;
        Jf = GRD ? rec : (rec & 0xFFFF);  // under the guard the record keeps the fallback flag
    }
    if (GRD) {
        const int fb = Jf & (1 << 16);
        Jf &= 0xFFFF;
        const double Nl = (double)(m * P);
        const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        const double c0p = -rr * c[0] * c[1];
        const bool flag = guard_flag<M>(a, c, m, Jf, c0p, aN, x0, xN, eps, detL, detR);
        const int PS = P * S, o = p * S + sl;
        double* gCN = lds;  // the exchange arrays hold cN, dN, a0', d0'
        double* gDN = gCN + PS;
        double* gA = gDN + PS;
        double* gD0 = gA + PS;
        double* gC0 = gD0 + PS;
        double* gAN = gC0 + PS;
        int* gF = (int*)(gAN + PS);
        gC0[o] = c0p;
        gAN[o] = aN;
        gF[o] = flag ? 1 : 0;
        __syncthreads();
        int sys = 0;
        for (int q = 0; q < P; ++q) sys |= gF[q * S + sl];
        if (sys && p == 0) {
            if (cyclic)
                guard_reduced_cyclic(gA, gC0, gD0, gAN, gCN, gDN, P, S, sl);
            else
                guard_reduced_thomas(gA, gC0, gD0, gAN, gCN, gDN, P, S, sl);
        }
        __syncthreads();
        if (sys) {
            x0 = gD0[o];
            xN = gDN[o];
            Jf = m - 1;
        }
        if (Jsys && live) Jsys[(size_t)p * ld + i] = Jf | fb | (sys ? (1 << 17) : 0);
    }

    // k_final_pass: truncated correction
    const int jL = Jf < m - 2 ? Jf : m - 2;
    const int jR = (m - 1) - Jf > 1 ? (m - 1) - Jf : 1;
    if (live) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (j < m) {
                double x;
                if (j == 0) {
                    x = x0;
                } else if (j == m - 1) {
                    x = xN;
                } else {
                    x = d[j];
                    if (j <= jL) x = x - a[j] * x0;
                    if (j >= jR) x = x - c[j] * xN;
                }
                FTD_ST(x, D + base + (size_t)j * ld);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-streamed Filtered v1/v2 (persistent): the arithmetic of
// k_filtered_fused over the StreamRing of k_pascal_stream. Same shape
// requirements (S = 32, two partitions per wave, n_sys % 32 == 0).
//   PASS 0: v2 -> x stored (M stores per tile).
//   PASS 1: v1 first pass -> x0, xN to xbuf, per-tile max to partial
//           (3 stores per tile).
//   PASS 2: v1 second pass -> x0, xN from xbuf (one more DMA per tile into
//           an LDS block, issued ahead of the sweep), x stored (M stores per
//           tile).
// ---------------------------------------------------------------------------
template <int M, int Q, int R, int VERSION, int PASS, int WPE, int GRD>
__global__ __launch_bounds__(Q * 32) __attribute__((amdgpu_waves_per_eu(WPE))) void k_filtered_stream(
    const double* __restrict__ A, const double* __restrict__ B, const double* __restrict__ C,
    double* D, int n_sys, int n_row, int n_tiles, int cyclic, const int* __restrict__ Jp,
    double* __restrict__ xbuf, double* __restrict__ partial, double eps_value, int eps_direct,
    int* __restrict__ Jsys) {
    constexpr int S = 32, NW = Q / 2;
    static_assert(Q % 2 == 0 && M % R == 0 && R >= 2, "stream shape");
    static_assert(!GRD || VERSION == 1, "the guard runs the v1 sweeps");
    // learned J under the guard spills the rows held in registers, and the
    // compiler has placed some of those copies under lane-divergent branch
    // masks (stale rows in the lanes that skipped the branch): not built
    static_assert(!(GRD && PASS == 3), "learned J under the guard: k_filtered_fused");
    __shared__ __attribute__((aligned(16))) double ring[NW][R][2][128];
    // C, D of row m-1; A, D of row 0 (packed); under the guard also c0' and
    // a'_{m-1}, and the guard flags
    __shared__ double ex[GRD ? 6 : 4][Q][S];
    __shared__ int gflag[GRD ? Q : 1][S];
    __shared__ __attribute__((aligned(16))) double xin[PASS == 2 ? NW : 1][128];

    const int t = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int l = t & 63;
    const int p = 2 * w + (l >> 5);
    const int sl = l & 31;
    const size_t ld = (size_t)n_sys;
    int r0, m0;
    part_of(n_row, Q, p, r0, m0);
    int J0 = (VERSION == 2 || PASS == 2 || (GRD && PASS == 0)) ? Jp[p] : m0;
    asm volatile("" : "+v"(J0));  // J is loaded before the ring starts
    // PASS 2: rows 4w..4w+3 of xbuf (x0, xN of partitions 2w, 2w+1), two
    // systems per lane, as the ring's sub-blocks
    const double* xsrc = xbuf + (size_t)(4 * w + (l >> 4)) * ld + 2 * (l & 15);
    StreamRing<M, Q, R> rg;
    int dcnt0;
    rg.init(ring[w], A, B, C, D, n_sys, n_row, w, l, dcnt0);

    int tile = blockIdx.x;
    rg.prologue(tile, dcnt0);
    bool first = true;
    for (; tile < n_tiles; tile += gridDim.x) {
        int m = m0, dcnt = dcnt0, J = J0;
        asm volatile("" : "+v"(m), "+v"(dcnt), "+v"(J));
        const int next = tile + (int)gridDim.x;
        const int ntile = next < n_tiles ? next : tile;
        const int i = tile * S + sl;
        const size_t base = (size_t)r0 * ld + (size_t)i;
        double x0 = 0.0, xN = 0.0;
        if (PASS == 2)
            __builtin_amdgcn_global_load_lds((const void*)(xsrc + (size_t)tile * S),
                                             (lds_ptr_t)&xin[w][0], 16, 0, FTD_DMA_AUX);

        double a[M], c[M], d[M];
        double aN = 0.0, cN = 0.0, dN = 0.0;
        double b1 = 1.0;
        double f_rho = 0.0, f_d0 = 0.0, f_dN = 0.0, f_dmax = 0.0;  // PASS 3 features
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double aj, bj, cj, dj;
            rg.take(j, first, tile, ntile, dcnt, l, aj, bj, cj, dj);
            if (j < m) {
                if (PASS == 3) {
                    // max(|a/b|, |c/b|) in one division, the same value
                    if (j >= 1) f_rho = fmax(f_rho, fmax(fabs(aj), fabs(cj)) / fabs(bj));
                    if (j == 0) f_d0 = dj;
                    if (j == m - 1) f_dN = dj;
                    f_dmax = fmax(f_dmax, fabs(dj));
                }
                if (j < 2) {
                    const double r = 1.0 / bj;
                    a[j] = aj * r;
                    c[j] = cj * r;
                    d[j] = dj * r;
                    if (j == 1) b1 = bj;
                } else if (VERSION == 1 || j <= J) {
                    const double inv = 1.0 / (bj - aj * c[j - 1]);
                    d[j] = inv * (dj - aj * d[j - 1]);
                    c[j] = inv * cj;
                    a[j] = -inv * aj * a[j - 1];
                } else {
                    const double inv = 1.0 / (bj - aj * c[j - 1]);
                    d[j] = inv * (dj - aj * d[j - 1]);
                    c[j] = inv * cj;
                    a[j] = aj;  // left raw, as the skip-A loop leaves A[j]
                }
                if (VERSION == 2 && j == 1 && J == 0) {
                    // k_fwd_pass_v2 with J == 0 eliminates the normalised row 1
                    // against itself
                    const double inv = 1.0 / (b1 - a[1] * c[1]);
                    d[1] = inv * (d[1] - a[1] * d[1]);
                    c[1] = inv * c[1];
                }
                if (j == m - 1) {
                    aN = a[j];
                    cN = c[j];
                    dN = d[j];
                }
            }
        }
        const int lo = (m - 2) - J;
#pragma unroll
        for (int j = M - 3; j >= 1; --j) {
            if (j <= m - 3) {
                const double cj = c[j];
                d[j] = d[j] - cj * d[j + 1];
                if (VERSION == 1) {
                    a[j] = a[j] - cj * a[j + 1];
                    c[j] = -cj * c[j + 1];
                } else {
                    if (J >= 2 && j < J) {
                        const double an = (J == m - 1 && j == m - 3) ? aN : a[j + 1];
                        a[j] = a[j] - cj * an;
                    }
                    if (j > lo) c[j] = -cj * c[j + 1];
                }
            }
        }
        const double rr = 1.0 / (1.0 - a[1] * c[0]);
        const double d0p = rr * (d[0] - c[0] * d[1]);
        const double a0p = rr * a[0];

        const bool has_left = cyclic || p > 0;
        const bool has_right = cyclic || p < Q - 1;
        const int pl = (p + Q - 1) % Q, pr = (p + 1) % Q;
        double detL = 1.0, detR = 1.0;  // the interfaces' 2x2 determinants (the guard's)
        if (PASS == 2) {
            // the 2M row DMAs of this tile were issued after the xbuf DMA
            wait_vmcnt<(2 * M < 63 ? 2 * M : 63)>();
            xin_read((lds_cdptr_t)&xin[w][(l >> 5) * 64 + sl], x0, xN);
            if (GRD) {
                lds_barrier();  // the previous tile's readers of ex[] are done
                ex[0][p][sl] = cN;
                ex[1][p][sl] = dN;
                ex[2][p][sl] = a0p;
                ex[3][p][sl] = d0p;
                lds_barrier();
                const double dl = 1.0 - ex[0][pl][sl] * a0p;
                const double dr = 1.0 - ex[2][pr][sl] * cN;
                detL = has_left ? dl : 1.0;
                detR = has_right ? dr : 1.0;
            }
        } else {
            lds_barrier();  // the previous tile's readers of ex[] are done
            ex[0][p][sl] = cN;
            ex[1][p][sl] = dN;
            ex[2][p][sl] = a0p;
            ex[3][p][sl] = d0p;
            lds_barrier();
            // both interface solves on every lane, then selected: under a
            // lane-divergent branch here the compiler has placed copies of the
            // register-resident rows before the exec restore of the join,
            // leaving stale values in the lanes that skipped the branch
            const double dl = 1.0 - ex[0][pl][sl] * a0p;
            const double xl = (d0p - a0p * ex[1][pl][sl]) / dl;
            const double dr = 1.0 - ex[2][pr][sl] * cN;
            const double xr = (dN - cN * ex[3][pr][sl]) / dr;
            x0 = has_left ? xl : d0p;
            xN = has_right ? xr : dN;
            detL = has_left ? dl : 1.0;
            detR = has_right ? dr : 1.0;
        }
        // the next tile's first R rows have landed before this tile's stores
        wait_vmcnt<0>();
        if (PASS == 1) {
            xbuf[(size_t)(2 * p) * ld + i] = x0;
            xbuf[(size_t)(2 * p + 1) * ld + i] = xN;
            const double v = group_max(fmax(fabs(x0), fabs(xN)), S);
            if (sl == 0) partial[(size_t)tile * Q + p] = v;
        } else {
            int Jf = J;
            if (PASS == 3) {
                // learned J of this (system, partition), checked against eps_
                // by the first dropped term of each side (as k_filtered_fused)
                const double Nl = (double)(m * Q);
                const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
                int Jl = mlp_J((float)fmin(f_rho, 1.0), lg_feature(f_d0), lg_feature(f_dN),
                               lg_feature(f_dmax), lg_feature(eps));
                Jl = Jl < 0 ? 0 : (Jl > m - 1 ? m - 1 : Jl);
                const int rec = learned_J_checked<M>(a, c, m, Jl, x0, xN, eps);
                if (!GRD) Jsys[(size_t)p * ld + i] = rec;
                Jf = GRD ? rec : (rec & 0xFFFF);  // under the guard the record keeps the flag
            }
            if (GRD) {
                // the guard (as k_filtered_fused): the dropped couplings and the
                // first truncated terms against eps_/4, the exact reduced solve
                // of a flagged system in LDS by its partition-0 lane
                const int fb = Jf & (1 << 16);
                Jf &= 0xFFFF;
                const double Nl = (double)(m * Q);
                const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
                const double c0p = -rr * c[0] * c[1];
                const bool flag = guard_flag<M>(a, c, m, Jf, c0p, aN, x0, xN, eps, detL, detR);
                ex[4][p][sl] = c0p;
                ex[5][p][sl] = aN;
                gflag[p][sl] = flag ? 1 : 0;
                lds_barrier();
                int sys = 0;
#pragma unroll
                for (int q = 0; q < Q; ++q) sys |= gflag[q][sl];
                if (sys && p == 0) {
                    if (cyclic)
                        guard_reduced_cyclic(&ex[2][0][0], &ex[4][0][0], &ex[3][0][0],
                                             &ex[5][0][0], &ex[0][0][0], &ex[1][0][0], Q, S, sl);
                    else
                        guard_reduced_thomas(&ex[2][0][0], &ex[4][0][0], &ex[3][0][0],
                                             &ex[5][0][0], &ex[0][0][0], &ex[1][0][0], Q, S, sl);
                }
                lds_barrier();
                if (sys) {
                    x0 = ex[3][p][sl];
                    xN = ex[1][p][sl];
                    Jf = m - 1;
                }
                Jsys[(size_t)p * ld + i] = Jf | fb | (sys ? (1 << 17) : 0);
            }
            const int jL = Jf < m - 2 ? Jf : m - 2;
            const int jR = (m - 1) - Jf > 1 ? (m - 1) - Jf : 1;
            double* const xlast = D + base + (size_t)(m - 1) * ld;
#pragma unroll
            for (int j = 0; j < M; ++j) {
                double x;
                if (j == 0) {
                    x = x0;
                } else if (j >= m - 1) {
                    x = xN;
                } else {
                    x = d[j];
                    if (j <= jL) x = x - a[j] * x0;
                    if (j >= jR) x = x - c[j] * xN;
                }
                FTD_ST(x, j < m - 1 ? D + base + (size_t)j * ld : xlast);
            }
        }
        first = false;
    }
    wait_vmcnt<0>();  // no DMA into LDS outlives the wave
}

// ---------------------------------------------------------------------------
// general kernels: one lane per (system, partition), rows walked in global
// memory, in place (A and C clobbered, as the reference allows)
// ---------------------------------------------------------------------------

// plain Thomas, P == 1 (tdma_many_kernel, tdma_local_cuda.cu:32-64)
__global__ __launch_bounds__(256) void k_gen_thomas(const double* __restrict__ A,
                                                    const double* __restrict__ B,
                                                    double* __restrict__ C, double* __restrict__ D,
                                                    int n_sys, int n_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sys) return;
    const size_t ld = (size_t)n_sys;
    double b = B[i], cp = C[i] / b, dp = D[i] / b;
    C[i] = cp;
    D[i] = dp;
    for (int j = 1; j < n_row; ++j) {
        const size_t o = (size_t)j * ld + i;
        const double aj = A[o];
        const double r = 1.0 / (B[o] - aj * cp);
        dp = r * (D[o] - aj * dp);
        cp = r * C[o];
        C[o] = cp;
        D[o] = dp;
    }
    double xn = dp;
    for (int j = n_row - 2; j >= 0; --j) {
        const size_t o = (size_t)j * ld + i;
        xn = D[o] - C[o] * xn;
        D[o] = xn;
    }
}

// Sherman-Morrison cyclic Thomas, P == 1 (tdma_cyclic_many_kernel,
// tdma_local_cuda.cu:67-128); E is [n_row x n_sys] scratch
__global__ __launch_bounds__(256) void k_gen_thomas_cyclic(const double* __restrict__ A,
                                                           const double* __restrict__ B,
                                                           double* __restrict__ C,
                                                           double* __restrict__ D,
                                                           double* __restrict__ E, int n_sys,
                                                           int n_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sys) return;
    const size_t ld = (size_t)n_sys;
    for (int j = 0; j < n_row; ++j) E[(size_t)j * ld + i] = 0.0;
    E[ld + i] = -A[ld + i];
    E[(size_t)(n_row - 1) * ld + i] = -C[(size_t)(n_row - 1) * ld + i];
    {
        const double inv_b = 1.0 / B[ld + i];
        D[ld + i] *= inv_b;
        E[ld + i] *= inv_b;
        C[ld + i] *= inv_b;
    }
    for (int j = 2; j < n_row; ++j) {
        const size_t o = (size_t)j * ld + i, om = o - ld;
        const double aj = A[o];
        const double r = 1.0 / (B[o] - aj * C[om]);
        D[o] = r * (D[o] - aj * D[om]);
        E[o] = r * (E[o] - aj * E[om]);
        C[o] = r * C[o];
    }
    for (int j = n_row - 2; j >= 1; --j) {
        const size_t o = (size_t)j * ld + i, op = o + ld;
        const double cj = C[o];
        D[o] = D[o] - cj * D[op];
        E[o] = E[o] - cj * E[op];
    }
    const size_t oN = (size_t)(n_row - 1) * ld + i;
    D[i] = (D[i] - A[i] * D[oN] - C[i] * D[ld + i]) / (B[i] + A[i] * E[oN] + C[i] * E[ld + i]);
    const double x0 = D[i];
    for (int j = 1; j < n_row; ++j) {
        const size_t o = (size_t)j * ld + i;
        D[o] = D[o] + x0 * E[o];
    }
}

// PaScaL partition sweep in place (modified_thomas_kernel,
// tdma_local_cuda.cu:134-203); reduced rows to R[(2p+{0,1})*3 + {a,c,d}][n_sys]
__global__ __launch_bounds__(256) void k_gen_pascal_sweep(double* __restrict__ A,
                                                          const double* __restrict__ B,
                                                          double* __restrict__ C,
                                                          double* __restrict__ D, int n_sys,
                                                          int n_row, int P,
                                                          double* __restrict__ R) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= n_sys) return;
    int s, m;
    part_of(n_row, P, p, s, m);
    const size_t ld = (size_t)n_sys;
    double* Ap = A + (size_t)s * ld + i;
    const double* Bp = B + (size_t)s * ld + i;
    double* Cp = C + (size_t)s * ld + i;
    double* Dp = D + (size_t)s * ld + i;
    for (int j = 0; j < 2; ++j) {
        const size_t o = (size_t)j * ld;
        const double b = Bp[o];
        Ap[o] = Ap[o] / b;
        Cp[o] = Cp[o] / b;
        Dp[o] = Dp[o] / b;
    }
    double am = Ap[ld], cm = Cp[ld], dm = Dp[ld];
    for (int j = 2; j < m; ++j) {
        const size_t o = (size_t)j * ld;
        const double aj = Ap[o];
        const double inv = 1.0 / (Bp[o] - aj * cm);
        dm = inv * (Dp[o] - aj * dm);
        cm = inv * Cp[o];
        am = -inv * aj * am;
        Ap[o] = am;
        Cp[o] = cm;
        Dp[o] = dm;
    }
    const double aN = am, cN = cm, dN = dm;
    double an = Ap[(size_t)(m - 2) * ld], cn = Cp[(size_t)(m - 2) * ld], dn = Dp[(size_t)(m - 2) * ld];
    for (int j = m - 3; j >= 1; --j) {
        const size_t o = (size_t)j * ld;
        const double cj = Cp[o];
        dn = Dp[o] - cj * dn;
        an = Ap[o] - cj * an;
        cn = -cj * cn;
        Ap[o] = an;
        Cp[o] = cn;
        Dp[o] = dn;
    }
    const double r = 1.0 / (1.0 - an * Cp[0]);
    const double d0 = r * (Dp[0] - Cp[0] * dn);
    const double a0 = r * Ap[0];
    const double c0 = -r * Cp[0] * cn;
    double* Rp = R + (size_t)(6 * p) * ld + i;
    Rp[0 * ld] = a0;
    Rp[1 * ld] = c0;
    Rp[2 * ld] = d0;
    Rp[3 * ld] = aN;
    Rp[4 * ld] = cN;
    Rp[5 * ld] = dN;
}

// reduced 2P-row solve per system, B == 1, in R (x left in the d slots);
// E is [2P x n_sys] scratch for the cyclic case
__global__ __launch_bounds__(256) void k_gen_pascal_reduced(double* __restrict__ R,
                                                            double* __restrict__ E, int n_sys,
                                                            int P, int cyclic) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sys) return;
    const size_t ld = (size_t)n_sys;
    const int NR = 2 * P;
    // reduced row k lives at R[(3k + {0,1,2}) * ld]
#define RA(k) R[(size_t)(3 * (k) + 0) * ld + i]
#define RC(k) R[(size_t)(3 * (k) + 1) * ld + i]
#define RD(k) R[(size_t)(3 * (k) + 2) * ld + i]
#define RE(k) E[(size_t)(k) * ld + i]
    if (!cyclic) {
        for (int k = 1; k < NR; ++k) {
            const double inv = 1.0 / (1.0 - RA(k) * RC(k - 1));
            RD(k) = inv * (RD(k) - RA(k) * RD(k - 1));
            RC(k) = inv * RC(k);
        }
        for (int k = NR - 2; k >= 0; --k) RD(k) = RD(k) - RC(k) * RD(k + 1);
    } else {
        for (int k = 0; k < NR; ++k) RE(k) = 0.0;
        RE(1) = -RA(1);
        RE(NR - 1) = -RC(NR - 1);
        for (int k = 2; k < NR; ++k) {
            const double inv = 1.0 / (1.0 - RA(k) * RC(k - 1));
            RD(k) = inv * (RD(k) - RA(k) * RD(k - 1));
            RE(k) = inv * (RE(k) - RA(k) * RE(k - 1));
            RC(k) = inv * RC(k);
        }
        for (int k = NR - 2; k >= 1; --k) {
            RD(k) = RD(k) - RC(k) * RD(k + 1);
            RE(k) = RE(k) - RC(k) * RE(k + 1);
        }
        RD(0) = (RD(0) - RA(0) * RD(NR - 1) - RC(0) * RD(1)) /
                (1.0 + RA(0) * RE(NR - 1) + RC(0) * RE(1));
        const double x0 = RD(0);
        for (int k = 1; k < NR; ++k) RD(k) = RD(k) + x0 * RE(k);
    }
#undef RA
#undef RC
#undef RD
#undef RE
}

// interior update (update_solution_kernel, tdma_local_cuda.cu:205-224)
__global__ __launch_bounds__(256) void k_gen_pascal_update(const double* __restrict__ A,
                                                           const double* __restrict__ C,
                                                           double* __restrict__ D, int n_sys,
                                                           int n_row, int P,
                                                           const double* __restrict__ R) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= n_sys) return;
    int s, m;
    part_of(n_row, P, p, s, m);
    const size_t ld = (size_t)n_sys;
    const double x0 = R[(size_t)(6 * p + 2) * ld + i];
    const double xN = R[(size_t)(6 * p + 5) * ld + i];
    double* Dp = D + (size_t)s * ld + i;
    const double* Ap = A + (size_t)s * ld + i;
    const double* Cp = C + (size_t)s * ld + i;
    Dp[0] = x0;
    Dp[(size_t)(m - 1) * ld] = xN;
    for (int j = 1; j < m - 1; ++j) {
        const size_t o = (size_t)j * ld;
        Dp[o] = Dp[o] - (Ap[o] * x0 + Cp[o] * xN);
    }
}

// Filtered sweeps in place: k_fwd_pass_v{1,2} + k_bwd_pass_v{1,2} with the
// fused pack (filtered_tdma_kernels.cuh:69-194). FEAT (learned J, v1
// sweeps): the network's features are taken from the raw rows as the
// forward pass reads them; its J of this (system, partition), clamped to
// [0, m-1], goes to Jsys[p][i] for k_gen_filtered_check.
template <int VERSION, bool FEAT>
__global__ __launch_bounds__(256) void k_gen_filtered_sweep(double* __restrict__ A,
                                                            const double* __restrict__ B,
                                                            double* __restrict__ C,
                                                            double* __restrict__ D, int n_sys,
                                                            int n_row, int P,
                                                            const int* __restrict__ Jp,
                                                            double eps_value, int eps_direct,
                                                            int* __restrict__ Jsys) {
    static_assert(!FEAT || VERSION == 1, "the learned J runs the v1 sweeps");
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (i >= n_sys) return;
    int s, m;
    part_of(n_row, P, p, s, m);
    const size_t ld = (size_t)n_sys;
    double* Ap = A + (size_t)s * ld + i;
    const double* Bp = B + (size_t)s * ld + i;
    double* Cp = C + (size_t)s * ld + i;
    double* Dp = D + (size_t)s * ld + i;
    const int J = (VERSION == 2) ? Jp[p] : m;
    // features: max |a/b|, |c/b| over rows 1..m-1, raw d of rows 0 and m-1,
    // max |d| (as k_filtered_fused takes them)
    double f_rho = 0.0, f_d0 = 0.0, f_dN = 0.0, f_dmax = 0.0;
    for (int j = 0; j < 2; ++j) {
        const size_t o = (size_t)j * ld;
        if (FEAT) {
            const double aj = Ap[o], bj = Bp[o], cj = Cp[o], dj = Dp[o];
            if (j >= 1) f_rho = fmax(f_rho, fmax(fabs(aj), fabs(cj)) / fabs(bj));
            if (j == 0) f_d0 = dj;
            f_dmax = fmax(f_dmax, fabs(dj));
        }
        const double r0 = 1.0 / Bp[o];
        Ap[o] *= r0;
        Cp[o] *= r0;
        Dp[o] *= r0;
    }
    double am = Ap[ld], cm = Cp[ld], dm = Dp[ld];
    // Row j + 1's loads are issued before row j's stores: the compiler cannot
    // move them there itself (the same pointers one row apart, and it cannot
    // rule out ld == 0), and a lane then waits for every row's loads in turn.
    // The next row is clamped to the partition's last row (a reload, unused).
    int j = 2;
    const int j2 = 2 < m ? 2 : m - 1;
    double na = Ap[(size_t)j2 * ld], nb = Bp[(size_t)j2 * ld], nc = Cp[(size_t)j2 * ld],
           nd = Dp[(size_t)j2 * ld];
    for (; j <= J && j < m; ++j) {
        const size_t o = (size_t)j * ld;
        const double aj = na, bj = nb, cj = nc, dj = nd;
        const size_t on = (size_t)(j + 1 < m ? j + 1 : m - 1) * ld;
        na = Ap[on];
        nb = Bp[on];
        nc = Cp[on];
        nd = Dp[on];
        if (FEAT) {
            f_rho = fmax(f_rho, fmax(fabs(aj), fabs(cj)) / fabs(bj));
            if (j == m - 1) f_dN = dj;
            f_dmax = fmax(f_dmax, fabs(dj));
        }
        const double inv = 1.0 / (bj - aj * cm);
        dm = inv * (dj - aj * dm);
        cm = inv * cj;
        am = -inv * aj * am;
        Ap[o] = am;
        Cp[o] = cm;
        Dp[o] = dm;
    }
    if (FEAT) {
        const double Nl = (double)(m * P);
        const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        int Jl = mlp_J((float)fmin(f_rho, 1.0), lg_feature(f_d0), lg_feature(f_dN),
                       lg_feature(f_dmax), lg_feature(eps));
        Jsys[(size_t)p * ld + i] = Jl < 0 ? 0 : (Jl > m - 1 ? m - 1 : Jl);
    }
    if (VERSION == 2) {
        // the skip-A rows, loads one row ahead as above
        j = J + 1 > 1 ? J + 1 : 1;
        if (j < m) {
            na = Ap[(size_t)j * ld];
            nb = Bp[(size_t)j * ld];
            nc = Cp[(size_t)j * ld];
            nd = Dp[(size_t)j * ld];
        }
        for (; j < m; ++j) {
            const size_t o = (size_t)j * ld;
            const double aj = na, bj = nb, cj = nc, dj = nd;
            const size_t on = (size_t)(j + 1 < m ? j + 1 : m - 1) * ld;
            na = Ap[on];
            nb = Bp[on];
            nc = Cp[on];
            nd = Dp[on];
            const double inv = 1.0 / (bj - aj * cm);
            dm = inv * (dj - aj * dm);
            cm = inv * cj;
            Cp[o] = cm;
            Dp[o] = dm;
        }
    }
    // backward, row j - 1's loads issued before row j's stores
    const size_t onm2 = (size_t)(m - 2) * ld;
    double dn = Dp[onm2], cn = Cp[onm2];
    double an = (VERSION == 1) ? Ap[onm2] : (J >= 2 ? Ap[(size_t)J * ld] : 0.0);
    const int lo = (m - 2) - J;
    // (v2 reads a row's A only where it updates a, rows below J: J is the
    // partition's, so the test is uniform across the wave)
    j = m - 3;
    double pa = 0.0, pc = 0.0, pd = 0.0;
    if (j >= 1) {
        if (VERSION == 1 || (J >= 2 && j < J)) pa = Ap[(size_t)j * ld];
        pc = Cp[(size_t)j * ld];
        pd = Dp[(size_t)j * ld];
    }
    for (; j >= 1; --j) {
        const size_t o = (size_t)j * ld;
        const double aj = pa, cj = pc, dj = pd;
        const int jp = j > 1 ? j - 1 : 1;
        const size_t op = (size_t)jp * ld;
        if (VERSION == 1 || (J >= 2 && jp < J)) pa = Ap[op];
        pc = Cp[op];
        pd = Dp[op];
        dn = dj - cj * dn;
        Dp[o] = dn;
        if (VERSION == 1 || (J >= 2 && j < J)) {
            an = aj - cj * an;
            Ap[o] = an;
        }
        if (VERSION == 1 || j > lo) {
            cn = -cj * cn;
            Cp[o] = cn;
        }
    }
    const double A1 = Ap[ld], C0 = Cp[0], D1 = Dp[ld];
    const double r = 1.0 / (1.0 - A1 * C0);
    Dp[0] = r * (Dp[0] - C0 * D1);
    Ap[0] = r * Ap[0];
}

// interface 2x2 solves (k_solve_both) -> xbuf rows (2p, 2p+1), plus per-block
// max of |x0|,|xN| for v1's J
__global__ __launch_bounds__(256) void k_gen_filtered_iface(const double* __restrict__ A,
                                                            const double* __restrict__ C,
                                                            const double* __restrict__ D,
                                                            int n_sys, int n_row, int P,
                                                            int cyclic, double* __restrict__ xbuf,
                                                            double* __restrict__ partial) {
    __shared__ double red[4];
    const int p = blockIdx.y;
    const int i_raw = blockIdx.x * 256 + threadIdx.x;
    const bool live = i_raw < n_sys;
    const int i = live ? i_raw : n_sys - 1;
    const size_t ld = (size_t)n_sys;
    int s, m, sl, ml, sr, mr;
    part_of(n_row, P, p, s, m);
    const int pl = (p + P - 1) % P, pr = (p + 1) % P;
    part_of(n_row, P, pl, sl, ml);
    part_of(n_row, P, pr, sr, mr);
    const double a0 = A[(size_t)s * ld + i], d0 = D[(size_t)s * ld + i];
    const double cN = C[(size_t)(s + m - 1) * ld + i], dN = D[(size_t)(s + m - 1) * ld + i];
    double x0 = d0, xN = dN;
    if (cyclic || p > 0) {
        const double C_left = C[(size_t)(sl + ml - 1) * ld + i];
        const double D_left = D[(size_t)(sl + ml - 1) * ld + i];
        x0 = (d0 - a0 * D_left) / (1.0 - C_left * a0);
    }
    if (cyclic || p < P - 1) {
        const double A_right = A[(size_t)sr * ld + i];
        const double D0_right = D[(size_t)sr * ld + i];
        xN = (dN - cN * D0_right) / (1.0 - A_right * cN);
    }
    if (live) {
        xbuf[(size_t)(2 * p) * ld + i] = x0;
        xbuf[(size_t)(2 * p + 1) * ld + i] = xN;
    }
    double v = group_max(fmax(fabs(x0), fabs(xN)), 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[(size_t)blockIdx.x * P + p] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Per-system J of the general path (one thread per system, its partitions
// in turn), after k_gen_filtered_iface wrote x0, xN to xbuf:
//   LEARNED: the network's J (Jsys, from k_gen_filtered_sweep<1, true>)
//     checked as learned_J_checked does. Otherwise J = Jp[p].
//   GRD: guard_flag's test on the values the sweeps left in A, C, D (a', c'
//     after the backward sweep, a0', d0' of row 0, the forward aN, cN, dN of
//     row m-1, c0' = -r c0 c1 recomputed); a flagged system is solved
//     exactly, guard_reduced_thomas over its 2P reduced rows (c', d' kept in
//     gbuf [4P x n_sys]; periodic systems: guard_reduced_cyclic, e in xbuf),
//     its x0, xN replaced in xbuf and J = m - 1.
// Jsys[p][i] = J | fallback << 16 | guard << 17 for the final pass.
template <bool LEARNED, bool GRD>
__global__ __launch_bounds__(256) void k_gen_filtered_check(
    const double* __restrict__ A, const double* __restrict__ C, const double* __restrict__ D,
    int n_sys, int n_row, int P, int cyclic, const int* __restrict__ Jp, double* __restrict__ xbuf,
    int* __restrict__ Jsys, double eps_value, int eps_direct, double* __restrict__ gbuf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sys) return;
    const size_t ld = (size_t)n_sys;
    auto at = [&](const double* X, int row) { return X[(size_t)row * ld + i]; };
    bool sys = false;
    for (int p = 0; p < P; ++p) {
        int s, m;
        part_of(n_row, P, p, s, m);
        const double x0 = xbuf[(size_t)(2 * p) * ld + i];
        const double xN = xbuf[(size_t)(2 * p + 1) * ld + i];
        const double Nl = (double)(m * P);
        const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        int Jf = LEARNED ? Jsys[(size_t)p * ld + i] : Jp[p];
        int fb = 0;
        if (LEARNED) {
            // learned_J_checked on the coefficients in memory
            auto fails = [&](int J) {
                const int kL = (J < m - 2 ? J : m - 2) + 1;
                const int kR = ((m - 1) - J > 1 ? (m - 1) - J : 1) - 1;
                const double ta = kL <= m - 2 ? fabs(at(A, s + kL) * x0) : 0.0;
                const double tc = kR >= 1 ? fabs(at(C, s + kR) * xN) : 0.0;
                return ta > eps || tc > eps;
            };
            if (fails(Jf)) {
                Jf = m - 1;
                fb = 1 << 16;
            }
        }
        if (GRD) {
            // guard_flag's bounds; the interface determinants as
            // k_gen_filtered_iface forms them
            double detL = 1.0, detR = 1.0;
            if (cyclic || p > 0) {
                int sl, ml;
                part_of(n_row, P, (p + P - 1) % P, sl, ml);
                detL = 1.0 - at(C, sl + ml - 1) * at(A, s);
            }
            if (cyclic || p < P - 1) {
                int sr, mr;
                part_of(n_row, P, (p + 1) % P, sr, mr);
                detR = 1.0 - at(A, sr) * at(C, s + m - 1);
            }
            const double thr = eps * 0.25;
            const double thrL = thr * fmin(1.0, fabs(detL));
            const double thrR = thr * fmin(1.0, fabs(detR));
            const int kL = (Jf < m - 2 ? Jf : m - 2) + 1;
            const int kR = ((m - 1) - Jf > 1 ? (m - 1) - Jf : 1) - 1;
            const double ta = kL <= m - 2 ? fabs(at(A, s + kL) * x0) : 0.0;
            const double tc = kR >= 1 ? fabs(at(C, s + kR) * xN) : 0.0;
            const double c0 = at(C, s), c1 = at(C, s + 1);
            const double rr = 1.0 / (1.0 - at(A, s + 1) * c0);
            const double c0p = -rr * c0 * c1;
            const double aN = at(A, s + m - 1);
            sys = sys || fabs(c0p * xN) > thrL || fabs(aN * x0) > thrR || ta > thr || tc > thr;
        }
        Jsys[(size_t)p * ld + i] = Jf | fb;
    }
    if (!GRD || !sys) return;
    // the exact reduced solve: rows 2q (a0', 1, c0' | d0') and 2q + 1
    // (aN, 1, cN | dN), in guard_reduced_thomas's order
    auto g = [&](int q, int k) -> double& { return gbuf[(size_t)(4 * q + k) * ld + i]; };
    if (cyclic) {
        // periodic, in guard_reduced_cyclic's order: c', y of row k in gbuf,
        // its e in xbuf row k (whose x0 / xN the solve replaces)
        auto xb = [&](int k) -> double& { return xbuf[(size_t)k * ld + i]; };
        auto rrow = [&](int k, double& a, double& c, double& d) {
            int s, m;
            part_of(n_row, P, k >> 1, s, m);
            if (k & 1) {
                a = at(A, s + m - 1);
                c = at(C, s + m - 1);
                d = at(D, s + m - 1);
            } else {
                const double c0 = at(C, s), c1 = at(C, s + 1);
                const double rr = 1.0 / (1.0 - at(A, s + 1) * c0);
                a = at(A, s);
                c = -rr * c0 * c1;
                d = at(D, s);
            }
        };
        const int n = 2 * P;
        double a, c, d;
        rrow(1, a, c, d);
        double cp = c, dp = d, e = -a;
        g(0, 2) = cp;
        g(0, 3) = dp;
        xb(1) = e;
        for (int k = 2; k < n; ++k) {
            rrow(k, a, c, d);
            const double e0 = k == n - 1 ? -c : 0.0;
            const double r = 1.0 / (1.0 - a * cp);
            dp = r * (d - a * dp);
            e = r * (e0 - a * e);
            cp = r * c;
            g(k >> 1, 2 * (k & 1)) = cp;
            g(k >> 1, 2 * (k & 1) + 1) = dp;
            xb(k) = e;
        }
        const double yl = dp, el = e;
        for (int k = n - 2; k >= 1; --k) {
            const double ck = g(k >> 1, 2 * (k & 1));
            dp = g(k >> 1, 2 * (k & 1) + 1) - ck * dp;
            e = xb(k) - ck * e;
            g(k >> 1, 2 * (k & 1) + 1) = dp;
            xb(k) = e;
        }
        rrow(0, a, c, d);
        const double x0 = (d - a * yl - c * dp) / (1.0 + a * el + c * e);
        xb(0) = x0;
        for (int k = 1; k < n; ++k) xb(k) = g(k >> 1, 2 * (k & 1) + 1) + x0 * xb(k);
        for (int q = 0; q < P; ++q) {
            int s, m;
            part_of(n_row, P, q, s, m);
            const int fb = Jsys[(size_t)q * ld + i] & (1 << 16);
            Jsys[(size_t)q * ld + i] = (m - 1) | fb | (1 << 17);
        }
        return;
    }
    double cp = 0.0, dp = 0.0;
    for (int q = 0; q < P; ++q) {
        int s, m;
        part_of(n_row, P, q, s, m);
        const double c0 = at(C, s), c1 = at(C, s + 1);
        const double rr = 1.0 / (1.0 - at(A, s + 1) * c0);
        const double c0p = -rr * c0 * c1;
        if (q == 0) {
            cp = c0p;
            dp = at(D, s);
        } else {
            const double ak = at(A, s);
            const double inv = 1.0 / (1.0 - ak * cp);
            dp = inv * (at(D, s) - ak * dp);
            cp = inv * c0p;
        }
        g(q, 0) = cp;
        g(q, 1) = dp;
        const double ak = at(A, s + m - 1);
        const double inv = 1.0 / (1.0 - ak * cp);
        dp = inv * (at(D, s + m - 1) - ak * dp);
        cp = inv * at(C, s + m - 1);
        g(q, 2) = cp;
        g(q, 3) = dp;
    }
    double xn = dp;
    for (int q = P - 1; q >= 0; --q) {
        int s, m;
        part_of(n_row, P, q, s, m);
        if (q < P - 1) xn = g(q, 3) - g(q, 2) * xn;
        xbuf[(size_t)(2 * q + 1) * ld + i] = xn;
        xn = g(q, 1) - g(q, 0) * xn;
        xbuf[(size_t)(2 * q) * ld + i] = xn;
        const int fb = Jsys[(size_t)q * ld + i] & (1 << 16);
        Jsys[(size_t)q * ld + i] = (m - 1) | fb | (1 << 17);
    }
}

__device__ inline int wave_max_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ inline int wave_min_i(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

// truncated final correction (k_final_pass, filtered_tdma_kernels.cuh:196-218);
// J per partition from Jp, or per (partition, system) from Jsys (low 16
// bits) where the general path's check kernel ran.
// One J per partition: the left rows 1..jL, then the right rows jR..m-2.
// A J per system: one walk over the rows, each read and written once, the
// two corrections predicated per lane ((d - a x0) - c xN where both apply,
// as the two walks give), and the rows no lane of the wave corrects skipped
// at once. A lane that falls back to J = m-1 then costs its wave one walk
// over the partition instead of two.
__global__ __launch_bounds__(256) void k_gen_filtered_final(const double* __restrict__ A,
                                                            const double* __restrict__ C,
                                                            double* __restrict__ D, int n_sys,
                                                            int n_row, int P,
                                                            const int* __restrict__ Jp,
                                                            const double* __restrict__ xbuf,
                                                            const int* __restrict__ Jsys) {
    const int i_raw = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    const bool live = i_raw < n_sys;
    if (!Jsys && !live) return;
    const int i = live ? i_raw : n_sys - 1;
    int s, m;
    part_of(n_row, P, p, s, m);
    const size_t ld = (size_t)n_sys;
    const double x0 = xbuf[(size_t)(2 * p) * ld + i];
    const double xN = xbuf[(size_t)(2 * p + 1) * ld + i];
    double* Dp = D + (size_t)s * ld + i;
    const double* Ap = A + (size_t)s * ld + i;
    const double* Cp = C + (size_t)s * ld + i;
    const int J = Jsys ? (Jsys[(size_t)p * ld + i] & 0xFFFF) : Jp[p];
    int jL = J < m - 2 ? J : m - 2;
    int jR = (m - 1) - J > 1 ? (m - 1) - J : 1;
    if (!Jsys) {
        Dp[0] = x0;
        Dp[(size_t)(m - 1) * ld] = xN;
        for (int j = 1; j <= jL; ++j) {
            const size_t o = (size_t)j * ld;
            Dp[o] = Dp[o] - Ap[o] * x0;
        }
        for (int j = jR; j < m - 1; ++j) {
            const size_t o = (size_t)j * ld;
            Dp[o] = Dp[o] - Cp[o] * xN;
        }
        return;
    }
    if (live) {
        Dp[0] = x0;
        Dp[(size_t)(m - 1) * ld] = xN;
    } else {
        jL = 0;  // no row
        jR = m - 1;
    }
    const int wjL = wave_max_i(jL), wjR = wave_min_i(jR);
    for (int j = 1; j <= m - 2; ++j) {
        if (j > wjL && j < wjR) j = wjR;  // wave-uniform: no lane corrects rows wjL+1..wjR-1
        if (j > m - 2) break;
        const bool l = j <= jL, r = j >= jR;
        if (l || r) {
            const size_t o = (size_t)j * ld;
            double x = Dp[o];
            if (l) x = x - Ap[o] * x0;
            if (r) x = x - Cp[o] * xN;
            Dp[o] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline int log2i(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

hipError_t launch_set_rho(const double* A, const double* B, const double* C, int n_sys, int n_row,
                          double* rho, hipStream_t st) {
    const int n = n_row + 1;
    hipLaunchKernelGGL(k_set_rho, dim3((n + 255) / 256), dim3(256), 0, st, A, B, C, n_sys, n_row,
                       rho, rho + n);
    return hipGetLastError();
}

hipError_t launch_boundary_maxabs(const double* D, int n_sys, int n_row, int P, int csh,
                                  double* partial, int* n_tiles, hipStream_t st) {
    const int nt = (n_sys + 255) / 256;
    *n_tiles = nt;
    hipLaunchKernelGGL(k_boundary_maxabs, dim3(nt, P), dim3(256), 0, st, D, n_sys, n_row, P, csh,
                       partial);
    return hipGetLastError();
}

hipError_t launch_compute_J(const double* partial, int n_tiles, int P, int n_row, const double* rho,
                            double eps_value, int eps_direct, int csh, int* J, hipStream_t st) {
    hipLaunchKernelGGL(k_compute_J, dim3(P << csh), dim3(256), 0, st, partial, n_tiles, P, n_row,
                       rho, rho + (n_row + 1), eps_value, eps_direct, J);
    return hipGetLastError();
}

// -------- fused PaScaL dispatch over the compiled (M, MAXT) variants --------

// (rows per lane M, max threads T, prefetch rows PF, target waves per SIMD)
#ifndef FTD_PASCAL_VARIANTS
#define FTD_PASCAL_VARIANTS(X) \
    X(8, 512, 8, 4)             \
    X(16, 512, 8, 2)            \
    X(32, 256, 8, 2)            \
    X(32, 512, 4, 2)            \
    X(64, 256, 4, 1)
#endif

int pascal_fused_supported(int M, int maxt) {
#define X(m_, t_, pf_, w_) \
    if (M == m_ && maxt == t_) return 1;
    FTD_PASCAL_VARIANTS(X)
#undef X
    return 0;
}

size_t pascal_fused_lds(int Q, int S, int cyclic) {
    return sizeof(double) * (size_t)(cyclic ? 4 : 3) * (size_t)(2 * Q) * (size_t)S;
}

template <typename TC, typename TD>
static hipError_t launch_pascal_fused_t(int M, int maxt, const TC* A, const TC* B, const TC* C,
                                        const TD* Din, double* Xout, int n_sys, int n_row, int Q,
                                        int S, int cyclic, int csh, hipStream_t st) {
    const int lgS = log2i(S);
    const dim3 grid((n_sys + S - 1) / S), block(S * Q);
    const size_t lds = pascal_fused_lds(Q, S, cyclic);
#define X(m_, t_, pf_, w_)                                                                  \
    if (M == m_ && maxt == t_) {                                                            \
        if (cyclic)                                                                         \
            hipLaunchKernelGGL((k_pascal_fused<m_, t_, true, pf_, w_, TC, TD>), grid, block, \
                               lds, st, A, B, C, Din, Xout, n_sys, n_row, Q, lgS);          \
        else if (csh)                                                                       \
            hipLaunchKernelGGL((k_pascal_fused<m_, t_, false, pf_, w_, TC, TD, 1>), grid,   \
                               block, lds, st, A, B, C, Din, Xout, n_sys, n_row, Q, lgS);   \
        else                                                                                \
            hipLaunchKernelGGL((k_pascal_fused<m_, t_, false, pf_, w_, TC, TD>), grid,      \
                               block, lds, st, A, B, C, Din, Xout, n_sys, n_row, Q, lgS);   \
        return hipGetLastError();                                                           \
    }
    FTD_PASCAL_VARIANTS(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_pascal_fused(int M, int maxt, const double* A, const double* B, const double* C,
                               double* D, int n_sys, int n_row, int Q, int S, int cyclic, int csh,
                               hipStream_t st) {
    if (csh && cyclic) return hipErrorInvalidValue;
    return launch_pascal_fused_t<double, double>(M, maxt, A, B, C, D, D, n_sys, n_row, Q, S,
                                                 cyclic, csh, st);
}

hipError_t launch_pascal_fused_f32c(int M, int maxt, const float* A, const float* B,
                                    const float* C, double* D, int n_sys, int n_row, int Q, int S,
                                    int cyclic, hipStream_t st) {
    return launch_pascal_fused_t<float, double>(M, maxt, A, B, C, D, D, n_sys, n_row, Q, S,
                                                cyclic, 0, st);
}

hipError_t launch_pascal_fused_f32(int M, int maxt, const float* A, const float* B, const float* C,
                                   const float* D, double* X, int n_sys, int n_row, int Q, int S,
                                   int cyclic, hipStream_t st) {
    return launch_pascal_fused_t<float, float>(M, maxt, A, B, C, D, X, n_sys, n_row, Q, S, cyclic,
                                               0, st);
}

hipError_t launch_pascal_ir(int M, int maxt, const float* A, const float* B, const float* C,
                            const float* D, double* X, int n_sys, int n_row, int Q, int S, int res,
                            hipStream_t st) {
    const int lgS = log2i(S);
    const dim3 grid((n_sys + S - 1) / S), block(S * Q);
    const size_t lds = pascal_fused_lds(Q, S, 0);
#define X(m_, t_, pf_, w_)                                                                    \
    if (M == m_ && maxt == t_) {                                                              \
        if (res)                                                                              \
            hipLaunchKernelGGL((k_pascal_ir<m_, t_, pf_, w_, true>), grid, block, lds, st, A, B, \
                               C, D, X, n_sys, n_row, Q, lgS);                                \
        else                                                                                  \
            hipLaunchKernelGGL((k_pascal_ir<m_, t_, pf_, w_, false>), grid, block, lds, st, A,  \
                               B, C, D, X, n_sys, n_row, Q, lgS);                             \
        return hipGetLastError();                                                             \
    }
    FTD_PASCAL_VARIANTS(X)
#undef X
    return hipErrorInvalidValue;
}

int thomas_fused_rows(int n_row) {
    return n_row <= 8 ? 8 : n_row <= 16 ? 16 : n_row <= 32 ? 32 : n_row <= 64 ? 64 : 0;
}

template <int CSH>
static hipError_t launch_thomas_fused_t(const double* A, const double* B, const double* C,
                                        double* D, int n_sys, int n_row, hipStream_t st) {
    const dim3 grid((n_sys + 255) / 256), block(256);
    switch (thomas_fused_rows(n_row)) {
        case 8: hipLaunchKernelGGL((k_thomas_fused<8, 8, CSH>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 16: hipLaunchKernelGGL((k_thomas_fused<16, 8, CSH>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 32: hipLaunchKernelGGL((k_thomas_fused<32, 8, CSH>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 64: hipLaunchKernelGGL((k_thomas_fused<64, 4, CSH>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_thomas_fused(const double* A, const double* B, const double* C, double* D,
                               int n_sys, int n_row, int csh, hipStream_t st) {
    return csh ? launch_thomas_fused_t<1>(A, B, C, D, n_sys, n_row, st)
               : launch_thomas_fused_t<0>(A, B, C, D, n_sys, n_row, st);
}

hipError_t launch_thomas_cyclic_fused(const double* A, const double* B, const double* C, double* D,
                                      int n_sys, int n_row, hipStream_t st) {
    const dim3 grid((n_sys + 255) / 256), block(256);
    switch (thomas_fused_rows(n_row)) {
        case 8: hipLaunchKernelGGL((k_thomas_cyclic_fused<8, 8>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 16: hipLaunchKernelGGL((k_thomas_cyclic_fused<16, 8>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 32: hipLaunchKernelGGL((k_thomas_cyclic_fused<32, 4>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        case 64: hipLaunchKernelGGL((k_thomas_cyclic_fused<64, 2>), grid, block, 0, st, A, B, C, D, n_sys, n_row); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// -------- fused Filtered dispatch --------

#ifndef FTD_FILTERED_VARIANTS
#define FTD_FILTERED_VARIANTS(X) \
    X(8, 512, 8, 4)               \
    X(16, 512, 8, 2)              \
    X(32, 256, 8, 2)              \
    X(64, 256, 4, 1)
#endif

int filtered_fused_supported(int M, int maxt) {
#define X(m_, t_, pf_, w_) \
    if (M == m_ && maxt == t_) return 1;
    FTD_FILTERED_VARIANTS(X)
#undef X
    return 0;
}

size_t filtered_fused_lds(int P, int S, int guard) {
    return guard ? (sizeof(double) * 6 + sizeof(int)) * (size_t)P * (size_t)S
                 : sizeof(double) * 4 * (size_t)P * (size_t)S;
}

hipError_t launch_filtered_fused(int M, int maxt, int version, int pass, const double* A,
                                 const double* B, const double* C, double* D, int n_sys, int n_row,
                                 int P, int S, int cyclic, int csh, const int* Jp, double* xbuf,
                                 double* partial, double eps_value, int eps_direct, int* Jsys,
                                 int guard, hipStream_t st) {
    const int lgS = log2i(S);
    const dim3 grid((n_sys + S - 1) / S), block(S * P);
    // the guard runs after the first pass of v1 only
    if (guard && pass == 1) return hipErrorInvalidValue;
    const size_t lds = guard ? filtered_fused_lds(P, S, 1)
                             : (pass == 2) ? 0 : filtered_fused_lds(P, S, 0);
#define LAUNCH(m_, t_, v_, p_, pf_, w_, g_)                                                   \
    do {                                                                                     \
        if (csh)                                                                             \
            hipLaunchKernelGGL((k_filtered_fused<m_, t_, v_, p_, pf_, w_, 1, g_>), grid, block, \
                               lds, st, A, B, C, D, n_sys, n_row, P, lgS, cyclic, Jp, xbuf,   \
                               partial, eps_value, eps_direct, Jsys);                        \
        else                                                                                 \
            hipLaunchKernelGGL((k_filtered_fused<m_, t_, v_, p_, pf_, w_, 0, g_>), grid, block, \
                               lds, st, A, B, C, D, n_sys, n_row, P, lgS, cyclic, Jp, xbuf,   \
                               partial, eps_value, eps_direct, Jsys);                        \
    } while (0)
#define X(m_, t_, pf_, w_)                                \
    if (M == m_ && maxt == t_) {                          \
        if (guard) {                                      \
            if (version == 2)                             \
                LAUNCH(m_, t_, 1, 0, pf_, 1, 1);          \
            else if (pass == 3)                           \
                LAUNCH(m_, t_, 1, 3, pf_, 1, 1);          \
            else                                          \
                LAUNCH(m_, t_, 1, 2, pf_, 1, 1);          \
        } else if (version == 2)                          \
            LAUNCH(m_, t_, 2, 0, pf_, w_, 0);             \
        else if (pass == 1)                               \
            LAUNCH(m_, t_, 1, 1, pf_, w_, 0);             \
        else if (pass == 3)                               \
            LAUNCH(m_, t_, 1, 3, pf_, w_, 0);             \
        else                                              \
            LAUNCH(m_, t_, 1, 2, pf_, w_, 0);             \
        return hipGetLastError();                         \
    }
    FTD_FILTERED_VARIANTS(X)
#undef X
#undef LAUNCH
    return hipErrorInvalidValue;
}

// -------- streamed (LDS ring, persistent) dispatch --------

#ifndef FTD_STREAM_VARIANTS
// X(M rows per partition, Q partitions, R ring depth, waves per SIMD)
#define FTD_STREAM_VARIANTS(X) \
    X(64, 8, 16, 1)            \
    X(32, 8, 8, 2)             \
    X(64, 4, 16, 1)            \
    X(32, 4, 8, 2)             \
    X(32, 16, 8, 2)
#endif

bool stream_guard_supported(int Q) { return Q <= 8; }

int stream_rows(int m_max, int Q) {
    int best = 0;
#define X(m_, q_, r_, w_) \
    if (Q == q_ && m_max <= m_ && (best == 0 || m_ < best)) best = m_;
    FTD_STREAM_VARIANTS(X)
#undef X
    return best;
}

int stream_ring_depth(int M, int Q) {
#define X(m_, q_, r_, w_) \
    if (M == m_ && Q == q_) return r_;
    FTD_STREAM_VARIANTS(X)
#undef X
    return 0;
}

bool stream_eligible(int n_sys, const void* A, const void* B, const void* C, const void* D) {
    const uintptr_t mis = ((uintptr_t)A | (uintptr_t)B | (uintptr_t)C | (uintptr_t)D) & 15u;
    return n_sys > 0 && n_sys % 32 == 0 && mis == 0;
}

// Workgroups of `kernel` resident on the whole device at once: the
// persistent grid. Cached per (kernel, device).
static int resident_grid(const void* kernel, int threads, int n_tiles) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    int total = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find({kernel, dev});
        if (it != cache.end()) total = it->second;
    }
    if (!total) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus < 1)
            cus = 1;
        total = per_cu * cus;
        std::lock_guard<std::mutex> lock(mu);
        cache[{kernel, dev}] = total;
    }
    return total < n_tiles ? total : n_tiles;
}

hipError_t launch_pascal_stream(int M, int Q, const double* A, const double* B, const double* C,
                                double* D, int n_sys, int n_row, int cyclic, hipStream_t st) {
    if (!stream_eligible(n_sys, A, B, C, D)) return hipErrorInvalidValue;
    const int n_tiles = n_sys / 32;
#define X(m_, q_, r_, w_)                                                                    \
    if (M == m_ && Q == q_) {                                                                \
        const void* k = cyclic ? (const void*)k_pascal_stream<m_, q_, r_, true, w_>          \
                               : (const void*)k_pascal_stream<m_, q_, r_, false, w_>;        \
        const dim3 grid(resident_grid(k, q_ * 32, n_tiles)), block(q_ * 32);                 \
        if (cyclic)                                                                          \
            hipLaunchKernelGGL((k_pascal_stream<m_, q_, r_, true, w_>), grid, block, 0, st, A, \
                               B, C, D, n_sys, n_row, n_tiles);                              \
        else                                                                                 \
            hipLaunchKernelGGL((k_pascal_stream<m_, q_, r_, false, w_>), grid, block, 0, st,   \
                               A, B, C, D, n_sys, n_row, n_tiles);                           \
        return hipGetLastError();                                                            \
    }
    FTD_STREAM_VARIANTS(X)
#undef X
    return hipErrorInvalidValue;
}

// FP32 rows are half the bytes of FP64 ones, so the ring is twice as deep
// (in rows) to keep as many bytes in flight
#ifndef FTD_STREAM_F32_VARIANTS
#define FTD_STREAM_F32_VARIANTS(X) \
    X(64, 8, 32, 1)                \
    X(32, 8, 16, 2)                \
    X(64, 4, 32, 1)                \
    X(32, 4, 16, 2)                \
    X(32, 16, 16, 2)
#endif

int stream_rows_f32(int m_max, int Q) {
    int best = 0;
#define X_(m_, q_, r_, w_) \
    if (Q == q_ && m_max <= m_ && (best == 0 || m_ < best)) best = m_;
    FTD_STREAM_F32_VARIANTS(X_)
#undef X_
    return best;
}

int stream_ring_depth_f32(int M, int Q) {
#define X_(m_, q_, r_, w_) \
    if (M == m_ && Q == q_) return r_;
    FTD_STREAM_F32_VARIANTS(X_)
#undef X_
    return 0;
}

hipError_t launch_pascal_stream_f32(int M, int Q, const float* A, const float* B, const float* C,
                                    const float* D, double* X, int n_sys, int n_row,
                                    hipStream_t st) {
    if (!stream_eligible(n_sys, A, B, C, D) || ((uintptr_t)X & 15u)) return hipErrorInvalidValue;
    const int n_tiles = n_sys / 32;
#define X_(m_, q_, r_, w_)                                                                   \
    if (M == m_ && Q == q_) {                                                                \
        const void* k = (const void*)k_pascal_stream_f32<m_, q_, r_, w_>;                    \
        const dim3 grid(resident_grid(k, q_ * 32, n_tiles)), block(q_ * 32);                 \
        hipLaunchKernelGGL((k_pascal_stream_f32<m_, q_, r_, w_>), grid, block, 0, st, A, B, C, D, \
                           X, n_sys, n_row, n_tiles);                                        \
        return hipGetLastError();                                                            \
    }
    FTD_STREAM_F32_VARIANTS(X_)
#undef X_
    return hipErrorInvalidValue;
}

hipError_t launch_filtered_stream(int M, int Q, int version, int pass, const double* A,
                                  const double* B, const double* C, double* D, int n_sys,
                                  int n_row, int cyclic, const int* Jp, double* xbuf,
                                  double* partial, double eps_value, int eps_direct, int* Jsys,
                                  int guard, hipStream_t st) {
    if (!stream_eligible(n_sys, A, B, C, D)) return hipErrorInvalidValue;
    if ((pass == 3 || guard) && !Jsys) return hipErrorInvalidValue;
    if (guard && pass == 1) return hipErrorInvalidValue;
    const int n_tiles = n_sys / 32;
#define LAUNCH(m_, q_, r_, w_, v_, p_, g_)                                                    \
    do {                                                                                     \
        const void* k = (const void*)k_filtered_stream<m_, q_, r_, v_, p_, w_, g_>;          \
        const dim3 grid(resident_grid(k, q_ * 32, n_tiles)), block(q_ * 32);                 \
        hipLaunchKernelGGL((k_filtered_stream<m_, q_, r_, v_, p_, w_, g_>), grid, block, 0, st, \
                           A, B, C, D, n_sys, n_row, n_tiles, cyclic, Jp, xbuf, partial,     \
                           eps_value, eps_direct, Jsys);                                     \
    } while (0)
    // the guard's variants: one wave per SIMD (its LDS flags and exact reduced
    // solve make them larger), none for 16 partitions (their ring fills the
    // LDS) and none for learned J (the network, the guard and 64 rows in
    // registers outgrow the register file; the fused kernel takes that case)
#define X(m_, q_, r_, w_)                         \
    if (M == m_ && Q == q_) {                     \
        if (guard) {                              \
            if (q_ > 8 || pass == 3)              \
                return hipErrorInvalidValue;      \
            else if (version == 2)                \
                LAUNCH(m_, q_, r_, 1, 1, 0, (q_ <= 8));  \
            else                                  \
                LAUNCH(m_, q_, r_, 1, 1, 2, (q_ <= 8));  \
        } else if (version == 2)                  \
            LAUNCH(m_, q_, r_, w_, 2, 0, 0);      \
        else if (pass == 1)                       \
            LAUNCH(m_, q_, r_, w_, 1, 1, 0);      \
        else if (pass == 3)                       \
            LAUNCH(m_, q_, r_, w_, 1, 3, 0);      \
        else                                      \
            LAUNCH(m_, q_, r_, w_, 1, 2, 0);      \
        return hipGetLastError();                 \
    }
    FTD_STREAM_VARIANTS(X)
#undef X
#undef LAUNCH
    return hipErrorInvalidValue;
}

// -------- general kernels --------

hipError_t launch_gen_thomas(const double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, hipStream_t st) {
    hipLaunchKernelGGL(k_gen_thomas, dim3((n_sys + 255) / 256), dim3(256), 0, st, A, B, C, D,
                       n_sys, n_row);
    return hipGetLastError();
}

hipError_t launch_gen_thomas_cyclic(const double* A, const double* B, double* C, double* D,
                                    double* E, int n_sys, int n_row, hipStream_t st) {
    hipLaunchKernelGGL(k_gen_thomas_cyclic, dim3((n_sys + 255) / 256), dim3(256), 0, st, A, B, C,
                       D, E, n_sys, n_row);
    return hipGetLastError();
}

hipError_t launch_gen_pascal(double* A, const double* B, double* C, double* D, int n_sys,
                             int n_row, int P, int cyclic, double* R, double* E, hipStream_t st) {
    const dim3 g2((n_sys + 255) / 256, P), g1((n_sys + 255) / 256);
    hipLaunchKernelGGL(k_gen_pascal_sweep, g2, dim3(256), 0, st, A, B, C, D, n_sys, n_row, P, R);
    hipLaunchKernelGGL(k_gen_pascal_reduced, g1, dim3(256), 0, st, R, E, n_sys, P, cyclic);
    hipLaunchKernelGGL(k_gen_pascal_update, g2, dim3(256), 0, st, A, C, D, n_sys, n_row, P, R);
    return hipGetLastError();
}

hipError_t launch_gen_filtered_sweep(int version, double* A, const double* B, double* C, double* D,
                                     int n_sys, int n_row, int P, const int* Jp, int learned,
                                     double eps_value, int eps_direct, int* Jsys, hipStream_t st) {
    const dim3 g2((n_sys + 255) / 256, P);
    if (learned && version != 1) return hipErrorInvalidValue;
    if (version == 2)
        hipLaunchKernelGGL((k_gen_filtered_sweep<2, false>), g2, dim3(256), 0, st, A, B, C, D,
                           n_sys, n_row, P, Jp, eps_value, eps_direct, Jsys);
    else if (learned)
        hipLaunchKernelGGL((k_gen_filtered_sweep<1, true>), g2, dim3(256), 0, st, A, B, C, D,
                           n_sys, n_row, P, Jp, eps_value, eps_direct, Jsys);
    else
        hipLaunchKernelGGL((k_gen_filtered_sweep<1, false>), g2, dim3(256), 0, st, A, B, C, D,
                           n_sys, n_row, P, Jp, eps_value, eps_direct, Jsys);
    return hipGetLastError();
}

hipError_t launch_gen_filtered_check(int learned, int guard, const double* A, const double* C,
                                     const double* D, int n_sys, int n_row, int P, int cyclic,
                                     const int* Jp, double* xbuf, int* Jsys, double eps_value,
                                     int eps_direct, double* gbuf, hipStream_t st) {
    const dim3 g((n_sys + 255) / 256);
#define LAUNCH(l_, g_)                                                                          \
    hipLaunchKernelGGL((k_gen_filtered_check<l_, g_>), g, dim3(256), 0, st, A, C, D, n_sys, n_row, \
                       P, cyclic, Jp, xbuf, Jsys, eps_value, eps_direct, gbuf)
    if (learned && guard)
        LAUNCH(true, true);
    else if (learned)
        LAUNCH(true, false);
    else if (guard)
        LAUNCH(false, true);
    else
        return hipErrorInvalidValue;
#undef LAUNCH
    return hipGetLastError();
}

hipError_t launch_gen_filtered_iface(const double* A, const double* C, const double* D, int n_sys,
                                     int n_row, int P, int cyclic, double* xbuf, double* partial,
                                     int* n_tiles, hipStream_t st) {
    const int nt = (n_sys + 255) / 256;
    *n_tiles = nt;
    hipLaunchKernelGGL(k_gen_filtered_iface, dim3(nt, P), dim3(256), 0, st, A, C, D, n_sys, n_row,
                       P, cyclic, xbuf, partial);
    return hipGetLastError();
}

hipError_t launch_gen_filtered_final(const double* A, const double* C, double* D, int n_sys,
                                     int n_row, int P, const int* Jp, const double* xbuf,
                                     const int* Jsys, hipStream_t st) {
    const dim3 g2((n_sys + 255) / 256, P);
    hipLaunchKernelGGL(k_gen_filtered_final, g2, dim3(256), 0, st, A, C, D, n_sys, n_row, P, Jp,
                       xbuf, Jsys);
    return hipGetLastError();
}

}  // namespace ftd
