/*
 * ftdma_oracle.c -- CPU restatement of the reference's batched tridiagonal
 * solvers: plain and cyclic Thomas, PaScaL (modified Thomas + exact reduced
 * solve) and Filtered TDMA v1/v2, each with the GPU and the CPU semantics of
 * the reference.
 *
 * TEST INFRASTRUCTURE ONLY. tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library as the checker or as the
 * timed CPU baseline; the product (filtered_tdma_amd/) never links it.
 *
 * Every function below names the reference lines it restates. Operation
 * order follows the reference statement by statement (same products, same
 * association, reciprocal-then-multiply where the reference does so and
 * division where it divides), so that with -ffp-contract=off the result is
 * bit-for-bit the reference's for the same inputs.
 *
 * Where the reference is undefined (float->int conversion of +-inf/NaN in
 * the J estimators, J<=0 in the CPU v2 driver, which walks past the array),
 * the restatement defines it and says how next to the code.
 */
#include "ftdma_oracle.h"

#include <limits.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[256];

static int fail(int code, const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}

const char* ftdo_last_error(void) { return g_err; }

/* Row j of array X with leading dimension ld. */
#define ROW(X, j) ((X) + (size_t)(j) * ld)

/* Column chunk used to split work over OpenMP threads. Systems are
 * independent, so every step below runs per chunk with no interaction. */
#define CHUNK 512

static int nchunks(int n_sys) { return (n_sys + CHUNK - 1) / CHUNK; }

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* libs/pascal_tdma/cpu/para_range.cpp:4-9 */
void ftdo_para_range(int start, int end, int nproc, int rank, int* ista, int* iend) {
    int len = end - start + 1;
    int base = len / nproc;
    int rem = len % nproc;
    *ista = start + rank * base + (rank < rem ? rank : rem);
    *iend = *ista + base - 1 + (rank < rem ? 1 : 0);
}

static void part_rows(int n_row, int P, int p, int* s, int* m) {
    int a, b;
    ftdo_para_range(0, n_row - 1, P, p, &a, &b);
    *s = a;
    *m = b - a + 1;
}

/* ------------------------------------------------------------------------ */
/* Plain Thomas on columns [c0,c1).                                          */
/* CPU: libs/pascal_tdma/cpu/tdma_local.cpp:6-52 (row 0 scaled by 1/b).     */
/* GPU: libs/pascal_tdma/gpu/tdma_local_cuda.cu:32-64 (row 0 divided by b). */
/* ------------------------------------------------------------------------ */
static void thomas_cols(const double* A, const double* B, double* C, double* D,
                        size_t ld, int c0, int c1, int n_row, int sem) {
    for (int i = c0; i < c1; ++i) {
        if (sem == FTDO_SEM_GPU) {
            D[i] = D[i] / B[i];
            C[i] = C[i] / B[i];
        } else {
            double inv_b = 1.0 / B[i];
            D[i] *= inv_b;
            C[i] *= inv_b;
        }
    }
    for (int j = 1; j < n_row; ++j) {
        const double* Aj = ROW(A, j);
        const double* Bj = ROW(B, j);
        double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        const double* Cm = ROW(C, j - 1);
        const double* Dm = ROW(D, j - 1);
        for (int i = c0; i < c1; ++i) {
            double r = 1.0 / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = r * (Dj[i] - Aj[i] * Dm[i]);
            Cj[i] = r * Cj[i];
        }
    }
    for (int j = n_row - 2; j >= 0; --j) {
        const double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        const double* Dp = ROW(D, j + 1);
        for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] - Cj[i] * Dp[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Sherman-Morrison cyclic Thomas on columns [c0,c1); E is [n_row x ld].     */
/* tdma_local.cpp:108-208 and tdma_local_cuda.cu:67-128 (same arithmetic).   */
/* ------------------------------------------------------------------------ */
static void cyclic_cols(const double* A, const double* B, double* C, double* D,
                        double* E, size_t ld, int c0, int c1, int n_row) {
    for (int j = 0; j < n_row; ++j) {
        double* Ej = ROW(E, j);
        for (int i = c0; i < c1; ++i) Ej[i] = 0.0;
    }
    {
        const double* A1 = ROW(A, 1);
        const double* CN = ROW(C, n_row - 1);
        double* E1 = ROW(E, 1);
        double* EN = ROW(E, n_row - 1);
        for (int i = c0; i < c1; ++i) {
            E1[i] = -A1[i];
            EN[i] = -CN[i];
        }
    }
    {
        const double* B1 = ROW(B, 1);
        double* C1 = ROW(C, 1);
        double* D1 = ROW(D, 1);
        double* E1 = ROW(E, 1);
        for (int i = c0; i < c1; ++i) {
            double inv_b = 1.0 / B1[i];
            D1[i] *= inv_b;
            E1[i] *= inv_b;
            C1[i] *= inv_b;
        }
    }
    for (int j = 2; j < n_row; ++j) {
        const double* Aj = ROW(A, j);
        const double* Bj = ROW(B, j);
        double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        double* Ej = ROW(E, j);
        const double* Cm = ROW(C, j - 1);
        const double* Dm = ROW(D, j - 1);
        const double* Em = ROW(E, j - 1);
        for (int i = c0; i < c1; ++i) {
            double r = 1.0 / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = r * (Dj[i] - Aj[i] * Dm[i]);
            Ej[i] = r * (Ej[i] - Aj[i] * Em[i]);
            Cj[i] = r * Cj[i];
        }
    }
    for (int j = n_row - 2; j >= 1; --j) {
        const double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        double* Ej = ROW(E, j);
        const double* Dp = ROW(D, j + 1);
        const double* Ep = ROW(E, j + 1);
        for (int i = c0; i < c1; ++i) {
            Dj[i] = Dj[i] - Cj[i] * Dp[i];
            Ej[i] = Ej[i] - Cj[i] * Ep[i];
        }
    }
    {
        const double* D1 = ROW(D, 1);
        const double* DN = ROW(D, n_row - 1);
        const double* E1 = ROW(E, 1);
        const double* EN = ROW(E, n_row - 1);
        for (int i = c0; i < c1; ++i) {
            D[i] = (D[i] - A[i] * DN[i] - C[i] * D1[i]) / (B[i] + A[i] * EN[i] + C[i] * E1[i]);
        }
    }
    for (int j = 1; j < n_row; ++j) {
        double* Dj = ROW(D, j);
        const double* Ej = ROW(E, j);
        for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] + D[i] * Ej[i];
    }
}

int ftdo_thomas(double* A, const double* B, double* C, double* D,
                int n_sys, int n_row, int sem, int nthreads) {
    if (n_sys < 0 || n_row < 1) return fail(-1, "thomas: need n_sys >= 0 and n_row >= 1");
    const size_t ld = (size_t)n_sys;
    set_threads(nthreads);
    const int nc = nchunks(n_sys);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        thomas_cols(A, B, C, D, ld, c0, c1, n_row, sem);
    }
    return 0;
}

int ftdo_thomas_cyclic(double* A, const double* B, double* C, double* D,
                       int n_sys, int n_row, int sem, int nthreads) {
    (void)sem; /* CPU and GPU cyclic Thomas share one operation order */
    if (n_sys < 0 || n_row < 3) return fail(-1, "thomas_cyclic: need n_row >= 3");
    const size_t ld = (size_t)n_sys;
    double* E = (double*)malloc(sizeof(double) * (size_t)n_row * ld + 8);
    if (!E) return fail(-3, "thomas_cyclic: out of memory");
    set_threads(nthreads);
    const int nc = nchunks(n_sys);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        cyclic_cols(A, B, C, D, E, ld, c0, c1, n_row);
    }
    free(E);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PaScaL modified Thomas on one partition (pointers at its first row).      */
/* CPU: pascal_tdma.cpp:131-183. GPU: tdma_local_cuda.cu:134-203 (rows 0,1  */
/* divided by b instead of scaled by 1/b). Writes the two reduced rows      */
/* (B == 1 implied) into R[0..5][ld]: a0,c0,d0 (row 0) and aN,cN,dN.         */
/* ------------------------------------------------------------------------ */
static void modthomas_cols(double* A, const double* B, double* C, double* D,
                           size_t ld, int c0, int c1, int m, int sem, double* R) {
    double* A0 = A; double* C0 = C; double* D0 = D;
    double* A1 = ROW(A, 1); double* C1 = ROW(C, 1); double* D1 = ROW(D, 1);
    const double* B0 = B; const double* B1 = ROW(B, 1);
    for (int i = c0; i < c1; ++i) {
        if (sem == FTDO_SEM_GPU) {
            A0[i] = A0[i] / B0[i]; C0[i] = C0[i] / B0[i]; D0[i] = D0[i] / B0[i];
            A1[i] = A1[i] / B1[i]; C1[i] = C1[i] / B1[i]; D1[i] = D1[i] / B1[i];
        } else {
            double r0 = 1.0 / B0[i];
            A0[i] *= r0; C0[i] *= r0; D0[i] *= r0;
            double r1 = 1.0 / B1[i];
            A1[i] *= r1; C1[i] *= r1; D1[i] *= r1;
        }
    }
    for (int j = 2; j < m; ++j) {
        double* Aj = ROW(A, j); const double* Bj = ROW(B, j);
        double* Cj = ROW(C, j); double* Dj = ROW(D, j);
        const double* Am = ROW(A, j - 1);
        const double* Cm = ROW(C, j - 1);
        const double* Dm = ROW(D, j - 1);
        for (int i = c0; i < c1; ++i) {
            double inv = 1.0 / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = inv * (Dj[i] - Aj[i] * Dm[i]);
            Cj[i] = inv * Cj[i];
            Aj[i] = -inv * Aj[i] * Am[i];
        }
    }
    for (int j = m - 3; j >= 1; --j) {
        double* Aj = ROW(A, j); double* Cj = ROW(C, j); double* Dj = ROW(D, j);
        const double* Ap = ROW(A, j + 1);
        const double* Cp = ROW(C, j + 1);
        const double* Dp = ROW(D, j + 1);
        for (int i = c0; i < c1; ++i) {
            Dj[i] = Dj[i] - Cj[i] * Dp[i];
            Aj[i] = Aj[i] - Cj[i] * Ap[i];
            Cj[i] = -Cj[i] * Cp[i];
        }
    }
    const double* AN = ROW(A, m - 1);
    const double* CN = ROW(C, m - 1);
    const double* DN = ROW(D, m - 1);
    for (int i = c0; i < c1; ++i) {
        double r = 1.0 / (1.0 - A1[i] * C0[i]);
        D0[i] = r * (D0[i] - C0[i] * D1[i]);
        A0[i] = r * A0[i];
        C0[i] = -r * C0[i] * C1[i];
        R[0 * ld + i] = A0[i]; R[1 * ld + i] = C0[i]; R[2 * ld + i] = D0[i];
        R[3 * ld + i] = AN[i]; R[4 * ld + i] = CN[i]; R[5 * ld + i] = DN[i];
    }
}

/* Reduced system of 2P rows ([row 0 of p, row m-1 of p] for p = 0..P-1,
 * pascal_tdma_kernels.cuh:50-62), B == 1, solved with tdma_many (or
 * tdma_cyclic_many) as pascal_tdma.cpp:201,322 and pascal_tdma_cuda.cu:277,
 * 389 do. Rows r: Ar, Cr, Dr each [2P x ld]; scratch E for the cyclic case. */
static void reduced_cols(double* Ar, double* Br, double* Cr, double* Dr, double* Er,
                         size_t ld, int c0, int c1, int n_rd, int cyclic, int sem) {
    if (cyclic)
        cyclic_cols(Ar, Br, Cr, Dr, Er, ld, c0, c1, n_rd);
    else
        thomas_cols(Ar, Br, Cr, Dr, ld, c0, c1, n_rd, sem);
}

/* Interior update x_j = d_j - (a_j x0 + c_j xN): pascal_tdma.cpp:212-226,
 * tdma_local_cuda.cu:205-224. */
static void update_cols(const double* A, const double* C, double* D, size_t ld,
                        int c0, int c1, int m, const double* x0, const double* xN) {
    double* DN = ROW(D, m - 1);
    for (int i = c0; i < c1; ++i) {
        D[i] = x0[i];
        DN[i] = xN[i];
    }
    for (int j = 1; j < m - 1; ++j) {
        const double* Aj = ROW(A, j);
        const double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] - (Aj[i] * x0[i] + Cj[i] * xN[i]);
    }
}

static int check_parts(int n_row, int P) {
    if (P < 1) return fail(-1, "partitions must be >= 1");
    if (P > 1) {
        int s, m;
        part_rows(n_row, P, P - 1, &s, &m); /* the last partitions are the shortest */
        if (m < 3) return fail(-2, "every partition needs at least 3 rows");
    }
    return 0;
}

/* partition p: rows starts[p] .. starts[p + 1] - 1 where the caller gives
 * the P + 1 starts, else the para_range split */
static void part_of_rows(int n_row, int P, const int* starts, int p, int* s, int* m) {
    if (starts) {
        *s = starts[p];
        *m = starts[p + 1] - starts[p];
    } else {
        part_rows(n_row, P, p, s, m);
    }
}

/* The 2P-row reduced system solved itself by PaScaL over G groups of rows
 * (para_range split, B == 1): modified Thomas per group, the 2G-row system
 * of the groups' end rows by Thomas, then the groups' interior update --
 * the order of the uneven-batch kernel's reduced solve for long systems
 * (ftdma_batched.hip). A2 holds 4 arrays of 2G rows and R2 6G rows. */
static void reduced_nested_cols(double* Ar, double* Br, double* Cr, double* Dr, size_t ld,
                                int c0, int c1, int n_rd, int G, int sem, double* A2,
                                double* R2) {
    double* B2 = A2 + ld * (size_t)(2 * G);
    double* C2 = B2 + ld * (size_t)(2 * G);
    double* D2 = C2 + ld * (size_t)(2 * G);
    for (int g = 0; g < G; ++g) {
        int s, m;
        part_rows(n_rd, G, g, &s, &m);
        double* R = R2 + (size_t)6 * g * ld;
        modthomas_cols(ROW(Ar, s), ROW(Br, s), ROW(Cr, s), ROW(Dr, s), ld, c0, c1, m, sem, R);
        for (int i = c0; i < c1; ++i) {
            A2[(size_t)(2 * g) * ld + i] = R[0 * ld + i];
            C2[(size_t)(2 * g) * ld + i] = R[1 * ld + i];
            D2[(size_t)(2 * g) * ld + i] = R[2 * ld + i];
            A2[(size_t)(2 * g + 1) * ld + i] = R[3 * ld + i];
            C2[(size_t)(2 * g + 1) * ld + i] = R[4 * ld + i];
            D2[(size_t)(2 * g + 1) * ld + i] = R[5 * ld + i];
            B2[(size_t)(2 * g) * ld + i] = 1.0;
            B2[(size_t)(2 * g + 1) * ld + i] = 1.0;
        }
    }
    thomas_cols(A2, B2, C2, D2, ld, c0, c1, 2 * G, sem);
    for (int g = 0; g < G; ++g) {
        int s, m;
        part_rows(n_rd, G, g, &s, &m);
        update_cols(ROW(Ar, s), ROW(Cr, s), ROW(Dr, s), ld, c0, c1, m,
                    D2 + (size_t)(2 * g) * ld, D2 + (size_t)(2 * g + 1) * ld);
    }
}

static int pascal_impl(double* A, const double* B, double* C, double* D, int n_sys, int n_row,
                       int P, int cyclic, int sem, int nthreads, int G, const int* starts);

int ftdo_pascal(double* A, const double* B, double* C, double* D,
                int n_sys, int n_row, int P, int cyclic, int sem, int nthreads) {
    return pascal_impl(A, B, C, D, n_sys, n_row, P, cyclic, sem, nthreads, 1, NULL);
}

int ftdo_pascal_nested(double* A, const double* B, double* C, double* D,
                       int n_sys, int n_row, int P, int G, int sem, int nthreads) {
    if (G < 1) return fail(-1, "pascal_nested: need G >= 1");
    if (G > 1 && P > 1) {
        int s, m;
        part_rows(2 * P, G, G - 1, &s, &m);
        if (m < 3) return fail(-2, "pascal_nested: every group needs at least 3 reduced rows");
    }
    return pascal_impl(A, B, C, D, n_sys, n_row, P, 0, sem, nthreads, G, NULL);
}

int ftdo_pascal_starts(double* A, const double* B, double* C, double* D, int n_sys, int n_row,
                       int P, const int* starts, int G, int sem, int nthreads) {
    if (P < 2) return fail(-1, "pascal_starts: need P >= 2");
    if (starts[0] != 0 || starts[P] != n_row)
        return fail(-1, "pascal_starts: the starts must run from 0 to n_row");
    for (int p = 0; p < P; ++p)
        if (starts[p + 1] - starts[p] < 3)
            return fail(-2, "pascal_starts: every partition needs at least 3 rows");
    if (G < 1) return fail(-1, "pascal_starts: need G >= 1");
    if (G > 1) {
        int s, m;
        part_rows(2 * P, G, G - 1, &s, &m);
        if (m < 3) return fail(-2, "pascal_starts: every group needs at least 3 reduced rows");
    }
    return pascal_impl(A, B, C, D, n_sys, n_row, P, 0, sem, nthreads, G, starts);
}

static int pascal_impl(double* A, const double* B, double* C, double* D, int n_sys, int n_row,
                       int P, int cyclic, int sem, int nthreads, int G, const int* starts) {
    int rc = starts ? 0 : check_parts(n_row, P);
    if (rc) return rc;
    if (P == 1) {
        /* pascal_tdma.cpp:116-119, pascal_tdma_cuda.cu:266-270 */
        return cyclic ? ftdo_thomas_cyclic(A, B, C, D, n_sys, n_row, sem, nthreads)
                      : ftdo_thomas(A, B, C, D, n_sys, n_row, sem, nthreads);
    }
    const size_t ld = (size_t)n_sys;
    const int n_rd = 2 * P;
    const int nested = G > 1 && !cyclic;
    const size_t extra = nested ? (size_t)(8 * G + 6 * G) : 0;
    double* buf =
        (double*)malloc(sizeof(double) * ld * ((size_t)(4 * n_rd + 6 * P) + extra) + 8);
    if (!buf) return fail(-3, "pascal: out of memory");
    double* Ar = buf;
    double* Br = Ar + ld * n_rd;
    double* Cr = Br + ld * n_rd;
    double* Dr = Cr + ld * n_rd;
    double* Rp = Dr + ld * n_rd; /* 6 rows per partition */
    double* A2 = Rp + ld * (size_t)(6 * P); /* nested: 4 arrays of 2G rows, then 6G rows */
    double* R2 = A2 + ld * (size_t)(8 * G);
    set_threads(nthreads);
    const int nc = nchunks(n_sys);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        for (int p = 0; p < P; ++p) {
            int s, m;
            part_of_rows(n_row, P, starts, p, &s, &m);
            double* R = Rp + (size_t)6 * p * ld;
            modthomas_cols(ROW(A, s), ROW(B, s), ROW(C, s), ROW(D, s), ld, c0, c1, m, sem, R);
            for (int i = c0; i < c1; ++i) {
                Ar[(size_t)(2 * p) * ld + i] = R[0 * ld + i];
                Cr[(size_t)(2 * p) * ld + i] = R[1 * ld + i];
                Dr[(size_t)(2 * p) * ld + i] = R[2 * ld + i];
                Ar[(size_t)(2 * p + 1) * ld + i] = R[3 * ld + i];
                Cr[(size_t)(2 * p + 1) * ld + i] = R[4 * ld + i];
                Dr[(size_t)(2 * p + 1) * ld + i] = R[5 * ld + i];
                Br[(size_t)(2 * p) * ld + i] = 1.0;
                Br[(size_t)(2 * p + 1) * ld + i] = 1.0;
            }
        }
        /* the cyclic reduced solve needs an E scratch of 2P rows: reuse Rp
         * (6P >= 2P rows), which is no longer read after the copy above */
        if (nested)
            reduced_nested_cols(Ar, Br, Cr, Dr, ld, c0, c1, n_rd, G, sem, A2, R2);
        else
            reduced_cols(Ar, Br, Cr, Dr, Rp, ld, c0, c1, n_rd, cyclic, sem);
        for (int p = 0; p < P; ++p) {
            int s, m;
            part_of_rows(n_row, P, starts, p, &s, &m);
            update_cols(ROW(A, s), ROW(C, s), ROW(D, s), ld, c0, c1, m,
                        Dr + (size_t)(2 * p) * ld, Dr + (size_t)(2 * p + 1) * ld);
        }
    }
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP32 elimination with FP64 refinement (no counterpart in the reference:  */
/* the mixed-precision mode of the engine's FP32-storage path). One solve:  */
/* the GPU-semantics modified Thomas of modthomas_cols in float arithmetic  */
/* on float rows, the reduced rows widened to double and solved as in       */
/* ftdo_pascal, the interior update in double on the widened values.        */
/* ------------------------------------------------------------------------ */
static void modthomas_f32_cols(float* A, const float* B, float* C, float* D,
                               size_t ld, int c0, int c1, int m, double* R) {
    float* A1 = ROW(A, 1); float* C1 = ROW(C, 1); float* D1 = ROW(D, 1);
    const float* B1 = ROW(B, 1);
    for (int i = c0; i < c1; ++i) {
        A[i] = A[i] / B[i]; C[i] = C[i] / B[i]; D[i] = D[i] / B[i];
        A1[i] = A1[i] / B1[i]; C1[i] = C1[i] / B1[i]; D1[i] = D1[i] / B1[i];
    }
    for (int j = 2; j < m; ++j) {
        float* Aj = ROW(A, j); const float* Bj = ROW(B, j);
        float* Cj = ROW(C, j); float* Dj = ROW(D, j);
        const float* Am = ROW(A, j - 1);
        const float* Cm = ROW(C, j - 1);
        const float* Dm = ROW(D, j - 1);
        for (int i = c0; i < c1; ++i) {
            float inv = 1.0f / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = inv * (Dj[i] - Aj[i] * Dm[i]);
            Cj[i] = inv * Cj[i];
            Aj[i] = -inv * Aj[i] * Am[i];
        }
    }
    for (int j = m - 3; j >= 1; --j) {
        float* Aj = ROW(A, j); float* Cj = ROW(C, j); float* Dj = ROW(D, j);
        const float* Ap = ROW(A, j + 1);
        const float* Cp = ROW(C, j + 1);
        const float* Dp = ROW(D, j + 1);
        for (int i = c0; i < c1; ++i) {
            Dj[i] = Dj[i] - Cj[i] * Dp[i];
            Aj[i] = Aj[i] - Cj[i] * Ap[i];
            Cj[i] = -Cj[i] * Cp[i];
        }
    }
    const float* AN = ROW(A, m - 1);
    const float* CN = ROW(C, m - 1);
    const float* DN = ROW(D, m - 1);
    for (int i = c0; i < c1; ++i) {
        float r = 1.0f / (1.0f - A1[i] * C[i]);
        float d0 = r * (D[i] - C[i] * D1[i]);
        float a0 = r * A[i];
        float cc0 = -r * C[i] * C1[i];
        R[0 * ld + i] = (double)a0; R[1 * ld + i] = (double)cc0; R[2 * ld + i] = (double)d0;
        R[3 * ld + i] = (double)AN[i]; R[4 * ld + i] = (double)CN[i]; R[5 * ld + i] = (double)DN[i];
        D[i] = d0; A[i] = a0; C[i] = cc0;
    }
}

int ftdo_pascal_ir(const float* A, const float* B, const float* C, const float* D, double* X,
                   int n_sys, int n_row, int Q, int sweeps, int nthreads) {
    if (Q < 2) return fail(-1, "pascal_ir: need Q >= 2 partitions");
    if (sweeps < 0) return fail(-1, "pascal_ir: sweeps must be >= 0");
    int rc = check_parts(n_row, Q);
    if (rc) return rc;
    const size_t ld = (size_t)n_sys;
    const size_t nn = ld * (size_t)n_row;
    const int n_rd = 2 * Q;
    float* fbuf = (float*)malloc(sizeof(float) * nn * 3 + 16);
    double* buf = (double*)malloc(sizeof(double) * ld * ((size_t)(4 * n_rd + 6 * Q)) + 8);
    if (!fbuf || !buf) {
        free(fbuf);
        free(buf);
        return fail(-3, "pascal_ir: out of memory");
    }
    float* Af = fbuf;
    float* Cf = Af + nn;
    float* Df = Cf + nn;
    double* Ar = buf;
    double* Br = Ar + ld * n_rd;
    double* Cr = Br + ld * n_rd;
    double* Dr = Cr + ld * n_rd;
    double* Rp = Dr + ld * n_rd;
    set_threads(nthreads);
    const int nc = nchunks(n_sys);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        for (int sw = 0; sw <= sweeps; ++sw) {
            /* right-hand side: D, then the FP64 residual of the current x,
             * rounded to float (every row's before any x changes) */
            for (int j = 0; j < n_row; ++j) {
                for (int i = c0; i < c1; ++i) {
                    const size_t o = (size_t)j * ld + i;
                    Af[o] = A[o];
                    Cf[o] = C[o];
                    if (sw == 0) {
                        Df[o] = D[o];
                    } else {
                        double t = (double)B[o] * X[o];
                        if (j > 0) t = (double)A[o] * X[o - ld] + t;
                        if (j < n_row - 1) t = t + (double)C[o] * X[o + ld];
                        Df[o] = (float)((double)D[o] - t);
                    }
                }
            }
            for (int p = 0; p < Q; ++p) {
                int s, m;
                part_rows(n_row, Q, p, &s, &m);
                double* R = Rp + (size_t)6 * p * ld;
                modthomas_f32_cols(ROW(Af, s), ROW(B, s), ROW(Cf, s), ROW(Df, s), ld, c0, c1, m, R);
                for (int i = c0; i < c1; ++i) {
                    Ar[(size_t)(2 * p) * ld + i] = R[0 * ld + i];
                    Cr[(size_t)(2 * p) * ld + i] = R[1 * ld + i];
                    Dr[(size_t)(2 * p) * ld + i] = R[2 * ld + i];
                    Ar[(size_t)(2 * p + 1) * ld + i] = R[3 * ld + i];
                    Cr[(size_t)(2 * p + 1) * ld + i] = R[4 * ld + i];
                    Dr[(size_t)(2 * p + 1) * ld + i] = R[5 * ld + i];
                    Br[(size_t)(2 * p) * ld + i] = 1.0;
                    Br[(size_t)(2 * p + 1) * ld + i] = 1.0;
                }
            }
            thomas_cols(Ar, Br, Cr, Dr, ld, c0, c1, n_rd, FTDO_SEM_GPU);
            for (int p = 0; p < Q; ++p) {
                int s, m;
                part_rows(n_row, Q, p, &s, &m);
                const double* x0 = Dr + (size_t)(2 * p) * ld;
                const double* xN = Dr + (size_t)(2 * p + 1) * ld;
                for (int j = 0; j < m; ++j) {
                    const float* Aj = ROW(Af, s + j);
                    const float* Cj = ROW(Cf, s + j);
                    const float* Dj = ROW(Df, s + j);
                    double* Xj = ROW(X, s + j);
                    for (int i = c0; i < c1; ++i) {
                        double v;
                        if (j == 0)
                            v = x0[i];
                        else if (j == m - 1)
                            v = xN[i];
                        else
                            v = (double)Dj[i] - ((double)Aj[i] * x0[i] + (double)Cj[i] * xN[i]);
                        Xj[i] = sw == 0 ? v : Xj[i] + v;
                    }
                }
            }
        }
    }
    free(fbuf);
    free(buf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Spectral bounds from column 0 (tdma_backend.cpp:33-44,                   */
/* filtered_tdma_kernels.cuh:46-67): |A/B|, |C/B| per row, 0 where B == 0,  */
/* trailing slot n_row = 0.                                                  */
/* ------------------------------------------------------------------------ */
void ftdo_set_rho(const double* A, const double* B, const double* C,
                  int n_sys, int n_row, double* A_rho, double* C_rho) {
    const size_t ld = (size_t)n_sys;
    for (int k = 0; k < n_row; ++k) {
        double bk = B[(size_t)k * ld];
        A_rho[k] = (bk != 0.0) ? fabs(A[(size_t)k * ld] / bk) : 0.0;
        C_rho[k] = (bk != 0.0) ? fabs(C[(size_t)k * ld] / bk) : 0.0;
    }
    A_rho[n_row] = 0.0;
    C_rho[n_row] = 0.0;
}

/* float -> int truncation toward zero. The reference writes static_cast<int>
 * / (int) of values that can be +-inf or NaN (eps/B with B == 0), which C++
 * leaves undefined; here it saturates (NaN -> 0), which is what the CUDA and
 * AMD conversion instructions do. Returned as int64 so that "+ 1" cannot
 * overflow. */
static int64_t trunc_sat(double x) {
    if (x != x) return 0;
    if (x >= 2147483647.0) return INT_MAX;
    if (x <= -2147483648.0) return INT_MIN;
    return (int64_t)(int)x;
}

/* ftdo_cal_J: filtered_tdma.cpp:63-115 (CPU) and filtered_tdma_kernels.cuh:
 * 245-259 (GPU). rho==0 or rho>=0.5 -> n_row-1 (full propagation). */
int ftdo_cal_J(double rho, double bound, double eps, int n_row, int which) {
    if (rho == 0.0 || rho >= 0.5) return n_row - 1;
    double lambda_p = (which == FTDO_J_GPU_RHS) ? (1.0 + sqrt(1.0 - 4.0 * rho * rho)) * 0.5
                                                : (1.0 + sqrt(1.0 - 4.0 * rho * rho)) / 2.0;
    double q = rho / lambda_p;
    double Bv;
    if (which == FTDO_J_CPU_V1)
        Bv = bound * q / (1.0 - q);
    else
        Bv = q * (2.0 + q) / ((1.0 - q) * (1.0 - 2.0 * rho)) * bound;
    int64_t J = trunc_sat(log(eps / Bv) / log(q)) + 1;
    if (J > n_row - 1) J = n_row - 1;
    if (which == FTDO_J_GPU_RHS && J < 0) J = 0;
    return (int)J;
}

static double part_rho(const double* A_rho, const double* C_rho, int s, int m) {
    /* k in [1, m-1] of the partition's local rows (filtered_tdma.cpp:68,
     * filtered_tdma_kernels.cuh:232) */
    double rho = 0.0;
    if (!A_rho || !C_rho) return 0.0;
    for (int k = 1; k < m; ++k) {
        double a = fabs(A_rho[s + k]), c = fabs(C_rho[s + k]);
        double v = a > c ? a : c;
        if (v > rho) rho = v;
    }
    return rho;
}

/* ---------------------------- Filtered steps ---------------------------- */

/* Forward sweep. v1: filtered_tdma_kernels.hpp:35-65 / .cuh:69-94 (full).
 * v2 GPU: .cuh:120-155 literally, including J == 0, where the second loop
 * starts at row 1 and eliminates the already-normalised row 1 once more.
 * v2 CPU: filtered_tdma_kernels.hpp:68-110 (the reference's solve keeps J >= 1 here). */
static void ffwd_cols(double* A, const double* B, double* C, double* D, size_t ld,
                      int c0, int c1, int m, int J_full /* last row with a-update */,
                      int gpu_v2) {
    double* A1 = ROW(A, 1); double* C1 = ROW(C, 1); double* D1 = ROW(D, 1);
    const double* B1 = ROW(B, 1);
    for (int i = c0; i < c1; ++i) {
        double r0 = 1.0 / B[i];
        A[i] *= r0; C[i] *= r0; D[i] *= r0;
        double r1 = 1.0 / B1[i];
        A1[i] *= r1; C1[i] *= r1; D1[i] *= r1;
    }
    int j = 2;
    for (; j <= J_full && j < m; ++j) {
        double* Aj = ROW(A, j); const double* Bj = ROW(B, j);
        double* Cj = ROW(C, j); double* Dj = ROW(D, j);
        const double* Am = ROW(A, j - 1);
        const double* Cm = ROW(C, j - 1);
        const double* Dm = ROW(D, j - 1);
        for (int i = c0; i < c1; ++i) {
            double inv = 1.0 / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = inv * (Dj[i] - Aj[i] * Dm[i]);
            Cj[i] = inv * Cj[i];
            Aj[i] = -inv * Aj[i] * Am[i];
        }
    }
    if (gpu_v2 && J_full == 0) {
        /* k_fwd_pass_v2 with J == 0: its skip-A loop runs j = 1..m-1 with the
         * sliding window already holding row 1 */
        const double* Bj = B1;
        for (int i = c0; i < c1; ++i) {
            double c_prev = C1[i], d_prev = D1[i];
            double Aj = A1[i];
            double inv = 1.0 / (Bj[i] - Aj * c_prev);
            D1[i] = inv * (D1[i] - Aj * d_prev);
            C1[i] = inv * C1[i];
        }
        j = 2;
    }
    for (; j < m; ++j) {
        const double* Aj = ROW(A, j); const double* Bj = ROW(B, j);
        double* Cj = ROW(C, j); double* Dj = ROW(D, j);
        const double* Cm = ROW(C, j - 1);
        const double* Dm = ROW(D, j - 1);
        for (int i = c0; i < c1; ++i) {
            double inv = 1.0 / (Bj[i] - Aj[i] * Cm[i]);
            Dj[i] = inv * (Dj[i] - Aj[i] * Dm[i]);
            Cj[i] = inv * Cj[i];
        }
    }
}

/* Backward sweep + pack_reduced (row 0 keeps d0, a0; c0 coupling dropped).
 * v1: filtered_tdma_kernels.hpp:117-135,178-189 / .cuh:96-118.
 * v2 GPU: .cuh:157-194: a-update only for j < J seeded from A[J] (J >= 2),
 *   c-update only for j > lo = (m-2)-J.
 * v2 CPU: filtered_tdma_kernels.hpp:138-172 (J <= m-2 by the reference's clamp). */
static void fbwd_cols(double* A, double* C, double* D, size_t ld, int c0, int c1, int m,
                      int version, int J, int sem) {
    if (version == 1) {
        for (int j = m - 3; j >= 1; --j) {
            double* Aj = ROW(A, j); double* Cj = ROW(C, j); double* Dj = ROW(D, j);
            const double* Ap = ROW(A, j + 1);
            const double* Cp = ROW(C, j + 1);
            const double* Dp = ROW(D, j + 1);
            for (int i = c0; i < c1; ++i) {
                double cj = Cj[i];
                Dj[i] = Dj[i] - cj * Dp[i];
                Aj[i] = Aj[i] - cj * Ap[i];
                Cj[i] = -cj * Cp[i];
            }
        }
    } else if (sem == FTDO_SEM_GPU) {
        const int lo = (m - 2) - J;
        for (int i = c0; i < c1; ++i) {
            double d_next = ROW(D, m - 2)[i];
            double c_next = ROW(C, m - 2)[i];
            double a_next = (J >= 2) ? ROW(A, J)[i] : 0.0;
            for (int j = m - 3; j >= 1; --j) {
                double cj = ROW(C, j)[i];
                double d_new = ROW(D, j)[i] - cj * d_next;
                ROW(D, j)[i] = d_new;
                d_next = d_new;
                if (J >= 2 && j < J) {
                    double a_new = ROW(A, j)[i] - cj * a_next;
                    ROW(A, j)[i] = a_new;
                    a_next = a_new;
                }
                if (j > lo) {
                    double c_new = -cj * c_next;
                    ROW(C, j)[i] = c_new;
                    c_next = c_new;
                }
            }
        }
    } else {
        const int lo = (m - 2) - J;
        for (int j = m - 3; j >= 1; --j) {
            const double* Cj = ROW(C, j); double* Dj = ROW(D, j);
            const double* Dp = ROW(D, j + 1);
            for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] - Cj[i] * Dp[i];
        }
        if (J >= 2) {
            for (int j = J - 1; j >= 1; --j) {
                double* Aj = ROW(A, j); const double* Cj = ROW(C, j);
                const double* Ap = ROW(A, j + 1);
                for (int i = c0; i < c1; ++i) Aj[i] = Aj[i] - Cj[i] * Ap[i];
            }
        }
        for (int j = m - 3; j >= lo + 1; --j) {
            double* Cj = ROW(C, j); const double* Cp = ROW(C, j + 1);
            for (int i = c0; i < c1; ++i) Cj[i] = Cj[i] * -Cp[i];
        }
    }
    /* pack_reduced: filtered_tdma_kernels.hpp:178-189, .cuh:114-117,187-193 */
    const double* A1 = ROW(A, 1);
    const double* D1 = ROW(D, 1);
    for (int i = c0; i < c1; ++i) {
        double r = 1.0 / (1.0 - A1[i] * C[i]);
        D[i] = r * (D[i] - C[i] * D1[i]);
        A[i] = r * A[i];
    }
}

/* Truncated correction: filtered_tdma_kernels.hpp:227-250 / .cuh:196-218. */
static void ffinal_cols(const double* A, const double* C, double* D, size_t ld,
                        int c0, int c1, int m, int left_end, int right_beg) {
    const double* DN = ROW(D, m - 1);
    for (int j = 1; j <= left_end; ++j) {
        const double* Aj = ROW(A, j);
        double* Dj = ROW(D, j);
        for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] - Aj[i] * D[i];
    }
    for (int j = right_beg; j < m - 1; ++j) {
        const double* Cj = ROW(C, j);
        double* Dj = ROW(D, j);
        for (int i = c0; i < c1; ++i) Dj[i] = Dj[i] - Cj[i] * DN[i];
    }
}

/* Sweeps of every partition, the exchange rows and the interface 2x2
 * solve (rows 0 and m-1 of each partition hold x_0, x_{m-1} afterwards).
 * Jp: the P values of J (v2 only). */
static void sweeps_iface(double* A, const double* B, double* C, double* D, size_t ld,
                         int n_sys, int P, const int* sv, const int* mv, int version,
                         const int* Jp, int sem, int cyclic, double* xbuf) {
    const int nc = nchunks(n_sys);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            int Jfull = (version == 1) ? m : Jp[p];
            ffwd_cols(ROW(A, s), ROW(B, s), ROW(C, s), ROW(D, s), ld, c0, c1, m, Jfull,
                      version == 2 && sem == FTDO_SEM_GPU);
            fbwd_cols(ROW(A, s), ROW(C, s), ROW(D, s), ld, c0, c1, m, version,
                      version == 2 ? Jp[p] : 0, sem);
            double* sr = xbuf + (size_t)(4 * p) * ld;
            const double* CN = ROW(C, s + m - 1);
            const double* DN = ROW(D, s + m - 1);
            const double* A0 = ROW(A, s);
            const double* D0 = ROW(D, s);
            for (int i = c0; i < c1; ++i) {
                sr[0 * ld + i] = CN[i]; /* send right: C */
                sr[1 * ld + i] = DN[i]; /* send right: D */
                sr[2 * ld + i] = A0[i]; /* send left : A */
                sr[3 * ld + i] = D0[i]; /* send left : D */
            }
        }
        /* interface 2x2 blocks: filtered_tdma_kernels.hpp:197-218, .cuh:284-303 */
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            int has_left = cyclic || p > 0;
            int has_right = cyclic || p < P - 1;
            int pl = (p + P - 1) % P, pr = (p + 1) % P;
            const double* C_left = xbuf + (size_t)(4 * pl + 0) * ld;
            const double* D_left = xbuf + (size_t)(4 * pl + 1) * ld;
            const double* A_right = xbuf + (size_t)(4 * pr + 2) * ld;
            const double* D0_right = xbuf + (size_t)(4 * pr + 3) * ld;
            double* A0 = ROW(A, s);
            double* D0 = ROW(D, s);
            const double* CN = ROW(C, s + m - 1);
            double* DN = ROW(D, s + m - 1);
            if (has_left)
                for (int i = c0; i < c1; ++i)
                    D0[i] = (D0[i] - A0[i] * D_left[i]) / (1.0 - C_left[i] * A0[i]);
            if (has_right)
                for (int i = c0; i < c1; ++i)
                    DN[i] = (DN[i] - CN[i] * D0_right[i]) / (1.0 - A_right[i] * CN[i]);
        }
    }
}

int ftdo_filtered(double* A, const double* B, double* C, double* D,
                  int n_sys, int n_row, int P, int version, int cyclic, int sem,
                  const double* A_rho, const double* C_rho,
                  double eps_value, int eps_direct,
                  const int* J_force, int* J_out, int nthreads) {
    int rc = check_parts(n_row, P);
    if (rc) return rc;
    if (version != 1 && version != 2) return fail(-1, "filtered: version must be 1 or 2");
    if (P == 1) {
        /* filtered_tdma.cpp:172,194,219,242; filtered_tdma_cuda.cu:108,135,170,199 */
        if (J_out) J_out[0] = -1;
        return cyclic ? ftdo_thomas_cyclic(A, B, C, D, n_sys, n_row, sem, nthreads)
                      : ftdo_thomas(A, B, C, D, n_sys, n_row, sem, nthreads);
    }
    const size_t ld = (size_t)n_sys;
    /* exchange buffers: per partition [C_{m-1} | D_{m-1}] (to the right) and
     * [A_0 | D_0] (to the left): filtered_tdma.cpp:151-163, .cuh:263-282 */
    double* xbuf = (double*)malloc(sizeof(double) * ld * (size_t)(4 * P) + 8);
    int* Jp = (int*)malloc(sizeof(int) * (size_t)P);
    double* maxb = (double*)malloc(sizeof(double) * (size_t)P);
    if (!xbuf || !Jp || !maxb) {
        free(xbuf); free(Jp); free(maxb);
        return fail(-3, "filtered: out of memory");
    }
    set_threads(nthreads);
    const int nc = nchunks(n_sys);
    int sv[4096], mv[4096];
    if (P > 4096) {
        free(xbuf); free(Jp); free(maxb);
        return fail(-1, "filtered: at most 4096 partitions");
    }
    for (int p = 0; p < P; ++p) part_rows(n_row, P, p, &sv[p], &mv[p]);
    double epsv[4096];
    for (int p = 0; p < P; ++p) {
        /* eps_ = eps_constant / (n_row_local * nprocs)^2 (filtered_tdma.cpp:50-51) */
        double Nl = (double)(mv[p] * P);
        epsv[p] = eps_direct ? eps_value : eps_value / (Nl * Nl);
    }

    /* v2: J before the sweeps */
    if (version == 2) {
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            double mb = 0.0;
            if (sem == FTDO_SEM_GPU) {
                /* max over systems of |D row 0| and |D row m-1| (.cuh:235-244) */
                const double* D0 = ROW(D, s);
                const double* DN = ROW(D, s + m - 1);
                for (int i = 0; i < n_sys; ++i) {
                    double a = fabs(D0[i]), b = fabs(DN[i]);
                    if (a > mb) mb = a;
                    if (b > mb) mb = b;
                }
            } else {
                /* max over every local row and system (filtered_tdma.cpp:75-80) */
                for (int j = 0; j < m; ++j) {
                    const double* Dj = ROW(D, s + j);
                    for (int i = 0; i < n_sys; ++i) {
                        double a = fabs(Dj[i]);
                        if (a > mb) mb = a;
                    }
                }
            }
            int J;
            if (J_force) {
                J = J_force[p];
            } else if (sem == FTDO_SEM_GPU) {
                J = ftdo_cal_J(part_rho(A_rho, C_rho, s, m), mb, epsv[p], m, FTDO_J_GPU_RHS);
            } else {
                J = ftdo_cal_J(part_rho(A_rho, C_rho, s, m), mb, epsv[p], m, FTDO_J_CPU_RHS);
                if (J > m - 2) J = m - 2; /* filtered_tdma.cpp:224 */
                if (J < 1) J = 1;         /* the reference is undefined for J < 1 here */
            }
            Jp[p] = J;
        }
    }

    sweeps_iface(A, B, C, D, ld, n_sys, P, sv, mv, version, Jp, sem, cyclic, xbuf);

    /* v1: J after the interface solve */
    if (version == 1) {
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            const double* D0 = ROW(D, s);
            const double* DN = ROW(D, s + m - 1);
            double m0 = 0.0, mN = 0.0;
            for (int i = 0; i < n_sys; ++i) {
                double a = fabs(D0[i]), b = fabs(DN[i]);
                if (a > m0) m0 = a;
                if (b > mN) mN = b;
            }
            int J;
            if (J_force)
                J = J_force[p];
            else if (sem == FTDO_SEM_GPU)
                J = ftdo_cal_J(part_rho(A_rho, C_rho, s, m), m0 > mN ? m0 : mN, epsv[p], m,
                               FTDO_J_GPU_RHS);
            else
                J = ftdo_cal_J(part_rho(A_rho, C_rho, s, m), m0 + mN, epsv[p], m, FTDO_J_CPU_V1);
            Jp[p] = J;
        }
    }

#pragma omp parallel for schedule(static)
    for (int k = 0; k < nc; ++k) {
        int c0 = k * CHUNK, c1 = c0 + CHUNK < n_sys ? c0 + CHUNK : n_sys;
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p], J = Jp[p];
            int left_end, right_beg;
            if (version == 2 && sem == FTDO_SEM_CPU) {
                /* filtered_tdma.cpp:236-237 */
                left_end = J;
                right_beg = (m - 2) - J + 1 > 1 ? (m - 2) - J + 1 : 1;
            } else {
                /* filtered_tdma.cpp:185-186 and k_final_pass .cuh:206-208 */
                left_end = J < m - 2 ? J : m - 2;
                right_beg = (m - 1) - J > 1 ? (m - 1) - J : 1;
            }
            ffinal_cols(ROW(A, s), ROW(C, s), ROW(D, s), ld, c0, c1, m, left_end, right_beg);
        }
    }
    if (J_out)
        for (int p = 0; p < P; ++p) J_out[p] = Jp[p];
    free(xbuf);
    free(Jp);
    free(maxb);
    return 0;
}

/* ------------------------- learned-J filtered mode ------------------------
 * No counterpart in the reference: the FILTERED solve (v1 sweeps and
 * interface, filtered_tdma_kernels.cuh:69-118,284-303) with J chosen per
 * (system, partition) by a 5 -> H -> H -> 1 ReLU network in FP32 instead
 * of one cal_J per partition, the first dropped correction term of each side
 * checked against eps_ (J = m-1 where it exceeds it), then the truncated
 * correction of k_final_pass (.cuh:196-218) with that J. The network's
 * weights come in as one flat array (W1[5][H], b1[H], W2[H][H], b2[H],
 * W3[H], b3) and every multiply-add is one fmaf, each sum in index order. */

/* lg|x| = e + f for |x| = (1 + f) 2^e, read off frexp; lg 0 = -1100. */
static float lg_feature(double x) {
    x = fabs(x);
    if (x == 0.0) return -1100.0f;
    int e;
    const double f = frexp(x, &e);
    return (float)((double)(e - 1) + (2.0 * f - 1.0));
}

static int mlp_J(const float* w, int H, int margin, const float f[5]) {
    const float* W1 = w;
    const float* b1 = W1 + 5 * H;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const float* b3 = W3 + H;
    float acc[64];
    for (int l = 0; l < H; ++l) acc[l] = b2[l];
    for (int k = 0; k < H; ++k) {
        float h = b1[k];
        for (int q = 0; q < 5; ++q) h = fmaf(W1[q * H + k], f[q], h);
        h = h > 0.0f ? h : 0.0f;
        for (int l = 0; l < H; ++l) acc[l] = fmaf(W2[k * H + l], h, acc[l]);
    }
    float out = b3[0];
    for (int l = 0; l < H; ++l) out = fmaf(W3[l], acc[l] > 0.0f ? acc[l] : 0.0f, out);
    const float cj = ceilf(out);
    const float lim = 1.0e6f;
    return (int)(cj > lim ? lim : (cj < -lim ? -lim : cj)) + margin;
}

int ftdo_mlp_J(const float* weights, int hidden, int margin, const float* features) {
    if (hidden < 1 || hidden > 64) return INT_MIN;
    return mlp_J(weights, hidden, margin, features);
}

float ftdo_lg(double x) { return lg_feature(x); }

/* Whether truncating partition (s, m) of system i at depth J drops a term
 * above eps: the first a-term |a'_{kL} x0| or the first c-term |c'_{kR} xN|
 * (kL = min(J, m-2) + 1, kR = max(m-1-J, 1) - 1), on the swept A, C. */
static int dropped_exceeds(const double* A, const double* C, size_t ld, int s, int m, size_t i,
                           int J, double x0, double xN, double eps) {
    const int kL = (J < m - 2 ? J : m - 2) + 1;
    const int kR = ((m - 1) - J > 1 ? (m - 1) - J : 1) - 1;
    const double ta = kL <= m - 2 ? fabs(ROW(A, s + kL)[i] * x0) : 0.0;
    const double tc = kR >= 1 ? fabs(ROW(C, s + kR)[i] * xN) : 0.0;
    return ta > eps || tc > eps;
}

int ftdo_filtered_learned(double* A, const double* B, double* C, double* D,
                          int n_sys, int n_row, int P, int cyclic,
                          double eps_value, int eps_direct,
                          const float* weights, int hidden, int margin, int* J_out,
                          int nthreads) {
    int rc = check_parts(n_row, P);
    if (rc) return rc;
    if (!weights) return fail(-1, "filtered_learned: weights needed");
    if (hidden < 1 || hidden > 64) return fail(-1, "filtered_learned: hidden width in [1, 64]");
    if (P == 1) {
        if (J_out)
            for (int i = 0; i < n_sys; ++i) J_out[i] = -1;
        return cyclic ? ftdo_thomas_cyclic(A, B, C, D, n_sys, n_row, FTDO_SEM_GPU, nthreads)
                      : ftdo_thomas(A, B, C, D, n_sys, n_row, FTDO_SEM_GPU, nthreads);
    }
    if (P > 4096) return fail(-1, "filtered_learned: at most 4096 partitions");
    const size_t ld = (size_t)n_sys;
    int sv[4096], mv[4096];
    for (int p = 0; p < P; ++p) part_rows(n_row, P, p, &sv[p], &mv[p]);
    double* xbuf = (double*)malloc(sizeof(double) * ld * (size_t)(4 * P) + 8);
    float* feat = (float*)malloc(sizeof(float) * ld * (size_t)P * 4 + 8);
    if (!xbuf || !feat) {
        free(xbuf); free(feat);
        return fail(-3, "filtered_learned: out of memory");
    }
    set_threads(nthreads);
    /* features of the raw rows: max over rows 1..m-1 of |a/b|, |c/b|
     * (clamped to 1), lg|d_0|, lg|d_m-1|, lg max|d| */
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; ++p) {
        const int s = sv[p], m = mv[p];
        for (size_t i = 0; i < ld; ++i) {
            double rho = 0.0, dmax = 0.0;
            for (int j = 0; j < m; ++j) {
                const double aj = ROW(A, s + j)[i], bj = ROW(B, s + j)[i];
                const double cj = ROW(C, s + j)[i], dj = ROW(D, s + j)[i];
                if (j >= 1) rho = fmax(rho, fmax(fabs(aj / bj), fabs(cj / bj)));
                dmax = fmax(dmax, fabs(dj));
            }
            float* f = feat + ((size_t)p * ld + i) * 4;
            f[0] = (float)fmin(rho, 1.0);
            f[1] = lg_feature(ROW(D, s)[i]);
            f[2] = lg_feature(ROW(D, s + m - 1)[i]);
            f[3] = lg_feature(dmax);
        }
    }
    sweeps_iface(A, B, C, D, ld, n_sys, P, sv, mv, 1, NULL, FTDO_SEM_GPU, cyclic, xbuf);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; ++p) {
        const int s = sv[p], m = mv[p];
        const double Nl = (double)(m * P);
        const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
        const float leps = lg_feature(eps);
        const double* D0 = ROW(D, s);
        const double* DN = ROW(D, s + m - 1);
        for (size_t i = 0; i < ld; ++i) {
            const float* f = feat + ((size_t)p * ld + i) * 4;
            const float fv[5] = {f[0], f[1], f[2], f[3], leps};
            int Jl = mlp_J(weights, hidden, margin, fv);
            Jl = Jl < 0 ? 0 : (Jl > m - 1 ? m - 1 : Jl);
            const double x0 = D0[i], xN = DN[i];
            /* the network's J where the first a-term and the first c-term it
             * drops are at most eps_; else the full correction, m-1 */
            const int fallback = dropped_exceeds(A, C, ld, s, m, i, Jl, x0, xN, eps);
            const int J = fallback ? m - 1 : Jl;
            if (J_out) J_out[(size_t)p * ld + i] = J | (fallback << 16);
            const int left_end = J < m - 2 ? J : m - 2;
            const int right_beg = (m - 1) - J > 1 ? (m - 1) - J : 1;
            for (int j = 1; j < m - 1; ++j) {
                double x = ROW(D, s + j)[i];
                if (j <= left_end) x = x - ROW(A, s + j)[i] * x0;
                if (j >= right_beg) x = x - ROW(C, s + j)[i] * xN;
                ROW(D, s + j)[i] = x;
            }
        }
    }
    free(xbuf);
    free(feat);
    return 0;
}

/* ------------------------------ guard mode --------------------------------
 * No counterpart in the reference. Besides the truncation by J, the filtered
 * kinds drop two couplings of each partition: row 0's to the partition's own
 * last unknown (pack_reduced, filtered_tdma_kernels.hpp:178-189) and row
 * m-1's to its own row 0 (solve_interfaces, :197-218). The guard evaluates,
 * per (partition, system) and with the interface values x0, xN, the terms
 * those drops leave out -- |c0' xN|, |a'_{m-1} x0| -- and the first term the
 * truncation leaves out on each side -- |a'_{kL} x0|, |c'_{kR} xN|. The
 * interface 2x2 solve passes a dropped coupling u, v on as (u - cN v)/det,
 * (v - a0' u)/det, det = 1 - cN_left a0'; the interior rows are convex
 * combinations of x0, xN (|a'_j| + |c'_j| <= 1 for diagonally dominant rows)
 * and drop at most the first truncated term each. So the couplings are held
 * to eps_/4 min(1, |det|) (the partition's left interface for c0' xN, its
 * right one for a'_{m-1} x0), the truncated terms to eps_/4, which keeps
 * x within about eps_ of the exact solve. A system
 * where any of them exceeds its bound (in any partition) is solved exactly: its
 * 2P-row reduced system, c0' kept, by Thomas (B == 1, row 0 "divided by 1"
 * left as it is), then the full correction. Every kind runs the v1 (full)
 * sweeps under the guard.
 * J[p * n_sys + i]: the J of (partition p, system i) before the guard (v1,
 * v2: the partition's J in every column; learned: the J after its own
 * check). rec_out[p * n_sys + i] = the J used | guard << 17.
 * Periodic systems (cyclic): the interfaces wrap (partition 0's left
 * neighbour is partition P-1), and the exact reduced system is periodic too
 * -- row 0 of partition 0 couples to row m-1 of partition P-1 -- solved as
 * cyclic_cols solves one (Thomas over rows 1..2P-1 with a second right-hand
 * side e, then row 0, then x_k = y_k + x_0 e_k), B == 1. */
int ftdo_filtered_guard(double* A, const double* B, double* C, double* D,
                        int n_sys, int n_row, int P, int cyclic, const int* J,
                        double eps_value, int eps_direct, int* rec_out, int nthreads) {
    int rc = check_parts(n_row, P);
    if (rc) return rc;
    if (P < 2) return fail(-1, "filtered_guard: needs at least 2 partitions");
    if (P > 512) return fail(-1, "filtered_guard: at most 512 partitions");
    if (!J) return fail(-1, "filtered_guard: J needed");
    const size_t ld = (size_t)n_sys;
    int sv[512], mv[512];
    for (int p = 0; p < P; ++p) part_rows(n_row, P, p, &sv[p], &mv[p]);
    /* reduced rows a0', c0', d0', aN, cN, dN per (partition, system) */
    double* R = (double*)malloc(sizeof(double) * ld * (size_t)(6 * P) + 8);
    if (!R) return fail(-3, "filtered_guard: out of memory");
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; ++p) {
        const int s = sv[p], m = mv[p];
        ffwd_cols(ROW(A, s), ROW(B, s), ROW(C, s), ROW(D, s), ld, 0, n_sys, m, m, 0);
        fbwd_cols(ROW(A, s), ROW(C, s), ROW(D, s), ld, 0, n_sys, m, 1, 0, FTDO_SEM_GPU);
        double* Rp = R + (size_t)(6 * p) * ld;
        const double* A0 = ROW(A, s);
        const double* A1 = ROW(A, s + 1);
        const double* C0 = ROW(C, s);
        const double* C1 = ROW(C, s + 1);
        for (size_t i = 0; i < ld; ++i) {
            /* pack_reduced's r, and the c0' it drops (pascal_tdma.cpp:177-183) */
            const double r = 1.0 / (1.0 - A1[i] * C0[i]);
            Rp[0 * ld + i] = A0[i];
            Rp[1 * ld + i] = -r * C0[i] * C1[i];
            Rp[2 * ld + i] = ROW(D, s)[i];
            Rp[3 * ld + i] = ROW(A, s + m - 1)[i];
            Rp[4 * ld + i] = ROW(C, s + m - 1)[i];
            Rp[5 * ld + i] = ROW(D, s + m - 1)[i];
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n_sys; ++ii) {
        const size_t i = (size_t)ii;
        double x0[512], xN[512];
        int guard = 0;
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            const double* Rp = R + (size_t)(6 * p) * ld;
            const double a0p = Rp[0 * ld + i], c0p = Rp[1 * ld + i], d0p = Rp[2 * ld + i];
            const double aN = Rp[3 * ld + i], cN = Rp[4 * ld + i], dN = Rp[5 * ld + i];
            /* interface 2x2 solves (filtered_tdma_kernels.cuh:284-303) */
            double v0 = d0p, vN = dN, detL = 1.0, detR = 1.0;
            if (cyclic || p > 0) {
                const double* Rl = R + (size_t)(6 * ((p + P - 1) % P)) * ld;
                detL = 1.0 - Rl[4 * ld + i] * a0p;
                v0 = (d0p - a0p * Rl[5 * ld + i]) / detL;
            }
            if (cyclic || p < P - 1) {
                const double* Rr = R + (size_t)(6 * ((p + 1) % P)) * ld;
                detR = 1.0 - Rr[0 * ld + i] * cN;
                vN = (dN - cN * Rr[2 * ld + i]) / detR;
            }
            x0[p] = v0;
            xN[p] = vN;
            const double Nl = (double)(m * P);
            const double eps = eps_direct ? eps_value : eps_value / (Nl * Nl);
            const double thr = eps * 0.25;
            /* a dropped coupling reaches the interface values divided by the
             * interface's 2x2 determinant (1 - cN_left a0', < 1 for M-matrices) */
            const double thrL = thr * fmin(1.0, fabs(detL));
            const double thrR = thr * fmin(1.0, fabs(detR));
            const int Jp = J[(size_t)p * ld + i];
            const int kL = (Jp < m - 2 ? Jp : m - 2) + 1;
            const int kR = ((m - 1) - Jp > 1 ? (m - 1) - Jp : 1) - 1;
            const double ta = kL <= m - 2 ? fabs(ROW(A, s + kL)[i] * v0) : 0.0;
            const double tc = kR >= 1 ? fabs(ROW(C, s + kR)[i] * vN) : 0.0;
            const double t0 = fabs(c0p * vN), tN = fabs(aN * v0);
            if (t0 > thrL || tN > thrR || ta > thr || tc > thr) guard = 1;
        }
        if (guard && !cyclic) {
            /* the exact reduced solve: rows [row 0 of p, row m-1 of p] */
            double cp[1024], dp[1024];
            cp[0] = R[1 * ld + i];
            dp[0] = R[2 * ld + i];
            for (int k = 1; k < 2 * P; ++k) {
                const double* Rp = R + (size_t)(6 * (k >> 1)) * ld;
                const double ak = Rp[((k & 1) ? 3 : 0) * ld + i];
                const double ck = Rp[((k & 1) ? 4 : 1) * ld + i];
                const double dk = Rp[((k & 1) ? 5 : 2) * ld + i];
                const double inv = 1.0 / (1.0 - ak * cp[k - 1]);
                dp[k] = inv * (dk - ak * dp[k - 1]);
                cp[k] = inv * ck;
            }
            for (int k = 2 * P - 2; k >= 0; --k) dp[k] = dp[k] - cp[k] * dp[k + 1];
            for (int p = 0; p < P; ++p) {
                x0[p] = dp[2 * p];
                xN[p] = dp[2 * p + 1];
            }
        } else if (guard) {
            /* the periodic reduced system: y and e over rows 1..n-1 (e's
             * right-hand side -a_1 at row 1, -c_{n-1} at row n-1), then
             * x_0 from row 0 and x_k = y_k + x_0 e_k (cyclic_cols, B == 1) */
            const int n = 2 * P;
            double cp[1024], dp[1024], e[1024];
            const double a0 = R[0 * ld + i], c0 = R[1 * ld + i], d0 = R[2 * ld + i];
            cp[1] = R[4 * ld + i];
            dp[1] = R[5 * ld + i];
            e[1] = -R[3 * ld + i];
            for (int k = 2; k < n; ++k) {
                const double* Rp = R + (size_t)(6 * (k >> 1)) * ld;
                const double ak = Rp[((k & 1) ? 3 : 0) * ld + i];
                const double ck = Rp[((k & 1) ? 4 : 1) * ld + i];
                const double dk = Rp[((k & 1) ? 5 : 2) * ld + i];
                const double e0 = k == n - 1 ? -ck : 0.0;
                const double r = 1.0 / (1.0 - ak * cp[k - 1]);
                dp[k] = r * (dk - ak * dp[k - 1]);
                e[k] = r * (e0 - ak * e[k - 1]);
                cp[k] = r * ck;
            }
            for (int k = n - 2; k >= 1; --k) {
                dp[k] = dp[k] - cp[k] * dp[k + 1];
                e[k] = e[k] - cp[k] * e[k + 1];
            }
            dp[0] = (d0 - a0 * dp[n - 1] - c0 * dp[1]) / (1.0 + a0 * e[n - 1] + c0 * e[1]);
            for (int k = 1; k < n; ++k) dp[k] = dp[k] + dp[0] * e[k];
            for (int p = 0; p < P; ++p) {
                x0[p] = dp[2 * p];
                xN[p] = dp[2 * p + 1];
            }
        }
        for (int p = 0; p < P; ++p) {
            const int s = sv[p], m = mv[p];
            const int Jf = guard ? m - 1 : J[(size_t)p * ld + i];
            const int left_end = Jf < m - 2 ? Jf : m - 2;
            const int right_beg = (m - 1) - Jf > 1 ? (m - 1) - Jf : 1;
            for (int j = 1; j < m - 1; ++j) {
                double x = ROW(D, s + j)[i];
                if (j <= left_end) x = x - ROW(A, s + j)[i] * x0[p];
                if (j >= right_beg) x = x - ROW(C, s + j)[i] * xN[p];
                ROW(D, s + j)[i] = x;
            }
            ROW(D, s)[i] = x0[p];
            ROW(D, s + m - 1)[i] = xN[p];
            if (rec_out) rec_out[(size_t)p * ld + i] = Jf | (guard << 17);
        }
    }
    free(R);
    return 0;
}
