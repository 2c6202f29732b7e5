/*
 * heat_oracle.c -- CPU restatement of the reference's 3D heat ADI harness
 * (apps/heat_cpu), the caller of the tridiagonal path. TEST INFRASTRUCTURE
 * ONLY, like ftdma_oracle.c.
 *
 * One address space replaces the reference's (px x py x pz) Cartesian MPI
 * layout: the sweep in direction d splits each system's rows into p_d
 * partitions (the ranks along d). The global array with one ghost layer per
 * face is what the subdomains plus their halo exchange hold, so the
 * arithmetic per cell is the reference's. One difference: the reference
 * takes the filtered J of a partition over the systems of one rank's
 * subdomain, this restatement over all systems of the partition (the same
 * when the other two directions are not split).
 *   grid and mesh      global.cpp:68-78, mpi_subdomain.cpp:141-162
 *   initial field      mpi_subdomain.cpp:164-173
 *   Dirichlet ghosts   mpi_subdomain.cpp:180-236
 *   periodic ghosts    mpi_subdomain.cpp:90-111 (wrap of the face planes)
 *   time step          solve_theta.cpp:92-334 (RHS, then Z, Y, X sweeps with
 *                      the implicit Dirichlet fold-in of the wall rows)
 *   L2 error           main.cpp:59-81 (note: divided by (nx+1)(ny+1)(nz+1))
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "ftdma_oracle.h"

static const double Pi = 3.14159265358979323846;

typedef struct {
    int nx, ny, nz;       /* interior cells per direction */
    int nx1, ny1, nz1;    /* with one ghost layer per face */
    double *x, *y, *z;    /* cell centres, index 0..n+1 */
} grid_t;

static size_t IDX(const grid_t* g, int i, int j, int k) {
    return ((size_t)k * g->ny1 + j) * g->nx1 + i;
}

static void centres(double* c, int n) {
    /* dd = l/(n_nodes-1) with n_nodes = n+1; c_i = x0 + dd/2 + (ista-2+i) dd, ista = 1 */
    double dd = 2.0 / (double)n;
    for (int i = 0; i <= n + 1; ++i) c[i] = -1.0 + dd / 2.0 + (double)(i - 1) * dd;
}

/* u_ghost = 2 u_BC - u_adjacent on the non-periodic faces */
static void dirichlet_ghosts(const grid_t* g, double* th, const int* periodic) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (!periodic[0])
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j) {
                double ubc = -cos(Pi * g->y[j]) * cos(Pi * g->z[k]);
                th[IDX(g, 0, j, k)] = 2.0 * ubc - th[IDX(g, 1, j, k)];
                th[IDX(g, nx + 1, j, k)] = 2.0 * ubc - th[IDX(g, nx, j, k)];
            }
    if (!periodic[1])
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double ubc = -cos(Pi * g->x[i]) * cos(Pi * g->z[k]);
                th[IDX(g, i, 0, k)] = 2.0 * ubc - th[IDX(g, i, 1, k)];
                th[IDX(g, i, ny + 1, k)] = 2.0 * ubc - th[IDX(g, i, ny, k)];
            }
    if (!periodic[2])
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double ubc = -cos(Pi * g->x[i]) * cos(Pi * g->y[j]);
                th[IDX(g, i, j, 0)] = 2.0 * ubc - th[IDX(g, i, j, 1)];
                th[IDX(g, i, j, nz + 1)] = 2.0 * ubc - th[IDX(g, i, j, nz)];
            }
}

/* periodic faces: ghost planes take the opposite interior planes */
static void periodic_ghosts(const grid_t* g, double* th, const int* periodic) {
    if (periodic[0])
        for (int k = 0; k < g->nz1; ++k)
            for (int j = 0; j < g->ny1; ++j) {
                th[IDX(g, 0, j, k)] = th[IDX(g, g->nx, j, k)];
                th[IDX(g, g->nx + 1, j, k)] = th[IDX(g, 1, j, k)];
            }
    if (periodic[1])
        for (int k = 0; k < g->nz1; ++k)
            for (int i = 0; i < g->nx1; ++i) {
                th[IDX(g, i, 0, k)] = th[IDX(g, i, g->ny, k)];
                th[IDX(g, i, g->ny + 1, k)] = th[IDX(g, i, 1, k)];
            }
    if (periodic[2])
        for (int j = 0; j < g->ny1; ++j)
            for (int i = 0; i < g->nx1; ++i) {
                th[IDX(g, i, j, 0)] = th[IDX(g, i, j, g->nz)];
                th[IDX(g, i, j, g->nz + 1)] = th[IDX(g, i, j, 1)];
            }
}

/* One direction's batch: kind 0/1 filtered v1/v2 with P partitions (P == 1
 * -> Thomas), kind 2 pascal. */
static int solve_batch(int kind, int P, int sem, int cyclic, double dt, double* A, double* B,
                       double* C, double* D, int n_sys, int n_row, double* rho, int nthreads) {
    if (kind == 2 || P == 1) return ftdo_pascal(A, B, C, D, n_sys, n_row, P, cyclic, sem, nthreads);
    ftdo_set_rho(A, B, C, n_sys, n_row, rho, rho + n_row + 1);
    return ftdo_filtered(A, B, C, D, n_sys, n_row, P, kind + 1, cyclic, sem, rho,
                         rho + n_row + 1, dt, 0, NULL, NULL, nthreads);
}

int ftdo_heat_run(int nx, int ny, int nz, const int* parts, int kind, int sem,
                  const int* periodic, double dt, int Nt, double* theta_out, double* l2_out,
                  int nthreads) {
    const int px = parts[0], py = parts[1], pz = parts[2];
    grid_t g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nx1 = nx + 2; g.ny1 = ny + 2; g.nz1 = nz + 2;
    const size_t ncell = (size_t)g.nx1 * g.ny1 * g.nz1;
    const size_t nin = (size_t)nx * ny * nz;
    int nmax = nx > ny ? nx : ny;
    nmax = nmax > nz ? nmax : nz;
    double* th = (double*)malloc(sizeof(double) * ncell);
    double* rhs = (double*)calloc(ncell, sizeof(double));
    double* buf = (double*)malloc(sizeof(double) * (4 * nin + 2 * (size_t)(nmax + 1)));
    g.x = (double*)malloc(sizeof(double) * (size_t)(nx + 2));
    g.y = (double*)malloc(sizeof(double) * (size_t)(ny + 2));
    g.z = (double*)malloc(sizeof(double) * (size_t)(nz + 2));
    if (!th || !rhs || !buf || !g.x || !g.y || !g.z) {
        free(th); free(rhs); free(buf); free(g.x); free(g.y); free(g.z);
        return -3;
    }
    double *A = buf, *B = A + nin, *C = B + nin, *D = C + nin, *rho = D + nin;
    centres(g.x, nx);
    centres(g.y, ny);
    centres(g.z, nz);
    const double ddx = 2.0 / nx, ddy = 2.0 / ny, ddz = 2.0 / nz;

    for (int k = 0; k < g.nz1; ++k)
        for (int j = 0; j < g.ny1; ++j)
            for (int i = 0; i < g.nx1; ++i)
                th[IDX(&g, i, j, k)] = sin(Pi * g.x[i]) * sin(Pi * g.y[j]) * sin(Pi * g.z[k]) +
                                       cos(Pi * g.x[i]) * cos(Pi * g.y[j]) * cos(Pi * g.z[k]);
    periodic_ghosts(&g, th, periodic);
    dirichlet_ghosts(&g, th, periodic);

    /* compute_stencil (stencil_coeffs.hpp:10-13): a = c = base, b = -2 base */
    const double bx = dt / (2.0 * ddx * ddx), by = dt / (2.0 * ddy * ddy), bz = dt / (2.0 * ddz * ddz);
    const double sxa = bx, sxb = -2.0 * bx, sxc = bx;
    const double sya = by, syb = -2.0 * by, syc = by;
    const double sza = bz, szb = -2.0 * bz, szc = bz;
    int rc = 0;

    for (int t = 0; t < Nt && rc == 0; ++t) {
        dirichlet_ghosts(&g, th, periodic);
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const size_t c0 = IDX(&g, i, j, k);
                    rhs[c0] = (sxc * th[IDX(&g, i + 1, j, k)] + sxa * th[IDX(&g, i - 1, j, k)]) +
                              (syc * th[IDX(&g, i, j + 1, k)] + sya * th[IDX(&g, i, j - 1, k)]) +
                              (szc * th[IDX(&g, i, j, k + 1)] + sza * th[IDX(&g, i, j, k - 1)]) +
                              (1.0 + sxb + syb + szb) * th[c0] +
                              dt * 3.0 * Pi * Pi * cos(Pi * g.x[i]) * cos(Pi * g.y[j]) *
                                  cos(Pi * g.z[k]);
                }
        /* Z: row kk, system jj*nx + ii */
        {
            const int ns = nx * ny;
            for (int kk = 0; kk < nz; ++kk)
                for (int jj = 0; jj < ny; ++jj)
                    for (int ii = 0; ii < nx; ++ii) {
                        size_t o = (size_t)kk * ns + (size_t)jj * nx + ii;
                        A[o] = -sza; B[o] = 1.0 - szb; C[o] = -szc;
                        D[o] = rhs[IDX(&g, ii + 1, jj + 1, kk + 1)];
                    }
            if (!periodic[2])
                for (int jj = 0; jj < ny; ++jj)
                    for (int ii = 0; ii < nx; ++ii) {
                        size_t o = (size_t)jj * nx + ii;
                        D[o] += sza * (th[IDX(&g, ii + 1, jj + 1, 0)] + th[IDX(&g, ii + 1, jj + 1, 1)]);
                        B[o] += sza;
                        o = (size_t)(nz - 1) * ns + (size_t)jj * nx + ii;
                        D[o] += szc * (th[IDX(&g, ii + 1, jj + 1, nz + 1)] + th[IDX(&g, ii + 1, jj + 1, nz)]);
                        B[o] += szc;
                    }
            rc = solve_batch(kind, pz, sem, periodic[2], dt, A, B, C, D, ns, nz, rho, nthreads);
            for (int kk = 0; kk < nz; ++kk)
                for (int jj = 0; jj < ny; ++jj)
                    for (int ii = 0; ii < nx; ++ii)
                        rhs[IDX(&g, ii + 1, jj + 1, kk + 1)] = D[(size_t)kk * ns + (size_t)jj * nx + ii];
        }
        if (rc) break;
        /* Y: row jj, system kk*nx + ii */
        {
            const int ns = nx * nz;
            for (int jj = 0; jj < ny; ++jj)
                for (int kk = 0; kk < nz; ++kk)
                    for (int ii = 0; ii < nx; ++ii) {
                        size_t o = (size_t)jj * ns + (size_t)kk * nx + ii;
                        A[o] = -sya; B[o] = 1.0 - syb; C[o] = -syc;
                        D[o] = rhs[IDX(&g, ii + 1, jj + 1, kk + 1)];
                    }
            if (!periodic[1])
                for (int kk = 0; kk < nz; ++kk)
                    for (int ii = 0; ii < nx; ++ii) {
                        size_t o = (size_t)kk * nx + ii;
                        D[o] += sya * (th[IDX(&g, ii + 1, 0, kk + 1)] + th[IDX(&g, ii + 1, 1, kk + 1)]);
                        B[o] += sya;
                        o = (size_t)(ny - 1) * ns + (size_t)kk * nx + ii;
                        D[o] += syc * (th[IDX(&g, ii + 1, ny + 1, kk + 1)] + th[IDX(&g, ii + 1, ny, kk + 1)]);
                        B[o] += syc;
                    }
            rc = solve_batch(kind, py, sem, periodic[1], dt, A, B, C, D, ns, ny, rho, nthreads);
            for (int jj = 0; jj < ny; ++jj)
                for (int kk = 0; kk < nz; ++kk)
                    for (int ii = 0; ii < nx; ++ii)
                        rhs[IDX(&g, ii + 1, jj + 1, kk + 1)] = D[(size_t)jj * ns + (size_t)kk * nx + ii];
        }
        if (rc) break;
        /* X: row ii, system kk*ny + jj */
        {
            const int ns = ny * nz;
            for (int ii = 0; ii < nx; ++ii)
                for (int kk = 0; kk < nz; ++kk)
                    for (int jj = 0; jj < ny; ++jj) {
                        size_t o = (size_t)ii * ns + (size_t)kk * ny + jj;
                        A[o] = -sxa; B[o] = 1.0 - sxb; C[o] = -sxc;
                        D[o] = rhs[IDX(&g, ii + 1, jj + 1, kk + 1)];
                    }
            if (!periodic[0])
                for (int kk = 0; kk < nz; ++kk)
                    for (int jj = 0; jj < ny; ++jj) {
                        size_t o = (size_t)kk * ny + jj;
                        D[o] += sxa * (th[IDX(&g, 0, jj + 1, kk + 1)] + th[IDX(&g, 1, jj + 1, kk + 1)]);
                        B[o] += sxa;
                        o = (size_t)(nx - 1) * ns + (size_t)kk * ny + jj;
                        D[o] += sxc * (th[IDX(&g, nx + 1, jj + 1, kk + 1)] + th[IDX(&g, nx, jj + 1, kk + 1)]);
                        B[o] += sxc;
                    }
            rc = solve_batch(kind, px, sem, periodic[0], dt, A, B, C, D, ns, nx, rho, nthreads);
            for (int ii = 0; ii < nx; ++ii)
                for (int kk = 0; kk < nz; ++kk)
                    for (int jj = 0; jj < ny; ++jj)
                        th[IDX(&g, ii + 1, jj + 1, kk + 1)] = D[(size_t)ii * ns + (size_t)kk * ny + jj];
        }
        periodic_ghosts(&g, th, periodic);
    }

    if (rc == 0 && l2_out) {
        double err = 0.0;
        const double decay = exp(-3.0 * Pi * Pi * (double)Nt * dt);
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double ex = sin(Pi * g.x[i]) * sin(Pi * g.y[j]) * sin(Pi * g.z[k]) * decay +
                                cos(Pi * g.x[i]) * cos(Pi * g.y[j]) * cos(Pi * g.z[k]);
                    double d = th[IDX(&g, i, j, k)] - ex;
                    err += d * d;
                }
        *l2_out = sqrt(err / (double)(nx + 1) / (double)(ny + 1) / (double)(nz + 1));
    }
    if (rc == 0 && theta_out) memcpy(theta_out, th, sizeof(double) * ncell);
    free(th); free(rhs); free(buf); free(g.x); free(g.y); free(g.z);
    return rc;
}
