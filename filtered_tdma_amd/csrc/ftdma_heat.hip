// ftdma_heat.hip -- the 3D heat ADI harness on one MI355X: the caller of the
// tridiagonal path in the reference (apps/heat_gpu/solve_theta.cu:391-650,
// SolveTheta::profile; apps/heat_cpu/solve_theta.cpp:18-334), with the
// ghost-cell Dirichlet / periodic boundaries of MPISubdomain
// (apps/heat_cpu/mpi_subdomain.cpp:90-236) and the L2 error of main.cpp:59-81.
//
// One address space holds the whole (nx+2)(ny+2)(nz+2) field with one ghost
// layer per face; the reference's (px x py x pz) ranks become partitions of
// the x, y and z sweeps (each sweep is one batched solve of this library).
// Per step (solve_theta.cpp:97-334):
//   Dirichlet ghosts -> explicit RHS -> Z: build, solve, copy -> Y: build,
//   solve, copy -> X: build, solve, copy into theta -> periodic ghosts.
// Where a sweep can, it runs fused: one kernel generates a, b, c from the
// stencil, reads d from the field (or, in the z sweep, forms it from theta
// in place of the separate RHS pass), solves and writes x into the field. The
// materialized build + solve + scatter stays as the other path and gives
// the same bits.
// Transcendental factors (cos/sin of the cell centres) are tabulated on the
// host with the C library, so the device arithmetic is exactly the CPU
// restatement's; the whole step can be captured once into a hipGraph and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ftdma.h"
#include "ftdma_kernels.h"
#include "ftdma_xcd.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// The field on the device: (nz+2) planes of (ny+2) lines, each line of
// nx+2 cells (ghosts included) in a pitch px (a multiple of 16 doubles),
// starting at kOrigin so that cell i = 1 of every line -- the first
// interior cell -- sits on a 128-byte boundary. A wave's 64 consecutive
// interior cells are then four whole cache lines, not five partial ones.
constexpr size_t kOrigin = 15;
struct Dims {
    int nx, ny, nz, nx1, ny1;  // interior cells; lines and planes with ghosts
    int px;                    // line pitch in doubles
};

inline int line_pitch(int nx) { return (nx + 2 + 15) / 16 * 16; }

__host__ __device__ __forceinline__ size_t cell(const Dims& g, int i, int j, int k) {
    return kOrigin + ((size_t)k * g.ny1 + j) * g.px + i;
}

// doubles of a device field
inline size_t field_doubles(const Dims& g, int nz) {
    return kOrigin + (size_t)g.px * g.ny1 * (nz + 2) + 1;
}

// u_ghost = 2 u_BC - u_adjacent, u_BC = -cos(pi a) cos(pi b) on each face
// (mpi_subdomain.cpp:180-236); face f = blockIdx.z: 0 = x, 1 = y, 2 = z, the
// faces in bit f of `faces`. The faces write disjoint ghost cells (no edges)
// and read only interior cells, so one launch serves all three.
__global__ void k_dirichlet(double* __restrict__ th, Dims g, int faces,
                            const double* __restrict__ cx, const double* __restrict__ cy,
                            const double* __restrict__ cz) {
    const int face = blockIdx.z;
    if (!((faces >> face) & 1)) return;
    const int u = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int v = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (face == 0) {
        if (u > g.ny || v > g.nz) return;
        const double ubc = -cy[u] * cz[v];
        th[cell(g, 0, u, v)] = 2.0 * ubc - th[cell(g, 1, u, v)];
        th[cell(g, g.nx + 1, u, v)] = 2.0 * ubc - th[cell(g, g.nx, u, v)];
    } else if (face == 1) {
        if (u > g.nx || v > g.nz) return;
        const double ubc = -cx[u] * cz[v];
        th[cell(g, u, 0, v)] = 2.0 * ubc - th[cell(g, u, 1, v)];
        th[cell(g, u, g.ny + 1, v)] = 2.0 * ubc - th[cell(g, u, g.ny, v)];
    } else {
        if (u > g.nx || v > g.ny) return;
        const double ubc = -cx[u] * cy[v];
        th[cell(g, u, v, 0)] = 2.0 * ubc - th[cell(g, u, v, 1)];
        th[cell(g, u, v, g.nz + 1)] = 2.0 * ubc - th[cell(g, u, v, g.nz)];
    }
}

// periodic wrap of one face pair over the whole plane (mpi_subdomain.cpp:90-111)
__global__ void k_periodic(double* __restrict__ th, Dims g, int face) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y * blockDim.y + threadIdx.y;
    const int nz1 = g.nz + 2;
    if (face == 0) {
        if (u >= g.ny1 || v >= nz1) return;
        th[cell(g, 0, u, v)] = th[cell(g, g.nx, u, v)];
        th[cell(g, g.nx + 1, u, v)] = th[cell(g, 1, u, v)];
    } else if (face == 1) {
        if (u >= g.nx1 || v >= nz1) return;
        th[cell(g, u, 0, v)] = th[cell(g, u, g.ny, v)];
        th[cell(g, u, g.ny + 1, v)] = th[cell(g, u, 1, v)];
    } else {
        if (u >= g.nx1 || v >= g.ny1) return;
        th[cell(g, u, v, 0)] = th[cell(g, u, v, g.nz)];
        th[cell(g, u, v, g.nz + 1)] = th[cell(g, u, v, 1)];
    }
}

struct Stencil {
    double xa, xb, xc, ya, yb, yc, za, zb, zc, c0;
};

// explicit RHS of the factored Crank-Nicolson step (solve_theta.cpp:114-134),
// marching in z: a thread owns one (i, j) column of a 64 x 4 tile and walks
// kb planes with th at k-1, k, k+1 in registers, so each value of th is
// fetched from HBM about once (one thread per cell would fetch every plane
// again for its neighbours above and below, from other workgroups). kb is
// kRhsKB, or shorter where a grid of kRhsKB-plane runs would not cover the
// CUs (rhs_planes); every cell's value is the same expression either way.
constexpr int kRhsKB = 32;
__global__ __launch_bounds__(256) void k_rhs_z(const double* __restrict__ th,
                                               double* __restrict__ rhs, Dims g, Stencil s,
                                               const double* __restrict__ cx,
                                               const double* __restrict__ cy,
                                               const double* __restrict__ cz, int kb) {
    // Workgroups go round-robin to the 8 XCDs, each with its own L2. Each
    // XCD takes a contiguous run of tiles, y fastest, so the rows above and
    // below a 64 x 4 tile were just read by its neighbours on the same XCD
    // (dealt in launch order, y-adjacent tiles landed on different XCDs and
    // every halo row came from memory again).
    const int nbx = (g.nx + 63) / 64, nby = (g.ny + 3) / 4;
    const int T = ftd::xcd_tile(blockIdx.x, gridDim.x);
    const int by = T % nby, bx = (T / nby) % nbx, bz = T / (nby * nbx);
    const int i = bx * 64 + threadIdx.x + 1;
    const int j = by * 4 + threadIdx.y + 1;
    const int k0 = bz * kb + 1;
    if (i > g.nx || j > g.ny) return;
    const int k1 = k0 + kb - 1 < g.nz ? k0 + kb - 1 : g.nz;
    const size_t sy = (size_t)g.px, sz = (size_t)g.px * g.ny1;
    const double diag = 1.0 + s.xb + s.yb + s.zb;
    const double cxy = s.c0 * cx[i] * cy[j];
    size_t c = cell(g, i, j, k0);
    double below = th[c - sz], here = th[c];
#pragma unroll 4
    for (int k = k0; k <= k1; ++k, c += sz) {
        const double above = th[c + sz];
        rhs[c] = (s.xc * th[c + 1] + s.xa * th[c - 1]) + (s.yc * th[c + sy] + s.ya * th[c - sy]) +
                 (s.zc * above + s.za * below) + diag * here + cxy * cz[k];
        below = here;
        here = above;
    }
}

// Coefficients of one sweep (solve_theta.cpp:146-186 for Z, 212-247 Y,
// 274-309 X): a = -s.a, b = 1 - s.b, c = -s.c, d = rhs; on a Dirichlet wall
// row d += s * (ghost + adjacent), b += s. Sweep dir: 0 = x (row ii, system
// kk*ny + jj), 1 = y (row jj, system kk*nx + ii), 2 = z (row kk, system
// jj*nx + ii). One thread per cell, indexed in the batch's own order so the
// writes are coalesced.
__global__ void k_build(const double* __restrict__ th, const double* __restrict__ rhs, Dims g,
                        int dir, double sa, double sb, double sc, int wall,
                        double* __restrict__ A, double* __restrict__ B, double* __restrict__ C,
                        double* __restrict__ D) {
    const size_t n_in = (size_t)g.nx * g.ny * g.nz;
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_in) return;
    int ii, jj, kk, row, n_row;
    if (dir == 2) {
        const size_t ns = (size_t)g.nx * g.ny;
        kk = (int)(o / ns);
        const int r = (int)(o % ns);
        jj = r / g.nx;
        ii = r % g.nx;
        row = kk;
        n_row = g.nz;
    } else if (dir == 1) {
        const size_t ns = (size_t)g.nx * g.nz;
        jj = (int)(o / ns);
        const int r = (int)(o % ns);
        kk = r / g.nx;
        ii = r % g.nx;
        row = jj;
        n_row = g.ny;
    } else {
        const size_t ns = (size_t)g.ny * g.nz;
        ii = (int)(o / ns);
        const int r = (int)(o % ns);
        kk = r / g.ny;
        jj = r % g.ny;
        row = ii;
        n_row = g.nx;
    }
    const int i = ii + 1, j = jj + 1, k = kk + 1;
    double a = -sa, b = 1.0 - sb, c = -sc, d = rhs[cell(g, i, j, k)];
    if (wall) {
        const size_t st = dir == 0 ? 1 : dir == 1 ? (size_t)g.px : (size_t)g.px * g.ny1;
        const size_t here = cell(g, i, j, k);
        if (row == 0) {
            d += sa * (th[here - st] + th[here]);
            b += sa;
        }
        if (row == n_row - 1) {
            d += sc * (th[here + st] + th[here]);
            b += sc;
        }
    }
    A[o] = a;
    B[o] = b;
    C[o] = c;
    D[o] = d;
}

// ---------------------------------------------------------------------------
// Fused sweeps (PaScaL kind, Dirichlet walls): the coefficients of a sweep
// are the stencil constants, so they are generated in the kernel instead of
// being written to A, B, C and read back; the right-hand side is read from
// the field and x is written straight into it. A sweep then moves 16 B per
// row (read d, write x) instead of build + solve + scatter (~96 B per row).
// The arithmetic is k_build's followed by the solve's, operation for
// operation, with the partition count Q the solver would pick for the same
// sweep (Q = 1: plain Thomas), so the field is bit-identical.
//
// Row j of partition q of a system: a = -sa, b = 1 - sb, c = -sc,
// d = rhs; on the system's first row b += sa, d += sa (ghost + cell); on its
// last row b += sc, d += sc (ghost + cell) (solve_theta.cpp:146-186).
// ---------------------------------------------------------------------------

// the generated row; first / last: it is the system's first / last row,
// lo / hi the cells of those rows, rs the row stride in the field, th the
// field holding the ghost values
__device__ __forceinline__ void adi_row(bool first, bool last, double sa, double sb, double sc,
                                        double d_in, const double* __restrict__ th, size_t lo,
                                        size_t hi, size_t rs, double& a, double& b, double& c,
                                        double& d) {
    a = -sa;
    b = 1.0 - sb;
    c = -sc;
    d = d_in;
    if (first) {
        d += sa * (th[lo - rs] + th[lo]);
        b += sa;
    }
    if (last) {
        d += sc * (th[hi + rs] + th[hi]);
        b += sc;
    }
}

// d of row j of one partition (rows past the partition repeat its last
// row, so the load carries no branch and can be issued ahead of its use)
struct RowAt {
    const double* p;
    size_t rs;
    int cnt;
    __device__ __forceinline__ double operator()(int j) const {
        const int jj = j < cnt ? j : cnt - 1;
        return p[(size_t)jj * rs];
    }
};

// d of row j of one z partition formed from th as k_rhs_z forms it (the
// z sweep then needs no rhs pass): c0 is the cell of the partition's first
// row, k0 its z index; rows past the partition repeat its last row
// Rows are formed in order (j = 0, 1, ...), so th at z - 1, z, z + 1 rolls
// through zm, z0, zp: one z load per row instead of three. Rows past the
// partition (j >= cnt) repeat its last row, as the clamped index does; the
// roll is by selects, not a branch (the lanes of a wave may hold partitions
// of different lengths).
struct RowRhsZ {
    const double* th;
    const double* cz;
    size_t c0, sy, sz;
    int cnt, k0;
    double xa, xc, ya, yc, za, zc, diag, cxy;
    double zm = 0.0, z0 = 0.0, zp = 0.0;
    __device__ __forceinline__ double operator()(int j) {
        const int jj = j < cnt ? j : cnt - 1;
        const size_t c = c0 + (size_t)jj * sz;
        if (j == 0) {
            zm = th[c - sz];
            z0 = th[c];
            zp = th[c + sz];
        } else {
            const bool adv = j < cnt;
            const double nzp = th[c + sz];
            zm = adv ? z0 : zm;
            z0 = adv ? zp : z0;
            zp = adv ? nzp : zp;
        }
        return (xc * th[c + 1] + xa * th[c - 1]) + (yc * th[c + sy] + ya * th[c - sy]) +
               (zc * zp + za * zm) + diag * z0 + cxy * cz[k0 + jj];
    }
};

// One lane per (system, partition) with the partition's rows in registers.
// rows(j) returns d of local row j (clamped to the partition), Store(j, x)
// writes x of local row j. The d loads run PF rows ahead of their use. Restates k_thomas_fused (Q == 1) and
// k_pascal_fused (Q >= 2, non-cyclic): rows 0, 1 of a partition divided by
// b, modified elimination, 2Q-row reduced system in LDS, interior update.
template <int M, int PF, bool THOMAS, typename Rows, typename Store>
__device__ __forceinline__ void adi_solve_lane(int N, int Q, int q, int r0, int cnt, int S, int sl,
                                               double sa, double sb, double sc,
                                               const double* __restrict__ th, size_t lo,
                                               size_t hi, size_t rs, bool live, double* lds,
                                               Rows& rows, Store store) {
    double rd[M];
#define FTD_ADI_LOAD(j) rd[(j)] = rows(j)
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_ADI_LOAD(j);
    if (THOMAS) {
        double c[M], d[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            if (j + PF < M) FTD_ADI_LOAD(j + PF);
            if (j < cnt) {
                double aj, bj, cj, dj;
                adi_row(j == 0, j == cnt - 1, sa, sb, sc, rd[j], th, lo, hi, rs, aj, bj, cj, dj);
                if (j == 0) {
                    d[0] = dj / bj;
                    c[0] = cj / bj;
                } else {
                    const double r = 1.0 / (bj - aj * c[j - 1]);
                    d[j] = r * (dj - aj * d[j - 1]);
                    c[j] = r * cj;
                }
            }
        }
#pragma unroll
        for (int j = M - 2; j >= 0; --j)
            if (j <= cnt - 2) d[j] = d[j] - c[j] * d[j + 1];
        if (live) {
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < cnt) store(j, d[j]);
        }
        return;
    }
    double a[M], c[M], d[M];
    double aN = 0.0, cN = 0.0, dN = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_ADI_LOAD(j + PF);
        if (j < cnt) {
            double aj, bj, cj, dj;
            // the system's first row is row 0 of partition 0, its last row the
            // last row of partition Q-1
            adi_row(j == 0 && q == 0, j == cnt - 1 && q == Q - 1, sa, sb, sc, rd[j], th, lo, hi,
                    rs, aj, bj, cj, dj);
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
    const int NR = 2 * Q;
    double* rA = lds;
    double* rC = rA + NR * S;
    double* rD = rC + NR * S;
    rA[(2 * q) * S + sl] = a0;
    rC[(2 * q) * S + sl] = c0;
    rD[(2 * q) * S + sl] = d0;
    rA[(2 * q + 1) * S + sl] = aN;
    rC[(2 * q + 1) * S + sl] = cN;
    rD[(2 * q + 1) * S + sl] = dN;
    __syncthreads();
    if (q == 0) {
        // the reduced rows 4 at a time: their values are read into registers
        // first and c', d' are carried in registers, so the serial chain is
        // the arithmetic, not a round trip through LDS per row
        double* Ar = rA + sl;
        double* Cr = rC + sl;
        double* Dr = rD + sl;
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
                    x = d[j] - (a[j] * x0 + c[j] * xN);
                store(j, x);
            }
        }
    }
#undef FTD_ADI_LOAD
}

// Z (dir 2) and Y (dir 1) sweeps: systems consecutive in x, so a wave's row
// loads are contiguous; rows strided by a plane (Z) or a line (Y). The
// solution goes back into rhs. RHSZ (the z sweep, first of the step): d is
// formed from th by the explicit stencil (RowRhsZ) instead of read from rhs,
// so the step has no separate rhs pass; th is not written until the x sweep.
template <int M, bool THOMAS, bool RHSZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M >= 64 ? 1 : 2))) void k_adi_yz(const double* __restrict__ th, double* rhs, Dims g,
                                                int nz, int dir, double sa, double sb, double sc,
                                                int Q, int lgS, Stencil st,
                                                const double* __restrict__ cx,
                                                const double* __restrict__ cy,
                                                const double* __restrict__ cz) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int sl = t & (S - 1);
    const int q = t >> lgS;
    const int n_sys = g.nx * (dir == 2 ? g.ny : nz);
    const int N = dir == 2 ? nz : g.ny;
    // RHSZ reads th at x +- 1 and y +- 1: each XCD takes a contiguous run of
    // system blocks (whole x lines, consecutive y), so the halo lines its
    // workgroups read are the ones their neighbours read, in the same L2
    const int blk = RHSZ ? ftd::xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const int s_raw = blk * S + sl;
    const bool live = s_raw < n_sys;
    const int s = live ? s_raw : n_sys - 1;
    const int ii = s % g.nx + 1, oo = s / g.nx + 1;
    const size_t rs = dir == 2 ? (size_t)g.px * g.ny1 : (size_t)g.px;
    const size_t here0 = dir == 2 ? cell(g, ii, oo, 1) : cell(g, ii, 1, oo);
    int r0, cnt;
    {
        const int base = N / Q, rem = N % Q;
        r0 = q * base + (q < rem ? q : rem);
        cnt = base + (q < rem ? 1 : 0);
    }
    double* const pbase = rhs + here0 + (size_t)r0 * rs;
    if (RHSZ) {
        RowRhsZ rows{th, cz, here0 + (size_t)r0 * rs, (size_t)g.px, rs, cnt, r0 + 1,
                     st.xa, st.xc, st.ya, st.yc, st.za, st.zc, 1.0 + st.xb + st.yb + st.zb,
                     st.c0 * cx[ii] * cy[oo]};
        adi_solve_lane<M, (M >= 32 ? 2 : 4), THOMAS>(N, Q, q, r0, cnt, S, sl, sa, sb, sc, th, here0,
                                     here0 + (size_t)(N - 1) * rs, rs, live, lds, rows,
                                     [&](int j, double x) { pbase[(size_t)j * rs] = x; });
    } else {
        RowAt rows{pbase, rs, cnt};
        adi_solve_lane<M, (M == 16 ? 16 : M == 32 ? 8 : 4), THOMAS>(
            N, Q, q, r0, cnt, S, sl, sa, sb, sc, th, here0, here0 + (size_t)(N - 1) * rs, rs,
            live, lds, rows, [&](int j, double x) { pbase[(size_t)j * rs] = x; });
    }
}

// ---------------------------------------------------------------------------
// Fused y or z sweep of the filtered kinds over the caller's partitions in
// that direction (Dirichlet walls). The rows are generated as k_build forms
// them (a = -s.a, b = 1 - s.b, c = -s.c of the direction; d = for z the
// explicit rhs of th as k_rhs_z forms it, for y the rhs array; the wall
// fold-in on the system's first and last rows), and the solve is
// k_filtered_fused's (ftdma_kernels.hip), statement for statement: the bits
// of the materialized build + solve + scatter, x written into rhs.
//   VERSION 2, PASS 0: one pass with J[q] (k_bound_maxabs + k_compute_J).
//   VERSION 1, PASS 1: x0, xN of every (partition, system) to xbuf, the
//     tile's largest |x0|, |xN| per partition to partial[tile * Q + q];
//   VERSION 1, PASS 2: the same sweeps, x0, xN from xbuf, J[q] applied.
// One lane per (system, partition q); S systems per workgroup.
// ---------------------------------------------------------------------------
// rows is taken by value: a reference kept the row source's rolling state
// in scratch memory (1.5 KB per lane in the 64-row filtered_v2 z variant)
template <int M, int VERSION, int PASS, typename Rows>
__device__ __forceinline__ void filt_lane(Rows rows, const double* __restrict__ th, int Q, int q,
                                          int S, int sl, int m, int n_sys, int s, bool live,
                                          double sa, double sb, double sc, size_t lo, size_t hi,
                                          size_t rs, double* pbase, const int* __restrict__ Jp,
                                          double* __restrict__ xbuf, double* __restrict__ partial,
                                          int tile, double* lds) {
    const int J = (VERSION == 2 || PASS == 2) ? Jp[q] : m;
    constexpr int PF = M >= 32 ? 2 : 4;
    double rd[M];
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) rd[j] = rows(j);

    double a[M], c[M], d[M];
    double aN = 0.0, cN = 0.0, dN = 0.0;
    double b1 = 1.0;
    // forward sweep (k_fwd_pass_v1 / k_fwd_pass_v2)
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) rd[j + PF] = rows(j + PF);
        if (j < m) {
            double aj, bj, cj, dj;
            adi_row(j == 0 && q == 0, j == m - 1 && q == Q - 1, sa, sb, sc, rd[j], th, lo, hi, rs,
                    aj, bj, cj, dj);
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
    // backward sweep (k_bwd_pass_v1 / k_bwd_pass_v2)
    const int lo_c = (m - 2) - J;
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
                if (j > lo_c) c[j] = -cj * c[j + 1];
            }
        }
    }
    // pack_reduced: row 0 keeps d0, a0 (its c0 coupling dropped)
    const double rr = 1.0 / (1.0 - a[1] * c[0]);
    const double d0p = rr * (d[0] - c[0] * d[1]);
    const double a0p = rr * a[0];

    double x0, xN;
    if (PASS == 2) {
        x0 = xbuf[(size_t)(2 * q) * n_sys + s];
        xN = xbuf[(size_t)(2 * q + 1) * n_sys + s];
    } else {
        // neighbour exchange through LDS and the interface 2x2 solves, both
        // formed on every lane and selected (no lane-divergent branch while
        // the rows are held in registers)
        double* sC = lds;
        double* sDr = sC + Q * S;
        double* sA = sDr + Q * S;
        double* sDl = sA + Q * S;
        sC[q * S + sl] = cN;
        sDr[q * S + sl] = dN;
        sA[q * S + sl] = a0p;
        sDl[q * S + sl] = d0p;
        __syncthreads();
        const int ql = q > 0 ? q - 1 : 0, qr = q < Q - 1 ? q + 1 : Q - 1;
        const double xl = (d0p - a0p * sDr[ql * S + sl]) / (1.0 - sC[ql * S + sl] * a0p);
        const double xr = (dN - cN * sDl[qr * S + sl]) / (1.0 - sA[qr * S + sl] * cN);
        x0 = q > 0 ? xl : d0p;
        xN = q < Q - 1 ? xr : dN;
    }
    if (PASS == 1) {
        if (live) {
            xbuf[(size_t)(2 * q) * n_sys + s] = x0;
            xbuf[(size_t)(2 * q + 1) * n_sys + s] = xN;
        }
        // dead lanes shadow a live system of this workgroup: the max holds
        double v = fmax(fabs(x0), fabs(xN));
        for (int off = S >> 1; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        if (sl == 0) partial[(size_t)tile * Q + q] = v;
        return;
    }
    // k_final_pass: truncated correction
    const int jL = J < m - 2 ? J : m - 2;
    const int jR = (m - 1) - J > 1 ? (m - 1) - J : 1;
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
                pbase[(size_t)j * rs] = x;
            }
        }
    }
}

// DIR 2 (z: systems (x, y), rows strided by a plane, d formed from th) or
// 1 (y: systems (x, z), rows strided by a line, d read from rhs)
template <int M, int VERSION, int PASS, int DIR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M >= 64 ? 1 : 2))) void k_adi_filt(
    const double* __restrict__ th, double* rhs, Dims g, int nz, double sa, double sb, double sc,
    int Q, int lgS, Stencil st, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, const int* __restrict__ Jp, double* __restrict__ xbuf,
    double* __restrict__ partial) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int sl = t & (S - 1);
    const int q = t >> lgS;
    const int n_sys = g.nx * (DIR == 2 ? g.ny : nz);
    const int N = DIR == 2 ? nz : g.ny;
    // z: each XCD takes a contiguous run of system blocks, so the halo lines
    // of th its workgroups read are the ones their neighbours read
    const int blk = DIR == 2 ? ftd::xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const int s_raw = blk * S + sl;
    const bool live = s_raw < n_sys;
    const int s = live ? s_raw : n_sys - 1;  // dead lanes shadow the last system
    const int ii = s % g.nx + 1, oo = s / g.nx + 1;
    const size_t rs = DIR == 2 ? (size_t)g.px * g.ny1 : (size_t)g.px;
    const size_t here0 = DIR == 2 ? cell(g, ii, oo, 1) : cell(g, ii, 1, oo);
    int r0, m;
    {
        const int base = N / Q, rem = N % Q;
        r0 = q * base + (q < rem ? q : rem);
        m = base + (q < rem ? 1 : 0);
    }
    const size_t lo = here0, hi = here0 + (size_t)(N - 1) * rs;
    double* const pbase = rhs + here0 + (size_t)r0 * rs;
    if constexpr (DIR == 2) {
        RowRhsZ rows{th, cz, here0 + (size_t)r0 * rs, (size_t)g.px, rs, m, r0 + 1,
                     st.xa, st.xc, st.ya, st.yc, st.za, st.zc, 1.0 + st.xb + st.yb + st.zb,
                     st.c0 * cx[ii] * cy[oo]};
        filt_lane<M, VERSION, PASS>(rows, th, Q, q, S, sl, m, n_sys, s, live, sa, sb, sc, lo, hi,
                                    rs, pbase, Jp, xbuf, partial, blk, lds);
    } else {
        RowAt rows{pbase, rs, m};
        filt_lane<M, VERSION, PASS>(rows, th, Q, q, S, sl, m, n_sys, s, live, sa, sb, sc, lo, hi,
                                    rs, pbase, Jp, xbuf, partial, blk, lds);
    }
}

// |d| of rows 0 and m-1 of every partition of direction DIR (2: z, d from th
// as k_rhs_z forms it; 1: y, d from rhs), d as k_build forms it with the
// wall fold-in on the system's first and last rows, the largest per
// partition over each block of 256 systems: the tile maxima that
// k_boundary_maxabs takes from the built D, for the fused filtered_v2 sweep
template <int DIR>
__global__ __launch_bounds__(256) void k_bound_maxabs(const double* __restrict__ th,
                                                      const double* __restrict__ rhs, Dims g,
                                                      int nz, int Q, double sa, double sc,
                                                      Stencil st, const double* __restrict__ cx,
                                                      const double* __restrict__ cy,
                                                      const double* __restrict__ cz,
                                                      double* __restrict__ partial) {
    __shared__ double red[4];
    const int q = blockIdx.y;
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int n_sys = g.nx * (DIR == 2 ? g.ny : nz);
    const int N = DIR == 2 ? nz : g.ny;
    double v = 0.0;
    if (s < n_sys) {
        const int ii = s % g.nx + 1, oo = s / g.nx + 1;
        const size_t sy = (size_t)g.px, sz = (size_t)g.px * g.ny1;
        const size_t rs = DIR == 2 ? sz : sy;
        const int base = N / Q, rem = N % Q;
        const int r0 = q * base + (q < rem ? q : rem);
        const int m = base + (q < rem ? 1 : 0);
        const size_t lo = DIR == 2 ? cell(g, ii, oo, 1) : cell(g, ii, 1, oo);
        const size_t hi = lo + (size_t)(N - 1) * rs;
        auto d_of = [&](int row) {
            const size_t c = lo + (size_t)row * rs;
            double d;
            if constexpr (DIR == 2) {
                const double diag = 1.0 + st.xb + st.yb + st.zb;
                const double cxy = st.c0 * cx[ii] * cy[oo];
                d = (st.xc * th[c + 1] + st.xa * th[c - 1]) +
                    (st.yc * th[c + sy] + st.ya * th[c - sy]) +
                    (st.zc * th[c + sz] + st.za * th[c - sz]) + diag * th[c] + cxy * cz[row + 1];
            } else {
                d = rhs[c];
            }
            if (row == 0) d += sa * (th[lo - rs] + th[lo]);
            if (row == N - 1) d += sc * (th[hi + rs] + th[hi]);
            return d;
        };
        v = fmax(fabs(d_of(r0)), fabs(d_of(r0 + m - 1)));
    }
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[(size_t)blockIdx.x * Q + q] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// line stride of the x sweep's LDS tile, in doubles: even, so that every
// line starts on a 16-byte boundary for the 16-byte DMAs
__host__ __device__ inline int x_tile_stride(int N) { return N + 2 - (N & 1); }

// X sweep (dir 0): the rows of a system are contiguous in x, so the
// workgroup first copies its S lines of rhs into an LDS tile (coalesced,
// line stride x_tile_stride(N)), solves from the tile, writes x back into the tile and copies the tile to
// theta (coalesced).
template <int M, bool THOMAS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M >= 64 ? 1 : 2))) void k_adi_x(double* th, const double* __restrict__ rhs, Dims g,
                                               int nz, double sa, double sb, double sc, int Q,
                                               int lgS) {
    extern __shared__ double lds[];
    const int S = 1 << lgS;
    const int t = threadIdx.x;
    const int T = blockDim.x;
    const int sl = t & (S - 1);
    const int q = t >> lgS;
    const int n_sys = g.ny * nz;
    const int N = g.nx;
    const int LS = x_tile_stride(N);
    double* tile = lds;
    double* red = tile + (size_t)S * LS;
    const int first = blockIdx.x * S;
    const int nl = n_sys - first < S ? n_sys - first : S;
    // a group of `step` threads (a wave, or the whole block if smaller)
    // copies whole lines (contiguous in x); threads past the last full group
    // only take part in the solve (its lane index is the wave's lane index).
    // The copy-in is direct global->LDS DMA of 16 bytes per lane (interior
    // cells start on 128-byte boundaries, tile lines on 16-byte ones), so
    // every line of the tile is in flight at once with no register staging
    // and a wave's DMAs stay under the count of operations it can track.
    const int step = T < 64 ? T : 64, nw = T / step, w = t / step, lane = t - w * step;
    const int nu = N >> 1;  // 16-byte units per line; an odd N leaves 8 bytes
    for (int l = w < nw ? w : nl; l < nl; l += nw) {
        const int s = first + l;
        const char* src = (const char*)(rhs + cell(g, 1, s % g.ny + 1, s / g.ny + 1));
        char* dst = (char*)(tile + (size_t)l * LS);
        for (int e = 0; e < nu; e += step) {
            if (e + lane < nu)
                __builtin_amdgcn_global_load_lds((const void*)(src + 16 * (e + lane)),
                                                 (__attribute__((address_space(3))) void*)(dst + 16 * e),
                                                 16, 0, 0);
        }
        if ((N & 1) && lane < 2)
            __builtin_amdgcn_global_load_lds((const void*)(src + 16 * nu + 4 * lane),
                                             (__attribute__((address_space(3))) void*)(dst + 16 * nu),
                                             4, 0, 0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // threads past S * Q (a Thomas tile of fewer than 64 lines is launched as
    // one full wave, so that its copies use 64 lanes) only copy: the Thomas
    // solve has no barrier
    if (q < Q) {
        const int s_raw = first + sl;
        const bool live = s_raw < n_sys;
        const int s = live ? s_raw : n_sys - 1;
        const size_t here0 = cell(g, 1, s % g.ny + 1, s / g.ny + 1);
        int r0, cnt;
        {
            const int base = N / Q, rem = N % Q;
            r0 = q * base + (q < rem ? q : rem);
            cnt = base + (q < rem ? 1 : 0);
        }
        const int lt = live ? sl : 0;  // dead lanes read a loaded line
        double* myline = tile + (size_t)lt * LS + r0;
        RowAt rows{myline, 1, cnt};
        adi_solve_lane<M, 2, THOMAS>(
            N, Q, q, r0, cnt, S, sl, sa, sb, sc, th, here0, here0 + (size_t)(N - 1), 1, live, red,
            rows, [&](int j, double x) { myline[j] = x; });
    }
    __syncthreads();
    for (int l = w < nw ? w : nl; l < nl; l += nw) {
        const int s2 = first + l;
        double* dst = th + cell(g, 1, s2 % g.ny + 1, s2 / g.ny + 1);
        const double* src = tile + l * LS;
#pragma unroll 4
        for (int i = lane; i < N; i += step) dst[i] = src[i];
    }
}

// solution of one sweep back into the field (rhs for Z and Y, theta for X)
__global__ void k_scatter(const double* __restrict__ D, double* __restrict__ dst, Dims g, int dir) {
    const size_t n_in = (size_t)g.nx * g.ny * g.nz;
    const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_in) return;
    int ii, jj, kk;
    if (dir == 2) {
        const size_t ns = (size_t)g.nx * g.ny;
        kk = (int)(o / ns);
        const int r = (int)(o % ns);
        jj = r / g.nx;
        ii = r % g.nx;
    } else if (dir == 1) {
        const size_t ns = (size_t)g.nx * g.nz;
        jj = (int)(o / ns);
        const int r = (int)(o % ns);
        kk = r / g.nx;
        ii = r % g.nx;
    } else {
        const size_t ns = (size_t)g.ny * g.nz;
        ii = (int)(o / ns);
        const int r = (int)(o % ns);
        kk = r / g.ny;
        jj = r % g.ny;
    }
    dst[cell(g, ii + 1, jj + 1, kk + 1)] = D[o];
}

thread_local std::string g_heat_err;

ftd_status herr(ftd_status s, const std::string& m) {
    g_heat_err = m;
    return s;
}

#define HH(call)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return herr(FTD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define HT(expr)                                               \
    do {                                                       \
        ftd_status s_ = (expr);                                \
        if (s_ != FTD_OK) return herr(s_, ftd_last_error_string()); \
    } while (0)

}  // namespace

struct ftd_heat {
    Dims g{};
    int nz = 0;
    int rhs_kb = kRhsKB;  // planes per z run of k_rhs_z
    int parts[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    int kind = FTD_KIND_PASCAL;
    double dt = 0.0;
    Stencil s{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *th = nullptr, *rhs = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;
    std::vector<double> hx, hy, hz, sx_, sy_, sz_, cx_, cy_, cz_;
    ftd_handle* solver[3] = {nullptr, nullptr, nullptr};  // x, y, z sweeps
    int rho_set[3] = {0, 0, 0};  // the solver's column-0 bounds are set (materialized sweeps)
    long steps_done = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    // fused sweep per direction (x, y, z): enabled, rows per lane, Thomas,
    // partitions, log2 systems per workgroup, dynamic LDS bytes
    struct Fused {
        int on = 0, M = 0, thomas = 0, Q = 0, lgS = 0;
        size_t lds = 0;
    } fused[3];
    // fused y, z sweeps of the filtered kinds over the caller's partitions
    // (k_adi_filt): enabled, rows per lane, partitions, log2 systems per
    // workgroup, dynamic LDS bytes of the exchange (x stays materialized)
    struct Filt {
        int on = 0, M = 0, Q = 0, lgS = 0;
        size_t lds = 0;
    } filt[3];
    int use_fused = 1;
    int fuse_rhs = 0;  // the fused z sweep forms the rhs (no k_rhs_z pass) this step
    char plan[256] = "";  // how the last enqueued step ran each sweep (ftd_heat_last_plan)
};

namespace {

// The partition count the sweep's solver picks for (N rows, P partitions):
// P == 1 -> plain Thomas when N <= 64, else ceil(N / 64) internal
// partitions (up to 16); P > 1 -> P. Then the fused sweep's shape, or off.
// rows_per_lane (16, 32, 64; 0 = automatic) sets how many rows one lane
// holds, and with one caller partition (P == 1) thereby the number of
// internal partitions Q; systems_per_tile (power of two; 0 = automatic)
// the systems of one workgroup (S * Q <= 256)
// compute units of the current device (1 if it cannot be read)
int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        return 1;
    return cus;
}

// planes per z run of k_rhs_z: kRhsKB, halved (down to 4) while the grid of
// 64 x 4 x kb tiles has fewer workgroups than the device has CUs (64^3: 32
// tiles of 32 planes would leave most CUs idle)
int rhs_planes(int nx, int ny, int nz) {
    const int cus = device_cus();
    const long xy = (long)((nx + 63) / 64) * ((ny + 3) / 4);
    int kb = kRhsKB;
    while (kb > 4 && xy * ((nz + kb - 1) / kb) < cus) kb >>= 1;
    return kb;
}

void plan_fused(ftd_heat* h, int d, int N, int n_sys, int rows_per_lane = 0,
                int systems_per_tile = 0) {
    ftd_heat::Fused f;
    const int P = h->parts[d];
    // with one caller partition every kind solves alike (the filtered kinds'
    // single-partition solve is the exact one), so their sweeps fuse too
    if ((h->kind != FTD_KIND_PASCAL && P != 1) || h->periodic[d] || N < 3 || n_sys < 1) {
        h->fused[d] = f;
        return;
    }
    // automatic: one lane per system up to 64 rows (plain Thomas, the
    // reference's single-partition solve); longer systems in partitions of
    // 32 rows, the fastest shape measured at 256^3 in every direction
    const int Mc = rows_per_lane ? rows_per_lane : N <= 64 ? 64 : 32;
    int Q;
    if (P == 1)
        Q = N <= Mc ? 1 : (N + Mc - 1) / Mc;
    else
        Q = P;
    const int m_max = (N + Q - 1) / Q, m_min = N / Q;
    if (Q > 16 || m_max > Mc || (Q > 1 && m_min < 3)) {
        h->fused[d] = f;
        return;
    }
    f.Q = Q;
    f.thomas = Q == 1;
    f.M = rows_per_lane ? rows_per_lane : m_max <= 16 ? 16 : m_max <= 32 ? 32 : 64;
    int S = systems_per_tile ? systems_per_tile : 64;
    while (S * Q > 256) S >>= 1;
    // x tiles: each workgroup copies its lines in, solves and copies them out
    // in one pass, so a grid of fewer tiles than CUs leaves CUs idle. Halve
    // the tile down to 16 lines until the grid covers the CUs (64^3: 256
    // tiles of 16 lines, 0.073 ms per step against 0.087 with 64 tiles of 64;
    // 8-line tiles were slower again). The bits do not depend on the tile.
    if (d == 0 && !systems_per_tile) {
        const int cus = device_cus();
        while (S > 16 && (n_sys + S - 1) / S < cus) S >>= 1;
    }
    auto lds_of = [&](int S_) {
        const size_t red = f.thomas ? 0 : (size_t)3 * 2 * Q * S_;
        const size_t tile = d == 0 ? (size_t)S_ * x_tile_stride(N) : 0;
        return sizeof(double) * (red + tile);
    };
    while (S > 1 && lds_of(S) > 160 * 1024) S >>= 1;
    if (lds_of(S) > 160 * 1024) {
        h->fused[d] = f;
        return;
    }
    int lg = 0;
    while ((1 << lg) < S) ++lg;
    f.lgS = lg;
    f.lds = lds_of(S);
    f.on = 1;
    h->fused[d] = f;
}

// the filtered kinds' fused sweep of direction d (1 = y, 2 = z): a filtered
// kind split there into 2..16 Dirichlet partitions of 3..64 rows. The tile
// is halved (down to 16 systems) until the grid covers the CUs; the bits do
// not depend on it.
void plan_filt(ftd_heat* h, int d) {
    ftd_heat::Filt f;
    const int Q = h->parts[d], N = d == 2 ? h->nz : h->g.ny;
    const int n_sys = h->g.nx * (d == 2 ? h->g.ny : h->nz);
    h->filt[d] = f;
    if (d == 0 || h->kind == FTD_KIND_PASCAL || Q < 2 || Q > 16 || h->periodic[d] || n_sys < 1)
        return;
    const int m_max = (N + Q - 1) / Q, m_min = N / Q;
    if (m_max > 64 || m_min < 3) return;
    f.M = m_max <= 16 ? 16 : m_max <= 32 ? 32 : 64;
    int S = 64;
    while (S * Q > 256) S >>= 1;
    const int cus = device_cus();
    while (S > 16 && (n_sys + S - 1) / S < cus) S >>= 1;
    int lg = 0;
    while ((1 << lg) < S) ++lg;
    f.Q = Q;
    f.lgS = lg;
    f.lds = sizeof(double) * 4 * (size_t)Q * S;
    f.on = 1;
    h->filt[d] = f;
}

// whether this step's sweep of direction d can run fused: planned, fused
// sweeps on, the solver's column-0 bounds set (by the first, materialized
// step) and the solver in its plain mode (no learned J, no guard)
bool filt_ready(ftd_heat* h, int d) {
    if (!h->use_fused || !h->filt[d].on || !h->rho_set[d]) return false;
    int* J;
    const double* rho;
    double ev;
    int ed;
    return ftd::handle_filtered_state(h->solver[d], &J, &rho, &ev, &ed) == FTD_OK;
}

template <int M, int DIR>
hipError_t launch_filt_t(ftd_heat* h, hipStream_t st) {
    const ftd_heat::Filt& f = h->filt[DIR];
    int* J;
    const double* rho;
    double ev;
    int ed;
    if (ftd::handle_filtered_state(h->solver[DIR], &J, &rho, &ev, &ed) != FTD_OK)
        return hipErrorInvalidValue;
    const Dims g = h->g;
    const int S = 1 << f.lgS, n_sys = g.nx * (DIR == 2 ? g.ny : h->nz);
    const int N = DIR == 2 ? h->nz : g.ny;
    const dim3 grid((n_sys + S - 1) / S), block(S * f.Q);
    // scratch: the materialized sweeps' A (x0, xN: 2Q x n_sys) and B (tile
    // maxima); sweeps run one after the other on the stream
    double* const xbuf = h->A;
    double* const partial = h->B;
    const Stencil& s = h->s;
    const double sa = DIR == 2 ? s.za : s.ya, sb = DIR == 2 ? s.zb : s.yb,
                 sc = DIR == 2 ? s.zc : s.yc;
    hipError_t e;
    if (h->kind == FTD_KIND_FILTERED_V2) {
        const dim3 gb((n_sys + 255) / 256, f.Q);
        hipLaunchKernelGGL((k_bound_maxabs<DIR>), gb, dim3(256), 0, st, h->th, h->rhs, g, h->nz,
                           f.Q, sa, sc, s, h->cx, h->cy, h->cz, partial);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = ftd::launch_compute_J(partial, (int)gb.x, f.Q, N, rho, ev, ed, 0, J, st)) !=
            hipSuccess)
            return e;
        hipLaunchKernelGGL((k_adi_filt<M, 2, 0, DIR>), grid, block, f.lds, st, h->th, h->rhs, g,
                           h->nz, sa, sb, sc, f.Q, f.lgS, s, h->cx, h->cy, h->cz, J, xbuf,
                           partial);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_adi_filt<M, 1, 1, DIR>), grid, block, f.lds, st, h->th, h->rhs, g, h->nz,
                       sa, sb, sc, f.Q, f.lgS, s, h->cx, h->cy, h->cz, J, xbuf, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = ftd::launch_compute_J(partial, (int)grid.x, f.Q, N, rho, ev, ed, 0, J, st)) !=
        hipSuccess)
        return e;
    hipLaunchKernelGGL((k_adi_filt<M, 1, 2, DIR>), grid, block, 0, st, h->th, h->rhs, g, h->nz,
                       sa, sb, sc, f.Q, f.lgS, s, h->cx, h->cy, h->cz, J, xbuf, partial);
    return hipGetLastError();
}

hipError_t launch_filt(ftd_heat* h, int d, hipStream_t st) {
#define FTD_FILT_CASE(m_)                                                              \
    if (h->filt[d].M == m_) return d == 2 ? launch_filt_t<m_, 2>(h, st) : launch_filt_t<m_, 1>(h, st);
    FTD_FILT_CASE(16)
    FTD_FILT_CASE(32)
    FTD_FILT_CASE(64)
#undef FTD_FILT_CASE
    return hipErrorInvalidValue;
}

template <int M, bool TH>
hipError_t launch_adi_t(ftd_heat* h, int d, hipStream_t st) {
    const ftd_heat::Fused& f = h->fused[d];
    const Dims g = h->g;
    const int S = 1 << f.lgS;
    const int n_sys = d == 0 ? g.ny * h->nz : d == 1 ? g.nx * h->nz : g.nx * g.ny;
    // an x tile of the Thomas kind with fewer than 64 lines runs as one full
    // wave: the extra threads take part in the copies only
    const int T = d == 0 && TH && S * f.Q < 64 ? 64 : S * f.Q;
    const dim3 grid((n_sys + S - 1) / S), block(T);
    const double sa[3] = {h->s.xa, h->s.ya, h->s.za}, sb[3] = {h->s.xb, h->s.yb, h->s.zb},
                 sc[3] = {h->s.xc, h->s.yc, h->s.zc};
    if (d == 0) {
        hipLaunchKernelGGL((k_adi_x<M, TH>), grid, block, f.lds, st, h->th, h->rhs, g, h->nz,
                           sa[0], sb[0], sc[0], f.Q, f.lgS);
    } else if (d == 2 && h->fuse_rhs) {
        hipLaunchKernelGGL((k_adi_yz<M, TH, true>), grid, block, f.lds, st, h->th, h->rhs, g,
                           h->nz, d, sa[d], sb[d], sc[d], f.Q, f.lgS, h->s, h->cx, h->cy, h->cz);
    } else {
        hipLaunchKernelGGL((k_adi_yz<M, TH, false>), grid, block, f.lds, st, h->th, h->rhs, g,
                           h->nz, d, sa[d], sb[d], sc[d], f.Q, f.lgS, h->s, h->cx, h->cy, h->cz);
    }
    return hipGetLastError();
}

hipError_t launch_adi(ftd_heat* h, int d, hipStream_t st) {
    const ftd_heat::Fused& f = h->fused[d];
#define FTD_ADI_CASE(m_)                                                              \
    if (f.M == m_) return f.thomas ? launch_adi_t<m_, true>(h, d, st)                   \
                                   : launch_adi_t<m_, false>(h, d, st);
    FTD_ADI_CASE(16)
    FTD_ADI_CASE(32)
    FTD_ADI_CASE(64)
#undef FTD_ADI_CASE
    return hipErrorInvalidValue;
}

// dynamic LDS above the default limit must be declared for the kernel
template <int M, bool TH>
hipError_t allow_lds_t(int d, size_t bytes) {
    if (d == 0)
        return hipFuncSetAttribute((const void*)k_adi_x<M, TH>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    const hipError_t e = hipFuncSetAttribute((const void*)k_adi_yz<M, TH, false>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess || d != 2) return e;
    return hipFuncSetAttribute((const void*)k_adi_yz<M, TH, true>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t allow_lds(const ftd_heat::Fused& f, int d) {
#define FTD_ADI_CASE(m_)                                                                     \
    if (f.M == m_) return f.thomas ? allow_lds_t<m_, true>(d, f.lds) : allow_lds_t<m_, false>(d, f.lds);
    FTD_ADI_CASE(16)
    FTD_ADI_CASE(32)
    FTD_ADI_CASE(64)
#undef FTD_ADI_CASE
    return hipErrorInvalidValue;
}

// Dirichlet ghosts of every non-periodic face, one launch (grid z = face)
void launch_dirichlet(ftd_heat* h, hipStream_t st) {
    const Dims g = h->g;
    int faces = 0, mu = 0, mv = 0;
    for (int f = 0; f < 3; ++f) {
        if (h->periodic[f]) continue;
        faces |= 1 << f;
        mu = std::max(mu, f == 0 ? g.ny : g.nx);
        mv = std::max(mv, f == 2 ? g.ny : g.nz);
    }
    if (faces)
        hipLaunchKernelGGL(k_dirichlet, dim3((mu + 15) / 16, (mv + 15) / 16, 3), dim3(16, 16), 0, st,
                           h->th, g, faces, h->cx, h->cy, h->cz);
}

ftd_status enqueue_step(ftd_heat* h) {
    const Dims g = h->g;
    hipStream_t st = h->stream;
    const dim3 b2(16, 16);
    launch_dirichlet(h, st);
    // the fused z sweep forms the rhs itself (k_adi_yz<..., RHSZ>), except a
    // one-partition (Thomas) sweep with fewer workgroups than CUs: that one
    // is a serial chain per lane on part of the chip, and the separate rhs
    // pass, spread over every CU, costs less than the stencil loads added to
    // the chain (64^3: 0.057 against 0.0595 ms per step)
    const ftd_heat::Fused& fz = h->fused[2];
    const int z_groups = (g.nx * g.ny + (1 << fz.lgS) - 1) >> fz.lgS;
    h->fuse_rhs = h->use_fused && fz.on && !(fz.thomas && z_groups < device_cus());
    // the filtered kinds' fused z sweep forms the rhs itself as well
    const bool z_filt = filt_ready(h, 2);
    if (!h->fuse_rhs && !z_filt) {
        const int kb = h->rhs_kb;
        hipLaunchKernelGGL(k_rhs_z,
                           dim3(((g.nx + 63) / 64) * ((g.ny + 3) / 4) * ((h->nz + kb - 1) / kb)),
                           dim3(64, 4), 0, st, h->th, h->rhs, g, h->s, h->cx, h->cy, h->cz, kb);
        HH(hipGetLastError());
    }
    const size_t n_in = (size_t)g.nx * g.ny * h->nz;
    const dim3 b1(256), g1((unsigned)((n_in + 255) / 256));
    const int order[3] = {2, 1, 0};
    const int n_sys[3] = {g.ny * h->nz, g.nx * h->nz, g.nx * g.ny};
    const double sa[3] = {h->s.xa, h->s.ya, h->s.za}, sb[3] = {h->s.xb, h->s.yb, h->s.zb},
                 sc[3] = {h->s.xc, h->s.yc, h->s.zc};
    size_t np = 0;
    auto note_plan = [&](int d, const char* how, int M, int Q, int lgS) {
        const char* sep = np ? "; " : "";
        const char dn = "xyz"[d];
        int w = M ? std::snprintf(h->plan + np, sizeof h->plan - np, "%s%c:%s M=%d Q=%d S=%d", sep,
                                  dn, how, M, Q, 1 << lgS)
                  : std::snprintf(h->plan + np, sizeof h->plan - np, "%s%c:%s", sep, dn, how);
        if (w > 0) np = std::min(sizeof h->plan - 1, np + (size_t)w);
    };
    h->plan[0] = '\0';
    for (int q = 0; q < 3; ++q) {
        const int d = order[q];
        if (h->use_fused && h->fused[d].on) {
            // coefficients generated, rhs read and x written in place
            HH(launch_adi(h, d, st));
            note_plan(d, "fused", h->fused[d].M, h->fused[d].Q, h->fused[d].lgS);
            continue;
        }
        if ((d == 2 && z_filt) || (d == 1 && filt_ready(h, 1))) {
            // the filtered kinds over the caller's partitions: rows
            // generated (z: the rhs formed from th), x written into rhs
            HT(ftd_set_eps_constant(h->solver[d], h->dt));
            HH(launch_filt(h, d, st));
            note_plan(d, "filtered_fused", h->filt[d].M, h->filt[d].Q, h->filt[d].lgS);
            continue;
        }
        note_plan(d, "materialized", 0, 0, 0);
        hipLaunchKernelGGL(k_build, g1, b1, 0, st, h->th, h->rhs, g, d, sa[d], sb[d], sc[d],
                           h->periodic[d] ? 0 : 1, h->A, h->B, h->C, h->D);
        HH(hipGetLastError());
        ftd_handle* sv = h->solver[d];
        HT(ftd_set_eps_constant(sv, h->dt));
        // the coefficients are the stencil's constants in every step, and so
        // are their column-0 bounds (solve_theta.cu calls set_rho_device per
        // step on the same values): set once per solver, in the first step,
        // which always runs eagerly, so no captured step holds the launch
        if (!h->rho_set[d]) {
            HT(ftd_set_rho(sv, h->A, h->B, h->C));
            h->rho_set[d] = 1;
        }
        if (h->periodic[d])
            HT(ftd_solve_cyclic(sv, h->A, h->B, h->C, h->D));
        else
            HT(ftd_solve(sv, h->A, h->B, h->C, h->D));
        (void)n_sys;
        hipLaunchKernelGGL(k_scatter, g1, b1, 0, st, h->D, d == 0 ? h->th : h->rhs, g, d);
        HH(hipGetLastError());
    }
    // periodic ghosts (ghostcellUpdate), x then y then z
    for (int f = 0; f < 3; ++f) {
        if (!h->periodic[f]) continue;
        const int u = f == 0 ? g.ny1 : g.nx1, v = f == 2 ? g.ny1 : h->nz + 2;
        hipLaunchKernelGGL(k_periodic, dim3((u + 15) / 16, (v + 15) / 16), b2, 0, st, h->th, g, f);
    }
    HH(hipGetLastError());
    return FTD_OK;
}

// the materialized sweep of direction d runs the fused plan's partitioning
// (with one caller partition the plan picks it), so both give the same bits
ftd_status sync_solver_plan(ftd_heat* h, int d) {
    const ftd_heat::Fused& f = h->fused[d];
    if (f.on && !f.thomas && h->parts[d] == 1) {
        const int T = f.M == 64 ? 256 : 512;
        if (ftd_set_tuning(h->solver[d], f.M, T, 1 << f.lgS, f.Q) != FTD_OK)
            return herr(FTD_ERR_UNSUPPORTED, std::string("sweep solver: ") + ftd_last_error_string());
        return FTD_OK;
    }
    ftd_set_tuning(h->solver[d], 0, 0, 0, 0);
    return FTD_OK;
}

void centres(std::vector<double>& c, int n) {
    c.resize(n + 2);
    const double dd = 2.0 / (double)n;
    for (int i = 0; i <= n + 1; ++i) c[i] = -1.0 + dd / 2.0 + (double)(i - 1) * dd;
}

}  // namespace

extern "C" {

ftd_status ftd_heat_create(ftd_heat** out, int nx, int ny, int nz, const int* parts, int kind,
                           const int* periodic, double dt, void* stream) {
    if (!out) return herr(FTD_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3) return herr(FTD_ERR_INVALID, "need at least 3 cells per direction");
    if (kind < FTD_KIND_FILTERED || kind > FTD_KIND_PASCAL) return herr(FTD_ERR_INVALID, "bad kind");
    ftd_heat* h = new (std::nothrow) ftd_heat();
    if (!h) return herr(FTD_ERR_ALLOC, "out of host memory");
    h->g = Dims{nx, ny, nz, nx + 2, ny + 2, line_pitch(nx)};
    h->nz = nz;
    for (int d = 0; d < 3; ++d) {
        h->parts[d] = parts ? parts[d] : 1;
        h->periodic[d] = periodic ? (periodic[d] != 0) : 0;
    }
    h->kind = kind;
    h->dt = dt;
    const double ddx = 2.0 / nx, ddy = 2.0 / ny, ddz = 2.0 / nz;
    const double bx = dt / (2.0 * ddx * ddx), by = dt / (2.0 * ddy * ddy), bz = dt / (2.0 * ddz * ddz);
    h->s = Stencil{bx, -2.0 * bx, bx, by, -2.0 * by, by, bz, -2.0 * bz, bz, dt * 3.0 * kPi * kPi};
    centres(h->hx, nx);
    centres(h->hy, ny);
    centres(h->hz, nz);
    auto tab = [](const std::vector<double>& x, std::vector<double>& s, std::vector<double>& c) {
        s.resize(x.size());
        c.resize(x.size());
        for (size_t i = 0; i < x.size(); ++i) {
            s[i] = std::sin(kPi * x[i]);
            c[i] = std::cos(kPi * x[i]);
        }
    };
    tab(h->hx, h->sx_, h->cx_);
    tab(h->hy, h->sy_, h->cy_);
    tab(h->hz, h->sz_, h->cz_);
    const size_t ncell = field_doubles(h->g, nz), nin = (size_t)nx * ny * nz;
    hipError_t e = hipSuccess;
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        h->own_stream = true;
    }
    double** bufs[] = {&h->th, &h->rhs};
    for (double** b : bufs)
        if (e == hipSuccess) e = hipMalloc(b, ncell * sizeof(double));
    double** ins[] = {&h->A, &h->B, &h->C, &h->D};
    for (double** b : ins)
        if (e == hipSuccess) e = hipMalloc(b, nin * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->cx, (nx + 2) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->cy, (ny + 2) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->cz, (nz + 2) * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(h->cx, h->cx_.data(), (nx + 2) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->cy, h->cy_.data(), (ny + 2) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->cz, h->cz_.data(), (nz + 2) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->rhs, 0, ncell * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&h->e0);
    if (e == hipSuccess) e = hipEventCreate(&h->e1);
    if (e != hipSuccess) {
        ftd_heat_destroy(h);
        return herr(FTD_ERR_HIP, std::string("ftd_heat_create: ") + hipGetErrorString(e));
    }
    h->rhs_kb = rhs_planes(nx, ny, nz);
    // x, y, z sweeps: n_sys = ny*nz, nx*nz, nx*ny
    const int ns[3] = {ny * nz, nx * nz, nx * ny}, nr[3] = {nx, ny, nz};
    plan_filt(h, 1);
    plan_filt(h, 2);
    for (int d = 0; d < 3; ++d) {
        plan_fused(h, d, nr[d], ns[d]);
        if (h->fused[d].on && h->fused[d].lds > 64 * 1024) {
            hipError_t ea = allow_lds(h->fused[d], d);
            if (ea != hipSuccess) h->fused[d].on = 0;
        }
    }
    for (int d = 0; d < 3; ++d) {
        ftd_status st = ftd_create(&h->solver[d], kind, ns[d], nr[d], h->parts[d], dt, h->stream);
        if (st == FTD_OK) st = sync_solver_plan(h, d);
        if (st != FTD_OK) {
            std::string m = ftd_last_error_string();
            ftd_heat_destroy(h);
            return herr(st, "sweep solver: " + m);
        }
    }
    *out = h;
    return ftd_heat_init(h);
}

ftd_status ftd_heat_destroy(ftd_heat* h) {
    if (!h) return FTD_OK;
    if (h->exec) hipGraphExecDestroy(h->exec);
    if (h->graph) hipGraphDestroy(h->graph);
    for (ftd_handle* s : h->solver) ftd_destroy(s);
    for (double* p : {h->th, h->rhs, h->A, h->B, h->C, h->D, h->cx, h->cy, h->cz})
        if (p) hipFree(p);
    if (h->e0) hipEventDestroy(h->e0);
    if (h->e1) hipEventDestroy(h->e1);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
    delete h;
    return FTD_OK;
}

// manufactured initial field at every cell (mpi_subdomain.cpp:164-173),
// periodic wrap, Dirichlet ghosts (main.cpp:45-49)
ftd_status ftd_heat_init(ftd_heat* h) {
    if (!h) return herr(FTD_ERR_INVALID, "null handle");
    const int nx1 = h->g.nx1, ny1 = h->g.ny1, nz1 = h->nz + 2;
    std::vector<double> f((size_t)nx1 * ny1 * nz1);
    for (int k = 0; k < nz1; ++k)
        for (int j = 0; j < ny1; ++j)
            for (int i = 0; i < nx1; ++i)
                f[((size_t)k * ny1 + j) * nx1 + i] =
                    h->sx_[i] * h->sy_[j] * h->sz_[k] + h->cx_[i] * h->cy_[j] * h->cz_[k];
    const Dims g = h->g;
    HH(hipMemcpy2DAsync(h->th + kOrigin, sizeof(double) * g.px, f.data(), sizeof(double) * nx1,
                        sizeof(double) * nx1, (size_t)ny1 * nz1, hipMemcpyHostToDevice, h->stream));
    const dim3 b2(16, 16);
    for (int d = 0; d < 3; ++d) {
        if (!h->periodic[d]) continue;
        const int u = d == 0 ? g.ny1 : g.nx1, v = d == 2 ? g.ny1 : h->nz + 2;
        hipLaunchKernelGGL(k_periodic, dim3((u + 15) / 16, (v + 15) / 16), b2, 0, h->stream, h->th, g, d);
    }
    launch_dirichlet(h, h->stream);
    HH(hipGetLastError());
    HH(hipStreamSynchronize(h->stream));
    h->steps_done = 0;
    return FTD_OK;
}

// nsteps ADI steps; use_graph != 0 captures one step into a hipGraph on first
// use and replays it (the first step always runs eagerly so that every
// solver scratch buffer exists before capture). Waits for the last step and
// records its device time.
ftd_status ftd_heat_step(ftd_heat* h, int nsteps, int use_graph) {
    if (!h) return herr(FTD_ERR_INVALID, "null handle");
    for (int t = 0; t < nsteps; ++t) {
        const bool last = t == nsteps - 1;
        if (last) HH(hipEventRecord(h->e0, h->stream));
        if (use_graph && h->steps_done > 0) {
            if (!h->exec) {
                HH(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                ftd_status s = enqueue_step(h);
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamEndCapture(h->stream, &graph);
                if (s != FTD_OK) {
                    if (graph) hipGraphDestroy(graph);
                    return s;
                }
                HH(e);
                h->graph = graph;
                HH(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
            }
            HH(hipGraphLaunch(h->exec, h->stream));
        } else {
            ftd_status s = enqueue_step(h);
            if (s != FTD_OK) return s;
        }
        if (last) HH(hipEventRecord(h->e1, h->stream));
        ++h->steps_done;
    }
    if (nsteps > 0) {
        HH(hipEventSynchronize(h->e1));
        float ms = 0.f;
        HH(hipEventElapsedTime(&ms, h->e0, h->e1));
        h->last_ms = ms;
    }
    return FTD_OK;
}

// fused sweeps on (default) or off; a captured step is dropped so that the
// next graph step captures the new sequence
ftd_status ftd_heat_set_plan(ftd_heat* h, int dir, int rows_per_lane, int systems_per_tile) {
    if (!h) return herr(FTD_ERR_INVALID, "null handle");
    if (dir < 0 || dir > 2) return herr(FTD_ERR_INVALID, "dir must be 0, 1 or 2");
    if (rows_per_lane != 0 && rows_per_lane != 16 && rows_per_lane != 32 && rows_per_lane != 64)
        return herr(FTD_ERR_INVALID, "rows_per_lane must be 0, 16, 32 or 64");
    if (systems_per_tile < 0 || systems_per_tile > 256 ||
        (systems_per_tile & (systems_per_tile - 1)))
        return herr(FTD_ERR_INVALID, "systems_per_tile must be 0 or a power of two <= 256");
    const Dims g = h->g;
    const int ns[3] = {g.ny * h->nz, g.nx * h->nz, g.nx * g.ny}, nr[3] = {g.nx, g.ny, h->nz};
    plan_fused(h, dir, nr[dir], ns[dir], rows_per_lane, systems_per_tile);
    ftd_heat::Fused& f = h->fused[dir];
    if (f.on && f.lds > 64 * 1024 && allow_lds(f, dir) != hipSuccess) f.on = 0;
    if ((rows_per_lane || systems_per_tile) && !f.on)
        return herr(FTD_ERR_UNSUPPORTED, "no fused sweep with this plan for this direction");
    ftd_status st = sync_solver_plan(h, dir);
    if (st != FTD_OK) return st;
    return ftd_heat_set_fused(h, h->use_fused);  // drops a captured step
}

ftd_status ftd_heat_set_fused(ftd_heat* h, int enable) {
    if (!h) return herr(FTD_ERR_INVALID, "null handle");
    h->use_fused = enable != 0;
    if (h->exec) {
        HH(hipStreamSynchronize(h->stream));
        hipGraphExecDestroy(h->exec);
        h->exec = nullptr;
    }
    if (h->graph) {
        hipGraphDestroy(h->graph);
        h->graph = nullptr;
    }
    return FTD_OK;
}

// copies the field with ghosts ((nz+2)(ny+2)(nx+2), x fastest) to host memory
ftd_status ftd_heat_field(ftd_heat* h, double* host_out) {
    if (!h || !host_out) return herr(FTD_ERR_INVALID, "null argument");
    const Dims& g = h->g;
    HH(hipMemcpy2DAsync(host_out, sizeof(double) * g.nx1, h->th + kOrigin, sizeof(double) * g.px,
                        sizeof(double) * g.nx1, (size_t)g.ny1 * (h->nz + 2),
                        hipMemcpyDeviceToHost, h->stream));
    HH(hipStreamSynchronize(h->stream));
    return FTD_OK;
}

// L2 error against the manufactured solution at t = steps*dt (main.cpp:59-81;
// the sum is divided by (nx+1)(ny+1)(nz+1) as the reference does)
ftd_status ftd_heat_l2(ftd_heat* h, double* l2) {
    if (!h || !l2) return herr(FTD_ERR_INVALID, "null argument");
    const int nx = h->g.nx, ny = h->g.ny, nz = h->nz, nx1 = h->g.nx1, ny1 = h->g.ny1;
    std::vector<double> f((size_t)nx1 * ny1 * (nz + 2));
    ftd_status s = ftd_heat_field(h, f.data());
    if (s != FTD_OK) return s;
    const double decay = std::exp(-3.0 * kPi * kPi * (double)h->steps_done * h->dt);
    double err = 0.0;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double ex = std::sin(kPi * h->hx[i]) * std::sin(kPi * h->hy[j]) *
                                      std::sin(kPi * h->hz[k]) * decay +
                                  std::cos(kPi * h->hx[i]) * std::cos(kPi * h->hy[j]) *
                                      std::cos(kPi * h->hz[k]);
                const double d = f[((size_t)k * ny1 + j) * nx1 + i] - ex;
                err += d * d;
            }
    *l2 = std::sqrt(err / (double)(nx + 1) / (double)(ny + 1) / (double)(nz + 1));
    return FTD_OK;
}

double ftd_heat_last_step_ms(const ftd_heat* h) { return h ? h->last_ms : 0.0; }

const char* ftd_heat_last_plan(const ftd_heat* h) { return h ? h->plan : ""; }

const char* ftd_heat_last_error_string(void) { return g_heat_err.c_str(); }

}  // extern "C"
