// Channel pressure Poisson solve on one GPU: the reference's
// PressureSolver::solve (apps/channel_cpu/PressureSolver.cpp:493-533) with
// its pencil transposes gone. The field stays in HBM as [z][y][x]; one
// batched 2-D real-to-complex FFT of every z plane (hipFFT) leaves the
// spectra as [z][ky][kx] complex, which is already the row-major
// [n_row = nz x n_sys = ny * (nx/2+1)] layout of the complex tridiagonal
// solve (mode s = kx + (nx/2+1) * ky, the reference's system order), so
// the wall-normal solve runs on it in place (ftd_solve_complex), and the
// inverse FFT and the projection follow.
//
//   divergence   compute_rhs_ (:193-212)
//   FFT x, y     plan_fwd_x_, plan_fwd_y_ (:110-135) as one 2-D D2Z
//   z solve      solve_tdma_z_ (:400-464): a, c from the z metrics, the
//                modified wavenumbers of :72-78, the (0, 0) mode pinned
//   inverse FFT  plan_bwd_y_, plan_bwd_x_ as one 2-D Z2D, scale 1/(nx ny)
//   projection   project_ (:470-490) with dp's ghosts (Neumann in z,
//                periodic in x and y, :512-526) read by index
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ftdma.h"
#include "ftdma_xcd.h"

struct ftd_poisson {
    int nx = 0, ny = 0, nz = 0, nxh = 0, nm = 0;
    hipStream_t stream = nullptr;  // the caller's
    hipfftHandle fwd = 0, bwd = 0;
    int have_fwd = 0, have_bwd = 0;
    double* A = nullptr;    // [nz][nm] coefficients, built once
    double* B = nullptr;
    double* C = nullptr;
    double* rhs = nullptr;  // [nz][ny][nx]: divergence, then dp (unscaled)
    double* hat = nullptr;  // [nz][ny][nxh] complex, interleaved (re, im)
    double* met = nullptr;  // dx, dmx [nx+2]; dy, dmy [ny+2]; dz, dmz [nz+2]
    ftd_handle* solver = nullptr;
    double scale = 1.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int profiling = 0;  // time each step (waits for it)
    double last_ms = 0.0;
};

namespace {

thread_local std::string g_perr;

ftd_status perr(ftd_status s, const std::string& m) {
    g_perr = m;
    return s;
}

#define PH(call)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return perr(FTD_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)
#define PF(call)                                                                              \
    do {                                                                                      \
        hipfftResult r_ = (call);                                                             \
        if (r_ != HIPFFT_SUCCESS) return perr(FTD_ERR_HIP, std::string(#call ": hipFFT error ") + std::to_string((int)r_)); \
    } while (0)

// One axis of Grid::build_axis (Grid.cpp:47-95) on the whole domain:
// dx[i] = face spacing, dmx[i] = centre spacing, i = 0..n+1, mirrored
// ghosts at walls, wrapped ones when periodic.
void axis_metrics(int n, int uniform, double gamma, double L, bool periodic, double* dx,
                  double* dmx) {
    std::vector<double> xf(n + 1), xc(n + 2);
    if (uniform == 1 || std::fabs(gamma) < 1.0e-14) {
        for (int i = 0; i <= n; ++i) xf[i] = L * static_cast<double>(i) / static_cast<double>(n);
    } else {
        const double th = std::tanh(gamma);
        for (int i = 0; i <= n; ++i) {
            const double zeta = static_cast<double>(i) / static_cast<double>(n);
            xf[i] = 0.5 * L * (1.0 - std::tanh(gamma * (1.0 - 2.0 * zeta)) / th);
        }
    }
    for (int i = 1; i <= n; ++i) xc[i] = 0.5 * (xf[i - 1] + xf[i]);
    if (periodic) {
        xc[0] = xc[n] - L;
        xc[n + 1] = xc[1] + L;
    } else {
        xc[0] = xf[0] - (xc[1] - xf[0]);
        xc[n + 1] = xf[n] + (xf[n] - xc[n]);
    }
    for (int i = 1; i <= n; ++i) dx[i] = xf[i] - xf[i - 1];
    dx[0] = periodic ? dx[n] : dx[1];
    dx[n + 1] = periodic ? dx[1] : dx[n];
    for (int i = 1; i <= n + 1; ++i) dmx[i] = xc[i] - xc[i - 1];
    dmx[0] = periodic ? dmx[n] : dmx[1];
}

struct Dims3 {
    int nx, ny, nz;
};

__device__ inline size_t fidx(const Dims3& g, int i, int j, int k) {
    return (size_t)i + (size_t)(g.nx + 2) * ((size_t)j + (size_t)(g.ny + 2) * (size_t)k);
}

// (1/dt) div u* of every interior cell into rhs [z][y][x]
__global__ void k_divergence(const double* __restrict__ U, const double* __restrict__ V,
                             const double* __restrict__ W, double* __restrict__ rhs, Dims3 g,
                             const double* __restrict__ dx, const double* __restrict__ dy,
                             const double* __restrict__ dz, double inv_dt) {
    // one wave per 64 cells of a line; lines in (x, y, z) order dealt to the
    // XCDs in contiguous runs, so V's row j + 1 is the next line's own row,
    // read on the same XCD
    const int nbx = (g.nx + 63) / 64;
    const int T = ftd::xcd_tile(blockIdx.x, gridDim.x);
    const int i = (T % nbx) * 64 + threadIdx.x + 1;
    const int j = (T / nbx) % g.ny + 1, k = T / (nbx * g.ny) + 1;
    if (i > g.nx) return;
    const size_t c = fidx(g, i, j, k);
    const double d = (U[c + 1] - U[c]) / dx[i] + (V[c + (g.nx + 2)] - V[c]) / dy[j] +
                     (W[c + (size_t)(g.nx + 2) * (g.ny + 2)] - W[c]) / dz[k];
    rhs[((size_t)(k - 1) * g.ny + (j - 1)) * g.nx + (i - 1)] = d * inv_dt;
}

// the (0, 0) mode's wall row: A = C = 0, B = 1 (built in), right-hand side 0
__global__ void k_pin_zero_mode(double* hat) {
    if (threadIdx.x < 2) hat[threadIdx.x] = 0.0;
}

// P += dp; u -= dt grad dp on the faces (project_), dp = 0 + raw * scale as
// the reference's accumulation into a zeroed field gives it
__global__ void k_project(double* __restrict__ U, double* __restrict__ V, double* __restrict__ W,
                          double* __restrict__ P, const double* __restrict__ raw, Dims3 g,
                          double scale, double dt, const double* __restrict__ dmx,
                          const double* __restrict__ dmy, const double* __restrict__ dmz) {
    // tile order as in k_divergence: dp's row j - 1 was the previous line's
    const int nbx = (g.nx + 63) / 64;
    const int T = ftd::xcd_tile(blockIdx.x, gridDim.x);
    const int i = (T % nbx) * 64 + threadIdx.x + 1;
    const int j = (T / nbx) % g.ny + 1, k = T / (nbx * g.ny) + 1;
    if (i > g.nx) return;
    auto dp = [&](int ii, int jj, int kk) {
        return 0.0 + raw[((size_t)(kk - 1) * g.ny + (jj - 1)) * g.nx + (ii - 1)] * scale;
    };
    const double p = dp(i, j, k);
    const double pw = dp(i > 1 ? i - 1 : g.nx, j, k);  // periodic ghost in x
    const double ps = dp(i, j > 1 ? j - 1 : g.ny, k);  // periodic ghost in y
    const double pb = dp(i, j, k > 1 ? k - 1 : 1);     // Neumann ghost at the lower wall
    const size_t c = fidx(g, i, j, k);
    P[c] += p;
    U[c] -= dt * (p - pw) / dmx[i];
    V[c] -= dt * (p - ps) / dmy[j];
    W[c] -= dt * (p - pb) / dmz[k];
}

__global__ void k_unpack_dp(const double* __restrict__ raw, double* __restrict__ dp, int n,
                            double scale) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dp[t] = 0.0 + raw[t] * scale;
}

ftd_status poisson_spectral(ftd_poisson* p) {
    PF(hipfftExecD2Z(p->fwd, p->rhs, reinterpret_cast<hipfftDoubleComplex*>(p->hat)));
    hipLaunchKernelGGL(k_pin_zero_mode, dim3(1), dim3(64), 0, p->stream, p->hat);
    PH(hipGetLastError());
    ftd_status s = ftd_solve_complex(p->solver, p->A, p->B, p->C, p->hat);
    if (s != FTD_OK) return perr(s, std::string("z solve: ") + ftd_last_error_string());
    PF(hipfftExecZ2D(p->bwd, reinterpret_cast<hipfftDoubleComplex*>(p->hat), p->rhs));
    return FTD_OK;
}

}  // namespace

extern "C" {

ftd_status ftd_poisson_create(ftd_poisson** out, int nx, int ny, int nz, double Lx, double Ly,
                              double Lz, double gamma_z, int kind, int partitions,
                              double eps_constant, void* stream) {
    if (!out) return perr(FTD_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (nx < 2 || ny < 1 || nz < 3 || (nx & 1))
        return perr(FTD_ERR_INVALID, "need even nx >= 2, ny >= 1, nz >= 3");
    if (!(Lx > 0.0) || !(Ly > 0.0) || !(Lz > 0.0)) return perr(FTD_ERR_INVALID, "lengths must be > 0");
    const size_t cells = (size_t)nx * ny * nz;
    if (cells >= ((size_t)1 << 31)) return perr(FTD_ERR_INVALID, "grid too large");
    ftd_poisson* p = new ftd_poisson();
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->nxh = nx / 2 + 1;
    p->nm = p->nxh * ny;
    p->scale = 1.0 / (static_cast<double>(nx) * ny);
    auto bail = [&](ftd_status s) {
        ftd_poisson_destroy(p);
        return s;
    };
    // NULL is the null stream, as for ftd_create: ordered with the caller's
    // work there (torch's default stream among it)
    hipError_t e = hipSuccess;
    p->stream = (hipStream_t)stream;
    // metrics: x and y uniform periodic, z tanh-stretched walls
    std::vector<double> met(2 * (nx + 2) + 2 * (ny + 2) + 2 * (nz + 2));
    double* dx = met.data();
    double* dmx = dx + (nx + 2);
    double* dy = dmx + (nx + 2);
    double* dmy = dy + (ny + 2);
    double* dz = dmy + (ny + 2);
    double* dmz = dz + (nz + 2);
    axis_metrics(nx, 1, 0.0, Lx, true, dx, dmx);
    axis_metrics(ny, 1, 0.0, Ly, true, dy, dmy);
    axis_metrics(nz, 0, gamma_z, Lz, false, dz, dmz);
    // coefficients (solve_tdma_z_), modified wavenumbers (:72-78)
    const int nxh = p->nxh, nm = p->nm;
    std::vector<double> kx2(nxh), ky2(ny);
    const double hx = Lx / nx, hy = Ly / ny;
    for (int i = 0; i < nxh; ++i) kx2[i] = 2.0 * (1.0 - std::cos(2.0 * M_PI * i / nx)) / (hx * hx);
    for (int j = 0; j < ny; ++j) {
        const int jj = (j <= ny / 2) ? j : j - ny;
        ky2[j] = 2.0 * (1.0 - std::cos(2.0 * M_PI * jj / ny)) / (hy * hy);
    }
    std::vector<double> hA((size_t)nz * nm), hB((size_t)nz * nm), hC((size_t)nz * nm);
    for (int k = 0; k < nz; ++k) {
        const double a = k == 0 ? 0.0 : 1.0 / (dz[k + 1] * dmz[k + 1]);
        const double c = k == nz - 1 ? 0.0 : 1.0 / (dz[k + 1] * dmz[k + 2]);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nxh; ++i) {
                const size_t idx = (size_t)k * nm + i + (size_t)nxh * j;
                hA[idx] = a;
                hC[idx] = c;
                hB[idx] = -(a + c) - (kx2[i] + ky2[j]);
            }
    }
    hA[0] = 0.0;
    hB[0] = 1.0;
    hC[0] = 0.0;
    const size_t cb = sizeof(double) * (size_t)nz * nm;
    if (e == hipSuccess) e = hipMalloc(&p->A, cb);
    if (e == hipSuccess) e = hipMalloc(&p->B, cb);
    if (e == hipSuccess) e = hipMalloc(&p->C, cb);
    if (e == hipSuccess) e = hipMalloc(&p->rhs, sizeof(double) * cells);
    if (e == hipSuccess) e = hipMalloc(&p->hat, 2 * cb);
    if (e == hipSuccess) e = hipMalloc(&p->met, sizeof(double) * met.size());
    if (e == hipSuccess) e = hipMemcpy(p->A, hA.data(), cb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->B, hB.data(), cb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->C, hC.data(), cb, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->met, met.data(), sizeof(double) * met.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&p->e0);
    if (e == hipSuccess) e = hipEventCreate(&p->e1);
    if (e != hipSuccess)
        return bail(perr(FTD_ERR_HIP, std::string("ftd_poisson_create: ") + hipGetErrorString(e)));
    int n2[2] = {ny, nx};
    if (hipfftPlanMany(&p->fwd, 2, n2, nullptr, 1, nx * ny, nullptr, 1, nm, HIPFFT_D2Z, nz) !=
        HIPFFT_SUCCESS)
        return bail(perr(FTD_ERR_HIP, "hipfftPlanMany (D2Z) failed"));
    p->have_fwd = 1;
    if (hipfftPlanMany(&p->bwd, 2, n2, nullptr, 1, nm, nullptr, 1, nx * ny, HIPFFT_Z2D, nz) !=
        HIPFFT_SUCCESS)
        return bail(perr(FTD_ERR_HIP, "hipfftPlanMany (Z2D) failed"));
    p->have_bwd = 1;
    if (hipfftSetStream(p->fwd, p->stream) != HIPFFT_SUCCESS ||
        hipfftSetStream(p->bwd, p->stream) != HIPFFT_SUCCESS)
        return bail(perr(FTD_ERR_HIP, "hipfftSetStream failed"));
    ftd_status s = ftd_create(&p->solver, kind, nm, nz, partitions, eps_constant, p->stream);
    if (s != FTD_OK) return bail(perr(s, std::string("z solver: ") + ftd_last_error_string()));
    *out = p;
    return FTD_OK;
}

ftd_status ftd_poisson_destroy(ftd_poisson* p) {
    if (!p) return FTD_OK;
    if (p->have_fwd) hipfftDestroy(p->fwd);
    if (p->have_bwd) hipfftDestroy(p->bwd);
    if (p->solver) ftd_destroy(p->solver);
    for (double* b : {p->A, p->B, p->C, p->rhs, p->hat, p->met})
        if (b) hipFree(b);
    if (p->e0) hipEventDestroy(p->e0);
    if (p->e1) hipEventDestroy(p->e1);
    delete p;
    return FTD_OK;
}

ftd_handle* ftd_poisson_solver(ftd_poisson* p) { return p ? p->solver : nullptr; }

ftd_status ftd_poisson_solve(ftd_poisson* p, double* U, double* V, double* W, double* P, double dt) {
    if (!p) return perr(FTD_ERR_INVALID, "null handle");
    if (!U || !V || !W || !P) return perr(FTD_ERR_INVALID, "null field pointer");
    if (!(dt > 0.0)) return perr(FTD_ERR_INVALID, "dt must be > 0");
    const Dims3 g{p->nx, p->ny, p->nz};
    const double* dx = p->met;
    const double* dmx = dx + (p->nx + 2);
    const double* dy = dmx + (p->nx + 2);
    const double* dmy = dy + (p->ny + 2);
    const double* dz = dmy + (p->ny + 2);
    const double* dmz = dz + (p->nz + 2);
    const dim3 block(64), grid(((p->nx + 63) / 64) * p->ny * p->nz);
    if (p->profiling) PH(hipEventRecord(p->e0, p->stream));
    hipLaunchKernelGGL(k_divergence, grid, block, 0, p->stream, U, V, W, p->rhs, g, dx, dy, dz,
                       1.0 / dt);
    PH(hipGetLastError());
    ftd_status s = poisson_spectral(p);
    if (s != FTD_OK) return s;
    hipLaunchKernelGGL(k_project, grid, block, 0, p->stream, U, V, W, P, p->rhs, g, p->scale, dt,
                       dmx, dmy, dmz);
    PH(hipGetLastError());
    if (p->profiling) {
        PH(hipEventRecord(p->e1, p->stream));
        PH(hipEventSynchronize(p->e1));
        float ms = 0.0f;
        PH(hipEventElapsedTime(&ms, p->e0, p->e1));
        p->last_ms = ms;
    }
    return FTD_OK;
}

ftd_status ftd_poisson_set_profiling(ftd_poisson* p, int enable) {
    if (!p) return perr(FTD_ERR_INVALID, "null handle");
    p->profiling = enable != 0;
    return FTD_OK;
}

ftd_status ftd_poisson_solve_rhs(ftd_poisson* p, const double* rhs, double* dp) {
    if (!p) return perr(FTD_ERR_INVALID, "null handle");
    if (!rhs || !dp) return perr(FTD_ERR_INVALID, "null array pointer");
    const size_t cells = (size_t)p->nx * p->ny * p->nz;
    PH(hipMemcpyAsync(p->rhs, rhs, sizeof(double) * cells, hipMemcpyDeviceToDevice, p->stream));
    ftd_status s = poisson_spectral(p);
    if (s != FTD_OK) return s;
    hipLaunchKernelGGL(k_unpack_dp, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, p->stream,
                       p->rhs, dp, (int)cells, p->scale);
    PH(hipGetLastError());
    PH(hipStreamSynchronize(p->stream));
    return FTD_OK;
}

double ftd_poisson_last_ms(const ftd_poisson* p) { return p ? p->last_ms : 0.0; }

const char* ftd_poisson_last_error_string(void) { return g_perr.c_str(); }

}  // extern "C"
