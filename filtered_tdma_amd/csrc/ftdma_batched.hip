// ftdma_batched.hip -- uneven batches: thousands of independent systems of
// different lengths in one persistent launch.
//
// Layout: system s occupies elements off_s .. off_s + N_s - 1 of each of A,
// B, C, D (rows of one system contiguous, systems anywhere in the arrays).
// D is overwritten with x; A, B, C are read only.
//
// Work split (planned on the host once per batch layout):
//   * a system of N_s <= 64 rows is one lane: plain Thomas with row 0
//     divided by b (the reference's single-rank path, tdma_local_cuda.cu:32-64);
//   * a longer system is cut by para_range into Q_s = ceil(N_s / 64) chunks,
//     one lane each, on consecutive lanes of ONE wave: modified Thomas per
//     chunk, the 2Q_s-row reduced system (B == 1) solved in the wave's LDS by
//     the lane of chunk 0, then the interior update -- PaScaL with P = Q_s
//     (tdma_local_cuda.cu:134-224) for that system.
// A work item is one wave of 64 lane descriptors, of one kind only (chunked
// systems, or single-chunk systems), so the two code paths never diverge
// inside a wave. The chunked items come first, then the single-chunk ones,
// each range ordered by estimated cost, longest first, and dealt round-robin
// to the waves of a persistent grid. Waves are independent: no workgroup
// barrier anywhere.
//
// Systems longer than 64 * 64 rows take a lane-per-system Thomas in global
// memory (C is used as scratch for c', which the reference contract allows).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "ftdma_kernels.h"

#ifndef FTD_LD
#define FTD_LD(p) __builtin_nontemporal_load(p)
#define FTD_ST(v, p) __builtin_nontemporal_store((v), (p))
#endif

namespace ftd {

namespace {

constexpr int kChunk = 64;  // rows per lane

__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

#define FTD_BLOAD(j)                                  \
    do {                                              \
        const int jj_ = (j) < cl ? (j) : cl - 1;      \
        const size_t o_ = base + (size_t)jj_;         \
        ra[(j)] = FTD_LD(A + o_);                     \
        rb[(j)] = FTD_LD(B + o_);                     \
        rc[(j)] = FTD_LD(C + o_);                     \
        rd[(j)] = FTD_LD(D + o_);                     \
    } while (0)

// single-chunk systems: plain Thomas, the whole system in one lane
template <int M, int PF>
__device__ __forceinline__ void thomas_item(const double* __restrict__ A,
                                            const double* __restrict__ B,
                                            const double* __restrict__ C, double* __restrict__ D,
                                            const BatchLane& L) {
    int cnt = L.m;
    asm volatile("" : "+v"(cnt));  // keeps the row clamps inside the item loop
    const int cl = cnt > 0 ? cnt : 1;
    const size_t base = (size_t)L.off;
    double ra[M], rb[M], rc[M], rd[M];
    double c[M], d[M];
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_BLOAD(j);
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_BLOAD(j + PF);
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
        if (j < cnt) FTD_ST(d[j], D + base + (size_t)j);
}

// one chunk of a chunked system: modified Thomas (rows 0, 1 divided by b),
// the system's reduced rows solved in this wave's LDS by its chunk-0 lane,
// interior update
template <int M, int PF>
__device__ __forceinline__ void pascal_item(const double* __restrict__ A,
                                            const double* __restrict__ B,
                                            const double* __restrict__ C, double* __restrict__ D,
                                            const BatchLane& L, int l, double* rA, double* rC,
                                            double* rD) {
    int cnt = L.m;
    asm volatile("" : "+v"(cnt));
    const int cl = cnt > 0 ? cnt : 1;
    const size_t base = (size_t)L.off;
    const int q = L.q, Q = L.Q, lead = L.lead;
    double ra[M], rb[M], rc[M], rd[M];
    double a[M], c[M], d[M];
    double aN = 0.0, cN = 0.0, dN = 0.0;
#pragma unroll
    for (int j = 0; j < PF && j < M; ++j) FTD_BLOAD(j);
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j + PF < M) FTD_BLOAD(j + PF);
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
#pragma unroll
    for (int j = M - 3; j >= 1; --j) {
        if (j <= cnt - 3) {
            const double cj = c[j];
            d[j] = d[j] - cj * d[j + 1];
            a[j] = a[j] - cj * a[j + 1];
            c[j] = -cj * c[j + 1];
        }
    }
    const bool active = cnt > 0;
    if (active) {
        const double r = 1.0 / (1.0 - a[1] * c[0]);
        rA[2 * l] = r * a[0];
        rC[2 * l] = -r * c[0] * c[1];
        rD[2 * l] = r * (d[0] - c[0] * d[1]);
        rA[2 * l + 1] = aN;
        rC[2 * l + 1] = cN;
        rD[2 * l + 1] = dN;
    }
    wave_lds_sync();
    if (active && q == 0) {
        // reduced system of this system: rows 2*lead .. 2*(lead+Q)-1
        const int r0 = 2 * lead, NR = 2 * Q;
        double cp = rC[r0], dp = rD[r0];
        for (int k = 1; k < NR; ++k) {
            const double ak = rA[r0 + k];
            const double inv = 1.0 / (1.0 - ak * cp);
            dp = inv * (rD[r0 + k] - ak * dp);
            cp = inv * rC[r0 + k];
            rD[r0 + k] = dp;
            rC[r0 + k] = cp;
        }
        double dn = dp;
        for (int k = NR - 2; k >= 0; --k) {
            dn = rD[r0 + k] - rC[r0 + k] * dn;
            rD[r0 + k] = dn;
        }
    }
    wave_lds_sync();
    if (active) {
        const double x0 = rD[2 * l], xN = rD[2 * l + 1];
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
                FTD_ST(x, D + base + (size_t)j);
            }
        }
    }
    wave_lds_sync();  // the reduced rows are read before the next item writes them
}
#undef FTD_BLOAD

// Items [0, n_chunked) hold chunked systems, [n_chunked, n_items) single-
// chunk ones; wave g of the grid takes items g, g + G, g + 2G, ... of each
// range in turn.
template <int M, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1))) void k_batched(
    const double* __restrict__ A, const double* __restrict__ B, const double* __restrict__ C,
    double* __restrict__ D, const BatchLane* __restrict__ lanes, int n_chunked, int n_items) {
    __shared__ double red[4][3][2 * 64];
    const int w = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int G = gridDim.x * 4;
    const int g = blockIdx.x * 4 + w;
    for (int item = g; item < n_chunked; item += G)
        pascal_item<M, PF>(A, B, C, D, lanes[(size_t)item * 64 + l], l, red[w][0], red[w][1],
                           red[w][2]);
    // the single-chunk items continue the deal where the chunked ones ended
    const int start = n_chunked + ((g - n_chunked % G) + G) % G;
    for (int item = start; item < n_items; item += G)
        thomas_item<M, PF>(A, B, C, D, lanes[(size_t)item * 64 + l]);
}

// one lane per long system, c' kept in C
__global__ __launch_bounds__(256) void k_batched_long(const double* __restrict__ A,
                                                      const double* __restrict__ B,
                                                      double* __restrict__ C,
                                                      double* __restrict__ D,
                                                      const int64_t* __restrict__ off,
                                                      const int* __restrict__ n, int count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const size_t o = (size_t)off[k];
    const int N = n[k];
    double cp = C[o] / B[o];
    double dp = D[o] / B[o];
    C[o] = cp;
    D[o] = dp;
    for (int j = 1; j < N; ++j) {
        const size_t p = o + (size_t)j;
        const double r = 1.0 / (B[p] - A[p] * cp);
        dp = r * (D[p] - A[p] * dp);
        cp = r * C[p];
        C[p] = cp;
        D[p] = dp;
    }
    double xn = dp;
    for (int j = N - 2; j >= 0; --j) {
        const size_t p = o + (size_t)j;
        xn = D[p] - C[p] * xn;
        D[p] = xn;
    }
}

void split(int n, int P, int p, int& s, int& m) {
    const int base = n / P, rem = n % P;
    s = p * base + (p < rem ? p : rem);
    m = base + (p < rem ? 1 : 0);
}

}  // namespace

int batch_plan(int n_systems, const int64_t* row_offsets, const int* n_rows,
               std::vector<BatchLane>& lanes, std::vector<int64_t>& long_off,
               std::vector<int>& long_n, int& n_chunked) {
    n_chunked = 0;
    lanes.clear();
    long_off.clear();
    long_n.clear();
    std::vector<int> multi, single;
    for (int s = 0; s < n_systems; ++s) {
        const int N = n_rows[s];
        if (N < 1) return -1;
        const int Q = (N + kChunk - 1) / kChunk;
        if (Q > 64) {
            long_off.push_back(row_offsets[s]);
            long_n.push_back(N);
        } else if (Q >= 2) {
            multi.push_back(s);
        } else {
            single.push_back(s);
        }
    }
    auto chunks = [&](int s) { return (n_rows[s] + kChunk - 1) / kChunk; };
    // longest first; ties keep the caller's order
    std::stable_sort(multi.begin(), multi.end(), [&](int x, int y) { return chunks(x) > chunks(y); });
    std::stable_sort(single.begin(), single.end(),
                     [&](int x, int y) { return n_rows[x] > n_rows[y]; });

    struct Item {
        std::vector<BatchLane> l;
        double cost;
    };
    std::vector<Item> items;
    auto idle = [](int64_t off) {
        BatchLane b{};
        b.off = off;
        b.m = 0;
        return b;
    };
    // multi-chunk waves: biggest systems from the front, gaps filled with the
    // smallest from the back
    size_t lo = 0, hi = multi.size();
    while (lo < hi) {
        Item it;
        int used = 0, maxQ = 0, maxm = 0;
        auto put = [&](int s) {
            const int Q = chunks(s), N = n_rows[s];
            for (int q = 0; q < Q; ++q) {
                int st, m;
                split(N, Q, q, st, m);
                BatchLane b{};
                b.off = row_offsets[s] + st;
                b.m = m;
                b.q = (uint8_t)q;
                b.Q = (uint8_t)Q;
                b.lead = (uint8_t)used;
                b.kind = 1;
                it.l.push_back(b);
                maxm = std::max(maxm, m);
            }
            used += Q;
            maxQ = std::max(maxQ, Q);
        };
        put(multi[lo++]);
        while (lo < hi && used + chunks(multi[lo]) <= 64) put(multi[lo++]);
        while (lo < hi && used + chunks(multi[hi - 1]) <= 64) put(multi[--hi]);
        const int64_t pad = it.l.front().off;
        while ((int)it.l.size() < 64) it.l.push_back(idle(pad));
        for (auto& b : it.l) b.kind = 1;
        it.cost = maxm + 6.0 * maxQ;  // streamed rows + serial reduced rows
        items.push_back(std::move(it));
    }
    for (size_t k = 0; k < single.size(); k += 64) {
        Item it;
        int maxm = 0;
        for (size_t u = k; u < std::min(single.size(), k + 64); ++u) {
            const int s = single[u];
            BatchLane b{};
            b.off = row_offsets[s];
            b.m = n_rows[s];
            b.q = 0;
            b.Q = 1;
            b.lead = (uint8_t)(u - k);
            b.kind = 0;
            it.l.push_back(b);
            maxm = std::max(maxm, n_rows[s]);
        }
        const int64_t pad = it.l.front().off;
        while ((int)it.l.size() < 64) {
            BatchLane b = idle(pad);
            b.kind = 0;
            it.l.push_back(b);
        }
        it.cost = maxm;
        items.push_back(std::move(it));
    }
    // chunked items first, each range longest first
    const size_t nch = items.size() - (single.size() + 63) / 64;
    auto by_cost = [](const Item& x, const Item& y) { return x.cost > y.cost; };
    std::stable_sort(items.begin(), items.begin() + nch, by_cost);
    std::stable_sort(items.begin() + nch, items.end(), by_cost);
    n_chunked = (int)nch;
    lanes.reserve(items.size() * 64);
    for (auto& it : items) lanes.insert(lanes.end(), it.l.begin(), it.l.end());
    return (int)items.size();
}

int batched_resident_blocks() {
    static int cached[64] = {};
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev]) return cached[dev];
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batched<kChunk, 4>,
                                                     256, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 1;
    cached[dev] = per_cu * cus;
    return cached[dev];
}

hipError_t launch_batched(const double* A, const double* B, double* C, double* D,
                          const BatchLane* lanes, int n_chunked, int n_items,
                          const int64_t* long_off, const int* long_n, int n_long,
                          hipStream_t st) {
    if (n_items > 0) {
        const int blocks = std::min(batched_resident_blocks(), (n_items + 3) / 4);
        hipLaunchKernelGGL((k_batched<kChunk, 4>), dim3(blocks), dim3(256), 0, st, A, B, C, D,
                           lanes, n_chunked, n_items);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_long > 0) {
        hipLaunchKernelGGL(k_batched_long, dim3((n_long + 255) / 256), dim3(256), 0, st, A, B, C, D,
                           long_off, long_n, n_long);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace ftd
