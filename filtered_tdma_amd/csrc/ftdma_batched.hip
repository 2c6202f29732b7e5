// ftdma_batched.hip -- uneven batches: thousands of independent systems of
// different lengths in one persistent launch.
//
// Layout: system s occupies elements off_s .. off_s + N_s - 1 of each of A,
// B, C, D (rows of one system contiguous, systems anywhere in the arrays).
// D is overwritten with x; A and B are read only.
//
// Work split (planned on the host once per batch layout), with chunks of
// at most K = 48 rows (the default) or 32:
//   * a system of N_s <= K rows is one lane: plain Thomas with row 0
//     divided by b (the reference's single-rank path, tdma_local_cuda.cu:32-64);
//   * a longer system is cut by para_range into Q_s = ceil(N_s / K) chunks,
//     one lane each, on consecutive lanes of ONE work item: modified Thomas
//     per chunk, the 2Q_s-row reduced system (B == 1) solved in the item's
//     LDS (Thomas in the lane of chunk 0 below 16 rows, else PaScaL over
//     groups of its rows, one lane per group), then the interior update --
//     PaScaL with P = Q_s (tdma_local_cuda.cu:134-224) for that system.
// A work item is one workgroup of 128 lanes (two waves) of one kind only
// (chunked systems, or single-chunk systems). The chunked items come first,
// then the single-chunk ones, each range ordered by estimated cost, longest
// first, and dealt round-robin to the workgroups of a persistent grid.
//
// Rows of a chunk are contiguous in memory, so the lanes of a wave reading
// row j of their chunks would touch 64 scattered cache lines per load. The
// rows are therefore staged through LDS in coalesced blocks (below), and the
// elimination reads them from there.
//
// Systems longer than 128 * K rows take a lane-per-system Thomas in global
// memory (C is used as scratch for c', which the reference contract allows).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "ftdma_kernels.h"

// Cached access on purpose: a lane's stores (one row of its chunk each) and
// a block's 64-byte row segments fill 128-byte lines piecewise, and the L2
// merges them into whole lines (nontemporal access measured 6.8x the
// algorithmic write bytes and 2.7x the read bytes here).
#define FTD_BST(v, p) (*(p) = (v))

// two consecutive rows of x in one 16-byte store (8-byte aligned)
typedef double ftd_dbl2 __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ void store_pair(double* p, double x0, double x1) {
    ftd_dbl2 v;
    v.x = x0;
    v.y = x1;
    *(ftd_dbl2*)p = v;
}
constexpr int kDmaAux = 0;

namespace ftd {

namespace {

// Rows per lane, a plan parameter with two compiled shapes. 48 is the longest
// chunk whose elimination state (a, c, d of every row) fits the register
// file without scratch (64 does not), and the default: systems of up to
// 128 * 48 rows stay on the chunked kernels. 32-row chunks give more items
// and shorter serial parts per item; with the reduced systems solved over
// groups of rows they are the faster shape on the 20000-system batch (0.232
// against 0.245 ms). ftd_autotune_batched times both on the caller's layout.
constexpr int kChunk = 48;  // the default
constexpr int kLanes = 128;  // lanes per work item (two waves)

__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// workgroup barrier for LDS traffic only (DMAs in flight are left alone)
__device__ __forceinline__ void group_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ---------------------------------------------------------------------------
// Row staging. Blocks of RB = 8 rows of a wave's 64 chunks are brought into
// LDS by direct global->LDS loads of 16 bytes per lane (two rows of one
// chunk): one DMA covers 16 chunks x 8 rows, 4 DMAs per array and block, 16
// per block in all. DMA j of array X lands at
//   stage + X * ARR + j * SEG  (SEG = 1040: 1024 bytes + 16 of padding);
// its lane i = 16 k + cc brings rows 2k, 2k + 1 of chunk 16 j + cc, so row r
// of chunk c sits at (c >> 4) * SEG + (c & 15) * 16 + (r >> 1) * 256 +
// (r & 1) * 8: each lane reads its chunk's rows at immediate offsets from one
// base, the 64 bases spread over the banks (at most 2-way conflicts).
// Two buffers per wave: block b + 1 is in flight while block b is used, so
// the wait for block b is vmcnt(16) (block b + 1's 16 DMAs were issued after
// it), or vmcnt(0) for the last block of an item.
// A pair starting on a chunk's last row also reads the element after it,
// which is another chunk's or unused -- except after the batch's last
// element `last`, which is never read: that one pair is read one element
// lower, and its lane moves the row into place once the block has landed.
// ---------------------------------------------------------------------------
constexpr int kRB = 8;           // rows per staged block
constexpr int kSEG = 1040;       // bytes per DMA slot (16 chunks x 64 bytes + padding)
constexpr int kARR = 4 * kSEG;   // bytes per array per buffer
constexpr int kBUF = 4 * kARR;   // bytes per buffer
constexpr int kDMA = 16;         // DMAs per wave and block
static_assert(kBUF + 3 * 256 + 8 + 3 * kARR < 65536, "ds_read immediate offsets");

// FP32 storage: the same blocks of 8 rows, 16 bytes per lane being four rows
// of one chunk. DMA j of array X (j = 0, 1) lands at
//   stage + X * ARR4 + j * SEG  (32 chunks x 32 bytes + 16 of padding);
// its lane i = 32 k + cc brings rows 4k .. 4k + 3 of chunk 32 j + cc, so row
// r of chunk c sits at (c >> 5) * SEG + (c & 31) * 16 + (r >> 2) * 512 +
// (r & 3) * 4. 8 DMAs per block. A quad that would run past `last` is read
// ending at `last` (up to three elements lower), and its lane moves its rows
// back into place once the block has landed.
constexpr int kARR4 = 2 * kSEG;   // bytes per array per buffer
constexpr int kBUF4 = 4 * kARR4;  // bytes per buffer
constexpr int kDMA4 = 8;          // DMAs per wave and block

template <typename TS>
constexpr int buf_bytes() { return sizeof(TS) == 8 ? kBUF : kBUF4; }

typedef __attribute__((address_space(3))) void* lds_vptr;
typedef const __attribute__((address_space(3))) double* lds_cdptr;
typedef int ftd_int4 __attribute__((ext_vector_type(4)));

struct Stage {
    unsigned char* buf0;    // this wave's buffer 0 (buffer 1 follows at + kBUF)
    const BatchLane* desc;  // this wave's 64 lane descriptors (LDS)
    lds_cdptr lane;         // this lane's chunk slot in buffer 0
    int64_t last;           // the batch's last element
    int fix_blk;            // block of this chunk's pair / quad read low (-1: none)
    int fix_off;            // byte offset of that pair / quad from the lane's slot
    int fix_shift;          // elements it was read low by (FP32: 1..3)
};

// The chunk's last pair (FP64) or quad (FP32) is read low when it would run
// past `last`: then its rows sit `fix_shift` elements further on in the slot.
template <typename TS>
__device__ __forceinline__ void stage_set_fix(Stage& st, const BatchLane& L) {
    st.fix_blk = -1;
    st.fix_off = 0;
    st.fix_shift = 0;
    if (L.m < 1) return;
    const int r = L.m - 1;  // the chunk's last row
    if (sizeof(TS) == 8) {
        if (!(r & 1) && L.off + r == st.last) {
            st.fix_blk = r >> 3;
            st.fix_off = ((r & 7) >> 1) * 256;
            st.fix_shift = 1;
        }
    } else {
        const int64_t e0 = L.off + (r & ~3);  // first element of its last quad
        if (e0 + 3 > st.last) {
            st.fix_blk = r >> 3;
            st.fix_off = ((r >> 2) & 1) * 512;
            st.fix_shift = (int)(e0 + 3 - st.last);
        }
    }
}

// DMA of block b (rows 8b .. 8b+7 of every chunk of the wave, clamped to the
// chunk) into buffer t
__device__ __forceinline__ void stage_issue(const Stage& st, int t, int b, int l,
                                            const double* A, const double* B, const double* C,
                                            const double* D) {
    const int k = l >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // descriptor of chunk 16j + l%16, read in place (opaque, so that the
        // descriptors are not hoisted into registers for the whole sweep)
        ftd_int4 cd;
        asm volatile(
            "ds_read_b128 %0, %1\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(cd)
            : "v"((const __attribute__((address_space(3))) BatchLane*)(st.desc + 16 * j + (l & 15)))
            : "memory");
        const int64_t off = (int64_t)(((uint64_t)(uint32_t)cd.y << 32) | (uint32_t)cd.x);
        const int mm = cd.z > 0 ? cd.z : 1;
        const int row = kRB * b + 2 * k < mm ? kRB * b + 2 * k : mm - 1;
        int64_t e = off + row;
        if (e >= st.last) e = st.last - 1;
        unsigned char* dst = st.buf0 + t * kBUF + j * kSEG;
        __builtin_amdgcn_global_load_lds((const void*)(A + e), (lds_vptr)(dst + 0 * kARR), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(B + e), (lds_vptr)(dst + 1 * kARR), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(C + e), (lds_vptr)(dst + 2 * kARR), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(D + e), (lds_vptr)(dst + 3 * kARR), 16, 0,
                                         kDmaAux);
    }
}

// FP32: DMA of block b into buffer t, four rows per lane
__device__ __forceinline__ void stage_issue(const Stage& st, int t, int b, int l, const float* A,
                                            const float* B, const float* C, const float* D) {
    const int k = l >> 5;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        // descriptor of chunk 32j + l%32, read in place
        ftd_int4 cd;
        asm volatile(
            "ds_read_b128 %0, %1\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=v"(cd)
            : "v"((const __attribute__((address_space(3))) BatchLane*)(st.desc + 32 * j + (l & 31)))
            : "memory");
        const int64_t off = (int64_t)(((uint64_t)(uint32_t)cd.y << 32) | (uint32_t)cd.x);
        const int mm = cd.z > 0 ? cd.z : 1;
        const int row = kRB * b + 4 * k < mm ? kRB * b + 4 * k : mm - 1;
        int64_t e = off + row;
        if (e > st.last - 3) e = st.last - 3;
        unsigned char* dst = st.buf0 + t * kBUF4 + j * kSEG;
        __builtin_amdgcn_global_load_lds((const void*)(A + e), (lds_vptr)(dst + 0 * kARR4), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(B + e), (lds_vptr)(dst + 1 * kARR4), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(C + e), (lds_vptr)(dst + 2 * kARR4), 16, 0,
                                         kDmaAux);
        __builtin_amdgcn_global_load_lds((const void*)(D + e), (lds_vptr)(dst + 3 * kARR4), 16, 0,
                                         kDmaAux);
    }
}

// FP32: the quad read fix_shift elements low: element p + fix_shift of the
// slot goes to p, for every array
template <int T>
__device__ __forceinline__ void stage_fix_f32(const Stage& st) {
    typedef const __attribute__((address_space(3))) unsigned char* lds_cbptr;
    const lds_cbptr p = (lds_cbptr)st.lane + st.fix_off;
    ftd_int4 v0, v1, v2, v3;
    asm volatile(
        "ds_read_b128 %0, %4 offset:%5\n\t"
        "ds_read_b128 %1, %4 offset:%6\n\t"
        "ds_read_b128 %2, %4 offset:%7\n\t"
        "ds_read_b128 %3, %4 offset:%8\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
        : "v"(p), "n"(T * kBUF4), "n"(T * kBUF4 + kARR4), "n"(T * kBUF4 + 2 * kARR4),
          "n"(T * kBUF4 + 3 * kARR4)
        : "memory");
    const int s = st.fix_shift;
    auto shift = [s](ftd_int4 v) {
        ftd_int4 o = v;
        o.x = s == 1 ? v.y : s == 2 ? v.z : v.w;
        o.y = s == 1 ? v.z : v.w;
        o.z = v.w;
        return o;
    };
    v0 = shift(v0);
    v1 = shift(v1);
    v2 = shift(v2);
    v3 = shift(v3);
    asm volatile(
        "ds_write_b128 %0, %1 offset:%5\n\t"
        "ds_write_b128 %0, %2 offset:%6\n\t"
        "ds_write_b128 %0, %3 offset:%7\n\t"
        "ds_write_b128 %0, %4 offset:%8\n\t"
        "s_waitcnt lgkmcnt(0)"
        :
        : "v"(p), "v"(v0), "v"(v1), "v"(v2), "v"(v3), "n"(T * kBUF4), "n"(T * kBUF4 + kARR4),
          "n"(T * kBUF4 + 2 * kARR4), "n"(T * kBUF4 + 3 * kARR4)
        : "memory");
}

// the pair read one element low: its second row is this chunk's last
template <int T>
__device__ __forceinline__ void stage_fix(const Stage& st) {
    const lds_cdptr p = (lds_cdptr)((const __attribute__((address_space(3))) unsigned char*)st.lane +
                                    st.fix_off);
    double a, b, c, d;
    asm volatile(
        "ds_read_b64 %0, %4 offset:%5\n\t"
        "ds_read_b64 %1, %4 offset:%6\n\t"
        "ds_read_b64 %2, %4 offset:%7\n\t"
        "ds_read_b64 %3, %4 offset:%8\n\t"
        "s_waitcnt lgkmcnt(0)\n\t"
        "ds_write_b64 %4, %0 offset:%9\n\t"
        "ds_write_b64 %4, %1 offset:%10\n\t"
        "ds_write_b64 %4, %2 offset:%11\n\t"
        "ds_write_b64 %4, %3 offset:%12\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
        : "v"(p), "n"(T * kBUF + 8), "n"(T * kBUF + kARR + 8), "n"(T * kBUF + 2 * kARR + 8),
          "n"(T * kBUF + 3 * kARR + 8), "n"(T * kBUF), "n"(T * kBUF + kARR),
          "n"(T * kBUF + 2 * kARR), "n"(T * kBUF + 3 * kARR)
        : "memory");
}

// Row reads of the staged blocks: one base address per lane and the buffer,
// row and array as immediate offsets, so that no per-row address lives in a
// register; inline asm (waiting inside the same block) keeps the compiler's
// own draining waits out of the staging.

// rows 2K and 2K + 1 of buffer T (FP64): they are adjacent in each array's
// slot, so one 16-byte read per array brings both, and one wait serves two
// rows
typedef double ftd_d2 __attribute__((ext_vector_type(2)));
template <int T, int K>
__device__ __forceinline__ void stage_read_pair(lds_cdptr lane, ftd_d2& a, ftd_d2& b, ftd_d2& c,
                                                ftd_d2& d) {
    constexpr int OFF = T * kBUF + K * 256;
    asm volatile(
        "ds_read_b128 %0, %4 offset:%5\n\t"
        "ds_read_b128 %1, %4 offset:%6\n\t"
        "ds_read_b128 %2, %4 offset:%7\n\t"
        "ds_read_b128 %3, %4 offset:%8\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
        : "v"(lane), "n"(OFF), "n"(OFF + kARR), "n"(OFF + 2 * kARR), "n"(OFF + 3 * kARR)
        : "memory");
}

// rows 4K .. 4K + 3 of buffer T (FP32), widened exactly: one 16-byte read
// per array, one wait for four rows
typedef float ftd_f4 __attribute__((ext_vector_type(4)));
template <int T, int K>
__device__ __forceinline__ void stage_read_quad_f32(lds_cdptr lane, ftd_f4& a, ftd_f4& b,
                                                    ftd_f4& c, ftd_f4& d) {
    constexpr int OFF = T * kBUF4 + K * 512;
    asm volatile(
        "ds_read_b128 %0, %4 offset:%5\n\t"
        "ds_read_b128 %1, %4 offset:%6\n\t"
        "ds_read_b128 %2, %4 offset:%7\n\t"
        "ds_read_b128 %3, %4 offset:%8\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
        : "v"(lane), "n"(OFF), "n"(OFF + kARR4), "n"(OFF + 2 * kARR4), "n"(OFF + 3 * kARR4)
        : "memory");
}

// the 8 rows of block B from buffer B & 1 through step(j, a, b, c, d)
template <typename TS, int B, typename F>
__device__ __forceinline__ void block_rows(lds_cdptr lane, F& step) {
    if constexpr (sizeof(TS) == 8) {
#pragma unroll
        for (int k = 0; k < kRB / 2; ++k) {
            ftd_d2 a, b, c, d;
            switch (k) {
                case 0: stage_read_pair<B & 1, 0>(lane, a, b, c, d); break;
                case 1: stage_read_pair<B & 1, 1>(lane, a, b, c, d); break;
                case 2: stage_read_pair<B & 1, 2>(lane, a, b, c, d); break;
                default: stage_read_pair<B & 1, 3>(lane, a, b, c, d); break;
            }
            step(B * kRB + 2 * k, a.x, b.x, c.x, d.x);
            step(B * kRB + 2 * k + 1, a.y, b.y, c.y, d.y);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kRB / 4; ++k) {
            ftd_f4 a, b, c, d;
            if (k == 0)
                stage_read_quad_f32<B & 1, 0>(lane, a, b, c, d);
            else
                stage_read_quad_f32<B & 1, 1>(lane, a, b, c, d);
            step(B * kRB + 4 * k, (double)a.x, (double)b.x, (double)c.x, (double)d.x);
            step(B * kRB + 4 * k + 1, (double)a.y, (double)b.y, (double)c.y, (double)d.y);
            step(B * kRB + 4 * k + 2, (double)a.z, (double)b.z, (double)c.z, (double)d.z);
            step(B * kRB + 4 * k + 3, (double)a.w, (double)b.w, (double)c.w, (double)d.w);
        }
    }
}

template <typename TS>
__device__ __forceinline__ void vm_wait_block(bool more) {
    if (more)
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(sizeof(TS) == 8 ? kDMA : kDMA4) : "memory");
    else
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

__device__ __forceinline__ void vm_drain() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Forward sweep over the staged blocks, block by block: Block<B>::run calls
// step(j, a, b, c, d) for the 8 rows of block B and recurses, so that every
// row index, buffer and LDS offset is a compile-time constant. Blocks 0 and 1
// have landed before the sweep starts (the previous item waited for them
// before its stores); block B >= 2 was issued once block B - 2 was read.
template <typename TS, int B, int NB>
struct Sweep {
    template <typename F>
    __device__ __forceinline__ static void run(const Stage& st, int l, int nblk, const TS* A,
                                               const TS* Bm, const TS* C, const TS* D, F& step) {
        if (B < nblk) {
            if (B >= 2) vm_wait_block<TS>(B + 1 < nblk);
            if (st.fix_blk == B) {
                if constexpr (sizeof(TS) == 8)
                    stage_fix<B & 1>(st);
                else
                    stage_fix_f32<B & 1>(st);
            }
            block_rows<TS, B>(st.lane, step);
            if (B + 2 < nblk) stage_issue(st, B & 1, B + 2, l, A, Bm, C, D);
            Sweep<TS, B + 1, NB>::run(st, l, nblk, A, Bm, C, D, step);
        }
    }
};
template <typename TS, int NB>
struct Sweep<TS, NB, NB> {
    template <typename F>
    __device__ __forceinline__ static void run(const Stage&, int, int, const TS*, const TS*,
                                               const TS*, const TS*, F&) {}
};

// blocks 0 and 1 of an item; the caller issues them one item ahead
template <typename TS>
__device__ __forceinline__ void stage_lead(const Stage& st, int l, int nblk, const TS* A,
                                          const TS* B, const TS* C, const TS* D) {
    stage_issue(st, 0, 0, l, A, B, C, D);
    if (nblk > 1) stage_issue(st, 1, 1, l, A, B, C, D);
}

// single-chunk systems: plain Thomas, the whole system in one lane
template <int M>
struct ThomasStep {
    double c[M], d[M];
    int cnt;
    __device__ __forceinline__ void operator()(int j, double aj, double bj, double cj, double dj) {
        if (j < cnt) {
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
};

// ahead(): issues the next item's leading blocks once this item's staged
// rows are all read
template <int M, typename TS, typename H>
__device__ __forceinline__ void thomas_item(const Stage& st, const TS* A, const TS* B, const TS* C,
                                            const TS* D, double* X, const BatchLane& L, int l,
                                            int nblk, H& ahead) {
    ThomasStep<M> s;
    s.cnt = L.m;
    asm volatile("" : "+v"(s.cnt));  // keeps the row tests inside the item loop
    const size_t base = (size_t)L.off;
    Sweep<TS, 0, M / kRB>::run(st, l, nblk, A, B, C, D, s);
    ahead();
#pragma unroll
    for (int j = M - 2; j >= 0; --j)
        if (j <= s.cnt - 2) s.d[j] = s.d[j] - s.c[j] * s.d[j + 1];
    vm_drain();  // the next item's leading blocks, before this item's stores
#pragma unroll
    for (int j = 0; j < M; j += 2) {
        if (j + 1 < s.cnt)
            store_pair(X + base + (size_t)j, s.d[j], s.d[j + 1]);
        else if (j < s.cnt)
            FTD_BST(s.d[j], X + base + (size_t)j);
    }
}

// ---------------------------------------------------------------------------
// Reduced systems in LDS (B == 1): rows r of a system's reduced system are
// (rA[r], 1, rC[r] | rD[r]). Statement for statement the oracle's
// thomas_cols / modthomas_cols / update_cols with B == 1 (rows 0 and 1 of a
// group "divided by b = 1" are left as they are, which is exact).
// ---------------------------------------------------------------------------

// para_range over [0, n) in P parts
__device__ __forceinline__ void split_rows(int n, int P, int p, int& s, int& m) {
    const int base = n / P, rem = n % P;
    s = p * base + (p < rem ? p : rem);
    m = base + (p < rem ? 1 : 0);
}

// groups of a reduced system of NR rows: 1 below 16 rows, else the largest
// G with 2 G^2 <= NR (groups of at least 2G >= 4 rows)
__host__ __device__ __forceinline__ int red_groups(int NR) {
    if (NR < 16) return 1;
    int G = 1;
    while (2 * (G + 1) * (G + 1) <= NR) ++G;
    return G;
}

// Thomas on rows r0 .. r0 + NR - 1, x into rD; the rows' values are read
// into registers 8 at a time, so the LDS latency is paid once per 8 rows
// and the serial chain is the arithmetic only
__device__ __forceinline__ void red_thomas(double* rA, double* rC, double* rD, int r0, int NR) {
    double cp = rC[r0], dp = rD[r0];
    int k = 1;
    for (; k + 8 <= NR; k += 8) {
        double ak[8], ck[8], dk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ak[u] = rA[r0 + k + u];
            ck[u] = rC[r0 + k + u];
            dk[u] = rD[r0 + k + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double inv = 1.0 / (1.0 - ak[u] * cp);
            dp = inv * (dk[u] - ak[u] * dp);
            cp = inv * ck[u];
            rD[r0 + k + u] = dp;
            rC[r0 + k + u] = cp;
        }
    }
    for (; k < NR; ++k) {
        const double ak = rA[r0 + k];
        const double inv = 1.0 / (1.0 - ak * cp);
        dp = inv * (rD[r0 + k] - ak * dp);
        cp = inv * rC[r0 + k];
        rD[r0 + k] = dp;
        rC[r0 + k] = cp;
    }
    double dn = dp;
    k = NR - 2;
    for (; k >= 7; k -= 8) {
        double ck[8], dk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ck[u] = rC[r0 + k - u];
            dk[u] = rD[r0 + k - u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            dn = dk[u] - ck[u] * dn;
            rD[r0 + k - u] = dn;
        }
    }
    for (; k >= 0; --k) {
        dn = rD[r0 + k] - rC[r0 + k] * dn;
        rD[r0 + k] = dn;
    }
}

// modified Thomas of one group of m >= 4 rows at b (pascal_tdma.cpp:131-183
// with B == 1): forward j = 2..m-1, backward j = m-3..1, then the row-0
// reduction; the rows' values are read 4 at a time ahead of their use
__device__ __forceinline__ void red_modthomas(double* rA, double* rC, double* rD, int b, int m) {
    double am = rA[b + 1], cm = rC[b + 1], dm = rD[b + 1];
    int j = 2;
    for (; j + 4 <= m; j += 4) {
        double aj[4], cj[4], dj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            aj[u] = rA[b + j + u];
            cj[u] = rC[b + j + u];
            dj[u] = rD[b + j + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double inv = 1.0 / (1.0 - aj[u] * cm);
            dm = inv * (dj[u] - aj[u] * dm);
            cm = inv * cj[u];
            am = -inv * aj[u] * am;
            rD[b + j + u] = dm;
            rC[b + j + u] = cm;
            rA[b + j + u] = am;
        }
    }
    for (; j < m; ++j) {
        const double aj = rA[b + j], cj = rC[b + j], dj = rD[b + j];
        const double inv = 1.0 / (1.0 - aj * cm);
        dm = inv * (dj - aj * dm);
        cm = inv * cj;
        am = -inv * aj * am;
        rD[b + j] = dm;
        rC[b + j] = cm;
        rA[b + j] = am;
    }
    double dn = rD[b + m - 2], an = rA[b + m - 2], cn = rC[b + m - 2];
    j = m - 3;
    for (; j >= 4; j -= 4) {
        double aj[4], cj[4], dj[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            aj[u] = rA[b + j - u];
            cj[u] = rC[b + j - u];
            dj[u] = rD[b + j - u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            dn = dj[u] - cj[u] * dn;
            an = aj[u] - cj[u] * an;
            cn = -cj[u] * cn;
            rD[b + j - u] = dn;
            rA[b + j - u] = an;
            rC[b + j - u] = cn;
        }
    }
    for (; j >= 1; --j) {
        const double cj = rC[b + j];
        dn = rD[b + j] - cj * dn;
        an = rA[b + j] - cj * an;
        cn = -cj * cn;
        rD[b + j] = dn;
        rA[b + j] = an;
        rC[b + j] = cn;
    }
    const double c0 = rC[b], c1 = rC[b + 1], d1 = rD[b + 1], a1 = rA[b + 1];
    const double r = 1.0 / (1.0 - a1 * c0);
    rD[b] = r * (rD[b] - c0 * d1);
    rA[b] = r * rA[b];
    rC[b] = -r * c0 * c1;
}

// Thomas on the 2G rows of the groups' ends (row 0 and row m - 1 of each
// group of the NR rows at r0), x into their rD. The groups are walked in
// order with their starts kept incrementally (no division per row), and both
// end rows of a group are read together, so the LDS latency is paid once per
// group
__device__ __forceinline__ void red_thomas_ends(double* rA, double* rC, double* rD, int r0, int NR,
                                                int G) {
    const int base = NR / G, rem = NR % G;
    // group 0: its row 0 starts the elimination, its row m - 1 follows
    int m = base + (0 < rem ? 1 : 0);
    double cp = rC[r0], dp = rD[r0];
    {
        const int rN = r0 + m - 1;
        const double aN = rA[rN], cN = rC[rN], dN = rD[rN];
        const double inv = 1.0 / (1.0 - aN * cp);
        dp = inv * (dN - aN * dp);
        cp = inv * cN;
        rD[rN] = dp;
        rC[rN] = cp;
    }
    int s = m;
    for (int g = 1; g < G; ++g) {
        m = base + (g < rem ? 1 : 0);
        const int r1 = r0 + s, rN = r0 + s + m - 1;
        const double a1 = rA[r1], c1 = rC[r1], d1 = rD[r1];
        const double aN = rA[rN], cN = rC[rN], dN = rD[rN];
        double inv = 1.0 / (1.0 - a1 * cp);
        dp = inv * (d1 - a1 * dp);
        cp = inv * c1;
        rD[r1] = dp;
        rC[r1] = cp;
        inv = 1.0 / (1.0 - aN * cp);
        dp = inv * (dN - aN * dp);
        cp = inv * cN;
        rD[rN] = dp;
        rC[rN] = cp;
        s += m;
    }
    // backward: x of the last row is dp; then row 0 of the last group, row
    // m - 1 and row 0 of each earlier group
    double dn = dp;
    for (int g = G - 1; g >= 0; --g) {
        m = base + (g < rem ? 1 : 0);
        s = g * base + (g < rem ? g : rem);
        const int r1 = r0 + s, rN = r0 + s + m - 1;
        const double c1 = rC[r1], d1 = rD[r1];
        if (g < G - 1) {
            const double cN = rC[rN], dN = rD[rN];
            dn = dN - cN * dn;
            rD[rN] = dn;
        }
        dn = d1 - c1 * dn;
        rD[r1] = dn;
    }
}

// interior update of one group: x_j = d_j - (a_j x_0 + c_j x_{m-1}), j = 1..m-2
__device__ __forceinline__ void red_update(const double* rA, const double* rC, double* rD, int b,
                                           int m) {
    const double x0 = rD[b], xN = rD[b + m - 1];
    for (int j = 1; j < m - 1; ++j) rD[b + j] = rD[b + j] - (rA[b + j] * x0 + rC[b + j] * xN);
}

// one chunk of a chunked system: modified Thomas (rows 0, 1 divided by b)
template <int M>
struct PascalStep {
    double a[M], c[M], d[M];
    double aN, cN, dN;
    int cnt;
    __device__ __forceinline__ void operator()(int j, double aj, double bj, double cj, double dj) {
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
};

// t: lane within the item (0..127); the system's reduced rows are solved in
// the item's LDS by its chunk-0 lane, then every chunk updates its interior
template <int M, typename TS, typename H>
__device__ __forceinline__ void pascal_item(const Stage& st, const TS* A, const TS* B, const TS* C,
                                            const TS* D, double* X, const BatchLane& L, int l,
                                            int t, int nblk, double* rA, double* rC, double* rD,
                                            H& ahead) {
    PascalStep<M> s;
    s.aN = s.cN = s.dN = 0.0;
    s.cnt = L.m;
    asm volatile("" : "+v"(s.cnt));
    const size_t base = (size_t)L.off;
    const int q = L.q, Q = L.Q, lead = L.lead;
    Sweep<TS, 0, M / kRB>::run(st, l, nblk, A, B, C, D, s);
    ahead();
    const int cnt = s.cnt;
#pragma unroll
    for (int j = M - 3; j >= 1; --j) {
        if (j <= cnt - 3) {
            const double cj = s.c[j];
            s.d[j] = s.d[j] - cj * s.d[j + 1];
            s.a[j] = s.a[j] - cj * s.a[j + 1];
            s.c[j] = -cj * s.c[j + 1];
        }
    }
    const bool active = cnt > 0;
    if (active) {
        const double r = 1.0 / (1.0 - s.a[1] * s.c[0]);
        rA[2 * t] = r * s.a[0];
        rC[2 * t] = -r * s.c[0] * s.c[1];
        rD[2 * t] = r * (s.d[0] - s.c[0] * s.d[1]);
        rA[2 * t + 1] = s.aN;
        rC[2 * t + 1] = s.cN;
        rD[2 * t + 1] = s.dN;
    }
    group_lds_sync();
    // The reduced system of this system, rows r0 .. r0 + NR - 1 of the
    // item's LDS (NR = 2Q, up to 256). Short ones (G == 1): Thomas in the
    // lane of chunk 0. Longer ones: PaScaL over G groups of rows, one lane
    // each (modified Thomas in LDS), the 2G rows of the groups' ends by
    // Thomas in the lane of chunk 0, then each group's interior update --
    // a serial chain of about NR / G + 2G rows instead of NR.
    const int NR = 2 * Q, r0 = 2 * lead;
    const int G = red_groups(NR);
    const bool grouped = active && G > 1 && q < G;
    int gs = 0, gm = 0;
    if (grouped) {
        split_rows(NR, G, q, gs, gm);
        red_modthomas(rA, rC, rD, r0 + gs, gm);
    }
    group_lds_sync();
    if (active && q == 0) {
        if (G == 1)
            red_thomas(rA, rC, rD, r0, NR);
        else
            red_thomas_ends(rA, rC, rD, r0, NR, G);
    }
    group_lds_sync();
    if (grouped) red_update(rA, rC, rD, r0 + gs, gm);
    group_lds_sync();
    vm_drain();  // the next item's leading blocks, before this item's stores
    if (active) {
        const double x0 = rD[2 * t], xN = rD[2 * t + 1];
        auto x_at = [&](int j) {
            return j == 0 ? x0 : j == cnt - 1 ? xN : s.d[j] - (s.a[j] * x0 + s.c[j] * xN);
        };
#pragma unroll
        for (int j = 0; j < M; j += 2) {
            if (j + 1 < cnt)
                store_pair(X + base + (size_t)j, x_at(j), x_at(j + 1));
            else if (j < cnt)
                FTD_BST(x_at(j), X + base + (size_t)j);
        }
    }
    group_lds_sync();  // the reduced rows are read before the next item writes them
}

// Items [0, n_chunked) hold chunked systems, [n_chunked, n_items) single-
// chunk ones; workgroup g takes items g, g + G, g + 2G, ... of each range.
// Each item's descriptors and first two blocks are requested while the
// previous item finishes (backward sweep, reduced solve), and waited for
// (vmcnt(0)) before that item's stores: an item starts on rows already in
// LDS, and the stores' completion, which vmcnt also counts, overlaps its first
// two blocks. The wait for block B >= 2 is then vmcnt(16): the staging DMAs
// complete in issue order, so while any of block B's is pending all 16 of
// block B + 1's are too, whatever the stores issued before them do.
template <int M, typename TS>
__global__ __launch_bounds__(kLanes) __attribute__((amdgpu_waves_per_eu(1))) void k_batched(
    const TS* A, const TS* B, const TS* C, const TS* D, double* X,
    const BatchLane* __restrict__ lanes, int n_chunked, int n_items, int64_t last) {
    __shared__ __attribute__((aligned(16))) unsigned char stage[2][2 * buf_bytes<TS>()];
    __shared__ __attribute__((aligned(16))) BatchLane desc[2][kLanes];
    __shared__ double red[3][2 * kLanes];
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int l = t & 63;
    const int G = gridDim.x;
    const int g = blockIdx.x;
    // the single-chunk items continue the deal where the chunked ones ended
    const int start1 = n_chunked + ((g - n_chunked % G) + G) % G;
    int item = g < n_chunked ? g : start1;
    if (item >= n_items) return;
    Stage st;
    st.buf0 = stage[w];
    st.desc = desc[0] + 64 * w;
    st.lane = (lds_cdptr)(stage[w] + (sizeof(TS) == 8 ? (l >> 4) * kSEG + (l & 15) * 16
                                                      : (l >> 5) * kSEG + (l & 31) * 16));
    st.last = last;
    // each wave writes and reads only its own 64 descriptors
    BatchLane L = lanes[(size_t)item * kLanes + t];
    desc[0][t] = L;
    int nblk = __builtin_amdgcn_readfirstlane((wave_max(L.m) + kRB - 1) / kRB);
    wave_lds_sync();
    stage_lead(st, l, nblk, A, B, C, D);
    vm_drain();
    for (int u = 0;; u ^= 1) {
        st.desc = desc[u] + 64 * w;
        stage_set_fix<TS>(st, L);
        int nxt = item + G;
        if (item < n_chunked && nxt >= n_chunked) nxt = start1;
        const bool more = nxt < n_items;
        // the next item's descriptors go straight to LDS, waited for when
        // this item's staged rows are all read
        if (more)
            __builtin_amdgcn_global_load_lds((const void*)(lanes + (size_t)nxt * kLanes + t),
                                             (lds_vptr)(desc[u ^ 1] + 64 * w), 16, 0, 0);
        BatchLane Ln = L;
        int nblk_n = 0;
        auto ahead = [&]() {
            if (more) {
                vm_drain();
                Ln = desc[u ^ 1][t];
                nblk_n = __builtin_amdgcn_readfirstlane((wave_max(Ln.m) + kRB - 1) / kRB);
                Stage sn = st;
                sn.desc = desc[u ^ 1] + 64 * w;
                stage_lead(sn, l, nblk_n, A, B, C, D);
            }
        };
        if (item < n_chunked)
            pascal_item<M>(st, A, B, C, D, X, L, l, t, nblk, red[0], red[1], red[2], ahead);
        else
            thomas_item<M>(st, A, B, C, D, X, L, l, nblk, ahead);
        if (!more) break;
        item = nxt;
        L = Ln;
        nblk = nblk_n;
    }
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

// Chunk starts (local rows, Q + 1 of them) of a system of N > chunk rows at
// element offset off, cut so that every chunk after the first starts on a
// 16-element boundary of the batch arrays: a 128-byte line of FP64 data, so
// a chunk's staged blocks of 8 rows are whole halves of lines instead of
// segments straddling two (the 20000-system batch laid out line-aligned ran
// in 0.194 ms against 0.232 at arbitrary offsets). The first chunk runs to
// chunk - 16 rows past the first boundary (to chunk rows where the system
// starts on one), the others are chunk rows, and a last chunk under 16 rows
// takes 16 from the one before. Returns Q, or 0 where the system keeps the
// para_range split: a remainder under 16 rows after the first chunk, or
// more than kLanes chunks.
int aligned_starts(int64_t off, int N, int chunk, int* starts) {
    const int a = (int)((16 - off % 16) % 16);
    const int f = a ? a + chunk - 16 : chunk;
    const int rem = N - f;
    if (rem < 16) return 0;
    const int Q = 1 + (rem + chunk - 1) / chunk;
    if (Q > kLanes) return 0;
    starts[0] = 0;
    for (int q = 1; q < Q; ++q) starts[q] = f + (q - 1) * chunk;
    starts[Q] = N;
    if (Q >= 3 && N - starts[Q - 1] < 16) starts[Q - 1] -= 16;
    return Q;
}

}  // namespace

int batch_plan(int n_systems, const int64_t* row_offsets, const int* n_rows, int chunk,
               std::vector<BatchLane>& lanes, std::vector<int64_t>& long_off,
               std::vector<int>& long_n, int& n_chunked, int64_t& last) {
    if (chunk != 32 && chunk != 48) return -1;
    n_chunked = 0;
    lanes.clear();
    long_off.clear();
    long_n.clear();
    last = 0;
    for (int s = 0; s < n_systems; ++s) {
        if (n_rows[s] < 1) return -1;
        last = std::max(last, row_offsets[s] + n_rows[s] - 1);
    }
    std::vector<int> multi, single, nq(n_systems);
    std::vector<int> tmp(kLanes + 1);
    for (int s = 0; s < n_systems; ++s) {
        const int N = n_rows[s];
        const int Q = (N + chunk - 1) / chunk;
        const int Qa = Q >= 2 ? aligned_starts(row_offsets[s], N, chunk, tmp.data()) : 0;
        nq[s] = Qa ? Qa : Q;
        // the staging reads two elements at a time, so a batch spanning one
        // element (a single 1-row system) takes the lane-per-system kernel
        if (Q > kLanes || last == 0) {
            long_off.push_back(row_offsets[s]);
            long_n.push_back(N);
        } else if (Q >= 2) {
            multi.push_back(s);
        } else {
            single.push_back(s);
        }
    }
    auto chunks = [&](int s) { return nq[s]; };
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
            const bool al = aligned_starts(row_offsets[s], N, chunk, tmp.data()) == Q;
            for (int q = 0; q < Q; ++q) {
                int st, m;
                if (al) {
                    st = tmp[q];
                    m = tmp[q + 1] - tmp[q];
                } else {
                    split(N, Q, q, st, m);
                }
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
        while (lo < hi && used + chunks(multi[lo]) <= kLanes) put(multi[lo++]);
        while (lo < hi && used + chunks(multi[hi - 1]) <= kLanes) put(multi[--hi]);
        const int64_t pad = it.l.front().off;
        while ((int)it.l.size() < kLanes) it.l.push_back(idle(pad));
        for (auto& b : it.l) b.kind = 1;
        it.cost = maxm + 6.0 * maxQ;  // streamed rows + serial reduced rows
        items.push_back(std::move(it));
    }
    for (size_t k = 0; k < single.size(); k += kLanes) {
        Item it;
        int maxm = 0;
        for (size_t u = k; u < std::min(single.size(), k + kLanes); ++u) {
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
        while ((int)it.l.size() < kLanes) {
            BatchLane b = idle(pad);
            b.kind = 0;
            it.l.push_back(b);
        }
        it.cost = maxm;
        items.push_back(std::move(it));
    }
    // chunked items first, each range longest first
    const size_t nch = items.size() - (single.size() + kLanes - 1) / kLanes;
    auto by_cost = [](const Item& x, const Item& y) { return x.cost > y.cost; };
    std::stable_sort(items.begin(), items.begin() + nch, by_cost);
    std::stable_sort(items.begin() + nch, items.end(), by_cost);
    n_chunked = (int)nch;
    lanes.reserve(items.size() * kLanes);
    for (auto& it : items) lanes.insert(lanes.end(), it.l.begin(), it.l.end());
    return (int)items.size();
}

template <int M, typename TS>
int batched_resident_blocks() {
    static int cached[64] = {};
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev]) return cached[dev];
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)k_batched<M, TS>, kLanes, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
        cus = 1;
    cached[dev] = per_cu * cus;
    return cached[dev];
}

template <int M, typename TS>
hipError_t launch_items(const TS* A, const TS* B, const TS* C, const TS* D, double* X,
                        const BatchLane* lanes, int n_chunked, int n_items, int64_t last,
                        hipStream_t st) {
    const int blocks = std::min(batched_resident_blocks<M, TS>(), n_items);
    hipLaunchKernelGGL((k_batched<M, TS>), dim3(blocks), dim3(kLanes), 0, st, A, B, C, D, X, lanes,
                       n_chunked, n_items, last);
    return hipGetLastError();
}

hipError_t launch_batched(const double* A, const double* B, double* C, double* D, int chunk,
                          const BatchLane* lanes, int n_chunked, int n_items, int64_t last,
                          const int64_t* long_off, const int* long_n, int n_long,
                          hipStream_t st) {
    if (n_items > 0) {
        const double* Dc = D;
        hipError_t e = chunk == 32
                           ? launch_items<32>(A, B, (const double*)C, Dc, D, lanes, n_chunked,
                                              n_items, last, st)
                           : launch_items<48>(A, B, (const double*)C, Dc, D, lanes, n_chunked,
                                              n_items, last, st);
        if (e != hipSuccess) return e;
    }
    if (n_long > 0) {
        hipLaunchKernelGGL(k_batched_long, dim3((n_long + 255) / 256), dim3(256), 0, st, A, B, C, D,
                           long_off, long_n, n_long);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_batched_f32(const float* A, const float* B, const float* C, const float* D,
                              double* X, int chunk, const BatchLane* lanes, int n_chunked,
                              int n_items, int64_t last, hipStream_t st) {
    if (n_items < 1) return hipSuccess;
    return chunk == 32 ? launch_items<32>(A, B, C, D, X, lanes, n_chunked, n_items, last, st)
                       : launch_items<48>(A, B, C, D, X, lanes, n_chunked, n_items, last, st);
}

int batched_default_chunk() { return kChunk; }

int batched_max_rows(int chunk) { return kLanes * chunk; }

}  // namespace ftd
