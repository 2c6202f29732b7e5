// ftdma_xcd.h -- tile order for stencil kernels on a part with several L2s.
#pragma once

#include <hip/hip_runtime.h>

namespace ftd {

// MI355X has 8 accelerator dies (XCDs), each with its own L2, and hands the
// workgroups of a launch to them round-robin (workgroup b to XCD b % 8).
constexpr int kXcds = 8;

// Tile of workgroup b among nt: each XCD takes a contiguous run of tiles
// (the runs differ in length by at most one), so tiles adjacent in the
// caller's order run one after the other on the same XCD and find the rows
// they share in its L2. A bijection of [0, nt).
__device__ __forceinline__ int xcd_tile(int b, int nt) {
    const int per = nt / kXcds, rem = nt % kXcds, x = b % kXcds;
    return x * per + (x < rem ? x : rem) + b / kXcds;
}

}  // namespace ftd
