// tile_map.h -- XCD-aware workgroup-to-tile mapping (host and device).
#pragma once

#ifndef ZK_HD
#ifdef __HIPCC__
#define ZK_HD __host__ __device__ __forceinline__
#else
#define ZK_HD inline
#endif
#endif

namespace zk {

// Block index -> tile index so that consecutive tiles land on the same XCD.
// The hardware deals workgroups round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  XCD x receives the blocks b = x, x+8, ...: c_x = nb/8 + (x <
// nb%8) of them; they are given the contiguous tile range starting at
// x*(nb/8) + min(x, nb%8), so neighbouring tiles share an L2.  A bijection of
// [0, nb) for every nb.
ZK_HD int xcd_tile(int b, int nb)
{
    constexpr int NXCD = 8;
    const int q = nb / NXCD, r = nb % NXCD;
    const int x = b % NXCD, k = b / NXCD;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace zk
