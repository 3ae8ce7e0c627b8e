// fv_map.h -- the finite-volume ghost-fill map and ghost-ring order in closed
// form, one cell at a time (host and device).
//
// zk_load builds the map by replaying the reference's fill loops on indices
// (build_src_map in zk_api.hip).  These functions give the same entries for
// a cell on their own, so that a kernel can build the map with one thread per
// cell and no sequential pass.
#pragma once

#include <cstdint>

#include "tile_map.h"

namespace zk {

// Map entry of cell (i, j) of the finite-volume path without slip walls, in
// the encoding source << 3 | flags (bit 2: an updated cell, bctype 0 inside
// [3,ny-3)x[3,nx-3)).  The reference fill (src_WENO/fluid.cpp:338-370) visits
// the updated cells row by row and copies each into the bctype -1 cells up to
// 3 away, first along its column (the row pass), then along its row (the
// column pass), so the last writer of a -1 cell is:
//   * the updated cell of its own row in range [3, nx-4] with the largest
//     column within 3 of it, when the cell's row lies in [3, ny-4];
//   * otherwise the updated cell of its own column with the largest row
//     within 3 of it, when the cell's column lies in [3, nx-4];
//   * otherwise none (the cell keeps its own value).
// A periodic ghost (-2, an extension) takes its periodic image, which lies
// inside [3, n-3) for ny, nx >= 9; every other code is never filled.
ZK_HD int32_t fv_src_entry(const int32_t *bct, int ny, int nx, int i, int j)
{
    const int64_t k = (int64_t)i * nx + j;
    const int32_t bt = bct[k];
    int64_t s = k;
    if (bt == -1) {
        bool found = false;
        if (i >= 3 && i < ny - 3)
            for (int jj = j + 3; jj >= j - 3 && !found; --jj)
                if (jj != j && jj >= 3 && jj < nx - 3 && bct[(int64_t)i * nx + jj] == 0) {
                    s = (int64_t)i * nx + jj;
                    found = true;
                }
        if (!found && j >= 3 && j < nx - 3)
            for (int ii = i + 3; ii >= i - 3 && !found; --ii)
                if (ii != i && ii >= 3 && ii < ny - 3 && bct[(int64_t)ii * nx + j] == 0) {
                    s = (int64_t)ii * nx + j;
                    found = true;
                }
    } else if (bt == -2) {
        const int pi = i < 3 ? i + (ny - 6) : (i >= ny - 3 ? i - (ny - 6) : i);
        const int pj = j < 3 ? j + (nx - 6) : (j >= nx - 3 ? j - (nx - 6) : j);
        s = (int64_t)pi * nx + pj;
    }
    int32_t e = (int32_t)(s << 3);
    if (bt == 0 && i >= 3 && i < ny - 3 && j >= 3 && j < nx - 3) e |= 4;
    return e;
}

// Number of cells outside [3,ny-3)x[3,nx-3) (ny, nx >= 7).
ZK_HD int64_t ring_count(int ny, int nx) { return 6 * (int64_t)nx + 6 * (int64_t)(ny - 6); }

// Position of cell (i, j) outside [3,ny-3)x[3,nx-3) in the row-major list of
// those cells (the order zk_load lists them in): a bijection onto
// [0, ring_count); -1 for a cell inside.
ZK_HD int64_t ring_slot(int ny, int nx, int i, int j)
{
    if (i < 3) return (int64_t)i * nx + j;
    if (i >= ny - 3) return 3 * (int64_t)nx + 6 * (int64_t)(ny - 6) + (int64_t)(i - (ny - 3)) * nx + j;
    if (j < 3) return 3 * (int64_t)nx + 6 * (int64_t)(i - 3) + j;
    if (j >= nx - 3) return 3 * (int64_t)nx + 6 * (int64_t)(i - 3) + (j - (nx - 6));
    return -1;
}

}  // namespace zk
