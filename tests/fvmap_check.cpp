// fvmap_check.cpp -- host check of the closed-form finite-volume ghost map
// (zuikakucfd_amd/csrc/fv_map.h) against the oracle's fill, which replays the
// reference's fill loops (src_WENO/fluid.cpp:338-370) with last-writer-wins.
//
// For random masks (transmissive or periodic ring, interior obstacles of
// codes -1 and 1) every cell carries its own index as a marker; after the
// oracle's fill each cell holds the index of the cell it was filled from.
// fv_src_entry must name that cell, flag exactly the updated cells, and
// ring_slot must be a bijection of the cells outside [3,ny-3)x[3,nx-3) onto
// [0, ring_count).  Prints one line and exits 0 when every case passes.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "fv_map.h"
#include "zk_oracle.h"

static int check(int ny, int nx, int ring_code, double p_obst, std::mt19937 &rng)
{
    const size_t np = (size_t)ny * nx;
    std::vector<int32_t> b(np, 0);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const bool ring = i < 3 || i >= ny - 3 || j < 3 || j >= nx - 3;
            int32_t code = ring ? ring_code : 0;
            // periodic images lie in rows/columns [3,6) and [n-6,n-3): keep those
            // updated so the oracle's copy and the map name the same cell
            const bool near = ring_code == -2 && (i < 6 || i >= ny - 6 || j < 6 || j >= nx - 6);
            if (!ring && !near && U(rng) < p_obst) code = U(rng) < 0.8 ? -1 : 1;
            if (ring && ring_code == -1 && U(rng) < 0.05) code = 0;  // fixed cells in the ring (jet inflow)
            b[(size_t)i * nx + j] = code;
        }
    std::vector<double> Q(4 * np, 0.0);
    for (size_t k = 0; k < np; ++k) Q[k] = (double)k;
    zko_fv_fill_ghosts(ny, nx, b.data(), Q.data());
    std::vector<int> seen(zk::ring_count(ny, nx), 0);
    int bad = 0;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const size_t k = (size_t)i * nx + j;
            const int32_t e = zk::fv_src_entry(b.data(), ny, nx, i, j);
            const long long s = (long long)(e >> 3);
            const bool in = i >= 3 && i < ny - 3 && j >= 3 && j < nx - 3;
            const bool upd = in && b[k] == 0;
            if (s < 0 || s >= (long long)np || (double)s != Q[k] || ((e & 4) != 0) != upd || (e & 3) != 0) {
                if (bad < 5)
                    std::fprintf(stderr, "mismatch %dx%d ring %d at (%d,%d): map %lld, oracle %.0f, flags %d\n", ny,
                                 nx, ring_code, i, j, s, Q[k], e & 7);
                ++bad;
            }
            const long long t = zk::ring_slot(ny, nx, i, j);
            if (in != (t < 0) || t >= (long long)seen.size()) {
                ++bad;
                continue;
            }
            if (t >= 0) ++seen[t];
        }
    for (int c : seen)
        if (c != 1) ++bad;
    return bad;
}

int main()
{
    std::mt19937 rng(12345);
    int cases = 0, bad = 0;
    for (int ny = 7; ny <= 31; ny += 3)
        for (int nx = 7; nx <= 34; nx += 3)
            for (int ring : {-1, -2})
                for (double p : {0.0, 0.1, 0.35}) {
                    // a periodic ring's images lie inside [3, n-3) only from n = 9 on;
                    // below that the oracle copies ghost to ghost
                    if (ring == -2 && (ny < 9 || nx < 9)) continue;
                    bad += check(ny, nx, ring, p, rng);
                    ++cases;
                }
    for (int r = 0; r < 20; ++r) {
        bad += check(64 + r, 48 + 2 * r, -1, 0.2, rng);
        ++cases;
    }
    std::printf("fvmap_check: %d cases, %d mismatches\n", cases, bad);
    return bad ? 1 : 0;
}
