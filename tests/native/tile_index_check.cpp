// Host-side enumeration of the bracket kernel's index arithmetic
// (vivlasov1d_amd/csrc/tile_index.h, the functions the kernel itself calls):
// for every grid nx x nv in the given ranges, every workgroup's tile, every
// staged halo cell, every x-factor and v-factor load and every stencil read
// of every output point is checked against the buffer bounds and against the
// periodic-in-x / bounded-in-v neighbour it must be.
// Usage: tile_index_check NX_LO NX_HI NV_LO NV_HI [NX NV ...]
// Prints one line per failure (at most 20) and "checked G grids" at the end;
// exit status 1 on any failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../vivlasov1d_amd/csrc/tile_index.h"

using namespace vv;

static long failures = 0;

static void fail(int nx, int nv, const char *what, long a, long b)
{
    if (failures++ < 20) std::printf("FAIL nx=%d nv=%d %s (%ld, %ld)\n", nx, nv, what, a, b);
}

static void check_grid(int nx, int nv)
{
    const long N = (long)nx * nv;
    const int nb = br_tiles(nx, nv);
    std::vector<int> seen_tile(nb, 0);
    std::vector<int> covered(N, 0);
    for (int block = 0; block < nb; ++block) {
        const int t = xcd_tile(block, nb);
        if (t < 0 || t >= nb) { fail(nx, nv, "xcd_tile out of range", block, t); continue; }
        seen_tile[t]++;
        const BrTile tile = br_tile(block, nx, nv);
        if (tile.i0 < 0 || tile.i0 >= nx || tile.j0 < 0 || tile.j0 >= nv) fail(nx, nv, "tile origin", tile.i0, tile.j0);
        // x-factor loads (threads 0..HX-1) and v-factor loads (HX..HX+HY-1)
        for (int c = 0; c < kBrHX; ++c) {
            const int gi = br_halo_column(tile, c, nx);
            const int want = ((tile.i0 - 2 + c) % nx + nx) % nx;
            if (gi < 0 || gi >= nx) fail(nx, nv, "x-factor column", c, gi);
            else if (gi != want) fail(nx, nv, "x-factor column is not the periodic neighbour", gi, want);
        }
        for (int r = 0; r < kBrHY; ++r) {
            const int gj = br_halo_row(tile, r, nv);
            const int want = tile.j0 - 2 + r;
            if (gj >= nv || gj < -1) fail(nx, nv, "v-factor row", r, gj);
            if (want >= 0 && want < nv ? gj != want : gj != -1) fail(nx, nv, "v-factor row mapping", gj, want);
        }
        // staged cells
        for (int r = 0; r < kBrHY; ++r)
            for (int c = 0; c < kBrHX; ++c) {
                const long o = (long)br_halo_offset(tile, r, c, nx, nv);
                if (o >= N) fail(nx, nv, "halo offset past the field", o, N);
                if (o < -1) fail(nx, nv, "halo offset negative", o, r);
            }
        // outputs: lane l of wave w computes column i0 + l, rows j0 + w*TY/4 + q;
        // its stencil reads staged cells (rr + 2 + dj, l + 2 + di), |di|, |dj| <= 2
        for (int lane = 0; lane < kBrTX; ++lane) {
            const int i = tile.i0 + lane;
            if (i >= nx) continue;
            for (int rr = 0; rr < kBrTY; ++rr) {
                const int j = tile.j0 + rr;
                if (j >= nv) continue;
                covered[(long)j * nx + i]++;
                for (int dj = -2; dj <= 2; ++dj)
                    for (int di = -2; di <= 2; ++di) {
                        const int r = rr + 2 + dj, c = lane + 2 + di;
                        if (r < 0 || r >= kBrHY || c < 0 || c >= kBrHX) {
                            fail(nx, nv, "stencil read outside the staged tile", r, c);
                            continue;
                        }
                        const long o = (long)br_halo_offset(tile, r, c, nx, nv);
                        const int jj = j + dj;
                        const long want = (jj >= 0 && jj < nv) ? (long)jj * nx + ((i + di) % nx + nx) % nx : -1;
                        if (o != want) fail(nx, nv, "stencil cell is not the neighbour", o, want);
                    }
            }
        }
    }
    for (int t = 0; t < nb; ++t)
        if (seen_tile[t] != 1) fail(nx, nv, "tile not visited exactly once", t, seen_tile[t]);
    for (long e = 0; e < N; ++e)
        if (covered[e] != 1) fail(nx, nv, "output not written exactly once", e, covered[e]);
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::printf("usage: %s NX_LO NX_HI NV_LO NV_HI [NX NV ...]\n", argv[0]);
        return 2;
    }
    long grids = 0;
    const int nx0 = std::atoi(argv[1]), nx1 = std::atoi(argv[2]), nv0 = std::atoi(argv[3]), nv1 = std::atoi(argv[4]);
    for (int nx = nx0; nx <= nx1; ++nx)
        for (int nv = nv0; nv <= nv1; ++nv, ++grids) check_grid(nx, nv);
    for (int k = 5; k + 1 < argc; k += 2, ++grids) check_grid(std::atoi(argv[k]), std::atoi(argv[k + 1]));
    std::printf("checked %ld grids, %ld failures\n", grids, failures);
    return failures ? 1 : 0;
}
