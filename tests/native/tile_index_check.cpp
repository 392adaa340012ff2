// Host-side enumeration of the index arithmetic of the bracket kernel and of
// the fused J u + rFFT kernel (vivlasov1d_amd/csrc/tile_index.h, the
// functions the kernels themselves call): for every grid in the given ranges,
// every workgroup's tile or row block, every staged halo cell, every x-factor
// and v-factor load and every stencil read of every output point is checked
// against the buffer bounds and against the periodic-in-x / bounded-in-v
// neighbour it must be.
// Usage: tile_index_check NX_LO NX_HI NV_LO NV_HI [NX NV ...]   (bracket)
//        tile_index_check jr NV_LO NV_HI [NV ...]             (fused J u + rFFT,
//        nx = 256, 512, 1024, rows per workgroup 2, 4, 8)
// Prints one line per failure (at most 20) and "checked G grids" at the end;
// exit status 1 on any failure.
#include <cstdio>
#include <cstdlib>
#include <string>
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

// The fused kernel: R output rows per workgroup, strips of kJrSW columns.
static void check_jr(int nx, int nv, int R)
{
    const long N = (long)nx * nv;
    const int RH = R + 4, SP = kJrSW + 4, nb = (nv + R - 1) / R, ns = nx / kJrSW;
    const int logn = nx == 256 ? 8 : nx == 512 ? 9 : 10;
    const long fft_elems = (long)(R / 2) * nx;                 // complex elements of the FFT buffer
    std::vector<int> seen(nb, 0);
    std::vector<int> covered(N, 0);
    for (int block = 0; block < nb; ++block) {
        const int tb = xcd_tile(block, nb);
        if (tb < 0 || tb >= nb) { fail(nx, nv, "jr: row block out of range", block, tb); continue; }
        seen[tb]++;
        const int j0 = tb * R;
        for (int s = 0; s < ns; ++s) {
            const int cb = s * kJrSW;
            // staged cell -> (grid row or -1, grid column); every cell written once
            std::vector<long> cell((long)RH * SP, -3), gcol((long)RH * SP, -3);
            std::vector<int> writes((long)RH * SP, 0);
            for (int t = 0; t < kJrSW; ++t) {
                const int i = cb + t;
                if (i < 0 || i >= nx) fail(nx, nv, "jr: own column", t, i);
                for (int r = 0; r < RH; ++r) {
                    const int gj = jr_row(j0, r, nv);
                    if (gj >= 0 && (long)gj * nx + i >= N) fail(nx, nv, "jr: own load past the field", gj, i);
                    cell[(long)r * SP + t + 2] = gj;
                    gcol[(long)r * SP + t + 2] = i;
                    writes[(long)r * SP + t + 2]++;
                }
                if (t < 4 * RH) {
                    int hr, tc;
                    jr_halo_cell(t, RH, &hr, &tc);
                    if (hr < 0 || hr >= RH || !(tc == 0 || tc == 1 || tc == kJrSW + 2 || tc == kJrSW + 3))
                        fail(nx, nv, "jr: halo cell", hr, tc);
                    const int gi = jr_col(cb, tc, nx);
                    if (gi < 0 || gi >= nx) fail(nx, nv, "jr: halo column", tc, gi);
                    const int gj = jr_row(j0, hr, nv);
                    if (gj >= 0 && (long)gj * nx + gi >= N) fail(nx, nv, "jr: halo load past the field", gj, gi);
                    cell[(long)hr * SP + tc] = gj;
                    gcol[(long)hr * SP + tc] = gi;
                    writes[(long)hr * SP + tc]++;
                }
            }
            for (long e = 0; e < (long)RH * SP; ++e)
                if (writes[e] != 1) fail(nx, nv, "jr: staged cell not written exactly once", e, writes[e]);
            // the walk: thread t, output rows j0 + q, window cells (q + 2 + dj, t + 2 + di)
            for (int t = 0; t < kJrSW; ++t)
                for (int q = 0; q < R; ++q) {
                    const int j = j0 + q, i = cb + t;
                    if (j < nv) covered[(long)j * nx + i]++;
                    for (int dj = -2; dj <= 2; ++dj)
                        for (int di = -2; di <= 2; ++di) {
                            const int r = q + 2 + dj, c = t + 2 + di;
                            if (r < 0 || r >= RH || c < 0 || c >= SP) {
                                fail(nx, nv, "jr: window read outside the strip", r, c);
                                continue;
                            }
                            if (j >= nv) continue;          // its result is the rFFT's zero row
                            const int jj = j + dj;
                            const long want_row = (jj >= 0 && jj < nv) ? jj : -1;
                            const long want_col = ((i + di) % nx + nx) % nx;
                            if (cell[(long)r * SP + c] != want_row ||
                                (want_row >= 0 && gcol[(long)r * SP + c] != want_col))
                                fail(nx, nv, "jr: window cell is not the neighbour", cell[(long)r * SP + c], want_row);
                        }
                    // where the result goes: the FFT buffer (complex rows p = q / 2)
                    if (ns == 4) {
                        for (int rr = 0; rr < 4; ++rr) {
                            const long o = ((long)(q / 2) << logn) + 4L * t + rr;
                            if (o < 0 || o >= fft_elems) fail(nx, nv, "jr: stage-0 store outside the buffer", o, fft_elems);
                        }
                    } else {
                        const long o = (long)(q / 2) * nx + s * kJrSW + t;
                        if (o < 0 || o >= fft_elems) fail(nx, nv, "jr: row store outside the buffer", o, fft_elems);
                    }
                }
        }
    }
    for (int b = 0; b < nb; ++b)
        if (seen[b] != 1) fail(nx, nv, "jr: row block not visited exactly once", b, seen[b]);
    for (long e = 0; e < N; ++e)
        if (covered[e] != 1) fail(nx, nv, "jr: output not formed exactly once", e, covered[e]);
}

int main(int argc, char **argv)
{
    if (argc >= 4 && std::string(argv[1]) == "jr") {
        long grids = 0;
        std::vector<int> nvs;
        for (int nv = std::atoi(argv[2]); nv <= std::atoi(argv[3]); ++nv) nvs.push_back(nv);
        for (int k = 4; k < argc; ++k) nvs.push_back(std::atoi(argv[k]));
        for (int nx : {256, 512, 1024})
            for (int R : {2, 4, 8})
                for (int nv : nvs) {
                    check_jr(nx, nv, R);
                    ++grids;
                }
        std::printf("checked %ld grids, %ld failures\n", grids, failures);
        return failures ? 1 : 0;
    }
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
