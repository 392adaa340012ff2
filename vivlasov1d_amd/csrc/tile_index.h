// Index arithmetic of the tiled kernels, shared by the device code and a
// host-side checker (tests/native/tile_index_check.cpp) that enumerates every
// tile, halo load and output of the bracket kernel and of the fused J u +
// rFFT kernel for whole ranges of grid sizes and asserts that each index
// stays inside its buffer.  A halo load that
// reads past a buffer feeds no stored output, so no parity test can see it;
// this checker can.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vv {

// XCD-aware block order: spread consecutive tiles so that each of the 8 XCDs
// works on one contiguous range of tiles (halo lines stay in that XCD's L2).
// A permutation of [0, nb) for every nb.
__host__ __device__ __forceinline__ int xcd_tile(int b, int nb)
{
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// Geometry of the bracket kernel: one workgroup computes a TX (x) x TY (v)
// tile; the tile of the first argument and of H with a 2-cell halo,
// HX x HY cells, is staged in LDS.
constexpr int kBrTX = 64;
constexpr int kBrTY = 16;
constexpr int kBrHX = kBrTX + 4;
constexpr int kBrHY = kBrTY + 4;

// Periodic x index of a tile column: the tile with its halo spans columns
// i0-2 .. i0+TX+1, which on grids narrower than TX+2 (nx < 66) reaches past
// nx + nx - 1; the extra columns feed no stored output but must still load
// from inside the row.  Returns a value in [0, nx) for gi in [-2, 2 nx + TX).
__host__ __device__ __forceinline__ int wrap_x(int gi, int nx)
{
    if (gi < 0) gi += nx;
    if (gi >= nx) gi -= nx;
    if (gi >= nx) gi %= nx;
    return gi;
}

struct BrTile {
    int i0, j0;       // first column and row of the tile's outputs
};

__host__ __device__ __forceinline__ int br_tiles_x(int nx) { return (nx + kBrTX - 1) / kBrTX; }
__host__ __device__ __forceinline__ int br_tiles(int nx, int nv)
{
    return br_tiles_x(nx) * ((nv + kBrTY - 1) / kBrTY);
}

// Tile of workgroup `block` (0 <= block < br_tiles(nx, nv)).
__host__ __device__ __forceinline__ BrTile br_tile(int block, int nx, int nv)
{
    const int ntx = br_tiles_x(nx);
    const int t = xcd_tile(block, br_tiles(nx, nv));
    BrTile r;
    r.i0 = (t % ntx) * kBrTX;
    r.j0 = (t / ntx) * kBrTY;
    return r;
}

// Grid row of staged row r, or -1 outside the grid (staged as zero).
__host__ __device__ __forceinline__ int br_halo_row(const BrTile &t, int r, int nv)
{
    const int gj = t.j0 - 2 + r;
    return (gj >= 0 && gj < nv) ? gj : -1;
}

// Staged cell (r, c), 0 <= r < HY, 0 <= c < HX, of a tile: the field offset
// gj * nx + gi it loads, or -1 for rows outside the grid (staged as zero).
__host__ __device__ __forceinline__ int64_t br_halo_offset(const BrTile &t, int r, int c, int nx, int nv)
{
    const int gj = br_halo_row(t, r, nv);
    if (gj < 0) return -1;
    return (int64_t)gj * nx + wrap_x(t.i0 - 2 + c, nx);
}

// x index of staged column c (the 1D x-factors of H), always in [0, nx).
__host__ __device__ __forceinline__ int br_halo_column(const BrTile &t, int c, int nx)
{
    return wrap_x(t.i0 - 2 + c, nx);
}

// Geometry of the fused J u + rFFT kernel: a workgroup's R output rows
// j0 .. j0+R-1 (j0 = xcd_tile(block) R) are computed in strips of kJrSW
// columns starting at cb; the strip with a 2-cell halo, (R+4) x (kJrSW+4)
// cells, is staged; thread t owns staged column t + 2 (grid column cb + t),
// threads t < 4 (R+4) also load one halo cell.
constexpr int kJrSW = 256;

// Grid row of staged row r (0 <= r < R+4), or -1 outside the grid (zero).
__host__ __device__ __forceinline__ int jr_row(int j0, int r, int nv)
{
    const int gj = j0 - 2 + r;
    return (gj >= 0 && gj < nv) ? gj : -1;
}

// Staged cell (row, column) that halo thread t (< 4 RH, RH = R+4) loads:
// columns 0, 1, kJrSW+2, kJrSW+3, one row each.
__host__ __device__ __forceinline__ void jr_halo_cell(int t, int RH, int *row, int *col)
{
    const int hc = t / RH;
    *row = t - hc * RH;
    *col = hc < 2 ? hc : kJrSW + hc;
}

// Grid column of staged column c of the strip starting at cb, in [0, nx).
__host__ __device__ __forceinline__ int jr_col(int cb, int c, int nx) { return wrap_x(cb - 2 + c, nx); }

}  // namespace vv
