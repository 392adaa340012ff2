// Arakawa J1 / J2 / J4 Poisson-bracket kernels: the implicit-midpoint
// residual F(f), the Jacobian-vector product J u and the plain accumulation
// y += factor {x, H} (or y = 0 + factor {x, H}; the double-bracket
// dissipation terms are chains of these).
//
// This translation unit is compiled with -ffp-contract=off: every stencil
// expression below keeps the reference's IEEE operation order
// (PoissonBracket.pyx:100-113, 221-234, 360-389; C left-to-right), so the
// result is bit-identical to the reference's unfused double arithmetic.
//
// Layout: natural order, element (i, j) at i + nx*j (x contiguous).
// Tiling: one workgroup computes a 64 (x) x TY (v) tile (TY = 16).  The tile
// of the bracket's first argument and of H_ave, both with a 2-cell halo, is
// staged in LDS; every global load of a thread is issued before the first
// LDS store.  H_ave is composed on the fly from its 1D factors
// ((((h0_j + .5 h1p_i) + .5 h1h_i) + .5 h2p_i) + .5 h2h_i), exactly the
// value the reference's 2D Have vector holds (PETScVlasovSolver.pyx:145-149),
// so it costs no HBM traffic.  Each wavefront owns TY/4 consecutive rows;
// lane l computes column i0 + l (global loads and stores are 512-byte rows)
// and walks down its rows with the stencil values in a sliding register
// window.

#include "arakawa.h"
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int TX = kBrTX;
constexpr int TY = kBrTY;
constexpr int HX = kBrHX;
constexpr int HY = kBrHY;
constexpr int NT = 256;
constexpr int ROWS_PER_WAVE = TY / (NT / 64);

// MODE_TRANSPOSED: y = alpha u + factor {u~, H} on every row, u~ = u with the
// rows outside 2..nv-3 zeroed.  The Arakawa stencils are skew-symmetric in
// their first argument ({f, H} summed against g equals -{g, H} summed
// against f for fields that vanish outside the grid), so for the masked
// operator M u = R {u, H} (R: keep rows 2..nv-3) the transpose is
// M^T g = -{R g, H} evaluated on all rows: this mode with factor = -1.
enum { MODE_RESIDUAL = 0, MODE_JACOBIAN = 1, MODE_ACCUMULATE = 2, MODE_TRANSPOSED = 3 };

// T is the field type: double (bit-exact FP64 path) or float (Jacobian-apply
// of the FP32 GMRES cycle of the mixed-precision solver; H_ave is composed in
// FP64 and rounded once, the stencil runs in float).
// Instance b of a batched context is blockIdx.y: its fields start at b * N,
// its x-factors of H at b * nx, its GMRES state b * gstride bytes after
// instance 0's (uscale, done), its hx-dependent factors in inst[b].
template <typename T>
struct BrArgs {
    int nx, nv;
    int64_t N;
    const T *f;            // residual: f; jacobian: u; accumulate: x
    const T *fh;           // residual: f at the previous time level
    const double *uscale;  // optional device scale of the first argument (unnormalised Krylov vector)
    T *y;
    const double *h0, *h1p, *h1h, *h2p, *h2h;
    const double *h2d;     // optional second argument as a 2D field (instead of H_ave)
    const double *hxv;     // optional second argument that depends on x only
    const InstConst *inst;
    double ht_inv, factor, alpha;
    const int32_t *done;   // optional early-exit flag (GMRES cycle finished)
    int64_t gstride;
    int store;             // accumulate mode: y = 0 + factor {x, H} (the reference's
                           // Vec.set(0.) followed by PoissonBracket.function)
};

template <int KIND, int MODE, typename T>
__global__ __launch_bounds__(NT) void bracket_kernel(BrArgs<T> a)
{
    const int b = blockIdx.y;
    __shared__ T sF[HY][HX];
    __shared__ T sH[HY][HX];
    __shared__ double sX[4][HX];      // h1p, h1h, h2p, h2h (or the x-only H) of the tile's columns
    __shared__ double sV[HY];         // h0 of the tile's rows

    const int nx = a.nx, nv = a.nv;
    const BrTile tile = br_tile(blockIdx.x, nx, nv);
    const int i0 = tile.i0, j0 = tile.j0;
    const size_t fo = (size_t)b * a.N, xo = (size_t)b * nx;
    const T *__restrict__ fin = a.f + fo;
    const T *__restrict__ fhin = (MODE == MODE_RESIDUAL) ? a.fh + fo : nullptr;
    T *__restrict__ yout = a.y + fo;
    const T half = T(0.5);
    const bool factors = (a.h2d == nullptr && a.hxv == nullptr);

    // Phase 1: every global load of the block is issued before anything waits
    // on one (the early-exit flag, the scale, the 1D factors and the tile), so
    // a block pays one memory latency, not one per dependent step.
    const int32_t dflag =
        (a.done != nullptr) ? *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(a.done) + b * a.gstride)
                            : 0;
    // batched: a finished instance skips its loads (one instance keeps them
    // in flight with the flag's)
    if (gridDim.y > 1 && dflag) return;
    const double usd = (a.uscale != nullptr)
                           ? *reinterpret_cast<const double *>(reinterpret_cast<const char *>(a.uscale) + b * a.gstride)
                           : 1.0;
    double xf0 = 0.0, xf1 = 0.0, xf2 = 0.0, xf3 = 0.0, vf = 0.0;
    if (threadIdx.x < HX) {
        const int gi = br_halo_column(tile, (int)threadIdx.x, nx);
        if (a.hxv != nullptr) {
            xf0 = a.hxv[xo + gi];
        } else if (factors) {
            xf0 = a.h1p[xo + gi];
            xf1 = a.h1h[xo + gi];
            xf2 = a.h2p[xo + gi];
            xf3 = a.h2h[xo + gi];
        }
    } else if (threadIdx.x < HX + HY) {
        const int gj = br_halo_row(tile, (int)threadIdx.x - HX, nv);
        if (factors && gj >= 0) vf = a.h0[gj];
    }
    constexpr int NQ = (HY * HX + NT - 1) / NT;
    T fa[NQ], fb[NQ];
    double hq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int idx = threadIdx.x + q * NT;
        fa[q] = T(0);
        fb[q] = T(0);
        hq[q] = 0.0;
        if (idx < HY * HX) {
            const int r = idx / HX;
            const int c = idx - r * HX;
            const int64_t o = br_halo_offset(tile, r, c, nx, nv);
            if (o >= 0) {
                fa[q] = fin[o];
                if (MODE == MODE_RESIDUAL) fb[q] = fhin[o];
                if (a.h2d != nullptr) hq[q] = a.h2d[fo + o];
            }
        }
    }

    // Phase 2: stage the 1D factors, then the tile of the first argument and of H
    if (dflag) return;                 // uniform over the block, before any barrier
    if (threadIdx.x < HX) {
        sX[0][threadIdx.x] = xf0;
        sX[1][threadIdx.x] = xf1;
        sX[2][threadIdx.x] = xf2;
        sX[3][threadIdx.x] = xf3;
    } else if (threadIdx.x < HX + HY) {
        sV[threadIdx.x - HX] = vf;
    }
    __syncthreads();

    const T us = (T)usd;
    // FP32: H_ave enters every Arakawa term only through differences, so the
    // float tile holds H_ave - h0(j0 + TY/2) (h0 at the centre row of the
    // tile's 16-row band, 0 past the grid), formed in FP64 and rounded once;
    // the rounding error is then relative to H's variation over the band
    // (h0's over its rows plus the x-dependent part), not to |H| (which would
    // cost ~|H|/(hv dH/dv) ulps on fine v grids).  The offset depends on the
    // band only, so a kernel that stages other tiles of the same rows (the
    // fused J u + rFFT) forms the same values.
    const double href = (sizeof(T) == 4 && factors) ? sV[HY / 2] : 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < HY * HX) {
            const int r = idx / HX;
            const int c = idx - r * HX;
            const int gj = br_halo_row(tile, r, nv);
            const bool in = gj >= 0;
            T fv;
            if (MODE == MODE_RESIDUAL) {
                fv = half * fb[q] + half * fa[q];            // Fave (PETScNLVlasovMP.pyx:44-46)
            } else if (MODE == MODE_TRANSPOSED) {
                fv = (gj >= 2 && gj < nv - 2) ? fa[q] * us : T(0);
            } else if (MODE == MODE_JACOBIAN) {
                fv = fa[q] * us;                             // == VecScale'd Krylov vector
            } else {
                fv = (a.uscale != nullptr) ? fa[q] * us : fa[q];
            }
            sF[r][c] = fv;
            // ((((h0_j + .5 h1p_i) + .5 h1h_i) + .5 h2p_i) + .5 h2h_i)  (PETScVlasovSolver.pyx:145-149);
            // outside the grid H is 0: those cells only meet rows whose result
            // is not formed (j < 2 or j >= nv - 2) except in the transposed mode
            if (a.h2d != nullptr) {
                sH[r][c] = in ? (T)hq[q] : T(0);
            } else if (a.hxv != nullptr) {
                sH[r][c] = in ? (T)sX[0][c] : T(0);
            } else if (sizeof(T) == 8) {
                sH[r][c] = in ? (T)compose_have(sV[r], sX[0][c], sX[1][c], sX[2][c], sX[3][c]) : T(0);
            } else {
                sH[r][c] = in ? (T)(compose_have(sV[r], sX[0][c], sX[1][c], sX[2][c], sX[3][c]) - href) : T(0);
            }
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i = i0 + lane;
    if (i >= nx) return;
    const int cx = lane + 2;
    const T ht_inv = (T)a.ht_inv, af1 = (T)a.inst[b].af1, af2 = (T)a.inst[b].af2, factor = (T)a.factor;

    // Each lane walks down its column with the stencil values of F and H in
    // a sliding register window (arakawa.h).  Same values, same expressions.
    const int rr0 = wave * ROWS_PER_WAVE;
    const TileRd<T, HX> rF{&sF[0][0]}, rH{&sH[0][0]};
    ArakawaWin<T> win;
    win.load(rF, rH, rr0 + 2, cx);
#pragma unroll
    for (int q = 0; q < ROWS_PER_WAVE; ++q) {
        const int rr = rr0 + q;
        const int j = j0 + rr;
        if (j < nv) {
            const size_t o = (size_t)j * nx + i;
            T base;
            if (MODE == MODE_RESIDUAL) {
                const T fder = fin[o] - fhin[o];                 // Fder = (0 + F) - Fh
                base = T(0) + ht_inv * fder;                     // Y = 0; Y += Fder/dt
            } else if (MODE == MODE_JACOBIAN) {
                base = jacobian_base(ht_inv, win.f0[2]);         // Y = 0; Y += u/dt
            } else if (MODE == MODE_TRANSPOSED) {
                base = (T)a.alpha * (fin[o] * us);
            } else {
                base = a.store ? T(0) : yout[o];
            }
            T out = base;
            if (MODE == MODE_TRANSPOSED || (j >= 2 && j < nv - 2))
                out = arakawa_add<KIND>(win, base, factor, af1, af2);
            yout[o] = out;
        }
        // move the window one row down (rows stay inside the staged tile)
        if (q + 1 < ROWS_PER_WAVE) win.advance(rF, rH, rr + 3, cx);
    }
}

template <int MODE, typename T>
hipError_t launch_mode(vv_ctx *c, BrArgs<T> a)
{
    dim3 grid(br_tiles(c->p.nx, c->p.nv), c->B), block(NT);
    switch (c->p.bracket) {
    case VV_BRACKET_J1:
        launch_timed(c, bracket_kernel<VV_BRACKET_J1, MODE, T>, grid, block, 0, a);
        break;
    case VV_BRACKET_J2:
        launch_timed(c, bracket_kernel<VV_BRACKET_J2, MODE, T>, grid, block, 0, a);
        break;
    default:
        launch_timed(c, bracket_kernel<VV_BRACKET_J4, MODE, T>, grid, block, 0, a);
        break;
    }
    return hipGetLastError();
}

template <typename T>
BrArgs<T> base_args(vv_ctx *c)
{
    BrArgs<T> a{};
    a.nx = c->p.nx;
    a.nv = c->p.nv;
    a.N = c->N;
    a.inst = c->inst;
    a.gstride = (int64_t)sizeof(GmresState);
    a.h0 = c->h0;
    a.h1p = c->h1p;
    a.h1h = c->h1h;
    a.h2p = c->h2p;
    a.h2h = c->h2h;
    a.ht_inv = c->ht_inv;
    return a;
}

}  // namespace

hipError_t launch_residual(vv_ctx *c, const double *f, const double *fh, double *y)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = f;
    a.fh = fh;
    a.y = y;
    a.factor = 1.0;                 // B(Fave, H, 1.0)  (PETScNLVlasovMP.pyx:55)
    return launch_mode<MODE_RESIDUAL>(c, a);
}

hipError_t launch_jacobian(vv_ctx *c, const double *u, const double *uscale, double *y)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = u;
    a.uscale = uscale;
    a.y = y;
    a.factor = 0.5;                 // B(u, H, 0.5)  (PETScNLVlasovMP.pyx:37)
    a.done = uscale != nullptr ? &c->gs->st.done : nullptr;
    return launch_mode<MODE_JACOBIAN>(c, a);
}

hipError_t launch_bracket_add(vv_ctx *c, const double *x, double *y, double factor)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = x;
    a.y = y;
    a.factor = factor;
    return launch_mode<MODE_ACCUMULATE>(c, a);
}

hipError_t launch_bracket_t(vv_ctx *c, const double *u, const double *uscale, const double *h2d, const double *hxv,
                            double *y, double alpha, double factor)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = u;
    a.uscale = uscale;
    a.h2d = h2d;
    a.hxv = hxv;
    a.y = y;
    a.alpha = alpha;
    a.factor = factor;
    a.done = uscale != nullptr ? &c->gs->st.done : nullptr;
    return launch_mode<MODE_TRANSPOSED>(c, a);
}

hipError_t launch_bracket_gen(vv_ctx *c, const double *x, const double *h2d, const double *hxv, double *y,
                              double factor)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = x;
    a.h2d = h2d;
    a.hxv = hxv;
    a.y = y;
    a.factor = factor;
    return launch_mode<MODE_ACCUMULATE>(c, a);
}

hipError_t launch_bracket_acc(vv_ctx *c, const double *x, const double *xscale, const double *h2d, double *y,
                              double factor, bool store, const int32_t *done)
{
    BrArgs<double> a = base_args<double>(c);
    a.f = x;
    a.uscale = xscale;
    a.h2d = h2d;
    a.y = y;
    a.factor = factor;
    a.store = store ? 1 : 0;
    a.done = done;
    return launch_mode<MODE_ACCUMULATE>(c, a);
}

hipError_t launch_jacobian32(vv_ctx *c, const float *u, const double *uscale, float *y)
{
    BrArgs<float> a = base_args<float>(c);
    a.f = u;
    a.uscale = uscale;
    a.y = y;
    a.factor = 0.5;
    a.done = uscale != nullptr ? &c->gs->st.done : nullptr;
    return launch_mode<MODE_JACOBIAN>(c, a);
}

}  // namespace vv
