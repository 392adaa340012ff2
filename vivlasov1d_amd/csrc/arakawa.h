// The Arakawa J1 / J2 / J4 stencils on one lane's sliding register window,
// shared by the bracket kernels (bracket.hip) and the Jacobian-vector product
// fused with the forward rFFT (spectral.hip).
//
// Every expression keeps the reference's IEEE operation order
// (PoissonBracket.pyx:100-113, 221-234, 360-389; C left-to-right) and is
// evaluated without contraction whatever the including unit's setting (the
// pragma below; the spectral unit is built with -ffp-contract=fast-honor-
// pragmas), so both kernels produce the reference's unfused double results
// bit for bit.
//
// The window of a lane at column cx, output row cy of a tile (read through
// a functor of (row, column, column offset)) holds the first argument F and the second
// argument H at column 0: rows cy-2..cy+2, columns +-1: rows cy-1..cy+1,
// columns +-2: row cy; it moves one row down per output row, so a row costs
// 5 + 5 LDS reads instead of 25.
#pragma once

#include "vv1d_internal.h"

namespace vv {
namespace {

// Cell (r, cx + d) of an LDS tile of pitch P (the column offset d is a
// compile-time constant at every call, so it folds into the LDS address).
template <typename T, int P>
struct TileRd {
    const T *p;
    __device__ __forceinline__ T operator()(int r, int cx, int d) const { return p[r * P + cx + d]; }
};

template <typename T>
struct ArakawaWin {
    T f0[5], fm[3], fp[3], fm2, fp2;
    T g0[5], gm[3], gp[3], gm2, gp2;

    template <typename RF, typename RH>
    __device__ __forceinline__ void load(const RF &F, const RH &H, int cy, int cx)
    {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            f0[k] = F(cy - 2 + k, cx, 0);
            g0[k] = H(cy - 2 + k, cx, 0);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            fm[k] = F(cy - 1 + k, cx, -1);
            fp[k] = F(cy - 1 + k, cx, 1);
            gm[k] = H(cy - 1 + k, cx, -1);
            gp[k] = H(cy - 1 + k, cx, 1);
        }
        fm2 = F(cy, cx, -2);
        fp2 = F(cy, cx, 2);
        gm2 = H(cy, cx, -2);
        gp2 = H(cy, cx, 2);
    }

    // move to output row cn (= the current row + 1)
    template <typename RF, typename RH>
    __device__ __forceinline__ void advance(const RF &F, const RH &H, int cn, int cx)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f0[k] = f0[k + 1];
            g0[k] = g0[k + 1];
        }
        f0[4] = F(cn + 2, cx, 0);
        g0[4] = H(cn + 2, cx, 0);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            fm[k] = fm[k + 1];
            fp[k] = fp[k + 1];
            gm[k] = gm[k + 1];
            gp[k] = gp[k + 1];
        }
        fm[2] = F(cn + 1, cx, -1);
        fp[2] = F(cn + 1, cx, 1);
        gm[2] = H(cn + 1, cx, -1);
        gp[2] = H(cn + 1, cx, 1);
        fm2 = F(cn, cx, -2);
        fp2 = F(cn, cx, 2);
        gm2 = H(cn, cx, -2);
        gp2 = H(cn, cx, 2);
    }
};

// base + factor {F, H} at the window's centre (the reference's
// Y += factor * bracket with Y = base).
template <int KIND, typename T>
__device__ __forceinline__ T arakawa_add(const ArakawaWin<T> &w, T base, T factor, T af1, T af2)
{
#pragma clang fp contract(off)
#define F(di, dj) ((di) == 0 ? w.f0[(dj) + 2] : (di) == -1 ? w.fm[(dj) + 1] : (di) == 1 ? w.fp[(dj) + 1] : (di) == -2 ? w.fm2 : w.fp2)
#define H(di, dj) ((di) == 0 ? w.g0[(dj) + 2] : (di) == -1 ? w.gm[(dj) + 1] : (di) == 1 ? w.gp[(dj) + 1] : (di) == -2 ? w.gm2 : w.gp2)
    if (KIND == VV_BRACKET_J1 || KIND == VV_BRACKET_J4) {
        // PoissonBracket.pyx:100-111 (J1 / second order)
        const T jpp = (F(1, 0) - F(-1, 0)) * (H(0, 1) - H(0, -1))
                    - (F(0, 1) - F(0, -1)) * (H(1, 0) - H(-1, 0));
        const T jpc = F(1, 0) * (H(1, 1) - H(1, -1))
                    - F(-1, 0) * (H(-1, 1) - H(-1, -1))
                    - F(0, 1) * (H(1, 1) - H(-1, 1))
                    + F(0, -1) * (H(1, -1) - H(-1, -1));
        const T jcp = F(1, 1) * (H(0, 1) - H(1, 0))
                    - F(-1, -1) * (H(-1, 0) - H(0, -1))
                    - F(-1, 1) * (H(0, 1) - H(-1, 0))
                    + F(1, -1) * (H(1, 0) - H(0, -1));
        if (KIND == VV_BRACKET_J1) return base + factor * (jpp + jpc + jcp) * af1;
        // PoissonBracket.pyx:373-389 (J4 = 2 J1 - J2)
        const T jcc2 = (F(1, 1) - F(-1, -1)) * (H(-1, 1) - H(1, -1))
                     - (F(-1, 1) - F(1, -1)) * (H(1, 1) - H(-1, -1));
        const T jpc2 = F(2, 0) * (H(1, 1) - H(1, -1))
                     - F(-2, 0) * (H(-1, 1) - H(-1, -1))
                     - F(0, 2) * (H(1, 1) - H(-1, 1))
                     + F(0, -2) * (H(1, -1) - H(-1, -1));
        const T jcp2 = F(1, 1) * (H(0, 2) - H(2, 0))
                     - F(-1, -1) * (H(-2, 0) - H(0, -2))
                     - F(-1, 1) * (H(0, 2) - H(-2, 0))
                     + F(1, -1) * (H(2, 0) - H(0, -2));
        const T r1 = (jpp + jpc + jcp) * af1;
        const T r2 = (jcc2 + jpc2 + jcp2) * af2;
        return base + factor * (T(2) * r1 - r2);
    } else {
        // PoissonBracket.pyx:221-234 (J2)
        const T jcc = (F(1, 1) - F(-1, -1)) * (H(-1, 1) - H(1, -1))
                    - (F(-1, 1) - F(1, -1)) * (H(1, 1) - H(-1, -1));
        const T jpc = F(2, 0) * (H(1, 1) - H(1, -1))
                    - F(-2, 0) * (H(-1, 1) - H(-1, -1))
                    - F(0, 2) * (H(1, 1) - H(-1, 1))
                    + F(0, -2) * (H(1, -1) - H(-1, -1));
        const T jcp = F(1, 1) * (H(0, 2) - H(2, 0))
                    - F(-1, -1) * (H(-2, 0) - H(0, -2))
                    - F(-1, 1) * (H(0, 2) - H(-2, 0))
                    + F(1, -1) * (H(2, 0) - H(0, -2));
        return base + factor * (jcc + jpc + jcp) * af2;
    }
#undef F
#undef H
}

// H_ave at (i, j) from its 1D factors: ((((h0_j + .5 h1p_i) + .5 h1h_i) +
// .5 h2p_i) + .5 h2h_i), exactly the value the reference's 2D Have vector
// holds (PETScVlasovSolver.pyx:145-149).
__device__ __forceinline__ double compose_have(double h0, double x0, double x1, double x2, double x3)
{
#pragma clang fp contract(off)
    return (((h0 + 0.5 * x0) + 0.5 * x1) + 0.5 * x2) + 0.5 * x3;
}

// The Jacobian's Y = 0; Y += u/dt (PETScNLVlasovMP.pyx:35-36).
template <typename T>
__device__ __forceinline__ T jacobian_base(T ht_inv, T u)
{
#pragma clang fp contract(off)
    return T(0) + ht_inv * u;
}

}  // namespace
}  // namespace vv
