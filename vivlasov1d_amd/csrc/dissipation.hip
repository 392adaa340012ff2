// Collision operators and the collision factor of the dissipation terms
// (CollisionOperator.pyx:68-250, Toolbox.pyx:153-165).  The double brackets
// (DoubleBracket.pyx:67-97) are chains of the bracket kernel's accumulate mode
// and live with the other host logic.
//
// This translation unit is compiled with -ffp-contract=off: each expression
// keeps the reference's operation order, so the results are bit-identical to
// its unfused double arithmetic.
//
// Layout: natural order, element (i, j) at i + nx*j.  One thread per point
// of rows 2..nv-3; a wavefront covers 64 consecutive x of one row, so the
// three rows j-1, j, j+1 it reads are coalesced and shared through L2 with
// the neighbouring rows' wavefronts.  Per point: 3 reads of f (two of them
// L2 hits), one read and one write of y, and the x-vectors n, U, a.
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;

struct CollArgs {
    const double *f;
    const double *fscale;   // optional device scale of f (unnormalised Krylov vector)
    double *y;
    const double *v, *n, *U, *a;
    int nx, nv;
    double drag_fac, diff_fac;
    const int32_t *done;    // optional early-exit flag (GMRES cycle finished)
};

// CollisionOperatorT.function (CollisionOperator.pyx:98-110) when E == false,
// CollisionOperatorE.function (:190-202) when E == true.
template <bool E>
__global__ __launch_bounds__(NT) void collision_kernel(CollArgs p)
{
    if (p.done != nullptr && *p.done) return;
    const int i = blockIdx.x * NT + threadIdx.x;
    const int j = 2 + blockIdx.y;                  // rows 2..nv-3 (grid.stencil = 2)
    if (i >= p.nx) return;
    const size_t o = (size_t)j * p.nx + i;
    double fm = p.f[o - p.nx], f0 = p.f[o], fp = p.f[o + p.nx];
    if (p.fscale != nullptr) {
        const double s = *p.fscale;
        fm = fm * s;
        f0 = f0 * s;
        fp = fp * s;
    }
    const double u = p.U[i] / p.n[i];              // u = U.getArray() / n
    const double vp = p.v[j + 1], vm = p.v[j - 1];
    double drag, diff;
    if (!E) {
        drag = ((vp - u) * fp - (vm - u) * fm) * p.a[i];
        diff = (fp - 2. * f0) + fm;
    } else {
        drag = (vp - u) * fp - (vm - u) * fm;
        diff = ((fp - 2. * f0) + fm) / p.a[i];
    }
    p.y[o] = p.y[o] + (p.drag_fac * drag + p.diff_fac * diff);
}

// Fave = (0 + .5 Fh) + .5 F (PETScNLVlasovMP.pyx:44-46), the first argument of
// the double brackets
__global__ __launch_bounds__(NT) void average_kernel(const double *f, const double *fh, double *out, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k < n) out[k] = 0.5 * fh[k] + 0.5 * f[k];
}

// Toolbox.compute_collision_factor (Toolbox.pyx:164-165)
__global__ __launch_bounds__(NT) void collision_factor_kernel(const double *n, const double *u, const double *e,
                                                              double *a, int nx)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nx) return;
    a[i] = n[i] * n[i] / (n[i] * e[i] - u[i] * u[i]);
}

// The FP32 GMRES cycle of the mixed solver takes the dissipation term in
// FP64: its FP32 Krylov vector is widened, the FP64 kernels above add the
// term into a zeroed FP64 field, and that is folded into the FP32 J u with
// one rounding.
__global__ __launch_bounds__(NT) void widen_kernel(const float *__restrict__ x, double *__restrict__ y, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k < n) y[k] = (double)x[k];
}

__global__ __launch_bounds__(NT) void narrow_add_kernel(float *y, const double *__restrict__ d, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k < n) y[k] = (float)((double)y[k] + d[k]);
}

}  // namespace

hipError_t launch_widen(vv_ctx *c, const float *x, double *y)
{
    const int64_t n = c->N * c->B;
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, c->stream, x, y, n);
    return hipGetLastError();
}

hipError_t launch_narrow_add(vv_ctx *c, float *y, const double *d)
{
    const int64_t n = c->N * c->B;
    hipLaunchKernelGGL(narrow_add_kernel, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, c->stream, y, d, n);
    return hipGetLastError();
}

hipError_t launch_collision(vv_ctx *c, const double *f, const double *fscale, double *y, const double *n,
                            const double *U, const double *a, double factor, const int32_t *done)
{
    const double nu = c->coll_freq;
    if (nu == 0. || factor == 0.) return hipSuccess;       // CollisionOperator.pyx:87
    if (c->p.nv <= 4) return hipSuccess;
    CollArgs p{};
    p.f = f;
    p.fscale = fscale;
    p.y = y;
    p.v = c->vgrid;
    p.n = n;
    p.U = U;
    p.a = a;
    p.nx = c->p.nx;
    p.nv = c->p.nv;
    // - factor * coll_freq * coll_drag * hv_inv * 0.5 and
    // - factor * coll_freq * coll_diff * hv2_inv, left to right (:80-81)
    p.drag_fac = -factor * nu * c->coll_drag * c->hv_inv * 0.5;
    p.diff_fac = -factor * nu * c->coll_diff * c->hv2_inv;
    p.done = done;
    dim3 grid((p.nx + NT - 1) / NT, p.nv - 4), block(NT);
    if (c->diss_kind == VV_DISSIPATION_COLLE) launch_timed(c, collision_kernel<true>, grid, block, 0, p);
    else launch_timed(c, collision_kernel<false>, grid, block, 0, p);
    return hipGetLastError();
}

hipError_t launch_average(vv_ctx *c, const double *f, const double *fh, double *out)
{
    const int64_t n = c->N * c->B;
    hipLaunchKernelGGL(average_kernel, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, c->stream, f, fh, out, n);
    return hipGetLastError();
}

hipError_t launch_collision_factor(vv_ctx *c, const double *n, const double *u, const double *e, double *a)
{
    hipLaunchKernelGGL(collision_factor_kernel, dim3((c->p.nx + NT - 1) / NT), dim3(NT), 0, c->stream, n, u, e, a,
                       c->p.nx);
    return hipGetLastError();
}

}  // namespace vv
