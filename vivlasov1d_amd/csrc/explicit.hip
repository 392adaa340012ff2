// Explicit Arakawa steppers of the initial guess (run_base.py:786-927;
// PETScArakawaSymplectic.pyx:62-97, PETScArakawaRungeKutta.pyx:20-98): the
// Hamiltonian of a stepper as a 2D field and the elementwise stage
// combinations.  The tendencies -{x, H} are the bracket kernel's store mode
// with factor -1.
//
// This translation unit is compiled with -ffp-contract=off: the combinations
// keep NumPy's evaluation order of the reference's expressions (scalar
// coefficients formed first on the host), so they are bit-identical to it.
//
// Layout: natural order, element (i, j) at i + nx*j; every kernel is one
// thread per element, HBM-bound (8 B per operand per point).
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;

// H[j][i] = h0[j], h1[i] + h2[i] or (h0[j] + h1[i]) + h2[i] (absent terms null):
// the reference's 2D Vecs H0, H1 + H2 and H0 + H1 + H2
__global__ __launch_bounds__(NT) void compose_h_kernel(const double *h0, const double *h1, const double *h2,
                                                       double *H, int nx, int64_t n)
{
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k >= n) return;
    const int i = (int)(k % nx);
    const int j = (int)(k / nx);
    double h;
    if (h0 != nullptr) {
        h = h0[j];
        if (h1 != nullptr) h = (h + h1[i]) + h2[i];
    } else {
        h = h1[i] + h2[i];
    }
    H[k] = h;
}

enum { C_AXPY = 0, C_RK18_S3 = 1, C_RK18_S4 = 2, C_RK16_END = 3, C_RK18_END = 4 };

struct Combo {
    const double *x, *t1, *t2, *t3, *t4;
    double *out;
    double s, c;
    int64_t n;
};

template <int OP>
__global__ __launch_bounds__(NT) void combo_kernel(Combo a)
{
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k >= a.n) return;
    const double x = a.x[k];
    double r;
    if (OP == C_AXPY) {                 // x + s * t1
        r = x + a.s * a.t1[k];
    } else if (OP == C_RK18_S3) {       // x - c * t1 + s * t2
        r = (x - a.c * a.t1[k]) + a.s * a.t2[k];
    } else if (OP == C_RK18_S4) {       // x + s * t1 - s * t2 + s * t3
        r = ((x + a.s * a.t1[k]) - a.s * a.t2[k]) + a.s * a.t3[k];
    } else if (OP == C_RK16_END) {      // x + s * (t1 + 2 t2 + 2 t3 + t4) / 6
        r = x + (a.s * (((a.t1[k] + 2. * a.t2[k]) + 2. * a.t3[k]) + a.t4[k])) / 6.;
    } else {                            // x + s * (t1 + 3 t2 + 3 t3 + t4) / 8
        r = x + (a.s * (((a.t1[k] + 3. * a.t2[k]) + 3. * a.t3[k]) + a.t4[k])) / 8.;
    }
    a.out[k] = r;
}

template <int OP>
hipError_t combo(vv_ctx *c, Combo a)
{
    a.n = c->N;
    hipLaunchKernelGGL(combo_kernel<OP>, dim3((unsigned)((a.n + NT - 1) / NT)), dim3(NT), 0, c->stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_compose_h(vv_ctx *c, const double *h0, const double *h1, const double *h2, double *H)
{
    const int64_t n = c->N;
    hipLaunchKernelGGL(compose_h_kernel, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, c->stream, h0, h1, h2,
                       H, c->p.nx, n);
    return hipGetLastError();
}

// out = x + s t
hipError_t launch_stage(vv_ctx *c, const double *x, const double *t, double s, double *out)
{
    Combo a{};
    a.x = x;
    a.t1 = t;
    a.s = s;
    a.out = out;
    return combo<C_AXPY>(c, a);
}

hipError_t launch_rk18_stage3(vv_ctx *c, const double *x, const double *t1, const double *t2, double cc, double s,
                              double *out)
{
    Combo a{};
    a.x = x;
    a.t1 = t1;
    a.t2 = t2;
    a.c = cc;
    a.s = s;
    a.out = out;
    return combo<C_RK18_S3>(c, a);
}

hipError_t launch_rk18_stage4(vv_ctx *c, const double *x, const double *t1, const double *t2, const double *t3,
                              double s, double *out)
{
    Combo a{};
    a.x = x;
    a.t1 = t1;
    a.t2 = t2;
    a.t3 = t3;
    a.s = s;
    a.out = out;
    return combo<C_RK18_S4>(c, a);
}

hipError_t launch_rk4_end(vv_ctx *c, bool three_eighths, double *x, const double *t1, const double *t2,
                          const double *t3, const double *t4, double s)
{
    Combo a{};
    a.x = x;
    a.t1 = t1;
    a.t2 = t2;
    a.t3 = t3;
    a.t4 = t4;
    a.s = s;
    a.out = x;
    return three_eighths ? combo<C_RK18_END>(c, a) : combo<C_RK16_END>(c, a);
}

}  // namespace vv
