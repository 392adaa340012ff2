// Moments, the periodic Poisson solve and the invariants.
//
//  density_kernel    column partial sums of f over chunks of v-rows
//                    (n_i = hv * sum_j f_ij, Toolbox.pyx:79-98); coalesced
//                    along x, one chunk of rows per block row.
//  poisson_*_kernel  prep (one workgroup): finishes n, forms
//                    b = -(n - mean n) q (PETScPoissonSolver.pyx:58-74);
//                    conv: solves A phi = b by the mean-free circulant inverse
//                    (the periodic CFD4/CFD2 matrix is circulant; its
//                    pseudo-inverse's first row is precomputed on the host);
//                    post (one workgroup): h1 = phi - mean phi
//                    (Toolbox.pyx:38-74) and the Poisson residual norm
//                    ||A phi + (n - mean n) q|| of run_base.py:631-637.
//  velocity_moments_kernel
//                    u_i = hv sum_j v_j f_ij and e_i = hv sum_j v_j^2 f_ij
//                    (Toolbox.pyx:103-148), the output moments: one thread
//                    per column, rows in order, as the reference's loop.
//  invariants_kernel mass, L2, kinetic / potential energy over f, plus the
//                    field energy over phi (distribution_function.pyx:130-199,
//                    hamiltonian.pyx:67-100, potential.pyx:48-59).
//
// In a batched context instance b is blockIdx.y (blockIdx.z for the density
// partials); its fields start at b * N, its x-vectors at b * nx.
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;
constexpr int PT = 1024;     // threads of the single-workgroup Poisson kernel
constexpr int DCH = kDensityChunks;   // row chunks of the density reduction

__global__ __launch_bounds__(NT) void density_kernel(const double *__restrict__ f, int nx, int nv,
                                                     double *__restrict__ npart)
{
    f += (size_t)blockIdx.z * nx * nv;
    npart += (size_t)blockIdx.z * DCH * nx;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nx) return;
    const int rows = (nv + DCH - 1) / DCH;
    const int j0 = blockIdx.y * rows;
    const int j1 = min(nv, j0 + rows);
    double s = 0.0;
    for (int j = j0; j < j1; ++j) s += f[(size_t)j * nx + i];
    npart[(size_t)blockIdx.y * nx + i] = s;
}

__global__ __launch_bounds__(NT) void velocity_moments_kernel(const double *__restrict__ f,
                                                              const double *__restrict__ v, int nx, int nv,
                                                              double hv, double *__restrict__ u, double *__restrict__ e)
{
    f += (size_t)blockIdx.y * nx * nv;
    u += (size_t)blockIdx.y * nx;
    e += (size_t)blockIdx.y * nx;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nx) return;
    double su = 0.0, se = 0.0;
    for (int j = 0; j < nv; ++j) {
        const double fv = f[(size_t)j * nx + i];
        su += v[j] * fv;
        se += (v[j] * v[j]) * fv;          // v2 = v**2 (Grid), then v2_j * f_ij
    }
    u[i] = su * hv;
    e[i] = se * hv;
}

__device__ double block_sum_big(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < PT / 64; ++w) s += red[w];
        red[PT / 64] = s;
    }
    __syncthreads();
    return red[PT / 64];    // valid in every thread
}

struct PoiArgs {
    int nx, laplace, from_part;
    double hv, charge;
    const InstConst *inst;
    const double *npart;   // density partials (from_part = 1)
    const double *n_in;    // density (from_part = 0)
    const double *pinv;
    double *work;          // 3*nx + 1 scratch: b | phi | n | mean n
    double *n_out, *phi, *h1, *pres;
    int pres_stride;       // doubles between the instances' residual norms in pres
};

// A x for the periodic CFD4 / CFD2 stencils, in the reference's order
// (PETScPoissonCFD4.pyx:78, PETScPoissonCFD2.pyx:78).
__device__ __forceinline__ double lap(const double *x, int i, int nx, int laplace, double hx2_inv)
{
    const int im2 = i >= 2 ? i - 2 : i - 2 + nx, im1 = i >= 1 ? i - 1 : i - 1 + nx;
    const int ip1 = i + 1 < nx ? i + 1 : i + 1 - nx, ip2 = i + 2 < nx ? i + 2 : i + 2 - nx;
    const double xm2 = x[im2], xm1 = x[im1], x0 = x[i], xp1 = x[ip1], xp2 = x[ip2];
    if (laplace == VV_LAPLACE_CFD4)
        return (5. * xm2 - 32. * xm1 + 54. * x0 - 32. * xp1 + 5. * xp2) * hx2_inv / 12.;
    return (2. * x0 - xm1 - xp1) * hx2_inv;
}

// per-instance views of the Poisson buffers (pres: 16 doubles per instance)
__device__ __forceinline__ void poi_instance(PoiArgs &a)
{
    const size_t b = blockIdx.y, nx = a.nx;
    if (a.npart) a.npart += b * DCH * nx;
    if (a.n_in) a.n_in += b * nx;
    a.pinv += b * nx;
    a.work += b * (3 * nx + 1);
    if (a.n_out) a.n_out += b * nx;
    if (a.phi) a.phi += b * nx;
    if (a.h1) a.h1 += b * nx;
    if (a.pres) a.pres += b * a.pres_stride;
}

// one workgroup: n (from the density partials), mean n, b = -(n - mean n) q
__global__ __launch_bounds__(PT) void poisson_prep_kernel(PoiArgs a)
{
    __shared__ double red[PT / 64 + 1];
    poi_instance(a);
    const int nx = a.nx;
    double *b = a.work, *wn = a.work + 2 * nx;
    double part = 0.0;
    for (int i = threadIdx.x; i < nx; i += PT) {
        double n;
        if (a.from_part) {
            double v[DCH];
#pragma unroll
            for (int ch = 0; ch < DCH; ++ch) v[ch] = a.npart[(size_t)ch * nx + i];
            double s = 0.0;
#pragma unroll
            for (int ch = 0; ch < DCH; ++ch) s += v[ch];
            n = s * a.hv;
        } else {
            n = a.n_in[i];
        }
        wn[i] = n;
        if (a.n_out) a.n_out[i] = n;
        part += n;
    }
    const double nmean = block_sum_big(part, red) / nx;
    for (int i = threadIdx.x; i < nx; i += PT) b[i] = -(wn[i] - nmean) * a.charge;
    if (threadIdx.x == 0) a.work[3 * nx] = nmean;
}

// phi_i = sum_m c_m b_{i-m}: C the mean-free circulant inverse of A.  A
// workgroup forms 64 outputs (one per lane); its 4 wavefronts take a quarter
// of the m range each, with 4 interleaved accumulators per lane, and the
// quarters are added in order: fixed summation order, chains of nx/16 terms.
constexpr int CONV_OUT = 64;
__global__ __launch_bounds__(NT) void poisson_conv_kernel(PoiArgs a)
{
    extern __shared__ double sm[];
    __shared__ double part[NT / 64][CONV_OUT];
    poi_instance(a);
    const int nx = a.nx;
    double *sb = sm, *sc = sm + nx;
    for (int i = threadIdx.x; i < nx; i += NT) {
        sb[i] = a.work[i];
        sc[i] = a.pinv[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * CONV_OUT + lane;
    const int q = (nx + NT / 64 - 1) / (NT / 64);
    const int m0 = wave * q, m1 = m0 + q < nx ? m0 + q : nx;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (i < nx) {
        int m = m0;
        int d = i - m0;                          // (i - m) mod nx, kept in [0, nx)
        d = d < 0 ? d + nx : d;
        for (; m + 4 <= m1; m += 4) {
            const int d1 = d - 1 < 0 ? d - 1 + nx : d - 1;
            const int d2 = d1 - 1 < 0 ? d1 - 1 + nx : d1 - 1;
            const int d3 = d2 - 1 < 0 ? d2 - 1 + nx : d2 - 1;
            s0 += sc[m] * sb[d];
            s1 += sc[m + 1] * sb[d1];
            s2 += sc[m + 2] * sb[d2];
            s3 += sc[m + 3] * sb[d3];
            d = d3 - 1 < 0 ? d3 - 1 + nx : d3 - 1;
        }
        for (; m < m1; ++m) {
            s0 += sc[m] * sb[d];
            d = d - 1 < 0 ? d - 1 + nx : d - 1;
        }
    }
    part[wave][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wave == 0 && i < nx) {
        double s = part[0][lane];
        for (int w = 1; w < NT / 64; ++w) s += part[w][lane];
        a.work[nx + i] = s;
    }
}

// one workgroup: h1 = phi - mean phi and ||A phi + (n - mean n) q||
__global__ __launch_bounds__(PT) void poisson_post_kernel(PoiArgs a)
{
    extern __shared__ double sm[];
    __shared__ double red[PT / 64 + 1];
    poi_instance(a);
    const double hx2_inv = a.inst[blockIdx.y].hx2_inv;
    const int nx = a.nx;
    const double *wphi = a.work + nx, *wn = a.work + 2 * nx;
    const double nmean = a.work[3 * nx];
    double part = 0.0;
    for (int i = threadIdx.x; i < nx; i += PT) {
        sm[i] = wphi[i];
        part += wphi[i];
    }
    const double phiave = block_sum_big(part, red) / nx;
    part = 0.0;
    for (int i = threadIdx.x; i < nx; i += PT) {
        if (a.phi) a.phi[i] = sm[i];
        if (a.h1) a.h1[i] = sm[i] - phiave;
        const double r = lap(sm, i, nx, a.laplace, hx2_inv) + (wn[i] - nmean) * a.charge;
        part += r * r;
    }
    const double rr = block_sum_big(part, red);
    if (threadIdx.x == 0 && a.pres) *a.pres = sqrt(rr);
}

// y = A x (+ (n - mean n) q when n != null)
__global__ __launch_bounds__(PT) void poisson_mult_kernel(const double *x, const double *n, double *y, int nx,
                                                          int laplace, const InstConst *inst, double charge)
{
    extern __shared__ double sm[];
    __shared__ double red[PT / 64 + 1];
    const double hx2_inv = inst[blockIdx.y].hx2_inv;
    x += (size_t)blockIdx.y * nx;
    y += (size_t)blockIdx.y * nx;
    if (n) n += (size_t)blockIdx.y * nx;
    double part = 0.0;
    for (int i = threadIdx.x; i < nx; i += PT) {
        sm[i] = x[i];
        if (n) part += n[i];
    }
    const double nmean = block_sum_big(part, red) / nx;
    for (int i = threadIdx.x; i < nx; i += PT) {
        double v = lap(sm, i, nx, laplace, hx2_inv);
        if (n) v = v + (n[i] - nmean) * charge;
        y[i] = v;
    }
}

__global__ __launch_bounds__(PT) void pot_to_ham_kernel(const double *phi, double *h1, int nx)
{
    __shared__ double red[PT / 64 + 1];
    phi += (size_t)blockIdx.y * nx;
    h1 += (size_t)blockIdx.y * nx;
    double part = 0.0;
    for (int i = threadIdx.x; i < nx; i += PT) part += phi[i];
    const double ave = block_sum_big(part, red) / nx;
    for (int i = threadIdx.x; i < nx; i += PT) h1[i] = phi[i] - ave;
}

__global__ __launch_bounds__(NT) void density_finish_kernel(const double *npart, int nch, int nx, double hv,
                                                            double *n)
{
    npart += (size_t)blockIdx.y * DCH * nx;
    n += (size_t)blockIdx.y * nx;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nx) return;
    double s = 0.0;
    for (int ch = 0; ch < nch; ++ch) s += npart[(size_t)ch * nx + i];
    n[i] = s * hv;
}

struct InvArgs {
    const double *f, *h0, *h1, *h2, *phi;
    int nx, nv;
    double hv, charge;
    const InstConst *inst;
    double *partial;      // [gridDim.x][5] per instance, pstride apart
    int64_t pstride;
    double *out;          // 6 doubles per instance
};

__device__ __forceinline__ void inv_instance(InvArgs &a)
{
    const size_t b = blockIdx.y;
    a.f += b * a.nx * a.nv;
    a.h1 += b * a.nx;
    a.h2 += b * a.nx;
    a.phi += b * a.nx;
    a.partial += b * a.pstride;
    a.out += b * 6;
}

__global__ __launch_bounds__(NT) void invariants_kernel(InvArgs a)
{
    inv_instance(a);
    __shared__ double red[5][NT / 64];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int64_t N = (int64_t)a.nx * a.nv;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < N; e += (int64_t)gridDim.x * NT) {
        const int j = (int)(e / a.nx);
        const int i = (int)(e - (int64_t)j * a.nx);
        const double fv = a.f[e];
        if (j < a.nv - 1) {
            const int ip = (i + 1 == a.nx) ? 0 : i + 1;
            s[0] += (fv + a.f[(size_t)j * a.nx + ip] + a.f[(size_t)(j + 1) * a.nx + ip] + a.f[e + a.nx]);
            s[1] += fv * fv;
        }
        s[2] += fv * a.h0[j];
        s[3] += fv * a.h1[i];
        s[4] += fv * a.h2[i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 5; ++q) {
        double v = s[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane == 0) red[q][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double v = 0.0;
        for (int w = 0; w < NT / 64; ++w) v += red[threadIdx.x][w];
        a.partial[(size_t)blockIdx.x * 5 + threadIdx.x] = v;
    }
}

// one workgroup: sums the per-block partials (fixed order) and the field energy
__global__ __launch_bounds__(NT) void invariants_finish_kernel(InvArgs a, int nb)
{
    __shared__ double red[5][NT];
    inv_instance(a);
    const double hx = a.inst[blockIdx.y].hx;
    double t5[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += NT)
        for (int q = 0; q < 5; ++q) t5[q] += a.partial[(size_t)b * 5 + q];
    for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = t5[q];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < NT; ++b)
        for (int q = 0; q < 5; ++q) t[q] += red[q][b];
    const double hxhv = hx * a.hv;
    const double mass = t[0] * 0.25 * hx * a.hv;
    const double l2 = hxhv * t[1];
    const double ekin = t[2] * hx * a.hv;
    const double epot = (t[3] + t[4]) * hx * a.hv;
    double ef = 0.0;
    for (int i = 0; i < a.nx; ++i) {
        const double d = a.phi[(i + 1) % a.nx] - a.phi[i];
        ef += d * d;
    }
    a.out[0] = mass;
    a.out[1] = l2;
    a.out[2] = ekin + 0.5 * epot;
    a.out[3] = ekin;
    a.out[4] = epot;
    a.out[5] = 0.5 * a.charge * ef / hx;
}

}  // namespace

hipError_t launch_density_partial(vv_ctx *c, const double *f, double *npart)
{
    dim3 grid((c->p.nx + NT - 1) / NT, DCH, c->B);
    hipLaunchKernelGGL(density_kernel, grid, dim3(NT), 0, c->stream, f, c->p.nx, c->p.nv,
                       npart != nullptr ? npart : c->npart);
    return hipGetLastError();
}

hipError_t launch_density(vv_ctx *c, const double *f, double *n)
{
    hipError_t e = launch_density_partial(c, f, nullptr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(density_finish_kernel, dim3((c->p.nx + NT - 1) / NT, c->B), dim3(NT), 0, c->stream, c->npart,
                       DCH,
                       c->p.nx, c->p.hv, n);
    return hipGetLastError();
}

hipError_t launch_velocity_moments(vv_ctx *c, const double *f, double *u, double *e)
{
    hipLaunchKernelGGL(velocity_moments_kernel, dim3((c->p.nx + NT - 1) / NT, c->B), dim3(NT), 0, c->stream, f,
                       c->vgrid, c->p.nx, c->p.nv, c->p.hv, u, e);
    return hipGetLastError();
}

hipError_t launch_poisson(vv_ctx *c, const double *n_in, double *n_out, double *phi, double *h1, double *pres,
                          int from_part, const double *npart, double *work, int pres_stride)
{
    PoiArgs a{c->p.nx, c->p.laplace, from_part, c->p.hv, c->p.charge, c->inst,
              from_part ? (npart != nullptr ? npart : c->npart) : nullptr, n_in, c->pinv,
              work != nullptr ? work : c->pwork, n_out, phi, h1, pres, pres_stride};
    const int nx = c->p.nx;
    hipLaunchKernelGGL(poisson_prep_kernel, dim3(1, c->B), dim3(PT), 0, c->stream, a);
    hipLaunchKernelGGL(poisson_conv_kernel, dim3((nx + CONV_OUT - 1) / CONV_OUT, c->B), dim3(NT),
                       2 * nx * sizeof(double), c->stream, a);
    hipLaunchKernelGGL(poisson_post_kernel, dim3(1, c->B), dim3(PT), nx * sizeof(double), c->stream, a);
    return hipGetLastError();
}

hipError_t launch_poisson_mult(vv_ctx *c, const double *x, const double *n, double *y)
{
    const size_t lds = (size_t)c->p.nx * sizeof(double);
    hipLaunchKernelGGL(poisson_mult_kernel, dim3(1, c->B), dim3(PT), lds, c->stream, x, n, y, c->p.nx, c->p.laplace,
                       c->inst, c->p.charge);
    return hipGetLastError();
}

hipError_t launch_pot_to_ham(vv_ctx *c, const double *phi, double *h1)
{
    hipLaunchKernelGGL(pot_to_ham_kernel, dim3(1, c->B), dim3(PT), 0, c->stream, phi, h1, c->p.nx);
    return hipGetLastError();
}

hipError_t launch_invariants(vv_ctx *c, const double *f, const double *h0, const double *h1, const double *h2,
                             const double *phi, double *out_dev)
{
    InvArgs a{f, h0, h1, h2, phi, c->p.nx, c->p.nv, c->p.hv, c->p.charge, c->inst, c->partial, c->part_stride,
              out_dev};
    hipLaunchKernelGGL(invariants_kernel, dim3(kReduceBlocks, c->B), dim3(NT), 0, c->stream, a);
    hipLaunchKernelGGL(invariants_finish_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, a, kReduceBlocks);
    return hipGetLastError();
}

}  // namespace vv
