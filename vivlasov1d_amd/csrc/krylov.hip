// GMRES(m) with classical Gram-Schmidt -- the PETSc KSPGMRES defaults the
// reference runs (run_nonlinear_matrixfree_split_mp.py:47-53: restart 30,
// CGS without refinement, Givens rotations, zero initial guess) -- and the
// BLAS-1 reductions around it.
//
// The basis V_0..V_m is stored unnormalised: v_j = scale[j] * V_j, which is
// bit-identical to the reference's VecNormalize (x *= 1/||x||) because every
// reader multiplies V_j by scale[j] exactly as VecScale would have.  This
// saves the separate normalisation pass.
//
// All reductions have a fixed grid and a fixed summation tree, so results
// are reproducible run to run.  Cross-block sums use the last-block pattern:
// each block publishes its partial, fences, takes a ticket; the block that
// draws the last ticket fences again and finishes the sum.
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;

__device__ __forceinline__ double block_sum(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; ++w) s += red[w];
    return s;   // valid in thread 0
}

// Returns true in every thread of the block that drew the last ticket.
__device__ __forceinline__ bool last_block(uint32_t *counter, int *s_flag)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int t = atomicAdd(counter, 1u);
        *s_flag = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (*s_flag) __threadfence();
    return *s_flag != 0;
}

__device__ __forceinline__ int ksp_test(const GmresState *g, double res)
{
    if (!isfinite(res)) return VV_KSP_DIVERGED_NANORINF;
    if (res <= g->ttol) return res < g->atol ? VV_KSP_CONVERGED_ATOL : VV_KSP_CONVERGED_RTOL;
    if (res >= g->dtol * g->st.rnorm0) return VV_KSP_DIVERGED_DTOL;
    return 0;
}

__global__ void gmres_init_kernel(GmresState *g, double rtol, double atol, double dtol, int max_it, int max_k)
{
    if (threadIdx.x != 0) return;
    g->st.it = 0;
    g->st.its = 0;
    g->st.reason = 0;
    g->st.done = 0;
    g->st.res = 0.0;
    g->st.rnorm0 = 0.0;
    g->rtol = rtol;
    g->atol = atol;
    g->dtol = dtol;
    g->haptol = 1.0e-30;     // PETSc gmres->haptol
    g->max_it = max_it;
    g->max_k = max_k;
    g->first = 1;
    for (int i = 0; i < 8; ++i) g->counter[i] = 0u;
}

// Sum of squares of x (len elements) into per-block partials; the last block
// finishes.  mode 0: *out = sqrt(sum).  mode 1: GMRES cycle start on V_0.
__global__ __launch_bounds__(NT) void norm_kernel(const double *__restrict__ x, int64_t len, double *partial,
                                                  GmresState *g, double *out, int mode)
{
    __shared__ double red[NT / 64];
    __shared__ int s_last;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT) {
        const double v = x[e];
        acc += v * v;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (!last_block(&g->counter[mode == 1 ? CNT_START : CNT_NORM], &s_last)) return;
    if (threadIdx.x != 0) return;
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    const double nrm = sqrt(tot);
    g->counter[mode == 1 ? CNT_START : CNT_NORM] = 0u;
    if (mode == 0) {
        *out = nrm;
        return;
    }
    // KSPGMRESCycle entry: VecNormalize(VV(0)), zero-residual and convergence tests
    g->st.it = 0;
    g->st.done = 0;
    g->scale[0] = nrm != 0.0 ? 1.0 / nrm : 1.0;
    g->grs[0] = nrm;
    g->st.res = nrm;
    if (nrm == 0.0) {
        g->st.reason = VV_KSP_CONVERGED_ATOL;
        g->st.done = 1;
        return;
    }
    if (g->first) {
        g->st.rnorm0 = nrm;
        g->ttol = fmax(g->rtol * nrm, g->atol);
        g->first = 0;
    }
    const int rsn = ksp_test(g, nrm);
    if (rsn != 0) {
        g->st.reason = rsn;
        g->st.done = 1;
    }
}

// CGS update  V_{it+1} <- w - sum_j h_j v_j  (VecMAXPY with lhh = -h), its
// norm (VecNormalize), and KSPGMRESUpdateHessenberg + the convergence test.
__global__ __launch_bounds__(NT) void maxpy_norm_kernel(double *basis, int64_t N, int it, double *partial,
                                                        GmresState *g)
{
    if (g->st.done) return;
    __shared__ double red[NT / 64];
    __shared__ int s_last;
    __shared__ double s_coef[kMaxRestart + 1];
    const int nv = it + 1;
    __shared__ double s_scale[kMaxRestart + 1];
    if (threadIdx.x < nv) s_coef[threadIdx.x] = g->coef[threadIdx.x];
    if (threadIdx.x < nv) s_scale[threadIdx.x] = g->scale[threadIdx.x];
    __syncthreads();
    double2 *w2 = reinterpret_cast<double2 *>(basis + (size_t)nv * N);
    const int64_t n2 = N >> 1;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < n2; e += (int64_t)gridDim.x * NT) {
        double2 w = w2[e];
        for (int j = 0; j < nv; ++j) {
            const double2 v = reinterpret_cast<const double2 *>(basis + (size_t)j * N)[e];
            const double sj = s_scale[j], cj = s_coef[j];
            w.x = w.x + cj * (v.x * sj);
            w.y = w.y + cj * (v.y * sj);
        }
        w2[e] = w;
        acc += w.x * w.x;
        acc += w.y * w.y;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (!last_block(&g->counter[CNT_MAXPY], &s_last)) return;
    if (threadIdx.x != 0) return;
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    g->counter[CNT_MAXPY] = 0u;
    const double tt = sqrt(tot);
    double *hh = g->hh + (size_t)it * (kMaxRestart + 1);
    hh[it + 1] = tt;
    g->scale[it + 1] = tt != 0.0 ? 1.0 / tt : 1.0;
    // happy breakdown
    double hapbnd = fabs(tt / g->grs[it]);
    if (hapbnd > g->haptol) hapbnd = g->haptol;
    const bool hapend = tt < hapbnd;
    // apply the previous rotations to the new column
    for (int j = 1; j <= it; ++j) {
        const double t0 = hh[j - 1];
        hh[j - 1] = g->cc[j - 1] * t0 + g->ss[j - 1] * hh[j];
        hh[j] = g->cc[j - 1] * hh[j] - g->ss[j - 1] * t0;
    }
    double res;
    int reason = 0;
    if (!hapend) {
        const double d = sqrt(hh[it] * hh[it] + hh[it + 1] * hh[it + 1]);
        if (d == 0.0) {
            reason = VV_KSP_DIVERGED_BREAKDOWN;
            res = g->st.res;
        } else {
            g->cc[it] = hh[it] / d;
            g->ss[it] = hh[it + 1] / d;
            g->grs[it + 1] = -(g->ss[it] * g->grs[it]);
            g->grs[it] = g->cc[it] * g->grs[it];
            hh[it] = g->cc[it] * hh[it] + g->ss[it] * hh[it + 1];
            res = fabs(g->grs[it + 1]);
        }
    } else {
        res = 0.0;
    }
    g->st.it = it + 1;
    g->st.its += 1;
    g->st.res = res;
    if (reason == 0) reason = ksp_test(g, res);
    if (hapend && reason == 0) reason = VV_KSP_DIVERGED_BREAKDOWN;
    g->st.reason = reason;
    if (reason != 0 || g->st.it >= g->max_k || g->st.its >= g->max_it) g->st.done = 1;
}

// KSPGMRESBuildSoln: back substitution with the rotated Hessenberg matrix;
// the coefficients go to coef[0..it).
__global__ void backsolve_kernel(GmresState *g)
{
    if (threadIdx.x != 0) return;
    const int n = g->st.it;
    if (n <= 0) return;
    const int k = n - 1;
    const double *hh = g->hh;
    auto H = [&](int r, int col) { return hh[(size_t)col * (kMaxRestart + 1) + r]; };
    double y[kMaxRestart];
    if (H(k, k) != 0.0) {
        y[k] = g->grs[k] / H(k, k);
    } else {
        g->st.reason = VV_KSP_DIVERGED_BREAKDOWN;
        y[k] = 0.0;
    }
    for (int ii = 1; ii <= k; ++ii) {
        const int kk = k - ii;
        double tt = g->grs[kk];
        for (int j = kk + 1; j <= k; ++j) tt = tt - H(kk, j) * y[j];
        if (H(kk, kk) == 0.0) {
            g->st.reason = VV_KSP_DIVERGED_BREAKDOWN;
            y[kk] = 0.0;
        } else {
            y[kk] = tt / H(kk, kk);
        }
    }
    for (int j = 0; j <= k; ++j) g->coef[j] = y[j];
}

// x += sum_j coef_j v_j  (VEC_TEMP = 0 + sum_j nrs_j v_j; x = x + VEC_TEMP)
__global__ __launch_bounds__(NT) void update_kernel(double *__restrict__ x, const double *basis, int64_t N,
                                                    const GmresState *g)
{
    const int n = g->st.it;
    if (n <= 0) return;
    __shared__ double s_coef[kMaxRestart + 1], s_scale[kMaxRestart + 1];
    if (threadIdx.x < n) {
        s_coef[threadIdx.x] = g->coef[threadIdx.x];
        s_scale[threadIdx.x] = g->scale[threadIdx.x];
    }
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < N; e += (int64_t)gridDim.x * NT) {
        double t = 0.0;
        for (int j = 0; j < n; ++j) t = t + s_coef[j] * (basis[(size_t)j * N + e] * s_scale[j]);
        x[e] = x[e] + t;
    }
}

__global__ __launch_bounds__(NT) void dot_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                                 int64_t len, double *partial, GmresState *g, double *out)
{
    __shared__ double red[NT / 64];
    __shared__ int s_last;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        acc += x[e] * y[e];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (!last_block(&g->counter[CNT_DOT], &s_last)) return;
    if (threadIdx.x != 0) return;
    double tot = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) tot += partial[b];
    g->counter[CNT_DOT] = 0u;
    *out = tot;
}

// y = y + alpha*x (VecAXPY).  mode 1: only when the GMRES reason is positive
// (SNESSolve_KSPONLY takes the step unless the linear solve diverged).
__global__ __launch_bounds__(NT) void axpy_kernel(double alpha, const double *__restrict__ x, double *y, int64_t len,
                                                  const GmresState *g, int mode)
{
    if (mode == 1 && g->st.reason <= 0) return;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        y[e] = y[e] + alpha * x[e];
}

// y = b - y   (KSPInitialResidual after a restart: r = b - A x)
__global__ __launch_bounds__(NT) void rsub_kernel(const double *__restrict__ b, double *y, int64_t len)
{
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        y[e] = b[e] + (-1.0) * y[e];
}

__global__ __launch_bounds__(NT) void copy_kernel(const double *__restrict__ x, double *__restrict__ y, int64_t len)
{
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT) y[e] = x[e];
}

int grid_for(int64_t len)
{
    const int64_t want = (len + NT - 1) / NT;
    return (int)(want < kReduceBlocks ? (want > 0 ? want : 1) : kReduceBlocks);
}

}  // namespace

hipError_t launch_gmres_init(vv_ctx *c, double rtol, double atol, double dtol, int max_it)
{
    hipLaunchKernelGGL(gmres_init_kernel, dim3(1), dim3(64), 0, c->stream, c->gs, rtol, atol, dtol, max_it,
                       c->p.restart);
    return hipGetLastError();
}

hipError_t launch_norm_start(vv_ctx *c)
{
    hipLaunchKernelGGL(norm_kernel, dim3(grid_for(c->N)), dim3(NT), 0, c->stream, c->basis, c->N, c->partial,
                       c->gs, (double *)nullptr, 1);
    return hipGetLastError();
}

hipError_t launch_maxpy_norm(vv_ctx *c, int it)
{
    hipLaunchKernelGGL(maxpy_norm_kernel, dim3(grid_for(c->N / 2)), dim3(NT), 0, c->stream, c->basis, c->N, it,
                       c->partial, c->gs);
    return hipGetLastError();
}

hipError_t launch_build_solution(vv_ctx *c, double *x)
{
    hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(64), 0, c->stream, c->gs);
    hipLaunchKernelGGL(update_kernel, dim3(grid_for(c->N)), dim3(NT), 0, c->stream, x, c->basis, c->N, c->gs);
    return hipGetLastError();
}

hipError_t launch_residual_restart(vv_ctx *c, const double *b)
{
    hipLaunchKernelGGL(rsub_kernel, dim3(grid_for(c->N)), dim3(NT), 0, c->stream, b, c->basis, c->N);
    return hipGetLastError();
}

hipError_t launch_norm2(vv_ctx *c, const double *x, int64_t n, double *out_dev)
{
    hipLaunchKernelGGL(norm_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, x, n, c->partial, c->gs,
                       out_dev, 0);
    return hipGetLastError();
}

hipError_t launch_dot(vv_ctx *c, const double *x, const double *y, int64_t n, double *out_dev)
{
    hipLaunchKernelGGL(dot_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, x, y, n, c->partial, c->gs, out_dev);
    return hipGetLastError();
}

hipError_t launch_axpy(vv_ctx *c, double alpha, const double *x, double *y, int64_t n)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, alpha, x, y, n, c->gs, 0);
    return hipGetLastError();
}

hipError_t launch_axpy_cond(vv_ctx *c, double alpha, const double *x, double *y, int64_t n)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, alpha, x, y, n, c->gs, 1);
    return hipGetLastError();
}

hipError_t launch_copy(vv_ctx *c, const double *x, double *y, int64_t n)
{
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, x, y, n);
    return hipGetLastError();
}

}  // namespace vv
