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
// Bandwidth layout of the CGS kernels (mdot, maxpy + norm, solution update):
// one thread owns EPT double2 elements (coalesced, NT apart) of every vector;
// vectors are streamed in groups of two whose 2*EPT loads are issued back to
// back, so each wavefront keeps 16 x 16 B in flight per lane.  The grid
// covers N exactly once (no grid-stride loop), EPT in {1,2,4,8} chosen so the
// launch has >= 256 workgroups.
//
// All reductions have a fixed grid and a fixed summation tree, so results
// are reproducible run to run.  Each reduction kernel writes one partial per
// workgroup; a one-workgroup "finish" kernel sums them (and, for GMRES, does
// the Hessenberg / Givens bookkeeping).  The kernel boundary orders the
// partials before the finish, so no device-scope fence (an L2 write-back on
// this GPU) is needed inside the streaming kernels.
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;
constexpr int NW = NT / 64;

// Sum over the block, returned in every thread (fixed tree).
__device__ __forceinline__ double block_sum_all(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < NW; ++w) s += red[w];
        red[NW] = s;
    }
    __syncthreads();
    const double s = red[NW];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;   // lane 0
}

// Sum of part[0..nb) by the whole block (thread t adds t, t+NT, ...).
__device__ __forceinline__ double final_sum_1(const double *part, int nb, double *red)
{
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) a += part[b];
    return block_sum_all(a, red);
}

// Column sums of part[b*PV + v] (b < nb) for v < PV, by the whole block:
// thread t owns column v = t % PV and rows b = t / PV (mod NT / PV), loads
// issued 16 at a time.  Result in out[v].
template <int PV>
__device__ void final_sum_cols(const double *part, int nb, double *out, double *red /*[NT]*/)
{
    constexpr int RS = NT / PV;      // row streams
    const int v = threadIdx.x % PV, r0 = threadIdx.x / PV;
    double acc = 0.0;
    for (int b0 = r0; b0 < nb; b0 += RS * 16) {
        double x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int b = b0 + q * RS;
            x[q] = (b < nb) ? part[(size_t)b * PV + v] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) acc += x[q];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < PV) {
        double s = 0.0;
        for (int r = 0; r < RS; ++r) s += red[r * PV + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
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

// Per-block partial sums of x*y over len elements (x == y: sum of squares).
// Cross-block sums are finished by a one-block kernel: the kernel boundary
// makes the partials visible without device-scope fences.
__global__ __launch_bounds__(NT) void dot_partial_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                                         int64_t len, double *partial)
{
    __shared__ double red[NW + 1];
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        acc += x[e] * y[e];
    const double s = block_sum_all(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// mode 0: *out = sqrt(sum) (VecNorm); mode 2: *out = sum (VecDot);
// mode 1: GMRES cycle start on V_0 with sum = ||V_0||^2.
__global__ __launch_bounds__(NT) void reduce_finish_kernel(const double *partial, int nb, GmresState *g, double *out,
                                                           int mode)
{
    __shared__ double red[NW + 1];
    const double tot = final_sum_1(partial, nb, red);
    if (threadIdx.x != 0) return;
    if (mode == 2) {
        *out = tot;
        return;
    }
    const double nrm = sqrt(tot);
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

struct CgsArgs {
    const double *basis;
    int64_t N;
    double *part;
    GmresState *g;
};

template <int EPT, int G>
__device__ __forceinline__ void load_group(const double *basis, int64_t N, int v0, int64_t e0, int64_t n2,
                                           double2 (&x)[G][EPT])
{
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        const double2 *V = reinterpret_cast<const double2 *>(basis + (size_t)(v0 + gg) * N);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int64_t e = e0 + (int64_t)q * NT;
            x[gg][q] = (e < n2) ? V[e] : make_double2(0.0, 0.0);
        }
    }
}

// VecMDot of w = V_{col+1} against v_0..v_col, written as HH(:, col) and the
// VecMAXPY coefficients -HH(:, col).
template <int EPT, int PV>
__global__ __launch_bounds__(NT) void mdot_kernel(CgsArgs a, int col)
{
    GmresState *g = a.g;
    if (g->st.done) return;
    __shared__ double red[NW][PV];
    __shared__ double s_scale[PV];
    const int nvec = col + 1;
    if (threadIdx.x < nvec) s_scale[threadIdx.x] = g->scale[threadIdx.x];
    __syncthreads();
    const int64_t n2 = a.N >> 1;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    const double2 *W = reinterpret_cast<const double2 *>(a.basis + (size_t)nvec * a.N);
    double2 w[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        w[q] = (e < n2) ? W[e] : make_double2(0.0, 0.0);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = 0;
    for (; v + 2 <= nvec; v += 2) {
        double2 x[2][EPT];
        load_group<EPT, 2>(a.basis, a.N, v, e0, n2, x);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sv = s_scale[v + gg];
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                acc += w[q].x * (x[gg][q].x * sv);
                acc += w[q].y * (x[gg][q].y * sv);
            }
            acc = wave_sum(acc);
            if (lane == 0) red[wave][v + gg] = acc;
        }
    }
    if (v < nvec) {
        double2 x[1][EPT];
        load_group<EPT, 1>(a.basis, a.N, v, e0, n2, x);
        const double sv = s_scale[v];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            acc += w[q].x * (x[0][q].x * sv);
            acc += w[q].y * (x[0][q].y * sv);
        }
        acc = wave_sum(acc);
        if (lane == 0) red[wave][v] = acc;
    }
    __syncthreads();
    if (threadIdx.x < PV) {
        double s = 0.0;
        if (threadIdx.x < nvec)
            for (int wv = 0; wv < NW; ++wv) s += red[wv][threadIdx.x];
        a.part[(size_t)blockIdx.x * PV + threadIdx.x] = s;
    }
}

// Column sums of the mdot partials: HH(:, col) = <w, v_v> and the VecMAXPY
// coefficients lhh = -HH(:, col).
template <int PV>
__global__ __launch_bounds__(NT) void mdot_finish_kernel(const double *part, int nb, int col, GmresState *g)
{
    if (g->st.done) return;
    __shared__ double s_red[NT];
    __shared__ double s_h[PV];
    final_sum_cols<PV>(part, nb, s_h, s_red);
    if (threadIdx.x < col + 1) {
        const double h = s_h[threadIdx.x];
        g->hh[col * (kMaxRestart + 1) + threadIdx.x] = h;
        g->coef[threadIdx.x] = -h;
    }
}

// CGS update  V_{it+1} <- w - sum_j h_j v_j  (VecMAXPY with lhh = -h) and
// per-block partial sums of its squared norm.
template <int EPT>
__global__ __launch_bounds__(NT) void maxpy_norm_kernel(CgsArgs a, int it)
{
    GmresState *g = a.g;
    if (g->st.done) return;
    __shared__ double red[NW + 1];
    __shared__ double s_coef[kMaxRestart + 1], s_scale[kMaxRestart + 1];
    const int nvec = it + 1;
    if (threadIdx.x < nvec) {
        s_coef[threadIdx.x] = g->coef[threadIdx.x];
        s_scale[threadIdx.x] = g->scale[threadIdx.x];
    }
    __syncthreads();
    const int64_t n2 = a.N >> 1;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    double2 *W = reinterpret_cast<double2 *>(const_cast<double *>(a.basis) + (size_t)nvec * a.N);
    double2 w[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        w[q] = (e < n2) ? W[e] : make_double2(0.0, 0.0);
    }
    int j = 0;
    for (; j + 2 <= nvec; j += 2) {
        double2 x[2][EPT];
        load_group<EPT, 2>(a.basis, a.N, j, e0, n2, x);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sj = s_scale[j + gg], cj = s_coef[j + gg];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                w[q].x = w[q].x + cj * (x[gg][q].x * sj);
                w[q].y = w[q].y + cj * (x[gg][q].y * sj);
            }
        }
    }
    if (j < nvec) {
        double2 x[1][EPT];
        load_group<EPT, 1>(a.basis, a.N, j, e0, n2, x);
        const double sj = s_scale[j], cj = s_coef[j];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            w[q].x = w[q].x + cj * (x[0][q].x * sj);
            w[q].y = w[q].y + cj * (x[0][q].y * sj);
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        if (e < n2) {
            W[e] = w[q];
            acc += w[q].x * w[q].x;
            acc += w[q].y * w[q].y;
        }
    }
    const double s = block_sum_all(acc, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = s;
}

// VecNormalize of V_{it+1} from the maxpy partials, then
// KSPGMRESUpdateHessenberg (rotations, happy breakdown, residual estimate)
// and the convergence test; sets done when the cycle ends.
__global__ __launch_bounds__(NT) void gmres_update_kernel(const double *part, int nb, int it, GmresState *g)
{
    if (g->st.done) return;
    __shared__ double red[NW + 1];
    const double tot = final_sum_1(part, nb, red);
    if (threadIdx.x != 0) return;
    const double tt = sqrt(tot);
    double *hh = g->hh + (size_t)it * (kMaxRestart + 1);
    hh[it + 1] = tt;
    g->scale[it + 1] = tt != 0.0 ? 1.0 / tt : 1.0;
    // happy breakdown
    double hapbnd = fabs(tt / g->grs[it]);
    if (hapbnd > g->haptol) hapbnd = g->haptol;
    const bool hapend = tt < hapbnd;
    // apply the previous rotations to the new column
    for (int q = 1; q <= it; ++q) {
        const double t0 = hh[q - 1];
        hh[q - 1] = g->cc[q - 1] * t0 + g->ss[q - 1] * hh[q];
        hh[q] = g->cc[q - 1] * hh[q] - g->ss[q - 1] * t0;
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
// the coefficients go to coef[0..it).  The triangle is staged in LDS first.
__global__ __launch_bounds__(NT) void backsolve_kernel(GmresState *g)
{
    __shared__ double sh[kMaxRestart][kMaxRestart];
    __shared__ double sg[kMaxRestart];
    const int n = g->st.it;
    if (n <= 0) return;
    for (int idx = threadIdx.x; idx < n * n; idx += NT) {
        const int col = idx / n, r = idx - col * n;
        sh[r][col] = g->hh[(size_t)col * (kMaxRestart + 1) + r];
    }
    if (threadIdx.x < n) sg[threadIdx.x] = g->grs[threadIdx.x];
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int k = n - 1;
    double y[kMaxRestart];
    if (sh[k][k] != 0.0) {
        y[k] = sg[k] / sh[k][k];
    } else {
        g->st.reason = VV_KSP_DIVERGED_BREAKDOWN;
        y[k] = 0.0;
    }
    for (int ii = 1; ii <= k; ++ii) {
        const int kk = k - ii;
        double tt = sg[kk];
        for (int j = kk + 1; j <= k; ++j) tt = tt - sh[kk][j] * y[j];
        if (sh[kk][kk] == 0.0) {
            g->st.reason = VV_KSP_DIVERGED_BREAKDOWN;
            y[kk] = 0.0;
        } else {
            y[kk] = tt / sh[kk][kk];
        }
    }
    for (int j = 0; j <= k; ++j) g->coef[j] = y[j];
}

// x += sum_j coef_j v_j  (VEC_TEMP = 0 + sum_j nrs_j v_j; x = x + VEC_TEMP)
template <int EPT>
__global__ __launch_bounds__(NT) void update_kernel(double *__restrict__ x, CgsArgs a)
{
    const GmresState *g = a.g;
    const int n = g->st.it;
    if (n <= 0) return;
    __shared__ double s_coef[kMaxRestart + 1], s_scale[kMaxRestart + 1];
    if (threadIdx.x < n) {
        s_coef[threadIdx.x] = g->coef[threadIdx.x];
        s_scale[threadIdx.x] = g->scale[threadIdx.x];
    }
    __syncthreads();
    const int64_t n2 = a.N >> 1;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    double2 t[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) t[q] = make_double2(0.0, 0.0);
    int j = 0;
    for (; j + 2 <= n; j += 2) {
        double2 v[2][EPT];
        load_group<EPT, 2>(a.basis, a.N, j, e0, n2, v);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sj = s_scale[j + gg], cj = s_coef[j + gg];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
                t[q].x = t[q].x + cj * (v[gg][q].x * sj);
                t[q].y = t[q].y + cj * (v[gg][q].y * sj);
            }
        }
    }
    if (j < n) {
        double2 v[1][EPT];
        load_group<EPT, 1>(a.basis, a.N, j, e0, n2, v);
        const double sj = s_scale[j], cj = s_coef[j];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            t[q].x = t[q].x + cj * (v[0][q].x * sj);
            t[q].y = t[q].y + cj * (v[0][q].y * sj);
        }
    }
    double2 *X = reinterpret_cast<double2 *>(x);
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        if (e < n2) {
            double2 xv = X[e];
            xv.x = xv.x + t[q].x;
            xv.y = xv.y + t[q].y;
            X[e] = xv;
        }
    }
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

// elements per thread of the CGS kernels: the largest of 8,4,2,1 that still
// gives >= 256 workgroups
int cgs_ept(int64_t N)
{
    const int64_t n2 = N >> 1;
    for (int ept = 8; ept > 1; ept >>= 1)
        if ((n2 + (int64_t)NT * ept - 1) / ((int64_t)NT * ept) >= 256) return ept;
    return 1;
}

}  // namespace

int cgs_blocks(int64_t N)
{
    const int ept = cgs_ept(N);
    return (int)(((N >> 1) + (int64_t)NT * ept - 1) / ((int64_t)NT * ept));
}

hipError_t launch_gmres_init(vv_ctx *c, double rtol, double atol, double dtol, int max_it)
{
    hipLaunchKernelGGL(gmres_init_kernel, dim3(1), dim3(64), 0, c->stream, c->gs, rtol, atol, dtol, max_it,
                       c->p.restart);
    return hipGetLastError();
}

hipError_t launch_norm_start(vv_ctx *c)
{
    const int nb = grid_for(c->N);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(NT), 0, c->stream, c->basis, c->basis, c->N, c->partial);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, nb, c->gs,
                       (double *)nullptr, 1);
    return hipGetLastError();
}

namespace {

template <int E>
void cgs_launch(vv_ctx *c, int which, int arg, double *x)
{
    CgsArgs a{c->basis, c->N, c->partial, c->gs};
    const dim3 grid(cgs_blocks(c->N)), block(NT);
    if (which == 0) {
        if (c->p.restart < 32) hipLaunchKernelGGL((mdot_kernel<E, 32>), grid, block, 0, c->stream, a, arg);
        else hipLaunchKernelGGL((mdot_kernel<E, 64>), grid, block, 0, c->stream, a, arg);
    } else if (which == 1) {
        hipLaunchKernelGGL((maxpy_norm_kernel<E>), grid, block, 0, c->stream, a, arg);
    } else {
        hipLaunchKernelGGL((update_kernel<E>), grid, block, 0, c->stream, x, a);
    }
}

hipError_t cgs_dispatch(vv_ctx *c, int which, int arg, double *x)
{
    switch (cgs_ept(c->N)) {
    case 8: cgs_launch<8>(c, which, arg, x); break;
    case 4: cgs_launch<4>(c, which, arg, x); break;
    case 2: cgs_launch<2>(c, which, arg, x); break;
    default: cgs_launch<1>(c, which, arg, x); break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_mdot(vv_ctx *c, int col)
{
    hipError_t e = cgs_dispatch(c, 0, col, nullptr);
    if (e != hipSuccess) return e;
    const int nb = cgs_blocks(c->N);
    if (c->p.restart < 32)
        hipLaunchKernelGGL((mdot_finish_kernel<32>), dim3(1), dim3(NT), 0, c->stream, c->partial, nb, col, c->gs);
    else
        hipLaunchKernelGGL((mdot_finish_kernel<64>), dim3(1), dim3(NT), 0, c->stream, c->partial, nb, col, c->gs);
    return hipGetLastError();
}

hipError_t launch_maxpy_norm(vv_ctx *c, int it)
{
    hipError_t e = cgs_dispatch(c, 1, it, nullptr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gmres_update_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, cgs_blocks(c->N), it, c->gs);
    return hipGetLastError();
}

hipError_t launch_build_solution(vv_ctx *c, double *x)
{
    hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(NT), 0, c->stream, c->gs);
    return cgs_dispatch(c, 2, 0, x);
}

hipError_t launch_residual_restart(vv_ctx *c, const double *b)
{
    hipLaunchKernelGGL(rsub_kernel, dim3(grid_for(c->N)), dim3(NT), 0, c->stream, b, c->basis, c->N);
    return hipGetLastError();
}

hipError_t launch_norm2(vv_ctx *c, const double *x, int64_t n, double *out_dev)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(NT), 0, c->stream, x, x, n, c->partial);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, nb, c->gs, out_dev, 0);
    return hipGetLastError();
}

hipError_t launch_dot(vv_ctx *c, const double *x, const double *y, int64_t n, double *out_dev)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(NT), 0, c->stream, x, y, n, c->partial);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, nb, c->gs, out_dev, 2);
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
