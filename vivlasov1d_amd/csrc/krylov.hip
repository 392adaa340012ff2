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
#include "cgs_vec.h"

namespace vv {
namespace {

constexpr int NT = kCgsNT;
constexpr int NW = kCgsNW;

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

// Sum of part[0..nb) by the whole block (thread t adds t, t+NT, ...).
__device__ __forceinline__ double final_sum_1(const double *part, int nb, double *red)
{
    double a = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) a += part[b];
    return block_sum_all(a, red);
}

// Column sums of part[b*PV + v] (b < nb) for v < ncols (<= PV), by the whole
// block: the columns are padded to cols = max(8, next power of two), thread t
// owns column t % cols and rows b = t / cols (mod NT / cols), loads issued 16
// at a time (one round for nb <= 16 NT / cols).  Result in out[v].
template <int PV>
__device__ void final_sum_cols(const double *part, int nb, int ncols, double *out, double *red /*[NT]*/)
{
    int cols = 8;
    while (cols < ncols) cols <<= 1;
    const int RS = NT / cols;        // row streams
    const int v = threadIdx.x & (cols - 1), r0 = threadIdx.x / cols;
    double acc = 0.0;
    if (v < ncols) {
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
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < ncols) {
        double s = 0.0;
        for (int r = 0; r < RS; ++r) s += red[r * cols + threadIdx.x];
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

// Instance b of a batched context is blockIdx.y in every kernel below.
__global__ void gmres_init_kernel(GmresState *g, const int32_t *active, double rtol, double atol, double dtol,
                                  int max_it, int max_k)
{
    if (threadIdx.x != 0) return;
    g += blockIdx.y;
    const int32_t act = active != nullptr ? active[blockIdx.y] : 1;
    g->st.it = 0;
    g->st.its = 0;
    g->st.reason = 0;
    g->st.done = act ? 0 : 1;
    g->st.active = act;
    g->st.res = 0.0;
    g->st.rnorm0 = 0.0;
    g->rtol = rtol;
    g->atol = atol;
    g->dtol = dtol;
    g->haptol = 1.0e-30;     // PETSc gmres->haptol
    g->max_it = max_it;
    g->max_k = max_k;
    g->first = 1;
}

// Per-block partial sums of x*y over len elements (x == y: sum of squares).
// Cross-block sums are finished by a one-block kernel: the kernel boundary
// makes the partials visible without device-scope fences.
__global__ __launch_bounds__(NT) void dot_partial_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                                         int64_t len, int64_t xstride, double *partial,
                                                         int64_t pstride)
{
    __shared__ double red[NW + 1];
    x += blockIdx.y * xstride;
    y += blockIdx.y * xstride;
    partial += blockIdx.y * pstride;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        acc += x[e] * y[e];
    const double s = block_sum_all(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// mode 0: *out = sqrt(sum) (VecNorm); mode 2: *out = sum (VecDot);
// mode 1: GMRES cycle start on V_0 with sum = ||V_0||^2.
__global__ __launch_bounds__(NT) void reduce_finish_kernel(const double *partial, int nb, int64_t pstride,
                                                           GmresState *g, double *out, int mode)
{
    __shared__ double red[NW + 1];
    partial += blockIdx.y * pstride;
    g += blockIdx.y;
    if (out != nullptr) out += blockIdx.y;
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
    // an instance that is not solving, or whose solve has ended, stays done
    if (!g->st.active || g->st.reason != 0) {
        g->st.it = 0;
        g->st.done = 1;
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

// TAIL: N is not a multiple of the vector width L (odd grids such as 9 x 9;
// one instance only, a batched context needs N % 4 == 0).  Thread t < N - L
// nvv of workgroup 0 then also owns the scalar element L nvv + t.
template <typename S, bool TAIL>
__device__ __forceinline__ int64_t tail_element(int64_t N, int64_t nvv)
{
    constexpr int L = V16<S>::L;
    return (TAIL && blockIdx.x == 0 && (int64_t)threadIdx.x < N - nvv * L) ? nvv * L + threadIdx.x : -1;
}
// VecMDot of w = V_{col+1} against v_0..v_col, written as HH(:, col) and the
// VecMAXPY coefficients -HH(:, col).
template <int EPT, int PV, typename S, bool TAIL>
__global__ __launch_bounds__(NT) void mdot_kernel(CgsArgs<S> a, int col)
{
    using V = typename V16<S>::type;
    constexpr int L = V16<S>::L;
    GmresState *g = a.g + blockIdx.y;
    a.basis += blockIdx.y * a.N;
    a.part += blockIdx.y * a.pstride;
    if (g->st.done) return;
    __shared__ double red[NW][PV];
    __shared__ double s_scale[PV];
    const int nvec = col + 1;
    if (threadIdx.x < nvec) s_scale[threadIdx.x] = g->scale[threadIdx.x];
    __syncthreads();
    const int64_t nvv = a.N / L;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    const V *W = reinterpret_cast<const V *>(a.basis + (size_t)nvec * a.vstride);
    const int64_t et = tail_element<S, TAIL>(a.N, nvv);
    const double wt = (TAIL && et >= 0) ? (double)a.basis[(size_t)nvec * a.vstride + et] : 0.0;
    V w[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        w[q] = (e < nvv) ? W[e] : V{};
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = 0;
    for (; v + 2 <= nvec; v += 2) {
        V x[2][EPT];
        load_group<EPT, 2, S>(a.basis, a.vstride, v, e0, nvv, x);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sv = s_scale[v + gg];
            double acc = dot_slice<EPT, S>(w, x[gg], sv);
            if (TAIL && et >= 0) acc = __builtin_fma(wt, (double)a.basis[(size_t)(v + gg) * a.vstride + et] * sv, acc);
            acc = wave_sum(acc);
            if (lane == 0) red[wave][v + gg] = acc;
        }
    }
    if (v < nvec) {
        V x[1][EPT];
        load_group<EPT, 1, S>(a.basis, a.vstride, v, e0, nvv, x);
        const double sv = s_scale[v];
        double acc = dot_slice<EPT, S>(w, x[0], sv);
        if (TAIL && et >= 0) acc = __builtin_fma(wt, (double)a.basis[(size_t)v * a.vstride + et] * sv, acc);
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

// CGS update  V_{it+1} <- w - sum_j h_j v_j  (VecMAXPY with lhh = -h) and
// per-block partial sums of its squared norm (of the stored values).
template <int EPT, int PV, typename S, bool TAIL>
__global__ __launch_bounds__(NT) void maxpy_norm_kernel(CgsArgs<S> a, int it)
{
    using V = typename V16<S>::type;
    constexpr int L = V16<S>::L;
    GmresState *g = a.g + blockIdx.y;
    a.basis += blockIdx.y * a.N;
    a.part += blockIdx.y * a.pstride;
    if (g->st.done) return;
    __shared__ double red[NW + 1];
    __shared__ double s_coef[kMaxRestart + 1], s_scale[kMaxRestart + 1];
    __shared__ double s_red[NT], s_h[PV];
    const int nvec = it + 1;
    const int64_t nvv = a.N / L;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    S *Ws = const_cast<S *>(a.basis) + (size_t)nvec * a.vstride;
    V *W = reinterpret_cast<V *>(Ws);
    const int64_t et = tail_element<S, TAIL>(a.N, nvv);
    double wt = (TAIL && et >= 0) ? (double)Ws[et] : 0.0;
    // w's loads are issued first, in flight while the dot partials are summed
    V wl[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        wl[q] = (e < nvv) ? W[e] : V{};
    }
    if (threadIdx.x < nvec) s_scale[threadIdx.x] = g->scale[threadIdx.x];
    // HH(:, it) = <w, v_j>: every workgroup sums the mdot partials in the same
    // order (so all hold the same values); workgroup 0 records the column
    final_sum_cols<PV>(a.part, a.nbd, nvec, s_h, s_red);
    if (threadIdx.x < nvec) {
        const double h = s_h[threadIdx.x];
        s_coef[threadIdx.x] = -h;            // VecMAXPY coefficients lhh = -HH(:, it)
        if (blockIdx.x == 0) {
            g->hh[it * (kMaxRestart + 1) + threadIdx.x] = h;
            g->coef[threadIdx.x] = -h;
        }
    }
    __syncthreads();
    double w[EPT][L];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
#pragma unroll
        for (int l = 0; l < L; ++l) w[q][l] = vget(wl[q], l);
    }
    int j = 0;
    for (; j + 2 <= nvec; j += 2) {
        V x[2][EPT];
        load_group<EPT, 2, S>(a.basis, a.vstride, j, e0, nvv, x);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sj = s_scale[j + gg], cj = s_coef[j + gg];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
#pragma unroll
                for (int l = 0; l < L; ++l) w[q][l] = w[q][l] + cj * (vget(x[gg][q], l) * sj);
            }
            if (TAIL && et >= 0) wt = wt + cj * ((double)a.basis[(size_t)(j + gg) * a.vstride + et] * sj);
        }
    }
    if (j < nvec) {
        V x[1][EPT];
        load_group<EPT, 1, S>(a.basis, a.vstride, j, e0, nvv, x);
        const double sj = s_scale[j], cj = s_coef[j];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
#pragma unroll
            for (int l = 0; l < L; ++l) w[q][l] = w[q][l] + cj * (vget(x[0][q], l) * sj);
        }
        if (TAIL && et >= 0) wt = wt + cj * ((double)a.basis[(size_t)j * a.vstride + et] * sj);
    }
    double acc = 0.0;
    if (TAIL && et >= 0) {
        const S o = (S)wt;
        Ws[et] = o;
        acc += (double)o * (double)o;
    }
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        if (e < nvv) {
            V o;
#pragma unroll
            for (int l = 0; l < L; ++l) vset(o, l, w[q][l]);
            W[e] = o;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const double r = vget(o, l);
                acc += r * r;
            }
        }
    }
    const double s = block_sum_all(acc, red);
    if (threadIdx.x == 0) a.part[a.noff + blockIdx.x] = s;
}

// VecNormalize of V_{it+1} from the maxpy partials, then
// KSPGMRESUpdateHessenberg (rotations, happy breakdown, residual estimate)
// and the convergence test; sets done when the cycle ends.  The Hessenberg
// column and the previous rotations are staged in LDS by the whole block
// (loads in flight together with the norm's partials), so the sequential
// rotation chain of thread 0 reads LDS, not dependent global loads.
// hout (optional, pinned host memory mapped for the device): the instance's
// status is also stored there, so the host polls it without a copy.
__device__ __forceinline__ void put_status(GmresStatus *hout, const GmresStatus &st)
{
    // per-field vector stores by one lane
    hout->it = st.it;
    hout->its = st.its;
    hout->reason = st.reason;
    hout->done = st.done;
    hout->active = st.active;
    hout->res = st.res;
    hout->rnorm0 = st.rnorm0;
}

__global__ __launch_bounds__(NT) void gmres_update_kernel(const double *part, int nb, int64_t pstride, int it,
                                                          GmresState *g, GmresStatus *hout)
{
    g += blockIdx.y;
    part += blockIdx.y * pstride;
    if (hout) hout += blockIdx.y;
    // every load of the block is issued before the early-exit test waits on
    // the flag's (reading the state of a finished cycle is harmless)
    const int32_t done = g->st.done;
    __shared__ double red[NW + 1];
    __shared__ double s_h[kMaxRestart + 2], s_c[kMaxRestart], s_s[kMaxRestart];
    double *hh = g->hh + (size_t)it * (kMaxRestart + 1);
    const double h_t = threadIdx.x <= it ? hh[threadIdx.x] : 0.0;
    const double c_t = threadIdx.x < it ? g->cc[threadIdx.x] : 0.0;
    const double s_t = threadIdx.x < it ? g->ss[threadIdx.x] : 0.0;
    const double grs_it = g->grs[it];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) acc += part[b];
    if (done) {                        // uniform over the block, before any barrier
        if (hout && threadIdx.x == 0) put_status(hout, g->st);
        return;
    }
    if (threadIdx.x <= it) s_h[threadIdx.x] = h_t;
    if (threadIdx.x < it) {
        s_c[threadIdx.x] = c_t;
        s_s[threadIdx.x] = s_t;
    }
    const double tot = block_sum_all(acc, red);        // barriers: the LDS stores above are visible
    if (threadIdx.x != 0) return;
    const double tt = sqrt(tot);
    s_h[it + 1] = tt;
    g->scale[it + 1] = tt != 0.0 ? 1.0 / tt : 1.0;
    // happy breakdown
    double hapbnd = fabs(tt / grs_it);
    if (hapbnd > g->haptol) hapbnd = g->haptol;
    const bool hapend = tt < hapbnd;
    // apply the previous rotations to the new column
    for (int q = 1; q <= it; ++q) {
        const double t0 = s_h[q - 1];
        s_h[q - 1] = s_c[q - 1] * t0 + s_s[q - 1] * s_h[q];
        s_h[q] = s_c[q - 1] * s_h[q] - s_s[q - 1] * t0;
    }
    double res;
    int reason = 0;
    if (!hapend) {
        const double d = sqrt(s_h[it] * s_h[it] + s_h[it + 1] * s_h[it + 1]);
        if (d == 0.0) {
            reason = VV_KSP_DIVERGED_BREAKDOWN;
            res = g->st.res;
        } else {
            const double cc = s_h[it] / d, ss = s_h[it + 1] / d;
            g->cc[it] = cc;
            g->ss[it] = ss;
            g->grs[it + 1] = -(ss * grs_it);
            g->grs[it] = cc * grs_it;
            s_h[it] = cc * s_h[it] + ss * s_h[it + 1];
            res = fabs(-(ss * grs_it));
        }
    } else {
        res = 0.0;
    }
    for (int q = 0; q <= it + 1; ++q) hh[q] = s_h[q];
    g->st.it = it + 1;
    g->st.its += 1;
    g->st.res = res;
    if (reason == 0) reason = ksp_test(g, res);
    if (hapend && reason == 0) reason = VV_KSP_DIVERGED_BREAKDOWN;
    if (reason == 0 && g->st.its >= g->max_it) reason = VV_KSP_DIVERGED_ITS;
    g->st.reason = reason;
    if (reason != 0 || g->st.it >= g->max_k || g->st.its >= g->max_it) g->st.done = 1;
    if (hout) put_status(hout, g->st);
}

// KSPGMRESBuildSoln: back substitution with the rotated Hessenberg matrix;
// the coefficients go to coef[0..it).  The triangle is staged in LDS first.
__global__ __launch_bounds__(NT) void backsolve_kernel(GmresState *g)
{
    g += blockIdx.y;
    __shared__ double sh[kMaxRestart][kMaxRestart];
    __shared__ double sg[kMaxRestart];
    __shared__ double y[kMaxRestart];     // in LDS: a runtime-indexed register array would live in scratch
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

// x += sum_j coef_j v_j  (VEC_TEMP = 0 + sum_j nrs_j v_j; x = x + VEC_TEMP);
// x is FP64 for either basis type.
template <int EPT, typename S, bool TAIL>
__global__ __launch_bounds__(NT) void update_kernel(double *__restrict__ x, CgsArgs<S> a)
{
    using V = typename V16<S>::type;
    constexpr int L = V16<S>::L;
    const GmresState *g = a.g + blockIdx.y;
    a.basis += blockIdx.y * a.N;
    x += blockIdx.y * a.N;
    const int n = g->st.it;
    if (n <= 0) return;
    __shared__ double s_coef[kMaxRestart + 1], s_scale[kMaxRestart + 1];
    if (threadIdx.x < n) {
        s_coef[threadIdx.x] = g->coef[threadIdx.x];
        s_scale[threadIdx.x] = g->scale[threadIdx.x];
    }
    __syncthreads();
    const int64_t nvv = a.N / L;
    const int64_t e0 = (int64_t)blockIdx.x * NT * EPT + threadIdx.x;
    const int64_t et = tail_element<S, TAIL>(a.N, nvv);
    double tt = 0.0;
    double t[EPT][L];
#pragma unroll
    for (int q = 0; q < EPT; ++q)
#pragma unroll
        for (int l = 0; l < L; ++l) t[q][l] = 0.0;
    int j = 0;
    for (; j + 2 <= n; j += 2) {
        V v[2][EPT];
        load_group<EPT, 2, S>(a.basis, a.vstride, j, e0, nvv, v);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            const double sj = s_scale[j + gg], cj = s_coef[j + gg];
#pragma unroll
            for (int q = 0; q < EPT; ++q) {
#pragma unroll
                for (int l = 0; l < L; ++l) t[q][l] = t[q][l] + cj * (vget(v[gg][q], l) * sj);
            }
            if (TAIL && et >= 0) tt = tt + cj * ((double)a.basis[(size_t)(j + gg) * a.vstride + et] * sj);
        }
    }
    if (j < n) {
        V v[1][EPT];
        load_group<EPT, 1, S>(a.basis, a.vstride, j, e0, nvv, v);
        const double sj = s_scale[j], cj = s_coef[j];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
#pragma unroll
            for (int l = 0; l < L; ++l) t[q][l] = t[q][l] + cj * (vget(v[0][q], l) * sj);
        }
        if (TAIL && et >= 0) tt = tt + cj * ((double)a.basis[(size_t)j * a.vstride + et] * sj);
    }
    if (TAIL && et >= 0) x[et] = x[et] + tt;
    double2 *X = reinterpret_cast<double2 *>(x);
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int64_t e = e0 + (int64_t)q * NT;
        if (e < nvv) {
#pragma unroll
            for (int h = 0; h < L / 2; ++h) {
                double2 xv = X[e * (L / 2) + h];
                xv.x = xv.x + t[q][2 * h];
                xv.y = xv.y + t[q][2 * h + 1];
                X[e * (L / 2) + h] = xv;
            }
        }
    }
}

// Cycle start of the FP32 GMRES cycle: V_0 = (float) r and per-block partial
// sums of the squares of the rounded values.
__global__ __launch_bounds__(NT) void cvt_partial_kernel(const double *__restrict__ r, float *__restrict__ v0,
                                                         int64_t len, int64_t bstride, double *partial,
                                                         int64_t pstride)
{
    __shared__ double red[NW + 1];
    r += blockIdx.y * len;
    v0 += blockIdx.y * bstride;
    partial += blockIdx.y * pstride;
    double acc = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT) {
        const double d = r[e];
        v0[e] = (float)d;
        acc += d * d;          // ||r|| of the FP64 right-hand side
    }
    const double s = block_sum_all(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y = y + alpha*x (VecAXPY).  mode 1: only when the GMRES reason is positive
// (SNESSolve_KSPONLY takes the step unless the linear solve diverged).
__global__ __launch_bounds__(NT) void axpy_kernel(double alpha, const double *__restrict__ x, double *y, int64_t len,
                                                  const GmresState *g, int mode)
{
    x += blockIdx.y * len;
    y += blockIdx.y * len;
    if (mode == 1 && g[blockIdx.y].st.reason <= 0) return;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        y[e] = y[e] + alpha * x[e];
}

// y = b - y   (KSPInitialResidual after a restart: r = b - A x); instance
// blockIdx.y at b + blockIdx.y len, y + blockIdx.y ystride
__global__ __launch_bounds__(NT) void rsub_kernel(const double *__restrict__ b, double *y, int64_t len,
                                                  int64_t ystride)
{
    b += blockIdx.y * len;
    y += blockIdx.y * ystride;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT)
        y[e] = b[e] + (-1.0) * y[e];
}

__global__ __launch_bounds__(NT) void copy_kernel(const double *__restrict__ x, double *__restrict__ y, int64_t len)
{
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT) y[e] = x[e];
}

// y[b][j][i] += alpha w[b][i] (an x-only field added to every v-row)
__global__ __launch_bounds__(NT) void xadd_kernel(double *__restrict__ y, const double *__restrict__ w, int nx,
                                                  int64_t N, double alpha)
{
    y += blockIdx.y * N;
    w += (size_t)blockIdx.y * nx;
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < N; e += (int64_t)gridDim.x * NT)
        y[e] = y[e] + alpha * w[e % nx];
}

// y = 0 (a kernel rather than a memset node, so captured graphs replay it)
__global__ __launch_bounds__(NT) void zero_kernel(double *__restrict__ y, int64_t len)
{
    for (int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x; e < len; e += (int64_t)gridDim.x * NT) y[e] = 0.0;
}

int grid_for(int64_t len)
{
    const int64_t want = (len + NT - 1) / NT;
    return (int)(want < kReduceBlocks ? (want > 0 ? want : 1) : kReduceBlocks);
}

// 16-byte vectors per thread of the CGS kernels: the context's tuned value,
// else the largest of 8,4,2,1 that still gives >= 256 workgroups (nvv:
// 16-byte vectors per basis vector)
int cgs_ept_v(const vv_ctx *c, int64_t nvv)
{
    if (c != nullptr && c->tune_cgs_ept > 0) return c->tune_cgs_ept;
    for (int ept = 8; ept > 1; ept >>= 1)
        if ((nvv + (int64_t)NT * ept - 1) / ((int64_t)NT * ept) >= 256) return ept;
    return 1;
}

int cgs_blocks_v(const vv_ctx *c, int64_t nvv)
{
    const int ept = cgs_ept_v(c, nvv);
    return (int)((nvv + (int64_t)NT * ept - 1) / ((int64_t)NT * ept));
}

// the maxpy norm partials sit after the dot partials (rows x kMaxRestart) in
// each instance's part_stride = rows x (kMaxRestart + 2) doubles
int64_t norm_partial_offset(const vv_ctx *c) { return c->part_stride / (kMaxRestart + 2) * kMaxRestart; }

}  // namespace

int cgs_ept(const vv_ctx *c, int64_t nvv) { return cgs_ept_v(c, nvv); }

// partial-sum rows the CGS kernels can need (one vector per thread)
int cgs_blocks_max(int64_t N) { return (int)(((N >> 1) + NT - 1) / NT); }

hipError_t launch_gmres_init(vv_ctx *c, double rtol, double atol, double dtol, int max_it)
{
    hipLaunchKernelGGL(gmres_init_kernel, dim3(1, c->B), dim3(64), 0, c->stream, c->gs, c->active, rtol, atol, dtol,
                       max_it, c->p.restart);
    return hipGetLastError();
}

hipError_t launch_norm_start(vv_ctx *c)
{
    const int nb = grid_for(c->N);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb, c->B), dim3(NT), 0, c->stream, c->basis, c->basis, c->N, c->N,
                       c->partial, c->part_stride);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, c->partial, nb, c->part_stride,
                       c->gs, (double *)nullptr, 1);
    return hipGetLastError();
}

namespace {

template <int E, typename S, bool TAIL>
void cgs_launch(vv_ctx *c, const S *basis, int which, int arg, double *x, int dot_rows)
{
    const int nb = cgs_blocks_v(c, c->N / V16<S>::L);
    CgsArgs<S> a{basis, c->N, c->vs, c->partial, c->part_stride, c->gs, norm_partial_offset(c),
                 dot_rows > 0 ? dot_rows : nb};
    const dim3 grid(nb, c->B), block(NT);
    if (which == 0) {
        if (c->p.restart < 32) launch_timed(c, mdot_kernel<E, 32, S, TAIL>, grid, block, 0, a, arg);
        else launch_timed(c, mdot_kernel<E, 64, S, TAIL>, grid, block, 0, a, arg);
    } else if (which == 1) {
        if (c->p.restart < 32) launch_timed(c, maxpy_norm_kernel<E, 32, S, TAIL>, grid, block, 0, a, arg);
        else launch_timed(c, maxpy_norm_kernel<E, 64, S, TAIL>, grid, block, 0, a, arg);
    } else {
        launch_timed(c, update_kernel<E, S, TAIL>, grid, block, 0, x, a);
    }
}

template <typename S, bool TAIL>
void cgs_dispatch_e(vv_ctx *c, const S *basis, int which, int arg, double *x, int dot_rows)
{
    switch (cgs_ept_v(c, c->N / V16<S>::L)) {
    case 8: cgs_launch<8, S, TAIL>(c, basis, which, arg, x, dot_rows); break;
    case 4: cgs_launch<4, S, TAIL>(c, basis, which, arg, x, dot_rows); break;
    case 2: cgs_launch<2, S, TAIL>(c, basis, which, arg, x, dot_rows); break;
    default: cgs_launch<1, S, TAIL>(c, basis, which, arg, x, dot_rows); break;
    }
}

template <typename S>
hipError_t cgs_dispatch(vv_ctx *c, const S *basis, int which, int arg, double *x, int dot_rows = 0)
{
    if (c->N % V16<S>::L) cgs_dispatch_e<S, true>(c, basis, which, arg, x, dot_rows);
    else cgs_dispatch_e<S, false>(c, basis, which, arg, x, dot_rows);
    return hipGetLastError();
}

// CGS dots of w with V_0..V_col: per-workgroup partials only; their column
// sums are formed by every workgroup of the following maxpy_norm launch
template <typename S>
hipError_t mdot_t(vv_ctx *c, const S *basis, int col)
{
    return cgs_dispatch(c, basis, 0, col, nullptr);
}

template <typename S>
hipError_t maxpy_norm_t(vv_ctx *c, const S *basis, int it, int dot_rows)
{
    hipError_t e = cgs_dispatch(c, basis, 1, it, nullptr, dot_rows);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gmres_update_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, c->partial + norm_partial_offset(c),
                       cgs_blocks_v(c, c->N / V16<S>::L), c->part_stride, it, c->gs, c->status_out);
    return hipGetLastError();
}

template <typename S>
hipError_t build_solution_t(vv_ctx *c, const S *basis, double *x)
{
    hipLaunchKernelGGL(backsolve_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, c->gs);
    return cgs_dispatch(c, basis, 2, 0, x);
}

}  // namespace

hipError_t launch_mdot(vv_ctx *c, int col) { return mdot_t<double>(c, c->basis, col); }
hipError_t launch_maxpy_norm(vv_ctx *c, int it, int dot_rows)
{
    return maxpy_norm_t<double>(c, c->basis, it, dot_rows);
}
hipError_t launch_build_solution(vv_ctx *c, double *x) { return build_solution_t<double>(c, c->basis, x); }
hipError_t launch_mdot32(vv_ctx *c, int col) { return mdot_t<float>(c, c->basis32, col); }
hipError_t launch_maxpy_norm32(vv_ctx *c, int it, int dot_rows)
{
    return maxpy_norm_t<float>(c, c->basis32, it, dot_rows);
}
hipError_t launch_build_solution32(vv_ctx *c, double *x) { return build_solution_t<float>(c, c->basis32, x); }

hipError_t launch_start32(vv_ctx *c, const double *r)
{
    const int nb = grid_for(c->N);
    hipLaunchKernelGGL(cvt_partial_kernel, dim3(nb, c->B), dim3(NT), 0, c->stream, r, c->basis32, c->N, c->N,
                       c->partial, c->part_stride);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, c->partial, nb, c->part_stride,
                       c->gs, (double *)nullptr, 1);
    return hipGetLastError();
}

hipError_t launch_residual_restart(vv_ctx *c, const double *b, double *y)
{
    hipLaunchKernelGGL(rsub_kernel, dim3(grid_for(c->N), c->B), dim3(NT), 0, c->stream, b, y, c->N, c->N);
    return hipGetLastError();
}

hipError_t launch_norms(vv_ctx *c, const double *x, double *out_dev)
{
    const int nb = grid_for(c->N);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb, c->B), dim3(NT), 0, c->stream, x, x, c->N, c->N, c->partial,
                       c->part_stride);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1, c->B), dim3(NT), 0, c->stream, c->partial, nb, c->part_stride,
                       c->gs, out_dev, 0);
    return hipGetLastError();
}

hipError_t launch_norm2(vv_ctx *c, const double *x, int64_t n, double *out_dev)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(NT), 0, c->stream, x, x, n, (int64_t)0, c->partial,
                       c->part_stride);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, nb, c->part_stride, c->gs,
                       out_dev, 0);
    return hipGetLastError();
}

hipError_t launch_dot(vv_ctx *c, const double *x, const double *y, int64_t n, double *out_dev)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb), dim3(NT), 0, c->stream, x, y, n, (int64_t)0, c->partial,
                       c->part_stride);
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(NT), 0, c->stream, c->partial, nb, c->part_stride, c->gs,
                       out_dev, 2);
    return hipGetLastError();
}

hipError_t launch_axpy(vv_ctx *c, double alpha, const double *x, double *y, int64_t n)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, alpha, x, y, n, c->gs, 0);
    return hipGetLastError();
}

hipError_t launch_axpy_cond(vv_ctx *c, double alpha, const double *x, double *y)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(c->N), c->B), dim3(NT), 0, c->stream, alpha, x, y, c->N, c->gs, 1);
    return hipGetLastError();
}

hipError_t launch_copy(vv_ctx *c, const double *x, double *y, int64_t n)
{
    hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, x, y, n);
    return hipGetLastError();
}

hipError_t launch_xadd(vv_ctx *c, double *y, const double *w, double alpha)
{
    hipLaunchKernelGGL(xadd_kernel, dim3(grid_for(c->N), c->B), dim3(NT), 0, c->stream, y, w, c->p.nx, c->N, alpha);
    return hipGetLastError();
}

hipError_t launch_zero(vv_ctx *c, double *y, int64_t n)
{
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n)), dim3(NT), 0, c->stream, y, n);
    return hipGetLastError();
}

}  // namespace vv
