/*
 * vv1d_oracle.c -- CPU restatement (C99 + OpenMP) of the viVlasov1D implicit
 * timestep.  TEST INFRASTRUCTURE: used only as the checker in tests/ and as
 * the host-core CPU baseline timed by bench.py; never by the product.
 *
 * Same algorithm as oracle/vlasov_oracle.py, restated from the reference:
 *   brackets         vlasov/solvers/components/PoissonBracket.pyx:80-113,
 *                    200-234, 335-389 (unfused; build with -ffp-contract=off)
 *   residual / J u   vlasov/solvers/vlasov/PETScNLVlasovMP.pyx:33-58
 *   H_ave            vlasov/solvers/vlasov/PETScVlasovSolver.pyx:142-157
 *   preconditioner   vlasov/solvers/preconditioner/TensorProductKinetic.pyx:146-160,
 *                    TensorProductKineticFast.pyx:61-90, 102-272 (FFTW and
 *                    LAPACK zgbtrf/zgbtrs replaced by an iterative radix-2 FFT
 *                    -- a direct DFT for lengths that are not powers of two --
 *                    and a pivot-free banded LU: the reference's LU never pivots
 *                    for these diagonally dominant T_k)
 *   Poisson          vlasov/solvers/poisson/PETScPoissonSolver.pyx:58-74,
 *                    PETScPoissonCFD4.pyx:64-100 (exact mean-free circulant solve)
 *   moments          vlasov/toolbox/Toolbox.pyx:38-98
 *   GMRES            PETSc KSPGMRES defaults: restart, CGS, Givens, x0 = 0
 *   Picard loop      vlasov/run/run_nonlinear_matrixfree_split_mp.py:111-174,
 *                    vlasov/run/run_base.py:631-763
 *
 * Layout: natural order, element (i, j) at i + nx*j.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

typedef struct {
    int nx, nv, bracket, laplace, restart, nmodes, logn;
    double hx, hv, ht, charge, ht_inv, hx_inv, hv_inv, hx2_inv, af1, af2;
    double ksp_rtol, ksp_atol, snes_atol;
    int ksp_max_it, snes_max_iter;
    double *h0, *h1c, *h1h, *h2c, *h2h;   /* Hamiltonian factors            */
    double *hp1, *hp2, *hp3, *hp4;        /* H_ave factors held by the solver */
    double *fc, *fh, *fl, *fb, *X, *D;
    double *n, *phi, *pinv;
    cplx *tw;                             /* exp(-2 pi i m / nx)            */
    cplx *lu_c, *lu_r, *lu_a;             /* per mode: super c_j, 1/pivot, sub a_j */
    cplx *spec;                           /* nmodes x nv                    */
    double *V;                            /* (restart+1) x N Krylov basis   */
    double pnorm;
    long N;
} vvo_run;

/* ------------------------------------------------------------------ brackets */
static inline double Hat(const vvo_run *r, int i, int j)
{
    return (((r->h0[j] + 0.5 * r->hp1[i]) + 0.5 * r->hp2[i]) + 0.5 * r->hp3[i]) + 0.5 * r->hp4[i];
}

/* y (overwritten): mode 0 residual F(f) with fh; mode 1 Jacobian J u */
static void bracket_op(const vvo_run *r, int mode, const double *f, const double *fh, double *y)
{
    const int nx = r->nx, nv = r->nv, kind = r->bracket;
    const double fac = mode == 0 ? 1.0 : 0.5;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nv; ++j) {
        for (int i = 0; i < nx; ++i) {
            const long o = (long)j * nx + i;
            double base = mode == 0 ? 0.0 + r->ht_inv * (f[o] - fh[o]) : 0.0 + r->ht_inv * f[o];
            if (j < 2 || j >= nv - 2) {
                y[o] = base;
                continue;
            }
            const int im1 = (i - 1 + nx) % nx, ip1 = (i + 1) % nx, im2 = (i - 2 + nx) % nx, ip2 = (i + 2) % nx;
#define XV(ii, jj) (mode == 0 ? 0.5 * fh[(long)(jj) * nx + (ii)] + 0.5 * f[(long)(jj) * nx + (ii)] : f[(long)(jj) * nx + (ii)])
            const double x_pc = XV(ip1, j), x_mc = XV(im1, j), x_cp = XV(i, j + 1), x_cm = XV(i, j - 1);
            const double x_pp = XV(ip1, j + 1), x_mm = XV(im1, j - 1), x_mp = XV(im1, j + 1), x_pm = XV(ip1, j - 1);
            const double h_cp = Hat(r, i, j + 1), h_cm = Hat(r, i, j - 1), h_pc = Hat(r, ip1, j), h_mc = Hat(r, im1, j);
            const double h_pp = Hat(r, ip1, j + 1), h_pm = Hat(r, ip1, j - 1), h_mp = Hat(r, im1, j + 1),
                         h_mm = Hat(r, im1, j - 1);
            double out = base;
            double s1 = 0.0, s2 = 0.0;
            if (kind == 1 || kind == 4) {
                const double jpp = (x_pc - x_mc) * (h_cp - h_cm) - (x_cp - x_cm) * (h_pc - h_mc);
                const double jpc = x_pc * (h_pp - h_pm) - x_mc * (h_mp - h_mm) - x_cp * (h_pp - h_mp) + x_cm * (h_pm - h_mm);
                const double jcp = x_pp * (h_cp - h_pc) - x_mm * (h_mc - h_cm) - x_mp * (h_cp - h_mc) + x_pm * (h_pc - h_cm);
                s1 = jpp + jpc + jcp;
            }
            if (kind == 2 || kind == 4) {
                const double x_2c = XV(ip2, j), x_m2 = XV(im2, j), x_c2 = XV(i, j + 2), x_cm2 = XV(i, j - 2);
                const double h_c2 = Hat(r, i, j + 2), h_2c = Hat(r, ip2, j), h_m2c = Hat(r, im2, j), h_cm2 = Hat(r, i, j - 2);
                const double jcc = (x_pp - x_mm) * (h_mp - h_pm) - (x_mp - x_pm) * (h_pp - h_mm);
                const double jpc = x_2c * (h_pp - h_pm) - x_m2 * (h_mp - h_mm) - x_c2 * (h_pp - h_mp) + x_cm2 * (h_pm - h_mm);
                const double jcp = x_pp * (h_c2 - h_2c) - x_mm * (h_m2c - h_cm2) - x_mp * (h_c2 - h_m2c) + x_pm * (h_2c - h_cm2);
                s2 = jcc + jpc + jcp;
            }
#undef XV
            if (kind == 1) out = base + fac * s1 * r->af1;
            else if (kind == 2) out = base + fac * s2 * r->af2;
            else out = base + fac * (2. * (s1 * r->af1) - s2 * r->af2);
            y[o] = out;
        }
    }
}

/* ----------------------------------------------------------- preconditioner */
/* z <- sum_m z_m w^(k m mod n), w = exp(-+2 pi i/n): the textbook DFT, for
 * lengths that are not powers of two (work: one n-vector of scratch) */
static void dft_inplace(cplx *z, int n, const cplx *tw, int inverse)
{
    cplx *t = malloc(sizeof(cplx) * n);
    for (int k = 0; k < n; ++k) {
        cplx s = 0.0;
        for (int m = 0, q = 0; m < n; ++m) {
            s += z[m] * (inverse ? conj(tw[q]) : tw[q]);
            q += k;
            if (q >= n) q -= n;
        }
        t[k] = s;
    }
    memcpy(z, t, sizeof(cplx) * n);
    free(t);
}

static void fft_inplace(cplx *z, int n, int logn, const cplx *tw, int inverse)
{
    if (n & (n - 1)) {
        dft_inplace(z, n, tw, inverse);
        return;
    }
    for (int i = 1, j = 0; i < n; ++i) {          /* bit reversal */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cplx t = z[i];
            z[i] = z[j];
            z[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; ++k) {
                cplx w = tw[k * step];
                if (inverse) w = conj(w);
                const cplx a = z[s + k], b = z[s + k + half] * w;
                z[s + k] = a + b;
                z[s + k + half] = a - b;
            }
    }
    (void)logn;
}

static void precond(vvo_run *r, const double *x, double *y)
{
    const int nx = r->nx, nv = r->nv, nm = r->nmodes, nh = nx / 2;
    const int nyq = (nx & 1) ? -1 : nh;       /* the real Nyquist mode of an even length */
    /* rfft of row pairs */
#pragma omp parallel
    {
        cplx *z = malloc(sizeof(cplx) * nx);
#pragma omp for schedule(static)
        for (int j = 0; j < nv; j += 2) {
            const int jb = j + 1 < nv ? j + 1 : -1;
            for (int i = 0; i < nx; ++i) z[i] = x[(long)j * nx + i] + I * (jb >= 0 ? x[(long)jb * nx + i] : 0.0);
            fft_inplace(z, nx, r->logn, r->tw, 0);
            for (int k = 0; k <= nh; ++k) {
                const cplx zk = z[k], zm = conj(z[k == 0 ? 0 : nx - k]);
                r->spec[(long)k * nv + j] = 0.5 * (zk + zm);
                if (jb >= 0) r->spec[(long)k * nv + jb] = -0.5 * I * (zk - zm);
            }
        }
        free(z);
    }
    /* T_k c = b per mode: forward / backward substitution with the stored LU */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nm; ++k) {
        cplx *b = r->spec + (long)k * nv;
        const cplx *c = r->lu_c + (long)k * nv, *rp = r->lu_r + (long)k * nv, *a = r->lu_a + (long)k * nv;
        b[0] = b[0] * rp[0];
        for (int j = 1; j < nv; ++j) b[j] = (b[j] - a[j] * b[j - 1]) * rp[j];
        for (int j = nv - 2; j >= 0; --j) b[j] = b[j] - c[j] * b[j + 1];
    }
    /* irfft of row pairs */
#pragma omp parallel
    {
        cplx *z = malloc(sizeof(cplx) * nx);
#pragma omp for schedule(static)
        for (int j = 0; j < nv; j += 2) {
            const int jb = j + 1 < nv ? j + 1 : -1;
            for (int k = 0; k <= nh; ++k) {
                cplx A = r->spec[(long)k * nv + j];
                cplx B = jb >= 0 ? r->spec[(long)k * nv + jb] : 0.0;
                if (k == 0 || k == nyq) {
                    A = creal(A);
                    B = creal(B);
                }
                z[k] = A + I * B;
                if (k != 0 && k != nyq) z[nx - k] = conj(A) + I * conj(B);
            }
            fft_inplace(z, nx, r->logn, r->tw, 1);
            const double s = 1.0 / nx;
            for (int i = 0; i < nx; ++i) {
                y[(long)j * nx + i] = creal(z[i]) * s;
                if (jb >= 0) y[(long)jb * nx + i] = cimag(z[i]) * s;
            }
        }
        free(z);
    }
}

/* ------------------------------------------------------------- BLAS-1 / CGS */
static double dot(const double *a, const double *b, long n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long e = 0; e < n; ++e) s += a[e] * b[e];
    return s;
}

static void apply_A(vvo_run *r, const double *u, double *w)
{
    bracket_op(r, 1, u, NULL, r->X);
    precond(r, r->X, w);
}

/* GMRES(restart) on P^-1 J from x = 0; returns its, sets *reason */
static int gmres(vvo_run *r, const double *b, double *x, int *reason)
{
    const long N = r->N;
    const int m = r->restart;
    double *H = calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = calloc(m, sizeof(double)), *sn = calloc(m, sizeof(double)), *g = calloc(m + 1, sizeof(double));
    double *y = calloc(m, sizeof(double)), *hcol = calloc(m + 1, sizeof(double));
    memset(x, 0, sizeof(double) * N);
    int its = 0, first = 1;
    double rnorm0 = 0.0, ttol = 0.0;
    *reason = 0;
    for (;;) {
        double *v0 = r->V;
        if (first) {
            memcpy(v0, b, sizeof(double) * N);
        } else {
            apply_A(r, x, v0);
#pragma omp parallel for schedule(static)
            for (long e = 0; e < N; ++e) v0[e] = b[e] - v0[e];
        }
        const double beta = sqrt(dot(v0, v0, N));
        if (beta == 0.0) {
            *reason = 3;
            break;
        }
        if (first) {
            rnorm0 = beta;
            ttol = fmax(r->ksp_rtol * beta, r->ksp_atol);
        }
        first = 0;
        if (beta <= ttol) {
            *reason = beta < r->ksp_atol ? 3 : 2;
            break;
        }
        {
            const double s = 1.0 / beta;
#pragma omp parallel for schedule(static)
            for (long e = 0; e < N; ++e) v0[e] *= s;
        }
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int it = 0;
        while (*reason == 0 && it < m && its < r->ksp_max_it) {
            double *w = r->V + (long)(it + 1) * N;
            apply_A(r, r->V + (long)it * N, w);
            for (int j = 0; j <= it; ++j) hcol[j] = dot(w, r->V + (long)j * N, N);   /* CGS: dots first */
#pragma omp parallel for schedule(static)
            for (long e = 0; e < N; ++e) {
                double t = w[e];
                for (int j = 0; j <= it; ++j) t = t + (-hcol[j]) * r->V[(long)j * N + e];
                w[e] = t;
            }
            const double tt = sqrt(dot(w, w, N));
            if (tt != 0.0) {
                const double s = 1.0 / tt;
#pragma omp parallel for schedule(static)
                for (long e = 0; e < N; ++e) w[e] *= s;
            }
            double *h = H + (long)it * (m + 1);
            for (int j = 0; j <= it; ++j) h[j] = hcol[j];
            h[it + 1] = tt;
            double hapbnd = fabs(tt / g[it]);
            if (hapbnd > 1e-30) hapbnd = 1e-30;
            const int hapend = tt < hapbnd;
            for (int j = 1; j <= it; ++j) {
                const double t0 = h[j - 1];
                h[j - 1] = cs[j - 1] * t0 + sn[j - 1] * h[j];
                h[j] = cs[j - 1] * h[j] - sn[j - 1] * t0;
            }
            double res = 0.0;
            if (!hapend) {
                const double d = sqrt(h[it] * h[it] + h[it + 1] * h[it + 1]);
                cs[it] = h[it] / d;
                sn[it] = h[it + 1] / d;
                g[it + 1] = -(sn[it] * g[it]);
                g[it] = cs[it] * g[it];
                h[it] = cs[it] * h[it] + sn[it] * h[it + 1];
                res = fabs(g[it + 1]);
            }
            ++it;
            ++its;
            if (!isfinite(res)) *reason = -9;
            else if (res <= ttol) *reason = res < r->ksp_atol ? 3 : 2;
            else if (res >= 1e5 * rnorm0) *reason = -4;
            if (hapend && *reason == 0) *reason = -5;
        }
        if (it > 0) {   /* back substitution and x += V y */
            for (int k = it - 1; k >= 0; --k) {
                double t = g[k];
                for (int j = k + 1; j < it; ++j) t -= H[(long)j * (m + 1) + k] * y[j];
                y[k] = t / H[(long)k * (m + 1) + k];
            }
#pragma omp parallel for schedule(static)
            for (long e = 0; e < N; ++e) {
                double t = 0.0;
                for (int j = 0; j < it; ++j) t = t + y[j] * r->V[(long)j * N + e];
                x[e] = x[e] + t;
            }
        }
        if (*reason != 0) break;
        if (its >= r->ksp_max_it) {
            *reason = -3;
            break;
        }
    }
    free(H);
    free(cs);
    free(sn);
    free(g);
    free(y);
    free(hcol);
    return its;
}

/* ----------------------------------------------------------- moments/Poisson */
static void moments(vvo_run *r)
{
    const int nx = r->nx, nv = r->nv;
    double *b = malloc(sizeof(double) * nx);
    for (int i = 0; i < nx; ++i) {
        double s = 0.0;
        for (int j = 0; j < nv; ++j) s += r->fc[(long)j * nx + i];
        r->n[i] = s * r->hv;
    }
    double ns = 0.0;
    for (int i = 0; i < nx; ++i) ns += r->n[i];
    const double nmean = ns / nx;
    for (int i = 0; i < nx; ++i) b[i] = -(r->n[i] - nmean) * r->charge;
    double ps = 0.0;
    for (int i = 0; i < nx; ++i) {
        /* phi_i = sum_m pinv_m b_{(i-m) mod nx} in the device kernel's order:
           four quarters of the m range, each with four interleaved partial
           sums, the quarters added in order (a summation order of this
           restatement's own; the reference solves with a PETSc KSP) */
        const int q = (nx + 3) / 4;
        double p[4];
        for (int w = 0; w < 4; ++w) {
            const int m0 = w * q, m1 = m0 + q < nx ? m0 + q : nx;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int m = m0;
            for (; m + 4 <= m1; m += 4) {
                s0 += r->pinv[m] * b[(i - m + nx) % nx];
                s1 += r->pinv[m + 1] * b[(i - m - 1 + nx) % nx];
                s2 += r->pinv[m + 2] * b[(i - m - 2 + nx) % nx];
                s3 += r->pinv[m + 3] * b[(i - m - 3 + nx) % nx];
            }
            for (; m < m1; ++m) s0 += r->pinv[m] * b[(i - m + nx) % nx];
            p[w] = (s0 + s1) + (s2 + s3);
        }
        const double s = ((p[0] + p[1]) + p[2]) + p[3];
        r->phi[i] = s;
        ps += s;
    }
    const double pave = ps / nx;
    double rr = 0.0;
    for (int i = 0; i < nx; ++i) {
        r->h1c[i] = r->phi[i] - pave;
        const double *x = r->phi;
        const double xm2 = x[(i - 2 + nx) % nx], xm1 = x[(i - 1 + nx) % nx], xp1 = x[(i + 1) % nx], xp2 = x[(i + 2) % nx];
        double ax = r->laplace == 4 ? (5. * xm2 - 32. * xm1 + 54. * x[i] - 32. * xp1 + 5. * xp2) * r->hx2_inv / 12.
                                    : (2. * x[i] - xm1 - xp1) * r->hx2_inv;
        const double res = ax + (r->n[i] - nmean) * r->charge;
        rr += res * res;
    }
    r->pnorm = sqrt(rr);
    free(b);
}

static void update_previous(vvo_run *r)
{
    memcpy(r->hp1, r->h1c, sizeof(double) * r->nx);
    memcpy(r->hp2, r->h1h, sizeof(double) * r->nx);
    memcpy(r->hp3, r->h2c, sizeof(double) * r->nx);
    memcpy(r->hp4, r->h2h, sizeof(double) * r->nx);
}

static double calc_residual(vvo_run *r)
{
    bracket_op(r, 0, r->fc, r->fh, r->fb);
    return sqrt(dot(r->fb, r->fb, r->N)) + r->pnorm;
}

/* ------------------------------------------------------------------- public */
void vvo_destroy(vvo_run *r);

vvo_run *vvo_create(int nx, int nv, double hx, double hv, double ht, int bracket, int laplace, double charge,
                    const double *v, const double *eigen, const double *f0, const double *h0, double ksp_rtol,
                    double ksp_atol, int ksp_max_it, double snes_atol, int snes_max_iter, int restart, int nthreads)
{
    if (nx < 4 || nv < 8 || restart < 1) return NULL;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    vvo_run *r = calloc(1, sizeof(vvo_run));
    r->nx = nx;
    r->nv = nv;
    r->N = (long)nx * nv;
    r->nmodes = nx / 2 + 1;
    for (r->logn = 0; (1 << r->logn) < nx; ++r->logn) {
    }
    r->hx = hx;
    r->hv = hv;
    r->ht = ht;
    r->bracket = bracket;
    r->laplace = laplace;
    r->charge = charge;
    r->restart = restart;
    r->ht_inv = 1. / ht;
    r->hx_inv = 1. / hx;
    r->hv_inv = 1. / hv;
    r->hx2_inv = 1. / (hx * hx);
    r->af1 = r->hx_inv * r->hv_inv / 12.;
    r->af2 = r->hx_inv * r->hv_inv / 24.;
    r->ksp_rtol = ksp_rtol;
    r->ksp_atol = ksp_atol;
    r->ksp_max_it = ksp_max_it;
    r->snes_atol = snes_atol;
    r->snes_max_iter = snes_max_iter;
    const long N = r->N;
    double **xs[] = {&r->h1c, &r->h1h, &r->h2c, &r->h2h, &r->hp1, &r->hp2, &r->hp3, &r->hp4, &r->n, &r->phi, &r->pinv};
    for (unsigned q = 0; q < sizeof(xs) / sizeof(xs[0]); ++q) *xs[q] = calloc(nx, sizeof(double));
    r->h0 = malloc(sizeof(double) * nv);
    memcpy(r->h0, h0, sizeof(double) * nv);
    double **fs[] = {&r->fc, &r->fh, &r->fl, &r->fb, &r->X, &r->D};
    for (unsigned q = 0; q < sizeof(fs) / sizeof(fs[0]); ++q) *fs[q] = calloc(N, sizeof(double));
    r->V = malloc(sizeof(double) * N * (restart + 1));
    r->spec = malloc(sizeof(cplx) * r->nmodes * nv);
    r->tw = malloc(sizeof(cplx) * nx);
    for (int m = 0; m < nx; ++m) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / nx;
        r->tw[m] = (double)cosl(ang) + I * (double)sinl(ang);
    }
    /* band LU of T_k (KineticFast.pyx:248-272), no pivoting */
    r->lu_c = malloc(sizeof(cplx) * r->nmodes * nv);
    r->lu_r = malloc(sizeof(cplx) * r->nmodes * nv);
    r->lu_a = malloc(sizeof(cplx) * r->nmodes * nv);
    const double af = 0.5 * r->hx_inv * r->hv_inv / 12.;
    const double hv2 = hv * hv;
    for (int k = 0; k < r->nmodes; ++k) {
        const double er = eigen[2 * k], ei = eigen[2 * k + 1];
        cplx *c = r->lu_c + (long)k * nv, *rp = r->lu_r + (long)k * nv, *a = r->lu_a + (long)k * nv;
        cplx cprev = 0.0;
        for (int j = 0; j < nv; ++j) {
            const double gm = 2. * hv * v[j] - hv2, g0 = 4. * hv * v[j], gp = 2. * hv * v[j] + hv2;
            const cplx aj = j > 0 ? ((er * 0.5) * gm * af) + I * ((ei * 0.5) * gm * af) : 0.0;
            const cplx bj = ((er * 1.0) * g0 * af + r->ht_inv) + I * ((ei * 1.0) * g0 * af);
            const cplx cj = j < nv - 1 ? ((er * 0.5) * gp * af) + I * ((ei * 0.5) * gp * af) : 0.0;
            const cplx piv = bj - aj * cprev;
            rp[j] = 1.0 / piv;
            c[j] = cj * rp[j];
            a[j] = aj;
            cprev = c[j];
        }
    }
    /* mean-free circulant pseudo-inverse of the periodic Laplacian */
    {
        long double *ct = malloc(sizeof(long double) * nx), *lam = malloc(sizeof(long double) * nx);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int q = 0; q < nx; ++q) ct[q] = cosl(2.0L * pi * q / nx);
        for (int k = 1; k < nx; ++k) {
            const long double s1 = sinl(pi * k / nx), s2 = sinl(2.0L * pi * k / nx);
            lam[k] = laplace == 4 ? (128.0L * s1 * s1 - 20.0L * s2 * s2) / (12.0L * hx * hx) : 4.0L * s1 * s1 / (hx * hx);
        }
        for (int m = 0; m < nx; ++m) {
            long double s = 0.0L;
            for (int k = 1; k < nx; ++k) s += ct[((long)k * m) % nx] / lam[k];
            r->pinv[m] = (double)(s / nx);
        }
        free(ct);
        free(lam);
    }
    memcpy(r->fc, f0, sizeof(double) * N);
    moments(r);
    return r;
}

void vvo_destroy(vvo_run *r)
{
    if (!r) return;
    void *ps[] = {r->h0, r->h1c, r->h1h, r->h2c, r->h2h, r->hp1, r->hp2, r->hp3, r->hp4, r->fc, r->fh, r->fl,
                  r->fb, r->X, r->D, r->n, r->phi, r->pinv, r->tw, r->lu_c, r->lu_r, r->lu_a, r->spec, r->V};
    for (unsigned q = 0; q < sizeof(ps) / sizeof(ps[0]); ++q) free(ps[q]);
    free(r);
}

/* One timestep; stats (3 + snes_max_iter entries): [0] Picard iterations, [1] GMRES
 * iterations, [2] final residual, [3 + q] the residual after Picard iteration q. */
int vvo_step(vvo_run *r, double *stats)
{
    const long N = r->N;
    memcpy(r->fh, r->fc, sizeof(double) * N);
    memcpy(r->h1h, r->h1c, sizeof(double) * r->nx);
    memcpy(r->h2h, r->h2c, sizeof(double) * r->nx);
    update_previous(r);
    double pred = calc_residual(r);
    int i = 0, gm = 0;
    for (;;) {
        ++i;
        memcpy(r->fl, r->fc, sizeof(double) * N);
        bracket_op(r, 0, r->fc, r->fh, r->X);
        precond(r, r->X, r->fb);
        double *b = malloc(sizeof(double) * N);
        memcpy(b, r->fb, sizeof(double) * N);
        int reason = 0;
        gm += gmres(r, b, r->D, &reason);
        free(b);
        if (reason > 0) {
#pragma omp parallel for schedule(static)
            for (long e = 0; e < N; ++e) r->fc[e] = r->fc[e] + (-1.0) * r->D[e];
        }
        moments(r);
        update_previous(r);
        const double prev = pred;
        pred = calc_residual(r);
        if (stats) stats[2 + i] = pred;
        if ((pred > prev && i > 1) || pred < r->snes_atol || i >= r->snes_max_iter) {
            if (pred > prev) {
                memcpy(r->fc, r->fl, sizeof(double) * N);
                moments(r);
            }
            break;
        }
    }
    if (stats) {
        stats[0] = i;
        stats[1] = gm;
        stats[2] = pred;
    }
    return pred > 10. ? -1 : 0;
}

void vvo_get_f(const vvo_run *r, double *out) { memcpy(out, r->fc, sizeof(double) * r->N); }

/* replace the state (natural order) and recompute its moments, as after set-up */
void vvo_set_f(vvo_run *r, const double *f)
{
    memcpy(r->fc, f, sizeof(double) * r->N);
    moments(r);
}

/* component entry points for cross-checks (H_ave from the given factors) */
void vvo_set_h(vvo_run *r, const double *h1p, const double *h1h, const double *h2p, const double *h2h)
{
    memcpy(r->hp1, h1p, sizeof(double) * r->nx);
    memcpy(r->hp2, h1h, sizeof(double) * r->nx);
    memcpy(r->hp3, h2p, sizeof(double) * r->nx);
    memcpy(r->hp4, h2h, sizeof(double) * r->nx);
}

void vvo_residual(vvo_run *r, const double *f, const double *fh, double *y) { bracket_op(r, 0, f, fh, y); }
void vvo_jacobian(vvo_run *r, const double *u, double *y) { bracket_op(r, 1, u, NULL, y); }
void vvo_precond(vvo_run *r, const double *x, double *y) { precond(r, x, y); }
void vvo_moments(vvo_run *r, double *n, double *phi, double *h1, double *pnorm)
{
    moments(r);
    memcpy(n, r->n, sizeof(double) * r->nx);
    memcpy(phi, r->phi, sizeof(double) * r->nx);
    memcpy(h1, r->h1c, sizeof(double) * r->nx);
    *pnorm = r->pnorm;
}
void vvo_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int vvo_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
