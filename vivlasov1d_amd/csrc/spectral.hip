// Kinetic tensor-product preconditioner  P^-1 x = irfft_x( T_k^-1 rfft_x(x) )
// (TensorProductKinetic.pyx:146-160; TensorProductKineticFast.pyx:102-224):
//
//  rfft_kernel    real rows x[j][0..nx) -> spectrum S[k][j], k = 0..nx/2.
//                 Two real rows are packed as one complex row and transformed
//                 by an in-LDS Stockham radix-4/2 FFT; the output is written
//                 mode-major (the transpose the reference does with a
//                 VecScatter is folded into the store).
//  tridiag_kernel one workgroup per mode k solves T_k c = S[k][.] along v:
//                 partitioned Thomas (one chunk of M rows per thread) whose
//                 reduced system (one unknown per chunk) is solved by parallel
//                 cyclic reduction in LDS.  T_k's coefficients are rebuilt
//                 from the mode's eigenvalue and three per-row factors, so
//                 no factor arrays are streamed from HBM.
//  irfft_kernel   spectrum -> real rows, backward-normalised like
//                 numpy.fft.irfft / pyFFTW (imaginary parts of the k = 0 and
//                 k = nx/2 coefficients are ignored, as in a c2r transform).
#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int NT = 256;
constexpr int MAXE = 4096;   // complex elements of one FFT workgroup (64 KiB of LDS)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double2 cdiv(double2 z, double2 w)
{
    const double inv = 1.0 / (w.x * w.x + w.y * w.y);
    return make_double2((z.x * w.x + z.y * w.y) * inv, (z.y * w.x - z.x * w.y) * inv);
}
__device__ __forceinline__ double2 crecip(double2 w)
{
    const double inv = 1.0 / (w.x * w.x + w.y * w.y);
    return make_double2(w.x * inv, -w.y * inv);
}

template <bool INV>
__device__ __forceinline__ double2 twiddle(const double2 *__restrict__ tw, int m)
{
    const double2 t = tw[m];
    return INV ? make_double2(t.x, -t.y) : t;
}

// Batched in-LDS Stockham FFT of `nb` transforms of length n = 2^logn stored
// back to back in buf.  Forward uses exp(-2 pi i m/n), inverse the conjugate
// (unnormalised).  Each stage loads every butterfly of the thread into
// registers, synchronises, then stores, so one buffer suffices.
template <bool INV>
__device__ void lds_fft(double2 *buf, int n, int logn, int nb, const double2 *__restrict__ tw)
{
    int Ns = 1;
    const int nq = n >> 2;
    for (int s = 0; s < (logn >> 1); ++s) {
        const int total = nb * nq;
        double2 v[MAXE / 4 / NT][4];
        int dst[MAXE / 4 / NT];
#pragma unroll
        for (int q = 0; q < MAXE / 4 / NT; ++q) {
            const int t = threadIdx.x + q * NT;
            dst[q] = -1;
            if (t < total) {
                const int b = t / nq;
                const int j = t - b * nq;
                const int k = j & (Ns - 1);
                const double2 *src = buf + b * n;
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][r] = src[j + r * nq];
                const int st = (n / (Ns * 4)) * k;
                v[q][1] = cmul(v[q][1], twiddle<INV>(tw, st));
                v[q][2] = cmul(v[q][2], twiddle<INV>(tw, 2 * st));
                v[q][3] = cmul(v[q][3], twiddle<INV>(tw, 3 * st));
                const double2 a0 = cadd(v[q][0], v[q][2]);
                const double2 a1 = csub(v[q][0], v[q][2]);
                const double2 a2 = cadd(v[q][1], v[q][3]);
                const double2 d = csub(v[q][1], v[q][3]);
                // forward: (v1 - v3) * (-i); inverse: (v1 - v3) * (+i)
                const double2 a3 = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
                v[q][0] = cadd(a0, a2);
                v[q][1] = cadd(a1, a3);
                v[q][2] = csub(a0, a2);
                v[q][3] = csub(a1, a3);
                dst[q] = b * n + (j - k) * 4 + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXE / 4 / NT; ++q) {
            if (dst[q] >= 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) buf[dst[q] + r * Ns] = v[q][r];
            }
        }
        __syncthreads();
        Ns <<= 2;
    }
    if (logn & 1) {
        const int nh = n >> 1;
        const int total = nb * nh;
        double2 v[MAXE / 2 / NT][2];
        int dst[MAXE / 2 / NT];
#pragma unroll
        for (int q = 0; q < MAXE / 2 / NT; ++q) {
            const int t = threadIdx.x + q * NT;
            dst[q] = -1;
            if (t < total) {
                const int b = t / nh;
                const int j = t - b * nh;
                const int k = j & (Ns - 1);
                const double2 *src = buf + b * n;
                const double2 x0 = src[j];
                const double2 x1 = cmul(src[j + nh], twiddle<INV>(tw, (n / (Ns * 2)) * k));
                v[q][0] = cadd(x0, x1);
                v[q][1] = csub(x0, x1);
                dst[q] = b * n + (j - k) * 2 + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXE / 2 / NT; ++q) {
            if (dst[q] >= 0) {
                buf[dst[q]] = v[q][0];
                buf[dst[q] + Ns] = v[q][1];
            }
        }
        __syncthreads();
    }
}

struct FftArgs {
    int nx, logn, nv, rows;     // rows: v-rows per workgroup (even)
    const double2 *tw;
    const int32_t *done;
};

__global__ __launch_bounds__(NT) void rfft_kernel(FftArgs a, const double *__restrict__ x,
                                                  double2 *__restrict__ spec)
{
    if (a.done != nullptr && *a.done) return;
    extern __shared__ double2 buf[];
    const int nx = a.nx, R = a.rows;
    const int j0 = blockIdx.x * R;
    // every global load of the thread is issued before the LDS stores
    constexpr int PER = 2 * MAXE / NT;   // R * nx <= 2 * MAXE real inputs
    double val[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int r = idx >> a.logn;
        const int i = idx & (nx - 1);
        val[q] = (idx < R * nx && j0 + r < a.nv) ? x[(size_t)(j0 + r) * nx + i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < R * nx) {
            const int r = idx >> a.logn;
            const int i = idx & (nx - 1);
            double *slot = reinterpret_cast<double *>(&buf[(r >> 1) * nx + i]);
            slot[r & 1] = val[q];
        }
    }
    __syncthreads();
    lds_fft<false>(buf, nx, a.logn, R >> 1, a.tw);
    const int nk = (nx >> 1) + 1;
    for (int idx = threadIdx.x; idx < nk * R; idx += NT) {
        const int k = idx / R;
        const int r = idx - k * R;
        if (j0 + r >= a.nv) continue;
        const double2 *z = buf + (r >> 1) * nx;
        const double2 zk = z[k];
        const double2 zm = z[(nx - k) & (nx - 1)];
        double2 out;
        if ((r & 1) == 0) out = make_double2((zk.x + zm.x) * 0.5, (zk.y - zm.y) * 0.5);
        else out = make_double2((zk.y + zm.y) * 0.5, (zm.x - zk.x) * 0.5);
        spec[(size_t)k * a.nv + j0 + r] = out;
    }
}

__global__ __launch_bounds__(NT) void irfft_kernel(FftArgs a, const double2 *__restrict__ spec,
                                                   double *__restrict__ y)
{
    if (a.done != nullptr && *a.done) return;
    extern __shared__ double2 buf[];
    const int nx = a.nx, R = a.rows, nh = nx >> 1;
    const int j0 = blockIdx.x * R;
    const int npair = R >> 1;
    // Z_k = A_k + i B_k (k <= nx/2) and its Hermitian extension for k > nx/2;
    // every global load of the thread is issued before the LDS stores.
    const int total = (nh + 1) * npair;
    constexpr int PER = (MAXE / 2 + 64 + NT - 1) / NT;   // (nx/2 + 1) R/2 <= MAXE/2 + 64
    double2 va[PER], vb[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int k = idx / npair;
        const int ja = j0 + 2 * (idx - k * npair);
        va[q] = (idx < total && ja < a.nv) ? spec[(size_t)k * a.nv + ja] : make_double2(0.0, 0.0);
        vb[q] = (idx < total && ja + 1 < a.nv) ? spec[(size_t)k * a.nv + ja + 1] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < total) {
            const int k = idx / npair;
            const int p = idx - k * npair;
            double2 A = va[q], B = vb[q];
            if (k == 0 || k == nh) {
                A.y = 0.0;
                B.y = 0.0;
            }
            double2 *z = buf + p * nx;
            z[k] = make_double2(A.x - B.y, A.y + B.x);
            if (k != 0 && k != nh) z[nx - k] = make_double2(A.x + B.y, B.x - A.y);
        }
    }
    __syncthreads();
    lds_fft<true>(buf, nx, a.logn, npair, a.tw);
    const double inv_n = 1.0 / (double)nx;   // exact: nx is a power of two
    for (int idx = threadIdx.x; idx < R * nx; idx += NT) {
        const int r = idx >> a.logn;
        const int i = idx & (nx - 1);
        if (j0 + r < a.nv) {
            const double2 z = buf[(r >> 1) * nx + i];
            y[(size_t)(j0 + r) * nx + i] = ((r & 1) ? z.y : z.x) * inv_n;
        }
    }
}

struct TriArgs {
    int nv;           // system size
    int P;            // chunks (threads), power of two
    const double2 *eigen;
    const double *gm, *g0, *gp;
    double af, ht_inv;
    const int32_t *done;
};

// Row j of T_k (KineticFast.pyx:248-272): sub a_j = diagm[j] (j >= 1),
// diagonal b_j = diag[j], super c_j = diagp[j] (j <= nv-2); rows past nv are
// identity rows that pad the system to P*M unknowns.
__device__ __forceinline__ void tri_row(const TriArgs &a, double2 e, int j, double2 &aj, double2 &bj, double2 &cj)
{
    if (j < a.nv) {
        const double g = a.gm[j];
        aj = (j > 0) ? make_double2(((e.x * 0.5) * g) * a.af, ((e.y * 0.5) * g) * a.af) : make_double2(0.0, 0.0);
        const double g0 = a.g0[j];
        bj = make_double2(((e.x * 1.0) * g0) * a.af + a.ht_inv, ((e.y * 1.0) * g0) * a.af);
        const double gp = a.gp[j];
        cj = (j < a.nv - 1) ? make_double2(((e.x * 0.5) * gp) * a.af, ((e.y * 0.5) * gp) * a.af)
                            : make_double2(0.0, 0.0);
    } else {
        aj = make_double2(0.0, 0.0);
        bj = make_double2(1.0, 0.0);
        cj = make_double2(0.0, 0.0);
    }
}

template <int M>
__global__ __launch_bounds__(NT) void tridiag_kernel(TriArgs a, double2 *__restrict__ spec)
{
    if (a.done != nullptr && *a.done) return;
    extern __shared__ double2 sd[];          // P*(M+1): right-hand sides / solution, chunk c at c*(M+1)
    __shared__ double2 sa[NT], sb[NT], sc[NT], se[NT];
    const int k = blockIdx.x;
    const double2 e = a.eigen[k];
    const int P = a.P, n = P * M;
    double2 *S = spec + (size_t)k * a.nv;
    {
        // n = P*M <= NT*M: every load of the thread in flight before the LDS stores
        double2 v[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = threadIdx.x + q * NT;
            v[q] = (j < a.nv) ? S[j] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = threadIdx.x + q * NT;
            if (j < n) sd[j + j / M] = v[q];      // one pad slot per chunk: no bank conflicts
        }
    }
    __syncthreads();

    const int c = threadIdx.x;
    double2 Am[M], Cm[M], Dm[M];
    double2 ae, be, ce;          // coefficients of the chunk's last row
    if (c < P) {
        const int s = c * M, ls = c * (M + 1);
        // forward sweep: y_j + c'_j y_{j+1} = d'_j - a'_j y_{s-1}
        double2 aj, bj, cj;
        tri_row(a, e, s, aj, bj, cj);
        double2 r = crecip(bj);
        Cm[0] = cmul(cj, r);
        Dm[0] = cmul(sd[ls], r);
        Am[0] = cmul(aj, r);
#pragma unroll
        for (int q = 1; q < M; ++q) {
            tri_row(a, e, s + q, aj, bj, cj);
            if (q == M - 1) {
                ae = aj;
                be = bj;
                ce = cj;
            }
            if (q < M - 1) {
                r = crecip(csub(bj, cmul(aj, Cm[q - 1])));
                Cm[q] = cmul(cj, r);
                Dm[q] = cmul(csub(sd[ls + q], cmul(aj, Dm[q - 1])), r);
                Am[q] = cneg(cmul(cmul(aj, Am[q - 1]), r));
            }
        }
        // backward sweep over rows s..e-1: y_j = D_j - A_j y_{s-1} - C_j y_e
#pragma unroll
        for (int q = M - 3; q >= 0; --q) {
            Dm[q] = csub(Dm[q], cmul(Cm[q], Dm[q + 1]));
            Am[q] = csub(Am[q], cmul(Cm[q], Am[q + 1]));
            Cm[q] = cneg(cmul(Cm[q], Cm[q + 1]));
        }
        // publish the first row of the chunk for the neighbour below
        sa[c] = Am[0];
        sc[c] = Cm[0];
        se[c] = Dm[0];
    }
    __syncthreads();
    // reduced equation of chunk c in the chunk-last unknowns E:
    //   alpha E_{c-1} + beta E_c + gamma E_{c+1} = delta
    double2 al = make_double2(0.0, 0.0), bt = make_double2(1.0, 0.0), gm = make_double2(0.0, 0.0),
            dl = make_double2(0.0, 0.0);
    if (c < P) {
        const double2 dE = sd[c * (M + 1) + M - 1];
        al = cneg(cmul(ae, Am[M - 2]));
        bt = csub(be, cmul(ae, Cm[M - 2]));
        dl = csub(dE, cmul(ae, Dm[M - 2]));
        if (c + 1 < P) {
            bt = csub(bt, cmul(ce, sa[c + 1]));
            gm = cneg(cmul(ce, sc[c + 1]));
            dl = csub(dl, cmul(ce, se[c + 1]));
        }
    }
    __syncthreads();
    // parallel cyclic reduction on the P reduced unknowns
    for (int st = 1; st < P; st <<= 1) {
        if (c < P) {
            sa[c] = al;
            sb[c] = bt;
            sc[c] = gm;
            se[c] = dl;
        }
        __syncthreads();
        if (c < P) {
            double2 nal = make_double2(0.0, 0.0), ngm = make_double2(0.0, 0.0);
            double2 nbt = bt, ndl = dl;
            if (c - st >= 0) {
                const double2 k1 = cdiv(al, sb[c - st]);
                nal = cneg(cmul(sa[c - st], k1));
                nbt = csub(nbt, cmul(sc[c - st], k1));
                ndl = csub(ndl, cmul(se[c - st], k1));
            }
            if (c + st < P) {
                const double2 k2 = cdiv(gm, sb[c + st]);
                ngm = cneg(cmul(sc[c + st], k2));
                nbt = csub(nbt, cmul(sa[c + st], k2));
                ndl = csub(ndl, cmul(se[c + st], k2));
            }
            al = nal;
            gm = ngm;
            bt = nbt;
            dl = ndl;
        }
        __syncthreads();
    }
    if (c < P) se[c] = cdiv(dl, bt);     // E_c = y at the chunk's last row
    __syncthreads();
    if (c < P) {
        const int ls = c * (M + 1);
        const double2 Eprev = (c > 0) ? se[c - 1] : make_double2(0.0, 0.0);
        const double2 Eown = se[c];
#pragma unroll
        for (int q = 0; q < M - 1; ++q)
            sd[ls + q] = csub(csub(Dm[q], cmul(Am[q], Eprev)), cmul(Cm[q], Eown));
        sd[ls + M - 1] = Eown;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < a.nv; j += NT) S[j] = sd[j + j / M];
}

int fft_rows(int nx, int nv)
{
    int R = 8192 / nx;                 // LDS: R/2 complex rows of nx = R*nx*8 bytes <= 64 KiB
    if (R < 2) R = 2;
    while (R > 2 && (nv + R - 1) / R < 512) R >>= 1;
    return R;
}

}  // namespace

int fft_blocks(vv_ctx *c)
{
    const int R = fft_rows(c->p.nx, c->p.nv);
    return (c->p.nv + R - 1) / R;
}

hipError_t launch_rfft(vv_ctx *c, const double *x, double2 *spec, const int32_t *done)
{
    FftArgs a{c->p.nx, c->logn, c->p.nv, fft_rows(c->p.nx, c->p.nv), c->twiddle, done};
    const int nb = (a.nv + a.rows - 1) / a.rows;
    const size_t lds = (size_t)(a.rows >> 1) * a.nx * sizeof(double2);
    hipLaunchKernelGGL(rfft_kernel, dim3(nb), dim3(NT), lds, c->stream, a, x, spec);
    return hipGetLastError();
}

hipError_t launch_irfft(vv_ctx *c, const double2 *spec, double *y, const int32_t *done)
{
    FftArgs a{c->p.nx, c->logn, c->p.nv, fft_rows(c->p.nx, c->p.nv), c->twiddle, done};
    const int nb = (a.nv + a.rows - 1) / a.rows;
    const size_t lds = (size_t)(a.rows >> 1) * a.nx * sizeof(double2);
    hipLaunchKernelGGL(irfft_kernel, dim3(nb), dim3(NT), lds, c->stream, a, spec, y);
    return hipGetLastError();
}

void tridiag_shape(int nv, int *P, int *M)
{
    if (nv <= 512) {
        *M = 2;
        int p = 1;
        while (p * 2 < nv) p <<= 1;          // P = next power of two >= nv/2
        *P = p;
    } else {
        *P = 256;
        int m = (nv + 255) / 256;
        *M = m <= 4 ? 4 : (m <= 8 ? 8 : 16);
    }
}

hipError_t launch_tridiag_k(vv_ctx *c, double2 *spec, const int32_t *done)
{
    int P, M;
    tridiag_shape(c->p.nv, &P, &M);
    TriArgs a{c->p.nv, P, c->eigen, c->gm, c->g0, c->gp, c->afpc, c->ht_inv, done};
    const size_t lds = (size_t)P * (M + 1) * sizeof(double2);
    dim3 grid(c->nmodes), block(NT);
    switch (M) {
    case 2: hipLaunchKernelGGL(tridiag_kernel<2>, grid, block, lds, c->stream, a, spec); break;
    case 4: hipLaunchKernelGGL(tridiag_kernel<4>, grid, block, lds, c->stream, a, spec); break;
    case 8: hipLaunchKernelGGL(tridiag_kernel<8>, grid, block, lds, c->stream, a, spec); break;
    default: hipLaunchKernelGGL(tridiag_kernel<16>, grid, block, lds, c->stream, a, spec); break;
    }
    return hipGetLastError();
}

}  // namespace vv
