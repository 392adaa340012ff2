// Kinetic tensor-product preconditioner  P^-1 x = irfft_x( T_k^-1 rfft_x(x) )
// (TensorProductKinetic.pyx:146-160; TensorProductKineticFast.pyx:102-224):
//
//  rfft_kernel    real rows x[j][0..nx) -> spectrum S[k][j], k = 0..nx/2.
//                 Two real rows are packed as one complex row and transformed
//                 by an in-LDS Stockham FFT (radix-4/2; length 1024 as
//                 4 x 8 x 8 x 4); the output is written
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
//
// The potential form (TensorProductPotential.pyx:157-173) runs the same three
// kernels with the roles of x and v exchanged: the field is transposed
// (transpose_kernel), rFFT along v, one tridiagonal system along x per
// v-mode with bands from phi differences (potential_bands_kernel), irFFT,
// transposed back.  The transform geometry (length, rows, modes, band
// constants) comes from the context.
//
// Every kernel is a template on the real type: double for the FP64 path,
// float for the FP32 preconditioner inside the mixed-precision GMRES cycle
// (complex64 spectrum, float twiddles; T_k's coefficients are formed in FP64
// and rounded).
#include "arakawa.h"
#include "cgs_vec.h"

namespace vv {
namespace {

constexpr int NT = 256;
constexpr int MAXE = 4096;   // complex elements of one FFT workgroup (64 KiB of LDS)

// complex arithmetic on double2 / float2
template <typename C> __device__ __forceinline__ C cmk(typename C::value_type re, typename C::value_type im)
{
    C r;
    r.x = re;
    r.y = im;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cneg(C a) { return cmk<C>(-a.x, -a.y); }
// 1/d for the complex divisions of the tridiagonal solve: the hardware
// reciprocal estimate refined by two Newton steps r <- r (2 - d r) instead of
// the IEEE division sequence.  The dependent chain of the Thomas sweep is
// shorter (21.5 -> 20.9 us per solve at 1024 x 2048); P^-1 moves by about
// one ulp (2.8e-16 relative).  The pivots are O(1/dt), far from the
// denormal and overflow ranges.
__device__ __forceinline__ double rcp_of(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = r * (2.0 - d * r);
    r = r * (2.0 - d * r);
    return r;
}
__device__ __forceinline__ float rcp_of(float d) { return 1.0f / d; }

template <typename C> __device__ __forceinline__ C cdiv(C z, C w)
{
    const auto inv = rcp_of(w.x * w.x + w.y * w.y);
    return cmk<C>((z.x * w.x + z.y * w.y) * inv, (z.y * w.x - z.x * w.y) * inv);
}
template <typename C> __device__ __forceinline__ C crecip(C w)
{
    const auto inv = rcp_of(w.x * w.x + w.y * w.y);
    return cmk<C>(w.x * inv, -w.y * inv);
}

// p[i] when ok, else zero.  The address is clamped to p[0] instead of
// selecting between p + i and a local zero, so the load stays an
// unconditional global load that can be in flight with the others.
template <typename C> __device__ __forceinline__ C load_or_zero(const C *__restrict__ p, int64_t i, bool ok)
{
    const C v = p[ok ? i : 0];
    return ok ? v : cmk<C>(typename C::value_type(0), typename C::value_type(0));
}

// complex type of a real type
template <typename T> struct CxOf;
template <> struct CxOf<double> { using type = double2; };
template <> struct CxOf<float> { using type = float2; };
template <typename T> using cx_t = typename CxOf<T>::type;

template <bool INV, typename C>
__device__ __forceinline__ C twiddle(const C *__restrict__ tw, int m)
{
    const C t = tw[m];
    return INV ? cmk<C>(t.x, -t.y) : t;
}

constexpr int MAXS = 6;     // radix-4 stages of the longest transform (MAXE = 4^6)

// Twiddles of one thread's butterflies, loaded once before a transform of
// length n <= 4 NT.  There every butterfly the thread owns, in any transform
// of the block and any stage, has the index j = threadIdx.x mod n/4 (mod n/2
// in the radix-2 stage of odd logn, where n <= 2 NT), so 3 twiddles per stage
// cover them; their loads are in flight with the tile's own loads instead of
// being waited on after every barrier of the transform.
// n = 1024 runs the plan 4 x 8 x 8 x 4 (see lds_fft1024): its 17 twiddles
// take the same registers, flat index i (always a compile-time index) at
// w[i / 3][i % 3] for i < 15, then w2 and w8 (0..6 and 7..13 for the radix-8
// stages with sub-transforms of 4 and 32, 14..16 for the last radix-4 stage;
// the last row of w, used only by n = 4096, stays out of it).
constexpr int N8 = 1024;
template <bool INV, typename C>
struct FftTw {
    C w[MAXS][3];
    C w2, w8;
    // P8: the n = 1024 plan's twiddles (n must be 1024), else the radix-4 ones
    template <bool P8 = false>
    __device__ __forceinline__ void load(const C *__restrict__ tw, int n, int logn)
    {
        if (P8) {
            const int t = threadIdx.x, ka = t & 3, kb = t & 31, kc = t & 255;
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                w[(r - 1) / 3][(r - 1) % 3] = twiddle<INV>(tw, r * ka * 32);
                w[(6 + r) / 3][(6 + r) % 3] = twiddle<INV>(tw, r * kb * 4);
            }
            w[4][2] = twiddle<INV>(tw, kc);
            w2 = twiddle<INV>(tw, 2 * kc);
            w8 = twiddle<INV>(tw, 3 * kc);
            return;
        }
        const int ns = logn >> 1;
        const int j = threadIdx.x & ((n >> 2) - 1);
#pragma unroll
        for (int s = 0; s < MAXS; ++s) {
            if (s < ns) {
                const int st = (j & ((1 << (2 * s)) - 1)) << (logn - 2 - 2 * s);
                w[s][0] = twiddle<INV>(tw, st);
                w[s][1] = twiddle<INV>(tw, 2 * st);
                w[s][2] = twiddle<INV>(tw, 3 * st);
            }
        }
        if (logn & 1) {
            const int jj = threadIdx.x & ((n >> 1) - 1);
            w2 = twiddle<INV>(tw, (jj & ((1 << (2 * ns)) - 1)) << (logn - 1 - 2 * ns));
        }
    }
};

// Where the last stage of lds_fft puts its outputs: back into the LDS buffer
// (the default), or wherever a kernel wants them (the irFFT stores them to
// its output rows directly).  pos = transform * n + natural index.
template <typename C>
struct LdsSink {
    C *buf;
    __device__ __forceinline__ void operator()(int pos, C v) const { buf[pos] = v; }
};

// The radix-4 butterfly of the Stockham stages (outputs in natural order of
// the 4-point DFT; forward (v1 - v3)(-i), inverse (v1 - v3)(+i)).
template <bool INV, typename C>
__device__ __forceinline__ void dft4(C &v0, C &v1, C &v2, C &v3)
{
    const C a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), d = csub(v1, v3);
    const C a3 = INV ? cmk<C>(-d.y, d.x) : cmk<C>(d.y, -d.x);
    v0 = cadd(a0, a2);
    v1 = cadd(a1, a3);
    v2 = csub(a0, a2);
    v3 = csub(a1, a3);
}

// 8-point DFT in place: the radix-4 DFTs of the even and the odd samples,
// combined with W8^m (forward W8 = (1 - i)/sqrt 2, inverse its conjugate).
template <bool INV, typename C>
__device__ __forceinline__ void dft8(C (&v)[8])
{
    using R = typename C::value_type;
    const R h = R(0.70710678118654752440);
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    o1 = INV ? cmk<C>((o1.x - o1.y) * h, (o1.x + o1.y) * h) : cmk<C>((o1.x + o1.y) * h, (o1.y - o1.x) * h);
    o2 = INV ? cmk<C>(-o2.y, o2.x) : cmk<C>(o2.y, -o2.x);
    o3 = INV ? cmk<C>(-(o3.x + o3.y) * h, (o3.x - o3.y) * h) : cmk<C>((o3.y - o3.x) * h, -(o3.x + o3.y) * h);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// One radix-8 Stockham stage of the n = 1024 plan: sub-transforms of length
// NS (4 or 32) combined into 8 NS; butterfly j takes elements j + 128 r,
// twiddled by W_1024^(r k 1024 / (8 NS)), k = j mod NS (= threadIdx.x mod NS
// for every butterfly of the thread, so the thread's twiddles are preloaded:
// flat index OFF + r - 1 of FftTw), and stores output m at 8 (j - k) + k + m NS.
template <bool INV, int NS, int OFF, typename C>
__device__ __forceinline__ void stage8_1024(C *buf, int nb, const FftTw<INV, C> &pt)
{
    constexpr int nq = N8 / 8;
    const int total = nb * nq;
    C v[2][8];
    int dst[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int bq = threadIdx.x + q * NT;
        dst[q] = -1;
        if (bq < total) {
            const int b = bq >> 7, j = bq & (nq - 1), k = j & (NS - 1);
            const C *src = buf + b * N8;
#pragma unroll
            for (int r = 0; r < 8; ++r) v[q][r] = src[j + r * nq];
#pragma unroll
            for (int r = 1; r < 8; ++r) v[q][r] = cmul(v[q][r], pt.w[(OFF + r - 1) / 3][(OFF + r - 1) % 3]);
            dft8<INV>(v[q]);
            dst[q] = b * N8 + ((j - k) << 3) + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q)
        if (dst[q] >= 0) {
#pragma unroll
            for (int m = 0; m < 8; ++m) buf[dst[q] + m * NS] = v[q][m];
        }
    __syncthreads();
}

// n = 1024 as 4 x 8 x 8 x 4: the first radix-4 stage (from LDS when s0 == 0;
// s0 == 1: the caller ran it in registers), two radix-8 stages and a last
// radix-4 stage whose outputs go to `last` -- four passes over LDS instead of
// the five of the radix-4 plan.  nb <= 4 transforms (MAXE elements).
template <bool INV, typename C, typename Sink>
__device__ __forceinline__ void lds_fft1024(C *buf, int nb, const FftTw<INV, C> &pt, int s0, const Sink &last)
{
    constexpr int nq = N8 / 4;
    const int total = nb * nq;
    if (s0 == 0) {
        C v[4][4];
        int dst[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int bq = threadIdx.x + q * NT;
            dst[q] = -1;
            if (bq < total) {
                const int b = bq >> 8, j = bq & (nq - 1);
                const C *src = buf + b * N8;
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][r] = src[j + r * nq];
                dft4<INV>(v[q][0], v[q][1], v[q][2], v[q][3]);
                dst[q] = b * N8 + (j << 2);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (dst[q] >= 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) buf[dst[q] + r] = v[q][r];
            }
        __syncthreads();
    }
    stage8_1024<INV, 4, 0>(buf, nb, pt);
    stage8_1024<INV, 32, 7>(buf, nb, pt);
    // last stage: sub-transforms of 256, k = j = threadIdx.x, outputs j + 256 m
    C v[4][4];
    int dst[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int bq = threadIdx.x + q * NT;
        dst[q] = -1;
        if (bq < total) {
            const int b = bq >> 8, j = bq & (nq - 1);
            const C *src = buf + b * N8;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[q][r] = src[j + r * nq];
            v[q][1] = cmul(v[q][1], pt.w[4][2]);
            v[q][2] = cmul(v[q][2], pt.w2);
            v[q][3] = cmul(v[q][3], pt.w8);
            dft4<INV>(v[q][0], v[q][1], v[q][2], v[q][3]);
            dst[q] = b * N8 + j;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (dst[q] >= 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m) last(dst[q] + m * nq, v[q][m]);
        }
    __syncthreads();
}

// Batched in-LDS Stockham FFT of `nb` transforms of length n = 2^logn stored
// back to back in buf.  Forward uses exp(-2 pi i m/n), inverse the conjugate
// (unnormalised).  Each stage loads every butterfly of the thread into
// registers, synchronises, then stores, so one buffer suffices.  C is
// double2 (FP64 path) or float2 (FP32 preconditioner of the mixed solver).
// PRE: twiddles from pt (n <= 4 NT), else from the global table tw.  P8:
// n = 1024 by the 4 x 8 x 8 x 4 plan (pt loaded with load<true>).
// s0: first radix-4 stage to run (1 when the caller did stage 0 in
// registers); the last stage's outputs go to `last`.
// Inlined so that buf is known to be LDS (ds_read/ds_write, not flat
// accesses); index arithmetic by shifts (n is a power of two).
template <bool INV, bool PRE, bool P8 = false, typename C, typename Sink>
__device__ __forceinline__ void lds_fft(C *buf, int n, int logn, int nb, const C *__restrict__ tw,
                                        const FftTw<INV, C> &pt, int s0, const Sink &last)
{
    if (P8) {                              // n = 1024, twiddles loaded for the plan
        lds_fft1024<INV>(buf, nb, pt, s0, last);
        return;
    }
    const int nq = n >> 2, lq = logn - 2, ns = logn >> 1;
    const int total4 = nb * nq;
#pragma unroll
    for (int s = 0; s < MAXS; ++s) {
        if (s >= ns) break;                // uniform over the block
        if (s < s0) continue;
        const int Ns = 1 << (2 * s);
        C v[MAXE / 4 / NT][4];
        int dst[MAXE / 4 / NT];
#pragma unroll
        for (int q = 0; q < MAXE / 4 / NT; ++q) {
            const int t = threadIdx.x + q * NT;
            dst[q] = -1;
            if (t < total4) {
                const int b = t >> lq;
                const int j = t & (nq - 1);
                const int k = j & (Ns - 1);
                const C *src = buf + (b << logn);
#pragma unroll
                for (int r = 0; r < 4; ++r) v[q][r] = src[j + r * nq];
                if (PRE) {
                    v[q][1] = cmul(v[q][1], pt.w[s][0]);
                    v[q][2] = cmul(v[q][2], pt.w[s][1]);
                    v[q][3] = cmul(v[q][3], pt.w[s][2]);
                } else {
                    const int st = k << (lq - 2 * s);
                    v[q][1] = cmul(v[q][1], twiddle<INV>(tw, st));
                    v[q][2] = cmul(v[q][2], twiddle<INV>(tw, 2 * st));
                    v[q][3] = cmul(v[q][3], twiddle<INV>(tw, 3 * st));
                }
                const C a0 = cadd(v[q][0], v[q][2]);
                const C a1 = csub(v[q][0], v[q][2]);
                const C a2 = cadd(v[q][1], v[q][3]);
                const C d = csub(v[q][1], v[q][3]);
                // forward: (v1 - v3) * (-i); inverse: (v1 - v3) * (+i)
                const C a3 = INV ? cmk<C>(-d.y, d.x) : cmk<C>(d.y, -d.x);
                v[q][0] = cadd(a0, a2);
                v[q][1] = cadd(a1, a3);
                v[q][2] = csub(a0, a2);
                v[q][3] = csub(a1, a3);
                dst[q] = (b << logn) + ((j - k) << 2) + k;
            }
        }
        __syncthreads();
        const bool fin = (s == ns - 1) && !(logn & 1);
#pragma unroll
        for (int q = 0; q < MAXE / 4 / NT; ++q) {
            if (dst[q] >= 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (fin) last(dst[q] + r * Ns, v[q][r]);
                    else buf[dst[q] + r * Ns] = v[q][r];
                }
            }
        }
        __syncthreads();
    }
    if (logn & 1) {
        const int Ns = 1 << (2 * ns);
        const int nh = n >> 1;
        const int total = nb * nh;
        C v[MAXE / 2 / NT][2];
        int dst[MAXE / 2 / NT];
#pragma unroll
        for (int q = 0; q < MAXE / 2 / NT; ++q) {
            const int t = threadIdx.x + q * NT;
            dst[q] = -1;
            if (t < total) {
                const int b = t >> (logn - 1);
                const int j = t & (nh - 1);
                const int k = j & (Ns - 1);
                const C *src = buf + (b << logn);
                const C x0 = src[j];
                const C w = PRE ? pt.w2 : twiddle<INV>(tw, k << (logn - 1 - 2 * ns));
                const C x1 = cmul(src[j + nh], w);
                v[q][0] = cadd(x0, x1);
                v[q][1] = csub(x0, x1);
                dst[q] = (b << logn) + ((j - k) << 1) + k;
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXE / 2 / NT; ++q) {
            if (dst[q] >= 0) {
                last(dst[q], v[q][0]);
                last(dst[q] + Ns, v[q][1]);
            }
        }
        __syncthreads();
    }
}

// Instance b of a batched context is blockIdx.y: real fields at b * N,
// spectra at b * S, GMRES state b * gstride bytes after instance 0's.
template <typename T>
struct FftArgs {
    int nx, logn, nv, rows;     // rows: v-rows per workgroup (even)
    int64_t N, S, gstride;
    const cx_t<T> *tw;
    const int32_t *done;
};

__device__ __forceinline__ int32_t cycle_done(const int32_t *done, int64_t gstride)
{
    return done != nullptr ? *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(done) + blockIdx.y * gstride)
                           : 0;
}

// The R real rows' spectra from their packed transforms in buf (complex row
// p = real rows 2p + i 2p+1), stored mode-major: S[k][j0 + r], k <= nx/2.
template <typename C>
__device__ __forceinline__ void store_real_spectrum(const C *buf, int nx, int R, int j0, int nv, C *__restrict__ spec)
{
    using T = typename C::value_type;
    const int nk = (nx >> 1) + 1;
    const T half = T(0.5);
    for (int idx = threadIdx.x; idx < nk * R; idx += NT) {
        const int k = idx / R;
        const int r = idx - k * R;
        if (j0 + r >= nv) continue;
        const C *z = buf + (r >> 1) * nx;
        const C zk = z[k];
        const C zm = z[(nx - k) & (nx - 1)];
        C out;
        if ((r & 1) == 0) out = cmk<C>((zk.x + zm.x) * half, (zk.y - zm.y) * half);
        else out = cmk<C>((zk.y + zm.y) * half, (zm.x - zk.x) * half);
        spec[(size_t)k * nv + j0 + r] = out;
    }
}

// PRE (nx <= 4 NT): the thread's twiddles are loaded with its tile; P8:
// nx = 1024 by the radix-8 plan.
template <typename T, bool PRE, bool P8>
__global__ __launch_bounds__(NT) void rfft_kernel(FftArgs<T> a, const T *__restrict__ x, cx_t<T> *__restrict__ spec)
{
    using C = cx_t<T>;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    x += blockIdx.y * a.N;
    spec += blockIdx.y * a.S;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw);
    const int nx = a.nx, R = a.rows;
    // XCD-aware row-block order: the mode-major segments of neighbouring row
    // blocks share L2 lines, so they are kept on one XCD
    const int j0 = xcd_tile(blockIdx.x, gridDim.x) * R;
    // Stage 0 of the FFT runs in registers: butterfly (p, j) of complex row p
    // (real rows 2p, 2p+1) takes z_p[j + m nx/4], m = 0..3, exactly the
    // elements the thread loads (coalesced along j); its twiddles are 1.
    // Every global load of the thread is issued before anything waits.
    constexpr int QB = MAXE / 4 / NT;    // butterflies per thread
    const int nq = nx >> 2, lq = a.logn - 2;
    const int total4 = (R >> 1) * nq;
    T xr[QB][4], xi[QB][4];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int bq = threadIdx.x + q * NT;
        const int p = bq >> lq, j = bq & (nq - 1);
        const bool ok = bq < total4;
        const bool r0 = ok && j0 + 2 * p < a.nv, r1 = ok && j0 + 2 * p + 1 < a.nv;
        const T *x0 = x + (size_t)(j0 + 2 * p) * nx + j;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            xr[q][m] = r0 ? x0[m * nq] : T(0);
            xi[q][m] = r1 ? x0[nx + m * nq] : T(0);
        }
    }
    FftTw<false, C> pt;
    if (PRE) pt.template load<P8>(a.tw, nx, a.logn);
    if (dflag) return;                 // uniform over the block, before any barrier
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int bq = threadIdx.x + q * NT;
        if (bq < total4) {
            const int p = bq >> lq, j = bq & (nq - 1);
            const C v0 = cmk<C>(xr[q][0], xi[q][0]), v1 = cmk<C>(xr[q][1], xi[q][1]);
            const C v2 = cmk<C>(xr[q][2], xi[q][2]), v3 = cmk<C>(xr[q][3], xi[q][3]);
            const C a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), d = csub(v1, v3);
            const C a3 = cmk<C>(d.y, -d.x);             // (v1 - v3) * (-i)
            C *z = buf + (p << a.logn) + (j << 2);      // stage 0 outputs: 4j + r
            z[0] = cadd(a0, a2);
            z[1] = cadd(a1, a3);
            z[2] = csub(a0, a2);
            z[3] = csub(a1, a3);
        }
    }
    __syncthreads();
    lds_fft<false, PRE, P8>(buf, nx, a.logn, R >> 1, a.tw, pt, 1, LdsSink<C>{buf});
    store_real_spectrum(buf, nx, R, j0, a.nv, spec);
}

// Jacobian-vector product fused with the forward rFFT of its rows:
// X = J u (the bracket kernel's MODE_JACOBIAN, PETScNLVlasovMP.pyx:33-37:
// Y = 0; Y += u/dt; Y += 0.5 {u, H_ave} on rows 2..nv-3) is formed in LDS row
// block by row block and transformed there, so X never goes to HBM.
// The workgroup's R rows are computed in strips of NT columns: the strip of
// u (times the Krylov vector's scale) and of H_ave with a 2-cell halo,
// (R+4) x (NT+4) cells each, is staged in LDS; thread t walks column t of
// the strip down the R rows with the bracket kernel's stencil window
// (arakawa.h), so every value is the bracket kernel's, bit for bit.  The next
// strip's loads are in flight during the current strip's walk, and the
// results stay in registers until the last walk is done, so the FFT buffer
// reuses the strips' LDS.  nx = 4 NT: thread t's columns over the 4 strips,
// t + m NT, are exactly the inputs of radix-4 butterfly t of the first FFT
// stage, which then runs in registers as in rfft_kernel (the same
// arithmetic, bit for bit).  nx = NT, 2 NT: the rows go to the FFT buffer and
// the transform starts from LDS (its unit twiddles can change only the sign
// of an exact zero).  R in 2, 4, 8.
struct JacRows {
    const double *uscale;                       // optional device scale of u (GMRES state)
    const double *h0, *h1p, *h1h, *h2p, *h2h;   // 1D factors of H_ave
    const InstConst *inst;
    double ht_inv, factor;
};

template <int R, typename T>
struct JacStrip {
    T u[R + 4];          // own column, rows j0-2 .. j0+R+1 (0 outside the grid)
    double x[4];         // own column's x-factors
    T eu;                // halo cell (threads t < 4 (R+4))
    double ex[4], eh0;
};

template <int R, typename T>
__device__ __forceinline__ void jac_strip_load(const T *__restrict__ u, const JacRows &jr, size_t xo, int nx, int nv,
                                               int j0, int cb, JacStrip<R, T> &g)
{
    constexpr int RH = R + 4;
    const int t = threadIdx.x;
    const int i = cb + t;                              // staged column t + 2 (inside the row: nx >= cb + NT)
#pragma unroll
    for (int r = 0; r < RH; ++r) {
        const int gj = jr_row(j0, r, nv);
        g.u[r] = gj >= 0 ? u[(size_t)gj * nx + i] : T(0);
    }
    g.x[0] = jr.h1p[xo + i];
    g.x[1] = jr.h1h[xo + i];
    g.x[2] = jr.h2p[xo + i];
    g.x[3] = jr.h2h[xo + i];
    if (t < 4 * RH) {
        int hr, tc;
        jr_halo_cell(t, RH, &hr, &tc);
        const int gi = jr_col(cb, tc, nx);
        const int gj = jr_row(j0, hr, nv);
        const bool in = gj >= 0;
        g.eu = in ? u[(size_t)gj * nx + gi] : T(0);
        g.eh0 = in ? jr.h0[gj] : 0.0;
        g.ex[0] = jr.h1p[xo + gi];
        g.ex[1] = jr.h1h[xo + gi];
        g.ex[2] = jr.h2p[xo + gi];
        g.ex[3] = jr.h2h[xo + gi];
    }
}

// H_ave from its factors in the tile's type: FP64 as is, FP32 relative to
// href (formed in FP64, rounded once), as the bracket kernel stages it
template <typename T>
__device__ __forceinline__ T have_rel(double h0, const double (&x)[4], double href)
{
    if constexpr (sizeof(T) == 8) return compose_have(h0, x[0], x[1], x[2], x[3]);
    else return (T)(compose_have(h0, x[0], x[1], x[2], x[3]) - href);
}

// Thread t's column of the staged strip down the R output rows (tile rows
// 2 .. R+1): J u, or 0 on rows past nv (the rFFT's zero rows).
template <int KIND, int R, typename T, typename RF, typename RH>
__device__ __forceinline__ void jac_walk(const RF &rF, const RH &rH, int j0, int nv, T ht_inv, T factor, T af1,
                                         T af2, T (&res)[R])
{
    const int cx = threadIdx.x + 2;
    ArakawaWin<T> win;
    win.load(rF, rH, 2, cx);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int j = j0 + q;
        T out = T(0);
        if (j < nv) {
            const T base = jacobian_base(ht_inv, win.f0[2]);
            out = (j >= 2 && j < nv - 2) ? arakawa_add<KIND>(win, base, factor, af1, af2) : base;
        }
        res[q] = out;
        if (q + 1 < R) win.advance(rF, rH, q + 3, cx);
    }
}

template <int KIND, int R, int NS, typename T>
__device__ __forceinline__ void jac_rfft_body(FftArgs<T> a, JacRows jr, const T *__restrict__ u,
                                              cx_t<T> *__restrict__ spec)
{
    using C = cx_t<T>;
    static_assert(NT == kJrSW, "a strip is one column per thread");
    constexpr int RH = R + 4, SP = NT + 4, nx = NS * NT;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    const int b = blockIdx.y;
    u += b * a.N;
    spec += b * a.S;
    const size_t xo = (size_t)b * nx;
    // the FFT buffer reuses the strips' LDS: the strips' results wait in
    // registers until the last walk is done
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw);
    T *sF = reinterpret_cast<T *>(lds_raw);
    T *sH = sF + RH * SP;
    const int nv = a.nv, t = threadIdx.x;
    const int j0 = xcd_tile(blockIdx.x, gridDim.x) * R;
    // FP32: H_ave - href, href = h0 at the centre row of the bracket kernel's
    // 16-row band that holds the workgroup's rows (R divides 16), as the FP32
    // bracket kernel stages it
    const int jc = ((j0 >> 4) << 4) + 8;
    const double href = (sizeof(T) == 4 && jc < nv) ? jr.h0[jc] : 0.0;
    const double usd =
        jr.uscale != nullptr ? *reinterpret_cast<const double *>(reinterpret_cast<const char *>(jr.uscale) + b * a.gstride)
                             : 1.0;
    double h0r[RH];
#pragma unroll
    for (int r = 0; r < RH; ++r) {
        const int gj = jr_row(j0, r, nv);
        h0r[r] = gj >= 0 ? jr.h0[gj] : 0.0;
    }
    JacStrip<R, T> g;
    jac_strip_load<R>(u, jr, xo, nx, nv, j0, 0, g);
    if (dflag) return;                 // uniform over the block, before any barrier
    const T us = (T)usd, ht_inv = (T)jr.ht_inv, factor = (T)jr.factor;
    const T af1 = (T)jr.inst[b].af1, af2 = (T)jr.inst[b].af2;
    T res[NS][R];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        // stage the strip: u times its scale (== the VecScale'd Krylov
        // vector) and H_ave (0 outside the grid), as the bracket kernel does
#pragma unroll
        for (int r = 0; r < RH; ++r) {
            sF[r * SP + t + 2] = g.u[r] * us;
            sH[r * SP + t + 2] = jr_row(j0, r, nv) >= 0 ? have_rel<T>(h0r[r], g.x, href) : T(0);
        }
        if (t < 4 * RH) {
            int hr, tc;
            jr_halo_cell(t, RH, &hr, &tc);
            sF[hr * SP + tc] = g.eu * us;
            sH[hr * SP + tc] = jr_row(j0, hr, nv) >= 0 ? have_rel<T>(g.eh0, g.ex, href) : T(0);
        }
        __syncthreads();
        if (s + 1 < NS) jac_strip_load<R>(u, jr, xo, nx, nv, j0, (s + 1) * NT, g);
        jac_walk<KIND, R>(TileRd<T, SP>{sF}, TileRd<T, SP>{sH}, j0, nv, ht_inv, factor, af1, af2, res[s]);
        __syncthreads();
    }
    // twiddles in flight during the first stage (not held through the walks)
    FftTw<false, C> pt;
    pt.template load<NS == 4>(a.tw, nx, a.logn);
    if (NS == 4) {
        // stage 0 of the transform in registers, as rfft_kernel: butterfly
        // (p, t) of complex row p (real rows 2p, 2p+1) takes columns t + m NT
#pragma unroll
        for (int p = 0; p < (R >> 1); ++p) {
            const C v0 = cmk<C>(res[0][2 * p], res[0][2 * p + 1]), v1 = cmk<C>(res[1 % NS][2 * p], res[1 % NS][2 * p + 1]);
            const C v2 = cmk<C>(res[2 % NS][2 * p], res[2 % NS][2 * p + 1]);
            const C v3 = cmk<C>(res[3 % NS][2 * p], res[3 % NS][2 * p + 1]);
            const C a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), d = csub(v1, v3);
            const C a3 = cmk<C>(d.y, -d.x);             // (v1 - v3) * (-i)
            C *z = buf + (p << a.logn) + (t << 2);      // stage 0 outputs: 4t + r
            z[0] = cadd(a0, a2);
            z[1] = cadd(a1, a3);
            z[2] = csub(a0, a2);
            z[3] = csub(a1, a3);
        }
    } else {
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int p = 0; p < (R >> 1); ++p) buf[p * nx + s * NT + t] = cmk<C>(res[s][2 * p], res[s][2 * p + 1]);
    }
    __syncthreads();
    lds_fft<false, true, NS == 4>(buf, nx, a.logn, R >> 1, a.tw, pt, NS == 4 ? 1 : 0, LdsSink<C>{buf});
    store_real_spectrum(buf, nx, R, j0, nv, spec);
}

// LDS max(R/2 nx 16, 2 (R+4)(NT+4) 8) bytes (33 KB at R = 4), so three
// workgroups per CU fit when the registers do (R <= 4; four cost register
// spills and were slower)
template <int KIND, int R, int NS, typename T>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(R <= 4 ? 3 : 2))) void jac_rfft_kernel(
    FftArgs<T> a, JacRows jr, const T *__restrict__ u, cx_t<T> *__restrict__ spec)
{
    jac_rfft_body<KIND, R, NS>(a, jr, u, spec);
}

// Real rows 2b, 2b+1 of an irFFT workgroup from its complex transform b
// (backward-normalised by 1/nx, exact: nx is a power of two).
template <typename T>
struct IrfftSink {
    T *y;
    int j0, nx, logn, nv;
    __device__ __forceinline__ void operator()(int pos, cx_t<T> z) const
    {
        const int b = pos >> logn, i = pos & (nx - 1);
        const T inv_n = T(1) / (T)nx;
        const int r = j0 + 2 * b;
        if (r < nv) y[(size_t)r * nx + i] = z.x * inv_n;
        if (r + 1 < nv) y[(size_t)(r + 1) * nx + i] = z.y * inv_n;
    }
};

template <typename T, bool PRE, bool P8>
__global__ __launch_bounds__(NT) void irfft_kernel(FftArgs<T> a, const cx_t<T> *__restrict__ spec, T *__restrict__ y)
{
    using C = cx_t<T>;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    spec += blockIdx.y * a.S;
    y += blockIdx.y * a.N;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw);
    const int nx = a.nx, R = a.rows, nh = nx >> 1;
    const int j0 = xcd_tile(blockIdx.x, gridDim.x) * R;   // as in rfft_kernel
    const int npair = R >> 1;
    // Z_k = A_k + i B_k (k <= nx/2) and its Hermitian extension for k > nx/2;
    // every global load of the thread is issued before the LDS stores.
    const int total = (nh + 1) * npair;
    constexpr int PER = (MAXE / 2 + 64 + NT - 1) / NT;   // (nx/2 + 1) R/2 <= MAXE/2 + 64
    C va[PER], vb[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int k = idx / npair;
        const int ja = j0 + 2 * (idx - k * npair);
        va[q] = load_or_zero(spec, (int64_t)k * a.nv + ja, idx < total && ja < a.nv);
        vb[q] = load_or_zero(spec, (int64_t)k * a.nv + ja + 1, idx < total && ja + 1 < a.nv);
    }
    FftTw<true, C> pt;
    if (PRE) pt.template load<P8>(a.tw, nx, a.logn);
    if (dflag) return;                 // uniform over the block, before any barrier
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < total) {
            const int k = idx / npair;
            const int p = idx - k * npair;
            C A = va[q], B = vb[q];
            if (k == 0 || k == nh) {
                A.y = T(0);
                B.y = T(0);
            }
            C *z = buf + p * nx;
            z[k] = cmk<C>(A.x - B.y, A.y + B.x);
            if (k != 0 && k != nh) z[nx - k] = cmk<C>(A.x + B.y, B.x - A.y);
        }
    }
    __syncthreads();
    // the last stage's outputs (natural order) go straight to the output rows
    lds_fft<true, PRE, P8>(buf, nx, a.logn, npair, a.tw, pt, 0, IrfftSink<T>{y, j0, nx, a.logn, a.nv});
}

// Real rows of the irFFT left in the workgroup's LDS buffer itself, natural
// order (row r of the block at wr[r nx + i]), backward-normalised as in
// IrfftSink.  The last FFT stage reads every input before its barrier, so
// the buffer can take its outputs.
template <typename T>
struct RealRowsSink {
    T *wr;
    int nx, logn;
    __device__ __forceinline__ void operator()(int pos, cx_t<T> z) const
    {
        const int b = pos >> logn, i = pos & (nx - 1);
        const T inv_n = T(1) / (T)nx;
        wr[(2 * b) * nx + i] = z.x * inv_n;
        wr[(2 * b + 1) * nx + i] = z.y * inv_n;
    }
};

// The Arnoldi step's irFFT fused with the classical Gram-Schmidt dots that
// follow it: w = V_{col+1} = irfft(spec) is written to the basis and, from
// the workgroup's LDS, dotted with v_0..v_col.  The workgroup's R rows are
// one CGS slice of NT EPT 16-byte vectors (R nx = NT EPT L, nv % R == 0);
// thread t takes the slice's vectors t + q NT, and the per-slice partials
// part[slice PV + j] are summed by maxpy_norm_kernel as mdot_kernel's are.
// Where EPT equals the CGS launches' own, w and the partials are
// bit-identical to irfft_kernel followed by mdot_kernel.  What the fusion
// saves: the second read of w, one launch, and the wait for the first basis
// vectors, whose loads are in flight while the FFT stages run.
template <typename T, int EPT, int PV, bool P8>
__global__ __launch_bounds__(NT) void irfft_mdot_kernel(FftArgs<T> a, const cx_t<T> *__restrict__ spec,
                                                        CgsArgs<T> g, int col)
{
    using C = cx_t<T>;
    using V = typename V16<T>::type;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    spec += blockIdx.y * a.S;
    const T *basis = g.basis + blockIdx.y * g.N;
    const GmresState *gs = g.g + blockIdx.y;
    double *part = g.part + blockIdx.y * g.pstride;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *buf = reinterpret_cast<C *>(lds_raw);
    __shared__ double red[kCgsNW][PV];
    __shared__ double s_scale[PV];
    const int nx = a.nx, R = a.rows, nh = nx >> 1;
    const int slice = xcd_tile(blockIdx.x, gridDim.x);
    const int j0 = slice * R;
    const int npair = R >> 1;
    const int nvec = col + 1;
    const int total = (nh + 1) * npair;
    constexpr int PER = (MAXE / 2 + 64 + NT - 1) / NT;
    C va[PER], vb[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int k = idx / npair;
        const int ja = j0 + 2 * (idx - k * npair);
        va[q] = load_or_zero(spec, (int64_t)k * a.nv + ja, idx < total);
        vb[q] = load_or_zero(spec, (int64_t)k * a.nv + ja + 1, idx < total);
    }
    const double sc = (int)threadIdx.x < nvec ? gs->scale[threadIdx.x] : 0.0;
    FftTw<true, C> pt;
    pt.template load<P8>(a.tw, nx, a.logn);
    if (dflag) return;                 // uniform over the block, before any barrier
    // the first pair of basis vectors is in flight while the FFT stages run
    const int64_t nvv = g.N / V16<T>::L;
    const int64_t e0 = (int64_t)slice * NT * EPT + threadIdx.x;
    V x[2][EPT];
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
        const V *Vp = reinterpret_cast<const V *>(basis + (size_t)gg * g.vstride);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int64_t e = e0 + (int64_t)q * NT;
            x[gg][q] = (gg < nvec && e < nvv) ? Vp[e] : V{};
        }
    }
    if ((int)threadIdx.x < PV) s_scale[threadIdx.x] = sc;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < total) {
            const int k = idx / npair;
            const int p = idx - k * npair;
            C A = va[q], B = vb[q];
            if (k == 0 || k == nh) {
                A.y = T(0);
                B.y = T(0);
            }
            C *z = buf + p * nx;
            z[k] = cmk<C>(A.x - B.y, A.y + B.x);
            if (k != 0 && k != nh) z[nx - k] = cmk<C>(A.x + B.y, B.x - A.y);
        }
    }
    __syncthreads();
    T *wr = reinterpret_cast<T *>(buf);
    lds_fft<true, true, P8>(buf, nx, a.logn, npair, a.tw, pt, 0, RealRowsSink<T>{wr, nx, a.logn});

    // w of the slice: to the basis, and into registers for the dots
    const V *wl = reinterpret_cast<const V *>(wr);
    V *W = reinterpret_cast<V *>(const_cast<T *>(basis) + (size_t)nvec * g.vstride);
    V w[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        w[q] = wl[threadIdx.x + q * NT];
        W[e0 + (int64_t)q * NT] = w[q];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = 0;
    for (; v + 2 <= nvec; v += 2) {
        if (v > 0) load_group<EPT, 2, T>(basis, g.vstride, v, e0, nvv, x);
#pragma unroll
        for (int gg = 0; gg < 2; ++gg) {
            double acc = dot_slice<EPT, T>(w, x[gg], s_scale[v + gg]);
            acc = wave_sum(acc);
            if (lane == 0) red[wave][v + gg] = acc;
        }
    }
    if (v < nvec) {
        V y[1][EPT];
        if (v > 0) {
            load_group<EPT, 1, T>(basis, g.vstride, v, e0, nvv, y);
        } else {
#pragma unroll
            for (int q = 0; q < EPT; ++q) y[0][q] = x[0][q];
        }
        double acc = dot_slice<EPT, T>(w, y[0], s_scale[v]);
        acc = wave_sum(acc);
        if (lane == 0) red[wave][v] = acc;
    }
    __syncthreads();
    if (threadIdx.x < PV) {
        double s = 0.0;
        if ((int)threadIdx.x < nvec)
            for (int wv = 0; wv < kCgsNW; ++wv) s += red[wv][threadIdx.x];
        part[(size_t)slice * PV + threadIdx.x] = s;
    }
}

// Transforms of a length n that is not a power of two (n = 100, 201, ...):
// a direct real DFT with exact twiddle indices (k m mod n into the table of
// exp(-2 pi i m/n)), same spectrum layout and c2r semantics as the FFT
// kernels above.  O(n) work per output; the workgroup's rows and, when they
// fit 32 KiB, the twiddles are staged in LDS.  Rows are not packed in pairs,
// so `rows` may be odd.
template <typename T, bool TWL>
__global__ __launch_bounds__(NT) void rdft_kernel(FftArgs<T> a, const T *__restrict__ x, cx_t<T> *__restrict__ spec)
{
    using C = cx_t<T>;
    if (cycle_done(a.done, a.gstride)) return;      // uniform over the block, before any barrier
    x += blockIdx.y * a.N;
    spec += blockIdx.y * a.S;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int n = a.nx, R = a.rows, ld = n | 1, nk = n / 2 + 1;
    C *stw = reinterpret_cast<C *>(lds_raw);
    T *sx = reinterpret_cast<T *>(lds_raw + (TWL ? (size_t)n * sizeof(C) : 0));
    const int j0 = blockIdx.x * R;
    const int nr = a.nv - j0 < R ? a.nv - j0 : R;   // rows of this block
    for (int e = threadIdx.x; e < nr * n; e += NT) {
        const int r = e / n;
        sx[r * ld + (e - r * n)] = x[(size_t)j0 * n + e];
    }
    if (TWL)
        for (int e = threadIdx.x; e < n; e += NT) stw[e] = a.tw[e];
    __syncthreads();
    const C *tw = TWL ? stw : a.tw;
    for (int o = threadIdx.x; o < nk * nr; o += NT) {
        const int k = o / nr, r = o - k * nr;
        const T *xr = sx + r * ld;
        T re = T(0), im = T(0);
        for (int m = 0, t = 0; m < n; ++m) {
            const C w = tw[t];
            re += xr[m] * w.x;
            im += xr[m] * w.y;
            t += k;
            if (t >= n) t -= n;
        }
        spec[(size_t)k * a.nv + j0 + r] = cmk<C>(re, im);
    }
}

// y[j][m] = (Re X_0 + 2 sum_{k=1}^{(n-1)/2} Re(X_k e^{2 pi i k m/n}) [+ Re X_{n/2} (-1)^m]) / n
template <typename T, bool TWL>
__global__ __launch_bounds__(NT) void irdft_kernel(FftArgs<T> a, const cx_t<T> *__restrict__ spec, T *__restrict__ y)
{
    using C = cx_t<T>;
    if (cycle_done(a.done, a.gstride)) return;      // uniform over the block, before any barrier
    spec += blockIdx.y * a.S;
    y += blockIdx.y * a.N;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int n = a.nx, R = a.rows, nk = n / 2 + 1, K = (n - 1) / 2;
    C *stw = reinterpret_cast<C *>(lds_raw);
    C *ss = reinterpret_cast<C *>(lds_raw + (TWL ? (size_t)n * sizeof(C) : 0));
    const int j0 = blockIdx.x * R;
    const int nr = a.nv - j0 < R ? a.nv - j0 : R;
    for (int e = threadIdx.x; e < nk * nr; e += NT) {
        const int k = e / nr, r = e - k * nr;
        C v = spec[(size_t)k * a.nv + j0 + r];
        if (k == 0 || 2 * k == n) v.y = T(0);
        ss[r * nk + k] = v;
    }
    if (TWL)
        for (int e = threadIdx.x; e < n; e += NT) stw[e] = a.tw[e];
    __syncthreads();
    const C *tw = TWL ? stw : a.tw;
    const T inv_n = T(1) / (T)n;
    for (int o = threadIdx.x; o < nr * n; o += NT) {
        const int r = o / n, m = o - r * n;
        const C *X = ss + r * nk;
        T acc = T(0);
        for (int k = 1, t = m; k <= K; ++k) {
            const C w = tw[t];
            acc += X[k].x * w.x + X[k].y * w.y;
            t += m;
            if (t >= n) t -= n;
        }
        T v = X[0].x + T(2) * acc;
        if (!(n & 1)) v += (m & 1) ? -X[n / 2].x : X[n / 2].x;
        y[(size_t)j0 * n + o] = v * inv_n;
    }
}

struct TriArgs {
    int nv;           // system size
    int P;            // chunks (threads), power of two
    int64_t S, gstride;
    const double2 *eigen;
    const double *gm, *g0, *gp;
    int64_t cstride;         // instance stride of gm/g0/gp (0: shared)
    double clo, cmid, chi;   // band constants (kinetic 0.5, 1.0, 0.5; potential 1.0, 2.0, 1.0)
    const InstConst *inst;   // band factor afpc of instance blockIdx.y
    double ht_inv;
    const int32_t *done;
};

// Row j of T_k (KineticFast.pyx:248-272): sub a_j = diagm[j] (j >= 1),
// diagonal b_j = diag[j], super c_j = diagp[j] (j <= nv-2); rows past nv are
// identity rows that pad the system to P*M unknowns.  The coefficients are
// formed in FP64 (as the reference forms its band matrix) and rounded to C.
// TR: row j of the conjugate transpose T_k^H (sub conj(c_{j-1}), diagonal
// conj(b_j), super conj(a_{j+1})), the mode-k block of the transposed real
// operator P^T (P^-T x = irfft(T_k^-H rfft(x)): transposing a real
// x-circulant operator maps its symbol at k to the conjugate at -k).
// the three real per-row factors row j of T_k (or of T_k^H, TR) uses
template <bool TR>
__device__ __forceinline__ void tri_load(const TriArgs &a, int j, double &lo, double &mid, double &hi)
{
    // Clamped addresses and no selects on the loaded values: nothing here
    // waits on a load, so the loads of all rows of a chunk are in flight
    // together.  Values loaded for rows outside the system are never used:
    // tri_row makes those rows identity rows and drops the sub-diagonal of
    // row 0 and the super-diagonal of row nv-1.
    const int jc = j < a.nv ? j : a.nv - 1;
    if (!TR) {
        lo = a.gm[jc];
        hi = a.gp[jc];
    } else {
        lo = a.gp[jc > 0 ? jc - 1 : 0];
        hi = a.gm[jc < a.nv - 1 ? jc + 1 : jc];
    }
    mid = a.g0[jc];
}

template <bool TR, typename C>
__device__ __forceinline__ void tri_row(const TriArgs &a, double af, double2 e, int j, double lo, double mid,
                                        double hi, C &aj, C &bj, C &cj)
{
    using R = typename C::value_type;
    const double ey = TR ? -e.y : e.y;
    if (j < a.nv) {
        aj = (j > 0) ? cmk<C>((R)(((e.x * a.clo) * lo) * af), (R)(((ey * a.clo) * lo) * af)) : cmk<C>(R(0), R(0));
        bj = cmk<C>((R)(((e.x * a.cmid) * mid) * af + a.ht_inv), (R)(((ey * a.cmid) * mid) * af));
        cj = (j < a.nv - 1) ? cmk<C>((R)(((e.x * a.chi) * hi) * af), (R)(((ey * a.chi) * hi) * af))
                            : cmk<C>(R(0), R(0));
    } else {
        aj = cmk<C>(R(0), R(0));
        bj = cmk<C>(R(1), R(0));
        cj = cmk<C>(R(0), R(0));
    }
}

template <int M, typename T, bool TR>
__global__ __launch_bounds__(NT) void tridiag_kernel(TriArgs a, cx_t<T> *__restrict__ spec)
{
    using C = cx_t<T>;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    spec += blockIdx.y * a.S;
    a.gm += blockIdx.y * a.cstride;
    a.g0 += blockIdx.y * a.cstride;
    a.gp += blockIdx.y * a.cstride;
    const double af = a.inst[blockIdx.y].afpc;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *sd = reinterpret_cast<C *>(lds_raw);   // P*(M+1): right-hand sides / solution, chunk c at c*(M+1)
    __shared__ C sa[NT], sb[NT], sc[NT], se[NT];
    const int k = blockIdx.x;
    const double2 e = a.eigen[k];
    const int P = a.P, n = P * M;
    const C zero = cmk<C>(T(0), T(0));
    C *S = spec + (size_t)k * a.nv;
    const int c = threadIdx.x;
    double clo[M], cmid[M], chi[M];   // per-row factors of the thread's chunk
    {
        // n = P*M <= NT*M: every load of the thread in flight before the LDS stores
        C v[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = threadIdx.x + q * NT;
            v[q] = load_or_zero(S, j, j < a.nv);
        }
        if (dflag) return;             // uniform over the block, before any barrier
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = threadIdx.x + q * NT;
            if (j < n) sd[j + j / M] = v[q];      // one pad slot per chunk: no bank conflicts
        }
    }
    __syncthreads();

    C Am[M], Cm[M], Dm[M];
    C ae, be, ce;          // coefficients of the chunk's last row
    if (c < P) {
        const int s = c * M, ls = c * (M + 1);
        // forward sweep: y_j + c'_j y_{j+1} = d'_j - a'_j y_{s-1}
        C aj, bj, cj;
        tri_load<TR>(a, s, clo[0], cmid[0], chi[0]);
        tri_row<TR>(a, af, e, s, clo[0], cmid[0], chi[0], aj, bj, cj);
        C r = crecip(bj);
        Cm[0] = cmul(cj, r);
        Dm[0] = cmul(sd[ls], r);
        Am[0] = cmul(aj, r);
#pragma unroll
        for (int q = 1; q < M; ++q) {
            tri_load<TR>(a, s + q, clo[q], cmid[q], chi[q]);
            tri_row<TR>(a, af, e, s + q, clo[q], cmid[q], chi[q], aj, bj, cj);
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
    C al = zero, bt = cmk<C>(T(1), T(0)), gm = zero, dl = zero;
    if (c < P) {
        const C dE = sd[c * (M + 1) + M - 1];
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
            C nal = zero, ngm = zero;
            C nbt = bt, ndl = dl;
            if (c - st >= 0) {
                const C k1 = cdiv(al, sb[c - st]);
                nal = cneg(cmul(sa[c - st], k1));
                nbt = csub(nbt, cmul(sc[c - st], k1));
                ndl = csub(ndl, cmul(se[c - st], k1));
            }
            if (c + st < P) {
                const C k2 = cdiv(gm, sb[c + st]);
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
        const C Eprev = (c > 0) ? se[c - 1] : zero;
        const C Eown = se[c];
#pragma unroll
        for (int q = 0; q < M - 1; ++q)
            sd[ls + q] = csub(csub(Dm[q], cmul(Am[q], Eprev)), cmul(Cm[q], Eown));
        sd[ls + M - 1] = Eown;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < a.nv; j += NT) S[j] = sd[j + j / M];
}

// ---- kinetic solve with precomputed pivots ---------------------------------
// The kinetic T_k do not change during a run, so their LU factors are formed
// once per context (KineticFast.pyx:88-90 factors once with zgbtrf as well):
// without pivoting (the bands are diagonally dominant: |Im b_j| is about
// twice |a_j| + |c_j|, so zgbtrf's partial pivoting swaps no rows),
//   u_0 = b_0,  u_j = b_j - (a_j / u_{j-1}) c_{j-1},   rp[k][j] = 1 / u_j,
// in FP64, one thread per mode (and instance), stored as C (FP64, or rounded
// for the FP32 solve).  The solve then has no divisions:
//   forward   y_j = d_j - (a_j rp_{j-1}) y_{j-1},
//   backward  x_j = rp_j (y_j - c_j x_{j+1}),
// a_j and c_j being rebuilt from e_k and the per-row factors as in tri_row.
template <typename T>
__global__ __launch_bounds__(NT) void tridiag_factor_kernel(TriArgs a, int nmodes, int M, cx_t<T> *__restrict__ rp)
{
    const int k = blockIdx.x * NT + threadIdx.x;
    if (k >= nmodes) return;
    const double af = a.inst[blockIdx.y].afpc;
    const double2 e = a.eigen[k];
    // chunk-interleaved: the pivot of row j = t M + q at q NT + t, so that the
    // solve's thread t, which owns rows t M .. t M + M - 1, loads them
    // coalesced; rows nv .. NT M - 1 hold zero
    const int L = NT * M;
    cx_t<T> *R = rp + ((size_t)blockIdx.y * nmodes + k) * L;
    double lo, mid, hi;
    double2 aj, bj, cj;
    tri_load<false>(a, 0, lo, mid, hi);
    tri_row<false>(a, af, e, 0, lo, mid, hi, aj, bj, cj);
    double2 r = crecip(bj);
    R[0] = cmk<cx_t<T>>((T)r.x, (T)r.y);
    for (int j = 1; j < a.nv; ++j) {
        const double2 cprev = cj;
        tri_load<false>(a, j, lo, mid, hi);
        tri_row<false>(a, af, e, j, lo, mid, hi, aj, bj, cj);
        r = crecip(csub(bj, cmul(cmul(aj, r), cprev)));
        R[(j % M) * NT + j / M] = cmk<cx_t<T>>((T)r.x, (T)r.y);
    }
    for (int j = a.nv; j < L; ++j) R[(j % M) * NT + j / M] = cmk<cx_t<T>>(T(0), T(0));
}

template <typename C> __device__ __forceinline__ C shfl_up_cx(C v, int off)
{
    return cmk<C>(__shfl_up(v.x, off), __shfl_up(v.y, off));
}
template <typename C> __device__ __forceinline__ C shfl_down_cx(C v, int off)
{
    return cmk<C>(__shfl_down(v.x, off), __shfl_down(v.y, off));
}

// One workgroup per mode (blockIdx.x) and instance (blockIdx.y); thread t
// owns the chunk of M rows [t M, t M + M).  The spectrum row and its pivots
// are staged in LDS (coalesced loads).  Across chunks the recurrences are
// affine maps y_out = A + B y_in, composed by a scan over the wavefront's
// lanes (shuffles) and over the 4 wavefronts (one LDS exchange), first
// forward, then backward; no thread waits on a division.
template <int M, typename T>
__global__ __launch_bounds__(NT) void tridiag_lu_kernel(TriArgs a, const cx_t<T> *__restrict__ rp,
                                                        cx_t<T> *__restrict__ spec)
{
    using C = cx_t<T>;
    const int32_t dflag = cycle_done(a.done, a.gstride);
    if (gridDim.y > 1 && dflag) return;   // batched: a finished instance skips its loads
    const int k = blockIdx.x;
    const double af = a.inst[blockIdx.y].afpc;
    const double2 e = a.eigen[k];
    C *S = spec + blockIdx.y * a.S + (size_t)k * a.nv;
    // pivots in the chunk-interleaved order of tridiag_factor_kernel: thread
    // t's rows t M + q at q NT + t, loaded coalesced straight into registers
    const C *Rg = rp + ((size_t)blockIdx.y * gridDim.x + k) * (NT * M);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *sd = reinterpret_cast<C *>(lds_raw);      // NT (M+1): right-hand side, then y, then x
    __shared__ C wa[NT / 64], wb[NT / 64];
    const C zero = cmk<C>(T(0), T(0)), one = cmk<C>(T(1), T(0));
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    C rpv[M];                                     // pivots of rows t M .. t M + M - 1
    C rprev;                                      // pivot of row t M - 1 (last row of chunk t - 1)
    {
        C v[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = t + q * NT;
            v[q] = load_or_zero(S, j, j < a.nv);
            rpv[q] = Rg[q * NT + t];
        }
        rprev = load_or_zero(Rg, (M - 1) * NT + t - 1, t > 0);
        if (dflag) return;             // uniform over the block, before any barrier
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = t + q * NT;
            sd[j + j / M] = v[q];     // one pad slot per chunk
        }
    }
    __syncthreads();
    const int s = t * M, ls = t * (M + 1);
    // sub- and super-diagonals of the chunk's rows (rows past nv: identity rows)
    C av[M], cv[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
        double lo, mid, hi;
        C bj;
        tri_load<false>(a, s + q, lo, mid, hi);
        tri_row<false>(a, af, e, s + q, lo, mid, hi, av[q], bj, cv[q]);
    }
    // forward map of the chunk: y_{s+M-1} = A + B y_{s-1}
    C A = sd[ls], B = cneg(cmul(av[0], rprev));
#pragma unroll
    for (int q = 1; q < M; ++q) {
        const C m = cmul(av[q], rpv[q - 1]);
        A = csub(sd[ls + q], cmul(m, A));
        B = cneg(cmul(m, B));
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {     // inclusive scan, earlier lanes first
        const C pa = shfl_up_cx(A, off), pb = shfl_up_cx(B, off);
        if (lane >= off) {
            A = cadd(A, cmul(B, pa));
            B = cmul(B, pb);
        }
    }
    if (lane == 63) {
        wa[wave] = A;
        wb[wave] = B;
    }
    C ea = shfl_up_cx(A, 1), eb = shfl_up_cx(B, 1);   // lanes 0 .. lane-1
    if (lane == 0) {
        ea = zero;
        eb = one;
    }
    __syncthreads();
    C y = zero;                                   // y at the end of the previous wavefronts' rows
    for (int w = 0; w < wave; ++w) y = cadd(wa[w], cmul(wb[w], y));
    y = cadd(ea, cmul(eb, y));                    // y_{s-1}
    y = csub(sd[ls], cmul(cmul(av[0], rprev), y));
    sd[ls] = y;
#pragma unroll
    for (int q = 1; q < M; ++q) {
        y = csub(sd[ls + q], cmul(cmul(av[q], rpv[q - 1]), y));
        sd[ls + q] = y;
    }
    // backward map of the chunk: x_s = G + D x_{s+M}
    C G = cmul(rpv[M - 1], sd[ls + M - 1]), D = cneg(cmul(rpv[M - 1], cv[M - 1]));
#pragma unroll
    for (int q = M - 2; q >= 0; --q) {
        const C rc = cmul(rpv[q], cv[q]);
        G = csub(cmul(rpv[q], sd[ls + q]), cmul(rc, G));
        D = cneg(cmul(rc, D));
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {     // inclusive scan, later lanes first
        const C na = shfl_down_cx(G, off), nb = shfl_down_cx(D, off);
        if (lane + off < 64) {
            G = cadd(G, cmul(D, na));
            D = cmul(D, nb);
        }
    }
    __syncthreads();                              // every thread has read wa / wb
    if (lane == 0) {
        wa[wave] = G;
        wb[wave] = D;
    }
    C xa = shfl_down_cx(G, 1), xb = shfl_down_cx(D, 1);   // lanes lane+1 .. 63
    if (lane == 63) {
        xa = zero;
        xb = one;
    }
    __syncthreads();
    C x = zero;                                   // x at the first row of the next wavefronts
    for (int w = NT / 64 - 1; w > wave; --w) x = cadd(wa[w], cmul(wb[w], x));
    x = cadd(xa, cmul(xb, x));                    // x_{s+M}
#pragma unroll
    for (int q = M - 1; q >= 0; --q) {
        x = cmul(rpv[q], csub(sd[ls + q], cmul(cv[q], x)));
        sd[ls + q] = x;
    }
    __syncthreads();
    for (int j = t; j < a.nv; j += NT) S[j] = sd[j + j / M];
}

// out = in^T for every instance (blockIdx.z): in holds R rows of C, out C
// rows of R.  32x32 tiles through LDS (one pad column), loads and stores
// coalesced along the contiguous axis of each side; clamped addresses for the
// edge tiles.  The potential preconditioner transforms along v, so it works on
// the transposed field.
template <typename T>
__global__ __launch_bounds__(NT) void transpose_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int C,
                                                       int64_t N, const int32_t *done, int64_t gstride)
{
    const int b = blockIdx.z;
    if (done != nullptr && *reinterpret_cast<const int32_t *>(reinterpret_cast<const char *>(done) + b * gstride))
        return;
    in += (size_t)b * N;
    out += (size_t)b * N;
    __shared__ T tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;      // 32 x 8
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + ty + 8 * q, cc = c0 + tx;
        const bool ok = r < R && cc < C;
        const T t = in[(size_t)(ok ? r : 0) * C + (ok ? cc : 0)];
        v[q] = ok ? t : T(0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) tile[ty + 8 * q][tx] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int oc = c0 + ty + 8 * q, orr = r0 + tx;
        if (oc < C && orr < R) out[(size_t)oc * R + orr] = tile[tx][ty + 8 * q];
    }
}

// The phi differences of the potential bands (TensorProductPotentialFast.pyx:
// 243-252): gm_i = phi_i - phi_{i-1}, g0_i = phi_{i+1} - phi_{i-1},
// gp_i = phi_{i+1} - phi_i, indices mod nx, per instance (blockIdx.y).
__global__ __launch_bounds__(NT) void potential_bands_kernel(const double *__restrict__ phi, int nx,
                                                             double *__restrict__ gm, double *__restrict__ g0,
                                                             double *__restrict__ gp)
{
    const size_t o = (size_t)blockIdx.y * nx;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= nx) return;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    gm[o + i] = phi[o + i] - phi[o + im];
    g0[o + i] = phi[o + ip] - phi[o + im];
    gp[o + i] = phi[o + ip] - phi[o + i];
}

template <typename T>
hipError_t transpose_t(vv_ctx *c, const T *in, T *out, bool back, const int32_t *done)
{
    const int R = back ? c->p.nx : c->p.nv, C = back ? c->p.nv : c->p.nx;
    dim3 grid((C + 31) / 32, (R + 31) / 32, c->B);
    hipLaunchKernelGGL(transpose_kernel<T>, grid, dim3(NT), 0, c->stream, in, out, R, C, c->N, done,
                       (int64_t)sizeof(GmresState));
    return hipGetLastError();
}

// v-rows per FFT workgroup: the context's tuned value, else the largest that
// fits 64 KiB of LDS while the launch keeps >= 512 workgroups
int fft_rows(const vv_ctx *c)
{
    if (c->tune_fft_rows > 0) return c->tune_fft_rows;
    const int nx = c->fft_n, nv = c->fft_rows;
    int R = 2;
    while (2 * R * nx <= 8192) R *= 2;
    if (R < 2) R = 2;
    while (R > 2 && (nv + R - 1) / R < 512) R >>= 1;
    return R;
}

// The DFT kernels' launch: rows per workgroup = fft_rows(c) capped so that
// the rows take <= 32 KiB of LDS; twiddles in LDS when they fit 32 KiB.
template <typename T>
FftArgs<T> dft_args(vv_ctx *c, const cx_t<T> *tw, const int32_t *done, bool *twl, size_t *lds)
{
    const int n = c->fft_n;
    const size_t row_bytes = (size_t)(n + 2) * sizeof(T);   // >= (n | 1) reals and >= (n/2 + 1) complex
    int R = fft_rows(c);
    const int cap = (int)(32768 / row_bytes);
    if (R > cap) R = cap > 1 ? cap : 1;
    *twl = (size_t)n * sizeof(cx_t<T>) <= 32768;
    *lds = (*twl ? (size_t)n * sizeof(cx_t<T>) : 0) + (size_t)R * row_bytes;
    return FftArgs<T>{n, 0, c->fft_rows, R, c->N, spec_stride(c), (int64_t)sizeof(GmresState), tw, done};
}

template <typename T>
hipError_t rdft_t(vv_ctx *c, const cx_t<T> *tw, const T *x, cx_t<T> *spec, const int32_t *done)
{
    bool twl;
    size_t lds;
    const FftArgs<T> a = dft_args<T>(c, tw, done, &twl, &lds);
    const dim3 grid((a.nv + a.rows - 1) / a.rows, c->B);
    if (twl) launch_timed(c, rdft_kernel<T, true>, grid, dim3(NT), (uint32_t)lds, a, x, spec);
    else launch_timed(c, rdft_kernel<T, false>, grid, dim3(NT), (uint32_t)lds, a, x, spec);
    return hipGetLastError();
}

template <typename T>
hipError_t irdft_t(vv_ctx *c, const cx_t<T> *tw, const cx_t<T> *spec, T *y, const int32_t *done)
{
    bool twl;
    size_t lds;
    const FftArgs<T> a = dft_args<T>(c, tw, done, &twl, &lds);
    const dim3 grid((a.nv + a.rows - 1) / a.rows, c->B);
    if (twl) launch_timed(c, irdft_kernel<T, true>, grid, dim3(NT), (uint32_t)lds, a, spec, y);
    else launch_timed(c, irdft_kernel<T, false>, grid, dim3(NT), (uint32_t)lds, a, spec, y);
    return hipGetLastError();
}

template <typename T>
hipError_t rfft_t(vv_ctx *c, const cx_t<T> *tw, const T *x, cx_t<T> *spec, const int32_t *done)
{
    if (c->fft_n & (c->fft_n - 1)) return rdft_t<T>(c, tw, x, spec, done);
    FftArgs<T> a{c->fft_n, c->logn, c->fft_rows, fft_rows(c), c->N, spec_stride(c),
                 (int64_t)sizeof(GmresState), tw, done};
    const int nb = (a.nv + a.rows - 1) / a.rows;
    const size_t lds = (size_t)(a.rows >> 1) * a.nx * sizeof(cx_t<T>);
    if (a.nx == N8) launch_timed(c, rfft_kernel<T, true, true>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, x, spec);
    else if (a.nx <= 4 * NT)
        launch_timed(c, rfft_kernel<T, true, false>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, x, spec);
    else launch_timed(c, rfft_kernel<T, false, false>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, x, spec);
    return hipGetLastError();
}

template <int KIND, int R, int NS, typename T>
void jac_rfft_launch(vv_ctx *c, const FftArgs<T> &a, const JacRows &jr, const T *u, cx_t<T> *spec)
{
    const size_t fft = (size_t)(R >> 1) * a.nx * sizeof(cx_t<T>);
    const size_t strip = 2 * (size_t)(R + 4) * (NT + 4) * sizeof(T);
    const dim3 grid((a.nv + R - 1) / R, c->B);
    launch_timed(c, jac_rfft_kernel<KIND, R, NS, T>, grid, dim3(NT), (uint32_t)(fft > strip ? fft : strip), a, jr, u,
                 spec);
}

template <int KIND, int R, typename T>
void jac_rfft_width(vv_ctx *c, const FftArgs<T> &a, const JacRows &jr, const T *u, cx_t<T> *spec)
{
    if (a.nx == 4 * NT) jac_rfft_launch<KIND, R, 4>(c, a, jr, u, spec);
    else if (a.nx == 2 * NT) jac_rfft_launch<KIND, R, 2>(c, a, jr, u, spec);
    else jac_rfft_launch<KIND, R, 1>(c, a, jr, u, spec);
}

template <int R, typename T>
void jac_rfft_kind(vv_ctx *c, const FftArgs<T> &a, const JacRows &jr, const T *u, cx_t<T> *spec)
{
    switch (c->p.bracket) {
    case VV_BRACKET_J1: jac_rfft_width<VV_BRACKET_J1, R>(c, a, jr, u, spec); break;
    case VV_BRACKET_J2: jac_rfft_width<VV_BRACKET_J2, R>(c, a, jr, u, spec); break;
    default: jac_rfft_width<VV_BRACKET_J4, R>(c, a, jr, u, spec); break;
    }
}

template <typename T>
hipError_t jac_rfft_t(vv_ctx *c, const cx_t<T> *tw, const T *u, const double *uscale, cx_t<T> *spec,
                      const int32_t *done)
{
    const int R = jac_rfft_rows(c);
    const FftArgs<T> a{c->fft_n, c->logn, c->fft_rows, R, c->N, spec_stride(c), (int64_t)sizeof(GmresState), tw, done};
    const JacRows jr{uscale, c->h0, c->h1p, c->h1h, c->h2p, c->h2h, c->inst, c->ht_inv, 0.5};
    switch (R) {
    case 2: jac_rfft_kind<2>(c, a, jr, u, spec); break;
    case 4: jac_rfft_kind<4>(c, a, jr, u, spec); break;
    case 8: jac_rfft_kind<8>(c, a, jr, u, spec); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t irfft_t(vv_ctx *c, const cx_t<T> *tw, const cx_t<T> *spec, T *y, const int32_t *done)
{
    if (c->fft_n & (c->fft_n - 1)) return irdft_t<T>(c, tw, spec, y, done);
    FftArgs<T> a{c->fft_n, c->logn, c->fft_rows, fft_rows(c), c->N, spec_stride(c),
                 (int64_t)sizeof(GmresState), tw, done};
    const int nb = (a.nv + a.rows - 1) / a.rows;
    const size_t lds = (size_t)(a.rows >> 1) * a.nx * sizeof(cx_t<T>);
    if (a.nx == N8) launch_timed(c, irfft_kernel<T, true, true>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, spec, y);
    else if (a.nx <= 4 * NT)
        launch_timed(c, irfft_kernel<T, true, false>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, spec, y);
    else launch_timed(c, irfft_kernel<T, false, false>, dim3(nb, c->B), dim3(NT), (uint32_t)lds, a, spec, y);
    return hipGetLastError();
}

// 16-byte vectors per thread of the slice one irFFT workgroup covers when the
// fused irFFT + dots kernel applies (power-of-two nx <= 4 NT, the kinetic
// form, full workgroups, R nx = NT EPT L with EPT in 1, 2, 4, 8), else 0
template <typename T>
int irfft_mdot_ept(const vv_ctx *c)
{
    const int nx = c->fft_n;
    constexpr int L = V16<T>::L;
    if (!c->fuse || (nx & (nx - 1)) || nx > 4 * NT || c->tb != nullptr) return 0;
    const int R = fft_rows(c);
    if (c->fft_rows % R || c->N % L) return 0;
    const int64_t pts = (int64_t)R * nx;
    if (pts % (NT * L)) return 0;
    const int64_t ept = pts / (NT * L);
    return (ept == 1 || ept == 2 || ept == 4 || ept == 8) ? (int)ept : 0;
}

template <typename T, int EPT>
void irfft_mdot_launch(vv_ctx *c, const FftArgs<T> &a, dim3 grid, size_t lds, const cx_t<T> *spec,
                       const CgsArgs<T> &g, int col)
{
    const bool p8 = a.nx == N8;
    if (c->p.restart < 32) {
        if (p8) launch_timed(c, irfft_mdot_kernel<T, EPT, 32, true>, grid, dim3(NT), (uint32_t)lds, a, spec, g, col);
        else launch_timed(c, irfft_mdot_kernel<T, EPT, 32, false>, grid, dim3(NT), (uint32_t)lds, a, spec, g, col);
    } else {
        if (p8) launch_timed(c, irfft_mdot_kernel<T, EPT, 64, true>, grid, dim3(NT), (uint32_t)lds, a, spec, g, col);
        else launch_timed(c, irfft_mdot_kernel<T, EPT, 64, false>, grid, dim3(NT), (uint32_t)lds, a, spec, g, col);
    }
}

template <typename T>
hipError_t irfft_mdot_t(vv_ctx *c, const cx_t<T> *tw, const cx_t<T> *spec, const T *basis, int col,
                        const int32_t *done)
{
    const int ept = irfft_mdot_ept<T>(c);
    FftArgs<T> a{c->fft_n, c->logn, c->fft_rows, fft_rows(c), c->N, spec_stride(c),
                 (int64_t)sizeof(GmresState), tw, done};
    const int nb = a.nv / a.rows;
    const size_t lds = (size_t)(a.rows >> 1) * a.nx * sizeof(cx_t<T>);
    const CgsArgs<T> g{basis, c->N, c->vs, c->partial, c->part_stride, c->gs, 0, nb};
    const dim3 grid(nb, c->B);
    switch (ept) {
    case 8: irfft_mdot_launch<T, 8>(c, a, grid, lds, spec, g, col); break;
    case 4: irfft_mdot_launch<T, 4>(c, a, grid, lds, spec, g, col); break;
    case 2: irfft_mdot_launch<T, 2>(c, a, grid, lds, spec, g, col); break;
    case 1: irfft_mdot_launch<T, 1>(c, a, grid, lds, spec, g, col); break;
    default: return hipErrorInvalidValue;
    }
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


template <typename T>
const cx_t<T> *lu_factors(const vv_ctx *c);
template <> const double2 *lu_factors<double>(const vv_ctx *c) { return c->lu64; }
template <> const float2 *lu_factors<float>(const vv_ctx *c) { return c->lu32; }

template <typename T>
hipError_t tridiag_lu_t(vv_ctx *c, cx_t<T> *spec, const int32_t *done)
{
    TriArgs a{c->fft_rows, NT, spec_stride(c), (int64_t)sizeof(GmresState), c->eigen, c->gm, c->g0, c->gp,
              c->coef_stride, c->cf_lo, c->cf_mid, c->cf_hi, c->inst, c->ht_inv, done};
    const int M = lu_rows(c->fft_rows);
    const size_t lds = (size_t)NT * (M + 1) * sizeof(cx_t<T>);
    const cx_t<T> *rp = lu_factors<T>(c);
    dim3 grid(c->nmodes, c->B), block(NT);
    switch (M) {
    case 1: launch_timed(c, tridiag_lu_kernel<1, T>, grid, block, (uint32_t)lds, a, rp, spec); break;
    case 2: launch_timed(c, tridiag_lu_kernel<2, T>, grid, block, (uint32_t)lds, a, rp, spec); break;
    case 4: launch_timed(c, tridiag_lu_kernel<4, T>, grid, block, (uint32_t)lds, a, rp, spec); break;
    case 8: launch_timed(c, tridiag_lu_kernel<8, T>, grid, block, (uint32_t)lds, a, rp, spec); break;
    default: launch_timed(c, tridiag_lu_kernel<16, T>, grid, block, (uint32_t)lds, a, rp, spec); break;
    }
    return hipGetLastError();
}

template <typename T, bool TR = false>
hipError_t tridiag_t(vv_ctx *c, cx_t<T> *spec, const int32_t *done)
{
    if (!TR && lu_factors<T>(c) != nullptr) return tridiag_lu_t<T>(c, spec, done);
    int P, M;
    tridiag_shape(c->fft_rows, &P, &M);
    TriArgs a{c->fft_rows, P, spec_stride(c), (int64_t)sizeof(GmresState), c->eigen, c->gm, c->g0, c->gp,
              c->coef_stride, c->cf_lo, c->cf_mid, c->cf_hi, c->inst, c->ht_inv, done};
    const size_t lds = (size_t)P * (M + 1) * sizeof(cx_t<T>);
    dim3 grid(c->nmodes, c->B), block(NT);
    switch (M) {
    case 2: launch_timed(c, tridiag_kernel<2, T, TR>, grid, block, (uint32_t)lds, a, spec); break;
    case 4: launch_timed(c, tridiag_kernel<4, T, TR>, grid, block, (uint32_t)lds, a, spec); break;
    case 8: launch_timed(c, tridiag_kernel<8, T, TR>, grid, block, (uint32_t)lds, a, spec); break;
    default: launch_timed(c, tridiag_kernel<16, T, TR>, grid, block, (uint32_t)lds, a, spec); break;
    }
    return hipGetLastError();
}

}  // namespace

// rows per thread of the pivoted solve (NT threads per mode), and the pivots
// stored per mode (NT x rows per thread, chunk-interleaved)
int lu_rows(int nv)
{
    const int m = (nv + NT - 1) / NT;
    return m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : 16;
}

int lu_pivots_per_mode(int nv) { return NT * lu_rows(nv); }

int fft_blocks(vv_ctx *c)
{
    const int R = fft_rows(c);
    return (c->fft_rows + R - 1) / R;
}

hipError_t launch_rfft(vv_ctx *c, const double *x, double2 *spec, const int32_t *done)
{
    return rfft_t<double>(c, c->twiddle, x, spec, done);
}

// Rows per workgroup of the fused J u + rFFT (its own choice: the spectrum's
// layout does not depend on it): the largest of 8, 4, 2 that still leaves
// >= 6 workgroups per CU (fewer halo rows per output row; batched contexts),
// else the FFT kernels' rows when those are 2, 4 or 8, else 2.  Measured at
// 512x1024 with 22 instances: 108 / 82 / 80 us at 2 / 4 / 8 rows; a single
// 1024x2048 instance: 23.4 / 23.2 / 28.9 us.
int jac_rfft_rows(const vv_ctx *c)
{
    const int nx = c->fft_n;
    if (!c->fuse || c->tb != nullptr || nx != c->p.nx || (nx & (nx - 1)) || nx < NT || nx > 4 * NT) return 0;
    for (int R = 8; R >= 2; R >>= 1)
        if ((int64_t)((c->fft_rows + R - 1) / R) * c->B >= 6 * 256) return R;
    const int R = fft_rows(c);
    return (R == 2 || R == 4 || R == 8) ? R : 2;
}

hipError_t launch_jac_rfft(vv_ctx *c, const double *u, const double *uscale, double2 *spec, const int32_t *done)
{
    return jac_rfft_t<double>(c, c->twiddle, u, uscale, spec, done);
}

hipError_t launch_jac_rfft32(vv_ctx *c, const float *u, const double *uscale, float2 *spec, const int32_t *done)
{
    return jac_rfft_t<float>(c, c->twiddle32, u, uscale, spec, done);
}

hipError_t launch_irfft(vv_ctx *c, const double2 *spec, double *y, const int32_t *done)
{
    return irfft_t<double>(c, c->twiddle, spec, y, done);
}

int irfft_mdot_rows(const vv_ctx *c, bool fp32)
{
    return (fp32 ? irfft_mdot_ept<float>(c) : irfft_mdot_ept<double>(c)) > 0 ? c->fft_rows / fft_rows(c) : 0;
}

hipError_t launch_irfft_mdot(vv_ctx *c, const double2 *spec, int col, const int32_t *done)
{
    return irfft_mdot_t<double>(c, c->twiddle, spec, c->basis, col, done);
}

hipError_t launch_irfft_mdot32(vv_ctx *c, const float2 *spec, int col, const int32_t *done)
{
    return irfft_mdot_t<float>(c, c->twiddle32, spec, c->basis32, col, done);
}

hipError_t launch_tridiag_k(vv_ctx *c, double2 *spec, const int32_t *done)
{
    return tridiag_t<double>(c, spec, done);
}

hipError_t launch_tridiag_t(vv_ctx *c, double2 *spec, const int32_t *done)
{
    return tridiag_t<double, true>(c, spec, done);
}

hipError_t launch_rfft32(vv_ctx *c, const float *x, float2 *spec, const int32_t *done)
{
    return rfft_t<float>(c, c->twiddle32, x, spec, done);
}

hipError_t launch_irfft32(vv_ctx *c, const float2 *spec, float *y, const int32_t *done)
{
    return irfft_t<float>(c, c->twiddle32, spec, y, done);
}

hipError_t launch_tridiag32(vv_ctx *c, float2 *spec, const int32_t *done)
{
    return tridiag_t<float>(c, spec, done);
}

hipError_t launch_tridiag_factor(vv_ctx *c, double2 *rp64, float2 *rp32)
{
    TriArgs a{c->fft_rows, NT, spec_stride(c), (int64_t)sizeof(GmresState), c->eigen, c->gm, c->g0, c->gp,
              c->coef_stride, c->cf_lo, c->cf_mid, c->cf_hi, c->inst, c->ht_inv, nullptr};
    dim3 grid((c->nmodes + NT - 1) / NT, c->B);
    const int M = lu_rows(c->fft_rows);
    if (rp64) hipLaunchKernelGGL(tridiag_factor_kernel<double>, grid, dim3(NT), 0, c->stream, a, c->nmodes, M, rp64);
    if (rp32) hipLaunchKernelGGL(tridiag_factor_kernel<float>, grid, dim3(NT), 0, c->stream, a, c->nmodes, M, rp32);
    return hipGetLastError();
}

hipError_t launch_transpose(vv_ctx *c, const double *in, double *out, bool back, const int32_t *done)
{
    return transpose_t<double>(c, in, out, back, done);
}

hipError_t launch_transpose32(vv_ctx *c, const float *in, float *out, bool back, const int32_t *done)
{
    return transpose_t<float>(c, in, out, back, done);
}

hipError_t launch_potential_bands(vv_ctx *c, const double *phi)
{
    hipLaunchKernelGGL(potential_bands_kernel, dim3((c->p.nx + NT - 1) / NT, c->B), dim3(NT), 0, c->stream, phi,
                       c->p.nx, c->gm, c->g0, c->gp);
    return hipGetLastError();
}

}  // namespace vv
