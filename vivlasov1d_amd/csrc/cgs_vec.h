// Vector access and per-slice dot products of the classical Gram-Schmidt
// kernels, shared by krylov.hip (mdot, maxpy + norm, solution update) and
// spectral.hip (the irFFT fused with the CGS dots of the Arnoldi step).
//
// A CGS launch covers the N elements of a basis vector once: workgroup s owns
// the slice of NT*EPT 16-byte vectors starting at s*NT*EPT, thread t the
// vectors t + q*NT (q < EPT) of it.  The dot of a slice is accumulated per
// thread in (q, lane-element) order without contraction, reduced over the
// wavefront by shuffles and over the NT/64 wavefronts in a fixed order, so a
// kernel that forms the same slice's w by other means (the fused irFFT)
// produces bit-identical partials.
#pragma once

#include "vv1d_internal.h"

namespace vv {
namespace {

constexpr int kCgsNT = 256;
constexpr int kCgsNW = kCgsNT / 64;

// 16-byte vector of the basis storage type: double2 (FP64 cycle) or float4
// (FP32 cycle of the mixed-precision solver).  Arithmetic is FP64 either way.
template <typename S> struct V16;
template <> struct V16<double> {
    using type = double2;
    static constexpr int L = 2;
};
template <> struct V16<float> {
    using type = float4;
    static constexpr int L = 4;
};
__device__ __forceinline__ double vget(const double2 &v, int l) { return l == 0 ? v.x : v.y; }
__device__ __forceinline__ double vget(const float4 &v, int l)
{
    return (double)(l == 0 ? v.x : (l == 1 ? v.y : (l == 2 ? v.z : v.w)));
}
__device__ __forceinline__ void vset(double2 &v, int l, double x)
{
    if (l == 0) v.x = x;
    else v.y = x;
}
__device__ __forceinline__ void vset(float4 &v, int l, double x)
{
    const float f = (float)x;
    if (l == 0) v.x = f;
    else if (l == 1) v.y = f;
    else if (l == 2) v.z = f;
    else v.w = f;
}

// The basis is vector-major across instances: V_j of instance b at
// basis + j * vstride + b * N (vstride = B N rounded up to a multiple of 4
// elements, so every V_j starts on a 32-byte boundary), so V_j of all
// instances is one contiguous field and the operator kernels see instance
// stride N.
template <typename S>
struct CgsArgs {
    const S *basis;
    int64_t N;
    int64_t vstride;
    double *part;
    int64_t pstride;
    GmresState *g;
    int64_t noff;      // maxpy: offset of its norm partials in part (the dot partials stay at 0)
    int nbd;           // maxpy: rows of dot partials (the mdot grid)
};

template <int EPT, int G, typename S>
__device__ __forceinline__ void load_group(const S *basis, int64_t vstride, int v0, int64_t e0, int64_t nvv,
                                           typename V16<S>::type (&x)[G][EPT])
{
    using V = typename V16<S>::type;
#pragma unroll
    for (int gg = 0; gg < G; ++gg) {
        const V *Vp = reinterpret_cast<const V *>(basis + (size_t)(v0 + gg) * vstride);
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int64_t e = e0 + (int64_t)q * kCgsNT;
            x[gg][q] = (e < nvv) ? Vp[e] : V{};
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    return v;   // lane 0
}

// sum_q sum_l w[q]_l (x[q]_l sv): one thread's share of <w, sv V> over a
// slice.  The accumulation is an explicit fused multiply-add, so the result
// does not depend on the contraction setting of the unit that inlines it.
template <int EPT, typename S>
__device__ __forceinline__ double dot_slice(const typename V16<S>::type (&w)[EPT],
                                            const typename V16<S>::type (&x)[EPT], double sv)
{
    constexpr int L = V16<S>::L;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
#pragma unroll
        for (int l = 0; l < L; ++l) acc = __builtin_fma(vget(w[q], l), vget(x[q], l) * sv, acc);
    }
    return acc;
}

}  // namespace
}  // namespace vv
