// Internal declarations shared by the HIP translation units of libvv1d.so.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/vv1d.h"

namespace vv {

constexpr int kMaxRestart = 64;     // GMRES restart limit (basis of restart+1 vectors)
constexpr int kMaxNx = 4096;        // largest x size (power of two)
constexpr int kMaxNv = 8192;
constexpr int kReduceBlocks = 512;  // fixed block count of N-sized reductions

// Host-visible part of the GMRES state; copied to pinned host memory after
// each chunk of iterations.
struct GmresStatus {
    int32_t it;        // iterations in the current cycle
    int32_t its;       // total iterations
    int32_t reason;    // KSP converged reason (0 = iterating)
    int32_t done;      // current cycle finished: every Krylov kernel returns at once
    double res;        // current residual estimate
    double rnorm0;     // norm of the right-hand side
};

// Device-resident GMRES state.  Written only by the one-workgroup control
// kernels (cycle start, mdot finish, Hessenberg update, back substitution);
// read by every Krylov kernel.
struct GmresState {
    GmresStatus st;
    double hh[(kMaxRestart + 1) * kMaxRestart];  // rotated Hessenberg, column it at hh[it*(kMaxRestart+1)]
    double cc[kMaxRestart];
    double ss[kMaxRestart];
    double grs[kMaxRestart + 1];
    double scale[kMaxRestart + 1];   // v_j = scale[j] * V_j (V_j stored unnormalised)
    double coef[kMaxRestart + 1];    // coefficients of the pending maxpy
    double ttol, rtol, atol, dtol, haptol;
    int32_t max_it;
    int32_t max_k;
    int32_t first;                   // first cycle: sets rnorm0 and ttol
    int32_t pad;
};

}  // namespace vv

struct vv_ctx {
    vv_params p;
    hipStream_t stream;
    int64_t N;          // nx * nv
    int nmodes;         // nx/2 + 1
    int logn;           // log2(nx)
    // grid-derived constants
    double ht_inv, hx_inv, hv_inv, hx2_inv, af1, af2, afpc;
    // Hamiltonian factors (device, ctx-owned copies)
    double *h0, *h1p, *h1h, *h2p, *h2h;
    // preconditioner data (device)
    double2 *eigen;     // nmodes used (k = 0..nx/2)
    double *gm, *g0, *gp;   // per-row real factors of the bands
    double2 *twiddle;   // exp(-2 pi i m / nx), m = 0..nx-1
    float2 *twiddle32;  // the same, rounded to float (FP32 preconditioner)
    double2 *spec;      // nmodes * nv complex spectrum, mode-major
    // Poisson
    double *pinv;       // first row of the mean-free circulant inverse (nx)
    double *pwork;      // Poisson scratch (3*nx + 1)
    // work vectors (N)
    double *X;          // bracket / residual output before P^-1
    double *B;          // SNES right-hand side P^-1 F
    double *D;          // Newton update
    double *basis;      // (restart+1) * N
    // mixed precision: the FP32 cycle reuses the FP64 basis allocation
    float *basis32;     // (restart+1) * N floats at the start of basis
    double *R;          // FP64 true residual of the refinement loop, after basis32
    // reductions
    double *partial;    // kReduceBlocks * (kMaxRestart+2)
    double *dscalar;    // device scalars
    double *npart;      // density partial sums
    vv::GmresState *gs;
    vv::GmresStatus *host_status;   // pinned, 2 slots
    hipEvent_t ev[2];
    // profiling (per-kernel HIP event pairs)
    int prof;           // bit k set: time kernel kind k
    void *prof_state;
    hipEvent_t prof_a, prof_b;   // events of the next main-kernel launch (null: none)
    std::string err;
};

namespace vv {

// ---- launchers -------------------------------------------------------------
// bracket.hip
hipError_t launch_residual(vv_ctx *c, const double *f, const double *fh, double *y);
hipError_t launch_jacobian(vv_ctx *c, const double *u, const double *uscale, double *y);
hipError_t launch_bracket_add(vv_ctx *c, const double *x, double *y, double factor);
hipError_t launch_jacobian32(vv_ctx *c, const float *u, const double *uscale, float *y);
// spectral.hip
int fft_blocks(vv_ctx *c);
hipError_t launch_rfft(vv_ctx *c, const double *x, double2 *spec, const int32_t *done);
hipError_t launch_tridiag_k(vv_ctx *c, double2 *spec, const int32_t *done);
hipError_t launch_irfft(vv_ctx *c, const double2 *spec, double *y, const int32_t *done);
hipError_t launch_rfft32(vv_ctx *c, const float *x, float2 *spec, const int32_t *done);
hipError_t launch_tridiag32(vv_ctx *c, float2 *spec, const int32_t *done);
hipError_t launch_irfft32(vv_ctx *c, const float2 *spec, float *y, const int32_t *done);
// krylov.hip
hipError_t launch_gmres_init(vv_ctx *c, double rtol, double atol, double dtol, int max_it);
hipError_t launch_norm_start(vv_ctx *c);                  // cycle start on basis[0]
int cgs_blocks(int64_t N);
hipError_t launch_mdot(vv_ctx *c, int col);               // CGS dots of basis[col+1] vs 0..col
hipError_t launch_maxpy_norm(vv_ctx *c, int it);          // CGS update of basis[it+1] + Givens
hipError_t launch_build_solution(vv_ctx *c, double *x);  // x += V y
hipError_t launch_start32(vv_ctx *c, const double *r);    // FP32 cycle start: V_0 = (float) r
hipError_t launch_mdot32(vv_ctx *c, int col);
hipError_t launch_maxpy_norm32(vv_ctx *c, int it);
hipError_t launch_build_solution32(vv_ctx *c, double *x);  // x (FP64) += V y (FP32 basis)
hipError_t launch_residual_restart(vv_ctx *c, const double *b, double *y);  // y = b - y
hipError_t launch_norm2(vv_ctx *c, const double *x, int64_t n, double *out_dev);
hipError_t launch_dot(vv_ctx *c, const double *x, const double *y, int64_t n, double *out_dev);
hipError_t launch_axpy(vv_ctx *c, double alpha, const double *x, double *y, int64_t n);
hipError_t launch_axpy_cond(vv_ctx *c, double alpha, const double *x, double *y, int64_t n);
hipError_t launch_copy(vv_ctx *c, const double *x, double *y, int64_t n);
// moments.hip
hipError_t launch_density_partial(vv_ctx *c, const double *f);
hipError_t launch_density(vv_ctx *c, const double *f, double *n);
hipError_t launch_poisson(vv_ctx *c, const double *n_in, double *n_out, double *phi, double *h1, double *pres,
                          int from_part);
hipError_t launch_poisson_mult(vv_ctx *c, const double *x, const double *n, double *y);
hipError_t launch_pot_to_ham(vv_ctx *c, const double *phi, double *h1);
hipError_t launch_invariants(vv_ctx *c, const double *f, const double *h0, const double *h1,
                             const double *h2, const double *phi, double *out_dev);

// Launch a timed kernel: when the profiler armed c->prof_a / c->prof_b, the
// events are attached to this kernel's own dispatch packet (start / end
// timestamps of the kernel, as a kernel trace reports them).
template <typename K, typename... Args>
inline void launch_timed(vv_ctx *c, K kernel, dim3 grid, dim3 block, uint32_t shmem, Args... args)
{
    if (c->prof_a != nullptr) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, c->stream, c->prof_a, c->prof_b, 0u, args...);
        c->prof_a = c->prof_b = nullptr;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, c->stream, args...);
    }
}

// XCD-aware block order: spread consecutive tiles so that each of the 8 XCDs
// works on one contiguous range of tiles (halo lines stay in that XCD's L2).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

}  // namespace vv
